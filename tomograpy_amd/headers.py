"""FITS-style header geometry for maps and image stacks (host side, numpy).

Restates the header helpers of tomograpy/siddon.py (297-453, 457-572,
703-787) and the observer placement of tomograpy/simu.py (82-137, 158-210):
map borders, rotation matrices, centred map/image headers, field of view and
circular trajectories.  Headers are plain dicts; an image stack carries a list
of per-view dicts (images on the last array axis, siddon.py:51-55).
"""
import copy

import numpy as np

from .infoarray import InfoArray

# fitsarray's BITPIX <-> dtype name tables (siddon.py:20-26).
bitpix = {"8": "int8", "16": "int16", "32": "int32", "-32": "float32", "-64": "float64"}
bitpix_inv = {v: int(k) for k, v in bitpix.items()}

INF = 100000  # siddon.py:82

# Keys of one view record, in the order the native library reads them
# (C_siddon.c.template:1077-1101).
VIEW_KEYS = ("CRPIX1", "CRPIX2", "CDELT1", "CDELT2", "CRVAL1", "CRVAL2",
             "R1_1", "R1_2", "R1_3", "R2_1", "R2_2", "R2_3", "R3_1", "R3_2", "R3_3",
             "M1", "M2", "M3")


def _dtype_name(dtype):
    return np.dtype(dtype).name


def rotation_matrix(lon, lat, rol):
    """Rotation matrix of an observer at (lon, lat) with roll rol (siddon.py:457-496)."""
    cosln, sinln = np.cos(lon), np.sin(lon)
    coslt, sinlt = np.cos(lat), np.sin(lat)
    cosrl, sinrl = np.cos(rol), np.sin(rol)
    R = np.empty((3, 3))
    R[0, 0] = -cosln * coslt
    R[0, 1] = -sinln * cosrl - cosln * sinlt * sinrl
    R[0, 2] = sinln * sinrl - cosln * sinlt * cosrl
    R[1, 0] = -sinln * coslt
    R[1, 1] = cosln * cosrl - sinln * sinlt * sinrl
    R[1, 2] = -cosln * sinrl - sinln * sinlt * cosrl
    R[2, 0] = -sinlt
    R[2, 1] = coslt * sinrl
    R[2, 2] = coslt * cosrl
    return R


def array_to_dict(indict, name, arr, imin=1):
    """Stores a 1-d or 2-d array as keys name{i} / name{i}_{j} (siddon.py:703-725)."""
    arr = np.asarray(arr)
    if arr.ndim == 1:
        for i in range(arr.shape[0]):
            indict[name + str(i + imin)] = arr[i].item()
    elif arr.ndim == 2:
        for i in range(arr.shape[0]):
            for j in range(arr.shape[1]):
                indict[name + "%i_%i" % (i + imin, j + imin)] = arr[i, j].item()
    else:
        raise ValueError("Not implemented for arr.ndim > 2")


def dict_to_array(h, name):
    """Reads keys name{i} / name{i}_{j} back into an array (siddon.py:727-787)."""
    import re
    l = len(name)
    idx1, idx2 = [], []
    for k in h.keys():
        if k[:l] != name:
            continue
        s = k[l:]
        if re.fullmatch(r"[0-9]+_[0-9]+", s):
            i, j = (int(a) for a in s.split("_"))
            idx2.append((i, j))
        elif re.fullmatch(r"[0-9]+", s):
            idx1.append(int(s))
    if idx2:
        imin = min(i for i, _ in idx2)
        jmin = min(j for _, j in idx2)
        imax = max(i for i, _ in idx2)
        jmax = max(j for _, j in idx2)
        arr = np.empty((imax + 1 - imin, jmax + 1 - jmin))
        for i in range(arr.shape[0]):
            for j in range(arr.shape[1]):
                arr[i, j] = h[name + "%i_%i" % (i + imin, j + jmin)]
        return arr
    if idx1:
        imin, imax = min(idx1), max(idx1)
        arr = np.empty(imax + 1 - imin)
        for i in range(arr.shape[0]):
            arr[i] = h[name + str(i + imin)]
        return arr
    raise ValueError("header does not contain this array.")


def header_rotation_matrix(h):
    """Adds R{i}_{j} computed from LON/LAT/ROL (siddon.py:513-529)."""
    R = rotation_matrix(h["LON"], h["LAT"], h["ROL"])
    array_to_dict(h, "R", R)


def full_rotation_matrix(data):
    """Adds the rotation matrix to every view header of a stack (siddon.py:498-511)."""
    for h in data.header:
        header_rotation_matrix(h)


def map_borders(h):
    """PSHAPE = CDELT*NAXIS, MMIN = -CRPIX*CDELT, MMAX = MMIN + PSHAPE (siddon.py:559-572).

    CRVAL is ignored, as in the reference.
    """
    cdelt = np.asarray([h["CDELT%d" % (i + 1)] for i in range(3)], dtype=np.float64)
    crpix = np.asarray([h["CRPIX%d" % (i + 1)] for i in range(3)], dtype=np.float64)
    naxis = np.asarray([h["NAXIS%d" % (i + 1)] for i in range(3)], dtype=np.float64)
    pshape = cdelt * naxis
    mmin = -crpix * cdelt
    mmax = mmin + pshape
    array_to_dict(h, "PSHAPE", pshape)
    array_to_dict(h, "MMIN", mmin)
    array_to_dict(h, "MMAX", mmax)


def _triple(x, what, allowed=(1, 3)):
    x = np.asarray([x] if np.isscalar(x) else x)
    if x.size == 1:
        x = np.asarray(list(x) * 3)
    elif x.size not in allowed:
        raise ValueError("%s should be a 1 or 3 tuple" % what)
    return x


def centered_cubic_map_header(pshape, shape, dtype=np.float64):
    """Header of a map cube centred on the origin (siddon.py:307-346)."""
    shape = _triple(shape, "shape")
    pshape = _triple(pshape, "physical shape")
    header = dict()
    header["SIMPLE"] = True
    header["BITPIX"] = bitpix_inv[_dtype_name(dtype)]
    header["NAXIS"] = 3
    for i in range(3):
        header["NAXIS%d" % (i + 1)] = int(shape[i])
        header["CRPIX%d" % (i + 1)] = shape[i] / 2.0
        header["CDELT%d" % (i + 1)] = pshape[i] / float(shape[i])
        header["CRVAL%d" % (i + 1)] = 0.0
    map_borders(header)
    return header


def centered_image_header(pshape, shape, dtype=np.float64):
    """Header of a centred image (siddon.py:368-405)."""
    shape = np.asarray([shape] if np.isscalar(shape) else shape)
    if shape.size == 1:
        shape = np.asarray(list(shape) * 2)
    elif shape.size > 2:
        raise ValueError("shape should be a 1 or 2 tuple")
    pshape = np.asarray([pshape] if np.isscalar(pshape) else pshape)
    if pshape.size == 1:
        pshape = np.asarray(list(pshape) * 2)
    elif pshape.size > 2:
        raise ValueError("physical shape should be a 1 or 2 tuple")
    header = dict()
    header["SIMPLE"] = True
    header["BITPIX"] = bitpix_inv[_dtype_name(dtype)]
    header["NAXIS"] = 2
    for i in range(2):
        header["NAXIS%d" % (i + 1)] = int(shape[i])
        header["CRPIX%d" % (i + 1)] = shape[i] / 2.0
        header["CDELT%d" % (i + 1)] = pshape[i] / float(shape[i])
        header["CRVAL%d" % (i + 1)] = 0.0
    return header


def fov(obj, radius):
    """Half-angle field of view that encloses a map seen from `radius` (siddon.py:440-453)."""
    h = obj.header if hasattr(obj, "header") else obj
    pshape = [h["NAXIS%d" % (i + 1)] * h["CDELT%d" % (i + 1)] for i in range(3)]
    pdiag = np.sqrt(np.sum(np.asarray(pshape) ** 2))
    return np.arctan2(pdiag, radius)


def update_observer(h, lon=None, lat=None, d=None):
    """Sets LON/LAT/D and the observer position M1..3 (simu.py:115-128)."""
    if lon is not None:
        h["LON"] = lon
    if lat is not None:
        h["LAT"] = lat
    if d is not None:
        h["D"] = d
    lon, lat, d = h["LON"], h["LAT"], h["D"]
    h["M1"] = d * np.cos(lat) * np.cos(lon)
    h["M2"] = d * np.cos(lat) * np.sin(lon)
    h["M3"] = d * np.sin(lat)


default_image_dict = {"NAXIS": 2, "NAXIS1": 1, "NAXIS2": 1, "BITPIX": -64, "SIMPLE": True,
                      "CRPIX1": 0.5, "CRPIX2": 0.5, "CDELT1": 1.0, "CDELT2": 1.0,
                      "LON": 0.0, "LAT": 0.0, "ROL": 0.0, "D": 1.0,
                      "M1": 1.0, "M2": 0.0, "M3": 0.0}  # simu.py:82-86


def image_header(header=None):
    """Completes a view header with the mandatory keywords of simu.Image (simu.py:98-113)."""
    h = copy.copy(default_image_dict) if header is None else dict(header)
    for k, v in default_image_dict.items():
        h.setdefault(k, v)
    return h


def circular_trajectory_header(n_images=1, radius=1.0, min_lon=0.0, max_lon=2 * np.pi,
                               longitudes=None, latitudes=None, **kargs):
    """Per-view headers of a circular trajectory (simu.py:158-210).

    As in the reference the default longitudes are
    ``numpy.linspace(min_lon, max_lon, n_images)``, endpoint included.
    ``longitudes`` / ``latitudes`` override them explicitly.
    """
    if longitudes is None:
        longitudes = np.linspace(min_lon, max_lon, n_images)
    longitudes = np.asarray(longitudes, dtype=np.float64)
    if latitudes is None:
        latitudes = [None] * len(longitudes)
    headers = []
    for lon, lat in zip(longitudes, latitudes):
        h = image_header(kargs)
        update_observer(h, lon=float(lon), lat=None if lat is None else float(lat), d=radius)
        header_rotation_matrix(h)
        headers.append(h)
    return headers


def circular_trajectory_data(**kargs):
    """Zeroed image stack with circular-trajectory headers (simu.py:158-210)."""
    kargs = dict(kargs)
    dtype = kargs.pop("dtype", np.float64)
    headers = circular_trajectory_header(**kargs)
    shape = (int(headers[0]["NAXIS1"]), int(headers[0]["NAXIS2"]), len(headers))
    return InfoArray(np.zeros(shape, dtype=dtype), header=headers)


def dataarray_from_header(header):
    """Zeroed image stack for a list of view headers (siddon.py:297-305)."""
    shape = [int(header[0]["NAXIS%d" % (i + 1)]) for i in range(int(header[0]["NAXIS"]))]
    shape.append(len(header))
    dtype = bitpix[str(int(header[0]["BITPIX"]))]
    return InfoArray(np.zeros(shape, dtype=dtype), header=list(header))


def fitsarray_from_header(header):
    """Zeroed array described by a FITS-like header (the fitsarray helper)."""
    n = int(header["NAXIS"])
    shape = [int(header["NAXIS%d" % (i + 1)]) for i in range(n)]
    dtype = bitpix[str(int(header.get("BITPIX", -64)))]
    return InfoArray(np.zeros(shape, dtype=dtype), header=dict(header))


def centered_cubic_map(pshape, shape, fill=0.0, dtype=np.float64):
    """Map cube with a centred header (siddon.py:348-366)."""
    header = centered_cubic_map_header(pshape, shape, dtype=dtype)
    m = fitsarray_from_header(header)
    m[:] = fill
    return m


def centered_image(pshape, shape, fill=0.0, dtype=np.float64):
    """Image with a centred header (siddon.py:407-423)."""
    header = centered_image_header(pshape, shape, dtype=dtype)
    im = fitsarray_from_header(header)
    im[:] = fill
    return im


def centered_stack(pshape, shape, n_images=1, radius=1.0, min_lon=0.0, max_lon=np.pi,
                   fill=0.0, dtype=np.float64):
    """Image stack with centred images on a circular trajectory (siddon.py:425-438)."""
    header = centered_image_header(pshape, shape, dtype=dtype)
    data = circular_trajectory_data(n_images=n_images, radius=radius, min_lon=min_lon,
                                    max_lon=max_lon, dtype=dtype, **header)
    data[:] = fill
    return data


def object_from_header(header, **kwargs):
    """Map cube described by a header, filled with `fill` (simu.py:212-227)."""
    header = dict(header)
    try:
        dtype = bitpix[str(header["BITPIX"])]
    except KeyError:
        dtype = "float64"
    dtype = header.pop("dtype", dtype)
    dtype = kwargs.pop("dtype", dtype)
    fill = kwargs.pop("fill", 0.0)
    map_borders(header)
    shape = (int(header["NAXIS1"]), int(header["NAXIS2"]), int(header["NAXIS3"]))
    obj = InfoArray(np.zeros(shape, dtype=dtype), header=header)
    obj[:] = fill
    return obj


def complete_view_headers(data_header):
    """Converts view headers to dicts and adds R{i}_{j} where missing (siddon.py:252-262)."""
    if not isinstance(data_header, list):
        raise ValueError("data.header should be a list.")
    for i in range(len(data_header)):
        if not isinstance(data_header[i], dict):
            data_header[i] = dict(data_header[i])
        if "R1_1" not in data_header[i]:
            header_rotation_matrix(data_header[i])
    return data_header


def check_headers(data_header, cube_header):
    """Completes headers as check_projector_inputs does (siddon.py:245-270).

    Returns (list of view dicts, map dict); adds R{i}_{j} when missing and the
    map borders when MMAX1 is missing.
    """
    complete_view_headers(data_header)
    if "MMAX1" not in cube_header:
        map_borders(cube_header)
    view_records(data_header)
    map_record(cube_header)
    return data_header, cube_header


MAP_KEYS = [p + str(i) for p in ("CDELT", "PSHAPE", "MMIN", "MMAX") for i in (1, 2, 3)]


def _header_values(h, keys, what):
    """The values of `keys` in header dict h as float64, checked before any
    device work.  The reference reads each key with PyDict_GetItemString and
    PyFloat_AsDouble (C_siddon.c.template:1051-1056), which fails obscurely on a
    missing key, and a NaN in the geometry makes its voxel() cascade
    (C:811-912) take no branch and loop forever.  Here a missing key raises
    KeyError and a non-finite value ValueError, naming the key."""
    out = np.empty(len(keys), dtype=np.float64)
    for k, key in enumerate(keys):
        if key not in h:
            raise KeyError("%s header has no %s" % (what, key))
        out[k] = float(h[key])
    bad = ~np.isfinite(out)
    if bad.any():
        k = int(np.argmax(bad))
        raise ValueError("%s header: %s = %r is not finite" % (what, keys[k], h[keys[k]]))
    return out


def view_records(data_header):
    """(nv, 18) float64 array of the values the projector reads per view
    (C_siddon.c.template:1077-1101); raises on missing or non-finite values,
    as _header_values does, naming the first such view and key."""
    try:
        rec = np.array([[float(h[key]) for key in VIEW_KEYS] for h in data_header],
                       dtype=np.float64).reshape(len(data_header), len(VIEW_KEYS))
    except KeyError:
        rec = None
    if rec is None or not np.isfinite(rec).all():
        # the slow path names the offending view and key
        for t, h in enumerate(data_header):
            _header_values(h, VIEW_KEYS, "view %d" % t)
    return rec


def map_record(cube_header):
    """(12,) float64: CDELT1-3, PSHAPE1-3, MMIN1-3, MMAX1-3 (C_siddon.c.template:1058-1075);
    raises on missing or non-finite values."""
    return _header_values(cube_header, MAP_KEYS, "map")


def pixel_angles(h, n1, n2):
    """Physical angles (gamma_i, lambda_j) of the pixel grid, in float64.

    The projector's convention (C_siddon.c.template:1043-1046): pixel index
    i -> (i - CRPIX + 1) * CDELT + CRVAL.
    """
    g = (np.arange(n1) - h["CRPIX1"] + 1) * h["CDELT1"] + h.get("CRVAL1", 0.0)
    l = (np.arange(n2) - h["CRPIX2"] + 1) * h["CDELT2"] + h.get("CRVAL2", 0.0)
    return g, l


def map_axes(cube_header, convention="fits"):
    """Coordinates of the voxel grid along each axis.

    ``"fits"``: (i + 1 - CRPIX) * CDELT + CRVAL, the FITS pixel convention the
    projector uses for image pixels.  ``"center"``: MMIN + (i + 1/2) * CDELT,
    the centre of the voxel the Siddon traversal assigns to index i.
    """
    axes = []
    for k in range(3):
        n = int(cube_header["NAXIS%d" % (k + 1)])
        cdelt = cube_header["CDELT%d" % (k + 1)]
        if convention == "fits":
            axes.append((np.arange(n) + 1 - cube_header["CRPIX%d" % (k + 1)]) * cdelt
                        + cube_header.get("CRVAL%d" % (k + 1), 0.0))
        elif convention == "center":
            mmin = -cube_header["CRPIX%d" % (k + 1)] * cdelt
            axes.append(mmin + (np.arange(n) + 0.5) * cdelt)
        else:
            raise ValueError("convention must be 'fits' or 'center'")
    return axes
