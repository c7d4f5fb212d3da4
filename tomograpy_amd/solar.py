"""Solar data front end: FITS images and headers -> the view descriptors of
the projector (tomograpy/solar.py:16-392).

``update_header`` turns an instrument header (Carrington longitude/latitude
of the observer, its distance, the roll from the PC matrix or CROTA2, angular
units) into the LON/LAT/ROL/D/M keys the projector reads; ``read_data``
reads a directory of FITS images into an (n1, n2, n_images) InfoArray with
one updated header per image.  ``pyfits`` is not available here, so a minimal
reader/writer of primary FITS HDUs (the only HDU the reference reads,
solar.py:35) is included.

Times (DATE_OBS) are converted as UTC seconds since 1970.  The reference uses
``time.mktime``, i.e. the host's local time zone; differences between two
times, which is what sorting, filtering and grouping use, are the same except
across a daylight-saving change.
"""
import calendar
import copy
import os
import time

import numpy as np

from .infoarray import InfoArray

# constants (solar.py:11-13)
solar_radius = 695000000.0  # m
arcsecond_to_radian = np.pi / 648000


def arcsec2radians(val):
    return val * arcsecond_to_radian


def no_conversion(val):
    return val


cunit_to_conversion = {"arcsec": arcsec2radians, "degree": np.radians,
                       None: no_conversion, "": no_conversion}


def convert_time(time_str):
    """'YYYY-MM-DDTHH:MM:SS[.fff][Z]' -> seconds (solar.py:126-143)."""
    s = time_str.rstrip("Z").strip()
    dpos = s.rfind(".")
    frac = 0.0
    if dpos != -1:
        s, frac = s[:dpos], float(s[dpos:])
    return calendar.timegm(time.strptime(s, "%Y-%m-%dT%H:%M:%S")) + frac


def update_header(header):
    """Adds LON, LAT, ROL, D, M1-3 and TIME and converts CDELT/CRVAL to radians
    (solar.py:59-115).  `header` is a dict; it is updated and returned."""
    h = header
    lon = np.radians(h["CRLN_OBS"])
    lat = np.radians(h["CRLT_OBS"])
    if all(k in h for k in ("PC2_1", "PC1_1", "CDELT1", "CDELT2")):
        rol = np.arctan2(h["PC2_1"] * h["CDELT2"], h["PC1_1"] * h["CDELT1"])
    else:
        rol = np.radians(h["CROTA2"])
    if "DSUN_OBS" in h:
        d = h["DSUN_OBS"] / solar_radius
    else:
        x = h["HAEX_OBS"] if "HAEX_OBS" in h else h["HAEX"]
        y = h["HAEY_OBS"] if "HAEY_OBS" in h else h["HAEY"]
        z = h["HAEZ_OBS"] if "HAEZ_OBS" in h else h["HAEZ"]
        d = np.sqrt(x ** 2 + y ** 2 + z ** 2) / solar_radius
    h["LON"] = float(lon)
    h["LAT"] = float(lat)
    h["ROL"] = float(rol)
    h["D"] = float(d)
    h["M1"] = float(d * np.cos(lat) * np.cos(lon))
    h["M2"] = float(d * np.cos(lat) * np.sin(lon))
    h["M3"] = float(d * np.sin(lat))
    for ax in ("1", "2"):
        conv = cunit_to_conversion.get(h.get("CUNIT" + ax), no_conversion)
        h["CDELT" + ax] = float(conv(h["CDELT" + ax]))
        h["CRVAL" + ax] = float(conv(h["CRVAL" + ax]))
    h["TIME"] = convert_time(h["DATE_OBS"])
    return h


def filter_headers(headers, instrume=None, telescop=None, time_window=None, time_step=None,
                   tmin=None, tmax=None):
    """Indices of the headers kept by solar.py:145-199, in time order: instrument
    and telescope names, a time window (tmin/tmax take precedence) and a
    minimal time step between consecutive kept images."""
    order = sorted(range(len(headers)), key=lambda k: convert_time(headers[k]["DATE_OBS"]))
    time_min = time_max = None
    if time_window is not None:
        time_min, time_max = convert_time(time_window[0]), convert_time(time_window[1])
        if time_min > time_max:
            time_min, time_max = time_max, time_min
    if tmin is not None:
        time_min = convert_time(tmin)
    if tmax is not None:
        time_max = convert_time(tmax)
    step = None
    if time_step is not None:
        step = convert_time(time_step) if isinstance(time_step, str) else float(time_step)
    kept = []
    for k in order:
        h = headers[k]
        if instrume is not None and h["INSTRUME"] not in instrume:
            continue
        if telescop is not None and h["TELESCOP"] not in telescop:
            continue
        t = convert_time(h["DATE_OBS"])
        if time_min is not None and t < time_min:
            continue
        if time_max is not None and t > time_max:
            continue
        if step is not None and kept and abs(convert_time(headers[kept[-1]]["DATE_OBS"]) - t) < step:
            continue
        kept.append(k)
    return kept


# ---- minimal FITS primary HDU ------------------------------------------------------------
_BITPIX = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}
_BITPIX_OF = {np.dtype(v).newbyteorder("="): k for k, v in _BITPIX.items()}


def _card_value(text):
    v = text.split("/", 1)[0].strip() if not text.lstrip().startswith("'") else text.strip()
    if v.startswith("'"):
        end = v.find("'", 1)
        while end != -1 and end + 1 < len(v) and v[end + 1] == "'":  # doubled quote
            end = v.find("'", end + 2)
        return v[1:end].replace("''", "'").rstrip()
    if v in ("T", "F"):
        return v == "T"
    try:
        return int(v)
    except ValueError:
        return float(v.replace("D", "E"))


def read_fits(path):
    """(data, header) of the primary HDU; data in numpy axis order (NAXISn .. NAXIS1)."""
    with open(path, "rb") as f:
        header = {}
        done = False
        nblocks = 0
        while not done:
            block = f.read(2880)
            if len(block) < 2880:
                raise ValueError("%s: truncated FITS header" % path)
            nblocks += 1
            for c in range(36):
                card = block[80 * c: 80 * (c + 1)].decode("ascii")
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= ":
                    header[key] = _card_value(card[10:])
        bitpix = int(header["BITPIX"])
        shape = tuple(int(header["NAXIS%d" % (k + 1)]) for k in range(int(header["NAXIS"])))[::-1]
        count = int(np.prod(shape)) if shape else 0
        raw = np.frombuffer(f.read(count * abs(bitpix) // 8), dtype=_BITPIX[bitpix]).reshape(shape)
    data = raw.astype(raw.dtype.newbyteorder("="))
    if "BSCALE" in header or "BZERO" in header:
        data = data * header.get("BSCALE", 1.0) + header.get("BZERO", 0.0)
    return data, header


def _card(key, value):
    if isinstance(value, bool):
        v = "%20s" % ("T" if value else "F")
    elif isinstance(value, (int, np.integer)):
        v = "%20d" % value
    elif isinstance(value, (float, np.floating)):
        v = "%20s" % repr(float(value)).upper()
    else:
        v = "'%-8s'" % str(value).replace("'", "''")
    return ("%-8s= %s" % (key[:8], v))[:80].ljust(80)


def write_fits(path, data, header=None):
    """Writes `data` (numpy axis order) as a primary HDU with the extra header keys."""
    data = np.asarray(data)
    bitpix = _BITPIX_OF[data.dtype.newbyteorder("=")]
    cards = [_card("SIMPLE", True), _card("BITPIX", bitpix), _card("NAXIS", data.ndim)]
    cards += [_card("NAXIS%d" % (k + 1), n) for k, n in enumerate(data.shape[::-1])]
    skip = {"SIMPLE", "BITPIX", "NAXIS", "END", "BSCALE", "BZERO"}
    for k, v in (header or {}).items():
        if k.upper() in skip or (k.upper().startswith("NAXIS") and k[5:].isdigit()):
            continue
        cards.append(_card(k.upper(), v))
    cards.append("END".ljust(80))
    text = "".join(cards)
    text += " " * (-len(text) % 2880)
    body = data.astype(_BITPIX[bitpix]).tobytes()
    body += b"\0" * (-len(body) % 2880)
    with open(path, "wb") as f:
        f.write(text.encode("ascii"))
        f.write(body)


def bin_image(image, header, factor):
    """Block mean by `factor` along both image axes, with CDELT x factor and
    CRPIX moved so that it keeps naming the same physical point.  (The
    reference delegates this to fitsarray's bin, which is not vendored: this
    convention is parity unpinned.)"""
    f = int(factor)
    n1, n2 = image.shape[0] // f, image.shape[1] // f
    out = image[: n1 * f, : n2 * f].reshape(n1, f, n2, f).mean(axis=(1, 3))
    h = dict(header)
    for ax, n in (("1", n1), ("2", n2)):
        h["CDELT" + ax] = h["CDELT" + ax] * f
        h["CRPIX" + ax] = (h["CRPIX" + ax] - 0.5) / f + 0.5
        h["NAXIS" + ax] = n
    return out, h


def read_data(path, dtype=np.float64, bin_factor=None, **filters):
    """Reads every FITS file of a directory (solar.py:16-57): filters and sorts
    them (filter_headers), transposes each image to (NAXIS1, NAXIS2), updates
    its header and stacks them as an (n1, n2, n_images) InfoArray."""
    if not os.path.isdir(path):
        raise ValueError("Directory does not exist")
    images, headers = [], []
    for fname in sorted(os.listdir(path)):
        try:
            d, h = read_fits(os.path.join(path, fname))
        except (OSError, ValueError, KeyError, UnicodeDecodeError):
            continue
        images.append(d)
        headers.append(h)
    keep = filter_headers(headers, **filters)
    if not keep:
        return None
    stack, hs = [], []
    for k in keep:
        img, h = images[k].T, dict(headers[k])
        if bin_factor is not None:
            img, h = bin_image(img, h, bin_factor)
        update_header(h)
        h["BITPIX"] = -32 if np.dtype(dtype) == np.float32 else -64
        stack.append(np.asarray(img, dtype=dtype))
        hs.append(h)
    return InfoArray(np.stack(stack, axis=-1), header=hs)


# ---- image stacks in time (solar.py:336-392) ----------------------------------------------
def get_times(data):
    return [convert_time(h["DATE_OBS"]) for h in data.header]


def slice_data(data, s):
    out = InfoArray(np.asarray(data)[:, :, s], header=copy.copy(list(data.header)[s]))
    return out


def concatenate(data_list):
    header = []
    for d in data_list:
        header += list(d.header)
    return InfoArray(np.concatenate([np.asarray(d) for d in data_list], axis=-1), header=header)


def sort_data_array(data):
    ind = np.argsort(get_times(data), kind="stable")
    return concatenate([slice_data(data, slice(i, i + 1)) for i in ind])


def temporal_groups_indexes(data, dt_min):
    """Indexes i where image i+1 follows image i by less than dt_min
    (solar.py:368-372, the start indexes temporal_groups slices from)."""
    return list(np.where(np.diff(get_times(data)) < dt_min)[0])


def temporal_groups(data, dt_min):
    """solar.py:381-388: slices of the stack between consecutive group indexes."""
    ind1 = temporal_groups_indexes(data, dt_min)
    ind2 = ind1[1:] + [None]
    return [slice_data(data, slice(i, j, None)) for i, j in zip(ind1, ind2)]
