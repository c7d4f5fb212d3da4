"""Solar tomography models around the projector pair.

Restates the parts of tomograpy/models.py and tomograpy/solar.py that shape
the hot path:
  * Thomson-scattering (pB) coefficients, models.py:237-309 (Billings 1966):
    ``thomson_lo`` = diag(rho^2) . P_sun . diag(f(r)) (models.py:223-235),
    fused into the kernels as pixel and voxel weights;
  * data and map masks by radius, solar.py:211-334;
  * the srt / stsrt / thomson model builders, models.py:19-221, returning the
    device operator, the regularisation spec for ``cg.reconstruct`` and the
    masks.  ``lo``'s operator algebra is not vendored: the smoothness priors
    lo.diff(axis=k) are the forward differences ``cg.NormalOperator`` applies.

Coordinates: the reference takes pixel and voxel coordinates from
fitsarray.axes() (not vendored); here they follow the FITS convention
(i + 1 - CRPIX) * CDELT + CRVAL, the one the projector itself uses for image
pixels (C_siddon.c.template:1043-1046).  ``convention="center"`` uses voxel
centres MMIN + (i + 1/2) CDELT instead.
"""
import numpy as np

from .headers import map_axes, pixel_angles
from .lo_wrapper import Siddon, Siddon4d

sigma = 1.0  # models.py:17


def billings_ab(r):
    """Billings' geometric factors A(r) and B(r) of Thomson scattering by an
    electron r solar radii from the Sun's centre (Billings 1966, "A Guide to
    the Solar Corona", ch. 6; the C1 and C2 of models.py:264-275).  With
    s = sin(Omega) = 1/r, Omega the Sun's angular radius seen from the
    electron, and c = cos(Omega) = sqrt(1 - s^2):

        A = c s^2
        B = -(1 - 3 s^2 - (1 + 3 s^2) (c^2 / s) ln((1 + s) / c)) / 8

    Both are NaN inside the Sun (r < 1) and at r = 0; the map coefficient
    zeroes them there, as models.py:307-308 does.
    """
    r = np.asarray(r, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = 1.0 / r
        s2 = s * s
        c = np.sqrt(1.0 - s2)
        a = c * s2
        log_term = (c * c / s) * np.log((1.0 + s) / c)
        b = -0.125 * (1.0 - 3.0 * s2 - (1.0 + 3.0 * s2) * log_term)
    return a, b


def impact_parameter_sq(alpha, beta, d):
    """Squared distance from the Sun's centre to the line of sight of the pixel
    at angles (alpha, beta) from an observer at distance d, at least 1: the
    line makes the angle arccos(cos alpha cos beta) with the Sun-observer line,
    so rho^2 = d^2 (1 - (cos alpha cos beta)^2) (models.py:244-249, whose rho is
    clamped below at 1 before squaring at models.py:289)."""
    t = np.cos(alpha) * np.cos(beta)
    rho2 = np.asarray(d * d * (1.0 - t * t), dtype=np.float64)
    return np.maximum(rho2, 1.0)


def pb_data_coef(data_header, n1, n2):
    """rho^2 per pixel, shape (n1, n2, n_images) (models.py:277-289)."""
    out = np.zeros((n1, n2, len(data_header)))
    for t, h in enumerate(data_header):
        a, b = pixel_angles(h, n1, n2)
        A, B = np.meshgrid(a, b, indexing="ij")
        out[..., t] = impact_parameter_sq(A, B, h["D"])
    return out


def pb_map_coef(cube_header, u=0.5, convention="fits"):
    """pi sigma / 2 ((1-u) C1 + u C2) / r^4 per voxel, non-finite -> 0 (models.py:291-309)."""
    x, y, z = map_axes(cube_header, convention)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    R2 = X * X + Y * Y + Z * Z
    a, b = billings_ab(np.sqrt(R2))
    with np.errstate(divide="ignore", invalid="ignore"):
        coefs = ((1 - u) * a + u * b) / (R2 * R2)
    coefs[~np.isfinite(coefs)] = 0.0
    return coefs * (np.pi * sigma / 2.0)


def map_radius(cube_header, convention="fits"):
    """Distance of every voxel to the Sun centre (solar.py:324-334)."""
    x, y, z = map_axes(cube_header, convention)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    return np.sqrt(X * X + Y * Y + Z * Z)


def define_map_mask(cube_header, obj_rmin=None, obj_rmax=None, obj_cylinder=None,
                    remove_nan=False, cube=None, convention="fits", **kwargs):
    """Boolean voxel mask, True = excluded (solar.py:295-322)."""
    shape = tuple(int(cube_header["NAXIS%d" % (k + 1)]) for k in range(3))
    if remove_nan and cube is not None:
        mask = np.isnan(np.asarray(cube))
    else:
        mask = np.zeros(shape, dtype=bool)
    if obj_rmin is not None or obj_rmax is not None:
        R = map_radius(cube_header, convention)
        if obj_rmin is not None:
            mask[R < obj_rmin] = True
        if obj_rmax is not None:
            mask[R > obj_rmax] = True
    if obj_cylinder is not None:
        x, y, _ = map_axes(cube_header, convention)
        X, Y = np.meshgrid(x, y, indexing="ij")
        mask[(X ** 2 + Y ** 2 > obj_cylinder ** 2)] = True
    return mask


def compute_rsun(data_header):
    """Solar radius in pixels per image along both axes (solar.py:286-293)."""
    r1 = np.array([np.arctan(1.0 / h["D"]) / h["CDELT1"] for h in data_header])
    r2 = np.array([np.arctan(1.0 / h["D"]) / h["CDELT2"] for h in data_header])
    return r1, r2


def distance_to_sun_center(data_header, n1, n2):
    """Distance of every pixel to the disc centre in solar radii (solar.py:254-284)."""
    R = np.zeros((n1, n2, len(data_header)))
    r1, r2 = compute_rsun(data_header)
    for t, h in enumerate(data_header):
        y = (np.arange(n1) - h["CRPIX1"]) / r1[t]
        x = (np.arange(n2) - h["CRPIX2"]) / r2[t]
        Y, X = np.meshgrid(y, x, indexing="ij")
        R[..., t] = np.sqrt(X ** 2 + Y ** 2)
    return R


def radius_data_mask(data_header, n1, n2, data_rmin=None, data_rmax=None, ring=None):
    """The radius part of define_data_mask from headers alone (solar.py:240-247)."""
    mask = np.zeros((n1, n2, len(data_header)), dtype=bool)
    if data_rmin is not None or data_rmax is not None or ring is not None:
        R = distance_to_sun_center(data_header, n1, n2)
        if data_rmin is not None:
            mask[R < data_rmin] = True
        if data_rmax is not None:
            mask[R > data_rmax] = True
        if ring is not None:
            mask[(ring[0] < R) & (R < ring[1])] = True
    return mask


def define_data_mask(data, data_rmin=None, data_rmax=None, ring=None, mask_negative=False,
                     mask_nan=True, **kwargs):
    """Boolean pixel mask, True = skipped (solar.py:211-252)."""
    arr = np.asarray(data.cpu() if hasattr(data, "cpu") else data)
    mask = radius_data_mask(data.header, arr.shape[0], arr.shape[1], data_rmin, data_rmax, ring)
    if mask_negative:
        mask[arr < 0.0] = True
    if mask_nan:
        mask[np.isnan(arr)] = True
    return mask


def thomson_lo(data_header, cube_header, u=0.5, mask=None, convention="fits", **kwargs):
    """pB Thomson operator O . P_sun . M (models.py:223-235) with O, M fused in the kernels."""
    n1 = int(data_header[0]["NAXIS1"])
    n2 = int(data_header[0]["NAXIS2"])
    O = pb_data_coef(data_header, n1, n2)
    M = pb_map_coef(cube_header, u, convention)
    return Siddon(data_header, cube_header, obstacle="sun", mask=mask, pix_weight=O,
                  vox_weight=M, **kwargs)


def smoothness_axes(ndim):
    """The lo.diff priors of smoothness_prior (models.py:85-94): one per axis."""
    return tuple(range(ndim))


class SmoothnessPrior(tuple):
    """The smoothness priors of a model: the axes k of the differences D_k (a
    tuple, as ``cg.solve(axes=...)`` takes them) and, for the height prior,
    the per-voxel weight ``weight`` that each D_k is composed with (None
    without it)."""

    def __new__(cls, axes, weight=None):
        self = super().__new__(cls, tuple(axes))
        self.weight = weight
        return self


def radius_sq_map(cube_header, convention="fits"):
    """Squared distance of every voxel to the Sun's centre: the _radius_map of
    models.py:96-104, which returns X^2 + Y^2 + Z^2 without a square root."""
    x, y, z = map_axes(cube_header, convention)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    return X * X + Y * Y + Z * Z


def smoothness_prior(cube_header, ndim=3, height_prior=False, convention="fits"):
    """Priors D_k = lo.diff(axis=k), k < ndim (models.py:85-94); with
    ``height_prior`` each composed with diag(r^2), r^2 from radius_sq_map
    (models.py:90-93).  The reference's _radius_map unpacks three axes, so a
    height prior on a 4-D map fails there (ValueError); it raises here too."""
    if not height_prior:
        return SmoothnessPrior(range(ndim))
    if ndim != 3:
        raise ValueError("the height prior is defined for 3-D maps (models.py:96-104)")
    return SmoothnessPrior(range(3), radius_sq_map(cube_header, convention))


def time_bins(data_header, dt_min=None):
    """Time bin of every image of a stack sorted by DATE_OBS: a new bin starts
    where an image follows the previous one by dt_min seconds or more, so
    images closer in time than dt_min share a bin (the grouping
    solar.temporal_groups describes, solar.py:381-388).  Without dt_min every
    image has its own bin, as the comment at models.py:125 intends.

    The reference's stsrt counts the short gaps instead (solar.py:368-372 with
    models.py:126-131): images taken at one time give nv - 1 bins for nv
    images, and the last image's bin index lies beyond the cube's time axis.
    """
    from .solar import convert_time
    nv = len(data_header)
    if dt_min is None:
        return np.arange(nv, dtype=np.int32)
    times = np.array([convert_time(h["DATE_OBS"]) for h in data_header], dtype=np.float64)
    if np.any(np.diff(times) < 0):
        raise ValueError("stsrt expects images sorted by DATE_OBS (solar.sort_data_array)")
    starts = np.concatenate([[0], (np.diff(times) >= dt_min).astype(np.int32)])
    return np.cumsum(starts).astype(np.int32)


def _object_mask(cube, kwargs):
    if kwargs.get("obj_rmin") is not None or kwargs.get("obj_rmax") is not None:
        return define_map_mask(cube.header, **kwargs)
    return None


def srt(data, cube, height_prior=False, **kwargs):
    """Solar rotational tomography model (models.py:19-54).

    Returns (P, prior, obj_mask, data_mask): P the data-masked projector with
    the Sun obstacle and the voxel mask of obj_rmin/obj_rmax fused in; prior a
    SmoothnessPrior (the axes for ``cg.solve(axes=...)``, with the r^2 weight
    of the height prior when ``height_prior`` is set).
    """
    data_mask = define_data_mask(data, **kwargs)
    obj_mask = _object_mask(cube, kwargs)
    P = Siddon(data.header, cube.header, mask=data_mask, obstacle="sun", voxmask=obj_mask,
               dtype=kwargs.get("dtype"))
    prior = smoothness_prior(cube.header, 3, height_prior, kwargs.get("convention", "fits"))
    return P, prior, obj_mask, data_mask


def stsrt(data, cube, n_groups=None, tindex=None, dt_min=None, height_prior=False, **kwargs):
    """Smooth temporal SRT (models.py:106-153): a 4-D map with one bin per group.

    The bins come from an explicit per-image ``tindex``, or ``n_groups``
    groups of consecutive images of equal size, or the DATE_OBS gaps of the
    stack with ``dt_min`` (``time_bins``), or else one bin per image.
    """
    if height_prior:
        raise ValueError("the height prior is defined for 3-D maps (models.py:96-104)")
    data_mask = define_data_mask(data, **kwargs)
    nv = len(data.header)
    if tindex is None:
        if n_groups is not None:
            ng = nv // n_groups
            tindex = np.minimum(np.arange(nv) // ng, n_groups - 1)
        else:
            tindex = time_bins(data.header, dt_min)
    obj_mask = _object_mask(cube, kwargs)
    P = Siddon4d(data.header, cube.header, tindex=tindex, mask=data_mask, obstacle="sun",
                 voxmask=obj_mask, dtype=kwargs.get("dtype"))
    return P, smoothness_prior(cube.header, 4), obj_mask, data_mask


def thomson(data, cube, u=0.5, **kwargs):
    """Thomson scattering model for white-light coronagraphs (models.py:190-221);
    its priors are the plain differences (models.py:217), whatever
    ``height_prior`` says."""
    data_mask = define_data_mask(data, **kwargs)
    pb = kwargs.get("pb", "pb")
    if pb != "pb":
        raise ValueError("Only pb implemented for now.")
    obj_mask = _object_mask(cube, kwargs)
    P = thomson_lo(data.header, cube.header, u=u, mask=data_mask, voxmask=obj_mask,
                   dtype=kwargs.get("dtype"))
    return P, smoothness_prior(cube.header, 3), obj_mask, data_mask
