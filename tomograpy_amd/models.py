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


def _r2omega(r):
    """Omega = arcsin(1/r), 0 where r == 0 (models.py:237-242)."""
    r = np.asarray(r, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        omega = np.arcsin(1.0 / r)
    omega = np.where(r == 0, 0.0, omega)
    return omega


def _impact_parameter(alpha, beta, d):
    """rho = d sin(arccos(cos a cos b)), clamped below at 1 (models.py:244-249)."""
    rho = d * np.sin(np.arccos(np.cos(alpha) * np.cos(beta)))
    rho = np.asarray(rho, dtype=np.float64)
    rho[rho < 1.0] = 1.0
    return rho


def _pb_thomson_coef(omega):
    """Billings' C1 = cos(O) sin(O)^2 and C2 (models.py:264-275)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        sino = np.sin(omega)
        sino2 = sino ** 2
        coso = np.cos(omega)
        coso2 = coso ** 2
        lno = np.log((1.0 + sino) / coso) * (coso2 / sino)
        C1 = coso * sino2
        C2 = -(1.0 / 8.0) * (1.0 - 3.0 * sino2 - (1.0 + 3.0 * sino2) * lno)
    return C1, C2


def _thomson_coef(omega):
    """All four Billings coefficients C1..C4 (models.py:251-262)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        sino = np.sin(omega)
        sino2 = sino ** 2
        coso = np.cos(omega)
        coso2 = coso ** 2
        lno = np.log((1.0 + sino) / coso) * (coso2 / sino)
        C1 = coso * sino2
        C2 = -(1.0 / 8.0) * (1.0 - 3.0 * sino2 - (1.0 + 3.0 * sino2) * lno)
        C3 = 4.0 / 3.0 - coso * (1.0 + coso2 / 3.0)
        C4 = (1.0 / 8.0) * (5.0 + sino2 - (5.0 - sino2) * lno)
    return C1, C2, C3, C4


def pb_data_coef(data_header, n1, n2):
    """rho^2 per pixel, shape (n1, n2, n_images) (models.py:277-289)."""
    out = np.zeros((n1, n2, len(data_header)))
    for t, h in enumerate(data_header):
        a, b = pixel_angles(h, n1, n2)
        A, B = np.meshgrid(a, b, indexing="ij")
        out[..., t] = _impact_parameter(A, B, h["D"]) ** 2
    return out


def pb_map_coef(cube_header, u=0.5, convention="fits"):
    """pi sigma / 2 ((1-u) C1 + u C2) / r^4 per voxel, non-finite -> 0 (models.py:291-309)."""
    x, y, z = map_axes(cube_header, convention)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    R2 = X * X + Y * Y + Z * Z
    R = np.sqrt(R2)
    O = _r2omega(R)
    C1, C2 = _pb_thomson_coef(O)
    with np.errstate(divide="ignore", invalid="ignore"):
        coefs = ((1 - u) * C1 + u * C2) / (R2 ** 2)
    coefs[~np.isfinite(coefs)] = 0.0
    return coefs * np.pi * sigma / 2.0


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


def srt(data, cube, **kwargs):
    """Solar rotational tomography model (models.py:19-54).

    Returns (P, prior_axes, obj_mask, data_mask): P the data-masked projector
    with the Sun obstacle and the voxel mask of obj_rmin/obj_rmax fused in.
    """
    data_mask = define_data_mask(data, **kwargs)
    obj_mask = None
    if kwargs.get("obj_rmin") is not None or kwargs.get("obj_rmax") is not None:
        obj_mask = define_map_mask(cube.header, **kwargs)
    P = Siddon(data.header, cube.header, mask=data_mask, obstacle="sun", voxmask=obj_mask,
               dtype=kwargs.get("dtype"))
    return P, smoothness_axes(3), obj_mask, data_mask


def stsrt(data, cube, n_groups=None, tindex=None, **kwargs):
    """Smooth temporal SRT (models.py:106-153): a 4-D map with one bin per group.

    ``n_groups`` consecutive-image groups (the reference's temporal_groups with
    equal group sizes), or an explicit per-image ``tindex``.
    """
    data_mask = define_data_mask(data, **kwargs)
    nv = len(data.header)
    if tindex is None:
        n_groups = n_groups or nv
        ng = nv // n_groups
        tindex = np.minimum(np.arange(nv) // ng, n_groups - 1)
    obj_mask = None
    if kwargs.get("obj_rmin") is not None or kwargs.get("obj_rmax") is not None:
        obj_mask = define_map_mask(cube.header, **kwargs)
    P = Siddon4d(data.header, cube.header, tindex=tindex, mask=data_mask, obstacle="sun",
                 voxmask=obj_mask, dtype=kwargs.get("dtype"))
    return P, smoothness_axes(4), obj_mask, data_mask


def thomson(data, cube, u=0.5, **kwargs):
    """Thomson scattering model for white-light coronagraphs (models.py:190-221)."""
    data_mask = define_data_mask(data, **kwargs)
    pb = kwargs.get("pb", "pb")
    if pb != "pb":
        raise ValueError("Only pb implemented for now.")
    obj_mask = None
    if kwargs.get("obj_rmin") is not None or kwargs.get("obj_rmax") is not None:
        obj_mask = define_map_mask(cube.header, **kwargs)
    P = thomson_lo(data.header, cube.header, u=u, mask=data_mask, voxmask=obj_mask,
                   dtype=kwargs.get("dtype"))
    return P, smoothness_axes(3), obj_mask, data_mask
