"""Synthetic solar-corona workloads: seeded Gaussian-blob phantoms and the
observer geometries of the benchmark configurations.

  C1  64^3 cube over [-1.5, 1.5]^3, 32 views of 128^2, D = 214, LON = 2 pi k / 32
  C2  256^3, 128 views of 512^2, LON = 2 pi k / 128, LAT ~ U(-7, 7) deg,
      40 blobs, images = P(phantom) + N(0, (0.01 max)^2)
  C3  128^3 x 8 time bins, 64 views of 256^2, view k -> bin k mod 8,
      voxel mask r < 1.05
  C4  64 uneven jobs: N ~ U{32..192}, V ~ U{8..96}, images in {64, 128, 256, 512}^2,
      D ~ U(150, 250), half Thomson-weighted
Config k draws from numpy.random.default_rng(20261203 + k).  Longitudes are
listed explicitly (no repeated endpoint, unlike simu.circular_trajectory_data's
linspace).  The field of view is +-1.6 R_sun at the observer distance.
"""
import numpy as np

from .headers import (centered_cubic_map_header, centered_image_header, header_rotation_matrix,
                      image_header, update_observer)

SEED0 = 20261203


def rng_for(config):
    return np.random.default_rng(SEED0 + int(config))


def cube_header(n, pshape=3.0, dtype=np.float32):
    """Header of an n^3 cube spanning [-pshape/2, pshape/2]^3."""
    return centered_cubic_map_header(pshape, n, dtype=dtype)


def view_headers(n_pix, lons, lats=None, d=214.0, half_fov=1.6, dtype=np.float32):
    """Per-view headers of n_pix^2 images seeing +-half_fov R_sun from distance d."""
    cdelt = 2.0 * np.arctan(half_fov / d) / n_pix
    base = centered_image_header(cdelt * n_pix, n_pix, dtype=dtype)
    lats = np.zeros(len(lons)) if lats is None else np.asarray(lats)
    ds = np.broadcast_to(np.asarray(d, dtype=np.float64), (len(lons),))
    hs = []
    for lon, lat, dd in zip(lons, lats, ds):
        h = image_header(base)
        update_observer(h, lon=float(lon), lat=float(lat), d=float(dd))
        header_rotation_matrix(h)
        hs.append(h)
    return hs


def voxel_radius(header):
    """Distance of each voxel centre MMIN + (i + 1/2) CDELT to the origin, (N0, N1, N2)."""
    axes = []
    for k in range(3):
        n = int(header["NAXIS%d" % (k + 1)])
        cd = header["CDELT%d" % (k + 1)]
        mmin = -header["CRPIX%d" % (k + 1)] * cd
        axes.append(mmin + (np.arange(n) + 0.5) * cd)
    x, y, z = np.meshgrid(*axes, indexing="ij")
    return np.sqrt(x * x + y * y + z * z)


def blob_phantom(header, n_blobs, rng, shell=(1.05, 1.4), sigma=(0.05, 0.2), amp=(0.5, 2.0),
                 r_zero=1.0, dtype=np.float32):
    """Sum of Gaussian blobs with centres uniform in volume in the shell, zero for r < r_zero."""
    axes = []
    for k in range(3):
        n = int(header["NAXIS%d" % (k + 1)])
        cd = header["CDELT%d" % (k + 1)]
        mmin = -header["CRPIX%d" % (k + 1)] * cd
        axes.append(mmin + (np.arange(n) + 0.5) * cd)
    x, y, z = np.meshgrid(*axes, indexing="ij")
    out = np.zeros(x.shape, dtype=np.float64)
    r3 = rng.uniform(shell[0] ** 3, shell[1] ** 3, n_blobs)
    rad = np.cbrt(r3)
    v = rng.normal(size=(n_blobs, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    centres = v * rad[:, None]
    sig = rng.uniform(sigma[0], sigma[1], n_blobs)
    amps = rng.uniform(amp[0], amp[1], n_blobs)
    for c, s, a in zip(centres, sig, amps):
        d2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
        out += a * np.exp(-0.5 * d2 / (s * s))
    out[np.sqrt(x * x + y * y + z * z) < r_zero] = 0.0
    return out.astype(dtype)


def blob_phantom_torch(header, n_blobs, rng, shell=(1.05, 1.4), sigma=(0.05, 0.2),
                       amp=(0.5, 2.0), r_zero=1.0, device="cuda"):
    """blob_phantom evaluated with torch on the device (same draws from rng), float64."""
    import torch
    axes = []
    for k in range(3):
        n = int(header["NAXIS%d" % (k + 1)])
        cd = header["CDELT%d" % (k + 1)]
        mmin = -header["CRPIX%d" % (k + 1)] * cd
        axes.append(torch.as_tensor(mmin + (np.arange(n) + 0.5) * cd, device=device))
    x, y, z = torch.meshgrid(*axes, indexing="ij")
    out = torch.zeros(x.shape, dtype=torch.float64, device=device)
    r3 = rng.uniform(shell[0] ** 3, shell[1] ** 3, n_blobs)
    rad = np.cbrt(r3)
    v = rng.normal(size=(n_blobs, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    centres = v * rad[:, None]
    sig = rng.uniform(sigma[0], sigma[1], n_blobs)
    amps = rng.uniform(amp[0], amp[1], n_blobs)
    for c, s, a in zip(centres, sig, amps):
        d2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
        out += a * torch.exp(-0.5 * d2 / (s * s))
    out[torch.sqrt(x * x + y * y + z * z) < r_zero] = 0.0
    return out


def config1(dtype=np.float32):
    """(cube_header, cube, view_headers) of C1."""
    rng = rng_for(1)
    ch = cube_header(64, dtype=dtype)
    cube = blob_phantom(ch, 8, rng, dtype=dtype)
    lons = 2 * np.pi * np.arange(32) / 32
    return ch, cube, view_headers(128, lons, dtype=dtype)


def config2_geometry(dtype=np.float32, n=256, n_views=128, n_pix=512):
    """(cube_header, view_headers, rng) of C2 (sizes overridable for scaled tests)."""
    rng = rng_for(2)
    ch = cube_header(n, dtype=dtype)
    lons = 2 * np.pi * np.arange(n_views) / n_views
    lats = np.deg2rad(rng.uniform(-7.0, 7.0, n_views))
    return ch, view_headers(n_pix, lons, lats, dtype=dtype), rng


def config3_geometry(dtype=np.float32, n=128, n_views=64, n_pix=256, n_bins=8):
    """(cube_header, view_headers, tindex, voxmask) of C3."""
    rng = rng_for(3)
    ch = cube_header(n, dtype=dtype)
    lons = 2 * np.pi * np.arange(n_views) / n_views
    hs = view_headers(n_pix, lons, dtype=dtype)
    tindex = (np.arange(n_views) % n_bins).astype(np.int32)
    voxmask = voxel_radius(ch) < 1.05
    return ch, hs, tindex, voxmask, rng


def config5_mix(n_requests, rng, dtype=np.float32, n_range=(32, 256), v_range=(8, 128),
                pix=(64, 128, 256, 512), max_bins=8):
    """Size mix of the served stream: list of dicts with the cube edge, view
    count, image size, mask flag and time-bin count of each request."""
    out = []
    for _ in range(n_requests):
        out.append({"n": int(rng.integers(n_range[0], n_range[1] + 1)),
                    "nv": int(rng.integers(v_range[0], v_range[1] + 1)),
                    "npix": int(rng.choice(pix)),
                    "mask": bool(rng.integers(0, 2)),
                    "bins": int(rng.integers(1, max_bins + 1)),
                    "d": float(rng.uniform(150.0, 250.0))})
    return out


def config4_jobs(n_jobs=64, dtype=np.float32, max_n=192, max_views=96, pix=(64, 128, 256, 512)):
    """List of job dicts of C4: cube header, view headers, thomson flag."""
    rng = rng_for(4)
    jobs = []
    for _ in range(n_jobs):
        n = int(rng.integers(32, max_n + 1))
        nv = int(rng.integers(8, max_views + 1))
        npx = int(rng.choice(pix))
        d = float(rng.uniform(150.0, 250.0))
        ch = cube_header(n, dtype=dtype)
        lons = rng.uniform(0, 2 * np.pi, nv)
        lats = np.deg2rad(rng.uniform(-7.0, 7.0, nv))
        hs = view_headers(npx, lons, lats, d=d, dtype=dtype)
        jobs.append({"cube_header": ch, "headers": hs, "thomson": bool(rng.integers(0, 2)),
                     "n": n, "nv": nv, "npix": npx})
    return jobs


def config3_problem(dtype=np.float32, **sizes):
    """Differentiation workload: (operator, data y, start x0) on the C3 geometry
    (128^3 x 8 time bins, 64 views of 256^2 round-robin to the bins, voxels with
    r < 1.05 excluded).  y = P x_true + 1% noise with a time-scaled blob corona."""
    import torch
    from .lo_wrapper import Siddon4d
    ch, hs, tindex, voxmask, rng = config3_geometry(dtype, **sizes)
    nt = int(tindex.max()) + 1
    P = Siddon4d(hs, ch, tindex=tindex, voxmask=voxmask, dtype=dtype)
    base = blob_phantom(ch, 16, rng, dtype=np.float64)
    x_true = np.stack([base * (1 + 0.05 * b) for b in range(nt)], axis=-1).astype(dtype)
    y = P.matvec(torch.from_numpy(x_true).to(P.dev)).clone()
    y += 0.01 * float(y.abs().max()) * torch.from_numpy(
        rng.standard_normal(tuple(y.shape)).astype(dtype)).to(P.dev)
    x0 = torch.from_numpy(rng.uniform(0, 1, x_true.shape).astype(dtype)).to(P.dev)
    return P, y, x0


def config3_loss(P, y, lam=1e-2, lam_t=1e-2):
    """x -> ||P x - y||^2 + lam ||grad_xyz x||^2 + lam_t ||d x / dt||^2 (differentiable;
    the two priors in one fused pass)."""
    from .autograd import project, smoothness

    def loss(x):
        r = project(x, P) - y
        return (r * r).sum() + smoothness(x, (lam, lam, lam, lam_t))

    return loss


def config3_loss_torch(P, y, lam=1e-2, lam_t=1e-2):
    """config3_loss with the priors written as torch.diff terms (a check of the
    fused prior)."""
    import torch
    from .autograd import project

    def loss(x):
        r = project(x, P) - y
        out = (r * r).sum()
        for ax in range(3):
            out = out + lam * (torch.diff(x, dim=ax) ** 2).sum()
        return out + lam_t * (torch.diff(x, dim=3) ** 2).sum()

    return loss


def config4_problem(dtype=np.float32, **kw):
    """Batched-jobs workload: operators and phantoms of the C4 job list (half of
    the jobs Thomson-weighted)."""
    import torch
    from . import models as Mo
    from .lo_wrapper import Siddon
    rng = rng_for(40)
    ops, xs = [], []
    for job in config4_jobs(dtype=dtype, **kw):
        if job["thomson"]:
            P = Mo.thomson_lo(job["headers"], job["cube_header"], dtype=dtype)
        else:
            P = Siddon(job["headers"], job["cube_header"], dtype=dtype)
        ops.append(P)
        xs.append(torch.from_numpy(blob_phantom(job["cube_header"], 8, rng, dtype=dtype)).to(P.dev))
    return ops, xs
