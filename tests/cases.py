"""Shared test geometries.

Restates the fixture parameters of the reference's tests/test_cases.py:7-32:
float32/float64, cubes of physical size 3 with 1^3 and 16^3 voxels, observers
at radius 200 with 1x1 and 32x32 images spanning fov(cube, 200), longitudes up
to 2 pi (numpy.linspace, endpoint included).
"""
import numpy as np

import tomograpy_amd as T
from tomograpy_amd.headers import map_record, view_records

DTYPES = [np.float32, np.float64]
RADIUS = 200.0
MAX_LON = 2 * np.pi


def object_headers(dtype):
    return [T.centered_cubic_map_header(3.0, s, dtype=dtype) for s in (1, 16)]


def image_headers(dtype):
    im_pshapes = [T.fov(h, RADIUS) for h in object_headers(np.float64)]
    hs = [T.centered_image_header(ps, s, dtype=dtype) for ps, s in zip(im_pshapes, (1, 32))]
    for h in hs:
        h["radius"] = RADIUS
        h["max_lon"] = MAX_LON
    return hs


def stack(im_h, dtype, n_images=1, fill=0.0, **kw):
    """circular_trajectory_data(**im_h) as the reference tests call it."""
    args = dict(im_h)
    args.update(kw)
    args["n_images"] = n_images
    d = T.circular_trajectory_data(dtype=dtype, **args)
    d[:] = fill
    return d


def records(data):
    T.headers.complete_view_headers(data.header)
    return view_records(data.header)


def maprec(cube_header):
    return map_record(cube_header)


def small_case(n=16, npix=24, nv=7, dtype=np.float64, seed=0, lon_offset=0.1, lat=0.05,
               radius=RADIUS):
    """A non-degenerate small problem: cube header, random cube, stack with headers."""
    rng = np.random.default_rng(seed)
    ch = T.centered_cubic_map_header(3.0, n, dtype=dtype)
    ih = T.centered_image_header(T.fov(ch, radius), npix, dtype=dtype)
    lons = lon_offset + 2 * np.pi * np.arange(nv) / nv
    lats = lat * np.sin(np.arange(nv) + 1.0)
    hs = T.circular_trajectory_header(radius=radius, longitudes=lons, latitudes=lats, **ih)
    data = T.InfoArray(np.zeros((npix, npix, nv), dtype=dtype), header=hs)
    cube = T.InfoArray(rng.uniform(0.0, 1.0, (n, n, n)).astype(dtype), header=dict(ch))
    return ch, cube, data
