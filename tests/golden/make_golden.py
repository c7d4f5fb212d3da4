"""Writes tests/golden/siddon_exact.npz: exact projections of tiny geometries.

Test infrastructure only.  The expected values come from the exact-rational
traversal in oracle/brute.py (Siddon 1985 restated with fractions.Fraction on
the exact values of the rounded inputs), not from the oracle's floating-point
code.  Each sum is exact and rounded once to float64.

The unit vectors are the closed form of C_siddon.c.template:1013-1046,
u0 = R (cos(lambda) cos(gamma), cos(lambda) sin(gamma), sin(lambda)) with
gamma = (i - CRPIX1 + 1) CDELT1 + CRVAL1 and lambda likewise on axis 2,
evaluated in numpy float64.

Geometry: an 8^3 cube spanning [-1.5, 1.5]^3, 4 views of 7 x 9 pixels from
radius 4 (so rays cross the unit sphere), longitudes 0.37 + k pi / 2, latitudes
0.2 sin(k), CRVAL1 = 0.013 k.  Two cases: no obstacle and the Sun obstacle.

Rays left out (mask False, their data zeroed in the backprojection inputs):
- sun case, float64 set: a voxel entry point within 1e-9 of r^2 = 1, where
  rounding may decide the stop differently;
- float32 set: the same within 1e-4 (float32 geometry moves entry points by
  about 1e-6), or a unit-vector component below 1e-6 in magnitude.

Run: python tests/golden/make_golden.py  (a few seconds)
"""
import os
import sys
from fractions import Fraction as F

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import tomograpy_amd as T  # noqa: E402
from tomograpy_amd.headers import map_record, view_records  # noqa: E402
from oracle import brute  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "siddon_exact.npz")
N = 8
N1, N2 = 7, 9
NV = 4
RADIUS = 4.0


def geometry():
    """(cube header, view headers) of the fixture."""
    ch = T.centered_cubic_map_header(3.0, N, dtype=np.float64)
    fv = T.fov(ch, RADIUS)
    ih = T.centered_image_header((fv, fv), (N1, N2), dtype=np.float64)
    lons = 0.37 + np.pi / 2 * np.arange(NV)
    lats = 0.2 * np.sin(np.arange(NV))
    hs = T.circular_trajectory_header(radius=RADIUS, longitudes=lons, latitudes=lats, **ih)
    for t, h in enumerate(hs):
        h["CRVAL1"] = 0.013 * t
    return ch, hs


def unit_vector(rec, i, j):
    g = (i - rec[0] + 1) * rec[2] + rec[4]
    lam = (j - rec[1] + 1) * rec[3] + rec[5]
    R = rec[6:15].reshape(3, 3)
    v = np.array([np.cos(lam) * np.cos(g), np.cos(lam) * np.sin(g), np.sin(lam)])
    return R @ v


def sphere_gap(M, u, segs, mrec):
    """Smallest |r^2 - 1| over the entry points of the exact segments."""
    if not segs:
        return np.inf
    lo, hi = mrec[6:9], mrec[9:12]
    amin = max((F(float(lo[k] if u[k] > 0 else hi[k])) - F(M[k])) / F(u[k])
               for k in range(3) if u[k] != 0)
    gap = None
    a = amin
    for *_, ln in segs:
        r2 = sum((F(M[k]) + a * F(u[k])) ** 2 for k in range(3))
        d = abs(r2 - 1)
        gap = d if gap is None or d < gap else gap
        a += ln
    return float(gap)


def case(sun, rng):
    ch, hs = geometry()
    mrec = map_record(ch)
    recs = view_records(hs)
    cube = rng.uniform(0.5, 1.5, (N, N, N))
    data = rng.normal(size=(N1, N2, NV))
    segs = {}
    gap = np.full((N1, N2, NV), np.inf)
    small_u = np.zeros((N1, N2, NV), bool)
    for t in range(NV):
        M = recs[t][15:18]
        for i in range(N1):
            for j in range(N2):
                u = unit_vector(recs[t], i, j)
                small_u[i, j, t] = np.any((u != 0) & (np.abs(u) < 1e-6))
                s = brute.exact_traversal(M, u, mrec[6:9], mrec[9:12], mrec[0:3], (N, N, N), sun=sun)
                segs[i, j, t] = s
                if sun:
                    # the unclipped traversal's entry points decide where the ray stops
                    full = brute.exact_traversal(M, u, mrec[6:9], mrec[9:12], mrec[0:3], (N, N, N))
                    gap[i, j, t] = sphere_gap(M, u, full, mrec)
    ok64 = (gap >= 1e-9) & ~small_u
    ok32 = (gap >= 1e-4) & ~small_u
    cube_f = [[[F(float(x)) for x in row] for row in plane] for plane in cube]
    fwd = np.zeros((N1, N2, NV))
    for (i, j, t), s in segs.items():
        fwd[i, j, t] = float(sum((ln * cube_f[a][b][c] for a, b, c, ln in s), F(0)))

    def backproject(keep):
        acc = {}
        for (i, j, t), s in segs.items():
            if not keep[i, j, t]:
                continue
            w = F(float(data[i, j, t]))
            for a, b, c, ln in s:
                acc[a, b, c] = acc.get((a, b, c), F(0)) + ln * w
        out = np.zeros((N, N, N))
        for (a, b, c), v in acc.items():
            out[a, b, c] = float(v)
        return out

    pre = "sun_" if sun else "plain_"
    return {
        pre + "cube": cube,
        pre + "data64": np.where(ok64, data, 0.0),
        pre + "data32": np.where(ok32, data, 0.0),
        pre + "ok64": ok64,
        pre + "ok32": ok32,
        pre + "fwd": fwd,
        pre + "bpj64": backproject(ok64),
        pre + "bpj32": backproject(ok32),
        pre + "nsegs": np.array([[[len(segs[i, j, t]) for t in range(NV)] for j in range(N2)]
                                 for i in range(N1)], np.int32),
    }


def main():
    rng = np.random.default_rng(20261203)
    ch, hs = geometry()
    out = {"map_record": map_record(ch), "view_records": view_records(hs)}
    out.update(case(False, rng))
    out.update(case(True, rng))
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
