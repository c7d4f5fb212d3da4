"""TEST INFRASTRUCTURE ONLY -- exact-rational brute-force ray/voxel traversal.

An independent statement of what a Siddon traversal computes (Siddon 1985):
the line x(a) = M + a u, clipped to the box [min, max], crosses the planes
min_k + m CDELT_k; between consecutive crossings it lies in one voxel, whose
index is floor((x(mid) - min) / CDELT) and whose length is the difference of
the parameters.  Everything is evaluated with fractions.Fraction on the exact
values of the (already rounded) floating-point inputs, so the result has no
rounding at all.  Used on tiny cubes to pin the oracle's floating-point
traversal (tomograpy/C_siddon.c.template:723-912) up to rounding.
"""
from fractions import Fraction as F


def exact_traversal(M, u, mmin, mmax, cdelt, n, sun=False):
    """Returns [(i0, i1, i2, length_fraction), ...] of nonzero-length segments.

    M, u, mmin, mmax, cdelt: 3 floats each (exact values are used); n: 3 ints.
    With sun=True the traversal stops before the first voxel whose entry point
    lies on or inside the unit sphere (the check C:575 makes before each step).
    """
    M = [F(x) for x in M]
    u = [F(x) for x in u]
    lo = [F(x) for x in mmin]
    hi = [F(x) for x in mmax]
    p = [F(x) for x in cdelt]
    amin, amax = None, None
    for k in range(3):
        if u[k] == 0:
            if not (lo[k] < M[k] < hi[k]):
                return []
            continue
        a1 = (lo[k] - M[k]) / u[k]
        a2 = (hi[k] - M[k]) / u[k]
        e, x = (a1, a2) if a1 < a2 else (a2, a1)
        amin = e if amin is None or e > amin else amin
        amax = x if amax is None or x < amax else amax
    if amin is None or not amin < amax:
        return []
    cuts = {amin, amax}
    for k in range(3):
        if u[k] == 0:
            continue
        for m in range(n[k] + 1):
            a = (lo[k] + m * p[k] - M[k]) / u[k]
            if amin < a < amax:
                cuts.add(a)
    cuts = sorted(cuts)
    out = []
    for a0, a1 in zip(cuts[:-1], cuts[1:]):
        if sun:
            x = [M[k] + a0 * u[k] for k in range(3)]
            if x[0] * x[0] + x[1] * x[1] + x[2] * x[2] <= 1:
                break
        mid = (a0 + a1) / 2
        idx = []
        for k in range(3):
            x = M[k] + mid * u[k]
            q = (x - lo[k]) / p[k]
            idx.append(q.numerator // q.denominator)
        if all(0 <= idx[k] < n[k] for k in range(3)):
            out.append((idx[0], idx[1], idx[2], a1 - a0))
    return out
