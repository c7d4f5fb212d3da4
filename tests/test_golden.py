"""The oracle against the exact-rational fixtures of tests/golden/siddon_exact.npz.

The fixtures hold exact projections (fractions, rounded once) of tiny 8^3
geometries with 7 x 9 images, with and without the Sun obstacle; the generator
is tests/golden/make_golden.py.  Float64 must agree to rounding, float32 to its
geometry's rounding.
"""
import os

import numpy as np
import pytest

from tests.golden import make_golden as G

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "siddon_exact.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(FIXTURE)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_fixture_headers_match_the_header_builders(gold):
    """The header code still produces the geometry the fixtures were made from."""
    from tomograpy_amd.headers import map_record, view_records
    ch, hs = G.geometry()
    assert np.array_equal(map_record(ch), gold["map_record"])
    assert np.array_equal(view_records(hs), gold["view_records"])


@pytest.mark.parametrize("pre", ["plain_", "sun_"])
def test_fixture_adjoint_identity(gold, pre):
    """<P x, y> == <x, P^T y> holds on the exact values (a check of the fixture)."""
    lhs = float(np.sum(gold[pre + "fwd"] * gold[pre + "data64"]))
    rhs = float(np.sum(gold[pre + "cube"] * gold[pre + "bpj64"]))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


@pytest.mark.parametrize("pre", ["plain_", "sun_"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_matches_exact_fixture(oracle, gold, pre, dtype):
    sun = pre == "sun_"
    recs, mrec = gold["view_records"], gold["map_record"]
    cube = gold[pre + "cube"].astype(dtype)
    out = np.zeros(gold[pre + "fwd"].shape, dtype)
    oracle.siddon(oracle.PJ, recs, mrec, cube, out, sun=sun)
    key = "64" if dtype == np.float64 else "32"
    ok = gold[pre + "ok" + key]
    fwd = gold[pre + "fwd"]
    if dtype == np.float64:
        np.testing.assert_allclose(out[ok], fwd[ok], rtol=1e-13, atol=1e-14)
    else:
        np.testing.assert_allclose(out[ok], fwd[ok], rtol=2e-5, atol=1e-5)
    # the rays that miss the cube stay zero
    assert np.all(out[gold[pre + "nsegs"] == 0] == 0)
    data = gold[pre + "data" + key].astype(dtype)
    back = np.zeros(cube.shape, dtype)
    oracle.siddon(oracle.BPJ, recs, mrec, back, data, sun=sun, nthread=1)
    tol = 1e-14 if dtype == np.float64 else 1e-5
    assert _rel(back, gold[pre + "bpj" + key]) <= tol
