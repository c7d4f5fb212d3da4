"""Every GPU path against the exact-rational fixtures of tests/golden/siddon_exact.npz.

The fixtures (tests/golden/make_golden.py) hold exact forward and backward
projections of an 8^3 cube seen in 4 views of 7 x 9 pixels, with and without
the Sun obstacle.  Tolerances: float64 agrees to rounding (1e-13 per pixel,
1e-13 relative L2 for the backward); float32 to the rounding of its geometry
(2e-5 per pixel, 1e-5 relative L2).
"""
import os

import numpy as np
import pytest
import torch

import tomograpy_amd as T
from tests.golden import make_golden as G

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "siddon_exact.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(FIXTURE)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("pre", ["plain_", "sun_"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gpu_paths_match_exact_fixture(gold, pre, dtype):
    from tomograpy_amd.headers import map_record, view_records
    from tomograpy_amd.lo_wrapper import siddon_lo
    print("golden", pre, np.dtype(dtype).name, flush=True)
    ch, hs = G.geometry()
    assert np.array_equal(view_records(hs), gold["view_records"])
    assert np.array_equal(map_record(ch), gold["map_record"])
    obstacle = "sun" if pre == "sun_" else None
    key = "64" if dtype == np.float64 else "32"
    ok = gold[pre + "ok" + key]
    fwd = gold[pre + "fwd"]
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    cube = torch.from_numpy(gold[pre + "cube"].astype(dtype)).cuda()
    cube.header = dict(ch)
    y = torch.from_numpy(gold[pre + "data" + key].astype(dtype)).cuda()

    def check_fwd(got, what):
        got = got.cpu().numpy()
        if dtype == np.float64:
            np.testing.assert_allclose(got[ok], fwd[ok], rtol=1e-13, atol=1e-14, err_msg=what)
        else:
            np.testing.assert_allclose(got[ok], fwd[ok], rtol=2e-5, atol=1e-5, err_msg=what)
        assert np.all(got[gold[pre + "nsegs"] == 0] == 0), what

    def check_bwd(got, what):
        tol = 1e-13 if dtype == np.float64 else 1e-5
        err = _rel(got.cpu().numpy(), gold[pre + "bpj" + key])
        assert err <= tol, (what, err)

    # reference-compatible functions: per-lane loop, atomic and deterministic backward
    data = torch.zeros(fwd.shape, dtype=tdt, device="cuda")
    data.header = hs
    T.projector(data, cube, obstacle=obstacle)
    check_fwd(data, "projector")
    for det in (False, True):
        out = torch.zeros(cube.shape, dtype=tdt, device="cuda")
        out.header = dict(ch)
        yd = y.clone()
        yd.header = hs
        T.backprojector(yd, out, obstacle=obstacle, deterministic=det)
        check_bwd(out, "backprojector deterministic=%s" % det)
    # the operator: precounted forward; brick-plan and atomic backward
    for mode in ("plan", "atomic"):
        P = siddon_lo(hs, ch, dtype=dtype, obstacle=obstacle, bpj_mode=mode)
        check_fwd(P.matvec(cube).reshape(fwd.shape), "siddon_lo.matvec " + mode)
        check_bwd(P.rmatvec(y).reshape(cube.shape), "siddon_lo.rmatvec " + mode)
