"""GPU checks of the paths around the projector's entry points: images read
from FITS files by the data front end (solar.read_data, solar.py:16-57 with
update_header, solar.py:59-115) projected and backprojected through siddon_lo,
and the one-image backprojector (siddon.py:233-243), against the oracle."""
import os

import numpy as np
import pytest
import torch

import tomograpy_amd as T
from tomograpy_amd import solar as SO
from tomograpy_amd.headers import map_record, view_records
from tomograpy_amd.lo_wrapper import siddon_lo
from tests import cases as C

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _instrument_header(k, n):
    """An EUVI-like header: Carrington longitude/latitude of the observer in
    degrees, its distance in metres, pixel scale in arcsec, roll from the PC
    matrix."""
    roll = 0.05 * k
    return {"CRLN_OBS": 10.0 + 47.0 * k, "CRLT_OBS": 3.0 * np.sin(k), "DSUN_OBS": 1.49e11,
            "DATE_OBS": "2010-06-0%dT12:00:00.000Z" % (k + 1),
            "CDELT1": 60.0, "CDELT2": 60.0, "CRVAL1": 0.0, "CRVAL2": 0.0,
            "CRPIX1": (n + 1) / 2.0, "CRPIX2": (n + 1) / 2.0, "CUNIT1": "arcsec",
            "CUNIT2": "arcsec", "INSTRUME": "EUVI", "TELESCOP": "STEREO_A",
            "PC1_1": np.cos(roll), "PC2_1": np.sin(roll)}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_read_data_through_operator_matches_oracle(oracle, tmp_path, dtype):
    """FITS files on disk -> read_data (transpose, arcsec -> rad, observer
    position and roll) -> siddon_lo on the GPU: the backprojection of the read
    images and the projection of a cube against the oracle on the same
    headers; forward bit-exact, backward within 1e-5 (fp32) / 1e-12 (fp64)."""
    print("front end", np.dtype(dtype).name, flush=True)
    n, nv = 40, 5
    rng = np.random.default_rng(12)
    for k in range(nv):
        img = rng.uniform(0.0, 1.0, (n, n))
        SO.write_fits(os.path.join(tmp_path, "img%02d.fits" % k), img, _instrument_header(k, n))
    data = SO.read_data(str(tmp_path), dtype=dtype)
    assert data.shape == (n, n, nv)
    ch = T.centered_cubic_map_header(3.0, 32, dtype=dtype)
    P = siddon_lo(data.header, ch, dtype=dtype, obstacle="sun")
    hs = P.data_header
    # backprojection of the images as read
    b = P.rmatvec(torch.from_numpy(np.ascontiguousarray(data)).cuda()).cpu().numpy()
    bref = np.zeros((32, 32, 32), dtype)
    oracle.siddon(oracle.BPJ, view_records(hs), map_record(ch), bref,
                  np.ascontiguousarray(data), sun=True, nthread=1)
    assert np.abs(bref).max() > 0
    assert rel_l2(b, bref) <= (1e-5 if dtype == np.float32 else 1e-12)
    # projection of a cube into the read geometry
    x = rng.uniform(0.0, 1.0, (32, 32, 32)).astype(dtype)
    y = P.matvec(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = np.zeros((n, n, nv), dtype)
    oracle.siddon(oracle.PJ, view_records(hs), map_record(ch), x, ref, sun=True)
    assert np.array_equal(y, ref)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_conic_image_backprojector_matches_oracle(oracle, dtype):
    """conic_image_backprojector(data, cube, t) adds the backprojection of image t
    alone to the cube (siddon.py:233-243, C:207-248); the oracle's view range
    [t, t + 1) gives the reference.  Images before and after t are ignored."""
    ch, cube, data = C.small_case(n=20, npix=28, nv=6, dtype=dtype)
    rng = np.random.default_rng(13)
    data[:] = rng.uniform(0.0, 1.0, data.shape).astype(dtype)
    for t in (0, 3, 5):
        print("image backprojector", np.dtype(dtype).name, t, flush=True)
        start = rng.uniform(0.0, 1.0, cube.shape).astype(dtype)
        out = T.InfoArray(start.copy(), header=dict(ch))
        ret = T.conic_image_backprojector(data, out, t, obstacle="sun")
        assert ret is data
        ref = start.copy()
        oracle.siddon(oracle.BPJ, view_records(data.header), map_record(ch), ref, data, sun=True,
                      nthread=1, t0=t, t1=t + 1)
        assert rel_l2(np.asarray(out) - start, ref - start) <= (1e-5 if dtype == np.float32
                                                                 else 1e-12)
        assert rel_l2(out, ref) <= (1e-6 if dtype == np.float32 else 1e-13)
