"""Host-side rejection of missing or non-finite header values, before any
device work (runs without a GPU).  The reference reads each key with
PyDict_GetItemString / PyFloat_AsDouble (C_siddon.c.template:1051-1056), and a
NaN in the geometry makes voxel()'s cascade (C:811-912) take no branch, so its
traversal never ends."""
import numpy as np
import pytest

import tomograpy_amd as T
from tomograpy_amd.headers import map_record, view_records
from tomograpy_amd.lo_wrapper import siddon_lo
from tests import cases as C


def _case():
    ch, cube, data = C.small_case(n=8, npix=6, nv=3, dtype=np.float64)
    return ch, cube, data


@pytest.mark.parametrize("key", ["M1", "R2_3", "CDELT1", "CRPIX2"])
@pytest.mark.parametrize("value", [np.nan, np.inf, -np.inf])
def test_non_finite_view_value_rejected(key, value):
    ch, cube, data = _case()
    data.header[1][key] = value
    with pytest.raises(ValueError, match="view 1 header: %s" % key):
        view_records(data.header)
    # the operator checks before it touches a device (none here)
    with pytest.raises(ValueError, match=key):
        siddon_lo(data.header, ch, dtype=np.float64)
    with pytest.raises(ValueError, match=key):
        T.projector(data, cube)


def test_missing_view_key_rejected():
    ch, cube, data = _case()
    del data.header[2]["M3"]
    with pytest.raises(KeyError, match="view 2 header has no M3"):
        view_records(data.header)
    with pytest.raises(KeyError, match="M3"):
        siddon_lo(data.header, ch, dtype=np.float64)


@pytest.mark.parametrize("key", ["MMIN1", "CDELT3", "MMAX2"])
def test_non_finite_map_value_rejected(key):
    ch, cube, data = _case()
    ch = dict(ch)
    ch[key] = np.nan
    with pytest.raises(ValueError, match="map header: %s" % key):
        map_record(ch)
    with pytest.raises(ValueError, match=key):
        siddon_lo(data.header, ch, dtype=np.float64)


def test_missing_map_key_rejected():
    ch, cube, data = _case()
    ch = dict(ch)
    del ch["PSHAPE2"]
    with pytest.raises(KeyError, match="map header has no PSHAPE2"):
        map_record(ch)


def test_finite_headers_pass():
    ch, cube, data = _case()
    rec = view_records(data.header)
    assert rec.shape == (3, 18) and np.all(np.isfinite(rec))
    assert np.all(np.isfinite(map_record(ch)))
