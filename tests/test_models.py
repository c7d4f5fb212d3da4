"""CPU pins of the model-level pieces: Thomson coefficients, impact parameter,
masks and the regularisation stencil reference used by the GPU tests."""
import numpy as np

import tomograpy_amd as T
from tomograpy_amd import models as Mo
from tomograpy_amd import workloads as W


def test_thomson_coefficients_limits():
    """Closed-form limits of Billings' coefficients (models.py:264-275):
    Omega -> pi/2 (r -> 1): C1 -> 0, C2 -> 1/4;
    Omega -> 0 (r -> inf):  C1 = s^2 - s^4/2 + O(s^6), C2 = 2/3 s^2 - 4/15 s^4 + O(s^6), s = sin Omega."""
    C1, C2 = Mo._pb_thomson_coef(np.array([np.pi / 2 - 1e-9]))
    assert abs(C1[0]) < 1e-8 and abs(C2[0] - 0.25) < 1e-7
    s = 1e-2
    om = np.arcsin(s)
    C1, C2 = Mo._pb_thomson_coef(np.array([om]))
    assert abs(C1[0] - (s ** 2 - s ** 4 / 2)) < 1e-10
    assert abs(C2[0] - (2.0 / 3.0 * s ** 2 - 4.0 / 15.0 * s ** 4)) < 1e-9
    # the 4-coefficient variant agrees on C1, C2
    a = Mo._thomson_coef(np.array([0.3, 0.7]))
    b = Mo._pb_thomson_coef(np.array([0.3, 0.7]))
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_r2omega_and_map_coef():
    assert Mo._r2omega(np.array([0.0]))[0] == 0.0
    np.testing.assert_allclose(Mo._r2omega(np.array([2.0])), [np.pi / 6])
    ch = T.centered_cubic_map_header(3.0, 8)
    f = Mo.pb_map_coef(ch, u=0.5)
    R = Mo.map_radius(ch)
    assert np.all(f[R < 1] == 0)  # arcsin(1/r) undefined inside the Sun -> 0
    assert np.all(np.isfinite(f))
    k = np.argmax(R)
    om = np.arcsin(1 / R.flat[k])
    C1, C2 = Mo._pb_thomson_coef(np.array([om]))
    expect = np.pi / 2 * (0.5 * C1[0] + 0.5 * C2[0]) / R.flat[k] ** 4
    np.testing.assert_allclose(f.flat[k], expect, rtol=1e-14)


def test_impact_parameter_closed_form():
    a = np.array([[0.0, 1e-2], [2e-3, 0.0]])
    b = np.array([[0.0, 2e-2], [0.0, 1e-4]])
    rho = Mo._impact_parameter(a, b, 200.0)
    assert rho[0, 0] == 1.0  # centre of the disc: clamped to 1
    # small angles: rho ~ D * sqrt(a^2 + b^2), to O(angle^2)
    np.testing.assert_allclose(rho[0, 1], 200 * np.sqrt(1e-4 + 4e-4), rtol=1e-3)
    np.testing.assert_allclose(rho[1, 1], 1.0)  # 200 * 1e-4 < 1 -> clamped


def test_masks():
    ch = T.centered_cubic_map_header(3.0, 16)
    m = Mo.define_map_mask(ch, obj_rmin=1.0, obj_rmax=1.4)
    R = Mo.map_radius(ch)
    assert np.array_equal(m, (R < 1.0) | (R > 1.4))
    hs = W.view_headers(32, [0.0, 1.0], d=214.0, dtype=np.float64)
    data = T.InfoArray(np.ones((32, 32, 2)), header=hs)
    data[0, 0, 1] = np.nan
    data[1, 1, 0] = -1.0
    dm = Mo.define_data_mask(data, data_rmin=1.1, mask_negative=True)
    Rd = Mo.distance_to_sun_center(hs, 32, 32)
    assert dm[0, 0, 1] and dm[1, 1, 0]
    assert np.array_equal(dm & ~np.isnan(data) & (data >= 0), (Rd < 1.1) & ~np.isnan(data) & (data >= 0))
    # solar radius in pixels: atan(1/D)/CDELT
    r1, _ = Mo.compute_rsun(hs)
    np.testing.assert_allclose(r1, np.arctan(1 / 214.0) / hs[0]["CDELT1"])


def regularizer_reference(x, axes, lam, identity):
    """lam (I + sum_k D_k^T D_k) x with D_k = numpy.diff along k, via explicit transposes."""
    out = x.copy() if identity else np.zeros_like(x)
    for ax in axes:
        d = np.diff(x, axis=ax)
        # D^T d: (D^T d)_m = d_{m-1} - d_m with out-of-range terms zero
        pad = [(0, 0)] * x.ndim
        pad[ax] = (1, 1)
        dp = np.pad(d, pad)
        sl_hi = [slice(None)] * x.ndim
        sl_lo = [slice(None)] * x.ndim
        sl_hi[ax] = slice(0, -1)
        sl_lo[ax] = slice(1, None)
        out += dp[tuple(sl_hi)] - dp[tuple(sl_lo)]
    return lam * out


def test_regularizer_reference_matches_matrix():
    """The numpy stencil reference equals D^T D built as explicit matrices."""
    rng = np.random.default_rng(0)
    x = rng.normal(size=(5, 4, 3))
    for ax in range(3):
        n = x.shape[ax]
        D = np.diff(np.eye(n), axis=0)  # (n-1, n) forward difference
        DtD = D.T @ D
        expect = np.moveaxis(np.tensordot(DtD, np.moveaxis(x, ax, 0), axes=1), 0, ax)
        np.testing.assert_allclose(regularizer_reference(x, [ax], 1.0, False), expect, atol=1e-12)
