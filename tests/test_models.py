"""CPU pins of the model-level pieces: Thomson coefficients, impact parameter,
masks and the regularisation stencil reference used by the GPU tests."""
import numpy as np

import tomograpy_amd as T
from tomograpy_amd import models as Mo
from tomograpy_amd import workloads as W


def test_billings_factors_limits():
    """Closed-form limits of Billings' A and B (the C1, C2 of models.py:264-275):
    r -> 1 (Omega -> pi/2): A -> 0, B -> 1/4;
    r -> inf (s = 1/r -> 0): A = s^2 - s^4/2 + O(s^6), B = 2/3 s^2 - 4/15 s^4 + O(s^6)."""
    a, b = Mo.billings_ab(np.array([1.0 + 1e-12]))
    assert abs(a[0]) < 1e-5 and abs(b[0] - 0.25) < 1e-5
    s = 1e-2
    a, b = Mo.billings_ab(np.array([1.0 / s]))
    assert abs(a[0] - (s ** 2 - s ** 4 / 2)) < 1e-10
    assert abs(b[0] - (2.0 / 3.0 * s ** 2 - 4.0 / 15.0 * s ** 4)) < 1e-9
    # inside the Sun and at the centre: not finite (zeroed by the map coefficient)
    a, b = Mo.billings_ab(np.array([0.5, 0.0]))
    assert not np.any(np.isfinite(b))


def test_billings_factors_against_angle_form():
    """A and B written with the angle Omega = arcsin(1/r) (Billings' own
    variable) agree with the 1/r form to rounding, over 1 < r < 10."""
    r = np.linspace(1.001, 10.0, 97)
    om = np.arcsin(1.0 / r)
    so, co = np.sin(om), np.cos(om)
    ln = np.log((1.0 + so) / co)
    a_ref = co * so ** 2
    b_ref = -(1.0 - 3.0 * so ** 2 - (1.0 + 3.0 * so ** 2) * ln * co ** 2 / so) / 8.0
    a, b = Mo.billings_ab(r)
    np.testing.assert_allclose(a, a_ref, rtol=1e-11)
    np.testing.assert_allclose(b, b_ref, rtol=1e-9)


def test_pb_map_coef():
    ch = T.centered_cubic_map_header(3.0, 8)
    f = Mo.pb_map_coef(ch, u=0.5)
    R = Mo.map_radius(ch)
    assert np.all(f[R < 1] == 0)  # arcsin(1/r) undefined inside the Sun -> 0
    assert np.all(np.isfinite(f))
    k = np.argmax(R)
    om = np.arcsin(1 / R.flat[k])
    so, co = np.sin(om), np.cos(om)
    A = co * so ** 2
    B = -(1.0 - 3.0 * so ** 2 - (1.0 + 3.0 * so ** 2) * np.log((1 + so) / co) * co ** 2 / so) / 8
    expect = np.pi / 2 * (0.5 * A + 0.5 * B) / R.flat[k] ** 4
    np.testing.assert_allclose(f.flat[k], expect, rtol=1e-9)
    # u weights A against B
    f0 = Mo.pb_map_coef(ch, u=0.0)
    np.testing.assert_allclose(f0.flat[k], np.pi / 2 * A / R.flat[k] ** 4, rtol=1e-9)


def test_impact_parameter_closed_form():
    a = np.array([[0.0, 1e-2], [2e-3, 0.0]])
    b = np.array([[0.0, 2e-2], [0.0, 1e-4]])
    rho2 = Mo.impact_parameter_sq(a, b, 200.0)
    assert rho2[0, 0] == 1.0  # centre of the disc: clamped to 1
    # small angles: rho ~ D * sqrt(a^2 + b^2), to O(angle^2)
    np.testing.assert_allclose(rho2[0, 1], 200 ** 2 * (1e-4 + 4e-4), rtol=2e-3)
    np.testing.assert_allclose(rho2[1, 1], 1.0)  # 200 * 1e-4 < 1 -> clamped
    # exact: the line of sight at angle a from the Sun-observer line, b = 0
    np.testing.assert_allclose(Mo.impact_parameter_sq(np.array(0.3), np.array(0.0), 5.0),
                               (5.0 * np.sin(0.3)) ** 2, rtol=1e-14)


def test_height_prior_weight():
    """height_prior composes every D_k with diag(r^2), r^2 = x^2 + y^2 + z^2 of
    the voxel coordinates (models.py:85-104); a 4-D map has none."""
    ch = T.centered_cubic_map_header(3.0, 6)
    p = Mo.smoothness_prior(ch, 3, height_prior=True)
    assert tuple(p) == (0, 1, 2)
    x, y, z = T.headers.map_axes(ch)
    i, j, k = 1, 4, 2
    assert p.weight[i, j, k] == x[i] ** 2 + y[j] ** 2 + z[k] ** 2
    assert Mo.smoothness_prior(ch, 3).weight is None
    assert tuple(Mo.smoothness_prior(ch, 4)) == (0, 1, 2, 3)
    import pytest
    with pytest.raises(ValueError):
        Mo.smoothness_prior(ch, 4, height_prior=True)


def test_time_bins():
    """Images closer in time than dt_min share a bin; default one bin per image."""
    hs = [{"DATE_OBS": d} for d in ("2010-01-01T00:00:00Z", "2010-01-01T00:00:30Z",
                                    "2010-01-01T01:00:00Z", "2010-01-01T01:00:10Z",
                                    "2010-01-02T00:00:00Z")]
    assert list(Mo.time_bins(hs, dt_min=60.0)) == [0, 0, 1, 1, 2]
    assert list(Mo.time_bins(hs, dt_min=1e6)) == [0, 0, 0, 0, 0]
    assert list(Mo.time_bins(hs)) == [0, 1, 2, 3, 4]
    same = [{"DATE_OBS": "2010-01-01T00:00:00Z"}] * 4
    assert list(Mo.time_bins(same, dt_min=1.0)) == [0, 0, 0, 0]
    import pytest
    with pytest.raises(ValueError):
        Mo.time_bins(hs[::-1], dt_min=60.0)


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


def weighted_regularizer_reference(x, w, axes, lam, identity):
    """lam (x + sum_k (D_k W)^T (D_k W) x) with W = diag(w) and D_k the numpy.diff
    matrix along axis k, built as explicit matrices (the height prior of
    models.py:85-104: each D_k composed with diag(r^2))."""
    n = x.size
    W = np.diag(w.ravel())
    total = np.eye(n) if identity else np.zeros((n, n))
    for ax in axes:
        # D_k on the flattened array: difference of index m + e_k and m along axis k
        idx = np.arange(n).reshape(x.shape)
        hi = np.take(idx, np.arange(1, x.shape[ax]), axis=ax).ravel()
        lo = np.take(idx, np.arange(0, x.shape[ax] - 1), axis=ax).ravel()
        D = np.zeros((hi.size, n))
        D[np.arange(hi.size), hi] = 1.0
        D[np.arange(hi.size), lo] = -1.0
        DW = D @ W
        total += DW.T @ DW
    return lam * (total @ x.ravel()).reshape(x.shape)


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
    # the weighted matrix form reduces to it with unit weights
    np.testing.assert_allclose(weighted_regularizer_reference(x, np.ones_like(x), [0, 1, 2], 0.5,
                                                              True),
                               regularizer_reference(x, [0, 1, 2], 0.5, True), atol=1e-12)
