"""Pins the oracle's set-up kernels to closed forms.

forward_modeling.py:203-241 (sponge) and :243-307 (averages, reciprocals).
"""
import numpy as np
import pytest

import oracle

A = 0.0053


def sponge_closed_form(nx, nz, fw, a=A):
    """Cerjan taper by distance to the left, right and bottom edges (no top
    strip): exp(-(a*(fw-d))^2) for d < fw, else 1, d = min(i, nx-1-i, nz-1-j)."""
    i = np.arange(nx)[:, None]
    j = np.arange(nz)[None, :]
    d = np.minimum(np.minimum(i, nx - 1 - i), nz - 1 - j)
    val = np.exp(-(a * a) * (fw - d).astype(np.float64) ** 2)
    return np.where(d < fw, val, 1.0)


@pytest.mark.parametrize("nx,nz,fw", [(50, 40, 10), (30, 30, 3), (12, 10, 3), (80, 61, 40), (9, 9, 1)])
def test_sponge_matches_distance_form(nx, nz, fw):
    got64 = oracle.sponge(nx, nz, fw, dtype=np.float64)
    got32 = oracle.sponge(nx, nz, fw, dtype=np.float32)
    # float64 scheme: a = 0.0053 in double
    np.testing.assert_allclose(got64, sponge_closed_form(nx, nz, fw), rtol=1e-14, atol=0)
    # float32 scheme: a is a ti.f32 kernel argument, so narrowed first
    ref32 = sponge_closed_form(nx, nz, fw, a=float(np.float32(A)))
    np.testing.assert_allclose(got32, ref32, rtol=3e-7, atol=0)
    # the taper never touches the top strip away from the side strips
    assert np.all(got32[fw:nx - fw, :max(0, nz - fw)] == 1.0)


def test_sponge_edge_values():
    c = oracle.sponge(200, 200, 20, dtype=np.float64)
    assert c[0, 100] == pytest.approx(np.exp(-(A * A) * 400.0), rel=1e-12)
    assert c[0, 100] == pytest.approx(0.98883, abs=5e-6)
    assert oracle.sponge(200, 200, 40, dtype=np.float64)[0, 50] == pytest.approx(0.95605, abs=5e-6)
    assert c[19, 100] == pytest.approx(0.99997, abs=5e-6)
    assert c[20, 100] == 1.0


def test_sponge_overlap_order_right_wins():
    """nx < 2*fw: the right strip is written after the left one, and the bottom
    strip after both (three sequential loops in one Taichi kernel)."""
    nx, nz, fw = 7, 20, 5
    c = oracle.sponge(nx, nz, fw, dtype=np.float64)
    e = lambda k: np.exp(-(A * A) * k * k)
    # cell i=2 lies in both side strips (i < fw and nx-1-i < fw) away from the bottom
    for i in range(nx):
        for j in range(0, 6):
            in_left = i < fw and j < nz - i - 1
            ir = nx - 1 - i
            in_right = ir < fw and j < nz - ir - 1
            if in_right:
                assert c[i, j] == pytest.approx(e(fw - ir), rel=1e-14)
            elif in_left:
                assert c[i, j] == pytest.approx(e(fw - i), rel=1e-14)
    assert c[2, 3] == pytest.approx(e(fw - 4), rel=1e-14)  # right (ir=4) overwrote left (i=2)
    assert c[2, 3] != pytest.approx(e(fw - 2), rel=1e-9)
    # bottom row wins over both
    jb = nz - 1
    for i in range(nx):
        assert c[i, jb] == pytest.approx(e(fw - 0), rel=1e-14)


def test_prepare_harmonic_and_arithmetic_means():
    rng = np.random.default_rng(3)
    nx, nz = 17, 13
    mu = rng.uniform(1e6, 1e7, (nx, nz)).astype(np.float32)
    rho = rng.uniform(1.5, 2.6, (nx, nz)).astype(np.float32)
    out = oracle.prepare(mu, rho, np.float32)
    m = mu.astype(np.float64)
    r = rho.astype(np.float64)
    I = (slice(1, nx - 1), slice(1, nz - 1))
    a, b, c, d = m[1:-1, 1:-1], m[2:, 1:-1], m[1:-1, 2:], m[2:, 2:]
    np.testing.assert_allclose(out["myx"][I], 2.0 / (1 / a + 1 / b), rtol=5e-7)
    np.testing.assert_allclose(out["myz"][I], 2.0 / (1 / a + 1 / c), rtol=5e-7)
    np.testing.assert_allclose(out["mxz"][I], 4.0 / (1 / a + 1 / b + 1 / c + 1 / d), rtol=1e-6)
    np.testing.assert_allclose(out["rho_u"][I], 0.5 * (r[1:-1, 1:-1] + r[2:, 1:-1]), rtol=3e-7)
    np.testing.assert_allclose(out["rho_w"][I], 0.5 * (r[1:-1, 1:-1] + r[1:-1, 2:]), rtol=3e-7)
    np.testing.assert_allclose(out["inv_rho"], 1.0 / r, rtol=3e-7)
    np.testing.assert_allclose(out["inv_rho_u"][I], 1.0 / out["rho_u"][I].astype(np.float64), rtol=3e-7)
    # borders: averages stay zero, their reciprocals are +inf
    for name in ("myx", "myz", "mxz", "rho_u", "rho_w"):
        f = out[name]
        assert np.all(f[0, :] == 0) and np.all(f[-1, :] == 0)
        assert np.all(f[:, 0] == 0) and np.all(f[:, -1] == 0)
    assert np.all(np.isinf(out["inv_rho_u"][0, :])) and np.all(np.isinf(out["inv_rho_w"][:, -1]))


def test_prepare_product_overflow_is_reference_behaviour():
    """The reference forms a*b*c*d in float32 (forward_modeling.py:271-276);
    for mu above ~4.3e9 the product overflows and mxz becomes inf while the
    float64 scheme and the two-point means stay finite."""
    mu = np.full((6, 6), 5.0e9, np.float32)
    rho = np.full((6, 6), 2200.0, np.float32)
    o32 = oracle.prepare(mu, rho, np.float32)
    o64 = oracle.prepare(mu, rho, np.float64)
    assert np.all(np.isinf(o32["mxz"][1:-1, 1:-1]))
    assert np.all(np.isfinite(o32["myx"][1:-1, 1:-1]))
    np.testing.assert_allclose(o64["mxz"][1:-1, 1:-1], 5.0e9, rtol=1e-12)
