"""Pins of the oracle's host pieces: scaling, damping kernels, moment counts, reconstruction.

Each check uses an independent fact: the worked value of the scale factor for energy_range
(-8.6, 8.6) with float bounds and the float tolerance 1 + 0.01f; the Jackson kernel's defining
convolution sum (Weisse et al., RMP 78, 275, Eq. 71); the Lorentz kernel's end points and its
Lorentzian line shape; the Chebyshev orthogonality sum rule of the reconstruction; and
LDOS = -Im G / pi (cppcore/tests/test_kpm.cpp:262).
"""
import numpy as np
import pytest


def test_scale_factors_closed_form(oracle):
    a, b = oracle.scale_factors(-8.6, 8.6)
    assert a == 8.686000303268429
    assert b == 0.0
    # b snaps to zero below |b/a| < 1e-4 (float), survives above
    a2, b2 = oracle.scale_factors(-8.6, 8.7)
    assert b2 != 0.0 and abs(b2 - 0.5 * (float(np.float32(8.7)) - float(np.float32(8.6)))) < 1e-12
    a3, b3 = oracle.scale_factors(-1.0, 1.00001)
    assert b3 == 0.0


def test_jackson_matches_convolution_definition(oracle):
    # g_n = sum_{v=0}^{N-n} a_v a_{v+n} / sum_v a_v^2 with a_v = sin(pi (v + 1) / (N + 1))
    for N in (10, 64, 2050):
        v = np.arange(N + 1)
        a = np.sin(np.pi * (v + 1) / (N + 1))
        ref = np.array([np.dot(a[: N + 1 - n], a[n:]) for n in range(N)]) / np.dot(a, a)
        g = oracle.jackson(N)
        np.testing.assert_allclose(g, ref, rtol=0, atol=5e-7)  # float pi vs exact pi
        assert g[0] == pytest.approx(1.0, abs=1e-12)


def test_lorentz_kernel(oracle):
    g = oracle.lorentz(100, 4.0)
    assert g[0] == pytest.approx(1.0)
    assert abs(oracle.lorentz(101, 4.0)[-1] - np.sinh(4 / 101) / np.sinh(4)) < 1e-15
    with pytest.raises(ValueError):
        oracle.lorentz(10, 0.0)


def test_round_and_required_num_moments(oracle):
    assert [oracle.round_num_moments(n) for n in (0, 1, 2, 3, 6, 7, 2048)] == [2, 2, 2, 6, 6, 10, 2050]
    # Jackson: round(int(pi_f / s) + 1)
    assert oracle.required_num_moments("jackson", 0.15 / 8.686000303268429) == oracle.round_num_moments(
        int(oracle.F32_PI / (0.15 / 8.686000303268429)) + 1)
    assert oracle.required_num_moments("lorentz", 0.1, lam=4.0) == 42


def test_spectral_density_sum_rule(oracle):
    # integral over E of the reconstructed density = mu_0_true * pi * k_f / 2 for any moments,
    # k_f = 2/pi evaluated in float (cppcore/include/kpm/reconstruct.hpp:28)
    # (orthogonality of T_n under 1/(pi sqrt(1-x^2))), with mu_0 stored halved
    rng = np.random.default_rng(0)
    M = 64
    lam = rng.uniform(-0.9, 0.9, 20)
    mu = np.cos(np.arange(M)[:, None] * np.arccos(lam)[None, :]).sum(axis=1)
    mu = mu * oracle.jackson(M)
    mu_stored = mu.copy()
    mu_stored[0] *= 0.5
    a, b = 3.0, 0.5
    x = np.cos(np.pi * (np.arange(4000) + 0.5) / 4000)  # Chebyshev-Gauss nodes
    e = a * x + b
    f = oracle.spectral_density(mu_stored, e, a, b)
    # Gauss-Chebyshev quadrature of f(E) dE = a f(ax+b) dx with weight sqrt(1-x^2) restored
    integral = np.sum(a * f * np.sqrt(1 - x * x)) * np.pi / 4000
    k_f = float(np.float32(2) / np.float32(np.pi))
    assert integral == pytest.approx(mu[0] * np.pi * k_f / 2, rel=1e-9)


def test_greens_imag_is_ldos(oracle):
    rng = np.random.default_rng(1)
    mu = rng.standard_normal(30)
    mu[0] = 0.7
    e = np.linspace(-2, 2, 11)
    g = oracle.greens_function(mu, e, 3.0, 0.1)
    ldos = oracle.spectral_density(mu, e, 3.0, 0.1)
    k_f = float(np.float32(2) / np.float32(np.pi))
    np.testing.assert_allclose(-g.imag / np.pi, ldos * (2 / np.pi) / k_f, rtol=1e-10)


def test_velocity_operator(oracle):
    import scipy.sparse as sp
    h = sp.csr_matrix(np.array([[0, 1.0, 0], [1.0, 0, 2.0], [0, 2.0, 0]]))
    x = np.array([0.0, 0.5, 2.0], dtype=np.float32)
    v = oracle.velocity(h, x).toarray()
    np.testing.assert_allclose(v, [[0, -0.5, 0], [0.5, 0, -3.0], [0, 3.0, 0]])
