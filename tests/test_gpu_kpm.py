"""GPU parity of the full `kpm()` facade (host C++ core + HIP backend) against the oracle.

The engine BFS-reorders H2 and permutes the starters (cppcore/src/kpm/OptimizedHamiltonian.cpp
:74-152); the oracle works in the original order, so agreement also pins the reorder.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import pybinding_amd as pb
from pybinding_amd import generators

pytestmark = pytest.mark.gpu


def flake(dtype=np.float64):
    h = generators.honeycomb_flake(6, 5, dtype=np.float64)
    d = np.random.default_rng(0).uniform(-0.3, 0.3, h.shape[0])
    return (h + sp.diags(d)).tocsr().astype(dtype)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-4), (np.complex128, 1e-10),
                                       (np.complex64, 1e-4)])
@pytest.mark.parametrize("fmt", ["ELL", "CSR"])
def test_dos_moments_match_oracle(oracle, dtype, tol, fmt):
    h = flake(dtype)
    n = h.shape[0]
    kpm = pb.kpm(h, energy_range=(-9, 9), silent=True, matrix_format=fmt)
    M, R = 202, 5
    mu = kpm.dos_moments(M, R)
    a, b = oracle.scale_factors(-9, 9)
    assert kpm.scaling_factors == pytest.approx((a, b), rel=1e-15)
    h2 = oracle.scale_hamiltonian(h, a, b, dtype)
    r0 = oracle.random_starter(n, R, dtype)
    ref = oracle.diag_moments(h2, r0, M).mean(axis=0)
    assert np.max(np.abs(mu - ref)) <= tol * abs(ref[0])


def test_autotuned_dos_moments_match_oracle(oracle):
    """kpm(..., autotune=True): the first block of each width is timed over the step variants
    before the moments are taken; the moments stay on the oracle."""
    h = generators.honeycomb_supercell(40, 30, disorder=0.5, dtype=np.float64)
    n = h.shape[0]
    kpm = pb.kpm(h, energy_range=(-9, 9), silent=True, autotune=True)
    M, R = 130, 16
    mu = kpm.dos_moments(M, R)
    a, b = oracle.scale_factors(-9, 9)
    h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
    ref = oracle.diag_moments(h2, oracle.random_starter(n, R, np.float64), M).mean(axis=0)
    assert np.max(np.abs(mu - ref)) <= 1e-10 * abs(ref[0])


def test_calc_dos_matches_oracle_reconstruction(oracle):
    h = flake()
    n = h.shape[0]
    kpm = pb.kpm(h, energy_range=(-9, 9), kernel=pb.lorentz_kernel(), silent=True)
    energy = np.linspace(-2, 2, 25)
    dos = kpm.calc_dos(energy, broadening=0.15, num_random=3)
    a, b = oracle.scale_factors(-9, 9)
    M = oracle.required_num_moments("lorentz", 0.15 / a)
    h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
    mu = oracle.diag_moments(h2, oracle.random_starter(n, 3, np.float64), M).mean(axis=0)
    mu = mu * oracle.lorentz(M)
    ref = oracle.spectral_density(mu, energy, a, b)
    np.testing.assert_allclose(dos.data, ref, rtol=1e-9, atol=1e-12 * abs(ref).max())


def test_moments_api_and_generic(oracle):
    h = flake()
    n = h.shape[0]
    kpm = pb.kpm(h, energy_range=(-9, 9), silent=True)
    a, b = kpm.scaling_factors
    h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
    alpha = np.zeros(n)
    alpha[5] = 1
    mu = kpm.moments(40, alpha)
    M = oracle.round_num_moments(40)
    ref = (oracle.diag_moments(h2, alpha[None, :], M)[0] * oracle.jackson(M))[:40]
    np.testing.assert_allclose(mu.real, ref, atol=1e-12)
    beta = np.random.default_rng(3).standard_normal(n)
    op = sp.random(n, n, density=3 / n, random_state=4, format="csr")
    mu = kpm.moments(40, alpha, beta, op)
    ref = (oracle.offdiag_moments(h2, alpha, M, "generic", beta=beta, op=op) * oracle.jackson(M))[:40]
    np.testing.assert_allclose(mu.real, ref, atol=1e-11 * np.abs(ref).max())
    with pytest.raises(RuntimeError, match="Size mismatch"):
        kpm.moments(10, [1, 2, 3])
    with pytest.raises(RuntimeError, match="Hamiltonian is real, but the given argument 'alpha' is complex"):
        kpm.moments(10, 1j * alpha)


def test_greens_vector_and_invariants(oracle):
    h = flake()
    n = h.shape[0]
    kpm = pb.kpm(h, energy_range=(-9, 9), silent=True)
    energy = np.linspace(-0.3, 0.3, 10)
    i, j = n // 2, n // 4
    gs = kpm.calc_greens(i, [j, j + 1, j + 2], energy, 0.8)
    g = kpm.calc_greens(j, i, energy, 0.8)
    np.testing.assert_allclose(gs[0], g, rtol=1e-9, atol=1e-12)  # g_ij = g_ji for real H
    gii = kpm.calc_greens(i, i, energy, 0.8)
    # against the oracle: unit starter at i, multi-unit moments at j
    a, b = kpm.scaling_factors
    h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
    M = oracle.required_num_moments("jackson", 0.8 / a)
    e = np.zeros(n)
    e[i] = 1
    mu = oracle.offdiag_moments(h2, e, M, "multi_unit", dest=[j, i])
    ref_ij = oracle.greens_function(mu[0] * oracle.jackson(M), energy, a, b)
    ref_ii = oracle.greens_function(mu[1] * oracle.jackson(M), energy, a, b)
    np.testing.assert_allclose(gs[0], ref_ij, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gii, ref_ii, rtol=1e-9, atol=1e-12)


def test_lanczos_bounds_used_when_no_range():
    h = generators.honeycomb_supercell(10, 8, disorder=0.4)
    ev = np.linalg.eigvalsh(h.toarray())
    kpm = pb.kpm(h, silent=True)
    a, b = kpm.scaling_factors
    lo, hi = kpm.impl.energy_range
    assert lo == pytest.approx(ev[0], rel=2e-3)
    assert hi == pytest.approx(ev[-1], rel=2e-3)


@pytest.mark.parametrize("order", ["auto", "bfs"])
def test_site_order_does_not_change_moments(oracle, order):
    """The Hilbert-curve site order (whole-system calculations) and the reference's BFS order give
    the same moments as the unpermuted oracle."""
    h, pos = generators.honeycomb_flake(8, 6, return_positions=True)
    d = np.random.default_rng(5).uniform(-0.4, 0.4, h.shape[0])
    h = (h + sp.diags(d)).tocsr()
    model = pb.Model(h, pb.model.System(pos[:, 0], pos[:, 1]))
    kpm = pb.kpm(model, energy_range=(-9, 9), silent=True, site_order=order)
    mu = kpm.dos_moments(130, 4)
    a, b = oracle.scale_factors(-9, 9)
    h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
    ref = oracle.diag_moments(h2, oracle.random_starter(h.shape[0], 4, np.float64), 130).mean(axis=0)
    assert np.max(np.abs(mu - ref)) <= 1e-10 * abs(ref[0])


def test_conductivity_matches_oracle(oracle):
    """calc_conductivity against the oracle: dense panels L = T_n(H~) v_x r and R = v_x T_m(H~) r,
    mu_nm = mean_r L R^H, 2-D Jackson damping, Kubo-Bastin reconstruction."""
    h, pos = generators.honeycomb_flake(5, 4, return_positions=True)
    d = np.random.default_rng(7).uniform(-0.5, 0.5, h.shape[0])
    h = (h + sp.diags(d)).tocsr()
    n = h.shape[0]
    model = pb.Model(h, pb.model.System(pos[:, 0], pos[:, 1]))
    kpm = pb.kpm(model, energy_range=(-9, 9), silent=True)
    mu_e = np.linspace(-2, 2, 9)
    sigma = kpm.calc_conductivity(mu_e, broadening=0.8, temperature=0, num_random=3, num_points=50)
    a, b = oracle.scale_factors(-9, 9)
    M = oracle.required_num_moments("jackson", 0.8 / a)
    h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
    x = pos[:, 0].astype(np.float32)
    v = oracle.velocity(h, x)
    r0 = oracle.random_starter(n, 3, np.float64)
    mu = np.zeros((M, M))
    for k in range(3):
        L = oracle.offdiag_moments(h2, v @ r0[k], M, "dense")
        R = oracle.offdiag_moments(h2, r0[k], M, "dense", op=v)
        mu += L @ R.T
    mu /= 3
    g = oracle.jackson(M)
    mu = mu * g[:, None] * g[None, :]
    lo, hi = float(np.float32(-9)), float(np.float32(9))
    samples = np.linspace(lo, hi, 50)
    ref = oracle.kubo_bastin(mu, mu_e, samples, 0.0, a, b).real
    np.testing.assert_allclose(sigma.data, ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())
