"""Pins of the oracle's Chebyshev recurrences and collectors against closed forms.

T_n(H~) = V diag(cos(n arccos l)) V^H on small dense matrices (oracle/brute.py), for every scalar
type and every collector of cppcore/src/kpm/default/collectors.cpp.  Also: the 1-D chain whose
moments are known analytically, mu_0 = |r|^2 / 2 (stored halved), |mu_n| <= mu_0, permutation
invariance, and the diagonal path equal to the generic path with beta = alpha.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle.brute as brute

DTYPES = [np.float32, np.float64, np.complex64, np.complex128]
TOL = {np.float32: 2e-5, np.float64: 1e-12, np.complex64: 2e-5, np.complex128: 1e-12}


def random_h2(n, dtype, seed=1, density=0.2):
    rng = np.random.default_rng(seed)
    a = sp.random(n, n, density=density, random_state=rng)
    a = a + a.T
    if np.dtype(dtype).kind == "c":
        b = sp.random(n, n, density=density / 2, random_state=rng)
        a = a + 1j * (b - b.T)
    ev = np.abs(np.linalg.eigvalsh(a.toarray())).max()
    return sp.csr_matrix(a * (2 / (1.05 * ev))).astype(dtype)


def vec(n, dtype, seed=2):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(n)
    if np.dtype(dtype).kind == "c":
        v = v + 1j * rng.standard_normal(n)
    return v.astype(dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_diagonal_matches_eigendecomposition(oracle, dtype):
    h2 = random_h2(40, dtype)
    r = np.stack([vec(40, dtype, s) for s in (2, 3, 4)])
    mu = oracle.diag_moments(h2, r, 60)
    for k in range(3):
        ref = brute.moments_eig(h2, r[k], 60)
        assert np.abs(mu[k] - ref).max() <= TOL[dtype] * abs(ref[0]) * 10
        assert abs(mu[k][0] - 0.5 * np.vdot(r[k], r[k]).real) <= TOL[dtype] * abs(ref[0])


@pytest.mark.parametrize("dtype", DTYPES)
def test_generic_matches_eigendecomposition(oracle, dtype):
    h2 = random_h2(35, dtype, seed=5)
    a, b = vec(35, dtype, 6), vec(35, dtype, 7)
    op = sp.random(35, 35, density=0.3, random_state=8).astype(dtype).tocsr()
    mu = oracle.offdiag_moments(h2, a, 50, "generic", beta=b, op=op)
    ref = brute.moments_eig(h2, a, 50, beta=b, op=op)
    assert np.abs(mu - ref).max() <= TOL[dtype] * np.abs(ref).max() * 10


@pytest.mark.parametrize("dtype", DTYPES)
def test_dense_and_multi_unit(oracle, dtype):
    h2 = random_h2(30, dtype, seed=9)
    a = vec(30, dtype, 10)
    d = oracle.offdiag_moments(h2, a, 40, "dense")
    ref = brute.vectors_eig(h2, a, 40)
    assert np.abs(d - ref).max() <= TOL[dtype] * np.abs(ref).max() * 10
    mu = oracle.offdiag_moments(h2, a, 40, "multi_unit", dest=[0, 7, 29])
    assert np.abs(mu - ref[:, [0, 7, 29]].T).max() <= TOL[dtype] * np.abs(ref).max() * 10


def test_chain_analytic_moments(oracle):
    # closed ring of N sites, hopping 1/2 in H~: eigenvalues cos(2 pi k / N), unit starter
    # mu_n = (1/N) sum_k cos(2 pi k n / N)  (= 1 if N | n else 0), before halving mu_0
    n = 64
    i = np.arange(n)
    h2 = sp.csr_matrix((np.ones(2 * n), (np.r_[i, i], np.r_[(i + 1) % n, (i - 1) % n])), shape=(n, n))
    e0 = np.zeros(n)
    e0[0] = 1
    mu = oracle.diag_moments(h2, e0[None, :], 130)[0]
    expected = np.array([1.0 if (k % n == 0) else 0.0 for k in range(130)])
    expected[0] = 0.5
    np.testing.assert_allclose(mu, expected, atol=1e-12)


def test_diagonal_equals_generic_and_is_permutation_invariant(oracle):
    h2 = random_h2(50, np.float64, seed=11)
    a = vec(50, np.float64, 12)
    d = oracle.diag_moments(h2, a[None, :], 40)[0]
    g = oracle.offdiag_moments(h2, a, 40, "generic")
    np.testing.assert_allclose(d, g, atol=1e-12 * abs(d[0]))
    p = np.random.default_rng(0).permutation(50)
    hp = h2[p][:, p]
    dp = oracle.diag_moments(hp, a[p][None, :], 40)[0]
    np.testing.assert_allclose(d, dp, atol=1e-12 * abs(d[0]))
    assert np.all(np.abs(d[1:]) <= 2 * d[0] * (1 + 1e-12))


def test_kpm_spmv_matches_dense(oracle):
    # kpm_spmv semantics y = H x - y (cppcore/tests/test_compute.cpp:82-148 style)
    for dtype in DTYPES:
        h2 = random_h2(100, dtype, seed=13)
        x, y = vec(100, dtype, 14), vec(100, dtype, 15)
        out = oracle.kpm_spmv(h2, x, y)
        ref = h2.toarray().astype(np.complex128) @ x - y
        assert np.abs(out - ref).max() <= TOL[dtype] * 100


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-4)])
@pytest.mark.parametrize("interleaved", [False, True])
def test_cpu_baseline_matches_diag_moments(oracle, dtype, tol, interleaved):
    """The CPU baseline (threaded ELL restatement of the reference's default compute, BFS order,
    W-wide batches plus promoted singles) gives the mean of the oracle's CSR diagonal moments over
    the same reference starters."""
    from pybinding_amd import generators
    h = generators.honeycomb_supercell(12, 9, disorder=0.8, seed=5, dtype=np.float64)
    n = h.shape[0]
    a, b = oracle.scale_factors(-9.0, 9.0)
    h2 = oracle.scale_hamiltonian(h, a, b, dtype)
    rmap, borders = oracle.bfs_reorder(h2, [0])
    inv = np.argsort(rmap)
    h2_opt = h2[inv][:, inv].tocsr()
    M, R = 42, 11   # two 4-wide (or one 8-wide) batches and leftover singles
    mu, _ = oracle.cpu_baseline_dos(oracle.to_ell(h2_opt), n, M, R, 3, perm=rmap, borders=borders,
                                    interleaved=interleaved)
    ref = oracle.diag_moments(h2, oracle.random_starter(n, R, dtype), M).mean(axis=0)
    assert np.max(np.abs(mu.astype(np.float64) - ref)) <= tol * abs(ref[0])
