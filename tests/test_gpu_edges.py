"""GPU parity at the edges: systems smaller than one 64-row slice or just past it, rows with no
entries (isolated sites), an all-zero Hamiltonian, the fewest moments each entry point accepts,
and blocks of 1 to 64 vectors on both matrix formats.  Same tolerances as the other parity tests
(1e-10 of mu_0 in fp64/c128, 1e-4 in fp32/c64)."""
import numpy as np
import pytest
import scipy.sparse as sp

import abi

pytestmark = pytest.mark.gpu

TOL = {np.dtype(np.float64): 1e-10, np.dtype(np.complex128): 1e-10,
       np.dtype(np.float32): 1e-4, np.dtype(np.complex64): 1e-4}


@pytest.fixture(scope="module")
def ctx():
    c = abi.Ctx(0)
    yield c
    c.close()


def hermitian(n, dtype, seed, empty_fraction=0.0, density=0.05):
    """A random Hermitian H2 scaled into (-2, 2), with a fraction of its sites isolated (their
    rows and columns emptied)."""
    rng = np.random.default_rng(seed)
    a = sp.random(n, n, density=min(1.0, max(density, 2.0 / n)), random_state=rng, format="csr")
    a = a + a.T
    if np.dtype(dtype).kind == "c":
        b = sp.random(n, n, density=min(1.0, max(density, 2.0 / n)), random_state=rng, format="csr")
        a = a + 1j * (b - b.T)
    if empty_fraction:
        keep = sp.diags((rng.random(n) >= empty_fraction).astype(float))
        a = keep @ a @ keep
    a = sp.csr_matrix(a)
    a.eliminate_zeros()
    bound = max(abs(a).sum(axis=1).max(), 1e-30)
    h2 = sp.csr_matrix(a * (2.0 / (1.01 * bound))).astype(dtype)
    h2.sort_indices()
    return h2


def normwise(mu, ref):
    mu, ref = np.atleast_2d(mu), np.atleast_2d(ref)
    return np.max(np.abs(mu - ref) / np.abs(ref[:, :1]))


@pytest.mark.parametrize("dtype", [np.float64, np.complex64])
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 130])
@pytest.mark.parametrize("nvec", [1, 4, 64])
@pytest.mark.parametrize("fmt", [0, 1], ids=["csr", "sell"])
def test_tiny_systems(ctx, oracle, dtype, n, nvec, fmt):
    dtype = np.dtype(dtype)
    h2 = hermitian(n, dtype, seed=n * 7 + nvec)
    r0 = oracle.random_starter(n, nvec, dtype)
    mat, blk = ctx.mat(h2, fmt=fmt), ctx.block(r0)
    for M in (2, 4, 34):
        mu = abi.moments_diagonal(mat, blk, M)
        assert normwise(mu, oracle.diag_moments(h2, r0, M)) <= TOL[dtype], M
    blk.close()
    mat.close()


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
@pytest.mark.parametrize("fmt", [0, 1], ids=["csr", "sell"])
@pytest.mark.parametrize("nvec", [1, 8])
def test_isolated_sites(ctx, oracle, dtype, fmt, nvec):
    """40% of the sites carry no entry: whole 64-row slices of zero width appear in SELL."""
    dtype = np.dtype(dtype)
    h2 = hermitian(700, dtype, seed=21, empty_fraction=0.4, density=0.01)
    assert np.any(np.diff(h2.indptr) == 0)
    r0 = oracle.random_starter(700, nvec, dtype)
    mat, blk = ctx.mat(h2, fmt=fmt), ctx.block(r0)
    mu = abi.moments_diagonal(mat, blk, 66)
    blk.close()
    mat.close()
    assert normwise(mu, oracle.diag_moments(h2, r0, 66)) <= TOL[dtype]


@pytest.mark.parametrize("fmt", [0, 1], ids=["csr", "sell"])
def test_zero_hamiltonian(ctx, oracle, fmt):
    """H2 = 0 (no stored entries): T_n(0) = cos(n pi / 2), so mu_n = mu_0 * 2 cos(n pi / 2) for n > 0."""
    n = 150
    h2 = sp.csr_matrix((n, n), dtype=np.float64)
    r0 = oracle.random_starter(n, 4, np.float64)
    mat, blk = ctx.mat(h2, fmt=fmt), ctx.block(r0)
    mu = abi.moments_diagonal(mat, blk, 10)
    blk.close()
    mat.close()
    mu0 = 0.5 * np.sum(r0 * r0, axis=1)
    expect = np.outer(mu0, np.r_[1.0, 2 * np.cos(np.arange(1, 10) * np.pi / 2)])
    np.testing.assert_allclose(mu, expect, rtol=0, atol=1e-12 * mu0.max())
    assert normwise(mu, oracle.diag_moments(h2, r0, 10)) <= 1e-12


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("n", [1, 3, 65])
def test_offdiagonal_collectors_small(ctx, oracle, dtype, n):
    """Generic, multi-unit and dense collectors at the fewest moments they accept (2) and a few
    more, on systems below and just past one slice."""
    dtype = np.dtype(dtype)
    h2 = hermitian(n, dtype, seed=40 + n)
    rng = np.random.default_rng(n)
    alpha = oracle.random_starter(n, 1, dtype)[0]
    beta = rng.standard_normal(n).astype(dtype)
    mat = ctx.mat(h2)
    blk = ctx.block(alpha[None, :])
    for M in (2, 3, 17):
        mu = abi.moments_generic(mat, blk, M, beta, None)
        ref = oracle.offdiag_moments(h2, alpha, M, "generic", beta=beta)
        assert np.abs(mu - ref).max() <= 1e-10 * max(np.abs(ref).max(), 1.0), ("generic", M)
        dest = [0, n - 1, n // 2]
        mu = abi.moments_multi_unit(mat, blk, M, dest)
        ref = oracle.offdiag_moments(h2, alpha, M, "multi_unit", dest=dest)
        assert np.abs(mu - ref).max() <= 1e-10 * max(np.abs(ref).max(), 1.0), ("multi_unit", M)
        d = abi.moments_dense(mat, blk, M, None)
        ref = oracle.offdiag_moments(h2, alpha, M, "dense")
        assert np.abs(d[0] - ref).max() <= 1e-10 * max(np.abs(ref).max(), 1.0), ("dense", M)
    blk.close()
    mat.close()
