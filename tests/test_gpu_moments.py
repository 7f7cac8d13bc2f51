"""GPU parity: the HIP engine (through its C-ABI) against the oracle on identical seeded inputs.

Tolerances (normwise, |mu_n| <= mu_0 makes them meaningful for every n):
    fp64 / c128:  max_n |dmu_n| <= 1e-10 * |mu_0|
    fp32 / c64 :  max_n |dmu_n| <= 1e-4  * |mu_0|
"""
import numpy as np
import pytest
import scipy.sparse as sp

import abi
from pybinding_amd import generators

pytestmark = pytest.mark.gpu

TOL = {np.dtype(np.float64): 1e-10, np.dtype(np.complex128): 1e-10,
       np.dtype(np.float32): 1e-4, np.dtype(np.complex64): 1e-4}
DTYPES = [np.float64, np.float32, np.complex128, np.complex64]


@pytest.fixture(scope="module")
def ctx():
    c = abi.Ctx(0)
    yield c
    c.close()


def scaled_random(n, dtype, seed=0, density=None):
    """A random Hermitian sparse matrix scaled into (-1, 1) (times 2, the H2 convention)."""
    rng = np.random.default_rng(seed)
    density = density or min(1.0, 6.0 / n)
    a = sp.random(n, n, density=density, random_state=rng, format="csr")
    a = a + a.T
    if np.dtype(dtype).kind == "c":
        b = sp.random(n, n, density=density / 2, random_state=rng, format="csr")
        a = a + 1j * (b - b.T)
    bound = abs(a).sum(axis=1).max()  # Gershgorin
    h2 = sp.csr_matrix(a * (2.0 / (1.01 * bound))).astype(dtype)
    h2.sort_indices()
    return h2


def graphene_h2(n1, n2, dtype, disorder=0.5, random_hoppings=False):
    h = generators.honeycomb_supercell(n1, n2, disorder=disorder, dtype=np.float64)
    if random_hoppings:  # every bond its own hopping (symmetric), so hundreds of distinct values
        upper = sp.triu(h, 1).tocsr()
        upper.data = upper.data * np.random.default_rng(n1 * n2).uniform(0.8, 1.2, upper.nnz)
        h = (upper + upper.T + sp.diags(h.diagonal())).tocsr()
        h.sort_indices()
    a = 0.5 * 2 * (3 * 2.8 * 1.2 + disorder / 2) * 1.01
    return sp.csr_matrix(h * (2 / a)).astype(dtype)


def normwise(mu, ref):
    mu = np.atleast_2d(mu)
    ref = np.atleast_2d(ref)
    return np.max(np.abs(mu - ref) / np.abs(ref[:, :1]))


@pytest.mark.parametrize("nvec", [4, 16, 64])
def test_diagonal_five_entry_rows(ctx, oracle, nvec):
    """Square lattice with onsite disorder: every interior row has exactly 5 entries (the SELL
    block kernel's 5-entry staging)."""
    h = generators.square_flake(40, 37, t=-1.0, disorder=1.5, seed=6)
    a, b = oracle.scale_factors(-4.8, 4.8)
    h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
    h2.sort_indices()
    assert np.diff(h2.indptr).max() == 5
    r0 = oracle.random_starter(h2.shape[0], nvec, np.float64)
    mat = ctx.mat(h2, fmt=1)
    blk = ctx.block(r0)
    mu = abi.moments_diagonal(mat, blk, 130)
    ref = oracle.diag_moments(h2, r0, 130)
    blk.close()
    mat.close()
    assert normwise(mu, ref) <= 1e-10


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("nvec", [1, 2, 3, 8, 16, 32, 64, 70])
@pytest.mark.parametrize("fmt", [0, 1], ids=["csr", "sell"])
def test_diagonal_block_vs_oracle(ctx, oracle, dtype, nvec, fmt):
    dtype = np.dtype(dtype)
    h2 = scaled_random(2000 + 37 * nvec, dtype, seed=nvec)  # ragged last slice
    r0 = oracle.random_starter(h2.shape[0], nvec, dtype)
    M = 130
    mat = ctx.mat(h2, fmt=fmt)
    blk = ctx.block(r0)
    mu = abi.moments_diagonal(mat, blk, M)
    blk.close()
    mat.close()
    ref = oracle.diag_moments(h2, r0, M)
    print(dtype, nvec, normwise(mu, ref))
    assert normwise(mu, ref) <= TOL[dtype]


@pytest.mark.parametrize("dtype,nvec", [(np.float64, 64), (np.float64, 32), (np.float64, 16), (np.float64, 8),
                                         (np.complex64, 8), (np.float32, 16), (np.complex128, 4)])
def test_every_step_variant(ctx, oracle, dtype, nvec):
    """All 60 variants of the diagonal step kernel (tile shape x grid x LDS, wave-shuffle, fp64
    column-pair staging, and column pairs on the value-table matrix form), pinned one by one on
    full-width column tiles, against the oracle: a graphene supercell (4 entries per row, one
    hopping value: the value-table form applies), a graphene supercell with random hoppings (more
    than 256 distinct values: it does not) and a random matrix (rows of 1 to 27 entries: the
    8-entry staging plus the wide-row loop)."""
    dtype = np.dtype(dtype)
    cases = ((30, graphene_h2(40, 30, dtype)), (32, graphene_h2(36, 20, dtype, random_hoppings=True)),
             (31, scaled_random(3000, dtype, seed=31)))
    default_width = ctx.tile_width(dtype)
    ctx.set_tile_width(dtype, 64)
    try:
        for seed, h2 in cases:
            r0 = oracle.random_starter(h2.shape[0], nvec, dtype)
            ref = oracle.diag_moments(h2, r0, 66)
            mat, blk = ctx.mat(h2, fmt=1), ctx.block(r0)
            try:
                for variant in range(60):
                    ctx.set_step_variant(dtype, nvec, variant)
                    mu = abi.moments_diagonal(mat, blk, 66)
                    assert normwise(mu, ref) <= TOL[dtype], (variant, seed)
            finally:
                ctx.set_step_variant(dtype, nvec, 0)
                blk.close()
                mat.close()
    finally:
        ctx.set_tile_width(dtype, default_width)


@pytest.mark.parametrize("dtype", DTYPES)
def test_diagonal_graph_replays_match_eager(ctx, oracle, dtype):
    """The captured recurrence against step-by-step launches, bit for bit, on every replay, with
    eager work on the same stream and synchronisations between the replays; then a new number of
    moments, a re-tuned variant and a new matrix each force a fresh capture."""
    dtype = np.dtype(dtype)
    h2 = graphene_h2(24, 20, dtype)
    r0 = oracle.random_starter(h2.shape[0], 8, dtype)
    mat = ctx.mat(h2, fmt=1)
    blk = ctx.block(r0)
    other = ctx.block(r0[:3])
    try:
        ctx.set_graph_policy(0)
        eager = abi.moments_diagonal(mat, blk, 258)
        eager_short = abi.moments_diagonal(mat, blk, 130)
        eager_other = abi.moments_diagonal(mat, other, 130)
        assert normwise(eager, oracle.diag_moments(h2, r0, 258)) <= TOL[dtype]
        for policy in (2, 1):
            ctx.set_graph_policy(policy)
            for rep in range(4):
                abi.moments_diagonal_run(mat, blk, 258)
                abi.moments_diagonal_run(mat, other, 130)  # other work queued behind the replay
                ctx.synchronize()
                assert np.array_equal(abi.moments_diagonal_fetch(mat, blk, 258), eager), (policy, rep)
                assert np.array_equal(abi.moments_diagonal_fetch(mat, other, 130), eager_other), (policy, rep)
            assert np.array_equal(abi.moments_diagonal(mat, blk, 130), eager_short)
            assert np.array_equal(abi.moments_diagonal(mat, blk, 258), eager)
        abi.autotune_diagonal(mat, blk)
        ctx.set_graph_policy(2)
        tuned = abi.moments_diagonal(mat, blk, 258)
        assert normwise(tuned, eager) <= TOL[dtype]
        mat2 = ctx.mat(h2 * dtype.type(0.5), fmt=1)
        try:
            half = abi.moments_diagonal(mat2, blk, 130)
            ctx.set_graph_policy(0)
            assert np.array_equal(half, abi.moments_diagonal(mat2, blk, 130))
        finally:
            mat2.close()
    finally:
        ctx.set_graph_policy(0)
        other.close()
        blk.close()
        mat.close()


@pytest.mark.parametrize("dtype", DTYPES)
def test_diagonal_sell_single_vector(ctx, oracle, dtype):
    dtype = np.dtype(dtype)
    h2 = graphene_h2(40, 30, dtype)
    r0 = oracle.random_starter(h2.shape[0], 1, dtype)
    mat = ctx.mat(h2, fmt=1)
    blk = ctx.block(r0)
    mu = abi.moments_diagonal(mat, blk, 258)
    ref = oracle.diag_moments(h2, r0, 258)
    blk.close()
    mat.close()
    assert normwise(mu, ref) <= TOL[dtype]


def test_diagonal_graphene_long_fp64(ctx, oracle):
    """2048 moments on a disordered periodic graphene supercell, 8 Rademacher vectors."""
    h2 = graphene_h2(64, 48, np.float64)
    r0 = oracle.random_starter(h2.shape[0], 8, np.float64)
    mat = ctx.mat(h2)
    blk = ctx.block(r0)
    mu = abi.moments_diagonal(mat, blk, 2050)
    ref = oracle.diag_moments(h2, r0, 2050)
    blk.close()
    mat.close()
    assert normwise(mu, ref) <= 1e-10


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("with_op", [False, True])
def test_generic_vs_oracle(ctx, oracle, dtype, with_op):
    dtype = np.dtype(dtype)
    n = 1500
    h2 = scaled_random(n, dtype, seed=3)
    rng = np.random.default_rng(4)
    alpha = oracle.random_starter(n, 1, dtype)[0]
    beta = rng.standard_normal(n).astype(dtype)
    op = sp.random(n, n, density=3.0 / n, random_state=5, format="csr").astype(dtype) if with_op else None
    mat = ctx.mat(h2)
    blk = ctx.block(alpha[None, :])
    mu = abi.moments_generic(mat, blk, 101, beta, op)
    ref = oracle.offdiag_moments(h2, alpha, 101, "generic", beta=beta, op=op)
    blk.close()
    mat.close()
    scale = np.abs(ref).max()
    assert np.abs(mu - ref).max() <= TOL[dtype] * scale * 10


@pytest.mark.parametrize("dtype", [np.float64, np.complex64])
def test_multi_unit_vs_oracle(ctx, oracle, dtype):
    dtype = np.dtype(dtype)
    n = 1200
    h2 = scaled_random(n, dtype, seed=6)
    e = np.zeros(n, dtype=dtype)
    e[17] = 1
    dest = [3, 17, 500, 3, 1199]
    mat = ctx.mat(h2)
    blk = ctx.block(e[None, :])
    mu = abi.moments_multi_unit(mat, blk, 80, dest)
    ref = oracle.offdiag_moments(h2, e, 80, "multi_unit", dest=dest)
    blk.close()
    mat.close()
    assert np.abs(mu - ref).max() <= TOL[dtype] * 10


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
@pytest.mark.parametrize("with_op", [False, True])
def test_dense_vs_oracle(ctx, oracle, dtype, with_op):
    dtype = np.dtype(dtype)
    n = 700
    h2 = scaled_random(n, dtype, seed=7)
    r0 = oracle.random_starter(n, 2, dtype)
    op = sp.random(n, n, density=4.0 / n, random_state=8, format="csr").astype(dtype) if with_op else None
    mat = ctx.mat(h2)
    opm = ctx.mat(op) if op is not None else None
    blk = ctx.block(r0)
    d = abi.moments_dense(mat, blk, 40, opm)
    for c in range(2):
        ref = oracle.offdiag_moments(h2, r0[c], 40, "dense", op=op)
        assert np.abs(d[c] - ref).max() <= TOL[dtype] * np.abs(ref).max() * 10
    blk.close()
    mat.close()
    if opm is not None:
        opm.close()


@pytest.mark.parametrize("M", [64, 70, 200])
@pytest.mark.parametrize("dtype", [np.complex128, np.float64, np.complex64, np.float32])
def test_kubo_contraction_vs_fp64_gemm(ctx, oracle, dtype, M):
    """mu_nm = (1/R) sum_r L_r R_r^H against an FP64 numpy GEMM of oracle panels; the FP32 split
    and bf16x3 split MFMA contractions are checked against the same oracle at fp32 accuracy.
    M = 70 and 200 leave partial output tiles; n = 900 leaves partial k panels."""
    dtype = np.dtype(dtype)
    n, R = 900, 3
    h2 = scaled_random(n, dtype, seed=9)
    x = np.random.default_rng(10).uniform(-5, 5, n).astype(np.float32)
    y = np.random.default_rng(11).uniform(-5, 5, n).astype(np.float32)
    vl = oracle.velocity(h2, x).astype(dtype)
    vr = oracle.velocity(h2, y).astype(dtype)
    r0 = oracle.random_starter(n, R, dtype)
    ref = np.zeros((M, M), dtype=np.complex128)
    for k in range(R):
        L = oracle.offdiag_moments(h2, vl @ r0[k], M, "dense")
        Rr = oracle.offdiag_moments(h2, r0[k], M, "dense", op=vr)
        ref += L.astype(np.complex128) @ Rr.astype(np.complex128).conj().T
    ref /= R
    mat, ml, mr = ctx.mat(h2), ctx.mat(vl), ctx.mat(vr)
    blk = ctx.block(r0)
    mus = {mode: abi.kubo_moments(mat, ml, mr, blk, M, mode=mode) for mode in (0, 1, 2)}
    for h in (blk, mat, ml, mr):
        h.close()
    scale = np.abs(ref).max()
    err = {mode: np.abs(mu - ref).max() / scale for mode, mu in mus.items()}
    print("kubo M=%d %s: fp64 err %.3e  fp32-split err %.3e  bf16x3 err %.3e" % (M, dtype, err[0], err[1], err[2]))
    double = dtype in (np.float64, np.complex128)
    assert err[0] <= (1e-11 if double else 1e-5)
    assert err[1] <= 1e-5
    assert err[2] <= 1e-5


@pytest.mark.parametrize("dtype", [np.complex128, np.float64])
def test_kubo_contraction_sixteen_vectors(ctx, oracle, dtype):
    """16 random vectors advanced as one group (the panel budget allows it at this size): the
    16-wide sweeps, split-plane writes and per-vector contractions of every mode."""
    dtype = np.dtype(dtype)
    n, R, M = 700, 16, 66
    h2 = scaled_random(n, dtype, seed=12)
    x = np.random.default_rng(13).uniform(-5, 5, n).astype(np.float32)
    vl = oracle.velocity(h2, x).astype(dtype)
    r0 = oracle.random_starter(n, R, dtype)
    ref = np.zeros((M, M), dtype=np.complex128)
    for k in range(R):
        L = oracle.offdiag_moments(h2, vl @ r0[k], M, "dense")
        Rr = oracle.offdiag_moments(h2, r0[k], M, "dense", op=vl)
        ref += L.astype(np.complex128) @ Rr.astype(np.complex128).conj().T
    ref /= R
    mat, ml = ctx.mat(h2, fmt=1), ctx.mat(vl, fmt=1)
    blk = ctx.block(r0)
    mus = {mode: abi.kubo_moments(mat, ml, ml, blk, M, mode=mode) for mode in (0, 1, 2)}
    for h in (blk, mat, ml):
        h.close()
    scale = np.abs(ref).max()
    err = {mode: np.abs(mu - ref).max() / scale for mode, mu in mus.items()}
    print("kubo 16 vectors %s: %s" % (dtype, err))
    assert err[0] <= 1e-11 and err[1] <= 1e-5 and err[2] <= 1e-5


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
def test_lanczos_bounds(ctx, dtype):
    h = generators.honeycomb_supercell(12, 10, disorder=1.0, dtype=np.float64)
    if np.dtype(dtype).kind == "c":
        h = h.astype(np.complex128)
    h = h.astype(dtype)
    ev = np.linalg.eigvalsh(h.toarray().astype(np.complex128))
    mat = ctx.mat(h)
    lo, hi, loops = abi.lanczos_bounds(mat, 0.002)
    mat.close()
    assert loops > 0
    assert abs(lo - ev[0]) <= 1e-3 * abs(ev[0])
    assert abs(hi - ev[-1]) <= 1e-3 * abs(ev[-1])


def test_errors_are_reported(ctx):
    h2 = scaled_random(100, np.float64)
    mat = ctx.mat(h2)
    blk = ctx.block(np.ones((1, 50)))
    with pytest.raises(RuntimeError, match="Size mismatch"):
        abi.moments_diagonal(mat, blk, 10)
    blk.close()
    blk = ctx.block(np.ones((1, 100)))
    with pytest.raises(RuntimeError, match="even"):
        abi.moments_diagonal(mat, blk, 11)
    blk.close()
    mat.close()
