"""The two-step diagonal recurrence (kpmx_ctx_set_two_step): each launch computes r_{k+1} on a
tile of rows and its halo and r_{k+2} on the tile.  Against the oracle (oracle.diag_moments,
cppcore/src/kpm/default/collectors.cpp:5-34) and against single steps, for every scalar type at
widths whose rows are whole 128-byte lines, odd and even step counts, tiles smaller than a slice
band and larger than the matrix, Hilbert-ordered and BFS-like site orders, and graph replay."""
import numpy as np
import pytest
import scipy.sparse as sp

import abi
from pybinding_amd import generators

pytestmark = pytest.mark.gpu


def matrix(dtype, n1=40, n2=28, seed=6, hilbert=True):
    import oracle
    h = generators.honeycomb_supercell(n1, n2, disorder=0.9, seed=seed, dtype=np.float64)
    if hilbert:
        q = generators.honeycomb_cell_order(n1, n2)
        inv = np.argsort(q)
        h = h[inv][:, inv].tocsr()
    a, b = oracle.scale_factors(-9.3, 9.3)
    return oracle.scale_hamiltonian(h, a, b, dtype)


CASES = [(np.float64, 16, 1e-10), (np.float64, 64, 1e-10), (np.float64, 20, 1e-10), (np.complex128, 8, 1e-10),
         (np.complex128, 16, 1e-10), (np.float32, 32, 1e-4), (np.float32, 64, 1e-4), (np.complex64, 16, 1e-4)]


@pytest.mark.parametrize("dtype,nvec,tol", CASES)
@pytest.mark.parametrize("M", [130, 132])
def test_two_step_matches_oracle_and_single_steps(oracle, dtype, nvec, tol, M):
    h2 = matrix(dtype)
    r0 = oracle.random_starter(h2.shape[0], nvec, dtype)
    ref = oracle.diag_moments(h2, r0, M)
    ctx = abi.Ctx(0)
    try:
        ctx.set_tile_width(dtype, 64)
        mat, blk = ctx.mat(h2, fmt=1), ctx.block(r0)
        single = abi.moments_diagonal(mat, blk, M)
        out = {}
        for rows, wg in ((64, 0), (256, 0), (1024, 0), (1 << 20, 0), (128, 1), (256, 3)):
            ctx.set_two_step(dtype, rows, wg)
            assert ctx.two_step(dtype) == rows
            out[rows, wg] = abi.moments_diagonal(mat, blk, M)
        blk.close()
        mat.close()
    finally:
        ctx.close()
    scale = np.abs(ref[:, :1])
    assert np.max(np.abs(single - ref) / scale) <= tol
    for rows, mu in out.items():
        assert np.max(np.abs(mu - ref) / scale) <= tol, rows
        assert np.max(np.abs(mu - single) / scale) <= (1e-13 if tol < 1e-6 else 1e-5), rows


def test_two_step_bfs_order_and_graph_replay(oracle):
    """A row order with far halos (the reference's BFS order of a periodic supercell) and the
    whole recurrence replayed from a HIP graph several times."""
    h2 = matrix(np.float64, 48, 32, seed=9, hilbert=False)
    r0 = oracle.random_starter(h2.shape[0], 32, np.float64)
    M = 98
    ref = oracle.diag_moments(h2, r0, M)
    ctx = abi.Ctx(0)
    try:
        ctx.set_two_step(np.float64, 512)
        ctx.set_graph_policy(2)
        mat, blk = ctx.mat(h2, fmt=1), ctx.block(r0)
        runs = [abi.moments_diagonal(mat, blk, M) for _ in range(3)]
        blk.close()
        mat.close()
    finally:
        ctx.close()
    for mu in runs:
        np.testing.assert_array_equal(mu, runs[0])
    assert np.max(np.abs(runs[0] - ref) / np.abs(ref[:, :1])) <= 1e-10


def test_two_step_off_for_unsupported_blocks(oracle):
    """Rows narrower than a 128-byte line (fp64 with 8 vectors) and CSR matrices keep single
    steps: the results are those of the single-step path."""
    h2 = matrix(np.float64, 20, 14)
    r0 = oracle.random_starter(h2.shape[0], 8, np.float64)
    M = 40
    ctx = abi.Ctx(0)
    try:
        mat8 = ctx.mat(h2, fmt=1)
        blk = ctx.block(r0)
        base = abi.moments_diagonal(mat8, blk, M)
        ctx.set_two_step(np.float64, 256)
        np.testing.assert_array_equal(abi.moments_diagonal(mat8, blk, M), base)
        csr = ctx.mat(h2, fmt=0)
        wide = ctx.block(oracle.random_starter(h2.shape[0], 16, np.float64))
        ref = oracle.diag_moments(h2, oracle.random_starter(h2.shape[0], 16, np.float64), M)
        assert np.max(np.abs(abi.moments_diagonal(csr, wide, M) - ref) / np.abs(ref[:, :1])) <= 1e-10
        for hnd in (wide, csr, blk, mat8):
            hnd.close()
        with pytest.raises(RuntimeError, match="two-step tile rows"):
            ctx.set_two_step(np.float64, 100)
        with pytest.raises(RuntimeError, match="workgroups per CU"):
            ctx.set_two_step(np.float64, 128, 65)
    finally:
        ctx.close()


def random_hoppings(h2, seed=3):
    """The same pattern with every off-diagonal value perturbed (no value-table form)."""
    h = h2.tocsr(copy=True)
    rng = np.random.default_rng(seed)
    off = h.indices != np.repeat(np.arange(h.shape[0]), np.diff(h.indptr))
    h.data[off] *= 1 + 0.05 * rng.random(int(off.sum()))
    h = ((h + h.T) * 0.5).tocsr()
    h.sort_indices()
    return h


LDS_CASES = [(np.float64, nvec) for nvec in (16, 8, 12, 20)] + [(np.complex128, nvec) for nvec in (4, 8, 6, 12)]


@pytest.mark.parametrize("shape", [(40, 28), (21, 13)])
@pytest.mark.parametrize("coded", [True, False])
@pytest.mark.parametrize("dtype,nvec", LDS_CASES)
@pytest.mark.parametrize("M", [130, 132])
def test_two_step_lds_matches_oracle(oracle, shape, coded, dtype, nvec, M):
    """The LDS form (kpmx_ctx_set_two_step_lds): 256-row tiles with the intermediate vector in LDS,
    on the value-table and the standard matrix form, tiles that end inside the matrix's last slice
    or leave waves without a slice, fp64 with 8 and 16 columns (20 vectors: a 16-column tile in the
    LDS form and a 4-column one on single steps) and complex128 with 4 and 8 (12 vectors: an
    8-column tile and a 4-column one), odd and even step counts, the resident grid and one
    workgroup per CU and the register budgets for 3 and 4: within rounding of single steps and of
    the oracle."""
    h2 = matrix(dtype, *shape)
    if not coded:
        h2 = random_hoppings(h2)
    r0 = oracle.random_starter(h2.shape[0], nvec, dtype)
    ref = oracle.diag_moments(h2, r0, M)
    ctx = abi.Ctx(0)
    try:
        if dtype == np.complex128:
            ctx.set_tile_width(dtype, 8)
        mat, blk = ctx.mat(h2, fmt=1), ctx.block(r0)
        assert (mat.coded_bytes() > 0) == coded
        single = abi.moments_diagonal(mat, blk, M)
        out = {}
        for wg in (0, 1, 3, 4):  # 1, 3, 4: register budgets for that many workgroups per CU
            ctx.set_two_step_lds(wg, dtype)
            assert ctx.two_step(dtype) == 256 and ctx.two_step_lds(dtype) == (True, wg)
            out[wg] = abi.moments_diagonal(mat, blk, M)
        ctx.set_two_step(dtype, 0)
        np.testing.assert_array_equal(abi.moments_diagonal(mat, blk, M), single)
        blk.close()
        mat.close()
    finally:
        ctx.close()
    scale = np.abs(ref[:, :1])
    assert np.max(np.abs(single - ref) / scale) <= 1e-10
    for wg, mu in out.items():
        assert np.max(np.abs(mu - ref) / scale) <= 1e-10, wg
        assert np.max(np.abs(mu - single) / scale) <= 1e-13, wg


def test_two_step_lds_bfs_order_and_graph_replay(oracle):
    """Far halos (row-major order of a periodic supercell: about 300 halo rows per 256-row tile)
    in the LDS form with 8 columns, replayed from a HIP graph several times; with 16 columns the
    halo does not fit 64 KB of LDS and the recurrence keeps single steps (the single-step result
    bit for bit); other scalar types keep single steps."""
    h2 = matrix(np.float64, 48, 32, seed=9, hilbert=False)
    r0 = oracle.random_starter(h2.shape[0], 8, np.float64)
    M = 98
    ref = oracle.diag_moments(h2, r0, M)
    r16 = oracle.random_starter(h2.shape[0], 16, np.float64)
    ctx = abi.Ctx(0)
    try:
        mat, blk, b16 = ctx.mat(h2, fmt=1), ctx.block(r0), ctx.block(r16)
        single16 = abi.moments_diagonal(mat, b16, M)
        ctx.set_two_step_lds()
        ctx.set_graph_policy(2)
        runs = [abi.moments_diagonal(mat, blk, M) for _ in range(3)]
        np.testing.assert_array_equal(abi.moments_diagonal(mat, b16, M), single16)
        b16.close()
        blk.close()
        mat.close()
        h32 = h2.astype(np.float32)
        r32 = oracle.random_starter(h2.shape[0], 32, np.float32)
        m32, b32 = ctx.mat(h32, fmt=1), ctx.block(r32)
        mu32 = abi.moments_diagonal(m32, b32, M)
        b32.close()
        m32.close()
        import ctypes
        from pybinding_amd import abi as pabi
        with pytest.raises(RuntimeError, match="fp64 or complex128 only"):
            pabi._check(pabi.lib().kpmx_ctx_set_two_step_lds(ctx.h, ctypes.c_int(0), ctypes.c_int(0)))
    finally:
        ctx.close()
    for mu in runs:
        np.testing.assert_array_equal(mu, runs[0])
    assert np.max(np.abs(runs[0] - ref) / np.abs(ref[:, :1])) <= 1e-10
    ref32 = oracle.diag_moments(h32, r32, M)
    assert np.max(np.abs(mu32 - ref32) / np.abs(ref32[:, :1])) <= 1e-4


@pytest.mark.parametrize("dtype,widths,default", [(np.float64, (8, 16), 16), (np.complex128, (4, 8), 64)])
def test_autotuner_two_step_choice(oracle, dtype, widths, default):
    """The autotuner times the LDS two-step form against single steps and keeps whichever is
    faster; either way the moments match the oracle, a later autotuning starts again from the
    defaults, and with the choice switched off the settings stay single steps."""
    h2 = matrix(dtype, 96, 64)
    r0 = oracle.random_starter(h2.shape[0], 32, dtype)
    M = 66
    ref = oracle.diag_moments(h2, r0, M)
    ctx = abi.Ctx(0)
    try:
        mat, blk = ctx.mat(h2, fmt=1), ctx.block(r0)
        abi.autotune_diagonal(mat, blk)
        chosen = ctx.two_step_lds(dtype)
        assert ctx.tile_width(dtype) in (widths if chosen[0] else (default,))
        mu = abi.moments_diagonal(mat, blk, M)
        ctx.set_autotune_two_step(False)
        abi.autotune_diagonal(mat, blk)
        assert ctx.two_step_lds(dtype) == (False, 0) and ctx.tile_width(dtype) == default
        single = abi.moments_diagonal(mat, blk, M)
        blk.close()
        mat.close()
    finally:
        ctx.close()
    scale = np.abs(ref[:, :1])
    assert np.max(np.abs(mu - ref) / scale) <= 1e-10
    assert np.max(np.abs(single - ref) / scale) <= 1e-10
