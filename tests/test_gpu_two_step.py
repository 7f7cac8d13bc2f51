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
