"""The benchmark workload at full size (periodic honeycomb 2048 x 1024 cells, N = 4,194,304,
Anderson W = 1 eV, fp64, 64 reference random vectors), checked through what holds at any size:
- one vector against the oracle, at a moment count the oracle finishes in seconds;
- mu_0 = N / 2 exactly for Rademacher starters, and |mu_n| <= 2 mu_0 (|T_n| <= 1 on the spectrum);
- the same columns computed in a 64-wide and in a 32-wide block agree (block-width invariance);
- the Hilbert and the BFS site orders agree, the starters being the same per original site;
- fp32 agrees with fp64 to the fp32 tolerance."""
import numpy as np
import pytest
import scipy.sparse as sp

import abi
from pybinding_amd import generators

pytestmark = pytest.mark.gpu

N1, N2, W = 2048, 1024, 1.0
BOUNDS = (-8.9, 8.9)


@pytest.fixture(scope="module")
def workload():
    from pybinding_amd._native import load_module
    native = load_module()
    h = generators.honeycomb_supercell(N1, N2, disorder=W, seed=2, dtype=np.float64)
    q = generators.honeycomb_cell_order(N1, N2)
    inv = np.argsort(q)
    hp = h[inv][:, inv].tocsr()
    hp.sort_indices()
    ip, ix, dv, _, _, _, _, _ = native.optimize(hp.shape, hp.indptr, hp.indices, hp.data, BOUNDS, False, [0], [0])
    h2 = sp.csr_matrix((dv, ix, ip), shape=h.shape)
    ip, ix, dv, rmap, _, _, _, _ = native.optimize(h.shape, h.indptr, h.indices, h.data, BOUNDS, True, [0], [0])
    h2_bfs = sp.csr_matrix((dv, ix, ip), shape=h.shape)
    ctx = abi.Ctx(0)
    yield native, ctx, h2, q.astype(np.int32), h2_bfs, np.asarray(rmap, dtype=np.int32)
    ctx.close()


def test_one_vector_against_oracle(workload, oracle):
    native, ctx, h2, q, _, _ = workload
    n = h2.shape[0]
    assert n == 2 * N1 * N2 and h2.nnz == 4 * n
    r0 = native.random_starters(n, 1, "float64", q)
    mat, blk = ctx.mat(h2, fmt=1), ctx.block(r0)
    mu = abi.moments_diagonal(mat, blk, 64)
    blk.close()
    mat.close()
    ref = oracle.diag_moments(h2, r0, 64)
    assert np.max(np.abs(mu - ref) / np.abs(ref[:, :1])) <= 1e-10


def test_block_properties(workload):
    native, ctx, h2, q, h2_bfs, rmap = workload
    n, M = h2.shape[0], 256
    r0 = native.random_starters(n, 64, "float64", q)
    mat = ctx.mat(h2, fmt=1)
    blk = ctx.block(r0)
    mu = abi.moments_diagonal(mat, blk, M)
    blk.close()
    assert np.all(mu[:, 0] == n / 2)                                  # |r|^2 / 2 of +-1 entries
    assert np.all(np.abs(mu[:, 1:]) <= 2 * mu[:, :1] * (1 + 1e-12))   # Chebyshev bound
    half = ctx.block(np.ascontiguousarray(r0[:32]))
    mu32 = abi.moments_diagonal(mat, half, M)
    half.close()
    mat.close()
    assert np.max(np.abs(mu32 - mu[:32]) / mu[:32, :1]) <= 1e-12
    # the BFS order: same starter values per original site, permuted by the BFS map
    r_bfs = native.random_starters(n, 8, "float64", rmap)
    mat_b, blk_b = ctx.mat(h2_bfs, fmt=1), ctx.block(r_bfs)
    mu_b = abi.moments_diagonal(mat_b, blk_b, M)
    blk_b.close()
    mat_b.close()
    assert np.max(np.abs(mu_b - mu[:8]) / mu[:8, :1]) <= 1e-10


def test_fp32_against_fp64(workload):
    native, ctx, h2, q, _, _ = workload
    n, M = h2.shape[0], 256
    r64 = native.random_starters(n, 16, "float64", q)
    mat, blk = ctx.mat(h2, fmt=1), ctx.block(r64)
    mu64 = abi.moments_diagonal(mat, blk, M)
    blk.close()
    mat.close()
    r32 = r64.astype(np.float32)
    mat, blk = ctx.mat(h2.astype(np.float32), fmt=1), ctx.block(r32)
    mu32 = abi.moments_diagonal(mat, blk, M)
    blk.close()
    mat.close()
    assert np.max(np.abs(mu32.astype(np.float64) - mu64) / mu64[:, :1]) <= 1e-4


def dos_window_weights(M, a, b, energy=0.5, sigma=0.1, points=4096):
    """w_n of L = integral of the Jackson-damped DOS against a Gaussian window, L = sum_n w_n mu_n."""
    xq = np.cos(np.pi * (np.arange(points) + 0.5) / points)
    win = np.exp(-0.5 * ((a * xq + b - energy) / sigma) ** 2)
    n = np.arange(M)
    g = ((M + 1 - n) * np.cos(np.pi * n / (M + 1)) + np.sin(np.pi * n / (M + 1)) / np.tan(np.pi / (M + 1))) / (M + 1)
    return g * (2 / np.pi) * np.pi / points * (np.cos(n[:, None] * np.arccos(xq)[None, :]) @ win)


def test_adjoint_full_size():
    """Strained graphene flake of ~1 M sites, 1024 moments, 16 vectors: the gradient with checkpoints
    every 64 steps equals the stored-tape gradient, and its contraction with H2 equals the central
    difference of the loss along H2 itself (two forward passes at (1 +- 1e-5) H2)."""
    from pybinding_amd.autograd import TightBinding
    n, bonds, hop, onsite, pos = generators.strained_graphene_flake(161.8, 161.8, return_positions=True)
    M, R = 1024, 16
    tb = TightBinding(n, bonds, (-9.0, 9.0), M, R, checkpoint=64, positions=(pos[:, 0], pos[:, 1]))
    try:
        h2 = tb.scaled(onsite, hop)
        w = dos_window_weights(M, *tb.scale)
        mat = tb.ctx.mat(h2, fmt=1)
        g_ck, mu = abi.adjoint_diagonal(mat, tb.block, M, w, checkpoint=64)
        g_all, _ = abi.adjoint_diagonal(mat, tb.block, M, w, checkpoint=0)
        mat.close()
        assert np.abs(g_ck - g_all).max() <= 1e-12 * np.abs(g_all).max()
        eps = 1e-5
        loss = []
        for s in (1.0, -1.0):
            m_s = tb.ctx.mat(h2 * (1.0 + s * eps), fmt=1)
            loss.append(np.dot(w, abi.moments_diagonal(m_s, tb.block, M).mean(axis=0)))
            m_s.close()
        fd = (loss[0] - loss[1]) / (2 * eps)
        ad = float(np.dot(g_all, h2.data))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)
    finally:
        tb.close()


def test_batched_full_size(oracle):
    """256 square flakes of 10^4 to 5 x 10^5 sites (31 M in all), 1024 moments, 4 vectors each, one
    graph: every system has mu_0 = N_s / 2 exactly and |mu_n| <= 2 mu_0, and the three smallest
    systems match the oracle."""
    from pybinding_amd.batch import kpm_batch
    rs = np.random.default_rng(3)
    sizes = np.exp(rs.uniform(np.log(1e4), np.log(5e5), 256))
    rw = np.random.default_rng(4)
    Ws = rw.uniform(0, 4, 256)
    hs, ranges = [], []
    for sz, Wf in zip(sizes, Ws):
        side = max(2, int(round(np.sqrt(sz))))
        hs.append(generators.square_flake(side, side, t=-1.0, disorder=Wf, rng=rw))
        ranges.append((-(4 + Wf / 2), 4 + Wf / 2))
    kb = kpm_batch(hs, ranges)
    try:
        M = 1024
        mu = kb.moments(M, 4)
    finally:
        kb.close()
    for s, h in enumerate(hs):
        assert np.all(mu[s, :, 0] == h.shape[0] / 2), s
        assert np.all(np.abs(mu[s, :, 1:]) <= 2 * mu[s, :, :1] * (1 + 1e-12)), s
    for s in np.argsort([h.shape[0] for h in hs])[:3]:
        a, b = oracle.scale_factors(*ranges[s])
        h2 = oracle.scale_hamiltonian(hs[s], a, b, np.float64)
        ref = oracle.diag_moments(h2, oracle.random_starter(hs[s].shape[0], 4, np.float64), M)
        assert np.max(np.abs(mu[s] - ref) / np.abs(ref[:, :1])) <= 1e-10, s
