"""GPU adjoint gradient against central finite differences of the fp64 oracle.

L = sum_n w_n mean_c mu_n^c; the engine's dL/dH2 (kpmx_adjoint_*) and the torch gradients with
respect to onsite energies and hoppings must agree with finite differences of the oracle's
moments to 1e-6 relative; stored-everything and checkpointed runs must agree with each other.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from pybinding_amd import abi, generators

pytestmark = pytest.mark.gpu


def strained_system(seed=0):
    h = generators.honeycomb_flake(4, 3)
    coo = sp.triu(h, k=1).tocoo()
    rng = np.random.default_rng(seed)
    bonds = np.stack([coo.row, coo.col], axis=1)
    hop = -2.8 * np.exp(-3.37 * rng.uniform(-0.05, 0.05, len(bonds)))
    onsite = rng.uniform(-0.3, 0.3, h.shape[0])
    return h.shape[0], bonds, onsite, hop


def loss_oracle(oracle, n, bonds, onsite, hop, w, M, R, erange):
    hm = sp.coo_matrix((np.r_[onsite, hop, hop], (np.r_[np.arange(n), bonds[:, 0], bonds[:, 1]],
                                                   np.r_[np.arange(n), bonds[:, 1], bonds[:, 0]])), shape=(n, n)).tocsr()
    a, b = oracle.scale_factors(*erange)
    h2 = oracle.scale_hamiltonian(hm, a, b, np.float64)
    mu = oracle.diag_moments(h2, oracle.random_starter(n, R, np.float64), M).mean(axis=0)
    return float(np.dot(w, mu)), mu


@pytest.mark.parametrize("checkpoint", [0, 4, 64])
def test_torch_gradient_matches_finite_differences(oracle, checkpoint):
    torch = pytest.importorskip("torch")
    from pybinding_amd.autograd import TightBinding, kpm_moments
    n, bonds, onsite, hop = strained_system()
    M, R, erange = 34, 3, (-9.0, 9.0)
    w = np.random.default_rng(1).standard_normal(M)
    tb = TightBinding(n, bonds, erange, M, R, checkpoint=checkpoint)
    on_t = torch.tensor(onsite, requires_grad=True)
    hop_t = torch.tensor(hop, requires_grad=True)
    mu = kpm_moments(tb, on_t, hop_t)
    L = (mu * torch.from_numpy(w)).sum()
    L.backward()
    _, mu_ref = loss_oracle(oracle, n, bonds, onsite, hop, w, M, R, erange)
    assert np.max(np.abs(mu.detach().numpy() - mu_ref)) <= 1e-10 * abs(mu_ref[0])
    eps = 1e-5
    rng = np.random.default_rng(2)
    for which, idx in [("onsite", i) for i in rng.choice(n, 6, replace=False)] + \
                      [("hop", i) for i in rng.choice(len(hop), 6, replace=False)]:
        p_on, m_on = onsite.copy(), onsite.copy()
        p_hop, m_hop = hop.copy(), hop.copy()
        if which == "onsite":
            p_on[idx] += eps
            m_on[idx] -= eps
            got = on_t.grad[idx].item()
        else:
            p_hop[idx] += eps
            m_hop[idx] -= eps
            got = hop_t.grad[idx].item()
        lp, _ = loss_oracle(oracle, n, bonds, p_on, p_hop, w, M, R, erange)
        lm, _ = loss_oracle(oracle, n, bonds, m_on, m_hop, w, M, R, erange)
        fd = (lp - lm) / (2 * eps)
        assert abs(got - fd) <= 1e-6 * max(abs(fd), 1e-3), (which, idx, got, fd)
    tb.close()


def test_checkpointed_equals_stored(oracle):
    n, bonds, onsite, hop = strained_system(3)
    M = 66
    hm = sp.coo_matrix((np.r_[onsite, hop, hop], (np.r_[np.arange(n), bonds[:, 0], bonds[:, 1]],
                                                   np.r_[np.arange(n), bonds[:, 1], bonds[:, 0]])), shape=(n, n)).tocsr()
    a, b = oracle.scale_factors(-9, 9)
    h2 = oracle.scale_hamiltonian(hm, a, b, np.float64)
    h2.sort_indices()
    r0 = oracle.random_starter(n, 5, np.float64)
    w = np.random.default_rng(4).standard_normal(M)
    ctx = abi.Ctx(0)
    mat = ctx.mat(h2, fmt=1)
    blk = ctx.block(r0)
    g_all, mu = abi.adjoint_diagonal(mat, blk, M, w, checkpoint=0)
    g_ck, _ = abi.adjoint_diagonal(mat, blk, M, w, checkpoint=5)
    ref = oracle.diag_moments(h2, r0, M)
    assert np.max(np.abs(mu - ref) / np.abs(ref[:, :1])) <= 1e-10
    np.testing.assert_allclose(g_ck, g_all, rtol=1e-11, atol=1e-13 * np.abs(g_all).max())
    # a checkpointed tape run backward repeatedly (its rebuilt vectors reuse three slots), with
    # another loss in between, gives the same gradient each time; the checkpoint interval that
    # leaves the last segment one step long (S = 33 = 4 * 8 + 1) is covered too
    for C in (5, 8, 2):
        tape = abi.Tape(mat, blk, M, C)
        w2 = np.random.default_rng(C).standard_normal(M)
        g1 = tape.backward(w)
        g_other = tape.backward(w2)
        g2 = tape.backward(w)
        tape.close()
        assert np.array_equal(g1, g2), C
        np.testing.assert_allclose(g1, g_all, rtol=1e-11, atol=1e-13 * np.abs(g_all).max())
        g_other_ref, _ = abi.adjoint_diagonal(mat, blk, M, w2, checkpoint=0)
        np.testing.assert_allclose(g_other, g_other_ref, rtol=1e-11, atol=1e-13 * np.abs(g_other_ref).max())
    # dL/dH2 entry-wise against finite differences of the oracle on a few stored entries
    eps = 1e-6
    for k in np.random.default_rng(5).choice(h2.nnz, 5, replace=False):
        hp, hm2 = h2.copy(), h2.copy()
        hp.data[k] += eps
        hm2.data[k] -= eps
        lp = np.dot(w, oracle.diag_moments(hp, r0, M).mean(axis=0))
        lm = np.dot(w, oracle.diag_moments(hm2, r0, M).mean(axis=0))
        fd = (lp - lm) / (2 * eps)
        assert abs(g_all[k] - fd) <= 1e-6 * max(abs(fd), 1e-3), (k, g_all[k], fd)
    blk.close()
    mat.close()
    ctx.close()


def test_site_order_does_not_change_gradient():
    """TightBinding(positions=...) runs in a Hilbert order of the sites; the moments and the
    gradients per original site and bond are the same as in the given order."""
    torch = pytest.importorskip("torch")
    from pybinding_amd.autograd import TightBinding, kpm_moments
    h, pos = generators.honeycomb_flake(5, 4, return_positions=True)
    coo = sp.triu(h, k=1).tocoo()
    bonds = np.stack([coo.row, coo.col], axis=1)
    rng = np.random.default_rng(8)
    hop = -2.8 * np.exp(-3.37 * rng.uniform(-0.05, 0.05, len(bonds)))
    onsite = rng.uniform(-0.3, 0.3, h.shape[0])
    M, R = 42, 4
    w = torch.from_numpy(np.random.default_rng(9).standard_normal(M))
    out = []
    for positions in (None, (pos[:, 0], pos[:, 1])):
        tb = TightBinding(h.shape[0], bonds, (-9.0, 9.0), M, R, checkpoint=8, positions=positions)
        on_t = torch.tensor(onsite, requires_grad=True)
        hop_t = torch.tensor(hop, requires_grad=True)
        mu = kpm_moments(tb, on_t, hop_t)
        (mu * w).sum().backward()
        out.append((mu.detach().numpy(), on_t.grad.numpy().copy(), hop_t.grad.numpy().copy()))
        tb.close()
    np.testing.assert_allclose(out[1][0], out[0][0], rtol=0, atol=1e-12 * abs(out[0][0][0]))
    for k in (1, 2):
        np.testing.assert_allclose(out[1][k], out[0][k], rtol=1e-10, atol=1e-12 * np.abs(out[0][k]).max())


@pytest.mark.parametrize("nvec", [8, 16, 64])
def test_adjoint_variants_agree(oracle, nvec):
    """The adjoint step's kernel variants (LDS staging, wave staging, fp64 column pairs; several
    grids and tile shapes) give the same gradient, checkpointed and stored, and the forward moments
    of the oracle."""
    n, bonds, onsite, hop = strained_system(7)
    M = 42
    hm = sp.coo_matrix((np.r_[onsite, hop, hop], (np.r_[np.arange(n), bonds[:, 0], bonds[:, 1]],
                                                   np.r_[np.arange(n), bonds[:, 1], bonds[:, 0]])), shape=(n, n)).tocsr()
    a, b = oracle.scale_factors(-9, 9)
    h2 = oracle.scale_hamiltonian(hm, a, b, np.float64)
    h2.sort_indices()
    r0 = oracle.random_starter(n, nvec, np.float64)
    w = np.random.default_rng(11).standard_normal(M)
    ctx = abi.Ctx(0)
    mat, blk = ctx.mat(h2, fmt=1), ctx.block(r0)
    try:
        ref_g, mu = abi.adjoint_diagonal(mat, blk, M, w, checkpoint=0)
        ref = oracle.diag_moments(h2, r0, M)
        assert np.max(np.abs(mu - ref) / np.abs(ref[:, :1])) <= 1e-10
        scale = np.abs(ref_g).max()
        for variant in (0, 1, 2, 15, 24, 30, 31, 32, 39, 44):
            ctx.set_adjoint_variant(variant)
            for ck in (0, 6):
                g, _ = abi.adjoint_diagonal(mat, blk, M, w, checkpoint=ck)
                assert np.abs(g - ref_g).max() <= 1e-11 * scale, (variant, ck)
    finally:
        ctx.set_adjoint_variant(-1)
        blk.close()
        mat.close()
        ctx.close()
