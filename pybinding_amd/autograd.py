"""Differentiable KPM moments: a torch.autograd.Function over the engine's adjoint recurrence.

`TightBinding(num_sites, bonds)` fixes the sparsity pattern of a real symmetric Hamiltonian
H = diag(onsite) + sum_e hopping[e] (|i_e><j_e| + |j_e><i_e|).  `kpm_moments(model, onsite,
hopping)` returns the mean diagonal KPM moments of the model's random starters (the reference's
DOS moments: random_starter, cppcore/src/kpm/Starter.cpp:48-83; collectors, cppcore/src/kpm/
default/collectors.cpp:5-34, mu_0 halved) as a torch tensor that carries gradients back to
`onsite` and `hopping`.  The backward pass is the engine's adjoint Chebyshev recurrence
(kpmx_adjoint_backward) with vectors recomputed between checkpoints; no autodiff graph of the
recurrence is built.  Any loss computed from the moments in torch (e.g. a DOS window, which is
linear in the moments) differentiates through it.
"""
import numpy as np
import torch

from . import abi
from ._native import load_module

__all__ = ["TightBinding", "kpm_moments", "KPMMomentsFunction"]


class TightBinding:
    """Sparsity pattern, scaling and starters of a real symmetric tight-binding Hamiltonian.

    Parameters
    ----------
    num_sites : int
    bonds : (E, 2) int array of site pairs (i, j), i != j, each pair once
    energy_range : (float, float)  explicit spectrum bounds (fixed scale factors)
    num_moments, num_random : int
    checkpoint : int   steps between stored vector pairs in the adjoint (<= 0: store all)
    dtype : np.float64 or np.float32
    positions : optional (x, y) site coordinates; when given, the GPU works in a Hilbert-curve
        order of the sites (gathers hit the L2 cache), with the starters permuted accordingly, so
        the moments and the gradients (returned per original site and bond) are unchanged
    """

    def __init__(self, num_sites, bonds, energy_range, num_moments, num_random=1, checkpoint=64,
                 dtype=np.float64, device=0, positions=None):
        self.n = int(num_sites)
        bonds = np.asarray(bonds, dtype=np.int64).reshape(-1, 2)
        self.num_bonds = len(bonds)
        self.energy_range = (float(energy_range[0]), float(energy_range[1]))
        self.num_moments = int(num_moments)
        self.num_random = int(num_random)
        self.checkpoint = int(checkpoint)
        self.dtype = np.dtype(dtype)
        # site order on the GPU: q[old site] = new index
        if positions is not None:
            from .ordering import locality_order
            q = np.asarray(locality_order(positions[0], positions[1]), dtype=np.int64)
        else:
            q = np.arange(self.n)
        self.reorder_map = q
        old_of_new = np.empty(self.n, dtype=np.int64)
        old_of_new[q] = np.arange(self.n)
        # entries: diagonal (kind 0), bond i->j (kind 1), bond j->i (kind 2); src = original site / bond
        d = np.arange(self.n)
        rows = np.concatenate([d, q[bonds[:, 0]], q[bonds[:, 1]]])
        cols = np.concatenate([d, q[bonds[:, 1]], q[bonds[:, 0]]])
        self.src = np.concatenate([old_of_new, np.arange(self.num_bonds), np.arange(self.num_bonds)])
        self.kind = np.concatenate([np.zeros(self.n, int), np.ones(self.num_bonds, int), np.full(self.num_bonds, 2)])
        order = np.lexsort((cols, rows))
        self.rows, self.cols = rows[order], cols[order]
        self.src, self.kind = self.src[order], self.kind[order]
        if np.any((np.diff(self.rows) == 0) & (np.diff(self.cols) == 0)):
            raise ValueError("duplicate bonds")
        self.indptr = np.concatenate([[0], np.cumsum(np.bincount(self.rows, minlength=self.n))]).astype(np.int32)
        self.indices = self.cols.astype(np.int32)
        self.native = load_module()
        self.ctx = abi.Ctx(device)
        r0 = self.native.random_starters(self.n, self.num_random, self.dtype.name,
                                         [] if positions is None else [int(v) for v in q])
        self.block = self.ctx.block(r0)
        self.scale = None

    def values(self, onsite, hopping):
        onsite = np.asarray(onsite, dtype=np.float64)
        hopping = np.asarray(hopping, dtype=np.float64)
        out = np.empty(len(self.src), dtype=np.float64)
        diag = self.kind == 0
        out[diag] = onsite[self.src[diag]]
        out[~diag] = hopping[self.src[~diag]]
        return out.astype(self.dtype)

    def scaled(self, onsite, hopping):
        """H2 = (H - b) 2/a, in the pattern's CSR order (the host core's create_scaled)."""
        data = self.values(onsite, hopping)
        ip, ix, dv, _, _, _, _, (a, b) = self.native.optimize((self.n, self.n), self.indptr, self.indices, data,
                                                              self.energy_range, False, [0], [0])
        assert np.array_equal(ix, self.indices), "scaling changed the sparsity pattern"
        self.scale = (a, b)
        import scipy.sparse as sp
        return sp.csr_matrix((dv, ix, ip), shape=(self.n, self.n))

    def close(self):
        self.block.close()
        self.ctx.close()


class KPMMomentsFunction(torch.autograd.Function):
    # The tape (device memory: every vector, or the checkpoint pairs) lives as long as this node
    # of the autograd graph: it is freed when the graph is released -- after a backward pass
    # without retain_graph, or when the output is dropped without one -- and right away when no
    # input needs a gradient.
    @staticmethod
    def forward(ctx, onsite, hopping, tb):
        h2 = tb.scaled(onsite.detach().cpu().numpy(), hopping.detach().cpu().numpy())
        mat = tb.ctx.mat(h2, fmt=1)
        tape = abi.Tape(mat, tb.block, tb.num_moments, tb.checkpoint)
        mu = torch.from_numpy(tape.mu.astype(np.float64).mean(axis=0))
        if any(ctx.needs_input_grad[:2]):
            ctx.tb, ctx.mat, ctx.tape = tb, mat, tape
            ctx.scale = tb.scale
        else:
            tape.close()
            mat.close()
        return mu

    @staticmethod
    def backward(ctx, grad_mu):
        tb = ctx.tb
        g2 = ctx.tape.backward(grad_mu.detach().cpu().numpy().astype(np.float64))
        a, _ = ctx.scale
        gh = g2 * (2.0 / a)  # H2 = (H - b) 2/a
        g_on = np.zeros(tb.n)
        g_hop = np.zeros(tb.num_bonds)
        diag = tb.kind == 0
        np.add.at(g_on, tb.src[diag], gh[diag])
        np.add.at(g_hop, tb.src[~diag], gh[~diag])
        return torch.from_numpy(g_on), torch.from_numpy(g_hop), None


def kpm_moments(tb, onsite, hopping):
    """Mean diagonal moments (num_moments,) of the model `tb` at (onsite, hopping), differentiable."""
    return KPMMomentsFunction.apply(onsite, hopping, tb)
