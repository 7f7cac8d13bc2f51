"""Many independent KPM calculations (e.g. disorder realizations of uneven size) in one engine run.

`kpm_batch(hamiltonians, energy_ranges)` builds each system's scaled H2 exactly like `kpm()` does
for one system (the reference's Scale and create_scaled, cppcore/include/kpm/Bounds.hpp:10-33,
cppcore/src/kpm/OptimizedHamiltonian.cpp:54-71), draws each system's random starters from its own
default-seeded generator (the reference's random_starter, cppcore/src/kpm/Starter.cpp:48-83), and
computes every system's DOS moments with one block-diagonal recurrence on the GPU (the whole step
sequence captured in one HIP graph).  Per system the result equals a separate `kpm(...)` run.

Like `kpm()` for whole-system results, each system whose site positions are known (a `Model`, or
`positions` given) is run in the Hilbert-curve order of its sites (`ordering.locality_order`), its
starters drawn in the original site order and moved along; the moments do not depend on the order
beyond rounding, and the gathered rows of every sweep stay in the XCD's L2.
"""
import numpy as np
import scipy.sparse as sp

from . import abi
from ._native import load_module

__all__ = ["kpm_batch", "KPMBatch"]


def _system_parts(system, pos):
    """(Hamiltonian, positions or None) of one system: a Model or a sparse matrix."""
    if hasattr(system, "hamiltonian") and hasattr(system, "system"):
        st = system.system
        x, y, _ = st.expanded_positions if st.is_multiorbital else st.positions
        return sp.csr_matrix(system.hamiltonian), (x, y) if pos is None else pos
    return sp.csr_matrix(system), pos


class KPMBatch:
    def __init__(self, hamiltonians, energy_ranges, device=0, positions=None, site_order="auto"):
        """hamiltonians: Models or square sparse matrices, one per system.  positions: optional
        (x, y) per system (None entries allowed) for matrices.  site_order: "auto" (Hilbert order
        of the positions where known) or None (the given order)."""
        native = load_module()
        self.native = native
        if positions is None:
            positions = [None] * len(hamiltonians)
        if len(positions) != len(hamiltonians):
            raise ValueError("one positions entry per system is required")
        if site_order not in ("auto", None):
            raise ValueError("site_order must be 'auto' or None")
        parts = [_system_parts(h, p) for h, p in zip(hamiltonians, positions)]
        hs = [h for h, _ in parts]
        if len(energy_ranges) != len(hs):
            raise ValueError("one energy range per system is required")
        from .ordering import locality_order
        self.orders = []  # per system: new index of every site, or None
        for h, pos in parts:
            if site_order == "auto" and pos is not None:
                q = locality_order(pos[0], pos[1])
                if q.size != h.shape[0]:
                    raise ValueError("positions do not match the Hamiltonian's size")
                self.orders.append(q)
            else:
                self.orders.append(None)
        self.scales = []
        h2s = []
        for h, (lo, hi), q in zip(hs, energy_ranges, self.orders):
            if q is not None:
                inv = np.argsort(q)
                h = h[inv][:, inv].tocsr()
            else:
                h = h.copy()
            h.sort_indices()
            ip, ix, dv, _, _, _, _, (a, b) = native.optimize(h.shape, h.indptr, h.indices, h.data, (lo, hi), False,
                                                             [0], [0])
            h2s.append(sp.csr_matrix((dv, ix, ip), shape=h.shape))
            self.scales.append((a, b))
        self.dtype = np.dtype(h2s[0].dtype)
        self.sizes = [h.shape[0] for h in h2s]
        self.ctx = abi.Ctx(device)
        self.batch = abi.Batch(self.ctx, h2s)

    def set_random_starters(self, num_random):
        """Each system's first `num_random` reference random starters."""
        r0 = np.concatenate([self.native.random_starters(n, num_random, self.dtype.name, q if q is not None else [])
                             for n, q in zip(self.sizes, self.orders)], axis=1)
        self.batch.starters(r0)

    def autotune(self, num_random=1):
        """Time the batched step over several slice partitions and keep the fastest; returns the
        chosen (workgroups, walk_chunk), walk_chunk 0 for contiguous ranges and > 0 for the
        XCD-aware walk."""
        if self.batch.nvec != num_random:
            self.set_random_starters(num_random)
        return self.batch.autotune()

    def set_partition(self, num_workgroups, walk_chunk=0):
        """Fix the slice partition of the batched step (see `abi.Batch.set_partition`)."""
        self.batch.set_partition(num_workgroups, walk_chunk)

    def set_tile_width(self, width):
        """Run the vectors as column tiles of at most `width` vectors (0: the library default),
        each tile its own recurrence over the whole batch; re-uploads the current starters."""
        self.batch.set_tile_width(width)
        if self.batch.nvec:
            self.set_random_starters(self.batch.nvec)

    def set_value_table(self, enable):
        """Run the step on the value-table matrix form (1-byte codes into per-system tables of
        off-diagonal values, dense diagonal; the default where every system fits it) or on the
        standard form (see `abi.Batch.set_value_table`)."""
        self.batch.set_value_table(enable)

    @property
    def matrix_format(self):
        """(value table available, value table in use, matrix bytes one step reads)."""
        return self.batch.format()

    def moments(self, num_moments, num_random=1, use_graph=True):
        """Undamped diagonal moments, shape (num_systems, num_random, num_moments).  use_graph:
        True (the step sequence replayed from one HIP graph), False (one launch per step) or
        "persistent" (one kernel launch for every step, grid barriers between them)."""
        if self.batch.nvec != num_random:
            self.set_random_starters(num_random)
        self.batch.run(num_moments, use_graph)
        return self.batch.fetch(num_moments)

    def dos_moments(self, num_moments, num_random=1, use_graph=True):
        """Mean over the random vectors of each system: (num_systems, num_moments)."""
        return self.moments(num_moments, num_random, use_graph).mean(axis=1)

    def close(self):
        self.batch.close()
        self.ctx.close()


def kpm_batch(hamiltonians, energy_ranges, device=0, positions=None, site_order="auto"):
    return KPMBatch(hamiltonians, energy_ranges, device, positions, site_order)
