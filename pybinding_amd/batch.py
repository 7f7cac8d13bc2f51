"""Many independent KPM calculations (e.g. disorder realizations of uneven size) in one engine run.

`kpm_batch(hamiltonians, energy_ranges)` builds each system's scaled H2 exactly like `kpm()` does
for one system (the reference's Scale and create_scaled, cppcore/include/kpm/Bounds.hpp:10-33,
cppcore/src/kpm/OptimizedHamiltonian.cpp:54-71), draws each system's random starters from its own
default-seeded generator (the reference's random_starter, cppcore/src/kpm/Starter.cpp:48-83), and
computes every system's DOS moments with one block-diagonal recurrence on the GPU (the whole step
sequence captured in one HIP graph).  Per system the result equals a separate `kpm(...)` run.
"""
import numpy as np
import scipy.sparse as sp

from . import abi
from ._native import load_module

__all__ = ["kpm_batch", "KPMBatch"]


class KPMBatch:
    def __init__(self, hamiltonians, energy_ranges, device=0):
        native = load_module()
        self.native = native
        hs = [sp.csr_matrix(h) for h in hamiltonians]
        if len(energy_ranges) != len(hs):
            raise ValueError("one energy range per system is required")
        self.scales = []
        h2s = []
        for h, (lo, hi) in zip(hs, energy_ranges):
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
        r0 = np.concatenate([self.native.random_starters(n, num_random, self.dtype.name, []) for n in self.sizes],
                            axis=1)
        self.batch.starters(r0)

    def autotune(self, num_random=1):
        """Time the batched step over several slice partitions and keep the fastest; returns the
        chosen workgroup count."""
        if self.batch.nvec != num_random:
            self.set_random_starters(num_random)
        return self.batch.autotune()

    def moments(self, num_moments, num_random=1, use_graph=True):
        """Undamped diagonal moments, shape (num_systems, num_random, num_moments)."""
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


def kpm_batch(hamiltonians, energy_ranges, device=0):
    return KPMBatch(hamiltonians, energy_ranges, device)
