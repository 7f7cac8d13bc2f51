"""Tight-binding model container used by `kpm()` (pybinding/model.py:14-162).

`Model` holds the Hamiltonian CSR and the `System` (site positions, sublattice of each site,
orbitals per site) that the KPM position-based methods need.
"""
import numpy as np
import scipy.sparse as sp

__all__ = ['Model', 'System']


class System:
    """Site positions and sublattice ids of a built model (pybinding/system.py)."""

    def __init__(self, x, y, z=None, sublattices=None, sublattice_names=None, norb=None):
        self.hopping_blocks = {}  # hopping family name -> (rows, cols) of the site pairs
        self.boundaries = []      # periodic models: (shift in nm, {family: (rows, cols)}) per translation
        self.x = np.asarray(x, dtype=np.float32)
        self.y = np.asarray(y, dtype=np.float32)
        self.z = np.zeros_like(self.x) if z is None else np.asarray(z, dtype=np.float32)
        n = self.x.size
        self.sublattices = np.zeros(n, dtype=np.int32) if sublattices is None else np.asarray(sublattices, np.int32)
        self.sublattice_names = list(sublattice_names or ["A"])
        self.norb = np.ones(n, dtype=np.int32) if norb is None else np.asarray(norb, np.int32)
        self._offsets = np.concatenate([[0], np.cumsum(self.norb)])

    @property
    def num_sites(self):
        return self.x.size

    @property
    def positions(self):
        return self.x, self.y, self.z

    @property
    def is_multiorbital(self):
        return bool(np.any(self.norb > 1))

    @property
    def sublattices_str(self):
        return np.array(self.sublattice_names, dtype=object)[self.sublattices]

    @property
    def expanded_positions(self):
        """Positions repeated once per orbital (System::expanded_positions)"""
        rep = self.norb
        return np.repeat(self.x, rep), np.repeat(self.y, rep), np.repeat(self.z, rep)

    @property
    def hamiltonian_size(self):
        return int(self._offsets[-1])

    def find_nearest(self, position, sublattice=""):
        """Index of the site nearest to `position` (System::find_nearest,
        cppcore/src/system/System.cpp:67-81: float32 distances, the first site wins ties),
        optionally restricted to one sublattice."""
        p = np.zeros(3, dtype=np.float32)
        p[:len(position)] = np.asarray(position, dtype=np.float32)
        dx, dy, dz = self.x - p[0], self.y - p[1], self.z - p[2]
        d = np.sqrt(dx * dx + dy * dy + dz * dz)
        if sublattice:
            sid = self.sublattice_names.index(sublattice)
            d = np.where(self.sublattices == sid, d, np.float32(np.inf))
        return int(np.argmin(d))

    def to_hamiltonian_indices(self, site):
        """Hamiltonian indices of all orbitals of `site`"""
        return list(range(int(self._offsets[site]), int(self._offsets[site + 1])))

    def with_data(self, data):
        """The sites of this system with one value each (a StructureMap)"""
        from .results import StructureMap
        return StructureMap(data, self)

    @property
    def sub(self):
        """Sublattice name of every site"""
        return self.sublattices_str

    def __getitem__(self, mask):
        idx = np.nonzero(mask)[0] if np.asarray(mask).dtype == bool else np.asarray(mask)
        return System(self.x[idx], self.y[idx], self.z[idx], self.sublattices[idx], self.sublattice_names,
                      self.norb[idx])


class Model:
    """A built tight-binding model: Hamiltonian + System.

    `Model(lattice, shape, *modifiers)` builds it like `pb.Model` (pybinding/model.py:14-70) in
    the reference's site order; `Model(hamiltonian, system)` wraps an existing CSR matrix.
    """

    def __init__(self, hamiltonian, *args):
        from .lattice import Lattice, Polygon, Primitive, build
        if isinstance(hamiltonian, Lattice):
            self.lattice = hamiltonian
            shapes = [a for a in args if isinstance(a, (Polygon, Primitive))]
            if len(shapes) > 1:
                raise RuntimeError("a model takes one shape")
            # no shape: a single unit cell, as the reference's default Primitive(1, 1, 1)
            self.shape = shapes[0] if shapes else Primitive()
            self.modifiers = [a for a in args if not isinstance(a, (Polygon, Primitive))]
            self.hamiltonian, self.system = build(self.lattice, self.shape, *self.modifiers)
        else:
            self.lattice = None
            self.hamiltonian = sp.csr_matrix(hamiltonian)
            self.system = args[0] if args else None

    @property
    def raw_hamiltonian(self):
        return self.hamiltonian
