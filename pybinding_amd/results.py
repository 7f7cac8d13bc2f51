"""Result containers returned by the KPM methods (pybinding/results.py:160-215: `Series`)."""
from copy import copy

import numpy as np

__all__ = ['Series', 'StructureMap', 'Deferred']


class Series:
    """Values sampled along one independent axis, e.g. a DOS over energies.

    Attributes
    ----------
    variable : array_like
        The sample points (energies, chemical potentials, ...).
    data : array_like
        One value per sample point, or one row per sample point with several columns
        (orbitals, sites).
    labels : dict
        Axis names for plotting: 'variable', 'data', 'title' and 'columns'.
    """

    def __init__(self, variable, data, labels=None):
        self.variable = np.atleast_1d(variable)
        self.data = np.atleast_1d(data)
        defaults = dict(variable="x", data="y", columns="")
        defaults.update(labels or {})
        self.labels = defaults

    def with_data(self, data):
        """Shallow copy with `data` replaced (same variable and labels)"""
        result = copy(self)
        result.data = data
        return result

    def reduced(self):
        """Shallow copy whose data is the row-wise sum over its columns"""
        return self.with_data(self.data.sum(axis=1))


class StructureMap:
    """Per-site values over the sites of a system: the data beside the sites' positions and
    sublattices (the interface of pybinding/results.py:217-300,387-420 that `SpatialLDOS` uses).
    Indexing (a mask, indices, or a shape) selects sites and their values together."""

    def __init__(self, data, system):
        data = np.atleast_1d(data)
        if data.size != system.num_sites:
            raise RuntimeError("Data size doesn't match number of sites")
        self._data = data
        self._system = system

    @property
    def data(self):
        """One value per site"""
        return self._data

    @data.setter
    def data(self, value):
        self._data = value

    @property
    def num_sites(self):
        return self._system.num_sites

    @property
    def positions(self):
        return self._system.positions

    @property
    def x(self):
        return self._system.x

    @property
    def y(self):
        return self._system.y

    @property
    def z(self):
        return self._system.z

    @property
    def sublattices(self):
        return self._system.sublattices

    @property
    def sub(self):
        """Sublattice name of every site"""
        return self._system.sublattices_str

    def with_data(self, data):
        """Shallow copy mapping `data` to the same sites"""
        result = copy(self)
        result._data = np.atleast_1d(data)
        return result

    def find_nearest(self, position, sublattice=""):
        return self._system.find_nearest(position, sublattice)

    def __getitem__(self, idx):
        if hasattr(idx, "contains"):
            idx = idx.contains(*self.positions)
        return StructureMap(self._data[idx], self._system[idx])


class Deferred:
    """A computation evaluated on demand (the reference's `Deferred` job for parallel sweeps)."""

    def __init__(self, fn):
        self._fn = fn
        self._result = None

    def compute(self):
        if self._result is None:
            self._result = self._fn()

    @property
    def result(self):
        self.compute()
        return self._result
