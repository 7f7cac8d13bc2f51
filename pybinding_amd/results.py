"""Result containers returned by the KPM methods (pybinding/results.py:160-215: `Series`)."""
from copy import copy

import numpy as np

__all__ = ['Series', 'Deferred']


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
