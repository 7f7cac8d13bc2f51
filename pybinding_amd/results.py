"""Result containers returned by the KPM methods (pybinding/results.py:160-215: `Series`)."""
from copy import copy

import numpy as np

__all__ = ['Series', 'Deferred']


class Series:
    """A series of data points determined by a common relation, i.e. :math:`y = f(x)`

    Attributes
    ----------
    variable : array_like
        Independent variable for which the data was computed.
    data : array_like
        An array of values computed as a function of `variable` (1D, or 2D with one column per
        function).
    labels : dict
        Plot labels: 'variable', 'data', 'title' and 'columns'.
    """

    def __init__(self, variable, data, labels=None):
        self.variable = np.atleast_1d(variable)
        self.data = np.atleast_1d(data)
        defaults = dict(variable="x", data="y", columns="")
        defaults.update(labels or {})
        self.labels = defaults

    def with_data(self, data):
        """Return a copy of this result object with different data"""
        result = copy(self)
        result.data = data
        return result

    def reduced(self):
        """Return a copy where the data is summed over the columns"""
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
