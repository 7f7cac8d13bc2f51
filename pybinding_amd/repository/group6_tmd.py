"""Group 6 transition metal dichalcogenides, nearest-neighbour 3-band model
(pybinding/repository/group6_tmd.py:7-101; parameters from doi:10.1103/PhysRevB.88.085433, kept
identical to the reference's table)."""
import math
import re

from ..lattice import Lattice

_params = {  # a, eps1, eps2, t0, t1, t2, t11, t12, t22
    "MoS2": [0.3190, 1.046, 2.104, -0.184, 0.401, 0.507, 0.218, 0.338, 0.057],
    "WS2": [0.3191, 1.130, 2.275, -0.206, 0.567, 0.536, 0.286, 0.384, -0.061],
    "MoSe2": [0.3326, 0.919, 2.065, -0.188, 0.317, 0.456, 0.211, 0.290, 0.130],
    "WSe2": [0.3325, 0.943, 2.179, -0.207, 0.457, 0.486, 0.263, 0.329, 0.034],
    "MoTe2": [0.3557, 0.605, 1.972, -0.169, 0.228, 0.390, 0.207, 0.239, 0.252],
    "WTe2": [0.3560, 0.606, 2.102, -0.175, 0.342, 0.410, 0.233, 0.270, 0.190],
}


def monolayer_3band(name, override_params=None):
    """One metal site with three d orbitals per cell; 3x3 hopping matrices to the nearest metals"""
    params = dict(_params)
    if override_params:
        params.update(override_params)
    a, eps1, eps2, t0, t1, t2, t11, t12, t22 = params[name]
    rt3 = math.sqrt(3)
    lat = Lattice(a1=[a, 0], a2=[1 / 2 * a, rt3 / 2 * a])
    metal, _ = re.findall("[A-Z][a-z]*", name)
    lat.add_one_sublattice(metal, [0, 0], [eps1, eps2, eps2])
    h1 = [[t0, -t1, t2],
          [t1, t11, -t12],
          [t2, t12, t22]]
    h2 = [[t0, 1 / 2 * t1 + rt3 / 2 * t2, rt3 / 2 * t1 - 1 / 2 * t2],
          [-1 / 2 * t1 + rt3 / 2 * t2, 1 / 4 * t11 + 3 / 4 * t22, rt3 / 4 * (t11 - t22) - t12],
          [-rt3 / 2 * t1 - 1 / 2 * t2, rt3 / 4 * (t11 - t22) + t12, 3 / 4 * t11 + 1 / 4 * t22]]
    h3 = [[t0, -1 / 2 * t1 - rt3 / 2 * t2, rt3 / 2 * t1 - 1 / 2 * t2],
          [1 / 2 * t1 - rt3 / 2 * t2, 1 / 4 * t11 + 3 / 4 * t22, rt3 / 4 * (t22 - t11) + t12],
          [-rt3 / 2 * t1 - 1 / 2 * t2, rt3 / 4 * (t22 - t11) - t12, 3 / 4 * t11 + 1 / 4 * t22]]
    lat.add_hoppings(([1, 0], metal, metal, h1), ([0, -1], metal, metal, h2), ([1, -1], metal, metal, h3))
    return lat
