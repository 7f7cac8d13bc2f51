"""Graphene lattice and modifiers (pybinding/repository/graphene/lattice.py:6-63,
pybinding/repository/graphene/modifiers.py:54-137, constants.py).  The parameter values and the
hopping table are the reference's, kept identical so that models built here are the same models."""
from math import sqrt, pi

import numpy as np

from ..lattice import Lattice, hopping_energy_modifier, site_position_modifier

a = 0.24595     # [nm] unit cell length
a_cc = 0.142    # [nm] carbon-carbon distance
t = -2.8        # [eV] nearest neighbour hopping
t_nn = 0.1      # [eV] next-nearest neighbour hopping
beta = 3.37     # strain hopping modulation
hbar = 6.58211899 * 10**-16   # [eV*s] (pybinding/constants.py)
phi0 = 2 * pi * hbar          # [V*s] magnetic flux quantum


# Hopping shells of the honeycomb lattice (lattice vectors a1 = (a, 0), a2 = (a/2, a sqrt(3)/2),
# A below B on the y axis): shell s -> (energy name, [(relative index, from, to), ...]).
# Shell 1: the 3 A-B bonds of length a_cc; shell 2: the 6 A-A / B-B bonds of length a (each listed
# once, the opposite direction is implied); shell 3: the 3 A-B bonds of length 2 a_cc.
HOPPING_SHELLS = {
    1: ("t", [((0, 0), "A", "B"), ((1, -1), "A", "B"), ((0, -1), "A", "B")]),
    2: ("t_nn", [(rel, s, s) for rel in ((0, -1), (1, -1), (1, 0)) for s in ("A", "B")]),
    3: ("t_nnn", [((1, -2), "A", "B"), ((1, 0), "A", "B"), ((-1, 0), "A", "B")]),
}


def monolayer(nearest_neighbors=1, onsite=(0, 0), **kwargs):
    """Monolayer graphene with hopping shells 1 .. `nearest_neighbors` (at most 3).

    With the second shell the onsite energies are raised by 3 t_nn, which moves the Dirac point
    back to zero energy (the reference's convention).  Energies: `t`, `t_nn`, `t_nnn` keywords."""
    if not 1 <= nearest_neighbors <= len(HOPPING_SHELLS):
        raise ValueError("nearest_neighbors must be 1, 2 or 3")
    energies = {"t": kwargs.get("t", t), "t_nn": kwargs.get("t_nn", t_nn), "t_nnn": kwargs.get("t_nnn", 0.05)}
    shift = 3 * energies["t_nn"] if nearest_neighbors >= 2 else 0
    lat = Lattice(a1=[a, 0], a2=[a / 2, a / 2 * sqrt(3)])
    lat.add_sublattices(("A", [0, -a_cc / 2], onsite[0] + shift), ("B", [0, a_cc / 2], onsite[1] + shift))
    lat.register_hopping_energies(energies)
    for shell in range(1, nearest_neighbors + 1):
        name, terms = HOPPING_SHELLS[shell]
        lat.add_hoppings(*[(rel, src, dst, name) for rel, src, dst in terms])
    lat.min_neighbors = 2
    return lat


def monolayer_4atom(onsite=(0, 0)):
    """Graphene with a rectangular unit cell of 4 atoms (a1 = (a, 0), a2 = (0, 3 a_cc)): A, B and
    their aliases A2, B2 shifted by (a/2, 3 a_cc/2); nearest-neighbour hoppings only."""
    lat = Lattice(a1=[a, 0], a2=[0, 3 * a_cc])
    lat.add_sublattices(("A", [0, -a_cc / 2], onsite[0]), ("B", [0, a_cc / 2], onsite[1]))
    lat.add_aliases(("A2", "A", [a / 2, a_cc]), ("B2", "B", [a / 2, 2 * a_cc]))
    inside = [((0, 0), "A", "B"), ((0, 0), "B", "A2"), ((0, 0), "A2", "B2")]
    across = [((-1, -1), "A", "B2"), ((0, -1), "A", "B2"), ((-1, 0), "B", "A2")]
    lat.add_hoppings(*[(rel, src, dst, t) for rel, src, dst in inside + across])
    lat.min_neighbors = 2
    return lat


def constant_magnetic_field(magnitude):
    """Peierls phase of a field `magnitude` [T] along z, vector potential along x; evaluated in
    float32/complex64 like the reference's numpy modifier."""
    const = 1e-18 * 2 * pi / phi0

    @hopping_energy_modifier(is_complex=True)
    def func(energy, x1, y1, x2, y2):
        vp_x = 0.5 * magnitude * (y1 + y2)
        peierls = vp_x * (x1 - x2)
        return energy * np.exp(1j * const * peierls)
    return func


@hopping_energy_modifier
def strained_hopping(energy, x1, y1, z1, x2, y2, z2):
    """t * exp(-beta (d / a_cc - 1))"""
    d = np.sqrt((x1 - x2) ** 2 + (y1 - y2) ** 2 + (z1 - z2) ** 2)
    return energy * np.exp(-beta * (d / a_cc - 1))


def gaussian_bump(height, sigma, center=(0, 0)):
    """Out-of-plane bump z = height * exp(-r^2 / sigma^2) (positions only; combine with
    `strained_hopping` for the strain)"""
    @site_position_modifier
    def displacement(x, y, z):
        r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2
        return x, y, z + height * np.exp(-r2 / sigma ** 2)
    return displacement
