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


def monolayer(nearest_neighbors=1, onsite=(0, 0), **kwargs):
    """Monolayer graphene with up to third-nearest-neighbour hoppings"""
    lat = Lattice(a1=[a, 0], a2=[a / 2, a / 2 * sqrt(3)])
    onsite_offset = 0 if nearest_neighbors < 2 else 3 * kwargs.get('t_nn', t_nn)
    lat.add_sublattices(('A', [0, -a_cc / 2], onsite[0] + onsite_offset),
                        ('B', [0, a_cc / 2], onsite[1] + onsite_offset))
    lat.register_hopping_energies({'t': kwargs.get('t', t), 't_nn': kwargs.get('t_nn', t_nn),
                                   't_nnn': kwargs.get('t_nnn', 0.05)})
    lat.add_hoppings(([0, 0], 'A', 'B', 't'), ([1, -1], 'A', 'B', 't'), ([0, -1], 'A', 'B', 't'))
    if nearest_neighbors >= 2:
        lat.add_hoppings(([0, -1], 'A', 'A', 't_nn'), ([0, -1], 'B', 'B', 't_nn'),
                         ([1, -1], 'A', 'A', 't_nn'), ([1, -1], 'B', 'B', 't_nn'),
                         ([1, 0], 'A', 'A', 't_nn'), ([1, 0], 'B', 'B', 't_nn'))
    if nearest_neighbors >= 3:
        lat.add_hoppings(((1, -2), 'A', 'B', 't_nnn'), ((1, 0), 'A', 'B', 't_nnn'),
                         ((-1, 0), 'A', 'B', 't_nnn'))
    if nearest_neighbors >= 4:
        raise RuntimeError("No more")
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
