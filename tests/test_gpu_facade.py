"""The rest of the `kpm()` facade surface against the oracle: `calc_spatial_ldos`
(cppcore/src/KPM.cpp:67-94, pybinding/chebyshev.py:203-227), `deferred_ldos`
(chebyshev.py:253-275) and the statistics behind `report` (cppcore/src/kpm/Stats.cpp:31-65,
OptimizedHamiltonian.cpp:177-207)."""
import numpy as np
import pytest

import pybinding_amd as pb
from pybinding_amd.repository import graphene

pytestmark = pytest.mark.gpu

ENERGY = np.linspace(-1.5, 1.5, 31)


@pytest.fixture(scope="module")
def flake():
    return pb.Model(graphene.monolayer(), pb.rectangle(3.0), pb.force_double_precision())


def ldos_reference(oracle, h, site, a, b, M):
    h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
    e = np.zeros((1, h.shape[0]))
    e[0, site] = 1
    mu = oracle.diag_moments(h2, e, M)[0] * oracle.jackson(M)
    return oracle.spectral_density(mu, ENERGY, a, b)


def test_spatial_ldos(flake, oracle):
    kpm = pb.kpm(flake, energy_range=(-9, 9), silent=True)
    a, b = kpm.scaling_factors
    shape = pb.rectangle(0.9, 0.6)
    sldos = kpm.calc_spatial_ldos(ENERGY, broadening=0.2, shape=shape)
    inside = np.nonzero(shape.contains(flake.system.x, flake.system.y))[0]
    assert sldos.data.shape == (ENERGY.size, inside.size) and inside.size >= 4
    np.testing.assert_array_equal(sldos.structure.x, flake.system.x[inside])
    np.testing.assert_array_equal(sldos.structure.y, flake.system.y[inside])
    M = oracle.required_num_moments("jackson", 0.2 / a)
    for k, site in enumerate(inside):
        ref = ldos_reference(oracle, flake.hamiltonian, site, a, b, M)
        np.testing.assert_allclose(sldos.data[:, k], ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    # one sublattice only, and the LDOS of one site picked back out of the map
    sub = kpm.calc_spatial_ldos(ENERGY, broadening=0.2, shape=shape, sublattice="B")
    keep = inside[flake.system.sublattices_str[inside] == "B"]
    assert sub.data.shape == (ENERGY.size, keep.size) and 0 < keep.size < inside.size
    site = keep[0]
    pos = (flake.system.x[site], flake.system.y[site])
    np.testing.assert_allclose(sub.ldos(pos, "B").data, sub.data[:, 0], rtol=0, atol=0)
    np.testing.assert_allclose(sub.structure_map(0.5).data, sub.data[np.argmin(abs(ENERGY - 0.5))], rtol=0, atol=0)


def test_deferred_ldos(flake, oracle):
    kpm = pb.kpm(flake, energy_range=(-9, 9), silent=True)
    deferred = kpm.deferred_ldos(ENERGY, broadening=0.2, position=[0.1, 0.2])
    direct = kpm.calc_ldos(ENERGY, broadening=0.2, position=[0.1, 0.2]).data
    np.testing.assert_array_equal(deferred.result, direct)
    site = flake.system.find_nearest([0.1, 0.2])
    a, b = kpm.scaling_factors
    ref = ldos_reference(oracle, flake.hamiltonian, site, a, b, oracle.required_num_moments("jackson", 0.2 / a))
    np.testing.assert_allclose(direct, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("fmt", ["ELL", "CSR"])
def test_dos_stats_and_report(flake, oracle, fmt):
    """DOS statistics: the reference counts M x (ELL width x N, or nnz) matrix entries, halved for
    the diagonal algorithm, times the number of random vectors, over the moment time."""
    kpm = pb.kpm(flake, energy_range=(-9, 9), silent=True, matrix_format=fmt)
    kpm.calc_dos(ENERGY, broadening=0.2, num_random=3)
    s = kpm.stats
    a, _ = kpm.scaling_factors
    M = oracle.required_num_moments("jackson", 0.2 / a)
    h = flake.hamiltonian
    n = h.shape[0]
    per_step = n * int(np.diff(h.indptr).max()) if fmt == "ELL" else h.nnz
    assert s.num_moments == M and s.multiplier == 3
    assert s.nnz == M * per_step // 2 and s.opt_nnz == s.nnz
    assert s.vec == M * n // 2 and s.opt_vec == s.vec
    assert s.moments_time > 0
    assert s.eps == pytest.approx(3 * s.opt_nnz / s.moments_time, rel=1e-12)
    long_form = kpm.report()
    short = kpm.report(shortform=True)
    assert str(M) in long_form and str(M) in short
    assert "{:.2f}".format(-9.0) in short


def test_ldos_stats_light_cone(flake, oracle):
    """LDOS statistics with optimal_size: the processed entries follow the slice map of the BFS
    order from the target site (the light cone), step by step."""
    kpm = pb.kpm(flake, energy_range=(-9, 9), silent=True, matrix_format="CSR", optimal_size=True)
    a, _ = kpm.scaling_factors
    kpm.calc_ldos(ENERGY, broadening=0.6, position=[0, 0])
    s = kpm.stats
    M = oracle.required_num_moments("jackson", 0.6 / a)
    h = flake.hamiltonian
    site = flake.system.find_nearest([0, 0])
    rmap, borders = oracle.bfs_reorder(h, [site])
    inv = np.argsort(rmap)
    row_nnz = np.diff(h.indptr)[inv]
    prefix = np.concatenate([[0], np.cumsum(row_nnz)])
    opt = [borders[oracle.slice_map_index(borders, 0, 0, k, M)] for k in range(M)]
    assert s.num_moments == M
    assert s.opt_nnz == sum(int(prefix[o]) for o in opt) // 2
    assert s.opt_vec == sum(opt) // 2
    assert s.nnz == M * h.nnz // 2


@pytest.mark.parametrize("fmt", ["CSR", "ELL"])
def test_light_cone_matches_full_sweeps(fmt):
    """optimal_size: each step computes only the rows of the BFS slices the recurrence has reached
    (LDOS, diagonal Green's function) or that can still reach the collected rows (off-diagonal
    Green's functions); the results equal the full sweeps'."""
    model = pb.Model(graphene.monolayer(), pb.rectangle(14.0), pb.force_double_precision())
    energy = np.linspace(-2, 2, 41)
    res = {}
    for opt in (True, False):
        kpm = pb.kpm(model, energy_range=(-9, 9), silent=True, matrix_format=fmt, optimal_size=opt,
                     interleaved=False)
        ldos = kpm.calc_ldos(energy, broadening=0.1, position=[0.3, -0.2]).data
        spatial = kpm.calc_spatial_ldos(energy, broadening=0.1, shape=pb.rectangle(0.8)).data
        i = model.system.find_nearest([0, 0])
        greens = kpm.calc_greens(i, [i, i + 1, i + 7], energy, broadening=0.1)
        res[opt] = (ldos, spatial, greens)
    for a, b in zip(res[True], res[False]):
        a, b = np.asarray(a), np.asarray(b)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
