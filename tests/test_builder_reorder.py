"""Integer known answers of the reference's reorder tests (cppcore/tests/test_kpm.cpp:19-184) on
the same model: graphene 0.6 x 0.8 nm + constant_potential(1), built by the model builder.

They pin three things at once: the builder's site order (find_nearest must return the same
sites), the BFS reorder with its slice borders (oracle restatement and the engine's C++ host
core), and SliceMap::index.
"""
import numpy as np
import pytest

from pybinding_amd import lattice as L
from pybinding_amd.model import Model


def fixture_monolayer():
    """The C++ tests' graphene lattice (cppcore/tests/fixtures.cpp:110-122): float32 constants,
    nearest neighbours only and no dangling-site pruning (min_neighbors stays 1)."""
    f = np.float32
    a, a_cc = f(0.24595), f(0.142)
    lat = L.Lattice(a1=[a, 0], a2=[a / f(2), a / f(2) * np.sqrt(f(3))])
    lat.add_sublattices(("A", [0, -a_cc / f(2)]), ("B", [0, a_cc / f(2)]))
    lat.register_hopping_energies({"t": f(-2.8)})
    lat.add_hoppings(([0, 0], "A", "B", "t"), ([1, -1], "A", "B", "t"), ([0, -1], "A", "B", "t"))
    return lat


@pytest.fixture(scope="module")
def model():
    return Model(fixture_monolayer(), L.rectangle(0.6, 0.8), L.constant_potential(1))


def nearest(model, x, y, sub):
    return model.system.find_nearest([x, y, 0], sub)


# (src sites, dest sites, optimized src, optimized dest, number of borders, offsets, {M: indices})
CASES = {
    "diagonal single": (
        [(0, 0.07, "B")], None, [0], [0], 5, (0, 0),
        {6: [0, 1, 2, 2, 1, 0], 9: [0, 1, 2, 3, 4, 3, 2, 1, 0], 12: [0, 1, 2, 3, 4, 4, 4, 4, 3, 2, 1, 0]}),
    "diagonal multi 1": (
        [(0, -0.07, "A"), (0, 0.07, "B")], None, [0, 3], [0, 3], 5, (1, 1),
        {6: [1, 2, 3, 3, 2, 1], 9: [1, 2, 3, 4, 4, 4, 3, 2, 1], 12: [1, 2, 3, 4, 4, 4, 4, 4, 4, 3, 2, 1]}),
    "diagonal multi 2": (
        [(0, 0.07, "B"), (0, -0.07, "A"), (0, 0.35, "A")], None, [0, 1, 15], [0, 1, 15], 5, (3, 3),
        {6: [3, 4, 4, 4, 4, 3], 9: [3, 4, 4, 4, 4, 4, 4, 4, 3], 12: [3] + [4] * 10 + [3]}),
    "off-diagonal single": (
        [(0, 0.35, "A")], [(0, 0.07, "B")], [0], [8], 8, (0, 3),
        {6: [0, 1, 2, 3, 4, 3], 9: [0, 1, 2, 3, 4, 5, 5, 4, 3], 12: [0, 1, 2, 3, 4, 5, 6, 7, 6, 5, 4, 3],
         14: [0, 1, 2, 3, 4, 5, 6, 7, 7, 7, 6, 5, 4, 3]}),
    "off-diagonal multi 1": (
        [(0, 0.35, "A")], [(0, 0.07, "B"), (0.12, 0.14, "A"), (0.12, 0.28, "B")], [0], [8, 5, 2], 8, (0, 3),
        {6: [0, 1, 2, 3, 4, 3], 9: [0, 1, 2, 3, 4, 5, 5, 4, 3], 12: [0, 1, 2, 3, 4, 5, 6, 7, 6, 5, 4, 3]}),
    "off-diagonal multi 2": (
        [(0, 0.35, "A"), (0, -0.35, "B")], [(0.12, 0.28, "B"), (-0.12, 0.28, "B")], [0, 18], [2, 1], 8, (7, 1),
        {6: [6, 5, 4, 3, 2, 1], 9: [7, 7, 7, 6, 5, 4, 3, 2, 1], 12: [7] * 6 + [6, 5, 4, 3, 2, 1]}),
}


@pytest.mark.parametrize("case", list(CASES))
def test_oracle_bfs_reorder_known_answers(oracle, model, case):
    srcs, dests, osrc, odest, nborders, offsets, indices = CASES[case]
    src = [nearest(model, *s) for s in srcs]
    dest = src if dests is None else [nearest(model, *d) for d in dests]
    h = model.hamiltonian
    reorder, borders = oracle.bfs_reorder(h, src)
    assert [int(reorder[i]) for i in src] == osrc
    assert [int(reorder[i]) for i in dest] == odest
    assert len(borders) == nborders
    assert borders[0] == 1 and borders[-1] == h.shape[0]
    so = oracle.slice_offset(borders, [reorder[i] for i in src])
    do = oracle.slice_offset(borders, [reorder[i] for i in dest])
    assert (so, do) == offsets
    for M, expect in indices.items():
        assert [oracle.slice_map_index(borders, so, do, n, M) for n in range(M)] == expect


@pytest.mark.parametrize("case", list(CASES))
def test_host_core_reorder_known_answers(oracle, model, case):
    """The engine's C++ OptimizedHamiltonian (BFS, slice borders, optimized indices, offsets)"""
    from pybinding_amd import _native
    native = _native.load_module()
    srcs, dests, osrc, odest, nborders, offsets, indices = CASES[case]
    src = [nearest(model, *s) for s in srcs]
    dest = src if dests is None else [nearest(model, *d) for d in dests]
    h = model.hamiltonian
    out = native.optimize(h.shape, h.indptr, h.indices, h.data, (-10.0, 10.0), True, src, dest)
    _, _, _, reorder_map, borders, (s, d), (so, do), _ = out
    assert list(s) == osrc and list(d) == odest
    assert len(borders) == nborders and borders[0] == 1 and borders[-1] == h.shape[0]
    assert (so, do) == offsets
    for M, expect in indices.items():
        assert [oracle.slice_map_index(list(borders), so, do, n, M) for n in range(M)] == expect


def test_scaling_adds_diagonal(oracle):
    """b != 0 gives H2 a full diagonal even when H has none (test_kpm.cpp:174-184)"""
    m = Model(fixture_monolayer(), L.rectangle(0.6, 0.8))
    h = m.hamiltonian
    assert h.diagonal().any() == False  # noqa: E712
    a, b = oracle.scale_factors(-12, 10)
    assert b != 0
    h2 = oracle.scale_hamiltonian(h, a, b, np.float32)
    assert h2.nnz == h.nnz + h.shape[0]
    from pybinding_amd import _native
    out = _native.load_module().optimize(h.shape, h.indptr, h.indices, h.data, (-12.0, 10.0), True, [0], [0])
    assert len(out[1]) == h.nnz + h.shape[0]
