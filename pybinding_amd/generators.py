"""Direct, vectorised generators of the tight-binding Hamiltonians used by the benchmarks.

They produce the same matrices the reference's model builder makes for these simple cases, in
the reference's site order (sublattice-major, then a2, then a1 -- cppcore/include/system/
Foundation.hpp:167), without going through the general `Model` machinery, so that multi-million
site systems are built in seconds:

- `honeycomb_supercell`: `pb.Model(graphene.monolayer(), pb.primitive(n1, n2),
  pb.translational_symmetry())` plus an optional Anderson onsite term;
- `square_flake`: a one-orbital square lattice L1 x L2 with open edges plus Anderson disorder;
- `honeycomb_flake`: graphene cells within a rectangle, with dangling sites (fewer than 2
  neighbours) removed like `Lattice.min_neighbors = 2` does.

Graphene constants follow pybinding/repository/graphene/constants.py:3-8 and the hopping table of
pybinding/repository/graphene/lattice.py:36-40.
"""
import numpy as np
import scipy.sparse as sp

from .ordering import hilbert_index

A_CC = 0.142        # nm, carbon-carbon distance
A_LAT = 0.24595     # nm, graphene lattice constant
T_GRAPHENE = -2.8   # eV

# graphene.monolayer() hoppings from A(i, j) to B(i + di, j + dj)
_HONEYCOMB_HOPS = ((0, 0), (1, -1), (0, -1))


def honeycomb_cell_order(n1, n2):
    """A locality order of the honeycomb supercell's sites: cells along a Hilbert curve, the two
    sublattice sites of a cell adjacent.  Returns q with q[original index] = new index."""
    i, j = np.meshgrid(np.arange(n1), np.arange(n2), indexing="xy")
    order = int(np.ceil(np.log2(max(n1, n2, 2))))
    key = hilbert_index(i.ravel(), j.ravel(), order)
    cell_rank = np.empty(n1 * n2, dtype=np.int64)
    cell_rank[np.argsort(key, kind="stable")] = np.arange(n1 * n2)
    ncell = n1 * n2
    q = np.empty(2 * ncell, dtype=np.int64)
    q[:ncell] = 2 * cell_rank
    q[ncell:] = 2 * cell_rank + 1
    return q


def _anderson(n, W, seed):
    if not W:
        return None
    rng = np.random.default_rng(seed)
    return rng.uniform(-W / 2, W / 2, size=n)


def _assemble(n, rows, cols, vals, onsite, dtype):
    r = np.concatenate([rows, cols])
    c = np.concatenate([cols, rows])
    v = np.concatenate([vals, np.conj(vals)])
    if onsite is not None:
        d = np.arange(n)
        r = np.concatenate([r, d])
        c = np.concatenate([c, d])
        v = np.concatenate([v, onsite])
    h = sp.csr_matrix((v.astype(dtype), (r, c)), shape=(n, n))
    h.sum_duplicates()
    h.sort_indices()
    return h


def honeycomb_supercell(n1, n2, t=T_GRAPHENE, disorder=0.0, seed=2, dtype=np.float64):
    """Periodic graphene supercell of n1 x n2 cells (2 n1 n2 sites), optional Anderson onsite
    energies uniform in [-W/2, W/2] (numpy PCG64 `seed`)."""
    i, j = np.meshgrid(np.arange(n1), np.arange(n2), indexing="xy")
    i = i.ravel()
    j = j.ravel()
    ncell = n1 * n2
    a_idx = j * n1 + i
    rows, cols = [], []
    for di, dj in _HONEYCOMB_HOPS:
        bi = (i + di) % n1
        bj = (j + dj) % n2
        rows.append(a_idx)
        cols.append(ncell + bj * n1 + bi)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = np.full(rows.size, t)
    return _assemble(2 * ncell, rows, cols, vals, _anderson(2 * ncell, disorder, seed), dtype)


def square_flake(l1, l2, t=-1.0, disorder=0.0, seed=None, dtype=np.float64, rng=None):
    """Square lattice l1 x l2 (one orbital, nearest neighbours `t`) with open edges, optional
    Anderson onsite energies uniform in [-W/2, W/2]."""
    n = l1 * l2
    idx = np.arange(n).reshape(l2, l1)
    rows = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    cols = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    vals = np.full(rows.size, t)
    onsite = None
    if disorder:
        g = rng if rng is not None else np.random.default_rng(seed)
        onsite = g.uniform(-disorder / 2, disorder / 2, size=n)
    return _assemble(n, rows, cols, vals, onsite, dtype)


def square_positions(l1, l2):
    """Positions (x, y) of the sites of `square_flake(l1, l2)` (lattice constant 1)."""
    i = np.arange(l1 * l2)
    return (i % l1).astype(np.float64), (i // l1).astype(np.float64)


def honeycomb_flake(width, height, t=T_GRAPHENE, dtype=np.float64, return_positions=False):
    """Graphene flake: sites of the monolayer lattice (a1 = [a, 0], a2 = [a/2, a sqrt(3)/2],
    A at [0, -a_cc/2], B at [0, a_cc/2]) inside the centred `width` x `height` rectangle, dangling
    sites with fewer than 2 neighbours removed repeatedly."""
    a1 = np.array([A_LAT, 0.0])
    a2 = np.array([A_LAT / 2, A_LAT / 2 * np.sqrt(3)])
    sub = np.array([[0.0, -A_CC / 2], [0.0, A_CC / 2]])
    # cell index range covering the rectangle
    inv = np.linalg.inv(np.stack([a1, a2], axis=1))
    corners = np.array([[sx * width / 2, sy * height / 2] for sx in (-1, 1) for sy in (-1, 1)])
    lc = corners @ inv.T
    lo = np.floor(lc.min(axis=0)).astype(int) - 1
    hi = np.ceil(lc.max(axis=0)).astype(int) + 1
    n1, n2 = hi - lo + 1
    i, j = np.meshgrid(np.arange(n1), np.arange(n2), indexing="xy")
    i = i.ravel()
    j = j.ravel()
    base = (lo[0] + i)[:, None] * a1 + (lo[1] + j)[:, None] * a2
    pos = np.concatenate([base + sub[0], base + sub[1]])
    ncell = n1 * n2
    inside = (np.abs(pos[:, 0]) < width / 2) & (np.abs(pos[:, 1]) < height / 2)
    a_idx = j * n1 + i
    edges = []
    for di, dj in _HONEYCOMB_HOPS:
        bi, bj = i + di, j + dj
        ok = (bi >= 0) & (bi < n1) & (bj >= 0) & (bj < n2)
        edges.append(np.stack([a_idx[ok], ncell + bj[ok] * n1 + bi[ok]], axis=1))
    edges = np.concatenate(edges)
    valid = inside.copy()
    while True:
        e_ok = valid[edges[:, 0]] & valid[edges[:, 1]]
        deg = np.bincount(edges[e_ok, 0], minlength=2 * ncell) + np.bincount(edges[e_ok, 1], minlength=2 * ncell)
        drop = valid & (deg < 2)
        if not drop.any():
            break
        valid &= ~drop
    new = -np.ones(2 * ncell, dtype=np.int64)
    new[valid] = np.arange(valid.sum())
    e_ok = valid[edges[:, 0]] & valid[edges[:, 1]]
    rows = new[edges[e_ok, 0]]
    cols = new[edges[e_ok, 1]]
    h = _assemble(int(valid.sum()), rows, cols, np.full(rows.size, t), None, dtype)
    if return_positions:
        return h, pos[valid]
    return h


def honeycomb_positions(n1, n2):
    """Site positions (x, y) of `honeycomb_supercell(n1, n2)` in its site order (float64)."""
    i, j = np.meshgrid(np.arange(n1), np.arange(n2), indexing="xy")
    i = i.ravel()
    j = j.ravel()
    x = i * A_LAT + j * A_LAT / 2
    y = j * A_LAT / 2 * np.sqrt(3)
    return np.concatenate([x, x]), np.concatenate([y - A_CC / 2, y + A_CC / 2])


def peierls_honeycomb_supercell(n1, n2, field, disorder=0.0, seed=5, t=T_GRAPHENE):
    """Periodic graphene supercell with a perpendicular magnetic field `field` (tesla) through
    Peierls phases exp(i c (y1 + y2)/2 B (x1 - x2)), c = 1e-18 * 2 pi / phi0 with the reference's
    float constants (pybinding/repository/graphene/modifiers.py:54-73), plus Anderson disorder.
    Returns (H complex128 CSR, x, y)."""
    x, y = honeycomb_positions(n1, n2)
    i, j = np.meshgrid(np.arange(n1), np.arange(n2), indexing="xy")
    i = i.ravel()
    j = j.ravel()
    ncell = n1 * n2
    a_idx = j * n1 + i
    rows, cols = [], []
    for di, dj in _HONEYCOMB_HOPS:
        rows.append(a_idx)
        cols.append(ncell + ((j + dj) % n2) * n1 + (i + di) % n1)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    hbar = np.float32(6.58211899e-16)
    phi0 = np.float32(2) * np.float32(np.pi) * hbar
    const = 1e-18 * 2 * np.pi / float(phi0)
    vp_x = 0.5 * field * (y[rows] + y[cols])
    vals = t * np.exp(1j * const * vp_x * (x[rows] - x[cols]))
    return _assemble(2 * ncell, rows, cols, vals, _anderson(2 * ncell, disorder, seed), np.complex128), x, y


def strained_graphene_flake(width, height, bump_height=1.0, bump_sigma=5.0, return_positions=False):
    """Graphene flake with a Gaussian bump z = h exp(-r^2 / sigma^2) and strained hoppings
    t = -2.8 exp(-3.37 (d / a_cc - 1)) (pybinding/repository/graphene/modifiers.py:76-82,
    118-137).  Returns (num_sites, bonds (E, 2), hoppings (E,), onsite zeros (N,)) and, with
    `return_positions`, the (N, 2) in-plane positions."""
    h, pos = honeycomb_flake(width, height, return_positions=True)
    coo = sp.triu(h, k=1).tocoo()
    bonds = np.stack([coo.row, coo.col], axis=1)
    z = bump_height * np.exp(-(pos[:, 0] ** 2 + pos[:, 1] ** 2) / bump_sigma ** 2)
    d = np.sqrt((pos[bonds[:, 0], 0] - pos[bonds[:, 1], 0]) ** 2 + (pos[bonds[:, 0], 1] - pos[bonds[:, 1], 1]) ** 2
                + (z[bonds[:, 0]] - z[bonds[:, 1]]) ** 2)
    hop = T_GRAPHENE * np.exp(-3.37 * (d / A_CC - 1))
    if return_positions:
        return h.shape[0], bonds, hop, np.zeros(h.shape[0]), pos[:, :2]
    return h.shape[0], bonds, hop, np.zeros(h.shape[0])
