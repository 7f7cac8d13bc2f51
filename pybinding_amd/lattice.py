"""Lattice, shapes, modifiers and the model builder: lattice + shape + modifiers -> Hamiltonian CSR
and site positions, in the reference's site order.

The order matters for parity: random starters are drawn per Hamiltonian row, so a DOS or a
conductivity computed with a few random vectors is only reproducible when every site sits at the
same row as in the reference.  The builder therefore follows, step by step:

* unit cell (cppcore/src/Lattice.cpp:19-120): sublattices get ids in insertion order, hopping
  families in registration order; onsite energies are made Hermitian from the upper triangle;
* optimized unit cell (cppcore/src/Lattice.cpp:288-318): stable sort by alias id, then by orbital
  count, each hopping term added to both ends (the second end with the opposite relative index,
  flagged as the conjugate);
* foundation (cppcore/src/system/Foundation.cpp:5-48): shape bounds in lattice coordinates,
  truncated to int and padded by one cell, positions generated sublattice-major in float32 with
  the reference's operation order;
* shape test (cppcore/src/system/Shape.cpp:49-86): polygon ray casting in float32;
* dangling-site removal (cppcore/src/system/Foundation.cpp:50-104);
* final indices (cppcore/src/system/Foundation.cpp:127-166): valid sites numbered sublattice by
  sublattice in foundation order;
* hoppings (cppcore/src/system/System.cpp:103-125) and the matrix
  (cppcore/include/hamiltonian/Hamiltonian.hpp:74-111, HamiltonianModifiers.hpp:106-290):
  H[row, col] = hopping, H[col, row] = its conjugate, zeros skipped, onsite terms only when the
  lattice has onsite energy or an onsite modifier is present;
* scalar type (cppcore/src/Model.cpp:91-98,180-204): float32 unless a modifier asks for double
  precision, complex when any energy or modifier is complex, or when the model is periodic;
* periodic models (cppcore/src/system/Foundation.cpp:107-114, Symmetry.cpp:6-98,
  System.cpp:128-158, cppcore/include/hamiltonian/Hamiltonian.hpp:118-131): a `primitive` block
  of unit cells (no shape test, no pruning), the cells outside the symmetry area dropped, and
  for every translation direction the hoppings that leave the area through one face re-entering
  through the opposite one, with the second site's position shifted back by the translation
  for the modifiers and the Bloch phase exp(i k.d) = 1 (k = 0).
"""
import inspect

import numpy as np
import scipy.sparse as sp

from .model import System

__all__ = ['Lattice', 'Polygon', 'Primitive', 'TranslationalSymmetry', 'rectangle', 'regular_polygon',
           'primitive', 'translational_symmetry', 'onsite_energy_modifier', 'hopping_energy_modifier',
           'site_position_modifier', 'constant_potential', 'force_double_precision', 'build']


def _cartesian(v):
    out = np.zeros(3, dtype=np.float32)
    v = np.atleast_1d(np.asarray(v, dtype=np.float64))
    out[:v.size] = v.astype(np.float32)
    return out


def _onsite_matrix(energy):
    """Scalar -> 1x1, vector -> diagonal, matrix -> Hermitian from its upper triangle
    (cppcore/src/Lattice.cpp:26-43)."""
    e = np.asarray(energy, dtype=np.complex128)
    if e.ndim == 0:
        return e.reshape(1, 1)
    if e.ndim == 1:
        return np.diag(e)
    if e.shape[0] != e.shape[1]:
        raise RuntimeError("The onsite hopping term must be a real vector or a square matrix")
    upper = np.triu(e)
    return upper + np.conj(np.triu(e, 1)).T


def _hopping_matrix(energy):
    e = np.asarray(energy, dtype=np.complex128)
    return e.reshape(1, 1) if e.ndim == 0 else e


class Lattice:
    """A Bravais lattice with sublattices and hopping terms (pybinding/lattice.py,
    cppcore/src/Lattice.cpp)."""

    def __init__(self, a1, a2=None, a3=None):
        self.vectors = [_cartesian(a1)]
        for v in (a2, a3):
            if v is not None and np.any(_cartesian(v) != 0):
                self.vectors.append(_cartesian(v))
        self.offset = np.zeros(3, dtype=np.float32)
        self.min_neighbors = 1
        self.sublattices = {}   # name -> dict(position, energy, unique_id, alias_id)
        self.hoppings = {}      # name -> dict(energy, family_id, terms[(rel, from_id, to_id)])

    @property
    def ndim(self):
        return len(self.vectors)

    @property
    def nsub(self):
        return len(self.sublattices)

    def vector(self, i):
        return self.vectors[i] if i < self.ndim else np.zeros(3, dtype=np.float32)

    def add_one_sublattice(self, name, position, onsite_energy=0.0, alias=None):
        if not name:
            raise RuntimeError("Sublattice name can't be blank")
        if name in self.sublattices:
            raise RuntimeError("Sublattice '{}' already exists".format(name))
        uid = len(self.sublattices)
        if alias is not None:
            original = self.sublattices[alias]
            energy, alias_id = original["energy"], original["unique_id"]
        else:
            energy, alias_id = _onsite_matrix(onsite_energy), uid
        self.sublattices[name] = dict(position=_cartesian(position), energy=energy, unique_id=uid,
                                      alias_id=alias_id)

    def add_sublattices(self, *sublattices):
        for sub in sublattices:
            self.add_one_sublattice(*sub)

    def add_one_alias(self, name, original, position):
        self.add_one_sublattice(name, position, alias=original)

    def add_aliases(self, *aliases):
        for a in aliases:
            self.add_one_alias(*a)

    def register_hopping_energies(self, mapping):
        for name, energy in mapping.items():
            if not name:
                raise RuntimeError("Hopping name can't be blank")
            if name in self.hoppings:
                raise RuntimeError("Hopping '{}' already exists".format(name))
            self.hoppings[name] = dict(energy=_hopping_matrix(energy), family_id=len(self.hoppings), terms=[])

    def add_one_hopping(self, relative_index, from_sub, to_sub, hop):
        rel = np.zeros(3, dtype=np.int64)
        ri = np.atleast_1d(np.asarray(relative_index, dtype=np.int64))
        rel[:ri.size] = ri
        if from_sub == to_sub and not rel.any():
            raise RuntimeError("Hoppings from/to the same sublattice must have a non-zero relative index "
                               "in at least one direction. Don't define onsite energy here.")
        if isinstance(hop, str):
            name = hop
        else:
            energy = _hopping_matrix(hop)
            name = next((k for k, f in self.hoppings.items()
                         if f["energy"].shape == energy.shape and np.array_equal(f["energy"], energy)), None)
            if name is None:
                name = "__anonymous__{}".format(len(self.hoppings))
                self.register_hopping_energies({name: energy})
        fam = self.hoppings[name]
        src, dst = self.sublattices[from_sub], self.sublattices[to_sub]
        if src["energy"].shape[0] != fam["energy"].shape[0] or dst["energy"].shape[1] != fam["energy"].shape[1]:
            raise RuntimeError("Hopping size mismatch: from '{}' to '{}' with matrix '{}'".format(from_sub, to_sub, name))
        term = (tuple(rel), src["unique_id"], dst["unique_id"])
        if any(term == t for f in self.hoppings.values() for t in f["terms"]):
            raise RuntimeError("The specified hopping already exists.")
        fam["terms"].append(term)

    def add_hoppings(self, *hoppings):
        for h in hoppings:
            self.add_one_hopping(*h)

    # -- geometry (cppcore/src/Lattice.cpp:190-216) --
    def calc_position(self, index):
        """offset + sum_i float(index[i]) * a_i, accumulated in float32 in the reference's order"""
        p = self.offset.copy()
        for i in range(self.ndim):
            p = p + np.float32(index[i]) * self.vectors[i]
        return p

    def translate_coordinates(self, position):
        """Solve [a_1 .. a_ndim] v = p for the lattice coordinates v of a Cartesian point"""
        nd = self.ndim
        m = np.stack([self.vectors[i][:nd].astype(np.float64) for i in range(nd)], axis=1)
        v = np.zeros(3)
        v[:nd] = np.linalg.solve(m, _cartesian(position)[:nd].astype(np.float64))
        return v

    def optimized_unit_cell(self):
        """Sublattices in unique-id order, stable-sorted by alias id then by orbital count; every
        hopping term listed at both ends (cppcore/src/Lattice.cpp:288-318)."""
        subs = sorted(self.sublattices.items(), key=lambda kv: kv[1]["unique_id"])
        sites = [dict(name=k, position=v["position"], norb=v["energy"].shape[0], unique_id=v["unique_id"],
                      alias_id=v["alias_id"], energy=v["energy"], hoppings=[]) for k, v in subs]
        sites.sort(key=lambda s: s["alias_id"])
        sites.sort(key=lambda s: s["norb"])
        where = {s["unique_id"]: i for i, s in enumerate(sites)}
        for fam in self.hoppings.values():
            for rel, src, dst in fam["terms"]:
                i1, i2 = where[src], where[dst]
                sites[i1]["hoppings"].append((np.array(rel), i2, fam["family_id"], False))
                sites[i2]["hoppings"].append((-np.array(rel), i1, fam["family_id"], True))
        return sites


# -- shapes (pybinding/shape.py, cppcore/src/system/Shape.cpp) --

class Polygon:
    """Shape defined by its vertices in the xy plane; `contains` is the reference's float32 ray
    casting test (cppcore/src/system/Shape.cpp:49-86)."""

    def __init__(self, vertices):
        if len(vertices) < 3:
            raise RuntimeError("A polygon must have at least 3 sides")
        self.vertices = [_cartesian(v) for v in vertices]

    def with_offset(self, vector):
        v = _cartesian(vector)
        return Polygon([p + v for p in self.vertices])

    def contains(self, x, y, z=None):
        x = np.asarray(x, dtype=np.float32)
        y = np.asarray(y, dtype=np.float32)
        vx = [p[0] for p in self.vertices]
        vy = [p[1] for p in self.vertices]
        inside = np.zeros(x.shape, dtype=bool)
        nv = len(vx)
        eps, tiny = np.finfo(np.float32).eps, np.finfo(np.float32).tiny
        for i in range(nv):
            j = (i - 1) % nv   # sides (i, j) with j the previous vertex
            x1, x2, y1, y2 = vx[i], vx[j], vy[i], vy[j]
            diff, scale = abs(y1 - y2), abs(y1 + y2)
            if diff <= eps * scale or diff <= tiny:   # num::approx_equal: side parallel to the ray
                continue
            k = (x2 - x1) / (y2 - y1)
            crosses_y = (y1 > y) != (y2 > y)
            crosses_x = x > k * (y - y1) + x1
            inside ^= crosses_y & crosses_x
        return inside


def rectangle(x, y=None):
    """A rectangle centred at the origin (pybinding/shape.py:263-279)"""
    y = y or x
    x0, y0 = x / 2, y / 2
    return Polygon([[x0, y0], [x0, -y0], [-x0, -y0], [-x0, y0]])


def regular_polygon(num_sides, radius, angle=0):
    """A polygon with equal sides (pybinding/shape.py:282-301)"""
    from math import pi, sin, cos
    angles = [angle + 2 * n * pi / num_sides for n in range(num_sides)]
    return Polygon([(radius * sin(a), radius * cos(a)) for a in angles])


class Primitive:
    """A block of a1 x a2 x a3 unit cells: every site kept, no dangling-site removal
    (pybinding/shape.py:233-246, cppcore/src/system/Shape.cpp:5-9)."""

    def __init__(self, a1=1, a2=1, a3=1):
        self.size = np.array([int(a1), int(a2), int(a3)], dtype=np.int64)
        if np.any(self.size <= 0):
            raise RuntimeError("Primitive: The size must be at least 1 in every direction.")


def primitive(a1=1, a2=1, a3=1):
    """Repeat the unit cell `a1`, `a2`, `a3` times along the lattice vectors"""
    return Primitive(a1, a2, a3)


class TranslationalSymmetry:
    """Periodicity along lattice vectors (pybinding/shape.py:322-344,
    cppcore/src/system/Symmetry.cpp:6-98): per direction a period length in nm, 0 for one unit
    cell, negative for none."""

    def __init__(self, a1=True, a2=True, a3=True):
        def length(v):
            if v is False:
                return np.float32(-1)
            if v is True:
                return np.float32(0)
            return np.float32(v)
        self.length = np.array([length(v) for v in (a1, a2, a3)], dtype=np.float32)
        self.enabled = self.length >= 0

    def area(self, lattice, size):
        """(left, right, middle) cell-index bounds of the periodic cell inside a foundation of
        `size` cells: `middle` cells per period, centred (Symmetry.cpp:13-40)."""
        left = np.zeros(3, dtype=np.int64)
        right = np.asarray(size, dtype=np.int64) - 1
        middle = np.zeros(3, dtype=np.int64)
        for i in range(lattice.ndim):
            if not self.enabled[i]:
                continue
            norm = np.float32(np.sqrt(np.sum(lattice.vector(i).astype(np.float32) ** 2, dtype=np.float32)))
            cells = int(np.round(np.float32(self.length[i] / norm)))
            cells = cells if cells > 0 else 1
            left[i] = int((size[i] - cells) / 2)   # C++ integer division (towards zero)
            right[i] = left[i] + cells - 1
            middle[i] = cells
        return left, right, middle

    def translations(self, lattice, uc):
        """Distinct translation directions: every unit-cell hopping's relative index restricted to
        each combination of the enabled directions (Symmetry.cpp:42-89, 106-127), in first-seen
        order, with the shift of each in cells and in nm (float32)."""
        dirs = [bool(self.enabled[i]) if i < lattice.ndim else False for i in range(3)]
        masks = [np.array([i, j, k]) for i in range(int(dirs[0]) + 1) for j in range(int(dirs[1]) + 1)
                 for k in range(int(dirs[2]) + 1)]
        out = []
        for site in uc:
            for rel, _, _, _ in site["hoppings"]:
                for mask in masks:
                    d = np.asarray(rel, dtype=np.int64) * mask
                    if not d.any() or any(np.array_equal(d, t) for t in out):
                        continue
                    out.append(d)
        return out


def translational_symmetry(a1=True, a2=True, a3=True):
    """Periodic boundaries along the lattice vectors.  Each argument: False (none), True (a period
    of one unit cell) or a period length in nm."""
    return TranslationalSymmetry(a1, a2, a3)


# -- modifiers (pybinding/modifier.py) --

class _SubId:
    """The `sub_id` argument of onsite modifiers: compares equal to sublattice names and gives
    the identity matrix of the sublattice's orbitals as `.eye`."""

    def __init__(self, name, norb):
        self.name = name
        self.eye = 1 if norb == 1 else np.eye(norb)

    def __eq__(self, other):
        return self.name == other

    def __ne__(self, other):
        return self.name != other


class _Modifier:
    def __init__(self, func, kind, is_double=False, is_complex=False):
        self.func = func
        self.kind = kind
        self.is_double = is_double
        self.is_complex = is_complex
        self.argnames = tuple(inspect.signature(func).parameters.keys())

    def __call__(self, **kwargs):
        return self.func(**{k: kwargs[k] for k in self.argnames})


def _decorator(kind, **flags):
    def wrap(func):
        return _Modifier(func, kind, **flags)
    return wrap


def onsite_energy_modifier(func=None, is_double=False):
    """Decorator for f(energy, x, y, z, sub_id) -> energy (pybinding/modifier.py:279-329)"""
    if func is not None:
        return _Modifier(func, "onsite")
    return _decorator("onsite", is_double=is_double)


def hopping_energy_modifier(func=None, is_double=False, is_complex=False):
    """Decorator for f(energy, x1, y1, z1, x2, y2, z2, hop_id) -> energy (pybinding/modifier.py:332-389)"""
    if func is not None:
        return _Modifier(func, "hopping")
    return _decorator("hopping", is_double=is_double, is_complex=is_complex)


def site_position_modifier(func=None):
    """Decorator for f(x, y, z, sub_id) -> (x, y, z), applied to the valid sites before the
    Hamiltonian is built (pybinding/modifier.py:236-276)"""
    if func is not None:
        return _Modifier(func, "position")
    return _decorator("position")


def constant_potential(magnitude):
    """Add `magnitude` to the onsite energy of every site (pybinding/modifier.py:391-402)"""
    @onsite_energy_modifier
    def f(energy, sub_id):
        return energy + sub_id.eye * magnitude
    return f


def force_double_precision():
    """Build the Hamiltonian in double precision (pybinding/modifier.py:405-410)"""
    @onsite_energy_modifier(is_double=True)
    def f(energy):
        return energy
    return f


# -- the builder --

def _foundation(lattice, shape, symmetry=None):
    """Positions (float32, foundation order), validity after the shape and dangling-site removal
    (or every site of a `Primitive` block) and after the symmetry area, and the spatial layout
    (cppcore/src/system/Foundation.cpp, cppcore/src/Model.cpp:128-131)."""
    uc = lattice.optimized_unit_cell()
    nsub = len(uc)
    nd = lattice.ndim
    if isinstance(shape, Primitive):
        if np.any(shape.size[nd:] != 1):
            raise RuntimeError("Primitive shape has more dimensions than the lattice")
        size = shape.size.copy()
        lo = -(size // 2)   # the block is centred: cells -size/2 .. (size-1)/2
    else:
        lo = np.full(3, np.iinfo(np.int32).max, dtype=np.int64)
        hi = np.full(3, np.iinfo(np.int32).min, dtype=np.int64)
        for p in shape.vertices:
            v = np.trunc(lattice.translate_coordinates(p)).astype(np.int64)   # cast<int>
            lo = np.minimum(lo, v)
            hi = np.maximum(hi, v)
        lo[:nd] -= 1
        hi[:nd] += 1
        size = hi - lo + 1
    s0, s1, s2 = (int(s) for s in size)
    origin = lattice.calc_position(lo)
    v0, v1, v2 = lattice.vector(0), lattice.vector(1), lattice.vector(2)
    fa = np.arange(s0, dtype=np.float32)[:, None] * v0[None, :]
    fb = np.arange(s1, dtype=np.float32)[:, None] * v1[None, :]
    fc = np.arange(s2, dtype=np.float32)[:, None] * v2[None, :]
    pos = np.empty((nsub, s2, s1, s0, 3), dtype=np.float32)
    for n, site in enumerate(uc):
        ps = origin + site["position"]
        pc = ps[None, :] + fc                                    # (s2, 3)
        pb = pc[:, None, :] + fb[None, :, :]                     # (s2, s1, 3)
        pos[n] = pb[:, :, None, :] + fa[None, None, :, :]
    pos = pos.reshape(-1, 3)
    block = s0 * s1 * s2
    grid = np.indices((s2, s1, s0)).reshape(3, -1)   # (c, b, a) for each spatial flat index
    if isinstance(shape, Primitive):
        valid = np.ones(nsub * block, dtype=bool)
    else:
        valid = shape.contains(pos[:, 0], pos[:, 1], pos[:, 2])

    # neighbours of every foundation site, one column per unit-cell hopping (both directions)
    nmax = max(len(s["hoppings"]) for s in uc) if nsub else 0
    nbr = np.full((nsub * block, max(nmax, 1)), -1, dtype=np.int64)
    for n, site in enumerate(uc):
        for h, (rel, to, _, _) in enumerate(site["hoppings"]):
            a = grid[2] + rel[0]
            b = grid[1] + rel[1]
            c = grid[0] + rel[2]
            inb = (a >= 0) & (a < s0) & (b >= 0) & (b < s1) & (c >= 0) & (c < s2)
            flat = ((to * s2 + c) * s1 + b) * s0 + a
            nbr[n * block:(n + 1) * block, h] = np.where(inb, flat, -1)
    inb = nbr >= 0
    count0 = inb.sum(axis=1)
    # remove_dangling: a valid site loses one count per invalid neighbour and is removed, with
    # the same effect on its own neighbours, when its count falls below min_neighbors.  Sites
    # whose count never changed are kept.  The end state does not depend on the visiting order.
    min_nb = lattice.min_neighbors
    safe = np.where(inb, nbr, 0)
    while not isinstance(shape, Primitive):
        bad = inb & ~valid[safe]
        nbad = bad.sum(axis=1)
        kill = valid & (nbad > 0) & (count0 - nbad < min_nb)
        if not kill.any():
            break
        valid &= ~kill
    if symmetry is not None:
        left, right, _ = symmetry.area(lattice, (s0, s1, s2))
        cell = np.stack([grid[2], grid[1], grid[0]])          # (a, b, c) of each spatial cell
        inside = np.all((cell >= left[:, None]) & (cell <= right[:, None]), axis=0)
        valid &= np.tile(inside, nsub)
    return uc, pos, valid, nbr, (s0, s1, s2)


def _boundary_hoppings(lattice, uc, symmetry, size, index):
    """Hoppings across the periodic boundaries (cppcore/src/system/System.cpp:128-158): for each
    translation direction d, the valid sites on the face the translation leaves from, moved by
    d x (cells per period), connect to every valid site their unit-cell hoppings (not the
    conjugate copies) then reach inside the foundation.  Returns [(shift in nm, {family id:
    (rows, cols)})], one entry per translation that has any hopping."""
    s0, s1, s2 = size
    block = s0 * s1 * s2
    grid = np.indices((s2, s1, s0)).reshape(3, -1)
    cell = np.stack([grid[2], grid[1], grid[0]])               # (a, b, c)
    dims = np.array([s0, s1, s2])[:, None]
    left, right, middle = symmetry.area(lattice, size)
    out = []
    for d in symmetry.translations(lattice, uc):
        face = np.ones(block, dtype=bool)
        for n in range(lattice.ndim):
            if d[n] > 0:
                face &= cell[n] == left[n]
            elif d[n] < 0:
                face &= cell[n] == right[n]
        cells = np.nonzero(face)[0]
        shift_index = d * middle
        shift = np.zeros(3, dtype=np.float32)
        for n in range(lattice.ndim):
            shift = shift + np.float32(d[n] * middle[n]) * lattice.vector(n)
        moved = cell[:, cells] + shift_index[:, None]
        per_family = {}
        for n_sub, site in enumerate(uc):
            src = index[n_sub * block + cells]
            for rel, to, fid, conj in site["hoppings"]:
                if conj:
                    continue
                nb = moved + np.asarray(rel)[:, None]
                inb = np.all((nb >= 0) & (nb < dims), axis=0)
                flat = ((to * s2 + nb[2]) * s1 + nb[1]) * s0 + nb[0]
                dst = np.where(inb, index[np.where(inb, flat, 0)], -1)
                ok = (src >= 0) & (dst >= 0)
                if ok.any():
                    per_family.setdefault(fid, []).append((src[ok], dst[ok]))
        if per_family:
            out.append((shift, {fid: (np.concatenate([p[0] for p in pairs]), np.concatenate([p[1] for p in pairs]))
                                for fid, pairs in sorted(per_family.items())}))
    return out


def _scalar_type(lattice, modifiers, periodic=False):
    is_double = any(m.is_double for m in modifiers)
    energies = [s["energy"] for s in lattice.sublattices.values()] + [f["energy"] for f in lattice.hoppings.values()]
    is_complex = (any(np.any(np.imag(e) != 0) for e in energies) or any(m.is_complex for m in modifiers)
                  or periodic)
    if is_complex:
        return np.complex128 if is_double else np.complex64
    return np.float64 if is_double else np.float32


def _cast(values, dtype):
    v = np.asarray(values)
    if not np.issubdtype(dtype, np.complexfloating):
        v = np.real(v)
    return v.astype(dtype)


def build(lattice, shape, *modifiers):
    """Build a model: returns (hamiltonian CSR, System) in the reference's site order.  `shape` is
    a polygon or a `primitive` block; `modifiers` may include one `translational_symmetry`."""
    modifiers = [m for m in modifiers if m is not None]
    symmetry = None
    for m in modifiers:
        if isinstance(m, TranslationalSymmetry):
            symmetry = m
        elif not isinstance(m, _Modifier):
            raise TypeError("unsupported model argument: {!r}".format(m))
    modifiers = [m for m in modifiers if isinstance(m, _Modifier)]
    periodic = symmetry is not None and bool(symmetry.enabled[:lattice.ndim].any())
    dtype = _scalar_type(lattice, modifiers, periodic)
    uc, pos, valid, nbr, (s0, s1, s2) = _foundation(lattice, shape, symmetry if periodic else None)
    block = s0 * s1 * s2
    nsub = len(uc)

    # final indices: valid sites, sublattice by sublattice in foundation order
    index = np.full(valid.size, -1, dtype=np.int64)
    index[valid] = np.arange(int(valid.sum()))
    num_sites = int(valid.sum())
    sub_of_flat = np.repeat(np.arange(nsub), block)
    site_sub = sub_of_flat[valid]
    x, y, z = (pos[valid, k].copy() for k in range(3))
    names = [s["name"] for s in uc]
    norb_of_sub = np.array([s["norb"] for s in uc], dtype=np.int64)
    norb = norb_of_sub[site_sub]
    ham_start = np.concatenate([[0], np.cumsum(norb)])[:-1]
    ham_size = int(norb.sum())

    for m in modifiers:
        if m.kind == "position":
            sub_name = np.array(names, dtype=object)[site_sub]
            x, y, z = (np.asarray(a, dtype=np.float32) for a in m(x=x, y=y, z=z, sub_id=sub_name))

    rows, cols, vals = [], [], []

    # onsite terms (HamiltonianModifiers.hpp:106-154)
    onsite_mods = [m for m in modifiers if m.kind == "onsite"]
    has_intrinsic = any(np.any(s["energy"] != 0) for s in lattice.sublattices.values())
    if has_intrinsic or onsite_mods:
        for n, s in enumerate(uc):
            sel = np.nonzero(site_sub == n)[0]
            if sel.size == 0:
                continue
            k = s["norb"]
            e0 = _cast(s["energy"], dtype)
            if k == 1:
                energy = np.full(sel.size, e0[0, 0], dtype=dtype)
            else:
                energy = np.broadcast_to(e0, (sel.size, k, k)).copy()
            for m in onsite_mods:
                energy = _cast(m(energy=energy, x=x[sel], y=y[sel], z=z[sel], sub_id=_SubId(s["name"], k)), dtype)
            energy = energy.reshape(sel.size, k, k)
            for p in range(k):
                for q in range(k):
                    v = energy[:, p, q]
                    nz = v != 0
                    rows.append(ham_start[sel[nz]] + p)
                    cols.append(ham_start[sel[nz]] + q)
                    vals.append(v[nz])

    # hoppings (System.cpp:103-125; HamiltonianModifiers.hpp:240-290)
    hop_mods = [m for m in modifiers if m.kind == "hopping"]
    fam_rows = {}
    for n, s in enumerate(uc):
        sites = np.arange(n * block, (n + 1) * block)
        for h, (rel, to, fid, conj) in enumerate(s["hoppings"]):
            if conj:
                continue
            nb = nbr[sites, h]
            ok = (nb >= 0) & valid[sites] & valid[np.where(nb >= 0, nb, 0)]
            fam_rows.setdefault(fid, []).append((index[sites[ok]], index[nb[ok]]))
    fams = {f["family_id"]: (name, f) for name, f in lattice.hoppings.items()}
    no_shift = np.zeros(3, dtype=np.float32)
    main = {fid: (np.concatenate([p[0] for p in fam_rows[fid]]), np.concatenate([p[1] for p in fam_rows[fid]]))
            for fid in sorted(fam_rows)}
    boundaries = _boundary_hoppings(lattice, uc, symmetry, (s0, s1, s2), index) if periodic else []
    groups = [(fid, r, c, no_shift) for fid, (r, c) in main.items()]
    for shift, fams_b in boundaries:
        groups += [(fid, r, c, shift) for fid, (r, c) in fams_b.items()]
    for fid, r, c, shift in groups:
        name, fam = fams[fid]
        if r.size == 0:
            continue
        e0 = _cast(fam["energy"], dtype)
        kr, kc = e0.shape
        if kr == 1 and kc == 1:
            energy = np.full(r.size, e0[0, 0], dtype=dtype)
        else:
            energy = np.broadcast_to(e0, (r.size, kr, kc)).copy()
        x2, y2, z2 = x[c] - shift[0], y[c] - shift[1], z[c] - shift[2]
        for m in hop_mods:
            energy = _cast(m(energy=energy, x1=x[r], y1=y[r], z1=z[r], x2=x2, y2=y2, z2=z2, hop_id=name), dtype)
        energy = energy.reshape(r.size, kr, kc)
        for p in range(kr):
            for q in range(kc):
                v = energy[:, p, q]
                nz = v != 0
                hr, hc = ham_start[r[nz]] + p, ham_start[c[nz]] + q
                rows += [hr, hc]
                cols += [hc, hr]
                vals += [v[nz], np.conj(v[nz])]

    if rows:
        rows, cols, vals = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals).astype(dtype)
    else:
        rows = cols = np.zeros(0, dtype=np.int64)
        vals = np.zeros(0, dtype=dtype)
    h = sp.csr_matrix((vals, (rows, cols)), shape=(ham_size, ham_size), dtype=dtype)
    h.sort_indices()
    if not np.all(np.isfinite(h.data)):
        raise RuntimeError("The Hamiltonian contains invalid values: NaN or INF.\n"
                           "Check the lattice and/or modifier functions.")
    # sites report the id of the sublattice they alias (their own unique id otherwise); names are
    # listed by unique id (CompressedSublattices, cppcore/src/system/CompressedSublattices.cpp)
    names_by_uid = [None] * len(uc)
    for site in uc:
        names_by_uid[site["unique_id"]] = site["name"]
    alias_of_sub = np.array([site["alias_id"] for site in uc], dtype=np.int32)
    system = System(x, y, z, sublattices=alias_of_sub[site_sub], sublattice_names=names_by_uid, norb=norb)
    fam_name = {fid: name for fid, (name, _) in fams.items()}
    system.hopping_blocks = {fam_name[fid]: rc for fid, rc in main.items()}
    system.boundaries = [(shift, {fam_name[fid]: rc for fid, rc in fb.items()}) for shift, fb in boundaries]
    return h, system
