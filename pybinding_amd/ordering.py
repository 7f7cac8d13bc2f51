"""Site orders that keep the GPU's gathered vector rows in cache.

The moments of the whole-system calculations (DOS, moments, conductivity) do not depend on the
order of the sites, so the engine may permute H2 and the starters freely.  The reference permutes
by BFS from site 0 (cppcore/src/kpm/OptimizedHamiltonian.cpp:74-152), whose levels on a 2-D
lattice are as wide as the system; ordering the sites along a space-filling (Hilbert) curve of
their positions instead makes every stretch of consecutive rows a compact patch, so the rows a
sweep gathers stay within a few patches that fit in one XCD's L2.
"""
import numpy as np

__all__ = ["hilbert_index", "locality_order"]


def hilbert_index(ix, iy, order):
    """Position along the Hilbert curve of the 2^order x 2^order grid for integer cells (ix, iy)
    (the rotate-and-reflect construction, vectorised)."""
    ix = np.asarray(ix, dtype=np.int64).copy()
    iy = np.asarray(iy, dtype=np.int64).copy()
    d = np.zeros_like(ix)
    s = 1 << (order - 1)
    while s > 0:
        rx = ((ix & s) > 0).astype(np.int64)
        ry = ((iy & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # reflect when (rx, ry) = (1, 0), transpose when ry = 0 (only the bits below s matter from
        # here on, so reflecting within s is enough)
        flip = ry == 0
        refl = flip & (rx == 1)
        ix = np.where(refl, s - 1 - ix, ix)
        iy = np.where(refl, s - 1 - iy, iy)
        ix, iy = np.where(flip, iy, ix), np.where(flip, ix, iy)
        s >>= 1
    return d


def locality_order(x, y):
    """Order of points (x, y) along a Hilbert curve: returns q with q[i] = new index of point i.
    Points are binned on a square grid of about one point per cell; ties keep their order."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = x.size
    if n == 0:
        return np.zeros(0, dtype=np.int32)
    order = max(1, int(np.ceil(np.log2(max(2.0, np.sqrt(n))))) + 1)
    cells = 1 << order
    span = max(x.max() - x.min(), y.max() - y.min(), 1e-30)
    gx = np.minimum(((x - x.min()) / span * cells).astype(np.int64), cells - 1)
    gy = np.minimum(((y - y.min()) / span * cells).astype(np.int64), cells - 1)
    key = hilbert_index(gx, gy, order)
    q = np.empty(n, dtype=np.int32)
    q[np.argsort(key, kind="stable")] = np.arange(n, dtype=np.int32)
    return q
