"""ORACLE -- test infrastructure only.

CPU restatement of the reference (pybinding) KPM moment engine, used as the checker for the HIP
product path and as the CPU baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
benchmark's ``cpu_baseline`` leg may import this package; the product (``pybinding_amd``) never
does, and it fails loudly when its HIP library is missing instead of falling back to this code.

Layout
------
- ``kpm_oracle.c`` (+ ``kpm_oracle_impl.h``): mt19937 stream, starters, CSR recurrences and
  collectors, for f32/f64/c64/c128 (wrapped by the functions below).
- ``kpm_cpu_baseline.c``: threaded ELL/SIMD-batch restatement of the reference default CPU compute
  (timed as the CPU baseline).
- This module: ctypes wrappers plus numpy restatements of the small host pieces (scaling,
  kernels, BFS reorder / slice map, reconstruction, velocity).
- ``brute.py``: dense eigendecomposition closed forms used to pin the recurrences.

Parity status of each function is stated in its docstring; a function with no independent pin
says "parity unpinned".
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

_DT = {
    "f32": (np.float32, "f32"),
    "f64": (np.float64, "f64"),
    "c64": (np.complex64, "c64"),
    "c128": (np.complex128, "c128"),
}

_lib = None
_blib = None


def build():
    """Compile the oracle C sources into oracle/_build (gcc, a few seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib, _blib
    if _lib is None:
        path = os.path.join(_BUILD, "libkpm_oracle.so")
        if not os.path.exists(path):
            build()
        _lib = ctypes.CDLL(path)
        _blib = ctypes.CDLL(os.path.join(_BUILD, "libkpm_cpu_baseline.so"))
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _sfx(dtype):
    dtype = np.dtype(dtype)
    return {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64",
            np.dtype(np.complex64): "c64", np.dtype(np.complex128): "c128"}[dtype]


# --------------------------------------------------------------------------------- RNG / starters
def mt19937_stream(n, seed=5489):
    """Raw std::mt19937 output.  Pinned by the C++ standard's known answer (10000th draw of a
    default-constructed engine is 4123659995) and by numpy's legacy MT19937 stream."""
    lib = _load()
    out = np.empty(n, dtype=np.uint32)
    lib.orc_mt19937_stream(ctypes.c_uint32(seed), ctypes.c_int64(n), _ptr(out))
    return out


def uniform(n, dtype=np.float64):
    """std::uniform_real_distribution<real_t> draws from mt19937(5489), libstdc++ rounding
    (cppcore/include/numeric/random.hpp:24-44).  Pinned against a direct restatement of
    generate_canonical over the numpy MT19937 stream (tests/test_oracle_rng.py)."""
    lib = _load()
    dtype = np.dtype(dtype)
    out = np.empty(n, dtype=dtype)
    fn = lib.orc_uniform_f64 if dtype == np.float64 else lib.orc_uniform_f32
    fn(ctypes.c_int64(n), _ptr(out))
    return out


def random_starter(n, nvec, dtype):
    """Random starter columns (cppcore/src/kpm/Starter.cpp:59-82), in ORIGINAL site order,
    returned as an (nvec, n) array.  Real: +-1 by `u < 0.5f`; complex: exp(i*2*pi_f*u)."""
    lib = _load()
    out = np.empty((nvec, n), dtype=np.dtype(dtype))
    getattr(lib, "orc_random_starter_" + _sfx(dtype))(ctypes.c_int64(n), ctypes.c_int64(nvec), _ptr(out))
    return out


# --------------------------------------------------------------------------------- recurrences
def _csr(h2, dtype):
    h2 = h2.tocsr()  # the given in-row order is kept (it only affects rounding)
    indptr = np.ascontiguousarray(h2.indptr, dtype=np.int32)
    indices = np.ascontiguousarray(h2.indices, dtype=np.int32)
    data = np.ascontiguousarray(h2.data, dtype=dtype)
    return indptr, indices, data


def kpm_spmv(h2, x, y):
    """y <- H2 x - y, CSR (cppcore/include/compute/kernel_polynomial.hpp:18-32).  Returns new y."""
    lib = _load()
    dtype = np.dtype(h2.dtype)
    indptr, indices, data = _csr(h2, dtype)
    x = np.ascontiguousarray(x, dtype=dtype)
    y = np.array(y, dtype=dtype, copy=True)
    getattr(lib, "orc_kpm_spmv_" + _sfx(dtype))(ctypes.c_int64(h2.shape[0]), _ptr(indptr), _ptr(indices),
                                                _ptr(data), _ptr(x), _ptr(y))
    return y


def diag_moments(h2, r0s, num_moments):
    """Raw diagonal moments mu_n = <r|T_n|r> (mu_0 halved, as the reference stores it) for each
    row of `r0s` (shape (nvec, n), already in the matrix's order).  Returns (nvec, M).

    Restates calc_moments::basic + DiagonalCollector (calc_moments.hpp:35-51,
    collectors.cpp:5-15).  Pinned by brute.moments_eig (dense eigendecomposition)."""
    lib = _load()
    dtype = np.dtype(h2.dtype)
    indptr, indices, data = _csr(h2, dtype)
    r0s = np.ascontiguousarray(np.atleast_2d(r0s), dtype=dtype)
    nvec, n = r0s.shape
    out = np.empty((nvec, num_moments), dtype=dtype)
    rc = getattr(lib, "orc_diag_moments_" + _sfx(dtype))(
        ctypes.c_int64(n), _ptr(indptr), _ptr(indices), _ptr(data), ctypes.c_int64(nvec), _ptr(r0s),
        ctypes.c_int64(num_moments), _ptr(out))
    if rc != 0:
        raise ValueError("diag_moments: num_moments must be even and >= 2")
    return out


def offdiag_moments(h2, r0, num_moments, kind, beta=None, op=None, dest=None):
    """Off-diagonal recurrence (calc_moments.hpp:102-114) with one collector
    (collectors.cpp:36-106): kind "generic" -> (M,), "multi_unit" -> (len(dest), M),
    "dense" -> (M, n).  Pinned by brute.moments_eig / brute.vectors_eig."""
    lib = _load()
    dtype = np.dtype(h2.dtype)
    n = h2.shape[0]
    indptr, indices, data = _csr(h2, dtype)
    r0 = np.ascontiguousarray(r0, dtype=dtype)
    k = {"generic": 0, "multi_unit": 1, "dense": 2}[kind]
    beta_a = np.ascontiguousarray(beta if beta is not None else r0, dtype=dtype) if k == 0 else None
    if op is not None:
        op_indptr, op_indices, op_data = _csr(op, dtype)
    else:
        op_indptr = op_indices = op_data = None
    dest_a = np.ascontiguousarray(dest, dtype=np.int32) if dest is not None else None
    ndest = len(dest_a) if dest_a is not None else 0
    if k == 0:
        out = np.empty(num_moments, dtype=dtype)
    elif k == 1:
        out = np.empty((ndest, num_moments), dtype=dtype)
    else:
        out = np.empty((num_moments, n), dtype=dtype)
    rc = getattr(lib, "orc_offdiag_moments_" + _sfx(dtype))(
        ctypes.c_int64(n), _ptr(indptr), _ptr(indices), _ptr(data), _ptr(r0), ctypes.c_int64(num_moments),
        ctypes.c_int(k), _ptr(beta_a), _ptr(op_indptr), _ptr(op_indices), _ptr(op_data),
        ctypes.c_int64(ndest), _ptr(dest_a), _ptr(out))
    if rc != 0:
        raise ValueError("offdiag_moments failed (%d)" % rc)
    return out


# --------------------------------------------------------------------------------- host pieces
F32_PI = float(np.float32(np.pi))            # constant::pi (cppcore/include/numeric/constant.hpp:8)
F32_TOL = float(np.float32(1) + np.float32(0.01))   # 1 + 0.01f (cppcore/include/kpm/Bounds.hpp:12,19)
F32_BSNAP = float(np.float32(0.01) * np.float32(0.01))  # 0.01f * tolerance (Bounds.hpp:21)
F32_KB = float(np.float32(8.6173303e-5))     # constant::kb (constant.hpp:20)


def scale_factors(min_energy, max_energy):
    """Scale<double>(min, max) from float config bounds (cppcore/include/kpm/Bounds.hpp:16-24;
    bounds are float in kpm::Config, cppcore/include/kpm/Config.hpp:24-25).  Pinned by the
    closed-form value a = 8.686000303268429 for (-8.6, 8.6) worked out independently."""
    lo = float(np.float32(min_energy))
    hi = float(np.float32(max_energy))
    a = 0.5 * (hi - lo) * F32_TOL
    b = 0.5 * (hi + lo)
    if abs(b / a) < F32_BSNAP:
        b = 0.0
    return a, b


def scale_hamiltonian(h, a, b, dtype):
    """H2 = (H - b I) * (2/a), unreordered path (cppcore/src/kpm/OptimizedHamiltonian.cpp:54-71),
    computed in the matrix's real type.  Returns a scipy CSR matrix of `dtype`."""
    import scipy.sparse as sp
    dtype = np.dtype(dtype)
    real = np.float32 if dtype in (np.float32, np.complex64) else np.float64
    ra, rb = real(a), real(b)
    h = sp.csr_matrix(h, dtype=dtype)
    if rb == 0:
        out = h.copy()
        out.data = (out.data * (real(2) / ra)).astype(dtype)
        return out
    eye = sp.identity(h.shape[0], dtype=dtype, format="csr")
    out = (h - eye * rb).tocsr()
    out.data = (out.data * (real(2) / ra)).astype(dtype)
    out.sort_indices()
    return out


def round_num_moments(n):
    """cppcore/include/kpm/Kernel.hpp:9-13"""
    if n < 2:
        return 2
    while (n - 2) % 4 != 0:
        n += 1
    return n


def jackson(num_moments):
    """Jackson damping with float pi (cppcore/src/kpm/Kernel.cpp:6-21).  Pinned against the
    kernel's defining convolution sum g_n = sum_v a_v a_{v+n} / sum_v a_v^2,
    a_v = sin(pi (v+1)/(N+1)) (tests/test_oracle_kernels.py)."""
    N = float(num_moments)
    Np = N + 1
    n = np.arange(num_moments, dtype=np.float64)
    pi = F32_PI
    return ((Np - n) * np.cos(pi * n / Np) + np.sin(pi * n / Np) / np.tan(pi / Np)) / Np


def lorentz(num_moments, lam=4.0):
    """Lorentz damping sinh(lam (1 - n/N)) / sinh(lam) (Kernel.cpp:24-38).  Pinned by
    g_0 = 1, g_N = 0 and by the Lorentzian it reconstructs (tests/test_oracle_kernels.py)."""
    if lam <= 0:
        raise ValueError("Lorentz kernel: lambda must be positive.")
    N = float(num_moments)
    n = np.arange(num_moments, dtype=np.float64)
    return np.sinh(lam * (1 - n / N)) / np.sinh(lam)


def dirichlet(num_moments):
    return np.ones(num_moments)


def required_num_moments(kind, scaled_broadening, lam=4.0):
    """Kernel::required_num_moments (Kernel.cpp:17-20,33-36,44-47)."""
    num = lam if kind == "lorentz" else F32_PI
    return round_num_moments(int(num / scaled_broadening) + 1)


def lanczos_bounds(h, precision_percent=0.002, seed=0, loop_limit=1000):
    """Lanczos min/max eigenvalues (cppcore/include/compute/lanczos.hpp:93-154): three-term
    recurrence, eigenvalues of the growing tridiagonal matrix, stop when both extremes change by
    less than `precision_percent` % between two iterations.  The start vector is numpy's, not
    Eigen's Random, so the result agrees with the reference's to the stopping precision only.
    Returns (min, max, loops).  Pinned through the reference's LDOS/DOS fixtures, which use
    these bounds for their scaling."""
    from scipy.linalg import eigvalsh_tridiagonal
    h = h.tocsr().astype(np.complex128 if np.iscomplexobj(h.data) else np.float64)
    rng = np.random.default_rng(seed)
    v1 = rng.uniform(-1, 1, h.shape[0])
    if np.iscomplexobj(h.data):
        v1 = v1 + 1j * rng.uniform(-1, 1, h.shape[0])
    v1 /= np.linalg.norm(v1)
    v0 = np.zeros_like(v1)
    alpha, beta = [], []
    prev_min, prev_max = np.inf, -np.inf
    prec = precision_percent / 100
    for i in range(loop_limit):
        b_prev = beta[-1] if beta else 0.0
        w = h @ v1 - b_prev * v0
        a = float(np.real(np.vdot(v1, w)))
        w = w - a * v1
        b = float(np.linalg.norm(w))
        v0, v1 = v1, w / b
        alpha.append(a)
        beta.append(b)
        ev = eigvalsh_tridiagonal(np.array(alpha), np.array(beta[:-1])) if len(alpha) > 1 else np.array(alpha)
        mn, mx = ev.min(), ev.max()
        if abs((prev_min - mn) / mn) < prec and abs((prev_max - mx) / mx) < prec:
            return float(mn), float(mx), i
        prev_min, prev_max = mn, mx
    raise RuntimeError("Lanczos algorithm did not converge for the min/max eigenvalues.")


def bfs_reorder(h, src, b_nonzero=False):
    """BFS reorder from `src[0]` and slice borders (OptimizedHamiltonian.cpp:74-152).
    Returns (reorder_map, borders): reorder_map[original] = optimized index.
    Pinned by the reference's integer known answers for the reorder tests (test_kpm.cpp:19-172)
    where the model builder reproduces that model; otherwise by permutation invariance."""
    h = h.tocsr()
    n = h.shape[0]
    indptr, indices = h.indptr, h.indices
    reorder = np.full(n, -1, dtype=np.int64)
    queue = [int(src[0])]
    reorder[src[0]] = 0
    borders = [1]
    for h2_row in range(n):
        row = queue[h2_row]
        for k in range(indptr[row], indptr[row + 1]):
            col = int(indices[k])
            if reorder[col] < 0:
                reorder[col] = len(queue)
                queue.append(col)
        if h2_row == borders[-1] - 1:
            borders.append(len(queue))
    borders.pop()
    return reorder, borders


def slice_map_index(borders, src_offset, dest_offset, n, num_moments):
    """SliceMap::index (cppcore/include/kpm/OptimizedHamiltonian.hpp:67-77)."""
    last = len(borders) - 1
    mid = (num_moments - 1 + dest_offset - src_offset) // 2
    mx = min(last, mid + src_offset)
    if n < mid:
        return min(mx, n + src_offset)
    return min(mx, num_moments - 1 - n + dest_offset)


def slice_offset(borders, idx):
    """SliceMap offset of the highest index (OptimizedHamiltonian.cpp:5-19)."""
    mx = max(idx)
    for k, v in enumerate(borders):
        if v > mx:
            return k
    raise ValueError("index beyond the last border")


def spectral_density(moments, energy, a, b, real=np.float64):
    """SpectralDensity (cppcore/include/kpm/reconstruct.hpp:16-54): moments (M,) or (M, K)
    already damped; returns (E,) or (E, K).  Pinned by the sum rule
    integral = mu_0 * pi / pi_f and by the Chebyshev closed form on exact moments."""
    real = np.dtype(real).type
    ra, rb = real(a), real(b)
    e = (np.asarray(energy, dtype=real) - rb) / ra
    mom = np.real(np.asarray(moments)).astype(real)
    ns = np.arange(mom.shape[0], dtype=real)
    k = real(np.float32(2) / np.float32(np.pi)) / ra
    t = np.cos(ns[None, :] * np.arccos(e)[:, None])
    pref = k / np.sqrt(1 - e * e)
    if mom.ndim == 1:
        return (pref * (t @ mom)).astype(np.float64)
    return (pref[:, None] * (t @ mom)).astype(np.float64)


def greens_function(moments, energy, a, b, real=np.float64):
    """GreensFunction (reconstruct.hpp:59-85): g(E) = -2i/(a sqrt(1-E^2)) sum mu_n e^{-i n acos E}.
    Pinned by LDOS = -Im(G)/pi against spectral_density (the reference's own invariant,
    cppcore/tests/test_kpm.cpp:262)."""
    real = np.dtype(real).type
    ra, rb = real(a), real(b)
    e = (np.asarray(energy, dtype=real) - rb) / ra
    ns = np.arange(len(moments), dtype=real)
    k = -2j / ra
    ex = np.exp(-1j * ns[None, :] * np.arccos(e)[:, None])
    return k / np.sqrt(1 - e * e) * (ex @ np.asarray(moments))


def kubo_bastin(mu_nm, chemical_potential, energy_samples, temperature, a, b, real=np.float64):
    """KuboBastin reconstruction (reconstruct.hpp:90-143), the direct M x M form.  Pinned by
    the reference's conductivity fixtures (tests/baseline_data/kpm/conductivity[*].pbz, matched
    to 4e-5 relative from moments this oracle computes on the rebuilt models)."""
    real = np.dtype(real).type
    ra, rb = real(a), real(b)
    M = mu_nm.shape[0]
    with np.errstate(divide="ignore"):
        inv_kbt = real(np.float64(ra) / (F32_KB * np.float64(temperature)))
    mus = (np.asarray(chemical_potential, dtype=real) - rb) / ra
    es = (np.asarray(energy_samples, dtype=real) - rb) / ra
    ns = np.arange(M, dtype=real)
    sum_nm = np.empty(len(es), dtype=np.complex128)
    for i, en in enumerate(es):
        ac = np.arccos(en)
        sqrt_n = en - 1j * ns[None, :] * np.sqrt(1 - en * en)   # varies along m (columns)
        exp_n = np.exp(1j * ac * ns[None, :])
        t_m = np.cos(ac * ns[:, None])                        # varies along n (rows)
        g_p = sqrt_n * exp_n * t_m
        gamma = g_p + g_p.conj().T
        k = 1 / ((1 - en * en) * (1 - en * en))
        sum_nm[i] = k * np.sum(mu_nm * gamma)
    coeff = (es.max() - es.min()) / (2 * len(es))
    out = np.empty(len(mus), dtype=np.complex128)
    with np.errstate(over="ignore", invalid="ignore"):
        for j, mu in enumerate(mus):
            fd = real(1) / (real(1) + np.exp((es - mu) * inv_kbt))
            func = fd * sum_nm
            out[j] = (4 / (ra * ra)) * coeff * (2 * func.sum() - func[0] - func[-1])
    return out


def velocity(h, coords):
    """v_ij = (x_i - x_j) * H_ij with float32 positions (cppcore/src/kpm/Moments.cpp:132-156)."""
    import scipy.sparse as sp
    h = sp.csr_matrix(h)
    coords = np.asarray(coords, dtype=np.float32)
    rows = np.repeat(np.arange(h.shape[0]), np.diff(h.indptr))
    diff = (coords[rows] - coords[h.indices]).astype(np.float32)
    out = h.copy()
    out.data = out.data * diff.astype(h.dtype)
    return out


# --------------------------------------------------------------------------------- CPU baseline
def to_ell(h2):
    """csr_to_ell (cppcore/include/numeric/ellmatrix.hpp:64-82): column-major, 32-byte aligned
    pitch, padding value 0 with the previous row's index."""
    h2 = h2.tocsr()
    n = h2.shape[0]
    itemsize = np.dtype(h2.dtype).itemsize
    per = 32 // itemsize
    pitch = ((n + per - 1) // per) * per
    counts = np.diff(h2.indptr)
    width = int(counts.max()) if n else 0
    data = np.zeros((width, pitch), dtype=h2.dtype)
    idx = np.zeros((width, pitch), dtype=np.int32)
    for k in range(width):
        has = counts > k
        rows = np.nonzero(has)[0]
        data[k, rows] = h2.data[h2.indptr[rows] + k]
        idx[k, rows] = h2.indices[h2.indptr[rows] + k]
        # padding: index repeated from the previous row, 0 for row 0
        miss = np.nonzero(~has)[0]
        if len(miss):
            col = idx[k, :n].copy()
            for r in miss:
                col[r] = col[r - 1] if r > 0 else 0
            idx[k, :n] = col
    return data, idx, pitch, width


def cpu_baseline_dos(h2_ell, n, num_moments, num_random, nthreads, perm=None, borders=None,
                     interleaved=False):
    """Time the restated default CPU compute for DOS moments.  Returns (mean_moments, seconds)."""
    _load()
    data, idx, pitch, width = h2_ell
    dtype = np.dtype(data.dtype)
    sfx = "f64" if dtype == np.float64 else "f32"
    fn = getattr(_blib, "bl_dos_moments_" + sfx)
    fn.restype = ctypes.c_double
    out = np.empty(num_moments, dtype=dtype)
    perm_a = np.ascontiguousarray(perm, dtype=np.int32) if perm is not None else None
    if borders is None:
        borders = [n]
    b_a = np.ascontiguousarray(borders, dtype=np.int32)
    t = fn(ctypes.c_int64(n), ctypes.c_int64(pitch), ctypes.c_int64(width), _ptr(np.ascontiguousarray(data)),
           _ptr(np.ascontiguousarray(idx)), _ptr(perm_a), _ptr(b_a), ctypes.c_int64(len(b_a)),
           ctypes.c_int64(num_moments), ctypes.c_int64(num_random), ctypes.c_int64(nthreads),
           ctypes.c_int(1 if interleaved else 0), _ptr(out))
    if t < 0:
        raise ValueError("cpu baseline: bad moment count")
    return out, t
