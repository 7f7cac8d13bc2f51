"""ctypes binding of the engine's C-ABI (include/kpmx.h).

This is the binding a caller uses to drive the engine directly through its drop-in boundary
(plain pointers and sizes), without the C++ host module: the GPU parity tests and the benchmark
use it, so that what they measure and check is the boundary itself.
"""
import ctypes
import os
import re
import weakref

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(_PKG)
HEADER = os.path.join(ROOT, "include", "kpmx.h")
LIB = os.path.join(_PKG, "libkpmx.so")

SCALAR = {np.dtype(np.float32): 0, np.dtype(np.float64): 1,
          np.dtype(np.complex64): 2, np.dtype(np.complex128): 3}


def declared_functions():
    """Names of every function declared in include/kpmx.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kpmx_[a-z0-9_]+)\s*\(", text)))


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(LIB)
        _lib.kpmx_last_error.restype = ctypes.c_char_p
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(rc):
    if rc != 0:
        raise RuntimeError("kpmx error %d: %s" % (rc, lib().kpmx_last_error().decode()))


def _destroy(fn_name, handle, *keep_alive):
    """Finalizer body: free one engine handle.  `keep_alive` holds the objects the handle depends
    on (its context, its matrix) so that they outlive it."""
    getattr(lib(), fn_name)(handle)


class _Handle:
    """An engine object owned by Python: freed by close(), or when it is garbage collected (or at
    interpreter exit), whichever comes first.  Objects made on a context keep that context alive,
    and closing a context first closes what is still open on it (tapes before the rest)."""

    _destroy_fn = None
    _close_rank = 1

    def _own(self, ctx=None, *depends_on):
        deps = (ctx,) + depends_on if ctx is not None else depends_on
        self._finalizer = weakref.finalize(self, _destroy, self._destroy_fn, ctypes.c_void_p(self.h.value), *deps)
        if ctx is not None:
            ctx._children.add(self)

    def close(self):
        fin = getattr(self, "_finalizer", None)
        if fin is not None:
            fin()
        self.h = ctypes.c_void_p()


class Ctx(_Handle):
    _destroy_fn = "kpmx_ctx_destroy"

    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        self._children = weakref.WeakSet()
        _check(lib().kpmx_ctx_create(ctypes.c_int(device), ctypes.byref(self.h)))
        self._own()

    def close(self):
        for child in sorted(list(self._children), key=lambda c: c._close_rank):
            child.close()
        _Handle.close(self)

    def mat(self, csr, fmt=0):
        return Mat(self, csr, fmt)

    def block(self, r0):
        return Block(self, r0)

    def synchronize(self):
        _check(lib().kpmx_ctx_synchronize(self.h))

    def timer_start(self):
        """Record a HIP event on the context's stream (start of a timed region)."""
        _check(lib().kpmx_ctx_timer_start(self.h))

    def timer_stop(self):
        """Record the closing event, wait for it, and return the elapsed milliseconds."""
        ms = ctypes.c_double()
        _check(lib().kpmx_ctx_timer_stop(self.h, ctypes.byref(ms)))
        return ms.value

    def set_step_variant(self, dtype, vector_width, variant):
        """Pin the diagonal step kernel's variant for one scalar type and block width."""
        _check(lib().kpmx_ctx_set_step_variant(self.h, ctypes.c_int(SCALAR[np.dtype(dtype)]), ctypes.c_int(vector_width),
                                               ctypes.c_int(variant)))

    def set_adjoint_variant(self, variant):
        """Pin the adjoint step kernel's variant (-1: the default)."""
        _check(lib().kpmx_ctx_set_adjoint_variant(self.h, ctypes.c_int(variant)))

    def set_tile_width(self, dtype, width):
        """Widest column tile the diagonal recurrence runs for one scalar type (wider starter tiles
        run as sub-tiles of this many columns)."""
        _check(lib().kpmx_ctx_set_tile_width(self.h, ctypes.c_int(SCALAR[np.dtype(dtype)]), ctypes.c_int(width)))

    def tile_width(self, dtype):
        w = ctypes.c_int()
        _check(lib().kpmx_ctx_get_tile_width(self.h, ctypes.c_int(SCALAR[np.dtype(dtype)]), ctypes.byref(w)))
        return w.value

    def set_two_step(self, dtype, tile_rows, workgroups_per_cu=0):
        """Two Chebyshev steps per launch over tiles of `tile_rows` rows for the diagonal
        recurrence of one scalar type (0: single steps), on a persistent grid of
        `workgroups_per_cu` workgroups per CU (0: one workgroup per tile)."""
        _check(lib().kpmx_ctx_set_two_step(self.h, ctypes.c_int(SCALAR[np.dtype(dtype)]), ctypes.c_int(tile_rows),
                                           ctypes.c_int(workgroups_per_cu)))

    def set_two_step_lds(self, workgroups_per_cu=0, dtype=np.float64):
        """The LDS form of the two-step recurrence (fp64 or complex128; 256-row tiles, the
        intermediate vector in LDS) on a persistent grid (kpmx_ctx_set_two_step_lds); set_two_step
        turns it off."""
        _check(lib().kpmx_ctx_set_two_step_lds(self.h, ctypes.c_int(SCALAR[np.dtype(dtype)]),
                                               ctypes.c_int(workgroups_per_cu)))

    def two_step_lds(self, dtype=np.float64):
        """(enabled, workgroups per CU) of the LDS two-step form (kpmx_ctx_get_two_step_lds)."""
        e, w = ctypes.c_int(), ctypes.c_int()
        _check(lib().kpmx_ctx_get_two_step_lds(self.h, ctypes.c_int(SCALAR[np.dtype(dtype)]), ctypes.byref(e),
                                               ctypes.byref(w)))
        return bool(e.value), w.value

    def set_autotune_two_step(self, enable):
        """Let the diagonal autotuner choose the LDS two-step form (kpmx_ctx_set_autotune_two_step)."""
        _check(lib().kpmx_ctx_set_autotune_two_step(self.h, ctypes.c_int(1 if enable else 0)))

    def two_step(self, dtype):
        p = ctypes.c_int()
        _check(lib().kpmx_ctx_get_two_step(self.h, ctypes.c_int(SCALAR[np.dtype(dtype)]), ctypes.byref(p)))
        return p.value

    def set_graph_policy(self, policy):
        """Graph policy of the diagonal recurrence: 0 never (default), 1 auto (small starter blocks,
        captured on the second identical run), 2 always."""
        _check(lib().kpmx_ctx_set_graph_policy(self.h, int(policy)))

    def stream_probe(self, nbytes=1 << 31):
        """(copy GB/s, read GB/s, triad GB/s) of the device's streaming probe (kpmx_stream_probe)."""
        c, r, t = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().kpmx_stream_probe(self.h, ctypes.c_int64(nbytes), ctypes.byref(c), ctypes.byref(r),
                                       ctypes.byref(t)))
        return c.value, r.value, t.value


class Mat(_Handle):
    _destroy_fn = "kpmx_mat_destroy"

    def __init__(self, ctx, csr, fmt=0):
        csr = csr.tocsr()
        self.dtype = np.dtype(csr.dtype)
        self.n = csr.shape[0]
        self.indptr = np.ascontiguousarray(csr.indptr, dtype=np.int32)
        self.indices = np.ascontiguousarray(csr.indices, dtype=np.int32)
        self.data = np.ascontiguousarray(csr.data)
        self.h = ctypes.c_void_p()
        _check(lib().kpmx_mat_upload(ctx.h, ctypes.c_int(SCALAR[self.dtype]), ctypes.c_int(fmt),
                                     ctypes.c_int64(self.n), ctypes.c_int64(len(self.data)), _p(self.indptr),
                                     _p(self.indices), _p(self.data), ctypes.byref(self.h)))
        self._ctx = ctx
        self._own(ctx)

    def coded_bytes(self):
        """Bytes of the value-table form of the SELL matrix (0: none; kpmx_mat_coded_bytes)."""
        b = ctypes.c_int64()
        _check(lib().kpmx_mat_coded_bytes(self.h, ctypes.byref(b)))
        return b.value


class Block(_Handle):
    _destroy_fn = "kpmx_block_destroy"

    def __init__(self, ctx, r0):
        r0 = np.ascontiguousarray(np.atleast_2d(r0))
        self.dtype = r0.dtype
        self.nvec, self.n = r0.shape
        self.h = ctypes.c_void_p()
        _check(lib().kpmx_block_upload(ctx.h, ctypes.c_int(SCALAR[self.dtype]), ctypes.c_int64(self.n),
                                       ctypes.c_int64(self.nvec), _p(r0), ctypes.byref(self.h)))
        self._own(ctx)


def moments_diagonal(mat, block, num_moments):
    out = np.empty((block.nvec, num_moments), dtype=mat.dtype)
    _check(lib().kpmx_moments_diagonal(mat.h, block.h, ctypes.c_int64(num_moments), _p(out)))
    return out


def moments_diagonal_noalloc(mat, block, num_moments):
    """kpmx_moments_diagonal_noalloc: the diagonal moments with the recurrence's buffers in the
    library's static device scratch array (no device allocation)."""
    out = np.empty((block.nvec, num_moments), dtype=mat.dtype)
    _check(lib().kpmx_moments_diagonal_noalloc(mat.h, block.h, ctypes.c_int64(num_moments), _p(out)))
    return out


def autotune_diagonal(mat, block):
    """Pick the fastest tile shape of the diagonal step kernel for this matrix and block."""
    v = ctypes.c_int()
    _check(lib().kpmx_autotune_diagonal(mat.h, block.h, ctypes.byref(v)))
    return v.value


def moments_diagonal_run(mat, block, num_moments):
    """Run the diagonal recurrence and keep the raw sums on the device (no host transfer)."""
    _check(lib().kpmx_moments_diagonal_run(mat.h, block.h, ctypes.c_int64(num_moments)))


def moments_diagonal_fetch(mat, block, num_moments):
    out = np.empty((block.nvec, num_moments), dtype=mat.dtype)
    _check(lib().kpmx_moments_diagonal_fetch(mat.h, block.h, ctypes.c_int64(num_moments), _p(out)))
    return out


def moments_generic(mat, block, num_moments, beta, op=None):
    out = np.empty(num_moments, dtype=mat.dtype)
    beta = np.ascontiguousarray(beta, dtype=mat.dtype)
    if op is not None:
        op = op.tocsr()
        ip = np.ascontiguousarray(op.indptr, dtype=np.int32)
        ix = np.ascontiguousarray(op.indices, dtype=np.int32)
        dv = np.ascontiguousarray(op.data, dtype=mat.dtype)
        _check(lib().kpmx_moments_generic(mat.h, block.h, ctypes.c_int64(num_moments), _p(beta),
                                          ctypes.c_int64(len(dv)), _p(ip), _p(ix), _p(dv), _p(out)))
    else:
        _check(lib().kpmx_moments_generic(mat.h, block.h, ctypes.c_int64(num_moments), _p(beta),
                                          ctypes.c_int64(0), None, None, None, _p(out)))
    return out


def moments_multi_unit(mat, block, num_moments, dest):
    dest = np.ascontiguousarray(dest, dtype=np.int32)
    out = np.empty((len(dest), num_moments), dtype=mat.dtype)
    _check(lib().kpmx_moments_multi_unit(mat.h, block.h, ctypes.c_int64(num_moments), ctypes.c_int64(len(dest)),
                                         _p(dest), _p(out)))
    return out


def moments_dense(mat, block, num_moments, op=None):
    out = np.empty((block.nvec, num_moments, mat.n), dtype=mat.dtype)
    _check(lib().kpmx_moments_dense(mat.h, block.h, ctypes.c_int64(num_moments), op.h if op is not None else None,
                                    _p(out)))
    return out


def kubo_moments(mat, vl, vr, block, num_moments, mode=0):
    out = np.empty((num_moments, num_moments), dtype=mat.dtype)
    _check(lib().kpmx_kubo_moments(mat.h, vl.h, vr.h, block.h, ctypes.c_int64(num_moments), ctypes.c_int(mode), _p(out)))
    return out


def lanczos_bounds(mat, precision_percent=0.002):
    lo, hi, loops = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    _check(lib().kpmx_lanczos_bounds(mat.h, ctypes.c_double(precision_percent), ctypes.byref(lo), ctypes.byref(hi),
                                     ctypes.byref(loops)))
    return lo.value, hi.value, loops.value


class Batch(_Handle):
    """Batched independent systems (kpmx_batch_*): `h2s` are the already scaled H2 of each
    system; they are laid out block-diagonally."""

    _destroy_fn = "kpmx_batch_destroy"

    def __init__(self, ctx, h2s):
        import scipy.sparse as sp
        h2s = [sp.csr_matrix(h) for h in h2s]
        self.dtype = np.dtype(h2s[0].dtype)
        self.sizes = [h.shape[0] for h in h2s]
        self.row_offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        big = sp.block_diag(h2s, format="csr", dtype=self.dtype)
        big.sort_indices()
        self.indptr = np.ascontiguousarray(big.indptr, dtype=np.int32)
        self.indices = np.ascontiguousarray(big.indices, dtype=np.int32)
        self.data = np.ascontiguousarray(big.data)
        self.n = big.shape[0]
        self.nvec = 0
        self.h = ctypes.c_void_p()
        _check(lib().kpmx_batch_upload(ctx.h, ctypes.c_int(SCALAR[self.dtype]), ctypes.c_int64(len(h2s)),
                                       _p(self.row_offsets), ctypes.c_int64(len(self.data)), _p(self.indptr),
                                       _p(self.indices), _p(self.data), ctypes.byref(self.h)))
        self._ctx = ctx
        self._own(ctx)

    def starters(self, r0):
        """r0: (num_vectors, N) over the concatenated systems."""
        r0 = np.ascontiguousarray(np.atleast_2d(r0), dtype=self.dtype)
        self.nvec = r0.shape[0]
        _check(lib().kpmx_batch_starters(self.h, ctypes.c_int64(self.nvec), _p(r0)))

    def run(self, num_moments, use_graph=True):
        """All num_moments / 2 steps: use_graph False (one launch per step), True (one HIP graph)
        or "persistent" (one persistent kernel launch with grid barriers between steps)."""
        mode = 2 if use_graph == "persistent" else (1 if use_graph else 0)
        _check(lib().kpmx_batch_run(self.h, ctypes.c_int64(num_moments), ctypes.c_int(mode)))

    def autotune(self):
        """Pick the fastest slice partition (kpmx_batch_autotune); returns (workgroups, walk_chunk),
        walk_chunk 0 for contiguous ranges."""
        wg, chunk = ctypes.c_int(), ctypes.c_int()
        _check(lib().kpmx_batch_autotune(self.h, ctypes.byref(wg), ctypes.byref(chunk)))
        return wg.value, chunk.value

    def set_partition(self, num_workgroups, walk_chunk=0):
        """Deal the slices to `num_workgroups` workgroups: contiguous ranges (walk_chunk 0) or the
        XCD-aware walk with runs of `walk_chunk` slices (kpmx_batch_set_partition)."""
        _check(lib().kpmx_batch_set_partition(self.h, ctypes.c_int(num_workgroups), ctypes.c_int(walk_chunk)))

    def set_tile_width(self, width):
        """Widest column tile (vectors) for the next `starters` upload; 0 = the library default
        (kpmx_batch_set_tile_width)."""
        _check(lib().kpmx_batch_set_tile_width(self.h, ctypes.c_int(width)))

    def tiles(self):
        """(number of column tiles, vectors per tile) of the uploaded starters (kpmx_batch_tiles)."""
        nt, tv = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().kpmx_batch_tiles(self.h, ctypes.byref(nt), ctypes.byref(tv)))
        return nt.value, tv.value

    def set_value_table(self, enable):
        """Select the value-table matrix form (True) or the standard one (False) for the step
        (kpmx_batch_set_value_table); RuntimeError when the batch has no value-table form."""
        _check(lib().kpmx_batch_set_value_table(self.h, ctypes.c_int(1 if enable else 0)))

    def format(self):
        """(value table available, value table in use, matrix bytes one step reads) of the
        batch (kpmx_batch_format)."""
        avail, used, nbytes = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        _check(lib().kpmx_batch_format(self.h, ctypes.byref(avail), ctypes.byref(used), ctypes.byref(nbytes)))
        return bool(avail.value), bool(used.value), nbytes.value

    def fetch(self, num_moments):
        out = np.empty((len(self.sizes), self.nvec, num_moments), dtype=self.dtype)
        _check(lib().kpmx_batch_fetch(self.h, ctypes.c_int64(num_moments), _p(out)))
        return out


class Tape(_Handle):
    """Forward record of an adjoint calculation (kpmx_adjoint_forward).  The tape is freed by
    close() or when it is garbage collected; it keeps its matrix and starter block alive."""

    _destroy_fn = "kpmx_tape_destroy"
    _close_rank = 0

    def __init__(self, mat, block, num_moments, checkpoint=64):
        self.mat = mat
        self.num_moments = num_moments
        self.mu = np.empty((block.nvec, num_moments), dtype=mat.dtype)
        self.h = ctypes.c_void_p()
        _check(lib().kpmx_adjoint_forward(mat.h, block.h, ctypes.c_int64(num_moments), ctypes.c_int64(checkpoint),
                                          _p(self.mu), ctypes.byref(self.h)))
        self._own(mat._ctx, mat, block)

    def backward(self, w):
        """dL/dH2 (CSR order of the uploaded matrix) for L = sum_n w[n] * mean over vectors of mu_n."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        grad = np.empty(len(self.mat.data), dtype=np.float64)
        if not self.h:
            raise RuntimeError("the adjoint tape was already released")
        _check(lib().kpmx_adjoint_backward(self.h, _p(w), _p(grad)))
        return grad


def chebyshev_series(ctx, mu, theta, pref):
    """out[e, c] = pref[e] * sum_n mu[c, n] * f(n theta[e]) with f = cos for real mu and pref
    (kpmx_chebyshev_cos_series) or exp(-i .) for complex ones (kpmx_chebyshev_exp_series); the sums
    run in mu's real precision."""
    mu = np.atleast_2d(np.ascontiguousarray(mu))
    cplx = np.iscomplexobj(mu) or np.iscomplexobj(pref)
    real = np.float32 if mu.real.dtype == np.float32 else np.float64
    code = 0 if real == np.float32 else 1
    cols, M = mu.shape
    theta = np.ascontiguousarray(theta, dtype=real)
    if cplx:
        mu = np.ascontiguousarray(mu, dtype=np.complex64 if code == 0 else np.complex128)
        pref = np.ascontiguousarray(pref, dtype=mu.dtype)
        out = np.empty((len(theta), cols), dtype=np.complex128)
        fn = lib().kpmx_chebyshev_exp_series
    else:
        mu = np.ascontiguousarray(mu, dtype=real)
        pref = np.ascontiguousarray(pref, dtype=real)
        out = np.empty((len(theta), cols), dtype=np.float64)
        fn = lib().kpmx_chebyshev_cos_series
    _check(fn(ctx.h, code, ctypes.c_int64(M), ctypes.c_int64(cols), _p(mu), ctypes.c_int64(len(theta)), _p(theta),
              _p(pref), _p(out)))
    return out


def kubo_bastin_sums(ctx, mu_nm, samples):
    """sum_nm mu_nm Gamma_nm(E_q) / (1 - E_q^2)^2 for every scaled sample E_q (kpmx_kubo_bastin_sums)."""
    mu = np.ascontiguousarray(mu_nm, dtype=np.complex128)
    samples = np.ascontiguousarray(samples, dtype=np.float64)
    out = np.empty(len(samples), dtype=np.complex128)
    _check(lib().kpmx_kubo_bastin_sums(ctx.h, ctypes.c_int64(mu.shape[0]), _p(mu), ctypes.c_int64(len(samples)),
                                       _p(samples), _p(out)))
    return out


def adjoint_diagonal(mat, block, num_moments, w, checkpoint=64):
    """(dL/dH2, moments) in one call (kpmx_adjoint_diagonal)."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    grad = np.empty(len(mat.data), dtype=np.float64)
    mu = np.empty((block.nvec, num_moments), dtype=mat.dtype)
    _check(lib().kpmx_adjoint_diagonal(mat.h, block.h, ctypes.c_int64(num_moments), _p(w), ctypes.c_int64(checkpoint),
                                       _p(grad), _p(mu)))
    return grad, mu
