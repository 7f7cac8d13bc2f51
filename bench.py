#!/usr/bin/env python3
"""Benchmark of the KPM moment engine on one MI355X.

Workload (one "step" = one full moment computation over one batch of synthetic input):
    Periodic honeycomb (graphene) supercell 2048 x 1024 cells, N = 4,194,304 sites, t = -2.8 eV,
    Anderson onsite disorder uniform in [-0.5, 0.5] eV (numpy PCG64 seed 2), real fp64,
    explicit energy range (-8.9, 8.9) eV, BFS-reordered H2 (the reference's default optimized
    Hamiltonian), 4096 moments, 64 Rademacher vectors from the reference's mt19937 starter stream
    computed as one 64-wide block SpMM recurrence (2048 fused Chebyshev steps).

Inputs (H2 and the starter block) are resident in HBM before the timed region; the timed region
runs K complete moment computations through the C-ABI (kpmx_moments_diagonal_run) and is
bracketed by stream synchronisation.  value = GNNZ*moments/s = nnz * M * R / t / 1e9.

roofline: the fused step kernel's algorithmic bytes per launch
    B_step = nnz*(s + 4) + 4*(N + 1) + 3*N*B*s
(the matrix term in the form the kernel reads: the value-table form's bytes where it reads that)
times the steps one launch does (two in the two-step form, which itself moves fewer bytes: the
counted traffic) divided by its average launch duration (HIP events on the engine's stream around
the timed region / number of launches), against 8 TB/s.

cpu_baseline: the oracle's restatement of the reference's default CPU compute (ELL, 4-wide fp64
vector batches, threads over batches, interleaved BFS slices) on a bounded sample of the same
workload: the same matrix and the same 64 vectors with fewer moments; reported in the same metric.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(n1, n2, disorder, dtype, order="bfs"):
    from pybinding_amd import generators
    from pybinding_amd._native import load_module
    import scipy.sparse as sp
    native = load_module()
    t0 = time.time()
    h = generators.honeycomb_supercell(n1, n2, disorder=disorder, seed=2, dtype=dtype)
    log("built H: n=%d nnz=%d (%.1fs)" % (h.shape[0], h.nnz, time.time() - t0))
    t0 = time.time()
    if order == "hilbert":
        q = generators.honeycomb_cell_order(n1, n2)
        inv = np.argsort(q)
        hp = h[inv][:, inv].tocsr()
        hp.sort_indices()
        ip, ix, dv, _, borders, _, _, (a, b) = native.optimize(hp.shape, hp.indptr, hp.indices, hp.data, (-8.9, 8.9),
                                                               False, [0], [0])
        rmap = q.astype(np.int32)
    else:
        ip, ix, dv, rmap, borders, _, _, (a, b) = native.optimize(h.shape, h.indptr, h.indices, h.data, (-8.9, 8.9),
                                                                   True, [0], [0])
    h2 = sp.csr_matrix((dv, ix, ip), shape=h.shape)
    log("optimized H2 (scale a=%.6f b=%.3g, %s order, %d slices) (%.1fs)" % (a, b, order, len(borders), time.time() - t0))
    return h2, rmap, borders, native


def starters(native, n, nvec, dtype, rmap):
    t0 = time.time()
    r0 = native.random_starters(n, nvec, np.dtype(dtype).name, rmap)
    log("drew %d reference starters (%.1fs)" % (nvec, time.time() - t0))
    return r0


def step_bytes(n, nnz, vw, s):
    return nnz * (s + 4) + 4 * (n + 1) + 3 * n * vw * s


def launch_widths(R, tile_width):
    """Vector widths of the step launches of one recurrence step over R vectors: starter tiles of
    up to 64 vectors (rounded up to a power of two), each run as sub-tiles of `tile_width`
    columns when wider (kpmx_ctx_set_tile_width)."""
    out = []
    for c0 in range(0, R, 64):
        cols = min(64, R - c0)
        vw = 1 << (cols - 1).bit_length()
        if vw <= tile_width:
            out.append(vw)
        else:
            out += [tile_width] * ((cols + tile_width - 1) // tile_width)
    return out


def periodic_graphene_model(n1, n2, disorder, seed, field=None):
    """The periodic honeycomb supercell through the model builder: graphene.monolayer() on a
    primitive n1 x n2 block with translational symmetry of the whole block (a complex128
    Hamiltonian, as the reference promotes periodic models), Anderson disorder uniform in
    [-W/2, W/2] from numpy PCG64(seed) and, optionally, a perpendicular field (Peierls phases)."""
    import pybinding_amd as pb
    from pybinding_amd.repository import graphene
    rng = np.random.default_rng(seed)

    @pb.onsite_energy_modifier(is_double=True)
    def anderson(energy):
        return energy + rng.uniform(-disorder / 2, disorder / 2, size=energy.shape)

    mods = [anderson] + ([graphene.constant_magnetic_field(field)] if field else [])
    return pb.Model(graphene.monolayer(), pb.primitive(n1, n2),
                    pb.translational_symmetry(a1=n1 * graphene.a, a2=n2 * graphene.a), *mods)


def model_leg(cells=(2048, 1024), M=4096, R=64):
    """The block-SpMM workload built by the model builder, in the reference's own scalar type for
    it (complex128, phase starters): (1) the recurrence through the C-ABI with inputs resident, in
    the headline metric; (2) `kpm(model).calc_dos` at 2000 energies end to end (site order,
    starters, upload, moments, damping, reconstruction), model already built."""
    import pybinding_amd as pb
    from pybinding_amd import abi
    from pybinding_amd._native import load_module
    import scipy.sparse as sp
    native = load_module()
    t0 = time.time()
    model = periodic_graphene_model(cells[0], cells[1], 1.0, 2)
    h = model.hamiltonian
    n, nnz = h.shape[0], h.nnz
    log("model: %d sites, %d nnz, %s (%.1fs)" % (n, nnz, h.dtype, time.time() - t0))
    out = {"workload": "periodic honeycomb %dx%d cells through the model builder (N=%d, nnz=%d, complex128), "
                       "Anderson W=1 eV, %d moments, %d phase starters" % (cells[0], cells[1], n, nnz, M, R)}
    # (1) C-ABI, Hilbert site order, resident inputs
    from pybinding_amd import generators
    q = generators.honeycomb_cell_order(cells[0], cells[1])
    inv = np.argsort(q)
    hp = h[inv][:, inv].tocsr()
    hp.sort_indices()
    ip, ix, dv, _, _, _, _, _ = native.optimize(hp.shape, hp.indptr, hp.indices, hp.data, (-8.9, 8.9), False, [0], [0])
    h2 = sp.csr_matrix((dv, ix, ip), shape=h.shape)
    t0 = time.time()
    r0 = native.random_starters(n, R, "complex128", q.astype(np.int32))
    out["starter_draw_s"] = time.time() - t0
    ctx = abi.Ctx(0)
    mat, blk = ctx.mat(h2, fmt=1), ctx.block(r0)
    del r0
    variant = abi.autotune_diagonal(mat, blk)
    abi.moments_diagonal_run(mat, blk, M)
    ctx.synchronize()
    ctx.timer_start()
    abi.moments_diagonal_run(mat, blk, M)
    ms = ctx.timer_stop()
    mu = abi.moments_diagonal_fetch(mat, blk, M)
    assert np.all(np.isfinite(mu))
    step_ms = ms / (M // 2)
    out.update({"value": nnz * M * R / (ms / 1000) / 1e9, "unit": "GNNZ*moments/s", "ms_per_computation": ms,
                "step_ms": step_ms, "achieved_GBps": step_bytes(n, nnz, 64, 16) / (step_ms / 1000) / 1e9,
                "kernel_variant": variant, "mu0_over_half_n": float(mu[:, 0].real.mean() / (n / 2))})
    blk.close()
    mat.close()
    ctx.close()
    log("model c128: %.1f ms per moment computation, %.4f ms per step, %.0f GB/s algorithmic; %d phase starters "
        "drawn in %.2f s" % (ms, step_ms, out["achieved_GBps"], R, out["starter_draw_s"]))
    # (2) the facade end to end
    t0 = time.perf_counter()
    kpm = pb.kpm(model, energy_range=(-8.9, 8.9), silent=True)
    a = kpm.scaling_factors[0]
    broadening = float(np.float32(np.pi)) * a / (M - 0.5)
    dos = kpm.calc_dos(np.linspace(-8.5, 8.5, 2000), broadening, num_random=R)
    total = time.perf_counter() - t0
    st = kpm.stats
    # the kernel asks for M moments; the count is rounded up so that (M - 2) % 4 == 0
    # (cppcore/include/kpm/Kernel.hpp:9-13)
    assert st.num_moments == M + (-(M - 2)) % 4 and np.all(np.isfinite(dos.data))
    out["facade_calc_dos"] = {"total_s": total, "moments_time_s": st.moments_time, "eps": st.eps,
                              "moments": st.num_moments, "vectors": R, "energies": 2000}
    log("kpm(model).calc_dos end to end: %.2f s (moment computation incl. starters and upload %.2f s)"
        % (total, st.moments_time))
    return out


def cpu_baseline(cells, dtype, nvec, target_seconds=15.0):
    """Oracle restatement of the reference CPU compute on a bounded sample (fewer moments), on the
    reference's own optimized Hamiltonian: BFS reorder + interleaved slices (its defaults)."""
    import oracle
    h2, rmap, borders, _ = build_workload(cells[0], cells[1], 0.5, dtype, "bfs")
    threads = max(1, min(16, os.cpu_count() or 1))
    ell = oracle.to_ell(h2)
    n = h2.shape[0]
    # calibrate on 10 and 106 moments (interleaved needs (M - 2) % 4 == 0): the difference is the
    # per-moment cost, the rest the fixed cost of drawing the starters
    _, t10 = oracle.cpu_baseline_dos(ell, n, 10, nvec, threads, perm=rmap, borders=borders, interleaved=True)
    _, t106 = oracle.cpu_baseline_dos(ell, n, 106, nvec, threads, perm=rmap, borders=borders, interleaved=True)
    per_moment = max((t106 - t10) / 96, t106 / 106 / 4, 1e-9)
    fixed = max(t10 - 10 * per_moment, 0.0)
    m = int((target_seconds - fixed) / per_moment)
    m = max(10, min(4098, m - (m - 2) % 4))
    _, t = oracle.cpu_baseline_dos(ell, n, m, nvec, threads, perm=rmap, borders=borders, interleaved=True)
    value = h2.nnz * m * nvec / t / 1e9
    return {"value": value, "unit": "GNNZ*moments/s", "cores": threads, "kind": "port",
            "sample": "same H2 (2048x1024 periodic honeycomb, BFS order) and the same %d vectors, %d moments "
                      "instead of 4096 (%.1f s); restated reference CPU KPM (not the original binary)"
                      % (nvec, m, t)}


def batched_realizations(steps=2, site_order="hilbert", partition=None, value_table=True):
    """256 square-lattice flakes (one orbital, t = -1), sizes log-uniform in [1e4, 5e5] sites
    (seed 3), Anderson W_f uniform in [0, 4] (seed 4), bounds +-(4 + W_f/2), fp64, 1024 moments,
    4 reference random vectors per flake, all flakes in one block-diagonal recurrence replayed from
    one HIP graph; each flake's sites in the Hilbert order of their positions ("hilbert") or in
    row-major order ("given").  partition: (workgroups, walk_chunk) to use instead of the
    autotuner's choice.  value_table: run on the value-table matrix form (1-byte codes into one
    table of hoppings per flake, dense diagonal) rather than the standard one.  Returns the
    GNNZ*moments/s figure of that run."""
    from pybinding_amd import generators
    from pybinding_amd.batch import kpm_batch
    t0 = time.time()
    rs = np.random.default_rng(3)
    sizes = np.exp(rs.uniform(np.log(1e4), np.log(5e5), 256))
    rw = np.random.default_rng(4)
    Ws = rw.uniform(0, 4, 256)
    hs, ranges, pos = [], [], []
    for sz, W in zip(sizes, Ws):
        side = max(2, int(round(np.sqrt(sz))))
        hs.append(generators.square_flake(side, side, t=-1.0, disorder=W, rng=rw))
        ranges.append((-(4 + W / 2), 4 + W / 2))
        pos.append(generators.square_positions(side, side))
    kb = kpm_batch(hs, ranges, positions=pos if site_order == "hilbert" else None)
    kb.set_random_starters(4)
    kb.set_value_table(value_table)
    n = sum(kb.sizes)
    nnz = sum(h.nnz for h in hs)
    nnz_diag = sum(int(np.count_nonzero(h.indices == np.repeat(np.arange(h.shape[0]), np.diff(h.indptr))))
                   for h in hs)
    log("batched: %d flakes, %d sites, nnz %d (%.1fs setup)" % (len(hs), n, nnz, time.time() - t0))
    M = 1024
    if partition is None:
        wg, chunk = kb.autotune(4)
    else:
        wg, chunk = partition
        kb.set_partition(wg, chunk)
    log("batched: %s partition over %d workgroups (%s)"
        % ("autotuned" if partition is None else "given", wg,
           "XCD-aware walk, runs of %d slices" % chunk if chunk else "contiguous ranges"))
    kb.batch.run(M, True)  # capture + warm-up
    kb.ctx.synchronize()
    kb.ctx.timer_start()
    for _ in range(steps):
        kb.batch.run(M, True)
    ms = kb.ctx.timer_stop() / steps
    mu = kb.batch.fetch(M)
    assert np.all(np.isfinite(mu))
    kb.close()
    value = nnz * M * 4 / (ms / 1000) / 1e9
    # algorithmic bytes per step over the whole batch (fp64, 4 vectors): the matrix (standard form:
    # 4-byte column + 8-byte value per entry; value-table form: 4-byte column + 1-byte code per
    # off-diagonal entry and an 8-byte diagonal per site) + 3 vector blocks of N x 4 x 8 bytes
    if value_table:
        bytes_step = (nnz - nnz_diag) * 5 + n * 8 + 3 * n * 4 * 8
    else:
        bytes_step = nnz * 12 + 4 * (n + 1) + 3 * n * 4 * 8
    gbps = bytes_step * (M // 2) / (ms / 1000) / 1e9
    log("batched: %.1f ms per run (one graph launch), %.1f GNNZ*moments/s, %.0f GB/s algorithmic" % (ms, value, gbps))
    return {"value": value, "unit": "GNNZ*moments/s", "ms_per_run": ms, "achieved_GBps": gbps, "workgroups": wg,
            "walk_chunk": chunk, "site_order": site_order,
            "matrix_format": "value table" if value_table else "standard",
            "workload": "256 square flakes, %d sites, 1024 moments, 4 vectors each, one HIP graph" % n}


def adjoint_cost(steps=2):
    """Strained graphene flake of ~1M sites (161.8 nm square, Gaussian bump h = 1 nm, sigma = 5 nm),
    fp64, 1024 moments, 16 vectors, checkpoint every 64 steps: time of one forward moment
    computation vs one adjoint gradient (forward with checkpoints + recompute + adjoint)."""
    from pybinding_amd import abi, generators
    from pybinding_amd.autograd import TightBinding
    t0 = time.time()
    n, bonds, hop, onsite, pos = generators.strained_graphene_flake(161.8, 161.8, return_positions=True)
    M, R = 1024, 16
    tb = TightBinding(n, bonds, (-9.0, 9.0), M, R, checkpoint=64, positions=(pos[:, 0], pos[:, 1]))
    h2 = tb.scaled(onsite, hop)
    mat = tb.ctx.mat(h2, fmt=1)
    # loss weights: the Jackson-damped DOS integrated against a Gaussian window at 0.5 eV, 0.1 eV wide
    a, b = tb.scale
    xq = np.cos(np.pi * (np.arange(4096) + 0.5) / 4096)
    win = np.exp(-0.5 * ((a * xq + b - 0.5) / 0.1) ** 2)
    g = ((M + 1 - np.arange(M)) * np.cos(np.pi * np.arange(M) / (M + 1)) +
         np.sin(np.pi * np.arange(M) / (M + 1)) / np.tan(np.pi / (M + 1))) / (M + 1)
    w = g * (2 / np.pi) * np.pi / 4096 * (np.cos(np.arange(M)[:, None] * np.arccos(xq)[None, :]) @ win)
    log("adjoint: %d sites, %d bonds (%.1fs setup)" % (n, len(bonds), time.time() - t0))

    def forward_ms():
        abi.moments_diagonal_run(mat, tb.block, M)
        tb.ctx.synchronize()
        tb.ctx.timer_start()
        for _ in range(steps):
            abi.moments_diagonal_run(mat, tb.block, M)
        return tb.ctx.timer_stop() / steps

    fwd_default = forward_ms()
    # the forward step kernel tuned for this matrix and block; the gradient's own forward pass
    # (with its checkpoint copies) runs the same tuned kernel
    variant = abi.autotune_diagonal(mat, tb.block)
    fwd = forward_ms()
    grad, _ = abi.adjoint_diagonal(mat, tb.block, M, w, checkpoint=64)
    tb.ctx.timer_start()
    for _ in range(steps):
        grad, _ = abi.adjoint_diagonal(mat, tb.block, M, w, checkpoint=64)
    adj = tb.ctx.timer_stop() / steps
    abi.adjoint_diagonal(mat, tb.block, M, w, checkpoint=0)  # the 65 GB tape is allocated once, then pooled
    tb.ctx.timer_start()
    for _ in range(steps):
        grad_all, _ = abi.adjoint_diagonal(mat, tb.block, M, w, checkpoint=0)
    adj_all = tb.ctx.timer_stop() / steps
    # the autograd split: the forward pass keeps a tape, the backward pass alone is timed
    tape = abi.Tape(mat, tb.block, M, 64)
    tape.backward(w)
    tb.ctx.timer_start()
    for _ in range(steps):
        tape.backward(w)
    bwd = tb.ctx.timer_stop() / steps
    tape.close()
    assert np.all(np.isfinite(grad))
    diff = float(np.abs(grad - grad_all).max() / np.abs(grad_all).max())
    mat.close()
    tb.close()
    log("adjoint: forward %.1f ms (tuned step variant %d; %.1f ms with the default one), gradient %.1f ms "
        "(checkpoint 64) = %.2fx, %.1f ms storing all = %.2fx, max |g_ck - g_all| / max |g_all| = %.1e"
        % (fwd, variant, fwd_default, adj, adj / fwd, adj_all, adj_all / fwd, diff))
    return {"forward_ms": fwd, "forward_step_variant": variant, "forward_ms_default_variant": fwd_default,
            "gradient_ms_checkpoint64": adj, "ratio_checkpoint64": adj / fwd,
            "gradient_ms_store_all": adj_all, "ratio_store_all": adj_all / fwd,
            "checkpoint64_vs_store_all_max_rel_diff": diff,
            "backward_only_ms_checkpoint64": bwd, "backward_only_ratio_checkpoint64": bwd / fwd,
            "workload": "strained graphene flake, %d sites, 1024 moments, 16 vectors, fp64" % n}


def reconstruction(reps=5):
    """Wall time of the reconstructions through the host core (host prep, device sums, transfers):
    the DOS of the 40 nm flake's 2050 moments at 2000 energies, the LDOS of 64 columns there, and
    KuboBastin of 512 x 512 moments at 1000 energy samples."""
    from pybinding_amd._native import load_module
    native = load_module()
    rng = np.random.default_rng(1)
    a, b = 8.686000303268429, 0.0
    energy = list(np.linspace(-8.5, 8.5, 2000))
    out = {}
    for name, cols in (("dos_2050x2000_ms", 1), ("ldos_64cols_2050x2000_ms", 64)):
        mu = rng.standard_normal((cols, 2050)) / (1.0 + np.arange(2050))
        native.spectral_density(mu, energy, (a, b))
        t0 = time.perf_counter()
        for _ in range(reps):
            native.spectral_density(mu, energy, (a, b))
        out[name] = (time.perf_counter() - t0) / reps * 1e3
    M = 512
    mu = (rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))) / np.outer(np.arange(1, M + 1), np.arange(1, M + 1))
    samples = list(np.linspace(-8.5, 8.5, 1000))
    chem = list(np.linspace(-1, 1, 50))
    native.kubo_bastin(mu, chem, samples, 0.0, (a, b))
    t0 = time.perf_counter()
    for _ in range(reps):
        native.kubo_bastin(mu, chem, samples, 0.0, (a, b))
    out["kubo_bastin_512x512_1000_samples_ms"] = (time.perf_counter() - t0) / reps * 1e3
    log("reconstruction: DOS %.2f ms, LDOS (64 columns) %.2f ms, KuboBastin %.2f ms"
        % (out["dos_2050x2000_ms"], out["ldos_64cols_2050x2000_ms"], out["kubo_bastin_512x512_1000_samples_ms"]))
    return out


def kubo_contraction(n1=724, n2=724, M=512, R=16):
    """Kubo-Bastin mu_nm of the periodic Peierls graphene supercell built by the model builder
    (B = 20 T, W = 1 eV, complex128), 512 x 512 moments, 16 vectors: device time of the FP64,
    FP32-split and bf16x3 contractions, and the split results' errors against the FP64 one."""
    from pybinding_amd import abi
    from pybinding_amd._native import load_module
    import scipy.sparse as sp
    native = load_module()
    t0 = time.time()
    model = periodic_graphene_model(n1, n2, 1.0, 5, field=20.0)
    h = model.hamiltonian
    x = model.system.x
    n = h.shape[0]
    ip, ix, dv, rmap, _, _, _, (a, b) = native.optimize(h.shape, h.indptr, h.indices, h.data, (-9.5, 9.5), True,
                                                        [0], [0])
    h2 = sp.csr_matrix((dv, ix, ip), shape=h.shape)
    # velocity v_ij = (x_i - x_j) H_ij in the optimized order
    perm = np.argsort(rmap)
    xo = x.astype(np.float32)[perm]
    hh = h2.tocoo()
    v = sp.csr_matrix(((xo[hh.row] - xo[hh.col]).astype(np.float64) * hh.data * (a / 2), (hh.row, hh.col)),
                      shape=h.shape)
    r0 = native.random_starters(n, R, "complex128", rmap)
    ctx = abi.Ctx(0)
    mat = ctx.mat(h2, fmt=1)
    vm = ctx.mat(v, fmt=1)
    blk = ctx.block(r0)
    log("kubo: %d sites (%.1fs setup)" % (n, time.time() - t0))
    out = {}
    results = {}
    for mode, name in ((0, "fp64"), (1, "fp32_split"), (2, "bf16x3")):
        ctx.timer_start()
        results[name] = abi.kubo_moments(mat, vm, vm, blk, M, mode=mode)
        out[name + "_s"] = ctx.timer_stop() / 1000
    ref = np.abs(results["fp64"]).max()
    for name in ("fp32_split", "bf16x3"):
        out[name + "_max_rel_error_vs_fp64"] = float(np.abs(results[name] - results["fp64"]).max() / ref)
    flops = 8.0 * M * M * n * R
    out.update({"contraction_flop": flops,
                "workload": "periodic Peierls graphene %dx%d cells through the model builder (%d sites), B=20 T, W=1 eV, "
                            "%dx%d moments, %d vectors"
                            % (n1, n2, n, M, M, R)})
    log("kubo: fp64 %.2f s, fp32-split %.2f s (error %.2e), bf16x3 %.2f s (error %.2e)"
        % (out["fp64_s"], out["fp32_split_s"], out["fp32_split_max_rel_error_vs_fp64"], out["bf16x3_s"],
           out["bf16x3_max_rel_error_vs_fp64"]))
    for hnd in (blk, vm, mat):
        hnd.close()
    ctx.close()
    return out


def ldos_light_cone(reps=3):
    """LDOS at the centre of a 1 M-site graphene flake (161.8 x 161.8 nm, fp64) through
    kpm(model).calc_ldos at 0.05 eV broadening: with the reference's optimal_size light cone (each
    step computes only the BFS slices reached so far) and with full sweeps.  Best of `reps`."""
    import pybinding_amd as pb
    from pybinding_amd.repository import graphene
    model = pb.Model(graphene.monolayer(), pb.rectangle(161.8, 161.8), pb.force_double_precision())
    n = model.hamiltonian.shape[0]
    energy = np.linspace(-1, 1, 200)
    out = {"sites": n, "broadening": 0.05}
    for opt in (True, False):
        kpm = pb.kpm(model, energy_range=(-9, 9), silent=True, optimal_size=opt, interleaved=False)
        kpm.calc_ldos(energy, broadening=0.05, position=[0, 0])   # BFS reorder and upload
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            kpm.calc_ldos(energy, broadening=0.05, position=[0, 0])
            best = min(best, time.perf_counter() - t0)
        st = kpm.stats
        out["cone_ms" if opt else "full_ms"] = best * 1000
        out["cone_fraction_of_rows" if opt else "moments"] = (st.opt_vec / st.vec) if opt else st.num_moments
    out["speedup"] = out["full_ms"] / out["cone_ms"]
    log("LDOS light cone: %d sites, %d moments: %.1f ms with the cone (%.2f of the rows), %.1f ms full sweeps"
        % (n, out["moments"], out["cone_ms"], out["cone_fraction_of_rows"], out["full_ms"]))
    return out


def plain_dos_flake(M=2050, R=8, reps=5):
    """The reference's plain DOS case: graphene 40 x 40 nm flake (the model builder, reference site
    order), fp64, energy range (-8.6, 8.6), 2050 moments (2048 requested, rounded as the
    reference does), 8 reference random vectors, BFS order.  Device time of the moments through
    the C-ABI (inputs resident), the restated reference CPU compute (interleaved, all threads) on the
    same H2 and vectors, and the largest difference of the two mean moment sets."""
    import oracle
    from pybinding_amd import abi, lattice as L
    from pybinding_amd._native import load_module
    from pybinding_amd.model import Model
    from pybinding_amd.repository import graphene
    import scipy.sparse as sp
    native = load_module()
    m = Model(graphene.monolayer(), L.rectangle(40, 40), L.force_double_precision())
    h = m.hamiltonian
    ip, ix, dv, rmap, borders, _, _, _ = native.optimize(h.shape, h.indptr, h.indices, h.data, (-8.6, 8.6), True,
                                                         [0], [0])
    h2 = sp.csr_matrix((dv, ix, ip), shape=h.shape)
    n = h.shape[0]
    r0 = native.random_starters(n, R, "float64", rmap)
    ctx = abi.Ctx(0)
    mat = ctx.mat(h2, fmt=1)
    blk = ctx.block(r0)
    variant = abi.autotune_diagonal(mat, blk)
    timings = {}
    for policy, name in ((1, "graph"), (0, "eager")):
        ctx.set_graph_policy(policy)
        abi.moments_diagonal_run(mat, blk, M)  # the automatic policy captures on the second run
        abi.moments_diagonal_run(mat, blk, M)
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(reps):
            abi.moments_diagonal_run(mat, blk, M)
        timings[name] = ctx.timer_stop() / reps
    gpu_ms = timings["eager"]
    mu_gpu = abi.moments_diagonal_fetch(mat, blk, M).mean(axis=0)
    blk.close()
    mat.close()
    ctx.close()
    # the drop-in end to end: pb.kpm(model).calc_dos at 2000 energies with the broadening that gives
    # 2050 moments (Jackson), 8 random vectors -- BFS reorder, starters, moments, damping and
    # reconstruction through the facade, the model already built
    from pybinding_amd import chebyshev
    kp = chebyshev.kpm(m, energy_range=(-8.6, 8.6), kernel=chebyshev.jackson_kernel(), silent=True)
    broadening = 8.686000303268429 * np.pi / 2048.5
    energy = np.linspace(-8.5, 8.5, 2000)
    kp.calc_dos(energy, broadening, num_random=R)
    t0 = time.perf_counter()
    kp.calc_dos(energy, broadening, num_random=R)
    facade_ms = (time.perf_counter() - t0) * 1e3
    assert kp.stats.num_moments == M, kp.stats.num_moments
    threads = max(1, min(16, os.cpu_count() or 1))
    mu_cpu, t_cpu = oracle.cpu_baseline_dos(oracle.to_ell(h2), n, M, R, threads, perm=rmap, borders=borders,
                                            interleaved=True)
    err = float(np.abs(mu_gpu - mu_cpu).max() / abs(mu_cpu[0]))
    out = {"sites": n, "nnz": int(h2.nnz), "moments": M, "vectors": R,
           "gpu_ms": gpu_ms, "gpu_value": h2.nnz * M * R / (gpu_ms / 1000) / 1e9, "gpu_us_per_step": gpu_ms * 1000 / (M // 2),
           "gpu_ms_step_launches": timings["eager"], "gpu_ms_graph_replay": timings["graph"],
           "facade_calc_dos_ms": facade_ms, "kernel_variant": variant,
           "cpu_ms": t_cpu * 1000, "cpu_value": h2.nnz * M * R / t_cpu / 1e9, "cpu_threads": threads,
           "max_rel_diff_vs_cpu": err, "unit": "GNNZ*moments/s",
           "workload": "graphene 40x40 nm flake (%d sites), fp64, 2050 moments, 8 vectors" % n}
    log("plain DOS flake: GPU %.2f ms (%.1f us/step), CPU %.1f ms on %d threads, max diff %.2e; "
        "kpm(model).calc_dos end to end %.1f ms" % (gpu_ms, out["gpu_us_per_step"], t_cpu * 1000, threads, err, facade_ms))
    return out


def fp32_variant(cells, M, R, fmt, order):
    """The same block-SpMM workload in fp32 (the reference's default precision for models without
    force_double_precision): one timed moment computation after autotuning."""
    from pybinding_amd import abi
    dtype = np.dtype(np.float32)
    h2, rmap, borders, native = build_workload(cells[0], cells[1], 0.5, dtype, order)
    n, nnz = h2.shape[0], h2.nnz
    r0 = starters(native, n, R, dtype, rmap)
    ctx = abi.Ctx(0)
    ctx_tw = ctx.tile_width(dtype)
    mat = ctx.mat(h2, fmt=1 if fmt == "SELL" else 0)
    blk = ctx.block(r0)
    ctx.synchronize()
    variant = abi.autotune_diagonal(mat, blk)
    abi.moments_diagonal_run(mat, blk, M)
    ctx.synchronize()
    ctx.timer_start()
    abi.moments_diagonal_run(mat, blk, M)
    ms = ctx.timer_stop()
    mu = abi.moments_diagonal_fetch(mat, blk, M)
    assert np.all(np.isfinite(mu)), "non-finite moments"
    blk.close()
    mat.close()
    ctx.close()
    widths = launch_widths(R, ctx_tw)
    step_ms = ms / (M // 2)
    out = {"value": nnz * M * R / (ms / 1000) / 1e9, "unit": "GNNZ*moments/s", "ms_per_computation": ms,
           "step_ms": step_ms, "launch_widths": widths,
           "achieved_GBps": sum(step_bytes(n, nnz, w, 4) for w in widths) / (step_ms / 1000) / 1e9,
           "kernel_variant": variant}
    log("fp32 variant: %.1f ms per moment computation, %.4f ms per step, %.0f GB/s algorithmic"
        % (ms, step_ms, out["achieved_GBps"]))
    return out


def counted_traffic(key):
    """HBM bytes per step launch measured by rocprofv3 counters (FETCH_SIZE and WRITE_SIZE, each
    scaled by its copy-kernel calibration factor) for this exact workload, read from
    profiles/hbm_traffic.json (the autotuner's tile-shape variants move the same bytes); None when
    that workload was not counted."""
    try:
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")) as f:
            return json.load(f)[key]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--moments", type=int, default=4096)
    ap.add_argument("--vectors", type=int, default=64)
    ap.add_argument("--cells", type=int, nargs=2, default=(2048, 1024))
    ap.add_argument("--dtype", choices=["float64", "float32"], default="float64")
    ap.add_argument("--format", choices=["CSR", "SELL"], default="SELL")
    ap.add_argument("--order", choices=["bfs", "hilbert"], default="hilbert",
                    help="site order of H2: the reference's BFS reorder, or cells along a Hilbert curve")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the single-vector (plain) comparison")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-realizations run")
    ap.add_argument("--no-flake", action="store_true", help="skip the plain DOS run on the 40 nm graphene flake")
    ap.add_argument("--no-fp32", action="store_true", help="skip the fp32 variant of the block-SpMM workload")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient cost run")
    ap.add_argument("--no-kubo", action="store_true", help="skip the Kubo contraction run")
    ap.add_argument("--no-ldos", action="store_true", help="skip the LDOS light-cone run")
    ap.add_argument("--no-model", action="store_true",
                    help="skip the complex128 run of the workload built by the model builder")
    ap.add_argument("--two-step", default="auto",
                    help="'auto' (the autotuner decides), 'off' (single steps), or 'W:C' to pin the LDS two-step "
                         "form with column tiles of W vectors on C workgroups per CU (e.g. 8:4); used to profile "
                         "one configuration")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 2 GiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()

    if args.copy_probe:
        import torch
        a = torch.ones(2**28, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        from pybinding_amd import abi
        pctx = abi.Ctx(0)
        cg, rg, tg = pctx.stream_probe(1 << 31)
        pctx.close()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copy_probe_bytes_written": b.numel() * 8,
                          "copy_ms": ms, "copy_GBps": 2 * a.numel() * 8 / (ms / 1000) / 1e9,
                          "stream_probe_copy_GBps": cg, "stream_probe_read_GBps": rg,
                          "stream_probe_triad_GBps": tg}))
        return

    if args.gpus != 1:
        print("bench.py: runs on one GPU only; a multi-GPU run (the random vectors split across GPUs with "
              "one moment all-reduce at the end) is not implemented.", file=sys.stderr)
        sys.exit(2)

    from pybinding_amd import abi
    dtype = np.dtype(args.dtype)
    s = dtype.itemsize
    h2, rmap, borders, native = build_workload(args.cells[0], args.cells[1], 0.5, dtype, args.order)
    n, nnz = h2.shape[0], h2.nnz
    M, R = args.moments, args.vectors
    t_draw = time.time()
    r0 = starters(native, n, R, dtype, rmap)
    t_draw = time.time() - t_draw

    ctx = abi.Ctx(0)
    mat = ctx.mat(h2, fmt=1 if args.format == "SELL" else 0)
    blk = ctx.block(r0)
    ctx.synchronize()
    if args.two_step != "auto":
        ctx.set_autotune_two_step(False)
    variant = abi.autotune_diagonal(mat, blk)
    if args.two_step not in ("auto", "off"):
        w, c = (int(x) for x in args.two_step.split(":"))
        ctx.set_tile_width(dtype, w)
        ctx.set_two_step_lds(c)
    tile_width = ctx.tile_width(dtype)
    two_step, two_step_wg = ctx.two_step_lds(dtype)
    log("autotuned step kernel variant: %d (column tiles of %d vectors)%s"
        % (variant, tile_width, ", LDS two-step form on %d workgroups per CU" % two_step_wg if two_step else ""))

    for _ in range(args.warmup):
        abi.moments_diagonal_run(mat, blk, M)
    ctx.synchronize()
    log("warmup done")

    t0 = time.perf_counter()
    ctx.timer_start()
    for _ in range(args.steps):
        abi.moments_diagonal_run(mat, blk, M)
    gpu_ms = ctx.timer_stop()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    mu = abi.moments_diagonal_fetch(mat, blk, M)
    assert np.all(np.isfinite(mu)), "non-finite moments"
    ms_per_step = wall * 1000 / args.steps
    value = nnz * M * R / (wall / args.steps) / 1e9

    # fused step launches per moment computation: one per two moments and column tile (one per
    # four moments in the two-step form, each reading r_k-1 and r_k and writing r_k+1 and r_k+2)
    widths = launch_widths(R, tile_width)
    steps_per_run = M // 2
    steps_per_launch = 2 if two_step else 1
    step_ms = gpu_ms / (args.steps * steps_per_run // steps_per_launch * len(widths))      # one launch
    # the variants 45..59 and the two-step form read the matrix's value-table form (8-bit codes of
    # the off-diagonal entries, the diagonal dense) when it has one: its bytes replace
    # nnz (s + 4) + 4 (N + 1) in a launch's count
    coded = mat.coded_bytes() if (variant >= 45 or two_step) else 0
    matrix_bytes = coded if coded else nnz * (s + 4) + 4 * (n + 1)
    # algorithmic bytes of the steps one launch does (a two-step launch moves fewer: it reads and
    # writes 4 vector blocks instead of 6, which the counted traffic shows)
    bytes_per_launch = sum(steps_per_launch * (matrix_bytes + 3 * n * w * s) for w in widths) / len(widths)
    achieved = bytes_per_launch / (step_ms / 1000)
    # the same step over all R vectors as one pass (the matrix read once): the algorithmic bytes of
    # the whole step over the time of its launches
    pass_ms = gpu_ms / (args.steps * steps_per_run)
    vw_all = min(64, 1 << (max(R, 1) - 1).bit_length())
    pass_gbps = step_bytes(n, nnz, vw_all, s) * ((R + 63) // 64) / (pass_ms / 1000) / 1e9
    log("GPU: %.1f ms per moment computation, %d launches of %s vectors per %d step(s), %.4f ms per launch, "
        "%.1f GB/s algorithmic per launch; %.4f ms per step over all %d vectors = %.1f GB/s of one-pass bytes"
        % (gpu_ms / args.steps, len(widths), "/".join(str(w) for w in sorted(set(widths))), steps_per_launch, step_ms,
           achieved / 1e9, pass_ms, R, pass_gbps))

    traffic_key = "%dx%d %s v%d launch%d %s%s %s%s" % (args.cells[0], args.cells[1], dtype.name, R, max(widths),
                                                       args.format, " coded" if coded else "", args.order,
                                                       " two-step" if two_step else "")
    extra = {"starter_draw_s": t_draw,
             "one_pass_equivalent": {"ms_per_step": pass_ms, "achieved_GBps": pass_gbps,
                                     "frac": pass_gbps * 1e9 / HBM_PEAK, "launch_widths": widths}}
    cg, rg, tg = ctx.stream_probe(1 << 31)
    step_traffic_gbps = (counted_traffic(traffic_key) or float("nan")) / (step_ms / 1000) / 1e9
    extra["stream_probe"] = {"copy_GBps": cg, "read_GBps": rg, "triad_GBps": tg,
                             "step_traffic_over_copy": step_traffic_gbps / cg,
                             "step_traffic_over_triad": step_traffic_gbps / tg}
    log("stream probe: copy %.0f GB/s, read %.0f GB/s, triad (b = a + b) %.0f GB/s; step traffic at %.2f of the triad"
        % (cg, rg, tg, step_traffic_gbps / tg))
    if two_step:
        # the same computation with single steps (the default column tiles, the autotuned single-step
        # variant), timed in the same run for comparison
        ctx.set_two_step(dtype, 0)
        ctx.set_tile_width(dtype, 16 if dtype == np.float64 else 64)
        abi.moments_diagonal_run(mat, blk, M)
        ctx.synchronize()
        ctx.timer_start()
        abi.moments_diagonal_run(mat, blk, M)
        ms_single = ctx.timer_stop()
        extra["single_steps"] = {"ms_per_step": ms_single / steps_per_run, "value": nnz * M * R / (ms_single / 1000) / 1e9,
                                 "two_step_speedup": ms_single / (gpu_ms / args.steps)}
        log("single steps on the same box: %.4f ms per step (the two-step form %.2fx faster)"
            % (ms_single / steps_per_run, ms_single / (gpu_ms / args.steps)))
    if not args.no_plain:
        # the plain single-vector path (one vector, SELL-64 kernel) on the same matrix, timed over
        # 512 moments and reported in the same metric
        mat1 = ctx.mat(h2, fmt=1)
        blk1 = ctx.block(r0[:1])
        abi.moments_diagonal_run(mat1, blk1, 512)
        ctx.synchronize()
        ctx.timer_start()
        abi.moments_diagonal_run(mat1, blk1, 512)
        ms1 = ctx.timer_stop()
        v1 = nnz * 512 * 1 / (ms1 / 1000) / 1e9
        b1 = step_bytes(n, nnz, 1, s)
        extra["plain_single_vector"] = {"value": v1, "unit": "GNNZ*moments/s",
                                        "step_ms": ms1 / 256, "achieved_GBps": b1 / (ms1 / 256 / 1000) / 1e9,
                                        "block_over_plain": value / v1}
        log("plain: %.1f GNNZ*moments/s; block/plain = %.2f" % (v1, value / v1))
        blk1.close()
        mat1.close()

    if not args.no_fp32 and dtype == np.float64:
        extra["fp32_variant"] = fp32_variant(args.cells, M, R, args.format, args.order)
    if not args.no_flake:
        extra["plain_dos_flake"] = plain_dos_flake()
    if not args.no_ldos:
        extra["ldos_light_cone"] = ldos_light_cone()
    if not args.no_model and dtype == np.float64:
        extra["model_complex128"] = model_leg(args.cells, M, R)
    if not args.no_batched:
        extra["batched_realizations"] = batched_realizations()
    if not args.no_adjoint:
        extra["adjoint_gradient"] = adjoint_cost()
    if not args.no_kubo:
        extra["kubo_contraction"] = kubo_contraction()
        extra["reconstruction"] = reconstruction()

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args.cells, dtype, R)
            log("cpu baseline: %.2f GNNZ*moments/s on %d threads" % (cpu["value"], cpu["cores"]))
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline failed: %s" % e)

    blk.close()
    mat.close()
    ctx.close()

    line = {
        "metric": "KPM GNNZ\u00b7moments/s and % HBM roofline; batched/adjoint vs plain GPU",
        "value": value,
        "unit": "GNNZ\u00b7moments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if dtype == np.float64 else "f32",
        "data": "synthetic (seeded disorder, reference mt19937 starters)",
        "config": {"workload": "periodic honeycomb %dx%d cells (N=%d, nnz=%d), Anderson W=1 eV, %d moments, "
                               "%d random vectors as one starter block (run in column tiles of %d)"
                               % (args.cells[0], args.cells[1], n, nnz, M, R, tile_width),
                   "matrix_format": args.format + (" (value-table form, %d bytes)" % coded if coded else ""),
                   "site_order": args.order, "kernel_variant": variant,
                   "moments": M, "vectors": R, "column_tile": tile_width,
                   "steps_per_launch": steps_per_launch,
                   "two_step": {"form": "LDS", "workgroups_per_cu": two_step_wg} if two_step else None},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": counted_traffic(traffic_key)},
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
