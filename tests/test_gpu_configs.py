"""GPU parity on two benchmark workloads at their own sizes:
- the 40 x 40 nm graphene flake (61,096 sites, fp64, 2050 moments, 8 reference random vectors,
  Jackson kernel, energy_range (-8.6, 8.6)), built with the model builder: moments through the
  C-ABI (CSR and SELL) against the oracle, and `kpm(model).calc_dos` at 2000 energies against
  the oracle's reconstruction;
- the Kubo-Bastin contraction at 512 x 512 moments: at reduced N against an FP64 numpy GEMM of
  oracle panels; through the periodic Peierls model (`Model -> kpm().calc_conductivity`) against
  the oracle's conductivity; at the full 1.05 M-site size the FP32-split and bf16x3 contractions
  against the FP64 one;
- the 64-bit gather path (blocks of 4 GiB and more) against the 32-bit one on the same columns.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import abi
import pybinding_amd as pb
from pybinding_amd import generators
from pybinding_amd.repository import graphene

pytestmark = pytest.mark.gpu

M1, R1 = 2050, 8


@pytest.fixture(scope="module")
def flake40(oracle):
    model = pb.Model(graphene.monolayer(), pb.rectangle(40, 40), pb.force_double_precision())
    h = model.hamiltonian
    a, b = oracle.scale_factors(-8.6, 8.6)
    h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
    r0 = oracle.random_starter(h.shape[0], R1, np.float64)
    ref = oracle.diag_moments(h2, r0, M1)
    return model, h2, r0, ref, (a, b)


@pytest.mark.parametrize("fmt", [0, 1])
def test_flake40_moments_through_abi(flake40, fmt):
    _, h2, r0, ref, (a, b) = flake40
    assert h2.shape[0] == 61096 and h2.nnz == 182590 and h2.dtype == np.float64
    assert (a, b) == (8.686000303268429, 0.0)
    print("40 nm flake: %d sites, %d moments, %d vectors, format %d" % (h2.shape[0], M1, R1, fmt), flush=True)
    ctx = abi.Ctx(0)
    try:
        mat, blk = ctx.mat(h2, fmt=fmt), ctx.block(r0)
        mu = abi.moments_diagonal(mat, blk, M1)
        blk.close()
        mat.close()
    finally:
        ctx.close()
    assert np.max(np.abs(mu - ref) / np.abs(ref[:, :1])) <= 1e-10


def test_flake40_calc_dos(flake40, oracle):
    model, _, _, ref, (a, b) = flake40
    # the broadening that asks the Jackson kernel for exactly 2050 moments
    broadening = float(np.float32(np.pi)) * a / 2049.5
    assert oracle.required_num_moments("jackson", broadening / a) == M1
    kpm = pb.kpm(model, energy_range=(-8.6, 8.6), silent=True)
    assert kpm.scaling_factors == pytest.approx((a, b), rel=1e-15)
    energy = np.linspace(-8.5, 8.5, 2000)
    dos = kpm.calc_dos(energy, broadening, num_random=R1)
    assert kpm.stats.num_moments == M1
    mu = ref.mean(axis=0) * oracle.jackson(M1)
    expect = oracle.spectral_density(mu, energy, a, b)
    # normwise: the moments agree to ~1e-13 mu_0, which the 2050-term sums carry into ~1e-12 of the
    # largest DOS value (near the band edges the DOS is 1e-6 of its peak)
    np.testing.assert_allclose(dos.data, expect, rtol=0, atol=1e-9 * np.abs(expect).max())
    mu_gpu = kpm.dos_moments(M1, R1)
    assert np.max(np.abs(mu_gpu - ref.mean(axis=0))) <= 1e-10 * abs(ref[0, 0])


def kubo_reference(oracle, h2, vl, vr, r0, M):
    ref = np.zeros((M, M), dtype=np.complex128)
    for r in r0:
        left = oracle.offdiag_moments(h2, vl @ r, M, "dense")
        right = oracle.offdiag_moments(h2, r, M, "dense", op=vr)
        ref += left.astype(np.complex128) @ right.astype(np.complex128).conj().T
    return ref / len(r0)


def test_kubo_512_reduced_n_vs_fp64_gemm(oracle):
    """Peierls graphene (20 T, W = 1 eV) of 1152 sites, 512 x 512 moments, 2 vectors: every
    contraction mode against an FP64 GEMM of the oracle's panels."""
    h, x, y = generators.peierls_honeycomb_supercell(24, 24, 20.0, disorder=1.0)
    a, b = oracle.scale_factors(-9.5, 9.5)
    h2 = oracle.scale_hamiltonian(h, a, b, np.complex128)
    v = oracle.velocity(h2, x.astype(np.float32))
    r0 = oracle.random_starter(h.shape[0], 2, np.complex128)
    M = 512
    ref = kubo_reference(oracle, h2, v, v, r0, M)
    ctx = abi.Ctx(0)
    try:
        mat, vm, blk = ctx.mat(h2, fmt=1), ctx.mat(v, fmt=1), ctx.block(r0)
        mus = {mode: abi.kubo_moments(mat, vm, vm, blk, M, mode=mode) for mode in (0, 1, 2)}
        for hnd in (blk, vm, mat):
            hnd.close()
    finally:
        ctx.close()
    scale = np.abs(ref).max()
    err = {mode: np.abs(mu - ref).max() / scale for mode, mu in mus.items()}
    print("kubo 512 x 512, 1152 sites: %s" % err, flush=True)
    assert err[0] <= 1e-11 and err[1] <= 1e-5 and err[2] <= 1e-5


def peierls_model(n1, n2, field=20.0, disorder=1.0, seed=5):
    rng = np.random.default_rng(seed)

    @pb.onsite_energy_modifier(is_double=True)
    def anderson(energy):
        return energy + rng.uniform(-disorder / 2, disorder / 2, size=energy.shape)

    return pb.Model(graphene.monolayer(), pb.primitive(n1, n2),
                    pb.translational_symmetry(a1=n1 * graphene.a, a2=n2 * graphene.a),
                    graphene.constant_magnetic_field(field), anderson)


def test_conductivity_periodic_peierls_model(oracle):
    """The Kubo workload's model built by the builder (periodic, complex128 by promotion, Peierls
    phases, Anderson disorder), through `kpm(model).calc_conductivity` with 2 random vectors,
    against the oracle's moments (the two starters of one vector share the reference stream,
    cppcore/src/kpm/Core.cpp:132-145) and its KuboBastin reconstruction."""
    model = peierls_model(16, 16)
    h = model.hamiltonian
    assert h.dtype == np.complex128 and h.nnz == 4 * h.shape[0]
    kpm = pb.kpm(model, energy_range=(-9.5, 9.5), kernel=pb.lorentz_kernel(), silent=True)
    mu_e = np.linspace(-2, 2, 21)
    sigma = kpm.calc_conductivity(mu_e, broadening=0.3, temperature=0, num_random=2, num_points=200)
    a, b = oracle.scale_factors(-9.5, 9.5)
    M = oracle.required_num_moments("lorentz", 0.3 / a)
    h2 = oracle.scale_hamiltonian(h, a, b, np.complex128)
    v = oracle.velocity(h, model.system.x)
    mu = kubo_reference(oracle, h2, v, v, oracle.random_starter(h.shape[0], 2, np.complex128), M)
    g = oracle.lorentz(M)
    mu = mu * g[:, None] * g[None, :]
    samples = np.linspace(float(np.float32(-9.5)), float(np.float32(9.5)), 200)
    ref = oracle.kubo_bastin(mu, mu_e, samples, 0.0, a, b).real
    np.testing.assert_allclose(sigma.data, ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())


def test_kubo_512_full_size_precisions():
    """The Kubo workload at full size: periodic Peierls graphene (B = 20 T, W = 1 eV, complex128)
    of 1.05 M sites, 16 vectors, 512 x 512 moments (the persistent bf16x3 grid runs 32 output
    tiles per K range): the FP32-split and bf16x3 contractions against the FP64 one."""
    from pybinding_amd._native import load_module
    native = load_module()
    h, x, _ = generators.peierls_honeycomb_supercell(724, 724, 20.0, disorder=1.0)
    ip, ix, dv, rmap, _, _, _, (a, b) = native.optimize(h.shape, h.indptr, h.indices, h.data, (-9.5, 9.5), True,
                                                        [0], [0])
    h2 = sp.csr_matrix((dv, ix, ip), shape=h.shape)
    xo = x.astype(np.float32)[np.argsort(rmap)]
    hh = h2.tocoo()
    v = sp.csr_matrix(((xo[hh.row] - xo[hh.col]).astype(np.float64) * hh.data * (a / 2), (hh.row, hh.col)),
                      shape=h.shape)
    r0 = native.random_starters(h.shape[0], 16, "complex128", rmap)
    print("kubo full size: %d sites, 512 x 512 moments, 16 vectors" % h.shape[0], flush=True)
    ctx = abi.Ctx(0)
    try:
        mat, vm, blk = ctx.mat(h2, fmt=1), ctx.mat(v, fmt=1), ctx.block(r0)
        mus = {mode: abi.kubo_moments(mat, vm, vm, blk, 512, mode=mode) for mode in (0, 1, 2)}
        for hnd in (blk, vm, mat):
            hnd.close()
    finally:
        ctx.close()
    scale = np.abs(mus[0]).max()
    assert np.all(np.isfinite(mus[0])) and scale > 0
    assert np.abs(mus[1] - mus[0]).max() <= 1e-6 * scale
    assert np.abs(mus[2] - mus[0]).max() <= 1e-6 * scale


def test_wide_gather_path_complex128():
    """64 complex128 vectors on the 4.19 M-site periodic honeycomb: the block is 4 GiB, so the
    step kernel takes its 64-bit gather offsets; 8 of its columns run again as an 8-wide block
    (32-bit offsets) and the first column against the oracle."""
    from pybinding_amd._native import load_module
    import oracle
    native = load_module()
    h = generators.honeycomb_supercell(2048, 1024, disorder=1.0, seed=2, dtype=np.float64).astype(np.complex128)
    q = generators.honeycomb_cell_order(2048, 1024)
    inv = np.argsort(q)
    hp = h[inv][:, inv].tocsr()
    hp.sort_indices()
    ip, ix, dv, _, _, _, _, _ = native.optimize(hp.shape, hp.indptr, hp.indices, hp.data, (-8.9, 8.9), False,
                                                [0], [0])
    h2 = sp.csr_matrix((dv, ix, ip), shape=h.shape)
    n, M = h2.shape[0], 32
    assert n * 64 * 16 >= 1 << 32
    r0 = native.random_starters(n, 64, "complex128", q.astype(np.int32))
    print("wide gathers: %d sites, 64 complex128 vectors (%.2f GiB per block)" % (n, n * 64 * 16 / 2**30), flush=True)
    ctx = abi.Ctx(0)
    try:
        mat = ctx.mat(h2, fmt=1)
        blk = ctx.block(r0)
        mu = abi.moments_diagonal(mat, blk, M)
        blk.close()
        sub = ctx.block(np.ascontiguousarray(r0[[0, 9, 17, 30, 41, 50, 58, 63]]))
        mu8 = abi.moments_diagonal(mat, sub, M)
        sub.close()
        mat.close()
    finally:
        ctx.close()
    assert np.max(np.abs(mu[:, 0] - n / 2)) <= 1e-12 * n
    assert np.max(np.abs(mu8 - mu[[0, 9, 17, 30, 41, 50, 58, 63]]) / np.abs(mu8[:, :1])) <= 1e-12
    ref = oracle.diag_moments(h2, r0[:1], M)
    assert np.max(np.abs(mu[:1] - ref) / np.abs(ref[:, :1])) <= 1e-10
