"""GPU parity of the batched independent-systems engine (kpmx_batch_*) against the oracle, each
system checked separately with its own scale and its own default-seeded starters."""
import numpy as np
import pytest

from pybinding_amd import generators
from pybinding_amd.batch import kpm_batch

pytestmark = pytest.mark.gpu


def systems(dtype=np.float64, count=7, seed=3):
    rng = np.random.default_rng(seed)
    hs, ranges = [], []
    for i in range(count):
        l1, l2 = int(rng.integers(5, 60)), int(rng.integers(3, 50))
        W = float(rng.uniform(0, 4))
        hs.append(generators.square_flake(l1, l2, t=-1.0, disorder=W, rng=rng, dtype=dtype))
        ranges.append((-(4 + W / 2), 4 + W / 2))
    return hs, ranges


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-4)])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("nvec", [4, 3, 8])
def test_batched_moments_match_oracle(oracle, dtype, tol, use_graph, nvec):
    hs, ranges = systems(dtype)
    kb = kpm_batch(hs, ranges)
    M = 130
    mu = kb.moments(M, nvec, use_graph=use_graph)
    mu2 = kb.moments(M, nvec, use_graph=use_graph)  # a replay must give the same result
    kb.close()
    np.testing.assert_array_equal(mu, mu2)
    for s, (h, (lo, hi)) in enumerate(zip(hs, ranges)):
        a, b = oracle.scale_factors(lo, hi)
        h2 = oracle.scale_hamiltonian(h, a, b, dtype)
        r0 = oracle.random_starter(h.shape[0], nvec, dtype)
        ref = oracle.diag_moments(h2, r0, M)
        err = np.max(np.abs(mu[s] - ref) / np.abs(ref[:, :1]))
        assert err <= tol, (s, err)


def test_batched_rejects_coupled_systems():
    import scipy.sparse as sp
    from pybinding_amd import abi
    ctx = abi.Ctx(0)
    h = sp.random(10, 10, density=0.5, random_state=1, format="csr") + sp.eye(10)
    with pytest.raises(RuntimeError, match="block diagonal"):
        b = abi.Batch.__new__(abi.Batch)
        abi.Batch.__init__(b, ctx, [h])
        big_rows = np.array([0, 5, 10], dtype=np.int64)
        import ctypes
        handle = ctypes.c_void_p()
        abi._check(abi.lib().kpmx_batch_upload(ctx.h, ctypes.c_int(1), ctypes.c_int64(2), abi._p(big_rows),
                                               ctypes.c_int64(len(b.data)), abi._p(b.indptr), abi._p(b.indices),
                                               abi._p(b.data), ctypes.byref(handle)))
    ctx.close()


def test_batched_autotune_keeps_results(oracle):
    """The autotuner's partitions (contiguous slice ranges over 256 ... 65536 workgroups, the
    XCD-aware walk) change only how each system's sums are grouped: every system still matches the
    oracle, and the tuned run agrees with the default partition to rounding."""
    hs, ranges = systems(np.float64, count=40, seed=8)
    kb = kpm_batch(hs, ranges)
    M = 66
    before = kb.moments(M, 4)
    wg, chunk = kb.autotune(4)
    after = kb.moments(M, 4)
    again = kb.moments(M, 4, use_graph=False)
    kb.close()
    assert wg >= 1 and chunk >= 0 and (chunk == 0 or wg % 8 == 0)
    np.testing.assert_array_equal(after, again)
    assert np.max(np.abs(after - before) / np.abs(before[:, :, :1])) <= 1e-13
    for s, (h, (lo, hi)) in enumerate(zip(hs, ranges)):
        a, b = oracle.scale_factors(lo, hi)
        h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
        ref = oracle.diag_moments(h2, oracle.random_starter(h.shape[0], 4, np.float64), M)
        assert np.max(np.abs(after[s] - ref) / np.abs(ref[:, :1])) <= 1e-10, s


@pytest.mark.parametrize("nvec", [4, 8])
def test_batched_partitions_agree(oracle, nvec):
    """Every slice partition -- contiguous ranges, and the XCD-aware walk with runs of 1, 2 and 5
    slices over grids with fewer, about as many and more workgroups than slices per region --
    visits every slice once: each system matches the oracle and the partitions agree to rounding
    (the grouping of each system's sums is all that changes)."""
    hs, ranges = systems(np.float64, count=30, seed=11)
    kb = kpm_batch(hs, ranges)
    nslices = sum((h.shape[0] + 63) // 64 for h in hs)
    M = 40
    res = {}
    try:
        for wg, chunk in [(1, 0), (7, 0), (nslices, 0), (8, 1), (16, 2), (64, 1), (64, 5),
                          (8 * (nslices // 8 + 3), 1), (256, 2)]:
            kb.set_partition(wg, chunk)
            res[wg, chunk] = kb.moments(M, nvec, use_graph=(chunk % 2 == 1))
    finally:
        kb.close()
    base = res[1, 0]
    for key, mu in res.items():
        assert np.max(np.abs(mu - base) / np.abs(base[:, :, :1])) <= 1e-13, key
    for s, (h, (lo, hi)) in enumerate(zip(hs, ranges)):
        a, b = oracle.scale_factors(lo, hi)
        h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
        ref = oracle.diag_moments(h2, oracle.random_starter(h.shape[0], nvec, np.float64), M)
        assert np.max(np.abs(res[64, 1][s] - ref) / np.abs(ref[:, :1])) <= 1e-10, s


def test_batched_partition_rejects_bad_walk():
    hs, ranges = systems(np.float64, count=3, seed=2)
    kb = kpm_batch(hs, ranges)
    try:
        with pytest.raises(RuntimeError, match="multiple of 8"):
            kb.set_partition(12, 1)
        with pytest.raises(RuntimeError, match="invalid partition"):
            kb.set_partition(0, 0)
    finally:
        kb.close()


def test_batched_hilbert_order_matches_oracle(oracle):
    """Systems with site positions run in the Hilbert order of their sites (starters drawn in the
    original order and moved along): every system still matches the oracle on its unpermuted
    Hamiltonian, and Models given directly are ordered the same way."""
    rng = np.random.default_rng(21)
    hs, ranges, pos = [], [], []
    for _ in range(9):
        l1, l2 = int(rng.integers(9, 70)), int(rng.integers(5, 40))
        W = float(rng.uniform(0, 3))
        hs.append(generators.square_flake(l1, l2, t=-1.0, disorder=W, rng=rng))
        ranges.append((-(4 + W / 2), 4 + W / 2))
        pos.append(generators.square_positions(l1, l2))
    M = 98
    kb = kpm_batch(hs, ranges, positions=pos)
    try:
        assert all(q is not None for q in kb.orders)
        assert any(np.any(q != np.arange(q.size)) for q in kb.orders)
        mu = kb.moments(M, 4)
    finally:
        kb.close()
    given = kpm_batch(hs, ranges, site_order=None)
    try:
        mu_given = given.moments(M, 4)
    finally:
        given.close()
    assert np.max(np.abs(mu - mu_given) / np.abs(mu_given[:, :, :1])) <= 1e-12
    for s, (h, (lo, hi)) in enumerate(zip(hs, ranges)):
        a, b = oracle.scale_factors(lo, hi)
        h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
        ref = oracle.diag_moments(h2, oracle.random_starter(h.shape[0], 4, np.float64), M)
        assert np.max(np.abs(mu[s] - ref) / np.abs(ref[:, :1])) <= 1e-10, s


def test_batched_models(oracle):
    """Models in the batch: their Hamiltonians (graphene flakes of two sizes) in the Hilbert order
    of their site positions, against the oracle."""
    import pybinding_amd as pb
    from pybinding_amd.repository import graphene
    models = [pb.Model(graphene.monolayer(), pb.rectangle(w), pb.force_double_precision()) for w in (2.0, 3.5)]
    ranges = [(-9, 9), (-9.5, 9.5)]
    kb = kpm_batch(models, ranges)
    try:
        mu = kb.moments(64, 2)
    finally:
        kb.close()
    for s, (m, (lo, hi)) in enumerate(zip(models, ranges)):
        h = m.hamiltonian
        a, b = oracle.scale_factors(lo, hi)
        h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
        ref = oracle.diag_moments(h2, oracle.random_starter(h.shape[0], 2, np.float64), 64)
        assert np.max(np.abs(mu[s] - ref) / np.abs(ref[:, :1])) <= 1e-10, s


@pytest.mark.parametrize("nvec,dtype,tol", [(4, np.float64, 1e-10), (8, np.float64, 1e-10), (4, np.float32, 1e-4),
                                            (4, np.complex128, 1e-10)])
def test_batched_persistent_run(oracle, nvec, dtype, tol):
    """All steps in one persistent launch (grid barriers between steps) for contiguous and
    XCD-aware partitions: every system matches the oracle and the per-step launches to rounding,
    and a partition wider than the resident grid is refused."""
    hs, ranges = systems(np.float64, count=25, seed=13)
    if dtype == np.complex128:
        hs = [h.astype(np.complex128) for h in hs]
    elif dtype == np.float32:
        hs = [h.astype(np.float32) for h in hs]
    kb = kpm_batch(hs, ranges)
    M = 70
    try:
        out = {}
        for wg, chunk in [(64, 0), (256, 2), (8, 1)]:
            kb.set_partition(wg, chunk)
            out[wg, chunk] = kb.moments(M, nvec, use_graph="persistent")
            again = kb.moments(M, nvec, use_graph="persistent")
            np.testing.assert_array_equal(again, out[wg, chunk])
            steps = kb.moments(M, nvec, use_graph=False)
            scale = np.abs(steps[:, :, :1])
            assert np.max(np.abs(out[wg, chunk] - steps) / scale) <= (1e-13 if tol < 1e-6 else 1e-5)
        kb.set_partition(2056, 1)  # more workgroups than 8 per CU: never resident at once
        with pytest.raises(RuntimeError, match="resident grid"):
            kb.moments(M, nvec, use_graph="persistent")
    finally:
        kb.close()
    mu = out[64, 0]
    for s, (h, (lo, hi)) in enumerate(zip(hs, ranges)):
        a, b = oracle.scale_factors(lo, hi)
        h2 = oracle.scale_hamiltonian(h, a, b, dtype)
        ref = oracle.diag_moments(h2, oracle.random_starter(h.shape[0], nvec, dtype), M)
        assert np.max(np.abs(mu[s] - ref) / np.abs(ref[:, :1])) <= tol, s


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-4), (np.complex128, 1e-10)])
def test_batched_value_table_form(oracle, dtype, tol):
    """The value-table form (1-byte codes into one table of off-diagonal values per system, dense
    diagonal) is built for flakes with one hopping value each (clean ones without diagonal
    entries, disordered ones with them), is the default, reads fewer bytes than the standard form,
    and matches it to rounding and every system the oracle, under both run modes."""
    hs, ranges = systems(np.float64, count=12, seed=17)
    hs[3] = generators.square_flake(30, 20, t=-1.0, disorder=0.0, dtype=np.float64)  # no onsite terms
    ranges[3] = (-4.0, 4.0)
    hs = [h.astype(dtype) for h in hs]
    kb = kpm_batch(hs, ranges)
    M = 90
    try:
        avail, used, coded_bytes = kb.matrix_format
        assert avail and used
        coded = kb.moments(M, 4)
        coded_steps = kb.moments(M, 4, use_graph=False)
        kb.set_value_table(False)
        avail, used, std_bytes = kb.matrix_format
        assert avail and not used and coded_bytes < std_bytes
        std = kb.moments(M, 4)
        kb.set_value_table(True)
        again = kb.moments(M, 4)
    finally:
        kb.close()
    np.testing.assert_array_equal(coded, again)
    scale = np.abs(std[:, :, :1])
    assert np.max(np.abs(coded - std) / scale) <= (1e-13 if tol < 1e-6 else 1e-5)
    assert np.max(np.abs(coded_steps - coded) / scale) <= (1e-13 if tol < 1e-6 else 1e-5)
    for s, (h, (lo, hi)) in enumerate(zip(hs, ranges)):
        a, b = oracle.scale_factors(lo, hi)
        h2 = oracle.scale_hamiltonian(h, a, b, dtype)
        ref = oracle.diag_moments(h2, oracle.random_starter(h.shape[0], 4, dtype), M)
        assert np.max(np.abs(coded[s] - ref) / np.abs(ref[:, :1])) <= tol, s


def test_batched_without_value_table(oracle):
    """A system with more than 255 distinct off-diagonal values leaves the batch on the standard
    form: the value-table form is not offered (selecting it raises) and the moments match the
    oracle."""
    hs, ranges = systems(np.float64, count=4, seed=19)
    rng = np.random.default_rng(5)
    h = generators.square_flake(40, 30, t=-1.0, disorder=1.0, rng=rng).tocsr()
    off = h.indices != np.repeat(np.arange(h.shape[0]), np.diff(h.indptr))
    h.data[off] *= 1 + 0.05 * rng.random(int(off.sum()))  # many distinct hoppings
    hs.append(((h + h.T) * 0.5).tocsr())
    ranges.append((-5.0, 5.0))
    kb = kpm_batch(hs, ranges)
    M = 60
    try:
        avail, used, _ = kb.matrix_format
        assert not avail and not used
        with pytest.raises(RuntimeError, match="no value-table form"):
            kb.set_value_table(True)
        mu = kb.moments(M, 2)
    finally:
        kb.close()
    for s, (hh, (lo, hi)) in enumerate(zip(hs, ranges)):
        a, b = oracle.scale_factors(lo, hi)
        h2 = oracle.scale_hamiltonian(hh, a, b, np.float64)
        ref = oracle.diag_moments(h2, oracle.random_starter(hh.shape[0], 2, np.float64), M)
        assert np.max(np.abs(mu[s] - ref) / np.abs(ref[:, :1])) <= 1e-10, s


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.complex128, 1e-10), (np.float32, 1e-4)])
def test_batched_column_tiles(oracle, dtype, tol):
    """Ten vectors per system as column tiles of 4, 4 and 2 vectors (three recurrences over the
    batch) against one 16-wide block: the same moments to rounding under every run mode, and
    every system's moments the oracle's."""
    hs, ranges = systems(np.float64, count=9, seed=23)
    hs = [h.astype(dtype) for h in hs]
    kb = kpm_batch(hs, ranges)
    M, R = 74, 10
    close = 1e-13 if tol < 1e-6 else 1e-5
    try:
        kb.set_tile_width(64)
        kb.set_random_starters(R)
        assert kb.batch.tiles() == (1, 10)
        one = kb.moments(M, R)
        kb.set_tile_width(4)
        assert kb.batch.tiles() == (3, 4)
        tiled = {mode: kb.moments(M, R, use_graph=mode) for mode in (True, False)}
        kb.set_partition(64, 0)
        tiled["persistent"] = kb.moments(M, R, use_graph="persistent")
    finally:
        kb.close()
    scale = np.abs(one[:, :, :1])
    for mode, mu in tiled.items():
        assert np.max(np.abs(mu - one) / scale) <= close, mode
    for s, (h, (lo, hi)) in enumerate(zip(hs, ranges)):
        a, b = oracle.scale_factors(lo, hi)
        h2 = oracle.scale_hamiltonian(h, a, b, dtype)
        ref = oracle.diag_moments(h2, oracle.random_starter(h.shape[0], R, dtype), M)
        assert np.max(np.abs(tiled[True][s] - ref) / np.abs(ref[:, :1])) <= tol, s
