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
    """The autotuner's partitions (256, 512 and 523 workgroups over 523 slices) change only how
    each system's sums are grouped: every system still matches the oracle, and the tuned run agrees
    with the default partition to rounding."""
    hs, ranges = systems(np.float64, count=40, seed=8)
    kb = kpm_batch(hs, ranges)
    M = 66
    before = kb.moments(M, 4)
    wg = kb.autotune(4)
    after = kb.moments(M, 4)
    again = kb.moments(M, 4, use_graph=False)
    kb.close()
    assert wg >= 1
    np.testing.assert_array_equal(after, again)
    assert np.max(np.abs(after - before) / np.abs(before[:, :, :1])) <= 1e-13
    for s, (h, (lo, hi)) in enumerate(zip(hs, ranges)):
        a, b = oracle.scale_factors(lo, hi)
        h2 = oracle.scale_hamiltonian(h, a, b, np.float64)
        ref = oracle.diag_moments(h2, oracle.random_starter(h.shape[0], 4, np.float64), M)
        assert np.max(np.abs(after[s] - ref) / np.abs(ref[:, :1])) <= 1e-10, s
