"""The allocation-free diagonal entry point (kpmx_moments_diagonal_noalloc) against
kpmx_moments_diagonal, through the C-ABI: the same kernels on buffers in the library's static
device scratch array must give the same moments, bit for bit, on consecutive calls (the second
one smaller, so it runs on what the first left in the array), for real and complex scalars and
for blocks of more than 64 vectors; sizes beyond the array are refused with a message."""
import numpy as np
import pytest

import abi
from pybinding_amd import generators

pytestmark = pytest.mark.gpu


def scaled(n1, n2, dtype, seed):
    import oracle
    h = generators.honeycomb_supercell(n1, n2, disorder=0.7, seed=seed, dtype=np.float64)
    a, b = oracle.scale_factors(-9.0, 9.0)
    return oracle.scale_hamiltonian(h, a, b, dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
def test_noalloc_matches_allocating_path(dtype):
    import oracle
    ctx = abi.Ctx(0)
    try:
        for (n1, n2, R, M, fmt) in [(40, 30, 16, 130, 1), (12, 9, 3, 42, 0), (20, 16, 70, 64, 1)]:
            print("noalloc: %s, %dx%d cells, %d vectors, %d moments, format %d" % (np.dtype(dtype), n1, n2, R, M, fmt),
                  flush=True)
            h2 = scaled(n1, n2, dtype, seed=n1)
            r0 = oracle.random_starter(h2.shape[0], R, dtype)
            mat, blk = ctx.mat(h2, fmt=fmt), ctx.block(r0)
            ref = abi.moments_diagonal(mat, blk, M)
            got = abi.moments_diagonal_noalloc(mat, blk, M)
            blk.close()
            mat.close()
            np.testing.assert_array_equal(got, ref)
    finally:
        ctx.close()


def test_noalloc_refuses_oversized_blocks():
    ctx = abi.Ctx(0)
    try:
        h2 = scaled(512, 256, np.float64, seed=3)   # 262144 sites
        mat = ctx.mat(h2, fmt=1)
        blk = ctx.block(np.ones((64, h2.shape[0])))  # 2 x 128 MiB of work vectors
        with pytest.raises(RuntimeError, match="limit is 67108864 bytes"):
            abi.moments_diagonal_noalloc(mat, blk, 16)
        blk.close()
        mat.close()
    finally:
        ctx.close()
