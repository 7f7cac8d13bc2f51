"""The host core's reference random starters (std::mt19937 default seed, libstdc++
uniform_real_distribution; cppcore/src/kpm/Starter.cpp:48-83, Starter.hpp:42-50) against the
oracle's, bit for bit, for every scalar type: several columns continue one stream, and each
column is permuted into the optimized order by the reorder map after it is drawn."""
import numpy as np
import pytest

from pybinding_amd._native import load_module


@pytest.mark.parametrize("dtype", ["float32", "float64", "complex64", "complex128"])
@pytest.mark.parametrize("n,nvec", [(1, 3), (977, 5), (4096, 2)])
def test_starters_bit_identical(oracle, dtype, n, nvec):
    native = load_module()
    rmap = np.random.default_rng(n).permutation(n).astype(np.int32)
    got = native.random_starters(n, nvec, dtype, rmap)
    ref = oracle.random_starter(n, nvec, np.dtype(dtype))
    expect = np.empty_like(ref)
    expect[:, rmap] = ref
    assert got.dtype == np.dtype(dtype)
    np.testing.assert_array_equal(got, expect)
    np.testing.assert_array_equal(native.random_starters(n, nvec, dtype, []), ref)


@pytest.mark.parametrize("dtype,group", [("float64", 0), ("float64", 2), ("float32", 3), ("complex128", 1),
                                         ("complex64", 0)])
def test_threaded_stream_bit_identical(oracle, dtype, group):
    """Blocks long enough to be drawn by several threads (chunks of 2^24 generator words, each
    started by the jump-ahead), drawn all at once or a few columns per call (each call continues
    the stream where the last one stopped)."""
    native = load_module()
    n, nvec = 3_000_001, 7
    got = native.random_starters(n, nvec, dtype, [], group)
    ref = oracle.random_starter(n, nvec, np.dtype(dtype))
    np.testing.assert_array_equal(got, ref)
