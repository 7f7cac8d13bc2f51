"""Column sub-tiles of the diagonal recurrence (kpmx_ctx_set_tile_width): a starter block wider
than the tile width runs as recurrences over sub-tiles of that many columns.  Every width gives
the oracle's moments (cppcore/src/kpm/default/collectors.cpp:5-34 via oracle.diag_moments), the
widths agree with each other to rounding, and the allocation-free entry point runs the same
sub-tiles bit for bit."""
import numpy as np
import pytest

import abi
from pybinding_amd import generators

pytestmark = pytest.mark.gpu


def scaled(n1, n2, dtype, seed=4):
    import oracle
    h = generators.honeycomb_supercell(n1, n2, disorder=0.8, seed=seed, dtype=np.float64)
    a, b = oracle.scale_factors(-9.2, 9.2)
    return oracle.scale_hamiltonian(h, a, b, dtype)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.complex128, 1e-10), (np.float32, 1e-4)])
@pytest.mark.parametrize("nvec", [64, 45, 70])
def test_tile_widths_match_oracle(oracle, dtype, tol, nvec):
    h2 = scaled(24, 18, dtype)
    r0 = oracle.random_starter(h2.shape[0], nvec, dtype)
    M = 66
    ref = oracle.diag_moments(h2, r0, M)
    ctx = abi.Ctx(0)
    try:
        default = ctx.tile_width(dtype)
        assert default == (16 if dtype == np.float64 else 64)
        mat = ctx.mat(h2, fmt=1)
        blk = ctx.block(r0)
        out = {}
        for w in (64, 32, 16, 8, 4):
            ctx.set_tile_width(dtype, w)
            out[w] = abi.moments_diagonal(mat, blk, M)
            again = abi.moments_diagonal(mat, blk, M)
            np.testing.assert_array_equal(again, out[w])
        ctx.set_tile_width(dtype, 16)
        noalloc = abi.moments_diagonal_noalloc(mat, blk, M)
        np.testing.assert_array_equal(noalloc, out[16])
        blk.close()
        mat.close()
    finally:
        ctx.close()
    for w, mu in out.items():
        err = np.max(np.abs(mu - ref) / np.abs(ref[:, :1]))
        assert err <= tol, (w, err)
        assert np.max(np.abs(mu - out[64]) / np.abs(out[64][:, :1])) <= (1e-13 if tol < 1e-6 else 1e-5), w


def test_tile_width_rejects_bad_values():
    ctx = abi.Ctx(0)
    try:
        for bad in (0, 3, 128, -1):
            with pytest.raises(RuntimeError, match="tile width"):
                ctx.set_tile_width(np.float64, bad)
    finally:
        ctx.close()
