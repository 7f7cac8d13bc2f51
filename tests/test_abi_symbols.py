"""The C-ABI library loads without a GPU and exports every function include/kpmx.h declares."""
import ctypes

import abi


def test_header_declares_the_boundary():
    names = abi.declared_functions()
    for required in ("kpmx_ctx_create", "kpmx_mat_upload", "kpmx_block_upload", "kpmx_moments_diagonal",
                     "kpmx_moments_generic", "kpmx_moments_multi_unit", "kpmx_moments_dense",
                     "kpmx_kubo_moments", "kpmx_lanczos_bounds", "kpmx_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB)
    missing = [n for n in abi.declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_string_without_gpu():
    lib = abi.lib()
    assert lib.kpmx_version() >= 100
    assert isinstance(lib.kpmx_last_error(), bytes)
