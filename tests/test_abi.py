"""The C-ABI library loads and exports every entry point include/khx.h declares
(no kernel is launched: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "khx.h")).read()
    return sorted(set(re.findall(r"\b(khx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "khx_eig" in names and "khx_redheffer" in names and len(names) >= 15


def test_library_exports_all_declared_symbols():
    from khepri_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_host_only_entry_points():
    from khepri_amd import _lib
    lib = _lib.load()
    assert lib.khx_abi_version() == 2
    assert lib.khx_eig_ws_bytes(50, 4) >= 2 * 50 * 50 * 16 * 4
    assert lib.khx_redheffer_ws_bytes(162, 2) >= 8 * 162 * 162 * 16 * 2
    # argument validation happens on the host, before any launch
    assert lib.khx_halfspace(3, None, None, 25, 1, None, None, None, None, None, None) == -1
    assert b"side" in lib.khx_last_error()
    # GEMM tile table: bad tile ids and shapes are rejected, recording is host-side
    assert lib.khx_gemm_force(8) == -1
    assert lib.khx_gemm_tune_set(10, 10, 10, 0, 0, 1, 8) == -1
    assert lib.khx_gemm_tune_set(10, 10, 10, 3, 0, 1, 0) == -1
    assert lib.khx_gemm_tune_set(10, 10, 10, 0, 0, 1, 2) == 0
    assert lib.khx_gemm_tune_set(10, 10, 10, 0, 0, 1, -1) == 0
    assert lib.khx_gemm_record(1, None, 0) == 0
    assert lib.khx_gemm_record(0, None, 0) == 0
    # the no-allocation eigensolver: a fixed 64 MiB device array, batches beyond it refused on the host
    assert lib.khx_eig_noalloc_bytes() == 64 << 20
    assert lib.khx_eig_noalloc(162, None, 80, None, None, None, None) == -1
    assert b"static workspace" in lib.khx_last_error()
