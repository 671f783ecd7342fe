"""CPU checks of the C-ABI boundary: the shared library loads and exports
every function include/fideslib_amd.h declares, the Python binding covers
exactly that set, and argument validation that needs no GPU rejects bad
parameters.  No compute call is made (there is no GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from fideslib_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(capi.HEADER_PATH).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fides_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(capi.LIB_PATH):
        from fideslib_amd import build
        build.build()
    return capi.load()


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("fides_context_create", "fides_ntt", "fides_intt", "fides_modup", "fides_moddown",
                 "fides_keyswitch", "fides_hmult", "fides_rescale", "fides_rotate", "fides_key_upload"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_compute_entry_points_open_roctx_ranges(lib):
    # the library imports the ROCTx range calls, and every compute entry point
    # of the key-switching path opens one range named after itself
    import subprocess
    dyn = subprocess.run(["nm", "-D", capi.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert re.search(r"\bU roctxRangePushA\b", dyn) and re.search(r"\bU roctxRangePop\b", dyn)
    src = open(os.path.join(ROOT, "fideslib_amd", "csrc", "context.hip")).read()
    for name in ("fides_ntt", "fides_intt", "fides_modup", "fides_moddown", "fides_keyswitch", "fides_hmult",
                 "fides_hmult_batch", "fides_rescale", "fides_rotate", "fides_rotate_batch", "fides_rotate_hoisted",
                 "fides_lincomb_batch", "fides_ptdot_batch", "fides_graph_launch"):
        m = re.search(r'extern "C" int ' + name + r"\([^{]*\{\n  FIDES_TRACE\(\);", src)
        assert m, name


def test_binding_matches_header():
    assert sorted(capi.SIGNATURES) == declared_functions()


def test_context_rejects_bad_parameters(lib):
    q = np.array([12289], np.uint64)  # 14-bit prime = 1 mod 2^12, not 1 mod 2^11 * 2 for logN=12? check below
    p = np.array([40961], np.uint64)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    h = ctypes.c_void_p()
    # logN out of range
    prm = capi.Params(9, 1, 1, 1, q.ctypes.data_as(u64p), p.ctypes.data_as(u64p), None)
    assert lib.fides_context_create(ctypes.byref(prm), ctypes.byref(h)) == -3
    # composite "prime"
    bad = np.array([2 ** 20 * 3 + 1 + 2 ** 11 * 0], np.uint64)
    prm = capi.Params(10, 1, 1, 1, bad.ctypes.data_as(u64p), p.ctypes.data_as(u64p), None)
    assert lib.fides_context_create(ctypes.byref(prm), ctypes.byref(h)) == -1
    assert b"prime" in lib.fides_last_error()
    # K = 0
    prm = capi.Params(10, 1, 0, 1, q.ctypes.data_as(u64p), p.ctypes.data_as(u64p), None)
    assert lib.fides_context_create(ctypes.byref(prm), ctypes.byref(h)) == -1


def test_tuning_null_arguments_and_dict_round_trip(lib):
    """Tuning entry points refuse null handles without touching a GPU; the
    binding's dict form round-trips and needs one entry per size class."""
    t = capi.Tuning()
    assert lib.fides_tuning_get(None, ctypes.byref(t)) == -1
    assert lib.fides_tuning_set(None, ctypes.byref(t)) == -1
    lv = (ctypes.c_uint32 * 1)(30)
    assert lib.fides_autotune(None, 1, lv, 2, ctypes.byref(t)) == -1
    d = {"conv_chunk": [8, 16, 32], "group_by_prime": [0, 1, 1], "fuse_row_dot": [1, 0, 1]}
    assert capi.Tuning.from_dict(d).as_dict() == d
    with pytest.raises(capi.FidesError):
        capi.Tuning.from_dict(dict(d, conv_chunk=[8, 16]))


def test_context_refuses_many_digits_with_wide_primes(lib):
    """More than 16 digits need every prime below 2^61 (the 128-bit column
    sums of the inner product hold 16 products of 62-bit values); refused
    before any device call."""
    from fideslib_amd import params
    q = np.ascontiguousarray(params.ntt_primes(62, 17, 10), np.uint64)
    p = np.ascontiguousarray(params.ntt_primes(62, 2, 10, exclude=list(q)), np.uint64)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    h = ctypes.c_void_p()
    prm = capi.Params(10, 17, 2, 17, q.ctypes.data_as(u64p), p.ctypes.data_as(u64p), None)
    assert lib.fides_context_create(ctypes.byref(prm), ctypes.byref(h)) == -3
    assert b"dnum" in lib.fides_last_error()
