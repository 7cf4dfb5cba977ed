"""The C ABI never lets a C++ exception reach its caller.  Host allocations
(std::vector, std::string) and std::thread throw on exhaustion; every
allocating entry point runs its body through c_guard
(gpluto_dev_amd/csrc/c_guard.hpp), built here for the host CPU
(tests/host/Makefile), which turns any exception into an error code."""
import ctypes as C
import os
import subprocess

import pytest

from gpluto_dev_amd._lib import GPL_STEP_API_ERROR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "_build", "libhost_guard.so")

GPL_ENOMEM = -3


@pytest.fixture(scope="module")
def hg():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "host")], check=True)
    lib = C.CDLL(LIB)
    lib.hg_guard.argtypes = [C.c_int]
    lib.hg_guard.restype = C.c_int
    lib.hg_guard_u32.argtypes = [C.c_int]
    lib.hg_guard_u32.restype = C.c_uint32
    return lib


def test_body_value_passes_through(hg):
    assert hg.hg_guard(0) == 7
    assert hg.hg_guard_u32(0) == 5


@pytest.mark.parametrize("kind", [1, 2, 3, 4], ids=["bad_alloc", "system_error", "length_error", "int"])
def test_exception_becomes_error_code(hg, kind):
    assert hg.hg_guard(kind) == GPL_ENOMEM


def test_stepping_calls_report_the_api_error_bit(hg):
    assert hg.hg_guard_u32(1) == GPL_STEP_API_ERROR


def test_library_entry_points_reject_a_null_handle():
    """Argument checks come before any HIP call, so these run without a GPU."""
    from gpluto_dev_amd import _lib
    L = _lib.load()
    assert L.gpl_next_time_step(None, 1e-3, 10.0) == 0.0
    assert L.gpl_run(None, 1, 1e-3, None, None) == GPL_STEP_API_ERROR
    assert L.gpl_advance_step(None, 1e-3, None) == GPL_STEP_API_ERROR
    for rc in (L.gpl_upload(None, None, None, 0), L.gpl_download(None, None, None, None, 0),
               L.gpl_wait_dbl(None), L.gpl_read_dbl(None, b".", 0),
               L.gpl_write_dbl_async(None, b".", 0, -1), L.gpl_destroy(None)):
        assert rc == _lib.GPL_EINVAL
    assert L.gpl_strerror(GPL_ENOMEM) == b"out of device or host memory"


def test_python_mirror_raises_on_a_failed_stepping_call():
    from gpluto_dev_amd import _lib
    assert _lib.check_step(0b101, "gpl_run") == 0b101          # numerics bits pass through
    with pytest.raises(RuntimeError, match="gpl_run: the call failed"):
        _lib.check_step(GPL_STEP_API_ERROR | 1, "gpl_run")
