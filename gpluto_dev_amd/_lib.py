"""ctypes binding of the C ABI in include/gpluto_amd.h (libgpluto_amd.so).

The product path: every call goes to the HIP library.  Loading fails loudly
when the library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpluto_amd.so")

GPL_HD, GPL_MHD, GPL_RMHD = 1, 2, 3
GPL_SOLVER_HLL, GPL_SOLVER_HLLC, GPL_SOLVER_HLLD = 1, 2, 3
GPL_VANLEER_LIM, GPL_MC_LIM = 26, 27
GPL_UCT_HLL, GPL_UCT_HLLD = 1, 2
GPL_RK2, GPL_RK3 = 2, 3
GPL_PERIODIC, GPL_OUTFLOW, GPL_REFLECTIVE = 1, 2, 3
GPL_FP64, GPL_MIXED = 0, 1
GPL_LAYOUT_ACTIVE, GPL_LAYOUT_ARY4D = 0, 1
GPL_EINVAL, GPL_EHIP, GPL_ENOMEM, GPL_EUNSUPPORTED, GPL_EIO, GPL_ESCRATCH = -1, -2, -3, -4, -5, -6
GPL_C2P_NOALLOC_MAX = 65536
GPL_STEP_API_ERROR = 0x80000000     # stepping calls: the call itself failed

# every symbol include/gpluto_amd.h declares
EXPORTS = ("gpl_default_config", "gpl_create", "gpl_destroy", "gpl_upload", "gpl_download",
           "gpl_advance_step", "gpl_next_time_step", "gpl_run", "gpl_set_timing",
           "gpl_kernel_count", "gpl_kernel_stats", "gpl_last_run_ms", "gpl_device_bytes",
           "gpl_strerror",
           "gpl_batch_create", "gpl_batch_destroy", "gpl_batch_upload", "gpl_batch_download",
           "gpl_batch_run", "gpl_batch_set_tuning", "gpl_batch_get_tuning",
           "gpl_batch_last_run_ms", "gpl_batch_device_bytes",
           "gpl_c2p_net_size", "gpl_c2p_net_make_data", "gpl_c2p_net_train", "gpl_c2p_net_grad",
           "gpl_c2p_net_infer", "gpl_c2p_solve", "gpl_set_c2p_network", "gpl_diagnostics",
           "gpl_set_step", "gpl_get_step", "gpl_write_dbl", "gpl_read_dbl",
           "gpl_write_dbl_async", "gpl_wait_dbl", "gpl_set_zone_reports", "gpl_get_zone_reports",
           "gpl_c2p_solve_noalloc", "gpl_c2p_net_pairs", "gpl_c2p_net_fit")


class GplConfig(C.Structure):
    _fields_ = [("physics", C.c_int), ("dimensions", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xbeg", C.c_double * 3), ("xend", C.c_double * 3),
                ("boundary", C.c_int), ("time_stepping", C.c_int), ("limiter", C.c_int),
                ("solver", C.c_int), ("ct_emf_average", C.c_int), ("precision", C.c_int),
                ("nghost", C.c_int), ("device", C.c_int),
                ("gamma", C.c_double), ("cfl", C.c_double), ("cfl_max_var", C.c_double),
                ("small_rho", C.c_double), ("small_prs", C.c_double), ("side_bc", C.c_int * 6)]


_lib = None


def load() -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gpluto_dev_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
    L.gpl_default_config.argtypes = [C.POINTER(GplConfig)]
    L.gpl_default_config.restype = None
    L.gpl_create.argtypes = [C.POINTER(GplConfig), C.POINTER(vp)]
    L.gpl_create.restype = ip
    L.gpl_destroy.argtypes = [vp]
    L.gpl_destroy.restype = ip
    L.gpl_upload.argtypes = [vp, vp, C.POINTER(vp), ip]
    L.gpl_upload.restype = ip
    L.gpl_download.argtypes = [vp, vp, C.POINTER(vp), vp, ip]
    L.gpl_download.restype = ip
    L.gpl_set_step.argtypes = [vp, C.c_double, C.c_double, C.c_longlong]
    L.gpl_set_step.restype = ip
    L.gpl_get_step.argtypes = [vp, dp, dp, C.POINTER(C.c_longlong)]
    L.gpl_get_step.restype = ip
    L.gpl_write_dbl.argtypes = [vp, C.c_char_p, ip, ip]
    L.gpl_write_dbl.restype = ip
    L.gpl_read_dbl.argtypes = [vp, C.c_char_p, ip]
    L.gpl_read_dbl.restype = ip
    L.gpl_write_dbl_async.argtypes = [vp, C.c_char_p, ip, ip]
    L.gpl_write_dbl_async.restype = ip
    L.gpl_wait_dbl.argtypes = [vp]
    L.gpl_wait_dbl.restype = ip
    L.gpl_advance_step.argtypes = [vp, C.c_double, dp]
    L.gpl_advance_step.restype = C.c_uint32
    L.gpl_next_time_step.argtypes = [vp, C.c_double, C.c_double]
    L.gpl_next_time_step.restype = C.c_double
    L.gpl_run.argtypes = [vp, ip, C.c_double, dp, dp]
    L.gpl_run.restype = C.c_uint32
    L.gpl_set_timing.argtypes = [vp, ip]
    L.gpl_set_timing.restype = ip
    L.gpl_kernel_count.argtypes = [vp]
    L.gpl_kernel_count.restype = ip
    L.gpl_kernel_stats.argtypes = [vp, ip, C.POINTER(C.c_char_p), C.POINTER(C.c_longlong), dp]
    L.gpl_kernel_stats.restype = ip
    L.gpl_last_run_ms.argtypes = [vp]
    L.gpl_last_run_ms.restype = C.c_double
    L.gpl_device_bytes.argtypes = [vp]
    L.gpl_device_bytes.restype = C.c_size_t
    L.gpl_strerror.argtypes = [ip]
    L.gpl_strerror.restype = C.c_char_p
    ipp = C.POINTER(C.c_int)
    L.gpl_batch_create.argtypes = [C.POINTER(GplConfig), ip, ipp, ipp, C.POINTER(vp)]
    L.gpl_batch_create.restype = ip
    L.gpl_batch_destroy.argtypes = [vp]
    L.gpl_batch_destroy.restype = ip
    L.gpl_batch_upload.argtypes = [vp, ip, vp, C.POINTER(vp), ip]
    L.gpl_batch_upload.restype = ip
    L.gpl_batch_download.argtypes = [vp, ip, vp, C.POINTER(vp), vp, ip]
    L.gpl_batch_download.restype = ip
    L.gpl_batch_run.argtypes = [vp, ip, dp, dp, dp]
    L.gpl_batch_run.restype = C.c_uint32
    L.gpl_batch_set_tuning.argtypes = [vp, ip, ip]
    L.gpl_batch_set_tuning.restype = ip
    L.gpl_batch_get_tuning.argtypes = [vp, ipp, ipp, dp]
    L.gpl_batch_get_tuning.restype = ip
    L.gpl_batch_last_run_ms.argtypes = [vp]
    L.gpl_batch_last_run_ms.restype = C.c_double
    L.gpl_batch_device_bytes.argtypes = [vp]
    L.gpl_batch_device_bytes.restype = C.c_size_t
    fp = C.POINTER(C.c_float)
    L.gpl_c2p_net_size.argtypes = []
    L.gpl_c2p_net_size.restype = ip
    L.gpl_c2p_net_make_data.argtypes = [ip, C.c_uint64, ip, C.c_double, dp, fp, fp]
    L.gpl_c2p_net_make_data.restype = ip
    L.gpl_c2p_net_train.argtypes = [ip, C.c_uint64, ip, ip, ip, C.c_float, C.c_double, fp, fp]
    L.gpl_c2p_net_train.restype = ip
    L.gpl_c2p_net_pairs.argtypes = [ip, dp, dp, ip, C.c_double, fp, fp]
    L.gpl_c2p_net_pairs.restype = ip
    L.gpl_c2p_net_fit.argtypes = [ip, C.c_uint64, fp, fp, ip, ip, ip, C.c_float, fp, fp]
    L.gpl_c2p_net_fit.restype = ip
    L.gpl_c2p_net_grad.argtypes = [ip, fp, fp, fp, ip, fp, fp]
    L.gpl_c2p_net_grad.restype = ip
    L.gpl_c2p_net_infer.argtypes = [ip, fp, fp, ip, fp]
    L.gpl_c2p_net_infer.restype = ip
    L.gpl_c2p_solve.argtypes = [ip, fp, dp, ip, C.c_double, dp, C.POINTER(C.c_int),
                                C.POINTER(C.c_uint32), dp, C.POINTER(C.c_int)]
    L.gpl_c2p_solve.restype = ip
    L.gpl_c2p_solve_noalloc.argtypes = [ip, fp, dp, ip, C.c_double, dp, C.POINTER(C.c_int),
                                        C.POINTER(C.c_uint32), C.POINTER(C.c_int)]
    L.gpl_c2p_solve_noalloc.restype = ip
    L.gpl_set_c2p_network.argtypes = [vp, fp]
    L.gpl_set_c2p_network.restype = ip
    L.gpl_diagnostics.argtypes = [vp, dp, dp, dp, dp]
    L.gpl_diagnostics.restype = ip
    L.gpl_set_zone_reports.argtypes = [vp, ip]
    L.gpl_set_zone_reports.restype = ip
    L.gpl_get_zone_reports.argtypes = [vp, vp, vp, C.POINTER(C.c_ulonglong)]
    L.gpl_get_zone_reports.restype = ip
    _lib = L
    return L


def check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what}: {load().gpl_strerror(rc).decode()} ({rc})")


def check_step(err: int, what: str) -> int:
    """Stepping calls return the OR of the ERR_* bits the numerics raised;
    GPL_STEP_API_ERROR (bit 31, never a numerics bit) says the call itself
    failed -- a HIP runtime error or host memory -- and raises here."""
    err = int(err)
    if err & GPL_STEP_API_ERROR:
        raise RuntimeError(f"{what}: the call failed (HIP runtime error or out of host memory)")
    return err
