"""The pointwise physics and the correctly rounded acos/sincos built for the GPU
(tests/host/dev_physics.hip, the library's FP64 flags) against the host build
of the same source with the same math (tests/host: host_crmath.cpp and
host_physics.cpp without GPL_CUBIC_HOST_LIBM): bit for bit.

The host build is in turn bitwise the oracle (tests/test_host_physics.py with
the host libm, tests/test_crmath.py for the routines), so these tests close
the chain device == host source == oracle for every interface solver and the
RMHD con2prim, each evaluated alone, away from the stage update's kernels."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from test_host_physics import _states

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "_build")
DP = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def libs():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "host")], check=True)
    dev = C.CDLL(os.path.join(BUILD, "libdev_physics.so"))
    dev.dp_crmath.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    dev.dp_face.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double,
                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    dev.dp_rmhd_c2p.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p]
    dev.dp_face_rmhd_hllc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    for f in (dev.dp_crmath, dev.dp_face, dev.dp_rmhd_c2p, dev.dp_face_rmhd_hllc):
        f.restype = C.c_int
    host = C.CDLL(os.path.join(BUILD, "libhost_physics_cr.so"))
    host.hp_face.argtypes = [C.c_int, C.c_int, DP, DP, C.c_int, C.c_double, DP, DP, DP]
    host.hp_face.restype = C.c_uint32
    host.hp_rmhd_c2p.argtypes = [DP, DP, C.c_double, C.c_double, C.c_double]
    host.hp_rmhd_c2p.restype = C.c_uint32
    crm = C.CDLL(os.path.join(BUILD, "libhost_crmath.so"))
    crm.crm_acos.argtypes = [C.c_double]
    crm.crm_acos.restype = C.c_double
    crm.crm_sincos.argtypes = [C.c_double, DP, DP]
    return dev, host, crm


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_crmath_device_equals_host(libs):
    dev, _, crm = libs
    rng = np.random.default_rng(3)
    n = 200000
    e = rng.integers(1, 53, n)
    xa = np.concatenate([rng.uniform(-1, 1, n), (1 - 2.0 ** -e) * rng.choice([-1.0, 1.0], n),
                         [1.0, -1.0, 0.0, 0.5, -0.5]])
    out = np.empty_like(xa)
    assert dev.dp_crmath(_p(xa), xa.size, 0, _p(out), None) == 0
    want = np.array([crm.crm_acos(float(x)) for x in xa])
    bad = np.nonzero(out != want)[0]
    print(f"acos: {bad.size} of {xa.size} differ", flush=True)
    assert bad.size == 0, [(xa[q], out[q], want[q]) for q in bad[:5]]
    xs = np.concatenate([rng.uniform(0, np.pi / 3, n), rng.uniform(0, 2.19, n), [0.0, 2.19]])
    s, c = np.empty_like(xs), np.empty_like(xs)
    assert dev.dp_crmath(_p(xs), xs.size, 1, _p(s), _p(c)) == 0
    hs, hc = C.c_double(), C.c_double()
    ws, wc = np.empty_like(xs), np.empty_like(xs)
    for q, x in enumerate(xs):
        crm.crm_sincos(float(x), C.byref(hs), C.byref(hc))
        ws[q], wc[q] = hs.value, hc.value
    bad = np.nonzero((s != ws) | (c != wc))[0]
    print(f"sincos: {bad.size} of {xs.size} differ", flush=True)
    assert bad.size == 0, [(xs[q], s[q], ws[q], c[q], wc[q]) for q in bad[:5]]


# rmhd-hllc-alone: the RMHD HLLC solver alone in its kernel, as the stage
# kernels instantiate it.  Its branching form was miscompiled (the HLL flux
# of faces whose star state turned out superluminal came back with the star
# velocities in it); these cases pin the branch-free form.
@pytest.mark.parametrize("phys,solver,gamma,alone", [(3, 1, 4 / 3, 0), (3, 2, 4 / 3, 0), (3, 2, 4 / 3, 1),
                                                      (3, 3, 4 / 3, 0), (2, 3, 5 / 3, 0), (2, 2, 5 / 3, 0),
                                                      (2, 1, 5 / 3, 0)],
                         ids=["rmhd-hll", "rmhd-hllc", "rmhd-hllc-alone", "rmhd-hlld", "mhd-hlld", "mhd-hllc",
                              "mhd-hll"])
def test_interface_solvers_device_equals_host(libs, phys, solver, gamma, alone):
    dev, host, _ = libs
    rng = np.random.default_rng(21 + 3 * solver + phys)
    n = 4000
    vl, vr = _states(rng, n), _states(rng, n)
    if phys == 2:
        vl[:, 1:4] *= 3.0
        vr[:, 1:4] *= 3.0
    vr[:, 4:7] = vl[:, 4:7] * (1 + 0.1 * rng.normal(size=(n, 3)))
    for dirn in range(3):
        vr[:, 4 + dirn] = vl[:, 4 + dirn]
        a, b = np.ascontiguousarray(vl), np.ascontiguousarray(vr)
        f = np.zeros((n, 8))
        sl, sr = np.zeros(n), np.zeros(n)
        err = np.zeros(n, dtype=np.uint32)
        if alone:
            assert dev.dp_face_rmhd_hllc(_p(a), _p(b), n, dirn, gamma, _p(f), _p(sl), _p(sr), _p(err)) == 0
        else:
            assert dev.dp_face(phys, solver, _p(a), _p(b), n, dirn, gamma, _p(f), _p(sl), _p(sr), _p(err)) == 0
        bad = []
        for k in range(n):
            f2, s2 = np.zeros(8), np.zeros(4)
            e2 = host.hp_face(phys, solver, a[k].ctypes.data_as(DP), b[k].ctypes.data_as(DP), dirn, gamma,
                              f2.ctypes.data_as(DP), s2[0:].ctypes.data_as(DP), s2[1:].ctypes.data_as(DP))
            if not (e2 == err[k] and np.array_equal(f[k], f2, equal_nan=True)
                    and np.array_equal(np.array([sl[k], sr[k]]), s2[:2], equal_nan=True)):
                bad.append(k)
                if len(bad) <= 3:
                    print(f"  face {k}: device err {err[k]} flux {f[k].tolist()} S {sl[k]!r} {sr[k]!r}\n"
                          f"           host   err {e2} flux {f2.tolist()} S {s2[0]!r} {s2[1]!r}\n"
                          f"           vl {a[k].tolist()}\n           vr {b[k].tolist()}", flush=True)
        print(f"phys {phys} solver {solver} dir {dirn}: {len(bad)} of {n} faces differ", flush=True)
        assert not bad, bad[:5]


def test_rmhd_con2prim_device_equals_host(libs):
    dev, host, _ = libs
    rng = np.random.default_rng(9)
    n = 8000
    p = _states(rng, n)
    u = np.empty_like(p)
    g2 = 1 / (1 - np.sum(p[:, 1:4] ** 2, 1))
    h = 1 + 4 * p[:, 7] / p[:, 0]
    u[:, 0] = np.sqrt(g2) * p[:, 0]
    u[:, 1:4] = (p[:, 0] * h * g2)[:, None] * p[:, 1:4]
    u[:, 4:7] = p[:, 4:7]
    u[:, 7] = p[:, 0] * h * g2 - p[:, 7] - u[:, 0] + 0.5 * np.sum(p[:, 4:7] ** 2, 1)
    u[rng.uniform(size=n) < 0.05, 7] *= -1.0
    ud = np.ascontiguousarray(u.copy())
    vd = np.zeros_like(ud)
    err = np.zeros(n, dtype=np.uint32)
    assert dev.dp_rmhd_c2p(_p(ud), _p(vd), n, 4 / 3, _p(err)) == 0
    bad = 0
    for k in range(n):
        a, v = u[k].copy(), np.zeros(8)
        e = host.hp_rmhd_c2p(a.ctypes.data_as(DP), v.ctypes.data_as(DP), 4 / 3, 1e-12, 1e-12)
        ok = e == err[k] and np.array_equal(a, ud[k], equal_nan=True)
        if e & (1 << 5) == 0:
            ok = ok and np.array_equal(v, vd[k], equal_nan=True)
        bad += not ok
    print(f"con2prim: {bad} of {n} differ", flush=True)
    assert bad == 0
