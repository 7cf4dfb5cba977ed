"""gpl_c2p_solve_noalloc, the entry point that allocates nothing, against
gpl_c2p_solve through the C ABI: the same states give the same bits (the same
kernel), on two calls in a row with the second smaller than the first (the
fixed scratch keeps the first call's contents beyond n), from the reference's
Eq. (A27) start and from a learned start; a request beyond the scratch is
refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from gpluto_dev_amd import _lib

pytestmark = pytest.mark.gpu


def _states(rng, n):
    from test_host_physics import _states as prim
    p = prim(rng, n)
    u = np.empty_like(p)
    g2 = 1 / (1 - np.sum(p[:, 1:4] ** 2, 1))
    h = 1 + 4 * p[:, 7] / p[:, 0]
    u[:, 0] = np.sqrt(g2) * p[:, 0]
    u[:, 1:4] = (p[:, 0] * h * g2)[:, None] * p[:, 1:4]
    u[:, 4:7] = p[:, 4:7]
    u[:, 7] = p[:, 0] * h * g2 - p[:, 7] - u[:, 0] + 0.5 * np.sum(p[:, 4:7] ** 2, 1)
    u[rng.uniform(size=n) < 0.05, 7] *= -1.0       # the pressure-fix branch
    return np.ascontiguousarray(u)


def _solve(L, fn, u, weights):
    n = u.shape[0]
    v = np.zeros((n, 8))
    passes = np.zeros(n, dtype=np.int32)
    err = np.zeros(n, dtype=np.uint32)
    rs = np.full(n, -1, dtype=np.int32)
    dp = C.POINTER(C.c_double)
    w = None if weights is None else weights.ctypes.data_as(C.POINTER(C.c_float))
    args = [0, w, u.ctypes.data_as(dp), n, 4.0 / 3.0, v.ctypes.data_as(dp),
            passes.ctypes.data_as(C.POINTER(C.c_int)), err.ctypes.data_as(C.POINTER(C.c_uint32))]
    if fn == "gpl_c2p_solve":
        ms = C.c_double(0.0)
        rc = L.gpl_c2p_solve(*args, C.byref(ms), rs.ctypes.data_as(C.POINTER(C.c_int)))
    else:
        rc = L.gpl_c2p_solve_noalloc(*args, rs.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == 0, L.gpl_strerror(rc)
    return v, passes, err, rs


@pytest.mark.parametrize("learned", [False, True])
def test_noalloc_equals_allocating_solve(learned):
    L = _lib.load()
    weights = None
    if learned:
        from gpluto_dev_amd import c2p_net
        weights, _ = c2p_net.train(nsamples=1 << 14, steps=200, batch=1024, lr=3e-3, seed=5)
        weights = np.ascontiguousarray(weights, dtype=np.float32)
    rng = np.random.default_rng(17)
    for n in (_lib.GPL_C2P_NOALLOC_MAX, 3001):       # the second call smaller than the first
        u = _states(rng, n)
        print(f"case learned={learned} n={n}", flush=True)
        a = _solve(L, "gpl_c2p_solve", u, weights)
        b = _solve(L, "gpl_c2p_solve_noalloc", u, weights)
        for x, y in zip(a, b):
            assert np.array_equal(x, y, equal_nan=True)


def test_noalloc_refuses_more_than_its_scratch():
    L = _lib.load()
    n = _lib.GPL_C2P_NOALLOC_MAX + 1
    u = np.zeros((n, 8))
    v = np.zeros((n, 8))
    passes = np.zeros(n, dtype=np.int32)
    err = np.zeros(n, dtype=np.uint32)
    dp = C.POINTER(C.c_double)
    rc = L.gpl_c2p_solve_noalloc(0, None, u.ctypes.data_as(dp), n, 4.0 / 3.0, v.ctypes.data_as(dp),
                                 passes.ctypes.data_as(C.POINTER(C.c_int)),
                                 err.ctypes.data_as(C.POINTER(C.c_uint32)), None)
    assert rc == _lib.GPL_ESCRATCH
    assert b"scratch" in L.gpl_strerror(rc)
