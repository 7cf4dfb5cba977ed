"""Learned initial guess for the RMHD con2prim Newton solve, over the HIP
library: training pairs (sampled states, or a run's own states), GPU
training, matrix-core inference and the batched solve (A27 guess, the
reference's, or the learned guess with the A27 restart)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


NF = 8          # input features


def size() -> int:
    return _lib.load().gpl_c2p_net_size()


def make_data(n: int, seed: int, gamma: float = 4.0 / 3.0, device: int = 0):
    """(u[n, 8], x[n, 8], y[n]): conserved states, raw features, log(W/W_A27)."""
    u = np.empty((n, 8))
    x = np.empty((n, NF), dtype=np.float32)
    y = np.empty(n, dtype=np.float32)
    _lib.check(_lib.load().gpl_c2p_net_make_data(device, seed, n, gamma, _d(u), _f(x), _f(y)),
               "gpl_c2p_net_make_data")
    return u, x, y


def pairs(u, v, gamma: float = 4.0 / 3.0, device: int = 0):
    """(x[n, 8], y[n]) of given (conserved u[n, 8], primitive v[n, 8]) pairs,
    e.g. the states of an earlier run."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    n = u.shape[0]
    x = np.empty((n, NF), dtype=np.float32)
    y = np.empty(n, dtype=np.float32)
    _lib.check(_lib.load().gpl_c2p_net_pairs(device, _d(u), _d(v), n, gamma, _f(x), _f(y)),
               "gpl_c2p_net_pairs")
    return x, y


def fit(x, y, steps: int = 3000, batch: int = 4096, lr: float = 3e-3, seed: int = 1, device: int = 0):
    """Train on given (x, y); (weights, last mini-batch loss)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    w = np.empty(size(), dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _lib.check(_lib.load().gpl_c2p_net_fit(device, seed, _f(x), _f(y), len(y), steps, batch, lr,
                                           _f(w), _f(loss)), "gpl_c2p_net_fit")
    return w, float(loss[0])


def train(nsamples: int = 1 << 18, steps: int = 3000, batch: int = 4096, lr: float = 3e-3,
          seed: int = 1, gamma: float = 4.0 / 3.0, device: int = 0):
    w = np.empty(size(), dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _lib.check(_lib.load().gpl_c2p_net_train(device, seed, nsamples, steps, batch, lr, gamma,
                                             _f(w), _f(loss)), "gpl_c2p_net_train")
    return w, float(loss[0])


def grad(w, x, y, device: int = 0):
    w = np.ascontiguousarray(w, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = np.empty(size(), dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _lib.check(_lib.load().gpl_c2p_net_grad(device, _f(w), _f(x), _f(y), len(y), _f(g), _f(loss)),
               "gpl_c2p_net_grad")
    return float(loss[0]), g


def infer(w, x, device: int = 0):
    w = np.ascontiguousarray(w, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape[0], dtype=np.float32)
    _lib.check(_lib.load().gpl_c2p_net_infer(device, _f(w), _f(x), x.shape[0], _f(out)),
               "gpl_c2p_net_infer")
    return out


def solve(u, w=None, gamma: float = 4.0 / 3.0, device: int = 0, restarts: bool = False):
    """Energy solve of every state: (v[n, 8], passes[n], err[n], kernel ms),
    and with ``restarts`` also restart[n] (1 where the learned start failed
    and the A27 restart ran)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    n = u.shape[0]
    v = np.empty((n, 8))
    passes = np.empty(n, dtype=np.int32)
    err = np.empty(n, dtype=np.uint32)
    rs = np.empty(n, dtype=np.int32) if restarts else None
    ms = np.zeros(1)
    wp = None if w is None else _f(np.ascontiguousarray(w, dtype=np.float32))
    _lib.check(_lib.load().gpl_c2p_solve(device, wp, _d(u), n, gamma, _d(v),
                                         passes.ctypes.data_as(C.POINTER(C.c_int)),
                                         err.ctypes.data_as(C.POINTER(C.c_uint32)), _d(ms),
                                         None if rs is None else rs.ctypes.data_as(C.POINTER(C.c_int))),
               "gpl_c2p_solve")
    if restarts:
        return v, passes, err, float(ms[0]), rs
    return v, passes, err, float(ms[0])
