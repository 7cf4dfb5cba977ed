"""Separate contexts are independent (include/gpluto_amd.h): two contexts on
one device, stepped at the same time from two host threads -- each capturing
and replaying its own graph -- give the bits they give when stepped one after
the other."""
import threading

import numpy as np
import pytest

from gpluto_dev_amd import ic
from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate

pytestmark = pytest.mark.gpu


def _ctx(n):
    return StageUpdate(Definitions(), RuntimeSetup(n=n, first_dt=1e-3))


@pytest.mark.parametrize("steps", [7, 9])
def test_two_contexts_in_two_threads(steps):
    n = (48, 40, 32)
    ics = [ic.mhd_turbulence(n, s) for s in (21, 22)]
    ref = []
    for vc, vs in ics:
        g = _ctx(n)
        g.upload(vc, vs)
        ref.append((g.run(steps, 1e-3), g.download()))
        g.close()
    ctxs = [_ctx(n) for _ in ics]
    for g, (vc, vs) in zip(ctxs, ics):
        g.upload(vc, vs)
    out = [None, None]
    fails = []

    def work(q):
        try:
            r = ctxs[q].run(steps, 1e-3)
            out[q] = (r, ctxs[q].download())
        except Exception as e:       # reported below, in the main thread
            fails.append(repr(e))

    th = [threading.Thread(target=work, args=(q,)) for q in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for g in ctxs:
        g.close()
    assert not fails, fails
    for q in range(2):
        (err, t, dt), (v, s, u) = out[q]
        (rerr, rt, rdt), (rv, rs, ru) = ref[q]
        assert err == rerr and t == rt and dt == rdt
        assert np.array_equal(v, rv) and np.array_equal(s, rs) and np.array_equal(u, ru)
