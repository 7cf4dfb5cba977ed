"""The OpenMP oracle gives the same bits for any thread count.

Its parallel loops split pencils, planes and cells among threads and reduce
only with max, bitwise or and integer sums, so no result may depend on the
thread count.  Each case runs in a fresh process with OMP_NUM_THREADS set
(OpenMP reads it once per process) and writes its state, time, dt, error bits
and zone reports (the HD case empties its centre and reaches the floors); the runs with 1, 3 and 8 threads must agree byte for byte.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
import numpy as np
sys.path[:0] = [sys.argv[1], sys.argv[1] + "/oracle"]
import oracle
from gpluto_dev_amd import ic
phys, solver, out = sys.argv[2], sys.argv[3], sys.argv[4]
if phys == "mhd":
    n = (18, 14, 22)
    vc, vs = ic.mhd_turbulence(n, 5)
    o = oracle.Oracle(phys="mhd", n=n, gamma=5 / 3, solver=solver, time_stepping=3, limiter="VANLEER_LIM")
elif phys == "rmhd":
    n = (14, 18, 12)
    vc, vs = ic.rmhd_turbulence(n, 12, lorentz_max=5.0, sigma=3.0)
    o = oracle.Oracle(phys="rmhd", n=n, gamma=4 / 3, solver=solver, time_stepping=2)
else:
    n = (48, 32, 1)
    vc, vs = ic.hd_vacuum(n, 20.0, 0.01), None
    o = oracle.Oracle(phys="hd", n=n, gamma=1.4, solver=solver, time_stepping=3,
                      limiter="VANLEER_LIM", periodic=False)
o.set_state(vc, vs)
err, t, dt = o.run(40, 1e-4) if phys == "hd" else o.run(6, 1e-3)
v, s, u = o.get_state()
zerr, zcount, ztotal = o.zone_errors()
np.savez(out, v=v, s=np.asarray(s if s is not None else []), u=u, t=t, dt=dt, err=err,
         zerr=zerr, zcount=zcount, ztotal=ztotal)
"""


def _run(tmp_path, phys, solver, threads):
    out = str(tmp_path / f"{phys}_{solver}_{threads}.npz")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run([sys.executable, "-c", SCRIPT, ROOT, phys, solver, out], check=True, env=env,
                   timeout=300)
    with np.load(out) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("phys,solver", [("mhd", "hlld"), ("rmhd", "hllc"), ("hd", "hllc")])
def test_oracle_bits_do_not_depend_on_thread_count(tmp_path, phys, solver):
    runs = [_run(tmp_path, phys, solver, t) for t in (1, 3, 8)]
    if phys == "hd":
        assert runs[0]["ztotal"] > 0, "the HD case is meant to reach the pressure floor"
    for r in runs[1:]:
        for k in runs[0]:
            assert runs[0][k].tobytes() == r[k].tobytes(), (phys, k)
