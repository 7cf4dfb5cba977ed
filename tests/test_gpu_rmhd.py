"""GPU parity of the relativistic MHD stage update (PLM + velocity limiter +
HLL with exact quartic fast speeds + UCT_HLL CT + energy-solve con2prim with
pressure fix), FP64.

The oracle is bitwise equal to the reference build on these fixtures
(tests/test_oracle_golden.py).  The GPU follows the same operation order,
but the quartic's resolvent cubic calls acos, cos and sin, whose last bit
differs between the host's libm and the device's math library; the bar is
therefore a relative L1 bound per variable (and the step's time), with the
bitwise share of cells reported.  tests/test_host_physics.py shows the same
device source built for the CPU is bitwise equal to the oracle, so what is
left is the device math library; its ulp-level differences grow most in the
pressure of strongly magnetised, fast cells (measured 2.2e-12 on the
Gamma_max 5, sigma 3 fixture, <= 1e-12 on the others)."""
import numpy as np
import pytest

from golden_io import cases, load
from gpluto_dev_amd import ic
from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate

pytestmark = pytest.mark.gpu

TOL = 5e-12
# HLLD solves for the total pressure by a secant iteration whose stop test
# (|dp| < 1e-5 p) and give-up tests see the quartic speeds: a last-bit
# difference can end the iteration one pass earlier or later at a face, so the
# differences grow further (measured 5.8e-12 on the Gamma_max 5, sigma 3
# reference fixture and 4.0e-11, the largest variable, on the same IC as the
# HLL case below, whose HLL and HLLC runs stay at <= 3e-14; the
# host build of the same source is bitwise the oracle,
# tests/test_host_physics.py)
TOL_SOLVER = {"hll": TOL, "hllc": TOL, "hlld": 5e-11}


def _stage(n, ts, lim, cfl=0.4, gamma=4 / 3, periodic=True, solver="hll", sides=None):
    return StageUpdate(Definitions(physics="RMHD", time_stepping=ts, limiter=lim),
                       RuntimeSetup(n=n, solver=solver, cfl=cfl, gamma=gamma,
                                    boundary="periodic" if periodic else "outflow",
                                    sides=tuple(sides) if sides else ()))


def _compare(gv, gs, ov, osv, label, tol=TOL):
    same = 0
    for v in range(8):
        den = np.sum(np.abs(ov[v]))
        l1 = np.sum(np.abs(gv[v] - ov[v])) / den if den > 0 else np.max(np.abs(gv[v] - ov[v]))
        same += int(np.count_nonzero(gv[v] == ov[v]))
        assert l1 <= tol, f"{label} var {v}: rel L1 {l1:.3e}"
    for d in range(gs.shape[0]):
        l1 = np.sum(np.abs(gs[d] - osv[d])) / np.sum(np.abs(osv[d]))
        assert l1 <= tol, f"{label} face {d}: rel L1 {l1:.3e}"
    worst = max(np.sum(np.abs(gv[v] - ov[v])) / max(np.sum(np.abs(ov[v])), 1e-300) for v in range(8))
    print(f"{label}: {same / (8 * ov[0].size):.4f} of primitive values bitwise, "
          f"max rel L1 {worst:.2e}", flush=True)


@pytest.mark.parametrize("name", cases("rmhd"))
def test_gpu_matches_reference_fixture(name):
    p, a = load(name)
    print(f"case {name}", flush=True)
    g = _stage(tuple(p["n"]), p["time_stepping"], p["limiter"], p["cfl"], p["gamma"], p["periodic"],
               p["solver"], p.get("sides"))
    g.upload(a["vc_in"], a["vs_in"])
    err, t, _ = g.run(p["steps"], p["first_dt"])
    vc, vs, _ = g.download()
    g.close()
    assert err & (1 << 15) == 0 and (err == 0) == (not p["err_log"])
    assert abs(t - p["time"]) <= 1e-6 * abs(p["time"])
    _compare(vc, vs, a["vc_out"], a["vs_out"], name, TOL_SOLVER[p["solver"]])


@pytest.mark.parametrize("n,seed,lmax,sigma,steps,ts,lim,solver", [
    ((24, 24, 24), 11, 4.0, 2.0, 8, "RK2", "MC_LIM", "hll"),
    ((20, 16, 28), 12, 5.0, 3.0, 6, "RK3", "VANLEER_LIM", "hll"),
    ((24, 20, 16), 13, 4.0, 2.0, 8, "RK2", "MC_LIM", "hllc"),
    ((20, 16, 28), 12, 5.0, 3.0, 6, "RK3", "VANLEER_LIM", "hllc"),   # the HLL case's IC
    ((24, 24, 24), 11, 4.0, 2.0, 8, "RK2", "MC_LIM", "hlld"),
    ((20, 16, 28), 12, 5.0, 3.0, 6, "RK3", "VANLEER_LIM", "hlld"),
])
def test_gpu_matches_oracle(n, seed, lmax, sigma, steps, ts, lim, solver):
    import oracle
    print(f"case n={n} seed={seed} Gamma_max={lmax} sigma={sigma} {ts} {lim} {solver}", flush=True)
    vc, vs = ic.rmhd_turbulence(n, seed, lorentz_max=lmax, sigma=sigma)
    o = oracle.Oracle(phys="rmhd", n=n, gamma=4 / 3, solver=solver, limiter=lim,
                      time_stepping=int(ts[-1]), cfl=0.4)
    o.set_state(vc, vs)
    oerr, ot, odt = o.run(steps, 1e-3)
    ov, osv, _ = o.get_state()
    g = _stage(n, ts, lim, solver=solver)
    g.upload(vc, vs)
    gerr, gt, gdt = g.run(steps, 1e-3)
    gv, gs, _ = g.download()
    g.close()
    assert gerr == oerr
    # a run the reference would stop (ERR_QUIT) is no parity case: past its
    # recovery branches a last-bit difference flips discrete choices
    assert oerr & (1 << 15) == 0
    assert abs(gt - ot) <= 1e-12 * ot and abs(gdt - odt) <= 1e-12 * odt
    _compare(gv, gs, ov, osv, f"n={n} {solver}", TOL_SOLVER[solver])


def test_rmhd_rejects_unsupported():
    with pytest.raises(RuntimeError, match="not supported"):
        StageUpdate(Definitions(physics="RMHD"), RuntimeSetup(n=(16, 16, 16), solver="hll"),
                    precision="mixed")
