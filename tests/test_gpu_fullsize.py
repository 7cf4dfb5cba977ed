"""Size-independent properties at the benchmark sizes, where the oracle is
too slow to run (3D MHD 256^3 FP64 and mixed, 3D RMHD 192^3):

* translation invariance, exact: every update is local and the periodic wrap
  is exact, so an initial condition shifted by whole cells gives the shifted
  result bit for bit (dt included, the maximum does not depend on order);
* conservation of mass, momentum, energy and of the face field's totals under
  periodic sides, to round-off;
* div B at round-off (constrained transport);
* the fused kernels against the per-interface split kernels, bit for bit.
"""
import numpy as np
import pytest

from gpluto_dev_amd import ic
from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate

pytestmark = pytest.mark.gpu

SHIFT = (5, 3, 7)          # cells along x, y, z


def _roll(a, shift=SHIFT):
    """Shift [..., nz, ny, nx] arrays by whole cells, periodically."""
    return np.roll(a, shift=(shift[2], shift[1], shift[0]), axis=(-3, -2, -1))


def _run(physics, n, vc, vs, steps, first_dt, precision="fp64"):
    g = StageUpdate(Definitions(physics=physics),
                    RuntimeSetup(n=n, solver="hll" if physics == "RMHD" else "hlld", first_dt=first_dt,
                                 gamma=4 / 3 if physics == "RMHD" else 5 / 3),
                    precision=precision)
    g.upload(vc, vs)
    _, _, u0 = g.download()
    err, t, dt = g.run(steps, first_dt)
    vco, vso, uco = g.download()
    g.close()
    return err, t, dt, u0, vco, vso, uco


def _totals_conserved(u0, u1, vs0, vs1, nvar_hydro=(0, 1, 2, 3, 7)):
    for q in nvar_hydro:
        a, b = np.sum(u0[q]), np.sum(u1[q])
        scale = np.sum(np.abs(u0[q]))
        assert abs(b - a) <= 1e-12 * scale, f"conserved variable {q}: {a!r} -> {b!r}"
    for d in range(3):
        a, b = np.sum(vs0[d]), np.sum(vs1[d])
        assert abs(b - a) <= 1e-12 * np.sum(np.abs(vs0[d])), f"face field {d}: {a!r} -> {b!r}"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_mhd256_translation_invariance_conservation_divb(precision):
    n, steps, first_dt = (256, 256, 256), 3, 5e-4
    print(f"case 3D MHD 256^3 {precision}", flush=True)
    vc, vs = ic.mhd_turbulence(n, 1)
    err, t, dt, u0, vco, vso, uco = _run("MHD", n, vc, vs, steps, first_dt, precision)
    assert err == 0
    _totals_conserved(u0, uco, vs, vso)
    assert ic.divb_max(vso) <= 1e-11 * np.max(np.abs(vso)) * n[0]
    errs, ts, dts, _, vcs, vss, _ = _run("MHD", n, _roll(vc), _roll(vs), steps, first_dt, precision)
    assert errs == 0 and ts == t and dts == dt
    assert np.array_equal(vcs, _roll(vco)), "primitives are not shifted bit for bit"
    assert np.array_equal(vss, _roll(vso)), "faces are not shifted bit for bit"


def test_mhd256_fused_equals_split(monkeypatch):
    n, steps, first_dt = (256, 256, 256), 2, 5e-4
    print("case 3D MHD 256^3 fused vs split", flush=True)
    vc, vs = ic.mhd_turbulence(n, 2)
    err, t, dt, _, vcf, vsf, ucf = _run("MHD", n, vc, vs, steps, first_dt)
    monkeypatch.setenv("GPL_SPLIT_KERNELS", "1")
    errs, ts, dts, _, vcs, vss, ucs = _run("MHD", n, vc, vs, steps, first_dt)
    assert err == errs == 0 and t == ts and dt == dts
    assert np.array_equal(vcf, vcs) and np.array_equal(vsf, vss) and np.array_equal(ucf, ucs)


def test_rmhd192_translation_invariance_conservation_divb():
    n, steps, first_dt = (192, 192, 192), 3, 1e-3
    print("case 3D RMHD 192^3", flush=True)
    # the benchmark's 192^3 IC: Lorentz factors up to 10, sigma log-uniform by seed
    sigma = float(10.0 ** np.random.Generator(np.random.PCG64(1 + 1000)).uniform(-1.0, 1.0))
    vc, vs = ic.rmhd_turbulence(n, 1, lorentz_max=10.0, sigma=sigma)
    err, t, dt, u0, vco, vso, uco = _run("RMHD", n, vc, vs, steps, first_dt)
    assert err == 0            # no recovery branch rewrites a zone: the totals must hold
    _totals_conserved(u0, uco, vs, vso)
    assert ic.divb_max(vso) <= 1e-11 * np.max(np.abs(vso)) * n[0]
    errs, ts, dts, _, vcs, vss, _ = _run("RMHD", n, _roll(vc), _roll(vs), steps, first_dt)
    assert errs == err and ts == t and dts == dt
    assert np.array_equal(vcs, _roll(vco)) and np.array_equal(vss, _roll(vso))
