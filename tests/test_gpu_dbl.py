"""The device dump writer and restart reader (gpl_write_dbl, gpl_read_dbl)
through the C ABI, against the reference's own files and its restarted runs
(tests/golden/*_restart.npz, written by tests/golden/make_golden.py), and
size-independent properties at a larger box.  The checker that parses files
is the oracle's restatement (oracle/dbl_io.py)."""
import os
import struct
import sys

import numpy as np
import pytest

from golden_io import load, restart_cases

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import dbl_io  # noqa: E402

# every restart fixture: 1D HD, 2D MHD (one box of the batched kernels behind
# gpl_create), 3D MHD and RMHD
CASES = restart_cases()


def _setup(p):
    phys = "RMHD" if p["cfg"].startswith("rmhd") else ("MHD" if p["cfg"].startswith("mhd") else "HD")
    n = tuple(p["n"])
    dims = 3 if n[2] > 1 else (2 if n[1] > 1 else 1)
    return phys, n, dims


def _stage(p, precision="fp64"):
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    phys, n, dims = _setup(p)
    return StageUpdate(Definitions(physics=phys, dimensions=dims, time_stepping=p["time_stepping"],
                                   limiter=p["limiter"]),
                       RuntimeSetup(n=n, solver=p["solver"], cfl=p["cfl"], first_dt=p["first_dt"],
                                    gamma=p["gamma"],
                                    boundary="periodic" if p["periodic"] else "outflow",
                                    sides=tuple(p.get("sides") or ())),
                       precision=precision)


def _records(raw):
    return [raw[k * 256:(k + 1) * 256] for k in range(len(raw) // 256)]


# RMHD's bar is a relative L1 bound, not bit equality: the device takes its
# quartic's acos, cos and sin correctly rounded, and the reference's host libm
# misrounds a few in a thousand of them (tests/test_gpu_rmhd.py)
RMHD_TOL = 1e-12


def _close_l1(mine, ref, label):
    for v in range(len(ref)):
        den = np.sum(np.abs(ref[v]))
        l1 = np.sum(np.abs(mine[v] - ref[v])) / den if den > 0 else np.max(np.abs(mine[v] - ref[v]))
        assert l1 <= RMHD_TOL, f"{label} {v}: rel L1 {l1:.3e}"


def _same_records(mine, ref, exact):
    """t, dt, param, nstep, nfile: the members the reference sets."""
    assert len(mine) == len(ref)
    for m, r in zip(mine, ref):
        if exact:
            assert m[:180] == r[:180]
        else:
            assert m[16:180] == r[16:180]
            for a, b in zip(struct.unpack_from("<2d", m), struct.unpack_from("<2d", r)):
                assert abs(a - b) <= 1e-12 * abs(b)


def _read(path):
    with open(path, "rb") as f:
        return f.read()


def _write_ref_files(tmp, a):
    for key, name in (("ref_dbl_out", "dbl.out"), ("ref_restart_out", "restart.out"),
                      ("ref_data_mid", "data.0001.dbl")):
        with open(os.path.join(tmp, name), "wb") as f:
            f.write(a[key].tobytes())


@pytest.mark.parametrize("name", CASES)
def test_dump_matches_reference_files(name, tmp_path):
    """Initial dump at t = 0, dt = first_dt, step 0; `steps` steps on the
    device; the post-loop dump with the reference's step counter: the data
    file, dbl.out and restart.out's set bytes equal the reference's."""
    p, a = load(name)
    print(f"case {name}", flush=True)
    phys, n, dims = _setup(p)
    g = _stage(p)
    g.upload(a["vc_in"], a.get("vs_in"))
    g.set_step(0.0, p["first_dt"], 0)
    g.write_dbl(str(tmp_path), 0)
    err, t, dt = g.run(p["steps"], 0.0)          # continue from the step state set above
    assert err == 0
    assert g.get_step() == (t, dt, p["steps"])
    g.write_dbl(str(tmp_path), 1, nstep=p["steps"] - 1)
    g.close()
    if phys != "RMHD":
        assert _read(tmp_path / "data.0001.dbl") == a["ref_data_mid"].tobytes()
    else:
        ref_dir = tmp_path / "ref"
        ref_dir.mkdir()
        _write_ref_files(str(ref_dir), a)
        mc, mf = dbl_io.read_dbl(str(tmp_path), 1, n, True, dims)
        rc, rf = dbl_io.read_dbl(str(ref_dir), 1, n, True, dims)
        _close_l1(mc, rc, "var")
        _close_l1(mf, rf, "face")
    assert _read(tmp_path / "dbl.out") == a["ref_dbl_out"].tobytes()
    mine, ref = _records(_read(tmp_path / "restart.out")), _records(a["ref_restart_out"].tobytes())
    assert len(mine) == 2
    _same_records(mine, ref, phys != "RMHD")


@pytest.mark.parametrize("name", CASES)
def test_restart_matches_reference_restart(name, tmp_path):
    """Restart from the reference's own dump files, `restart_steps` more
    steps: state (the upper boundary faces included), time and dt equal the
    reference's restarted run to the bit; the dump it then writes extends the
    reference's dbl.out and restart.out exactly as the reference did."""
    p, a = load(name)
    print(f"case {name}", flush=True)
    phys, n, dims = _setup(p)
    _write_ref_files(str(tmp_path), a)
    ref1 = dbl_io.unpack_restart(_records(a["ref_restart_out"].tobytes())[1])
    g = _stage(p)
    g.read_dbl(str(tmp_path), 1)
    assert g.get_step() == (ref1["t"], ref1["dt"], p["steps"] - 1)
    err, t, dt = g.run(p["restart_steps"], 0.0)
    assert err == 0
    vc, vs, _ = g.download()
    up = g.upper_faces() if phys != "HD" else None
    # the reference's counter after a restart reads steps - 1 + restart_steps - 1
    g.write_dbl(str(tmp_path), 2, nstep=p["steps"] + p["restart_steps"] - 2)
    g.close()
    exact = phys != "RMHD"
    if exact:
        for v in range(vc.shape[0]):
            assert np.array_equal(vc[v], a["vc_out"][v]), f"var {v}: max diff " \
                f"{np.max(np.abs(vc[v] - a['vc_out'][v])):.3e}"
    else:
        _close_l1(vc, a["vc_out"], "var")
    if phys != "HD":
        if exact:
            assert np.array_equal(vs[:dims], a["vs_out"][:dims])
            for d in range(dims):
                assert np.array_equal(up[d], a[f"vs_upper{d}"]), f"upper face {d}"
        else:
            _close_l1(vs[:dims], a["vs_out"][:dims], "face")
    final = _records(a["ref_restart_out_final"].tobytes())
    if exact:
        assert struct.unpack_from("<2d", final[2]) == (t, dt)
    assert _read(tmp_path / "dbl.out") == a["ref_dbl_out_final"].tobytes()
    mine = _records(_read(tmp_path / "restart.out"))
    assert len(mine) == 3
    _same_records(mine, final, exact)


@pytest.mark.parametrize("name", [c for c in CASES if c.startswith("mhd3d")])
def test_mixed_restart_from_reference_dump(name, tmp_path):
    """The mixed form restarts from an FP64 dump (its FP32 state is the
    rounded primitives) and stays within its stated accuracy of the
    reference's restarted run."""
    p, a = load(name)
    print(f"case {name}", flush=True)
    _write_ref_files(str(tmp_path), a)
    g = _stage(p, precision="mixed")
    g.read_dbl(str(tmp_path), 1)
    err, _, _ = g.run(p["restart_steps"], 0.0)
    vc, vs, _ = g.download()
    g.close()
    assert err == 0
    ref = a["vc_out"]
    for v in range(8):
        rel = np.sum(np.abs(vc[v] - ref[v])) / max(np.sum(np.abs(ref[v])), 1e-300)
        assert rel <= 1e-5, f"var {v}: relative L1 {rel:.3e}"


def _mhd_box(n, seed, sides=None, precision="fp64"):
    from gpluto_dev_amd import ic
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    vc, vs = ic.mhd_turbulence(n, seed)
    g = StageUpdate(Definitions(), RuntimeSetup(n=n, first_dt=1e-3, boundary="periodic",
                                                sides=tuple(sides or ())), precision=precision)
    g.upload(vc, vs)
    return g


@pytest.mark.parametrize("sides", [None, ("outflow",) * 2 + ("reflective",) * 2 + ("periodic",) * 2])
def test_dump_round_trip_128(sides, tmp_path):
    """At 128^3: the dump holds exactly the downloaded state, the faces 0..n
    along each direction included; a context restarted from it writes the
    same bytes back (read and write are inverse on an FP64 state), and
    resumes with the dumped time, dt and step counter."""
    n = (128, 128, 128)
    print(f"round trip 128^3 sides={sides}", flush=True)
    g = _mhd_box(n, 21, sides)
    err, t, dt = g.run(4, 1e-3)
    assert err == 0
    g.write_dbl(str(tmp_path), 0)
    vc, vs, _ = g.download()
    up = g.upper_faces()
    g.close()
    fvc, faces = dbl_io.read_dbl(str(tmp_path), 0, n, True, 3)
    assert np.array_equal(fvc, vc)
    for d in range(3):
        assert np.array_equal(np.take(faces[d], range(n[d]), axis=2 - d), vs[d])
        assert np.array_equal(np.take(faces[d], [n[d]], axis=2 - d), up[d])
    assert dbl_io.read_restart(str(tmp_path), 0)["nstep"] == 4
    first = _read(tmp_path / "data.0000.dbl")
    h = _mhd_box(n, 99, sides)              # another state, replaced by the restart
    h.read_dbl(str(tmp_path), 0)
    assert h.get_step() == (t, dt, 4)
    h.write_dbl(str(tmp_path), 1)
    h.close()
    assert _read(tmp_path / "data.0001.dbl") == first


def test_restart_continues_within_roundoff(tmp_path):
    """A run restarted from its own dump rebuilds Uc = PrimToCons(Vc), as
    the reference does, so it continues within round-off of the run that
    kept its Uc."""
    n = (64, 64, 64)
    a = _mhd_box(n, 5)
    a.run(3, 1e-3)
    a.write_dbl(str(tmp_path), 0)
    a.run(3, 0.0)
    va, sa, _ = a.download()
    a.close()
    b = _mhd_box(n, 6)
    b.read_dbl(str(tmp_path), 0)
    b.run(3, 0.0)
    vb, sb, _ = b.download()
    b.close()
    for v in range(8):
        scale = np.max(np.abs(va[v])) + 1e-300
        assert np.max(np.abs(va[v] - vb[v])) <= 1e-11 * scale, f"var {v}"
    assert np.max(np.abs(sa - sb)) <= 1e-11 * np.max(np.abs(sa))


def test_restart_rejects_foreign_or_missing_dumps(tmp_path):
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    g = _mhd_box((16, 16, 16), 3)
    with pytest.raises(RuntimeError, match="file error"):
        g.read_dbl(str(tmp_path / "absent"), 0)
    g.write_dbl(str(tmp_path), 0)
    with pytest.raises(RuntimeError, match="invalid argument"):
        g.read_dbl(str(tmp_path), 1)                 # no dump 1 in dbl.out
    other = StageUpdate(Definitions(), RuntimeSetup(n=(16, 16, 8)))
    with pytest.raises(RuntimeError, match="invalid argument"):
        other.read_dbl(str(tmp_path), 0)             # 16^3 dump, 16x16x8 box: wrong size
    other.close()
    hd = StageUpdate(Definitions(physics="HD", dimensions=3), RuntimeSetup(n=(16, 16, 16), solver="hllc",
                                                                           gamma=1.4))
    with pytest.raises(RuntimeError, match="invalid argument"):
        hd.read_dbl(str(tmp_path), 0)                # MHD variables in an HD context
    hd.close()
    g.close()


@pytest.mark.parametrize("kind", ["mhd", "mhd-sides", "mhd-mixed", "hd"])
def test_async_dump_snapshots_before_later_steps(kind, tmp_path):
    """gpl_write_dbl_async takes its snapshot behind the work already
    enqueued and writes while the next steps run: the files equal a
    synchronous dump of the same state, not the state the run reaches."""
    from gpluto_dev_amd import ic
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    print(f"async dump {kind}", flush=True)
    if kind == "hd":
        n = (4096, 1, 1)
        g = StageUpdate(Definitions(physics="HD", dimensions=1),
                        RuntimeSetup(n=n, solver="hllc", cfl=0.8, gamma=1.4, boundary="outflow"))
        g.upload(ic.hd_riemann(n[0], 3))
        mhd, first_dt = False, 1e-4
    else:
        n = (48, 40, 36)
        sides = ("outflow", "reflective", "periodic", "periodic", "reflective", "outflow") \
            if kind == "mhd-sides" else ()
        g = StageUpdate(Definitions(), RuntimeSetup(n=n, first_dt=1e-3, sides=sides),
                        precision="mixed" if kind == "mhd-mixed" else "fp64")
        g.upload(*ic.mhd_turbulence(n, 8))
        mhd, first_dt = True, 1e-3
    g.run(3, first_dt)
    g.write_dbl(str(tmp_path), 0)
    g.write_dbl_async(str(tmp_path), 1)
    err, _, _ = g.run(5, 0.0)
    g.wait_dbl()
    assert err == 0
    assert _read(tmp_path / "data.0001.dbl") == _read(tmp_path / "data.0000.dbl")
    lines = _read(tmp_path / "dbl.out").decode().splitlines()
    assert lines[1].split()[1:] == lines[0].split()[1:]          # same t, dt, step, names
    g.write_dbl(str(tmp_path), 2)
    vc, vs, _ = g.download()
    up = g.upper_faces() if mhd else None
    g.close()
    fvc, faces = dbl_io.read_dbl(str(tmp_path), 2, n, mhd, 3 if mhd else 1)
    assert np.array_equal(fvc, vc)
    if mhd:
        for d in range(3):
            assert np.array_equal(np.take(faces[d], range(n[d]), axis=2 - d), vs[d])
            assert np.array_equal(np.take(faces[d], [n[d]], axis=2 - d), up[d])
    assert dbl_io.read_restart(str(tmp_path), 2)["nstep"] == 8
