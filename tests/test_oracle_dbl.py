"""The oracle's dump and restart restatement (oracle/dbl_io.py, the oracle's
gplo_set_state_faces) against the reference's own files.

Fixtures: tests/golden/*_restart.npz, written by tests/golden/make_golden.py:
the reference runs `steps` steps from a seeded IC and dumps (data.0001.dbl,
dbl.out, restart.out, kept as bytes), then restarts from that dump
(``-restart 1``) and runs `restart_steps` more.
"""
import os
import struct
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
sys.path.insert(0, HERE)

import dbl_io  # noqa: E402
from golden_io import load, restart_cases  # noqa: E402
from oracle import Oracle  # noqa: E402

CASES = restart_cases()


def _setup(p):
    phys = "rmhd" if p["cfg"].startswith("rmhd") else ("mhd" if p["cfg"].startswith("mhd") else "hd")
    mhd = phys != "hd"
    n = tuple(p["n"])
    dims = 3 if n[2] > 1 else (2 if n[1] > 1 else 1)
    return phys, mhd, n, dims


def _oracle(p, phys):
    return Oracle(phys=phys, n=tuple(p["n"]), gamma=p["gamma"], solver=p["solver"],
                  limiter=p["limiter"], time_stepping=3 if p["time_stepping"] == "RK3" else 2,
                  periodic=p["periodic"], cfl=p["cfl"], cfl_max_var=p["cfl_max_var"],
                  sides=p.get("sides"))


def _faces(vs, upper, dims):
    """Full staggered components (index n along d included)."""
    return [np.concatenate([vs[d], upper[d]], axis=2 - d) for d in range(dims)]


def _records(raw):
    return [raw[k * 256:(k + 1) * 256] for k in range(len(raw) // 256)]


def _write_ref_files(tmp, a):
    for key, name in (("ref_dbl_out", "dbl.out"), ("ref_restart_out", "restart.out"),
                      ("ref_data_mid", "data.0001.dbl")):
        with open(os.path.join(tmp, name), "wb") as f:
            f.write(a[key].tobytes())


def test_restart_record_layout():
    """sizeof(Restart) on x86-64 (Src/structs.hpp:382-390) and its fields'
    offsets: nstep right after param[12], nfile after nstep."""
    rec = dbl_io.pack_restart(1.5, 0.25, 7, 3)
    assert len(rec) == 256
    assert struct.unpack_from("<d", rec, 8)[0] == 0.25
    assert struct.unpack_from("<i", rec, 112)[0] == 7
    assert struct.unpack_from("<16i", rec, 116)[1] == 3
    assert struct.unpack_from("<16i", rec, 116)[0] == -1


@pytest.mark.parametrize("name", CASES)
def test_writer_reproduces_reference_dump(name, tmp_path):
    """Oracle run of `steps` steps, then the oracle's writer: data.0001.dbl,
    dbl.out and the defined bytes of restart.out equal the reference's."""
    p, a = load(name)
    phys, mhd, n, dims = _setup(p)
    o = _oracle(p, phys)
    o.set_state(a["vc_in"], a.get("vs_in"))
    _, t, dt = o.run(p["steps"], p["first_dt"])
    ref_recs = _records(a["ref_restart_out"].tobytes())
    # time and the next dt to the bit (RestartDump keeps them in binary)
    assert struct.unpack_from("<2d", ref_recs[1]) == (t, dt)
    vc, vs, _ = o.get_state()
    np.testing.assert_array_equal(vc, a["vc_mid"])
    faces = None
    if mhd:
        np.testing.assert_array_equal(vs[:dims], a["vs_mid"][:dims])
        faces = _faces(a["vs_mid"], [a[f"vs_upper_mid{d}"] for d in range(dims)], dims)
    # the initial dump (Src/main.cpp:290-291) at t = 0, dt = first_dt, step 0
    ic_faces = None
    if mhd:
        ic_faces = _faces(a["vs_in"], [np.take(a["vs_in"][d], [0], axis=2 - d) for d in range(dims)], dims)
    dbl_io.write_dbl(str(tmp_path), 0, 0.0, p["first_dt"], 0, a["vc_in"], ic_faces, mhd=mhd, dims=dims)
    # the post-loop dump: the step counter reads steps - 1 (Src/main.cpp:336, 482, 497-498)
    dbl_io.write_dbl(str(tmp_path), 1, t, dt, p["steps"] - 1, vc, faces, mhd=mhd, dims=dims)
    with open(tmp_path / "data.0001.dbl", "rb") as f:
        assert f.read() == a["ref_data_mid"].tobytes()
    with open(tmp_path / "dbl.out", "rb") as f:
        assert f.read() == a["ref_dbl_out"].tobytes()
    with open(tmp_path / "restart.out", "rb") as f:
        mine = _records(f.read())
    assert len(mine) == len(ref_recs) == 2
    for m, r in zip(mine, ref_recs):
        # t, dt, param, nstep, nfile; nfileParticles, fill and the padding
        # are never set (Runtime lives on main()'s stack, Src/main.cpp:59)
        assert m[:180] == r[:180]


@pytest.mark.parametrize("name", CASES)
def test_reader_parses_reference_dump(name, tmp_path):
    p, a = load(name)
    phys, mhd, n, dims = _setup(p)
    _write_ref_files(str(tmp_path), a)
    r = dbl_io.read_restart(str(tmp_path), 1)
    # dbl.out prints t with 7 significant digits; the binary t is pinned above
    assert r["nstep"] == p["steps"] - 1 and abs(r["t"] - p["time_mid"]) <= 1e-6 * p["time_mid"]
    vc, faces = dbl_io.read_dbl(str(tmp_path), 1, n, mhd, dims)
    np.testing.assert_array_equal(vc, a["vc_mid"])
    if mhd:
        for d in range(dims):
            np.testing.assert_array_equal(faces[d], _faces(a["vs_mid"], [a[f"vs_upper_mid{e}"] for e in range(dims)], dims)[d])
    with pytest.raises(ValueError):
        dbl_io.read_restart(str(tmp_path), 5)
    with pytest.raises(ValueError):
        dbl_io.read_dbl(str(tmp_path), 1, n, not mhd, dims)


@pytest.mark.parametrize("name", CASES)
def test_oracle_restart_matches_reference(name, tmp_path):
    """RestartFromFile: U = PrimToCons(V) of the dump, every face of the dump
    (the upper boundary face too on a non-periodic side), t, dt and the step
    counter from restart.out; then `restart_steps` steps with the main-loop
    recurrence equal the reference's restarted run to the bit, state, time
    and next dt."""
    p, a = load(name)
    phys, mhd, n, dims = _setup(p)
    _write_ref_files(str(tmp_path), a)
    r = dbl_io.read_restart(str(tmp_path), 1)
    vc, faces = dbl_io.read_dbl(str(tmp_path), 1, n, mhd, dims)
    o = _oracle(p, phys)
    if mhd:
        lower = np.zeros((3,) + vc.shape[1:])
        upper = []
        for d in range(dims):
            sl = [slice(None)] * 3
            sl[2 - d] = slice(0, n[d])
            lower[d] = faces[d][tuple(sl)]
            sl[2 - d] = slice(n[d], n[d] + 1)
            upper.append(faces[d][tuple(sl)])
        o.set_state(vc, lower, upper)
    else:
        o.set_state(vc)
    t, dt = r["t"], r["dt"]
    for _ in range(p["restart_steps"]):
        err, inv = o.advance(dt)
        assert err == 0
        t += dt                                                       # Src/main.cpp:458
        dt_new = p["cfl"] / inv                                       # Src/next_time_step.cpp:80
        dt = dt_new if dt_new <= p["cfl_max_var"] * dt else p["cfl_max_var"] * dt   # :125
    vco, vso, _ = o.get_state()
    np.testing.assert_array_equal(vco, a["vc_out"])
    if mhd:
        np.testing.assert_array_equal(vso[:dims], a["vs_out"][:dims])
    rec2 = _records(a["ref_restart_out_final"].tobytes())[2]
    assert struct.unpack_from("<2d", rec2) == (t, dt)
    assert struct.unpack_from("<i", rec2, 112)[0] == p["steps"] + p["restart_steps"] - 2


def test_restart_needs_the_upper_faces():
    """On the outflow case the dump's upper boundary faces are not the
    periodic image of the lower ones: a restart that wraps them instead ends
    elsewhere."""
    name = "mhd3d_6x8x6_rk3_vl_outflow_restart"
    p, a = load(name)
    for d in range(3):
        lower0 = np.take(a["vs_mid"][d], [0], axis=2 - d)
        assert not np.array_equal(lower0, a[f"vs_upper_mid{d}"])
