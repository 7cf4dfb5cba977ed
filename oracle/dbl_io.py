"""CPU restatement of the reference's double-precision output and restart
files.  Test infrastructure only: the tests use it to write and parse the
files the device writer and reader exchange with the reference.

Three files per output directory (single-file mode, the serial build):

* ``data.NNNN.dbl`` -- each dumped variable's active zones in turn, i
  fastest, native-endian doubles (Src/write_data.cpp:136-159 with
  FillOutputBuffer :564-607).  The cell-centred variables come first in
  NVAR order, then, with a staggered field, Bx1s..Bx<D>s (Src/set_output.cpp:
  174-193), each carrying one extra face along its own direction: file index
  0 is the lower boundary face, index n the upper one (:594-596).
* ``dbl.out`` -- one line per dump, ``"%d %12.6e %12.6e %d "`` of nfile,
  time, dt and the step counter, then ``single_file``, the endianness and
  the variable names, each followed by a space (Src/write_data.cpp:476-505).
  Line nfile is written after the first nfile lines, which are kept
  (:478-484).
* ``restart.out`` -- one ``Restart`` record per dump (Src/structs.hpp:382-390,
  Src/restart.cpp:417-453): t, dt, param[12], nstep, nfile[16],
  nfileParticles[16], fill[10], 256 bytes with the trailing padding.  The
  output table is indexed by type, DBL_OUTPUT = 1 (Src/runtime_setup.cpp:252,
  Src/gpluto.hpp:139), and every type not yet written holds nfile = -1
  (Src/set_output.cpp:106).  A restart scans the records for the one whose
  nfile[DBL_OUTPUT] is the requested file (Src/restart.cpp:356-378) and takes
  t, dt and nstep from it (:388-392).
"""
from __future__ import annotations

import os
import struct

import numpy as np

DBL_OUTPUT = 1                 # Src/gpluto.hpp:139
MAX_OUTPUT_TYPES = 16          # Src/gpluto.hpp:184
MAX_RESTART_PARAMS = 12        # Src/gpluto.hpp:186
# t, dt, param[12], nstep, nfile[16], nfileParticles[16], fill[10]; the
# doubles make the record 8-byte aligned: 254 bytes of members, 256 in all
RESTART_FMT = "<2d12di16i16i10s2x"
RESTART_SIZE = struct.calcsize(RESTART_FMT)
assert RESTART_SIZE == 256

HD_NAMES = ["rho", "vx1", "vx2", "vx3", "prs"]                          # Src/set_var_names.cpp
MHD_NAMES = ["rho", "vx1", "vx2", "vx3", "Bx1", "Bx2", "Bx3", "prs"]


def var_names(mhd: bool, dims: int) -> list[str]:
    """Dumped names in file order (Src/set_output.cpp:129-193)."""
    if not mhd:
        return list(HD_NAMES)
    return MHD_NAMES + [f"Bx{d + 1}s" for d in range(dims)]


def dbl_out_line(nfile: int, time: float, dt: float, nstep: int, names) -> str:
    """One line of dbl.out (Src/write_data.cpp:488-505)."""
    s = "%d %12.6e %12.6e %d " % (nfile, time, dt, nstep)
    s += "single_file " + "little "
    for n in names:
        s += n + " "
    return s + "\n"


def pack_restart(time: float, dt: float, nstep: int, nfile: int,
                 params=None) -> bytes:
    """The Restart record a dbl dump appends (Src/restart.cpp:425-437);
    nfileParticles, fill and the padding, which the reference never sets
    (its Runtime lives uninitialised on main()'s stack, Src/main.cpp:59), as
    zeros."""
    nf = [-1] * MAX_OUTPUT_TYPES
    nf[DBL_OUTPUT] = nfile
    par = list(params) if params is not None else [0.0] * MAX_RESTART_PARAMS
    return struct.pack(RESTART_FMT, time, dt, *par, nstep, *nf, *([0] * MAX_OUTPUT_TYPES), b"\0" * 10)


def unpack_restart(rec: bytes) -> dict:
    f = struct.unpack(RESTART_FMT, rec)
    return {"t": f[0], "dt": f[1], "param": list(f[2:14]), "nstep": f[14],
            "nfile": list(f[15:31])}


def file_blocks(vc: np.ndarray, faces=None) -> list[np.ndarray]:
    """The arrays a dump holds, in file order: vc[NVAR, nz, ny, nx] and the
    staggered components faces[d] with shape n + 1 along d."""
    blocks = [vc[q] for q in range(vc.shape[0])]
    if faces is not None:
        blocks += list(faces)
    return blocks


def write_dbl(directory: str, nfile: int, time: float, dt: float, nstep: int,
              vc: np.ndarray, faces=None, mhd: bool | None = None, dims: int = 3):
    """WriteData(DBL_OUTPUT) followed by RestartDump (Src/main.cpp:673-687)."""
    mhd = faces is not None if mhd is None else mhd
    names = var_names(mhd, dims)
    raw = b"".join(np.ascontiguousarray(b, dtype="<f8").tobytes() for b in file_blocks(vc, faces))
    with open(os.path.join(directory, "data.%04d.dbl" % nfile), "wb") as f:
        f.write(raw)
    path = os.path.join(directory, "dbl.out")
    line = dbl_out_line(nfile, time, dt, nstep, names).encode()
    old = b""
    if nfile > 0:
        with open(path, "rb") as f:
            old = f.read()
    # skip nfile lines, then overwrite in place (fopen "r+"): later bytes stay
    pos = 0
    for _ in range(nfile):
        e = old.find(b"\n", pos)
        pos = len(old) if e < 0 else e + 1
    with open(path, "wb") as f:
        f.write(old[:pos] + line + old[pos + len(line):])
    rpath = os.path.join(directory, "restart.out")
    rec = pack_restart(time, dt, nstep, nfile)
    mode = "wb" if nfile == 0 or not os.path.exists(rpath) else "r+b"
    with open(rpath, mode) as f:
        f.seek(nfile * RESTART_SIZE)
        f.write(rec)


def read_restart(directory: str, nfile: int) -> dict:
    """RestartGet (Src/restart.cpp:342-392): the first record naming nfile."""
    if nfile < 0:
        raise ValueError("negative restart file")
    with open(os.path.join(directory, "restart.out"), "rb") as f:
        raw = f.read()
    for k in range(len(raw) // RESTART_SIZE):
        r = unpack_restart(raw[k * RESTART_SIZE:(k + 1) * RESTART_SIZE])
        if r["nfile"][DBL_OUTPUT] == nfile:
            return r
    raise ValueError(f"restart.out holds no record of file {nfile}")


def read_dbl(directory: str, nfile: int, n, mhd: bool, dims: int):
    """RestartFromFile's data read (Src/restart.cpp:124-172): (vc, faces)
    with faces[d] of shape n + 1 along d (None for HD)."""
    with open(os.path.join(directory, "dbl.out")) as f:
        lines = f.read().splitlines()
    if nfile > len(lines) - 1:
        raise ValueError(f"output #{nfile} does not exist in dbl.out")
    tok = lines[nfile].split()
    if tok[4] != "single_file":
        raise ValueError("only single_file dumps are read")
    dt = "<f8" if tok[5] == "little" else ">f8"
    names = var_names(mhd, dims)
    if tok[6:] != names:
        raise ValueError(f"dbl.out names {tok[6:]}, expected {names}")
    nx, ny, nz = n
    raw = np.fromfile(os.path.join(directory, "data.%04d.dbl" % nfile), dtype=dt).astype(np.float64)
    nvar = 8 if mhd else 5
    off = 0
    cells = raw[off:off + nvar * nx * ny * nz].reshape(nvar, nz, ny, nx)
    off += cells.size
    faces = None
    if mhd:
        faces = []
        for d in range(dims):
            shape = [nz, ny, nx]
            shape[2 - d] += 1
            cnt = shape[0] * shape[1] * shape[2]
            faces.append(raw[off:off + cnt].reshape(shape))
            off += cnt
    if off != raw.size:
        raise ValueError(f"data.{nfile:04d}.dbl holds {raw.size} doubles, expected {off}")
    return np.ascontiguousarray(cells), faces
