"""Load the committed reference fixtures (tests/golden/*.npz)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _all(prefix):
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz") and f.startswith(prefix))


def cases(prefix=""):
    """Fixtures of a plain run from the initial condition."""
    return [c for c in _all(prefix) if not c.endswith("_restart")]


def restart_cases(prefix=""):
    """Fixtures of a run, its dump and a restart from that dump."""
    return [c for c in _all(prefix) if c.endswith("_restart")]


def load(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    p = json.loads(str(d["params"]))
    arrays = {k: d[k] for k in d.files if k != "params"}
    return p, arrays


def expected_err(arrays) -> int:
    """The OR of the ERR_* bits the reference's run raised: from its
    zone-by-zone log where the fixture holds one (Where, WARNING_MESSAGES 3),
    else 0 (the fixture's run logged no error)."""
    z = arrays.get("zones")
    if z is None or len(z) == 0:
        return 0
    out = 0
    for b in z[:, 5]:
        out |= int(b)
    return out
