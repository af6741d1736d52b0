"""The C restatement against golden fixtures written by the reference's own solver
(tests/golden/make_golden.py): bit-identical states, max_U grids and per-step scalars
(the volume is not stored: the reference sums it over threads)."""
import json
import os

import numpy as np
import pytest

import pyoracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# single-grid run fixtures written by make_golden.py (nested_*: setup tables, fullsize_*:
# full-size runs; both tested elsewhere)
NAMES = sorted(f[:-4] for f in os.listdir(HERE)
               if f.endswith(".npz") and not f.startswith(("nested_", "fullsize_")))


def load(name):
    z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(oracle_built, name):
    z, meta = load(name)
    d = pyoracle.CpuDomain("oracle", meta["nx"], meta["ny"], meta["lw"], meta["ll"], method=meta["method"],
                           spherical=meta["spherical"], **meta["kw"])
    d.set_U(z["U_in"])
    if "manning_squared" in z:
        d.set_manning_squared(z["manning_squared"])
    if meta.get("stage_bc"):
        b = meta["stage_bc"]
        d.set_stage_bc(b["amp"], b["period"], 0.0, b["sides"])
    keys = json.loads(bytes(z["scalar_keys"]).decode())
    dt = None if meta["method"] == "rk2" else meta["dt"]
    for k in range(meta["nsteps"]):
        d.evolve(dt, 1)
        s = d.scalars()
        assert [s[q] for q in keys] == list(z["scalars"][k]), k
    assert np.array_equal(d.get_U(), z["U_out"])
    for q in pyoracle.MAXQ:
        assert np.array_equal(d.max_U(q), z["max_" + q]), q
