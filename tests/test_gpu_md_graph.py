"""GPU: the multidomain global step replayed from HIP graphs (swh_md_graph_capture /
swh_md_graph_launch) against eager global steps (swh_md_evolve_one_step), on the
three-level nest of the full-size nested workload ('linear' parent 3000 x 1800 taking
one sub-step per global step -- so its buffers alternate and two graphs are captured --,
'leapfrog_linear_plus_nonlinear_friction' child taking two, 'rk2' grandchild taking ten,
max_stage tracked on every domain).

Replays equal the eager steps bit for bit: state, max-stage grids and every domain's
time, with eager steps mixed in between replays (which moves the buffers' parity, so the
other graph continues). Capture is refused before the first step (max_U would be
recorded twice) and leaves the model where it was.
"""
import json
import os

import numpy as np
import pytest

from ptha_amd.backend import BackendError
from test_gpu_fullsize import GOLDEN, build_config3

pytestmark = pytest.mark.gpu


def _tables():
    g = np.load(os.path.join(GOLDEN, "nested_c3_tables.npz"))
    return g, json.loads(bytes(g["tables_json"]).decode())


def _snapshot(doms, md):
    md.synchronize()
    return ([d.U.copy() for d in doms], [d.max_U("max_stage").copy() for d in doms], [d.time for d in doms])


def _close(doms, md):
    md.close()
    for d in doms:
        d.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32k"])
def test_graph_replay_equals_eager_steps(precision):
    print("CASE multidomain graph replay", precision, flush=True)
    g, tables = _tables()
    dt = float(tables["dt"])
    plan = [("eager", 1), ("graph", 2), ("eager", 1), ("graph", 3), ("eager", 2), ("graph", 1)]
    total = sum(n for _, n in plan)
    doms, md = build_config3(precision, g, tables)
    md.evolve_steps(dt, total)
    want = _snapshot(doms, md)
    _close(doms, md)

    doms, md = build_config3(precision, g, tables)
    with pytest.raises(BackendError, match="first step"):
        md.graph_capture(dt)
    captured = False
    for kind, n in plan:
        if kind == "eager":
            md.evolve_steps(dt, n)
        else:
            if not captured:
                before = _snapshot(doms, md)
                md.graph_capture(dt)
                after = _snapshot(doms, md)
                for k in range(len(doms)):
                    assert np.array_equal(before[0][k], after[0][k]) and before[2][k] == after[2][k], k
                captured = True
            md.graph_launch(n)
    got = _snapshot(doms, md)
    for k in range(len(doms)):
        assert np.array_equal(got[0][k], want[0][k]), ("state", k)
        assert np.array_equal(got[1][k], want[1][k]), ("max_stage", k)
        assert got[2][k] == want[2][k], ("time", k)
    md.graph_destroy()
    with pytest.raises(BackendError, match="no captured step"):
        md.graph_launch(1)
    _close(doms, md)
