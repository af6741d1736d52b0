"""Full-size FP64 reference run of the three-level nested multidomain
(``ptha_amd.configs.nested_c3``: 'linear' parent 3000 x 1800 at 2 arcmin,
'leapfrog_linear_plus_nonlinear_friction' child at 1/3, 'rk2' grandchild at 1/15 with
the shore inside it; 3 m elliptical uplift; 6 h simulated), computed by the
reference's own multidomain_type compiled from the reference sources
(oracle/_ref), with the global step and timestepping refinement factors of
tests/golden/nested_c3_tables.npz. Writes a compact fixture: per domain, 9 gauge
series (cells at 1/4, 1/2, 3/4 of each axis) sampled every 20 global steps, the
max-stage grid and the final stage on every 10th cell, and the priority-domain
volumes at the start and the end.

    OMP_NUM_THREADS=6 python tests/golden/make_fullsize_config3.py
writes tests/golden/fullsize_config3.npz
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from ptha_amd import configs as C  # noqa: E402

EVERY = 20
SUB = 10


def gauge_ij(nx, ny):
    return np.array([(int(nx * fx), int(ny * fy)) for fy in (0.25, 0.5, 0.75) for fx in (0.25, 0.5, 0.75)],
                    dtype=np.int32)


def main():
    z = np.load(os.path.join(HERE, "nested_c3_tables.npz"))
    tables = json.loads(bytes(z["tables_json"]).decode())
    cfg = C.nested_c3(scale=tables["scale"])
    doms = cfg["domains"]
    for d, t in zip(doms, tables["trf"]):
        d["timestepping_refinement_factor"] = t
    md = pyoracle.RefMultidomain(doms, spherical=True, max_stage=True)
    dt = tables["dt"]
    nglobal = int(round(cfg["duration"] / dt))
    ijs = []
    for k in range(md.nd):
        d = md.dims(k)
        assert [d["nx"], d["ny"]] == [tables["domains"][k]["dims"]["nx"], tables["domains"][k]["dims"]["ny"]]
        lw, ll, _ = md.geometry(k)
        x, y = C.cell_centres(d["nx"], d["ny"], lw, ll)
        md.set_U(k, cfg["state"](x, y))
        if doms[k]["method"] != "linear":
            md.set_manning_squared(k, np.full((d["ny"], d["nx"]), cfg["manning_n"] ** 2))
        ijs.append(gauge_ij(d["nx"], d["ny"]))
    md.make_initial_conditions_consistent()
    vol0 = md.volumes()
    series = [[md.sample(k, ijs[k])] for k in range(md.nd)]
    t0 = time.time()
    for n in range(1, nglobal + 1):
        md.evolve(dt, 1)
        if n % EVERY == 0:
            for k in range(md.nd):
                series[k].append(md.sample(k, ijs[k]))
        if n % 200 == 0:
            print("global step", n, "of", nglobal, "wall %.0f s" % (time.time() - t0), flush=True)
    out = dict(dt=dt, nglobal=nglobal, every=EVERY, sub=SUB, vol0=vol0, vol1=md.volumes())
    for k in range(md.nd):
        out["gauge_ij_%d" % k] = ijs[k]
        out["series_%d" % k] = np.array(series[k])
        out["max_stage_%d" % k] = md.max_stage(k)[::SUB, ::SUB]
        U = md.get_U(k)
        out["stage_%d" % k] = U[0][::SUB, ::SUB]
        out["elev_%d" % k] = U[3][::SUB, ::SUB]
    np.savez_compressed(os.path.join(HERE, "fullsize_config3.npz"), **out)
    print("done: global steps", nglobal, "wall %.0f s" % (time.time() - t0))


if __name__ == "__main__":
    main()
