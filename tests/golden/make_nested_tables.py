"""Writes the nesting geometry of the three-level nested workload
(``ptha_amd.configs.nested_c3``) as decided by the reference's own ``md%setup``
(compiled from the reference sources into oracle/_ref/): per domain the grid after
setup, the send/receive boxes and the priority arrays, plus all_dx_md and
all_timestepping_methods_md. These tables are the input a SWALS host hands to the
device (the drop-in runs md%setup in Fortran); bench.py's nested workload reads
them, so the GPU run does not need the reference.

The time-stepping refinement factors are chosen here from the bathymetry: the
parent takes the 'linear' CFL step (0.7), and each nested domain the number of
sub-steps that keeps its step below 0.8 of its stationary CFL step.

    python tests/golden/make_nested_tables.py [scale]
writes tests/golden/nested_c3_tables[_sN].npz
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.pyoracle import RefMultidomain  # noqa: E402
from ptha_amd import configs as C  # noqa: E402
from ptha_amd.domain import build_metrics  # noqa: E402


def cfl_step(elev, m, cfl):
    """linear_timestep_max (domain_mod.f90:1906-1931) of a grid."""
    dbe, dle = m["distance_bottom_edge"], m["distance_left_edge"]
    dist = np.minimum((dbe[:-1] + dbe[1:])[:, None], (dle[:-1] + dle[1:])[None, :])
    return float((0.5 * dist / np.sqrt(9.8 * np.maximum(0.0 - elev, 1.0e-05))).min() * cfl)


def main(scale=1):
    cfg = C.nested_c3(scale=scale)
    doms = cfg["domains"]
    # geometry as match_geometry_to_parent will set it (before halos), for the CFL steps
    geo = [(doms[0]["nx"], doms[0]["ny"], doms[0]["lw"], doms[0]["ll"])]
    for d in doms[1:]:
        pnx, pny, plw, pll = geo[d["parent"]]
        pdx = (plw[0] / pnx, plw[1] / pny)
        ll = [pll[q] + round((d["ll"][q] - pll[q]) / pdx[q]) * pdx[q] for q in range(2)]
        ur = [pll[q] + round((d["ur"][q] - pll[q]) / pdx[q]) * pdx[q] for q in range(2)]
        dx = (pdx[0] / d["dx_refinement_factor"], pdx[1] / d["dx_refinement_factor"])
        nx, ny = int(round((ur[0] - ll[0]) / dx[0])), int(round((ur[1] - ll[1]) / dx[1]))
        geo.append((nx, ny, (ur[0] - ll[0], ur[1] - ll[1]), tuple(ll)))
    steps = []
    for k, (nx, ny, lw, ll) in enumerate(geo):
        x, y = C.cell_centres(nx, ny, lw, ll)
        m = build_metrics(nx, ny, lw, ll, True)
        e = cfg["elevation"](x, y)
        steps.append(cfl_step(e, m, 0.7) if doms[k]["method"] != "rk2" else cfl_step(e, m, 0.99) * 0.5)
    dt = steps[0]
    safety = [0.8 for d in doms]
    trf = [1] + [int(math.ceil(dt / (f * s))) for f, s in zip(safety[1:], steps[1:])]
    for d, t in zip(doms, trf):
        d["timestepping_refinement_factor"] = t
    md = RefMultidomain(doms, spherical=True)
    dx_all, meth = md.all_dx_and_methods()
    arrays = dict(all_dx=dx_all, all_methods=meth)
    tables = dict(dt=dt, trf=trf, scale=scale, methods=[d["method"] for d in doms], domains=[])

    def clean(c):
        return {kk: (vv.tolist() if isinstance(vv, np.ndarray) else vv) for kk, vv in c.items()}

    for k in range(md.nd):
        d = md.dims(k)
        lw, ll, dx = md.geometry(k)
        pdi, pnp = md.priority(k)
        arrays["pdi_%d" % k] = pdi.astype(np.int8)
        arrays["pnp_%d" % k] = pnp.astype(np.int8)
        tables["domains"].append(dict(dims=d, lw=lw.tolist(), ll=ll.tolist(), dx=dx.tolist(),
                                      send=[clean(c) for c in md.comms(k, False)],
                                      recv=[clean(c) for c in md.comms(k, True)]))
    arrays["tables_json"] = np.frombuffer(json.dumps(tables).encode(), dtype=np.uint8)
    name = "nested_c3_tables%s.npz" % ("" if scale == 1 else "_s%d" % scale)
    np.savez_compressed(os.path.join(HERE, name), **arrays)
    print(name, "dt", dt, "trf", trf, [md.dims(k)["nx"] for k in range(md.nd)])


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1)
