"""Runs a nested multidomain through the compiled reference (oracle/_ref) and
writes what the GPU parity tests compare against. Started as a child process by
tests/test_gpu_multidomain.py so that the reference (whose md%setup runs shell
commands to make its output folders) never shares a process with the GPU.

usage: python tests/md_ref_run.py OUT_PREFIX NSTEPS [planar | flow:METHOD]
writes OUT_PREFIX.json (setup tables) and OUT_PREFIX.npz (arrays).

With 'flow:METHOD' the grandchild uses that finite-volume method ('euler',
'midpoint', 'rk2n') instead of 'rk2'. With 'child:METHOD' the child uses that
leapfrog method (e.g. 'leapfrog_nonlinear') instead of
'leapfrog_linear_plus_nonlinear_friction'.

With 'planar', every domain is 'rk2' and holds the planar field of the
reference's own nesting test (multidomain_mod.f90:4088-4101): U(:, :, k) =
(x - x0) * 1e-4 + (y - y0) * 3e-5 + (domain index) + [9, 3, 6, 0](k) (the
reference's test adds [0, 3, 6, 9]; the depth is made positive here so that the
wet-dry limiter leaves the interpolation linear). Only the initial exchange runs.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle.pyoracle import RefMultidomain  # noqa: E402
from ptha_amd import configs  # noqa: E402


def centres(lw, ll, nx, ny):
    x = ll[0] + ((np.arange(1, nx + 1) - 0.5) / (nx * 1.0)) * lw[0]
    y = ll[1] + ((np.arange(1, ny + 1) - 0.5) / (ny * 1.0)) * lw[1]
    return x, y


def planar(x, y, k, x0, y0):
    X, Y = np.meshgrid(x, y)
    val = (X - x0) * 1.0e-4 + (Y - y0) * 3.0e-5 + (k + 1)
    return np.stack([val + 9.0, val + 3.0, val + 6.0, val + 0.0])


def main(prefix, nsteps, mode="flow"):
    cfg = configs.nested_small()
    if mode.startswith("flow:"):
        cfg["domains"][2]["method"] = mode.split(":", 1)[1]
    if mode.startswith("child:"):
        cfg["domains"][1]["method"] = mode.split(":", 1)[1]
    if mode == "planar":
        for d in cfg["domains"]:
            d["method"] = "rk2"
            d["timestepping_refinement_factor"] = 1
    md = RefMultidomain(cfg["domains"], spherical=cfg["spherical"])
    x0, y0 = cfg["domains"][0]["ll"]
    nd = md.nd
    arrays = {}
    tables = dict(nd=nd, methods=md.methods, dt=cfg["dt"], domains=[])
    dx_all, meth = md.all_dx_and_methods()
    arrays["all_dx"] = dx_all
    arrays["all_methods"] = meth
    for k in range(nd):
        d = md.dims(k)
        lw, ll, dx = md.geometry(k)
        x, y = centres(lw, ll, d["nx"], d["ny"])
        U = configs.nested_small_state(x, y) if mode != "planar" else planar(x, y, k, x0, y0)
        md.set_U(k, U)
        if cfg["domains"][k]["method"] != "linear" and mode != "planar":
            md.set_manning_squared(k, np.full((d["ny"], d["nx"]), cfg["manning_n"] ** 2))
        arrays["U0_%d" % k] = U
        for name, v in md.metrics(k).items():
            arrays["%s_%d" % (name, k)] = v
        pdi, pnp = md.priority(k)
        arrays["pdi_%d" % k] = pdi
        arrays["pnp_%d" % k] = pnp

        def clean(c):
            return {kk: (vv.tolist() if isinstance(vv, np.ndarray) else vv) for kk, vv in c.items()}

        tables["domains"].append(dict(dims=d, lw=lw.tolist(), ll=ll.tolist(), dx=dx.tolist(),
                                      send=[clean(c) for c in md.comms(k, False)],
                                      recv=[clean(c) for c in md.comms(k, True)]))
    md.make_initial_conditions_consistent()
    arrays["vol0"] = md.volumes()
    for k in range(nd):
        arrays["Ucons_%d" % k] = md.get_U(k)
        for c in range(md.dims(k)["n_recv"]):
            arrays["rbuf0_%d_%d" % (k, c)] = md.comm_buffer(k, True, c)
    tables["nsteps"] = 0
    if mode == "planar":
        with open(prefix + ".json", "w") as f:
            json.dump(tables, f)
        np.savez(prefix + ".npz", **arrays)
        return
    md.evolve(cfg["dt"], 1)
    for k in range(nd):
        arrays["U1_%d" % k] = md.get_U(k)
        for c in range(md.dims(k)["n_recv"]):
            arrays["rbuf1_%d_%d" % (k, c)] = md.comm_buffer(k, True, c)
    md.evolve(cfg["dt"], nsteps - 1)
    arrays["volN"] = md.volumes()
    for k in range(nd):
        arrays["UN_%d" % k] = md.get_U(k)
    tables["nsteps"] = nsteps
    with open(prefix + ".json", "w") as f:
        json.dump(tables, f)
    np.savez(prefix + ".npz", **arrays)


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]), sys.argv[3] if len(sys.argv) > 3 else "flow")
