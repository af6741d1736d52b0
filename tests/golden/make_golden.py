"""Writes the golden fixtures of tests/golden/ by running the reference's own Fortran
solver (compiled from the reference sources into oracle/_ref/ by oracle/Makefile) on
small seeded problems. Each .npz holds the inputs, the parameters and the reference's
outputs: U after the run, the max_U grids, and per-step scalars.

    python tests/golden/make_golden.py [name ...]     (default: every problem)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import helpers as H  # noqa: E402
import pyoracle  # noqa: E402


def problems():
    lin = H.small_cartesian_linear(nx=40, ny=32, dx=5000.0, sigma=30e3)
    sph = H.small_spherical(nx=48, ny=40, arcmin=15.0)
    llp = dict(sph, method="leapfrog_linear_plus_nonlinear_friction",
               manning_squared=np.full((40, 48), 0.03 ** 2))
    beach = H.small_beach(nx=48, ny=32)
    U = sph["U"].copy()
    U[3] = np.maximum(U[3], -300.0)
    U[0] = np.maximum(U[0], U[3])
    rk2s = dict(sph, U=U, method="rk2", manning_squared=np.full((40, 48), 0.025 ** 2))
    return {"linear_cart": (lin, {}, 30), "llpnf_sph": (llp, {}, 30), "rk2_beach": (beach, {}, 25),
            "rk2_sph": (rk2s, {}, 15), "linear_cart_ntl": (lin, {"truely_linear": False}, 30),
            "euler_beach": (dict(beach, method="euler"), {}, 25),
            "midpoint_beach": (dict(beach, method="midpoint"), {}, 25),
            "rk2n_sph": (dict(rk2s, method="rk2n"), {}, 12)}


def main(names=None):
    for name, (cfg, kw, nsteps) in problems().items():
        if names and name not in names:
            continue
        d, dt, hist = H.cpu_run("ref", cfg, nsteps, record=lambda x: x.scalars(), **kw)
        meta = dict(nx=cfg["nx"], ny=cfg["ny"], lw=list(cfg["lw"]), ll=list(cfg["ll"]), method=cfg["method"],
                    spherical=bool(cfg["spherical"]), nsteps=nsteps, dt=dt, kw=kw,
                    stage_bc=cfg.get("stage_bc"))
        arrays = dict(U_in=cfg["U"], U_out=d.get_U(), meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
        if "manning_squared" in cfg:
            arrays["manning_squared"] = cfg["manning_squared"]
        for q in pyoracle.MAXQ:
            arrays["max_" + q] = d.max_U(q)
        keys = sorted(hist[0])
        arrays["scalar_keys"] = np.frombuffer(json.dumps(keys).encode(), dtype=np.uint8)
        arrays["scalars"] = np.array([[h[k] for k in keys] for h in hist])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print("wrote", name)


if __name__ == "__main__":
    main(sys.argv[1:])
