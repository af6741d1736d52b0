"""Full-size FP64 reference run of the source-inversion workload
(``ptha_amd.configs.adjoint_inversion``: 2048 x 2048 spherical 'linear' grid at 1
arcmin, power-law bathymetry with a shelf ramp and walls, Coriolis; initial stage the
weighted sum of the 40 x 20 Gaussian unit sources with the hidden-truth weights),
4 h simulated, computed by the reference's own solver compiled from the reference
sources (oracle/_ref).

The run loop is the reference's unit-source driver
(propagation/SWALS/examples/generic_example/generic_model.f90:377-405): the fixed
step is domain%stationary_timestep_max(); the 32 gauges are written at the first step
on or after each 60 s write time; the run stops before the first step with
time > final_time.

Writes tests/golden/fullsize_config5.npz: the stage at the gauges at every write step,
the write steps and times, the final stage on every 16th cell, the step count, dt, the
gauge cells and SHA-256 digests of the elevation and of the initial stage.

    OMP_NUM_THREADS=8 python tests/golden/make_fullsize_config5.py   (about 2 minutes)
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from ptha_amd import adjoint as A  # noqa: E402
from ptha_amd import configs as C  # noqa: E402
from ptha_amd.batch import reference_write_schedule  # noqa: E402

SUB = 16


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def main():
    cfg = C.adjoint_inversion()
    nx, ny = cfg["nx"], cfg["ny"]
    dx = (cfg["lw"][0] / nx, cfg["lw"][1] / ny)
    ij = C.gauge_site_index(cfg["gauges_xy"], cfg["ll"], dx, nx, ny)
    GX, GY = A.unit_source_basis(cfg["x"], cfg["y"], cfg["src_lon"], cfg["src_lat"], cfg["sigma_km"] * 1e3)
    eta = A.stage_from_weights(cfg["w_true"], GX, GY, cfg["elev"])
    d = pyoracle.CpuDomain("ref", nx, ny, cfg["lw"], cfg["ll"], method="linear", spherical=True)
    d.set_record_max_U(False)
    U = np.zeros((4, ny, nx))
    U[3] = cfg["elev"]
    U[0] = eta
    d.set_U(U)
    dt = d.stationary_timestep_max()
    final_time = cfg["hours"] * 3600.0
    freq = cfg["gauge_interval"]
    last = -freq
    n, steps, times, series = 0, [], [], []
    t0 = time.time()
    while True:
        t = d.scalars()["time"]
        if t - last >= freq:
            last = last + freq
            steps.append(n)
            times.append(t)
            series.append(d.sample(ij)[0])
        if t > final_time:
            break
        d.evolve(dt, 1)
        n += 1
    nref, sref, tref = reference_write_schedule(dt, final_time, freq)
    assert nref == n and np.array_equal(sref, steps) and np.array_equal(tref, times)
    np.savez_compressed(os.path.join(HERE, "fullsize_config5.npz"), gauge_ij=ij, gauge_stage=np.array(series),
                        steps=np.array(steps, dtype=np.int64), times=np.array(times), nsteps=n, dt=dt,
                        final_time=final_time, write_freq=freq, sub=SUB, stage=d.get_U()[0][::SUB, ::SUB],
                        elev_sha256=digest(cfg["elev"]), init_sha256=digest(eta))
    print("steps", n, "samples", len(steps), "wall %.0f s" % (time.time() - t0))


if __name__ == "__main__":
    main()
