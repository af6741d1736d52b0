"""Full-size FP64 reference runs of four unit sources of the batched unit-source
workload (``ptha_amd.configs.unit_sources``: 1200 x 1200 spherical 'linear' grid at
1 arcmin, power-law bathymetry with a shelf ramp to land, walls, Coriolis; 1 m
Gaussian sources of sigma 25 km), each for its own full duration, computed by the
reference's own solver compiled from the reference sources (oracle/_ref).

The run loop is the reference's unit-source driver
(propagation/SWALS/examples/generic_example/generic_model.f90:377-405): the fixed
step is domain%stationary_timestep_max(); gauges are written when
time - last_write_time >= 60 s (last_write_time starting at -60 s and advancing by
60 s per write); the run stops before the first step with time > final_time.

Sources: the shortest and the longest run of the 512, one on the eastern lattice
column next to the shelf and coast, and the south-west lattice corner.

Writes tests/golden/fullsize_config4.npz: per source the stage at the 64 gauges at
every write step, the write steps and times, the final stage on every 10th cell, the
step count; plus dt, the gauge cells and SHA-256 digests of the elevation and of each
initial stage (so a test can confirm it rebuilt the same inputs).

    OMP_NUM_THREADS=8 python tests/golden/make_fullsize_config4.py   (about 4 minutes)
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
from ptha_amd import configs as C  # noqa: E402
from ptha_amd.batch import reference_write_schedule  # noqa: E402

SOURCES = {"shortest": 195, "longest": 294, "coastal": 287, "corner": 0}
SUB = 10
WRITE_FREQ = 60.0


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def main():
    cfg = C.unit_sources()
    nx, ny = cfg["nx"], cfg["ny"]
    dx = (cfg["lw"][0] / nx, cfg["lw"][1] / ny)
    ij = C.gauge_site_index(cfg["gauges_xy"], cfg["ll"], dx, nx, ny)
    out = dict(gauge_ij=ij, sub=SUB, write_freq=WRITE_FREQ, elev_sha256=digest(cfg["elev"]),
               names=np.array(list(SOURCES)), index=np.array(list(SOURCES.values()), dtype=np.int64))
    dt = None
    t_all = time.time()
    for name, k in SOURCES.items():
        d = pyoracle.CpuDomain("ref", nx, ny, cfg["lw"], cfg["ll"], method="linear", spherical=True)
        d.set_record_max_U(False)
        U = np.zeros((4, ny, nx))
        U[3] = cfg["elev"]
        U[0] = C.unit_source_stage(cfg, k)
        d.set_U(U)
        if dt is None:
            dt = d.stationary_timestep_max()
        final_time = float(cfg["durations"][k])
        # generic_model.f90:380-405, with the reference's own domain%time
        last = -WRITE_FREQ
        n, steps, times, series = 0, [], [], []
        t0 = time.time()
        while True:
            t = d.scalars()["time"]
            if t - last >= WRITE_FREQ:
                last = last + WRITE_FREQ
                steps.append(n)
                times.append(t)
                series.append(d.sample(ij)[0])
            if t > final_time:
                break
            d.evolve(dt, 1)
            n += 1
        # the host-side schedule used by the batch must reproduce the reference's own
        nref, sref, tref = reference_write_schedule(dt, final_time, WRITE_FREQ)
        assert nref == n and np.array_equal(sref, steps) and np.array_equal(tref, times), name
        out[name + "_gauge_stage"] = np.array(series)
        out[name + "_steps"] = np.array(steps, dtype=np.int64)
        out[name + "_times"] = np.array(times)
        out[name + "_nsteps"] = n
        out[name + "_final_time"] = final_time
        out[name + "_stage"] = d.get_U()[0][::SUB, ::SUB]
        out[name + "_init_sha256"] = digest(U[0])
        d.close()
        print(name, "source", k, "steps", n, "samples", len(steps), "wall %.0f s" % (time.time() - t0), flush=True)
    out["dt"] = dt
    np.savez_compressed(os.path.join(HERE, "fullsize_config4.npz"), **out)
    print("done in %.0f s" % (time.time() - t_all))


if __name__ == "__main__":
    main()
