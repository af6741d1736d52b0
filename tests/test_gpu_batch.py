"""GPU: batched unit sources (one launch per step over all running sources).

Each source of an FP64 batch must equal, bit for bit, the oracle run of that source
alone for its own number of steps (sources of different durations freeze at their
own end), and its gauge series must equal the oracle state sampled at the same steps.
"""
import numpy as np
import pytest

import helpers as H
import pyoracle
from ptha_amd import configs as C

pytestmark = pytest.mark.gpu


def _cfg():
    u = C.unit_sources(nx=96, ny=72, arcmin=10.0, n_src_x=3, n_src_y=1, spacing_km=250, sigma_km=60)
    return u


@pytest.mark.parametrize("precision", ["fp64", "fp32k"])
def test_batch_matches_single_source_runs(oracle_built, precision):
    print("CASE batch", precision, flush=True)
    from ptha_amd.batch import UnitSourceBatch
    u = _cfg()
    nx, ny = u["nx"], u["ny"]
    nsteps = np.array([50, 80, 30], dtype=np.int32)
    every = 10
    b = UnitSourceBatch(nx, ny, u["lw"], u["ll"], 3, precision=precision)
    b.set_elevation(u["elev"])
    b.set_nsteps(nsteps)
    stages = []
    for k in range(3):
        s = C.gaussian_on_sphere(u["x"], u["y"], *u["centres"][k], 60e3)
        s[u["elev"] >= 0] = 0.0
        stages.append(s)
        b.set_source(k, s)
    ij = b.setup_point_gauges(u["gauges_xy"], every=every, max_samples=20)
    o0 = pyoracle.CpuDomain("oracle", nx, ny, u["lw"], u["ll"], method="linear", spherical=True)
    U = np.zeros((4, ny, nx)); U[3] = u["elev"]
    o0.set_U(U)
    dt = o0.stationary_timestep_max()
    b.evolve(dt, 45)
    b.evolve(dt, 100)   # runs past every source's end: all freeze
    assert b.status() == (80, 0)
    for k in range(3):
        o = pyoracle.CpuDomain("oracle", nx, ny, u["lw"], u["ll"], method="linear", spherical=True)
        U = np.zeros((4, ny, nx)); U[3] = u["elev"]; U[0] = stages[k]
        o.set_U(U)
        ref_series = []
        for step in range(nsteps[k] + 1):
            if step % every == 0:
                V = o.get_U()
                ref_series.append(np.array([[V[q][j - 1, i - 1] for (i, j) in ij] for q in range(3)]))
            if step < nsteps[k]:
                o.evolve(dt, 1)
        got = b.get_source(k)
        ref = o.get_U()[:3]
        series = b.gauge_series(k)
        assert series.shape[0] == len(ref_series)
        if precision == "fp64":
            assert np.array_equal(got, ref), k
            assert np.array_equal(series, np.array(ref_series)), k
        else:
            wet = u["elev"] < 0
            assert np.abs(got[0] - ref[0])[wet].max() <= 1e-5, k
            assert np.abs(series[:, 0] - np.array(ref_series)[:, 0]).max() <= 1e-5, k
    b.close()
