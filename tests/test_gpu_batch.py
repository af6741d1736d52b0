"""GPU: batched unit sources (one launch per step, or per two steps, over all
running sources).

Each source of an FP64 batch must equal, bit for bit, the oracle run of that source
alone for its own number of steps (sources of different durations freeze at their
own end), and its gauge series must equal the oracle state sampled at the same steps.
The two-steps-per-launch schedule must give the same bits as one step per launch.
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


@pytest.mark.parametrize("precision,two_step", [("fp64", False), ("fp32k", False), ("fp64", True),
                                                ("fp32k", True), ("fp64", "xcd"), ("fp32k", "xcd"),
                                                ("fp64", "three"), ("fp32k", "three"), ("fp64", "four"),
                                                ("fp32k", "four"), ("fp64", "five"), ("fp32k", "five")])
def test_batch_matches_single_source_runs(oracle_built, precision, two_step):
    """two_step: False = one step per launch; True = two fused steps per launch (next-row
    loads in flight); 'xcd' = the same with the XCD-aware tile order; 'three' = three
    fused steps per launch where no gauge sample falls inside them (else two or one);
    'four', 'five' = up to four or five."""
    print("CASE batch", precision, {False: "one-step", True: "two-step", "xcd": "two-step xcd",
                                     "three": "three-step", "four": "four-step",
                                     "five": "five-step"}[two_step], flush=True)
    from ptha_amd.batch import UnitSourceBatch
    u = _cfg()
    nx, ny = u["nx"], u["ny"]
    # odd counts and a sampling period that splits two-step rounds
    nsteps = np.array([51, 80, 33], dtype=np.int32) if two_step else np.array([50, 80, 30], dtype=np.int32)
    every = 10
    b = UnitSourceBatch(nx, ny, u["lw"], u["ll"], 3, precision=precision)
    if two_step == "three":
        b.set_launch(128, 8, 16 | 64 | 2 | (1 << 7) | 1024)
    elif two_step == "four":
        b.set_launch(64, 16, 16 | 64 | 2 | (2 << 7) | 2048)
    elif two_step == "five":
        b.set_launch(128, 32, 16 | 64 | 2 | (4 << 7) | (3 << 10))
    elif two_step == "xcd":
        # XCD order with sources interleaved in groups of 2 (3 sources: a short last group)
        b.set_launch(128, 8, 16 | 64 | 2 | (1 << 7))
    elif two_step:
        b.set_launch(64, 16, 16 | (3 if precision == "fp32k" else 1))
    else:
        b.set_launch(128, 16, 3 if precision == "fp32k" else 1)
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


@pytest.mark.parametrize("precision", ["fp64", "fp32k"])
def test_two_step_friction_and_autotune_bitwise(precision):
    """'leapfrog_linear_plus_nonlinear_friction' batch (Manning n = 0.03): two to five
    steps per launch equal one step per launch bit for bit; running the autotuner first
    changes nothing in the results."""
    print("CASE batch llpnf two-step/autotune", precision, flush=True)
    from ptha_amd.batch import UnitSourceBatch
    u = _cfg()
    nx, ny = u["nx"], u["ny"]
    nsteps = np.array([41, 64, 27], dtype=np.int32)
    runs = []
    for mode in ["one", "two", "three", "four", "five", "tuned"]:
        b = UnitSourceBatch(nx, ny, u["lw"], u["ll"], 3, precision=precision,
                            timestepping_method="leapfrog_linear_plus_nonlinear_friction")
        b.set_manning_squared(np.full((ny, nx), 0.03 ** 2))
        b.set_elevation(u["elev"])
        b.set_nsteps(nsteps)
        for k in range(3):
            st = C.gaussian_on_sphere(u["x"], u["y"], *u["centres"][k], 60e3)
            st[u["elev"] >= 0] = 0.0
            b.set_source(k, st)
        b.setup_point_gauges(u["gauges_xy"], every=7, max_samples=20)
        dt = 20.0
        if mode == "one":
            b.set_launch(128, 16, 1)
        elif mode == "two":
            b.set_launch(128, 32, 16 | 64 | 2)
        elif mode == "three":
            b.set_launch(64, 16, 16 | 64 | 2 | 1024)
        elif mode == "four":
            b.set_launch(128, 8, 16 | 64 | 2 | 2048)
        elif mode == "five":
            b.set_launch(64, 7, 16 | 64 | 2 | (3 << 10))
        else:
            shape, ms = b.autotune(dt, reps=1)
            assert ms > 0 and len(shape) == 3
        b.evolve(dt, 30)
        b.evolve(dt, 50)
        assert b.status() == (64, 0)
        runs.append(([b.get_source(k) for k in range(3)], [b.gauge_series(k) for k in range(3)]))
        b.close()
    for other in runs[1:]:
        for k in range(3):
            assert np.array_equal(runs[0][0][k], other[0][k]), k
            assert np.array_equal(runs[0][1][k], other[1][k]), k
