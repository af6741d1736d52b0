"""Pins of the event-summation oracle (oracle/superpose.py) on the CPU.

The truly-linear leapfrog is linear in (stage - msl, uh, vh): the solver run from a
weighted sum of unit-source initial conditions equals the same weighted sum of the
unit-source runs to round-off. So an event built by the oracle from separate
unit-source runs (uniform slip: unit weights then the slip; stochastic slip:
per-source weights; msl taken out and put back) must equal the solver's run of the
event's own initial condition. A dropped slip, a wrong sign, msl added twice or
taken from the momenta fails this.
"""
import numpy as np
import pytest

import pyoracle
import superpose as SP
from ptha_amd import configs as C

NSTEPS, EVERY = 40, 5


def _grid():
    u = C.unit_sources(nx=60, ny=48, arcmin=10.0, n_src_x=3, n_src_y=1, spacing_km=200, sigma_km=60)
    return u


def _run(u, stage, msl, dt=None):
    nx, ny = u["nx"], u["ny"]
    o = pyoracle.CpuDomain("oracle", nx, ny, u["lw"], u["ll"], method="linear", spherical=True, msl_linear=msl)
    U = np.zeros((4, ny, nx))
    U[3] = u["elev"]
    U[0] = msl + stage
    o.set_U(U)
    if dt is None:
        dt = o.stationary_timestep_max()
    ij = C.gauge_site_index(u["gauges_xy"], u["ll"], (u["lw"][0] / nx, u["lw"][1] / ny), nx, ny)
    series = []
    for step in range(NSTEPS + 1):
        if step % EVERY == 0:
            V = o.get_U()
            series.append(np.array([[V[q][j - 1, i - 1] for (i, j) in ij] for q in range(3)]))
        if step < NSTEPS:
            o.evolve(dt, 1)
    return np.array(series), dt


@pytest.mark.parametrize("msl", [0.0, 0.7])
def test_events_equal_runs_of_event_initial_conditions(oracle_built, msl):
    u = _grid()
    bumps = []
    for k in range(3):
        b = C.gaussian_on_sphere(u["x"], u["y"], *u["centres"][k], 60e3)
        b[u["elev"] >= msl] = 0.0
        bumps.append(b)
    runs, dt = [], None
    for k in range(3):
        s, dt = _run(u, bumps[k], msl, dt)
        runs.append(s)
    ns = [len(r) for r in runs]
    scale = max(np.abs(r).max() for r in runs)
    # uniform slip 2.5 over sources 0 and 2
    ev = SP.event_series(runs, ns, [0, 2], slip=2.5, msl=msl)
    direct, _ = _run(u, 2.5 * (bumps[0] + bumps[2]), msl, dt)
    assert np.abs(ev - direct).max() <= 1e-12 * scale * 10
    # stochastic slip over all three
    w = [0.3, -1.2, 2.0]
    ev = SP.event_series(runs, ns, [0, 1, 2], weights=w, msl=msl)
    direct, _ = _run(u, w[0] * bumps[0] + w[1] * bumps[1] + w[2] * bumps[2], msl, dt)
    assert np.abs(ev - direct).max() <= 1e-12 * scale * 10
    # the summary is the maximum of the stage over time at each gauge
    mx = SP.event_max_stage(runs, ns, [0, 1, 2], weights=w, msl=msl)
    assert np.abs(mx - direct[:, 0, :].max(axis=0)).max() <= 1e-12 * scale * 10


def test_uniform_and_stochastic_forms_agree(oracle_built):
    """A uniform-slip event equals the stochastic form with every slip equal (the two
    orders of operations differ only by rounding), and shorter sources limit the
    event to their samples."""
    rng = np.random.default_rng(3)
    runs = [rng.normal(size=(9, 3, 5)) for _ in range(4)]
    ns = [9, 7, 9, 8]
    a = SP.event_series(runs, ns, [0, 1, 3], slip=1.7)
    b = SP.event_series(runs, ns, [0, 1, 3], weights=[1.7, 1.7, 1.7])
    assert a.shape == (7, 3, 5)
    np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-14)


def test_non_finite_source_reaches_only_its_events():
    """The reference sums each event's own unit sources (sum_tsunami_unit_sources.R:
    658-690, R arithmetic: NaN + x and 0 * NaN are NaN, and max over a vector holding
    NaN is NaN): a unit source whose series went non-finite makes NaN exactly the
    entries of the events that contain it where its series is non-finite (whatever its
    weight, zero included), and every other event stays finite."""
    rng = np.random.default_rng(5)
    ns, ng = [12, 12, 9], 4
    series = [rng.normal(size=(n, 3, ng)) for n in ns]
    series[1][7:, 0, 2] = np.nan      # stage at gauge 2 from sample 7
    series[1][9, 1, 0] = np.inf       # one uh value
    bad = ~np.isfinite(series[1])
    for src, w in (([0, 2], None), ([0, 1], None), ([1, 2], [0.0, 2.0]), ([1], [3.0])):
        ev = SP.event_series(series, ns, src, weights=w, slip=2.0 if w is None else None, msl=0.3)
        n = min(ns[k] for k in src)
        if 1 in src:
            assert np.array_equal(~np.isfinite(ev), bad[:n]), src
        else:
            assert np.isfinite(ev).all(), src
        mx = SP.event_max_stage(series, ns, src, weights=w, slip=2.0 if w is None else None, msl=0.3)
        want_nan = bad[:n, 0, :].any(axis=0) if 1 in src else np.zeros(ng, bool)
        assert np.array_equal(np.isnan(mx), want_nan), src
