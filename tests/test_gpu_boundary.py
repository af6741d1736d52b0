"""GPU: Flather boundaries (flather_boundary, boundary_mod.f90:502-720, through
swh_set_flather_boundary) and the other boundary subroutines (transmissive_boundary,
boundary_stage_radiation_momentum, boundary_stage_transmissive_momentum, :23-405, through
swh_set_boundary) against the oracle, whose restatements are pinned bit for bit to the
compiled reference (tests/test_oracle_vs_reference.py, the *_flather, *_transmissive and
*_radiation cases). The open basin has no land and no walls, a hump next to the west edge and
currents across the north and east edges, so every side carries flow.

* leapfrog FP64 ('linear' Cartesian and spherical with Coriolis, truely-linear llpnf;
  the sea at rest outside, or the sine stage outside on some sides): state, max_U
  grids and time bit-identical; boundary flux integral to 1e-12;
* the same with a forcing patch active (the stage forcing after the boundary, the
  elevation forcing before the momentum half-step): bit-identical to the compiled
  reference;
* FP32 + Kahan leapfrog: stage within 1e-5 m of FP64 (measured 4.2e-7 m), momenta
  within 1e-3 m^2/s (measured 1.1e-4; momenta reach 50 m^2/s);
* finite volume (rk2, midpoint, euler; the device sin for the boundary stage): FP64
  within 1e-9 of the oracle.
"""
import numpy as np
import pytest

import helpers as H
import pyoracle

pytestmark = pytest.mark.gpu

MAXQ = ["max_stage", "max_speed", "max_flux", "arrival_time", "time_of_max_stage", "min_stage"]

LF_CASES = {
    "linear_sph_transmissive": lambda: H.with_bc(H.open_basin(), "transmissive"),
    "linear_sph_radiation_fun": lambda: H.with_bc(H.open_basin(), "stage_radiation_momentum", amp=0.3),
    "lfnl_sph_transmissive": lambda: H.with_bc(dict(H.open_basin(method="leapfrog_nonlinear"),
                                                    manning_squared=np.zeros((80, 90))), "transmissive"),
    "linear_sph": lambda: H.with_flather(H.open_basin()),
    "linear_sph_fun": lambda: H.with_flather(H.open_basin(), sides=(1, 0, 0, 1), amp=0.3),
    "linear_cart": lambda: H.with_flather(dict(H.open_basin(), spherical=False)),
    "llpnf_sph_fun": lambda: H.with_flather(dict(H.open_basin(method="leapfrog_linear_plus_nonlinear_friction"),
                                                 manning_squared=np.full((80, 90), 0.03 ** 2)),
                                            sides=(0, 1, 1, 1), amp=0.2, period=900.0),
}

FV_CASES = {
    "rk2_sph_transmissive": (lambda: H.with_bc(H.open_basin(method="rk2"), "transmissive"), 30),
    "rk2_sph_radiation": (lambda: H.with_bc(H.open_basin(method="rk2"), "stage_radiation_momentum"), 30),
    "rk2_beach_radiation_fun": (lambda: H.with_bc(H.small_beach(), "stage_radiation_momentum", sides=(0, 0, 0, 1),
                                                  amp=1.0), 60),
    "midpoint_beach_transmissive_momentum": (lambda: H.with_bc(dict(H.small_beach(), method="midpoint"),
                                                               "stage_transmissive_momentum", sides=(0, 0, 0, 1),
                                                               amp=1.0), 40),
    "rk2_sph": (lambda: H.with_flather(H.open_basin(method="rk2")), 30),
    "midpoint_sph_fun": (lambda: H.with_flather(H.open_basin(method="midpoint"), amp=0.5), 30),
    "euler_beach_fun": (lambda: H.with_flather(dict(H.small_beach(), method="euler"), sides=(0, 0, 0, 1), amp=1.0),
                        40),
    "rk2_beach_fun": (lambda: H.with_flather(H.small_beach(), amp=1.0), 60),
}


@pytest.mark.parametrize("case", sorted(LF_CASES))
def test_leapfrog_flather_fp64_bitexact(oracle_built, case):
    print("CASE flather leapfrog fp64", case, flush=True)
    cfg = LF_CASES[case]()
    nsteps = 60
    o, dt, _ = H.cpu_run("oracle", cfg, nsteps)
    g = H.gpu_domain(cfg)
    g.evolve_steps(dt, 31)
    g.evolve_steps(dt, nsteps - 31)
    Ug, Uo = g.U, o.get_U()
    for k in range(4):
        assert np.array_equal(Ug[k], Uo[k]), (k, np.abs(Ug[k] - Uo[k]).max())
    for q in MAXQ:
        assert np.array_equal(g.max_U(q), o.max_U(q)), q
    sg, so = g.scalars(), o.scalars()
    assert sg["time"] == so["time"]
    assert sg["boundary_flux_time_integral"] == pytest.approx(so["boundary_flux_time_integral"], rel=1e-12, abs=1e-6)
    g.close()


@pytest.mark.parametrize("case", ["linear_sph", "llpnf_sph_fun"])
def test_leapfrog_flather_fp32k(oracle_built, case):
    print("CASE flather leapfrog fp32k", case, flush=True)
    cfg = LF_CASES[case]()
    nsteps = 60
    o, dt, _ = H.cpu_run("oracle", cfg, nsteps)
    g = H.gpu_domain(cfg, precision="fp32k")
    g.evolve_steps(dt, nsteps)
    Ug, Uo = g.U, o.get_U()
    errs = [float(np.abs(Ug[k] - Uo[k]).max()) for k in range(3)]
    print("fp32k errors", errs, flush=True)
    assert errs[0] <= 1e-5
    assert max(errs[1:]) <= 1e-3
    g.close()


def test_leapfrog_flather_with_forcing_matches_reference(oracle_built):
    print("CASE flather leapfrog + forcing fp64", flush=True)
    if not pyoracle.available("ref", True):
        pytest.skip("compiled reference not built (oracle/_ref)")
    cfg = LF_CASES["linear_sph_fun"]()
    nx, ny = cfg["nx"], cfg["ny"]
    i0, i1, j0, j1 = 30, 50, 20, 41
    xx = np.linspace(-1, 1, i1 - i0 + 1)[None, :]
    yy = np.linspace(-1, 1, j1 - j0 + 1)[:, None]
    work = np.zeros((4, j1 - j0 + 1, i1 - i0 + 1))
    work[0] = work[3] = 1.2 * np.exp(-3 * (xx ** 2 + yy ** 2))
    r = H.cpu_run("ref", cfg, 0)[0]
    dt = r.stationary_timestep_max()
    r.set_forcing_patch(3.3 * dt, 9.7 * dt, i0, i1, j0, j1, 1, 4, work)
    r.evolve(dt, 25)
    g = H.gpu_domain(cfg)
    g.set_forcing_patch(3.3 * dt, 9.7 * dt, i0, i1, j0, j1, 1, 4, work)
    g.evolve_steps(dt, 25)
    Ug, Ur = g.U, r.get_U()
    for k in range(4):
        assert np.array_equal(Ug[k], Ur[k]), (k, np.abs(Ug[k] - Ur[k]).max())
    assert np.array_equal(g.max_U("max_stage"), r.max_U("max_stage"))
    g.close()


@pytest.mark.parametrize("case", sorted(FV_CASES))
def test_fv_flather_fp64(oracle_built, case):
    print("CASE flather finite volume fp64", case, flush=True)
    make, nsteps = FV_CASES[case]
    cfg = make()
    o = H.cpu_run("oracle", cfg, 0)[0]
    g = H.gpu_domain(cfg)
    for _ in range(nsteps):
        o.evolve(None, 1)
        g.evolve_one_step()
    Ug, Uo = g.U, o.get_U()
    wet = np.abs(Uo[3]) < 1e5
    errs = [float(np.abs(Ug[k] - Uo[k])[wet].max()) for k in range(3)]
    print("fp64 errors", errs, flush=True)
    assert max(errs) <= 1e-9
    assert g.scalars()["time"] == pytest.approx(o.scalars()["time"], rel=1e-12)
    g.close()


def test_flather_argument_checks():
    from ptha_amd.backend import BackendError
    cfg = LF_CASES["linear_sph"]()
    g = H.gpu_domain(cfg)
    with pytest.raises(BackendError):
        g.set_flather_boundary(sides=(1, 1, 1, 1), amp=1.0, period=0.0)
    g.close()
