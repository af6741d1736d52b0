"""GPU: 'leapfrog_nonlinear' (the nonlinear staggered leapfrog,
domain_mod_leapfrog_nonlinear_include.f90) against the oracle, whose restatement is
bit-identical to the compiled reference on the same cases
(tests/test_oracle_vs_reference.py, lfnl_*).

* without friction (manning 0): FP64 bit-identical state, max_U grids and time;
* with Manning friction: friction_work is formed with the device pow every step, which
  may differ from glibc's by an ulp, so FP64 is held to round-off: 1e-12 m on the stage
  and 1e-11 m^2/s on the momenta over the run (measured: bit-identical on four of the
  five cases, 2.5e-15 / 1.1e-14 on the fifth; wet/dry fronts, a dam break, Flather
  boundaries, a stage boundary);
* a forcing patch on a frictionless basin: bit-identical to the compiled reference;
* FP32 + Kahan stage: within 1e-5 m of the FP64 oracle on the all-wet basin (measured
  7.1e-7 m), momenta within 2e-3 m^2/s (measured 1.6e-4; they reach 50 m^2/s).
"""
import numpy as np
import pytest

import helpers as H
import pyoracle

pytestmark = pytest.mark.gpu

MAXQ = ["max_stage", "max_speed", "max_flux", "arrival_time", "time_of_max_stage", "min_stage"]
CASES = H.nonlinear_leapfrog_cases()


def _no_friction(cfg):
    return dict(cfg, manning_squared=np.zeros((cfg["ny"], cfg["nx"])))


def _run_both(cfg, nsteps, kw, precision="fp64"):
    o, dt, _ = H.cpu_run("oracle", cfg, nsteps, **kw)
    g = H.gpu_domain(cfg, precision=precision, **kw)
    assert g.stationary_timestep_max() == pytest.approx(dt, rel=1e-15 if precision == "fp64" else 1e-6)
    g.evolve_steps(dt, nsteps // 2)
    g.evolve_steps(dt, nsteps - nsteps // 2)
    return o, g


@pytest.mark.parametrize("case", ["basin_sph", "land_sph", "beach_cart"])
def test_nonlinear_no_friction_fp64_bitexact(oracle_built, case):
    print("CASE leapfrog_nonlinear no friction fp64", case, flush=True)
    make, nsteps, kw = CASES[case]
    cfg = _no_friction(make())
    o, g = _run_both(cfg, nsteps, kw)
    Ug, Uo = g.U, o.get_U()
    for k in range(4):
        assert np.array_equal(Ug[k], Uo[k]), (k, np.abs(Ug[k] - Uo[k]).max())
    for q in MAXQ:
        assert np.array_equal(g.max_U(q), o.max_U(q)), q
    assert g.scalars()["time"] == o.scalars()["time"]
    assert g.scalars()["boundary_flux_time_integral"] == pytest.approx(o.scalars()["boundary_flux_time_integral"],
                                                                       rel=1e-12, abs=1e-9)
    g.close()


@pytest.mark.parametrize("nx,ny", [(62, 81), (63, 82), (123, 83)])
def test_nonlinear_strip_edges_fp64_bitexact(oracle_built, nx, ny):
    """The flux pass walks strips of 61 columns and 4 rows on grids this small: column
    nx-1 (62, 123) or nx-2 (63) is the first column of a strip, and the last strip
    begins at row ny-1 (81), ny-2 (82) or ny-3 (83), so the index maps at the east
    columns and north rows are taken across strip edges."""
    print("CASE leapfrog_nonlinear strip edges fp64", nx, ny, flush=True)
    cfg = _no_friction(H.open_basin(nx=nx, ny=ny, method="leapfrog_nonlinear"))
    o, g = _run_both(cfg, 40, {})
    Ug, Uo = g.U, o.get_U()
    for k in range(4):
        assert np.array_equal(Ug[k], Uo[k]), (k, np.abs(Ug[k] - Uo[k]).max())
    g.close()


@pytest.mark.parametrize("case", sorted(CASES))
def test_nonlinear_fp64(oracle_built, case):
    print("CASE leapfrog_nonlinear fp64", case, flush=True)
    make, nsteps, kw = CASES[case]
    o, g = _run_both(make(), nsteps, kw)
    Ug, Uo = g.U, o.get_U()
    errs = [float(np.abs(Ug[k] - Uo[k]).max()) for k in range(3)]
    print("errors", errs, flush=True)
    assert errs[0] <= 1e-12
    assert max(errs[1:]) <= 1e-11
    wet = np.abs(Uo[3]) < 1e5
    assert np.abs(g.max_U("max_stage") - o.max_U("max_stage"))[wet].max() <= 1e-12
    g.close()


def test_nonlinear_forcing_matches_reference(oracle_built):
    print("CASE leapfrog_nonlinear + forcing fp64", flush=True)
    if not pyoracle.available("ref", True):
        pytest.skip("compiled reference not built (oracle/_ref)")
    make, nsteps, kw = CASES["basin_sph"]
    cfg = _no_friction(make())
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
    g.close()


def test_nonlinear_fp32k(oracle_built):
    print("CASE leapfrog_nonlinear fp32k basin_sph", flush=True)
    make, nsteps, kw = CASES["basin_sph"]
    o, g = _run_both(make(), nsteps, kw, precision="fp32k")
    Ug, Uo = g.U, o.get_U()
    errs = [float(np.abs(Ug[k] - Uo[k]).max()) for k in range(3)]
    print("errors", errs, flush=True)
    assert errs[0] <= 1e-5
    assert max(errs[1:]) <= 2e-3
    g.close()
