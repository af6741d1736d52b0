"""The C restatement (oracle/swals_oracle.c) against the reference's own Fortran solver
compiled from the reference sources (oracle/_ref/libswals_ref_*.so). Both run on the
CPU here; the restatement follows the reference's operation order, so the states,
scalars, metric arrays and max_U grids agree bit for bit. Skipped where the compiled
reference is not present."""
import numpy as np
import pytest

import helpers as H
import pyoracle
from ptha_amd import configs as C

MAXQ = pyoracle.MAXQ


def _need_ref(spherical):
    if not pyoracle.available("ref", spherical):
        pytest.skip("compiled reference not built (oracle/_ref)")


CASES = {
    "linear_cart": (lambda: H.small_cartesian_linear(), {}, 50),
    "linear_cart_ntl": (lambda: H.small_cartesian_linear(), {"truely_linear": False}, 50),
    "linear_cart_lfc": (lambda: H.small_cartesian_linear(), {"linear_friction_coeff": 1e-4}, 50),
    "llpnf_cart": (lambda: dict(H.small_cartesian_linear(), method="leapfrog_linear_plus_nonlinear_friction",
                                manning_squared=np.full((50, 60), 0.03 ** 2)), {}, 50),
    "linear_sph": (lambda: H.small_spherical(), {}, 60),
    "llpnf_sph": (lambda: dict(H.small_spherical(), method="leapfrog_linear_plus_nonlinear_friction",
                               manning_squared=np.full((80, 90), 0.03 ** 2)), {}, 60),
    "llpnf_sph_chezy": (lambda: dict(H.small_spherical(), method="leapfrog_linear_plus_nonlinear_friction",
                                     manning_squared=np.full((80, 90), 1.0 / 40.0 ** 2)),
                        {"friction_type": "chezy"}, 40),
    "llpnf_sph_ntl": (lambda: dict(H.small_spherical(), method="leapfrog_linear_plus_nonlinear_friction",
                                   manning_squared=np.full((80, 90), 0.03 ** 2)), {"truely_linear": False}, 60),
    "rk2_beach": (lambda: H.small_beach(), {}, 60),
}


def _rk2_sph():
    cfg = H.small_spherical()
    U = cfg["U"].copy()
    U[3] = np.maximum(U[3], -300.0)
    U[0] = np.maximum(U[0], U[3])
    return dict(cfg, U=U, method="rk2", manning_squared=np.full((cfg["ny"], cfg["nx"]), 0.025 ** 2))


CASES["rk2_sph"] = (_rk2_sph, {}, 30)
# the other finite-volume schemes (timestepping_alternatives_include.f90:8-62, 149-336)
for _m in ("euler", "midpoint", "rk2n"):
    CASES[_m + "_beach"] = ((lambda m: lambda: dict(H.small_beach(), method=m))(_m), {}, 40)
    CASES[_m + "_sph"] = ((lambda m: lambda: dict(_rk2_sph(), method=m))(_m), {}, 20)


# Flather boundaries (boundary_mod.f90:502-720): sea at rest outside, or the sine stage
# outside (a boundary function), on all or some sides
CASES["linear_sph_flather"] = (lambda: H.with_flather(H.open_basin()), {}, 60)
CASES["linear_sph_flather_fun"] = (lambda: H.with_flather(H.open_basin(), sides=(1, 0, 0, 1), amp=0.3), {}, 60)
CASES["linear_cart_flather"] = (lambda: H.with_flather(dict(H.open_basin(), spherical=False)), {}, 60)
CASES["llpnf_sph_flather"] = (lambda: H.with_flather(dict(H.open_basin(method="leapfrog_linear_plus_nonlinear_friction"),
                                                          manning_squared=np.full((80, 90), 0.03 ** 2))), {}, 60)
CASES["rk2_beach_flather_fun"] = (lambda: H.with_flather(H.small_beach(), amp=1.0), {}, 60)
CASES["rk2_sph_flather"] = (lambda: H.with_flather(H.open_basin(method="rk2")), {}, 30)
CASES["midpoint_sph_flather_fun"] = (lambda: H.with_flather(H.open_basin(method="midpoint"), sides=(1, 1, 1, 1),
                                                            amp=0.5), {}, 30)
CASES["euler_beach_flather"] = (lambda: H.with_flather(dict(H.small_beach(), method="euler"), sides=(0, 0, 0, 1),
                                                      amp=1.0), {}, 40)


# the other boundary subroutines (boundary_mod.f90:23-75, 166-405)
CASES["linear_sph_transmissive"] = (lambda: H.with_bc(H.open_basin(), "transmissive"), {}, 60)
CASES["rk2_sph_transmissive"] = (lambda: H.with_bc(H.open_basin(method="rk2"), "transmissive"), {}, 30)
CASES["rk2_sph_radiation"] = (lambda: H.with_bc(H.open_basin(method="rk2"), "stage_radiation_momentum"), {}, 30)
CASES["rk2_beach_radiation_fun"] = (lambda: H.with_bc(H.small_beach(), "stage_radiation_momentum", sides=(0, 0, 0, 1),
                                                      amp=1.0), {}, 60)
CASES["midpoint_beach_transmissive_momentum"] = (lambda: H.with_bc(dict(H.small_beach(), method="midpoint"),
                                                                   "stage_transmissive_momentum", sides=(0, 0, 0, 1),
                                                                   amp=1.0), {}, 40)
CASES["linear_sph_radiation_fun"] = (lambda: H.with_bc(H.open_basin(), "stage_radiation_momentum", amp=0.3), {}, 60)
CASES["lfnl_sph_transmissive"] = (lambda: H.with_bc(dict(H.open_basin(method="leapfrog_nonlinear"),
                                                         manning_squared=np.full((80, 90), 0.025 ** 2)), "transmissive"),
                                  {}, 60)

# 'leapfrog_nonlinear' (domain_mod_leapfrog_nonlinear_include.f90)
for _k, (_make, _n, _kw) in H.nonlinear_leapfrog_cases().items():
    CASES["lfnl_" + _k] = (_make, _kw, _n)


@pytest.mark.parametrize("method", ["euler", "midpoint", "rk2n", "rk2"])
def test_fv_fixed_timestep_equals_compiled_reference(oracle_built, method):
    """The finite-volume schemes with a timestep argument (60 % of the stationary CFL
    step) instead of the CFL step."""
    cfg = dict(H.small_beach(), method=method)
    _need_ref(False)
    o = H.cpu_run("oracle", cfg, 0)[0]
    dt = 0.6 * o.stationary_timestep_max()
    o, _, ho = H.cpu_run("oracle", cfg, 30, dt=dt, record=lambda d: d.scalars())
    r, _, hr = H.cpu_run("ref", cfg, 30, dt=dt, record=lambda d: d.scalars())
    assert np.array_equal(o.get_U(), r.get_U())
    for a, b in zip(ho, hr):
        for k in a:
            assert a[k] == b[k], k


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_equals_compiled_reference(oracle_built, case):
    make, kw, nsteps = CASES[case]
    cfg = make()
    _need_ref(cfg["spherical"])
    o, dt, ho = H.cpu_run("oracle", cfg, nsteps, record=lambda d: d.scalars(), **kw)
    r, dt2, hr = H.cpu_run("ref", cfg, nsteps, record=lambda d: d.scalars(), **kw)
    assert dt == dt2
    assert np.array_equal(o.get_U(), r.get_U())
    for a, b in zip(ho, hr):
        for k in ["time", "max_dt", "evolve_step_dt", "boundary_flux_time_integral", "negative_depth_fix_counter",
                  "nsteps_advanced", "bfs_N", "bfs_E", "bfs_S", "bfs_W"]:
            assert a[k] == b[k], k
    for q in MAXQ:
        assert np.array_equal(o.max_U(q), r.max_U(q)), q
    # the reference sums the volume over OpenMP threads: agreement to round-off
    assert o.volume_interior() == pytest.approx(r.volume_interior(), rel=1e-14)


@pytest.mark.parametrize("spherical", [False, True])
def test_metrics_equal_reference(oracle_built, spherical):
    _need_ref(spherical)
    from ptha_amd.domain import build_metrics
    if spherical:
        nx, ny, lw, ll = 90, 80, (15.0, 13.3333), (140.0, -40.0)
    else:
        nx, ny, lw, ll = 60, 50, (3e5, 2.5e5), (0.0, 0.0)
    o = pyoracle.CpuDomain("oracle", nx, ny, lw, ll, spherical=spherical)
    r = pyoracle.CpuDomain("ref", nx, ny, lw, ll, spherical=spherical)
    mo, mr = o.metrics(), r.metrics()
    mp = build_metrics(nx, ny, lw, ll, spherical)
    for k in mo:
        assert np.array_equal(mo[k], mr[k]), k
        if k in mp:
            assert np.array_equal(mp[k], mr[k]), k


@pytest.mark.parametrize("case", sorted(H.dispersive_cases()))
def test_dispersion_equals_compiled_reference(oracle_built, case):
    """The staggered dispersive solve (linear_dispersive_solver_mod.f90: matrix-free
    explicit part, right-hand side, quadratic extrapolation in time, tridiagonal sweeps
    with LAPACK's dgtsv) driven as multidomain_mod.f90:1245-1410 drives one domain:
    bit-identical to the compiled reference, and different from the run without it."""
    make, nsteps, td1, td2, outer = H.dispersive_cases()[case]
    cfg = make()
    _need_ref(cfg["spherical"])
    runs = []
    for kind in ("oracle", "ref"):
        d, dt, _ = H.cpu_run(kind, cfg, 0)
        d.dispersion_setup(td1, td2, 2)
        for _ in range(nsteps):
            d.dispersive_step(dt, outer)
        runs.append(d.get_U())
    assert np.array_equal(runs[0], runs[1])
    plain = H.cpu_run("oracle", cfg, nsteps)[0].get_U()
    assert np.abs(plain[:3] - runs[0][:3]).max() > 1e-6
