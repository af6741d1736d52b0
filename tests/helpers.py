"""Shared set-up for parity tests: one problem run through the C restatement (or the
compiled reference) and through the HIP backend on identical inputs."""
import numpy as np

import pyoracle
from ptha_amd import configs as C


def small_cartesian_linear(nx=60, ny=50, dx=5000.0, sigma=40e3):
    return C.gaussian_hump_cartesian(nx=nx, ny=ny, dx=dx, sigma=sigma)


def small_spherical(nx=90, ny=80, arcmin=10.0, seed=20261202):
    u = C.unit_sources(nx=nx, ny=ny, arcmin=arcmin, n_src_x=2, n_src_y=2, spacing_km=200, sigma_km=60, seed=seed)
    U = np.zeros((4, ny, nx))
    U[3] = u["elev"]
    U[0] = C.gaussian_on_sphere(u["x"], u["y"], *u["centres"][0], 60e3)
    U[0][U[3] > 0] = 0.0
    return dict(U=U, lw=u["lw"], ll=u["ll"], nx=nx, ny=ny, method="linear", spherical=True)


def small_beach(nx=80, ny=40, dx=10.0):
    return C.beach_rk2(nx=nx, ny=ny, dx=dx)


def open_basin(nx=90, ny=80, arcmin=10.0, method="linear", depth=3000.0):
    """Spherical basin with no land and no walls, a 2 m hump 300 km from the west edge
    and a northward current near the north edge, so waves and flow cross every side
    within tens of steps (a test bed for open boundaries)."""
    cfg = small_spherical(nx=nx, ny=ny, arcmin=arcmin)
    x = np.linspace(0.0, 1.0, nx)[None, :]
    y = np.linspace(0.0, 1.0, ny)[:, None]
    U = np.zeros((4, ny, nx))
    U[3] = -depth + 0.2 * depth * np.sin(3.0 * x + 1.0) * np.cos(2.0 * y)
    U[0] = 2.0 * np.exp(-((x - 0.2) ** 2 + (y - 0.6) ** 2) / 0.02)
    U[2] = 50.0 * np.exp(-((x - 0.7) ** 2) / 0.05) * np.exp(-((y - 0.95) ** 2) / 0.01)
    U[1] = -30.0 * np.exp(-((y - 0.3) ** 2) / 0.05) * np.exp(-((x - 0.97) ** 2) / 0.01)
    return dict(cfg, U=U, method=method)


def nonlinear_leapfrog_cases():
    """Configurations for 'leapfrog_nonlinear' (the nonlinear staggered leapfrog):
    an all-wet spherical basin with currents, a spherical grid with land (wet/dry), the
    Cartesian beach (shore, stage boundary), a dam break onto a dry slope (its wetting
    front engages the Froude limiter), the basin with Flather boundaries.
    Each entry: (make config, steps, domain keywords)."""
    def manning(cfg, n=0.025):
        return dict(cfg, method="leapfrog_nonlinear", manning_squared=np.full((cfg["ny"], cfg["nx"]), n ** 2))

    def dambreak():
        # a 2 m column released onto a dry bed sloping gently up to the east: the thin
        # wetting front engages the Froude limiter (results change without it)
        cfg = C.gaussian_hump_cartesian(nx=96, ny=24, dx=1.0, sigma=10.0)
        x = np.arange(96)[None, :] * np.ones((24, 1))
        U = np.zeros((4, 24, 96))
        U[3] = 0.01 * (x - 30.0)
        U[0] = np.where(x < 30, 2.0, U[3])
        return dict(cfg, U=U)

    return {
        "basin_sph": (lambda: manning(open_basin()), 60, {}),
        "land_sph": (lambda: manning(small_spherical()), 60, {}),
        "beach_cart": (lambda: manning(small_beach()), 80, {}),
        "dambreak_cart": (lambda: manning(dambreak(), 0.02), 120, {"cfl": 0.2}),
        "basin_sph_flather": (lambda: with_flather(manning(open_basin())), 60, {}),
    }


def dispersive_cases():
    """Configurations for the staggered dispersive solve (linear_dispersive_solver_mod.f90):
    a 20 m deep Cartesian channel at 25 m cells (dispersion strong: d^2 / (3 dx^2) = 0.2)
    with a 1 m hump and land on its east side, a spherical basin, 'leapfrog_nonlinear'
    on the channel, and the cell-centred form with 'midpoint' (dispersion also in its
    predictor) on the channel. Each entry: (config, steps, td1, td2, outer
    iterations)."""
    def channel(method="linear"):
        cfg = C.gaussian_hump_cartesian(nx=72, ny=40, dx=25.0, sigma=150.0)
        x = np.arange(72)[None, :] * np.ones((40, 1))
        U = np.zeros((4, 40, 72))
        U[3] = -20.0 + 0.4 * np.maximum(x - 50.0, 0.0)
        U[0] = np.maximum(np.exp(-((x - 25.0) ** 2) / 40.0) * 1.0, U[3])
        cfg = dict(cfg, U=U, method=method)
        if method != "linear":
            cfg["manning_squared"] = np.zeros((40, 72))
        return cfg

    return {
        "channel_linear": (lambda: channel(), 40, 0.0, -1.0, 1),
        "channel_linear_taper": (lambda: channel(), 40, 15.0, 5.0, 2),
        "basin_sph_linear": (lambda: open_basin(), 30, 0.0, -1.0, 1),
        "channel_nonlinear": (lambda: channel("leapfrog_nonlinear"), 40, 0.0, -1.0, 2),
        "channel_midpoint": (lambda: channel("midpoint"), 40, 0.0, -1.0, 1),
        "channel_midpoint_taper": (lambda: channel("midpoint"), 30, 15.0, 5.0, 2),
    }


def with_bc(cfg, kind, sides=(1, 1, 1, 1), amp=None, period=600.0):
    """cfg with one of boundary_mod.f90's other boundary subroutines ('transmissive',
    'stage_radiation_momentum', 'stage_transmissive_momentum') instead of any stage
    boundary."""
    cfg = dict(cfg)
    cfg.pop("stage_bc", None)
    cfg["bc"] = dict(kind=kind, sides=sides, amp=amp, period=period)
    return cfg


def with_flather(cfg, sides=(1, 1, 1, 1), amp=None, period=600.0):
    """cfg with a Flather boundary (boundary_mod.f90:502-720) instead of any stage
    boundary: outside sea at rest, or the sine stage of amplitude amp."""
    cfg = dict(cfg)
    cfg.pop("stage_bc", None)
    cfg["flather"] = dict(sides=sides, amp=amp, period=period)
    return cfg


# finite-volume methods: without a timestep argument they take the CFL step
FV_METHODS = ("rk2", "rk2n", "midpoint", "euler")


def cpu_run(kind, cfg, nsteps, dt=None, record=None, **kw):
    """Run nsteps of evolve_one_step; returns (domain, per-step scalars list)."""
    d = pyoracle.CpuDomain(kind, cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], method=cfg["method"],
                           spherical=cfg["spherical"], **kw)
    d.set_U(cfg["U"])
    if "manning_squared" in cfg:
        d.set_manning_squared(cfg["manning_squared"])
    if "stage_bc" in cfg:
        b = cfg["stage_bc"]
        d.set_stage_bc(b["amp"], b["period"], 0.0, b["sides"])
    if "flather" in cfg:
        d.set_flather_bc(**cfg["flather"])
    if "bc" in cfg:
        d.set_bc(**cfg["bc"])
    if dt is None and cfg["method"] not in FV_METHODS:
        dt = d.stationary_timestep_max()
    hist = []
    for _ in range(nsteps):
        d.evolve(dt, 1)
        if record is not None:
            hist.append(record(d))
    return d, dt, hist


def gpu_domain(cfg, precision="fp64", max_U_variables=("max_stage", "max_speed", "max_flux", "arrival_time",
                                                       "time_of_max_stage", "min_stage"), **kw):
    from ptha_amd.domain import Domain
    g = Domain(cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], timestepping_method=cfg["method"], precision=precision,
               spherical=cfg["spherical"], max_U_variables=max_U_variables, **kw)
    if "manning_squared" in cfg:
        g.set_manning_squared(cfg["manning_squared"])
    g.U = cfg["U"]
    if "stage_bc" in cfg:
        b = cfg["stage_bc"]
        g.set_stage_boundary_sine(b["amp"], b["period"], 0.0, b["sides"])
    if "flather" in cfg:
        g.set_flather_boundary(**cfg["flather"])
    if "bc" in cfg:
        g.set_boundary(**cfg["bc"])
    return g


def cell_areas(cfg):
    """Cell areas (ny, nx) of a configuration, from the same metric build as Domain."""
    from ptha_amd.domain import build_metrics
    m = build_metrics(cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], cfg["spherical"])
    return np.repeat(m["area_cell_y"][:, None], cfg["nx"], axis=1)
