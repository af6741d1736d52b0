"""Synthetic workloads for the SWALS hot path (all seeds fixed).

Every generator returns plain numpy FP64 data in the reference's layout: ``U`` has
shape (4, ny, nx) with planes [stage, uh, vh, elevation] (bytes identical to the
reference's Fortran ``U(nx, ny, 4)``), plus the grid geometry (``lw``, ``ll``,
``nx``, ``ny``). The same data feeds the HIP backend, the C restatement and the
compiled reference, so all three see identical inputs.

Workloads (sizes as stated; the ``scale`` argument shrinks a grid for quick tests):

* ``gaussian_hump_cartesian``: 'linear', 400 x 400 cells at 1 km, flat bed 4000 m,
  1 m Gaussian hump (sigma 20 km) at the centre, reflective walls; 16 gauges on a
  4 x 4 lattice at 50/150/250/350 km.
* ``beach_rk2``: 'rk2' finite volume with wetting/drying on 2000 x 1000 cells at 10 m,
  a 1:50 planar beach from -200 m to +20 m with 0.5 m RMS Gaussian-correlated
  roughness (20 m correlation length), Manning n = 0.025, a sinusoidal west stage
  boundary (1 m, 600 s) and walls elsewhere.
* ``unit_sources``: 512 Gaussian humps (1 m, sigma 25 km) on a 32 x 16 lattice at
  50 km spacing over one 1200 x 1200 spherical 'linear' grid at 1 arcmin, run
  durations U(2 h, 12 h), 64 gauges.
* ``power_law_bathymetry``: the power-law-spectrum (slope -2) depth generator with a
  shelf ramp to land near the east edge, used by the spherical workloads.
"""
import math

import numpy as np

WALL_ELEVATION = 1.0e6  # global_mod.f90:44
GRAVITY = 9.8
RADIUS_EARTH = 6371000.0


def cell_centres(nx, ny, lw, ll):
    """Cell-centre coordinates as in domain_mod.f90:922-928."""
    x = ll[0] + ((np.arange(1, nx + 1) - 0.5) / (nx * 1.0)) * lw[0]
    y = ll[1] + ((np.arange(1, ny + 1) - 0.5) / (ny * 1.0)) * lw[1]
    return x, y


def add_walls(elev, sides=(1, 1, 1, 1)):
    """Reflective walls as an outer ring at wall_elevation (sides = N, E, S, W)."""
    e = elev.copy()
    if sides[0]:
        e[-1, :] = WALL_ELEVATION
    if sides[1]:
        e[:, -1] = WALL_ELEVATION
    if sides[2]:
        e[0, :] = WALL_ELEVATION
    if sides[3]:
        e[:, 0] = WALL_ELEVATION
    return e


def gaussian_hump_cartesian(nx=400, ny=400, dx=1000.0, depth=4000.0, amp=1.0, sigma=20e3):
    lw = (nx * dx, ny * dx)
    ll = (0.0, 0.0)
    x, y = cell_centres(nx, ny, lw, ll)
    X, Y = np.meshgrid(x, y)
    cx, cy = lw[0] / 2, lw[1] / 2
    U = np.zeros((4, ny, nx))
    U[3] = add_walls(np.full((ny, nx), -depth))
    U[0] = amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sigma ** 2))
    U[0][U[3] >= WALL_ELEVATION] = 0.0
    gx = np.array([50e3, 150e3, 250e3, 350e3]) * (lw[0] / 400e3)
    gy = np.array([50e3, 150e3, 250e3, 350e3]) * (lw[1] / 400e3)
    gauges_xy = np.array([(a, b) for b in gy for a in gx])
    return dict(U=U, lw=lw, ll=ll, nx=nx, ny=ny, gauges_xy=gauges_xy, method="linear",
                spherical=False, nsteps=1000, cfl=0.7)


def gaussian_random_field(nx, ny, dx, rms, corr_len, seed):
    """Gaussian-correlated field: white noise filtered by a Gaussian kernel in Fourier
    space (correlation exp(-r^2 / (2 L^2))), rescaled to the requested RMS."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((ny, nx))
    kx = np.fft.fftfreq(nx, d=dx) * 2 * math.pi
    ky = np.fft.fftfreq(ny, d=dx) * 2 * math.pi
    KX, KY = np.meshgrid(kx, ky)
    filt = np.exp(-(KX ** 2 + KY ** 2) * corr_len ** 2 / 4.0)
    f = np.real(np.fft.ifft2(np.fft.fft2(noise) * filt))
    f -= f.mean()
    f *= rms / max(f.std(), 1e-300)
    return f


def beach_rk2(nx=2000, ny=1000, dx=10.0, seed=20261202, manning_n=0.025, bc_amp=1.0, bc_period=600.0):
    lw = (nx * dx, ny * dx)
    ll = (0.0, 0.0)
    x, y = cell_centres(nx, ny, lw, ll)
    # 1:50 beach over the last 11 km of the 20 km domain (scaled with the domain length)
    toe = lw[0] - 11000.0 * (lw[0] / 20000.0)
    slope = 1.0 / 50.0 * (20000.0 / lw[0])
    base = np.clip(-200.0 + (x - toe) * slope, -200.0, 20.0)
    elev = np.broadcast_to(base, (ny, nx)) + gaussian_random_field(nx, ny, dx, 0.5, 20.0, seed)
    elev = add_walls(elev, sides=(1, 1, 1, 0))
    U = np.zeros((4, ny, nx))
    U[3] = elev
    U[0] = np.maximum(0.0, elev)
    manning_sq = np.full((ny, nx), manning_n ** 2)
    return dict(U=U, lw=lw, ll=ll, nx=nx, ny=ny, method="rk2", spherical=False,
                manning_squared=manning_sq, stage_bc=dict(amp=bc_amp, period=bc_period, sides=(0, 0, 0, 1)),
                duration=3600.0)


def power_law_bathymetry(nx, ny, seed, mean=-4000.0, std=1000.0, slope=-2.0, shelf=True, land=50.0,
                         shelf_fraction=0.08):
    """Random depth field with a power spectrum ~ k^slope (amplitude ~ k^(slope/2)),
    rescaled to the given mean/std, plus an eastward ramp to +land over the last
    shelf_fraction of the columns."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx)
    ky = np.fft.fftfreq(ny)
    KX, KY = np.meshgrid(kx, ky)
    k = np.sqrt(KX ** 2 + KY ** 2)
    k[0, 0] = 1.0
    amp = k ** (slope / 2.0)
    amp[0, 0] = 0.0
    spec = amp * np.exp(2j * math.pi * rng.random((ny, nx)))
    f = np.real(np.fft.ifft2(spec))
    f = (f - f.mean()) / max(f.std(), 1e-300)
    elev = mean + std * f
    elev = np.minimum(elev, -200.0)
    if shelf:
        ncol = max(2, int(round(nx * shelf_fraction)))
        w = np.clip((np.arange(nx) - (nx - ncol)) / float(ncol), 0.0, 1.0)
        elev = elev * (1 - w) + land * w
    return elev


def unit_sources(nx=1200, ny=1200, n_src_x=32, n_src_y=16, spacing_km=50.0, sigma_km=25.0, seed=20261202,
                 ll=(140.0, -40.0), arcmin=1.0, n_gauges_side=8, dur_range_h=(2.0, 12.0)):
    """Batched unit-source workload on one spherical grid (shared bathymetry)."""
    lw = (nx * arcmin / 60.0, ny * arcmin / 60.0)
    x, y = cell_centres(nx, ny, lw, ll)
    elev = add_walls(power_law_bathymetry(nx, ny, seed))
    cx0 = ll[0] + lw[0] / 2
    cy0 = ll[1] + lw[1] / 2
    coslat = math.cos(math.radians(cy0))
    dlon = spacing_km * 1e3 / (RADIUS_EARTH * coslat) * 180 / math.pi
    dlat = spacing_km * 1e3 / RADIUS_EARTH * 180 / math.pi
    cx = cx0 + (np.arange(n_src_x) - (n_src_x - 1) / 2) * dlon
    cy = cy0 + (np.arange(n_src_y) - (n_src_y - 1) / 2) * dlat
    centres = np.array([(a, b) for b in cy for a in cx])
    rng = np.random.default_rng(seed + 1)
    durations = rng.uniform(dur_range_h[0] * 3600.0, dur_range_h[1] * 3600.0, size=len(centres))
    gx = ll[0] + (np.arange(n_gauges_side) + 0.5) / n_gauges_side * lw[0]
    gy = ll[1] + (np.arange(n_gauges_side) + 0.5) / n_gauges_side * lw[1]
    gauges_xy = np.array([(a, b) for b in gy for a in gx])
    return dict(elev=elev, lw=lw, ll=ll, nx=nx, ny=ny, x=x, y=y, centres=centres, sigma_km=sigma_km,
                durations=durations, gauges_xy=gauges_xy, method="linear", spherical=True, cfl=0.7,
                gauge_interval=60.0)


def unit_source_stage(cfg, k):
    """Initial stage of unit source k of ``unit_sources``: a 1 m Gaussian of width
    sigma_km at the source centre (local equirectangular distance, evaluated as the
    product of its longitude and latitude factors), zero on land."""
    x, y = cfg["x"], cfg["y"]
    cx, cy = cfg["centres"][k]
    sig = cfg["sigma_km"] * 1e3
    gx = np.exp(-(np.radians(x - cx) * RADIUS_EARTH * math.cos(math.radians(cy))) ** 2 / (2 * sig ** 2))
    gy = np.exp(-(np.radians(y - cy) * RADIUS_EARTH) ** 2 / (2 * sig ** 2))
    stage = np.outer(gy, gx)
    stage[cfg["elev"] >= 0.0] = 0.0
    return stage


def gaussian_on_sphere(x, y, cx, cy, sigma_m, amp=1.0):
    """Gaussian hump of width sigma_m (metres) centred at (cx, cy) degrees, using the
    local equirectangular distance."""
    X, Y = np.meshgrid(x, y)
    dxm = np.radians(X - cx) * RADIUS_EARTH * np.cos(np.radians(cy))
    dym = np.radians(Y - cy) * RADIUS_EARTH
    return amp * np.exp(-(dxm ** 2 + dym ** 2) / (2 * sigma_m ** 2))


def gauge_site_index(xy, ll, dx, nx, ny):
    """Nearest-cell gauge index as in point_gauge_mod.f90:329-336 (1-based (i, j))."""
    xy = np.asarray(xy, dtype=np.float64)
    i = 1 + np.minimum(np.floor((xy[:, 0] - ll[0]) / dx[0]), nx - 1)
    j = 1 + np.minimum(np.floor((xy[:, 1] - ll[1]) / dx[1]), ny - 1)
    return np.stack([i, j], axis=1).astype(np.int32)


def adjoint_inversion(nx=2048, ny=2048, arcmin=1.0, ll=(140.0, -45.0), n_src_x=40, n_src_y=20, spacing_km=20.0,
                      sigma_km=15.0, n_gauges=32, hours=4.0, gauge_interval=60.0, seed=20261202, noise=0.01):
    """Source-inversion workload: 'linear' leapfrog on a spherical grid, initial stage a
    weighted sum of Gaussian unit sources on a lattice (weights ~ N(0, 1), the hidden
    truth), gauges on a ring around the source region, synthetic observations with 1%
    multiplicative Gaussian noise added by the caller (see ``noise``)."""
    lw = (nx * arcmin / 60.0, ny * arcmin / 60.0)
    x, y = cell_centres(nx, ny, lw, ll)
    elev = add_walls(power_law_bathymetry(nx, ny, seed + 7))
    cx0, cy0 = ll[0] + 0.45 * lw[0], ll[1] + 0.5 * lw[1]
    coslat = math.cos(math.radians(cy0))
    dlon = spacing_km * 1e3 / (RADIUS_EARTH * coslat) * 180 / math.pi
    dlat = spacing_km * 1e3 / RADIUS_EARTH * 180 / math.pi
    lon = cx0 + (np.arange(n_src_x) - (n_src_x - 1) / 2) * dlon
    lat = cy0 + (np.arange(n_src_y) - (n_src_y - 1) / 2) * dlat
    rng = np.random.default_rng(seed)
    w_true = rng.standard_normal((n_src_y, n_src_x))
    ang = 2 * math.pi * (np.arange(n_gauges) + 0.5) / n_gauges
    rad = 0.3 * min(lw)
    gauges_xy = np.stack([cx0 + rad * np.cos(ang), cy0 + rad * np.sin(ang)], axis=1)
    return dict(nx=nx, ny=ny, lw=lw, ll=ll, x=x, y=y, elev=elev, src_lon=lon, src_lat=lat, sigma_km=sigma_km,
                w_true=w_true, gauges_xy=gauges_xy, hours=hours, gauge_interval=gauge_interval, noise=noise,
                seed=seed)


def nested_small():
    """A three-level nest in the layout of the nested multidomain workload, scaled
    down for parity tests: a 'linear' parent (lon 100-120, lat 60S-48S at 20
    arcmin), a 'leapfrog_linear_plus_nonlinear_friction' child at 1/3 of its cell
    size (Manning n = 0.03) and an 'rk2' grandchild at 1/9, stepped 3 times per
    global step. The shore runs north-south at lon 109.5 inside the grandchild
    (4000 m deep to the west, land to +50 m to the east), and a 1 m Gaussian hump
    (sigma 50 km) sits in the child west of the grandchild. Global step 10 s."""
    domains = [
        dict(lw=(20.0, 12.0), ll=(100.0, -60.0), nx=60, ny=36, method="linear"),
        dict(parent=0, ll=(105.0, -57.0), ur=(113.0, -50.0), dx_refinement_factor=3,
             timestepping_refinement_factor=1, method="leapfrog_linear_plus_nonlinear_friction"),
        dict(parent=1, ll=(108.0, -55.0), ur=(110.0, -53.0), dx_refinement_factor=3,
             timestepping_refinement_factor=3, method="rk2"),
    ]
    return dict(domains=domains, dt=10.0, manning_n=0.03, spherical=True)


def nested_small_state(x, y):
    """Initial U (4, ny, nx) of ``nested_small`` at cell centres x (nx,), y (ny,)."""
    X, Y = np.meshgrid(x, y)
    elev = -4000.0 * np.clip((109.5 - X) / 1.0, 0.0, 1.0) + 50.0 * np.clip((X - 109.5) / 0.5, 0.0, 1.0)
    elev = np.where(X < 109.5, np.minimum(elev, -5.0), elev)
    hump = np.exp(-(((X - 107.0) * math.cos(math.radians(-54.0)) * 111e3) ** 2 + ((Y + 54.0) * 111e3) ** 2)
                  / (2 * 50e3 ** 2))
    U = np.zeros((4,) + X.shape)
    U[3] = elev
    U[0] = np.maximum(0.0, elev) + hump * (elev < 0)
    return U


def nested_c3(scale=1, seed=20261202):
    """The three-level nested multidomain workload: a 'linear' parent over lon
    100-200E, lat 60S-0 at 2 arcmin (3000 x 1800), a
    'leapfrog_linear_plus_nonlinear_friction' child at 1/3 of its cell size (600 x 600,
    Manning n = 0.03) and an 'rk2' grandchild at 1/5 of the child's (1000 x 1000, n =
    0.03), on the eastern shelf; power-law bathymetry (slope -2, mean -4000 m, std
    1000 m, no shallower than 200 m) with a shelf ramp over lon 191-197 to +50 m land
    (the shore lies inside the child); the grandchild (lon 193.0-195.2) lies wholly in
    water, at least 24 m deep, so its nesting boundary never crosses the coast; an
    elliptical Gaussian uplift of 3 m peak, sigma 200 km north-south by 50 km
    east-west, at lon 190, lat 26.5S; 6 h simulated.

    ``scale`` > 1 coarsens every grid by that factor (same extents) for quick runs.
    Returns the domain list for a multidomain setup and ``state(x, y)``, which gives
    the initial U of any domain from its cell-centre coordinates; the parent's
    random field is interpolated bilinearly for the nested grids.
    """
    nxp, nyp = 3000 // scale, 1800 // scale
    llp, lwp = (100.0, -60.0), (100.0, 60.0)
    field = power_law_bathymetry(nxp, nyp, seed, shelf=False)
    xp, yp = cell_centres(nxp, nyp, lwp, llp)

    def elevation(x, y):
        # bilinear interpolation of the parent's random field, then the shelf ramp
        fx = np.clip((np.asarray(x) - xp[0]) / (xp[1] - xp[0]), 0, nxp - 1 - 1e-9)
        fy = np.clip((np.asarray(y) - yp[0]) / (yp[1] - yp[0]), 0, nyp - 1 - 1e-9)
        i0, j0 = np.floor(fx).astype(int), np.floor(fy).astype(int)
        tx, ty = fx - i0, fy - j0
        F = (field[np.ix_(j0, i0)] * np.outer(1 - ty, 1 - tx) + field[np.ix_(j0, i0 + 1)] * np.outer(1 - ty, tx) +
             field[np.ix_(j0 + 1, i0)] * np.outer(ty, 1 - tx) + field[np.ix_(j0 + 1, i0 + 1)] * np.outer(ty, tx))
        w = np.clip((np.asarray(x) - 191.0) / 6.0, 0.0, 1.0)[None, :]
        return F * (1.0 - w) + 50.0 * w

    def state(x, y):
        X, Y = np.meshgrid(x, y)
        U = np.zeros((4,) + X.shape)
        U[3] = elevation(x, y)
        dxm = np.radians(X - 190.0) * RADIUS_EARTH * np.cos(np.radians(Y))
        dym = np.radians(Y + 26.5) * RADIUS_EARTH
        uplift = 3.0 * np.exp(-0.5 * ((dxm / 50e3) ** 2 + (dym / 200e3) ** 2))
        U[0] = np.where(U[3] < 0.0, uplift, U[3])
        return U

    a2 = (2.0 / 60.0) * scale
    domains = [
        dict(lw=lwp, ll=llp, nx=nxp, ny=nyp, method="linear"),
        dict(parent=0, ll=(192.0, -30.0), ur=(192.0 + 600 * a2 / 3, -30.0 + 600 * a2 / 3), dx_refinement_factor=3,
             method="leapfrog_linear_plus_nonlinear_friction"),
        dict(parent=1, ll=(193.0, -27.5), ur=(193.0 + 1000 * a2 / 15, -27.5 + 1000 * a2 / 15),
             dx_refinement_factor=5, method="rk2"),
    ]
    return dict(domains=domains, state=state, elevation=elevation, manning_n=0.03, duration=6 * 3600.0,
                spherical=True)
