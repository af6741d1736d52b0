"""Host-side mirror of SWALS's domain_type for the HIP backend.

``Domain`` keeps the reference's field names and call sequence
(propagation/SWALS/src/shallow_water/domain_mod.f90): construct it with the grid
geometry and ``timestepping_method``, which runs the metric part of
``allocate_quantities`` (:852-1014) on the host in FP64, set ``U`` (and
``manning_squared``), then call ``evolve_one_step(timestep)`` (:1679-1813) in a
loop. The state lives on the GPU; ``U`` is only copied back when read.
"""
import ctypes
import math

import numpy as np

from . import backend as B

GRAVITY = 9.8
RADIUS_EARTH = 6371000.0
EARTH_ANGULAR_FREQ = 7.292115e-05  # spherical_mod.f90:15
DEG2RAD = math.acos(-1.0) / 180.0  # spherical_mod.f90:8


def area_lonlat_rectangle_flat(lon1, lat1, dlon, dlat):
    """area_lonlat_rectangle(..., flat=.TRUE.) of spherical_mod.f90:46-50."""
    return (RADIUS_EARTH * RADIUS_EARTH * math.cos((lat1 + 0.5 * dlat) * DEG2RAD) * abs(dlat) * DEG2RAD *
            math.fmod(abs(dlon), 360.0) * DEG2RAD)


def build_metrics(nx, ny, lw, ll, spherical, coriolis=True):
    """Metric arrays of allocate_quantities (domain_mod.f90:918-1014), scalar libm
    calls in the reference's operation order so the values match it bit for bit."""
    dx = (lw[0] / (nx * 1.0), lw[1] / (ny * 1.0))
    x = np.array([ll[0] + ((i - 0.5) / (nx * 1.0)) * lw[0] for i in range(1, nx + 1)])
    y = np.array([ll[1] + ((j - 0.5) / (ny * 1.0)) * lw[1] for j in range(1, ny + 1)])
    m = dict(x=x, y=y, dx=dx)
    if spherical:
        cbe = [math.cos((y[j] - 0.5 * dx[1]) * DEG2RAD) for j in range(ny)]
        cbe.append(math.cos((y[ny - 1] + 0.5 * dx[1]) * DEG2RAD))
        m["tanlat_on_radius_earth"] = np.array([math.tan(y[j] * DEG2RAD) / RADIUS_EARTH for j in range(ny)])
        if coriolis:
            m["coriolis"] = np.array([2.0 * math.sin(y[j] * DEG2RAD) * EARTH_ANGULAR_FREQ for j in range(ny)])
            cb = [2.0 * EARTH_ANGULAR_FREQ * math.sin((y[j] - 0.5 * dx[1]) * DEG2RAD) for j in range(ny)]
            cb.append(2.0 * EARTH_ANGULAR_FREQ * math.sin((y[ny - 1] + 0.5 * dx[1]) * DEG2RAD))
            m["coriolis_bottom_edge"] = np.array(cb)
        m["distance_left_edge"] = np.full(nx + 1, dx[1] * DEG2RAD * RADIUS_EARTH)
        m["distance_bottom_edge"] = np.array([dx[0] * DEG2RAD * RADIUS_EARTH * c for c in cbe])
        m["area_cell_y"] = np.array([area_lonlat_rectangle_flat(x[0] - dx[0] * 0.5, y[j] - dx[1] * 0.5, dx[0], dx[1])
                                     for j in range(ny)])
    else:
        m["distance_left_edge"] = np.full(nx + 1, dx[1])
        m["distance_bottom_edge"] = np.full(ny + 1, dx[0])
        m["area_cell_y"] = np.full(ny, dx[0] * dx[1])
    return m


def _p(a):
    return a.ctypes.data_as(B._PD)


class Domain:
    """One single-grid SWALS domain advanced on the GPU.

    Arrays are numpy float64 of shape (4, ny, nx) for U (planes stage, uh, vh, elev)
    and (ny, nx) for 2-D fields: the byte layout of the reference's Fortran arrays.
    """

    def __init__(self, nx, ny, lw, lower_left, timestepping_method="linear", precision="fp64", spherical=False,
                 coriolis=None, linear_solver_is_truely_linear=True, msl_linear=0.0, linear_friction_coeff=0.0,
                 friction_type="manning", cfl=None, theta=None, max_U_variables=("max_stage",),
                 max_U_update_frequency=1, exterior_cells_width=2, arrival_stage_threshold_above_msl_linear=0.01):
        lib = B.load()
        if timestepping_method not in B.SWH_METHOD:
            raise ValueError("timestepping_method %r is not provided by this backend" % (timestepping_method,))
        if precision not in B.SWH_PREC:
            raise ValueError("precision must be 'fp64' or 'fp32k'")
        self.nx, self.ny = int(nx), int(ny)
        self.lw, self.lower_left = tuple(map(float, lw)), tuple(map(float, lower_left))
        self.timestepping_method = timestepping_method
        self.precision = precision
        self.spherical = bool(spherical)
        self.coriolis = self.spherical if coriolis is None else bool(coriolis)
        self.msl_linear = float(msl_linear)
        self.linear_solver_is_truely_linear = bool(linear_solver_is_truely_linear)
        self.max_U_variables = tuple(max_U_variables or ())
        self.metrics = build_metrics(self.nx, self.ny, self.lw, self.lower_left, self.spherical, self.coriolis)
        self.dx = self.metrics["dx"]
        self.x, self.y = self.metrics["x"], self.metrics["y"]
        d = B.DomainDesc()
        d.nx, d.ny = self.nx, self.ny
        d.timestepping_method = B.SWH_METHOD[timestepping_method]
        d.precision = B.SWH_PREC[precision]
        d.linear_solver_is_truely_linear = int(self.linear_solver_is_truely_linear)
        d.include_coriolis = int(self.coriolis and self.spherical)
        d.spherical = int(self.spherical)
        d.friction_type = 1 if friction_type == "chezy" else 0
        d.exterior_cells_width = int(exterior_cells_width)
        for k in range(4):
            d.boundary_exterior[k] = 1
        mask = 0
        for v in self.max_U_variables:
            mask |= B.MAXQ_BITS[v]
        d.max_U_mask = mask
        d.max_U_update_frequency = int(max_U_update_frequency)
        d.msl_linear = self.msl_linear
        d.cfl = float(cfl) if cfl else -1.0
        d.theta = float(theta) if theta else -1.0
        d.linear_friction_coeff = float(linear_friction_coeff)
        d.maximum_timestep = -1.0
        d.arrival_stage_threshold_above_msl_linear = float(arrival_stage_threshold_above_msl_linear)
        self._keep = {}
        for k in ["distance_bottom_edge", "distance_left_edge", "area_cell_y", "coriolis", "coriolis_bottom_edge",
                  "tanlat_on_radius_earth"]:
            if k in self.metrics:
                arr = np.ascontiguousarray(self.metrics[k], dtype=np.float64)
                self._keep[k] = arr
                setattr(d, k, _p(arr))
        self._desc = d
        h = ctypes.c_void_p()
        B.check(lib.swh_domain_create(ctypes.byref(d), ctypes.byref(h)), "swh_domain_create")
        self._h = h
        self._lib = lib
        self._gauge_n = 0

    # -- lifetime --
    def close(self):
        if getattr(self, "_h", None):
            self._lib.swh_domain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state --
    @property
    def U(self):
        out = np.empty((4, self.ny, self.nx))
        B.check(self._lib.swh_download_U(self._h, _p(out)), "swh_download_U")
        return out

    @U.setter
    def U(self, value):
        a = np.ascontiguousarray(value, dtype=np.float64)
        if a.shape != (4, self.ny, self.nx):
            raise ValueError("U must have shape (4, ny, nx) = %s" % ((4, self.ny, self.nx),))
        B.check(self._lib.swh_upload_U(self._h, _p(a)), "swh_upload_U")

    def set_state(self, stage, uh=None, vh=None):
        """A new initial state on the same grid: stage, UH, VH as (ny, nx) arrays (None:
        zeros); the elevation stays (swh_upload_state)."""
        def plane(v):
            if v is None:
                return None
            a = np.ascontiguousarray(v, dtype=np.float64)
            if a.shape != (self.ny, self.nx):
                raise ValueError("planes must have shape (ny, nx) = %s" % ((self.ny, self.nx),))
            return a
        planes = [plane(stage), plane(uh), plane(vh)]
        ptrs = [_p(a) if a is not None else ctypes.POINTER(ctypes.c_double)() for a in planes]
        B.check(self._lib.swh_upload_state(self._h, *ptrs), "swh_upload_state")

    def set_manning_squared(self, m):
        a = np.ascontiguousarray(np.broadcast_to(m, (self.ny, self.nx)), dtype=np.float64)
        B.check(self._lib.swh_upload_manning_squared(self._h, _p(a)), "swh_upload_manning_squared")

    def set_time(self, t):
        B.check(self._lib.swh_set_time(self._h, float(t)), "swh_set_time")

    def set_forcing_patch(self, start_time, end_time, i0, i1, j0, j1, k0, k1, work):
        """forcing_patch_type + apply_forcing_patch (forcing_mod.f90:24-56, 158-239):
        work (k1-k0+1, j1-j0+1, i1-i0+1) is added to U(i0:i1, j0:j1, k0:k1) (1-based,
        variables 1 stage .. 4 elevation) at a constant rate over [start_time, end_time]."""
        w = np.ascontiguousarray(work, dtype=np.float64)
        if w.shape != (k1 - k0 + 1, j1 - j0 + 1, i1 - i0 + 1):
            raise ValueError("work must have shape (k1-k0+1, j1-j0+1, i1-i0+1)")
        B.check(self._lib.swh_set_forcing_patch(self._h, float(start_time), float(end_time), int(i0), int(i1),
                                                int(j0), int(j1), int(k0), int(k1), _p(w)), "swh_set_forcing_patch")

    def clear_forcing_patch(self):
        B.check(self._lib.swh_clear_forcing_patch(self._h), "swh_clear_forcing_patch")

    def set_stage_boundary_sine(self, amp, period, msl=0.0, sides=(0, 0, 0, 1)):
        s = (B._I32 * 4)(*[int(bool(v)) for v in sides])
        B.check(self._lib.swh_set_stage_boundary_sine(self._h, float(amp), float(period), float(msl), s),
                "swh_set_stage_boundary_sine")

    BOUNDARY_KINDS = {"stage_transmissive_normal_momentum": 0, "flather": 1, "transmissive": 2,
                      "stage_radiation_momentum": 3, "stage_transmissive_momentum": 4}

    def set_boundary(self, kind, sides=(1, 1, 1, 1), amp=None, period=1.0, msl=0.0):
        """domain%boundary_subroutine => one of boundary_mod.f90's subroutines (kind: a key
        of BOUNDARY_KINDS) on sides [N, E, S, W]; amp None: no boundary_function, else the
        stage msl + amp*sin(2*pi*t/period)."""
        s = (B._I32 * 4)(*[int(bool(v)) for v in sides])
        B.check(self._lib.swh_set_boundary(self._h, self.BOUNDARY_KINDS[kind], s, int(amp is not None),
                                           float(amp or 0.0), float(period), float(msl)), "swh_set_boundary")

    def set_flather_boundary(self, sides=(1, 1, 1, 1), amp=None, period=1.0, msl=0.0):
        """domain%boundary_subroutine => flather_boundary (boundary_mod.f90:502-720) on
        sides [N, E, S, W]. amp None: no boundary_function (outside sea at rest at
        msl_linear); else the outside stage msl + amp*sin(2*pi*t/period), zero momenta."""
        s = (B._I32 * 4)(*[int(bool(v)) for v in sides])
        use_fun = amp is not None
        B.check(self._lib.swh_set_flather_boundary(self._h, s, int(use_fun), float(amp or 0.0), float(period),
                                                   float(msl)), "swh_set_flather_boundary")

    # -- stepping --
    def evolve_one_step(self, timestep=None):
        B.check(self._lib.swh_evolve_one_step(self._h, -1.0 if timestep is None else float(timestep)),
                "swh_evolve_one_step")

    def evolve_steps(self, timestep, nsteps):
        B.check(self._lib.swh_evolve_steps(self._h, -1.0 if timestep is None else float(timestep), int(nsteps)),
                "swh_evolve_steps")

    def synchronize(self):
        B.check(self._lib.swh_synchronize(self._h), "swh_synchronize")

    def set_launch(self, block_x=256, rows_per_block=32, variant=2):
        """Launch shape of the leapfrog step kernel (results do not depend on it);
        variant bits: 1 vector lanes, 2 row prefetch, 4 non-temporal state traffic,
        32 one step per launch in evolve_steps (no two-step launches)."""
        B.check(self._lib.swh_domain_set_launch(self._h, int(block_x), int(rows_per_block), int(variant)),
                "swh_domain_set_launch")

    def set_launch2(self, block_x=256, rows_per_block=0):
        """Launch shape of the two-step leapfrog kernel of evolve_steps (results do not
        depend on it); rows_per_block 0: chosen from the kernel's occupancy."""
        B.check(self._lib.swh_domain_set_launch2(self._h, int(block_x), int(rows_per_block)),
                "swh_domain_set_launch2")

    # -- diagnostics --
    def scalars(self):
        s = B.Scalars()
        B.check(self._lib.swh_get_scalars(self._h, ctypes.byref(s)), "swh_get_scalars")
        return dict(time=s.time, max_dt=s.max_dt, evolve_step_dt=s.evolve_step_dt, dt_last_update=s.dt_last_update,
                    boundary_flux_store=list(s.boundary_flux_store),
                    boundary_flux_evolve_integral=s.boundary_flux_evolve_integral,
                    boundary_flux_time_integral=s.boundary_flux_time_integral,
                    negative_depth_fix_counter=s.negative_depth_fix_counter, nsteps_advanced=s.nsteps_advanced)

    @property
    def time(self):
        return self.scalars()["time"]

    def volume_interior(self):
        v = ctypes.c_double()
        B.check(self._lib.swh_volume_interior(self._h, ctypes.byref(v)), "swh_volume_interior")
        return v.value

    def mass_balance_interior(self):
        """volume_interior + boundary_flux_time_integral (domain_mod.f90:1891-1904)."""
        return self.volume_interior() + self.scalars()["boundary_flux_time_integral"]

    STATS_KEYS = ("max_stage", "max_speed", "min_stage", "min_speed", "energy_potential_on_rho",
                  "energy_kinetic_on_rho", "energy_total_on_rho")

    def statistics(self):
        """compute_domain_statistics (domain_mod.f90:683-849) on the device, as a dict."""
        st = B.DomainStats()
        B.check(self._lib.swh_domain_statistics(self._h, ctypes.byref(st)), "swh_domain_statistics")
        return {k: getattr(st, k) for k in self.STATS_KEYS}

    def check_stability(self):
        """Number of interior values that are NaN or beyond +-HUGE(1.0)/2
        (check_multidomain_stability's scan, multidomain_mod.f90:348-455)."""
        n = ctypes.c_int64()
        B.check(self._lib.swh_domain_check_stability(self._h, ctypes.byref(n)), "swh_domain_check_stability")
        return n.value

    def stationary_timestep_max(self):
        v = ctypes.c_double()
        B.check(self._lib.swh_stationary_timestep_max(self._h, ctypes.byref(v)), "swh_stationary_timestep_max")
        return v.value

    # -- dispersion (linear_dispersive_solver_mod.f90; staggered methods) --
    def dispersion_enable(self, td1=0.0, td2=-1.0, niter=2):
        B.check(self._lib.swh_dispersion_enable(self._h, float(self.dx[0]), float(self.dx[1]), float(td1), float(td2),
                                                int(niter)), "swh_dispersion_enable")

    def dispersion_store(self):
        B.check(self._lib.swh_dispersion_store_last_U(self._h), "swh_dispersion_store_last_U")

    def dispersion_append(self):
        B.check(self._lib.swh_dispersion_append_extrapolation(self._h), "swh_dispersion_append_extrapolation")

    def dispersion_solve(self, rhs_is_up_to_date, estimate_solution_forward_in_time, forward_time):
        B.check(self._lib.swh_dispersion_solve(self._h, int(bool(rhs_is_up_to_date)),
                                               int(bool(estimate_solution_forward_in_time)), float(forward_time)),
                "swh_dispersion_solve")

    def dispersive_step(self, dt, outer_iterations=1):
        """One step of a dispersive domain as multidomain_mod.f90:1245-1410 runs it: the
        extrapolation level appended, U stored, the shallow-water step, then the outer
        dispersive iterations (the first with a new right-hand side and the solution
        extrapolated forward in time)."""
        self.dispersion_append()
        self.dispersion_store()
        self.evolve_one_step(dt)
        t = self.time
        for it in range(outer_iterations):
            self.dispersion_solve(it > 0, it == 0, t)

    def snapshot(self):
        """Pack the current state and time on the device for output (swh_snapshot) and
        return at once; fetch it later with snapshot_download while stepping goes on."""
        B.check(self._lib.swh_snapshot(self._h), "swh_snapshot")

    def snapshot_download(self, out=None):
        """(U (4, ny, nx) float64, time) of the last snapshot. out: an optional
        preallocated (ideally pinned) float64 array of that shape."""
        if out is None:
            out = np.empty((4, self.ny, self.nx))
        if out.shape != (4, self.ny, self.nx) or out.dtype != np.float64 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float64 array of shape (4, ny, nx)")
        t = ctypes.c_double(0.0)
        B.check(self._lib.swh_snapshot_download(self._h, _p(out), ctypes.byref(t)), "swh_snapshot_download")
        return out, t.value

    def max_U(self, name):
        out = np.empty((self.ny, self.nx))
        B.check(self._lib.swh_download_max_U(self._h, B.MAXQ_BITS[name], _p(out)), "swh_download_max_U")
        return out

    # -- gauges (point_gauge_mod.f90) --
    def setup_point_gauges(self, xy=None, max_samples=100000, ij=None):
        """Nearest-cell gauges: site index = 1 + min(floor((xy - ll)/dx), nx - 1)
        (point_gauge_mod.f90:329-336), or the 1-based cells ij (n, 2) directly."""
        if ij is None:
            xy = np.asarray(xy, dtype=np.float64).reshape(-1, 2)
            i = 1 + np.minimum(np.floor((xy[:, 0] - self.lower_left[0]) / self.dx[0]), self.nx - 1)
            j = 1 + np.minimum(np.floor((xy[:, 1] - self.lower_left[1]) / self.dx[1]), self.ny - 1)
            ij = np.stack([i, j], axis=1)
        ij = np.ascontiguousarray(np.asarray(ij).reshape(-1, 2).astype(np.int32))
        if (ij < 1).any() or (ij[:, 0] > self.nx).any() or (ij[:, 1] > self.ny).any():
            raise ValueError("gauge outside the domain")
        self.gauge_ij = ij
        self._gauge_n = len(ij)
        self._gauge_max = int(max_samples)
        B.check(self._lib.swh_gauges_set(self._h, len(ij), ij.ctypes.data_as(B._PI32), int(max_samples)),
                "swh_gauges_set")
        return ij

    def sample_gauges(self):
        B.check(self._lib.swh_gauges_sample(self._h), "swh_gauges_sample")

    def gauge_series(self):
        """Returns (times[nsamples], values[nsamples, 3, ngauges]) with vars [stage, uh, vh]."""
        n = self._gauge_n
        vals = np.empty(3 * max(n, 1) * self._gauge_max)
        times = np.empty(self._gauge_max)
        ns = B._I32()
        B.check(self._lib.swh_gauges_download(self._h, _p(vals), _p(times), ctypes.byref(ns)), "swh_gauges_download")
        k = ns.value
        return times[:k].copy(), vals[:3 * n * k].reshape(k, 3, n).copy()
