"""Batched unit sources on one shared grid (the PTHA unit-source workload).

The reference runs one serial 'linear' SWALS model per unit source, each over the same
grid and bathymetry, with its own duration, recording gauges
(R/examples/austptha_template/SOURCE_ZONES/TEMPLATE/TSUNAMI_UNIT_SOURCE/run_16.sh,
template/model_namelist.in). ``UnitSourceBatch`` advances all of them on one GPU: one
kernel launch per leapfrog step covers every source still running.
"""
import ctypes

import numpy as np

from . import backend as B
from .domain import build_metrics, _p


def leapfrog_times(dt, nsteps):
    """domain%time after 0..nsteps leapfrog steps of size dt, accumulated as the
    reference does: two half-steps per step (domain_mod_linear_solver_include.f90:119,
    571), in FP64."""
    t = np.empty(int(nsteps) + 1)
    x = 0.0
    t[0] = x
    h = float(dt) * 0.5
    for n in range(1, int(nsteps) + 1):
        x = x + h
        x = x + h
        t[n] = x
    return t


def reference_write_schedule(dt, final_time, writeout_frequency):
    """Steps taken and gauge write steps of the reference's unit-source driver
    (examples/generic_example/generic_model.f90:380-405): last_write_time starts at
    -frequency; before each step, if time - last_write_time >= frequency the gauges
    are written and last_write_time advances by the frequency; the run stops before the
    first step taken with time > final_time. Returns (nsteps, sample_steps,
    sample_times)."""
    dt = float(dt)
    freq = float(writeout_frequency)
    last = -freq
    steps, times = [], []
    t, n, h = 0.0, 0, dt * 0.5
    while True:
        if t - last >= freq:
            last = last + freq
            steps.append(n)
            times.append(t)
        if t > final_time:
            break
        t = t + h
        t = t + h
        n += 1
    return n, np.array(steps, dtype=np.int64), np.array(times)


class UnitSourceBatch:
    def __init__(self, nx, ny, lw, lower_left, nsources, timestepping_method="linear", precision="fp32k",
                 spherical=True, coriolis=None, msl_linear=0.0, linear_friction_coeff=0.0, friction_type="manning"):
        lib = B.load()
        self._lib = lib
        self.nx, self.ny, self.nsources = int(nx), int(ny), int(nsources)
        self.lw, self.lower_left = tuple(map(float, lw)), tuple(map(float, lower_left))
        self.precision = precision
        self.spherical = bool(spherical)
        self.coriolis = self.spherical if coriolis is None else bool(coriolis)
        self.metrics = build_metrics(self.nx, self.ny, self.lw, self.lower_left, self.spherical, self.coriolis)
        self.dx = self.metrics["dx"]
        d = B.DomainDesc()
        d.nx, d.ny = self.nx, self.ny
        d.timestepping_method = B.SWH_METHOD[timestepping_method]
        d.precision = B.SWH_PREC[precision]
        d.linear_solver_is_truely_linear = 1
        d.include_coriolis = int(self.coriolis and self.spherical)
        d.spherical = int(self.spherical)
        d.friction_type = 1 if friction_type == "chezy" else 0
        d.exterior_cells_width = 2
        for k in range(4):
            d.boundary_exterior[k] = 1
        d.max_U_mask = 0
        d.max_U_update_frequency = 1
        d.msl_linear = float(msl_linear)
        d.cfl = -1.0
        d.theta = -1.0
        d.linear_friction_coeff = float(linear_friction_coeff)
        d.maximum_timestep = -1.0
        d.arrival_stage_threshold_above_msl_linear = 0.01
        self._keep = {}
        for k in ["distance_bottom_edge", "distance_left_edge", "area_cell_y", "coriolis", "coriolis_bottom_edge",
                  "tanlat_on_radius_earth"]:
            if k in self.metrics:
                arr = np.ascontiguousarray(self.metrics[k], dtype=np.float64)
                self._keep[k] = arr
                setattr(d, k, _p(arr))
        h = ctypes.c_void_p()
        B.check(lib.swh_batch_create(ctypes.byref(d), self.nsources, ctypes.byref(h)), "swh_batch_create")
        self._h = h
        self._ng = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.swh_batch_destroy(self._h)
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

    def set_elevation(self, elev):
        a = np.ascontiguousarray(elev, dtype=np.float64)
        assert a.shape == (self.ny, self.nx)
        B.check(self._lib.swh_batch_upload_elevation(self._h, _p(a)), "swh_batch_upload_elevation")

    def set_manning_squared(self, m):
        a = np.ascontiguousarray(np.broadcast_to(m, (self.ny, self.nx)), dtype=np.float64)
        B.check(self._lib.swh_batch_upload_manning_squared(self._h, _p(a)), "swh_batch_upload_manning_squared")

    def set_nsteps(self, nsteps):
        a = np.ascontiguousarray(nsteps, dtype=np.int32)
        assert a.shape == (self.nsources,)
        B.check(self._lib.swh_batch_set_nsteps(self._h, a.ctypes.data_as(B._PI32)), "swh_batch_set_nsteps")

    def set_source(self, k, stage, uh=None, vh=None):
        U3 = np.zeros((3, self.ny, self.nx))
        U3[0] = stage
        if uh is not None:
            U3[1] = uh
        if vh is not None:
            U3[2] = vh
        B.check(self._lib.swh_batch_upload_source(self._h, int(k), _p(U3)), "swh_batch_upload_source")

    def get_source(self, k):
        U3 = np.empty((3, self.ny, self.nx))
        B.check(self._lib.swh_batch_download_source(self._h, int(k), _p(U3)), "swh_batch_download_source")
        return U3

    def setup_point_gauges(self, xy, every=None, max_samples=None, sample_steps=None):
        """Gauges at the cells nearest xy, sampled at steps 0, every, 2 every, ...
        (max_samples of them), or at the given sample_steps (for example the write
        steps of ``reference_write_schedule``)."""
        xy = np.asarray(xy, dtype=np.float64).reshape(-1, 2)
        i = 1 + np.minimum(np.floor((xy[:, 0] - self.lower_left[0]) / self.dx[0]), self.nx - 1)
        j = 1 + np.minimum(np.floor((xy[:, 1] - self.lower_left[1]) / self.dx[1]), self.ny - 1)
        ij = np.ascontiguousarray(np.stack([i, j], axis=1).astype(np.int32))
        self.gauge_ij = ij
        self._ng = len(ij)
        if sample_steps is not None:
            st = np.ascontiguousarray(sample_steps, dtype=np.int64)
            self._gmax = len(st)
            B.check(self._lib.swh_batch_gauges_set_schedule(self._h, len(ij), ij.ctypes.data_as(B._PI32), len(st),
                                                            st.ctypes.data_as(B._PI64)),
                    "swh_batch_gauges_set_schedule")
        else:
            self._gmax = int(max_samples)
            B.check(self._lib.swh_batch_gauges_set(self._h, len(ij), ij.ctypes.data_as(B._PI32), int(every),
                                                   int(max_samples)), "swh_batch_gauges_set")
        return ij

    def gauge_series(self, k):
        vals = np.empty(3 * max(self._ng, 1) * self._gmax)
        ns = B._I32()
        B.check(self._lib.swh_batch_gauges_download(self._h, int(k), _p(vals), ctypes.byref(ns)),
                "swh_batch_gauges_download")
        return vals[:3 * self._ng * ns.value].reshape(ns.value, 3, self._ng).copy()

    def set_launch(self, block_x, rows_per_block, variant=2):
        """Step-kernel launch shape; variant bits 1 vector lanes, 2 row prefetch,
        4 non-temporal state traffic, 8 sources-fastest block order."""
        B.check(self._lib.swh_batch_set_launch(self._h, int(block_x), int(rows_per_block), int(variant)),
                "swh_batch_set_launch")

    def autotune(self, dt, reps=3):
        """Pick the fastest step-kernel launch shape on this batch (state unchanged).
        Returns ((block_x, rows_per_block, variant), ms_per_step)."""
        best = np.zeros(3, dtype=np.int32)
        ms = ctypes.c_double()
        B.check(self._lib.swh_batch_autotune(self._h, float(dt), int(reps),
                                             best.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(ms)),
                "swh_batch_autotune")
        return tuple(int(v) for v in best), ms.value

    def evolve(self, dt, nsteps):
        B.check(self._lib.swh_batch_evolve(self._h, float(dt), int(nsteps)), "swh_batch_evolve")

    def evolve_timed(self, dt, nsteps):
        ms = ctypes.c_double()
        B.check(self._lib.swh_batch_evolve_timed(self._h, float(dt), int(nsteps), ctypes.byref(ms)),
                "swh_batch_evolve_timed")
        return ms.value

    def synchronize(self):
        B.check(self._lib.swh_batch_synchronize(self._h), "swh_batch_synchronize")

    def superpose(self, events, slip=None, weights=None, msl=0.0, summary="series", timed=False):
        """Earthquake events from the recorded unit-source gauge series, on the GPU
        (make_tsunami_event_from_unit_sources, sum_tsunami_unit_sources.R:555-711).

        events   list of source-index lists (0-based), one per event
        slip     per-event uniform slip (weights 1, sum scaled by the slip), or None
        weights  per-event lists of per-source slips (stochastic slip), or None
        summary  "series": (n_events, samples, 3, gauges) array;
                 "max_stage": (n_events, gauges) maximum stage over time
        Returns (values, nvalid) where nvalid[e] is the number of samples recorded by
        every source of event e (series past it are not meaningful); with timed=True
        also the device time of the products in ms."""
        ne = len(events)
        ptr = np.zeros(ne + 1, dtype=np.int32)
        for e, ev in enumerate(events):
            ptr[e + 1] = ptr[e] + len(ev)
        src = np.ascontiguousarray(np.concatenate([np.asarray(ev, dtype=np.int32) for ev in events])
                                   if ptr[-1] else np.zeros(1, dtype=np.int32))
        w = None
        if weights is not None:
            w = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in weights]))
            if len(w) != ptr[-1]:
                raise ValueError("weights must give one slip per event source")
        sl = None if slip is None else np.ascontiguousarray(np.asarray(slip, dtype=np.float64).reshape(ne))
        mode = {"series": 0, "max_stage": 1}[summary]
        ncol = self._gmax * 3 * self._ng if mode == 0 else self._ng
        out = np.empty((ne, ncol))
        nvalid = np.zeros(ne, dtype=np.int32)
        ms = ctypes.c_double()
        B.check(self._lib.swh_batch_superpose(self._h, ne, ptr.ctypes.data_as(B._PI32), src.ctypes.data_as(B._PI32),
                                              None if w is None else _p(w), None if sl is None else _p(sl),
                                              float(msl), mode, _p(out), nvalid.ctypes.data_as(B._PI32),
                                              ctypes.byref(ms)), "swh_batch_superpose")
        vals = out.reshape(ne, self._gmax, 3, self._ng) if mode == 0 else out
        return (vals, nvalid, ms.value) if timed else (vals, nvalid)

    def status(self):
        step = ctypes.c_int64()
        na = B._I32()
        B.check(self._lib.swh_batch_status(self._h, ctypes.byref(step), ctypes.byref(na)), "swh_batch_status")
        return step.value, na.value

    def launch_counts(self):
        """(launches, sources summed over launches, source steps summed over launches) of
        the step kernels so far, the autotuner's included."""
        v = [ctypes.c_int64() for _ in range(3)]
        B.check(self._lib.swh_batch_launch_counts(self._h, *[ctypes.byref(x) for x in v]), "swh_batch_launch_counts")
        return tuple(x.value for x in v)
