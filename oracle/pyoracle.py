"""ctypes front-end for the two CPU checkers (test infrastructure only).

* ``kind="oracle"``: the C restatement in ``oracle/swals_oracle.c``
  (``oracle/_ref/liboracle.so``).
* ``kind="ref"``: the reference's own Fortran solver, compiled from the reference
  sources by ``oracle/Makefile`` into ``oracle/_ref/libswals_ref_{cart,sph}.so``
  (its bind(C) shim is ``oracle/ref/swals_ref_capi.f90``).

Both expose the same operations, so a test can run one problem through both and
through the HIP product. Only ``tests/``, ``smoke()`` and the bench's CPU-baseline
leg import this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

METHODS = {"rk2": 1, "rk2n": 2, "midpoint": 3, "euler": 4, "linear": 5, "leapfrog_linear_plus_nonlinear_friction": 6,
           "leapfrog_nonlinear": 8}
MAXQ = ["max_stage", "max_speed", "max_flux", "arrival_time", "time_of_max_stage", "min_stage"]

_D = ctypes.c_double
_I = ctypes.c_int
_PD = ctypes.POINTER(_D)
_libs = {}


def lib_path(kind, spherical=False):
    if kind == "oracle":
        return os.path.join(REF_DIR, "liboracle.so")
    if kind == "ref_fast":
        # the spherical reference built with its own default optimisation (-O3 and
        # fast-math, as -Ofast), for CPU timing only: not a parity checker
        return os.path.join(REF_DIR, "libswals_ref_sph_fast.so")
    return os.path.join(REF_DIR, "libswals_ref_%s.so" % ("sph" if spherical else "cart"))


def available(kind, spherical=False):
    return os.path.exists(lib_path(kind, spherical))


def _load(kind, spherical):
    key = (kind, spherical if kind != "oracle" else False)
    if key in _libs:
        return _libs[key]
    path = lib_path(kind, spherical)
    if not os.path.exists(path):
        raise FileNotFoundError(path + " is not built (run `make -C oracle`)")
    lib = ctypes.CDLL(path)
    p = "or_" if kind == "oracle" else "ref_"
    fns = {}
    hd = ctypes.c_void_p if kind == "oracle" else _I

    def f(name, res, args):
        fn = getattr(lib, p + name)
        fn.restype = res
        fn.argtypes = args
        fns[name] = fn

    if kind == "oracle":
        f("domain_create", hd, [_I, _I, _PD, _PD, _I, _I, _I, _I, _D, _D, _I, _D, _D])
    else:
        f("domain_create", hd, [_I, _I, _PD, _PD, _I, _I, _D, _D, _I, _D, _D])
    f("domain_destroy", None, [hd])
    f("domain_set_U", None, [hd, _PD])
    f("domain_get_U", None, [hd, _PD])
    f("domain_set_manning_squared", None, [hd, _PD])
    f("domain_get_metrics", None, [hd] + [_PD] * 8)
    f("domain_evolve", _D, [hd, _D, _I])
    f("domain_get_scalars", None, [hd, _PD])
    f("domain_set_time", None, [hd, _D])
    f("domain_volume_interior", _D, [hd])
    f("domain_stationary_timestep_max", _D, [hd])
    f("domain_get_max_U", None, [hd, _I, _PD])
    f("domain_set_stage_bc", None, [hd, _D, _D, _D, ctypes.POINTER(_I)])
    f("domain_set_flather_bc", None, [hd, _I, _D, _D, _D, ctypes.POINTER(_I)])
    f("domain_set_bc", None, [hd, _I, _I, _D, _D, _D, ctypes.POINTER(_I)])
    f("domain_dispersion_setup", None, [hd, _D, _D, _I])
    f("domain_dispersion_store", None, [hd])
    f("domain_dispersion_append", None, [hd])
    f("domain_dispersion_solve", None, [hd, _I, _I, _D])
    f("domain_set_record_max_U", None, [hd, _I])
    if kind != "oracle":
        f("domain_statistics", None, [hd, _PD])
        f("domain_sample", None, [hd, _I, ctypes.POINTER(_I), _PD])
        f("domain_set_forcing", None, [hd, _D, _D, _I, _I, _I, _I, _I, _I, _PD])
    _libs[key] = fns
    return fns


def _p(a):
    return a.ctypes.data_as(_PD)


class CpuDomain:
    """One single-grid domain in a CPU checker.

    Arrays are exchanged as numpy float64 arrays of shape (4, ny, nx) (planes STG, UH,
    VH, ELV; row j, column i), which is the byte layout of the reference's Fortran
    U(nx, ny, 4).
    """

    def __init__(self, kind, nx, ny, lw, ll, method="linear", spherical=False, coriolis=True,
                 truely_linear=True, msl_linear=0.0, linear_friction_coeff=0.0,
                 friction_type="manning", cfl=-1.0, theta=-1.0):
        self.kind = kind
        self.nx, self.ny = int(nx), int(ny)
        self.spherical = bool(spherical)
        self.f = _load(kind, self.spherical)
        lw_a = (_D * 2)(*lw)
        ll_a = (_D * 2)(*ll)
        ft = 1 if friction_type == "chezy" else 0
        if kind == "oracle":
            self.h = self.f["domain_create"](self.nx, self.ny, lw_a, ll_a, METHODS[method], int(spherical),
                                             int(coriolis), int(truely_linear), msl_linear,
                                             linear_friction_coeff, ft, cfl, theta)
        else:
            if spherical and not coriolis:
                raise ValueError("the compiled reference's spherical build always includes Coriolis")
            self.h = self.f["domain_create"](self.nx, self.ny, lw_a, ll_a, METHODS[method],
                                             int(truely_linear), msl_linear, linear_friction_coeff, ft,
                                             cfl, theta)
            if self.h < 0:
                raise RuntimeError("reference domain pool exhausted")

    def close(self):
        if self.h is not None:
            self.f["domain_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_U(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (4, self.ny, self.nx)
        self.f["domain_set_U"](self.h, _p(U))

    def get_U(self):
        U = np.empty((4, self.ny, self.nx))
        self.f["domain_get_U"](self.h, _p(U))
        return U

    def set_manning_squared(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        assert m.shape == (self.ny, self.nx)
        self.f["domain_set_manning_squared"](self.h, _p(m))

    def metrics(self):
        nx, ny = self.nx, self.ny
        out = dict(x=np.empty(nx), y=np.empty(ny), distance_bottom_edge=np.empty(ny + 1),
                   distance_left_edge=np.empty(nx + 1), area_cell_y=np.empty(ny), coriolis=np.empty(ny),
                   coriolis_bottom_edge=np.empty(ny + 1), tanlat_on_radius_earth=np.empty(ny))
        self.f["domain_get_metrics"](self.h, *[_p(out[k]) for k in
                                              ["x", "y", "distance_bottom_edge", "distance_left_edge",
                                               "area_cell_y", "coriolis", "coriolis_bottom_edge",
                                               "tanlat_on_radius_earth"]])
        return out

    def evolve(self, dt=None, nsteps=1):
        return self.f["domain_evolve"](self.h, -1.0 if dt is None else float(dt), int(nsteps))

    def scalars(self):
        o = np.empty(12)
        self.f["domain_get_scalars"](self.h, _p(o))
        keys = ["time", "max_dt", "evolve_step_dt", "boundary_flux_time_integral",
                "boundary_flux_evolve_integral", "bfs_N", "bfs_E", "bfs_S", "bfs_W",
                "negative_depth_fix_counter", "nsteps_advanced", "dt_last_update"]
        return dict(zip(keys, o.tolist()))

    def set_time(self, t):
        self.f["domain_set_time"](self.h, float(t))

    def volume_interior(self):
        return self.f["domain_volume_interior"](self.h)

    def stationary_timestep_max(self):
        return self.f["domain_stationary_timestep_max"](self.h)

    def max_U(self, name):
        o = np.empty((self.ny, self.nx))
        self.f["domain_get_max_U"](self.h, MAXQ.index(name) + 1, _p(o))
        return o

    def set_record_max_U(self, on):
        self.f["domain_set_record_max_U"](self.h, int(bool(on)))

    def set_stage_bc(self, amp, period, msl=0.0, sides=(0, 0, 0, 1)):
        s = (_I * 4)(*sides)
        self.f["domain_set_stage_bc"](self.h, float(amp), float(period), float(msl), s)

    BC_KINDS = {"transmissive": 3, "stage_radiation_momentum": 4, "stage_transmissive_momentum": 5}

    def set_bc(self, kind, sides=(1, 1, 1, 1), amp=None, period=1.0, msl=0.0):
        """kind: 'transmissive', 'stage_radiation_momentum', 'stage_transmissive_momentum'
        (boundary_mod.f90); amp None: no boundary_function."""
        s = (_I * 4)(*sides)
        self.f["domain_set_bc"](self.h, self.BC_KINDS[kind], int(amp is not None), float(amp or 0.0), float(period),
                                float(msl), s)

    def set_flather_bc(self, sides=(1, 1, 1, 1), amp=None, period=1.0, msl=0.0):
        """Flather boundary on sides [N, E, S, W]; amp None: outside sea at rest at
        msl_linear, else the sine stage msl + amp*sin(2*pi*t/period)."""
        s = (_I * 4)(*sides)
        use_fun = amp is not None
        self.f["domain_set_flather_bc"](self.h, int(use_fun), float(amp or 0.0), float(period), float(msl), s)

    def dispersion_setup(self, td1=0.0, td2=-1.0, niter=2):
        self.f["domain_dispersion_setup"](self.h, float(td1), float(td2), int(niter))

    def dispersion_store(self):
        self.f["domain_dispersion_store"](self.h)

    def dispersion_append(self):
        self.f["domain_dispersion_append"](self.h)

    def dispersion_solve(self, rhs_up, estimate, forward_time):
        self.f["domain_dispersion_solve"](self.h, int(bool(rhs_up)), int(bool(estimate)), float(forward_time))

    def dispersive_step(self, dt, outer_iterations=2):
        """One step as multidomain_mod.f90:1245-1410 runs a single dispersive domain:
        append the extrapolation level, store U, the shallow-water step, then the outer
        dispersive iterations (the first with a new right-hand side and the solution
        extrapolated forward in time)."""
        self.dispersion_append()
        self.dispersion_store()
        self.evolve(dt, 1)
        t = self.scalars()["time"]
        for it in range(outer_iterations):
            self.dispersion_solve(it > 0, it == 0, t)

    def set_forcing_patch(self, start_time, end_time, i0, i1, j0, j1, k0, k1, work):
        """forcing_patch_type applied by apply_forcing_patch (compiled reference only):
        work has shape (k1-k0+1, j1-j0+1, i1-i0+1) (1-based inclusive indices)."""
        w = np.ascontiguousarray(work, dtype=np.float64)
        assert w.shape == (k1 - k0 + 1, j1 - j0 + 1, i1 - i0 + 1)
        self.f["domain_set_forcing"](self.h, float(start_time), float(end_time), i0, i1, j0, j1, k0, k1, _p(w))

    STATS_KEYS = ("max_stage", "max_speed", "min_stage", "min_speed", "energy_potential_on_rho",
                  "energy_kinetic_on_rho", "energy_total_on_rho")

    def statistics(self):
        """compute_domain_statistics (compiled reference only)."""
        out = np.zeros(7)
        self.f["domain_statistics"](self.h, _p(out))
        return dict(zip(self.STATS_KEYS, out.tolist()))

    def sample(self, ij):
        """[stage, uh, vh] (3, n) at 1-based (i, j) cells (compiled reference only)."""
        ij = np.ascontiguousarray(np.asarray(ij, dtype=np.int32).reshape(-1, 2))
        out = np.zeros(3 * len(ij))
        self.f["domain_sample"](self.h, len(ij), ij.ctypes.data_as(ctypes.POINTER(_I)), _p(out))
        return out.reshape(3, len(ij))


def _load_md(spherical):
    """The multidomain entry points of the compiled reference (ref kind only)."""
    key = ("ref_md", spherical)
    if key in _libs:
        return _libs[key]
    path = lib_path("ref", spherical)
    if not os.path.exists(path):
        raise FileNotFoundError(path + " is not built (run `make -C oracle ref`)")
    lib = ctypes.CDLL(path)
    _PI = ctypes.POINTER(_I)
    fns = {}

    def f(name, res, args):
        fn = getattr(lib, "ref_md_" + name)
        fn.restype = res
        fn.argtypes = args
        fns[name] = fn

    f("begin", None, [_I])
    f("set_domain", None, [_I, _I, _I, _PD, _PD, _I, _I, _D])
    f("set_nested", None, [_I, _I, _PD, _PD, _D, _I, _I])
    f("set_params", None, [_I, _I, _D, _D])
    f("setup", None, [_I])
    f("sample", None, [_I, _I, _PI, _PD])
    f("get_max_stage", None, [_I, _PD])
    f("dims", None, [_I, _PI])
    f("geometry", None, [_I, _PD, _PD, _PD])
    f("metrics", None, [_I] + [_PD] * 6)
    f("set_U", None, [_I, _PD])
    f("get_U", None, [_I, _PD])
    f("set_manning_squared", None, [_I, _PD])
    f("make_consistent", None, [])
    f("evolve", None, [_D, _I])
    f("time", None, [_I, _PD])
    f("priority", None, [_I, _PI, _PI])
    f("all", None, [_PD, _PI])
    f("comm", None, [_I, _I, _I, _PI, _PD])
    f("comm_buffer", None, [_I, _I, _I, _PD])
    f("volume", None, [_PD])
    f("statistics", None, [_I, _PD])
    _libs[key] = fns
    return fns


def _pi(a):
    return a.ctypes.data_as(ctypes.POINTER(_I))


class RefMultidomain:
    """The reference's own multidomain_type (multidomain_mod.f90), driven through the
    bind(C) shim: the same setup (md%setup builds the nesting halos, priority
    domains and send/recv boxes), initial exchange and evolve_one_step.

    ``domains`` is a list of dicts: the first (and any other top-level domain) has
    ``lw, ll, nx, ny``; nested ones have ``parent`` (0-based), ``ll``, ``ur`` and
    ``dx_refinement_factor``. Every dict has ``method`` and
    ``timestepping_refinement_factor`` (default 1), and optionally
    ``linear_solver_is_truely_linear``, ``msl_linear``, ``linear_friction_coeff``.

    The reference writes its log and output folders under the working directory,
    so ``setup`` runs inside ``workdir`` (a scratch folder)."""

    def __init__(self, domains, spherical=True, workdir=None, max_stage=False):
        import tempfile
        self.spherical = spherical
        self.f = _load_md(spherical)
        self.nd = len(domains)
        self.f["begin"](self.nd)
        for k, d in enumerate(domains):
            m = METHODS[d["method"]]
            trf = int(d.get("timestepping_refinement_factor", 1))
            if "parent" in d:
                ll = np.asarray(d["ll"], dtype=np.float64)
                ur = np.asarray(d["ur"], dtype=np.float64)
                self.f["set_nested"](k + 1, d["parent"] + 1, _p(ll), _p(ur), float(d["dx_refinement_factor"]), trf, m)
            else:
                lw = np.asarray(d["lw"], dtype=np.float64)
                ll = np.asarray(d["ll"], dtype=np.float64)
                self.f["set_domain"](k + 1, int(d["nx"]), int(d["ny"]), _p(lw), _p(ll), m, trf, 1.0)
            self.f["set_params"](k + 1, int(d.get("linear_solver_is_truely_linear", 1)), float(d.get("msl_linear", 0.0)),
                                 float(d.get("linear_friction_coeff", 0.0)))
        self.workdir = workdir or tempfile.mkdtemp(prefix="swals_md_")
        # the first domain's setup copies '*.f* make*' from the working directory
        # (domain_mod.f90:2061) and stops if nothing matches
        for name in ("make_placeholder", "placeholder.f90"):
            open(os.path.join(self.workdir, name), "a").close()
        cwd = os.getcwd()
        os.chdir(self.workdir)
        try:
            self.f["setup"](int(bool(max_stage)))
        finally:
            os.chdir(cwd)
        self.methods = [d["method"] for d in domains]

    def sample(self, k, ij):
        """[stage, uh, vh] (3, n) of domain k at 1-based (i, j) cells."""
        ij = np.ascontiguousarray(np.asarray(ij, dtype=np.int32).reshape(-1, 2))
        out = np.zeros(3 * len(ij))
        self.f["sample"](k + 1, len(ij), _pi(ij), _p(out))
        return out.reshape(3, len(ij))

    def max_stage(self, k):
        d = self.dims(k)
        out = np.zeros((d["ny"], d["nx"]))
        self.f["get_max_stage"](k + 1, _p(out))
        return out

    def dims(self, k):
        out = np.zeros(8, dtype=np.int32)
        self.f["dims"](k + 1, _pi(out))
        return dict(nx=int(out[0]), ny=int(out[1]), n_send=int(out[2]), n_recv=int(out[3]), staggered=int(out[4]),
                    method_index=int(out[5]), timestepping_refinement_factor=int(out[6]),
                    exterior_cells_width=int(out[7]))

    def geometry(self, k):
        lw, ll, dx = np.zeros(2), np.zeros(2), np.zeros(2)
        self.f["geometry"](k + 1, _p(lw), _p(ll), _p(dx))
        return lw, ll, dx

    def metrics(self, k):
        d = self.dims(k)
        nx, ny = d["nx"], d["ny"]
        dbe, dle, area = np.zeros(ny + 1), np.zeros(nx + 1), np.zeros(ny)
        cor, corbe, tl = np.zeros(ny), np.zeros(ny + 1), np.zeros(ny)
        self.f["metrics"](k + 1, _p(dbe), _p(dle), _p(area), _p(cor), _p(corbe), _p(tl))
        return dict(distance_bottom_edge=dbe, distance_left_edge=dle, area_cell_y=area, coriolis=cor,
                    coriolis_bottom_edge=corbe, tanlat_on_radius_earth=tl)

    def set_U(self, k, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        self.f["set_U"](k + 1, _p(U))

    def get_U(self, k):
        d = self.dims(k)
        U = np.zeros((4, d["ny"], d["nx"]))
        self.f["get_U"](k + 1, _p(U))
        return U

    def set_manning_squared(self, k, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        self.f["set_manning_squared"](k + 1, _p(m))

    def make_initial_conditions_consistent(self):
        self.f["make_consistent"]()

    def evolve(self, dt, nsteps=1):
        self.f["evolve"](float(dt), int(nsteps))

    def time(self, k):
        t = np.zeros(1)
        self.f["time"](k + 1, _p(t))
        return float(t[0])

    def priority(self, k):
        d = self.dims(k)
        pdi = np.zeros((d["ny"], d["nx"]), dtype=np.int32)
        pnp = np.zeros((d["ny"], d["nx"]), dtype=np.int32)
        self.f["priority"](k + 1, _pi(pdi), _pi(pnp))
        return pdi, pnp

    def statistics(self, k):
        """Domain k's compute_domain_statistics (priority cells): max/min stage, max/min
        speed, potential, kinetic, total energy / rho."""
        out = np.zeros(7)
        self.f["statistics"](k + 1, _p(out))
        return out

    def all_dx_and_methods(self):
        dx = np.zeros((self.nd, 2))
        m = np.zeros(self.nd, dtype=np.int32)
        self.f["all"](_p(dx), _pi(m))
        return dx, m

    def comms(self, k, recv):
        """The send (recv=False) or receive (recv=True) boxes of domain k, as dicts."""
        d = self.dims(k)
        n = d["n_recv"] if recv else d["n_send"]
        out = []
        for c in range(n):
            io = np.zeros(32, dtype=np.int32)
            do = np.zeros(4)
            self.f["comm"](k + 1, int(recv), c + 1, _pi(io), _p(do))
            out.append(dict(inds=io[0:6].reshape(3, 2).copy(), my_index=int(io[6]), neighbour_index=int(io[7]),
                            my_comms_index=int(io[8]), neighbour_comms_index=int(io[9]),
                            my_domain_is_finer=int(io[10]), equal_cell_ratios=int(io[11]),
                            my_staggered=int(io[12]), neighbour_staggered=int(io[13]), wetdry_limiting=int(io[14]),
                            buffer_size=int(io[15]), flux_start=int(io[16]), time_index=int(io[17]),
                            flux_shape=io[18:26].reshape(4, 2).copy(), active=int(io[26]),
                            cell_ratios=do[0:2].copy()))
        return out

    def comm_buffer(self, k, recv, c):
        info = self.comms(k, recv)[c]
        b = np.zeros(info["buffer_size"])
        self.f["comm_buffer"](k + 1, int(recv), c + 1, _p(b))
        return b

    def volumes(self):
        v = np.zeros(self.nd)
        self.f["volume"](_p(v))
        return v
