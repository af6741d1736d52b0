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

METHODS = {"rk2": 1, "euler": 4, "linear": 5, "leapfrog_linear_plus_nonlinear_friction": 6}
MAXQ = ["max_stage", "max_speed", "max_flux", "arrival_time", "time_of_max_stage", "min_stage"]

_D = ctypes.c_double
_I = ctypes.c_int
_PD = ctypes.POINTER(_D)
_libs = {}


def lib_path(kind, spherical=False):
    if kind == "oracle":
        return os.path.join(REF_DIR, "liboracle.so")
    return os.path.join(REF_DIR, "libswals_ref_%s.so" % ("sph" if spherical else "cart"))


def available(kind, spherical=False):
    return os.path.exists(lib_path(kind, spherical))


def _load(kind, spherical):
    key = (kind, spherical if kind == "ref" else False)
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
    f("domain_set_record_max_U", None, [hd, _I])
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
