"""Multidomain (nested grids) on one GPU over the C ABI (swh_md_*).

The nesting geometry is what the reference's ``md%setup`` decides
(multidomain_mod.f90:2663-2779): for every domain its send and receive boxes
(two_way_nesting_comms_type, nested_grid_comms_mod.f90:69-219), its
priority_domain_index / is_priority_domain_not_periodic arrays, and the
multidomain's all_dx_md / all_timestepping_methods_md. ``MultiDomain`` hands that
to the device once; each ``evolve_one_step(dt)`` then runs the whole multidomain
step of multidomain_mod.f90:1093-1212 on the GPU: every domain's
timestepping_refinement_factor sub-steps with flux-integral tracking, the halo
exchange (pack straight into the receiving buffer, unpack, staggered wet/dry
fix-ups) and flux correction. Domains step concurrently on their own HIP streams.
"""
import ctypes

import numpy as np

from . import backend as B
from .domain import _p

_I32 = ctypes.c_int32


def _pi(a):
    return a.ctypes.data_as(ctypes.POINTER(_I32))


def nest_box(info):
    """swh_nest_box from a dict with the fields of two_way_nesting_comms_type:
    inds (3, 2) [[iL, iU], [jL, jU], [kL, kU]], neighbour_index, neighbour_comms_index,
    my_domain_is_finer, equal_cell_ratios, my_staggered, neighbour_staggered,
    wetdry_limiting, buffer_size, flux_start, time_index, flux_shape (4, 2),
    cell_ratios (2,)."""
    b = B.NestBox()
    inds = np.asarray(info["inds"]).ravel()
    for q in range(6):
        b.inds[q] = int(inds[q])
    b.neighbour_domain_index = int(info["neighbour_index"])
    b.neighbour_comms_index = int(info["neighbour_comms_index"])
    b.my_domain_is_finer = int(info["my_domain_is_finer"])
    b.equal_cell_ratios = int(info["equal_cell_ratios"])
    b.my_domain_staggered_grid = int(info["my_staggered"])
    b.neighbour_domain_staggered_grid = int(info["neighbour_staggered"])
    b.use_wetdry_limiting = int(info["wetdry_limiting"])
    b.buffer_size = int(info["buffer_size"])
    b.flux_start = int(info["flux_start"])
    b.time_index = int(info["time_index"])
    fs = np.asarray(info["flux_shape"])
    for q in range(4):
        b.flux_rows[q] = int(fs[q, 0])
    b.flux_nvar = int(fs[0, 1])
    b.cell_ratios[0] = float(info["cell_ratios"][0])
    b.cell_ratios[1] = float(info["cell_ratios"][1])
    return b


class MultiDomain:
    """Nested domains advanced together on one GPU.

    domains      list of ``ptha_amd.domain.Domain`` (all one precision), in md%domains order
    all_dx       (n, 2) cell sizes (md%all_dx_md(:, :, 1))
    methods      (n,) timestepping method indices (md%all_timestepping_methods_md(:, 1))
    trf          (n,) timestepping_refinement_factor of each domain
    send, recv   per domain, the list of box dicts (see ``nest_box``) in send_comms /
                 recv_comms order
    priority     per domain, (priority_domain_index, is_priority_domain_not_periodic),
                 int arrays of shape (ny, nx)
    """

    def __init__(self, domains, all_dx, methods, trf, send, recv, priority, flux_correct_dry_cells=None):
        lib = B.load()
        self._lib = lib
        self.domains = list(domains)
        n = len(self.domains)
        hs = (ctypes.c_void_p * n)(*[d._h.value for d in self.domains])
        dx = np.ascontiguousarray(np.asarray(all_dx, dtype=np.float64).reshape(n, 2))
        m = np.ascontiguousarray(np.asarray(methods, dtype=np.int32))
        t = np.ascontiguousarray(np.asarray(trf, dtype=np.int32))
        h = ctypes.c_void_p()
        B.check(lib.swh_md_create(n, hs, _p(dx), _pi(m), _pi(t), ctypes.byref(h)), "swh_md_create")
        self._h = h
        try:
            for k in range(n):
                pdi, pnp = priority[k]
                pdi = np.ascontiguousarray(pdi, dtype=np.int32)
                pnp = np.ascontiguousarray(pnp, dtype=np.int32)
                B.check(lib.swh_md_set_priority(h, k + 1, _pi(pdi), _pi(pnp)), "swh_md_set_priority")
                if flux_correct_dry_cells is not None:
                    B.check(lib.swh_md_set_flux_correct_dry_cells(h, k + 1, int(flux_correct_dry_cells[k])),
                            "swh_md_set_flux_correct_dry_cells")
                for info in send[k]:
                    box = nest_box(info)
                    B.check(lib.swh_md_add_box(h, k + 1, 0, ctypes.byref(box)), "swh_md_add_box")
                for info in recv[k]:
                    box = nest_box(info)
                    B.check(lib.swh_md_add_box(h, k + 1, 1, ctypes.byref(box)), "swh_md_add_box")
            B.check(lib.swh_md_finalize(h), "swh_md_finalize")
        except Exception:
            lib.swh_md_destroy(h)
            self._h = None
            raise
        self._recv_sizes = [[int(info["buffer_size"]) for info in recv[k]] for k in range(n)]

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.swh_md_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def make_initial_conditions_consistent(self):
        B.check(self._lib.swh_md_make_initial_conditions_consistent(self._h),
                "swh_md_make_initial_conditions_consistent")

    def evolve_one_step(self, dt):
        B.check(self._lib.swh_md_evolve_one_step(self._h, float(dt)), "swh_md_evolve_one_step")

    def evolve_steps(self, dt, nsteps):
        B.check(self._lib.swh_md_evolve_steps(self._h, float(dt), int(nsteps)), "swh_md_evolve_steps")

    def graph_capture(self, dt):
        """Record one global step of dt as a HIP graph (not taken); see graph_launch."""
        B.check(self._lib.swh_md_graph_capture(self._h, float(dt)), "swh_md_graph_capture")

    def graph_launch(self, nsteps):
        """Replay the captured global step nsteps times (each equal to evolve_one_step)."""
        B.check(self._lib.swh_md_graph_launch(self._h, int(nsteps)), "swh_md_graph_launch")

    def graph_destroy(self):
        B.check(self._lib.swh_md_graph_destroy(self._h), "swh_md_graph_destroy")

    def synchronize(self):
        B.check(self._lib.swh_md_synchronize(self._h), "swh_md_synchronize")

    def volumes(self):
        """Water volume in each domain's priority cells (volume_in_priority_domain)."""
        v = np.zeros(len(self.domains))
        B.check(self._lib.swh_md_volumes(self._h, _p(v)), "swh_md_volumes")
        return v

    def statistics(self):
        """print_multidomain's statistics (multidomain_mod.f90:2301-2447): (per-domain list
        of dicts over each domain's priority cells, global dict, energy_is_finite)."""
        nd = len(self.domains)
        per = (B.DomainStats * nd)()
        glob = B.DomainStats()
        fin = ctypes.c_int32()
        B.check(self._lib.swh_md_statistics(self._h, per, ctypes.byref(glob), ctypes.byref(fin)), "swh_md_statistics")
        keys = [f[0] for f in B.DomainStats._fields_]
        return ([{k: getattr(per[i], k) for k in keys} for i in range(nd)], {k: getattr(glob, k) for k in keys},
                bool(fin.value))

    def check_stability(self):
        """Number of interior values over all domains that are NaN or beyond +-HUGE(1.0)/2
        (check_multidomain_stability, multidomain_mod.f90:348-455)."""
        n = ctypes.c_int64()
        B.check(self._lib.swh_md_check_stability(self._h, ctypes.byref(n)), "swh_md_check_stability")
        return n.value

    def recv_buffer(self, k, c):
        """Receive buffer of domain k's receive box c (0-based), from the last exchange."""
        b = np.zeros(self._recv_sizes[k][c])
        B.check(self._lib.swh_md_download_recv_buffer(self._h, k + 1, c + 1, _p(b)), "swh_md_download_recv_buffer")
        return b
