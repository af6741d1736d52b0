"""Source inversion with the adjoint of the 'linear' leapfrog solver.

The forward map takes an initial stage field eta0 to the stage recorded at a set of
gauges every ``every`` steps (the PTHA unit-source gauge series, point_gauge_mod.f90).
It is linear (truely-linear solver, walls), so with unit sources phi_k and weights w
the gauges are g = G w and, for loss = sum (g - obs)^2, d loss / d w = 2 G^T r. The
backward pass runs the exact transpose of each leapfrog step on the GPU
(swh_adjoint_run), injecting d loss / d g at the sampling steps, which yields
d loss / d eta0 in one sweep instead of one forward run per unit source.

Checkpoint/recompute: the adjoint of the truely-linear step does not depend on the
forward state, so the backward sweep needs no forward states. ``GaugeMap.forward``
still keeps the state every ``checkpoint_every`` steps in device memory (as a
state-dependent adjoint would need; a device copy, no host transfer);
``recompute_from_checkpoints`` rebuilds any sampled gauge value from them, which the
tests use to check the checkpoints.

``GaugeMapFunction`` is the thin torch.autograd.Function over the C ABI; torch only
carries tensors and gradients, all numerics run in libswals_hip.so.
"""
import ctypes

import numpy as np

from . import backend as B
from .domain import Domain, _p


class GaugeMap:
    """The forward map eta0 -> stage at the gauges, and its adjoint, on one grid.

    Gauges are sampled at ``sample_steps`` (or every ``every`` steps from step 0),
    recorded on the device in one call per checkpoint segment; the state is kept every
    ``checkpoint_every`` steps in device memory (swh_checkpoint_save), so a forward run
    never waits on the host. The gauge schedule of the reference's unit-source driver
    (the first step on or after each write time) comes from
    ``ptha_amd.batch.reference_write_schedule``."""

    def __init__(self, nx, ny, lw, ll, elev, dt, nsteps, every, gauges_xy, precision="fp64", spherical=True,
                 linear_friction_coeff=0.0, checkpoint_every=256, sample_steps=None):
        self.dom = Domain(nx, ny, lw, ll, "linear", precision=precision, spherical=spherical, max_U_variables=(),
                          linear_friction_coeff=linear_friction_coeff)
        self.nx, self.ny = nx, ny
        self.elev = np.asarray(elev, dtype=np.float64)
        self.dt, self.nsteps = float(dt), int(nsteps)
        if sample_steps is None:
            sample_steps = np.arange(0, self.nsteps + 1, int(every))
        self.sample_steps = np.ascontiguousarray(sample_steps, dtype=np.int64)
        assert self.sample_steps[-1] <= self.nsteps
        self.nsamples = len(self.sample_steps)
        self.gauges_xy = np.asarray(gauges_xy, dtype=np.float64)
        self.gauge_ij = self.dom.setup_point_gauges(self.gauges_xy, max_samples=self.nsamples)
        self.ngauges = len(self.gauge_ij)
        self.checkpoint_every = int(checkpoint_every)
        self.checkpoint_steps = []
        self._elev_set = False

    def _reset(self, eta0, uh0=None, vh0=None):
        # a fresh forward run: new state (the elevation stays), time zero, empty gauge buffer
        if not self._elev_set:
            U = np.zeros((4, self.ny, self.nx))
            U[3] = self.elev
            self.dom.U = U
            self._elev_set = True
        self.dom.set_state(eta0, uh0, vh0)
        self.dom.set_time(0.0)
        self.gauge_ij = self.dom.setup_point_gauges(self.gauges_xy, max_samples=self.nsamples)

    def _record(self, n0, n1, lib):
        """Steps n0 -> n1, sampling the gauges at the scheduled steps in (n0, n1] (and n0
        when n0 == 0)."""
        lo = 0 if n0 == 0 else n0 + 1
        sel = self.sample_steps[(self.sample_steps >= lo) & (self.sample_steps <= n1)]
        if len(sel):
            rel = np.ascontiguousarray(sel - n0, dtype=np.int64)
            B.check(lib.swh_evolve_record_gauges_schedule(self.dom._h, self.dt, len(rel),
                                                          rel.ctypes.data_as(B._PI64)),
                    "swh_evolve_record_gauges_schedule")
            n = int(sel[-1])
        else:
            n = n0
        if n1 > n:
            self.dom.evolve_steps(self.dt, n1 - n)

    def forward(self, eta0, uh0=None, vh0=None):
        """Gauge stage series (nsamples, ngauges) from the initial state."""
        self._reset(eta0, uh0, vh0)
        lib = self.dom._lib
        B.check(lib.swh_checkpoint_save(self.dom._h, 0), "swh_checkpoint_save")
        self.checkpoint_steps = [0]
        n = 0
        while n < self.nsteps:
            n1 = min(n + self.checkpoint_every, self.nsteps)
            self._record(n, n1, lib)
            n = n1
            if n < self.nsteps:
                B.check(lib.swh_checkpoint_save(self.dom._h, len(self.checkpoint_steps)), "swh_checkpoint_save")
                self.checkpoint_steps.append(n)
        if self.nsteps == 0:
            self._record(0, 0, lib)
        _, vals = self.dom.gauge_series()
        return vals[:, 0, :].copy()

    def recompute_from_checkpoints(self, sample):
        """Stage at the gauges for one sample, recomputed on the device from the nearest
        earlier checkpoint (forward recompute within a checkpoint segment)."""
        n = int(self.sample_steps[sample])
        k = max(i for i, c in enumerate(self.checkpoint_steps) if c <= n)
        B.check(self.dom._lib.swh_checkpoint_restore(self.dom._h, k), "swh_checkpoint_restore")
        self.dom.evolve_steps(self.dt, n - self.checkpoint_steps[k])
        S = self.dom.U[0]
        return np.array([S[j - 1, i - 1] for (i, j) in self.gauge_ij])

    def adjoint(self, grad_g):
        """d loss / d (eta0, uh0, vh0) given d loss / d g, shape (nsamples, ngauges)."""
        g = np.ascontiguousarray(grad_g, dtype=np.float64)
        assert g.shape == (self.nsamples, self.ngauges)
        lib = self.dom._lib
        B.check(lib.swh_adjoint_reset(self.dom._h), "swh_adjoint_reset")
        B.check(lib.swh_adjoint_run_schedule(self.dom._h, self.dt, self.nsteps, self.nsamples,
                                             self.sample_steps.ctypes.data_as(B._PI64), _p(g)),
                "swh_adjoint_run_schedule")
        lam = np.empty((3, self.ny, self.nx))
        B.check(lib.swh_adjoint_download(self.dom._h, _p(lam)), "swh_adjoint_download")
        return lam

    def transpose_steps(self, lam3, nsteps):
        """Applies (A^T)^nsteps to an adjoint state (3, ny, nx)."""
        lib = self.dom._lib
        a = np.ascontiguousarray(lam3, dtype=np.float64)
        B.check(lib.swh_adjoint_reset(self.dom._h), "swh_adjoint_reset")
        B.check(lib.swh_adjoint_upload(self.dom._h, _p(a)), "swh_adjoint_upload")
        B.check(lib.swh_adjoint_step(self.dom._h, self.dt, int(nsteps)), "swh_adjoint_step")
        out = np.empty((3, self.ny, self.nx))
        B.check(lib.swh_adjoint_download(self.dom._h, _p(out)), "swh_adjoint_download")
        return out

    def forward_steps(self, U3, nsteps):
        """Applies A^nsteps to a state (3, ny, nx) (elevation fixed)."""
        U = np.zeros((4, self.ny, self.nx))
        U[:3] = U3
        U[3] = self.elev
        self.dom.U = U
        self.dom.evolve_steps(self.dt, int(nsteps))
        return self.dom.U[:3]

    def close(self):
        self.dom.close()


def unit_source_basis(x, y, centres_lon, centres_lat, sigma_m, radius=6371000.0):
    """Separable Gaussian unit sources on a lon/lat lattice: phi_ab(x, y) =
    gx_ab(x) * gy_b(y), with the east-west distance scaled by cos(lat_b).
    Returns (GX[b, a, nx], GY[b, ny]) as numpy arrays."""
    x = np.asarray(x)
    y = np.asarray(y)
    cl = np.cos(np.radians(centres_lat))
    GX = np.exp(-((np.radians(x[None, None, :] - centres_lon[None, :, None]) * radius * cl[:, None, None]) ** 2)
                / (2 * sigma_m ** 2))
    GY = np.exp(-((np.radians(y[None, :] - centres_lat[:, None]) * radius) ** 2) / (2 * sigma_m ** 2))
    return GX, GY


def stage_from_weights(w, GX, GY, elev):
    """eta0 = sum_ab w[b, a] gx_ab(x) gy_b(y) on the wet cells (elevation < 0), zero on
    land (numpy; the same sum as sources_to_stage)."""
    w = np.asarray(w, dtype=np.float64)
    eta = GY.T @ np.stack([w[b] @ GX[b] for b in range(w.shape[0])])
    return np.where(np.asarray(elev) < 0, eta, 0.0)


def _torch_function():
    import torch

    class GaugeMapFunction(torch.autograd.Function):
        """eta0 (ny, nx) -> gauge stage series (nsamples, ngauges), with the
        backward pass on the device adjoint."""

        @staticmethod
        def forward(ctx, eta0, gmap):
            ctx.gmap = gmap
            g = gmap.forward(eta0.detach().cpu().double().numpy())
            return torch.from_numpy(g).to(eta0.dtype)

        @staticmethod
        def backward(ctx, grad_g):
            lam = ctx.gmap.adjoint(grad_g.detach().cpu().double().numpy())
            return torch.from_numpy(lam[0]).to(grad_g.dtype), None

    return GaugeMapFunction


def sources_to_stage(w, GX, GY):
    """eta0 = sum_ab w[b, a] gx_ab(x) gy_b(y) as torch ops (differentiable in w)."""
    import torch
    GXt = torch.as_tensor(GX, dtype=w.dtype)
    GYt = torch.as_tensor(GY, dtype=w.dtype)
    X = torch.bmm(w.unsqueeze(1), GXt).squeeze(1)  # (nb, nx)
    return GYt.t().contiguous() @ X                  # (ny, nx)


def gauge_map_apply(eta0, gmap):
    return _torch_function().apply(eta0, gmap)
