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
still stores the state every ``checkpoint_every`` steps (as the solver would need for
a state-dependent adjoint); ``recompute_from_checkpoints`` rebuilds any sampled
gauge value from them, which the tests use to check the checkpoints.

``GaugeMapFunction`` is the thin torch.autograd.Function over the C ABI; torch only
carries tensors and gradients, all numerics run in libswals_hip.so.
"""
import ctypes

import numpy as np

from . import backend as B
from .domain import Domain, _p


class GaugeMap:
    def __init__(self, nx, ny, lw, ll, elev, dt, nsteps, every, gauges_xy, precision="fp64", spherical=True,
                 linear_friction_coeff=0.0, checkpoint_every=256):
        self.dom = Domain(nx, ny, lw, ll, "linear", precision=precision, spherical=spherical, max_U_variables=(),
                          linear_friction_coeff=linear_friction_coeff)
        self.nx, self.ny = nx, ny
        self.elev = np.asarray(elev, dtype=np.float64)
        self.dt, self.nsteps, self.every = float(dt), int(nsteps), int(every)
        self.nsamples = self.nsteps // self.every + 1
        self.gauge_ij = self.dom.setup_point_gauges(gauges_xy, max_samples=self.nsamples)
        self.ngauges = len(self.gauge_ij)
        self.checkpoint_every = int(checkpoint_every)
        self.checkpoints = {}

    def _reset(self, U):
        # a fresh forward run: new state, time zero, empty gauge buffer
        self.dom.U = U
        self.dom.set_time(0.0)
        self.gauge_ij = self.dom.setup_point_gauges(self._gxy(), max_samples=self.nsamples)

    def _gxy(self):
        ll, dx = self.dom.lower_left, self.dom.dx
        return np.stack([ll[0] + (self.gauge_ij[:, 0] - 0.5) * dx[0], ll[1] + (self.gauge_ij[:, 1] - 0.5) * dx[1]], 1)

    def forward(self, eta0, uh0=None, vh0=None):
        """Gauge stage series (nsamples, ngauges) from the initial state."""
        U = np.zeros((4, self.ny, self.nx))
        U[0] = eta0
        if uh0 is not None:
            U[1] = uh0
        if vh0 is not None:
            U[2] = vh0
        U[3] = self.elev
        self._reset(U)
        self.checkpoints = {0: U[:3].copy()}
        lib = self.dom._lib
        n = 0
        B.check(lib.swh_gauges_sample(self.dom._h), "swh_gauges_sample")
        while n < self.nsteps:
            # advance to the next gauge sample or checkpoint, whichever comes first
            # (runs of steps go to the device as one call, two steps per launch)
            k = min(self.every - (n % self.every), self.checkpoint_every - (n % self.checkpoint_every),
                    self.nsteps - n)
            self.dom.evolve_steps(self.dt, k)
            n += k
            if n % self.every == 0:
                B.check(lib.swh_gauges_sample(self.dom._h), "swh_gauges_sample")
            if n % self.checkpoint_every == 0 and n < self.nsteps:
                self.checkpoints[n] = self.dom.U[:3].copy()
        _, vals = self.dom.gauge_series()
        return vals[:, 0, :].copy()

    def recompute_from_checkpoints(self, sample):
        """Stage at the gauges for one sample, recomputed from the nearest earlier
        checkpoint (forward recompute within a checkpoint segment)."""
        n = sample * self.every
        c = max(k for k in self.checkpoints if k <= n)
        U = np.zeros((4, self.ny, self.nx))
        U[:3] = self.checkpoints[c]
        U[3] = self.elev
        self.dom.U = U
        self.dom.evolve_steps(self.dt, n - c)
        S = self.dom.U[0]
        return np.array([S[j - 1, i - 1] for (i, j) in self.gauge_ij])

    def adjoint(self, grad_g):
        """d loss / d (eta0, uh0, vh0) given d loss / d g, shape (nsamples, ngauges)."""
        g = np.ascontiguousarray(grad_g, dtype=np.float64)
        assert g.shape == (self.nsamples, self.ngauges)
        lib = self.dom._lib
        B.check(lib.swh_adjoint_reset(self.dom._h), "swh_adjoint_reset")
        B.check(lib.swh_adjoint_run(self.dom._h, self.dt, self.nsteps, self.every, _p(g)), "swh_adjoint_run")
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
