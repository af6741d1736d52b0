"""GPU: the adjoint of the 'linear' leapfrog step.

* dot-product identity <l, A^n x> = <(A^T)^n l, x> for random states (FP64, Coriolis,
  linear drag, wet/dry masks from real bathymetry);
* d loss / d w from the adjoint sweep against central finite differences (the loss is
  quadratic in w, so central differences carry no truncation error);
* against the CPU superposition gradient 2 G^T r, with G built column by column from
  FP64 oracle runs of each unit source;
* the FP32 adjoint against the FP64 one; the torch.autograd path; checkpoints.
"""
import numpy as np
import pytest

import pyoracle
from ptha_amd import adjoint as A
from ptha_amd import configs as C
from ptha_amd.domain import build_metrics

pytestmark = pytest.mark.gpu


def _problem(nx=96, ny=80, nsx=4, nsy=2):
    cfg = C.adjoint_inversion(nx=nx, ny=ny, arcmin=10.0, n_src_x=nsx, n_src_y=nsy, spacing_km=60.0,
                              sigma_km=40.0, n_gauges=8)
    m = build_metrics(nx, ny, cfg["lw"], cfg["ll"], True)
    dbe, dle = m["distance_bottom_edge"], m["distance_left_edge"]
    dist = np.minimum((dbe[:-1] + dbe[1:])[:, None], (dle[:-1] + dle[1:])[None, :])
    dt = float((0.5 * dist / np.sqrt(9.8 * np.maximum(-cfg["elev"], 1e-5))).min() * 0.7)
    GX, GY = A.unit_source_basis(cfg["x"], cfg["y"], cfg["src_lon"], cfg["src_lat"], cfg["sigma_km"] * 1e3)
    return cfg, dt, GX, GY


def _eta(w, GX, GY, elev):
    eta = GY.T @ np.stack([w[b] @ GX[b] for b in range(w.shape[0])])
    return np.where(elev < 0, eta, 0.0)


def test_adjoint_dot_product():
    print("CASE adjoint dot", flush=True)
    cfg, dt, _, _ = _problem()
    gm = A.GaugeMap(cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], cfg["elev"], dt, 30, 10, cfg["gauges_xy"],
                    linear_friction_coeff=2e-4)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((3, cfg["ny"], cfg["nx"]))
    lam = rng.standard_normal((3, cfg["ny"], cfg["nx"]))
    Ax = gm.forward_steps(x, 30)
    Atl = gm.transpose_steps(lam, 30)
    lhs, rhs = float((lam * Ax).sum()), float((Atl * x).sum())
    print("dot", lhs, rhs)
    assert lhs == pytest.approx(rhs, rel=1e-12)
    gm.close()


def test_adjoint_gradient_fd_and_cpu_GTr(oracle_built):
    print("CASE adjoint fd + GTr", flush=True)
    cfg, dt, GX, GY = _problem()
    nsteps, every = 60, 6
    gm = A.GaugeMap(cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], cfg["elev"], dt, nsteps, every, cfg["gauges_xy"])
    elev = cfg["elev"]
    w_true = cfg["w_true"]
    obs = gm.forward(_eta(w_true, GX, GY, elev))
    obs = obs * (1 + 0.01 * np.random.default_rng(9).standard_normal(obs.shape))
    w = np.zeros_like(w_true)

    def loss(wv):
        g = gm.forward(_eta(wv, GX, GY, elev))
        return float(((g - obs) ** 2).sum()), g

    L0, g0 = loss(w)
    lam = gm.adjoint(2 * (g0 - obs))
    lam_s = np.where(elev < 0, lam[0], 0.0)
    grad = np.array([[float((lam_s * np.outer(GY[b], GX[b, a])).sum()) for a in range(w.shape[1])]
                     for b in range(w.shape[0])])
    # central finite differences (exact for a quadratic loss up to rounding)
    for (b, a) in [(0, 0), (1, 2), (0, 3)]:
        e = np.zeros_like(w)
        e[b, a] = 1e-2
        fd = (loss(w + e)[0] - loss(w - e)[0]) / 2e-2
        assert grad[b, a] == pytest.approx(fd, rel=1e-7), (b, a)
    # CPU superposition: G column k from an FP64 oracle run of unit source k
    nx, ny = cfg["nx"], cfg["ny"]
    ij = gm.gauge_ij
    cols = []
    for b in range(w.shape[0]):
        for a in range(w.shape[1]):
            o = pyoracle.CpuDomain("oracle", nx, ny, cfg["lw"], cfg["ll"], method="linear", spherical=True)
            o.set_record_max_U(False)
            U = np.zeros((4, ny, nx))
            U[3] = elev
            U[0] = np.where(elev < 0, np.outer(GY[b], GX[b, a]), 0.0)
            o.set_U(U)
            series = []
            for n in range(nsteps + 1):
                if n % every == 0:
                    S = o.get_U()[0]
                    series.append([S[j - 1, i - 1] for (i, j) in ij])
                if n < nsteps:
                    o.evolve(dt, 1)
            cols.append(np.array(series).ravel())
    G = np.stack(cols, axis=1)
    r = (G @ w.ravel() - obs.ravel())
    gtr = 2 * G.T @ r
    assert np.allclose(grad.ravel(), gtr, rtol=1e-9, atol=1e-9 * np.abs(gtr).max())
    # checkpoints rebuild a recorded sample exactly
    gm.checkpoint_every = 16
    g1 = gm.forward(_eta(w_true, GX, GY, elev))
    assert np.array_equal(gm.recompute_from_checkpoints(7), g1[7])
    gm.close()


def test_adjoint_fp32_and_torch():
    print("CASE adjoint fp32 + torch", flush=True)
    import torch
    cfg, dt, GX, GY = _problem()
    nsteps, every = 60, 6
    res = {}
    for prec in ["fp64", "fp32k"]:
        gm = A.GaugeMap(cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], cfg["elev"], dt, nsteps, every, cfg["gauges_xy"],
                        precision=prec)
        obs = gm.forward(_eta(cfg["w_true"], GX, GY, cfg["elev"]))
        w = torch.zeros(cfg["w_true"].shape, dtype=torch.float64, requires_grad=True)
        mask = torch.as_tensor((cfg["elev"] < 0).astype(np.float64))
        eta = A.sources_to_stage(w, GX, GY) * mask
        g = A.gauge_map_apply(eta, gm)
        loss = ((g - torch.as_tensor(obs)) ** 2).sum()
        loss.backward()
        res[prec] = w.grad.numpy().copy()
        gm.close()
    rel = np.abs(res["fp32k"] - res["fp64"]).max() / np.abs(res["fp64"]).max()
    print("fp32 adjoint relative difference", rel)
    assert rel < 1e-3
