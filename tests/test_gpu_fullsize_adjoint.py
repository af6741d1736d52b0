"""GPU: the source-inversion workload at full size (2048 x 2048 spherical 'linear' grid
at 1 arcmin, 4 h, 32 gauges written at 60 s by the reference driver's time rule,
initial stage the weighted sum of 40 x 20 Gaussian unit sources with the hidden-truth
weights) against the FP64 run of the compiled reference (fixture written by
tests/golden/make_fullsize_config5.py).

* FP64 forward on the device: bit-identical to the reference (gauge stage at every
  write step, final stage on every 16th cell).
* FP32 with a Kahan-compensated stage: stage error against the reference FP64 run
  within 1e-5 m at every gauge sample and on the final wet stage (printed).
* The gradient of loss = sum (g(w) - obs)^2 with respect to the 800 unit-source
  weights from one adjoint sweep, against central finite differences of full-size
  forward runs on a few weights (the loss is quadratic in w, so central differences
  carry no truncation error), in FP64; and the FP32 adjoint gradient against the FP64
  one. The forward with device checkpoints and the adjoint sweep are timed.
"""
import hashlib
import json
import os
import time

import numpy as np
import pytest

from ptha_amd import adjoint as A
from ptha_amd import configs as C

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FP32K_STAGE_BAR = 1e-5   # metres


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def problem():
    p = os.path.join(GOLDEN, "fullsize_config5.npz")
    if not os.path.exists(p):
        pytest.skip("fullsize_config5.npz not generated")
    z = dict(np.load(p))
    cfg = C.adjoint_inversion()
    assert _digest(cfg["elev"]) == str(z["elev_sha256"])
    GX, GY = A.unit_source_basis(cfg["x"], cfg["y"], cfg["src_lon"], cfg["src_lat"], cfg["sigma_km"] * 1e3)
    eta = A.stage_from_weights(cfg["w_true"], GX, GY, cfg["elev"])
    assert _digest(eta) == str(z["init_sha256"])
    return z, cfg, GX, GY, eta


def _gmap(z, cfg, precision):
    return A.GaugeMap(cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], cfg["elev"], float(z["dt"]), int(z["nsteps"]),
                      None, cfg["gauges_xy"], precision=precision, sample_steps=z["steps"])


@pytest.mark.parametrize("precision", ["fp64", "fp32k"])
def test_config5_forward_fullsize(problem, precision):
    z, cfg, GX, GY, eta = problem
    print("CASE config 5 forward full size", precision, flush=True)
    gm = _gmap(z, cfg, precision)
    assert np.array_equal(gm.gauge_ij, z["gauge_ij"])
    t0 = time.perf_counter()
    g = gm.forward(eta)
    wall = time.perf_counter() - t0
    sub = int(z["sub"])
    stage = gm.dom.U[0][::sub, ::sub]
    wet = cfg["elev"][::sub, ::sub] < 0
    res = dict(steps=int(z["nsteps"]), samples=int(len(z["steps"])), gauge_stage_err=float(np.abs(g - z["gauge_stage"]).max()),
               final_stage_err=float(np.abs(stage - z["stage"])[wet].max()),
               gauge_amp=float(np.abs(z["gauge_stage"]).max()),
               forward_gcell_per_s=round(cfg["nx"] * cfg["ny"] * int(z["nsteps"]) / wall / 1e9, 2))
    print("config5 forward " + precision, json.dumps(res), flush=True)
    gm.close()
    if precision == "fp64":
        assert np.array_equal(g, z["gauge_stage"])
        assert np.array_equal(stage[wet], z["stage"][wet])
    else:
        assert res["gauge_stage_err"] <= FP32K_STAGE_BAR
        assert res["final_stage_err"] <= FP32K_STAGE_BAR


def test_config5_gradient_fullsize(problem):
    z, cfg, GX, GY, eta = problem
    print("CASE config 5 gradient full size", flush=True)
    elev = cfg["elev"]
    w_true = cfg["w_true"]
    gm = _gmap(z, cfg, "fp64")
    obs = gm.forward(eta)
    obs = obs * (1 + 0.01 * np.random.default_rng(cfg["seed"] + 11).standard_normal(obs.shape))
    w = 0.5 * w_true

    def loss(wv):
        gg = gm.forward(A.stage_from_weights(wv, GX, GY, elev))
        return float(((gg - obs) ** 2).sum()), gg

    L0, g0 = loss(w)
    t0 = time.perf_counter()
    lam = gm.adjoint(2 * (g0 - obs))
    t_adj = time.perf_counter() - t0
    lam_s = np.where(elev < 0, lam[0], 0.0)
    # d loss / d w[b, a] = sum over cells of lambda_S * phi_ab (wet cells)
    grad = np.einsum("bax,bx->ba", GX, GY @ lam_s)
    res = dict(loss=L0, adjoint_s=round(t_adj, 3),
               adjoint_gcell_per_s=round(cfg["nx"] * cfg["ny"] * int(z["nsteps"]) / t_adj / 1e9, 2))
    worst = 0.0
    for (b, a) in [(10, 20), (3, 7), (17, 35)]:
        e = np.zeros_like(w)
        e[b, a] = 1e-2
        fd = (loss(w + e)[0] - loss(w - e)[0]) / 2e-2
        rel = abs(grad[b, a] - fd) / abs(fd)
        res["w_%d_%d" % (b, a)] = dict(adjoint=float(grad[b, a]), central_difference=fd, rel=rel)
        worst = max(worst, rel)
    gm.close()
    # FP32 with Kahan stage: forward and adjoint in FP32, against the FP64 gradient
    gm32 = _gmap(z, cfg, "fp32k")
    g32 = gm32.forward(A.stage_from_weights(w, GX, GY, elev))
    lam32 = gm32.adjoint(2 * (g32 - obs))
    grad32 = np.einsum("bax,bx->ba", GX, GY @ np.where(elev < 0, lam32[0], 0.0))
    res["fp32k_gradient_rel"] = float(np.abs(grad32 - grad).max() / np.abs(grad).max())
    gm32.close()
    print("config5 gradient", json.dumps(res), flush=True)
    assert worst < 1e-6
    assert res["fp32k_gradient_rel"] < 1e-3
