"""GPU: forcing patches (forcing_patch_type + apply_forcing_patch, forcing_mod.f90:24-56,
80-156, 158-239) against the reference's own solver compiled from the reference
sources (oracle/_ref), which applies the same patch through its own forcing hook.

* leapfrog ('linear' and truely-linear 'leapfrog_linear_plus_nonlinear_friction',
  spherical, Coriolis): a co-seismic style patch raising stage and elevation together
  over a rise time that starts and ends inside steps, and an instantaneous patch;
  FP64 bit-identical to the reference (state and max_U), FP32 + Kahan stage within
  1e-5 m and the FP32 elevation within 8 float spacings;
* rk2 on the beach (adaptive steps decided on the device): a stage patch; FP64 within
  round-off (1e-9) of the reference.
"""
import numpy as np
import pytest

import helpers as H
import pyoracle

pytestmark = pytest.mark.gpu


def _need_ref(spherical):
    if not pyoracle.available("ref", spherical):
        pytest.skip("compiled reference not built (oracle/_ref)")


def _patch(nx, ny, nk, scale):
    i0, i1, j0, j1 = nx // 3, nx // 3 + 17, ny // 4, ny // 4 + 23
    x = np.linspace(-1, 1, i1 - i0 + 1)[None, :]
    y = np.linspace(-1, 1, j1 - j0 + 1)[:, None]
    bump = scale * np.exp(-4 * (x ** 2 + y ** 2))
    work = np.zeros((nk, j1 - j0 + 1, i1 - i0 + 1))
    return (i0, i1, j0, j1), bump, work


@pytest.mark.parametrize("method", ["linear", "leapfrog_linear_plus_nonlinear_friction"])
@pytest.mark.parametrize("instant", [False, True])
def test_leapfrog_forcing_matches_reference(oracle_built, method, instant):
    print("CASE forcing leapfrog", method, "instant" if instant else "rise", flush=True)
    cfg = H.small_spherical()
    if method != "linear":
        cfg = dict(cfg, method=method, manning_squared=np.full((cfg["ny"], cfg["nx"]), 0.03 ** 2))
    _need_ref(True)
    (i0, i1, j0, j1), bump, work = _patch(cfg["nx"], cfg["ny"], 4, 1.5)
    work[0] = bump          # stage
    work[3] = bump          # elevation (uplift of the sea floor)
    r = pyoracle.CpuDomain("ref", cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], method=method, spherical=True)
    r.set_U(cfg["U"])
    if "manning_squared" in cfg:
        r.set_manning_squared(cfg["manning_squared"])
    dt = r.stationary_timestep_max()
    t0, t1 = (4.6 * dt, 4.6 * dt) if instant else (2.25 * dt, 7.6 * dt)
    r.set_forcing_patch(t0, t1, i0 + 1, i1 + 1, j0 + 1, j1 + 1, 1, 4, work)
    r.evolve(dt, 21)
    ref = r.get_U()
    for precision in ("fp64", "fp32k"):
        g = H.gpu_domain(cfg, precision=precision)
        g.set_forcing_patch(t0, t1, i0 + 1, i1 + 1, j0 + 1, j1 + 1, 1, 4, work)
        g.evolve_steps(dt, 10)
        g.evolve_steps(dt, 11)
        U = g.U
        if precision == "fp64":
            assert np.array_equal(U, ref)
            assert np.array_equal(g.max_U("max_stage"), r.max_U("max_stage"))
        else:
            wet = ref[3] < 0
            assert np.abs(U[0] - ref[0])[wet].max() <= 1e-5
            # the elevation is held in FP32: a few float spacings of its magnitude
            ulp = np.spacing(np.abs(ref[3]).astype(np.float32)).astype(np.float64)
            assert (np.abs(U[3] - ref[3]) <= 8 * ulp).all()
        g.close()


def test_rk2_forcing_matches_reference(oracle_built):
    print("CASE forcing rk2 beach", flush=True)
    cfg = H.small_beach()
    _need_ref(False)
    (i0, i1, j0, j1), bump, work = _patch(cfg["nx"], cfg["ny"], 1, 0.5)
    work[0] = bump
    r = H.cpu_run("ref", cfg, 0)[0]
    r.set_forcing_patch(1.0, 3.0, i0 + 1, i1 + 1, j0 + 1, j1 + 1, 1, 1, work)
    hist = []
    for _ in range(40):
        r.evolve(None, 1)
        hist.append(r.scalars())
    g = H.gpu_domain(cfg)
    g.set_forcing_patch(1.0, 3.0, i0 + 1, i1 + 1, j0 + 1, j1 + 1, 1, 1, work)
    for k in range(40):
        g.evolve_one_step()
        assert g.scalars()["time"] == pytest.approx(hist[k]["time"], rel=1e-12)
    U, ref = g.U, r.get_U()
    wet = cfg["U"][3] < 1e5
    for k in range(3):
        assert np.abs(U[k] - ref[k])[wet].max() <= 1e-9, k
    g.close()


def test_forcing_argument_checks():
    """UH/VH forcing on a leapfrog domain and elevation forcing on a finite-volume
    domain are rejected (the latter needs the reference's support_elevation_forcing)."""
    from ptha_amd.backend import BackendError as SwalsError
    cfg = H.small_spherical()
    g = H.gpu_domain(cfg)
    work = np.zeros((4, 3, 3))
    work[1] = 1.0
    with pytest.raises(SwalsError):
        g.set_forcing_patch(0.0, 1.0, 5, 7, 5, 7, 1, 4, work)
    with pytest.raises(SwalsError):
        g.set_forcing_patch(2.0, 1.0, 5, 7, 5, 7, 1, 1, work[:1])
    g.close()
    cb = H.small_beach()
    g = H.gpu_domain(cb)
    work = np.zeros((1, 3, 3)) + 1.0
    with pytest.raises(SwalsError):
        g.set_forcing_patch(0.0, 1.0, 5, 7, 5, 7, 4, 4, work)
    g.close()
