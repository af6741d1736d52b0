"""GPU: output snapshots (swh_snapshot / swh_snapshot_download) and the NetCDF grid
writer on them. A snapshot holds the state and time of the moment it was taken, also
after further steps: FP64 leapfrog bit-identical to the oracle at that step, FP32 +
Kahan equal to the synchronous download at that step, finite volume with the device
time; a grid file written from it reads back as its FP32 rounding."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import helpers as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32k"])
def test_snapshot_is_the_state_when_taken(oracle_built, precision):
    print("CASE snapshot leapfrog", precision, flush=True)
    cfg = H.small_spherical()
    o, dt, _ = H.cpu_run("oracle", cfg, 10)
    g = H.gpu_domain(cfg, precision=precision)
    g.evolve_steps(dt, 10)
    U10 = g.U
    t10 = g.scalars()["time"]
    g.snapshot()
    g.evolve_steps(dt, 7)
    U, t = g.snapshot_download()
    assert t == t10
    assert np.array_equal(U, U10)
    if precision == "fp64":
        assert np.array_equal(U, o.get_U())
    assert not np.array_equal(U, g.U)
    # a second snapshot replaces the first
    g.snapshot()
    U2, t2 = g.snapshot_download()
    assert t2 > t and np.array_equal(U2, g.U)
    g.close()


def test_snapshot_finite_volume_device_time(oracle_built):
    print("CASE snapshot rk2 beach", flush=True)
    cfg = H.small_beach()
    g = H.gpu_domain(cfg)
    for _ in range(12):
        g.evolve_one_step()
    g.snapshot()
    s = g.scalars()
    U0 = g.U
    for _ in range(5):
        g.evolve_one_step()
    U, t = g.snapshot_download()
    assert t == s["time"]
    assert np.array_equal(U, U0)
    g.close()


def test_grid_file_from_snapshot(oracle_built, tmp_path):
    print("CASE grid output from snapshot", flush=True)
    from ptha_amd.domain import build_metrics
    from ptha_amd.output import GridOutput, write_domain_grids
    cfg = H.small_spherical()
    g = H.gpu_domain(cfg)
    m = build_metrics(cfg["nx"], cfg["ny"], cfg["lw"], cfg["ll"], True)
    fn = str(tmp_path / "Grid_output_ID0000000001.nc")
    out = GridOutput(fn, m["x"], m["y"], spatial_stride=2)
    dt = g.stationary_timestep_max()
    frames = []
    for _ in range(3):
        g.evolve_steps(dt, 5)
        t = write_domain_grids(out, g)
        frames.append((t, g.U))
    out.store_static_variable("max_stage", g.max_U("max_stage"))
    out.close()
    with netcdf_file(fn, "r", mmap=False) as f:
        for k, (t, U) in enumerate(frames):
            assert f.variables["time"][k] == t
            np.testing.assert_array_equal(f.variables["stage"][k], U[0][1::2, 1::2].astype(np.float32))
            np.testing.assert_array_equal(f.variables["elev"][k], U[3][1::2, 1::2].astype(np.float32))
    g.close()
