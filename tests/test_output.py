"""NetCDF output (ptha_amd/output.py) against the reference's schemas, read back with
scipy.io.netcdf_file on the CPU: dimensions, variable names and types, the spatial
subsampling rule (every spatial_stride'th cell from spatial_stride/2 + 1, 1-based,
netcdf_util_mod.f90:216, 260-262), values, priority-domain grids and attributes
(netcdf_util_mod.f90:165-644; point_gauge_mod.f90:247-294, 411-639)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from ptha_amd import output as O


@pytest.mark.parametrize("stride", [1, 2, 3, 4])
def test_grid_output_schema_and_values(tmp_path, stride):
    nx, ny = 37, 23
    xs = 100.0 + 0.5 * np.arange(nx)
    ys = -30.0 + 0.25 * np.arange(ny)
    rng = np.random.default_rng(stride)
    fn = str(tmp_path / "Grid_output_ID0000000001.nc")
    g = O.GridOutput(fn, xs, ys, time_grids_to_store=["stage", "uh", "vh", ""],
                     nontemporal_grids_to_store=["max_stage", "elevation0", ""], spatial_stride=stride,
                     attributes={"timestepping_method": "linear", "myid": "0000000001"})
    frames = [rng.normal(size=(4, ny, nx)) for _ in range(3)]
    for k, U in enumerate(frames):
        g.write_grids(10.0 * (k + 1), U)
    pdi = rng.integers(1, 3, size=(ny, nx))
    pim = np.ones((ny, nx), dtype=int)
    g.store_priority_domain_cells(pdi, pim, my_index=1)
    ms = rng.normal(size=(ny, nx))
    g.store_static_variable("max_stage", ms)
    g.store_static_variable("elevation0", frames[0][3])
    with pytest.raises(ValueError):
        g.store_static_variable("max_speed", ms)
    g.close()

    # 1-based start = stride/2 + 1 (integer division)
    s1 = stride // 2 + 1
    ix = np.arange(s1 - 1, nx, stride)
    iy = np.arange(s1 - 1, ny, stride)
    with netcdf_file(fn, "r", mmap=False) as f:
        assert f.dimensions["x"] == len(ix) and f.dimensions["y"] == len(iy) and f.dimensions["time"] is None
        assert f.variables["x"].typecode() == "d" and f.variables["time"].typecode() == "d"
        np.testing.assert_array_equal(f.variables["x"][:], xs[ix])
        np.testing.assert_array_equal(f.variables["y"][:], ys[iy])
        np.testing.assert_array_equal(f.variables["time"][:], [10.0, 20.0, 30.0])
        assert "elev" not in f.variables
        for k, v in enumerate(["stage", "uh", "vh"]):
            var = f.variables[v]
            assert var.dimensions == ("time", "y", "x") and var.typecode() == "f"
            for t in range(3):
                np.testing.assert_array_equal(var[t], frames[t][k][np.ix_(iy, ix)].astype(np.float32))
        assert f.variables["is_priority_domain"].typecode() == "b"
        assert f.variables["priority_domain_index"].typecode() == "h"
        np.testing.assert_array_equal(f.variables["is_priority_domain"][:], (pdi[np.ix_(iy, ix)] == 1).astype(np.int8))
        np.testing.assert_array_equal(f.variables["priority_domain_index"][:], pdi[np.ix_(iy, ix)])
        np.testing.assert_array_equal(f.variables["max_stage"][:], ms[np.ix_(iy, ix)].astype(np.float32))
        assert f.timestepping_method == b"linear"


def test_grid_output_rejects_unknown_names(tmp_path):
    with pytest.raises(ValueError):
        O.GridOutput(str(tmp_path / "a.nc"), np.arange(5.0), np.arange(4.0), time_grids_to_store=["depth"])
    with pytest.raises(ValueError):
        O.GridOutput(str(tmp_path / "b.nc"), np.arange(5.0), np.arange(4.0), spatial_stride=5)


def test_gauge_output_schema(tmp_path):
    fn = str(tmp_path / "Gauges_data_ID0000000001.nc")
    xy = np.array([[150.1, -35.2], [151.3, -34.0], [152.0, -33.5]])
    g = O.GaugeOutput(fn, xy, [1.0, 2.0, 7.5], spherical=True, attributes={"run": "test"})
    g.store_static(np.array([[0.1, 0.2, 0.3], [0, 0, 0], [0, 0, 0], [-100.0, -200.0, -300.0]]))
    times = np.array([0.0, 30.0])
    series = np.arange(2 * 3 * 3, dtype=np.float64).reshape(2, 3, 3)
    g.write_series(times, series)
    g.close()
    with netcdf_file(fn, "r", mmap=False) as f:
        assert f.dimensions["station"] == 3 and f.dimensions["time"] is None
        assert f.dimensions["NoCharStationName"] == 32
        assert f.variables["lon"].units == b"degrees_east"
        assert f.variables["time"].long_name == b"time_from_start_of_simulation"
        assert f.variables["stage"].dimensions == ("time", "station")
        np.testing.assert_array_equal(f.variables["uh"][:], series[:, 1].astype(np.float32))
        np.testing.assert_array_equal(f.variables["elevation0"][:], np.float32([-100, -200, -300]))
        assert f.featureType == b"timeSeries" and f.priority_gauges_only == b"false"
        np.testing.assert_array_equal(f.variables["gaugeID"][:], np.float32([1, 2, 7.5]))
