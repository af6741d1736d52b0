"""NetCDF output in the reference's schemas, for runs driven from Python.

* GridOutput mirrors nc_grid_output_type (propagation/SWALS/src/util/netcdf_util_mod.f90:
  initialise :165-435, write_grids :437-506, store_priority_domain_cells :509-580,
  store_static_variable :583-629, finalise :631-644) as domain%create_output_files /
  write_to_output_files / write_max_quantities call it (domain_mod.f90:2090-2218,
  2425-2465): dimensions x, y and an unlimited time; x, y, time in FP64; the chosen
  time grids of 'stage', 'uh', 'vh', 'elev' and the static grids ('max_stage', ...,
  'elevation0') in output_precision (FP32 by default, global_mod.f90:22);
  is_priority_domain (byte), priority_domain_index and priority_domain_image (int16);
  every spatial_stride'th cell from spatial_stride // 2 + 1 (1-based).
* GaugeOutput mirrors setup_gauge_netcdf_output / write_current_time_series
  (point_gauge_mod.f90:247-294, 411-639): dimensions station, time (unlimited),
  NoCharStationName; lon, lat, time, gaugeID and the gauge series in FP32, with the
  reference's CF attributes.

Files are NetCDF classic (64-bit offset) written with scipy.io.netcdf_file, which the
reference's readers (R ncdf4, Python netCDF4/xarray) open unchanged. The grids come from
Domain.snapshot / snapshot_download, so the device keeps stepping while a file is
written.
"""
import numpy as np
from scipy.io import netcdf_file

TIME_VARS = ["stage", "uh", "vh", "elev"]                       # netcdf_util_mod.f90:36-37
MAX_U_VARS = ["max_stage", "max_speed", "max_flux", "arrival_time", "time_of_max_stage", "min_stage"]  # :39-40
OTHER_STATIC_VARS = ["manning_squared", "elevation0", "elevation_source_file_index"]                  # :43-44
STATIC_VARS = MAX_U_VARS + OTHER_STATIC_VARS


def _flag(names, allowed, what):
    for n in names:
        if n != "" and n not in allowed:
            raise ValueError("Unknown value in %s: %s" % (what, n))


class GridOutput:
    """Gridded output of one domain (nc_grid_output_type)."""

    def __init__(self, filename, xs, ys, time_grids_to_store=("stage", "uh", "vh", "elev"),
                 nontemporal_grids_to_store=("max_stage", "elevation0"), output_precision=np.float32,
                 spatial_stride=1, attributes=None):
        xs = np.asarray(xs, dtype=np.float64)
        ys = np.asarray(ys, dtype=np.float64)
        nx, ny = len(xs), len(ys)
        if spatial_stride > min(nx, ny):
            raise ValueError("Cannot have spatial_stride > minval(domain%nx)")
        _flag(time_grids_to_store, TIME_VARS, "time_grids_to_store")
        _flag(nontemporal_grids_to_store, STATIC_VARS, "nontemporal_grids_to_store")
        self.nx, self.ny = nx, ny
        self.stride = int(spatial_stride)
        self.start = self.stride // 2          # 0-based form of spatial_stride/2 + 1
        self.prec = np.dtype(output_precision)
        self.time_vars = [v for v in TIME_VARS if v in time_grids_to_store]
        self.static_vars = [v for v in STATIC_VARS if v in nontemporal_grids_to_store]
        self.num_output_steps = 0
        self.f = netcdf_file(filename, "w", version=2)
        f = self.f
        xsel, ysel = xs[self.start::self.stride], ys[self.start::self.stride]
        # the record dimension is defined first (the writer requires it; readers go by name)
        f.createDimension("time", None)
        f.createDimension("x", len(xsel))
        f.createDimension("y", len(ysel))
        f.createVariable("x", "d", ("x",))[:] = xsel
        f.createVariable("y", "d", ("y",))[:] = ysel
        f.createVariable("time", "d", ("time",))
        for v in self.time_vars:
            f.createVariable(v, self.prec.char, ("time", "y", "x"))
        f.createVariable("is_priority_domain", "b", ("y", "x"))
        f.createVariable("priority_domain_index", "h", ("y", "x"))
        f.createVariable("priority_domain_image", "h", ("y", "x"))
        for v in self.static_vars:
            f.createVariable(v, self.prec.char, ("y", "x"))
        for k, val in (attributes or {}).items():
            setattr(f, k, str(val))

    def _sub(self, a):
        return np.asarray(a)[self.start::self.stride, self.start::self.stride]

    def write_grids(self, time, U):
        """U: (4, ny, nx) = [stage, uh, vh, elev] (Domain.snapshot_download)."""
        U = np.asarray(U)
        if U.shape[0] != 4:
            raise ValueError("Incorrect size of 3 rank of U when writing grids to netcdf")
        k = self.num_output_steps
        self.num_output_steps += 1
        self.f.variables["time"][k] = time
        for v in self.time_vars:
            self.f.variables[v][k] = self._sub(U[TIME_VARS.index(v)]).astype(self.prec)

    def store_priority_domain_cells(self, priority_domain_index, priority_domain_image, my_index, my_image=1):
        pdi = self._sub(priority_domain_index)
        pim = self._sub(priority_domain_image)
        self.f.variables["is_priority_domain"][:] = ((pdi == my_index) & (pim == my_image)).astype(np.int8)
        self.f.variables["priority_domain_index"][:] = pdi.astype(np.int16)
        self.f.variables["priority_domain_image"][:] = pim.astype(np.int16)

    def store_static_variable(self, name, values):
        if name not in self.static_vars:
            raise ValueError("Cannot store static variable: %s" % name)
        self.f.variables[name][:] = self._sub(values).astype(self.prec)

    def close(self):
        self.f.close()


def write_domain_grids(out, domain):
    """write_to_output_files for a Domain: a snapshot taken now, downloaded, written."""
    domain.snapshot()
    U, t = domain.snapshot_download()
    out.write_grids(t, U)
    return t


class GaugeOutput:
    """Point-gauge time series of one domain (point_gauge_type's NetCDF file)."""

    _STATIC = {0: ("stage0", "water_surface_height_above_mean_sea_level_at_initial_condition", "m"),
               1: ("uh0", "x_velocity_multiplied_by_depth_at_initial_condition", "m^2/s"),
               2: ("vh0", "y_velocity_multiplied_by_depth_at_initial_condition", "m^2/s"),
               3: ("elevation0", "ground_level_altitude_above_mean_sea_level_at_initial_condition", "m")}
    _SERIES = {0: ("stage", "water_surface_height_above_mean_sea_level", "m"),
               1: ("uh", "x_velocity_multiplied_by_depth", "m^2/s"),
               2: ("vh", "y_velocity_multiplied_by_depth", "m^2/s"),
               3: ("elevation", "ground_level_altitude_above_mean_sea_level", "m")}

    def __init__(self, filename, xy, gauge_ids, spherical=True, time_series_var=(0, 1, 2), static_var=(3,),
                 attributes=None, priority_gauges_only=False):
        xy = np.asarray(xy, dtype=np.float64).reshape(-1, 2)
        self.n = len(xy)
        self.series_var = list(time_series_var)
        self.static_var = list(static_var)
        self.num_output_steps = 0
        self.f = f = netcdf_file(filename, "w", version=2)
        f.createDimension("time", None)
        f.createDimension("station", self.n)
        f.createDimension("NoCharStationName", 32)
        lon = f.createVariable("lon", "f", ("station",))
        lat = f.createVariable("lat", "f", ("station",))
        if spherical:
            lon.standard_name, lon.long_name, lon.units = "longitude", "station_longitude", "degrees_east"
            lat.standard_name, lat.long_name, lat.units = "latitude", "station_latitude", "degrees_north"
        else:
            lon.standard_name, lon.long_name, lon.units = "projection_x_coordinate", "station_x_coordinate", "m"
            lat.standard_name, lat.long_name, lat.units = "projection_y_coordinate", "station_y_coordinate", "m"
        t = f.createVariable("time", "f", ("time",))
        t.standard_name, t.long_name, t.units = "time", "time_from_start_of_simulation", "s"
        gid = f.createVariable("gaugeID", "f", ("station",))
        gid.long_name, gid.units = "real_ID_for_each_station", "-"
        for k in self.static_var:
            name, long_name, units = self._STATIC[k]
            v = f.createVariable(name, "f", ("station",))
            v.long_name, v.units = long_name, units
        for k in self.series_var:
            name, long_name, units = self._SERIES[k]
            v = f.createVariable(name, "f", ("time", "station"))
            v.long_name, v.units = long_name, units
        f.featureType = "timeSeries"
        for k, val in (attributes or {}).items():
            setattr(f, k, str(val))
        f.priority_gauges_only = "true" if priority_gauges_only else "false"
        lon[:] = xy[:, 0]
        lat[:] = xy[:, 1]
        gid[:] = np.asarray(gauge_ids, dtype=np.float64)

    def store_static(self, values):
        """values: (4, n) [stage, uh, vh, elev] at the gauges at the initial condition."""
        for k in self.static_var:
            self.f.variables[self._STATIC[k][0]][:] = np.asarray(values[k], dtype=np.float32)

    def write_current_time_series(self, time, values):
        """values: (4, n) or (3, n) [stage, uh, vh(, elev)] at the gauges at this time."""
        k = self.num_output_steps
        self.num_output_steps += 1
        self.f.variables["time"][k] = time
        for q in self.series_var:
            self.f.variables[self._SERIES[q][0]][k] = np.asarray(values[q], dtype=np.float32)

    def write_series(self, times, series):
        """A whole device gauge record: times (m,), series (m, 3, n) from Domain.gauge_series."""
        for t, v in zip(times, series):
            self.write_current_time_series(t, v)

    def close(self):
        self.f.close()
