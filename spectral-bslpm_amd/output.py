"""History output of the BVE driver without the NetCDF C library.

The reference writes one NetCDF-4 file per run (bve_time_evo.cpp:125-230,
300-350) either on a regular lat-lon grid (``interp_output``: every field
BLI-interpolated to the grid, interp_funcs.hpp:22-77) or on the unique
particle positions (``cubed_sphere_2d_to_1d``, initialize_cubed_sphere.cpp:
290-324, and ``vec_2d_to_1d``, initialize_cubed_sphere.hpp:58-86).  This
module keeps the variable names, dimensions and attributes and writes the
NetCDF classic format with ``scipy.io.netcdf_file`` (pure Python); the grid
interpolation and the (u, v) conversion run on the GPU through the library
(``Context.interp_points``, ``Context.xyz_vel_to_uv_vel``)."""
import math

import numpy as np


def latlon_grid(point_count):
    """Output grid of bve_time_evo.cpp:93-95, 206-211: spacing (degrees) =
    2^floor(log2(sqrt(4 pi / point_count) * 180/pi)), latitudes -90 + spacing*i
    (180/spacing + 1 of them), longitudes spacing*j (360/spacing)."""
    spacing = 2.0 ** math.floor(math.log2(math.sqrt(4 * math.pi / point_count) * 180.0 / math.pi))
    lat_count = int(180.0 / spacing + 1.0)
    lon_count = int(360.0 / spacing)
    return spacing, -90.0 + spacing * np.arange(lat_count), spacing * np.arange(lon_count)


def latlon_points(lats, lons):
    """Unit vectors of the grid points [lat][lon] as latlon_interp_host computes
    them (interp_funcs.hpp:38-46)."""
    lat = np.radians(np.asarray(lats, dtype=np.float64))[:, None]
    lon = np.radians(np.asarray(lons, dtype=np.float64))[None, :]
    x = np.cos(lat) * np.cos(lon)
    y = np.cos(lat) * np.sin(lon)
    z = np.sin(lat) * np.ones_like(lon)
    return x, y, z


def vec_2d_to_1d(vec, two_d_to_1d, n1d, add):
    """initialize_cubed_sphere.hpp:58-86: with ``add`` the sum of the values at
    coincident particles (areas); otherwise the first non-zero value in
    [leaf][node] order (the reference applies that rule from parallel threads;
    here the order is fixed)."""
    v = np.asarray(vec, dtype=np.float64).ravel()
    loc = np.asarray(two_d_to_1d).ravel()
    if add:
        return np.bincount(loc, weights=v, minlength=n1d)
    out = np.zeros(n1d)
    nz = np.nonzero(v)[0]
    first_loc, first = np.unique(loc[nz], return_index=True)
    out[first_loc] = v[nz[first]]
    return out


class HistoryWriter:
    """NetCDF classic history file with the reference's layout.

    ``grid`` is either ("latlon", lats, lons) or ("points", x1d, y1d, z1d,
    area1d).  Variables: time (int, "s"); latitude/longitude ("degrees_north"/
    "degrees_east"); for point output also point_index, x, y, z and area; the
    fields vor ("1/s"), u_vel, v_vel, vel_x, vel_y, vel_z ("m/s") and
    tracer_<name>, each [time][lat][lon] or [time][point_index]; global
    attributes as bve_time_evo.cpp:196-204."""

    FIELDS = (("vor", "1/s", "relative vorticity"), ("u_vel", "m/s", "zonal velocity"),
              ("v_vel", "m/s", "meridional velocity"), ("vel_x", "m/s", "x velocity"),
              ("vel_y", "m/s", "y velocity"), ("vel_z", "m/s", "z velocity"))

    def __init__(self, path, grid, tracers=(), attrs=None):
        from scipy.io import netcdf_file
        self.f = netcdf_file(path, "w", version=2)
        self.kind = grid[0]
        self.f.createDimension("time", None)
        tv = self.f.createVariable("time", "i4", ("time",))
        tv.units = "s"
        if self.kind == "latlon":
            lats, lons = grid[1], grid[2]
            self.f.createDimension("latitude", len(lats))
            self.f.createDimension("longitude", len(lons))
            la = self.f.createVariable("latitude", "f8", ("latitude",))
            lo = self.f.createVariable("longitude", "f8", ("longitude",))
            la[:] = lats
            lo[:] = lons
            dims = ("time", "latitude", "longitude")
        elif self.kind == "points":
            x1d, y1d, z1d, a1d = [np.asarray(v, dtype=np.float64) for v in grid[1:5]]
            n = len(x1d)
            self.f.createDimension("point_index", n)
            pi = self.f.createVariable("point_index", "i4", ("point_index",))
            pi[:] = np.arange(n, dtype=np.int32)
            for name, val in (("x", x1d), ("y", y1d), ("z", z1d), ("area", a1d)):
                v = self.f.createVariable(name, "f8", ("point_index",))
                v[:] = val
            self.f.variables["area"]._attributes["long name"] = "area on a sphere of radius 1"
            la = self.f.createVariable("latitude", "f8", ("point_index",))
            lo = self.f.createVariable("longitude", "f8", ("point_index",))
            # bve_time_evo.cpp:234-237
            la[:] = 90.0 - np.degrees(np.arccos(np.clip(z1d, -1.0, 1.0)))
            lo[:] = np.degrees(np.arctan2(y1d, x1d))
            dims = ("time", "point_index")
        else:
            raise ValueError("grid must be ('latlon', lats, lons) or ('points', x, y, z, area)")
        la.units = "degrees_north"
        lo.units = "degrees_east"
        for name, units, long in self.FIELDS:
            v = self.f.createVariable(name, "f8", dims)
            v.units = units
            v._attributes["long name"] = long
        self.tracers = list(tracers)
        for name in self.tracers:
            self.f.createVariable("tracer_" + name, "f8", dims)
        for k, val in (attrs or {}).items():
            self.f._attributes[k] = val
        self.n = 0

    def write(self, time_s, fields):
        """Appends one record: ``fields`` maps variable names (vor, u_vel, ...,
        tracer_<name>) to arrays already on this file's grid."""
        self.f.variables["time"][self.n] = int(time_s)
        for name, val in fields.items():
            v = self.f.variables[name]
            v[self.n] = np.asarray(val, dtype=np.float64).reshape(v.shape[1:])
        self.n += 1

    def close(self):
        self.f.close()


def bve_output_fields(ctx, x, y, z, vor, vel, tracers=None, tracer_names=(), latlon=None, dedupe=None):
    """The fields of one BVE history record (bve_time_evo.cpp:307-350), on the
    lat-lon grid (``latlon`` = (lats, lons)) or on the unique particles
    (``dedupe`` = (two_d_to_1d, n1d)).  x, y, z, vor, vel (3) are the context's
    device tensors [Pl, m2]; tracers [Pl, m2, T]."""
    u, v = ctx.xyz_vel_to_uv_vel(x.reshape(-1), y.reshape(-1), z.reshape(-1), [w.reshape(-1) for w in vel])
    named = [("vor", vor), ("u_vel", u.reshape(vor.shape)), ("v_vel", v.reshape(vor.shape)), ("vel_x", vel[0]),
             ("vel_y", vel[1]), ("vel_z", vel[2])]
    for k, name in enumerate(tracer_names):
        named.append(("tracer_" + name, tracers[..., k].contiguous()))
    if latlon is not None:
        torch = ctx.torch
        gx, gy, gz = [torch.from_numpy(np.ascontiguousarray(a.ravel())).cuda() for a in latlon_points(*latlon)]
        out = {}
        for i in range(0, len(named), 8):
            chunk = named[i:i + 8]
            vals = ctx.interp_points(gx, gy, gz, [t.contiguous() for _, t in chunk])
            for (name, _), val in zip(chunk, vals):
                out[name] = val.cpu().numpy().reshape(len(latlon[0]), len(latlon[1]))
        return out
    m, n1d = dedupe
    return {name: vec_2d_to_1d(t.cpu().numpy(), m, n1d, False) for name, t in named}
