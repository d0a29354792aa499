"""Host-side set-up of the lat-lon shallow-water path: the cell-centred grid,
its point areas, the reference's lat-lon initial conditions and topography.

Fields are [lat_count][lon_count] row-major (index i * lon_count + j), as the
reference's Kokkos::View<double**, LayoutRight> arrays.
"""
import numpy as np

EARTH_RADIUS = 6371000.0
GRAVITY = 9.81
OMEGA = 0.00007272205216643039  # run_config.hpp:31


def cell_grid(spacing):
    """swe_time_evo_ll.cpp:48-80 with latlon_point (general_utils_impl.hpp:252-276):
    lat_count = 180/spacing rows at -90 + spacing (i + 1/2) degrees, lon_count =
    360/spacing columns at spacing (j + 1/2) degrees; each point's area is its cell's
    (sin(lat + dx/2) - sin(lat - dx/2)) dx with the edges clipped to the poles.
    Returns (lats, lons, x, y, z, area), the last four [lat_count, lon_count]."""
    lat_count = int(180.0 / spacing)
    lon_count = int(360.0 / spacing)
    if lat_count < 3 or lon_count < 3:
        raise ValueError(f"grid spacing {spacing} gives fewer than 3 x 3 points")
    lats = -90.0 + spacing * np.arange(lat_count) + 0.5 * spacing
    lons = spacing * np.arange(lon_count) + 0.5 * spacing
    lat = np.repeat((lats * np.pi / 180.0)[:, None], lon_count, axis=1)
    lon = np.repeat((lons * np.pi / 180.0)[None, :], lat_count, axis=0)
    x = np.cos(lat) * np.cos(lon)
    y = np.cos(lat) * np.sin(lon)
    z = np.sin(lat)
    rad_dx = spacing * np.pi / 180.0
    lat1 = np.minimum(np.pi / 2.0, lat + 0.5 * rad_dx)
    lat2 = np.maximum(-np.pi / 2.0, lat - 0.5 * rad_dx)
    area = (np.sin(lat1) - np.sin(lat2)) * rad_dx
    return lats, lons, x, y, z, area


def cone_topography(lats, lons):
    """cone_mountain_ll (topography.cpp:111-131): a 2000 m cone (in Earth radii) of
    radius pi/9 centred at (30N, 90E)."""
    lat = (np.asarray(lats) * np.pi / 180.0)[:, None]
    lon = (np.asarray(lons) * np.pi / 180.0)[None, :]
    radius = np.pi / 9.0
    r = np.sqrt(np.fmin((lon - 0.5 * np.pi) ** 2 + (lat - np.pi / 6.0) ** 2, radius * radius))
    return 2000.0 / EARTH_RADIUS * (1.0 - r / radius)


def initial_state(case, lats, lons, topo=None):
    """swe_initialize_ll (initial_conditions.cpp:729-746) for "tc2" (:437-457) and "tc5"
    (:486-506): zonal flow in geostrophic balance, height in Earth radii, minus the
    topography (apply_topo, :696-709).  Returns (vors, divs, height)."""
    lat = np.repeat((np.asarray(lats) * np.pi / 180.0)[:, None], len(lons), axis=1)
    slat = np.sin(lat)
    g = GRAVITY / EARTH_RADIUS
    if case == "tc2":
        u0 = np.pi / (6.0 * 86400.0)
        h = 29400.0 / (GRAVITY * EARTH_RADIUS) - 1.0 / g * (2.0 * np.pi * u0 / 86400.0 + 0.5 * u0 * u0) * slat * slat
    elif case == "tc5":
        u0 = 20.0 / EARTH_RADIUS
        h = 5960.0 / EARTH_RADIUS - 1.0 / g * (2.0 * np.pi * u0 / 86400.0 + 0.5 * u0 * u0) * slat * slat
    else:
        raise ValueError(f"unknown lat-lon initial condition {case!r} (tc2, tc5)")
    vors = 2.0 * u0 * slat
    divs = np.zeros_like(vors)
    if topo is not None:
        h = h - topo
    return vors, divs, h
