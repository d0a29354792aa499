"""Output-side host logic (mirrors bve_time_evo.cpp:90-350):
unique particle positions (cubed_sphere_2d_to_1d), 2-D to 1-D vectors, the
lat-lon output grid and the NetCDF-classic history writer."""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("levels,m", [(1, 4), (2, 8), (3, 8), (3, 5)])
def test_dedupe_equals_reference_scan(sb, orc, levels, m):
    ht = sb.HostTree(sb.RunConfig(levels=levels, interp_degree=8, leaf_degree=m - 1))
    mp, cnt, x1, y1, z1 = ht.dedupe()
    x, y, z, _ = ht.spec_points()
    mo, co, xo, yo, zo = orc.dedupe(x, y, z)
    np.testing.assert_array_equal(mp.ravel(), mo)
    np.testing.assert_array_equal(cnt, co)
    for a, b in ((x1, xo), (y1, yo), (z1, zo)):
        np.testing.assert_array_equal(a, b)
    # closed form: a cubed sphere with K intervals per face edge has 6 K^2 + 2 vertices,
    # the 8 cube corners are shared by 3 faces, no point by more than 4 leaves
    K = (1 << (levels - 1)) * (m - 1)
    assert len(x1) == 6 * K * K + 2
    assert int(np.sum(cnt == 3)) == 8 and cnt.max() <= 4 and cnt.sum() == x.size
    # every particle maps to a unique point at its own position
    np.testing.assert_allclose(x1[mp.ravel()], x.ravel(), rtol=0, atol=1e-15)


def test_vec_2d_to_1d_rules(sb):
    out = sb.output
    loc = np.array([[0, 1, 1], [2, 0, 1]])
    v = np.array([[0.0, 5.0, 7.0], [3.0, 2.0, 9.0]])
    np.testing.assert_array_equal(out.vec_2d_to_1d(v, loc, 3, True), [2.0, 21.0, 3.0])
    # first non-zero value in order: point 0 takes 2.0 (its first entry is 0), point 1 takes 5.0
    np.testing.assert_array_equal(out.vec_2d_to_1d(v, loc, 3, False), [2.0, 5.0, 3.0])


def test_unique_areas_sum_to_sphere(sb):
    ht = sb.HostTree(sb.RunConfig(levels=3, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis"))
    mp, cnt, x1, y1, z1 = ht.dedupe()
    a = ht.spec_points()[3]
    a1 = sb.output.vec_2d_to_1d(a, mp, len(x1), True)
    assert abs(a1.sum() - 4 * math.pi) < 1e-12 and np.all(a1 > 0)


@pytest.mark.parametrize("n", [1178, 4706, 98306, 1572864])
def test_latlon_grid(sb, n):
    spacing, lats, lons = sb.output.latlon_grid(n)
    assert spacing == 2.0 ** math.floor(math.log2(math.sqrt(4 * math.pi / n) * 180 / math.pi))
    assert lats[0] == -90.0 and abs(lats[-1] - 90.0) < 1e-12 and len(lons) * spacing == 360.0
    x, y, z = sb.output.latlon_points(lats, lons)
    np.testing.assert_allclose(x * x + y * y + z * z, 1.0, rtol=0, atol=1e-15)


@pytest.mark.parametrize("kind", ["latlon", "points"])
def test_history_writer_round_trip(sb, tmp_path, kind):
    from scipy.io import netcdf_file
    out = sb.output
    if kind == "latlon":
        _, lats, lons = out.latlon_grid(1178)
        grid = ("latlon", lats, lons)
        shape = (len(lats), len(lons))
    else:
        ht = sb.HostTree(sb.RunConfig(levels=2, interp_degree=8, leaf_degree=7))
        mp, cnt, x1, y1, z1 = ht.dedupe()
        grid = ("points", x1, y1, z1, np.ones_like(x1))
        shape = (len(x1),)
    path = str(tmp_path / "output.nc")
    w = out.HistoryWriter(path, grid, tracers=["gv"], attrs={"tree levels": 2, "fmm theta": 0.7,
                                                             "initial condition": "gv"})
    rng = np.random.default_rng(0)
    recs = []
    for step in range(2):
        f = {name: rng.normal(size=shape) for name, _, _ in out.HistoryWriter.FIELDS}
        f["tracer_gv"] = rng.normal(size=shape)
        w.write(3600 * step, f)
        recs.append(f)
    w.close()
    with netcdf_file(path, "r", mmap=False) as r:
        assert list(r.variables["time"][:]) == [0, 3600]
        assert r.variables["vor"].units == b"1/s" and r.variables["u_vel"].units == b"m/s"
        assert r._attributes["tree levels"] == 2
        for step in range(2):
            for name, val in recs[step].items():
                np.testing.assert_array_equal(r.variables[name][step], val)
        if kind == "points":
            np.testing.assert_allclose(r.variables["latitude"][:], 90 - np.degrees(np.arccos(np.clip(z1, -1, 1))))
