"""Output operators on the GPU: BLI interpolation at arbitrary points (the
functor of latlon_interp_host, interp_funcs.hpp:22-60) and the Cartesian to
zonal/meridional velocity conversion (general_utils_impl.hpp:195-229), against
the oracle and closed forms; one history record written end to end."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(sb, orc, torch_gpu):
    print("\n[case output L3 m8]", flush=True)
    kw = dict(levels=3, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis")
    ctx = sb.Context(sb.RunConfig(**kw))
    o = orc.OracleFMM(3, 8, 7, 0.7, 1)
    return ctx, o


def test_interp_points_matches_oracle_and_closed_form(setup, sb, torch_gpu):
    t = torch_gpu
    ctx, o = setup
    x, y, z, a = ctx.spec_points()
    _, lats, lons = sb.output.latlon_grid(4706)
    px, py, pz = [v.ravel() for v in sb.output.latlon_points(lats, lons)]
    rng = np.random.default_rng(0)
    r = rng.normal(size=(3, 1000))
    r /= np.linalg.norm(r, axis=0)
    px, py, pz = [np.concatenate([p, q]) for p, q in zip((px, py, pz), r)]
    fields = [x * y + z, np.sin(3 * x) * np.cos(2 * z), rng.normal(size=x.shape)]
    got = [to_np(v) for v in ctx.interp_points(to_gpu(t, px), to_gpu(t, py), to_gpu(t, pz),
                                                [to_gpu(t, f) for f in fields])]
    ctx.synchronize()
    ref = o.interp_points(px, py, pz, fields)
    for g, rf in zip(got, ref):
        np.testing.assert_allclose(g, rf, rtol=0, atol=1e-13 * max(1.0, np.abs(rf).max()))
    assert np.abs(got[0] - (px * py + pz)).max() < 1e-6
    assert np.abs(got[1] - np.sin(3 * px) * np.cos(2 * pz)).max() < 1e-4


def test_uv_matches_oracle_and_solid_rotation(setup, sb, orc, torch_gpu):
    t = torch_gpu
    ctx, _ = setup
    x, y, z, _ = ctx.spec_points()
    # solid-body rotation about z with angular speed w: u = w cos(lat), v = 0
    w = 2.0
    vel = [-w * y, w * x, np.zeros_like(z)]
    u, v = ctx.xyz_vel_to_uv_vel(*[to_gpu(t, q.ravel()) for q in (x, y, z)], [to_gpu(t, q.ravel()) for q in vel])
    uo, vo = orc.xyz_vel_to_uv_vel(x, y, z, *vel)
    np.testing.assert_allclose(to_np(u), uo, rtol=0, atol=1e-15)
    np.testing.assert_allclose(to_np(v), vo, rtol=0, atol=1e-15)
    np.testing.assert_allclose(to_np(u), w * np.sqrt(x.ravel() ** 2 + y.ravel() ** 2), rtol=0, atol=1e-14)
    np.testing.assert_allclose(to_np(v), 0.0, rtol=0, atol=1e-14)


def test_history_record_end_to_end(setup, sb, torch_gpu, tmp_path):
    from scipy.io import netcdf_file
    t = torch_gpu
    ctx, _ = setup
    x, y, z, a = ctx.spec_points()
    g = [to_gpu(t, q) for q in (x, y, z, a)]
    vor = to_gpu(t, sb.workloads.gaussian_vortices(x, y, z, a, seed=3))
    vel = ctx.bve_velocity(g[3], vor)
    _, lats, lons = sb.output.latlon_grid(4706)
    rec = sb.output.bve_output_fields(ctx, g[0], g[1], g[2], vor, vel, latlon=(lats, lons))
    ctx.synchronize()
    ht = sb.HostTree(sb.RunConfig(levels=3, interp_degree=8, leaf_degree=7, weights="clenshaw_curtis"))
    mp, cnt, x1, y1, z1 = ht.dedupe()
    rec1d = sb.output.bve_output_fields(ctx, g[0], g[1], g[2], vor, vel, dedupe=(mp, len(x1)))
    for grid, fields in ((("latlon", lats, lons), rec), (("points", x1, y1, z1, np.ones_like(x1)), rec1d)):
        path = str(tmp_path / f"{grid[0]}.nc")
        w = sb.output.HistoryWriter(path, grid)
        w.write(0, fields)
        w.close()
        with netcdf_file(path, "r", mmap=False) as r:
            np.testing.assert_array_equal(r.variables["vor"][0], fields["vor"])
    # the grid vorticity is the field interpolated; the 1-D vorticity the particle values
    np.testing.assert_array_equal(rec1d["vor"][mp.ravel()[:64]], to_np(vor).ravel()[:64])
