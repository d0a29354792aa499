"""The lat-lon shallow-water step (swe_time_step_ll.cpp) on the GPU against the
oracle's restatement: each grid operator through the C-ABI, the velocity and
tendency sequence on the irregular tree, and whole steps -- at 5 and 3 degrees,
and one step on the reference's 1-degree grid (interp_degree 10, leaf size
4 (d+1)^2 = 484, swe_time_evo_ll.cpp:84-85).

Tolerance: 1e-12 of each field's largest magnitude for the grid operators (the
kernels round every operation as the oracle does; only libm's cos, tan, atan2
and the workgroup reduction of the polar filter differ), 1e-11 after the FMM
velocity and for whole steps."""
import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu

EPS = 1e-4
DT = 1800.0
OMEGA = 0.00007272205216643039  # run_config.hpp:31


def tangent(lat, lon, u, v):
    sl, cl, so, co = np.sin(lat), np.cos(lat), np.sin(lon), np.cos(lon)
    return [-u * so - v * sl * co, u * co - v * sl * so, v * cl]


def make_state(sb, dx, seed=0):
    L = sb.latlon
    lats, lons, x, y, z, area = L.cell_grid(dx)
    topo = L.cone_topography(lats, lons)
    vor, div, h = L.initial_state("tc5", lats, lons, topo)
    lat = np.radians(lats)[:, None] + 0 * x
    lon = np.radians(lons)[None, :] + 0 * x
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, 3)
    vor = vor + 2e-6 * np.cos(lat) * np.sin(2 * lon + ph[0])
    div = div + 1e-6 * np.cos(lat) ** 2 * np.cos(3 * lon + lat + ph[1])
    h = h + 2e-5 * np.cos(lat) * np.sin(lon + ph[2])
    u = 3e-6 * np.cos(lat) * (1 + 0.3 * np.sin(lon))
    v = 2e-6 * np.cos(2 * lon) * np.cos(lat)
    return dict(dx=dx, grid=(x.shape[0], x.shape[1], dx), lats=lats, x=x, y=y, z=z, area=area, topo=topo,
                vor=vor, div=div, h=h, vel=tangent(lat, lon, u, v), zonal=[-3e-6 * y, 3e-6 * x, 0 * x])


def close(got, ref, rel):
    ref = np.asarray(ref).ravel()
    np.testing.assert_allclose(to_np(got).ravel(), ref, rtol=0, atol=rel * max(np.abs(ref).max(), 1e-300))


@pytest.fixture(scope="module", params=[5.0, 3.0])
def case(request, sb, torch_gpu):
    dx = request.param
    print(f"\n[case lat-lon grid {dx} deg]", flush=True)
    s = make_state(sb, dx)
    ctx = sb.Context(sb.RunConfig(levels=2, interp_degree=6, kernel_eps=EPS))
    g = {k: to_gpu(torch_gpu, s[k]) for k in ("x", "y", "z", "area", "topo", "vor", "div", "h")}
    g["vel"] = [to_gpu(torch_gpu, v) for v in s["vel"]]
    g["zonal"] = [to_gpu(torch_gpu, v) for v in s["zonal"]]
    return ctx, s, g


def test_tendencies_match_oracle(case, orc):
    ctx, s, g = case
    nlat, nlon, dx = s["grid"]
    got = ctx.ll_tendencies(s["grid"], g["x"], g["y"], g["z"], g["vor"], g["div"], g["h"], g["topo"], g["vel"])
    ctx.synchronize()
    ref = orc.ll_tendencies(nlat, nlon, dx, OMEGA, s["x"], s["y"], s["z"], s["vor"], s["div"], s["h"],
                            s["topo"], s["vel"])
    assert np.abs(ref[1]).max() > 1e-10  # a divergence tendency well above the 2e-12 cut
    for a, b in zip(got, ref):
        close(a, b, 1e-12)


def test_polar_filter_matches_oracle(case, orc, torch_gpu):
    ctx, s, g = case
    nlat, nlon, dx = s["grid"]
    f = np.random.default_rng(5).uniform(-3e-11, 3e-11, size=(nlat, nlon))
    fg = to_gpu(torch_gpu, f)
    ctx.ll_polar_filter(s["grid"], g["x"], g["y"], g["z"], fg)
    ctx.synchronize()
    close(fg, orc.ll_polar_filter(nlat, nlon, dx, s["x"], s["y"], s["z"], f), 1e-12)


@pytest.mark.parametrize("flow", ["vel", "zonal"])
def test_arrival_values_match_oracle(case, orc, torch_gpu, flow):
    ctx, s, g = case
    nlat, nlon, dx = s["grid"]
    rng = np.random.default_rng(9)
    vt = 1e-10 * rng.normal(size=(nlat, nlon))
    dvt = 1e-10 * rng.normal(size=(nlat, nlon))
    got = ctx.ll_arrival_values(s["grid"], DT, g["x"], g["y"], g["z"], g["vor"], g["div"], g["h"], g[flow],
                                to_gpu(torch_gpu, vt), to_gpu(torch_gpu, dvt))
    ctx.synchronize()
    ref = orc.ll_arrival_values(nlat, nlon, dx, DT, s["x"], s["y"], s["z"], s["vor"], s["div"], s["h"], s[flow],
                                vt, dvt)
    for a, b in zip(got, ref):
        close(a, b, 1e-12)


def test_averaging_and_diffusion_match_oracle(case, orc):
    ctx, s, g = case
    nlat, nlon, dx = s["grid"]
    a = ctx.ll_field_averaging(s["grid"], g["x"], g["y"], g["z"], g["div"], 0.16)
    d = ctx.ll_laplacian_diffusion(s["grid"], DT, g["x"], g["y"], g["z"], a)
    ctx.synchronize()
    ra = orc.ll_field_averaging(nlat, nlon, s["x"], s["y"], s["z"], s["div"], 0.16)
    close(a, ra, 1e-12)
    close(d, orc.ll_laplacian_diffusion(nlat, nlon, dx, DT, s["x"], s["y"], s["z"], ra), 1e-12)


@pytest.fixture(scope="module")
def tree(case, orc):
    ctx, s, g = case
    irr = ctx.irregular(s["x"], s["y"], s["z"], 4 * 7 ** 2)
    o = orc.OracleIrreg(s["x"], s["y"], s["z"], 4 * 7 ** 2, 6, 0.7)
    return irr, o


def test_vel_tend_matches_oracle(case, tree, orc):
    ctx, s, g = case
    irr, o = tree
    nlat, nlon, dx = s["grid"]
    flat = lambda k: s[k].ravel()
    vel, tend = irr.swe_vel_tend_ll(s["grid"], g["x"], g["y"], g["z"], g["area"], g["vor"], g["div"], g["h"],
                                    g["topo"])
    ctx.synchronize()
    rvel, rvt, rdt, rht = o.swe_vel_tend_ll(nlat, nlon, dx, OMEGA, flat("area"), flat("vor"), flat("div"),
                                            flat("h"), flat("topo"), EPS)
    for a, b in zip(vel, rvel):
        close(a, b, 1e-12)
    for a, b in zip(tend, (rvt, rdt, rht)):
        close(a, b, 1e-11)


def test_two_steps_match_oracle(case, tree, orc, torch_gpu):
    ctx, s, g = case
    irr, o = tree
    nlat, nlon, dx = s["grid"]
    st = [g[k].clone() for k in ("vor", "div", "h")]
    ref = [s[k].ravel().copy() for k in ("vor", "div", "h")]
    for step in range(2):
        vel = irr.swe_back_rk4_step_ll(s["grid"], DT, g["x"], g["y"], g["z"], g["area"], *st, g["topo"])
        ctx.synchronize()
        *ref, rvel = o.swe_back_rk4_step_ll(nlat, nlon, dx, DT, OMEGA, s["area"].ravel(), *ref,
                                            s["topo"].ravel(), EPS)
        print(f"  step {step}: max |div| {np.abs(ref[1]).max():.3e}", flush=True)
        for a, b in zip(st, ref):
            close(a, b, 1e-11)
        for a, b in zip(vel, rvel):
            close(a, b, 1e-11)
    assert np.abs(ref[1]).max() > 0  # the divergence moved


def test_lat_lon_argument_checks(case, tree, sb, torch_gpu):
    ctx, s, g = case
    irr, _ = tree
    nlat, nlon, dx = s["grid"]
    out = [torch_gpu.empty_like(g["vor"]) for _ in range(3)]
    with pytest.raises(sb.SbslpmError, match="alias"):
        ctx.ll_tendencies(s["grid"], g["x"], g["y"], g["z"], g["vor"], g["div"], g["h"], g["topo"], g["vel"],
                          out=[g["vor"], out[1], out[2]])
    p = lambda t: t.data_ptr()
    # a grid under 3 x 3, and a grid that is not the tree's points: SBSLPM_E_ARG, nothing launched
    assert ctx._lib.sbslpm_ll_polar_filter(ctx.h, 2, nlon, dx, p(g["x"]), p(g["y"]), p(g["z"]), p(out[0])) == 1
    assert irr._lib.sbslpm_irreg_swe_back_rk4_step_ll(
        irr.h, nlat - 1, nlon, dx, DT, OMEGA, p(g["x"]), p(g["y"]), p(g["z"]), p(g["area"]), p(out[0]),
        p(out[1]), p(out[2]), p(g["topo"]), p(g["vel"][0]), p(g["vel"][1]), p(g["vel"][2]), 0.0) == 1
    ctx.synchronize()


def test_one_degree_step_matches_oracle(sb, orc, torch_gpu):
    """One step on the reference's grid: 1 degree (180 x 360 points), interp_degree 10,
    leaf size 484, theta 0.5, delta_t 50 s (namelist.txt), test case 5 with the cone."""
    print("\n[case lat-lon 1 degree, d=10]", flush=True)
    s = make_state(sb, 1.0, seed=4)
    ctx = sb.Context(sb.RunConfig(levels=2, interp_degree=10, kernel_eps=EPS, fmm_theta=0.5))
    irr = ctx.irregular(s["x"], s["y"], s["z"], 484)
    o = orc.OracleIrreg(s["x"], s["y"], s["z"], 484, 10, 0.5)
    g = {k: to_gpu(torch_gpu, s[k]) for k in ("x", "y", "z", "area", "topo", "vor", "div", "h")}
    st = [g["vor"], g["div"], g["h"]]
    vel = irr.swe_back_rk4_step_ll(s["grid"], 50.0, g["x"], g["y"], g["z"], g["area"], *st, g["topo"])
    ctx.synchronize()
    *ref, rvel = o.swe_back_rk4_step_ll(180, 360, 1.0, 50.0, OMEGA, s["area"].ravel(), s["vor"].ravel(),
                                        s["div"].ravel(), s["h"].ravel(), s["topo"].ravel(), EPS)
    for a, b in zip(st, ref):
        close(a, b, 1e-11)
    for a, b in zip(vel, rvel):
        close(a, b, 1e-11)
