"""The torch FP64 direct-sum restatement of the shallow-water step (tests/torch_ref_swe.py),
the reference gradient of the step adjoint, pinned on the CPU against the oracle in the
physical form (g/a^2, Gaussian mountain, regularised stage-1 Laplacian, free-surface
correction, leaf rule): its Biot-Savart part equals the oracle's direct sum, its velocity
agrees with the oracle's FMM to the FMM's truncation, its tendencies (surface gradient from
autograd label-coordinate gradients) equal the oracle's (xyzvec_from_xietavec conversion),
and its step agrees with the oracle's FMM step to the FMM's accuracy."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from torch_ref_swe import SweModel  # noqa: E402

OMEGA, DT, EPS = 7.292e-5, 900.0, 1e-4


@pytest.fixture(scope="module")
def setup(orc, sb):
    W = sb.workloads
    o = orc.OracleFMM(3, 5, 5, 0.7, 1)
    x, y, z, a = o.spec_points()
    mtn = W.gaussian_mountain(seed=5)
    G = W.GRAVITY / W.EARTH_RADIUS ** 2
    o.set_swe_physics(G, mtn, 3)
    vor, div, h = W.swe_zonal_flow(x, y, z, mtn)
    vor = vor + 0.3 * W.gaussian_vortices(x, y, z, a, seed=31)
    div = 0.05 * W.gaussian_vortices(x, y, z, a, seed=32)
    h = h + 2e6 * W.gaussian_vortices(x, y, z, a, seed=33)
    M = SweModel(x, y, z, a, 3, EPS, OMEGA, DT, G, mtn, 0)
    return o, W, M, x, y, z, a, vor, div, h


def T(v):
    return torch.tensor(np.ravel(v))


def test_velocity(orc, setup):
    o, W, M, x, y, z, a, vor, div, h = setup
    u = M.swe_velocity(T(a * vor), T(0 * a)).numpy()
    ds = np.stack([np.ravel(q) for q in orc.direct_sum_bve_vel(x.ravel(), y.ravel(), z.ravel(), a.ravel(),
                                                               vor.ravel(), EPS)], 1)
    assert np.linalg.norm(u - ds) / np.linalg.norm(ds) < 1e-13
    uf = np.stack([q.ravel() for q in o.swe_velocity(1, a, vor, div, EPS)], 1)
    u = M.swe_velocity(T(a * vor), T(a * div)).numpy()
    assert np.linalg.norm(u - uf) / np.linalg.norm(uf) < 5e-5


def test_tendency(setup):
    o, W, M, x, y, z, a, vor, div, h = setup
    uf = np.stack([q.ravel() for q in o.swe_velocity(1, a, vor, div, EPS)], 1)
    lap = np.ravel(1e3 * W.gaussian_vortices(x, y, z, a, seed=4) / 3e-5)
    vel = [uf[:, i].reshape(x.shape) for i in range(3)]
    to = o.swe_tendency(x, y, z, vel, vor, div, h, lap, OMEGA, 1)
    tt = M.tendency(torch.tensor(uf), T(vor), T(div), T(h), T(lap))
    for q in range(3):
        assert np.linalg.norm(tt[q].numpy() - np.ravel(to[q])) / np.linalg.norm(np.ravel(to[q])) < 1e-12


def test_step_agrees_with_oracle_fmm_step(setup):
    o, W, M, x, y, z, a, vor, div, h = setup
    out = M.step(T(vor), T(div), T(h))
    r = o.swe_back_rk4_step(x.ravel(), y.ravel(), z.ravel(), a, vor, div, h, DT, OMEGA, EPS, 0, 2, 1)
    for q, tol in ((0, 1e-4), (1, 1e-3), (2, 1e-4)):
        change = np.ravel(r[q]) - np.ravel((vor, div, h)[q])
        err = np.linalg.norm(out[q].numpy() - np.ravel(r[q])) / np.linalg.norm(change)
        assert err < tol, (q, err)
