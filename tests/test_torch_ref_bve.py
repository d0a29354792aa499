"""The torch FP64 direct-sum restatement of the BVE step (tests/torch_ref_bve.py),
used as the reference gradient of the step adjoint, pinned on the CPU:
its Biot-Savart sums equal the oracle's direct sum, its BLI reproduces smooth
fields, and its step agrees with the oracle's FMM step to the FMM's accuracy
(the displaced stages carry the reference scheme's interpolated proxy
displacements, about 1e-5 relative at d = 8)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from torch_ref_bve import StepModel, project  # noqa: E402

EPS, DT, OMEGA = 1e-4, 3600.0, 7.292e-5


@pytest.fixture(scope="module")
def setup(orc, sb):
    o = orc.OracleFMM(3, 8, 7, 0.7, 1)
    x, y, z, a = o.spec_points()
    w = sb.workloads.gaussian_vortices(x, y, z, a, seed=12)
    return o, x, y, z, a, w, StepModel(x, y, z, a, 3, EPS, OMEGA, DT)


def test_velocity_equals_oracle_direct_sum(orc, setup):
    o, x, y, z, a, w, M = setup
    vt = M.velocity(torch.tensor(w.ravel())).numpy()
    vd = np.stack([np.asarray(v).ravel() for v in orc.direct_sum_bve_vel(x.ravel(), y.ravel(), z.ravel(), a.ravel(),
                                                                          w.ravel(), EPS)], 1)
    assert np.linalg.norm(vd - vt) / np.linalg.norm(vd) < 1e-13


def test_interpolation_of_smooth_field(setup):
    o, x, y, z, a, w, M = setup
    X = M.X
    p = project(X + 1e-3 * torch.randn(X.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(1)))
    f = X[:, 0] * X[:, 1] + X[:, 2] ** 2
    got = M.interp([f], p)[0]
    assert (got - (p[:, 0] * p[:, 1] + p[:, 2] ** 2)).abs().max().item() < 1e-8


def test_step_agrees_with_oracle_fmm_step(setup):
    o, x, y, z, a, w, M = setup
    new_o = np.asarray(o.bve_back_rk4_step(x, y, z, a, w.copy(), DT, OMEGA, EPS, 0)[0]).ravel()
    new_t = M.step(torch.tensor(w.ravel())).numpy()
    change = new_o - w.ravel()
    assert np.linalg.norm(new_o - new_t) / np.linalg.norm(new_o) < 1e-6
    assert np.linalg.norm(change - (new_t - w.ravel())) / np.linalg.norm(change) < 1e-4
