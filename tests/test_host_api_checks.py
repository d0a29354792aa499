"""Argument checks of the Python host layer (no GPU needed).

Every pointer the wrappers hand to libsbslpm.so goes through a check of
device, dtype, layout and element count, so a short array, a host tensor or a
float32 tensor raises ValueError before any kernel could read or write outside
the caller's buffer.  The wrappers are exercised on a context object whose
library handle fails the test if it is ever called."""
import importlib

import pytest

torch = pytest.importorskip("torch")

api = importlib.import_module("spectral-bslpm_amd.api")


class _NoLib:
    def __getattr__(self, name):
        def call(*args):
            raise AssertionError(f"library function {name} called although an argument is invalid")
        return call


def _ctx(P=10, Pl=6, n=4, m2=9):
    c = api.Context.__new__(api.Context)
    c.sizes = api.Sizes([P, Pl, n, m2, 0, Pl * m2, P - Pl, 0])
    c._lib = _NoLib()
    c.h = None
    c.torch = torch
    c.cfg = api.RunConfig(levels=1, interp_degree=1)
    return c


@pytest.fixture
def host_as_device(monkeypatch):
    """Host tensors pass the device check, so the checks of later arguments are
    reached; the library is never called because one argument is always invalid."""
    monkeypatch.setattr(api, "_on_gpu", lambda t: True)


def part(c, dtype=torch.float64, k=None):
    shape = (c.sizes.Pl, c.sizes.m2) if k is None else (k, c.sizes.Pl, c.sizes.m2)
    return torch.zeros(shape, dtype=dtype)


def prox(c):
    return torch.zeros((c.sizes.P, c.sizes.n), dtype=torch.float64)


def test_ptr_checks_each_property():
    good = torch.zeros(12, dtype=torch.float64)
    with pytest.raises(ValueError, match="12 elements"):
        api._ptr(good[:11], 12, "a")
    with pytest.raises(ValueError, match="float64"):
        api._ptr(good.float(), 12, "a")
    with pytest.raises(ValueError, match="contiguous"):
        api._ptr(torch.zeros(24, dtype=torch.float64)[::2], 12, "a")
    with pytest.raises(ValueError, match="GPU"):
        api._ptr(good, 12, "a")  # a host tensor of the right size and dtype
    with pytest.raises(ValueError, match="torch tensor"):
        api._ptr([0.0] * 12, 12, "a")
    assert api._ptr(None, 12, "a") is None
    with pytest.raises(ValueError, match="int32"):
        api._iptr(torch.zeros(4, dtype=torch.int64), 4, "i")


def test_wrappers_reject_wrong_shapes_before_calling_the_library(host_as_device):
    c = _ctx()
    a, v = part(c), part(c)
    short = torch.zeros(c.sizes.Pl * c.sizes.m2 - 1, dtype=torch.float64)
    vel = [part(c) for _ in range(3)]
    with pytest.raises(ValueError, match="elements"):
        c.bve_velocity(short, v, vel=vel)
    with pytest.raises(ValueError, match="3 arrays"):
        c.bve_velocity(a, v, vel=vel[:2])
    with pytest.raises(ValueError, match="elements"):
        c.bve_velocity(a, v, disp_part=vel, disp_proxy=vel, vel=vel)  # particle-sized proxy displacement
    with pytest.raises(ValueError, match="elements"):
        c.upward_pass(a, v, q_part=part(c), q_proxy=part(c))
    with pytest.raises(ValueError, match="float64"):
        c.poisson_solve(a, part(c, torch.float32), psi=part(c))


def test_wrappers_reject_host_tensors():
    c = _ctx()
    a = part(c)
    with pytest.raises(ValueError, match="GPU"):
        c.poisson_solve(a, a, psi=part(c))
    with pytest.raises(ValueError, match="GPU"):
        c.bve_find_departure_points(a, a, a, [[part(c) for _ in range(3)] for _ in range(4)], 3600.0,
                                    dep=[part(c) for _ in range(3)])


def test_stage_velocity_and_departure_wrappers_check_every_pointer(host_as_device):
    c = _ctx()
    x = part(c)
    good = [[part(c) for _ in range(3)] for _ in range(4)]
    dep = [part(c) for _ in range(3)]
    with pytest.raises(ValueError, match="4 arrays"):
        c.bve_find_departure_points(x, x, x, good[:3], 3600.0, dep=dep)
    bad = [list(s) for s in good]
    bad[2][1] = torch.zeros(5, dtype=torch.float64)
    with pytest.raises(ValueError, match="elements"):
        c.bve_find_departure_points(x, x, x, bad, 3600.0, dep=dep)
    with pytest.raises(ValueError, match="elements"):
        c.swe_find_departure_points(x, x, x, bad, 3600.0, dep=dep)
    with pytest.raises(ValueError, match="elements"):
        c.bve_find_departure_points(x, x, x, good, 3600.0, dep=dep[:2] + [prox(c)])


def test_ensemble_and_record_wrappers_check_member_counts(host_as_device):
    c = _ctx()
    x = part(c)
    vors = part(c, k=4)
    with pytest.raises(ValueError, match="elements"):
        c.bve_back_rk4_step_ensemble(x, x, x, x, vors, [part(c, k=3) for _ in range(3)])
    with pytest.raises(ValueError, match="members"):
        c.bve_back_rk4_step_ensemble(x, x, x, x, torch.zeros(5, dtype=torch.float64), [x, x, x])
    with pytest.raises(ValueError, match="elements"):
        c.bve_back_rk4_step_record(x, x, x, x, x, vor_out=part(c), stages=part(c, k=11))


def test_interp_points_checks_fields_and_outputs(host_as_device):
    c = _ctx()
    pts = torch.zeros(7, dtype=torch.float64)
    with pytest.raises(ValueError, match="elements"):
        c.interp_points(pts, pts, pts, [part(c)], out=[torch.zeros(6, dtype=torch.float64)])
    with pytest.raises(ValueError, match="elements"):
        c.interp_points(pts, pts, pts, [torch.zeros(3, dtype=torch.float64)], out=[pts])
    with pytest.raises(ValueError, match="1 arrays"):
        c.interp_points(pts, pts, pts, [part(c)], out=[pts, pts])
    with pytest.raises(ValueError, match="elements"):
        c.interp_points_adjoint(pts, pts, pts, [pts], field_bar=[torch.zeros(3, dtype=torch.float64)])


def test_direct_sums_check_lengths(host_as_device):
    c = _ctx()
    s = torch.zeros(20, dtype=torch.float64)
    t = torch.zeros(8, dtype=torch.float64)
    with pytest.raises(ValueError, match="elements"):
        c.direct_sum_bve_vel_targets(t, t, t, s, s, s, s[:19], s, vel=[t, t, t])
    with pytest.raises(ValueError, match="elements"):
        c.direct_sum_bve_vel(s, s, s, s, s, vel=[s, s, t])
    with pytest.raises(ValueError, match="elements"):
        c.direct_sum_inv_lap_targets(t, t, t, s, s, s, s, s, soln=s)
