"""Allocation-free direct sum (sbslpm_direct_sum_inv_lap_noalloc) against the
allocating entry point sbslpm_direct_sum_inv_lap, both called through the C ABI."""
import ctypes

import numpy as np
import pytest

from conftest import to_gpu, to_np

pytestmark = pytest.mark.gpu


def _points(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((3, n))
    p /= np.linalg.norm(p, axis=0)
    area = rng.uniform(0.5, 1.5, n) * (4 * np.pi / n)
    pots = rng.standard_normal(n)
    return p[0], p[1], p[2], area, pots


def _call(lib, fn, ctx, arrs, soln, lb, ub):
    n = arrs[0].numel()
    ptrs = [ctypes.c_void_p(t.data_ptr()) for t in arrs] + [ctypes.c_void_p(soln.data_ptr())]
    rc = getattr(lib, fn)(ctx.h, n, *ptrs, lb, ub)
    return rc


def _err(lib, ctx):
    msg = lib.sbslpm_last_error(ctx.h)
    return msg.decode() if msg else ""


@pytest.fixture(scope="module")
def ctx(sb, torch_gpu):
    cfg = sb.RunConfig(levels=2, interp_degree=4, leaf_degree=4)
    return sb.Context(cfg)


def test_noalloc_matches_allocating_entry_point_twice(sb, torch_gpu, ctx):
    torch = torch_gpu
    lib = sb._lib.load()
    # two calls in a row, the second with a smaller input; the second also on a
    # slice, accumulating into a non-zero solution
    for n, seed, lb, ub, start in ((20000, 1, 0, 20000, 0.0), (3001, 2, 117, 2950, 0.25)):
        arrs = [to_gpu(torch, a) for a in _points(n, seed)]
        base = torch.full((n,), start, dtype=torch.float64, device="cuda")
        ref, out = base.clone(), base.clone()
        assert _call(lib, "sbslpm_direct_sum_inv_lap", ctx, arrs, ref, lb, ub) == 0, _err(lib, ctx)
        assert _call(lib, "sbslpm_direct_sum_inv_lap_noalloc", ctx, arrs, out, lb, ub) == 0, _err(lib, ctx)
        ctx.synchronize()
        r, o = to_np(ref), to_np(out)
        # the same chunking and summation order: equal to the bit
        assert np.array_equal(r, o), (n, np.max(np.abs(r - o)))
        assert np.all(o[:lb] == start) and np.all(o[ub:] == start)
        assert np.count_nonzero(o[lb:ub] != start) > 0.99 * (ub - lb)


def test_noalloc_rejects_slices_beyond_its_scratch(sb, torch_gpu, ctx):
    torch = torch_gpu
    lib = sb._lib.load()
    n = (1 << 20) + 64
    arrs = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(5)]
    soln = torch.zeros(n, dtype=torch.float64, device="cuda")
    rc = _call(lib, "sbslpm_direct_sum_inv_lap_noalloc", ctx, arrs, soln, 0, n)
    assert rc == sb._lib.SBSLPM_E_ARG
    assert "split" in _err(lib, ctx)
    ctx.synchronize()
    assert not torch.any(soln)
