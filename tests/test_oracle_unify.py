"""unify_boundary_vals (initialize_cubed_sphere.cpp:328-370) in the oracle: every
copy of a unique node takes the mean of its copies.  Pinned by the mean over
copies computed with numpy's bincount (brute force), by invariance of a field
that is a function of position, by idempotence and by equal copies."""
import numpy as np
import pytest


@pytest.mark.parametrize("levels,m", [(1, 4), (2, 8), (3, 5)])
def test_unify_equals_mean_over_copies(sb, orc, levels, m):
    ht = sb.HostTree(sb.RunConfig(levels=levels, interp_degree=8, leaf_degree=m - 1))
    mp, cnt, *_ = ht.dedupe()
    x, y, z, _ = ht.spec_points()
    rng = np.random.default_rng(levels * 10 + m)
    vals = rng.standard_normal(mp.shape)
    out = orc.unify_boundary_vals(mp, cnt, vals)
    flat = mp.ravel()
    sums = np.bincount(flat, weights=vals.ravel(), minlength=cnt.size)
    np.testing.assert_array_equal(np.bincount(flat, minlength=cnt.size), cnt)
    np.testing.assert_allclose(out.ravel(), (sums / cnt)[flat], rtol=1e-14, atol=1e-15)
    # every copy of a point holds the same value, and a second pass changes nothing
    for k in np.flatnonzero(cnt > 1)[:50]:
        assert np.ptp(out.ravel()[flat == k]) == 0.0
    np.testing.assert_allclose(orc.unify_boundary_vals(mp, cnt, out), out, rtol=1e-15, atol=0)
    # a field that is a function of position is left as it is (copies coincide within 1e-15)
    f = x + 2 * y - z ** 2
    np.testing.assert_allclose(orc.unify_boundary_vals(mp, cnt, f), f, rtol=0, atol=1e-14)
    # interior nodes (multiplicity 1) are untouched bit for bit
    single = cnt[flat] == 1
    np.testing.assert_array_equal(out.ravel()[single], vals.ravel()[single])


def test_unify_rejects_bad_map(orc):
    with pytest.raises(ValueError):
        orc.unify_boundary_vals(np.array([0, 2]), np.array([1, 1]), np.array([1.0, 2.0]))
