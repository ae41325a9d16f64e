"""dmpm_target_from_points_noalloc (the entry point that allocates nothing) against
dmpm_target_from_points, through the C ABI, on two calls in a row (the second with
fewer points, so stale staging from the first call would show), on the FP64 and FP32
paths.  Tolerances: FP64 1e-12 of the largest node mass; FP32 1e-5: the two entry
points add the same float contributions through HBM atomics in different orders,
about 1600 per node at 200k points on 32^3, whose float sums differ by about
sqrt(1600) ulp (measured 1.5e-6); both also against the FP64 oracle's P2G mass.
Sizes beyond the fixed staging array are rejected before any GPU work."""
import numpy as np
import pytest

from _scenes import maxrel, orc, pkg

pytestmark = pytest.mark.gpu


def _points(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.2, 0.8, size=(n, 3))
    m = rng.uniform(0.5, 1.5, size=n) * (1.0 / 64) ** 3
    return x, m


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-5)])
def test_noalloc_target_matches_allocating_entry_point(precision, tol):
    dim, dx = (32, 32, 32), 1.0 / 32
    mo = pkg.Morph(100, dim, dx, 3, precision=precision)
    try:
        for n, seed in ((200000, 1), (1234, 2)):
            x, m = _points(n, seed)
            print(f"noalloc target: precision {precision} n {n}", flush=True)
            mo.target_from_points(x, m)
            ref = mo.target_grid()
            mo.target_from_points_noalloc(x, m)
            got = mo.target_grid()
            assert maxrel(got, ref) <= tol, (n, maxrel(got, ref))
            want = orc.p2g_mass(orc.Params(n=n, dim=dim, dx=dx), x, m)
            assert maxrel(got, want) <= tol
            assert got.sum() == pytest.approx(m.sum(), rel=1e-12 if precision == 0 else 1e-5)
        # an empty cloud clears the target
        mo.target_from_points_noalloc(np.zeros((0, 3)), np.zeros(0))
        assert not mo.target_grid().any()
    finally:
        mo.close()


def test_noalloc_rejects_too_many_points():
    mo = pkg.Morph(10, (8, 8, 8), 0.125, 2)
    try:
        n = pkg.NOALLOC_MAX_POINTS + 1
        with pytest.raises(pkg.DmpmError, match="at most 1048576 points"):
            mo.target_from_points_noalloc(np.zeros((n, 3)), np.zeros(n))
    finally:
        mo.close()
