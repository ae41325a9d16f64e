"""Device order of the FP32 path (dmpm_sort_particles; dmpm_set_particles, dmpm_chain_episode and a stale
forward pass re-sort) against the FP64 oracle (needs an MI355X).

The order is invisible through the C ABI (every call keeps the API order), so these tests check
that results do not depend on it and that the re-sorting happens when it should:
* chained episodes (DiffMPMAnimator3D.cpp:424-429): the second episode of an FP32 context, with and
  without the re-sort at the chain, equals the oracle's episode started from the chained state the
  context returns (loss 1e-6, control gradient 1e-4, final positions 1e-6, F 2e-5 relative);
* a stale order (positions shuffled across device slots through dmpm_set_layer_state): the P2G
  workgroups leave the fixed window, results still equal the oracle's, and the next forward pass
  re-sorts (window statistics back on the fixed window, one more sort);
* an explicit re-sort of a sorted order changes no result beyond FP32 atomics' round-off.
"""
import numpy as np
import pytest

from _scenes import f32_positions, gpu_morph, make_case, maxrel, orc, pkg, rel

pytestmark = pytest.mark.gpu


def _episode(mo):
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    return loss, mo.control_grad(0)


def _oracle_from(case, x, v, F, C, controls):
    st = orc.State(x, v, C, F)
    return orc.episode(case.prm, st, case.cs, controls, case.target, case.L)


@pytest.mark.parametrize("resort", [1, 0])
def test_fp32_chained_episode_vs_oracle(resort):
    case = make_case(seed=3, L=11, vel=0.5)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT)
    try:
        mo.set_option(pkg.OPT_RESORT, resort)
        _episode(mo)
        sorts0 = mo.window_stats()["sorts"]
        mo.chain_episode()
        assert mo.window_stats()["sorts"] == sorts0 + (1 if resort == 1 else 0)
        x0, v0, F0, C0 = mo.layer_state(0)
        assert np.abs(mo.control(0)).max() == 0.0  # chaining clears the controls
        dFc = pkg.scenes.random_controls(case.sc.n, 1e-2, 77)
        mo.set_control(0, dFc)
        loss, g = _episode(mo)
        xl, _, Fl, _ = mo.layer_state(case.L - 1)
        controls = [dFc] + [np.zeros_like(dFc)] * (case.L - 2)
        ref_loss, layers, gd, _ = _oracle_from(case, x0, v0, F0, C0, controls)
        print(f"chained FP32 (resort {resort}): loss {loss:.10e} oracle {ref_loss:.10e}, grad rel {rel(g, gd[0]):.2e}")
        assert loss == pytest.approx(ref_loss, rel=1e-6)
        assert rel(g, gd[0]) < 1e-4
        assert maxrel(xl, f32_positions(layers[-1].x, case.sc.dim, case.sc.dx, case.sc.min_point)) < 1e-6
        assert maxrel(Fl, layers[-1].F) < 2e-5
    finally:
        mo.close()


def test_stale_order_is_detected_and_resorted():
    case = make_case(seed=4, L=6)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT)
    try:
        _episode(mo)
        st = mo.window_stats()
        assert st["general"] + st["hbm"] == 0 and st["fixed"] > 0  # a fresh order: every range on the fixed window
        sorts = st["sorts"]
        # the same particles, positions dealt to other device slots: every range now spans the sphere
        rng = np.random.default_rng(5)
        perm = rng.permutation(case.sc.n)
        x, v, F, C = (a[perm] for a in (case.st0.x, case.st0.v, case.st0.F, case.st0.C))
        mo.set_layer_state(0, x, v, F, C)
        controls = [c[perm] for c in case.controls]
        for i, c in enumerate(controls):
            mo.set_control(i, c)
        # the constants (m, vol, lam, mu) stay with their API index: the oracle gets the shuffled
        # states with the original constants
        ref_loss, _, gd, _ = orc.episode(case.prm, orc.State(x, v, C, F), case.cs, controls, case.target, case.L)
        loss, g = _episode(mo)
        st = mo.window_stats()
        print(f"stale order: {st}")
        assert st["general"] + st["hbm"] > 0.5 * (st["fixed"] + st["general"] + st["hbm"])
        assert st["sorts"] == sorts
        assert loss == pytest.approx(ref_loss, rel=1e-6)
        assert rel(g, gd[0]) < 1e-4
        loss2, g2 = _episode(mo)  # stale: this forward pass re-sorts first
        st2 = mo.window_stats()
        assert st2["sorts"] == sorts + 1 and st2["general"] + st2["hbm"] == 0
        assert loss2 == pytest.approx(ref_loss, rel=1e-6)
        assert rel(g2, gd[0]) < 1e-4
    finally:
        mo.close()


def test_chain_resorts_a_stale_order_only():
    """DMPM_OPT_RESORT 2 (the default): a chain after an episode on the fixed window keeps the
    order; a chain after an episode whose order was stale re-sorts."""
    case = make_case(seed=7, L=6)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT)
    try:
        assert mo.option(pkg.OPT_RESORT) == 2
        _episode(mo)
        s0 = mo.window_stats()["sorts"]
        mo.chain_episode()
        assert mo.window_stats()["sorts"] == s0
        perm = np.random.default_rng(8).permutation(case.sc.n)
        mo.set_layer_state(0, *(a[perm] for a in mo.layer_state(0)))
        _episode(mo)
        assert mo.window_stats()["general"] + mo.window_stats()["hbm"] > 0
        mo.chain_episode()
        assert mo.window_stats()["sorts"] == s0 + 1
    finally:
        mo.close()


def test_explicit_resort_changes_nothing():
    case = make_case(seed=6, L=6, vel=0.2)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_STORE_ALL)
    try:
        l1, g1 = _episode(mo)
        s1 = mo.layer_state(case.L - 1)
        sorts = mo.window_stats()["sorts"]
        mo.sort_particles()
        assert mo.window_stats()["sorts"] == sorts + 1
        l2, g2 = _episode(mo)
        s2 = mo.layer_state(case.L - 1)
        assert l2 == pytest.approx(l1, rel=1e-6)
        assert rel(g2, g1) < 1e-5
        for a, b in zip(s1, s2):
            assert maxrel(a, b) < 1e-5
        assert np.array_equal(mo.control(0), np.float32(case.controls[0]).astype(np.float64))
    finally:
        mo.close()
