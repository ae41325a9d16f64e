"""The fused forward step of the FP32 path (DMPM_OPT_FUSE, on by default: the G2P of a step
and the P2G of the next in one launch, k_step_blk) against separate launches and the FP64
oracle (needs an MI355X).

Both forms compute the same window sums (fixed-point LDS accumulation, independent of the
order of the additions); only the order of the HBM float atomics that flush the windows
differs, so the bars are those of two runs of one form:
* store-all tape, 10 steps: every layer's state and every step's grid 1e-5 (max-relative),
  loss 1e-6, every layer's control gradient 1e-5 (relative L2);
* checkpoint tape (recompute on the second stream): loss 1e-6, control gradients 1e-5;
* a stale order (the windows of both halves on the general and HBM paths): the fused form
  against the oracle at the FP32 bars (loss 1e-6, control gradient 1e-4);
* the kernel timings count L - 2 fused launches per forward pass, one P2G and one G2P;
* a batch of 4 instances: fused against separate launches, loss 1e-6, gradient 1e-5;
* the Fc layers of the checkpoint tape (no smoothing): against store-all at the bars above.
"""
import numpy as np
import pytest

from _scenes import gpu_morph, make_case, maxrel, orc, pkg, rel

pytestmark = pytest.mark.gpu


def _run(mo, L):
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    return loss, [mo.control_grad(i) for i in range(L - 1)]


def test_fuse_option():
    case = make_case(seed=1, L=4)
    mo = gpu_morph(case, precision=pkg.PREC_F32)
    try:
        assert mo.option(pkg.OPT_FUSE) == 1
        mo.set_option(pkg.OPT_FUSE, 0)
        assert mo.option(pkg.OPT_FUSE) == 0
        with pytest.raises(pkg.DmpmError):
            mo.set_option(pkg.OPT_FUSE, 2)
    finally:
        mo.close()


def test_fused_equals_separate_store_all():
    case = make_case(seed=11, L=11, vel=0.3, per_layer_ctrl=True)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_STORE_ALL)
    try:
        out = {}
        for fuse in (0, 1):
            mo.set_option(pkg.OPT_FUSE, fuse)
            loss, g = _run(mo, case.L)
            out[fuse] = (loss, g, [mo.layer_state(j) for j in range(case.L)], [mo.grid(i) for i in range(case.L - 1)],
                         mo.window_stats())
        (l0, g0, s0, gr0, w0), (l1, g1, s1, gr1, w1) = out[0], out[1]
        print(f"store-all fused vs separate: loss {abs(l1 - l0) / abs(l0):.2e}, grad {max(rel(a, b) for a, b in zip(g1, g0)):.2e}")
        assert l1 == pytest.approx(l0, rel=1e-6)
        for a, b in zip(g1, g0):
            assert rel(a, b) < 1e-5
        for j in range(case.L):
            for a, b in zip(s1[j], s0[j]):
                assert maxrel(a, b) < 1e-5, j
        for i in range(case.L - 1):
            for a, b in zip(gr1[i], gr0[i]):
                assert maxrel(a, b) < 1e-5, i
        assert {k: w1[k] for k in ("fixed", "general", "hbm")} == {k: w0[k] for k in ("fixed", "general", "hbm")}
    finally:
        mo.close()


def test_fused_equals_separate_checkpoint():
    case = make_case(seed=12, L=17, vel=0.3, per_layer_ctrl=True)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT)
    try:
        res = {}
        for fuse in (0, 1):
            mo.set_option(pkg.OPT_FUSE, fuse)
            res[fuse] = _run(mo, case.L)
        (l0, g0), (l1, g1) = res[0], res[1]
        assert l1 == pytest.approx(l0, rel=1e-6)
        for a, b in zip(g1, g0):
            assert rel(a, b) < 1e-5
    finally:
        mo.close()


def test_fused_stale_order_vs_oracle():
    case = make_case(seed=13, L=6)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT)
    try:
        mo.set_option(pkg.OPT_RESORT, 0)  # keep the stale order through the forward pass
        perm = np.random.default_rng(14).permutation(case.sc.n)
        x, v, F, C = (a[perm] for a in (case.st0.x, case.st0.v, case.st0.F, case.st0.C))
        mo.set_layer_state(0, x, v, F, C)
        controls = [c[perm] for c in case.controls]
        for i, c in enumerate(controls):
            mo.set_control(i, c)
        ref_loss, _, gd, _ = orc.episode(case.prm, orc.State(x, v, C, F), case.cs, controls, case.target, case.L)
        loss, g = _run(mo, case.L)
        st = mo.window_stats()
        print(f"stale order, fused: {st}")
        assert st["general"] + st["hbm"] > 0
        assert loss == pytest.approx(ref_loss, rel=1e-6)
        assert rel(g[0], gd[0]) < 1e-4
    finally:
        mo.close()


def test_fused_launch_counts():
    case = make_case(seed=15, L=9)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_STORE_ALL)
    try:
        mo.profile(True)
        mo.profile_reset()
        mo.forward_pass(0)
        t = mo.profile_read()
        assert t["step"][1] == case.L - 2 and t["p2g"][1] == 1 and t["g2p"][1] == 1, t
        mo.set_option(pkg.OPT_FUSE, 0)
        mo.set_control(0, case.controls[0])  # invalidates the pass
        mo.profile_reset()
        mo.forward_pass(0)
        t = mo.profile_read()
        assert t["step"][1] == 0 and t["p2g"][1] == case.L - 1 and t["g2p"][1] == case.L - 1, t
    finally:
        mo.close()


def test_checkpoint_recompute_starts_from_checkpoint_grids():
    """Checkpoint tape, 16 steps in segments of 4 (the forward pass leaves the last two in the tapes):
    the backward pass recomputes segments 0 and 1 from the grids and affine matrices the forward pass
    kept for their checkpoint steps, so it launches no P2G, 3 fused steps per segment (the G2P of a
    checkpoint step with the P2G of the next) and no G2P (each ends on a checkpoint); the gradients
    equal those of a store-all context (1e-5, the bar of two runs)."""
    case = make_case(seed=16, L=17, vel=0.3, per_layer_ctrl=True)
    ref = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_STORE_ALL)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT, ckpt_stride=4)
    try:
        _, gr = _run(ref, case.L)
        mo.set_option(pkg.OPT_OVERLAP, 0)  # every launch on the timed stream
        mo.profile(True)
        mo.profile_reset()
        mo.forward_pass(0)
        t = mo.profile_read()
        assert t["p2g"][1] == 1 and t["step"][1] == case.L - 2 and t["g2p"][1] == 1, t
        mo.end_layer_mass_loss()
        mo.profile_reset()
        mo.backward_pass(0)
        t = mo.profile_read()
        assert t["p2g"][1] == 0 and t["step"][1] == 2 * 3 and t["g2p"][1] == 0, t
        for i in range(case.L - 1):
            assert rel(mo.control_grad(i), gr[i]) < 1e-5, i
    finally:
        mo.close()
        ref.close()


def test_fused_batch_equals_separate():
    L = 6
    scs = pkg.scenes.batched_instances(4, dt=1e-5)
    vols = []
    for sc in scs:
        prm = orc.Params(n=sc.n, dim=sc.dim, dx=sc.dx, min_point=sc.min_point)
        vols.append(orc.point_volumes(prm, sc.x, sc.m) * sc.dx ** 3)
    ctrl = np.concatenate([pkg.scenes.random_controls(sc.n, 1e-2, 40 + i) for i, sc in enumerate(scs)])
    mb = pkg.MorphBatch.from_scenes(scs, num_layers=L, dt=1e-5, drag=0.0, smoothing=0.0, tape=pkg.TAPE_CHECKPOINT)
    try:
        st = [sc.state0() for sc in scs]
        cat = lambda k: np.concatenate([t[k] for t in st])  # noqa: E731
        mb.set_particles(cat(0), cat(1), cat(3), cat(2), np.concatenate([sc.m for sc in scs]), np.concatenate(vols),
                         np.concatenate([sc.lam for sc in scs]), np.concatenate([sc.mu for sc in scs]))
        mb.set_control(0, ctrl)
        mb.target_from_points([sc.target_x for sc in scs], [sc.target_m for sc in scs])
        res = {}
        for fuse in (0, 1):
            mb.set_option(pkg.OPT_FUSE, fuse)
            mb.forward_pass(0)
            total = mb.end_layer_mass_loss()
            mb.backward_pass(0)
            res[fuse] = (total, mb.instance_losses(), mb.control_grad(0))
        assert res[1][0] == pytest.approx(res[0][0], rel=1e-6)
        np.testing.assert_allclose(res[1][1], res[0][1], rtol=1e-6)
        assert rel(res[1][2], res[0][2]) < 1e-5
    finally:
        mb.close()


@pytest.mark.parametrize("smoothing", [0.0, 0.1])
def test_fc_layers_against_store_all(smoothing):
    """FP32 checkpoint tape, 12 steps in segments of 4.  Without smoothing the passes keep Fc = F + dFc in
    the F field of the layers between checkpoints (their readers skip the control); with smoothing they
    keep F.  Either way the loss (1e-6), every control gradient (1e-5), the state of one layer of each
    segment and the point diagnostics (J, psi) of a layer between checkpoints (1e-5) equal those of a
    store-all context.  Switching the fused step off drops the pass when it changes what the tapes
    hold (the backward pass then raises); the pass run again gives the same gradients."""
    case = make_case(seed=17, L=13, vel=0.3, per_layer_ctrl=True, smoothing=smoothing)
    ref = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_STORE_ALL)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT, ckpt_stride=4)
    try:
        lr, gr = _run(ref, case.L)
        lm, gm = _run(mo, case.L)
        assert lm == pytest.approx(lr, rel=1e-6)
        for i in range(case.L - 1):
            assert rel(gm[i], gr[i]) < 1e-5, i
        for j in (10, 5, 2):  # one layer of each segment
            for a, b in zip(mo.layer_state(j), ref.layer_state(j)):
                assert maxrel(a, b) < 1e-5, j
        for a, b in zip(mo.point_diagnostics(9), ref.point_diagnostics(9)):
            assert maxrel(a, b) < 1e-5
        mo.forward_pass(0)
        mo.end_layer_mass_loss()
        mo.set_option(pkg.OPT_FUSE, 0)
        if smoothing == 0.0:
            with pytest.raises(pkg.DmpmError):
                mo.backward_pass(0)
        else:
            mo.backward_pass(0)
            assert rel(mo.control_grad(0), gr[0]) < 1e-5
        mo.set_option(pkg.OPT_FUSE, 1)
        lm2, gm2 = _run(mo, case.L)
        assert lm2 == pytest.approx(lr, rel=1e-6)
        for i in range(case.L - 1):
            assert rel(gm2[i], gr[i]) < 1e-5, i
    finally:
        mo.close()
        ref.close()
