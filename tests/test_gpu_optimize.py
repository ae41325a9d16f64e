"""Device control updates (dmpm_descend), the device gradient norm and the
optimisation loops (needs an MI355X).

The update rules are pinned by their closed forms for a fixed gradient g
(normalised: h = g / |g|):
* gradient descent: dFc - alpha h;
* Adam and AMSGrad: with bias correction the first step, and every later step
  on the same gradient, moves by alpha h / (|h| + eps) per entry;
* Lion: the first step moves by alpha (1 - beta) sign(h).
Tolerance: float rounding of the reference's float arithmetic, 2e-6 relative
to alpha (FP32 path); 1e-12 for the norm in FP64.
"""
import numpy as np
import pytest

from _scenes import gpu_morph, make_case, orc, pkg, rel

pytestmark = pytest.mark.gpu


def _after_backward(precision):
    case = make_case(L=4, per_layer_ctrl=True)
    mo = gpu_morph(case, precision=precision)
    mo.forward_pass(0)
    mo.end_layer_mass_loss()
    mo.backward_pass(0)
    return case, mo


@pytest.mark.parametrize("precision", [0, 1])
def test_device_grad_norm(precision):
    _, mo = _after_backward(precision)
    g = mo.control_grad(0)
    assert mo.control_grad_norm(0) == pytest.approx(np.linalg.norm(g), rel=1e-12 if precision == 0 else 1e-6)
    assert mo.control_grad_norm(2) == pytest.approx(np.linalg.norm(mo.control_grad(2)), rel=1e-6)


@pytest.mark.parametrize("rule", ["gd", "adam", "amsgrad", "lion"])
@pytest.mark.parametrize("precision", [0, 1])
def test_descend_closed_forms(rule, precision):
    _, mo = _after_backward(precision)
    g = mo.control_grad(1)
    gn = mo.control_grad_norm(1)
    d0 = mo.control(1)
    h = g / gn
    alpha, beta1, eps = 1e-3, 0.9, 1e-5
    code = {"gd": pkg.DESCEND_GD, "adam": pkg.DESCEND_ADAM, "amsgrad": pkg.DESCEND_AMSGRAD, "lion": pkg.DESCEND_LION}[rule]
    if rule == "gd":
        step = alpha * h
    elif rule == "lion":
        step = alpha * (1 - beta1) * np.sign(h)
    else:
        step = alpha * h / (np.abs(h) + eps)
    mo.descend(1, code, alpha, gn, beta1, 0.999, eps, 1)
    d1 = mo.control(1)
    assert np.abs((d0 - d1) - step).max() < 2e-6 * alpha + 4 * np.abs(d0).max() * np.finfo(np.float32).eps
    if rule in ("adam", "amsgrad"):
        # same gradient, second step: the bias-corrected moments are again h and h^2
        mo.descend(1, code, alpha, gn, beta1, 0.999, eps, 2)
        d2 = mo.control(1)
        assert np.abs((d1 - d2) - step).max() < 2e-6 * alpha + 4 * np.abs(d1).max() * np.finfo(np.float32).eps
    # the control changed: later layers must be recomputed before a loss
    with pytest.raises(pkg.DmpmError):
        mo.end_layer_mass_loss()


F32_EPS = float(np.finfo(np.float32).eps)


def _adam_lion_rounds(h, alpha, b1, b2, eps, rounds, fail_trial):
    """PointCloud::Descend_Adam then Descend_Lion (PointCloud.cpp:83-99,134-162) on one shared
    momentum, as each line-search trial of CompGraph.cpp:391-394 applies them, for a constant
    normalised gradient h, in float64 with float32 parameters; optionally one failed trial and
    its undo at -alpha (:406-408) at the end.  Returns the total change of dFc."""
    a, b1, b2, eps = (np.float64(np.float32(v)) for v in (alpha, b1, b2, eps))
    m = np.zeros_like(h)
    v = np.zeros_like(h)
    d = np.zeros_like(h)
    sg = np.sign(h)

    def adam(al, t):
        nonlocal m, v, d
        m = b1 * m + (1 - b1) * h
        v = b2 * v + (1 - b2) * h * h
        d = d - al * (m / np.float64(np.float32(1 - np.float64(b1) ** t))) / (
            np.sqrt(v / np.float64(np.float32(1 - np.float64(b2) ** t))) + eps)

    def lion(al):
        nonlocal m, d
        m = b1 * m + (1 - b1) * sg
        d = d - al * m

    for t in range(1, rounds + 1):
        adam(a, t)
        lion(a)
    if fail_trial:
        adam(a, rounds + 1)
        lion(a)
        adam(-a, rounds + 1)
        lion(-a)
    return d


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-5)])
def test_adam_then_lion_on_shared_momentum(precision, tol):
    """Several rounds of the reference's line-search update (Adam then Lion on the same
    momentum) and a failed trial with its inexact undo, on a fixed gradient:
    * every round against the float64 restatement of PointCloud.cpp (relative to the largest
      entry of the total change: Lion's sign term enters Adam's numerator, so entries with
      |h| << eps step by up to alpha / eps);
    * the first round from zero moments in closed form: Adam moves dFc by alpha h / (|h| + eps)
      and leaves m = (1 - beta) h, which Lion then moves to beta (1 - beta) h + (1 - beta) sign h
      and steps by (Lion first would step by alpha (1 - beta) sign h);
    * the closed form the rounds approach: the shared momentum tends to
      m* = (beta h + sign h) / (1 + beta) after Lion, the bias-corrected second moment is h^2
      exactly, so a round moves dFc by alpha (beta m* + (1 - beta) h) / (|h| + eps) + alpha m*
      (separate moments would give alpha h / (|h| + eps) + alpha sign h)."""
    _, mo = _after_backward(precision)
    g = mo.control_grad(1)
    gn = mo.control_grad_norm(1)
    h = g / np.float64(np.float32(gn))  # dmpm_descend takes the norm as a float
    d0 = mo.control(1)
    # eps of the order of |h| keeps Adam's steps near alpha for the small entries of h
    alpha, b1, b2, eps = 1e-3, 0.5, 0.9, 1e-2
    rounds = 30
    prev = d0
    for t in range(1, rounds + 1):
        mo.descend(1, pkg.DESCEND_ADAM, alpha, gn, b1, b2, eps, t)
        mo.descend(1, pkg.DESCEND_LION, alpha, gn, b1)
        if t == 1:
            first = alpha * h / (np.abs(h) + eps) + alpha * (b1 * (1 - b1) * h + (1 - b1) * np.sign(h))
            assert np.abs((d0 - mo.control(1)) - first).max() < 2e-6 * alpha + 4 * np.abs(d0).max() * F32_EPS
        if t in (1, 2, 5, rounds):
            want = _adam_lion_rounds(h, alpha, b1, b2, eps, t, False)
            got = mo.control(1) - d0
            assert np.abs(got - want).max() < tol * np.abs(want).max() + 4 * np.abs(d0).max() * F32_EPS, t
        if t == rounds - 1:
            prev = mo.control(1)
    last = mo.control(1) - prev
    ms = (b1 * h + np.sign(h)) / (1 + b1)
    limit = alpha * (b1 * ms + (1 - b1) * h) / (np.abs(h) + eps) + alpha * ms
    assert np.abs(-last - limit).max() < 1e-6 * np.abs(limit).max() + 4 * np.abs(prev).max() * F32_EPS
    separate = alpha * h / (np.abs(h) + eps) + alpha * np.sign(h)
    assert np.abs(-last - separate).max() > 1e-1 * alpha
    # a failed trial and its undo at -alpha leave dFc moved (the moments moved twice)
    mo.descend(1, pkg.DESCEND_ADAM, alpha, gn, b1, b2, eps, rounds + 1)
    mo.descend(1, pkg.DESCEND_LION, alpha, gn, b1)
    mo.descend(1, pkg.DESCEND_ADAM, -alpha, gn, b1, b2, eps, rounds + 1)
    mo.descend(1, pkg.DESCEND_LION, -alpha, gn, b1)
    want = _adam_lion_rounds(h, alpha, b1, b2, eps, rounds, True)
    got = mo.control(1) - d0
    assert np.abs(got - want).max() < tol * np.abs(want).max() + 4 * np.abs(d0).max() * F32_EPS
    assert np.abs(got - _adam_lion_rounds(h, alpha, b1, b2, eps, rounds, False)).max() > 1e-3 * alpha


def test_gradient_descent_fp32_tracks_fp64():
    """Normalised GD with line search: the loss decreases monotonically and the
    FP32 curve follows the FP64 one (relative 1e-4 per iteration)."""
    curves = []
    for precision in (0, 1):
        case = make_case(L=6, per_layer_ctrl=False)
        mo = gpu_morph(case, precision=precision, tape=pkg.TAPE_CHECKPOINT)
        curves.append(np.array(pkg.optimize.gradient_descent(mo, 6, alpha=0.05)))
    c64, c32 = curves
    print("fp64", c64, "\nfp32", c32)
    assert np.all(np.diff(c64) < 0) and np.all(np.diff(c32) < 0)
    assert len(c64) == len(c32) and np.max(np.abs(c32 - c64) / c64) < 1e-4


def test_optimize_control_sequence_runs():
    """The reference's control-stride loop on the device.  Its Adam-then-Lion trials
    (per-entry steps of alpha = 0.05) drive some trial episodes to diverge, which
    the line search rejects; an accepted step lowers the loss of its pass, but the
    inexact undo of a rejected trial (CompGraph.cpp:406-408) can leave the end
    loss above the start, so the loop is checked for running through those
    episodes: its first gradient norm is the device norm of the initial control
    gradient, every norm is finite, the end loss is finite, the control moved."""
    case = make_case(L=4, per_layer_ctrl=False)
    mo = gpu_morph(case, precision=pkg.PREC_F32)
    mo.forward_pass(0)
    mo.end_layer_mass_loss()
    mo.backward_pass(0)
    g0 = np.linalg.norm(mo.control_grad(0))
    lines = []
    l0, l1, hist = pkg.optimize.optimize_control_sequence(mo, control_stride=1, max_gd_iters=2, max_line_search_iters=4,
                                                          initial_alpha=0.05, gd_tol=1e-3, temporal_iters=1,
                                                          log=lines.append)
    print("\n".join(lines))
    assert hist[0] == pytest.approx(g0, rel=1e-5)
    assert np.isfinite(l0) and np.isfinite(l1)
    assert len(hist) >= 3 and all(np.isfinite(hist))
    assert rel(mo.control(0), case.controls[0]) > 0  # the control moved


def _host_optimize(case, control_stride, max_gd_iters, max_ls, alpha0, gd_tol, temporal_iters, b1=0.9, b2=0.999,
                   eps=1e-5):
    """CompGraph::OptimizeDefGradControlSequence (CompGraph.cpp:280-444) restated on the host over the FP64
    oracle: every episode, loss and control gradient from oracle.episode, and Descend_Adam then Descend_Lion
    on one shared momentum per control layer (PointCloud.cpp:83-99,134-162) in float64 with the float32
    parameters the device takes, a failed trial undone at -alpha (:406-408).  Returns (initial loss, final
    loss, gradient-norm history, controls, accepted and rejected trials)."""
    L = case.L
    ctl = [c.copy() for c in case.controls]
    mom = [np.zeros_like(c) for c in ctl]
    vec = [np.zeros_like(c) for c in ctl]
    f32 = lambda v: np.float64(np.float32(v))  # noqa: E731
    a1, a2, ae = f32(b1), f32(b2), f32(eps)

    def ep(cf):
        loss, _, gd, _ = orc.episode(case.prm, case.st0, case.cs, ctl, case.target, L, control_layer=cf)
        return loss, gd[cf]

    def adam(cf, h, al, t):
        mom[cf] = a1 * mom[cf] + (1 - a1) * h
        vec[cf] = a2 * vec[cf] + (1 - a2) * h * h
        ctl[cf] = ctl[cf] - f32(al) * (mom[cf] / f32(1 - np.float64(a1) ** t)) / (
            np.sqrt(vec[cf] / f32(1 - np.float64(a2) ** t)) + ae)

    def lion(cf, h, al):
        mom[cf] = a1 * mom[cf] + (1 - a1) * np.sign(h)
        ctl[cf] = ctl[cf] - f32(al) * mom[cf]

    loss0, g0 = ep(0)
    init_norm = np.linalg.norm(g0)
    hist = []
    trials = [0, 0]
    t = 1
    for _ in range(temporal_iters):
        for cf in range(0, L - 1, control_stride):
            alpha = alpha0
            for _gd in range(max_gd_iters):
                gd_loss, g = ep(cf)
                gn = np.linalg.norm(g)
                hist.append(gn)
                if gn < gd_tol * init_norm:
                    break
                h = g / f32(gn)
                found = False
                for _ls in range(max_ls):
                    adam(cf, h, alpha, t)
                    lion(cf, h, alpha)
                    if orc.loss_only(case.prm, case.st0, case.cs, ctl, case.target, L) < gd_loss:
                        found = True
                        trials[0] += 1
                        break
                    trials[1] += 1
                    adam(cf, h, -alpha, t)
                    lion(cf, h, -alpha)
                    alpha /= 2.0
                if not found:
                    break
                t += 1
    return loss0, orc.loss_only(case.prm, case.st0, case.cs, ctl, case.target, L), hist, ctl, trials


def test_optimize_control_sequence_vs_host_restatement():
    """The device loop (FP64 path) against its host restatement over the FP64 oracle, every control layer
    (stride 1), 3 descent iterations each with up to 3 line-search trials at alpha 5e-3 and Adam's eps 1e-2
    (the trials move dFc by up to about alpha / eps per entry, so that episodes stay in the elastic regime):
    both accept and reject trials, with the same decisions (the trial losses differ by ~1e-7 relative, the
    two paths agree to ~1e-12); gradient norms within 1e-9, losses within 1e-10 and every layer's control
    within 1e-8 of the restatement."""
    case = make_case(L=4, per_layer_ctrl=True)
    kw = dict(control_stride=1, max_gd_iters=3, max_ls=3, alpha0=5e-3, gd_tol=1e-3, temporal_iters=1, eps=1e-2)
    l0, l1, hist, ctl, trials = _host_optimize(case, **kw)
    assert trials[0] > 0 and trials[1] > 0, trials  # accepted and rejected trials are both exercised
    mo = gpu_morph(case, precision=pkg.PREC_F64)
    lines = []
    d0, d1, dhist = pkg.optimize.optimize_control_sequence(
        mo, control_stride=1, max_gd_iters=kw["max_gd_iters"], max_line_search_iters=kw["max_ls"],
        initial_alpha=kw["alpha0"], gd_tol=kw["gd_tol"], temporal_iters=1, epsilon=kw["eps"], log=lines.append)
    print("\n".join(lines))
    print(f"host: loss {l0:.12g} -> {l1:.12g}, {len(hist)} gradient norms, trials accepted/rejected {trials}")
    assert d0 == pytest.approx(l0, rel=1e-10)
    assert d1 == pytest.approx(l1, rel=1e-10)
    assert len(dhist) == len(hist)
    assert np.allclose(dhist, hist, rtol=1e-9, atol=0)
    for i in range(case.L - 1):
        assert rel(mo.control(i), ctl[i]) < 1e-8, i
    mo.close()
