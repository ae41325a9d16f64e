"""Where the FP32 path runs the elastic half of the transposed P2G (DMPM_OPT_ELASTIC): 0 in a kernel
of its own after the gather (one thread per particle slot; with persistent launches, per block
range), 1 inside the gather kernel, 2 there with registers for 3 waves per SIMD.
The polar path is the closed form of H(Fc):Pbar written in the rotation's frame.  Particles whose Fc
needs the SVD (inverted or nearly flat elements) take it through an out-of-line call in mode 0's
one-thread-per-particle kernel, and are deferred to a small kernel after the step otherwise.

Bars: FP32 path vs the FP64 oracle as in test_gpu_parity (loss 1e-5, x 1e-6, F 2e-5, control gradient
1e-4 relative L2); the three modes against one another 1e-6 (the same FP64 arithmetic, compiled into
different kernels).
"""
import numpy as np
import pytest

from _scenes import gpu_morph, make_case, maxrel, orc, pkg, rel

pytestmark = pytest.mark.gpu

MODES = (0, 1, 2)


def test_elastic_option():
    case = make_case(L=3)
    mo = gpu_morph(case, precision=pkg.PREC_F32)
    assert mo.option(pkg.OPT_ELASTIC) in MODES
    for m in MODES:
        mo.set_option(pkg.OPT_ELASTIC, m)
        assert mo.option(pkg.OPT_ELASTIC) == m
    for bad in (-1, 3):
        with pytest.raises(pkg.DmpmError):
            mo.set_option(pkg.OPT_ELASTIC, bad)
    mo.close()


def _episode(case, mode, compat=0, tape=0, graphs=False, persistent=False, reps=1):
    mo = gpu_morph(case, precision=pkg.PREC_F32, compat=compat, tape=tape)
    mo.set_option(pkg.OPT_ELASTIC, mode)
    if persistent:
        mo.set_option(pkg.OPT_PERSISTENT, 1)
    mo.set_graphs(graphs)
    out = []
    for _ in range(reps):
        mo.forward_pass(0)
        loss = mo.end_layer_mass_loss()
        mo.backward_pass(0)
        out.append((loss, [mo.control_grad(i) for i in range(case.L - 1)], mo.layer_grads(0)))
    mo.close()
    return out


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("phys,compat", [(dict(), 0),
                                         (dict(drag=0.3, smoothing=0.1, f_ext=(0.3, -1.0, 0.5), vel=0.2), 0),
                                         (dict(drag=0.3, smoothing=0.1, f_ext=(0.3, -1.0, 0.5), vel=0.2),
                                          pkg.COMPAT_REFERENCE)],
                         ids=["rest", "phys", "phys-compat"])
def test_elastic_modes_vs_oracle(mode, phys, compat):
    """Every per-layer control gradient of a 6-layer sphere episode against the FP64 oracle, exact and
    reference-compat adjoints, with drag, smoothing and an external force."""
    case = make_case(L=6, per_layer_ctrl=True, **phys)
    case.prm.compat = compat
    (loss, g, _), = _episode(case, mode, compat=compat)
    ref_loss, _, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L)
    errs = [rel(g[i], gd[i]) for i in range(case.L - 1)]
    print(f"elastic mode {mode} compat {compat}: loss {abs(loss - ref_loss) / abs(ref_loss):.2e} grads {max(errs):.2e}")
    assert abs(loss - ref_loss) / abs(ref_loss) < 1e-5
    assert max(errs) < 1e-4


def _inverted_case(seed=9, L=4):
    """The sphere scene with one particle in 25 inverted (det F < 0) and one in 25 nearly flat
    (det F below 1e-6 |F|^3): both take the SVD path of the elastic adjoint, the deferred kernel."""
    rng = np.random.default_rng(seed)
    case = make_case(L=L, per_layer_ctrl=True, vel=0.1)
    n = case.sc.n
    F = case.st0.F.copy()
    idx = rng.permutation(n)
    inv, flat = idx[: n // 25], idx[n // 25: 2 * (n // 25)]
    for p in inv:
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        F[p] = Q @ np.diag([0.9, 1.0, -0.3]) @ Q.T
    F[flat] = np.diag([1.0, 1.0, 2e-6])
    f32 = lambda a: np.float32(a).astype(np.float64)  # noqa: E731
    case.st0 = orc.State(case.st0.x, case.st0.v, case.st0.C, f32(F))
    return case


@pytest.mark.parametrize("mode", MODES)
def test_elastic_modes_deferred_svd_vs_oracle(mode):
    """Inverted and nearly flat elements in every mode: the SVD path's gradients (inline in mode 0,
    deferred in modes 1 and 2) against the oracle; a graph-captured pass replayed twice gives the
    gradients of the first pass each time (the deferred list and its counters reset after every
    step), within 1e-6 (the float HBM atomics of the window flushes add in any order)."""
    case = _inverted_case()
    runs = _episode(case, mode, graphs=True, reps=3)
    ref_loss, layers, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L)
    for r, (loss, g, _) in enumerate(runs):
        errs = [rel(g[i], gd[i]) for i in range(case.L - 1)]
        print(f"deferred svd mode {mode} rep {r}: loss {abs(loss - ref_loss) / abs(ref_loss):.2e} grads {max(errs):.2e}")
        assert abs(loss - ref_loss) / abs(ref_loss) < 1e-5
        assert max(errs) < 1e-4
        for i in range(case.L - 1):
            assert rel(g[i], runs[0][1][i]) < 1e-6, (r, i)


def test_deferred_list_longer_than_its_workgroups():
    """Every particle inverted: in the fused modes about 8.8k particles on the deferred list of each
    backward step, more than the 16 x 256 threads of the deferred kernel take in one pass (the kernel
    strides over the list), and every particle through the SVD call of the split mode's kernel;
    gradients against the oracle in each mode, and the split mode's persistent launch (per block range,
    with the deferred list)."""
    rng = np.random.default_rng(3)
    case = make_case(L=3, per_layer_ctrl=True, vel=0.1)
    n = case.sc.n
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    F = np.broadcast_to(Q @ np.diag([0.9, 1.0, -0.3]) @ Q.T, (n, 3, 3)).copy()
    f32 = lambda a: np.float32(a).astype(np.float64)  # noqa: E731
    case.st0 = orc.State(case.st0.x, case.st0.v, case.st0.C, f32(F))
    ref_loss, _, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L)
    for mode, persistent in ((0, False), (0, True), (1, False), (2, False)):
        (loss, g, _), = _episode(case, mode, persistent=persistent)
        errs = [rel(g[i], gd[i]) for i in range(case.L - 1)]
        print(f"all inverted, mode {mode} persistent {persistent}: loss {abs(loss - ref_loss) / abs(ref_loss):.2e} "
              f"grads {max(errs):.2e}")
        assert abs(loss - ref_loss) / abs(ref_loss) < 1e-5
        assert max(errs) < 1e-4


def test_elastic_modes_agree():
    """The three modes on one episode (checkpoint tape, persistent launches on and off): gradients
    within 1e-6 of one another."""
    case = _inverted_case(L=6)
    for persistent in (False, True):
        runs = {m: _episode(case, m, tape=pkg.TAPE_CHECKPOINT, persistent=persistent)[0] for m in MODES}
        for m in MODES[1:]:
            assert runs[m][0] == pytest.approx(runs[0][0], rel=1e-7)
            for i in range(case.L - 1):
                assert rel(runs[m][1][i], runs[0][1][i]) < 1e-6, (persistent, m, i)
            gx, gv, gF, gC = runs[m][2]
            hx, hv, hF, hC = runs[0][2]
            assert maxrel(gF, hF) < 1e-6 and maxrel(gx, hx) < 1e-6, (persistent, m)


def test_elastic_modes_batch():
    """A batch of instances (one of them with inverted elements) in every mode against the single
    contexts of the default mode: the split mode's kernel finds each particle's instance from the slot
    prefix, the deferred list of the fused modes carries it."""
    cases = [make_case(seed=s, L=4, per_layer_ctrl=True) for s in (1, 2)] + [_inverted_case(L=4)]
    singles = [_episode(c, 0)[0] for c in cases]
    cat = lambda f: np.concatenate([f(c) for c in cases])  # noqa: E731
    geo = [(c.sc.n, c.sc.dim, c.sc.dx, c.sc.min_point) for c in cases]
    for mode in MODES:
        mb = pkg.MorphBatch(geo, 4, dt=cases[0].prm.dt, drag=0.0, smoothing=0.0, keep_control_grads=True)
        mb.set_option(pkg.OPT_ELASTIC, mode)
        mb.set_particles(cat(lambda c: c.st0.x), cat(lambda c: c.st0.v), cat(lambda c: c.st0.F), cat(lambda c: c.st0.C),
                         cat(lambda c: c.cs.m), cat(lambda c: c.cs.vol), cat(lambda c: c.cs.lam), cat(lambda c: c.cs.mu))
        for i in range(3):
            mb.set_control(i, cat(lambda c: c.controls[i]))
        mb.set_target_grid(cat(lambda c: c.target))
        mb.forward_pass(0)
        mb.end_layer_mass_loss()
        mb.backward_pass(0)
        for i in range(3):
            for b, g in enumerate(mb.split(mb.control_grad(i))):
                assert rel(g, singles[b][1][i]) < 1e-6, (mode, b, i)
        mb.close()
