"""HIP-graph replay of the passes, range splitting and the autotuner (needs an MI355X).

Graph replays must give what eager launches give.  Each pass graph is
captured on first use and replayed afterwards; the tests replay it several
times with different controls and eager work (control uploads, state reads)
in between, and check every replay against the eager result for the same
controls: FP64 1e-12 relative (order of FP64 atomics), FP32 1e-6 (order of
the float atomics of the LDS-window flush).  Range sizes and the autotuner
change only the launch shape: results stay within the FP32-vs-oracle bars.
"""
import numpy as np
import pytest

from _scenes import f32_positions, gpu_morph, make_case, maxrel, orc, pkg, rel

pytestmark = pytest.mark.gpu


def _episode(mo, control_layer=0):
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(control_layer)
    return loss, mo.control_grad(control_layer), mo.layer_state(mo.num_layers - 1)[0]


@pytest.mark.parametrize("precision,tape", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_graph_replays_match_eager(precision, tape):
    tol = 1e-12 if precision == 0 else 1e-6
    case = make_case(L=11, per_layer_ctrl=True, drag=0.1, smoothing=0.05)
    mo = gpu_morph(case, precision=precision, tape=tape, ckpt_stride=3)
    alt = [c * -0.5 for c in case.controls]
    eager = {}
    for name, ctr in (("a", case.controls), ("b", alt)):
        for i, c in enumerate(ctr):
            mo.set_control(i, c)
        eager[name] = _episode(mo)
        eager[name + "3"] = _episode(mo, 3)
    mo.set_graphs(True)
    assert mo.option(pkg.OPT_GRAPHS) == 1
    for rep, name in enumerate(["a", "b", "a", "b"]):
        ctr = case.controls if name == "a" else alt
        for i, c in enumerate(ctr):
            mo.set_control(i, c)
        for suffix, cl in (("", 0), ("3", 3)):
            loss, g, x = _episode(mo, cl)
            le, ge, xe = eager[name + suffix]
            print(f"replay {rep} {name}{suffix}: loss {abs(loss - le) / abs(le):.1e} grad {rel(g, ge):.1e}")
            assert abs(loss - le) <= tol * abs(le)
            assert rel(g, ge) < tol
            assert maxrel(x, xe) < tol
    # a partial forward pass (restart at layer 5) replays its own graph
    mo.set_control(5, case.controls[5] * 2.0)
    for _ in range(2):
        mo.forward_pass(5)
        loss_p = mo.end_layer_mass_loss()
    mo.set_graphs(False)
    mo.forward_pass(0)
    assert abs(mo.end_layer_mass_loss() - loss_p) <= tol * abs(loss_p)


@pytest.mark.parametrize("range_max", [64, 200, 512])
def test_fp32_range_sizes_vs_oracle(range_max):
    case = make_case(L=6, per_layer_ctrl=True)
    mo = gpu_morph(case, precision=pkg.PREC_F32)
    mo.set_option(pkg.OPT_RANGE_MAX, range_max)
    assert mo.option(pkg.OPT_RANGE_MAX) == range_max
    loss, g, x = _episode(mo)
    ref_loss, layers, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L)
    assert maxrel(x, layers[-1].x) < 1e-6
    assert abs(loss - ref_loss) / abs(ref_loss) < 1e-5
    assert rel(g, gd[0]) < 1e-4


def test_autotune_keeps_results():
    case = make_case(L=6, per_layer_ctrl=True)
    mo = gpu_morph(case, precision=pkg.PREC_F32)
    ms = mo.autotune()
    assert 0.0 < ms < 1e4
    assert mo.option(pkg.OPT_RANGE_MAX) in (128, 256, 512)
    assert mo.option(pkg.OPT_LDS_CAP) in (512, 1024, 2048)
    with pytest.raises(pkg.DmpmError):
        mo.layer_state(3)  # the trajectory was invalidated
    x0 = mo.layer_state(0)[0]
    assert np.array_equal(x0, f32_positions(case.st0.x, case.sc.dim, case.sc.dx, case.sc.min_point))
    mo.set_graphs(True)
    loss, g, x = _episode(mo)
    ref_loss, layers, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L)
    assert maxrel(x, layers[-1].x) < 1e-6
    assert abs(loss - ref_loss) / abs(ref_loss) < 1e-5
    assert rel(g, gd[0]) < 1e-4


def test_option_errors():
    case = make_case(L=3)
    mo = gpu_morph(case, precision=pkg.PREC_F64)
    with pytest.raises(pkg.DmpmError):
        mo.autotune()
    with pytest.raises(pkg.DmpmError):
        mo.set_option(pkg.OPT_RANGE_MAX, 1000)
    with pytest.raises(pkg.DmpmError):
        mo.set_option(pkg.OPT_LDS_CAP, 100)
    with pytest.raises(pkg.DmpmError):
        mo.set_option(99, 1)
