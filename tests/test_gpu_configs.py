"""The batched and the largest benchmark workloads at full size (needs an MI355X).

Batched uneven instances: 24 morphs from ``scenes.batched_instances(24)`` (5k-300k particles each,
log-uniform, about 1.2M in all; grids drawn from 32^3, 48^3, 64^3, 96^3; seeded sphere sources and
sphere-union targets), advanced by one FP32 checkpointed context:
* every instance against the same morph in a single FP32 context: loss 1e-6, final positions 1e-6
  (max-relative), control gradient 1e-5 (relative L2; only the order of float HBM atomics differs);
* the three smallest instances against the FP64 oracle: loss 1e-5, F 2e-5, gradient 1e-4.
Episodes of 5 steps at dt = 1e-5 with volumes from the reference formula scaled by dx^3 (the
configuration's dt = 1e-3 with dimensionless volumes exceeds the explicit CFL limit).

Largest single-GPU run: 4M particles, 256^3 grid, E log-uniform in [1e4, 1e5], at density 1000 (at
density 1 every node mass stays below the FLT_EPSILON threshold of the grid update, so nothing
moves and every gradient is exactly zero), dt = 1e-5, a 20-step episode (the configuration's 200
steps run in bench.py; the FP64 store-all tape of 200 steps does not fit in HBM):
* FP32 + checkpoint control gradient against the plain FP64 path's: 1e-4 relative, loss 1e-6;
* FP32 checkpoint + recompute against FP32 storing every layer: 2e-5;
* the first step's FP64 grid mass equals the particles' mass (every stencil inside the grid), 1e-12.
"""
import numpy as np
import pytest

from _scenes import maxrel, orc, pkg, rel

pytestmark = pytest.mark.gpu

DT = 1e-5


# ------------------------------------------------------------- 24 instances --
@pytest.fixture(scope="module")
def c4():
    L = 6
    scs = pkg.scenes.batched_instances(24, dt=DT)
    vols, ctrls = [], []
    for i, sc in enumerate(scs):
        prm = orc.Params(n=sc.n, dim=sc.dim, dx=sc.dx, min_point=sc.min_point)
        vols.append(orc.point_volumes(prm, sc.x, sc.m) * sc.dx ** 3)
        ctrls.append(pkg.scenes.random_controls(sc.n, 1e-2, 100 + i))
    mb = pkg.MorphBatch.from_scenes(scs, num_layers=L, dt=DT, drag=0.0, smoothing=0.0, tape=pkg.TAPE_CHECKPOINT)
    st = [sc.state0() for sc in scs]
    cat = lambda k: np.concatenate([t[k] for t in st])  # noqa: E731
    mb.set_particles(cat(0), cat(1), cat(3), cat(2), np.concatenate([sc.m for sc in scs]), np.concatenate(vols),
                     np.concatenate([sc.lam for sc in scs]), np.concatenate([sc.mu for sc in scs]))
    mb.set_control(0, np.concatenate(ctrls))
    mb.target_from_points([sc.target_x for sc in scs], [sc.target_m for sc in scs])
    mb.forward_pass(0)
    total = mb.end_layer_mass_loss()
    mb.backward_pass(0)
    out = dict(scs=scs, vols=vols, ctrls=ctrls, L=L, total=total, losses=mb.instance_losses(),
               x=mb.split(mb.layer_state(L - 1)[0]), F=mb.split(mb.layer_state(L - 1)[2]), g=mb.split(mb.control_grad(0)))
    print(f"24 instances: {mb.n} particles, grids {sorted({sc.dim[0] for sc in scs})}, stats {mb.window_stats()}", flush=True)
    mb.close()
    return out


def test_batched_24_instances_match_single_contexts(c4):
    assert c4["total"] == pytest.approx(c4["losses"].sum(), rel=1e-14)
    L = c4["L"]
    worst = [0.0, 0.0, 0.0]
    for i, sc in enumerate(c4["scs"]):
        mo = pkg.Morph.from_scene(sc, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT, num_layers=L, dt=DT, drag=0.0,
                                  smoothing=0.0)
        x, v, C, F = sc.state0()
        mo.set_particles(x, v, F, C, sc.m, c4["vols"][i], sc.lam, sc.mu)
        mo.set_control(0, c4["ctrls"][i])
        mo.target_from_points(sc.target_x, sc.target_m)
        mo.forward_pass(0)
        loss = mo.end_layer_mass_loss()
        mo.backward_pass(0)
        e = (abs(c4["losses"][i] - loss) / abs(loss), maxrel(c4["x"][i], mo.layer_state(L - 1)[0]),
             rel(c4["g"][i], mo.control_grad(0)))
        mo.close()
        worst = [max(a, b) for a, b in zip(worst, e)]
        assert e[0] < 1e-6 and e[1] < 1e-6 and e[2] < 1e-5, (sc.name, sc.n, e)
    print(f"batch vs single, worst over 24: loss {worst[0]:.2e} x {worst[1]:.2e} grad {worst[2]:.2e}")


def test_batched_24_smallest_instances_vs_oracle(c4):
    L = c4["L"]
    order = np.argsort([sc.n for sc in c4["scs"]])[:3]
    for i in order:
        sc = c4["scs"][i]
        prm = orc.Params(n=sc.n, dim=sc.dim, dx=sc.dx, min_point=sc.min_point, dt=DT)
        x, v, C, F = sc.state0()
        cs = orc.Consts(sc.m, c4["vols"][i], sc.lam, sc.mu)
        target = orc.p2g_mass(orc.Params(n=len(sc.target_m), dim=sc.dim, dx=sc.dx, min_point=sc.min_point), sc.target_x,
                              sc.target_m)
        controls = [c4["ctrls"][i]] + [np.zeros_like(c4["ctrls"][i])] * (L - 2)
        ref_loss, layers, gd, _ = orc.episode(prm, orc.State(x, v, C, F), cs, controls, target, L)
        el, eF, eg = abs(c4["losses"][i] - ref_loss) / abs(ref_loss), maxrel(c4["F"][i], layers[-1].F), rel(c4["g"][i], gd[0])
        print(f"{sc.name} ({sc.n} particles) vs oracle: loss {el:.2e} F {eF:.2e} grad {eg:.2e}")
        assert el < 1e-5 and eF < 2e-5 and eg < 1e-4


# ------------------------------------------------------------ 4M / 256^3 ----
C5_STEPS = 20
C5_RHO = 1000.0


@pytest.fixture(scope="module")
def c5():
    sc = pkg.scenes.morph_4m(dt=DT)
    sc.m = sc.m * C5_RHO
    sc.target_m = sc.target_m * C5_RHO
    mo = pkg.Morph.from_scene(sc, num_layers=2)
    x, v, C, F = sc.state0()
    mo.set_particles(x, v, F, C, sc.m, None, sc.lam, sc.mu)
    mo.compute_volumes()
    vol = mo.volumes() * sc.dx ** 3
    mo.close()
    dFc = pkg.scenes.random_controls(sc.n, 1e-2, 1)
    runs = {}
    for name, prec, tape in (("f64_all", pkg.PREC_F64, pkg.TAPE_STORE_ALL), ("f32_all", pkg.PREC_F32, pkg.TAPE_STORE_ALL),
                             ("f32_ckpt", pkg.PREC_F32, pkg.TAPE_CHECKPOINT)):
        mo = pkg.Morph.from_scene(sc, precision=prec, tape=tape, num_layers=C5_STEPS + 1, dt=DT)
        mo.set_particles(x, v, F, C, sc.m, vol, sc.lam, sc.mu)
        mo.set_control(0, dFc)
        mo.target_from_points(sc.target_x, sc.target_m)
        mo.forward_pass(0)
        loss = mo.end_layer_mass_loss()
        mo.backward_pass(0)
        runs[name] = dict(loss=loss, grad=mo.control_grad(0))
        if name == "f64_all":
            runs[name]["grid_m0"] = mo.grid(0)[0]
        mo.close()
        print(f"4M/256^3 {name}: loss {loss:.12g} |g| {np.linalg.norm(runs[name]['grad']):.6g}", flush=True)
    return sc, runs


def test_4m_fp32_checkpoint_gradient_vs_fp64(c5):
    _, r = c5
    assert np.linalg.norm(r["f64_all"]["grad"]) > 0  # density 1000: the morph moves
    e = rel(r["f32_ckpt"]["grad"], r["f64_all"]["grad"])
    print(f"4M FP32+ckpt control gradient vs FP64: {e:.3g}")
    assert e <= 1e-4
    assert r["f32_ckpt"]["loss"] == pytest.approx(r["f64_all"]["loss"], rel=1e-6)


def test_4m_checkpoint_equals_store_all(c5):
    _, r = c5
    assert rel(r["f32_ckpt"]["grad"], r["f32_all"]["grad"]) <= 2e-5


def test_4m_first_step_conserves_mass(c5):
    sc, r = c5
    x = sc.state0()[0]
    base = np.floor((x - np.asarray(sc.min_point)) / sc.dx) - 1
    assert base.min() >= 0 and (base + 3).max() < min(sc.dim)
    assert r["f64_all"]["grid_m0"].sum() == pytest.approx(sc.m.sum(), rel=1e-12)
