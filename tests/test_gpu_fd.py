"""Adjoint vs central finite differences on the 130k-particle, 64^3 scene
(needs an MI355X).

Scene: seeded sphere of ~130k particles, 64^3 grid, 20 steps of dt = 2e-4,
layer-0 control dFc ~ N(0, 1e-2), density 1000 (particle mass pdx^3 rho).
The grid update zeroes nodes with |m| <= FLT_EPSILON (ForwardSimulation.cpp:55);
at density 1 a particle weighs 4.8e-7, so rim nodes of a quarter particle
mass cross that threshold under any perturbation and the loss jumps: central
differences then disagree with every adjoint by ~0.5% at any h.  At density
1000 the threshold sits at 2.5e-4 of a particle mass and the loss is smooth at
the scale of h = 1e-4.

Checked, on 32 seeded random control entries with h = 1e-4: the plain FP64
path's adjoint (|FD - adj| <= 1e-5 max|FD|) and the FP32 + checkpoint adjoint
(1e-4), plus 8 random directions over all entries (1e-5 / 1e-4); the FP64
path's gradient against the FP64 oracle (1e-10 relative).
"""
import numpy as np
import pytest

from _scenes import orc, pkg, rel

pytestmark = pytest.mark.gpu

STEPS = 20
DT = 2e-4
H = 1e-4
RHO = 1000.0


@pytest.fixture(scope="module")
def scene():
    sc = pkg.scenes.morph_sphere(130_000, 64, STEPS, 1, 6, seed=0, dt=DT, name="morph_130k_64")
    sc.m = sc.m * RHO
    sc.target_m = sc.target_m * RHO
    prm = orc.Params(n=sc.n, dim=sc.dim, dx=sc.dx, dt=DT)
    vol = orc.point_volumes(prm, sc.x, sc.m) * sc.dx ** 3
    dFc = pkg.scenes.random_controls(sc.n, 1e-2, 1)
    return sc, prm, vol, dFc


def _morph(sc, vol, dFc, precision, tape):
    mo = pkg.Morph.from_scene(sc, num_layers=STEPS + 1, precision=precision, tape=tape, dt=DT)
    x, v, C, F = sc.state0()
    mo.set_particles(x, v, F, C, sc.m, vol, sc.lam, sc.mu)
    mo.set_control(0, dFc)
    mo.target_from_points(sc.target_x, sc.target_m)
    return mo


def _loss(mo, dFc):
    mo.set_control(0, dFc)
    mo.forward_pass(0)
    return mo.end_layer_mass_loss()


def test_adjoint_vs_fd_130k(scene):
    sc, prm, vol, dFc = scene
    rng = np.random.default_rng(42)
    m64 = _morph(sc, vol, dFc, pkg.PREC_F64, pkg.TAPE_STORE_ALL)
    m64.forward_pass(0)
    m64.end_layer_mass_loss()
    m64.backward_pass(0)
    g64 = m64.control_grad(0)
    m32 = _morph(sc, vol, dFc, pkg.PREC_F32, pkg.TAPE_CHECKPOINT)
    m32.forward_pass(0)
    m32.end_layer_mass_loss()
    m32.backward_pass(0)
    g32 = m32.control_grad(0)
    entries = rng.choice(dFc.size, 32, replace=False)
    dirs = [np.eye(1, dFc.size, k).reshape(dFc.shape) for k in entries] + [rng.normal(size=dFc.shape) for _ in range(8)]
    fd = np.array([(_loss(m64, dFc + H * u) - _loss(m64, dFc - H * u)) / (2 * H) for u in dirs])
    ad64 = np.array([np.sum(g64 * u) for u in dirs])
    ad32 = np.array([np.sum(g32 * u) for u in dirs])
    for name, sl in (("entries", slice(0, 32)), ("directions", slice(32, 40))):
        scale = np.abs(fd[sl]).max()
        e64 = np.abs(ad64[sl] - fd[sl]).max() / scale
        e32 = np.abs(ad32[sl] - fd[sl]).max() / scale
        print(f"n={sc.n} {name}: max|FD| {scale:.3e}, FP64 adjoint vs FD {e64:.2e}, FP32+ckpt adjoint vs FD {e32:.2e}")
        assert e64 < 1e-5
        assert e32 < 1e-4
    print(f"FP32 vs FP64 gradient {rel(g32, g64):.2e}")
    assert rel(g32, g64) < 1e-4


def test_fp64_gradient_vs_oracle_130k(scene):
    sc, prm, vol, dFc = scene
    m64 = _morph(sc, vol, dFc, pkg.PREC_F64, pkg.TAPE_CHECKPOINT)
    m64.forward_pass(0)
    loss = m64.end_layer_mass_loss()
    m64.backward_pass(0)
    x, v, C, F = sc.state0()
    cs = orc.Consts(sc.m, vol, sc.lam, sc.mu)
    target = orc.p2g_mass(prm, sc.target_x, sc.target_m)
    ref_loss, _, gd, _ = orc.episode(prm, orc.State(x, v, C, F), cs, [dFc] + [np.zeros_like(dFc)] * (STEPS - 1), target,
                                     STEPS + 1)
    assert loss == pytest.approx(ref_loss, rel=1e-12)
    assert rel(m64.control_grad(0), gd[0]) < 1e-10
