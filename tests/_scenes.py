"""Shared seeded test scenes: the same inputs go to the oracle and the HIP path."""
from dataclasses import dataclass

import numpy as np

from conftest import load_oracle, load_pkg

orc = load_oracle()
pkg = load_pkg()


@dataclass
class Case:
    sc: object
    prm: object          # oracle params
    cs: object           # oracle constants
    st0: object          # oracle layer-0 state
    controls: list       # per-layer dFc (L-1 arrays)
    target: np.ndarray   # target grid mass
    L: int


def make_case(seed=0, dt=1e-4, ctrl=1e-2, L=11, physical_vol=True, drag=0.0, smoothing=0.0, f_ext=(0.0, 0.0, 0.0),
              vel=0.0, per_layer_ctrl=False, scene=None):
    """The 8.8k-particle 32^3 sphere scene with seeded non-rest controls.

    ``physical_vol`` scales the reference's dimensionless volume
    (ForwardSimulation.cpp:288) by dx^3 so that dt below the CFL limit gives
    a stable trajectory."""
    sc = scene if scene is not None else pkg.scenes.morph_small(seed=seed, dt=dt)
    prm = orc.Params(n=sc.n, dim=sc.dim, dx=sc.dx, min_point=sc.min_point, dt=dt, drag=drag, smoothing=smoothing,
                     f_ext=f_ext)
    vol = orc.point_volumes(prm, sc.x, sc.m)
    if physical_vol:
        vol = vol * sc.dx ** 3
    cs = orc.Consts(sc.m.copy(), vol, sc.lam.copy(), sc.mu.copy())
    x, v, C, F = sc.state0()
    rng = np.random.default_rng(1000 + seed)
    if vel:
        v = vel * rng.normal(size=v.shape)
    st0 = orc.State(x, v, C, F)
    base = pkg.scenes.random_controls(sc.n, ctrl, seed + 1)
    if per_layer_ctrl:
        controls = [base] + [pkg.scenes.random_controls(sc.n, ctrl, seed + 2 + i) for i in range(L - 2)]
    else:
        controls = [base] + [np.zeros_like(base) for _ in range(L - 2)]
    target = orc.p2g_mass(prm, sc.target_x, sc.target_m)
    return Case(sc, prm, cs, st0, controls, target, L)


def gpu_morph(case: Case, precision=0, tape=0, compat=0, ckpt_stride=0, keep_control_grads=True):
    sc = case.sc
    mo = pkg.Morph(sc.n, sc.dim, sc.dx, case.L, min_point=sc.min_point, dt=case.prm.dt, drag=case.prm.drag,
                   smoothing=case.prm.smoothing, f_ext=case.prm.f_ext, precision=precision, compat=compat,
                   tape=tape, ckpt_stride=ckpt_stride, keep_control_grads=keep_control_grads)
    st = case.st0
    mo.set_particles(st.x, st.v, st.F, st.C, case.cs.m, case.cs.vol, case.cs.lam, case.cs.mu)
    for i, c in enumerate(case.controls):
        mo.set_control(i, c)
    mo.set_target_grid(case.target)
    return mo


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def maxrel(a, b):
    """max |a-b| / max |b|"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def f32_positions(x, dim, dx, min_point):
    """Positions as an FP32 context returns them: stored on the device relative
    to the grid centre (min_point + dim dx / 2) in float, the centre added back."""
    c = np.asarray(min_point, dtype=np.float64) + 0.5 * np.asarray(dim, dtype=np.float64) * dx
    return (np.asarray(x, dtype=np.float64) - c).astype(np.float32).astype(np.float64) + c
