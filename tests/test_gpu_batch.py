"""Batched contexts (dmpm_create_batch): uneven, independent morphs advanced by
one set of launches (needs an MI355X).

Each instance of a batch must give what the same morph gives on its own:
* against the FP64 oracle run on that instance alone, with the FP32 bars of
  the single-morph tests: positions 1e-6 (max-relative), F 2e-5, loss 1e-5,
  control gradient 1e-4 (relative L2);
* against a single-instance FP32 context: 1e-6 relative on the gradient (only
  the order of the float HBM atomics of the flush differs);
* an instance with no particles has the closed-form loss
  sum_i 1/2 (ln(1 + 1e-4) - ln(t_i + 1 + 1e-4))^2 + 10 (1e-3)^2 G.
"""
import ctypes

import numpy as np
import pytest

from _scenes import maxrel, orc, pkg, rel

pytestmark = pytest.mark.gpu

DT = 1e-4
L = 6
PHYS = dict(drag=0.1, smoothing=0.05, f_ext=(0.0, -9.8, 0.0))


def _instance(kind, seed):
    """(oracle params, layer-0 state, constants, per-layer controls, target grid)."""
    rng = np.random.default_rng(seed)
    if kind == "sphere32":
        dim, dx, mp = (32, 32, 32), 1.0 / 32, (0.0, 0.0, 0.0)
        x = pkg.scenes.lattice_spheres([0.5, 0.5, 0.5], 0.2, dx / 2)
        tx = pkg.scenes.lattice_box([0.35, 0.4, 0.3], [0.7, 0.7, 0.65], dx / 2)
    elif kind == "odd_grid":
        # node count 17*19*23 = 7429, not a multiple of the 32-node clear
        # tiles: the tiles at this instance's ends are shared with its
        # neighbours in the batch (the clears then look the instance up per node)
        dim, dx, mp = (17, 19, 23), 1.0 / 20, (0.0, 0.0, 0.0)
        x = pkg.scenes.lattice_spheres([0.4, 0.45, 0.55], 0.15, dx / 2)
        tx = pkg.scenes.lattice_spheres([0.45, 0.45, 0.5], 0.13, dx / 2)
    elif kind == "offset_box":
        # non-cubic grid with an offset origin
        dim, dx, mp = (20, 28, 24), 1.0 / 24, (-0.1, 0.05, -0.2)
        x = pkg.scenes.lattice_spheres([0.3, 0.55, 0.25], 0.15, dx / 2)
        tx = pkg.scenes.lattice_spheres([0.35, 0.5, 0.3], 0.13, dx / 2)
    elif kind == "clusters":
        # small lattice clusters spread over a 16^3 grid, one over the grid
        # corner (stencils cut by the boundary) and one outside the grid
        # (clamped into edge blocks: general windows and the HBM fallback).
        # Lattice density keeps dt below the CFL limit, which isolated random
        # particles would not.
        dim, dx, mp = (16, 16, 16), 1.0 / 16, (0.0, 0.0, 0.0)
        lo = [rng.uniform(0.1, 0.75, size=3) for _ in range(4)] + [np.full(3, -0.1), np.array([1.05, 0.4, 0.5])]
        x = np.concatenate([pkg.scenes.lattice_box(a, a + 0.14, dx / 2) for a in lo])
        tx = pkg.scenes.lattice_box([0.3, 0.3, 0.3], [0.6, 0.6, 0.6], dx / 2)
    else:
        raise ValueError(kind)
    n = len(x)
    pdx = dx / 2
    m = np.full(n, pdx ** 3)
    lam, mu = pkg.scenes.lame(np.exp(rng.uniform(np.log(5e3), np.log(5e4), size=n)), 0.3)
    prm = orc.Params(n=n, dim=dim, dx=dx, min_point=mp, dt=DT, **PHYS)
    vol = orc.point_volumes(prm, x, m) * dx ** 3
    cs = orc.Consts(m, vol, lam, mu)
    st = orc.State(x, 0.05 * rng.normal(size=(n, 3)), np.zeros((n, 3, 3)), np.broadcast_to(np.eye(3), (n, 3, 3)).copy())
    controls = [1e-2 * rng.normal(size=(n, 3, 3)) for _ in range(L - 1)]
    target = orc.p2g_mass(orc.Params(n=len(tx), dim=dim, dx=dx, min_point=mp), tx, np.full(len(tx), pdx ** 3))
    return prm, st, cs, controls, target


def _batch(insts, **kw):
    geo = [(p.n, p.dim, p.dx, p.min_point) for p, *_ in insts]
    mb = pkg.MorphBatch(geo, L, dt=DT, **PHYS, **kw)
    cat = lambda f: np.concatenate([f(t) for t in insts])  # noqa: E731
    mb.set_particles(cat(lambda t: t[1].x), cat(lambda t: t[1].v), cat(lambda t: t[1].F), cat(lambda t: t[1].C),
                     cat(lambda t: t[2].m), cat(lambda t: t[2].vol), cat(lambda t: t[2].lam), cat(lambda t: t[2].mu))
    for i in range(L - 1):
        mb.set_control(i, cat(lambda t: t[3][i]))
    mb.set_target_grid(cat(lambda t: t[4]))
    return mb


def _single(t, **kw):
    prm, st, cs, controls, target = t
    mo = pkg.Morph(prm.n, prm.dim, prm.dx, L, min_point=prm.min_point, dt=DT, precision=pkg.PREC_F32, **PHYS, **kw)
    mo.set_particles(st.x, st.v, st.F, st.C, cs.m, cs.vol, cs.lam, cs.mu)
    for i in range(L - 1):
        mo.set_control(i, controls[i])
    mo.set_target_grid(target)
    return mo


KINDS = ["sphere32", "odd_grid", "offset_box", "clusters"]


@pytest.fixture(scope="module")
def insts():
    return [_instance(k, 11 + i) for i, k in enumerate(KINDS)]


@pytest.mark.parametrize("tape", [0, 1])
def test_batch_episode_vs_oracle(insts, tape):
    mb = _batch(insts, tape=tape, keep_control_grads=True)
    mb.forward_pass(0)
    total = mb.end_layer_mass_loss()
    mb.backward_pass(0)
    losses = mb.instance_losses()
    assert total == pytest.approx(losses.sum(), rel=1e-15)
    xs = mb.split(mb.layer_state(L - 1)[0])
    Fs = mb.split(mb.layer_state(L - 1)[2])
    g0 = mb.split(mb.control_grad(0))
    g3 = mb.split(mb.control_grad(3))
    gx = mb.split(mb.layer_grads(0)[0])
    for k, (prm, st, cs, controls, target) in zip(KINDS, insts):
        ref_loss, layers, gd, g = orc.episode(prm, st, cs, controls, target, L)
        i = KINDS.index(k)
        ex, eF = maxrel(xs[i], layers[-1].x), maxrel(Fs[i], layers[-1].F)
        el = abs(losses[i] - ref_loss) / abs(ref_loss)
        eg0, eg3, egx = rel(g0[i], gd[0]), rel(g3[i], gd[3]), rel(gx[i], g.x)
        print(f"{k} tape {tape}: x {ex:.2e} F {eF:.2e} loss {el:.2e} grad dFc0 {eg0:.2e} dFc3 {eg3:.2e} x {egx:.2e}")
        assert ex < 1e-6 and eF < 2e-5
        assert el < 1e-5
        assert eg0 < 1e-4 and eg3 < 1e-4 and egx < 1e-4


def test_batch_matches_single_contexts(insts):
    mb = _batch(insts)
    mb.forward_pass(0)
    mb.end_layer_mass_loss()
    mb.backward_pass(0)
    losses = mb.instance_losses()
    gb = mb.split(mb.control_grad(0))
    xb = mb.split(mb.layer_state(L - 1)[0])
    mg, vg = mb.grid(L - 2)
    mgs, vgs = mb.split_grid(mg), mb.split_grid(vg)
    for i, t in enumerate(insts):
        mo = _single(t)
        mo.forward_pass(0)
        loss = mo.end_layer_mass_loss()
        mo.backward_pass(0)
        m1, v1 = mo.grid(L - 2)
        assert losses[i] == pytest.approx(loss, rel=1e-6)
        assert maxrel(xb[i], mo.layer_state(L - 1)[0]) < 1e-6
        assert maxrel(mgs[i], m1) < 1e-5 and maxrel(vgs[i], v1) < 1e-5
        assert rel(gb[i], mo.control_grad(0)) < 1e-6


def test_batch_point_diagnostics(insts):
    """det(F + dFc) and the elastic energy per particle of a batch layer, in API
    order across the instances, against the oracle on the layer's F and control
    as the batch returns them (FP64 from the stored floats): 1e-12 relative."""
    mb = _batch(insts)
    mb.forward_pass(0)
    cat = lambda f: np.concatenate([f(t) for t in insts])  # noqa: E731
    lam = np.float32(cat(lambda t: t[2].lam)).astype(float)
    mu = np.float32(cat(lambda t: t[2].mu)).astype(float)
    for layer in (2, L - 1):
        F = mb.layer_state(layer)[2]
        if layer < L - 1:
            F = F + mb.control(layer)
        J, psi = mb.point_diagnostics(layer)
        assert maxrel(J, np.linalg.det(F)) < 1e-12
        ref = np.array([orc.energy(F[k], lam[k], mu[k]) for k in range(len(F))])
        assert maxrel(psi, ref) < 1e-12


def test_batch_volumes_targets_and_empty_instance(insts):
    """GPU volumes and target P2G per instance; an instance with no particles."""
    geo = [(p.n, p.dim, p.dx, p.min_point) for p, *_ in insts] + [(0, (8, 12, 10), 0.1, (0.2, -0.3, 0.0))]
    mb = pkg.MorphBatch(geo, 2, dt=DT, **PHYS)
    cat = lambda f: np.concatenate([f(t) for t in insts])  # noqa: E731
    mb.set_particles(cat(lambda t: t[1].x), cat(lambda t: t[1].v), cat(lambda t: t[1].F), cat(lambda t: t[1].C),
                     cat(lambda t: t[2].m), None, cat(lambda t: t[2].lam), cat(lambda t: t[2].mu))
    mb.compute_volumes()
    vols = mb.split(mb.volumes())
    rng = np.random.default_rng(4)
    txs, tms = [], []
    for (prm, st, cs, *_), v in zip(insts, vols):
        ref = orc.point_volumes(prm, st.x, cs.m)
        assert maxrel(v, ref) < 1e-5
        txs.append(st.x + 0.01 * rng.normal(size=st.x.shape))
        tms.append(cs.m * 2.0)
    # empty instance: a target from 500 points of its own grid
    txs.append(rng.uniform([0.25, -0.2, 0.1], [0.9, 0.8, 0.9], size=(500, 3)))
    tms.append(np.full(500, 1e-3))
    mb.target_from_points(txs, tms)
    tgs = mb.split_grid(mb.target_grid())
    for i, g in enumerate(geo):
        prm = orc.Params(n=len(tms[i]), dim=g[1], dx=g[2], min_point=g[3])
        assert maxrel(tgs[i], orc.p2g_mass(prm, txs[i], tms[i])) < 1e-5
    mb.forward_pass(0)
    mb.end_layer_mass_loss()
    losses = mb.instance_losses()
    t = tgs[-1]
    closed = np.sum(0.5 * (np.log(1 + 1e-4) - np.log(t + 1 + 1e-4)) ** 2) + 10 * 1e-6 * len(t)
    assert losses[-1] == pytest.approx(closed, rel=1e-6)


def test_batch_errors():
    L_ = pkg.load_library()
    cfg = pkg._make_config(L_, 3, DT, 0.0, 0.0, (0, 0, 0), pkg.PREC_F64, 0, 0, 0, False, 0)
    arr = (pkg.Instance * 1)()
    arr[0].n_particles = 10
    for a in range(3):
        arr[0].dim[a] = 8
    arr[0].dx = 0.1
    h = ctypes.c_void_p()
    assert L_.dmpm_create_batch(ctypes.byref(cfg), 1, arr, ctypes.byref(h)) == -1
    assert b"FP32" in L_.dmpm_last_error()
    mb = pkg.MorphBatch([(10, (8, 8, 8), 0.1, (0, 0, 0)), (5, (4, 4, 4), 0.2, (0, 0, 0))], 3, dt=DT)
    with pytest.raises(pkg.DmpmError):
        pkg.Morph.target_from_points(mb, np.zeros((3, 3)), np.ones(3))
    with pytest.raises(pkg.DmpmError):
        mb.instance_losses()
