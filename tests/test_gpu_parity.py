"""HIP path vs the CPU FP64 oracle on the same seeded inputs (needs an MI355X).

Tolerances (stated per test):
* plain FP64 path vs oracle: 1e-12 relative on forward state and grids, 1e-10 on
  gradients (only the order of FP64 atomic additions differs);
* FP32 path vs FP64 oracle: 1e-6 on positions (a few float ulps), 2e-5 on F,
  1e-4 relative (L2 over all entries) on the control gradient.
"""
import numpy as np
import pytest

from _scenes import f32_positions, gpu_morph, make_case, maxrel, orc, pkg, rel

pytestmark = pytest.mark.gpu


def _rand_F(n, seed, scale):
    rng = np.random.default_rng(seed)
    return np.eye(3) + scale * rng.normal(size=(n, 3, 3))


# --------------------------------------------------------- constitutive ----
def test_svd_conventions_gpu():
    F = np.concatenate([_rand_F(500, 1, 0.5), np.diag([2.0, -0.5, 1.0])[None], np.eye(3)[None],
                        np.diag([3.0, 3.0, 1e-9])[None]])
    U, s, V = pkg.eval_svd(F)
    for k in range(len(F)):
        assert np.allclose(U[k] @ np.diag(s[k]) @ V[k].T, F[k], atol=1e-12)
        assert np.linalg.det(U[k]) == pytest.approx(1.0, abs=1e-12)
        assert np.linalg.det(V[k]) == pytest.approx(1.0, abs=1e-12)
        ref = np.linalg.svd(F[k], compute_uv=False)
        assert np.allclose(np.abs(s[k]), ref, rtol=1e-11, atol=1e-12)
        assert s[k][0] >= 0 and s[k][1] >= 0


def test_pk1_gpu_vs_oracle():
    F = _rand_F(2000, 2, 0.4)
    rng = np.random.default_rng(3)
    lam = rng.uniform(1e3, 1e5, 2000)
    mu = rng.uniform(1e3, 1e5, 2000)
    P = pkg.eval_pk1(F, lam, mu)
    ref = orc.pk1_batch(F, lam, mu)
    assert maxrel(P, ref) < 1e-12


@pytest.mark.parametrize("exact", [True, False])
def test_hessian_gpu_vs_oracle(exact):
    F = _rand_F(2000, 4, 0.4)
    rng = np.random.default_rng(5)
    lam = rng.uniform(1e3, 1e5, 2000)
    mu = rng.uniform(1e3, 1e5, 2000)
    dF = rng.normal(size=(2000, 3, 3))
    H = pkg.eval_hessian(F, lam, mu, dF, exact)
    ref = orc.hess_mult_batch(F, lam, mu, dF, exact)
    assert maxrel(H, ref) < 1e-10


def _energies(F, lam, mu):
    return np.array([orc.energy(F[k], lam[k], mu[k]) for k in range(len(F))])


def test_energy_gpu_vs_oracle():
    """FixedCorotatedElasticity: random F (polar path), det F < 0 and singular F
    (SVD path), near-rest F; 1e-12 relative to the batch's largest energy, 1e-9
    for the near-rest batch (energies ~1e-8 mu)."""
    rng = np.random.default_rng(6)
    F = np.concatenate([_rand_F(1000, 6, 0.4), -_rand_F(50, 7, 0.2), np.diag([3.0, 3.0, 0.0])[None], np.eye(3)[None]])
    lam = rng.uniform(1e3, 1e5, len(F))
    mu = rng.uniform(1e3, 1e5, len(F))
    assert maxrel(pkg.eval_energy(F, lam, mu), _energies(F, lam, mu)) < 1e-12
    Fr = _rand_F(200, 8, 1e-4)
    assert maxrel(pkg.eval_energy(Fr, lam[:200], mu[:200]), _energies(Fr, lam[:200], mu[:200])) < 1e-9
    # closed forms: F = s R gives sigma = (s, s, s); diag(2, -0.5, 1) has signed
    # sigma (2, 1, -0.5) and det -1 clamped to 1e-6; the identity is the rest state
    lam0, mu0, s = 58333.0, 38888.9, 0.8
    c, sn = np.cos(np.pi / 4), np.sin(np.pi / 4)
    R = np.array([[1, 0, 0], [0, c, -sn], [0, sn, c]])
    got = pkg.eval_energy(np.stack([s * R, np.diag([2.0, -0.5, 1.0]), np.eye(3)]), lam0, mu0)
    assert got[0] == pytest.approx(3 * mu0 * (s - 1) ** 2 + 0.5 * lam0 * (s ** 3 - 1) ** 2, rel=1e-13)
    assert got[1] == pytest.approx(3.25 * mu0 + 0.5 * lam0 * (1e-6 - 1) ** 2, rel=1e-13)
    assert got[2] == 0.0


@pytest.mark.parametrize("prec,tape,tol", [(0, 0, (1e-12, 1e-12)), (1, 0, (1e-12, 1e-12)), (1, 1, (1e-5, 1e-4))],
                         ids=["f64_store_all", "f32_store_all", "f32_checkpoint"])
def test_point_diagnostics_gpu(prec, tape, tol):
    """det(F + dFc) and the energy at F + dFc per particle of a layer, against the
    oracle on the layer's F and control as the context returns them (both in FP64
    from the stored values): 1e-12 relative on a stored tape.  The checkpoint tape
    recomputes layers 3 and 5 for each call, and an FP32 recompute is reproducible
    only to float rounding (workgroups add their node sums to the grid with float
    atomics in any order): 1e-5 on J, 1e-4 on the energy (|F - I| ~ 1e-2).  The
    last layer has no control."""
    case = make_case(L=6, per_layer_ctrl=True)
    mo = gpu_morph(case, precision=prec, tape=tape)
    mo.forward_pass(0)
    lam, mu = case.cs.lam, case.cs.mu
    if prec == 1:  # the FP32 path stores float Lame parameters
        lam, mu = np.float32(lam).astype(float), np.float32(mu).astype(float)
    for layer in (0, 3, case.L - 1):
        _, _, F, _ = mo.layer_state(layer)
        if layer < case.L - 1:
            F = F + mo.control(layer)
        J, psi = mo.point_diagnostics(layer)
        assert maxrel(J, np.linalg.det(F)) < tol[0]
        assert maxrel(psi, _energies(F, lam, mu)) < tol[1]
    mo.close()


# ----------------------------------------------------------- FP64 step -----
def test_volumes_and_target_gpu():
    case = make_case(physical_vol=False)
    sc = case.sc
    mo = pkg.Morph(sc.n, sc.dim, sc.dx, 2, precision=pkg.PREC_F64)
    x, v, C, F = sc.state0()
    mo.set_particles(x, v, F, C, sc.m, None, sc.lam, sc.mu)
    mo.compute_volumes()
    assert maxrel(mo.volumes(), case.cs.vol) < 1e-12
    mo.target_from_points(sc.target_x, sc.target_m)
    assert maxrel(mo.target_grid(), case.target) < 1e-12


@pytest.mark.parametrize("phys", [dict(), dict(drag=0.3, smoothing=0.1, f_ext=(0.3, -1.0, 0.5), vel=0.2)])
def test_forward_step_fp64(phys):
    case = make_case(L=3, per_layer_ctrl=True, **phys)
    mo = gpu_morph(case)
    mo.forward_pass(0)
    layers, grids = orc.forward_pass(case.prm, case.st0, case.cs, case.controls, case.L)
    for j in (1, 2):
        x, v, F, C = mo.layer_state(j)
        assert maxrel(x, layers[j].x) < 1e-14
        assert maxrel(v, layers[j].v) < 1e-12
        assert maxrel(F, layers[j].F) < 1e-12
        assert maxrel(C, layers[j].C) < 1e-12
        gm, gv = mo.grid(j - 1)
        assert maxrel(gm, grids[j - 1][0]) < 1e-12
        assert maxrel(gv, grids[j - 1][1]) < 1e-12


def test_reference_config_rest_and_first_step():
    """As written (reference volume formula, dt = 1e-3): the rest state is static, one step matches."""
    case = make_case(dt=1e-3, physical_vol=False, L=2, ctrl=0.0)
    mo = gpu_morph(case)
    mo.forward_pass(0)
    x, v, F, C = mo.layer_state(1)
    assert np.array_equal(x, case.st0.x) and np.abs(v).max() == 0.0 and np.abs(C).max() == 0.0
    assert np.abs(F - np.eye(3)).max() == 0.0
    case = make_case(dt=1e-3, physical_vol=False, L=2, ctrl=1e-2)
    mo = gpu_morph(case)
    mo.forward_pass(0)
    layers, _ = orc.forward_pass(case.prm, case.st0, case.cs, case.controls, case.L)
    x, v, F, C = mo.layer_state(1)
    assert maxrel(v, layers[1].v) < 1e-12 and maxrel(F, layers[1].F) < 1e-12


def test_loss_fp64():
    case = make_case(L=3)
    mo = gpu_morph(case)
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    layers, _ = orc.forward_pass(case.prm, case.st0, case.cs, case.controls, case.L)
    ref, _, _, dLdx = orc.end_mass_loss(case.prm, layers[-1].x, case.cs.m, case.target)
    assert loss == pytest.approx(ref, rel=1e-12)
    gx, gv, gF, gC = mo.layer_grads(case.L - 1)
    assert maxrel(gx, dLdx) < 1e-10
    assert np.abs(gv).max() == 0 and np.abs(gF).max() == 0 and np.abs(gC).max() == 0


@pytest.mark.parametrize("phys", [dict(), dict(drag=0.3, smoothing=0.1, f_ext=(0.3, -1.0, 0.5), vel=0.2)])
@pytest.mark.parametrize("compat", [0, pkg.COMPAT_REFERENCE])
def test_episode_gradient_fp64(phys, compat):
    case = make_case(L=6, per_layer_ctrl=True, **phys)
    case.prm.compat = compat
    mo = gpu_morph(case, compat=compat)
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    ref_loss, layers, gd, g0 = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L)
    assert loss == pytest.approx(ref_loss, rel=1e-12)
    for i in range(case.L - 1):
        assert rel(mo.control_grad(i), gd[i]) < 1e-10, i
    gx, gv, gF, gC = mo.layer_grads(0)
    assert rel(gx, g0.x) < 1e-10 and rel(gv, g0.v) < 1e-10
    assert rel(gF, g0.F) < 1e-10 and rel(gC, g0.C) < 1e-10


def test_single_step_api_matches_pass():
    case = make_case(L=4, per_layer_ctrl=True)
    a = gpu_morph(case)
    a.forward_pass(0)
    la = a.end_layer_mass_loss()
    a.backward_pass(0)
    b = gpu_morph(case)
    for i in range(case.L - 1):
        b.forward_step(i)
    lb = b.end_layer_mass_loss()
    for i in range(case.L - 2, -1, -1):
        b.backward_step(i)
    assert la == pytest.approx(lb, rel=1e-13)
    assert rel(a.control_grad(0), b.control_grad(0)) < 1e-12


# ------------------------------------------------------------ checkpoint ---
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 2e-5)])
@pytest.mark.parametrize("stride", [0, 1, 4, 7])
def test_checkpoint_matches_store_all(prec, tol, stride):
    case = make_case(L=12, per_layer_ctrl=True)
    a = gpu_morph(case, precision=prec)
    a.forward_pass(0)
    la = a.end_layer_mass_loss()
    a.backward_pass(0)
    b = gpu_morph(case, precision=prec, tape=pkg.TAPE_CHECKPOINT, ckpt_stride=stride)
    b.forward_pass(0)
    lb = b.end_layer_mass_loss()
    b.backward_pass(0)
    assert lb == pytest.approx(la, rel=1e-12 if prec == 0 else 1e-6)
    for i in range(case.L - 1):
        assert rel(b.control_grad(i), a.control_grad(i)) < tol, i
    xa = a.layer_state(5)[0]
    xb = b.layer_state(5)[0]
    assert maxrel(xb, xa) < (1e-12 if prec == 0 else 1e-6)


def test_checkpoint_layer_states_after_a_pass():
    """FP32 checkpoint tape, 16 steps in segments of 4: after a forward and a backward pass every layer
    read through the API (checkpoints, taped layers, and layers of recomputed segments, which the
    forward pass held only in rolling slots of x and F) has the full state (x, v, F, C) of a store-all
    context (1e-5 max-relative: two runs' float atomics)."""
    case = make_case(seed=22, L=17, vel=0.3, per_layer_ctrl=True)
    a = gpu_morph(case, precision=pkg.PREC_F32)
    b = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT, ckpt_stride=4)
    try:
        for mo in (a, b):
            mo.forward_pass(0)
            mo.end_layer_mass_loss()
            mo.backward_pass(0)
        for j in range(case.L):
            for f, (x, y) in enumerate(zip(b.layer_state(j), a.layer_state(j))):
                assert maxrel(x, y) < 1e-5, (j, f)
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 2e-5)])
def test_checkpoint_overlap_and_repeated_backward(prec, tol):
    """The checkpoint tape's backward pass with the recompute of the segment below on a second
    stream (default) equals the one-stream pass and the store-all tape; a second backward pass
    to another control layer after one loss reuses the kept loss seed and the tapes the first
    pass left (CompGraph's layers keep the last layer's gradients, CompGraph.cpp:258-267)."""
    case = make_case(L=14, per_layer_ctrl=True)
    a = gpu_morph(case, precision=prec)
    a.forward_pass(0)
    la = a.end_layer_mass_loss()
    a.backward_pass(0)
    for overlap in (1, 0):
        b = gpu_morph(case, precision=prec, tape=pkg.TAPE_CHECKPOINT, ckpt_stride=3)
        b.set_option(pkg.OPT_OVERLAP, overlap)
        b.set_graphs(bool(overlap))
        for rep in range(2):  # a captured pass and its replay
            b.forward_pass(0)
            lb = b.end_layer_mass_loss()
            assert lb == pytest.approx(la, rel=1e-12 if prec == 0 else 1e-6)
            b.backward_pass(0)
            for i in range(case.L - 1):
                assert rel(b.control_grad(i), a.control_grad(i)) < tol, (overlap, rep, i)
        b.backward_pass(7)  # no new loss: the seed of the last layer is kept
        for i in range(7, case.L - 1):
            assert rel(b.control_grad(i), a.control_grad(i)) < tol, (overlap, "again", i)
        b.backward_pass(0)
        assert rel(b.control_grad(0), a.control_grad(0)) < tol
        b.close()
    a.close()


def test_checkpoint_memory_smaller():
    case = make_case(L=41)
    a = gpu_morph(case, keep_control_grads=False)
    b = gpu_morph(case, tape=pkg.TAPE_CHECKPOINT, keep_control_grads=False)
    assert b.memory_bytes() < 0.6 * a.memory_bytes()


def test_checkpoint_control_layer_restart():
    """Optimiser pattern: change one control, forward from it, backward to it (CompGraph.cpp:365-373)."""
    case = make_case(L=9, per_layer_ctrl=True)
    b = gpu_morph(case, tape=pkg.TAPE_CHECKPOINT, ckpt_stride=3)
    b.forward_pass(0)
    b.end_layer_mass_loss()
    b.backward_pass(0)
    new = case.controls[4] + 1e-3
    case.controls[4] = new
    b.set_control(4, new)
    b.forward_pass(4)
    lb = b.end_layer_mass_loss()
    b.backward_pass(4)
    ref_loss, _, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L, control_layer=4)
    assert lb == pytest.approx(ref_loss, rel=1e-12)
    assert rel(b.control_grad(4), gd[4]) < 1e-10


@pytest.mark.parametrize("prec,tape", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_single_control_block_holds_the_control_layer(prec, tape):
    """keep_control_grads off: one block holds the control gradient, written by the backward step of the
    control layer only (the steps above it skip the store). For control layers 0 and 3 it equals the
    oracle's (1e-10 FP64, 1e-4 FP32) and every other layer's is refused."""
    case = make_case(seed=21, L=7, per_layer_ctrl=True)
    mo = gpu_morph(case, precision=prec, tape=tape, ckpt_stride=2 if tape else 0, keep_control_grads=False)
    try:
        mo.forward_pass(0)
        mo.end_layer_mass_loss()
        for control in (3, 0):
            mo.backward_pass(control)
            _, _, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L,
                                      control_layer=control)
            assert rel(mo.control_grad(control), gd[control]) < (1e-10 if prec == 0 else 1e-4), control
            with pytest.raises(pkg.DmpmError):
                mo.control_grad(control + 1)
    finally:
        mo.close()


# ------------------------------------------------------------------ FP32 ---
def test_fp32_episode_vs_fp64_oracle():
    case = make_case(L=11)
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT)
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    ref_loss, layers, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L)
    x, v, F, C = mo.layer_state(case.L - 1)
    ex, eF = maxrel(x, layers[-1].x), maxrel(F, layers[-1].F)
    eg = rel(mo.control_grad(0), gd[0])
    el = abs(loss - ref_loss) / abs(ref_loss)
    print(f"fp32 vs fp64 oracle: x {ex:.2e} F {eF:.2e} loss {el:.2e} control-grad {eg:.2e}")
    assert ex < 1e-6      # a few float ulps of |x| ~ 0.7
    assert eF < 2e-5
    assert el < 1e-5
    assert eg < 1e-4


# ------------------------------------------------------------- edge cases --
def _edge_episode(x, F, prec, seed=7, L=4, dt=1e-4):
    """FP64 oracle and the GPU path on one small edge-case episode: returns
    (loss errors, x, F and control-gradient errors) of the GPU path."""
    rng = np.random.default_rng(seed)
    sc = pkg.scenes.morph_small()
    n = len(x)
    m = np.full(n, sc.m[0])
    lam = np.full(n, sc.lam[0])
    mu = np.full(n, sc.mu[0])
    if prec == pkg.PREC_F32:  # give the oracle the float inputs the FP32 path stores
        f32 = lambda a: np.float32(a).astype(np.float64)  # noqa: E731
        x, F, m, lam, mu = f32(x), f32(F), f32(m), f32(lam), f32(mu)
    prm = orc.Params(n=n, dim=sc.dim, dx=sc.dx, dt=dt)
    vol = orc.point_volumes(prm, x, m) * sc.dx ** 3
    v = 0.1 * rng.normal(size=(n, 3))
    dFc = 1e-2 * rng.normal(size=(n, 3, 3))
    if prec == pkg.PREC_F32:
        vol, v, dFc = f32(vol), f32(v), f32(dFc)
    cs = orc.Consts(m, vol, lam, mu)
    st = orc.State(x, v, np.zeros((n, 3, 3)), F)
    target = orc.p2g_mass(prm, sc.target_x, sc.target_m)
    controls = [dFc] + [np.zeros_like(dFc)] * (L - 2)
    mo = pkg.Morph(n, sc.dim, sc.dx, L, dt=dt, drag=0.0, smoothing=0.0, precision=prec, keep_control_grads=True)
    mo.set_particles(st.x, st.v, st.F, st.C, m, vol, lam, mu)
    mo.set_control(0, dFc)
    mo.set_target_grid(target)
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    ref_loss, layers, gd, _ = orc.episode(prm, st, cs, controls, target, L)
    _, _, Fl, _ = mo.layer_state(L - 1)
    out = (abs(loss - ref_loss) / abs(ref_loss), maxrel(mo.layer_state(L - 1)[0], layers[-1].x), maxrel(Fl, layers[-1].F),
           rel(mo.control_grad(0), gd[0]))
    mo.close()
    return out


def _near_grid_points():
    sc = pkg.scenes.morph_small()
    # a lattice block over the grid corner (stencils cut by the boundary) and
    # a second one entirely outside the grid
    a = pkg.scenes.lattice_box([-0.06, -0.06, -0.06], [0.12, 0.12, 0.12], sc.point_dx)
    b = pkg.scenes.lattice_box([1.02, 0.4, 0.4], [1.1, 0.5, 0.5], sc.point_dx)
    x = np.concatenate([a, b])
    return x, np.broadcast_to(np.eye(3), (len(x), 3, 3)).copy()


def test_particles_near_and_outside_grid_fp64():
    """Stencil nodes outside the grid are skipped (ForwardSimulation.cpp:34-36)."""
    el, ex, eF, eg = _edge_episode(*_near_grid_points(), pkg.PREC_F64)
    assert el < 1e-12 and ex < 1e-12 and eF < 1e-12 and eg < 1e-10


def test_particles_near_and_outside_grid_fp32():
    """The FP32 block kernels on the same scene: windows clamped to the grid,
    blocks whose particles lie outside it.  FP32 tolerances as for the sphere
    episodes (x a few float ulps, F 2e-5, loss 1e-5, gradient 1e-4)."""
    el, ex, eF, eg = _edge_episode(*_near_grid_points(), pkg.PREC_F32)
    print(f"fp32 near/outside grid: loss {el:.2e} x {ex:.2e} F {eF:.2e} grad {eg:.2e}")
    assert el < 1e-5 and ex < 1e-6 and eF < 2e-5 and eg < 1e-4


# loss, x, F and control-gradient bars of the edge-case episodes: FP64 to rounding, FP32 as for the
# sphere episodes
_EDGE_BARS = {pkg.PREC_F64: (1e-12, 1e-12, 1e-12, 1e-10), pkg.PREC_F32: (1e-5, 1e-6, 2e-5, 1e-4)}
_PRECS = pytest.mark.parametrize("prec", [pkg.PREC_F64, pkg.PREC_F32], ids=["fp64", "fp32"])


def _edge_check(name, x, F, prec, dt=1e-4):
    el, ex, eF, eg = _edge_episode(x, F, prec, dt=dt)
    print(f"{name} ({'fp32' if prec == pkg.PREC_F32 else 'fp64'}): loss {el:.2e} x {ex:.2e} F {eF:.2e} grad {eg:.2e}")
    bl, bx, bF, bg = _EDGE_BARS[prec]
    assert el < bl and ex < bx and eF < bF and eg < bg


@_PRECS
def test_one_particle(prec):
    """A single particle: one block range of one particle, one 64-node stencil (dt 1e-5: a lone
    particle's volume is about 9 dx^3, so its own stress is stable over the episode only there)."""
    _edge_check("one particle", np.array([[0.5, 0.51, 0.49]]), _rand_F(1, 1, 0.01), prec, dt=1e-5)


@_PRECS
def test_coincident_particles(prec):
    """1 300 particles at one position beside a lattice ball: a block holding three ranges (512, 512,
    276 particles) whose stencils are all equal, so every LDS and HBM atomic of the scatters meets
    the same 64 nodes."""
    sc = pkg.scenes.morph_small()
    ball = pkg.scenes.lattice_spheres([[0.3, 0.6, 0.5]], [0.08], sc.point_dx)
    x = np.concatenate([np.tile([[0.55, 0.45, 0.52]], (1300, 1)), ball])
    _edge_check("coincident", x, _rand_F(len(x), 2, 0.01), prec)


@_PRECS
def test_particles_on_nodes_and_cell_centres(prec):
    """One particle on each grid node of an 8^3 patch (the stencil's last node then has weight 0) and
    one at each cell centre of a second patch (weights symmetric about the centre); one particle per
    cell, so dt 1e-5 keeps the episode stable."""
    dx = pkg.scenes.morph_small().dx
    g = np.arange(8, dtype=np.float64)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    x = np.concatenate([(lat + 10.0) * dx, (lat + [19.5, 10.5, 10.5]) * dx])
    _edge_check("nodes and centres", x, _rand_F(len(x), 3, 0.01), prec, dt=1e-5)


@pytest.mark.parametrize("n", [511, 512, 513, 1025])
def test_block_range_boundaries_fp32(n):
    """n particles in one 4^3-cell block of the FP32 path: one range below, at and one past the
    512-particle range size, and three ranges with a last range of one particle."""
    rng = np.random.default_rng(n)
    x = 0.5 + 0.125 * rng.uniform(0.0, 1.0, size=(n, 3))
    _edge_check(f"{n} in one block", x, _rand_F(n, 4, 0.01), pkg.PREC_F32)


@pytest.mark.parametrize("prec", [pkg.PREC_F64, pkg.PREC_F32], ids=["fp64", "fp32"])
def test_diverged_particles_touch_no_node(prec):
    """Particles of a diverged simulation (positions 1e12 cells out, +-inf, NaN) add to
    no node and read none: their stencil base is clamped before the integer
    conversion (a saturated base plus the node offset would overflow and could
    pass the grid bounds test as a wrapped index).  The sphere's particles then
    step exactly as without them: loss, positions and control gradient against
    the oracle on the same particle set (which clamps the same way), with the
    path's tolerances."""
    rng = np.random.default_rng(13)
    sc = pkg.scenes.morph_small()
    bad = np.array([[1e12 * sc.dx, 0.5, 0.5], [-1e12 * sc.dx, 0.5, 0.5], [np.inf, 0.5, 0.5], [0.5, -np.inf, 0.5],
                    [0.5, 0.5, np.nan], [np.nan, np.nan, np.nan]])
    x = np.concatenate([sc.x, bad])
    n, nb = len(x), len(bad)
    good = slice(0, n - nb)
    m = np.full(n, sc.m[0])
    lam = np.full(n, sc.lam[0])
    mu = np.full(n, sc.mu[0])
    prm = orc.Params(n=n - nb, dim=sc.dim, dx=sc.dx, dt=1e-4)
    vol = np.concatenate([orc.point_volumes(prm, sc.x, m[good]) * sc.dx ** 3, np.full(nb, (sc.dx / 2) ** 3)])
    v = np.concatenate([0.1 * rng.normal(size=(n - nb, 3)), np.zeros((nb, 3))])
    dFc = 1e-2 * rng.normal(size=(n, 3, 3))
    F = np.broadcast_to(np.eye(3), (n, 3, 3)).copy()
    if prec == pkg.PREC_F32:
        f32 = lambda a: np.float32(a).astype(np.float64)  # noqa: E731
        x, v, dFc, vol, m, lam, mu = f32(x), f32(v), f32(dFc), f32(vol), f32(m), f32(lam), f32(mu)
    L = 4
    prm = orc.Params(n=n, dim=sc.dim, dx=sc.dx, dt=1e-4)
    target = orc.p2g_mass(prm, sc.target_x, sc.target_m)
    st = orc.State(x, v, np.zeros((n, 3, 3)), F)
    cs = orc.Consts(m, vol, lam, mu)
    controls = [dFc] + [np.zeros_like(dFc)] * (L - 2)
    mo = pkg.Morph(n, sc.dim, sc.dx, L, dt=1e-4, drag=0.0, smoothing=0.0, precision=prec, keep_control_grads=True)
    mo.set_particles(x, v, F, st.C, m, vol, lam, mu)
    mo.set_control(0, dFc)
    mo.set_target_grid(target)
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    ref_loss, layers, gd, _ = orc.episode(prm, st, cs, controls, target, L)
    xl = mo.layer_state(L - 1)[0]
    g = mo.control_grad(0)
    tol = (1e-12, 1e-12, 1e-10) if prec == pkg.PREC_F64 else (1e-5, 1e-6, 1e-4)
    el, ex, eg = abs(loss - ref_loss) / abs(ref_loss), maxrel(xl[good], layers[-1].x[good]), rel(g[good], gd[0][good])
    print(f"diverged particles prec {prec}: loss {el:.2e} x {ex:.2e} grad {eg:.2e}")
    assert np.isfinite(loss) and el < tol[0] and ex < tol[1] and eg < tol[2]
    # the diverged particles gather no node: a finite one stays where it is, a non-finite
    # one stays non-finite (FP64 skips its nodes and keeps inf; the FP32 gather's
    # fixed window multiplies its NaN weights by the window's zeros)
    fin = np.isfinite(x[n - nb:]).all(axis=1)
    kept = x[n - nb:][fin]
    if prec == pkg.PREC_F32:  # as the FP32 context stores it (relative to the grid centre)
        kept = f32_positions(kept, sc.dim, sc.dx, sc.min_point)
    assert np.array_equal(xl[n - nb:][fin], kept)
    assert not np.isfinite(xl[n - nb:][~fin]).all(axis=1).any()
    mo.close()


def _inverted_points(seed=9):
    """The sphere scene with one particle in 25 inverted (F = Q diag(0.9, 1, -0.3) Q^T:
    det F < 0, the SVD path, J clamped to 1e-6) and one in 25 nearly flat
    (F = diag(1, 1, 2e-6): det F below 1e-6 |F|^3, the SVD path with J > 1e-6)."""
    rng = np.random.default_rng(seed)
    sc = pkg.scenes.morph_small()
    x = sc.x.copy()
    n = len(x)
    F = np.broadcast_to(np.eye(3), (n, 3, 3)).copy()
    idx = rng.permutation(n)
    inv, flat = idx[: n // 25], idx[n // 25: 2 * (n // 25)]
    for p in inv:
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        F[p] = Q @ np.diag([0.9, 1.0, -0.3]) @ Q.T
    F[flat] = np.diag([1.0, 1.0, 2e-6])
    return x, F


@pytest.mark.parametrize("prec,tol", [(pkg.PREC_F64, (1e-12, 1e-12, 1e-12, 1e-10)),
                                      (pkg.PREC_F32, (1e-5, 1e-6, 2e-5, 1e-4))],
                         ids=["fp64", "fp32"])
def test_inverted_and_flat_elements(prec, tol):
    """Inverted and nearly flat deformation gradients take the SVD path of the
    polar (Elasticity.cpp:88-98 through ImplicitQRSVD's sign convention) in
    P2G and in the elastic adjoint; both paths against the FP64 oracle, with the
    tolerances of the sphere episodes (FP64: 1e-12 state, 1e-10 gradient; FP32:
    loss 1e-5, x 1e-6, F 2e-5, gradient 1e-4)."""
    el, ex, eF, eg = _edge_episode(*_inverted_points(), prec)
    print(f"inverted/flat prec {prec}: loss {el:.2e} x {ex:.2e} F {eF:.2e} grad {eg:.2e}")
    assert el < tol[0] and ex < tol[1] and eF < tol[2] and eg < tol[3]


def test_chain_episode():
    case = make_case(L=4)
    mo = gpu_morph(case)
    mo.forward_pass(0)
    last = mo.layer_state(case.L - 1)
    mo.chain_episode()
    first = mo.layer_state(0)
    for a, b in zip(first, last):
        assert np.array_equal(a, b)
    assert np.abs(mo.control(0)).max() == 0.0


def test_error_paths():
    case = make_case(L=3)
    mo = gpu_morph(case)
    with pytest.raises(pkg.DmpmError):
        mo.backward_pass(0)  # before the loss
    with pytest.raises(pkg.DmpmError):
        mo.forward_pass(5)   # out of range
    with pytest.raises(pkg.DmpmError):
        pkg.Morph(10, (0, 4, 4), 0.1, 3)


@pytest.mark.parametrize("prec", [pkg.PREC_F64, pkg.PREC_F32])
def test_out_of_memory_then_a_working_context(prec):
    """A tape far beyond HBM (1M particles x 100 000 layers): dmpm_create returns
    DMPM_ERR_NOMEM and frees what it had; the failed allocation leaves no error
    behind, so the next context's episode runs and matches the oracle (1e-10
    FP64, 1e-5 FP32 on the loss)."""
    with pytest.raises(pkg.DmpmError, match="dmpm error -4"):
        pkg.Morph(1_000_000, (8, 8, 8), 0.125, 100_000, precision=prec)
    case = make_case(L=3)
    mo = gpu_morph(case, precision=prec)
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    ref_loss = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, case.L)[0]
    assert loss == pytest.approx(ref_loss, rel=1e-10 if prec == pkg.PREC_F64 else 1e-5)
    mo.close()


# ------------------------------------------------- FP32 tiled scatter paths --
def _one_step_case(x, seed=11):
    rng = np.random.default_rng(seed)
    sc = pkg.scenes.morph_small()
    n = len(x)
    m = np.full(n, sc.m[0])
    lam = np.full(n, sc.lam[0])
    mu = np.full(n, sc.mu[0])
    prm = orc.Params(n=n, dim=sc.dim, dx=sc.dx, dt=1e-4)
    vol = orc.point_volumes(prm, x, m) * sc.dx ** 3
    cs = orc.Consts(m, vol, lam, mu)
    st = orc.State(x, 0.1 * rng.normal(size=(n, 3)), rng.normal(size=(n, 3, 3)),
                   np.eye(3) + 0.01 * rng.normal(size=(n, 3, 3)))
    dFc = 1e-2 * rng.normal(size=(n, 3, 3))
    target = orc.p2g_mass(prm, sc.target_x, sc.target_m)
    return sc, prm, cs, st, dFc, target


@pytest.mark.parametrize("layout", ["scattered", "corner_blocks", "shuffled_sphere"])
def test_fp32_tiled_scatter_layouts(layout):
    """Window fits (LDS path) and does not fit (HBM-atomic fallback): both match the oracle."""
    rng = np.random.default_rng(3)
    sc0 = pkg.scenes.morph_small()
    if layout == "scattered":
        x = rng.uniform(-0.05, 1.05, size=(3000, 3))
    elif layout == "corner_blocks":
        x = np.concatenate([pkg.scenes.lattice_box([-0.06] * 3, [0.12] * 3, sc0.point_dx),
                            pkg.scenes.lattice_box([1.02, 0.4, 0.4], [1.1, 0.5, 0.5], sc0.point_dx)])
    else:
        x = sc0.x[rng.permutation(sc0.n)]
    sc, prm, cs, st, dFc, target = _one_step_case(x)
    n = len(x)
    L = 2
    ref_loss, layers, gd, g0 = orc.episode(prm, st, cs, [dFc], target, L)
    mo = pkg.Morph(n, sc.dim, sc.dx, L, dt=1e-4, drag=0.0, smoothing=0.0, precision=pkg.PREC_F32)
    mo.set_particles(st.x, st.v, st.F, st.C, cs.m, cs.vol, cs.lam, cs.mu)
    mo.set_control(0, dFc)
    mo.set_target_grid(target)
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    x1, v1, F1, C1 = mo.layer_state(1)
    ev, eC = maxrel(v1, layers[1].v), maxrel(C1, layers[1].C)
    eg, ex = rel(mo.control_grad(0), gd[0]), rel(mo.layer_grads(0)[0], g0.x)
    print(f"{layout}: v {ev:.2e} C {eC:.2e} grad dFc {eg:.2e} grad x {ex:.2e}")
    assert maxrel(x1, layers[1].x) < 1e-6
    assert ev < 1e-4 and eC < 1e-4
    assert abs(loss - ref_loss) / abs(ref_loss) < 1e-5
    assert eg < 1e-4 and ex < 1e-4


def test_fp32_api_order_roundtrip():
    """The FP32 path stores particles block-sorted; every API array comes back in API order."""
    case = make_case(L=3)
    rng = np.random.default_rng(5)
    perm = rng.permutation(case.sc.n)
    st = case.st0
    mo = pkg.Morph(case.sc.n, case.sc.dim, case.sc.dx, 3, dt=1e-4, precision=pkg.PREC_F32)
    mo.set_particles(st.x[perm], st.v[perm], st.F[perm], st.C[perm], case.cs.m[perm], case.cs.vol[perm],
                     case.cs.lam[perm], case.cs.mu[perm])
    c = case.controls[0][perm]
    mo.set_control(1, c)
    assert np.array_equal(mo.control(1), c.astype(np.float32).astype(np.float64))
    x0 = mo.layer_state(0)[0]
    assert np.array_equal(x0, f32_positions(st.x[perm], case.sc.dim, case.sc.dx, case.sc.min_point))
    assert np.array_equal(mo.volumes(), case.cs.vol[perm].astype(np.float32).astype(np.float64))


# ------------------------------------------------------------- phase calls --
def test_phase_calls_compose_to_the_forward_step():
    """dmpm_phase_* in ForwardTimeStep's order (P_op_1 via dmpm_eval_pk1, P2G, G_op, G2P,
    P_op_2; ForwardSimulation.cpp:136-157) = the oracle's step at 1e-12 (states and grid),
    with drag, smoothing, f_ext and non-zero v.  Also: P2G adds into the grid it is given,
    G_op leaves the velocity of nodes at or below FLT_EPSILON mass, and P2G mass of unit
    points = the oracle's."""
    case = make_case(L=2, drag=0.2, smoothing=0.05, f_ext=(0.4, -0.3, 0.2), vel=0.1)
    sc, st, cs, prm = case.sc, case.st0, case.cs, case.prm
    dFc = case.controls[0]
    grid = (sc.dim, sc.dx, sc.min_point)
    P = pkg.eval_pk1(st.F + dFc, cs.lam, cs.mu)
    gm, gp = pkg.phase_p2g(grid, st.x, st.v, st.C, st.F, dFc, P, cs.m, cs.vol, prm.dt, prm.drag)
    gv = pkg.phase_grid_op(grid, gm, gp, prm.dt, prm.f_ext)
    vn, Cn = pkg.phase_g2p(grid, st.x, gv)
    xn, Fn = pkg.phase_p_op2(st.x, st.F, dFc, vn, Cn, prm.smoothing, prm.dt)
    nxt, om, ov = orc.forward_step(prm, st, cs, dFc)
    assert maxrel(gm, om) < 1e-12 and maxrel(gv, ov) < 1e-12
    assert maxrel(vn, nxt.v) < 1e-12 and maxrel(Cn, nxt.C) < 1e-12
    assert maxrel(xn, nxt.x) < 1e-12 and maxrel(Fn, nxt.F) < 1e-12
    _, op = orc.p2g(prm, st, cs, dFc)
    assert maxrel(gp, op) < 1e-12
    gm2, gp2 = pkg.phase_p2g(grid, st.x, st.v, st.C, st.F, dFc, P, cs.m, cs.vol, prm.dt, prm.drag, gm, gp)
    assert maxrel(gm2, 2 * gm) < 1e-12 and maxrel(gp2, 2 * gp) < 1e-12
    keep = np.full((gm.size, 3), 7.0)
    gvk = pkg.phase_grid_op(grid, gm, gp, prm.dt, prm.f_ext, keep)
    dead = ~(np.abs(gm) > 1.1920928955078125e-07)
    assert dead.any() and np.all(gvk[dead] == 7.0) and np.array_equal(gvk[~dead], gv[~dead])
    assert maxrel(pkg.phase_p2g_mass(grid, st.x, np.ones(sc.n)), orc.p2g_mass(prm, st.x, np.ones(sc.n))) < 1e-12
