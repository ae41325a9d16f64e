"""Pins of the CPU FP64 oracle (no GPU).

The reference ships no golden vectors, fixtures or known-answer files for
this path (its checks are GUI buttons that print, DiffMPMAnimator3D.cpp:212-243,
CompGraph.cpp:457-504), and its sources need CUDA toolkit headers to build.
Each oracle function is therefore pinned by what the mathematics fixes:
closed forms, library routines (numpy/scipy SVD and polar), invariants,
brute-force enumeration and central finite differences.
"""
import math
import os

import numpy as np
import pytest
import scipy.linalg

from conftest import load_oracle, load_pkg

orc = load_oracle()
scenes = load_pkg().scenes


def rot(axis, ang):
    return scipy.linalg.expm(np.cross(np.eye(3), np.asarray(axis, float) / np.linalg.norm(axis)) * ang)


# ------------------------------------------------------------------ spline --
def test_bspline_closed_form_values():
    assert orc.bspline(0.0) == pytest.approx(2.0 / 3.0, abs=1e-15)
    assert orc.bspline(1.0) == pytest.approx(1.0 / 6.0, abs=1e-15)
    assert orc.bspline(-1.0) == pytest.approx(1.0 / 6.0, abs=1e-15)
    assert orc.bspline(2.0) == 0.0 and orc.bspline(2.5) == 0.0
    assert orc.bspline(0.5) == pytest.approx(23.0 / 48.0, abs=1e-15)
    assert orc.bspline(1.5) == pytest.approx(1.0 / 48.0, abs=1e-15)


@pytest.mark.parametrize("f", [0.0, 0.1234, 0.5, 0.87, 0.999999])
def test_bspline_moments(f):
    """sum w = 1, sum w d = 0, sum w d^2 = 1/3 (fixes D^-1 = 3/dx^2)."""
    d = np.arange(-1, 3) - f  # offsets of the 4 stencil nodes, base = floor - 1
    w = np.array([orc.bspline(u) for u in d])
    assert w.sum() == pytest.approx(1.0, abs=1e-14)
    assert (w * d).sum() == pytest.approx(0.0, abs=1e-14)
    assert (w * d * d).sum() == pytest.approx(1.0 / 3.0, abs=1e-14)
    g = np.array([orc.bspline_slope(u) for u in d])
    assert g.sum() == pytest.approx(0.0, abs=1e-13)


def test_bspline_slope_is_derivative():
    for x in np.linspace(-2.4, 2.4, 97):
        if min(abs(abs(x) - 1), abs(abs(x) - 2), abs(x)) < 1e-3:
            continue
        h = 1e-6
        fd = (orc.bspline(x + h) - orc.bspline(x - h)) / (2 * h)
        assert orc.bspline_slope(x) == pytest.approx(fd, abs=1e-8)


# -------------------------------------------------------------- svd/polar --
def _rand_mats(n, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return np.eye(3) + scale * rng.normal(size=(n, 3, 3))


def test_svd_against_numpy_and_conventions():
    mats = list(_rand_mats(200, 1, 0.6)) + [np.diag([2.0, -0.5, 1.0]), np.diag([1.0, 1.0, 1.0]),
                                              np.diag([3.0, 3.0, 0.0]), 0.8 * rot([1, 0, 0], math.pi / 4)]
    for A in mats:
        U, s, V = orc.svd3(A)
        ref = np.linalg.svd(A, compute_uv=False)
        assert np.allclose(np.abs(s), ref, rtol=1e-12, atol=1e-12)
        assert abs(s[0]) >= abs(s[1]) >= abs(s[2])
        assert s[0] >= 0 and s[1] >= 0
        assert np.linalg.det(U) == pytest.approx(1.0, abs=1e-12)
        assert np.linalg.det(V) == pytest.approx(1.0, abs=1e-12)
        assert np.allclose(U @ U.T, np.eye(3), atol=1e-12)
        assert np.allclose(U @ np.diag(s) @ V.T, A, atol=1e-12)
        assert np.sign(s[2]) == np.sign(np.linalg.det(A)) or abs(np.linalg.det(A)) < 1e-14


def test_polar_against_scipy_and_procrustes():
    for A in _rand_mats(200, 2, 0.6):
        R, S = orc.polar3(A)
        assert np.allclose(R @ S, A, atol=1e-12)
        assert np.allclose(S, S.T, atol=1e-12)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-12)
        if np.linalg.det(A) > 0:
            Rs, _ = scipy.linalg.polar(A)
            assert np.allclose(R, Rs, atol=1e-12)
        # closest rotation (Kabsch): maximises tr(R^T A) over SO(3)
        u, _, vt = np.linalg.svd(A)
        D = np.diag([1.0, 1.0, np.sign(np.linalg.det(u @ vt))])
        Rk = u @ D @ vt
        assert np.allclose(R, Rk, atol=1e-10)


# ----------------------------------------------------------------- PK1 / H --
LAM, MU = 58333.0, 38888.9  # MPMPolyscope.h:31-32, the reference's gradient-check material


def test_pk1_rest_and_reference_check_point():
    assert np.allclose(orc.pk1(np.eye(3), LAM, MU), 0.0, atol=1e-12)
    # DiffMPMAnimator3D.cpp:219-225: F = 0.8 Rx(45 deg); P(sR) = (2mu(s-1) + lam(s^3-1)s^2) R
    R = rot([1, 0, 0], math.pi / 4)
    s = 0.8
    P = orc.pk1(s * R, LAM, MU)
    assert np.allclose(P, (2 * MU * (s - 1) + LAM * (s ** 3 - 1) * s ** 2) * R, rtol=1e-12, atol=1e-9)


def test_pk1_is_energy_gradient():
    rng = np.random.default_rng(3)
    for F in _rand_mats(20, 4, 0.25):
        lam, mu = rng.uniform(1e3, 1e5, size=2)
        P = orc.pk1(F, lam, mu)
        h = 1e-6
        fd = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                E = np.zeros((3, 3))
                E[i, j] = h
                fd[i, j] = (orc.energy(F + E, lam, mu) - orc.energy(F - E, lam, mu)) / (2 * h)
        assert np.allclose(P, fd, rtol=1e-6, atol=1e-6 * np.abs(P).max())


def test_pk1_rotation_covariance():
    for k, F in enumerate(_rand_mats(20, 5, 0.3)):
        Q = rot([1, 2, 3 + k], 0.3 * k + 0.1)
        assert np.allclose(orc.pk1(Q @ F, LAM, MU), Q @ orc.pk1(F, LAM, MU), atol=1e-7)


def test_hessian_at_identity_closed_form():
    rng = np.random.default_rng(6)
    for _ in range(10):
        dF = rng.normal(size=(3, 3))
        H = orc.hess_mult(np.eye(3), LAM, MU, dF)
        assert np.allclose(H, 2 * MU * 0.5 * (dF + dF.T) + LAM * np.trace(dF) * np.eye(3), rtol=1e-12, atol=1e-8)


@pytest.mark.parametrize("scale", [0.05, 0.3, 0.6])
def test_hessian_matches_fd_of_pk1(scale):
    rng = np.random.default_rng(7)
    for F in _rand_mats(30, 8 + int(scale * 10), scale):
        if np.linalg.det(F) <= 1e-3:
            continue
        dF = rng.normal(size=(3, 3))
        h = 1e-6
        fd = (orc.pk1(F + h * dF, LAM, MU) - orc.pk1(F - h * dF, LAM, MU)) / (2 * h)
        H = orc.hess_mult(F, LAM, MU, dF)
        assert np.allclose(H, fd, rtol=1e-5, atol=1e-5 * np.abs(fd).max())


def test_hessian_major_symmetry():
    rng = np.random.default_rng(9)
    for F in _rand_mats(20, 10, 0.4):
        A, B = rng.normal(size=(2, 3, 3))
        a = np.sum(A * orc.hess_mult(F, LAM, MU, B))
        b = np.sum(B * orc.hess_mult(F, LAM, MU, A))
        assert a == pytest.approx(b, rel=1e-9)


def test_lame():
    lam, mu = orc.lame(1e4, 0.3)
    assert lam == pytest.approx(1e4 * 0.3 / (1.3 * 0.4))
    assert mu == pytest.approx(1e4 / 2.6)
    l2, m2 = scenes.lame(1e4, 0.3)
    assert l2 == pytest.approx(lam, rel=1e-15) and m2 == pytest.approx(mu, rel=1e-15)


# ------------------------------------------------------------------ step ----
def tiny_scene(n=60, grid=12, seed=0, vel=0.3, ctrl=0.02, jitter=True, lam_mu=(300.0, 200.0)):
    rng = np.random.default_rng(seed)
    dx = 1.0 / grid
    x = rng.uniform(0.3, 0.7, size=(n, 3)) if jitter else None
    m = rng.uniform(0.5, 1.5, size=n) * dx ** 3
    prm = orc.Params(n=n, dim=(grid, grid, grid), dx=dx, dt=2e-3, drag=0.3, smoothing=0.1, f_ext=(0.5, -1.0, 0.2))
    vol = orc.point_volumes(prm, x, m) * dx ** 3
    cs = orc.Consts(m, vol, np.full(n, lam_mu[0]), np.full(n, lam_mu[1]))
    st = orc.State(x, vel * rng.normal(size=(n, 3)), 2.0 * rng.normal(size=(n, 3, 3)),
                   np.eye(3) + 0.05 * rng.normal(size=(n, 3, 3)))
    dFc = ctrl * rng.normal(size=(n, 3, 3))
    return prm, st, cs, dFc


def _brute_weight(xp, xi, dx):
    def N(u):
        a = abs(u)
        return 0.5 * a ** 3 - a * a + 2 / 3 if a < 1 else ((2 - a) ** 3 / 6 if a < 2 else 0.0)
    return N((xi[0] - xp[0]) / dx) * N((xi[1] - xp[1]) / dx) * N((xi[2] - xp[2]) / dx)


def test_p2g_mass_brute_force_stencil():
    """Every grid node, no stencil: pins base = floor - 1, the 4^3 window and bounds."""
    rng = np.random.default_rng(11)
    grid = 6
    dx = 1.0 / grid
    x = rng.uniform(-0.05, 0.95, size=(25, 3))  # includes particles whose stencil leaves the grid
    m = rng.uniform(0.5, 1.5, size=25)
    prm = orc.Params(n=25, dim=(grid, grid + 1, grid - 1), dx=dx, min_point=(0.0, 0.0, 0.0))
    gm = orc.p2g_mass(prm, x, m)
    ref = np.zeros(prm.G)
    for i in range(grid):
        for j in range(grid + 1):
            for k in range(grid - 1):
                xi = np.array([i, j, k]) * dx
                ref[(i * (grid + 1) + j) * (grid - 1) + k] = sum(m[p] * _brute_weight(x[p], xi, dx) for p in range(25))
    assert np.allclose(gm, ref, rtol=1e-13, atol=1e-15)


def test_volumes_closed_form():
    """One particle on a node: vol = 1/sum w^2 = 1/(1/2)^3 = 8 (dimensionless, ForwardSimulation.cpp:288)."""
    prm = orc.Params(n=1, dim=(8, 8, 8), dx=0.125)
    vol = orc.point_volumes(prm, np.array([[0.5, 0.5, 0.5]]), np.array([2.0]))
    assert vol[0] == pytest.approx(8.0, rel=1e-13)


def test_forward_conservation():
    prm, st, cs, dFc = tiny_scene()
    gm, gp = orc.p2g(prm, st, cs, dFc)
    assert gm.sum() == pytest.approx(cs.m.sum(), rel=1e-13)
    mom = (cs.m[:, None] * st.v).sum(0) * (1 - prm.dt * prm.drag)
    assert np.allclose(gp.sum(0), mom, rtol=1e-11, atol=1e-14)


def test_forward_rest_state_is_static():
    prm, st, cs, _ = tiny_scene()
    st = orc.State(st.x, np.zeros_like(st.v), np.zeros_like(st.C), np.broadcast_to(np.eye(3), st.F.shape).copy())
    prm.f_ext = (0.0, 0.0, 0.0)
    nxt, _, gv = orc.forward_step(prm, st, cs, np.zeros_like(st.F))
    assert np.allclose(nxt.x, st.x, atol=0) and np.allclose(nxt.F, st.F, atol=1e-15)
    assert np.abs(nxt.v).max() == 0.0 and np.abs(gv).max() == 0.0


def test_forward_affine_reproduction():
    """v(x) = A x + b with C = A and no stress: G2P returns v' = A x_p + b and C' = A."""
    prm, st, cs, _ = tiny_scene(n=40)
    rng = np.random.default_rng(12)
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    st = orc.State(st.x, st.x @ A.T + b, np.broadcast_to(A, st.C.shape).copy(), np.broadcast_to(np.eye(3), st.F.shape).copy())
    prm.drag = 0.0
    prm.f_ext = (0.0, 0.0, 0.0)
    prm.smoothing = 0.0
    # heavy particles: nodes below the mass threshold (w m < FLT_EPSILON) carry w < 1e-13
    cs = orc.Consts(cs.m * 1e6 / cs.m.mean(), cs.vol, np.zeros(prm.n), np.zeros(prm.n))
    nxt, _, _ = orc.forward_step(prm, st, cs, np.zeros_like(st.F))
    assert np.allclose(nxt.v, st.x @ A.T + b, atol=1e-12)
    assert np.allclose(nxt.C, A, atol=1e-11)
    assert np.allclose(nxt.F, np.eye(3) + prm.dt * A, atol=1e-12)
    assert np.allclose(nxt.x, st.x + prm.dt * nxt.v, atol=1e-15)


# ------------------------------------------------------------------ loss ----
def test_loss_against_own_target():
    """Target = the cloud's own mass: the log term vanishes, leaving the soft-min penalty."""
    prm, st, cs, _ = tiny_scene(n=50)
    gm = orc.p2g_mass(prm, st.x, cs.m)
    loss, gm2, dLdm, _ = orc.end_mass_loss(prm, st.x, cs.m, gm)
    assert np.allclose(gm2, gm)
    low = gm < 1e-3
    assert loss == pytest.approx(10.0 * ((1e-3 - gm[low]) ** 2).sum(), rel=1e-12)
    assert np.allclose(dLdm[low], -20.0 * (1e-3 - gm[low]), atol=1e-15)
    assert np.allclose(dLdm[~low], 0.0, atol=1e-15)


def test_loss_gradient_fd():
    prm, st, cs, _ = tiny_scene(n=30, seed=13)
    rng = np.random.default_rng(14)
    target = orc.p2g_mass(prm, st.x + 0.03 * rng.normal(size=st.x.shape), cs.m) * 1e3
    loss, _, _, dLdx = orc.end_mass_loss(prm, st.x, cs.m * 1e3, target)
    h = 1e-7
    for p in range(0, 30, 7):
        for a in range(3):
            xp = st.x.copy()
            xm = st.x.copy()
            xp[p, a] += h
            xm[p, a] -= h
            fd = (orc.end_mass_loss(prm, xp, cs.m * 1e3, target)[0] - orc.end_mass_loss(prm, xm, cs.m * 1e3, target)[0]) / (2 * h)
            assert dLdx[p, a] == pytest.approx(fd, rel=1e-5, abs=1e-7 * np.abs(dLdx).max())


# --------------------------------------------------------------- adjoint ----
def _flat(st: "orc.State"):
    return np.concatenate([st.x.ravel(), st.v.ravel(), st.C.ravel(), st.F.ravel()])


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_step_dot_product(seed):
    """<J du, lambda> = <du, J^T lambda> for one step, J by FP64 central FD."""
    prm, st, cs, dFc = tiny_scene(n=40, seed=20 + seed)
    rng = np.random.default_rng(30 + seed)
    n = prm.n
    nxt, gm, gv = orc.forward_step(prm, st, cs, dFc)
    lam_ = orc.Grads(rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3, 3)), rng.normal(size=(n, 3, 3)))
    g = orc.backward_step(prm, st, cs, dFc, gm, gv, nxt.C, lam_)
    du = orc.State(1e-3 * rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3, 3)), rng.normal(size=(n, 3, 3)))
    ddFc = rng.normal(size=(n, 3, 3))
    lam_flat = np.concatenate([lam_.x.ravel(), lam_.v.ravel(), lam_.C.ravel(), lam_.F.ravel()])

    def J(h):
        s2 = orc.State(st.x + h * du.x, st.v + h * du.v, st.C + h * du.C, st.F + h * du.F)
        out, _, _ = orc.forward_step(prm, s2, cs, dFc + h * ddFc)
        return _flat(out) @ lam_flat

    h = 1e-6
    fd = (J(h) - J(-h)) / (2 * h)
    an = (np.sum(g.x * du.x) + np.sum(g.v * du.v) + np.sum(g.C * du.C) + np.sum(g.F * du.F) + np.sum(g.dFc * ddFc))
    assert an == pytest.approx(fd, rel=2e-6)


def test_backward_compat_difference_is_defect_terms():
    """Compat (a): the extra (I + dt C')^T F' of BackPropagation.cpp:121; (b): C' gets dt F' Fc^T without (1-s)."""
    prm, st, cs, dFc = tiny_scene(n=30, seed=40)
    prm.f_ext = (0.0, 0.0, 0.0)
    prm.drag = 0.0
    rng = np.random.default_rng(41)
    n = prm.n
    nxt, gm, gv = orc.forward_step(prm, st, cs, dFc)
    lam_ = orc.Grads(rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3, 3)), rng.normal(size=(n, 3, 3)))
    ex = orc.backward_step(prm, st, cs, dFc, gm, gv, nxt.C, lam_)
    prm_a = orc.Params(**{**prm.__dict__, "compat": orc.COMPAT_DUP_FBAR})
    ca = orc.backward_step(prm_a, st, cs, dFc, gm, gv, nxt.C, lam_)
    IC = np.eye(3) + prm.dt * nxt.C
    base = np.einsum("pki,pkj->pij", IC, lam_.F)
    assert np.allclose(ca.F - ex.F, base, rtol=1e-9, atol=1e-9 * np.abs(base).max())
    assert np.allclose(ca.x, ex.x) and np.allclose(ca.v, ex.v) and np.allclose(ca.C, ex.C)
    # with s = 0 defect (b) vanishes
    prm0 = orc.Params(**{**prm.__dict__, "smoothing": 0.0})
    nxt0, gm0, gv0 = orc.forward_step(prm0, st, cs, dFc)
    e0 = orc.backward_step(prm0, st, cs, dFc, gm0, gv0, nxt0.C, lam_)
    b0 = orc.backward_step(orc.Params(**{**prm0.__dict__, "compat": orc.COMPAT_NO_1MS}), st, cs, dFc, gm0, gv0, nxt0.C, lam_)
    assert np.allclose(e0.x, b0.x, rtol=1e-12) and np.allclose(e0.F, b0.F, rtol=1e-12)


def _compat_case(seed):
    """One step of the tiny scene (s = 0.1, drag = 0.3, f_ext != 0, v != 0) and a random seed lambda."""
    prm, st, cs, dFc = tiny_scene(n=30, seed=seed)
    rng = np.random.default_rng(seed + 1)
    n = prm.n
    nxt, gm, gv = orc.forward_step(prm, st, cs, dFc)
    lam_ = orc.Grads(rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3, 3)), rng.normal(size=(n, 3, 3)))
    return prm, st, cs, dFc, nxt, gm, gv, lam_


def _bs(t):
    """Cubic B-spline N and N' (Interpolation.h:5-30) from their closed forms, elementwise."""
    a = np.abs(t)
    N = np.where(a < 1, 0.5 * a ** 3 - a * a + 2.0 / 3.0, np.where(a < 2, (2 - a) ** 3 / 6.0, 0.0))
    dN = np.where(a < 1, 1.5 * t * a - 2.0 * t, np.where(a < 2, -0.5 * (2 - a) ** 2 * np.sign(t), 0.0))
    return N, dN


def _brute_weights(prm, x):
    """w_ip, dw_ip/dx_p (n,G,3) and d_ip = x_i - x_p (n,G,3) over EVERY node of the grid."""
    ii, jj, kk = np.meshgrid(*[np.arange(d) for d in prm.dim], indexing="ij")
    xi = np.asarray(prm.min_point) + prm.dx * np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)  # Grid.cpp:22
    d = xi[None, :, :] - x[:, None, :]
    N, dN = _bs(-d / prm.dx)  # argument (x_p - x_i)/dx: d/dx_p N = N'/dx
    w = N[..., 0] * N[..., 1] * N[..., 2]
    gw = np.stack([dN[..., 0] * N[..., 1] * N[..., 2], N[..., 0] * dN[..., 1] * N[..., 2],
                   N[..., 0] * N[..., 1] * dN[..., 2]], axis=-1) / prm.dx
    return w, gw, d


def _brute_pbar(prm, st, dFc, gm, lam_, cs_factor):
    """pbar_i = vbar_i / m_i (|m_i| > FLT_EPSILON) with vbar_i = sum_p w (vt + D Ct d),
    vt = v' + dt x', Ct = C' + cs dt F' Fc^T (BackPropagation.cpp:23-59), brute force."""
    w, _, d = _brute_weights(prm, st.x)
    Fc = st.F + dFc
    vt = lam_.v + prm.dt * lam_.x
    Ct = lam_.C + cs_factor * prm.dt * np.einsum("pik,pjk->pij", lam_.F, Fc)
    D = 3.0 / prm.dx ** 2
    vbar = np.einsum("pg,pa->ga", w, vt) + D * np.einsum("pg,pab,pgb->ga", w, Ct, d)
    ok = np.abs(gm) > 1.1920928955078125e-07
    return np.where(ok[:, None], vbar / np.where(ok, gm, 1.0)[:, None], 0.0)


def test_compat_no_1ms_is_exact_adjoint_of_shifted_cbar():
    """Defect (b), BackPropagation.cpp:24: dLdC_next = dt dLdF' Fc^T + dLdC' lacks the
    (1-s) of the exact Ct = C' + (1-s) dt F' Fc^T.  The adjoint is linear in the seed
    and C' enters only through Ct, so (b) alone must equal the exact adjoint with
    C' + s dt F' Fc^T in every output; a wrong factor (1-s, 2s) or Fc in place of
    Fc^T does not."""
    prm, st, cs, dFc, nxt, gm, gv, lam_ = _compat_case(60)
    assert prm.smoothing == pytest.approx(0.1)
    comp = orc.backward_step(orc.Params(**{**prm.__dict__, "compat": orc.COMPAT_NO_1MS}), st, cs, dFc, gm, gv, nxt.C, lam_)
    Fc = st.F + dFc

    def shifted(extra):
        lam2 = orc.Grads(lam_.x, lam_.v, lam_.F, lam_.C + extra)
        return orc.backward_step(prm, st, cs, dFc, gm, gv, nxt.C, lam2)

    good = shifted(prm.smoothing * prm.dt * np.einsum("pik,pjk->pij", lam_.F, Fc))
    for f in ("x", "v", "F", "C", "dFc"):
        a, b = getattr(comp, f), getattr(good, f)
        assert np.abs(a - b).max() <= 1e-11 * np.abs(b).max(), f
    ex = orc.backward_step(prm, st, cs, dFc, gm, gv, nxt.C, lam_)
    assert np.abs(comp.x - ex.x).max() > 1e-6 * np.abs(ex.x).max()  # the defect is visible here
    for bad in (shifted((1 - prm.smoothing) * prm.dt * np.einsum("pik,pjk->pij", lam_.F, Fc)),
                shifted(2 * prm.smoothing * prm.dt * np.einsum("pik,pjk->pij", lam_.F, Fc)),
                shifted(prm.smoothing * prm.dt * np.einsum("pik,pkj->pij", lam_.F, Fc))):
        assert np.abs(comp.x - bad.x).max() > 1e-6 * np.abs(comp.x).max()


def test_compat_no_drag_x_closed_form():
    """Defect (c), BackPropagation.cpp:109 vs ForwardSimulation.cpp:48: the x-gradient's
    weight-gradient term uses m v where the forward scatters m (1 - dt drag) v, so (c)
    alone changes xbar only, by sum_i grad w_ip (dt drag m_p v_p . pbar_i), with
    pbar and grad w by brute force over every node; every other output is unchanged."""
    prm, st, cs, dFc, nxt, gm, gv, lam_ = _compat_case(61)
    assert prm.drag != 0.0
    ex, pbar_o, _ = orc.backward_step(prm, st, cs, dFc, gm, gv, nxt.C, lam_, return_grid=True)
    comp = orc.backward_step(orc.Params(**{**prm.__dict__, "compat": orc.COMPAT_NO_DRAG_X}), st, cs, dFc, gm, gv, nxt.C, lam_)
    pbar = _brute_pbar(prm, st, dFc, gm, lam_, 1.0 - prm.smoothing)
    assert np.abs(pbar_o - pbar).max() <= 1e-10 * np.abs(pbar).max()
    _, gw, _ = _brute_weights(prm, st.x)
    db = prm.dt * prm.drag * cs.m[:, None] * st.v  # m v - m (1 - dt drag) v
    expect = np.einsum("pga,pg->pa", gw, np.einsum("pb,gb->pg", db, pbar))
    diff = comp.x - ex.x
    assert np.abs(diff).max() > 1e-6 * np.abs(ex.x).max()
    assert np.abs(diff - expect).max() <= 1e-10 * np.abs(expect).max()
    assert np.abs(diff - 2 * expect).max() > 0.5 * np.abs(expect).max()  # a wrong factor fails
    for f in ("v", "F", "C", "dFc"):  # the OpenMP scatter's order leaves round-off only
        assert np.abs(getattr(comp, f) - getattr(ex, f)).max() <= 1e-12 * np.abs(getattr(ex, f)).max(), f


def test_compat_dldm_no_fext_closed_form():
    """Defect (d), BackPropagation.cpp:57: dLdm = -(v.dLdv)/m uses the node velocity v,
    not v - dt f_ext, the velocity before G_op added dt f_ext; so mbar_i changes by
    -dt (f_ext . pbar_i) and (d) alone changes xbar only, by
    sum_i m_p grad w_ip (-dt f_ext . pbar_i) (brute force over every node)."""
    prm, st, cs, dFc, nxt, gm, gv, lam_ = _compat_case(62)
    f = np.asarray(prm.f_ext)
    assert np.abs(f).max() > 0
    ex, _, mb_ex = orc.backward_step(prm, st, cs, dFc, gm, gv, nxt.C, lam_, return_grid=True)
    comp, _, mb_c = orc.backward_step(orc.Params(**{**prm.__dict__, "compat": orc.COMPAT_DLDM_NO_FEXT}), st, cs, dFc, gm,
                                      gv, nxt.C, lam_, return_grid=True)
    pbar = _brute_pbar(prm, st, dFc, gm, lam_, 1.0 - prm.smoothing)
    dmb = -prm.dt * pbar @ f
    assert np.abs((mb_c - mb_ex) - dmb).max() <= 1e-10 * np.abs(dmb).max()
    _, gw, _ = _brute_weights(prm, st.x)
    expect = cs.m[:, None] * np.einsum("pga,g->pa", gw, dmb)
    diff = comp.x - ex.x
    assert np.abs(diff).max() > 1e-6 * np.abs(ex.x).max()
    assert np.abs(diff - expect).max() <= 1e-10 * np.abs(expect).max()
    assert np.abs(diff + expect).max() > 0.5 * np.abs(expect).max()  # the opposite sign fails
    for f_ in ("v", "F", "C", "dFc"):
        assert np.abs(getattr(comp, f_) - getattr(ex, f_)).max() <= 1e-12 * np.abs(getattr(ex, f_)).max(), f_


def test_compat_reference_composes_the_four_defects():
    """All four defects at once (the reference's Back_Timestep, compat = 15), from the
    exact adjoint and closed forms only: (b) as the exact adjoint with C' + s dt F' Fc^T,
    whose pbar (brute force, Ct without 1-s) then carries (c) and (d) into xbar; (a)
    adds (I + dt C')^T F' to the F-gradient; and the control gradient is the reference's
    dLdF, the F-gradient itself."""
    prm, st, cs, dFc, nxt, gm, gv, lam_ = _compat_case(63)
    full = orc.backward_step(orc.Params(**{**prm.__dict__, "compat": orc.COMPAT_DUP_FBAR | orc.COMPAT_NO_1MS | orc.COMPAT_NO_DRAG_X | orc.COMPAT_DLDM_NO_FEXT}), st, cs, dFc, gm, gv, nxt.C,
                             lam_)
    Fc = st.F + dFc
    lam2 = orc.Grads(lam_.x, lam_.v, lam_.F, lam_.C + prm.smoothing * prm.dt * np.einsum("pik,pjk->pij", lam_.F, Fc))
    gb = orc.backward_step(prm, st, cs, dFc, gm, gv, nxt.C, lam2)
    pbar = _brute_pbar(prm, st, dFc, gm, lam_, 1.0)
    _, gw, _ = _brute_weights(prm, st.x)
    db = prm.dt * prm.drag * cs.m[:, None] * st.v
    dmb = -prm.dt * pbar @ np.asarray(prm.f_ext)
    want_x = gb.x + np.einsum("pga,pg->pa", gw, np.einsum("pb,gb->pg", db, pbar)) + \
        cs.m[:, None] * np.einsum("pga,g->pa", gw, dmb)
    base = np.einsum("pki,pkj->pij", np.eye(3) + prm.dt * nxt.C, lam_.F)
    want_F = gb.F + base
    assert np.abs(full.x - want_x).max() <= 1e-10 * np.abs(want_x).max()
    assert np.abs(full.F - want_F).max() <= 1e-10 * np.abs(want_F).max()
    assert np.array_equal(full.dFc, full.F)
    for f in ("v", "C"):
        assert np.abs(getattr(full, f) - getattr(gb, f)).max() <= 1e-11 * np.abs(getattr(gb, f)).max(), f


def test_episode_gradient_vs_central_fd():
    """Whole-episode dL/d(dFc_0) against FP64 central differences (h = 1e-4, Richardson-extrapolated)."""
    prm, st, cs, dFc = tiny_scene(n=40, seed=50, vel=0.1, ctrl=0.01)
    prm.smoothing = 0.0
    L = 5
    rng = np.random.default_rng(51)
    controls = [dFc] + [0.01 * rng.normal(size=dFc.shape) for _ in range(L - 2)]
    target = orc.p2g_mass(prm, st.x + 0.05, cs.m) * 3.0
    loss, layers, gdFc, g0 = orc.episode(prm, st, cs, controls, target, L)
    idx = [(p, i, j) for p in range(0, 40, 5) for (i, j) in [(0, 0), (1, 2), (2, 1)]]

    def cfd(p, i, j, h):
        cp = [c.copy() for c in controls]
        cm = [c.copy() for c in controls]
        cp[0][p, i, j] += h
        cm[0][p, i, j] -= h
        return (orc.loss_only(prm, st, cs, cp, target, L) - orc.loss_only(prm, st, cs, cm, target, L)) / (2 * h)

    h = 1e-4
    for (p, i, j) in idx:
        fd = (4.0 * cfd(p, i, j, h / 2) - cfd(p, i, j, h)) / 3.0  # Richardson: O(h^4)
        assert gdFc[0][p, i, j] == pytest.approx(fd, rel=1e-6, abs=1e-6 * np.abs(gdFc[0]).max())


def test_diverged_particles_reach_no_node():
    """A particle 1e12 cells out, at +-inf or NaN adds nothing to the P2G mass: the
    stencil base is clamped to +-2^20 cells before its integer conversion, so the
    node offsets cannot overflow (brute force over the grid: the mass equals that
    of the finite particles alone)."""
    prm = orc.Params(n=3, dim=(8, 8, 8), dx=0.125)
    xs = np.array([[0.5, 0.5, 0.5], [0.3, 0.6, 0.2], [0.9, 0.1, 0.4]])
    ms = np.array([1.0, 2.0, 0.5])
    ref = orc.p2g_mass(prm, xs, ms)
    bad = np.array([[1e12, 0.5, 0.5], [-1e12, 0.5, 0.5], [np.inf, 0.5, 0.5], [0.5, -np.inf, 0.5], [0.5, 0.5, np.nan]])
    prm2 = orc.Params(n=len(xs) + len(bad), dim=(8, 8, 8), dx=0.125)
    got = orc.p2g_mass(prm2, np.concatenate([xs, bad]), np.concatenate([ms, np.ones(len(bad))]))
    assert np.array_equal(got, ref)


# ------------------------------------------------------- proxy-grid forward --
def test_proxy_forward_one_section_is_the_serial_step():
    """One section and one thread: the proxy-grid forward adds every particle into
    one grid in particle order, as the atomic P2G does serially, so with smoothing
    0 both give the same bits (pins the shared scatter, grid update and gather)."""
    prm, st, cs, dFc = tiny_scene(n=50, seed=3)
    prm.smoothing = 0.0
    orc.set_threads(1)
    try:
        a, gma, gva = orc.forward_step(prm, st, cs, dFc)
        b, gmb, gvb = orc.forward_step_proxy(prm, st, cs, dFc, 1)
    finally:
        orc.set_threads(os.cpu_count() or 1)
    for f in ("x", "v", "C", "F"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(gma, gmb) and np.array_equal(gva, gvb)


@pytest.mark.parametrize("sections", [2, 7, 50, 64])
def test_proxy_forward_sections(sections):
    """Ragged sections (50 particles into 7: 8 + 6 x 7; into 64: the first takes all
    50, as the reference's split gives it the remainder, and 63 are empty): the step equals the atomic one to summation order
    (1e-12), ignores the smoothing factor (MultiThreadForwardSimulation.cpp:220:
    the overload without it), and repeats bit for bit with a fixed section count
    whatever the thread schedule."""
    prm, st, cs, dFc = tiny_scene(n=50, seed=4)
    ref_prm = orc.Params(**{**prm.__dict__, "smoothing": 0.0})
    a, gma, gva = orc.forward_step(ref_prm, st, cs, dFc)
    b, gmb, gvb = orc.forward_step_proxy(prm, st, cs, dFc, sections)
    c, gmc, gvc = orc.forward_step_proxy(prm, st, cs, dFc, sections)
    for f in ("x", "v", "C", "F"):
        assert np.allclose(getattr(b, f), getattr(a, f), rtol=1e-12, atol=1e-14), f
        assert np.array_equal(getattr(b, f), getattr(c, f)), f
    assert np.allclose(gmb, gma, rtol=1e-12, atol=1e-18) and np.allclose(gvb, gva, rtol=1e-10, atol=1e-12)
    assert np.array_equal(gmb, gmc) and np.array_equal(gvb, gvc)
    smoothed, _, _ = orc.forward_step(prm, st, cs, dFc)
    assert not np.allclose(smoothed.F, b.F, rtol=1e-6)


def test_proxy_forward_rejects_no_sections():
    prm, st, cs, dFc = tiny_scene(n=5)
    with pytest.raises(ValueError):
        orc.forward_step_proxy(prm, st, cs, dFc, 0)
