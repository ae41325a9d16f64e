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
