"""Python driver for the CPU FP64 oracle (test infrastructure only).

Loads ``liboracle_dmpm.so`` (built from ``dmpm_oracle.c`` by ``oracle/Makefile``)
and strings its step functions into the reference's episode structure:

* ``forward_pass``  -- CompGraph::ComputeForwardPass (CompGraph.cpp:246-256),
  every layer kept (the reference's store-every-layer graph, :7-17).
* ``end_mass_loss`` -- CompGraph::EndLayerMassLoss (CompGraph.cpp:110-244).
* ``backward_pass`` -- CompGraph::ComputeBackwardPass (CompGraph.cpp:258-267).

Only tests, the smoke entry point and bench.py's CPU-baseline leg import this
module; the product never does.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

COMPAT_DUP_FBAR = 1
COMPAT_NO_1MS = 2
COMPAT_NO_DRAG_X = 4
COMPAT_DLDM_NO_FEXT = 8
COMPAT_ALL = 15


class _Params(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("dim", ctypes.c_int * 3),
        ("dx", ctypes.c_double),
        ("min_point", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("drag", ctypes.c_double),
        ("smoothing", ctypes.c_double),
        ("f_ext", ctypes.c_double * 3),
        ("compat", ctypes.c_int),
    ]


def build():
    """Compile the oracle library with its Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_dmpm.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(_Params)
        L.orc_cubic_bspline.restype = ctypes.c_double
        L.orc_cubic_bspline.argtypes = [ctypes.c_double]
        L.orc_cubic_bspline_slope.restype = ctypes.c_double
        L.orc_cubic_bspline_slope.argtypes = [ctypes.c_double]
        L.orc_svd3.restype = ctypes.c_int
        L.orc_svd3.argtypes = [dp, dp, dp, dp]
        L.orc_polar3.argtypes = [dp, dp, dp]
        L.orc_fc_energy.restype = ctypes.c_double
        L.orc_fc_energy.argtypes = [dp, ctypes.c_double, ctypes.c_double]
        L.orc_pk1.argtypes = [dp, ctypes.c_double, ctypes.c_double, dp]
        L.orc_hess_mult.argtypes = [dp, ctypes.c_double, ctypes.c_double, dp, dp, ctypes.c_int]
        L.orc_pk1_batch.argtypes = [ctypes.c_int, dp, dp, dp, dp]
        L.orc_hess_mult_batch.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_int]
        L.orc_lame.argtypes = [ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_p2g.argtypes = [pp] + [dp] * 9 + [ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_p2g_mass.argtypes = [pp, dp, dp, dp]
        L.orc_forward_step.argtypes = [pp] + [dp] * 9 + [dp, dp] + [dp] * 4
        L.orc_forward_step_proxy.argtypes = [pp, ctypes.c_int] + [dp] * 9 + [dp, dp] + [dp] * 4
        L.orc_forward_step_proxy.restype = ctypes.c_int
        L.orc_point_volumes.argtypes = [pp, dp, dp, dp]
        L.orc_end_mass_loss.restype = ctypes.c_double
        L.orc_end_mass_loss.argtypes = [pp, dp, dp, dp, dp, dp, dp]
        L.orc_backward_step.argtypes = [pp] + [dp] * 9 + [dp, dp] + [dp] * 5 + [dp, dp] + [dp] * 5
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _n(prm, *arrs):
    """Every per-particle array must hold prm.n rows (the C code trusts it)."""
    for a in arrs:
        assert np.shape(a)[0] == prm.n, (np.shape(a), prm.n)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


@dataclass
class Params:
    """Simulation parameters of one scene (CompGraph.h:55-58 semantics)."""

    n: int
    dim: tuple
    dx: float
    min_point: tuple = (0.0, 0.0, 0.0)
    dt: float = 1e-3
    drag: float = 0.0
    smoothing: float = 0.0
    f_ext: tuple = (0.0, 0.0, 0.0)
    compat: int = 0

    def c(self) -> _Params:
        p = _Params()
        p.n = int(self.n)
        for a in range(3):
            p.dim[a] = int(self.dim[a])
            p.min_point[a] = float(self.min_point[a])
            p.f_ext[a] = float(self.f_ext[a])
        p.dx = float(self.dx)
        p.dt = float(self.dt)
        p.drag = float(self.drag)
        p.smoothing = float(self.smoothing)
        p.compat = int(self.compat)
        return p

    @property
    def G(self) -> int:
        return int(self.dim[0]) * int(self.dim[1]) * int(self.dim[2])


# ---------------------------------------------------------------- scalars --
def bspline(x: float) -> float:
    return lib().orc_cubic_bspline(float(x))


def bspline_slope(x: float) -> float:
    return lib().orc_cubic_bspline_slope(float(x))


def svd3(A):
    A = _c(A).reshape(9)
    U = np.zeros(9)
    s = np.zeros(3)
    V = np.zeros(9)
    lib().orc_svd3(_p(A), _p(U), _p(s), _p(V))
    return U.reshape(3, 3), s, V.reshape(3, 3)


def polar3(A):
    A = _c(A).reshape(9)
    R = np.zeros(9)
    S = np.zeros(9)
    lib().orc_polar3(_p(A), _p(R), _p(S))
    return R.reshape(3, 3), S.reshape(3, 3)


def energy(F, lam, mu) -> float:
    return lib().orc_fc_energy(_p(_c(F).reshape(9)), float(lam), float(mu))


def pk1(F, lam, mu):
    out = np.zeros(9)
    lib().orc_pk1(_p(_c(F).reshape(9)), float(lam), float(mu), _p(out))
    return out.reshape(3, 3)


def hess_mult(F, lam, mu, dF, exact=True):
    out = np.zeros(9)
    lib().orc_hess_mult(_p(_c(F).reshape(9)), float(lam), float(mu), _p(_c(dF).reshape(9)), _p(out), int(exact))
    return out.reshape(3, 3)


def pk1_batch(F, lam, mu):
    F = _c(F).reshape(-1, 9)
    n = F.shape[0]
    out = np.zeros((n, 9))
    lib().orc_pk1_batch(n, _p(F), _p(_c(lam)), _p(_c(mu)), _p(out))
    return out.reshape(n, 3, 3)


def hess_mult_batch(F, lam, mu, dF, exact=True):
    F = _c(F).reshape(-1, 9)
    n = F.shape[0]
    out = np.zeros((n, 9))
    lib().orc_hess_mult_batch(n, _p(F), _p(_c(lam)), _p(_c(mu)), _p(_c(dF).reshape(-1, 9)), _p(out), int(exact))
    return out.reshape(n, 3, 3)


def lame(E: float, nu: float):
    lam = ctypes.c_double()
    mu = ctypes.c_double()
    lib().orc_lame(float(E), float(nu), ctypes.byref(lam), ctypes.byref(mu))
    return lam.value, mu.value


# ------------------------------------------------------------------ state --
@dataclass
class State:
    """One layer's particle state: x, v (n,3); C, F (n,3,3)."""

    x: np.ndarray
    v: np.ndarray
    C: np.ndarray
    F: np.ndarray

    def copy(self) -> "State":
        return State(self.x.copy(), self.v.copy(), self.C.copy(), self.F.copy())

    @staticmethod
    def zeros(n: int) -> "State":
        return State(np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3, 3)))


@dataclass
class Consts:
    """Per-particle constants: m, vol, lam, mu (n,)."""

    m: np.ndarray
    vol: np.ndarray
    lam: np.ndarray
    mu: np.ndarray


def p2g_mass(prm: Params, x, m):
    """P2G mass of an arbitrary point cloud (its own count, not prm.n)."""
    x = _c(x).reshape(-1, 3)
    m = _c(m)
    assert len(m) == len(x)
    out = np.zeros(prm.G)
    P = prm.c()
    P.n = len(x)
    lib().orc_p2g_mass(ctypes.byref(P), _p(_c(x)), _p(_c(m)), _p(out))
    return out


def point_volumes(prm: Params, x, m):
    x = _c(x).reshape(-1, 3)
    m = _c(m)
    assert len(m) == len(x)
    out = np.zeros(len(x))
    P = prm.c()
    P.n = len(x)
    lib().orc_point_volumes(ctypes.byref(P), _p(_c(x)), _p(_c(m)), _p(out))
    return out


def p2g(prm: Params, st: State, cs: Consts, dFc, dt=None, drag=None):
    _n(prm, st.x, st.v, st.C, st.F, dFc, cs.m, cs.vol, cs.lam, cs.mu)
    gm = np.zeros(prm.G)
    gp = np.zeros((prm.G, 3))
    P = prm.c()
    lib().orc_p2g(ctypes.byref(P), _p(_c(st.x)), _p(_c(st.v)), _p(_c(st.C)), _p(_c(st.F)), _p(_c(dFc)),
                  _p(_c(cs.m)), _p(_c(cs.vol)), _p(_c(cs.lam)), _p(_c(cs.mu)),
                  float(prm.dt if dt is None else dt), float(prm.drag if drag is None else drag), _p(gm), _p(gp))
    return gm, gp


def forward_step(prm: Params, st: State, cs: Consts, dFc):
    """ForwardTimeStep (ForwardSimulation.cpp:136-157). Returns (next, grid_m, grid_v)."""
    _n(prm, st.x, st.v, st.C, st.F, dFc, cs.m, cs.vol, cs.lam, cs.mu)
    n = prm.n
    nxt = State.zeros(n)
    gm = np.zeros(prm.G)
    gv = np.zeros((prm.G, 3))
    P = prm.c()
    lib().orc_forward_step(ctypes.byref(P), _p(_c(st.x)), _p(_c(st.v)), _p(_c(st.C)), _p(_c(st.F)),
                           _p(_c(dFc)), _p(_c(cs.m)), _p(_c(cs.vol)), _p(_c(cs.lam)), _p(_c(cs.mu)),
                           _p(gm), _p(gv), _p(nxt.x), _p(nxt.v), _p(nxt.C), _p(nxt.F))
    return nxt, gm, gv


def forward_step_proxy(prm: Params, st: State, cs: Consts, dFc, sections: int):
    """MultiThreadMPM::ForwardTimeStep (MultiThreadForwardSimulation.cpp:7-225): proxy
    grids per particle section, no smoothing. Returns (next, grid_m, grid_v)."""
    _n(prm, st.x, st.v, st.C, st.F, dFc, cs.m, cs.vol, cs.lam, cs.mu)
    n = prm.n
    nxt = State.zeros(n)
    gm = np.zeros(prm.G)
    gv = np.zeros((prm.G, 3))
    P = prm.c()
    rc = lib().orc_forward_step_proxy(ctypes.byref(P), int(sections), _p(_c(st.x)), _p(_c(st.v)), _p(_c(st.C)),
                                      _p(_c(st.F)), _p(_c(dFc)), _p(_c(cs.m)), _p(_c(cs.vol)), _p(_c(cs.lam)),
                                      _p(_c(cs.mu)), _p(gm), _p(gv), _p(nxt.x), _p(nxt.v), _p(nxt.C), _p(nxt.F))
    if rc == -1:
        raise ValueError("sections must be >= 1")
    if rc != 0:
        raise MemoryError("proxy grids could not be allocated")
    return nxt, gm, gv


def end_mass_loss(prm: Params, x, m, target):
    """EndLayerMassLoss. Returns (loss, grid_m, grid_dLdm, dLdx)."""
    _n(prm, x, m)
    assert np.size(target) == prm.G
    gm = np.zeros(prm.G)
    gd = np.zeros(prm.G)
    dLdx = np.zeros((prm.n, 3))
    P = prm.c()
    loss = lib().orc_end_mass_loss(ctypes.byref(P), _p(_c(x)), _p(_c(m)), _p(_c(target)), _p(gm), _p(gd), _p(dLdx))
    return loss, gm, gd, dLdx


@dataclass
class Grads:
    x: np.ndarray
    v: np.ndarray
    F: np.ndarray
    C: np.ndarray
    dFc: np.ndarray = field(default=None)

    @staticmethod
    def zeros(n: int) -> "Grads":
        return Grads(np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3, 3)))


def backward_step(prm: Params, st: State, cs: Consts, dFc, grid_m, grid_v, C1, g1: Grads, return_grid=False):
    """Back_Timestep (BackPropagation.cpp:6-125) for layer n from layer n+1's gradients.

    ``return_grid``: also return the adjoint grid after the transposed grid
    update, (pbar (G,3), mbar (G,)), as (grads, pbar, mbar)."""
    _n(prm, st.x, st.v, st.C, st.F, dFc, cs.m, cs.vol, cs.lam, cs.mu, C1, g1.x, g1.v, g1.F, g1.C)
    assert np.size(grid_m) == prm.G and np.size(grid_v) == 3 * prm.G
    n = prm.n
    out = Grads.zeros(n)
    wv = np.zeros((prm.G, 3))
    wm = np.zeros(prm.G)
    P = prm.c()
    lib().orc_backward_step(ctypes.byref(P), _p(_c(st.x)), _p(_c(st.v)), _p(_c(st.C)), _p(_c(st.F)),
                            _p(_c(dFc)), _p(_c(cs.m)), _p(_c(cs.vol)), _p(_c(cs.lam)), _p(_c(cs.mu)),
                            _p(_c(grid_m)), _p(_c(grid_v)), _p(_c(C1)), _p(_c(g1.x)), _p(_c(g1.v)),
                            _p(_c(g1.F)), _p(_c(g1.C)), _p(wv), _p(wm),
                            _p(out.x), _p(out.v), _p(out.F), _p(out.C), _p(out.dFc))
    if return_grid:
        return out, wv, wm
    return out


# ---------------------------------------------------------------- episode --
def forward_pass(prm: Params, st0: State, cs: Consts, controls, num_layers: int):
    """All layers and grids of one episode (CompGraph.cpp:246-256)."""
    layers = [st0.copy()]
    grids = []
    for i in range(num_layers - 1):
        nxt, gm, gv = forward_step(prm, layers[i], cs, controls[i])
        layers.append(nxt)
        grids.append((gm, gv))
    return layers, grids


def episode(prm: Params, st0: State, cs: Consts, controls, target, num_layers: int, control_layer: int = 0):
    """Forward pass, EndLayerMassLoss and backward pass to ``control_layer``.

    Returns (loss, layers, per-layer control gradients dL/d(dFc_i) for
    i in [control_layer, L-2], and the full gradients of ``control_layer``).
    """
    layers, grids = forward_pass(prm, st0, cs, controls, num_layers)
    loss, _, _, dLdx = end_mass_loss(prm, layers[-1].x, cs.m, target)
    n = prm.n
    g = Grads(dLdx, np.zeros((n, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3, 3)), np.zeros((n, 3, 3)))
    gdFc = {}
    for i in range(num_layers - 2, control_layer - 1, -1):
        gm, gv = grids[i]
        g = backward_step(prm, layers[i], cs, controls[i], gm, gv, layers[i + 1].C, g)
        gdFc[i] = g.dFc.copy()
    return loss, layers, gdFc, g


def loss_only(prm: Params, st0: State, cs: Consts, controls, target, num_layers: int) -> float:
    layers, _ = forward_pass(prm, st0, cs, controls, num_layers)
    return end_mass_loss(prm, layers[-1].x, cs.m, target)[0]


# ------------------------------------------------------- float build (timing) --
_LIB32 = None


def _float_field(t):
    if issubclass(t, ctypes.Array):
        return ctypes.c_float * t._length_ if t._type_ is ctypes.c_double else t
    return ctypes.c_float if t is ctypes.c_double else t


class _Params32(ctypes.Structure):
    _fields_ = [(name, _float_field(t)) for name, t in _Params._fields_]


def lib32():
    """``liboracle_dmpm_f32.so``: the same source with ORC_REAL=float.  Only bench.py's
    CPU-baseline leg times it; no test checks against it."""
    global _LIB32
    if _LIB32 is None:
        path = os.path.join(_HERE, "liboracle_dmpm_f32.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        fp = ctypes.POINTER(ctypes.c_float)
        pp = ctypes.POINTER(_Params32)
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_forward_step.argtypes = [pp] + [fp] * 9 + [fp, fp] + [fp] * 4
        L.orc_end_mass_loss.restype = ctypes.c_float
        L.orc_end_mass_loss.argtypes = [pp, fp, fp, fp, fp, fp, fp]
        L.orc_backward_step.argtypes = [pp] + [fp] * 9 + [fp, fp] + [fp] * 5 + [fp, fp] + [fp] * 5
        _LIB32 = L
    return _LIB32


def episode_f32(prm: Params, st0: State, cs: Consts, controls, target, num_layers: int, threads: int = 0):
    """``episode`` (forward pass keeping every layer, EndLayerMassLoss, backward pass to layer 0)
    through the float build.  Returns the loss."""
    L = lib32()
    if threads > 0:
        L.orc_set_threads(int(threads))
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
    P = _Params32()
    P.n = int(prm.n)
    for a in range(3):
        P.dim[a] = int(prm.dim[a])
        P.min_point[a] = float(prm.min_point[a])
        P.f_ext[a] = float(prm.f_ext[a])
    P.dx, P.dt, P.drag, P.smoothing, P.compat = prm.dx, prm.dt, prm.drag, prm.smoothing, int(prm.compat)
    n, G = prm.n, prm.G
    m, vol, lam, mu = f(cs.m), f(cs.vol), f(cs.lam), f(cs.mu)
    ctl = [f(c) for c in controls]
    lay = [[f(st0.x), f(st0.v), f(st0.C), f(st0.F)]]
    grids = []
    for i in range(num_layers - 1):
        x, v, C, F = lay[i]
        nx = [np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros((n, 9), np.float32),
              np.zeros((n, 9), np.float32)]
        gm, gv = np.zeros(G, np.float32), np.zeros(3 * G, np.float32)
        L.orc_forward_step(ctypes.byref(P), p(x), p(v), p(C), p(F), p(ctl[i]), p(m), p(vol), p(lam), p(mu), p(gm),
                           p(gv), *(p(a) for a in nx))
        lay.append(nx)
        grids.append((gm, gv))
    gx = np.zeros((n, 3), np.float32)
    loss = L.orc_end_mass_loss(ctypes.byref(P), p(lay[-1][0]), p(m), p(f(target)), p(np.zeros(G, np.float32)),
                               p(np.zeros(G, np.float32)), p(gx))
    g = [gx, np.zeros((n, 3), np.float32), np.zeros((n, 9), np.float32), np.zeros((n, 9), np.float32)]
    wv, wm = np.zeros(3 * G, np.float32), np.zeros(G, np.float32)
    for i in range(num_layers - 2, -1, -1):
        x, v, C, F = lay[i]
        out = [np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros((n, 9), np.float32),
               np.zeros((n, 9), np.float32), np.zeros((n, 9), np.float32)]
        gm, gv = grids[i]
        L.orc_backward_step(ctypes.byref(P), p(x), p(v), p(C), p(F), p(ctl[i]), p(m), p(vol), p(lam), p(mu), p(gm),
                            p(gv), p(lay[i + 1][2]), p(g[0]), p(g[1]), p(g[2]), p(g[3]), p(wv), p(wm),
                            *(p(a) for a in out))
        g = [out[0], out[1], out[2], out[3]]
    return float(loss)
