"""MI355X-native differentiable MLS-MPM step (the DiffMPMLib3D hot path).

Python binding of the C ABI in ``include/dmpm.h`` (``lib/libdmpm.so``, HIP
kernels for gfx950).  There is no CPU fallback: loading raises when the
library is missing or no GPU is present.

``Morph`` mirrors CompGraph (CompGraph.h:17-52): ``forward_pass`` =
ComputeForwardPass, ``end_layer_mass_loss`` = EndLayerMassLoss,
``backward_pass`` = ComputeBackwardPass, with layer states, controls and
gradients in the reference's per-particle layout (x, v (n,3); F, C, dFc
(n,3,3)).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import mpmbin, optimize, optinput, scenes  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdmpm.so")

PREC_F64 = 0
PREC_F32 = 1
TAPE_STORE_ALL = 0
TAPE_CHECKPOINT = 1
COMPAT_DUP_FBAR = 1
COMPAT_NO_1MS = 2
COMPAT_NO_DRAG_X = 4
COMPAT_DLDM_NO_FEXT = 8
COMPAT_REFERENCE = 15

DESCEND_ADAM, DESCEND_AMSGRAD, DESCEND_LION, DESCEND_GD = range(4)

OPT_GRAPHS = 1
OPT_RANGE_MAX = 2
OPT_LDS_CAP = 3
OPT_RESORT = 4
OPT_RESORT_PERMILLE = 5
OPT_OVERLAP = 6
OPT_FUSE = 7
OPT_RANGE_ORDER = 8
OPT_SEG_RESORT = 9
OPT_TILE = 10
OPT_PERSISTENT = 11
OPT_ELASTIC = 12

K_P2G, K_GRID_UPDATE, K_G2P, K_G2P_ADJ, K_GRID_ADJ, K_P2G_ADJ, K_LOSS, K_FILL = range(8)
KERNEL_NAMES = ["p2g", "grid_update", "g2p", "g2p_adj", "grid_adj", "p2g_adj", "loss", "fill", "step"]

# C-ABI symbols declared in include/dmpm.h
EXPORTS = [
    "dmpm_version", "dmpm_last_error", "dmpm_config_init", "dmpm_create", "dmpm_destroy",
    "dmpm_set_particles", "dmpm_compute_volumes", "dmpm_get_volumes", "dmpm_set_control", "dmpm_get_control",
    "dmpm_set_target_grid", "dmpm_target_from_points", "dmpm_get_target_grid", "dmpm_forward_pass",
    "dmpm_end_layer_mass_loss", "dmpm_backward_pass", "dmpm_forward_step", "dmpm_backward_step",
    "dmpm_get_layer_state", "dmpm_set_layer_state", "dmpm_get_layer_grads", "dmpm_set_layer_grads",
    "dmpm_get_control_grad",
    "dmpm_control_grad_norm", "dmpm_get_grid", "dmpm_chain_episode", "dmpm_synchronize", "dmpm_memory_bytes",
    "dmpm_profile_enable", "dmpm_profile_reset", "dmpm_profile_read", "dmpm_eval_pk1", "dmpm_eval_hessian",
    "dmpm_eval_svd", "dmpm_create_batch", "dmpm_num_instances", "dmpm_get_instance_losses",
    "dmpm_target_from_points_batch", "dmpm_set_option", "dmpm_get_option", "dmpm_autotune", "dmpm_descend",
    "dmpm_eval_energy", "dmpm_get_point_diagnostics", "dmpm_target_from_points_noalloc",
    "dmpm_phase_p2g", "dmpm_phase_grid_op", "dmpm_phase_grid_op_point", "dmpm_phase_g2p", "dmpm_phase_p_op2",
    "dmpm_phase_p2g_mass", "dmpm_phase_g2g_mass", "dmpm_sort_particles", "dmpm_get_window_stats",
    "dmpm_get_segment_window_stats",
]

# most points of one dmpm_target_from_points_noalloc call (DMPM_NOALLOC_MAX_POINTS)
NOALLOC_MAX_POINTS = 1 << 20


class Config(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32),
        ("n_particles", ctypes.c_int32),
        ("dim", ctypes.c_int32 * 3),
        ("num_layers", ctypes.c_int32),
        ("dx", ctypes.c_double),
        ("min_point", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("drag", ctypes.c_double),
        ("smoothing", ctypes.c_double),
        ("f_ext", ctypes.c_double * 3),
        ("compat", ctypes.c_int32),
        ("tape_mode", ctypes.c_int32),
        ("ckpt_stride", ctypes.c_int32),
        ("keep_control_grads", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class Instance(ctypes.Structure):
    """``dmpm_instance``: one morph of a batched context."""
    _fields_ = [
        ("n_particles", ctypes.c_int32),
        ("dim", ctypes.c_int32 * 3),
        ("dx", ctypes.c_double),
        ("min_point", ctypes.c_double * 3),
    ]


class GridDesc(ctypes.Structure):
    """``dmpm_grid_desc``: the grid of the phase calls."""
    _fields_ = [
        ("dim", ctypes.c_int32 * 3),
        ("reserved", ctypes.c_int32),
        ("dx", ctypes.c_double),
        ("min_point", ctypes.c_double * 3),
    ]


def grid_desc(dim, dx, min_point=(0.0, 0.0, 0.0)):
    g = GridDesc()
    for a in range(3):
        g.dim[a] = int(dim[a])
        g.min_point[a] = float(min_point[a])
    g.dx = float(dx)
    return g


class DmpmError(RuntimeError):
    pass


_LIB = None
_DP = ctypes.POINTER(ctypes.c_double)


def load_library(path: str = LIB_PATH):
    """Load ``libdmpm.so`` (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise DmpmError(f"HIP library not built: {path} (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    i32 = ctypes.c_int32
    sig = {
        "dmpm_version": ([], ctypes.c_int),
        "dmpm_last_error": ([], ctypes.c_char_p),
        "dmpm_config_init": ([ctypes.POINTER(Config)], ctypes.c_int),
        "dmpm_create": ([ctypes.POINTER(Config), ctypes.POINTER(vp)], ctypes.c_int),
        "dmpm_destroy": ([vp], ctypes.c_int),
        "dmpm_set_particles": ([vp] + [_DP] * 8, ctypes.c_int),
        "dmpm_compute_volumes": ([vp], ctypes.c_int),
        "dmpm_get_volumes": ([vp, _DP], ctypes.c_int),
        "dmpm_set_control": ([vp, i32, _DP], ctypes.c_int),
        "dmpm_get_control": ([vp, i32, _DP], ctypes.c_int),
        "dmpm_set_target_grid": ([vp, _DP], ctypes.c_int),
        "dmpm_target_from_points": ([vp, _DP, _DP, i32], ctypes.c_int),
        "dmpm_target_from_points_noalloc": ([vp, _DP, _DP, i32], ctypes.c_int),
        "dmpm_get_target_grid": ([vp, _DP], ctypes.c_int),
        "dmpm_forward_pass": ([vp, i32], ctypes.c_int),
        "dmpm_end_layer_mass_loss": ([vp, _DP], ctypes.c_int),
        "dmpm_backward_pass": ([vp, i32], ctypes.c_int),
        "dmpm_forward_step": ([vp, i32], ctypes.c_int),
        "dmpm_backward_step": ([vp, i32], ctypes.c_int),
        "dmpm_get_layer_state": ([vp, i32] + [_DP] * 4, ctypes.c_int),
        "dmpm_set_layer_state": ([vp, i32] + [_DP] * 4, ctypes.c_int),
        "dmpm_get_layer_grads": ([vp, i32] + [_DP] * 4, ctypes.c_int),
        "dmpm_set_layer_grads": ([vp, i32] + [_DP] * 4, ctypes.c_int),
        "dmpm_get_control_grad": ([vp, i32, _DP], ctypes.c_int),
        "dmpm_control_grad_norm": ([vp, i32, _DP], ctypes.c_int),
        "dmpm_get_grid": ([vp, i32, _DP, _DP], ctypes.c_int),
        "dmpm_chain_episode": ([vp], ctypes.c_int),
        "dmpm_synchronize": ([vp], ctypes.c_int),
        "dmpm_memory_bytes": ([vp, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
        "dmpm_profile_enable": ([vp, i32], ctypes.c_int),
        "dmpm_profile_reset": ([vp], ctypes.c_int),
        "dmpm_profile_read": ([vp, i32, _DP, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "dmpm_eval_pk1": ([i32, _DP, _DP, _DP, _DP], ctypes.c_int),
        "dmpm_eval_hessian": ([i32, _DP, _DP, _DP, _DP, _DP, i32], ctypes.c_int),
        "dmpm_eval_svd": ([i32, _DP, _DP, _DP, _DP], ctypes.c_int),
        "dmpm_eval_energy": ([i32, _DP, _DP, _DP, _DP], ctypes.c_int),
        "dmpm_get_point_diagnostics": ([vp, i32, _DP, _DP], ctypes.c_int),
        "dmpm_create_batch": ([ctypes.POINTER(Config), i32, ctypes.POINTER(Instance), ctypes.POINTER(vp)], ctypes.c_int),
        "dmpm_num_instances": ([vp, ctypes.POINTER(i32)], ctypes.c_int),
        "dmpm_get_instance_losses": ([vp, _DP], ctypes.c_int),
        "dmpm_target_from_points_batch": ([vp, _DP, _DP, ctypes.POINTER(i32)], ctypes.c_int),
        "dmpm_set_option": ([vp, i32, ctypes.c_int64], ctypes.c_int),
        "dmpm_get_option": ([vp, i32, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "dmpm_autotune": ([vp, _DP], ctypes.c_int),
        "dmpm_descend": ([vp, i32, i32] + [ctypes.c_float] * 5 + [i32], ctypes.c_int),
        "dmpm_phase_p2g": ([ctypes.POINTER(GridDesc), i32] + [_DP] * 8 + [ctypes.c_double] * 2 + [_DP] * 2, ctypes.c_int),
        "dmpm_phase_grid_op": ([ctypes.POINTER(GridDesc), _DP, _DP, ctypes.c_double, _DP, _DP], ctypes.c_int),
        "dmpm_phase_grid_op_point": ([ctypes.POINTER(GridDesc), _DP, _DP, ctypes.c_double, _DP, ctypes.c_double, _DP],
                                     ctypes.c_int),
        "dmpm_phase_g2p": ([ctypes.POINTER(GridDesc), i32, _DP, _DP, _DP, _DP], ctypes.c_int),
        "dmpm_phase_p_op2": ([i32] + [_DP] * 5 + [ctypes.c_double] * 2 + [_DP] * 2, ctypes.c_int),
        "dmpm_sort_particles": ([vp], ctypes.c_int),
        "dmpm_get_window_stats": ([vp, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "dmpm_get_segment_window_stats": ([vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "dmpm_phase_p2g_mass": ([ctypes.POINTER(GridDesc), i32, _DP, _DP, _DP], ctypes.c_int),
        "dmpm_phase_g2g_mass": ([ctypes.POINTER(GridDesc), _DP, ctypes.POINTER(GridDesc), _DP], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = L
    return L


def _check(rc: int):
    if rc != 0:
        msg = _LIB.dmpm_last_error().decode(errors="replace")
        raise DmpmError(f"dmpm error {rc}: {msg}")


def _in(a, size=None, name="array"):
    """Contiguous float64 copy of ``a`` and its pointer; ``size``: the number of values
    the C side reads from the pointer (ValueError when ``a`` holds another count)."""
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, dtype=np.float64)
    if size is not None and arr.size != size:
        raise ValueError(f"{name}: expected {size} values, got {arr.size} (shape {arr.shape})")
    return arr, arr.ctypes.data_as(_DP)


def _ptr(arr):
    return None if arr is None else arr.ctypes.data_as(_DP)


# ----------------------------------------------------- batched evaluators --
def eval_pk1(F, lam, mu):
    """PK_FixedCorotatedElasticity on the GPU for a batch of (n,3,3) F."""
    L = load_library()
    F, pF = _in(np.asarray(F).reshape(-1, 9))
    n = F.shape[0]
    lam, pl = _in(np.broadcast_to(np.asarray(lam, float), (n,)))
    mu, pm = _in(np.broadcast_to(np.asarray(mu, float), (n,)))
    out = np.zeros((n, 9))
    _check(L.dmpm_eval_pk1(n, pF, pl, pm, _ptr(out)))
    return out.reshape(n, 3, 3)


def eval_hessian(F, lam, mu, dF, exact=True):
    """d2_FCE_psi_dF2_mult_by_dF on the GPU for a batch."""
    L = load_library()
    F, pF = _in(np.asarray(F).reshape(-1, 9))
    n = F.shape[0]
    lam, pl = _in(np.broadcast_to(np.asarray(lam, float), (n,)))
    mu, pm = _in(np.broadcast_to(np.asarray(mu, float), (n,)))
    dF, pd = _in(dF, 9 * n, "dF")
    out = np.zeros((n, 9))
    _check(L.dmpm_eval_hessian(n, pF, pl, pm, pd, _ptr(out), int(exact)))
    return out.reshape(n, 3, 3)


def eval_svd(F):
    """3x3 SVD on the GPU: (U, sigma, V) with U, V rotations."""
    L = load_library()
    F, pF = _in(np.asarray(F).reshape(-1, 9))
    n = F.shape[0]
    U = np.zeros((n, 9))
    s = np.zeros((n, 3))
    V = np.zeros((n, 9))
    _check(L.dmpm_eval_svd(n, pF, _ptr(U), _ptr(s), _ptr(V)))
    return U.reshape(n, 3, 3), s, V.reshape(n, 3, 3)


def eval_energy(F, lam, mu):
    """FixedCorotatedElasticity (Elasticity.cpp:74-86) on the GPU for a batch of (n,3,3) F."""
    L = load_library()
    F, pF = _in(np.asarray(F).reshape(-1, 9))
    n = F.shape[0]
    lam, pl = _in(np.broadcast_to(np.asarray(lam, float), (n,)))
    mu, pm = _in(np.broadcast_to(np.asarray(mu, float), (n,)))
    out = np.zeros(n)
    _check(L.dmpm_eval_energy(n, pF, pl, pm, _ptr(out)))
    return out


# ------------------------------------------------------------------ phases --
# The reference's ForwardTimeStep phases one call each (SingleThreadMPM,
# ForwardSimulation.h:23-45), FP64 on the GPU over host arrays; grids are
# (dim, dx, min_point) with masses (G,) and momenta / velocities (G, 3).
def _grid_in(g):
    return ctypes.byref(grid_desc(*g))


def _G(g):
    return int(g[0][0]) * int(g[0][1]) * int(g[0][2])


def phase_p2g(grid, x, v, C, F, dFc, P, m, vol, dt, drag, grid_m=None, grid_p=None):
    """P2G (ForwardSimulation.cpp:11-51) added into grid_m, grid_p (zero when None); returns them."""
    L = load_library()
    n = np.asarray(m).size
    G = _G(grid)
    gm = np.zeros(G) if grid_m is None else np.array(grid_m, dtype=np.float64).reshape(G)
    gp = np.zeros((G, 3)) if grid_p is None else np.array(grid_p, dtype=np.float64).reshape(G, 3)
    arrs = [_in(a, sz, nm) for a, sz, nm in zip((x, v, C, F, dFc, P, m, vol), (3 * n, 3 * n, 9 * n, 9 * n, 9 * n, 9 * n, n, n),
                                                 ("x", "v", "C", "F", "dFc", "P", "m", "vol"))]
    _check(L.dmpm_phase_p2g(_grid_in(grid), n, *[p for _, p in arrs], float(dt), float(drag), _ptr(gm), _ptr(gp)))
    return gm, gp


def phase_grid_op(grid, grid_m, grid_p, dt, f_ext, grid_v=None, gravity_mag=None):
    """G_op (ForwardSimulation.cpp:53-58,184-200); with ``gravity_mag`` the gravity-point form
    (:202-224), f_ext then being the point.  Nodes at or below the mass threshold keep grid_v."""
    L = load_library()
    G = _G(grid)
    gv = np.zeros((G, 3)) if grid_v is None else np.array(grid_v, dtype=np.float64).reshape(G, 3)
    _, pm = _in(grid_m, G, "grid_m")
    _, pp = _in(grid_p, 3 * G, "grid_p")
    _, pf = _in(f_ext, 3, "f_ext")
    if gravity_mag is None:
        _check(L.dmpm_phase_grid_op(_grid_in(grid), pm, pp, float(dt), pf, _ptr(gv)))
    else:
        _check(L.dmpm_phase_grid_op_point(_grid_in(grid), pm, pp, float(dt), pf, float(gravity_mag), _ptr(gv)))
    return gv


def phase_g2p(grid, x, grid_v):
    """G2P (ForwardSimulation.cpp:60-106): (v', C') per particle."""
    L = load_library()
    xa, px = _in(x, None, "x")
    n = xa.size // 3
    _, pv = _in(grid_v, 3 * _G(grid), "grid_v")
    vn = np.zeros((n, 3))
    Cn = np.zeros((n, 3, 3))
    _check(L.dmpm_phase_g2p(_grid_in(grid), n, px, pv, _ptr(vn), _ptr(Cn)))
    return vn, Cn


def phase_p_op2(x, F, dFc, v_next, C_next, smoothing, dt):
    """P_op_2 (ForwardSimulation.cpp:108-124): (x', F')."""
    L = load_library()
    xa, px = _in(x, None, "x")
    n = xa.size // 3
    arrs = [_in(a, sz, nm) for a, sz, nm in zip((F, dFc, v_next, C_next), (9 * n, 9 * n, 3 * n, 9 * n),
                                                 ("F", "dFc", "v_next", "C_next"))]
    xn = np.zeros((n, 3))
    Fn = np.zeros((n, 3, 3))
    _check(L.dmpm_phase_p_op2(n, px, *[p for _, p in arrs], float(smoothing), float(dt), _ptr(xn), _ptr(Fn)))
    return xn, Fn


def phase_p2g_mass(grid, x, m, grid_m=None):
    """P2G_Mass's scatter (ForwardSimulation.cpp:292-327) added into grid_m."""
    L = load_library()
    ma, pm = _in(np.asarray(m, dtype=np.float64).reshape(-1), None, "m")
    _, px = _in(x, 3 * ma.size, "x")
    G = _G(grid)
    gm = np.zeros(G) if grid_m is None else np.array(grid_m, dtype=np.float64).reshape(G)
    _check(L.dmpm_phase_p2g_mass(_grid_in(grid), ma.size, px, pm, _ptr(gm)))
    return gm


def phase_g2g_mass(grid1, grid1_m, grid2):
    """G2G_Mass (ForwardSimulation.cpp:329-367): grid-1 masses gathered at grid 2's nodes."""
    L = load_library()
    _, p1 = _in(grid1_m, _G(grid1), "grid1_m")
    out = np.zeros(_G(grid2))
    _check(L.dmpm_phase_g2g_mass(_grid_in(grid1), p1, _grid_in(grid2), _ptr(out)))
    return out


# ------------------------------------------------------------------ Morph --
def _make_config(L, num_layers, dt, drag, smoothing, f_ext, precision, compat, tape, ckpt_stride,
                 keep_control_grads, device):
    cfg = Config()
    _check(L.dmpm_config_init(ctypes.byref(cfg)))
    cfg.precision = int(precision)
    for a in range(3):
        cfg.f_ext[a] = float(f_ext[a])
    cfg.num_layers = int(num_layers)
    cfg.dt = float(dt)
    cfg.drag = float(drag)
    cfg.smoothing = float(smoothing)
    cfg.compat = int(compat)
    cfg.tape_mode = int(tape)
    cfg.ckpt_stride = int(ckpt_stride)
    cfg.keep_control_grads = int(bool(keep_control_grads))
    cfg.device = int(device)
    return cfg


class Morph:
    """One morph (CompGraph) on one GPU, backed by a ``dmpm_ctx``."""

    def __init__(self, n, dim, dx, num_layers, *, min_point=(0.0, 0.0, 0.0), dt=1.0 / 120.0, drag=0.5,
                 smoothing=0.1, f_ext=(0.0, 0.0, 0.0), precision=PREC_F64, compat=0, tape=TAPE_STORE_ALL,
                 ckpt_stride=0, keep_control_grads=False, device=0):
        L = load_library()
        cfg = _make_config(L, num_layers, dt, drag, smoothing, f_ext, precision, compat, tape, ckpt_stride,
                           keep_control_grads, device)
        cfg.n_particles = int(n)
        for a in range(3):
            cfg.dim[a] = int(dim[a])
            cfg.min_point[a] = float(min_point[a])
        cfg.dx = float(dx)
        self.cfg = cfg
        self.n = int(n)
        self.G = int(dim[0]) * int(dim[1]) * int(dim[2])
        self.num_layers = int(num_layers)
        h = ctypes.c_void_p()
        _check(L.dmpm_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._L = L

    @classmethod
    def from_scene(cls, sc, **kw):
        kw.setdefault("dt", sc.dt)
        kw.setdefault("drag", sc.drag)
        kw.setdefault("smoothing", sc.smoothing)
        kw.setdefault("f_ext", sc.f_ext)
        kw.setdefault("min_point", sc.min_point)
        return cls(sc.n, sc.dim, sc.dx, kw.pop("num_layers", sc.num_layers), **kw)

    def close(self):
        if getattr(self, "_h", None):
            self._L.dmpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- inputs ---------------------------------------------------------
    def set_particles(self, x, v, F, C, m, vol, lam, mu):
        n = self.n
        sizes = (3 * n, 3 * n, 9 * n, 9 * n, n, n, n, n)
        names = ("x", "v", "F", "C", "m", "vol", "lam", "mu")
        arrs = [_in(a, sz, nm) for a, sz, nm in zip((x, v, F, C, m, vol, lam, mu), sizes, names)]
        _check(self._L.dmpm_set_particles(self._h, *[p for _, p in arrs]))

    def compute_volumes(self):
        _check(self._L.dmpm_compute_volumes(self._h))

    def volumes(self):
        out = np.zeros(self.n)
        _check(self._L.dmpm_get_volumes(self._h, _ptr(out)))
        return out

    def set_control(self, layer, dFc):
        a, p = _in(dFc, 9 * self.n, "dFc")
        _check(self._L.dmpm_set_control(self._h, int(layer), p))

    def control(self, layer):
        out = np.zeros((self.n, 3, 3))
        _check(self._L.dmpm_get_control(self._h, int(layer), _ptr(out)))
        return out

    def set_target_grid(self, mass):
        a, p = _in(mass, self.G, "target mass")
        _check(self._L.dmpm_set_target_grid(self._h, p))

    def target_from_points(self, x, m):
        ma, pm = _in(np.asarray(m, dtype=np.float64).reshape(-1), None, "m")
        xa, px = _in(x, 3 * ma.size, "x")
        _check(self._L.dmpm_target_from_points(self._h, px, pm, int(ma.size)))

    def target_from_points_noalloc(self, x, m):
        """``target_from_points`` staged in the library's fixed device array: allocates
        nothing, at most NOALLOC_MAX_POINTS points."""
        ma, pm = _in(np.asarray(m, dtype=np.float64).reshape(-1), None, "m")
        xa, px = _in(x, 3 * ma.size, "x")
        _check(self._L.dmpm_target_from_points_noalloc(self._h, px, pm, int(ma.size)))

    def target_grid(self):
        out = np.zeros(self.G)
        _check(self._L.dmpm_get_target_grid(self._h, _ptr(out)))
        return out

    # -- passes ---------------------------------------------------------
    def forward_pass(self, start_layer=0):
        _check(self._L.dmpm_forward_pass(self._h, int(start_layer)))

    def end_layer_mass_loss(self):
        loss = ctypes.c_double()
        _check(self._L.dmpm_end_layer_mass_loss(self._h, ctypes.byref(loss)))
        return loss.value

    def backward_pass(self, control_layer=0):
        _check(self._L.dmpm_backward_pass(self._h, int(control_layer)))

    def forward_step(self, layer):
        _check(self._L.dmpm_forward_step(self._h, int(layer)))

    def backward_step(self, layer):
        _check(self._L.dmpm_backward_step(self._h, int(layer)))

    def chain_episode(self):
        _check(self._L.dmpm_chain_episode(self._h))

    def synchronize(self):
        _check(self._L.dmpm_synchronize(self._h))

    # -- outputs --------------------------------------------------------
    def layer_state(self, layer):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        F = np.zeros((self.n, 3, 3))
        C = np.zeros((self.n, 3, 3))
        _check(self._L.dmpm_get_layer_state(self._h, int(layer), _ptr(x), _ptr(v), _ptr(F), _ptr(C)))
        return x, v, F, C

    def point_diagnostics(self, layer):
        """(J, psi) per particle of a layer: det(F + dFc) and the fixed-corotated energy at
        F + dFc (PointCloud::GetPointDeterminants / GetPointElasticEnergies,
        PointCloud.cpp:441-467), FP64 on the GPU."""
        J = np.zeros(self.n)
        psi = np.zeros(self.n)
        _check(self._L.dmpm_get_point_diagnostics(self._h, int(layer), _ptr(J), _ptr(psi)))
        return J, psi

    def set_layer_state(self, layer, x=None, v=None, F=None, C=None):
        n = self.n
        arrs = [_in(a, sz, nm) for a, sz, nm in zip((x, v, F, C), (3 * n, 3 * n, 9 * n, 9 * n), ("x", "v", "F", "C"))]
        _check(self._L.dmpm_set_layer_state(self._h, int(layer), *[p for _, p in arrs]))

    def layer_grads(self, layer):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        F = np.zeros((self.n, 3, 3))
        C = np.zeros((self.n, 3, 3))
        _check(self._L.dmpm_get_layer_grads(self._h, int(layer), _ptr(x), _ptr(v), _ptr(F), _ptr(C)))
        return x, v, F, C

    def set_layer_grads(self, layer, dLdx=None, dLdv=None, dLdF=None, dLdC=None):
        n = self.n
        arrs = [_in(a, sz, nm) for a, sz, nm in
                zip((dLdx, dLdv, dLdF, dLdC), (3 * n, 3 * n, 9 * n, 9 * n), ("dLdx", "dLdv", "dLdF", "dLdC"))]
        _check(self._L.dmpm_set_layer_grads(self._h, int(layer), *[p for _, p in arrs]))

    def control_grad(self, layer):
        out = np.zeros((self.n, 3, 3))
        _check(self._L.dmpm_get_control_grad(self._h, int(layer), _ptr(out)))
        return out

    def control_grad_norm(self, layer):
        out = ctypes.c_double()
        _check(self._L.dmpm_control_grad_norm(self._h, int(layer), ctypes.byref(out)))
        return out.value

    def descend(self, layer, rule, alpha, gradient_norm, beta1=0.9, beta2=0.999, epsilon=1e-5, timestep=1):
        """PointCloud::Descend_* on the device for the control of ``layer`` (rule: DESCEND_*)."""
        _check(self._L.dmpm_descend(self._h, int(layer), int(rule), float(alpha), float(gradient_norm), float(beta1),
                                    float(beta2), float(epsilon), int(timestep)))

    def grid(self, step):
        m = np.zeros(self.G)
        v = np.zeros((self.G, 3))
        _check(self._L.dmpm_get_grid(self._h, int(step), _ptr(m), _ptr(v)))
        return m, v

    def memory_bytes(self):
        b = ctypes.c_uint64()
        _check(self._L.dmpm_memory_bytes(self._h, ctypes.byref(b)))
        return int(b.value)

    # -- execution options ----------------------------------------------
    def set_option(self, option, value):
        _check(self._L.dmpm_set_option(self._h, int(option), int(value)))

    def option(self, option):
        v = ctypes.c_int64()
        _check(self._L.dmpm_get_option(self._h, int(option), ctypes.byref(v)))
        return int(v.value)

    def set_graphs(self, on=True):
        """Capture each pass into a HIP graph on first use and replay it afterwards."""
        self.set_option(OPT_GRAPHS, int(bool(on)))

    def sort_particles(self):
        """FP32: re-order the particles on the device by grid block (API order unchanged)."""
        _check(self._L.dmpm_sort_particles(self._h))

    def window_stats(self):
        """FP32: P2G workgroups on the fixed / general / HBM paths in the last forward pass that
        reached the loss, and the number of device sorts so far, as a dict."""
        a = (ctypes.c_int64 * 4)()
        _check(self._L.dmpm_get_window_stats(self._h, a))
        return {"fixed": int(a[0]), "general": int(a[1]), "hbm": int(a[2]), "sorts": int(a[3])}

    def segment_window_stats(self, segment):
        """FP32 checkpoint tape: one segment's P2G window modes in the last forward pass that
        reached the loss, whether it holds an order of its own, and the segment sorts so far."""
        a = (ctypes.c_int64 * 5)()
        _check(self._L.dmpm_get_segment_window_stats(self._h, int(segment), a))
        return {"fixed": int(a[0]), "general": int(a[1]), "hbm": int(a[2]), "own_order": bool(a[3]), "seg_sorts": int(a[4])}

    def autotune(self):
        """FP32: pick the fastest range size / LDS capacity; returns its step time (ms).
        Invalidates the layers after 0."""
        ms = ctypes.c_double()
        _check(self._L.dmpm_autotune(self._h, ctypes.byref(ms)))
        return ms.value

    # -- profiling ------------------------------------------------------
    def profile(self, on=True):
        _check(self._L.dmpm_profile_enable(self._h, int(bool(on))))

    def profile_reset(self):
        _check(self._L.dmpm_profile_reset(self._h))

    def profile_read(self):
        out = {}
        for k, name in enumerate(KERNEL_NAMES):
            ms = ctypes.c_double()
            n = ctypes.c_int64()
            _check(self._L.dmpm_profile_read(self._h, k, ctypes.byref(ms), ctypes.byref(n)))
            out[name] = (ms.value, int(n.value))
        return out


def morph_from_scene(sc, dFc=None, **kw):
    """Build a ``Morph`` for a ``scenes.Scene``: layer 0 particles, reference-formula
    volumes (computed on the GPU), controls and the target grid (GPU P2G)."""
    mo = Morph.from_scene(sc, **kw)
    x, v, C, F = sc.state0()
    mo.set_particles(x, v, F, C, sc.m, None, sc.lam, sc.mu)
    mo.compute_volumes()
    if dFc is not None:
        mo.set_control(-1, dFc)
    mo.target_from_points(sc.target_x, sc.target_m)
    return mo


# ------------------------------------------------------------- MorphBatch --
class MorphBatch(Morph):
    """Independent morphs advanced together (FP32 path): every launch covers
    all instances.  Particle arrays are the instances' particles concatenated
    in instance order and grid arrays their grids concatenated; ``split`` and
    ``split_grid`` cut them back per instance.  ``end_layer_mass_loss`` is the
    sum of the instance losses; ``instance_losses`` gives each."""

    def __init__(self, instances, num_layers, *, dt=1.0 / 120.0, drag=0.5, smoothing=0.1, f_ext=(0.0, 0.0, 0.0),
                 compat=0, tape=TAPE_STORE_ALL, ckpt_stride=0, keep_control_grads=False, device=0):
        """``instances``: sequence of (n, dim, dx, min_point)."""
        L = load_library()
        cfg = _make_config(L, num_layers, dt, drag, smoothing, f_ext, PREC_F32, compat, tape, ckpt_stride,
                           keep_control_grads, device)
        arr = (Instance * len(instances))()
        for i, (n, dim, dx, mp) in enumerate(instances):
            arr[i].n_particles = int(n)
            arr[i].dx = float(dx)
            for a in range(3):
                arr[i].dim[a] = int(dim[a])
                arr[i].min_point[a] = float(mp[a])
        self.cfg = cfg
        self.counts = [int(t[0]) for t in instances]
        self.gcounts = [int(t[1][0]) * int(t[1][1]) * int(t[1][2]) for t in instances]
        self.B = len(instances)
        self.n = sum(self.counts)
        self.G = sum(self.gcounts)
        self.num_layers = int(num_layers)
        h = ctypes.c_void_p()
        _check(L.dmpm_create_batch(ctypes.byref(cfg), len(instances), arr, ctypes.byref(h)))
        self._h = h
        self._L = L

    @classmethod
    def from_scenes(cls, scs, **kw):
        s0 = scs[0]
        for k in ("dt", "drag", "smoothing", "f_ext"):
            kw.setdefault(k, getattr(s0, k))
        inst = [(sc.n, sc.dim, sc.dx, sc.min_point) for sc in scs]
        return cls(inst, kw.pop("num_layers", s0.num_layers), **kw)

    def split(self, a):
        out, p = [], 0
        for n in self.counts:
            out.append(a[p:p + n])
            p += n
        return out

    def split_grid(self, a):
        out, q = [], 0
        for g in self.gcounts:
            out.append(a[q:q + g])
            q += g
        return out

    def target_from_points(self, xs, ms):
        """Target grids from one point cloud per instance (lists of (n_i,3) and (n_i,))."""
        if len(xs) != self.B or len(ms) != self.B:
            raise ValueError(f"target point clouds: expected {self.B} instances, got {len(xs)} and {len(ms)}")
        counts = (ctypes.c_int32 * self.B)(*[np.asarray(m).size for m in ms])
        ma, pm = _in(np.concatenate([np.asarray(m, float).reshape(-1) for m in ms]), None, "m")
        xa, px = _in(np.concatenate([np.asarray(x, float).reshape(-1) for x in xs]), 3 * ma.size, "x")
        _check(self._L.dmpm_target_from_points_batch(self._h, px, pm, counts))

    def instance_losses(self):
        out = np.zeros(self.B)
        _check(self._L.dmpm_get_instance_losses(self._h, _ptr(out)))
        return out


def batch_from_scenes(scs, dFcs=None, **kw):
    """``MorphBatch`` over ``scenes.Scene`` instances: layer 0 particles, GPU volumes,
    controls (one (n_i,3,3) array per instance, or None) and the target grids."""
    mb = MorphBatch.from_scenes(scs, **kw)
    st = [sc.state0() for sc in scs]
    cat = lambda k: np.concatenate([s[k] for s in st])  # noqa: E731
    mb.set_particles(cat(0), cat(1), cat(3), cat(2), np.concatenate([sc.m for sc in scs]), None,
                     np.concatenate([sc.lam for sc in scs]), np.concatenate([sc.mu for sc in scs]))
    mb.compute_volumes()
    if dFcs is not None:
        mb.set_control(-1, np.concatenate([np.asarray(d).reshape(-1, 3, 3) for d in dFcs]))
    mb.target_from_points([sc.target_x for sc in scs], [sc.target_m for sc in scs])
    return mb

