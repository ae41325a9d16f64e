"""``OptInput``: the reference's optimisation-input record and its JSON file.

Fields, types and defaults follow ``OptInput`` (MPMPolyscope.h:16-52).  The
file is cereal's JSON archive of that record (``serialize``,
MPMPolyscope.h:57-82; written and read at DiffMPMAnimator3D.cpp:252-279): a
top-level object whose one member ``"value0"`` holds the fields by name, in
declaration order.  A ``Vec3`` field goes through the Eigen vector serializer
(cereal_eigen.h:27-38), which writes the row count and then the coefficients
as unnamed values, so cereal names them ``"value0"`` (the count, 3) to
``"value3"``.  Floats are the record's float32 values written as doubles.

``grid`` is LoadScene's grid construction (MPMPolyscope.cpp:92-98), including
its use of the x extent for all three axes.  ``morph_from_opt_input`` builds a
device ``Morph`` the way LoadCompGraph and OptimizeDefGradControlSequence set
up the computation graph (MPMPolyscope.cpp:124-167, CompGraph.cpp:280-330):
layer 0 from the input point cloud, every control equal to its dFc
(SetUpCompGraph copies layer 0 into every layer), ``num_timesteps`` layers.
Only ``.mpmbin`` point clouds load here; the reference's ``.obj`` branch needs
mesh voxelisation, which is not part of this library.
"""
from __future__ import annotations

import dataclasses
import json
import math
import os

import numpy as np

from . import mpmbin

_STR, _FLOAT, _INT, _VEC3 = "str", "float", "int", "vec3"

# (name, kind), in the order of OptInput::serialize (MPMPolyscope.h:60-80)
FIELDS = [
    ("mpm_input_mesh_path", _STR),
    ("mpm_target_mesh_path", _STR),
    ("grid_dx", _FLOAT),
    ("points_per_cell_cuberoot", _INT),
    ("grid_min_point", _VEC3),
    ("grid_max_point", _VEC3),
    ("lam", _FLOAT),
    ("mu", _FLOAT),
    ("p_density", _FLOAT),
    ("dt", _FLOAT),
    ("drag", _FLOAT),
    ("f_ext", _VEC3),
    ("num_animations", _INT),
    ("num_timesteps", _INT),
    ("control_stride", _INT),
    ("max_gd_iters", _INT),
    ("max_ls_iters", _INT),
    ("initial_alpha", _FLOAT),
    ("gd_tol", _FLOAT),
    ("smoothing_factor", _FLOAT),
]


def _f32(x) -> float:
    return float(np.float32(x))


@dataclasses.dataclass
class OptInput:
    """The reference's defaults (MPMPolyscope.h:18-52), float fields rounded to float32."""
    mpm_input_mesh_path: str = "experiments\\big_sca_demo\\input\\bob.obj"
    mpm_target_mesh_path: str = "experiments\\big_sca_demo\\input\\spot.obj"
    grid_dx: float = 1.0
    points_per_cell_cuberoot: int = 2
    grid_min_point: tuple = (-16.0, -16.0, -16.0)
    grid_max_point: tuple = (16.0, 16.0, 16.0)
    lam: float = _f32(58333.0)
    mu: float = _f32(38888.9)
    p_density: float = 40.0
    dt: float = _f32(1.0 / 120.0)
    drag: float = 0.5
    f_ext: tuple = (0.0, 0.0, 0.0)
    num_animations: int = 40
    num_timesteps: int = 30
    control_stride: int = 10
    max_gd_iters: int = 25
    max_ls_iters: int = 15
    initial_alpha: float = _f32(0.1)
    gd_tol: float = _f32(1e-3)
    smoothing_factor: float = _f32(0.1)

    def __post_init__(self):
        for name, kind in FIELDS:
            v = getattr(self, name)
            if kind == _FLOAT:
                setattr(self, name, _f32(v))
            elif kind == _INT:
                setattr(self, name, int(v))
            elif kind == _VEC3:
                setattr(self, name, tuple(_f32(c) for c in v))
            else:
                setattr(self, name, str(v))


def to_json(opt: OptInput) -> str:
    """The cereal JSON text of ``opt`` (4-space indentation, as cereal writes it)."""
    body = {}
    for name, kind in FIELDS:
        v = getattr(opt, name)
        if kind == _VEC3:
            body[name] = {"value0": 3, "value1": v[0], "value2": v[1], "value3": v[2]}
        else:
            body[name] = v
    return json.dumps({"value0": body}, indent=4)


def from_json(text: str) -> OptInput:
    """Parse cereal's JSON of an OptInput.  Every field must be present (cereal's
    named reads throw on a missing name); a Vec3 must hold 3 coefficients."""
    doc = json.loads(text)
    if not isinstance(doc, dict) or "value0" not in doc or not isinstance(doc["value0"], dict):
        raise ValueError("OptInput JSON: no top-level \"value0\" object")
    body = doc["value0"]
    kw = {}
    for name, kind in FIELDS:
        if name not in body:
            raise ValueError(f"OptInput JSON: missing field {name!r}")
        v = body[name]
        if kind == _VEC3:
            if isinstance(v, dict):
                vals = [v[k] for k in sorted(v, key=lambda s: int(s[5:]) if s.startswith("value") else -1)]
            else:
                vals = list(v)
            if len(vals) != 4 or int(vals[0]) != 3:
                raise ValueError(f"OptInput JSON: {name!r} is not a 3-vector (rows, x, y, z)")
            kw[name] = tuple(float(c) for c in vals[1:])
        elif kind == _INT:
            if isinstance(v, bool) or not isinstance(v, int):
                raise ValueError(f"OptInput JSON: {name!r} is not an integer")
            kw[name] = v
        elif kind == _FLOAT:
            if isinstance(v, bool) or not isinstance(v, (int, float)):
                raise ValueError(f"OptInput JSON: {name!r} is not a number")
            kw[name] = float(v)
        else:
            if not isinstance(v, str):
                raise ValueError(f"OptInput JSON: {name!r} is not a string")
            kw[name] = v
    return OptInput(**kw)


def write(path, opt: OptInput) -> None:
    with open(path, "w") as fh:
        fh.write(to_json(opt))


def read(path) -> OptInput:
    with open(path) as fh:
        return from_json(fh.read())


def grid(opt: OptInput):
    """LoadScene's grid (MPMPolyscope.cpp:92-98): ``ceil((max.x - min.x) / dx)``
    nodes on every axis (the reference uses the x extent for all three), node
    spacing ``grid_dx``, origin ``grid_min_point``.  Returns (dims, dx, min_point)."""
    ext = np.float32(np.float32(opt.grid_max_point[0]) - np.float32(opt.grid_min_point[0]))
    n = int(math.ceil(float(np.float32(ext / np.float32(opt.grid_dx)))))
    if n <= 0:
        raise ValueError("OptInput: empty grid (grid_max_point.x <= grid_min_point.x)")
    return (n, n, n), opt.grid_dx, tuple(opt.grid_min_point)


def _resolve(path, base):
    p = path.replace("\\", os.sep)
    return p if os.path.isabs(p) or base is None else os.path.join(base, p)


def load_point_cloud(path, base=None) -> dict:
    """The fields of a ``.mpmbin`` point cloud (LoadScene's .mpmbin branch,
    MPMPolyscope.cpp:80-85).  A mesh path (.obj) raises ``ValueError``."""
    p = _resolve(path, base)
    if os.path.splitext(p)[1].lower() != ".mpmbin":
        raise ValueError(f"OptInput: {path!r} is not a .mpmbin point cloud (mesh import is not part of this library)")
    return mpmbin.read(p)


def morph_from_opt_input(opt: OptInput, base=None, target_x=None, target_m=None, **kw):
    """A device ``Morph`` set up from ``opt``: grid from ``grid(opt)``, ``num_timesteps``
    layers, dt, drag, f_ext and smoothing from the record, layer 0 and every
    layer's control from the input ``.mpmbin`` (volumes as stored, like
    LoadScene's .mpmbin branch).  The target grid is the P2G mass of
    ``target_x``/``target_m`` when given, else of the target ``.mpmbin``
    (MPMPolyscope.cpp:156-160).  ``base``: directory for relative paths.
    Other keyword arguments go to ``Morph`` (precision, tape, ...)."""
    from . import Morph  # the package's device context

    pc = load_point_cloud(opt.mpm_input_mesh_path, base)
    dims, dx, mn = grid(opt)
    n = len(pc["m"])
    mo = Morph(n, dims, dx, opt.num_timesteps, min_point=mn, dt=opt.dt, drag=opt.drag, smoothing=opt.smoothing_factor,
               f_ext=opt.f_ext, **kw)
    f64 = lambda a: np.asarray(a, dtype=np.float64)  # noqa: E731
    mo.set_particles(f64(pc["x"]), f64(pc["v"]), f64(pc["F"]), f64(pc["C"]), f64(pc["m"]), f64(pc["vol"]),
                     f64(pc["lam"]), f64(pc["mu"]))
    mo.set_control(-1, f64(pc["dFc"]))
    if target_x is None:
        tg = load_point_cloud(opt.mpm_target_mesh_path, base)
        target_x, target_m = tg["x"], tg["m"]
    mo.target_from_points(f64(target_x), f64(target_m))
    return mo
