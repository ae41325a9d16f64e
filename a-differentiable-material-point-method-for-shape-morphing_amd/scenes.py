"""Scene and target construction for the differentiable MLS-MPM step.

Host-side numpy code that builds the inputs of the hot path the way the
reference's loaders do, with analytic shapes in place of OBJ meshes:

* lattice sampling at ``point_dx`` (GeometryLoading.cpp:19-35: lattice
  ``min + point_dx*(i,j,k)``, ``ceil((max-min)/point_dx)+1`` samples per axis,
  points inside the shape kept),
* particle init ``m = point_dx^3 * rho``, ``F = I``, ``v = 0``, Lame parameters
  per particle (MPMPolyscope.cpp:45-53, Elasticity.cpp:231-235),
* ``point_dx = grid_dx / points_per_cell_cuberoot`` (MPMPolyscope.cpp:63),
* the control ``dFc`` starts at zero (MaterialPoint.h:31).

The named workloads below are the five benchmark scenes of this project
(sizes in their docstrings).  Volumes and target grids are computed on the
GPU by the library (``CalculatePointCloudVolumes`` and the target-grid P2G).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def lame(E, nu):
    """CalculateLameParameters (Elasticity.cpp:231-235)."""
    E = np.asarray(E, dtype=np.float64)
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 + 2.0 * nu)
    return lam, mu


def _lattice(lo, hi, point_dx):
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    dims = [int(math.ceil((hi[a] - lo[a]) / point_dx)) + 1 for a in range(3)]
    ii = [np.arange(d, dtype=np.float64) for d in dims]
    g = np.stack(np.meshgrid(*ii, indexing="ij"), axis=-1).reshape(-1, 3)
    return lo + point_dx * g


def lattice_spheres(centers, radii, point_dx, jitter=0.0, seed=0):
    """Lattice points inside the union of spheres (x-major order, as the
    reference's i-j-k loop emits them)."""
    centers = np.atleast_2d(np.asarray(centers, dtype=np.float64))
    radii = np.atleast_1d(np.asarray(radii, dtype=np.float64))
    lo = (centers - radii[:, None]).min(axis=0)
    hi = (centers + radii[:, None]).max(axis=0)
    pts = _lattice(lo, hi, point_dx)
    inside = np.zeros(len(pts), dtype=bool)
    for c, r in zip(centers, radii):
        inside |= ((pts - c) ** 2).sum(axis=1) <= r * r
    pts = pts[inside]
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        pts = pts + rng.uniform(-jitter, jitter, size=pts.shape) * point_dx
    return np.ascontiguousarray(pts)


def lattice_box(lo, hi, point_dx):
    """Lattice points of an axis-aligned box."""
    pts = _lattice(lo, hi, point_dx)
    lo = np.asarray(lo)
    hi = np.asarray(hi)
    keep = np.all((pts >= lo - 1e-12) & (pts <= hi + 1e-12), axis=1)
    return np.ascontiguousarray(pts[keep])


def sphere_radius_for_count(n, point_dx):
    """Radius whose lattice ball holds about ``n`` points."""
    return (3.0 * n / (4.0 * math.pi)) ** (1.0 / 3.0) * point_dx


@dataclass
class Scene:
    """Inputs of one morph: particles of layer 0, constants, grid and target points."""

    name: str
    x: np.ndarray                 # (n,3)
    m: np.ndarray                 # (n,)
    lam: np.ndarray               # (n,)
    mu: np.ndarray                # (n,)
    dim: tuple
    dx: float
    min_point: tuple
    dt: float
    drag: float
    smoothing: float
    f_ext: tuple
    num_layers: int               # layers per episode (steps + 1)
    episodes: int
    target_x: np.ndarray          # target point cloud (n_t,3)
    target_m: np.ndarray          # (n_t,)
    point_dx: float
    rho: float
    notes: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def G(self) -> int:
        return int(self.dim[0] * self.dim[1] * self.dim[2])

    def state0(self):
        n = self.n
        v = np.zeros((n, 3))
        C = np.zeros((n, 3, 3))
        F = np.broadcast_to(np.eye(3), (n, 3, 3)).copy()
        return self.x.copy(), v, C, F


def _particles(x, point_dx, rho, E, nu):
    n = len(x)
    m = np.full(n, point_dx ** 3 * rho)
    lam, mu = lame(np.broadcast_to(np.asarray(E, dtype=np.float64), (n,)), nu)
    return m, np.ascontiguousarray(lam), np.ascontiguousarray(mu)


def _sphere_union_targets(rng, count, rmin, rmax, lo=0.25, hi=0.75):
    centers = rng.uniform(lo, hi, size=(count, 3))
    radii = rng.uniform(rmin, rmax, size=count)
    return centers, radii


def morph_small(seed=0, dt=1e-3):
    """8.8k-particle FP64 scene: sphere r=0.2 at the box centre, 2x2x2 per
    cell of a 32^3 grid (dx=1/32), E=1e4, nu=0.3, rho=1, 10 steps, target an
    axis-aligned cube of side 0.35 at a seeded centre."""
    dx = 1.0 / 32
    pdx = dx / 2
    x = lattice_spheres([0.5, 0.5, 0.5], 0.2, pdx)
    m, lam, mu = _particles(x, pdx, 1.0, 1e4, 0.3)
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.45, 0.55, size=3)
    tx = lattice_box(c - 0.175, c + 0.175, pdx)
    return Scene("morph_8k_32", x, m, lam, mu, (32, 32, 32), dx, (0.0, 0.0, 0.0), dt, 0.0, 0.0,
                 (0.0, 0.0, 0.0), 11, 1, tx, np.full(len(tx), pdx ** 3), pdx, 1.0)


def morph_sphere(n_target, grid, steps, episodes, n_spheres, seed=0, E=1e4, E_range=None,
                 dt=1e-3, rmin=0.08, rmax=0.15, name=None):
    """Seeded sphere of about ``n_target`` particles in a ``grid^3`` box of
    side 1 at 2x2x2 per cell, target a union of ``n_spheres`` seeded spheres.
    ``E_range`` draws E log-uniform per particle."""
    dx = 1.0 / grid
    pdx = dx / 2
    r = sphere_radius_for_count(n_target, pdx)
    x = lattice_spheres([0.5, 0.5, 0.5], r, pdx)
    rng = np.random.default_rng(seed)
    if E_range is not None:
        E = np.exp(rng.uniform(math.log(E_range[0]), math.log(E_range[1]), size=len(x)))
    m, lam, mu = _particles(x, pdx, 1.0, E, 0.3)
    centers, radii = _sphere_union_targets(rng, n_spheres, rmin, rmax)
    tx = lattice_spheres(centers, radii, pdx)
    return Scene(name or f"morph_{len(x)}_{grid}", x, m, lam, mu, (grid, grid, grid), dx, (0.0, 0.0, 0.0),
                 dt, 0.0, 0.0, (0.0, 0.0, 0.0), steps + 1, episodes, tx, np.full(len(tx), pdx ** 3), pdx, 1.0)


def morph_130k(seed=0, dt=1e-3):
    """130k particles, 64^3, 5 episodes x 50 steps, target 6 spheres r~U[0.08,0.15]."""
    return morph_sphere(130_000, 64, 50, 5, 6, seed=seed, dt=dt, name="morph_130k_64")


def morph_1m(seed=0, dt=2e-4):
    """1M particles, 128^3, E log-uniform [5e3,5e4], 10 episodes x 100 steps, 12-sphere target."""
    return morph_sphere(1_000_000, 128, 100, 10, 12, seed=seed, E_range=(5e3, 5e4), dt=dt, name="morph_1m_128")


def morph_4m(seed=0, dt=2e-4):
    """4M particles, 256^3, E log-uniform [1e4,1e5], 5 episodes x 200 steps, 20-sphere target."""
    return morph_sphere(4_000_000, 256, 200, 5, 20, seed=seed, E_range=(1e4, 1e5), dt=dt, name="morph_4m_256")


def batched_instances(count=24, seed=0, steps=50, episodes=3, dt=1e-3):
    """Uneven instances: N log-uniform in [5k,300k], grid from {32,48,64,96}^3."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        grid = int(rng.choice([32, 48, 64, 96]))
        n = int(math.exp(rng.uniform(math.log(5e3), math.log(3e5))))
        # keep the sphere inside the box: r <= 0.3
        pdx = 0.5 / grid
        cap = int(0.9 * (4.0 / 3.0) * math.pi * (0.3 / pdx) ** 3)
        out.append(morph_sphere(min(n, cap), grid, steps, episodes, int(rng.integers(2, 7)),
                                seed=seed * 1000 + i, dt=dt, name=f"inst{i}_{grid}"))
    return out


def random_controls(n, scale, seed):
    """Seeded non-rest control increments dFc ~ N(0, scale^2)."""
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, size=(n, 3, 3))
