"""The C++ host mirror of the reference API (include/DiffMPMLib3D/DiffMPMLib3D.h).

CPU: the .mpmbin format round trip through the C++ reader/writer.
GPU: CompGraph passes, EndLayerMassLoss, Back_Timestep, ForwardTimeStep (and
MultiThreadMPM's in place, against the oracle's proxy-grid step), PK1,
the elastic energy and the per-point J / energy getters through the mirror,
against the FP64 oracle in reference-compat mode (the mirror's default).
Tolerances: the mirror stores float fields like the reference, so states
compare to 1e-6 and gradients to 1e-5 relative.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from _scenes import make_case, maxrel, orc, pkg, rel

BIN = os.path.join(ROOT, "tests", "cpp", "compgraph_selftest")
mpmbin = pkg.mpmbin


def _rand_cloud(n, seed):
    rng = np.random.default_rng(seed)
    return dict(x=rng.normal(size=(n, 3)), v=rng.normal(size=(n, 3)), F=rng.normal(size=(n, 3, 3)),
                C=rng.normal(size=(n, 3, 3)), P=rng.normal(size=(n, 3, 3)), m=rng.uniform(size=n),
                vol=rng.uniform(size=n), lam=rng.uniform(size=n), mu=rng.uniform(size=n),
                dFc=rng.normal(size=(n, 3, 3)), dLdx=rng.normal(size=(n, 3)), dLdv=rng.normal(size=(n, 3)),
                dLdF=rng.normal(size=(n, 3, 3)), momentum=rng.normal(size=(n, 3, 3)),
                vector=rng.normal(size=(n, 3, 3)), vector_max=rng.normal(size=(n, 3, 3)))


def test_mpmbin_layout_and_cpp_roundtrip(tmp_path):
    fields = _rand_cloud(37, 1)
    a = tmp_path / "a.mpmbin"
    b = tmp_path / "b.mpmbin"
    mpmbin.write(a, **fields)
    assert os.path.getsize(a) == 8 + 37 * 432
    raw = open(a, "rb").read()
    assert int.from_bytes(raw[:8], "little") == 37
    assert int.from_bytes(raw[8:12], "little") == 3  # x rows
    assert np.frombuffer(raw[12:24], "<f4").tolist() == np.float32(fields["x"][0]).tolist()
    subprocess.run([BIN, "roundtrip", str(a), str(b)], check=True, timeout=60)
    assert open(a, "rb").read() == open(b, "rb").read()
    back = mpmbin.read(b)
    for k, v in fields.items():
        assert np.array_equal(back[k], np.float32(v).astype(np.float64)), k


def test_mpmbin_rejects_truncated(tmp_path):
    a = tmp_path / "a.mpmbin"
    mpmbin.write(a, **_rand_cloud(3, 2))
    raw = open(a, "rb").read()
    open(a, "wb").write(raw[:-5])
    with pytest.raises(ValueError):
        mpmbin.read(a)
    r = subprocess.run([BIN, "roundtrip", str(a), str(tmp_path / "b")], timeout=60)
    assert r.returncode == 2


def _write_case(tmp_path):
    """The L=4 sphere case in the selftest's input files; returns the oracle's
    float-rounded inputs (case, state, constants, control, params, target)."""
    case = make_case(L=4, per_layer_ctrl=False, drag=0.2, smoothing=0.05)
    sc, st, cs = case.sc, case.st0, case.cs
    # the mirror reads float fields: give the oracle the same float-rounded inputs
    f32 = lambda a: np.float32(a).astype(np.float64)  # noqa: E731
    st = orc.State(f32(st.x), f32(st.v), f32(st.C), f32(st.F))
    cs = orc.Consts(f32(cs.m), f32(cs.vol), f32(cs.lam), f32(cs.mu))
    dFc = f32(case.controls[0])
    prm = orc.Params(**{**case.prm.__dict__, "compat": orc.COMPAT_ALL, "dx": float(np.float32(sc.dx)),
                        "dt": float(np.float32(case.prm.dt)), "drag": float(np.float32(0.2)),
                        "smoothing": float(np.float32(0.05))})
    target = f32(case.target)
    mpmbin.write(tmp_path / "in.mpmbin", x=st.x, v=st.v, F=st.F, C=st.C, m=cs.m, vol=cs.vol, lam=cs.lam, mu=cs.mu,
                 dFc=dFc)
    target.astype("<f8").tofile(tmp_path / "target.bin")
    with open(tmp_path / "params.txt", "w") as fh:
        fh.write(f"{sc.dim[0]} {sc.dim[1]} {sc.dim[2]} {sc.dx!r} 0 0 0 {prm.dt!r} {prm.drag!r} {prm.smoothing!r} 0 0 0 "
                 f"{case.L} {orc.COMPAT_ALL}\n")
    return case, st, cs, dFc, prm, target


@pytest.mark.gpu
def test_cpp_mirror_compgraph_vs_oracle(tmp_path):
    inputs = _write_case(tmp_path)
    subprocess.run([BIN, "run", str(tmp_path)], check=True, timeout=300)
    check_run_outputs(tmp_path, *inputs)


def check_run_outputs(tmp_path, case, st, cs, dFc, prm, target):
    """The selftest's ``run`` outputs against the oracle."""
    controls = [dFc] * (case.L - 1)  # SetUpCompGraph copies layer 0 (and its dFc) into every layer
    ref_loss, layers, gd, g0 = orc.episode(prm, st, cs, controls, target, case.L)
    loss = float(open(tmp_path / "loss.txt").read())
    assert loss == pytest.approx(ref_loss, rel=1e-6)
    last = mpmbin.read(tmp_path / "last.mpmbin")
    assert maxrel(last["x"], layers[-1].x) < 1e-6
    assert maxrel(last["F"], layers[-1].F) < 1e-6
    l0 = mpmbin.read(tmp_path / "l0.mpmbin")
    assert rel(l0["dLdF"], g0.F) < 1e-5
    assert rel(l0["dLdx"], g0.x) < 1e-5
    bt = mpmbin.read(tmp_path / "bt.mpmbin")
    assert rel(bt["dLdF"], g0.F) < 1e-5 and rel(bt["dLdx"], g0.x) < 1e-5
    step = mpmbin.read(tmp_path / "step.mpmbin")
    assert maxrel(step["x"], layers[1].x) < 1e-6 and maxrel(step["F"], layers[1].F) < 1e-6
    vals = np.array([float(t) for t in open(tmp_path / "pk1.txt").read().split()])
    P, psi0 = vals[:9].reshape(3, 3), vals[9]
    s = 0.8
    R = np.array([[1, 0, 0], [0, math.cos(math.pi / 4), -math.sin(math.pi / 4)], [0, math.sin(math.pi / 4), math.cos(math.pi / 4)]])
    lam, mu = 58333.0, 38888.9
    assert np.allclose(P, (2 * mu * (s - 1) + lam * (s ** 3 - 1) * s ** 2) * R, rtol=1e-5, atol=1e-2)
    assert psi0 == pytest.approx(3 * mu * (s - 1) ** 2 + 0.5 * lam * (s ** 3 - 1) ** 2, rel=1e-5)
    # GetPointDeterminants / GetPointElasticEnergies of layer 0: F + dFc summed in
    # float like the reference, then det in float and the energy in FP64 (printed
    # as float)
    diag = np.loadtxt(tmp_path / "diag.txt").reshape(-1, 2)
    Fe = (np.float32(st.F) + np.float32(dFc)).astype(np.float64)
    assert maxrel(diag[:, 0], np.linalg.det(Fe)) < 1e-6
    psi_ref = np.array([orc.energy(Fe[k], cs.lam[k], cs.mu[k]) for k in range(len(Fe))])
    assert maxrel(diag[:, 1], psi_ref) < 1e-6
    # MultiThreadMPM::ForwardTimeStep in place with 4 proxy grids: the oracle's
    # proxy-grid step with 4 sections (no smoothing), cloud and main grid
    rp, rgm, rgv = orc.forward_step_proxy(prm, st, cs, dFc, 4)
    replay = mpmbin.read(tmp_path / "replay.mpmbin")
    for f in ("x", "v", "C", "F"):
        assert maxrel(replay[f], getattr(rp, f)) < 1e-6, f
    assert np.array_equal(replay["dFc"], dFc)  # the control is not touched
    g = np.loadtxt(tmp_path / "replay_grid.txt").reshape(-1, 4)
    assert maxrel(g[:, 0], rgm) < 1e-6 and maxrel(g[:, 1:], rgv) < 1e-5


@pytest.mark.gpu
def test_cpp_mirror_optimize_runs(tmp_path):
    """CompGraph::OptimizeDefGradControlSequence through the mirror (CompGraph.cpp:280-444:
    five sweeps over the control layers, Adam then Lion per line-search trial).
    Some trial episodes diverge (the line search rejects them), and the inexact
    undo of a rejected trial can leave the end loss above the start, so the loop
    is checked for running through them: its first gradient norm is the
    oracle's |dL/dF| of layer 0 at the initial controls (1e-5, float fields),
    every norm and the end loss are finite, the layer-0 control moved."""
    case, st, cs, dFc, prm, target = _write_case(tmp_path)
    r = subprocess.run([BIN, "optimize", str(tmp_path)], check=True, timeout=600, capture_output=True, text=True)
    controls = [dFc] * (case.L - 1)
    ref_loss, _, _, g0 = orc.episode(prm, st, cs, controls, target, case.L)
    initial = float(r.stdout.split("Initial loss =")[1].split()[0])
    final = float(open(tmp_path / "final_loss.txt").read())
    hist = np.loadtxt(tmp_path / "history.txt").reshape(-1)
    print(f"mirror optimisation: loss {initial:.6g} -> {final:.6g}, {len(hist)} gradient evaluations")
    assert initial == pytest.approx(ref_loss, rel=1e-5)  # printed with 6 significant digits
    assert hist[0] == pytest.approx(np.linalg.norm(g0.F), rel=1e-5)
    assert np.all(np.isfinite(hist)) and len(hist) >= case.L - 1
    assert np.isfinite(final)
    l0 = mpmbin.read(tmp_path / "opt_l0.mpmbin")
    assert rel(l0["dFc"], dFc) > 0  # the layer-0 control moved


def _grid_txt(path, n=None):
    g = np.loadtxt(path).reshape(-1, 7)
    return g[:, 0], g[:, 1:4], g[:, 4:7]


def _cubic(t):
    a = np.abs(t)
    return np.where(a < 1, 0.5 * a ** 3 - a * a + 2.0 / 3.0, np.where(a < 2, (2 - a) ** 3 / 6.0, 0.0))


@pytest.mark.gpu
def test_cpp_mirror_phase_functions(tmp_path):
    """The ForwardSimulation.h phase functions through the mirror, called as reference code
    calls them (float containers, so 1e-6 relative unless stated):
    * P_op_1, G_Reset, P2G, G_op, G2P, P_op_2 in turn = one oracle ForwardTimeStep (state, node
      m and v, node p = m v where the node mass passes the FLT_EPSILON threshold);
    * the target grid of MPMPolyscope.cpp:158-160 (grid copy, ResetValues, P2G(cloud, grid,
      0.0, 0.0)) = the oracle's P2G at dt = drag = 0;
    * P2G_Mass(positions, grid, 1.0) = the oracle's P2G mass of unit points, and G2G_Mass onto
      a coarser grid = the separable closed form sum_ijk N_x N_y N_z m_ijk;
    * G_op with a gravity point: v = p/m + dt 3 (c - x_i)/|c - x_i| where the mass passes;
    * the single-particle forms on particles 0..4 = the cloud forms on those particles."""
    case, st, cs, dFc, prm, target = _write_case(tmp_path)
    subprocess.run([BIN, "phases", str(tmp_path)], check=True, timeout=300)
    eps = 1.1920928955078125e-07
    nxt, gm, gv = orc.forward_step(prm, st, cs, dFc)
    ph = mpmbin.read(tmp_path / "phases.mpmbin")
    for f in ("x", "v", "C", "F"):
        assert maxrel(ph[f], getattr(nxt, f)) < 1e-6, f
    m, p, v = _grid_txt(tmp_path / "phases_grid.txt")
    assert maxrel(m, gm) < 1e-6 and maxrel(v, gv) < 1e-5
    live = np.abs(gm) > eps
    assert maxrel(p[live], (gm[:, None] * gv)[live]) < 1e-5

    tm, tp, tv = _grid_txt(tmp_path / "target_grid.txt")
    om, op = orc.p2g(prm, st, cs, dFc, dt=0.0, drag=0.0)
    assert maxrel(tm, om) < 1e-6 and np.abs(tp - op).max() <= 1e-6 * max(np.abs(op).max(), 1e-30) + 1e-12
    assert not tv.any()

    pm, pp, pv = _grid_txt(tmp_path / "p2g_mass.txt")
    assert maxrel(pm, orc.p2g_mass(prm, st.x, np.ones(prm.n))) < 1e-6
    assert not pp.any() and not pv.any()
    cm, _, _ = _grid_txt(tmp_path / "g2g_mass.txt")
    dims = np.array(prm.dim)
    cdims = (dims + 1) // 2
    dx1 = float(np.float32(prm.dx))
    dx2 = float(np.float32(2.0) * np.float32(dx1))
    W = [_cubic((np.arange(dims[a])[None, :] * dx1 - np.arange(cdims[a])[:, None] * dx2) / dx1) for a in range(3)]
    want = np.einsum("Ii,Jj,Kk,ijk->IJK", W[0], W[1], W[2], pm.reshape(tuple(dims))).ravel()
    assert maxrel(cm, want) < 1e-6

    gm2, gp2, gv2 = _grid_txt(tmp_path / "gravity_grid.txt")
    c = (np.float32(dims) * (np.float32(0.5) * np.float32(dx1))).astype(np.float64)
    ijk = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, 3) * dx1
    r = c[None, :] - ijk
    nr = np.linalg.norm(r, axis=1)
    f = np.where(nr[:, None] > 0, 3.0 * r / np.where(nr > 0, nr, 1.0)[:, None], 0.0)
    live2 = np.abs(gm2) > eps
    want_v = np.where(live2[:, None], gp2 / np.where(live2, gm2, 1.0)[:, None] + prm.dt * f, v)
    assert maxrel(gv2, want_v) < 1e-5

    sg = mpmbin.read(tmp_path / "single.mpmbin")
    for f_ in ("x", "v", "C", "F"):
        assert maxrel(sg[f_], ph[f_][:5]) < 1e-6, f_
    sm, sp_, sv = _grid_txt(tmp_path / "single_grid.txt")
    sub = orc.Params(**{**prm.__dict__, "n": 5})
    st5 = orc.State(st.x[:5], st.v[:5], st.C[:5], st.F[:5])
    cs5 = orc.Consts(cs.m[:5], cs.vol[:5], cs.lam[:5], cs.mu[:5])
    om5, op5 = orc.p2g(sub, st5, cs5, dFc[:5])
    assert maxrel(sm, om5) < 1e-6 and maxrel(sp_, op5) < 1e-5
    live5 = np.abs(sm) > eps
    assert maxrel(sv, np.where(live5[:, None], sp_ / np.where(live5, sm, 1.0)[:, None], 0.0)) < 1e-6
