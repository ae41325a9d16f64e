#!/usr/bin/env python3
"""Forward+adjoint particle-steps/s of the differentiable MLS-MPM step on one MI355X.

Workload (default ``morph_1m_128``): 1M particles in a seeded sphere, 128^3
grid (dx = 1/128), E log-uniform in [5e3, 5e4], nu = 0.3, rho = 1, 100-step
episodes (101 layers), target the P2G mass of a seeded 12-sphere union, seeded
non-rest control dFc ~ N(0, 1e-2) on layer 0.  One bench step is one
gradient evaluation of an episode: ComputeForwardPass(0) + EndLayerMassLoss +
ComputeBackwardPass(0).  particle-steps/s = particles x 100 x steps / time.

Timed paths: FP32 state with FP64 polar/SVD/H:dF and checkpoint+recompute
(the reported value: autotuned range size / LDS capacity, passes replayed
from HIP graphs), the FP32 path storing every layer (memory and speed of the
two tapes), and the plain FP64 path with every layer stored; the
control-gradient error is FP32 vs FP64 on the same inputs.  Both paths report
their whole-step roofline fraction (the same algorithmic bytes, twice as wide
in FP64).  Per-kernel durations come from HIP events around every launch of
eager episodes run right after the timed region (one for the FP64 path;
graph replays cannot be bracketed per kernel, the kernels are the same).  A batched run advances 24 uneven morphs
(5k-300k particles, 32^3-96^3 grids, 50-step episodes) together.  The CPU
baseline is the FP64 oracle (a port) on a bounded sample of the same scene.

Stability: the scene's own dt (2e-4) with the reference's dimensionless
particle volume exceeds the explicit CFL limit many times over; the bench uses
volumes scaled by dx^3 and dt = 1e-5 (CFL ~0.33 at E = 5e4).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from __graft_entry__ import _load_pkg  # noqa: E402

OPT_RANGE_MAX, OPT_LDS_CAP, OPT_TILE, OPT_PERSISTENT = 2, 3, 10, 11

# Algorithmic HBM bytes per FP32 particle / per touched node / per dense grid node
# of each kernel (FP64: twice), the per-particle-step figures of the scope
# table.  Forward: P2G reads x, v, C, F, dFc, m, vol, lam, mu (148 B) and
# writes m, p (16 B per touched node); G2P + update reads x, F, dFc (84 B),
# writes x, v, C, F (96 B) and reads v (12 B per touched node).  The adjoint
# step (376 B + 72 B per touched node) split by what each kernel needs: the
# transposed G2P reads x, F, dFc, x', v', F', C' (180 B) and scatters vbar
# (12 B per node); the transposed P2G with its elastic half reads v, C, m, vol,
# lam, mu, C' and writes the x, v, C, F gradients (196 B), per node m, v read,
# pbar/mbar write and read, reset (60 B).  The fused launch of a step's G2P with
# the next step's P2G reads each particle once: the scope table's 244 B for a
# fused P2G + G2P (+ 28 B per touched node).  What the implementation adds (the
# cached affine matrix, vt/Ct and Gb round trips, window overlap) shows in the
# measured traffic, not here.
KBYTES = {
    "p2g": (148, 16, 0),
    "g2p": (180, 12, 0),
    "step": (244, 28, 0),
    "g2p_adj": (180, 12, 0),
    "p2g_adj": (196, 60, 0),
}
HBM_PEAK_GBS = 8000.0
# one forward + one adjoint particle-step: (328 + 376) B per particle, (56 + 72) B
# per touched node; no dense grid pass (resets cover the flagged tiles only)
STEP_BYTES = (704, 128, 0)
# the FP32 path fuses each step's G2P with the next P2G: (244 + 376) B per particle
STEP_BYTES_FUSED = (620, 128, 0)
# measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a
# device-to-device copy of known size) of the 1M-particle scene's kernels
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_morph_1m_128.txt")


def measured_traffic(kernel):
    """HBM bytes per launch of ``kernel`` from the committed counter summary, or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        for line in fh:
            f = line.split()
            if f and not f[0].startswith("#") and f[0] == kernel:
                return int(f[1])
    return None


def touched_nodes(sc):
    """Nodes inside the 4^3 stencil of at least one particle of layer 0."""
    mark = np.zeros(sc.dim, dtype=bool)
    base = np.floor((sc.x - np.asarray(sc.min_point)) / sc.dx).astype(np.int64) - 1
    for i in range(4):
        for j in range(4):
            for k in range(4):
                idx = base + np.array([i, j, k])
                ok = np.all((idx >= 0) & (idx < np.asarray(sc.dim)), axis=1)
                mark[idx[ok, 0], idx[ok, 1], idx[ok, 2]] = True
    return int(mark.sum())


def make_scene(pkg, name, dt):
    S = pkg.scenes
    if name == "morph_1m_128":
        return S.morph_1m(dt=dt)
    if name == "morph_130k_64":
        return S.morph_130k(dt=dt)
    if name == "morph_8k_32":
        return S.morph_small(dt=dt)
    if name == "morph_4m_256":
        return S.morph_4m(dt=dt)
    raise SystemExit(f"unknown workload {name}")


OVERLAP = True  # --no-overlap: checkpoint recompute on the main stream (isolated kernel times in a trace)
ELASTIC = None  # --elastic N: DMPM_OPT_ELASTIC of the FP32 contexts (None: the library's default)
FUSE = None  # --fuse N: DMPM_OPT_FUSE of the FP32 contexts (None: the library's default)


def setup(pkg, sc, vol, dFc, precision, tape, num_layers):
    mo = pkg.Morph.from_scene(sc, precision=precision, tape=tape, num_layers=num_layers, dt=sc.dt)
    if not OVERLAP and tape == pkg.TAPE_CHECKPOINT:
        mo.set_option(pkg.OPT_OVERLAP, 0)
    if ELASTIC is not None and precision == pkg.PREC_F32:
        mo.set_option(pkg.OPT_ELASTIC, ELASTIC)
    if FUSE is not None and precision == pkg.PREC_F32:
        mo.set_option(pkg.OPT_FUSE, FUSE)
    x, v, C, F = sc.state0()
    mo.set_particles(x, v, F, C, sc.m, vol, sc.lam, sc.mu)
    mo.set_control(0, dFc)
    mo.target_from_points(sc.target_x, sc.target_m)
    return mo


def episode(mo):
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    return loss


def timed_episodes(mo, warmup, steps, prof_steps=0, tune=False, graphs=False):
    """Warm up, time ``steps`` episodes, then (optionally) profile ``prof_steps`` eager episodes."""
    tuned = None
    if tune:
        tuned = {"step_ms": mo.autotune(), "range_max": mo.option(OPT_RANGE_MAX), "lds_cap": mo.option(OPT_LDS_CAP),
                 "tile": mo.option(OPT_TILE), "persistent": mo.option(OPT_PERSISTENT)}
    mo.set_graphs(graphs)
    for _ in range(warmup):
        episode(mo)
    mo.synchronize()
    t0 = time.perf_counter()
    loss = None
    for _ in range(steps):
        loss = episode(mo)
    mo.synchronize()
    dt = time.perf_counter() - t0
    prof = None
    if prof_steps:
        mo.profile(True)
        mo.profile_reset()
        for _ in range(prof_steps):
            episode(mo)
        prof = mo.profile_read()
        mo.profile(False)
    return dict(time=dt, loss=loss, prof=prof, tuned=tuned)


def run_path(pkg, sc, vol, dFc, precision, tape, L, warmup, steps, prof_steps=0, tune=False, graphs=False):
    mo = setup(pkg, sc, vol, dFc, precision, tape, L)
    r = timed_episodes(mo, warmup, steps, prof_steps, tune, graphs)
    r["grad"] = mo.control_grad(0)
    r["mem"] = mo.memory_bytes()
    mo.close()
    return r


def run_batched(pkg, count, L, dt, warmup, steps):
    """Uneven instances advanced together: one FP32 checkpointed context, graphs, autotuned."""
    scs = pkg.scenes.batched_instances(count, dt=dt)
    mb = pkg.MorphBatch.from_scenes(scs, num_layers=L, dt=dt, tape=pkg.TAPE_CHECKPOINT)
    if ELASTIC is not None:
        mb.set_option(pkg.OPT_ELASTIC, ELASTIC)
    if FUSE is not None:
        mb.set_option(pkg.OPT_FUSE, FUSE)
    st = [sc.state0() for sc in scs]
    cat = lambda k: np.concatenate([t[k] for t in st])  # noqa: E731
    m = np.concatenate([sc.m for sc in scs])
    lam = np.concatenate([sc.lam for sc in scs])
    mu = np.concatenate([sc.mu for sc in scs])
    mb.set_particles(cat(0), cat(1), cat(3), cat(2), m, None, lam, mu)
    mb.compute_volumes()
    vol = np.concatenate([v * sc.dx ** 3 for v, sc in zip(mb.split(mb.volumes()), scs)])
    mb.set_particles(cat(0), cat(1), cat(3), cat(2), m, vol, lam, mu)
    mb.set_control(0, np.concatenate([pkg.scenes.random_controls(sc.n, 1e-2, 7 + i) for i, sc in enumerate(scs)]))
    mb.target_from_points([sc.target_x for sc in scs], [sc.target_m for sc in scs])
    r = timed_episodes(mb, warmup, steps, 0, tune=True, graphs=True)
    n = mb.n
    grids = sorted({sc.dim[0] for sc in scs})
    mb.close()
    return {"value": n * (L - 1) * steps / r["time"], "unit": "particle-steps/s", "instances": count,
            "particles": n, "grids": grids, "ms_per_step": r["time"] / steps * 1e3, "episode_steps": L - 1,
            "tuned": r["tuned"], "loss": r["loss"]}


def run_chained(pkg, sc, vol, L, episodes, resort_mode, seed=11, alpha=1e-3):
    """Chained episodes as the reference's episode loop runs them (DiffMPMAnimator3D.cpp:421-471),
    all on the device: an episode is one descent iteration on the layer-0 control (fwd + loss +
    adjoint, the gradient norm, Descend_Adam) and the forward pass of the updated control, then
    the chain to the next (layer 0 <- last layer, controls <- 0; with resort_mode >= 1 the particles
    are re-sorted by block there).  FP32 + checkpoint, graphs, autotuned once; resort_mode is
    DMPM_OPT_RESORT.  Returns per-episode wall ms (chain included; chain_ms alone), the losses and
    the P2G window statistics of each episode's first forward pass."""
    x, v, C, F = sc.state0()
    mo = pkg.Morph.from_scene(sc, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT, num_layers=L, dt=sc.dt)
    mo.set_particles(x, v, F, C, sc.m, vol, sc.lam, sc.mu)
    mo.target_from_points(sc.target_x, sc.target_m)
    mo.set_control(0, pkg.scenes.random_controls(sc.n, 1e-2, seed))
    mo.autotune()
    mo.set_graphs(True)
    mo.set_option(pkg.OPT_RESORT, resort_mode)
    episode(mo)  # warm-up: graphs captured
    mo.synchronize()
    ms, chain_ms, stats, losses = [], [], [], []
    for e in range(episodes):
        t0 = time.perf_counter()
        if e > 0:
            mo.chain_episode()
            mo.synchronize()
        t1 = time.perf_counter()
        losses.append(episode(mo))
        stats.append(mo.window_stats())
        gn = mo.control_grad_norm(0)
        if gn > 0:
            mo.descend(0, pkg.DESCEND_ADAM, alpha, gn, timestep=e + 1)
        mo.forward_pass(0)
        mo.synchronize()
        t2 = time.perf_counter()
        ms.append((t2 - t0) * 1e3)
        chain_ms.append((t1 - t0) * 1e3)
    mo.close()
    tot = [s_["fixed"] + s_["general"] + s_["hbm"] for s_ in stats]
    return {"ms": [round(t, 3) for t in ms], "chain_ms": [round(t, 3) for t in chain_ms], "loss": losses,
            "fixed_share": [round(s_["fixed"] / max(t, 1), 4) for s_, t in zip(stats, tot)],
            "general_share": [round(s_["general"] / max(t, 1), 4) for s_, t in zip(stats, tot)],
            "hbm_share": [round(s_["hbm"] / max(t, 1), 4) for s_, t in zip(stats, tot)],
            "sorts": stats[-1]["sorts"] if stats else 0}


def loss_curves(pkg, sc, vol, dFc, L, args):
    """Descent on the layer-0 control from the same seed on the FP32+checkpoint and
    plain FP64 (checkpointed, to fit) paths; one JSON line with both loss curves."""
    out = {"workload": sc.name, "particles": sc.n, "grid": list(sc.dim), "episode_steps": L - 1, "dt": sc.dt,
           "rho": args.rho,
           "iterations": args.gd_iters, "rule": "Descend_Adam with halving line search", "alpha": args.gd_alpha}
    paths = {"f32": (pkg.PREC_F32, "fp32_ckpt"), "f64": (pkg.PREC_F64, "fp64_ckpt")}
    for key in (["f32", "f64"] if args.gd_precision == "both" else [args.gd_precision]):
        prec, name = paths[key]
        mo = setup(pkg, sc, vol, dFc, prec, pkg.TAPE_CHECKPOINT, L)
        if prec == pkg.PREC_F32:
            mo.autotune()
            mo.set_graphs(True)
        t0 = time.perf_counter()
        curve = pkg.optimize.gradient_descent(
            mo, args.gd_iters, args.gd_alpha, rule=pkg.DESCEND_ADAM,
            log=lambda m, n=name: print(f"# {n} {m}", file=sys.stderr, flush=True))
        mo.synchronize()
        out[name] = {"loss": curve, "seconds": time.perf_counter() - t0, "memory_bytes": mo.memory_bytes()}
        mo.close()
    if "fp32_ckpt" in out and "fp64_ckpt" in out:
        a, b = np.array(out["fp32_ckpt"]["loss"]), np.array(out["fp64_ckpt"]["loss"])
        k = min(len(a), len(b))
        out["max_rel_diff_fp32_vs_fp64"] = float(np.max(np.abs(a[:k] - b[:k]) / np.abs(b[:k])))
    print(json.dumps(out))


def cpu_baseline(sc, vol, dFc, threads, steps):
    """One ``steps``-step episode (forward keeping every layer, loss, adjoint) of the CPU
    restatement on ``threads`` OpenMP threads, in the float build (the reference's precision)
    and in double.  Returns {precision: (particle-steps/s, seconds)}."""
    from oracle import oracle as orc

    orc.set_threads(threads)
    prm = orc.Params(n=sc.n, dim=sc.dim, dx=sc.dx, min_point=sc.min_point, dt=sc.dt)
    x, v, C, F = sc.state0()
    st = orc.State(x, v, C, F)
    cs = orc.Consts(sc.m, vol, sc.lam, sc.mu)
    target = orc.p2g_mass(prm, sc.target_x, sc.target_m)
    controls = [dFc] + [np.zeros_like(dFc)] * (steps - 1)
    out = {}
    t0 = time.perf_counter()
    orc.episode_f32(prm, st, cs, controls, target, steps + 1, threads=threads)
    dt = time.perf_counter() - t0
    out["f32"] = (sc.n * steps / dt, dt)
    t0 = time.perf_counter()
    orc.episode(prm, st, cs, controls, target, steps + 1)
    dt = time.perf_counter() - t0
    out["f64"] = (sc.n * steps / dt, dt)
    return out


def copy_probe(nbytes):
    """One device-to-device copy of ``nbytes`` through the HIP runtime's copy kernel:
    the known-size reference for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a = ctypes.c_void_p()
    b = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes}))


def copy_bandwidth(nbytes=1 << 28, reps=5):
    """Achievable HBM bandwidth on this device: the HIP runtime's device-to-device
    copy of ``nbytes`` (read + write) timed with HIP events on the null stream,
    best of ``reps`` after one untimed copy."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    assert hip.hipMalloc(ctypes.byref(a), n) == 0 and hip.hipMalloc(ctypes.byref(b), n) == 0
    assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
    assert hip.hipMemset(a, 1, n) == 0
    best = None
    for r in range(reps + 1):
        assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpyAsync(b, a, n, 3, None) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        if r and (best is None or ms.value < best):
            best = ms.value
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2.0 * nbytes / (best / 1e3) / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="morph_1m_128")
    ap.add_argument("--episode-steps", type=int, default=100)
    ap.add_argument("--dt", type=float, default=1e-5)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=25)
    ap.add_argument("--prof-steps", type=int, default=2)
    ap.add_argument("--store-all-steps", type=int, default=3)
    ap.add_argument("--batched-steps", type=int, default=3)
    ap.add_argument("--episodes", type=int, default=10,
                    help="chained episodes after the timed region, under each re-sort policy; 0: skip")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-overlap", action="store_true",
                    help="recompute checkpoint segments on the main stream, not beside the adjoint (kernel traces)")
    ap.add_argument("--elastic", type=int, default=None, choices=[0, 1, 2],
                    help="where the FP32 transposed P2G's elastic half runs (DMPM_OPT_ELASTIC)")
    ap.add_argument("--fuse", type=int, default=None, choices=[0, 1],
                    help="FP32 forward steps: 0 separate G2P and P2G launches, 1 fused (DMPM_OPT_FUSE)")
    ap.add_argument("--gd-iters", type=int, default=0,
                    help="loss-curve mode: this many descent iterations on the layer-0 control (FP32+ckpt and FP64)")
    ap.add_argument("--gd-precision", default="both", choices=["both", "f32", "f64"])
    ap.add_argument("--gd-alpha", type=float, default=1e-3)
    ap.add_argument("--rho", type=float, default=1.0, help="density (scales particle and target masses)")
    ap.add_argument("--copy-probe", type=int, default=0, help="only run a D2D copy of this many bytes")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    global OVERLAP, ELASTIC, FUSE
    OVERLAP = not args.no_overlap
    ELASTIC = args.elastic
    FUSE = args.fuse

    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; --gpus must be 1 (sharding across GPUs is described, not implemented)"}))
        sys.exit(2)

    pkg = _load_pkg()
    sc = make_scene(pkg, args.workload, args.dt)
    if args.rho != 1.0:
        sc.m = sc.m * args.rho
        sc.target_m = sc.target_m * args.rho
    L = args.episode_steps + 1
    # reference-formula volumes (computed on the GPU) scaled by dx^3
    mo = pkg.Morph.from_scene(sc, num_layers=2)
    x, v, C, F = sc.state0()
    mo.set_particles(x, v, F, C, sc.m, None, sc.lam, sc.mu)
    mo.compute_volumes()
    vol = mo.volumes() * sc.dx ** 3
    mo.close()
    dFc = pkg.scenes.random_controls(sc.n, 1e-2, 1)
    if args.gd_iters:
        loss_curves(pkg, sc, vol, dFc, L, args)
        return
    touched = touched_nodes(sc)
    print(f"# workload {sc.name}: n={sc.n} grid={sc.dim} L={L} touched_nodes={touched}", file=sys.stderr, flush=True)

    r32 = run_path(pkg, sc, vol, dFc, pkg.PREC_F32, pkg.TAPE_CHECKPOINT, L, args.warmup, args.steps, args.prof_steps,
                   tune=not args.no_tune, graphs=not args.no_graphs)
    value = sc.n * args.episode_steps * args.steps / r32["time"]
    print(f"# fp32+ckpt: {r32['time'] / args.steps * 1e3:.2f} ms/episode, {value:.4e} particle-steps/s "
          f"(tuned {r32['tuned']})", file=sys.stderr, flush=True)

    rsa = None
    if args.store_all_steps:
        rsa = run_path(pkg, sc, vol, dFc, pkg.PREC_F32, pkg.TAPE_STORE_ALL, L, 1, args.store_all_steps,
                       tune=not args.no_tune, graphs=not args.no_graphs)
        vsa = sc.n * args.episode_steps * args.store_all_steps / rsa["time"]
        print(f"# fp32 store-all: {rsa['time'] / args.store_all_steps * 1e3:.2f} ms/episode, {vsa:.4e} particle-steps/s, "
              f"{rsa['mem'] / 2**30:.2f} GiB vs ckpt {r32['mem'] / 2**30:.2f} GiB", file=sys.stderr, flush=True)

    r64 = None
    if not args.no_fp64:
        tape64 = "store-all"
        try:
            r64 = run_path(pkg, sc, vol, dFc, pkg.PREC_F64, pkg.TAPE_STORE_ALL, L, 1, args.fp64_steps,
                           prof_steps=1 if args.prof_steps else 0)
        except pkg.DmpmError as e:
            if "dmpm error -4:" not in str(e):  # only an out-of-memory tape falls back
                raise
            # every FP64 layer does not fit in HBM: the checkpoint tape holds sqrt(L) of them
            print(f"# fp64 store-all does not fit ({e}); checkpoint tape instead", file=sys.stderr, flush=True)
            tape64 = "checkpoint"
            r64 = run_path(pkg, sc, vol, dFc, pkg.PREC_F64, pkg.TAPE_CHECKPOINT, L, 1, args.fp64_steps,
                           prof_steps=1 if args.prof_steps else 0)
        v64 = sc.n * args.episode_steps * args.fp64_steps / r64["time"]
        print(f"# fp64 plain: {r64['time'] / args.fp64_steps * 1e3:.2f} ms/episode, {v64:.4e} particle-steps/s", file=sys.stderr, flush=True)

    # roofline of the dominant kernel: the largest single kernel by total time
    # (the "p2g_adj" timing covers two launches, the transposed P2G gather and
    # the elastic adjoint, so it is not a candidate); every kernel's own
    # fraction goes to kernel_rooflines
    prof = r32["prof"]
    roof, kernels, kroof = None, None, None
    if prof is not None:
        dom = max((k for k in KBYTES if k != "p2g_adj"), key=lambda k: prof[k][0])
        ms, launches = prof[dom]
        pb, nb, gb = KBYTES[dom]
        bytes_per_launch = pb * sc.n + nb * touched + gb * sc.G
        avg_s = ms / 1e3 / max(launches, 1)
        achieved = bytes_per_launch / avg_s / 1e9
        traffic = measured_traffic(dom) if sc.name == "morph_1m_128" else None
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_launch": bytes_per_launch,
                "traffic_source": "profiles/traffic_morph_1m_128.txt (rocprofv3 FETCH_SIZE/WRITE_SIZE)" if traffic else None}
        kernels = {k: {"ms_total": round(prof[k][0], 3), "launches": prof[k][1],
                       "avg_us": round(prof[k][0] * 1e3 / max(prof[k][1], 1), 2)} for k in prof}
        kroof = {}
        for k, (pb_, nb_, gb_) in KBYTES.items():
            ms_, n_ = prof[k]
            if n_:
                gbs = (pb_ * sc.n + nb_ * touched + gb_ * sc.G) / (ms_ / 1e3 / n_) / 1e9
                kroof[k] = {"achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}

    copy_gbs = copy_bandwidth()
    print(f"# device-to-device copy: {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    if roof is not None:
        roof["frac_of_copy"] = roof["achieved"] / copy_gbs

    batched = None
    if args.batched_steps:
        batched = run_batched(pkg, 24, 51, args.dt, 1, args.batched_steps)
        print(f"# batched 24: {batched['ms_per_step']:.2f} ms/episode, {batched['value']:.4e} particle-steps/s "
              f"({batched['particles']} particles)", file=sys.stderr, flush=True)

    chained = None
    if args.episodes:
        chained = {"episodes": args.episodes, "episode_steps": args.episode_steps}
        for key, mode in (("resort_when_stale", 2), ("resort_every_chain", 1), ("no_resort", 0)):
            r = run_chained(pkg, sc, vol, L, args.episodes, mode)
            chained[key] = r
            print(f"# chained {key}: ms {r['ms']}, chain ms {r['chain_ms']}, fixed-window share {r['fixed_share']}, "
                  f"sorts {r['sorts']}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        # every core this process may run on, at most the 16 host cores that one GPU's share of the
        # machine provides (OMP_NUM_THREADS)
        nproc = os.cpu_count() or 1
        usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
        share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or 16
        threads = max(1, min(share, usable))
        cb = cpu_baseline(sc, vol, dFc, threads, args.cpu_steps)
        cpu = {"value": cb["f32"][0], "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "nproc": nproc, "usable_cores": usable,
               "sample": f"CPU restatement of the reference step, float build (the reference's precision), one "
                         f"{args.cpu_steps}-step episode fwd+loss+adjoint of {sc.name} ({cb['f32'][1]:.1f} s)",
               "f64": {"value": cb["f64"][0], "seconds": cb["f64"][1]}}
        print(f"# cpu baseline on {threads} of {nproc} cores ({usable} usable): f32 {cb['f32'][0]:.4g}, "
              f"f64 {cb['f64'][0]:.4g} particle-steps/s", file=sys.stderr, flush=True)

    step32 = STEP_BYTES_FUSED[0] * sc.n + STEP_BYTES_FUSED[1] * touched + STEP_BYTES_FUSED[2] * sc.G
    out = {
        "metric": "fwd+adjoint particle-steps/s (FP32+ckpt vs plain FP64) at 1M pts; grad err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["time"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{sc.name}: {sc.n} particles, {sc.dim[0]}^3 grid, {args.episode_steps}-step episodes, "
                               f"FP32 state + FP64 polar/SVD, checkpoint+recompute; step = one fwd+loss+adjoint episode",
                   "dt": args.dt, "volume": "reference formula x dx^3", "control": "dFc ~ N(0,1e-2) on layer 0"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "hbm_copy_probe": {"achieved_GBs": copy_gbs, "frac_of_peak": copy_gbs / HBM_PEAK_GBS,
                           "note": "HIP runtime device-to-device copy of 256 MiB, read + write bytes, best of 5"},
        "step_roofline": {"bytes_per_particle_step": step32 / sc.n,
                          "achieved_GBs": step32 * args.episode_steps * args.steps / r32["time"] / 1e9,
                          "frac": step32 * args.episode_steps * args.steps / r32["time"] / 1e9 / HBM_PEAK_GBS,
                          "note": "one fused forward (244 B) + one adjoint (376 B) per particle-step, 128 B per touched "
                                  "node; recompute counts in the time"},
        "kernels_fp32": kernels,
        "kernel_rooflines": kroof,
        "kernel_timing": f"HIP events around every launch of {args.prof_steps} eager episodes after the timed region",
        "tuned": r32["tuned"],
        "graphs": not args.no_graphs,
        "overlap": not args.no_overlap,
        "touched_nodes": touched,
        "memory_fp32_ckpt_bytes": r32["mem"],
    }
    if rsa is not None:
        out["fp32_store_all"] = {"value": sc.n * args.episode_steps * args.store_all_steps / rsa["time"],
                                 "ms_per_step": rsa["time"] / args.store_all_steps * 1e3, "memory_bytes": rsa["mem"]}
    if batched is not None:
        out["batched"] = batched
    if chained is not None:
        out["chained"] = chained
    if r64 is not None:
        v64 = sc.n * args.episode_steps * args.fp64_steps / r64["time"]
        g32, g64 = r32["grad"], r64["grad"]
        # the same algorithmic bytes as the FP32 step, every field twice as wide
        step64 = 2 * (STEP_BYTES[0] * sc.n + STEP_BYTES[1] * touched + STEP_BYTES[2] * sc.G)
        gbs64 = step64 * args.episode_steps * args.fp64_steps / r64["time"] / 1e9
        out["fp64_plain"] = {"value": v64, "ms_per_step": r64["time"] / args.fp64_steps * 1e3, "memory_bytes": r64["mem"],
                             "loss": r64["loss"], "tape": tape64,
                             "step_roofline": {"bytes_per_particle_step": step64 / sc.n, "achieved_GBs": gbs64,
                                               "frac": gbs64 / HBM_PEAK_GBS}}
        if r64["prof"] is not None:
            out["fp64_plain"]["kernels"] = {k: {"ms_total": round(v[0], 3), "launches": v[1],
                                                "avg_us": round(v[0] * 1e3 / max(v[1], 1), 2)}
                                            for k, v in r64["prof"].items() if v[1]}
        out["speedup_fp32_vs_fp64"] = value / v64
        out["grad_rel_err_fp32_vs_fp64"] = float(np.linalg.norm(g32 - g64) / max(np.linalg.norm(g64), 1e-300))
        out["loss_fp32"] = r32["loss"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
