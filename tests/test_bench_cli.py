"""bench.py's host side (no GPU): the one-GPU rule and the CPU-baseline leg.

* ``--gpus N`` with N > 1 prints a JSON error naming the one-GPU scope and
  exits 2 before any device call;
* the CPU baseline (the CPU restatement on a bounded episode of the workload's
  scene, float build and double) runs on the small sphere scene and returns
  particles x steps / seconds for each;
* the float build computes the episode the double oracle does: loss within
  1e-4 relative (float round-off over a 5-step episode of 8.8k particles).
"""
import json
import os
import subprocess
import sys

import numpy as np

from conftest import ROOT, load_oracle, load_pkg


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_cpu_baseline_leg():
    sys.path.insert(0, ROOT)
    import bench

    pkg = load_pkg()
    orc = load_oracle()
    sc = pkg.scenes.morph_small(dt=1e-4)
    prm = orc.Params(n=sc.n, dim=sc.dim, dx=sc.dx, min_point=sc.min_point, dt=sc.dt)
    vol = orc.point_volumes(prm, sc.x, sc.m) * sc.dx ** 3
    dFc = pkg.scenes.random_controls(sc.n, 1e-2, 1)
    out = bench.cpu_baseline(sc, vol, dFc, 2, 2)
    assert set(out) == {"f32", "f64"}
    for value, seconds in out.values():
        assert seconds > 0 and value == sc.n * 2 / seconds and np.isfinite(value)
    orc.set_threads(os.cpu_count() or 1)


def test_float_build_matches_double_oracle():
    pkg = load_pkg()
    orc = load_oracle()
    sc = pkg.scenes.morph_small(dt=1e-4)
    prm = orc.Params(n=sc.n, dim=sc.dim, dx=sc.dx, min_point=sc.min_point, dt=sc.dt)
    vol = orc.point_volumes(prm, sc.x, sc.m) * sc.dx ** 3
    dFc = pkg.scenes.random_controls(sc.n, 1e-2, 1)
    x, v, C, F = sc.state0()
    st, cs = orc.State(x, v, C, F), orc.Consts(sc.m, vol, sc.lam, sc.mu)
    target = orc.p2g_mass(prm, sc.target_x, sc.target_m)
    controls = [dFc] + [np.zeros_like(dFc)] * 4
    ref, _, _, _ = orc.episode(prm, st, cs, controls, target, 6)
    got = orc.episode_f32(prm, st, cs, controls, target, 6)
    assert abs(got - ref) <= 1e-4 * abs(ref), (got, ref)
