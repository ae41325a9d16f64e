"""bench.py's host side (no GPU): the one-GPU rule and the CPU-baseline leg.

* ``--gpus N`` with N > 1 prints a JSON error naming the one-GPU scope and
  exits 2 before any device call;
* the CPU baseline (the FP64 oracle on a bounded episode of the workload's
  scene) runs on the small sphere scene and returns particles x steps / seconds.
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
    value, seconds = bench.cpu_baseline(sc, vol, dFc, 2, 2)
    assert seconds > 0 and value == sc.n * 2 / seconds
    orc.set_threads(os.cpu_count() or 1)
    assert np.isfinite(value)
