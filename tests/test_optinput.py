"""OptInput records and their cereal JSON files (MPMPolyscope.h:16-82).

CPU: the defaults fixture (the reference's default record in cereal's JSON
layout: a top-level "value0" object, Vec3 fields as {"value0": 3, "value1".."value3"},
float fields as the doubles of their float32 values) reads to the dataclass
defaults and is what the writer produces; round trips; malformed files raise;
LoadScene's grid rule.  GPU: a Morph set up from an OptInput and .mpmbin files
steps like one set up directly from the same float32 inputs (1e-12: only the
order of FP64 atomic additions differs).
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT
from _scenes import maxrel, pkg

oi = pkg.optinput
GOLDEN = os.path.join(ROOT, "tests", "golden", "optinput_reference_defaults.json")


def test_defaults_fixture_reads_and_writes():
    opt = oi.read(GOLDEN)
    assert opt == oi.OptInput()
    assert opt.mu == float(np.float32(38888.9)) and opt.dt == float(np.float32(1.0 / 120.0))
    assert opt.grid_min_point == (-16.0, -16.0, -16.0) and opt.num_timesteps == 30
    assert json.loads(oi.to_json(opt)) == json.load(open(GOLDEN))


def test_roundtrip(tmp_path):
    opt = oi.OptInput(mpm_input_mesh_path="in.mpmbin", mpm_target_mesh_path="t.mpmbin", grid_dx=1.0 / 32,
                      grid_min_point=(0.0, -0.5, 0.25), grid_max_point=(1.0, 0.5, 1.25), lam=1234.5, mu=0.1,
                      p_density=1000.0, dt=1e-4, drag=0.2, f_ext=(0.0, -9.81, 0.0), num_timesteps=11,
                      control_stride=2, smoothing_factor=0.0)
    oi.write(tmp_path / "o.json", opt)
    back = oi.read(tmp_path / "o.json")
    assert back == opt
    assert back.f_ext[1] == float(np.float32(-9.81)) and back.dt == float(np.float32(1e-4))


@pytest.mark.parametrize("edit", ["drop_field", "vec_rows", "vec_short", "int_as_float", "no_top"])
def test_malformed_files_raise(edit):
    doc = json.load(open(GOLDEN))
    body = doc["value0"]
    if edit == "drop_field":
        del body["gd_tol"]
    elif edit == "vec_rows":
        body["f_ext"]["value0"] = 2
    elif edit == "vec_short":
        del body["grid_max_point"]["value3"]
    elif edit == "int_as_float":
        body["num_timesteps"] = 30.5
    else:
        doc = body
    with pytest.raises(ValueError):
        oi.from_json(json.dumps(doc))


def test_grid_uses_x_extent_on_every_axis():
    # MPMPolyscope.cpp:94-97 divides the x extent on all three axes
    opt = oi.OptInput(grid_dx=0.5, grid_min_point=(-1.0, -2.0, -3.0), grid_max_point=(1.0, 2.0, 3.0))
    dims, dx, mn = oi.grid(opt)
    assert dims == (4, 4, 4) and dx == 0.5 and mn == (-1.0, -2.0, -3.0)
    assert oi.grid(oi.OptInput())[0] == (32, 32, 32)
    opt = oi.OptInput(grid_dx=0.3, grid_min_point=(0.0, 0.0, 0.0), grid_max_point=(1.0, 1.0, 1.0))
    assert oi.grid(opt)[0] == (4, 4, 4)  # ceil(3.33)


def test_mesh_paths_are_rejected():
    with pytest.raises(ValueError):
        oi.load_point_cloud(oi.OptInput().mpm_input_mesh_path)


@pytest.mark.gpu
def test_morph_from_opt_input_matches_direct_setup(tmp_path):
    sc = pkg.scenes.morph_small(seed=3, dt=1e-4)
    x, v, C, F = sc.state0()
    f32 = lambda a: np.float32(a).astype(np.float64)  # noqa: E731
    dFc = f32(pkg.scenes.random_controls(sc.n, 1e-2, 5))
    vol = f32(np.full(sc.n, float(sc.dx) ** 3 / 8))
    pkg.mpmbin.write(tmp_path / "in.mpmbin", x=x, v=v, F=F, C=C, m=sc.m, vol=vol, lam=sc.lam, mu=sc.mu, dFc=dFc)
    pkg.mpmbin.write(tmp_path / "target.mpmbin", x=sc.target_x, m=sc.target_m)
    opt = oi.OptInput(mpm_input_mesh_path="in.mpmbin", mpm_target_mesh_path="target.mpmbin", grid_dx=sc.dx,
                      grid_min_point=(0.0, 0.0, 0.0), grid_max_point=(1.0, 1.0, 1.0), dt=1e-4, drag=0.1,
                      num_timesteps=4, smoothing_factor=0.05)
    oi.write(tmp_path / "opt.json", opt)
    opt = oi.read(tmp_path / "opt.json")
    assert oi.grid(opt)[0] == tuple(sc.dim)
    mo = oi.morph_from_opt_input(opt, base=str(tmp_path))
    ref = pkg.Morph(sc.n, sc.dim, opt.grid_dx, 4, dt=opt.dt, drag=opt.drag, smoothing=opt.smoothing_factor)
    ref.set_particles(f32(x), f32(v), f32(F), f32(C), f32(sc.m), vol, f32(sc.lam), f32(sc.mu))
    ref.set_control(-1, dFc)
    ref.target_from_points(f32(sc.target_x), f32(sc.target_m))
    for m in (mo, ref):
        m.forward_pass(0)
    la, lb = mo.end_layer_mass_loss(), ref.end_layer_mass_loss()
    # the same inputs; only the order of FP64 atomic additions differs
    assert la == pytest.approx(lb, rel=1e-12)
    for j in (1, 3):
        a, b = mo.layer_state(j), ref.layer_state(j)
        for u, w in zip(a, b):
            assert maxrel(u, w) < 1e-12
    mo.close()
    ref.close()
