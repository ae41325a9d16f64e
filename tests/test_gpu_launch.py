"""Launch configurations of the FP32 path against the FP64 oracle (needs an MI355X): block tile
(DMPM_OPT_TILE 4 or 8 cells per edge), persistent launches over the block ranges
(DMPM_OPT_PERSISTENT) and the autotuner that picks them with the range size and LDS capacity.

Every configuration computes the same sums (only the order of float HBM atomics differs), so
each is checked against the oracle at the FP32 bars (loss 1e-6, control gradient 1e-4, final
positions 1e-6, F 2e-5), on the single 8.8k-particle morph (checkpoint tape) and on a batch of
4 uneven instances (each instance's loss 1e-6 against its single-context run):
* tile 8 puts a 512-particle range across a whole 8^3-cell block: the P2G workgroups leave the
  fixed window for the general one;
* a persistent launch runs every range of every instance with as many workgroups as fit on the
  chip, and its counters reset themselves: repeated episodes (and graph replays) agree;
* the autotuner's choice is one of the candidates and leaves results unchanged.
"""
import numpy as np
import pytest

from _scenes import f32_positions, gpu_morph, make_case, maxrel, orc, pkg, rel

pytestmark = pytest.mark.gpu

L = 11


@pytest.fixture(scope="module")
def ref():
    case = make_case(seed=31, L=L, vel=0.2)
    loss, layers, gd, _ = orc.episode(case.prm, case.st0, case.cs, case.controls, case.target, L)
    return case, loss, layers, gd


def _episode(mo):
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    return loss


@pytest.mark.parametrize("tile,persistent", [(4, 0), (8, 0), (4, 1), (8, 1)])
def test_tile_and_persistent_vs_oracle(ref, tile, persistent):
    case, ref_loss, layers, gd = ref
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT)
    try:
        mo.set_option(pkg.OPT_TILE, tile)
        mo.set_option(pkg.OPT_PERSISTENT, persistent)
        assert mo.option(pkg.OPT_TILE) == tile and mo.option(pkg.OPT_PERSISTENT) == persistent
        losses = [_episode(mo) for _ in range(3)]
        st = mo.window_stats()
        print(f"tile {tile} persistent {persistent}: {st}")
        if tile == 8:
            assert st["general"] + st["hbm"] > 0
        for loss in losses:
            assert loss == pytest.approx(ref_loss, rel=1e-6)
        assert rel(mo.control_grad(0), gd[0]) < 1e-4
        x, _, F, _ = mo.layer_state(L - 1)
        assert maxrel(x, f32_positions(layers[-1].x, case.sc.dim, case.sc.dx, case.sc.min_point)) < 1e-6
        assert maxrel(F, layers[-1].F) < 2e-5
        mo.set_graphs(True)
        for _ in range(3):
            assert _episode(mo) == pytest.approx(ref_loss, rel=1e-6)
        assert rel(mo.control_grad(0), gd[0]) < 1e-4
    finally:
        mo.close()


def test_options_reject_other_values(ref):
    case = ref[0]
    mo = gpu_morph(case, precision=pkg.PREC_F32)
    try:
        for opt, bad in ((pkg.OPT_TILE, 6), (pkg.OPT_PERSISTENT, 2)):
            with pytest.raises(pkg.DmpmError):
                mo.set_option(opt, bad)
    finally:
        mo.close()


def test_autotune_choice_keeps_results(ref):
    case, ref_loss, _, gd = ref
    mo = gpu_morph(case, precision=pkg.PREC_F32, tape=pkg.TAPE_CHECKPOINT)
    try:
        ms = mo.autotune()
        cfg = {k: mo.option(o) for k, o in (("tile", pkg.OPT_TILE), ("persistent", pkg.OPT_PERSISTENT),
                                             ("range_max", pkg.OPT_RANGE_MAX), ("lds_cap", pkg.OPT_LDS_CAP))}
        print(f"autotune: {ms:.3f} ms, {cfg}")
        assert ms > 0 and cfg["tile"] in (4, 8) and cfg["persistent"] in (0, 1)
        assert cfg["range_max"] in (128, 256, 512) and cfg["lds_cap"] in (512, 1024, 2048)
        for i, c in enumerate(case.controls):  # the tuner leaves the controls in place; set them again anyway
            mo.set_control(i, c)
        assert _episode(mo) == pytest.approx(ref_loss, rel=1e-6)
        assert rel(mo.control_grad(0), gd[0]) < 1e-4
    finally:
        mo.close()


def _batch(scs, vols, ctrl, tile, persistent):
    mb = pkg.MorphBatch.from_scenes(scs, num_layers=6, dt=1e-5, drag=0.0, smoothing=0.0, tape=pkg.TAPE_CHECKPOINT)
    st = [sc.state0() for sc in scs]
    cat = lambda k: np.concatenate([t[k] for t in st])  # noqa: E731
    mb.set_particles(cat(0), cat(1), cat(3), cat(2), np.concatenate([sc.m for sc in scs]), np.concatenate(vols),
                     np.concatenate([sc.lam for sc in scs]), np.concatenate([sc.mu for sc in scs]))
    mb.set_option(pkg.OPT_TILE, tile)
    mb.set_option(pkg.OPT_PERSISTENT, persistent)
    mb.set_control(0, ctrl)
    mb.target_from_points([sc.target_x for sc in scs], [sc.target_m for sc in scs])
    return mb


@pytest.mark.parametrize("tile,persistent", [(8, 0), (4, 1), (8, 1)])
def test_batch_configurations_equal_default(tile, persistent):
    scs = pkg.scenes.batched_instances(4, dt=1e-5)
    vols = []
    for sc in scs:
        prm = orc.Params(n=sc.n, dim=sc.dim, dx=sc.dx, min_point=sc.min_point)
        vols.append(orc.point_volumes(prm, sc.x, sc.m) * sc.dx ** 3)
    ctrl = np.concatenate([pkg.scenes.random_controls(sc.n, 1e-2, 50 + i) for i, sc in enumerate(scs)])
    out = {}
    for cfg in ((4, 0), (tile, persistent)):
        mb = _batch(scs, vols, ctrl, *cfg)
        try:
            for _ in range(2):
                total = _episode(mb)
            out[cfg] = (total, mb.instance_losses(), mb.control_grad(0))
        finally:
            mb.close()
    a, b = out[(tile, persistent)], out[(4, 0)]
    assert a[0] == pytest.approx(b[0], rel=1e-6)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-6)
    assert rel(a[2], b[2]) < 1e-5
