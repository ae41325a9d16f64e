"""Host logic of the optimisation loops (no GPU): ``optimize.gradient_descent`` driven
against a stand-in morph whose loss, gradient and Descend_Adam update
(PointCloud.cpp:83-99) are small numpy functions, so that the loop's bookkeeping
can be checked exactly: every recorded loss is the loss of the control held at
that point, and the Adam timestep advances on accepted steps only
(CompGraph.cpp:440)."""
import numpy as np

from conftest import load_pkg

opt = load_pkg().optimize


class StubMorph:
    """One control vector c; loss(c) = sum((c - t)^2).  ``flip``: backward_pass hands
    the ascent direction, so every line-search trial fails and the -alpha undo (which
    also moves Adam's moments) leaves c changed."""

    def __init__(self, n=6, flip=False, seed=0):
        rng = np.random.default_rng(seed)
        self.c = rng.normal(size=n)
        self.t = rng.normal(size=n)
        self.flip = flip
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.g = None
        self.timesteps = []
        self.evaluated = None

    def loss_of(self, c):
        return float(np.sum((c - self.t) ** 2))

    def forward_pass(self, start):
        self.evaluated = self.c.copy()

    def end_layer_mass_loss(self):
        return self.loss_of(self.evaluated)

    def backward_pass(self, layer):
        g = 2.0 * (self.evaluated - self.t)
        self.g = -g if self.flip else g

    def control_grad_norm(self, layer):
        return float(np.linalg.norm(self.g))

    def descend(self, layer, rule, alpha, gn, beta1, beta2, eps, timestep):
        assert rule == opt.DESCEND_ADAM
        self.timesteps.append(timestep)
        g = self.g / gn
        self.m = beta1 * self.m + (1 - beta1) * g
        self.v = beta2 * self.v + (1 - beta2) * g * g
        mh = self.m / (1 - beta1 ** timestep)
        vh = self.v / (1 - beta2 ** timestep)
        self.c = self.c - alpha * mh / (np.sqrt(vh) + eps)


def _run(mo, iters, alpha=0.1):
    held = []
    losses = opt.gradient_descent(mo, iters, alpha, rule=opt.DESCEND_ADAM, max_line_search_iters=3,
                                  log=lambda _msg: held.append(mo.c.copy()))
    return losses, held


def test_recorded_losses_are_those_of_the_held_control():
    """Accepted and failed searches alike: each recorded loss equals a fresh evaluation of
    the control held after that iteration."""
    for flip in (False, True):
        mo = StubMorph(flip=flip, seed=3)
        c0 = mo.c.copy()
        losses, held = _run(mo, 5)
        assert losses[0] == mo.loss_of(c0)
        assert len(losses) == 6 and len(held) == 5
        for k, c in enumerate(held):
            assert losses[k + 1] == mo.loss_of(c), (flip, k)
        if flip:
            # the inexact undo moved the control: the pre-search loss would be wrong
            assert any(losses[k + 1] != losses[k] for k in range(5))


def test_adam_timestep_advances_on_accepted_steps_only():
    mo = StubMorph(flip=True, seed=4)
    _run(mo, 3)
    assert set(mo.timesteps) == {1}  # every search failed: the timestep never moves
    mo = StubMorph(flip=False, seed=5)
    losses, _ = _run(mo, 4)
    assert all(b < a for a, b in zip(losses, losses[1:]))  # every iteration accepted
    # each trial and its undo share the timestep; it rises by one per accepted step
    steps = sorted(set(mo.timesteps))
    assert steps == list(range(1, len(steps) + 1)) and len(steps) == 4
