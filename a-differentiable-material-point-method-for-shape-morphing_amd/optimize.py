"""Control optimisation loops over a device ``Morph``.

``optimize_control_sequence`` is CompGraph::OptimizeDefGradControlSequence
(CompGraph.cpp:280-444) with every pass, the gradient norm and the control
update (PointCloud::Descend_*) on the GPU; only the loop control runs on the
host.  ``gradient_descent`` is the plain loop used to compare loss curves of
the FP32 and FP64 paths: gradient, normalised step, backtracking line search.
"""
from __future__ import annotations

DESCEND_ADAM, DESCEND_AMSGRAD, DESCEND_LION, DESCEND_GD = range(4)


def optimize_control_sequence(mo, control_stride, max_gd_iters, max_line_search_iters, initial_alpha, gd_tol,
                              temporal_iters=5, beta1=0.9, beta2=0.999, epsilon=1e-5, log=None):
    """CompGraph.cpp:280-444 on the device.  The reference applies Descend_Adam
    then Descend_Lion at every line-search trial and undoes a failed trial with
    the same calls at -alpha (CompGraph.cpp:391-409); so does this loop.  One
    deliberate difference: the convergence test uses the norm of the control
    layer's own gradient, where the reference reads layer 0's (stale when the
    control layer is not 0).  Returns (initial loss, final loss, gradient-norm
    history)."""
    L = mo.num_layers
    mo.forward_pass(0)
    initial_loss = mo.end_layer_mass_loss()
    mo.backward_pass(0)
    initial_norm = mo.control_grad_norm(0)
    history = []
    timestep = 1
    for _ in range(temporal_iters):
        for cf in range(0, L - 1, control_stride):
            alpha = initial_alpha
            for _gd in range(max_gd_iters):
                mo.forward_pass(cf)
                gd_loss = mo.end_layer_mass_loss()
                mo.backward_pass(cf)
                gn = mo.control_grad_norm(cf)
                history.append(gn)
                if log:
                    log(f"control {cf} loss {gd_loss:.6g} |g| {gn:.6g}")
                if gn < gd_tol * initial_norm:
                    break
                found = False
                for _ls in range(max_line_search_iters):
                    mo.descend(cf, DESCEND_ADAM, alpha, gn, beta1, beta2, epsilon, timestep)
                    mo.descend(cf, DESCEND_LION, alpha, gn, beta1)
                    mo.forward_pass(cf)
                    if mo.end_layer_mass_loss() < gd_loss:
                        found = True
                        break
                    mo.descend(cf, DESCEND_ADAM, -alpha, gn, beta1, beta2, epsilon, timestep)
                    mo.descend(cf, DESCEND_LION, -alpha, gn, beta1)
                    alpha /= 2.0
                if not found:
                    break
                timestep += 1
    mo.forward_pass(0)
    final_loss = mo.end_layer_mass_loss()
    return initial_loss, final_loss, history


def gradient_descent(mo, iters, alpha, control_layer=0, max_line_search_iters=8, log=None, rule=DESCEND_GD,
                     beta1=0.9, beta2=0.999, epsilon=1e-5):
    """``iters`` iterations of gradient descent on the control of
    ``control_layer`` with a halving line search that keeps the last accepted
    step size.  ``rule`` DESCEND_GD is Descend_GradientDescent
    (PointCloud.cpp:164-179: step alpha g / |g|); DESCEND_ADAM is
    Descend_Adam (PointCloud.cpp:83-99) with the timestep advanced per accepted
    step only (CompGraph.cpp:440) and a failed trial undone by the reference's
    own -alpha call.  That undo also moves Adam's moments, so it does not restore
    the control exactly: after a failed search the loss is evaluated again and
    the recorded loss is always that of the control held.
    Returns the loss after every iteration (index 0: the initial loss)."""
    mo.forward_pass(0)
    loss = mo.end_layer_mass_loss()
    losses = [loss]
    timestep = 1
    for it in range(iters):
        mo.backward_pass(control_layer)
        gn = mo.control_grad_norm(control_layer)
        if gn == 0.0:
            break
        accepted = False
        for _ in range(max_line_search_iters):
            mo.descend(control_layer, rule, alpha, gn, beta1, beta2, epsilon, timestep)
            mo.forward_pass(control_layer)
            trial = mo.end_layer_mass_loss()
            if trial < loss:
                loss = trial
                accepted = True
                break
            mo.descend(control_layer, rule, -alpha, gn, beta1, beta2, epsilon, timestep)
            alpha *= 0.5
        if accepted:
            timestep += 1
        else:
            mo.forward_pass(control_layer)
            loss = mo.end_layer_mass_loss()
        losses.append(loss)
        if log:
            log(f"iter {it + 1}: loss {loss:.8g} alpha {alpha:.3g}")
    return losses
