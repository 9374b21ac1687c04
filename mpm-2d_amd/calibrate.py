"""Calibration of the g_local_mu2 constants mu_s and E from a column's runout.

The observables of a run are the runout front (the largest x of an active
particle) and the particle centroid (mean x, mean y of the active
particles) at the final step; the loss is the squared error of the three
against a target run,

    L = (front - front*)^2 + (cx - cx*)^2 + (cy - cy*)^2.

Its gradient comes from the library's reverse sweep (mpm_hip_adjoint_*):
the loss is a function of the final x and y only, so the reverse sweep is
seeded with dL/dx and dL/dy (one-hot on the front particle, uniform over
the active particles for the centroid) and returns dL/d(mu_s, E, ...) and
dL/d(initial state).  The front is max(x); its derivative is that of the
front particle's x wherever the maximum is unique, which is where the
reverse sweep and central differences agree.

`calibrate` fits (mu_s, E) to a target's observables by Gauss-Newton on
the three residuals in the variables (mu_s, ln E), with the Jacobian from
three reverse sweeps (one seed per residual) and a backtracking line search
on the squared residual.  The reference has no derivative of its step; the
constants are the ones it hard-codes in materialsrc/g_local_mu2.c:30-41 and
reads from `properties` at g_local_mu2.c:139-140.
"""
import numpy as np

import pympm

OBSERVABLES = ("front", "cx", "cy")


def observables(x, y, active):
    """(front, cx, cy) of a final state and the index of the front particle."""
    idx = np.flatnonzero(np.asarray(active) != 0)
    if idx.size == 0:
        raise ValueError("no active particle")
    k = int(idx[np.argmax(x[idx])])
    return np.array([x[k], x[idx].mean(), y[idx].mean()]), k


def observable_seeds(x, y, active):
    """d(front, cx, cy)/d(final x, y): three seeds for the reverse sweep."""
    idx = np.flatnonzero(np.asarray(active) != 0)
    _, k = observables(x, y, active)
    n = x.size
    sf = {"x": np.zeros(n), "y": np.zeros(n)}
    sf["x"][k] = 1.0
    sx = {"x": np.zeros(n), "y": np.zeros(n)}
    sx["x"][idx] = 1.0 / idx.size
    sy = {"x": np.zeros(n), "y": np.zeros(n)}
    sy["y"][idx] = 1.0 / idx.size
    return [sf, sx, sy]


def loss_and_seed(x, y, active, target):
    """L and dL/d(final x, y) for target observables `target` (length 3)."""
    obs, _ = observables(x, y, active)
    r = obs - np.asarray(target, dtype=np.float64)
    seeds = observable_seeds(x, y, active)
    seed = {f: sum(2.0 * r[j] * seeds[j][f] for j in range(3)) for f in ("x", "y")}
    return float(np.dot(r, r)), seed


def run_observables(adj, nsteps):
    """Forward run of the adjoint object and its observables."""
    adj.forward(nsteps)
    return observables(adj.get("x"), adj.get("y"), adj.active())[0]


def residual_jacobian(adj, nsteps, target):
    """Residuals r = obs - target and dr/d(mu_s, ln E) at the adjoint's
    current constants: one forward run and one reverse sweep per residual
    (a reverse sweep consumes its forward run)."""
    J = np.zeros((3, 2))
    r = None
    E = adj.cfg.material.E
    for j in range(3):
        adj.forward(nsteps)
        x, y, act = adj.get("x"), adj.get("y"), adj.active()
        obs, _ = observables(x, y, act)
        r = obs - np.asarray(target, dtype=np.float64)
        g, _ = adj.backward(observable_seeds(x, y, act)[j])
        J[j] = (g["mu_s"], g["E"] * E)
    return r, J


def calibrate(adj, nsteps, target, mu_s0, E0, max_iter=30, tol=1e-13, log=None):
    """Gauss-Newton fit of (mu_s, E) so that the run's observables match
    `target`.  Returns a dict with the fitted constants, the final residual
    norm, and the history of (mu_s, E, |r|) per iteration."""
    theta = np.array([float(mu_s0), float(np.log(E0))])
    adj.set_material(mu_s=theta[0], E=np.exp(theta[1]))
    history = []
    r, J = residual_jacobian(adj, nsteps, target)
    for it in range(max_iter):
        rn = float(np.linalg.norm(r))
        history.append((float(theta[0]), float(np.exp(theta[1])), rn))
        if log:
            log(f"calibrate: iter {it} mu_s={theta[0]:.10f} E={np.exp(theta[1]):.6e} |r|={rn:.3e}")
        if rn <= tol:
            break
        step, *_ = np.linalg.lstsq(J, -r, rcond=None)
        accepted = False
        for _ in range(12):
            trial = theta + step
            adj.set_material(mu_s=trial[0], E=np.exp(trial[1]))
            r_t = run_observables(adj, nsteps) - np.asarray(target, dtype=np.float64)
            if np.linalg.norm(r_t) < rn:
                accepted = True
                break
            step = 0.5 * step
        if not accepted:
            adj.set_material(mu_s=theta[0], E=np.exp(theta[1]))
            break
        theta = trial
        if np.max(np.abs(step)) <= 1e-14 * max(1.0, np.max(np.abs(theta))):
            r = r_t
            history.append((float(theta[0]), float(np.exp(theta[1])), float(np.linalg.norm(r))))
            break
        r, J = residual_jacobian(adj, nsteps, target)
    else:
        history.append((float(theta[0]), float(np.exp(theta[1])), float(np.linalg.norm(r))))
    return {"mu_s": float(theta[0]), "E": float(np.exp(theta[1])),
            "residual": history[-1][2], "iterations": len(history) - 1, "history": history}
