"""Adjoint (reverse mode) of the USL run.

CPU: the numpy restatement of the reverse sweep (oracle/adjoint_oracle.py)
against central finite differences of the C oracle's forward run -- the
derivative of the map the reference computes, pinned by the forward alone.
GPU: the library's reverse kernels against that restatement on the same
inputs, against finite differences of the library's own forward run, for
independence from the checkpoint interval, through the torch.autograd
bridge, and at the full adjoint workload (129^2 nodes, 65,536 particles,
200 steps, checkpoint every 20).
"""
import sys

import numpy as np
import pytest

import adjoint_oracle as AO
import pympm
import workloads

FD_REL_TOL = 1e-5    # central differences vs adjoint
PARITY_TOL = 1e-9    # GPU reverse sweep vs the CPU restatement


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_case(cells=16, seed=3):
    """A sheared block in the plastic regime: lithostatic stress (tau/p =
    0.43 > mu_s) and random velocities."""
    n = cells
    p = workloads.block(6, 10, n, seed=seed, cell_x0=2)
    p = workloads.lithostatic(p, lambda x: 10.0 / n + 0 * x, k0=0.4)
    rng = np.random.default_rng(seed)
    p["vx"] = 0.05 * rng.standard_normal(len(p["x"]))
    p["vy"] = 0.05 * rng.standard_normal(len(p["x"]))
    return dict(name="adjoint_small", N=n + 1, h=1.0 / n, dt=1e-5,
                material=dict(workloads.WORKLOAD_MATERIAL), g_mag=9.81, ramp_time=0.1,
                **workloads.BOX_FRICTION, particles=p)


SCALES = {"x": 1.0, "y": 1.0, "vx": 1.0, "vy": 1.0, "v": 1e4, "txx": 1e-3, "txy": 1e-3,
          "tyy": 1e-3, "tzz": 1e-3, "gammap": 1.0, "gammadotp": 1e-2, "ux": 1.0, "uy": 1.0}


def weights(np_, seed=0):
    rng = np.random.default_rng(seed)
    return {k: s * rng.standard_normal(np_) for k, s in SCALES.items()}


def linear_loss(wts, final):
    return sum(float(np.dot(wts[k], final[k])) for k in wts)


# FD steps: absolute for the initial state, relative for the constants
INIT_STEPS = {"x": 1e-6, "y": 1e-6, "vx": 1e-5, "vy": 1e-5, "v": 1e-9,
              "sxx": 1e-1, "sxy": 1e-1, "syy": 1e-1}
PARAM_STEP = 1e-6
FD_PARAMS = ["E", "nu", "mu_s", "mu_2", "grains_d"]


def fd_check(name, adj, fd):
    rel = abs(adj - fd) / max(abs(fd), 1e-300)
    say(f"  {name}: adjoint {adj:.10e} fd {fd:.10e} rel {rel:.2e}")
    assert rel <= FD_REL_TOL, f"{name}: adjoint {adj} vs finite difference {fd}"


# ---------------------------------------------------------------------------
# CPU

def test_oracle_adjoint_matches_finite_differences():
    w = small_case()
    steps = 20
    recs, final = AO.record(w, steps)
    assert final["gammap"].max() > 0, "case must reach the plastic branch"
    wts = weights(len(final["x"]))
    lam0, grad = AO.adjoint(recs, wts, w["material"], w["N"], w["h"], w["dt"])
    for k in FD_PARAMS:
        i = AO.PARAMS.index(k)
        base = w["material"][k]
        hs = PARAM_STEP * abs(base)
        vals = []
        for sgn in (1, -1):
            _, f2 = AO.record(w, steps, material=dict(w["material"], **{k: base + sgn * hs}))
            vals.append(linear_loss(wts, f2))
        fd_check(k, grad[i], (vals[0] - vals[1]) / (2 * hs))
    for k, hs in INIT_STEPS.items():
        field = {"sxx": "txx", "sxy": "txy", "syy": "tyy"}.get(k, k)
        for pi in (0, 40):
            vals = []
            for sgn in (1, -1):
                p2 = {q: a.copy() for q, a in w["particles"].items()}
                p2[k][pi] += sgn * hs
                _, f2 = AO.record(w, steps, particles=p2)
                vals.append(linear_loss(wts, f2))
            fd_check(f"{k}[{pi}]", lam0[field][pi], (vals[0] - vals[1]) / (2 * hs))


def test_material_reverse_branches():
    """Closed forms of the branches: the density cut-off passes nothing back;
    an elastic update is linear in the stress, so its adjoint of T is the
    transpose of the (rotation) map and carries no mu_s dependence."""
    c = AO.constants(workloads.WORKLOAD_MATERIAL)
    one = np.ones(1)
    T = [np.array([-100.0]), np.array([3.0]), np.array([-120.0]), np.array([-90.0])]
    Tb = [np.array([1.0]), np.array([0.5]), np.array([-2.0]), np.array([0.25])]
    # density cut-off: m / v < rho_crit
    r = AO.material_reverse(0.1 * one, 0.0 * one, 0.1 * one, 0.0 * one, 1.0 * one, 1.0 * one,
                            T, 1e-5, c, Tb, 0 * one, 0 * one)
    assert all(float(q[0]) == 0.0 for q in r[4]) and float(r[0][0]) == 0.0
    # elastic (tau small against mu_s p), zero strain rate: T passes through
    r = AO.material_reverse(0 * one, 0 * one, 0 * one, 0 * one, 2000.0 * one, 1.0 * one,
                            T, 1e-5, c, Tb, 0 * one, 0 * one)
    assert np.allclose([q[0] for q in r[4]], [q[0] for q in Tb], rtol=0, atol=1e-15)
    assert float(r[5]["mu_s"][0]) == 0.0


# ---------------------------------------------------------------------------
# GPU

def gpu_config(w, material=None, cap=0):
    return pympm.make_config(w["N"], w["h"], w["dt"], material=material or w["material"],
                             bc_kind=w["bc_kind"], bc_fp64=w["bc_fp64"], bc_int=w["bc_int"],
                             g_mag=w["g_mag"], ramp_time=w["ramp_time"], bin_capacity=cap)


def gpu_final(w, steps, material=None, particles=None):
    s = pympm.Sim(gpu_config(w, material), particles if particles is not None else w["particles"])
    s.step(steps)
    out = s.state(AO.FIELDS)
    s.close()
    return out


@pytest.mark.gpu
def test_gpu_adjoint_matches_oracle_adjoint():
    w = small_case()
    steps = 20
    say(f"adjoint parity: {len(w['particles']['x'])} particles, {steps} steps, checkpoint 7")
    recs, final = AO.record(w, steps)
    wts = weights(len(final["x"]))
    lam0, grad = AO.adjoint(recs, wts, w["material"], w["N"], w["h"], w["dt"])
    a = pympm.Adjoint(gpu_config(w), w["particles"], checkpoint_every=7)
    a.forward(steps)
    for k in ("x", "y", "txx", "gammap"):
        assert np.array_equal(a.get(k), final[k]), f"forward {k}"
    gp, gi = a.backward(wts)
    a.close()
    for i, k in enumerate(AO.PARAMS):
        rel = abs(gp[k] - grad[i]) / max(abs(grad[i]), 1e-300)
        say(f"  d/d{k}: gpu {gp[k]:.12e} cpu {grad[i]:.12e} rel {rel:.1e}")
        assert rel <= PARITY_TOL, k
    for k in ("x", "y", "vx", "vy", "v", "txx", "txy", "tyy", "tzz", "gammap", "ux", "uy"):
        scale = np.abs(lam0[k]).max()
        err = np.abs(gi[k] - lam0[k]).max()
        assert err <= PARITY_TOL * scale, f"d/d initial {k}: max err {err} scale {scale}"


@pytest.mark.gpu
def test_gpu_adjoint_dense_cells_match_oracle_adjoint():
    """36 particles per cell in bins of capacity 64: the reverse gather's
    long-bin path (sorted by insertion) against the CPU restatement."""
    n = 16
    p = workloads.block(4, 6, n, seed=9, cell_x0=5, ppc_side=6)
    p = workloads.lithostatic(p, lambda x: 6.0 / n + 0 * x, k0=0.4)
    rng = np.random.default_rng(9)
    p["vx"] = 0.05 * rng.standard_normal(len(p["x"]))
    w = dict(small_case(), particles=p)
    steps = 12
    say(f"adjoint parity, dense cells: {len(p['x'])} particles, {steps} steps")
    recs, final = AO.record(w, steps)
    wts = weights(len(final["x"]), seed=4)
    lam0, grad = AO.adjoint(recs, wts, w["material"], w["N"], w["h"], w["dt"])
    a = pympm.Adjoint(gpu_config(w, cap=64), p, checkpoint_every=5)
    a.forward(steps)
    gp, gi = a.backward(wts)
    a.close()
    for i, k in enumerate(AO.PARAMS):
        rel = abs(gp[k] - grad[i]) / max(abs(grad[i]), 1e-300)
        assert rel <= PARITY_TOL, (k, gp[k], grad[i])
    for k in ("x", "y", "vx", "vy", "v", "txx", "tyy", "gammap"):
        scale = np.abs(lam0[k]).max()
        assert np.abs(gi[k] - lam0[k]).max() <= PARITY_TOL * scale, k


@pytest.mark.gpu
def test_gpu_adjoint_checkpoint_interval_does_not_matter():
    w = small_case(seed=5)
    steps = 23
    wts = weights(len(w["particles"]["x"]), seed=2)
    res = []
    for K in (1, 6, 50):
        say(f"checkpoint every {K}")
        a = pympm.Adjoint(gpu_config(w), w["particles"], checkpoint_every=K)
        a.forward(steps)
        res.append(a.backward(wts))
        a.close()
    for gp, gi in res[1:]:
        assert gp == res[0][0]
        for k in gi:
            assert np.array_equal(gi[k], res[0][1][k]), k


@pytest.mark.gpu
def test_gpu_adjoint_matches_finite_differences():
    w = small_case(seed=7)
    steps = 30
    wts = weights(len(w["particles"]["x"]), seed=3)
    a = pympm.Adjoint(gpu_config(w), w["particles"], checkpoint_every=8)
    a.forward(steps)
    gp, gi = a.backward(wts)
    a.close()
    for k in FD_PARAMS:
        base = w["material"][k]
        hs = PARAM_STEP * abs(base)
        vals = [linear_loss(wts, gpu_final(w, steps, dict(w["material"], **{k: base + s * hs})))
                for s in (1, -1)]
        fd_check(k, gp[k], (vals[0] - vals[1]) / (2 * hs))
    for k in ("x", "vy", "sxy"):
        field = {"sxy": "txy"}.get(k, k)
        hs = INIT_STEPS[k]
        vals = []
        for s in (1, -1):
            p2 = {q: v.copy() for q, v in w["particles"].items()}
            p2[k][11] += s * hs
            vals.append(linear_loss(wts, gpu_final(w, steps, particles=p2)))
        fd_check(f"{k}[11]", gi[field][11], (vals[0] - vals[1]) / (2 * hs))


@pytest.mark.gpu
def test_torch_autograd_bridge():
    torch = pytest.importorskip("torch")
    import usl_autograd as ag
    w = small_case(seed=9)
    steps = 12
    p = w["particles"]
    a = pympm.Adjoint(gpu_config(w), p, checkpoint_every=5)
    mat = w["material"]
    params = torch.tensor([mat[k] for k in pympm.PARAMS], dtype=torch.float64,
                          requires_grad=True)
    ins = [torch.tensor(p[k], dtype=torch.float64, requires_grad=True)
           for k in ("x", "y", "vx", "vy", "v", "sxx", "sxy", "syy")]
    outs = ag.UslRun.apply(a, steps, torch.tensor(p["m"]), params, *ins)
    target = [o.detach() + 0.01 for o in outs[:2]]
    loss = ((outs[0] - target[0]) ** 2).sum() + ((outs[1] - target[1]) ** 2).sum()
    loss.backward()
    # the same derivative through the C ABI directly
    b = pympm.Adjoint(gpu_config(w), p, checkpoint_every=5)
    b.forward(steps)
    seed = {"x": 2 * (b.get("x") - target[0].numpy()), "y": 2 * (b.get("y") - target[1].numpy())}
    gp, gi = b.backward(seed)
    b.close()
    a.close()
    assert np.allclose(params.grad.numpy(), [gp[k] for k in pympm.PARAMS], rtol=1e-12, atol=0)
    assert np.allclose(ins[0].grad.numpy(), gi["x"], rtol=1e-12, atol=1e-300)
    assert np.allclose(ins[6].grad.numpy(), gi["txy"], rtol=1e-12, atol=1e-300)


@pytest.mark.gpu
def test_adjoint_workload_gradient():
    """Full adjoint workload: dL/dmu_s and dL/dE of L = |x - x*|^2 + |y - y*|^2
    (x* from the target constants) against central differences."""
    w = workloads.adjoint_block()
    steps = w["steps"]
    say(f"adjoint workload: {len(w['particles']['x'])} particles, {steps} steps")
    tgt = gpu_final(w, steps, workloads.ADJOINT_TARGET)
    a = pympm.Adjoint(gpu_config(w), w["particles"], checkpoint_every=20)
    a.forward(steps)
    x, y = a.get("x"), a.get("y")
    gp, _ = a.backward({"x": 2 * (x - tgt["x"]), "y": 2 * (y - tgt["y"])})
    a.close()

    def loss(mat):
        f = gpu_final(w, steps, mat)
        return float(np.sum((f["x"] - tgt["x"]) ** 2) + np.sum((f["y"] - tgt["y"]) ** 2))

    for k in ("mu_s", "E"):
        base = w["material"][k]
        hs = 1e-5 * abs(base)
        fd = (loss(dict(w["material"], **{k: base + hs}))
              - loss(dict(w["material"], **{k: base - hs}))) / (2 * hs)
        fd_check(k, gp[k], fd)


# ---------------------------------------------------------------------------
# The adjoint workload as specified: loss = squared error of the final
# runout front and particle centroid (calibrate.py), gradients with respect
# to mu_s, E and the initial velocities, and the calibration of mu_s and E.

def _c5_target(w):
    import calibrate
    s = pympm.Sim(gpu_config(w, workloads.ADJOINT_TARGET), w["particles"])
    s.step(w["steps"])
    target, _ = calibrate.observables(s.get("x"), s.get("y"), s.active())
    s.close()
    return target


def _c5_loss(a, w, target, material=None, particles=None):
    import calibrate
    if material is not None:
        a.set_material(**material)
    if particles is not None:
        a.set_particles(particles)
    a.forward(w["steps"])
    return calibrate.loss_and_seed(a.get("x"), a.get("y"), a.active(), target)


@pytest.mark.gpu
def test_c5_gradient_against_finite_differences():
    """dL/dmu_s and dL/dE against central differences with a relative step
    of 1e-5, and dL/d(initial vx, vy) of the front particle and of the
    particles with the largest components against steps of 1e-6 m/s; bound
    1e-5.  (At a relative step of 1e-6 in the constants the rounding of the
    200-step forward run is of the order of the bound: the differences of
    the observables are then a few ulps of positions near 0.5 m.)"""
    import calibrate
    w = workloads.adjoint_block()
    say(f"C5 gradient: {len(w['particles']['x'])} particles, {w['steps']} steps")
    target = _c5_target(w)
    a = pympm.Adjoint(gpu_config(w), w["particles"], checkpoint_every=20)
    loss, seed = _c5_loss(a, w, target)
    _, front = calibrate.observables(a.get("x"), a.get("y"), a.active())
    gp, gi = a.backward(seed)
    base = dict(w["material"])
    for k in ("mu_s", "E"):
        for rel_step in (1e-5, 1e-6):
            hs = rel_step * abs(base[k])
            lp, _ = _c5_loss(a, w, target, material=dict(base, **{k: base[k] + hs}))
            lm, _ = _c5_loss(a, w, target, material=dict(base, **{k: base[k] - hs}))
            fd = (lp - lm) / (2 * hs)
            if rel_step == 1e-5:
                fd_check(f"{k} (h = {rel_step:g} |{k}|)", gp[k], fd)
            else:
                say(f"  {k} (h = {rel_step:g} |{k}|): rel {abs(gp[k] - fd) / abs(fd):.2e} "
                    "(reported)")
    a.set_material(**base)
    picks = [("vx", front)] + [("vx", int(i)) for i in np.argsort(-np.abs(gi["vx"]))[:2]] \
        + [("vy", int(i)) for i in np.argsort(-np.abs(gi["vy"]))[:2]]
    for f, i in picks:
        vals = []
        for sgn in (1, -1):
            p2 = {q: v.copy() for q, v in w["particles"].items()}
            p2[f][i] += sgn * 1e-6
            vals.append(_c5_loss(a, w, target, particles=p2)[0])
        fd_check(f"{f}0[{i}]", gi[f][i], (vals[0] - vals[1]) / 2e-6)
    a.close()


@pytest.mark.gpu
def test_c5_gradient_matches_cpu_adjoint():
    """The first 20 steps of the workload: the GPU reverse sweep of the
    front-and-centroid loss against the FP64 CPU adjoint (numpy
    restatement over the oracle's recorded run)."""
    import calibrate
    w = dict(workloads.adjoint_block(), steps=20)
    say(f"C5 vs CPU adjoint: {len(w['particles']['x'])} particles, 20 steps")
    recs, final = AO.record(w, 20)
    target = np.array([0.75, 0.5, 0.25])
    _, seed = calibrate.loss_and_seed(final["x"], final["y"], final["active"], target)
    lam0, grad = AO.adjoint(recs, seed, w["material"], w["N"], w["h"], w["dt"])
    a = pympm.Adjoint(gpu_config(w), w["particles"], checkpoint_every=20)
    loss, gseed = _c5_loss(a, w, target)
    assert all(np.array_equal(gseed[k], seed[k]) for k in seed)
    gp, gi = a.backward(gseed)
    a.close()
    for k in ("mu_s", "E"):
        i = AO.PARAMS.index(k)
        rel = abs(gp[k] - grad[i]) / abs(grad[i])
        say(f"  d/d{k}: gpu {gp[k]:.12e} cpu {grad[i]:.12e} rel {rel:.1e}")
        assert rel <= PARITY_TOL
    for k in ("vx", "vy"):
        scale = np.abs(lam0[k]).max()
        assert np.abs(gi[k] - lam0[k]).max() <= PARITY_TOL * scale, k


@pytest.mark.gpu
def test_c5_calibration_recovers_mu_s_and_E():
    """Gauss-Newton on the front and centroid from mu_s = 0.38, E = 1e6
    recovers the target run's mu_s = 0.42 and E = 2e6."""
    import calibrate
    w = workloads.adjoint_block()
    target = _c5_target(w)
    a = pympm.Adjoint(gpu_config(w), w["particles"], checkpoint_every=20)
    out = calibrate.calibrate(a, w["steps"], target, w["material"]["mu_s"], w["material"]["E"],
                              log=say)
    a.close()
    tgt = workloads.ADJOINT_TARGET
    assert abs(out["mu_s"] / tgt["mu_s"] - 1) <= 1e-6
    assert abs(out["E"] / tgt["E"] - 1) <= 1e-6
    assert out["iterations"] <= 10


@pytest.mark.gpu
def test_torch_bridge_rejects_a_stale_backward():
    """A second forward through the same Adjoint before the first node's
    backward: that backward raises instead of differentiating the second
    run."""
    torch = pytest.importorskip("torch")
    import usl_autograd as ag
    w = small_case(seed=9)
    p = w["particles"]
    a = pympm.Adjoint(gpu_config(w), p, checkpoint_every=5)
    mat = w["material"]

    def run(scale):
        params = torch.tensor([mat[k] for k in pympm.PARAMS], dtype=torch.float64,
                              requires_grad=True)
        ins = [torch.tensor(p[k] * (scale if k == "vx" else 1.0), dtype=torch.float64,
                            requires_grad=True)
               for k in ("x", "y", "vx", "vy", "v", "sxx", "sxy", "syy")]
        outs = ag.UslRun.apply(a, 6, torch.tensor(p["m"]), params, *ins)
        return (outs[0] ** 2).sum()

    first = run(1.0)
    second = run(2.0)
    with pytest.raises(RuntimeError, match="forward again"):
        first.backward()
    second.backward()      # the live run's backward works
    a.close()


DEVICE_BRIDGE_SCRIPT = r"""
import sys
import torch                     # first: the library binds to torch's HIP runtime
assert torch.cuda.is_available()
sys.path[:0] = ["mpm-2d_amd", "oracle", "tests"]
import numpy as np
import pympm
import usl_autograd as ag
from test_adjoint import small_case, gpu_config
w = small_case(seed=9); p = w["particles"]; mat = w["material"]
res = {}
for dev in ("cpu", "cuda"):
    a = pympm.Adjoint(gpu_config(w), p, checkpoint_every=5)
    params = torch.tensor([mat[k] for k in pympm.PARAMS], dtype=torch.float64, device=dev,
                          requires_grad=True)
    ins = [torch.tensor(p[k], dtype=torch.float64, device=dev, requires_grad=True)
           for k in ("x", "y", "vx", "vy", "v", "sxx", "sxy", "syy")]
    outs = ag.UslRun.apply(a, 12, torch.tensor(p["m"], device=dev), params, *ins)
    assert all(o.device.type == dev for o in outs)
    loss = ((outs[0] - 0.3) ** 2).sum() + ((outs[1] - 0.2) ** 2).sum() + (outs[5] ** 2).sum()
    loss.backward()
    res[dev] = [params.grad] + [t.grad for t in ins] + [o.detach() for o in outs]
    assert all(t.device.type == dev for t in res[dev])
    a.close()
same = [bool(np.array_equal(a.cpu().numpy(), b.cpu().numpy()))
        for a, b in zip(res["cpu"], res["cuda"])]
print("device bridge bitwise equal to the host bridge:", same, flush=True)
assert all(same)
"""


@pytest.mark.gpu
def test_torch_bridge_device_resident():
    """The bridge with CUDA tensors (device-to-device copies) against the
    same run with CPU tensors (host staging): gradients and outputs bitwise
    equal.  In a fresh process that imports torch before the library, so
    both use one HIP runtime (this test process loaded the library first)."""
    pytest.importorskip("torch")
    import os
    import subprocess
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    say("case torch bridge with CUDA tensors, fresh process")
    r = subprocess.run([sys.executable, "-c", DEVICE_BRIDGE_SCRIPT], cwd=repo, timeout=300,
                       capture_output=True, text=True)
    say(r.stdout[-2000:] + r.stderr[-2000:])
    assert r.returncode == 0
