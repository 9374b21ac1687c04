"""GPU parity: libmpm_hip.so against the CPU oracle on identical inputs.

The bar is bitwise equality of every particle and node field (compared as
values, so +0 == -0).  That is stricter than the 1e-12 per-step / 1e-8
after-100-steps tolerance the workloads are specified with; the tolerance is
asserted as well so a failure reports both.
"""
import os
import sys

import numpy as np
import pytest

import pympm
import pyoracle
import workloads

pytestmark = pytest.mark.gpu

PFIELDS = ["x", "y", "vx", "vy", "v", "txx", "txy", "tyy", "tzz", "gammap", "gammadotp",
           "ux", "uy"]
AUX = ["exx", "exy", "eyy", "wxy", "ax", "ay", "xl", "yl"]
NFIELDS = ["m", "mx", "my", "fx", "fy", "vx", "vy"]
REL_TOL_1STEP = 1e-12
REL_TOL_100 = 1e-8


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_sim(w, particles=None, flags=0, cap=0, bc_kind=None):
    cfg = pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"],
                            bc_kind=w.get("bc_kind", 0) if bc_kind is None else bc_kind,
                            bc_fp64=w.get("bc_fp64", ()), bc_int=w.get("bc_int", ()),
                            g_mag=w.get("g_mag", 9.8), ramp_time=w.get("ramp_time", 0.1),
                            bin_capacity=cap, flags=flags,
                            material_kind=w.get("material_kind", 0))
    return pympm.Sim(cfg, particles if particles is not None else w["particles"])


def compare(gpu, ref, fields, what, tol=0.0):
    bad = []
    for k in fields:
        a, b = gpu[k], ref[k]
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        if not same.all():
            scale = np.maximum(np.abs(b), np.abs(b).max() * 1e-3 + 1e-300)
            rel = np.max(np.abs(a - b) / scale)
            bad.append(f"{k}: {int((~same).sum())} differ, max rel {rel:.3e}")
            if tol and rel > tol:
                raise AssertionError(f"{what}: {k} beyond tolerance {tol}: {bad[-1]}")
    assert not bad, f"{what}: not bitwise: " + "; ".join(bad)


def run_both(w, steps, threads=4, particles=None, flags=pympm.STORE_AUX, cap=0):
    parts = particles if particles is not None else w["particles"]
    say(f"case {w['name']}: np={len(parts['x'])} N={w['N']} steps={steps}")
    sim = make_sim(w, particles=parts, flags=flags, cap=cap)
    o = pyoracle.from_workload(w, num_threads=threads, particles=parts)
    if steps:
        sim.step(steps)
        sim.sync()
        o.run(steps)
    return sim, o


# ---------------------------------------------------------------------------

def test_device_arithmetic_is_ieee():
    say("device exp/div/sqrt vs host")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-2.0 ** -8, 2.0 ** -8, 200000) * 10.0 ** rng.uniform(-8, 0, 200000),
                        rng.uniform(-740, 709, 100000), [0.0, -0.0, 1e-320, -1e-320, 700.0]])
    dev = pympm.eval_ops(0, x)
    host = np.array([pympm.exp_host(v) for v in x[:20000]])
    assert np.array_equal(dev[:20000].view(np.uint64), host.view(np.uint64))
    a = rng.normal(size=300000) * 10.0 ** rng.uniform(-300, 300, 300000)
    b = rng.normal(size=300000) * 10.0 ** rng.uniform(-300, 300, 300000)
    assert np.array_equal(pympm.eval_ops(1, a, b).view(np.uint64), (a / b).view(np.uint64))
    s = np.abs(a)
    assert np.array_equal(pympm.eval_ops(2, s).view(np.uint64), np.sqrt(s).view(np.uint64))


def test_one_step_bitwise_all_fields():
    w = workloads.small_box(cells=16, bx=5, by=9)
    sim, o = run_both(w, 1)
    gp = {k: sim.get(k) for k in PFIELDS + AUX}
    rp = o.particles()
    compare(gp, rp, PFIELDS + AUX, "particles after 1 step", tol=REL_TOL_1STEP)
    gn = {k: sim.node(k) for k in NFIELDS}
    compare(gn, o.nodes(), NFIELDS, "nodes after 1 step", tol=REL_TOL_1STEP)
    assert np.array_equal(sim.active(), rp["active"])
    t, by = sim.time()
    assert t == o.time()
    assert by == rp["by"][0]


CASES = {
    "box_reference_material": lambda: workloads.small_box(cells=16, bx=5, by=9, steps=100),
    "box_baseline_material": lambda: workloads.small_box(cells=32, bx=9, by=20, steps=100,
                                                         material=workloads.WORKLOAD_MATERIAL,
                                                         g_mag=9.81),
    "silo_64": lambda: dict(workloads.silo(seed=5, cells=64), steps=100),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_100_steps_bitwise(case):
    w = CASES[case]()
    sim, o = run_both(w, w["steps"])
    rp = o.particles()
    compare({k: sim.get(k) for k in PFIELDS + AUX}, rp, PFIELDS + AUX,
            f"{case} after {w['steps']} steps", tol=REL_TOL_100)
    compare({k: sim.node(k) for k in NFIELDS}, o.nodes(), NFIELDS, f"{case} nodes")
    assert np.array_equal(sim.active(), rp["active"])


def test_shuffled_particle_order():
    """The fold order is by particle index inside each cell, whatever the
    memory order of the particles."""
    w = workloads.small_box(cells=16, bx=6, by=10, steps=60)
    rng = np.random.default_rng(4)
    perm = rng.permutation(len(w["particles"]["x"]))
    parts = {k: v[perm].copy() for k, v in w["particles"].items()}
    sim, o = run_both(w, w["steps"], particles=parts)
    compare({k: sim.get(k) for k in PFIELDS}, o.particles(), PFIELDS, "shuffled")


def test_tile_dispatch_order_leaves_results_unchanged():
    """P2G dispatches the tiles around the particles' bounding box, taken
    when the particles are uploaded, before the other tiles.  The order is a
    hint: particles that fly out of the box across tile rows and columns
    match the oracle bitwise, and so does the same run with the hint dropped
    (positions set again)."""
    w = workloads.small_box(cells=96, bx=20, by=12, steps=60)
    p = workloads.block(20, 12, 96, seed=11, cell_x0=40, cell_y0=50)
    p["vx"][:] = 0.4 * w["h"] / w["dt"]     # 24 cells over the run
    p["vy"][:] = -0.2 * w["h"] / w["dt"]
    sim, o = run_both(w, w["steps"], particles=p)
    rp, rn = o.particles(), o.nodes()
    compare({k: sim.get(k) for k in PFIELDS}, rp, PFIELDS, "hinted order")
    compare({k: sim.node(k) for k in NFIELDS}, rn, NFIELDS, "hinted order nodes")
    assert sim.get("x").min() > (40 + 20) * w["h"]   # all left the hinted box
    sim.close()
    sim = make_sim(w, particles=p)
    sim.set("x", p["x"])
    sim.step(w["steps"])
    sim.sync()
    compare({k: sim.get(k) for k in PFIELDS}, rp, PFIELDS, "row order")
    compare({k: sim.node(k) for k in NFIELDS}, rn, NFIELDS, "row order nodes")


def test_host_mask_equals_box_plugin():
    """MPM_BC_MASK with the mask the box plugin generates gives the same
    trajectory as the device-evaluated box."""
    w = workloads.small_box(cells=16, bx=5, by=9, steps=50)
    o = pyoracle.from_workload(w)
    o.run(1)
    m = o.bc_mask()
    mask = ((m[:, 0] != 0) * 1 + (m[:, 1] != 0) * 2).astype(np.uint8)
    say("case mask-vs-box")
    a = make_sim(w)
    b = make_sim(w, bc_kind=pympm.BC_MASK)
    b.set_bc_mask(mask)
    a.step(50)
    b.step(50)
    compare({k: b.get(k) for k in PFIELDS}, {k: a.get(k) for k in PFIELDS}, PFIELDS, "mask vs box")


def test_particles_leaving_the_grid():
    """Particles thrown out of [0,1)^2 are deactivated one step after they
    leave (process_usl.c:660-665) and stop contributing."""
    w = workloads.small_box(cells=16, bx=4, by=4, steps=40, bc_kind=0, g_mag=0.0)
    parts = {k: v.copy() for k, v in w["particles"].items()}
    parts["x"] += 0.02
    parts["y"] += 0.02
    parts["vx"][:] = -300.0        # 0.12 in 40 steps of 1e-5: many leave
    parts["vy"][::3] = -100.0
    sim, o = run_both(w, w["steps"], particles=parts)
    rp = o.particles()
    assert 0 < rp["active"].sum() < len(rp["active"])
    assert np.array_equal(sim.active(), rp["active"])
    compare({k: sim.get(k) for k in PFIELDS}, rp, PFIELDS, "leaving the grid")


def test_bin_overflow_is_reported():
    """More particles in one cell than the bin capacity: the step reports
    MPM_HIP_ERR_BIN_OVERFLOW instead of returning wrong sums."""
    w = workloads.small_box(cells=16, bx=1, by=1, steps=1)
    parts = workloads.block(1, 1, 16, ppc_side=5, seed=3, cell_x0=3, cell_y0=3)   # 25 in one cell
    say("case overflow cap=16")
    with pytest.raises(pympm.MpmHipError) as e:
        sim = make_sim(w, particles=parts, cap=16)
        sim.step(1)
        sim.sync()
    assert e.value.code == -4
    sim = make_sim(w, particles=parts, cap=32)
    sim.step(2)
    sim.sync()


SHAPE_CASES = {
    # 4 particles per cell: one staging batch per 8 x 4 tile
    "silo_64": lambda: dict(workloads.silo(seed=5, cells=64), steps=60),
    # 36 per cell: a full 9 x 5-cell region (1,620 records) spans batches
    "dense_36": lambda: dict(workloads.small_box(cells=32, bx=10, by=12, steps=30), cap=64,
                             particles=workloads.block(10, 12, 32, seed=13, ppc_side=6,
                                                       jitter=0.02, cell_x0=11, cell_y0=2)),
    # 64 per cell: one colour of a region (15 cells) spans several runs
    "dense_64": lambda: dict(workloads.small_box(cells=32, bx=8, by=8, steps=8), cap=64,
                             particles=workloads.block(8, 8, 32, seed=17, ppc_side=8,
                                                       jitter=0.02, cell_x0=12, cell_y0=3)),
}


@pytest.mark.parametrize("scheme", [pympm.SCHEME_USL, pympm.SCHEME_MUSL])
@pytest.mark.parametrize("tile", [1, 2, 3])
@pytest.mark.parametrize("case", sorted(SHAPE_CASES))
def test_p2g_tile_shapes_bitwise(case, tile, scheme):
    """Every P2G tile shape (16 x 8 nodes, colours staged one by one in
    batches of 256 or of 512 records; 8 x 4 nodes, colours merged into
    shared staging batches), forced through the config on the same inputs,
    in the USL and MUSL steps: bitwise against the oracle."""
    w = dict(SHAPE_CASES[case](), scheme=scheme)
    p = w["particles"]
    say(f"case {case}: tile {tile} scheme {scheme} np={len(p['x'])} steps={w['steps']}")
    cfg = pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"],
                            bc_kind=w["bc_kind"], bc_fp64=w["bc_fp64"], bc_int=w["bc_int"],
                            g_mag=w["g_mag"], ramp_time=w["ramp_time"], scheme=scheme,
                            bin_capacity=w.get("cap", 0), p2g_tile=tile)
    sim = pympm.Sim(cfg, p)
    sim.step(w["steps"])
    sim.sync()
    o = pyoracle.from_workload(w, num_threads=4, particles=p)
    o.run(w["steps"])
    compare({k: sim.get(k) for k in PFIELDS}, o.particles(), PFIELDS,
            f"{case} tile {tile} scheme {scheme}", tol=REL_TOL_100)
    compare({k: sim.node(k) for k in NFIELDS}, o.nodes(), NFIELDS, f"{case} nodes")


def test_p2g_tile_config_is_validated():
    w = workloads.small_box(cells=16, bx=5, by=9)
    cfg = pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"], p2g_tile=4)
    with pytest.raises(pympm.MpmHipError):
        pympm.Sim(cfg, w["particles"])


def test_column_collapse_1000_steps():
    """The 64 x 64 column collapse (8,192 particles), all 1000 steps."""
    w = workloads.column_collapse()
    sim, o = run_both(w, w["steps"], threads=8, flags=0)
    compare({k: sim.get(k) for k in PFIELDS}, o.particles(), PFIELDS, "column collapse",
            tol=REL_TOL_100)


def test_full_size_silo_five_steps():
    """The 512 x 512 silo with 1,048,576 particles at full size: five steps
    bitwise against the oracle on all host threads."""
    w = workloads.silo()
    threads = min(16, os.cpu_count() or 1)
    sim, o = run_both(w, 5, threads=threads, flags=0)
    compare({k: sim.get(k) for k in PFIELDS}, o.particles(), PFIELDS, "silo 1M, 5 steps")
    compare({k: sim.node(k) for k in NFIELDS}, o.nodes(), NFIELDS, "silo 1M nodes")


def test_full_size_silo_100_steps_bitwise():
    """The headline workload itself (512 x 512 cells, 1,048,576 particles)
    over 100 steps, every particle field bitwise against the oracle on all
    host threads (with the P2G dispatch order the full-size run uses)."""
    w = workloads.silo()
    threads = min(16, os.cpu_count() or 1)
    sim, o = run_both(w, 100, threads=threads, flags=0)
    compare({k: sim.get(k) for k in PFIELDS}, o.particles(), PFIELDS, "silo 1M, 100 steps",
            tol=REL_TOL_100)
    compare({k: sim.node(k) for k in NFIELDS}, o.nodes(), NFIELDS, "silo 1M nodes, 100 steps")
    assert np.array_equal(sim.active(), o.particles()["active"])


def test_full_size_silo_properties_after_200_steps():
    """Size-independent properties at full size over more steps: mass is
    conserved, a rerun is bitwise identical, nothing leaves the silo yet."""
    w = workloads.silo()
    say("case silo 1M x 200 steps twice")
    runs = []
    for _ in range(2):
        sim = make_sim(w)
        sim.step(200)
        sim.sync()
        runs.append({k: sim.get(k) for k in ("x", "y", "vy", "txx", "v")})
        act = sim.active()
        nm = sim.node("m")
        sim.close()
    for k in runs[0]:
        assert np.array_equal(runs[0][k].view(np.uint64), runs[1][k].view(np.uint64)), k
    assert act.sum() == len(act)
    assert nm.sum() == pytest.approx(w["particles"]["m"].sum(), rel=1e-12)
    assert np.all(runs[0]["vy"] <= 1e-9)


@pytest.mark.parametrize("case", ["usl", "musl", "nonlocal"])
def test_timed_steps_follow_the_same_trajectory(case):
    """Steps taken through the timing calls (the per-kernel pass, whose P2G
    and plain G2P dispatches record their own events, and the span) follow
    the trajectory of plain steps bitwise, and report consistent times:
    each part's time positive and their sum within the span."""
    if case == "nonlocal":
        w = workloads.heap(cells=32)
        extra = dict(nonlocal_material=w["nonlocal_material"], cg_rel_tol=1e-12,
                     cg_max_iter=50000)
    else:
        w = workloads.silo(seed=5, cells=64)
        extra = dict(scheme=pympm.SCHEME_MUSL if case == "musl" else pympm.SCHEME_USL)
    cfg = pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"],
                            bc_kind=w.get("bc_kind", 0), bc_fp64=w.get("bc_fp64", ()),
                            bc_int=w.get("bc_int", ()), g_mag=w["g_mag"],
                            ramp_time=w["ramp_time"], **extra)
    say(f"case timed steps, {case}")
    a = pympm.Sim(cfg, w["particles"])
    b = pympm.Sim(cfg, w["particles"])
    ms = a.time_steps(12)
    span = a.time_span(8)
    b.step(20)
    b.sync()
    compare({k: a.get(k) for k in PFIELDS}, {k: b.get(k) for k in PFIELDS}, PFIELDS,
            "timed vs plain steps")
    assert ms[1] > 0 and ms[2] > 0 and ms[1] + ms[2] <= ms[0] * 1.001
    assert span > 0
    a.close()
    b.close()
