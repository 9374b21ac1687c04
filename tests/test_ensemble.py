"""Ensembles: many independent jobs stepped by shared launches.

Every member of a batch must follow, bit for bit, the trajectory it has when
run alone -- on the CPU oracle (the reference step restated) and on a
single-job GPU context.  The batch replays its steps from a captured HIP
graph, so runs longer than one replay check every replay against the
eagerly stepped single-job contexts.
"""
import sys

import numpy as np
import pytest

import pympm
import pyoracle
import workloads

PFIELDS = ["x", "y", "vx", "vy", "v", "txx", "txy", "tyy", "tzz", "gammap", "gammadotp",
           "ux", "uy"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def member_config(w, cap=0, tile=0):
    return pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"],
                             bc_kind=w["bc_kind"], bc_fp64=w["bc_fp64"], bc_int=w["bc_int"],
                             g_mag=w["g_mag"], ramp_time=w["ramp_time"], bin_capacity=cap,
                             p2g_tile=tile)


def same(a, b):
    return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def assert_member_equal(batch, k, ref_state, ref_active, what):
    bad = [f for f in PFIELDS if not same(batch.get(k, f), ref_state[f])]
    assert not bad, f"{what}: member {k} differs in {bad}"
    assert np.array_equal(batch.active(k), ref_active), f"{what}: member {k} active flags"


# ---------------------------------------------------------------------------
# CPU: the workload generator

def test_ensemble_workload_shape():
    e = workloads.ensemble()
    counts = np.array([len(m["particles"]["x"]) for m in e])
    sides = np.array([m["N"] - 1 for m in e])
    assert len(e) == 256
    assert np.all(sides % 4 == 0)
    assert np.array_equal(counts, 2 * sides ** 2)           # n/4 x n/2 cells x 16
    assert 2000 <= counts.min() and counts.max() <= 66000
    assert 4.4e6 < counts.sum() < 4.9e6
    mu = np.array([m["material"]["mu_s"] for m in e])
    E = np.array([m["material"]["E"] for m in e])
    assert np.all((0.3 <= mu) & (mu < 0.5)) and np.all((5e5 <= E) & (E < 5e6))
    # members are reproducible and independent of the ensemble size
    e2 = workloads.ensemble(members=3)
    for a, b in zip(e[:3], e2):
        assert a["N"] == b["N"] and np.array_equal(a["particles"]["x"], b["particles"]["x"])


# ---------------------------------------------------------------------------
# GPU

@pytest.mark.gpu
def test_batch_matches_oracle_per_member():
    ens = workloads.ensemble(members=6, sides=[8, 12, 16, 20, 12, 8])
    steps = 61   # 7 graph replays of 8 steps plus eager steps on both sides
    say(f"batch of {len(ens)} vs oracle, {steps} steps")
    b = pympm.Batch([member_config(w) for w in ens], [w["particles"] for w in ens])
    b.step(steps)
    b.sync()
    for k, w in enumerate(ens):
        o = pyoracle.from_workload(w, num_threads=2)
        o.run(steps)
        ref = o.particles()
        assert_member_equal(b, k, ref, ref["active"], "oracle")
        t, _ = b.time(k)
        assert t == o.time()
        o.close()
    b.close()


@pytest.mark.gpu
def test_batch_matches_single_contexts():
    ens = workloads.ensemble(members=24, seed=5, p_max=16384)
    steps = 100
    say(f"batch of {len(ens)} vs single contexts, {steps} steps")
    b = pympm.Batch([member_config(w) for w in ens], [w["particles"] for w in ens])
    done = 0
    for chunk in (1, 8, 3, 16, 72):     # replays from both bin parities
        b.step(chunk)
        done += chunk
    assert done == steps
    b.sync()
    for k, w in enumerate(ens):
        s = pympm.Sim(member_config(w), w["particles"])
        s.step(steps)
        s.sync()
        assert_member_equal(b, k, s.state(PFIELDS), s.active(), "single context")
        assert b.time(k) == s.time()
        s.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [1, 2, 3])
def test_batch_particles_crossing_into_empty_tiles(tile):
    """Members whose blocks fly across tile rows and columns: tiles whose
    regions were empty at upload (checked sixteen at a time by the light-tile
    kernel) fill up and empty again, and tiles occupied at upload empty out;
    every member matches the oracle bitwise, with 16 x 8- and 8 x 4-node
    P2G tiles and 16 x 8-node tiles staged in 512-record batches."""
    members = []
    for k, (n, cx0, cy0, vx, vy) in enumerate([(48, 4, 20, 0.4, -0.1), (64, 30, 30, -0.35, 0.2),
                                               (40, 2, 2, 0.3, 0.3)]):
        w = workloads.small_box(cells=n, bx=6, by=5, steps=0)
        p = workloads.block(6, 5, n, seed=20 + k, cell_x0=cx0, cell_y0=cy0)
        p["vx"][:] = vx * w["h"] / w["dt"]      # cells per step
        p["vy"][:] = vy * w["h"] / w["dt"]
        members.append(dict(w, particles=p))
    steps = 50
    say(f"batch of {len(members)} crossing tiles, {steps} steps, p2g_tile {tile}")
    b = pympm.Batch([member_config(w, tile=tile) for w in members],
                    [w["particles"] for w in members])
    b.step(steps)
    b.sync()
    for k, w in enumerate(members):
        o = pyoracle.from_workload(w, num_threads=2)
        o.run(steps)
        ref = o.particles()
        assert_member_equal(b, k, ref, ref["active"], "crossing tiles")
        moved = np.abs(ref["x"] - w["particles"]["x"]).min() * (w["N"] - 1)
        assert moved > 12.0, f"member {k} moved only {moved} cells"
        o.close()
    b.close()


@pytest.mark.gpu
def test_autotune_keeps_the_trajectory():
    ens = workloads.ensemble(members=10, seed=8, p_max=8192)
    say(f"autotune on a batch of {len(ens)}")
    a = pympm.Batch([member_config(w) for w in ens], [w["particles"] for w in ens])
    ref = pympm.Batch([member_config(w) for w in ens], [w["particles"] for w in ens])
    a.step(10)
    cap, ms = a.autotune(trial_steps=5)
    assert cap in (16, 32, 64) and ms[[16, 32, 64].index(cap)] > 0
    a.step(30)
    ref.step(40)
    a.sync()
    ref.sync()
    for k in range(len(ens)):
        assert_member_equal(a, k, ref.state(k, PFIELDS), ref.active(k), "autotuned")
    a.close()
    ref.close()


@pytest.mark.gpu
def test_batch_rejects_unsupported_members():
    w = workloads.ensemble(members=2, sides=[8, 8])
    cfgs = [member_config(x) for x in w]
    cfgs[1].bc_kind = pympm.BC_MASK
    with pytest.raises(pympm.MpmHipError) as ei:
        pympm.Batch(cfgs, [x["particles"] for x in w])
    assert ei.value.code == -1
    cfgs = [member_config(x) for x in w]
    cfgs[0].material_kind = 1
    with pytest.raises(pympm.MpmHipError):
        pympm.Batch(cfgs, [x["particles"] for x in w])


@pytest.mark.gpu
def test_autotune_launch_shape_keeps_the_trajectory():
    """The search over bin capacity x threads per node restores the state
    after every trial and leaves the trajectory bitwise unchanged, whichever
    pair it keeps."""
    ens = workloads.ensemble(members=12, seed=9, p_max=16384)
    say(f"launch-shape autotune on a batch of {len(ens)}")
    a = pympm.Batch([member_config(w) for w in ens], [w["particles"] for w in ens])
    ref = pympm.Batch([member_config(w) for w in ens], [w["particles"] for w in ens])
    a.step(9)
    cap, roles, ms = a.autotune_launch(trial_steps=5)
    say(f"  kept capacity {cap}, {roles} threads per node; ms/step {ms.tolist()}")
    assert cap in (16, 32, 64) and roles in (2, 3)
    assert ms[[16, 32, 64].index(cap) * 2 + (roles - 2)] == ms[ms > 0].min()
    a.step(31)
    ref.step(40)
    a.sync()
    ref.sync()
    for k in range(len(ens)):
        assert_member_equal(a, k, ref.state(k, PFIELDS), ref.active(k), "launch autotuned")
    a.close()
    ref.close()


@pytest.mark.gpu
def test_batch_small_tiles_match_single_contexts():
    """The batch on 8 x 4-node P2G tiles (all four colours of a tile staged
    at once) against single-job contexts, over graph replays from both bin
    parities."""
    ens = workloads.ensemble(members=16, seed=6, p_max=16384)
    steps = 40
    say(f"batch of {len(ens)} on 8 x 4 tiles vs single contexts, {steps} steps")
    b = pympm.Batch([member_config(w, tile=2) for w in ens], [w["particles"] for w in ens])
    for chunk in (3, 8, 29):
        b.step(chunk)
    b.sync()
    for k, w in enumerate(ens):
        s = pympm.Sim(member_config(w), w["particles"])
        s.step(steps)
        s.sync()
        assert_member_equal(b, k, s.state(PFIELDS), s.active(), "8 x 4 batch")
        s.close()
    b.close()


@pytest.mark.gpu
def test_autotune_tiles_keeps_the_trajectory():
    """The search over P2G tile shape x bin capacity x threads per node
    (eighteen trials, each from the restored state; a shape switch resets the
    tiles' zero flags) leaves the trajectory bitwise unchanged."""
    ens = workloads.ensemble(members=12, seed=10, p_max=16384)
    say(f"tile-shape autotune on a batch of {len(ens)}")
    a = pympm.Batch([member_config(w) for w in ens], [w["particles"] for w in ens])
    ref = pympm.Batch([member_config(w) for w in ens], [w["particles"] for w in ens])
    a.step(11)
    tile, cap, roles, ms = a.autotune_tiles(trial_steps=5)
    say(f"  kept tile {tile}, capacity {cap}, {roles} threads per node; ms/step {ms.tolist()}")
    assert tile in (1, 2, 3) and cap in (16, 32, 64) and roles in (2, 3)
    assert ms[(tile - 1) * 6 + [16, 32, 64].index(cap) * 2 + (roles - 2)] == ms[ms > 0].min()
    a.step(29)
    ref.step(40)
    a.sync()
    ref.sync()
    for k in range(len(ens)):
        assert_member_equal(a, k, ref.state(k, PFIELDS), ref.active(k), "tile autotuned")
    a.close()
    ref.close()
