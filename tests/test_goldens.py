"""Committed fixtures (tests/golden/*.npz, written by make_oracle_goldens.py)
against the oracle (CPU) and against the HIP library (GPU)."""
import glob
import json
import os

import numpy as np
import pytest

import pympm
import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
PF = ["x", "y", "vx", "vy", "v", "txx", "txy", "tyy", "tzz", "gammap", "gammadotp", "ux", "uy"]
NF = ["m", "mx", "my", "fx", "fy", "vx", "vy"]


def _load(path):
    z = np.load(path, allow_pickle=False)
    cfg = json.loads(str(z["config"]))
    parts = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    return cfg, parts, z


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_fixture(path):
    cfg, parts, z = _load(path)
    lib = pyoracle.load()
    q = pyoracle.make_params(lib, cfg["N"], cfg["h"], cfg["dt"], material=cfg["material"],
                             bc_kind=cfg["bc_kind"], bc_fp64=cfg["bc_fp64"],
                             bc_int=cfg["bc_int"], g_mag=cfg["g_mag"],
                             ramp_time=cfg["ramp_time"], num_threads=3)
    o = pyoracle.Oracle(q, parts)
    o.run(cfg["steps"])
    p = o.particles()
    for k in PF:
        assert np.array_equal(p[k].view(np.uint64), z[f"out_{k}"].view(np.uint64)), k
    assert np.array_equal(p["active"], z["out_active"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_reproduces_fixture(path):
    cfg, parts, z = _load(path)
    print(f"fixture {os.path.basename(path)}", flush=True)
    c = pympm.make_config(cfg["N"], cfg["h"], cfg["dt"], material=cfg["material"],
                          bc_kind=cfg["bc_kind"], bc_fp64=cfg["bc_fp64"], bc_int=cfg["bc_int"],
                          g_mag=cfg["g_mag"], ramp_time=cfg["ramp_time"])
    sim = pympm.Sim(c, parts)
    sim.step(cfg["steps"])
    sim.sync()
    for k in PF:
        assert np.array_equal(sim.get(k), z[f"out_{k}"]), k
    for k in NF:
        assert np.array_equal(sim.node(k), z[f"node_{k}"]), k
    assert np.array_equal(sim.active(), z["out_active"])
