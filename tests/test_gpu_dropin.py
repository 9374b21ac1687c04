"""The drop-in libmpm (mpm-2d_amd/host/libmpm_gpu.c) end to end: input files
in the reference formats, plugins loaded by dlopen, T worker threads and the
reference driver's loop (mpm-2d_amd/host/mpm_usl_run.c), checked bitwise
against the CPU oracle.  Both the device path (plugins exporting GPU
descriptors) and the host fallbacks (plugins without them: host stress
update between the kernels, host bc_time_varying with mask upload) run."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
import workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "mpm-2d_amd")
RUN = os.path.join(PKG, "bin", "mpm_usl_run")
COLS = ["x", "y", "vx", "vy", "v", "sxx", "sxy", "syy", "tzz", "gammap", "gammadotp", "ux", "uy"]


def read_state(path):
    lines = open(path).read().split("\n")
    t = float.fromhex(lines[0].split()[0])
    rows = [ln.split() for ln in lines[1:] if ln.strip()]
    data = {k: np.array([float.fromhex(r[i]) for r in rows]) for i, k in enumerate(COLS)}
    data["active"] = np.array([int(r[-1]) for r in rows], dtype=np.int32)
    return t, data


def run_dropin(tmp_path, w, steps, plugin_dir, threads=3, frames=False):
    g, p = tmp_path / "grid.txt", tmp_path / "particles.txt"
    workloads.write_reference_inputs(w, g, p)
    mat = os.path.join(plugin_dir, "g_local_mu2.so")
    bc = os.path.join(plugin_dir, "box.so" if w["bc_kind"] == 1 else "square_silo.so")
    mprops = ",".join(repr(float(w["material"][k])) for k in
                      ("E", "nu", "mu_s", "mu_2", "I_0", "grains_d", "grains_rho", "rho_crit"))
    cmd = [RUN, "-g", str(g), "-p", str(p), "-u", mat, "-b", bc, "-t", str(threads),
           "--dt", repr(w["dt"]), "--steps", str(steps), "--mprops", mprops,
           "--bprops", ",".join(repr(float(v)) for v in w["bc_fp64"]),
           "--state", str(tmp_path / "state.txt")]
    if w["bc_int"]:
        cmd += ["--biprops", ",".join(str(int(v)) for v in w["bc_int"])]
    if frames:
        cmd += ["--sample-rate", "20000", "--frames", str(tmp_path / "frames.txt")]
    print("case", " ".join(cmd[1:4]), plugin_dir, flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    return read_state(tmp_path / "state.txt")


def oracle_state(w, steps):
    o = pyoracle.from_workload(w, num_threads=2)
    o.run(steps)
    return o.time(), o.particles()


def check(got, ref):
    t, d = got
    tr, r = ref
    assert t == tr
    ref_map = {"sxx": "txx", "sxy": "txy", "syy": "tyy"}
    for k in COLS:
        assert np.array_equal(d[k], r[ref_map.get(k, k)]), k
    assert np.array_equal(d["active"], r["active"])


# dropin runs with the reference's loading constants (g = 9.8, ramp 0.1 s)
CASES = {
    "box": lambda: workloads.small_box(cells=16, bx=5, by=9),
    "silo": lambda: dict(workloads.silo(seed=5, cells=32), g_mag=9.8),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_dropin_device_plugins(tmp_path, case):
    w = CASES[case]()
    check(run_dropin(tmp_path, w, 60, os.path.join(PKG, "lib"), frames=True),
          oracle_state(w, 60))
    frames = open(tmp_path / "frames.txt").read()
    assert frames.startswith("0 ")


@pytest.mark.parametrize("case", sorted(CASES))
def test_dropin_host_fallbacks(tmp_path, case):
    w = CASES[case]()
    check(run_dropin(tmp_path, w, 40, os.path.join(PKG, "lib", "host_only"), threads=4),
          oracle_state(w, 40))
