"""isolin (materialsrc/isolin.c): isotropic linear elasticity with the
Jaumann rate.

CPU: the oracle's update pinned by closed forms written independently of its
component form -- plane-strain Hooke's law in (E, nu) form for a spin-free
rate, and the Jaumann correction W s - s W as a 2x2 matrix product for a pure
spin.  GPU: the library's isolin path bitwise against the oracle, through the
C ABI and through the drop-in libmpm with the repository's isolin plugin
(device descriptor and host fallback).
"""
import os
import subprocess

import numpy as np
import pytest

import pympm
import pyoracle
import workloads

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "mpm-2d_amd")
RUN = os.path.join(PKG, "bin", "mpm_usl_run")


def isolin_point(s, e, dt, E, nu):
    lib = pyoracle.load()
    q = pyoracle.make_params(lib, 17, 1 / 16, dt, material=dict(E=E, nu=nu), material_kind=2)
    sin = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros(3)
    lib.oracle_isolin_point(pyoracle._dptr(sin), *[float(v) for v in e], dt, q,
                            pyoracle._dptr(out))
    return out


def test_isolin_hooke_plane_strain():
    rng = np.random.default_rng(0)
    for _ in range(200):
        E = 10 ** rng.uniform(5, 8)
        nu = rng.uniform(0.05, 0.45)
        dt = 10 ** rng.uniform(-7, -4)
        s = rng.normal(size=3) * 1e3
        exx, exy, eyy = rng.normal(size=3)
        got = isolin_point(s, (exx, exy, eyy, 0.0), dt, E, nu)
        # plane strain: sigma = E / ((1 + nu)(1 - 2 nu)) [(1 - nu) e + nu tr e]
        c = E / ((1 + nu) * (1 - 2 * nu))
        want = s + dt * np.array([c * ((1 - nu) * exx + nu * eyy),
                                  E / (1 + nu) * exy,
                                  c * ((1 - nu) * eyy + nu * exx)])
        assert np.allclose(got, want, rtol=1e-12, atol=1e-9 * np.abs(s).max())


def test_isolin_jaumann_spin():
    rng = np.random.default_rng(1)
    for _ in range(200):
        s = rng.normal(size=3) * 1e3
        w = rng.normal()
        dt = 1e-5
        got = isolin_point(s, (0.0, 0.0, 0.0, w), dt, 1e6, 0.3)
        S = np.array([[s[0], s[1]], [s[1], s[2]]])
        # spin tensor of L with L_xy - L_yx = 2 w (wxy = (L_xy - L_yx) / 2)
        W = np.array([[0.0, w], [-w, 0.0]])
        R = S + dt * (W @ S - S @ W)
        assert np.allclose(got, [R[0, 0], R[0, 1], R[1, 1]], rtol=1e-13, atol=1e-9)
        # the spin leaves the trace unchanged
        assert abs((got[0] + got[2]) - (s[0] + s[2])) <= 1e-12 * np.abs(s).sum()


def isolin_box(cells=16, seed=7, vx=0.3):
    w = workloads.small_box(cells=cells, bx=5, by=9, seed=seed)
    w["material_kind"] = 2
    rng = np.random.default_rng(seed)
    w["particles"]["vx"] = vx * rng.standard_normal(len(w["particles"]["x"]))
    w["particles"]["sxy"] = 50.0 * rng.standard_normal(len(w["particles"]["x"]))
    return w


FIELDS = ["x", "y", "vx", "vy", "v", "txx", "txy", "tyy", "ux", "uy"]


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [1, 100])
def test_gpu_isolin_bitwise(steps):
    w = isolin_box()
    print(f"isolin box, {steps} steps", flush=True)
    cfg = pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"],
                            bc_kind=w["bc_kind"], bc_fp64=w["bc_fp64"], bc_int=w["bc_int"],
                            g_mag=w["g_mag"], ramp_time=w["ramp_time"],
                            material_kind=pympm.MAT_ISOLIN)
    sim = pympm.Sim(cfg, w["particles"])
    sim.step(steps)
    o = pyoracle.from_workload(w, num_threads=3)
    o.run(steps)
    ref = o.particles()
    ref = dict(ref, txx=ref["sxx"], txy=ref["sxy"], tyy=ref["syy"])
    for k in FIELDS:
        assert np.array_equal(sim.get(k), ref[k]), k
    assert np.array_equal(sim.active(), ref["active"])
    assert np.abs(ref["sxx"]).max() > 0


def read_state(path):
    lines = open(path).read().split("\n")
    cols = ["x", "y", "vx", "vy", "v", "sxx", "sxy", "syy", "tzz", "gammap", "gammadotp",
            "ux", "uy"]
    rows = [ln.split() for ln in lines[1:] if ln.strip()]
    d = {k: np.array([float.fromhex(r[i]) for r in rows]) for i, k in enumerate(cols)}
    d["active"] = np.array([int(r[-1]) for r in rows], dtype=np.int32)
    return float.fromhex(lines[0].split()[0]), d


@pytest.mark.gpu
@pytest.mark.parametrize("plugin_dir", ["lib", "lib/host_only"])
def test_dropin_isolin_plugin(tmp_path, plugin_dir):
    w = isolin_box(vx=0.0)
    w["particles"]["sxy"] = np.zeros(len(w["particles"]["x"]))
    steps = 40
    g, p = tmp_path / "grid.txt", tmp_path / "particles.txt"
    workloads.write_reference_inputs(w, g, p)
    cmd = [RUN, "-g", str(g), "-p", str(p), "-u", os.path.join(PKG, plugin_dir, "isolin.so"),
           "-b", os.path.join(PKG, plugin_dir, "box.so"), "-t", "3", "--dt", repr(w["dt"]),
           "--steps", str(steps), "--mprops", f"{w['material']['E']!r},{w['material']['nu']!r}",
           "--bprops", ",".join(repr(float(v)) for v in w["bc_fp64"]),
           "--biprops", ",".join(str(int(v)) for v in w["bc_int"]),
           "--state", str(tmp_path / "state.txt")]
    print("case isolin drop-in", plugin_dir, flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    t, d = read_state(tmp_path / "state.txt")
    o = pyoracle.from_workload(dict(w, g_mag=9.8), num_threads=2)
    o.run(steps)
    r = o.particles()
    assert t == o.time()
    for k in ["x", "y", "vx", "vy", "v", "sxx", "sxy", "syy", "ux", "uy"]:
        rk = {"vx": "vx", "vy": "vy"}.get(k, k)
        assert np.array_equal(d[k], r[rk]), k
    assert np.array_equal(d["active"], r["active"])
