"""Other material laws through the device route: isolin
(materialsrc/isolin.c, isotropic linear elasticity with the Jaumann rate) and
dp_ri_mt (materialsrc/dp_ri_mt.c, rate-independent Drucker-Prager).

CPU: the oracle's updates pinned by closed forms written independently of
their component form -- plane-strain Hooke's law in (E, nu) form, the
Jaumann correction W s - s W as a 2x2 matrix product, and the
Drucker-Prager return (pressure kept, tau = mu_s p on the cone, deviator
direction kept, plastic rate (tau_tr - tau) / (G dt), cut-offs).  GPU: the
library's paths bitwise against the oracle, through the C ABI and through
the drop-in libmpm with the repository's plugins (device descriptor and host
fallback).
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


# ---------------------------------------------------------------------------
# dp_ri_mt (materialsrc/dp_ri_mt.c): rate-independent Drucker-Prager

def dp_point(s, m, v, e, dt, E=1e6, nu=0.3, mu_s=0.4):
    lib = pyoracle.load()
    q = pyoracle.make_params(lib, 17, 1 / 16, dt, material=dict(E=E, nu=nu, mu_s=mu_s),
                             material_kind=3)
    sin = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros(6)
    lib.oracle_dp_ri_mt_point(pyoracle._dptr(sin), m, v, *[float(x) for x in e], dt, q,
                              pyoracle._dptr(out))
    return out


def invariants(sxx, sxy, syy, szz):
    p = -(sxx + syy + szz) / 3.0
    d = np.array([sxx + p, syy + p, szz + p])
    tau = np.sqrt(0.5 * (d @ d + 2 * sxy * sxy))
    return p, tau


def test_dp_ri_mt_return_map():
    """Closed forms of the return map: inside the cone the update is the
    elastic (plane-strain Hooke + Jaumann) trial; outside it the stress
    lands on tau = mu_s p with the pressure of the trial state and the
    deviator's direction, and the plastic rate is (tau_tr - tau) / (G dt);
    loose or tensile states become stress-free."""
    rng = np.random.default_rng(3)
    E, nu, mu_s, dt = 1e6, 0.3, 0.4, 1e-5
    G = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    n_el = n_pl = 0
    for _ in range(400):
        p0 = rng.uniform(50, 500)
        s = np.array([-p0 + rng.normal() * 0.3 * p0, rng.normal() * 0.3 * p0,
                      -p0 + rng.normal() * 0.3 * p0, -p0, 0.0])
        e = rng.normal(size=4) * 0.5
        out = dp_point(s, 2000.0, 1.0, e, dt, E, nu, mu_s)
        # trial state, written with the plane-strain stiffness and W s - s W
        tr = e[0] + e[2]
        w = e[3]
        S = np.array([[s[0], s[1]], [s[1], s[2]]])
        W = np.array([[0.0, w], [-w, 0.0]])
        rate = np.array([[lam * tr + 2 * G * e[0], 2 * G * e[1]],
                         [2 * G * e[1], lam * tr + 2 * G * e[2]]]) + W @ S - S @ W
        T = S + dt * rate
        tzz = s[3] + dt * lam * tr
        p_tr, tau_tr = invariants(T[0, 0], T[0, 1], T[1, 1], tzz)
        p_new, tau_new = invariants(out[0], out[1], out[2], out[3])
        assert np.isclose(p_new, p_tr, rtol=1e-12)
        if tau_tr <= mu_s * p_tr:
            n_el += 1
            assert np.allclose(out[:4], [T[0, 0], T[0, 1], T[1, 1], tzz], rtol=1e-12,
                               atol=1e-9 * p0)
            assert out[5] == 0.0
        else:
            n_pl += 1
            assert np.isclose(tau_new, mu_s * p_tr, rtol=1e-12)
            assert np.isclose(out[5], (tau_tr - mu_s * p_tr) / (G * dt), rtol=1e-9)
            # deviator direction kept
            d_tr = np.array([T[0, 0] + p_tr, T[0, 1], T[1, 1] + p_tr])
            d_new = np.array([out[0] + p_tr, out[1], out[2] + p_tr])
            assert np.allclose(d_new, d_tr * (mu_s * p_tr / tau_tr), rtol=1e-10,
                               atol=1e-9 * p0)
        assert np.isclose(out[4], s[4] + out[5] * dt, rtol=1e-14)
    assert n_el > 20 and n_pl > 20
    # density cut-off and tension
    out = dp_point([-100.0, 5.0, -100.0, -100.0, 0.0], 1000.0, 1.0, (0, 0, 0, 0), dt)
    assert np.all(out[:4] == 0.0)
    out = dp_point([100.0, 0.0, 100.0, 100.0, 0.0], 2000.0, 1.0, (0, 0, 0, 0), dt)
    assert np.all(out[:4] == 0.0)


def dp_box(cells=16, seed=11):
    w = workloads.small_box(cells=cells, bx=5, by=9, seed=seed,
                            material=dict(workloads.REFERENCE_MATERIAL, mu_s=0.4))
    w["material_kind"] = 3
    p = w["particles"]
    p = workloads.lithostatic(p, lambda x: 9.0 / cells + 0 * x, rho=1500.0, g=9.8, k0=0.4)
    rng = np.random.default_rng(seed)
    p["vx"] = 0.2 * rng.standard_normal(len(p["x"]))
    w["particles"] = p
    return w


@pytest.mark.gpu
def test_gpu_dp_ri_mt_bitwise():
    w = dp_box()
    steps = 100
    print(f"dp_ri_mt box, {steps} steps", flush=True)
    cfg = pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"],
                            bc_kind=w["bc_kind"], bc_fp64=w["bc_fp64"], bc_int=w["bc_int"],
                            g_mag=w["g_mag"], ramp_time=w["ramp_time"],
                            material_kind=pympm.MAT_DP_RI_MT)
    sim = pympm.Sim(cfg, w["particles"])
    sim.step(steps)
    o = pyoracle.from_workload(w, num_threads=3)
    o.run(steps)
    r = o.particles()
    ref = dict(r, txx=r["sxx"], txy=r["sxy"], tyy=r["syy"], tzz=r["state1"])
    for k in ["x", "y", "vx", "vy", "v", "txx", "txy", "tyy", "tzz", "gammap", "gammadotp",
              "ux", "uy"]:
        assert np.array_equal(sim.get(k), ref[k]), k
    assert r["gammap"].max() > 0, "case must reach the plastic branch"


@pytest.mark.gpu
@pytest.mark.parametrize("plugin_dir", ["lib", "lib/host_only"])
def test_dropin_dp_ri_mt_plugin(tmp_path, plugin_dir):
    w = dp_box()
    w["g_mag"] = 9.8
    steps = 40
    g, p = tmp_path / "grid.txt", tmp_path / "particles.txt"
    workloads.write_reference_inputs(w, g, p)
    m = w["material"]
    cmd = [RUN, "-g", str(g), "-p", str(p), "-u", os.path.join(PKG, plugin_dir, "dp_ri_mt.so"),
           "-b", os.path.join(PKG, plugin_dir, "box.so"), "-t", "3", "--dt", repr(w["dt"]),
           "--steps", str(steps), "--mprops", f"{m['E']!r},{m['nu']!r},{m['mu_s']!r}",
           "--bprops", ",".join(repr(float(v)) for v in w["bc_fp64"]),
           "--biprops", ",".join(str(int(v)) for v in w["bc_int"]),
           "--state", str(tmp_path / "state.txt")]
    print("case dp_ri_mt drop-in", plugin_dir, flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    t, d = read_state(tmp_path / "state.txt")
    o = pyoracle.from_workload(w, num_threads=2)
    o.run(steps)
    r = o.particles()
    assert t == o.time()
    for k in ["x", "y", "vx", "vy", "v", "sxx", "sxy", "syy", "gammap", "gammadotp", "ux",
              "uy"]:
        assert np.array_equal(d[k], r[k]), k
