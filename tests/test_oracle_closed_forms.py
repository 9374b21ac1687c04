"""Closed-form pins of the oracle's strain rate, volume update, g_local_mu2
trial stress and the g_nonlocal point laws.

Each expected value is computed here by a different route from the oracle's
(matrix algebra in numpy, textbook forms of the granular-fluidity law, the
exactness of bilinear interpolation on linear fields), so that a dropped
term, a wrong sign or a swapped index in the oracle fails:

* a linear velocity field on an unjittered lattice: interior particles must
  see the exact velocity gradient (strain rate and spin, process_usl.c:764-808),
  the exact advection x + dt v(x) (process_usl.c:914-957) and the volume
  v exp(dt tr D) (process_usl.c:961-970);
* g_local_mu2's trial stress T + dt (2G D + lambda tr(D) I + W T - T W)
  from 3x3 matrices of the velocity gradient (g_local_mu2.c:205-239), in the
  elastic branch exactly and, in the plastic branch, through the pressure,
  the direction of the deviator and the slip of the return map;
* g_nonlocal's bulk fluidity against g_loc = sqrt(p / (rho d^2)) (mu - mu_s) /
  (b mu) and the cooperativity length xi = A d / sqrt|mu - mu_s|, with tau the
  radius of Mohr's circle (g_nonlocal.c:306-352);
* g_nonlocal's stress update in its limits (g_nonlocal.c:148-298).
"""
import ctypes
import math

import numpy as np
import pytest

import pyoracle
import workloads

D = ctypes.c_double


# ---------------------------------------------------------------------------
# Strain rate, advection and volume from a linear velocity field
# ---------------------------------------------------------------------------

def test_linear_velocity_field_gives_exact_gradient_advection_and_volume():
    """v = (a x + b y, c x + d y) on the unjittered 4 x 4 lattice: the
    lattice is symmetric about every node whose four cells are filled, so
    the mapped node velocity is v(node) there, and the bilinear gradient of a
    linear nodal field is exact.  The strain rate definitions are those of
    process_usl.c:764-808: exx = dvx/dx, eyy = dvy/dy, exy = (dvx/dy +
    dvy/dx) / 2, wxy = (dvx/dy - dvy/dx) / 2."""
    lib = pyoracle.load()
    n, c0, nb = 32, 12, 8
    pts = workloads.block(nb, nb, n, jitter=0.0, cell_x0=c0, cell_y0=c0)
    a, b, c, d = 0.31, -0.73, 0.47, -0.19
    pts["vx"] = a * pts["x"] + b * pts["y"]
    pts["vy"] = c * pts["x"] + d * pts["y"]
    dt = 1e-4
    q = pyoracle.make_params(lib, n + 1, 1.0 / n, dt, material=workloads.REFERENCE_MATERIAL,
                             g_mag=0.0)
    o = pyoracle.Oracle(q, pts)
    o.run(1)
    p = o.particles()
    h = 1.0 / n
    cx = np.floor(pts["x"] / h).astype(int)
    cy = np.floor(pts["y"] / h).astype(int)
    inner = (cx >= c0 + 1) & (cx <= c0 + nb - 2) & (cy >= c0 + 1) & (cy <= c0 + nb - 2)
    assert inner.sum() == 16 * (nb - 2) ** 2
    tol = 1e-13
    assert np.abs(p["exx"][inner] - a).max() <= tol
    assert np.abs(p["eyy"][inner] - d).max() <= tol
    assert np.abs(p["exy"][inner] - 0.5 * (b + c)).max() <= tol
    assert np.abs(p["wxy"][inner] - 0.5 * (b - c)).max() <= tol
    # no stress and no body force in the first step: no acceleration
    assert np.array_equal(p["vx"], pts["vx"]) and np.array_equal(p["vy"], pts["vy"])
    # advection with the interpolated (exact) node velocity
    x0, y0 = pts["x"][inner], pts["y"][inner]
    assert np.abs(p["x"][inner] - (x0 + dt * (a * x0 + b * y0))).max() <= 1e-16
    assert np.abs(p["y"][inner] - (y0 + dt * (c * x0 + d * y0))).max() <= 1e-16
    assert np.abs(p["ux"][inner] - dt * (a * x0 + b * y0)).max() <= 1e-18
    # volume: v exp(dt (exx + eyy)), i.e. the exact exponential of dt tr(D)
    want = pts["v"][inner] * math.exp(dt * (a + d))
    assert np.abs(p["v"][inner] / want - 1.0).max() <= 1e-15
    # the tolerance separates it from exp(dt (exx - eyy)) and from the
    # linearisation 1 + dt tr D
    assert abs(math.exp(dt * (a + d)) - (1 + dt * (a + d))) > 1e-11
    assert abs(math.exp(dt * (a + d)) - math.exp(dt * (a - d))) > 1e-6


# ---------------------------------------------------------------------------
# g_local_mu2 trial stress with the Jaumann rate
# ---------------------------------------------------------------------------

def _stress_point(lib, T, m, v, L, dt, q):
    out = (D * 6)()
    lib.oracle_stress_point((D * 4)(*T), m, v, 0.0, L[0], L[1], L[2], L[3], dt,
                            ctypes.byref(q), out)
    return np.array(list(out))


def _mats(T4, grad):
    """3x3 stress from (xx, xy, yy, zz) and the plane velocity gradient
    L_ij = dv_i/dx_j padded to 3x3; returns T, L, D = sym L, W = skw L and
    the oracle's rate arguments (exx, exy, eyy, wxy) as the strain-rate
    kernel defines them from the same gradient."""
    T = np.array([[T4[0], T4[1], 0.0], [T4[1], T4[2], 0.0], [0.0, 0.0, T4[3]]])
    Lm = np.zeros((3, 3))
    Lm[:2, :2] = grad
    Dm = 0.5 * (Lm + Lm.T)
    Wm = 0.5 * (Lm - Lm.T)
    rates = (grad[0, 0], 0.5 * (grad[0, 1] + grad[1, 0]), grad[1, 1],
             0.5 * (grad[0, 1] - grad[1, 0]))
    return T, Lm, Dm, Wm, rates


def _jaumann_trial(T, Dm, Wm, dt, E, nu):
    G = E / (2 * (1 + nu))
    K = E / (3 * (1 - 2 * nu))
    lam = K - 2 * G / 3
    return T + dt * (2 * G * Dm + lam * np.trace(Dm) * np.eye(3) + Wm @ T - T @ Wm)


def test_g_local_mu2_elastic_trial_with_jaumann_spin():
    """Elastic branch (tau_tr <= mu_s p_tr): the result is the Jaumann trial
    stress.  The states are deviatoric (Txy != 0, Txx != Tyy) and the spin is
    large, so W T - T W contributes far above the tolerance."""
    lib = pyoracle.load()
    mat = workloads.REFERENCE_MATERIAL
    q = pyoracle.make_params(lib, 9, 0.125, 1e-5, material=mat)
    rng = np.random.default_rng(31)
    checked = 0
    for _ in range(400):
        p0 = 10 ** rng.uniform(3, 5)
        T4 = [-p0 + 0.08 * p0 * rng.normal(), 0.08 * p0 * rng.normal(),
              -p0 + 0.08 * p0 * rng.normal(), -p0 + 0.05 * p0 * rng.normal()]
        grad = rng.normal(size=(2, 2)) * 0.5
        grad[0, 1] += rng.choice([-1, 1]) * 50.0     # strong spin
        grad[1, 0] -= grad[0, 1] - rng.normal()
        dt = 1e-4
        T, Lm, Dm, Wm, rates = _mats(T4, grad)
        Ttr = _jaumann_trial(T, Dm, Wm, dt, mat["E"], mat["nu"])
        p_tr = -np.trace(Ttr) / 3
        dev = Ttr + p_tr * np.eye(3)
        tau_tr = math.sqrt(0.5 * np.sum(dev * dev))
        if not (p_tr > 0 and tau_tr <= mat["mu_s"] * p_tr * (1 - 1e-9)):
            continue
        spin = dt * (Wm @ T - T @ Wm)
        assert np.abs(spin).max() > 1e-3 * p0 * 1e-2
        out = _stress_point(lib, T4, 1500.0, 1.0, rates, dt, q)
        want = np.array([Ttr[0, 0], Ttr[0, 1], Ttr[1, 1], Ttr[2, 2]])
        assert np.abs(out[:4] - want).max() <= 1e-13 * p0, (out[:4], want)
        assert out[5] == 0.0
        # the same state with the spin's sign flipped differs by twice the spin
        flipped = _stress_point(lib, T4, 1500.0, 1.0, rates[:3] + (-rates[3],), dt, q)
        assert np.abs((out[:4] - flipped[:4])
                      - 2 * np.array([spin[0, 0], spin[0, 1], spin[1, 1], spin[2, 2]])).max() \
            <= 1e-12 * p0
        checked += 1
    assert checked > 200


def test_g_local_mu2_plastic_return_keeps_trial_pressure_and_direction():
    """Plastic branch: p = p_tr, dev T' is dev T_tr scaled by tau'/tau_tr,
    tau' solves tau^2 - (mu_2 p + tau_tr + alpha) tau + mu_2 p tau_tr +
    mu_s p alpha = 0 (smaller root), and gammadot_p = (tau_tr - tau') / (G dt),
    with T_tr the Jaumann trial formed from matrices here."""
    lib = pyoracle.load()
    mat = workloads.REFERENCE_MATERIAL
    q = pyoracle.make_params(lib, 9, 0.125, 1e-5, material=mat)
    E, nu = mat["E"], mat["nu"]
    G = E / (2 * (1 + nu))
    rng = np.random.default_rng(32)
    checked = 0
    for _ in range(600):
        p0 = 10 ** rng.uniform(2, 4)
        T4 = [-p0 + 0.5 * p0 * rng.normal(), 0.5 * p0 * rng.normal(),
              -p0 + 0.5 * p0 * rng.normal(), -p0 + 0.2 * p0 * rng.normal()]
        grad = rng.normal(size=(2, 2)) * 10 ** rng.uniform(-1, 2)
        dt = 10 ** rng.uniform(-6, -4)
        T, Lm, Dm, Wm, rates = _mats(T4, grad)
        Ttr = _jaumann_trial(T, Dm, Wm, dt, E, nu)
        p_tr = -np.trace(Ttr) / 3
        dev_tr = Ttr + p_tr * np.eye(3)
        tau_tr = math.sqrt(0.5 * np.sum(dev_tr * dev_tr))
        if not (p_tr > 0 and tau_tr > mat["mu_s"] * p_tr * (1 + 1e-6)):
            continue
        out = _stress_point(lib, T4, 1500.0, 1.0, rates, dt, q)
        Tn = np.array([[out[0], out[1], 0.0], [out[1], out[2], 0.0], [0.0, 0.0, out[3]]])
        p_new = -np.trace(Tn) / 3
        dev_new = Tn + p_new * np.eye(3)
        tau_new = math.sqrt(0.5 * np.sum(dev_new * dev_new))
        scale = p0 + tau_tr
        assert p_new == pytest.approx(p_tr, rel=1e-12, abs=1e-13 * scale)
        assert np.abs(dev_new - dev_tr * (tau_new / tau_tr)).max() <= 1e-11 * scale
        alpha = G * mat["I_0"] * dt * math.sqrt(p_tr / mat["grains_rho"]) / mat["grains_d"]
        Bq = -(mat["mu_2"] * p_tr + tau_tr + alpha)
        Hq = mat["mu_2"] * p_tr * tau_tr + mat["mu_s"] * p_tr * alpha
        root = (-Bq - math.sqrt(Bq * Bq - 4 * Hq)) / 2
        assert tau_new == pytest.approx(root, rel=1e-9, abs=1e-12 * scale)
        assert out[5] * G * dt == pytest.approx(tau_tr - tau_new, rel=1e-8, abs=1e-11 * scale)
        checked += 1
    assert checked > 200


# ---------------------------------------------------------------------------
# g_nonlocal: bulk fluidity
# ---------------------------------------------------------------------------

def _nl_params(lib, **kw):
    return pyoracle.make_params(lib, 9, 0.125, 1e-5,
                                nonlocal_material=dict(workloads.REFERENCE_NONLOCAL, **kw))


def _bulk(lib, q, sxx, sxy, syy):
    out = (D * 3)()
    lib.oracle_nonlocal_bulk_point(sxx, sxy, syy, ctypes.byref(q), out)
    return list(out)   # eta, gf_bulk, xisq_inv


def _textbook_bulk(sxx, sxy, syy, nl):
    """Kamrin-Koval local fluidity in its textbook form: p the in-plane mean
    pressure (capped below at p_cap), tau the radius of Mohr's circle,
    mu = tau / p, g_loc = sqrt(p / (rho d^2)) (mu - mu_s) / (b mu) above
    yield, else 0; xi^-2 = |mu - mu_s| / (A d)^2; eta = b d sqrt(p rho)."""
    p = max(-(sxx + syy) / 2, nl["p_cap"])
    tau = math.hypot((sxx - syy) / 2, sxy)
    mu = tau / p
    g = math.sqrt(p / (nl["rho"] * nl["d"] ** 2)) * (mu - nl["mu_s"]) / (nl["B"] * mu) \
        if mu > nl["mu_s"] else 0.0
    return nl["B"] * nl["d"] * math.sqrt(p * nl["rho"]), g, abs(mu - nl["mu_s"]) / (nl["A"] * nl["d"]) ** 2


@pytest.mark.parametrize("state", [
    (-3000.0, 800.0, -1000.0),     # above yield: mu = 0.64
    (-1000.0, 10.0, -1100.0),      # below yield
    (-0.004, 0.001, -0.004),       # p under the cap, below yield at p_cap
    (0.05, 0.0, -0.05),            # p = 0: capped, above yield at p_cap
    (2.0, -0.3, 1.0),              # tension
    (-50.0, 0.0, -50.0),           # isotropic: tau = 0
])
def test_nonlocal_bulk_fluidity_hand_picked(state):
    lib = pyoracle.load()
    nl = workloads.REFERENCE_NONLOCAL
    q = _nl_params(lib)
    got = _bulk(lib, q, *state)
    want = _textbook_bulk(*state, nl)
    for g, w in zip(got, want):
        assert g == pytest.approx(w, rel=1e-13, abs=1e-300)


def test_nonlocal_bulk_fluidity_random_states_and_parameters():
    lib = pyoracle.load()
    rng = np.random.default_rng(41)
    above = 0
    for _ in range(3000):
        nl = dict(workloads.REFERENCE_NONLOCAL, mu_s=rng.uniform(0.2, 0.6),
                  A=rng.uniform(0.2, 1.0), B=rng.uniform(1.0, 30.0), d=10 ** rng.uniform(-3, -1),
                  rho=rng.uniform(1000, 3000))
        q = _nl_params(lib, **{k: nl[k] for k in ("mu_s", "A", "B", "d", "rho")})
        p0 = 10 ** rng.uniform(-3, 4)
        s = (-p0 * (1 + 0.6 * rng.normal()), p0 * 0.5 * rng.normal(), -p0 * (1 + 0.6 * rng.normal()))
        got = _bulk(lib, q, *s)
        want = _textbook_bulk(*s, nl)
        above += want[1] > 0
        for g, w in zip(got, want):
            assert g == pytest.approx(w, rel=1e-12, abs=1e-300)
    assert above > 500


# ---------------------------------------------------------------------------
# g_nonlocal: stress update in its limits
# ---------------------------------------------------------------------------

def _nl_stress(lib, q, s, gammap, rates, gf, t, dt):
    sa = (D * 3)(*s)
    gp = D(gammap)
    lib.oracle_nonlocal_stress_point(sa, ctypes.byref(gp), rates[0], rates[1], rates[2],
                                     rates[3], gf, t, dt, ctypes.byref(q))
    return np.array(list(sa)), gp.value


def _c_star(Dm, E, nu):
    """The reference's in-plane elastic rate: the isotropic plane-stress law
    E / (1 - nu^2) ((1 - nu) D + nu tr(D) I), except that its shear term is
    E / (2 (1 + nu)) exy (g_nonlocal.c:155, 266), half the isotropic law's
    E / (1 + nu) exy; the oracle keeps that."""
    S = E / (1 - nu * nu) * ((1 - nu) * Dm + nu * np.trace(Dm) * np.eye(2))
    S[0, 1] *= 0.5
    S[1, 0] *= 0.5
    return S


def _nl_expected(s, rates, dt, nl, Dp=None):
    """sigma' = sigma + dt (C*:(D - D^p) + sigma W - W sigma) in 2x2
    matrices.  The spin enters with the sign g_nonlocal.c:158-160 and
    264-266 give it (sigma W - W sigma, the opposite of g_local_mu2's
    W T - T W)."""
    sig = np.array([[s[0], s[1]], [s[1], s[2]]])
    exx, exy, eyy, wxy = rates
    Dm = np.array([[exx, exy], [exy, eyy]])
    Wm = np.array([[0.0, wxy], [-wxy, 0.0]])
    if Dp is not None:
        Dm = Dm - Dp
    out = sig + dt * (_c_star(Dm, nl["E"], nl["nu"]) + sig @ Wm - Wm @ sig)
    return np.array([out[0, 0], out[0, 1], out[1, 1]])


def _plastic_rate(s, nup):
    """Flow rule D^p = gammadot_p sigma' / (2 tau), sigma' the in-plane
    deviator and tau = sqrt(sigma':sigma' / 2)."""
    p = -(s[0] + s[2]) / 2
    dev = np.array([[s[0] + p, s[1]], [s[1], s[2] + p]])
    tau = math.sqrt(0.5 * np.sum(dev * dev))
    return nup * dev / (2 * tau), tau, p


def _shear_rate(rates):
    """gammadot = sqrt(2 D':D') of the in-plane strain-rate deviator."""
    exx, exy, eyy, _ = rates
    return 2 * math.hypot((exx - eyy) / 2, exy)


def _random_state(rng, p0):
    return (-p0 * (1 + 0.5 * rng.normal()), 0.4 * p0 * rng.normal(), -p0 * (1 + 0.5 * rng.normal()))


def test_nonlocal_stress_without_fluidity_is_elastic_with_spin():
    lib = pyoracle.load()
    nl = workloads.REFERENCE_NONLOCAL
    q = _nl_params(lib)
    rng = np.random.default_rng(51)
    n = 0
    for _ in range(500):
        p0 = 10 ** rng.uniform(1, 4)
        s = _random_state(rng, p0)
        if -(s[0] + s[2]) / 2 < 1.0:
            continue
        rates = tuple(rng.normal(size=3) * 1e-2) + (rng.normal() * 20.0,)
        dt = 1e-6
        for gf, t in ((0.0, 1e-3), (0.3, 0.0)):      # no fluidity / first step
            got, gp = _nl_stress(lib, q, s, 0.25, rates, gf, t, dt)
            want = _nl_expected(s, rates, dt, nl)
            assert np.abs(got - want).max() <= 1e-13 * p0, (gf, t)
            assert gp == 0.25
        n += 1
    assert n > 300


def test_nonlocal_stress_cap_below_cohesion_pressure():
    lib = pyoracle.load()
    nl = workloads.REFERENCE_NONLOCAL
    q = _nl_params(lib)
    rates = (4e-3, 2e-3, 1e-3, 5.0)     # extension lowers the pressure
    for s in ((-0.004, 0.001, -0.006), (3.0, 0.5, 1.0), (0.0, 0.0, 0.0)):
        trial = _nl_expected(s, rates, 1e-6, nl)
        assert -(trial[0] + trial[2]) / 2 < nl["c"]
        got, gp = _nl_stress(lib, q, s, 0.5, rates, 1.0, 1e-3, 1e-6)
        cap = 0.5 * nl["c"] / nl["mu_s"]
        assert got.tolist() == [cap, 0.0, cap]
        assert gp == 0.5


def test_nonlocal_stress_plastic_flow_and_its_clamp():
    """With fluidity g the plastic shear rate is g tau / p (mu times the
    fluidity), clamped to the total shear rate; it is a rate of the
    start-of-step stress direction and advances gammap by rate * dt."""
    lib = pyoracle.load()
    nl = workloads.REFERENCE_NONLOCAL
    q = _nl_params(lib)
    rng = np.random.default_rng(52)
    clamped = free = 0
    for _ in range(2000):
        p0 = 10 ** rng.uniform(0, 4)
        s = _random_state(rng, p0)
        if -(s[0] + s[2]) / 2 <= nl["c"] * 2:
            continue
        rates = tuple(rng.normal(size=3) * 10 ** rng.uniform(-2, 1)) + (rng.normal(),)
        dt = 1e-6
        if -(_nl_expected(s, rates, dt, nl)[[0, 2]].sum()) / 2 < nl["c"]:
            continue
        gf = 10 ** rng.uniform(-3, 3)
        Dp1, tau, p = _plastic_rate(s, 1.0)
        nup = min(gf * tau / p, _shear_rate(rates))
        clamped += nup < gf * tau / p
        free += nup == gf * tau / p
        got, gp = _nl_stress(lib, q, s, 0.0, rates, gf, 1e-3, dt)
        want = _nl_expected(s, rates, dt, nl, Dp=nup * Dp1)
        scale = p0 + tau + np.abs(want).max()
        assert np.abs(got - want).max() <= 1e-13 * scale, (gf, nup)
        assert gp == pytest.approx(nup * dt, rel=1e-13)
        # a negative fluidity gives no flow
        got0, gp0 = _nl_stress(lib, q, s, 0.0, rates, -gf, 1e-3, dt)
        assert np.abs(got0 - _nl_expected(s, rates, dt, nl)).max() <= 1e-13 * scale
        assert gp0 == 0.0
    assert clamped > 200 and free > 200


def test_nonlocal_stress_no_flow_at_or_below_cohesion_start_pressure():
    """p at the start of the step <= c but the trial pressure above c: the
    update stays elastic whatever the fluidity."""
    lib = pyoracle.load()
    nl = workloads.REFERENCE_NONLOCAL
    q = _nl_params(lib)
    s = (-0.012, 0.004, -0.006)       # p = 0.009 < c
    rates = (-2.0, 0.3, -1.0, 0.7)    # compression raises the trial pressure
    dt = 1e-6
    trial = _nl_expected(s, rates, dt, nl)
    assert -(trial[0] + trial[2]) / 2 > nl["c"]
    got, gp = _nl_stress(lib, q, s, 0.0, rates, 5.0, 1e-3, dt)
    assert np.abs(got - trial).max() <= 1e-13 * np.abs(trial).max()
    assert gp == 0.0
    # flow at that fluidity would have moved the stress far beyond that
    Dp1, tau, p = _plastic_rate(s, 1.0)
    flowing = _nl_expected(s, rates, dt, nl, Dp=min(5.0 * tau / p, _shear_rate(rates)) * Dp1)
    assert np.abs(flowing - trial).max() > 1e-3
