"""CPU restatement of the reverse-mode (adjoint) USL step, in numpy.

TEST INFRASTRUCTURE: imported only by tests/, as the checker of the GPU
adjoint.  The forward trajectory comes from the C oracle (usl_oracle.c),
stepped one step at a time; this module runs the reverse sweep over the
recorded steps.

The forward step it differentiates is explicit_mpm_step_usl_threaded
(libmpm/process_usl.c:465-592) with the g_local_mu2 law
(materialsrc/g_local_mu2.c:197-307):

  P2G    m_i = sum S_ip m_p, p_i = sum S_ip m_p v_p,
         f_i = sum S_ip m_p b - v_p sigma_p grad S_ip      (map.c:85-104)
  BC     p_i, f_i zeroed on fixed DOFs                      (box.c:183-224)
  grid   p~_i = p_i + dt f_i, v_i = p~_i / m_i  if m_i > TOL (process_usl.c:596-625)
  G2P    v_p += dt sum S f_i/m_i, x_p += dt sum S p~_i/m_i   (914-957)
  rate   L_p = sum v_i (x) grad S_ip                        (764-808)
  volume v_p *= exp(dt tr L)                                (961-970)
  stress T_p <- g_local_mu2(L_p, T_p)

Discrete choices are held fixed (cell of each particle, active flags, node
mass thresholds, boundary masks, the material branch: density cut-off /
elastic / plastic), so the result is the derivative of the piecewise-smooth
map on the recorded branch.  Mass m_p is treated as a constant (its adjoint
is not formed).

Adjoint state: one array per particle field, in the order of FIELDS (the
library's mpm_hip_field order).  Parameter gradient: d/dE, d/dnu, d/dmu_s,
d/dmu_2, d/dI_0, d/dgrains_d, d/dgrains_rho.
"""
import numpy as np

import pyoracle

FIELDS = ["x", "y", "vx", "vy", "m", "v", "txx", "txy", "tyy", "tzz", "gammap",
          "gammadotp", "ux", "uy"]
PARAMS = ["E", "nu", "mu_s", "mu_2", "I_0", "grains_d", "grains_rho"]
TOL = 1e-10
SGN = np.array([1.0, -1.0, 1.0, -1.0])


def record(w, steps, material=None, particles=None, num_threads=1):
    """Runs the oracle `steps` steps; returns (records, final state).  Record
    n holds the particle state at the start of step n (FIELDS + active, by),
    and the node values and fixed-DOF flags of step n."""
    wm = dict(w)
    if material is not None:
        wm["material"] = material
    o = pyoracle.from_workload(wm, num_threads=num_threads, particles=particles)
    N = w["N"]
    recs = []
    cur = o.particles()
    for _ in range(steps):
        st = {k: cur[k].copy() for k in FIELDS}
        st["active"] = cur["active"].copy()
        st["by"] = float(cur["by"][0]) if len(cur["by"]) else 0.0
        o.run(1)
        nodes = o.nodes()
        mask = o.bc_mask().reshape(N * N, 3)
        st["nodes"] = {k: nodes[k].copy() for k in ("m", "mx", "my", "fx", "fy", "vx", "vy")}
        st["fixed"] = mask[:, :2] != 0
        recs.append(st)
        cur = o.particles()
    final = {k: cur[k].copy() for k in FIELDS}
    final["active"] = cur["active"].copy()
    o.close()
    return recs, final


def constants(mat):
    E, nu = mat["E"], mat["nu"]
    G = E / (2.0 * (1.0 + nu))
    K = E / (3.0 * (1.0 - 2 * nu))
    return dict(G=G, lam=K - 2.0 * G / 3.0, mu_s=mat["mu_s"], mu_2=mat["mu_2"],
                I_0=mat["I_0"], d=mat["grains_d"], rho_s=mat["grains_rho"],
                rho_crit=mat["rho_crit"])


def _snap(l):
    s = np.copysign(1.0, l)
    a = s * l
    return np.where((a > 1.0) & ((a - 1.0) < TOL), s, l)


def shape(x, y, N, h):
    """Cell, nodes (BL, BR, TR, TL), S, b1, b2 of each particle."""
    Ne = N - 1
    inside = (x < 1.0) & (x >= 0.0) & (y < 1.0) & (y >= 0.0)
    e = np.where(inside, np.floor(x / h) + np.floor(y / h) * Ne, -1).astype(np.int64)
    e = np.where(e < Ne * Ne, e, -1)
    ee = np.maximum(e, 0)
    cc, rr = ee % Ne, ee // Ne
    xl = _snap((x - cc * h) / h)
    yl = _snap((y - rr * h) / h)
    ox, oy = 1.0 - xl, 1.0 - yl
    S = np.stack([ox * oy, xl * oy, xl * yl, ox * yl], axis=1)
    r = 1.0 / h
    b1 = np.stack([-oy * r, oy * r, yl * r, -yl * r], axis=1)
    b2 = np.stack([-ox * r, -xl * r, xl * r, ox * r], axis=1)
    n0 = rr * N + cc
    nid = np.stack([n0, n0 + 1, n0 + N + 1, n0 + N], axis=1)
    return e, nid, S, b1, b2


def material_reverse(exx, exy, eyy, wxy, m, v_new, T, dt, c, Tb, gpb, gdpb):
    """Reverse of g_local_mu2 for arrays of particles.  T = (txx, txy, tyy,
    tzz) at the start of the update, Tb = adjoint of the updated T, gpb and
    gdpb the adjoints of gammap and gammadotp after the update.  Returns
    (adjoints of exx, exy, eyy, wxy, adjoint of T (4), per-particle
    parameter adjoints dict over G, lam, mu_s, mu_2, I_0, d, rho_s)."""
    txx, txy, tyy, tzz = T
    G, lam = c["G"], c["lam"]
    two_G = 2.0 * G
    dxx, dxy, dyy, w = exx, exy, eyy, wxy
    trD = dxx + dyy
    ltr = lam * trD
    a = w * txy
    rxx = ((dxx * two_G + ltr) + 2 * a) * dt + txx
    rxy = ((dxy * two_G) + (w * tyy - w * txx)) * dt + txy
    ryy = ((dyy * two_G + ltr) - 2 * a) * dt + tyy
    rzz = ltr * dt + tzz
    cm = (rxx + ryy + rzz) / 3.0
    d_xx, d_yy, d_zz = rxx - cm, ryy - cm, rzz - cm
    p = -cm
    tau2 = d_xx ** 2 + 2 * rxy ** 2 + d_yy ** 2 + d_zz ** 2
    tau = np.sqrt(0.5 * tau2)
    zero = ((m / v_new) < c["rho_crit"]) | (p <= 0.0)
    elastic = ~zero & (tau <= c["mu_s"] * p)
    plastic = ~zero & ~elastic
    n = len(exx)
    z = np.zeros(n)
    pb = {k: np.zeros(n) for k in ("G", "lam", "mu_s", "mu_2", "I_0", "d", "rho_s")}
    rxxb, rxyb, ryyb, rzzb = (np.where(elastic, q, 0.0) for q in Tb)

    if plastic.any():
        sel = plastic
        P = p[sel]
        ta = tau[sel]
        S0 = c["mu_s"] * P
        S2 = c["mu_2"] * P
        alpha = G * c["I_0"] * dt * np.sqrt(P / c["rho_s"]) / c["d"]
        B = -(S2 + ta + alpha)
        H = S2 * ta + S0 * alpha
        disc = np.sqrt(B * B - 4.0 * H)
        taup = np.where(B > 0.0, (-B - disc) / 2.0, (2.0 * H) / (-B + disc))
        scale = taup / ta
        txb, txyb, tyb, tzb = (q[sel] for q in Tb)
        nupb = gdpb[sel] + gpb[sel] * dt
        dxx_, rxy_, dyy_, dzz_ = d_xx[sel], rxy[sel], d_yy[sel], d_zz[sel]
        scaleb = txb * dxx_ + txyb * rxy_ + tyb * dyy_ + tzb * dzz_
        dxxb, rxyb_p, dyyb, dzzb = txb * scale, txyb * scale, tyb * scale, tzb * scale
        Pb = -(txb + tyb + tzb)
        taub = nupb / (G * dt)
        taupb = -nupb / (G * dt)
        Gb = -nupb * (ta - taup) / (G * G * dt)
        taupb = taupb + scaleb / ta
        taub = taub - scaleb * taup / (ta * ta)
        den = 2.0 * taup + B
        Bb = -taupb * taup / den
        Hb = -taupb / den
        S2b = -Bb + Hb * ta
        taub = taub - Bb + Hb * S2
        alphab = -Bb + Hb * S0
        S0b = Hb * alpha
        Gb = Gb + alphab * alpha / G
        pb["I_0"][sel] = alphab * alpha / c["I_0"]
        pb["d"][sel] = -alphab * alpha / c["d"]
        pb["rho_s"][sel] = -alphab * alpha / (2.0 * c["rho_s"])
        Pb = Pb + alphab * alpha / (2.0 * P)
        pb["mu_s"][sel] = S0b * P
        Pb = Pb + S0b * c["mu_s"]
        pb["mu_2"][sel] = S2b * P
        Pb = Pb + S2b * c["mu_2"]
        pb["G"][sel] = Gb
        tau2b = taub * 0.25 / ta
        dxxb = dxxb + 2 * dxx_ * tau2b
        rxyb_p = rxyb_p + 4 * rxy_ * tau2b
        dyyb = dyyb + 2 * dyy_ * tau2b
        dzzb = dzzb + 2 * dzz_ * tau2b
        cmb = -Pb - (dxxb + dyyb + dzzb)
        rxxb[sel] = dxxb + cmb / 3.0
        ryyb[sel] = dyyb + cmb / 3.0
        rzzb[sel] = dzzb + cmb / 3.0
        rxyb[sel] = rxyb_p

    # trial stress (elastic and plastic branches)
    Tb_in = [rxxb.copy(), rxyb.copy(), ryyb.copy(), rzzb.copy()]
    dxxb = rxxb * dt * two_G
    dxyb = rxyb * dt * two_G
    dyyb = ryyb * dt * two_G
    ltrb = (rxxb + ryyb + rzzb) * dt
    jsxxb, jsxyb, jsyyb = rxxb * dt, rxyb * dt, ryyb * dt
    twoGb = (rxxb * dxx + rxyb * dxy + ryyb * dyy) * dt
    pb["G"] = pb["G"] + 2.0 * twoGb
    pb["lam"] = ltrb * trD
    trDb = ltrb * lam
    dxxb = dxxb + trDb
    dyyb = dyyb + trDb
    wb = 2 * txy * (jsxxb - jsyyb) + jsxyb * (tyy - txx)
    Tb_in[1] = Tb_in[1] + 2 * w * (jsxxb - jsyyb)
    Tb_in[2] = Tb_in[2] + jsxyb * w
    Tb_in[0] = Tb_in[0] - jsxyb * w
    return dxxb, dxyb, dyyb, wb, Tb_in, pb, z


def reverse_step(rec, lam_out, mat, N, h, dt):
    """Adjoint of one step: lam_out (dict over FIELDS) at the end of the step
    -> adjoint at its start, and the per-step parameter adjoints."""
    c = constants(mat)
    x, y = rec["x"], rec["y"]
    e, nid, S, b1, b2 = shape(x, y, N, h)
    upd = (rec["active"] != 0) & (e >= 0)
    nd = rec["nodes"]
    nn = N * N
    lam = {k: lam_out[k].copy() for k in FIELDS}
    lam["m"] = np.zeros_like(x)
    grads = {k: 0.0 for k in ("G", "lam", "mu_s", "mu_2", "I_0", "d", "rho_s")}
    if not upd.any():
        return lam, grads
    I = np.nonzero(upd)[0]
    S, b1, b2, nid = S[I], b1[I], b2[I], nid[I]
    m = rec["m"][I]
    v = rec["v"][I]
    T = [rec[k][I] for k in ("txx", "txy", "tyy", "tzz")]
    nm = nd["m"][nid]
    M = nm >= TOL
    safe = np.where(M, nm, 1.0)
    fx, fy, pxu, pyu = nd["fx"][nid], nd["fy"][nid], nd["mx"][nid], nd["my"][nid]
    vxn, vyn = nd["vx"][nid], nd["vy"][nid]

    # forward G2P quantities
    exx = np.sum(b1 * vxn, axis=1)
    eyy = np.sum(b2 * vyn, axis=1)
    dxy = np.sum(b2 * vxn, axis=1)
    dyx = np.sum(b1 * vyn, axis=1)
    exy = 0.5 * (dxy + dyx)
    wxy = 0.5 * (dxy - dyx)
    ex = np.exp(dt * (exx + eyy))
    v_new = v * ex

    lo = {k: lam_out[k][I] for k in FIELDS}
    Tb = [lo[k] for k in ("txx", "txy", "tyy", "tzz")]
    exxb, exyb, eyyb, wxyb, Tb_in, pb, _ = material_reverse(
        exx, exy, eyy, wxy, m, v_new, T, dt, c, Tb, lo["gammap"], lo["gammadotp"])
    for k in grads:
        grads[k] = float(np.sum(pb[k]))
    # volume
    vb = lo["v"] * ex
    exxb = exxb + lo["v"] * v_new * dt
    eyyb = eyyb + lo["v"] * v_new * dt
    dxyb = 0.5 * (exyb + wxyb)
    dyxb = 0.5 * (exyb - wxyb)
    # G2P
    A = lo["vx"] * dt
    B = lo["vy"] * dt
    Cx = (lo["x"] + lo["ux"]) * dt
    Cy = (lo["y"] + lo["uy"]) * dt
    Sb = np.where(M, (A[:, None] * fx + B[:, None] * fy + Cx[:, None] * pxu
                      + Cy[:, None] * pyu) / safe, 0.0)
    b1b = exxb[:, None] * vxn + dyxb[:, None] * vyn
    b2b = eyyb[:, None] * vyn + dxyb[:, None] * vxn

    # node adjoints: gather of the particle coefficients
    def scat(vals):
        out = np.zeros(nn)
        np.add.at(out, nid.ravel(), vals.ravel())
        return out
    Fx = scat(S * A[:, None])
    Fy = scat(S * B[:, None])
    Ux = scat(S * Cx[:, None])
    Uy = scat(S * Cy[:, None])
    Vx = scat(b1 * exxb[:, None] + b2 * dxyb[:, None])
    Vy = scat(b2 * eyyb[:, None] + b1 * dyxb[:, None])
    gm = nd["m"]
    Mn = gm >= TOL
    sm = np.where(Mn, gm, 1.0)
    fxb = np.where(Mn, Fx / sm, 0.0)
    fyb = np.where(Mn, Fy / sm, 0.0)
    pxub = np.where(Mn, Ux / sm, 0.0)
    pyub = np.where(Mn, Uy / sm, 0.0)
    mb = np.where(Mn, -(Fx * nd["fx"] + Fy * nd["fy"] + Ux * nd["mx"] + Uy * nd["my"])
                  / (sm * sm), 0.0)
    live = gm > TOL
    sl = np.where(live, gm, 1.0)
    pxub = pxub + np.where(live, Vx / sl, 0.0)
    pyub = pyub + np.where(live, Vy / sl, 0.0)
    mb = mb - np.where(live, (Vx * nd["mx"] + Vy * nd["my"]) / (sl * sl), 0.0)
    pxb = pxub
    pyb = pyub
    fxb = fxb + np.where(live, dt * pxub, 0.0)
    fyb = fyb + np.where(live, dt * pyub, 0.0)
    fixed = rec["fixed"]
    pxb = np.where(fixed[:, 0], 0.0, pxb)
    fxb = np.where(fixed[:, 0], 0.0, fxb)
    pyb = np.where(fixed[:, 1], 0.0, pyb)
    fyb = np.where(fixed[:, 1], 0.0, fyb)

    # P2G reverse
    vx, vy = rec["vx"][I], rec["vy"][I]
    sxx, sxy, syy = T[0], T[1], T[2]
    by = rec["by"]
    mbj, pxbj, pybj, fxbj, fybj = mb[nid], pxb[nid], pyb[nid], fxb[nid], fyb[nid]
    Sb = Sb + mbj * m[:, None] + pxbj * (vx * m)[:, None] + pybj * (vy * m)[:, None] \
        + fybj * (by * m)[:, None]
    b1b = b1b + fxbj * (-sxx * v)[:, None] + fybj * (-sxy * v)[:, None]
    b2b = b2b + fxbj * (-sxy * v)[:, None] + fybj * (-syy * v)[:, None]
    vxb = lo["vx"] + m * np.sum(S * pxbj, axis=1)
    vyb = lo["vy"] + m * np.sum(S * pybj, axis=1)
    Tb_in[0] = Tb_in[0] - v * np.sum(b1 * fxbj, axis=1)
    Tb_in[1] = Tb_in[1] - v * np.sum(b2 * fxbj + b1 * fybj, axis=1)
    Tb_in[2] = Tb_in[2] - v * np.sum(b2 * fybj, axis=1)
    vb = vb + np.sum(fxbj * (-sxx[:, None] * b1 - sxy[:, None] * b2)
                     + fybj * (-sxy[:, None] * b1 - syy[:, None] * b2), axis=1)
    r2 = 1.0 / h / h
    xb = lo["x"] + np.sum(Sb * b1, axis=1) + np.sum(b2b * SGN, axis=1) * r2
    yb = lo["y"] + np.sum(Sb * b2, axis=1) + np.sum(b1b * SGN, axis=1) * r2

    lam["x"][I] = xb
    lam["y"][I] = yb
    lam["vx"][I] = vxb
    lam["vy"][I] = vyb
    lam["v"][I] = vb
    for k, q in zip(("txx", "txy", "tyy", "tzz"), Tb_in):
        lam[k][I] = q
    lam["gammap"][I] = lo["gammap"]
    lam["gammadotp"][I] = 0.0
    lam["ux"][I] = lo["ux"]
    lam["uy"][I] = lo["uy"]
    return lam, grads


def params_gradient(g, mat):
    """Per-step adjoints of G, lambda, ... -> d/d(E, nu, mu_s, mu_2, I_0,
    grains_d, grains_rho)."""
    E, nu = mat["E"], mat["nu"]
    dG_dE = 1.0 / (2.0 * (1.0 + nu))
    dK_dE = 1.0 / (3.0 * (1.0 - 2 * nu))
    dG_dnu = -E / (2.0 * (1.0 + nu) ** 2)
    dK_dnu = 2.0 * E / (3.0 * (1.0 - 2 * nu) ** 2)
    dlam_dE = dK_dE - 2.0 / 3.0 * dG_dE
    dlam_dnu = dK_dnu - 2.0 / 3.0 * dG_dnu
    return np.array([g["G"] * dG_dE + g["lam"] * dlam_dE,
                     g["G"] * dG_dnu + g["lam"] * dlam_dnu,
                     g["mu_s"], g["mu_2"], g["I_0"], g["d"], g["rho_s"]])


def adjoint(recs, lam_final, mat, N, h, dt):
    """Reverse sweep: (adjoint of the initial state, dL/d params)."""
    lam = {k: np.asarray(lam_final.get(k, np.zeros_like(recs[0]["x"])), dtype=np.float64).copy()
           for k in FIELDS}
    tot = {k: 0.0 for k in ("G", "lam", "mu_s", "mu_2", "I_0", "d", "rho_s")}
    for rec in reversed(recs):
        lam, g = reverse_step(rec, lam, mat, N, h, dt)
        for k in tot:
            tot[k] += g[k]
    return lam, params_gradient(tot, mat)
