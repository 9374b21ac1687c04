"""ctypes binding of libmpm_hip.so (include/mpm_hip.h).

Host-side harness for tests and benchmarks.  It calls the C ABI only; there
is no CPU fallback: if the HIP library is missing or no gfx950 device is
present, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmpm_hip.so")

# mpm_hip_bc_kind
BC_NONE, BC_BOX, BC_SQUARE_SILO, BC_MASK = 0, 1, 2, 3
STORE_AUX = 0x1
HOST_MATERIAL = 0x2
# mpm_hip_material_kind
MAT_G_LOCAL_MU2, MAT_G_NONLOCAL, MAT_ISOLIN, MAT_DP_RI_MT = 0, 1, 2, 3
# mpm_hip_scheme
SCHEME_USL, SCHEME_MUSL = 0, 1

# mpm_hip_field
FIELDS = ["x", "y", "vx", "vy", "m", "v", "txx", "txy", "tyy", "tzz",
          "gammap", "gammadotp", "ux", "uy",
          "exx", "exy", "eyy", "wxy", "ax", "ay", "xl", "yl",
          "gf", "gf_bulk", "xisq_inv"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
AUX_FIELDS = FIELDS[14:]
NODE_FIELDS = ["m", "mx", "my", "fx", "fy", "vx", "vy"]

STATUS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error",
          -3: "no gfx950 device", -4: "cell bin capacity exceeded",
          -5: "device allocation failed"}


class MpmHipError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {STATUS.get(code, code)} ({code})")
        self.code = code


class Material(ctypes.Structure):
    _fields_ = [("E", ctypes.c_double), ("nu", ctypes.c_double),
                ("mu_s", ctypes.c_double), ("mu_2", ctypes.c_double),
                ("I_0", ctypes.c_double), ("grains_d", ctypes.c_double),
                ("grains_rho", ctypes.c_double), ("rho_crit", ctypes.c_double)]


class Nonlocal(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("E", "nu", "mu_s", "A", "B", "d", "rho",
                                                 "p_cap", "c")]


class Config(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("h", ctypes.c_double), ("dt", ctypes.c_double),
                ("t0", ctypes.c_double), ("loads_t0", ctypes.c_double),
                ("g_mag", ctypes.c_double),
                ("ramp_time", ctypes.c_double), ("material", Material),
                ("bc_kind", ctypes.c_int), ("bc_fp64", ctypes.c_double * 4),
                ("bc_int", ctypes.c_int * 4), ("bin_capacity", ctypes.c_int),
                ("flags", ctypes.c_int), ("device", ctypes.c_int),
                ("material_kind", ctypes.c_int), ("nl", Nonlocal),
                ("cg_rel_tol", ctypes.c_double), ("cg_max_iter", ctypes.c_int),
                ("scheme", ctypes.c_int), ("cg_max_blocks", ctypes.c_int),
                ("p2g_tile", ctypes.c_int)]


_lib = None


def load():
    """Loads libmpm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P, D, I, Z = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_size_t
    lib.mpm_hip_default_config.argtypes = [ctypes.POINTER(Config)]
    lib.mpm_hip_default_config.restype = None
    lib.mpm_hip_create.argtypes = [ctypes.POINTER(Config), Z, ctypes.POINTER(P)]
    lib.mpm_hip_set_particles.argtypes = [P] + [D] * 9
    lib.mpm_hip_set_field.argtypes = [P, I, D]
    lib.mpm_hip_get_field.argtypes = [P, I, D]
    lib.mpm_hip_get_active.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    lib.mpm_hip_get_node_field.argtypes = [P, I, D]
    lib.mpm_hip_set_bc_mask.argtypes = [P, ctypes.POINTER(ctypes.c_uint8)]
    lib.mpm_hip_step.argtypes = [P, I]
    lib.mpm_hip_set_body_force.argtypes = [P, ctypes.c_double, ctypes.c_double]
    lib.mpm_hip_set_store_aux.argtypes = [P, I]
    lib.mpm_hip_sync.argtypes = [P]
    lib.mpm_hip_get_time.argtypes = [P, D, D]
    lib.mpm_hip_get_crossings.argtypes = [P, ctypes.POINTER(ctypes.c_longlong)]
    lib.mpm_hip_time_steps.argtypes = [P, I, D]
    lib.mpm_hip_time_span.argtypes = [P, I, D]
    lib.mpm_hip_get_solver_stats.argtypes = [P, ctypes.POINTER(ctypes.c_int), D]
    lib.mpm_hip_destroy.argtypes = [P]
    lib.mpm_hip_status_string.argtypes = [I]
    lib.mpm_hip_status_string.restype = ctypes.c_char_p
    lib.mpm_hip_exp_host.argtypes = [ctypes.c_double]
    lib.mpm_hip_exp_host.restype = ctypes.c_double
    lib.mpm_hip_eval_ops.argtypes = [I, D, D, D, Z, I]
    lib.mpm_hip_batch_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(Z), I,
                                         ctypes.POINTER(P)]
    lib.mpm_hip_batch_set_particles.argtypes = [P, I] + [D] * 9
    lib.mpm_hip_batch_get_field.argtypes = [P, I, I, D]
    lib.mpm_hip_batch_get_active.argtypes = [P, I, ctypes.POINTER(ctypes.c_int)]
    lib.mpm_hip_batch_get_time.argtypes = [P, I, D, D]
    lib.mpm_hip_batch_step.argtypes = [P, I]
    lib.mpm_hip_batch_sync.argtypes = [P]
    lib.mpm_hip_batch_time_steps.argtypes = [P, I, D]
    lib.mpm_hip_batch_autotune.argtypes = [P, I, ctypes.POINTER(ctypes.c_int), D]
    lib.mpm_hip_batch_autotune_launch.argtypes = [P, I, ctypes.POINTER(ctypes.c_int),
                                                  ctypes.POINTER(ctypes.c_int), D]
    lib.mpm_hip_batch_autotune_tiles.argtypes = [P, I, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int), D]
    lib.mpm_hip_batch_destroy.argtypes = [P]
    DP = ctypes.POINTER(D)
    lib.mpm_hip_adjoint_create.argtypes = [ctypes.POINTER(Config), Z, I, ctypes.POINTER(P)]
    lib.mpm_hip_adjoint_set_particles.argtypes = [P] + [D] * 9
    lib.mpm_hip_adjoint_set_material.argtypes = [P, ctypes.POINTER(Material)]
    lib.mpm_hip_adjoint_forward.argtypes = [P, I]
    lib.mpm_hip_adjoint_get_field.argtypes = [P, I, D]
    lib.mpm_hip_adjoint_get_active.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    lib.mpm_hip_adjoint_backward.argtypes = [P, DP, D, DP]
    lib.mpm_hip_adjoint_destroy.argtypes = [P]
    lib.mpm_hip_adjoint_set_particles_device.argtypes = [P] + [P] * 9
    lib.mpm_hip_adjoint_get_field_device.argtypes = [P, I, P]
    lib.mpm_hip_adjoint_backward_device.argtypes = [P, ctypes.POINTER(P), D, ctypes.POINTER(P)]
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(code, what):
    if code != 0:
        raise MpmHipError(code, what)


def make_config(N, h, dt, material=None, bc_kind=BC_NONE, bc_fp64=(), bc_int=(),
                g_mag=9.8, ramp_time=0.1, t0=0.0, bin_capacity=0, flags=0, device=0,
                nonlocal_material=None, cg_rel_tol=1e-10, cg_max_iter=20000, material_kind=0,
                scheme=0, cg_max_blocks=0, p2g_tile=0):
    """nonlocal_material: g_nonlocal constants (E, nu, mu_s, A, B, d, rho,
    p_cap, c); selects the g_nonlocal law when given.  material_kind:
    MAT_G_LOCAL_MU2, MAT_G_NONLOCAL, MAT_ISOLIN or MAT_DP_RI_MT.  cg_max_blocks
    caps the solve's persistent workgroups (0 = as many as fit).  p2g_tile:
    nodes per P2G tile (0 chosen from the grid, 1 16 x 8, 2 8 x 4)."""
    lib = load()
    cfg = Config()
    lib.mpm_hip_default_config(ctypes.byref(cfg))
    cfg.material_kind = material_kind
    cfg.scheme = scheme
    if nonlocal_material is not None:
        cfg.material_kind = 1
        for k, val in nonlocal_material.items():
            setattr(cfg.nl, k, val)
    cfg.cg_rel_tol, cfg.cg_max_iter = cg_rel_tol, cg_max_iter
    cfg.cg_max_blocks = cg_max_blocks
    cfg.p2g_tile = p2g_tile
    cfg.N, cfg.h, cfg.dt, cfg.t0, cfg.loads_t0 = N, h, dt, t0, t0
    cfg.g_mag, cfg.ramp_time = g_mag, ramp_time
    for k, val in (material or {}).items():
        setattr(cfg.material, k, val)
    cfg.bc_kind = bc_kind
    for i, val in enumerate(bc_fp64):
        cfg.bc_fp64[i] = val
    for i, val in enumerate(bc_int):
        cfg.bc_int[i] = val
    cfg.bin_capacity, cfg.flags, cfg.device = bin_capacity, flags, device
    return cfg


class Sim:
    """One job on one GPU: the device-resident USL state."""

    def __init__(self, cfg, particles):
        self.lib = load()
        self.cfg = cfg
        self.np = len(particles["x"])
        self.ctx = ctypes.c_void_p()
        _check(self.lib.mpm_hip_create(ctypes.byref(cfg), self.np, ctypes.byref(self.ctx)),
               "mpm_hip_create")
        arrs = [np.ascontiguousarray(particles[k], dtype=np.float64)
                for k in ("m", "v", "x", "y", "vx", "vy")]
        stress = [particles.get(k) for k in ("sxx", "sxy", "syy")]
        stress = [None if s is None else np.ascontiguousarray(s, dtype=np.float64) for s in stress]
        ptrs = [_dptr(a) for a in arrs] + [None if s is None else _dptr(s) for s in stress]
        _check(self.lib.mpm_hip_set_particles(self.ctx, *ptrs), "mpm_hip_set_particles")

    def set_body_force(self, bx, by):
        _check(self.lib.mpm_hip_set_body_force(self.ctx, bx, by), "mpm_hip_set_body_force")

    def set_store_aux(self, on):
        _check(self.lib.mpm_hip_set_store_aux(self.ctx, 1 if on else 0), "mpm_hip_set_store_aux")

    def step(self, n=1):
        _check(self.lib.mpm_hip_step(self.ctx, n), "mpm_hip_step")

    def sync(self):
        _check(self.lib.mpm_hip_sync(self.ctx), "mpm_hip_sync")

    def get(self, name):
        out = np.empty(self.np, dtype=np.float64)
        _check(self.lib.mpm_hip_get_field(self.ctx, FIELD_ID[name], _dptr(out)),
               f"mpm_hip_get_field({name})")
        return out

    def set(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        _check(self.lib.mpm_hip_set_field(self.ctx, FIELD_ID[name], _dptr(a)),
               f"mpm_hip_set_field({name})")

    def active(self):
        out = np.empty(self.np, dtype=np.int32)
        _check(self.lib.mpm_hip_get_active(
            self.ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "mpm_hip_get_active")
        return out

    def node(self, name):
        nn = self.cfg.N * self.cfg.N
        out = np.empty(nn, dtype=np.float64)
        _check(self.lib.mpm_hip_get_node_field(self.ctx, NODE_FIELDS.index(name), _dptr(out)),
               f"mpm_hip_get_node_field({name})")
        return out

    def set_bc_mask(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        _check(self.lib.mpm_hip_set_bc_mask(
            self.ctx, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), "mpm_hip_set_bc_mask")

    def time(self):
        t, by = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.mpm_hip_get_time(self.ctx, ctypes.byref(t), ctypes.byref(by)),
               "mpm_hip_get_time")
        return t.value, by.value

    def crossings(self):
        c = ctypes.c_longlong()
        _check(self.lib.mpm_hip_get_crossings(self.ctx, ctypes.byref(c)), "mpm_hip_get_crossings")
        return c.value

    def solver_stats(self):
        it, res = ctypes.c_int(), ctypes.c_double()
        _check(self.lib.mpm_hip_get_solver_stats(self.ctx, ctypes.byref(it), ctypes.byref(res)),
               "mpm_hip_get_solver_stats")
        return it.value, res.value

    def time_steps(self, n):
        ms = np.zeros(3, dtype=np.float64)
        _check(self.lib.mpm_hip_time_steps(self.ctx, n, _dptr(ms)), "mpm_hip_time_steps")
        return ms

    def time_span(self, n):
        """Device time (ms) of n steps between two events, no others."""
        ms = np.zeros(1, dtype=np.float64)
        _check(self.lib.mpm_hip_time_span(self.ctx, n, _dptr(ms)), "mpm_hip_time_span")
        return float(ms[0])

    def state(self, names=None):
        names = names or FIELDS[:14]
        return {k: self.get(k) for k in names}

    def close(self):
        if self.ctx:
            self.lib.mpm_hip_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """An ensemble of independent jobs stepped together on one GPU
    (mpm_hip_batch_*): `cfgs[k]` and `particles[k]` describe member k."""

    def __init__(self, cfgs, particles):
        self.lib = load()
        self.n = len(cfgs)
        self.cfgs = (Config * self.n)(*cfgs)
        self.np = [len(p["x"]) for p in particles]
        counts = (ctypes.c_size_t * self.n)(*self.np)
        self.b = ctypes.c_void_p()
        _check(self.lib.mpm_hip_batch_create(self.cfgs, counts, self.n, ctypes.byref(self.b)),
               "mpm_hip_batch_create")
        for k, p in enumerate(particles):
            arrs = [np.ascontiguousarray(p[f], dtype=np.float64)
                    for f in ("m", "v", "x", "y", "vx", "vy")]
            stress = [p.get(f) for f in ("sxx", "sxy", "syy")]
            stress = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                      for a in stress]
            ptrs = [_dptr(a) for a in arrs] + [None if a is None else _dptr(a) for a in stress]
            _check(self.lib.mpm_hip_batch_set_particles(self.b, k, *ptrs),
                   f"mpm_hip_batch_set_particles({k})")

    def step(self, n=1):
        _check(self.lib.mpm_hip_batch_step(self.b, n), "mpm_hip_batch_step")

    def sync(self):
        _check(self.lib.mpm_hip_batch_sync(self.b), "mpm_hip_batch_sync")

    def get(self, member, name):
        out = np.empty(self.np[member], dtype=np.float64)
        _check(self.lib.mpm_hip_batch_get_field(self.b, member, FIELD_ID[name], _dptr(out)),
               f"mpm_hip_batch_get_field({member}, {name})")
        return out

    def active(self, member):
        out = np.empty(self.np[member], dtype=np.int32)
        _check(self.lib.mpm_hip_batch_get_active(
            self.b, member, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
            "mpm_hip_batch_get_active")
        return out

    def time(self, member):
        t, by = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.mpm_hip_batch_get_time(self.b, member, ctypes.byref(t),
                                               ctypes.byref(by)), "mpm_hip_batch_get_time")
        return t.value, by.value

    def state(self, member, names=None):
        names = names or FIELDS[:14]
        return {k: self.get(member, k) for k in names}

    def time_steps(self, n):
        ms = np.zeros(3, dtype=np.float64)
        _check(self.lib.mpm_hip_batch_time_steps(self.b, n, _dptr(ms)),
               "mpm_hip_batch_time_steps")
        return ms

    def autotune(self, trial_steps=4):
        """(chosen bin capacity, per-step trial ms for capacities 16, 32, 64)."""
        cap = ctypes.c_int()
        ms = np.zeros(3, dtype=np.float64)
        _check(self.lib.mpm_hip_batch_autotune(self.b, trial_steps, ctypes.byref(cap),
                                               _dptr(ms)), "mpm_hip_batch_autotune")
        return cap.value, ms

    def autotune_launch(self, trial_steps=4):
        """(capacity, threads per node, per-step trial ms [6] over capacity
        16/32/64 x 2/3 threads per node)."""
        cap, roles = ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(6, dtype=np.float64)
        _check(self.lib.mpm_hip_batch_autotune_launch(self.b, trial_steps, ctypes.byref(cap),
                                                      ctypes.byref(roles), _dptr(ms)),
               "mpm_hip_batch_autotune_launch")
        return cap.value, roles.value, ms

    def autotune_tiles(self, trial_steps=4):
        """(tile shape as p2g_tile codes it: 1 16 x 8, 2 8 x 4, 3 16 x 8 in
        512-record batches; capacity; threads per node; per-step trial ms
        [18]: the three shapes in that order, each over capacity 16/32/64 x
        2/3 threads per node)."""
        tile, cap, roles = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = np.zeros(18, dtype=np.float64)
        _check(self.lib.mpm_hip_batch_autotune_tiles(self.b, trial_steps, ctypes.byref(tile),
                                                     ctypes.byref(cap), ctypes.byref(roles),
                                                     _dptr(ms)),
               "mpm_hip_batch_autotune_tiles")
        return tile.value, cap.value, roles.value, ms

    def close(self):
        if self.b:
            self.lib.mpm_hip_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


STATE_FIELDS = FIELDS[:14]
PARAMS = ["E", "nu", "mu_s", "mu_2", "I_0", "grains_d", "grains_rho"]


class Adjoint:
    """Reverse-mode derivative of an n-step run (mpm_hip_adjoint_*).

    `generation` counts forward runs; a reverse sweep belongs to the forward
    run it follows (the library keeps one run's checkpoints)."""

    def __init__(self, cfg, particles, checkpoint_every=20):
        self.lib = load()
        self.cfg = cfg
        self.np = len(particles["x"])
        self.generation = 0
        self.a = ctypes.c_void_p()
        _check(self.lib.mpm_hip_adjoint_create(ctypes.byref(cfg), self.np, checkpoint_every,
                                               ctypes.byref(self.a)), "mpm_hip_adjoint_create")
        self.set_particles(particles)

    def set_particles(self, particles):
        arrs = [np.ascontiguousarray(particles[k], dtype=np.float64)
                for k in ("m", "v", "x", "y", "vx", "vy")]
        stress = [particles.get(k) for k in ("sxx", "sxy", "syy")]
        stress = [None if s is None else np.ascontiguousarray(s, dtype=np.float64) for s in stress]
        ptrs = [_dptr(a) for a in arrs] + [None if s is None else _dptr(s) for s in stress]
        _check(self.lib.mpm_hip_adjoint_set_particles(self.a, *ptrs),
               "mpm_hip_adjoint_set_particles")

    def set_material(self, **kw):
        mat = Material()
        for name, _ in Material._fields_:
            setattr(mat, name, getattr(self.cfg.material, name))
        for k, v in kw.items():
            setattr(mat, k, float(v))
        _check(self.lib.mpm_hip_adjoint_set_material(self.a, ctypes.byref(mat)),
               "mpm_hip_adjoint_set_material")
        self.cfg.material = mat

    def forward(self, nsteps):
        self.generation += 1
        _check(self.lib.mpm_hip_adjoint_forward(self.a, nsteps), "mpm_hip_adjoint_forward")

    def set_particles_device(self, ptrs):
        """ptrs: device addresses of m, v, x, y, vx, vy, sxx, sxy, syy."""
        _check(self.lib.mpm_hip_adjoint_set_particles_device(self.a, *ptrs),
               "mpm_hip_adjoint_set_particles_device")

    def get_device(self, name, ptr):
        """Copies a field of the final state to device address `ptr`."""
        _check(self.lib.mpm_hip_adjoint_get_field_device(self.a, FIELD_ID[name], ptr),
               f"mpm_hip_adjoint_get_field_device({name})")

    def backward_device(self, seed_ptrs, grad_ptrs):
        """seed_ptrs / grad_ptrs: dict field -> device address (missing =
        zero seed / not wanted).  Returns dL/d params as a dict."""
        n = len(STATE_FIELDS)
        seeds = (ctypes.c_void_p * n)(*[seed_ptrs.get(k) for k in STATE_FIELDS])
        outs = (ctypes.c_void_p * n)(*[grad_ptrs.get(k) for k in STATE_FIELDS])
        g = np.zeros(len(PARAMS))
        _check(self.lib.mpm_hip_adjoint_backward_device(self.a, seeds, _dptr(g), outs),
               "mpm_hip_adjoint_backward_device")
        return dict(zip(PARAMS, g))

    def get(self, name):
        out = np.empty(self.np, dtype=np.float64)
        _check(self.lib.mpm_hip_adjoint_get_field(self.a, FIELD_ID[name], _dptr(out)),
               f"mpm_hip_adjoint_get_field({name})")
        return out

    def active(self):
        out = np.empty(self.np, dtype=np.int32)
        _check(self.lib.mpm_hip_adjoint_get_active(
            self.a, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "mpm_hip_adjoint_get_active")
        return out

    def backward(self, seed):
        """seed: dict field -> dL/d(final field).  Returns (dL/d params as a
        dict over PARAMS, dict of dL/d(initial field) over STATE_FIELDS)."""
        n = len(STATE_FIELDS)
        keep = []
        seeds = (ctypes.POINTER(ctypes.c_double) * n)()
        for i, k in enumerate(STATE_FIELDS):
            if k in seed and seed[k] is not None:
                a = np.ascontiguousarray(seed[k], dtype=np.float64)
                keep.append(a)
                seeds[i] = _dptr(a)
        outs = {k: np.zeros(self.np) for k in STATE_FIELDS}
        optr = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(outs[k]) for k in STATE_FIELDS])
        g = np.zeros(len(PARAMS))
        _check(self.lib.mpm_hip_adjoint_backward(self.a, seeds, _dptr(g), optr),
               "mpm_hip_adjoint_backward")
        return dict(zip(PARAMS, g)), outs

    def close(self):
        if self.a:
            self.lib.mpm_hip_adjoint_destroy(self.a)
            self.a = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def eval_ops(op, a, b=None, device=0):
    """Device evaluation of mpm_exp (op 0), a / b (op 1), sqrt(a) (op 2)."""
    lib = load()
    a = np.ascontiguousarray(a, dtype=np.float64)
    bb = np.ascontiguousarray(b if b is not None else a, dtype=np.float64)
    out = np.empty_like(a)
    _check(lib.mpm_hip_eval_ops(op, _dptr(a), _dptr(bb), _dptr(out), a.size, device),
           "mpm_hip_eval_ops")
    return out


def exp_host(x):
    return load().mpm_hip_exp_host(float(x))
