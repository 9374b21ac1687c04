"""ctypes binding of the CPU oracle (oracle/usl_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker or the timed CPU baseline.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
REF_PIECES = os.path.join(HERE, "_ref", "libmpm_ref_pieces.so")

# particle field order of oracle_get_particles (usl_oracle.c: OF_*)
FIELDS = ["x", "y", "vx", "vy", "m", "v", "txx", "txy", "tyy", "tzz",
          "gammap", "gammadotp", "ux", "uy", "exx", "exy", "eyy", "wxy",
          "ax", "ay", "xl", "yl", "sxx", "sxy", "syy", "txz", "tyx", "tyz",
          "tzx", "tzy", "by", "gf", "gf_bulk", "xisq_inv", "state1"]
NODE_FIELDS = ["m", "mx", "my", "fx", "fy", "vx", "vy", "ax", "ay"]


class Params(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("h", ctypes.c_double), ("dt", ctypes.c_double),
                ("num_threads", ctypes.c_int),
                ("E", ctypes.c_double), ("nu", ctypes.c_double),
                ("mu_s", ctypes.c_double), ("mu_2", ctypes.c_double),
                ("I_0", ctypes.c_double), ("grains_d", ctypes.c_double),
                ("grains_rho", ctypes.c_double), ("rho_crit", ctypes.c_double),
                ("g_mag", ctypes.c_double), ("ramp_time", ctypes.c_double),
                ("bc_kind", ctypes.c_int), ("bc_fp64", ctypes.c_double * 4),
                ("bc_int", ctypes.c_int * 4), ("log_crossings", ctypes.c_int),
                ("log_path", ctypes.c_char_p), ("material_kind", ctypes.c_int),
                ("nl_E", ctypes.c_double), ("nl_nu", ctypes.c_double),
                ("nl_mu_s", ctypes.c_double), ("nl_A", ctypes.c_double),
                ("nl_B", ctypes.c_double), ("nl_d", ctypes.c_double),
                ("nl_rho", ctypes.c_double), ("nl_p_cap", ctypes.c_double),
                ("nl_c", ctypes.c_double), ("scheme", ctypes.c_int)]


_libs = {}


def load(variant="parity"):
    """variant: 'parity' (-ffp-contract=off, portable exp) or 'v4'/'v3'
    (reference compiler flags, glibc exp; for timing)."""
    if variant in _libs:
        return _libs[variant]
    name = {"parity": "liboracle.so", "v4": "liboracle_v4.so", "v3": "liboracle_v3.so"}[variant]
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (make -C oracle)")
    lib = ctypes.CDLL(path)
    D = ctypes.POINTER(ctypes.c_double)
    P = ctypes.c_void_p
    lib.oracle_params_size.restype = ctypes.c_size_t
    assert lib.oracle_params_size() == ctypes.sizeof(Params), "oracle_params_t mismatch"
    lib.oracle_default_params.argtypes = [ctypes.POINTER(Params)]
    lib.oracle_create.argtypes = [ctypes.POINTER(Params), ctypes.c_size_t] + [D] * 9
    lib.oracle_create.restype = P
    lib.oracle_run.argtypes = [P, ctypes.c_int]
    lib.oracle_get_particles.argtypes = [P, D, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]
    lib.oracle_get_nodes.argtypes = [P, D]
    lib.oracle_get_bc_mask.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_time.argtypes = [P]
    lib.oracle_time.restype = ctypes.c_double
    lib.oracle_crossings.argtypes = [P]
    lib.oracle_crossings.restype = ctypes.c_longlong
    lib.oracle_destroy.argtypes = [P]
    lib.oracle_exp.argtypes = [ctypes.c_double]
    lib.oracle_exp.restype = ctypes.c_double
    lib.oracle_stress_point.argtypes = [D, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double,
                                        ctypes.POINTER(Params), D]
    lib.oracle_isolin_point.argtypes = [D] + [ctypes.c_double] * 5 + [ctypes.POINTER(Params), D]
    lib.oracle_dp_ri_mt_point.argtypes = [D] + [ctypes.c_double] * 7 + [ctypes.POINTER(Params), D]
    lib.oracle_num_particle_fields.restype = ctypes.c_int
    assert lib.oracle_num_particle_fields() == len(FIELDS)
    for fn in ("oracle_global_to_local_coords",):
        getattr(lib, fn).argtypes = [D, D] + [ctypes.c_double] * 5
    lib.oracle_tent.argtypes = [D] * 4 + [ctypes.c_double] * 2
    lib.oracle_grad_tent.argtypes = [D] * 8 + [ctypes.c_double] * 3
    lib.oracle_nonlocal_system.argtypes = [D, D, ctypes.POINTER(ctypes.c_size_t)]
    lib.oracle_nonlocal_system.restype = ctypes.c_size_t
    lib.oracle_band_cholesky_solve.argtypes = [D, ctypes.c_size_t, ctypes.c_size_t, D]
    lib.oracle_band_cholesky_solve.restype = ctypes.c_int
    lib.oracle_negative_root.argtypes = [ctypes.c_double] * 3
    lib.oracle_nonlocal_bulk_point.argtypes = [ctypes.c_double] * 3 + [ctypes.POINTER(Params), D]
    lib.oracle_nonlocal_stress_point.argtypes = [D, D] + [ctypes.c_double] * 7 + [ctypes.POINTER(Params)]
    lib.oracle_negative_root.restype = ctypes.c_double
    _libs[variant] = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


NONLOCAL_KEYS = ("E", "nu", "mu_s", "A", "B", "d", "rho", "p_cap", "c")


def make_params(lib, N, h, dt, material=None, bc_kind=0, bc_fp64=(), bc_int=(),
                g_mag=9.8, ramp_time=0.1, num_threads=1, log_crossings=False,
                nonlocal_material=None, material_kind=0, scheme=0, log_path=None):
    """material: g_local_mu2 constants (E and nu also serve isolin);
    nonlocal_material: g_nonlocal constants (keys of NONLOCAL_KEYS) -- selects
    the g_nonlocal law when given; material_kind 2 selects isolin."""
    q = Params()
    lib.oracle_default_params(ctypes.byref(q))
    q.material_kind = material_kind
    q.scheme = scheme     # 0 USL, 1 MUSL
    if nonlocal_material is not None:
        q.material_kind = 1
        for k, val in nonlocal_material.items():
            setattr(q, "nl_" + k, val)
    q.N, q.h, q.dt, q.num_threads = N, h, dt, num_threads
    for k, val in (material or {}).items():
        setattr(q, k, val)
    q.g_mag, q.ramp_time = g_mag, ramp_time
    q.bc_kind = bc_kind
    for i, val in enumerate(bc_fp64):
        q.bc_fp64[i] = val
    for i, val in enumerate(bc_int):
        q.bc_int[i] = val
    q.log_crossings = 1 if log_crossings else 0
    if log_path is not None:
        q.log_path = log_path.encode()
    return q


class Oracle:
    """One reference-semantics job on the CPU."""

    def __init__(self, params, particles, variant="parity"):
        self.lib = load(variant)
        self.params = params
        self.np = len(particles["x"])
        arrs = [np.ascontiguousarray(particles[k], dtype=np.float64)
                for k in ("m", "v", "x", "y", "vx", "vy", "sxx", "sxy", "syy")]
        self._keep = arrs
        self.h = self.lib.oracle_create(ctypes.byref(params), self.np, *[_dptr(a) for a in arrs])

    def run(self, n):
        self.lib.oracle_run(self.h, n)

    def particles(self):
        out = np.empty((len(FIELDS), self.np), dtype=np.float64)
        active = np.empty(self.np, dtype=np.int32)
        elem = np.empty(self.np, dtype=np.int32)
        self.lib.oracle_get_particles(self.h, _dptr(out),
                                      active.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      elem.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        d = {k: out[i] for i, k in enumerate(FIELDS)}
        d["active"] = active
        d["in_element"] = elem
        return d

    def nodes(self):
        nn = self.params.N * self.params.N
        out = np.empty((len(NODE_FIELDS), nn), dtype=np.float64)
        self.lib.oracle_get_nodes(self.h, _dptr(out))
        return {k: out[i] for i, k in enumerate(NODE_FIELDS)}

    def nonlocal_system(self):
        """(K lower band [bw+1, n], f [n], bw) of the last nonlocal solve."""
        bw = ctypes.c_size_t()
        n = self.lib.oracle_nonlocal_system(None, None, ctypes.byref(bw))
        band = np.empty((bw.value + 1, n))
        f = np.empty(n)
        self.lib.oracle_nonlocal_system(_dptr(band), _dptr(f), ctypes.byref(bw))
        return band, f, bw.value

    def bc_mask(self):
        nn = self.params.N * self.params.N
        out = np.empty(3 * nn, dtype=np.int32)
        self.lib.oracle_get_bc_mask(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out.reshape(nn, 3)

    def time(self):
        return self.lib.oracle_time(self.h)

    def crossings(self):
        return self.lib.oracle_crossings(self.h)

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def from_workload(w, variant="parity", num_threads=1, log_crossings=False, particles=None,
                  log_path=None):
    lib = load(variant)
    q = make_params(lib, w["N"], w["h"], w["dt"], material=w["material"],
                    bc_kind=w.get("bc_kind", 0), bc_fp64=w.get("bc_fp64", ()),
                    bc_int=w.get("bc_int", ()), g_mag=w.get("g_mag", 9.8),
                    ramp_time=w.get("ramp_time", 0.1), num_threads=num_threads,
                    log_crossings=log_crossings,
                    nonlocal_material=w.get("nonlocal_material"),
                    material_kind=w.get("material_kind", 0), scheme=w.get("scheme", 0),
                    log_path=log_path)
    return Oracle(q, particles if particles is not None else w["particles"], variant=variant)
