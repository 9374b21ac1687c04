"""mpm_hip_step_noalloc against mpm_hip_step: the same USL steps on
caller-owned device arrays with the library's static scratch array instead
of a created context.

Two calls in a row on one device -- the first on a 96 x 96-cell column
collapse, the second on a smaller job (64 x 64 cells, fewer particles) that
reuses the scratch the first left behind -- each bitwise against a created
context (mpm_hip_create + set_particles + step) on the same inputs, with a
continuation call on the first job in between (the clock carried in the
caller's array).  Also: sizes beyond the scratch maximum are rejected.
"""
import ctypes
import sys

import numpy as np
import pytest

import pympm
import workloads

pytestmark = pytest.mark.gpu

STATE = pympm.FIELDS[:14]      # MPM_F_X .. MPM_F_UY


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _lib():
    lib = pympm.load()
    lib.mpm_hip_step_noalloc.argtypes = [ctypes.POINTER(pympm.Config), ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    # the caller's allocations, through the HIP runtime the library itself
    # is bound to (looked up through its handle)
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


H2D, D2H = 1, 2


def _config(w):
    return pympm.make_config(w["N"], w["h"], w["dt"], material=w["material"],
                             bc_kind=w["bc_kind"], bc_fp64=w["bc_fp64"], bc_int=w["bc_int"],
                             g_mag=w["g_mag"], ramp_time=w["ramp_time"])


class DeviceJob:
    """The caller's side: particle fields, active flags and clock in device
    memory it allocated itself, filled as mpm_hip_set_particles would."""

    def __init__(self, lib, w, active):
        p = w["particles"]
        n = len(p["x"])
        self.lib, self.n = lib, n
        init = {"x": p["x"], "y": p["y"], "vx": p["vx"], "vy": p["vy"], "m": p["m"],
                "v": p["v"], "txx": p["sxx"], "txy": p["sxy"], "tyy": p["syy"]}
        self.f = {k: self._upload(np.asarray(init[k], dtype=np.float64) if k in init
                                  else np.zeros(n)) for k in STATE}
        self.active = self._upload(np.asarray(active, dtype=np.int32))
        self.clock = self._upload(np.zeros(4))
        self.ptrs = (ctypes.c_void_p * len(STATE))(*[self.f[k] for k in STATE])

    def _upload(self, a):
        a = np.ascontiguousarray(a)
        ptr = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(ptr), max(a.nbytes, 8)) == 0
        assert self.lib.hipMemcpy(ptr, a.ctypes.data, a.nbytes, H2D) == 0
        return ptr.value

    def download(self, ptr, like):
        out = np.empty_like(like)
        assert self.lib.hipMemcpy(out.ctypes.data, ptr, out.nbytes, D2H) == 0
        return out

    def step(self, cfg, nsteps):
        rc = self.lib.mpm_hip_step_noalloc(ctypes.byref(cfg), self.n, self.ptrs, self.active,
                                           self.clock, nsteps, None)
        assert rc == 0, pympm.STATUS.get(rc, rc)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.download(self.clock, np.zeros(4))[3] == 0.0, "bin overflow"

    def free(self):
        for ptr in list(self.f.values()) + [self.active, self.clock]:
            self.lib.hipFree(ptr)


def _check(job, sim, what):
    bad = [k for k in STATE
           if not np.array_equal(job.download(job.f[k], np.zeros(job.n)), sim.get(k))]
    assert not bad, f"{what}: fields differ from the created context: {bad}"
    act = job.download(job.active, np.zeros(job.n, dtype=np.int32))
    assert np.array_equal(act, sim.active()), f"{what}: active"
    clock = job.download(job.clock, np.zeros(4))
    t, by = sim.time()
    assert clock[0] == t and clock[2] == by, f"{what}: clock"


def test_noalloc_step_matches_created_context_twice():
    lib = _lib()
    jobs = [workloads.small_box(cells=96, bx=24, by=48, seed=31, steps=0),
            workloads.small_box(cells=64, bx=10, by=20, seed=32, steps=0)]
    for i, w in enumerate(jobs):
        cfg = _config(w)
        sim = pympm.Sim(cfg, w["particles"])
        job = DeviceJob(lib, w, sim.active())
        say(f"case noalloc call {i}: N={w['N']} particles={job.n}")
        job.step(cfg, 25)
        sim.step(25)
        sim.sync()
        _check(job, sim, f"call {i}")
        if i == 0:
            # a second call continues the run from the caller's clock
            say("case noalloc call 0, continued for 15 steps")
            job.step(cfg, 15)
            sim.step(15)
            sim.sync()
            _check(job, sim, "call 0 continued")
        job.free()
        sim.close()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    lib = _lib()
    w = workloads.small_box(cells=256, bx=4, by=4, steps=0)
    cfg = _config(w)
    job = DeviceJob(lib, w, np.ones(len(w["particles"]["x"]), dtype=np.int32))
    rc = lib.mpm_hip_step_noalloc(ctypes.byref(cfg), job.n, job.ptrs, job.active, job.clock,
                                  1, None)
    job.free()
    assert rc == -1
