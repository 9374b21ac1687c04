"""The C-ABI library builds for gfx950, loads, and exports every entry point
declared in include/mpm_hip.h; its host-side exp is the oracle's bit for bit.
No compute call needs a GPU here."""
import ctypes
import os
import re

import numpy as np
import pytest

import pympm
import pyoracle

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    decl = r"^(?:int|void|double|const char \*)\s*\**\s*(mpm_hip_[a-z_0-9]+)\s*\("
    return sorted(set(re.findall(decl, text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = pympm.load()
    names = _declared("mpm_hip.h")
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name


def test_library_carries_gfx950_code_object():
    data = open(pympm.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_host_exp_is_the_oracle_exp():
    lib = pyoracle.load("parity")
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(-1e-3, 1e-3, 5000), rng.uniform(-50, 50, 5000)])
    for x in xs:
        assert pympm.exp_host(x) == lib.oracle_exp(float(x))


def test_create_without_a_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    cfg = pympm.make_config(N=9, h=0.125, dt=1e-5)
    ctx = ctypes.c_void_p()
    rc = pympm.load().mpm_hip_create(ctypes.byref(cfg), 10, ctypes.byref(ctx))
    assert rc == -3 and not ctx.value


def test_invalid_configs_rejected():
    lib = pympm.load()
    ctx = ctypes.c_void_p()
    for kw in (dict(N=1), dict(h=0.0), dict(dt=-1.0), dict(bin_capacity=48)):
        base = dict(N=9, h=0.125, dt=1e-5)
        base.update(kw)
        cfg = pympm.make_config(**base)
        assert lib.mpm_hip_create(ctypes.byref(cfg), 10, ctypes.byref(ctx)) == -1
