"""Field offsets of the plugin-ABI structs job_t and threadtask_t (and the
control structs inside job_t) in include/mpm_abi.h against the reference's
libmpm/process.h:60-249.

process.h cannot be compiled here (it includes <suitesparse/cs.h>, absent),
so its layout is computed from its text by oracle/layout_calc.py and kept
in tests/golden/ref_job_layout.txt.  The calculator is pinned against the
compiler: on the reference headers that do compile (particle.h, node.h,
element.h) it must reproduce every offset the compiled probe recorded in
tests/golden/ref_layout.txt.
"""
import os
import subprocess

import pytest

import layout_calc

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
REF = "/root/reference/libmpm"


def _golden(name):
    return [ln for ln in open(os.path.join(GOLDEN, name)).read().split("\n") if ln.strip()]


def _need_reference():
    if not os.path.exists(os.path.join(REF, "process.h")):
        pytest.skip("the reference tree is not mounted here")


def test_calculator_reproduces_the_compiler_on_compilable_headers():
    _need_reference()
    lay = layout_calc.layout_of([os.path.join(REF, h) for h in ("particle.h", "node.h",
                                                                 "element.h")])
    probe = _golden("ref_layout.txt")
    got = {}
    for t in ("particle_t", "node_t", "element_t"):
        for ln in lay.lines(t):
            got[tuple(ln.split()[:2])] = ln
    assert len(probe) == 99
    assert [got.get(tuple(ln.split()[:2])) for ln in probe] == probe


def test_golden_is_the_reference_header_text():
    _need_reference()
    lay = layout_calc.layout_of([os.path.join(REF, h) for h in ("particle.h", "node.h",
                                                                 "element.h", "process.h")])
    gold = _golden("ref_job_layout.txt")
    structs = sorted({ln.split()[0] for ln in gold}, key=[ln.split()[0] for ln in gold].index)
    assert [ln for s in structs for ln in lay.lines(s)] == gold


def test_mpm_abi_job_t_and_threadtask_t_offsets(tmp_path):
    gold = _golden("ref_job_layout.txt")
    assert "job_t sizeof 800" in gold and "threadtask_t sizeof 96" in gold
    body = ['#include <stdio.h>', '#include <stddef.h>', '#include "mpm_abi.h"',
            'int main(void){']
    for ln in gold:
        t, f, _ = ln.split()
        if f == "sizeof":
            body.append(f'printf("{t} sizeof %zu\\n", sizeof({t}));')
        else:
            body.append(f'printf("{t} {f} %zu\\n", offsetof({t}, {f}));')
    body.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(body))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=gnu99", "-I", os.path.join(REPO, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert [ln for ln in out.split("\n") if ln.strip()] == gold
