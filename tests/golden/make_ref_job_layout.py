"""Writes tests/golden/ref_job_layout.txt: the field offsets of the
reference's job_t, threadtask_t and the control structs inside job_t,
computed by oracle/layout_calc.py from the text of the reference's
libmpm/process.h (with particle.h, node.h and element.h before it).
process.h includes <suitesparse/cs.h>, absent here, so it cannot be
compiled; the calculator is pinned against the compiler on the headers
that do compile (tests/test_abi_layout.py).

    python tests/golden/make_ref_job_layout.py [/root/reference]
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))

import layout_calc  # noqa: E402

STRUCTS = ("timestep_control_t", "implicit_control_t", "output_control_t",
           "material_control_t", "boundary_control_t", "job_t", "threadtask_t")
HEADERS = ("particle.h", "node.h", "element.h", "process.h")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    lay = layout_calc.layout_of([os.path.join(ref, "libmpm", h) for h in HEADERS])
    with open(os.path.join(HERE, "ref_job_layout.txt"), "w") as f:
        for s in STRUCTS:
            f.write("\n".join(lay.lines(s)) + "\n")


if __name__ == "__main__":
    main()
