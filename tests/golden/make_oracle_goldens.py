"""Writes small regression fixtures from the CPU oracle (parity build).

Each fixture holds the inputs (particle arrays, configuration) and the
oracle's particle and node state after a number of steps, so the GPU path can
be checked on a box where the oracle is not built, and so any later change of
the oracle that moves a single bit is caught.

    python tests/golden/make_oracle_goldens.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "mpm-2d_amd"), os.path.join(REPO, "oracle")]

import pyoracle  # noqa: E402
import workloads  # noqa: E402

CASES = {
    "small_box_20": (lambda: workloads.small_box(cells=16, bx=5, by=9), 20),
    "silo_16_50": (lambda: workloads.silo(seed=5, cells=16), 50),
}


def main():
    for name, (make, steps) in CASES.items():
        w = make()
        o = pyoracle.from_workload(w, num_threads=1)
        o.run(steps)
        p, nd = o.particles(), o.nodes()
        cfg = {k: w[k] for k in ("N", "h", "dt", "material", "g_mag", "ramp_time")}
        cfg.update(bc_kind=w.get("bc_kind", 0), bc_fp64=list(w.get("bc_fp64", ())),
                   bc_int=list(w.get("bc_int", ())), steps=steps, name=name)
        arrays = {f"in_{k}": v for k, v in w["particles"].items()}
        keep = ["x", "y", "vx", "vy", "v", "txx", "txy", "tyy", "tzz", "gammap", "gammadotp",
                "ux", "uy"]
        arrays.update({f"out_{k}": p[k] for k in keep})
        arrays["out_active"] = p["active"]
        arrays.update({f"node_{k}": nd[k] for k in pyoracle.NODE_FIELDS[:7]})
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), config=json.dumps(cfg), **arrays)
        print(name, len(w["particles"]["x"]), "particles,", steps, "steps")


if __name__ == "__main__":
    main()
