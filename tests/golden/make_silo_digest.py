"""Writes tests/golden/silo_512_digest.json: SHA-256 digests of the oracle's
state on the full square-silo workload (512 x 512 cells, 1,048,576
particles, dt = 2e-6) after 1, 100, 1000 and 2000 steps.

The oracle's parity build (-ffp-contract=off, portable exp) is bitwise
independent of its thread count, and the GPU step is bitwise equal to it,
so a digest of each field's values stands in for the 117 MB state.  Values
are hashed as little-endian float64 after x + 0.0 (which maps -0.0 to +0.0:
the parity bar compares values, and +0 == -0).  Sums and a few sampled
values are kept beside each digest so that a mismatch says where it is.

    python tests/golden/make_silo_digest.py [threads]

Takes about ten minutes on eight cores.
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "mpm-2d_amd"), os.path.join(REPO, "oracle")]

import numpy as np  # noqa: E402

import pyoracle  # noqa: E402
import workloads  # noqa: E402

FIELDS = ["x", "y", "vx", "vy", "v", "txx", "txy", "tyy", "tzz", "gammap", "gammadotp",
          "ux", "uy"]
NODE_FIELDS = ["m", "mx", "my", "fx", "fy", "vx", "vy"]
CHECKPOINTS = (1, 100, 1000, 2000)
SAMPLE = np.array([0, 1, 4095, 65536, 262143, 524288, 786431, 1048575])


def digest(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64) + 0.0)
    return hashlib.sha256(a.astype("<f8").tobytes()).hexdigest()


def summary(a):
    a = np.asarray(a, dtype=np.float64)
    return {"sha256": digest(a), "sum": float(np.sum(a)),
            "sample": [float(v) for v in a[SAMPLE[SAMPLE < a.size]]]}


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    w = workloads.silo()
    o = pyoracle.from_workload(w, num_threads=threads)
    out = {"workload": "silo_512 (workloads.silo(), seed 2)", "steps": list(CHECKPOINTS),
           "checker": "oracle parity build", "states": {}}
    done, t0 = 0, time.time()
    for n in CHECKPOINTS:
        o.run(n - done)
        done = n
        p = o.particles()
        nd = o.nodes()
        st = {k: summary(p[k]) for k in FIELDS}
        st["active"] = {"count": int(p["active"].sum()),
                        "sha256": hashlib.sha256(p["active"].astype("<i4").tobytes()).hexdigest()}
        st["nodes"] = {k: summary(nd[k]) for k in NODE_FIELDS}
        st["time"] = o.time()
        out["states"][str(n)] = st
        print(f"step {n}: active {st['active']['count']}, {time.time() - t0:.0f} s", flush=True)
    with open(os.path.join(HERE, "silo_512_digest.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
