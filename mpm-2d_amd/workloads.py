"""Synthetic inputs for the USL workloads (seeded, reproducible).

Particles sit on a 4 x 4 lattice per cell at ((k + 1/2) / 4) h with a uniform
jitter of +-0.1 h per axis, so no particle starts outside its cell.  The
jitter comes from SplitMix64 keyed by (seed, particle index, axis), so any
subset of particles can be regenerated independently.  Volume is the
lattice area (h/4)^2 and mass rho * volume, as the reference's particle
generators do (generate_block_particles.py:21-62, gen_particles.m); initial
velocity and stress are zero.  Particles are numbered cell by cell (cell
rows bottom to top, cells left to right, then lattice rows and columns), so
memory order starts out cell-sorted.

The grid always spans [0, 1]^2 (process_usl.c:43-44 hard-codes the unit
domain): N = cells + 1 nodes per side, h = 1 / (N - 1).
"""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15)) & M64
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
    return z ^ (z >> np.uint64(31))


def uniform01(seed, ids, stream):
    """U[0,1) keyed by (seed, id, stream), 53-bit resolution."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0xD1342543DE82EF95)
               + ids.astype(np.uint64) * np.uint64(4) + np.uint64(stream)) & M64
        return (splitmix64(key) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def block(cells_x, cells_y, n_cells, seed=1, rho=1500.0, ppc_side=4, jitter=0.1,
          cell_x0=0, cell_y0=0):
    """Particles filling cells [cell_x0, cell_x0+cells_x) x [cell_y0, ...)."""
    h = 1.0 / n_cells
    k = ppc_side
    cy, cx, b, a = np.meshgrid(np.arange(cells_y) + cell_y0, np.arange(cells_x) + cell_x0,
                               np.arange(k), np.arange(k), indexing="ij")
    cy, cx, b, a = (q.ravel() for q in (cy, cx, b, a))
    ids = np.arange(cx.size, dtype=np.uint64)
    jx = (uniform01(seed, ids, 0) - 0.5) * (2.0 * jitter) * h
    jy = (uniform01(seed, ids, 1) - 0.5) * (2.0 * jitter) * h
    x = (cx + (a + 0.5) / k) * h + jx
    y = (cy + (b + 0.5) / k) * h + jy
    v = np.full(x.size, (h / k) ** 2)
    m = rho * v
    z = np.zeros(x.size)
    return {"m": m, "v": v, "x": x, "y": y, "vx": z.copy(), "vy": z.copy(),
            "sxx": z.copy(), "sxy": z.copy(), "syy": z.copy()}


def triangle_pile(base_cells, height_cells, n_cells, seed=3, rho=1500.0, ppc_side=4):
    """An isosceles pile: lattice points of the base_cells-wide block that lie
    under the line from (0, height) to (base/2, ...) mirrored about the
    centre, centred in x."""
    x0 = (n_cells - base_cells) // 2
    p = block(base_cells, height_cells, n_cells, seed=seed, rho=rho, ppc_side=ppc_side,
              cell_x0=x0)
    h = 1.0 / n_cells
    xc = (x0 + base_cells / 2.0) * h
    half = base_cells / 2.0 * h
    keep = p["y"] <= height_cells * h * (1.0 - np.abs(p["x"] - xc) / half)
    return {k: v[keep].copy() for k, v in p.items()}


# Material parameters of the workloads (E = 1e6, nu = 0.3, mu_s = 0.38,
# mu_2 = 0.64, I_0 = 0.279, d = 1e-3); the reference's g_local_mu2
# hard-codes its own values of the rest (materialsrc/g_local_mu2.c:30-41).
WORKLOAD_MATERIAL = dict(E=1e6, nu=0.3, mu_s=0.38, mu_2=0.64, I_0=0.279,
                         grains_d=1e-3, grains_rho=2450.0, rho_crit=1485.0)
REFERENCE_MATERIAL = dict(E=1e6, nu=0.3, mu_s=0.3819, mu_2=0.6435, I_0=0.278,
                          grains_d=0.001 * 5, grains_rho=2450.0, rho_crit=1485.0)

BOX_FRICTION = dict(bc_kind=1, bc_fp64=(0.0, 0.0, 1.0, 1.0), bc_int=(1, 3, 1, 0))


def column_collapse(seed=1):
    """Granular column collapse: 64 x 64 cells, 16 x 32 cell block, 8,192
    particles, box with frictionless walls and no-slip floor."""
    n = 64
    return dict(name="column_collapse_64", N=n + 1, h=1.0 / n, dt=1e-5, steps=1000,
                material=WORKLOAD_MATERIAL, g_mag=9.81, ramp_time=0.1, **BOX_FRICTION,
                particles=block(16, 32, n, seed=seed))


def silo(seed=2, cells=512):
    """Square-silo discharge: 512 x 512 cells, 256 x 256 cell block
    (1,048,576 particles), wall at 0.6, orifice radius 0.05 open from t=0."""
    n = cells
    blk = cells // 2
    return dict(name=f"silo_{n}", N=n + 1, h=1.0 / n, dt=2e-6, steps=2000,
                material=WORKLOAD_MATERIAL, g_mag=9.81, ramp_time=0.1,
                bc_kind=2, bc_fp64=(0.05, 0.0, 0.6), bc_int=(),
                particles=block(blk, blk, n, seed=seed))


def small_box(cells=16, bx=5, by=9, seed=7, material=None, **kw):
    """A small column for parity tests (seconds on the oracle)."""
    d = dict(name=f"small_{cells}", N=cells + 1, h=1.0 / cells, dt=1e-5, steps=20,
             material=material or REFERENCE_MATERIAL, g_mag=9.8, ramp_time=0.1,
             **BOX_FRICTION, particles=block(bx, by, cells, seed=seed))
    d.update(kw)
    return d


def write_reference_inputs(w, grid_path, particle_path):
    """Writes the workload in the reference's grid and particle file formats
    (src/reader.c:17-119), with 17 significant digits so every double
    survives the round trip."""
    with open(grid_path, "w") as f:
        f.write(f"{w['N']}\n1\n")
    p = w["particles"]
    cols = [p[k] for k in ("m", "v", "x", "y", "vx", "vy", "sxx", "sxy", "syy")]
    with open(particle_path, "w") as f:
        f.write(f"{len(p['x'])}\n")
        for row in zip(*cols):
            f.write(" ".join(repr(float(v)) for v in row) + "\n")


# g_nonlocal constants (materialsrc/g_nonlocal.c:31-47, 141, 313); the heap
# workload sets the nonlocal amplitude A and mu_s.
REFERENCE_NONLOCAL = dict(E=1e8, nu=0.3, mu_s=0.3819, A=0.48, B=0.9377 * 20, d=0.01 * 5,
                          rho=2450.0, p_cap=1e-2, c=1e-2)
HEAP_NONLOCAL = dict(REFERENCE_NONLOCAL, mu_s=0.38, A=0.48)


def lithostatic(p, surface, rho=1500.0, g=9.81, k0=0.4):
    """Sets a lithostatic stress: syy = -rho g (surface(x) - y), sxx = k0 syy,
    sxy = 0.  With k0 = 0.4, tau / p = (1 - k0) / (1 + k0) = 0.43 > mu_s, so
    every particle starts with a nonzero local fluidity."""
    depth = np.maximum(surface(p["x"]) - p["y"], 0.0)
    p["syy"] = -rho * g * depth
    p["sxx"] = k0 * p["syy"]
    p["sxy"] = np.zeros_like(depth)
    return p


def heap(cells=256, seed=3, **kw):
    """Nonlocal fluidity heap: cells x cells grid, an isosceles pile with a
    base of `cells` cells and a height of cells/2 (262,144 particles at 256),
    g_nonlocal, box with frictionless walls and no-slip floor, dt = 5e-6.
    The pile starts from the lithostatic state above (same input to the GPU
    and the CPU step)."""
    h = 1.0 / cells
    height, half, xc = (cells // 2) * h, 0.5, 0.5
    parts = lithostatic(triangle_pile(cells, cells // 2, cells, seed=seed),
                        lambda x: height * (1.0 - np.abs(x - xc) / half))
    d = dict(name=f"heap_{cells}", N=cells + 1, h=h, dt=5e-6, steps=500,
             material=WORKLOAD_MATERIAL, nonlocal_material=HEAP_NONLOCAL,
             g_mag=9.81, ramp_time=0.1, **BOX_FRICTION, particles=parts)
    d.update(kw)
    return d


def ensemble(members=256, seed=11, p_min=2048, p_max=65536, sides=None):
    """An ensemble of column collapses of uneven size: member k has
    log-uniform particle count in [p_min, p_max], a grid of n cells per side
    with n = 4 round(sqrt(count / 2) / 4) (so the n/4 x n/2 cell block holds
    2 n^2 particles), mu_s ~ U[0.3, 0.5] and E ~ log-U[5e5, 5e6]; box walls as
    in the column collapse, dt = 1e-5.  `sides` overrides the cell counts."""
    ids = np.arange(members, dtype=np.uint64)
    if sides is None:
        u = uniform01(seed, ids, 0)
        counts = np.exp(np.log(p_min) + u * (np.log(p_max) - np.log(p_min)))
        sides = np.clip(4 * np.rint(np.sqrt(counts / 2.0) / 4.0), 8, 46336).astype(int)
    sides = [int(n) for n in sides]
    mu_s = 0.3 + 0.2 * uniform01(seed, ids, 1)
    E = np.exp(np.log(5e5) + uniform01(seed, ids, 2) * (np.log(5e6) - np.log(5e5)))
    out = []
    for k, n in enumerate(sides):
        mat = dict(WORKLOAD_MATERIAL, mu_s=float(mu_s[k]), E=float(E[k]))
        out.append(dict(name=f"member_{k}_{n}", N=n + 1, h=1.0 / n, dt=1e-5,
                        material=mat, g_mag=9.81, ramp_time=0.1, **BOX_FRICTION,
                        particles=block(n // 4, n // 2, n, seed=seed * 1000 + k)))
    return out


def adjoint_block(seed=4, cells=128, material=None):
    """Adjoint workload: 128 x 128 cells (129^2 nodes), a 64 x 64 cell block
    (65,536 particles) at x in [0.25, 0.75), y in [0, 0.5) under full gravity
    from the first step, with a lithostatic initial stress (k0 = 0.4, so
    tau/p = 0.43 starts above mu_s) and an initial slide vx = 0.5 m/s over
    the no-slip floor; box walls as in the column collapse, dt = 1e-5, 200
    steps.  The target run uses mu_s = 0.42 and E = 2e6."""
    n = cells
    p = block(n // 2, n // 2, n, seed=seed, cell_x0=n // 4)
    top = 0.5
    p = lithostatic(p, lambda x: top + 0.0 * x, rho=1500.0, g=9.81, k0=0.4)
    p["vx"] = np.full(len(p["x"]), 0.5)
    return dict(name=f"adjoint_block_{n}", N=n + 1, h=1.0 / n, dt=1e-5, steps=200,
                material=material or dict(WORKLOAD_MATERIAL), g_mag=9.81, ramp_time=1e-9,
                **BOX_FRICTION, particles=p)


ADJOINT_TARGET = dict(WORKLOAD_MATERIAL, mu_s=0.42, E=2e6)
