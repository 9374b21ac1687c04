// mpm_hip.hip -- gfx950 implementation of the explicit USL MPM step behind
// the C ABI declared in include/mpm_hip.h.
//
// One step is two kernels on one stream:
//
//   k_p2g_grid   16x8 grid nodes per workgroup, two or three threads per
//                node.  For each colour in turn the workgroup stages, in LDS,
//                the local coordinates and mass/momentum/stress products of
//                the particles in that colour's cells among the 17x9 cells
//                around its tile, ranked by particle index within each cell.
//                Each node's chains then fold the contributions of its four
//                cells in colour order 0..3 and, inside a cell, its particles
//                in ascending index -- the summation order of the
//                reference's coloured, barrier-separated scatter
//                (libmpm/process_usl.c:812-910, libmpm/map.c:85-104) -- so
//                the node sums are bitwise those of the CPU step, with no
//                atomics.  The boundary conditions (bc_momentum/bc_force,
//                bcsrc/box.c:183-224) and the grid update (move_grid_split,
//                libmpm/process_usl.c:596-625) follow, and one 64-byte node
//                record per node is written.
//
//   k_g2p_stress one thread per particle: G2P (process_usl.c:914-957),
//                the next step's element (process_usl.c:43-44, 631-674) and
//                insertion into the next step's cell bins, strain rate
//                (764-808), volume update (961-970) and the g_local_mu2
//                stress update (materialsrc/g_local_mu2.c:197-307).
//
// Shape functions (process_usl.c:727-760, particle.c:14-30,
// interpolate.c:23-65) are recomputed where needed from the same inputs with
// the same operations, so both kernels see identical values.  Everything is
// FP64 and compiled with -ffp-contract=off; exp is the IEEE-only mpm_exp.
//
// The translation unit is this file and the parts it includes, in order:
//   mpm_types.inc             constants, device and table structs
//   mpm_point_ops.inc         shape functions, boundary masks, material laws,
//                             bin insertion
//   mpm_step_kernels.inc      binning, P2G (+ the ensembles' light tiles), G2P
//   mpm_nonlocal_kernels.inc  g_nonlocal assembly, CG solve, stress update
//   mpm_adjoint_kernels.inc   reverse kernels of the adjoint
//   mpm_job.inc               a single job's context, launches and C ABI
//   mpm_batch.inc             ensembles: batch, dispatch tables, graph replay
//   mpm_adjoint.inc           adjoint runs and the remaining C-ABI helpers
//   mpm_noalloc.inc           the step on caller-owned arrays with a static
//                             scratch array (no allocation per call)

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>

#include "mpm_exp.h"
#include "mpm_hip.h"

namespace {
#include "mpm_types.inc"
#include "mpm_point_ops.inc"
#include "mpm_step_kernels.inc"
#include "mpm_nonlocal_kernels.inc"
#include "mpm_adjoint_kernels.inc"

}  // namespace

#include "mpm_job.inc"
#include "mpm_batch.inc"
#include "mpm_adjoint.inc"
#include "mpm_noalloc.inc"
