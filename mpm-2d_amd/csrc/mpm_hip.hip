// mpm_hip.hip -- gfx950 implementation of the explicit USL MPM step behind
// the C ABI declared in include/mpm_hip.h.
//
// One step is two kernels on one stream:
//
//   k_p2g_grid   one thread per grid node, 16x8 nodes per workgroup.  For
//                each colour in turn the workgroup loads the particle lists of
//                that colour's cells among the 17x9 cells around its tile,
//                sorts each list by particle index, and stages the particles'
//                local coordinates and mass/momentum/stress products in LDS
//                (coalesced loads, each particle read once per workgroup).
//                Each thread then folds, for its own node, the contributions
//                of the four adjacent cells in colour order 0..3 and, inside
//                a cell, of its particles in ascending index -- the exact
//                summation order of the reference's coloured, barrier-
//                separated scatter (libmpm/process_usl.c:812-910,
//                libmpm/map.c:85-104) -- so the node sums are bitwise those
//                of the CPU step, with no atomics.  The same thread applies
//                the boundary conditions (bc_momentum/bc_force,
//                bcsrc/box.c:183-224) and the grid update (move_grid_split,
//                libmpm/process_usl.c:596-625) and writes one 64-byte node
//                record.
//
//   k_g2p_stress one thread per particle: G2P (process_usl.c:914-957),
//                strain rate (764-808), volume update (961-970), the
//                g_local_mu2 stress update (materialsrc/g_local_mu2.c:
//                197-307), the next step's element (process_usl.c:43-44,
//                631-674) and insertion into the next step's cell bins.
//
// Shape functions (process_usl.c:727-760, particle.c:14-30,
// interpolate.c:23-65) are recomputed where needed from the same inputs with
// the same operations, so both kernels see identical values.  Everything is
// FP64 and compiled with -ffp-contract=off; exp is the IEEE-only mpm_exp.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <type_traits>

#include "mpm_exp.h"
#include "mpm_hip.h"

namespace {

constexpr double kTol = 1e-10;     // TOL in process_usl.c:25 and particle.c:10
constexpr int kTileX = 16;         // nodes per tile in k_p2g_grid: 16 x 8
constexpr int kTileY = 8;
constexpr int kRegX = kTileX + 1;  // cells whose particles reach the tile: 17 x 9
constexpr int kRegY = kTileY + 1;
constexpr int kRegCells = kRegX * kRegY;
constexpr int kColourCellsMax = ((kRegX + 1) / 2) * ((kRegY + 1) / 2);   // 9 x 5
constexpr int kStageMax = 720;     // staged records per batch (64 B each); a multiple
                                   // of 16, the LDS bank swizzle's block
constexpr int kTileNodes = kTileX * kTileY;
// k_p2g_grid runs two or three threads per node (ROLES): the node's five
// sums are split between them.
constexpr int kG2PThreads = 256;
constexpr int kBatchTileChunk = 16;   // ensemble P2G tiles per XCD chunk
constexpr int kLightGroup = 32;       // ensemble P2G: empty-at-upload tiles per workgroup
constexpr int kFixBlocks = 256;       // ensemble P2G: workgroups for listed light tiles
// Particle count limit: byte offsets into a particle field fit the 31-bit
// range of a raw buffer load.
constexpr size_t kMaxParticles = (size_t)1 << 28;
// Word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data format,
// no swizzle): loads past num_records return 0.
constexpr int kBufferDword3 = 0x00020000;

enum : int { kStatusBinOverflow = 1 };
enum : int { kModeStoreAux = 1, kModeHostMaterial = 2, kModeNonlocal = 4, kModeIsolin = 8,
              kModeMuslMove = 16, kModeMuslFinish = 32, kModeDpRiMt = 64 };

struct DevState {
    double t;           // job->t
    double bx, by;      // body force of the next step (loading.c)
    double t0;          // job->step_start_time
    unsigned long long crossings;
    int status;
    int pad;
};

struct GridParams {
    int N;              // nodes per side
    int Ne;             // cells per side
    int ncells;
    int cap;            // bin capacity
    double h;
    double inv_h;       // grad_tent's 1.0 / h
    int h_pow2;         // h is a power of two
    int m_finite;       // every particle mass is finite (checked on the host)
    double dt;
    double g_mag, ramp;
    int bc_kind;
    int left_col, right_col, bottom_row, top_row;
    int left_fric, bottom_fric, right_fric, top_fric;
    int wall_col;
    double hole_radius, open_time;
};

struct Material {
    double G, two_G, lambda;
    double mu_s, mu_2, I_0, grains_d, grains_rho, rho_crit;
    // g_nonlocal (materialsrc/g_nonlocal.c:31-47)
    double nl_E, nl_nu, nl_mu_s, nl_A, nl_B, nl_d, nl_rho, nl_p_cap, nl_c;
};

struct Particles {
    double *x, *y, *vx, *vy, *m, *v;
    double *txx, *txy, *tyy, *tzz, *gp, *gdp, *ux, *uy;
    double *exx, *exy, *eyy, *wxy, *ax, *ay, *xl, *yl;   // auxiliary
    double *gf, *gf_bulk, *xisq_inv;                     // g_nonlocal state
    int *elem;          // element of the next step, -1 outside the grid
    int *active;
    int *elem_prev;     // g_nonlocal: element of the step just taken
    double *xn, *yn;    // MUSL: positions after G2P, committed after the stress
};

struct alignas(64) NodeRec {
    double m, mx, my, fx, fy, vx, vy, pad;
};

// One member of an ensemble stepped by shared launches: its own grid,
// constants and slices of the concatenated particle, node and cell arrays.
// Member particle ranges start at multiples of kG2PThreads, so a particle
// workgroup never spans two members.
struct MemberInfo {
    GridParams g;
    Material mt;
    int node_base, cell_base, part_base, np;
    int tile_base, tiles_x;
};

// Dispatch order of a single job's P2G tiles, in three ranges of
// workgroups: the core tiles of the particles (cx0.., cy0.., cw x ch tiles),
// then the rest of the box around them (bx0.., by0.., bw x bh tiles: every
// tile whose cell region held a particle when the particles were last
// uploaded; its rim holds the lightly loaded edge tiles), then the tiles
// outside the box.  The tiles with work start first instead of queueing
// behind empty ones, and the heavy tiles are dealt evenly to the CUs before
// the light ones take the spare workgroup slots.  Any order gives the same
// node sums (each tile is written by one workgroup), so a stale box only
// costs time.
struct TileOrder {
    int cx0, cy0, cw, ch;
    int bx0, by0, bw, bh;
    int blocks_a, blocks_b;   // workgroups of the core and of the rim
    int core_rows;            // core rows per XCD chunk
};

// Ensemble tables (device memory); unused by single-job launches.
struct BatchTables {
    const MemberInfo *mem;
    const int2 *tile_order;    // P2G workgroup -> (tile, member); tile -1: a spare one
    const int2 *light_tiles;   // tiles for k_p2g_light: (tile, member)
    const int *block_member;   // particle workgroup -> member
    DevState *st;              // per member
};

// ---------------------------------------------------------------------------
// Point operations shared by both kernels.
// ---------------------------------------------------------------------------

// a / h.  When h is a power of two the quotient is exact, and so is the
// product with 1 / h (a power of two too): the same value, for one
// multiplication instead of the division sequence.
__device__ __forceinline__ double div_h(double a, const GridParams &g)
{
    return g.h_pow2 ? a * g.inv_h : a / g.h;
}

// WHICH_ELEMENT4 (process_usl.c:43-44).  An index at or beyond the element
// count (an ulp-level rounding case the reference reads out of bounds for)
// is treated as outside.
__device__ __forceinline__ int which_element(double x, double y, const GridParams &g)
{
    if (x < 1.0 && x >= 0.0 && y < 1.0 && y >= 0.0) {
        double e = floor(div_h(x, g)) + floor(div_h(y, g)) * (double)g.Ne;
        int ei = (int)e;
        return ei < g.ncells ? ei : -1;
    }
    return -1;
}

// global_to_local_coords' snapping of |xl| in (1, 1 + TOL) to +-1
// (particle.c:14-30).
__device__ __forceinline__ double snap_local(double l)
{
    const double sgn = copysign(1.0, l);
    if ((sgn * l) > 1.0 && ((sgn * l) - 1.0) < kTol) {
        return sgn;
    }
    return l;
}

struct Shape {
    double s[4], b1[4], b2[4];
    double xl, yl;
};

// tent / grad_tent (interpolate.c:23-65) for the cell with lower-left node
// (cc, rr); node order BL, BR, TR, TL.
__device__ __forceinline__ void shape_functions(double x, double y, int cc, int rr,
                                                const GridParams &g, Shape &sh)
{
    const double xn = (double)cc * g.h;
    const double yn = (double)rr * g.h;
    const double xl = snap_local(div_h(x - xn, g));
    const double yl = snap_local(div_h(y - yn, g));
    const double ox = 1.0 - xl, oy = 1.0 - yl;
    sh.s[0] = ox * oy;
    sh.s[1] = xl * oy;
    sh.s[2] = xl * yl;
    sh.s[3] = ox * yl;
    sh.b1[0] = -oy * g.inv_h;
    sh.b1[1] = oy * g.inv_h;
    sh.b1[2] = yl * g.inv_h;
    sh.b1[3] = -yl * g.inv_h;
    sh.b2[0] = -ox * g.inv_h;
    sh.b2[1] = -xl * g.inv_h;
    sh.b2[2] = xl * g.inv_h;
    sh.b2[3] = ox * g.inv_h;
    sh.xl = xl;
    sh.yl = yl;
}

// Which DOFs of node (i, j) the boundary plugin fixes: bit 0 = X, bit 1 = Y.
//   box         bcsrc/box.c:106-169
//   square_silo bcsrc/square_silo.c:58-92 (trapdoor open once t > open_time)
__device__ __forceinline__ unsigned bc_fixed(const GridParams &g, int i, int j,
                                             double t_now, const uint8_t *mask)
{
    unsigned f = 0;
    switch (g.bc_kind) {
    case MPM_BC_BOX:
        if (i >= g.left_col && i <= g.right_col) {
            if (j == g.bottom_row) f |= (unsigned)g.bottom_fric & 3u;
            if (j == g.top_row) f |= (unsigned)g.top_fric & 3u;
        }
        if (j >= g.bottom_row && j <= g.top_row) {
            if (i == g.left_col) f |= (unsigned)g.left_fric & 3u;
            if (i == g.right_col) f |= (unsigned)g.right_fric & 3u;
        }
        break;
    case MPM_BC_SQUARE_SILO:
        if (j == 0) {
            const double xn = (double)i * g.h;
            if (!(xn <= g.hole_radius && t_now > g.open_time)) f |= 3u;
        }
        if (i == 0 || i == g.wall_col) f |= 1u;
        break;
    case MPM_BC_MASK:
        f = mask[(size_t)j * g.N + i] & 3u;
        break;
    default:
        break;
    }
    return f;
}

// g_local_mu2 calculate_stress_threaded body (g_local_mu2.c:197-307) on the
// four independent components of the plane-strain stress.  The reference
// carries full 3x3 tensors whose XZ/YZ/ZX/ZY entries stay zero and whose
// YX entry stays equal to XY; dropping those terms changes no value (only
// the sign of an exact zero could differ).  Every remaining operation is
// the reference's, in its order.
__device__ __forceinline__ void g_local_mu2_update(double exx, double exy, double eyy,
                                                   double wxy, double m, double v,
                                                   double dt, const Material &mt,
                                                   double &txx, double &txy,
                                                   double &tyy, double &tzz,
                                                   double &gp, double &gdp)
{
    // L, W = skw(L), D = sym(L)
    const double lxy = exy + wxy;
    const double lyx = exy - wxy;
    const double w = 0.5 * (lxy - lyx);
    const double dxx = 0.5 * (exx + exx);
    const double dxy = 0.5 * (lxy + lyx);
    const double dyy = 0.5 * (eyy + eyy);
    const double trD = (dxx + dyy) + 0.0;
    // Jaumann term JS = W T - T W
    const double a = w * txy;
    const double jsxx = a + a;
    const double jsxy = w * tyy - w * txx;
    const double jsyy = -(a + a);
    // trial stress Ttr = T + dt (2G D + lambda tr(D) I + JS)
    const double ltr = mt.lambda * trD;
    double rxx = ((dxx * mt.two_G + ltr) + jsxx) * dt + txx;
    double rxy = ((dxy * mt.two_G) + jsxy) * dt + txy;
    double ryy = ((dyy * mt.two_G + ltr) + jsyy) * dt + tyy;
    double rzz = ltr * dt + tzz;
    // deviator, pressure, equivalent shear stress
    const double c = ((rxx + ryy) + rzz) / 3.0;
    const double d_xx = rxx - c, d_yy = ryy - c, d_zz = rzz - c;
    const double p_tr = c * -1.0;
    double tau2 = d_xx * d_xx;
    tau2 += rxy * rxy;
    tau2 += rxy * rxy;
    tau2 += d_yy * d_yy;
    tau2 += d_zz * d_zz;
    const double tau_tr = sqrt(0.5 * tau2);

    const int density_flag = (m / v) < mt.rho_crit;
    double nup_tau;
    if (density_flag || p_tr <= 0.0) {
        nup_tau = 0.0;
        txx = 0.0;
        txy = 0.0;
        tyy = 0.0;
        tzz = 0.0;
    } else if (p_tr > 0.0) {
        const double S0 = mt.mu_s * p_tr;
        double tau_tau, scale;
        if (tau_tr <= S0) {
            tau_tau = tau_tr;
            scale = 1.0;
        } else {
            const double S2 = mt.mu_2 * p_tr;
            const double alpha = mt.G * mt.I_0 * dt * sqrt(p_tr / mt.grains_rho) / mt.grains_d;
            const double B = -(S2 + tau_tr + alpha);
            const double H = S2 * tau_tr + S0 * alpha;
            // negative_root(1, B, H) (g_local_mu2.c:90-99)
            if (B > 0.0) {
                tau_tau = (-B - sqrt(B * B - 4.0 * H)) / 2.0;
            } else {
                tau_tau = (2.0 * H) / (-B + sqrt(B * B - 4.0 * H));
            }
            scale = tau_tau / tau_tr;
        }
        nup_tau = ((tau_tr - tau_tau) / mt.G) / dt;
        txx = d_xx * scale - p_tr;
        txy = rxy * scale;
        tyy = d_yy * scale - p_tr;
        tzz = d_zz * scale - p_tr;
    } else {
        nup_tau = 0.0;   // NaN pressure: stress left unchanged
    }
    gp += nup_tau * dt;
    gdp = nup_tau;
}

// Appends particle p to the bin of cell `key` when `valid`.  The lanes of a
// wave that share a cell take consecutive slots in lane order with one
// atomic, so a bin is sorted by particle index whenever its particles come
// from one wave (the fold order readers need; they check and sort the rare
// bin filled from several waves).  When every cell of the wave is one run
// of consecutive lanes with cells rising from run to run (particles stored
// cell by cell), the runs are found with two ballots; otherwise one leader
// per distinct cell is picked in turn.  Must be reached by every lane of the
// wave that has not exited.
__device__ __forceinline__ void bin_insert(bool valid, int key, int p, int *cnt, int *bkt,
                                           int cap, int *status)
{
    const int lane = __lane_id();
    const unsigned long long lane_bit = 1ull << lane;
    const unsigned long long vmask = __ballot(valid);
    const int prev = __shfl_up(key, 1);
    const bool prev_same = lane > 0 && ((vmask >> (lane - 1)) & 1ull) && prev == key;
    const bool start = valid && !prev_same;
    const unsigned long long starts = __ballot(start);
    // the run start below this lane's own, and its cell
    const unsigned long long lower = starts & (lane_bit - 1ull);
    const int ps = lower ? 63 - __builtin_clzll(lower) : lane;
    const int pkey = __shfl(key, ps);
    const bool rising = __ballot(start && lower && pkey >= key) == 0;
    int base = 0, r = 0;
    if (rising) {
        if (!valid) {
            return;
        }
        // this lane's run: from the last start at or below it up to the
        // next start or invalid lane above that start
        const int s0 = 63 - __builtin_clzll(starts & ((2ull << lane) - 1ull));
        const unsigned long long above = (starts | ~vmask) & ~((2ull << s0) - 1ull);
        const int s1 = above ? __builtin_ctzll(above) : 64;
        if (start) {
            base = atomicAdd(&cnt[key], s1 - s0);
        }
        base = __shfl(base, s0);
        r = base + (lane - s0);
    } else {
        unsigned long long pending = vmask;
        while (pending) {
            const int leader = __builtin_ctzll(pending);
            const int lkey = __shfl(key, leader);
            const unsigned long long group = __ballot(valid && key == lkey) & pending;
            int b = 0;
            if (lane_bit == (1ull << leader)) {
                b = atomicAdd(&cnt[lkey], __popcll(group));
            }
            b = __shfl(b, leader);
            if (group & lane_bit) {
                r = b + __popcll(group & (lane_bit - 1));
            }
            pending &= ~group;
        }
        if (!valid) {
            return;
        }
    }
    if (r < cap) {
        bkt[(size_t)key * cap + r] = p;
    } else {
        atomicOr(status, kStatusBinOverflow);
    }
}

// calculate_bulk_granular_fluidity (g_nonlocal.c:306-352) for one particle:
// local fluidity gf_bulk and xi^-2 from the stress at the start of the
// material update.
__device__ __forceinline__ void nonlocal_bulk(double sxx, double sxy, double syy,
                                              const Material &mt, double &gf_bulk,
                                              double &xisq_inv)
{
    double p_t = -0.5 * (sxx + syy);
    const double sxx_tr0 = sxx + p_t;
    const double sxy_tr0 = sxy;
    const double syy_tr0 = syy + p_t;
    const double tau_t = sqrt(0.5 * (sxx_tr0 * sxx_tr0 + 2 * sxy_tr0 * sxy_tr0
                                     + syy_tr0 * syy_tr0));
    if (p_t < mt.nl_p_cap) {
        p_t = mt.nl_p_cap;
    }
    const double S0 = p_t * mt.nl_mu_s;
    const double eta = mt.nl_B * mt.nl_d * sqrt(p_t * mt.nl_rho);
    xisq_inv = fabs(tau_t / p_t - mt.nl_mu_s) / (mt.nl_d * mt.nl_d * mt.nl_A * mt.nl_A);
    gf_bulk = tau_t > S0 ? (p_t / eta) * (1 - S0 / tau_t) : 0.0;
}

__device__ __forceinline__ double gravity(double t, double t0, double g_mag, double ramp)
{
    // time_varying_loads (loading.c:34-54)
    if ((ramp + t0) > t) {
        return -g_mag * ((t - t0) / ramp);
    }
    return -g_mag;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Element assignment + binning of all particles (first step, or after the
// host rewrote positions).  init != 0 applies mpm_init's deactivation of
// particles outside the grid (process_usl.c:236-244); in an ensemble, only
// to member init - 1.
template <bool BATCH>
__global__ __launch_bounds__(kG2PThreads) void k_bin_particles(
    const Particles *__restrict__ Pd, int *__restrict__ cnt, int *__restrict__ bkt, GridParams g,
    DevState *__restrict__ st, int np, int init, BatchTables bt)
{
    const Particles &P = *Pd;   // field pointers, loaded as used
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (BATCH) {
        const int mb = bt.block_member[blockIdx.x];
        const MemberInfo &mi = bt.mem[mb];
        g = mi.g;
        np = mi.part_base + mi.np;
        cnt += mi.cell_base;
        bkt += (size_t)mi.cell_base * g.cap;
        st = bt.st + mb;
        init = init == mb + 1;
    }
    if (p >= np || (!init && !P.active[p])) {
        return;
    }
    const int e = which_element(P.x[p], P.y[p], g);
    P.elem[p] = e;
    if (init) {
        P.active[p] = e >= 0;
    }
    bin_insert(e >= 0, e, p, cnt, bkt, g.cap, &st->status);
}

// The v-th tile, in row-major order, of the tiles of the rectangle x0.., y0..
// (w tiles wide) that lie outside the inner rectangle ix0.., iy0.., iw x ih
// within it.
__device__ __forceinline__ int rect_rim_tile(int v, int x0, int y0, int w, int ix0, int iy0,
                                             int iw, int ih, int tiles_x)
{
    const int n0 = (iy0 - y0) * w;   // rows below the inner rectangle
    if (v < n0) {
        return (y0 + v / w) * tiles_x + x0 + v % w;
    }
    v -= n0;
    const int wr = w - iw;           // tiles beside it in each of its rows
    const int n1 = ih * wr;
    if (v < n1) {
        const int c = v % wr;
        return (iy0 + v / wr) * tiles_x + x0 + (c < ix0 - x0 ? c : c + iw);
    }
    v -= n1;
    return (iy0 + ih + v / w) * tiles_x + x0 + v % w;
}

// Tile of P2G workgroup b of a single job (TileOrder), -1 for a spare one.
// Within each range, chunks of consecutive tiles (a band of core rows, one
// tile of the rim, a grid row outside) go to the 8 XCDs in turn: workgroups
// are dealt to XCDs round-robin, so the tiles of a chunk share an XCD's L2
// (a band's tiles read their neighbours' edge cells from it) while any
// patch of the grid that holds particles is spread over all eight.
__device__ __forceinline__ int p2g_tile(int b, int tiles_x, int ntiles, const TileOrder &o)
{
    auto chunked = [](int b, int chunk) {
        const int xcd = b & 7, slot = b >> 3;
        return ((slot / chunk) * 8 + xcd) * chunk + slot % chunk;
    };
    if (b < o.blocks_a) {
        const int v = chunked(b, o.cw * o.core_rows);
        return v < o.cw * o.ch ? (o.cy0 + v / o.cw) * tiles_x + o.cx0 + v % o.cw : -1;
    }
    b -= o.blocks_a;
    if (b < o.blocks_b) {
        const int v = chunked(b, 1);
        return v < o.bw * o.bh - o.cw * o.ch
                   ? rect_rim_tile(v, o.bx0, o.by0, o.bw, o.cx0, o.cy0, o.cw, o.ch, tiles_x)
                   : -1;
    }
    b -= o.blocks_b;
    const int v = chunked(b, tiles_x);
    return v < ntiles - o.bw * o.bh
               ? rect_rim_tile(v, 0, 0, tiles_x, o.bx0, o.by0, o.bw, o.bh, tiles_x)
               : -1;
}

// P2G, boundary conditions and grid update of one tile (global index
// `tile`; member `mb` of an ensemble), by the whole workgroup.
// MUSL2: the second mapping of the MUSL step (process_usl.c:409-440): only
// the mass and momentum chains, then v = p / m without boundary conditions
// or the force update.
template <int CAP, bool BATCH, bool MUSL2, int ROLES>
__device__ __forceinline__ void p2g_tile_body(
    const Particles *__restrict__ Pd, const int *__restrict__ cnt_cur, const int *__restrict__ bkt_cur,
    int *__restrict__ cnt_next, NodeRec *__restrict__ nodes, uint8_t *__restrict__ tile_zero,
    const uint8_t *__restrict__ mask, GridParams g, const DevState *__restrict__ st,
    int tiles_x, int np, const BatchTables &bt, int tile, int mb)
{
    const Particles &P = *Pd;   // field pointers, loaded as used
    constexpr int kP2GThreads = ROLES * kTileNodes;
    constexpr int kWaves = kP2GThreads / 64;
    constexpr int kStageSlots = (kStageMax + kP2GThreads - 1) / kP2GThreads;   // per thread
    constexpr int kMaxBatches = 8;   // 45 cells of at most 64 particles: <= 5 batches
    static_assert(ROLES == 2 || ROLES == 3, "two or three threads per node");
    static_assert(kWaves >= 4, "one wave per colour scans the colour's cells");

    __shared__ int s_cnt[kRegCells];
    // per colour: exclusive scan of the cells' counts (compact list offsets),
    // staging batches of whole cells [s_bq[c][b], s_bq[c][b+1]) and their number
    __shared__ int s_off[4][kColourCellsMax + 1];
    __shared__ int s_bq[4][kMaxBatches + 1];
    __shared__ int s_nbat[4];
    // per colour: the cell holding list slot 16 b (slots are looked up from
    // there: one step for bins of 16 entries)
    __shared__ unsigned char s_blk[4][(kColourCellsMax * CAP + 15) / 16];
    __shared__ int s_inv[kWaves];   // a wave saw a bin out of index order
    // a batch's particle indices in bin order (used only to rank the
    // entries of bins that are out of index order)
    __shared__ int s_cmp[kStageSlots * kP2GThreads];
    // staged particles of one batch, paired for 16-byte LDS reads; slot s is
    // stored at s ^ ((s >> 4) & 15) so the cells folded side by side read
    // different LDS banks
    __shared__ double2 s_xy[kStageMax];     // local coordinates xl, yl
    __shared__ double2 s_msy[kStageMax];    // m, -sxy * v
    __shared__ double2 s_pxs[kStageMax];    // vx * m, -sxx * v
    __shared__ double2 s_pys[kStageMax];    // vy * m, -syy * v

    // 1: the tile's node records are all +0 (written so by the last launch,
    // or never written since the array was zeroed)
    const int zero_before = tile_zero[tile];
    const int flag_index = tile;
    if constexpr (BATCH) {
        const MemberInfo &mi = bt.mem[mb];
        g = mi.g;
        tile -= mi.tile_base;
        tiles_x = mi.tiles_x;
        cnt_cur += mi.cell_base;
        bkt_cur += (size_t)mi.cell_base * CAP;
        cnt_next += mi.cell_base;
        nodes += mi.node_base;
        st = bt.st + mb;
    }
    const int tx0 = (tile % tiles_x) * kTileX;
    const int ty0 = (tile / tiles_x) * kTileY;
    const int cx0 = tx0 - 1, cy0 = ty0 - 1;   // staged cell region origin
    const int tid = threadIdx.x;
    // Each chain is one node field, summed in the reference's order by
    // exactly one thread.  Two roles: role 0 folds the m, mx, fx chains,
    // role 1 the my, fy chains.  Three roles: role 0 the m, mx, my chains,
    // role 1 fx, role 2 fy (shorter chains, more waves per tile).
    const int role = __builtin_amdgcn_readfirstlane(tid / kTileNodes);   // uniform per wave
    const int nt = tid % kTileNodes;

    // This tile owns cells (tx0.., ty0..): clear their next-step bin counts.
    if (role == 0) {
        const int cc = tx0 + (nt % kTileX), rr = ty0 + (nt / kTileX);
        if (cc < g.Ne && rr < g.Ne) {
            cnt_next[rr * g.Ne + cc] = 0;
        }
    }
    // Bin counts of the 17 x 9 cells around the tile; a tile whose region
    // holds no particle skips to writing its (zero) node records, and writes
    // nothing when they are zero already.
    int occupied = 0;
    for (int cl = tid; cl < kRegCells; cl += kP2GThreads) {
        const int cc = cx0 + cl % kRegX, rr = cy0 + cl / kRegX;
        int n = 0;
        if (cc >= 0 && rr >= 0 && cc < g.Ne && rr < g.Ne) {
            n = cnt_cur[rr * g.Ne + cc];
            n = n < CAP ? n : CAP;
        }
        s_cnt[cl] = n;
        occupied |= n > 0;
    }
    const bool any_particles = __syncthreads_or(occupied) != 0;
    if (!any_particles && zero_before) {
        return;
    }
    if (tid == 0 && zero_before != (int)!any_particles) {
        tile_zero[flag_index] = !any_particles;
    }

    // Colour c's cells: region column a = 2*ax + ca holds cell column cx0 + a
    // (cx0 is odd, so colour parity pc <-> ca = 1 - pc); likewise rows.
    auto ncx_of = [](int c) { return (c & 1) ? (kRegX + 1) / 2 : kRegX / 2; };
    auto ncy_of = [](int c) { return (c >> 1) ? (kRegY + 1) / 2 : kRegY / 2; };

    if (any_particles && tid < 4 * 64) {
        // Wave w lays out colour w: list offsets (a scan of the counts) and
        // the staging batches.  Offsets never decrease and one cell holds at
        // most CAP <= kStageMax particles, so the cells fitting a batch form
        // a nonempty run, found with one ballot.
        const int w = tid >> 6, lane = tid & 63;
        const int ncx = ncx_of(w), nq = ncx * ncy_of(w);
        const int ca = (w & 1) ? 0 : 1, cb = (w >> 1) ? 0 : 1;
        int n = 0;
        if (lane < nq) {
            n = s_cnt[(2 * (lane / ncx) + cb) * kRegX + 2 * (lane % ncx) + ca];
        }
        int incl = n;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int up = __shfl_up(incl, d);
            if (lane >= d) incl += up;
        }
        if (lane < nq) {
            s_off[w][lane + 1] = incl;
            for (int blk = (incl - n + 15) >> 4; blk < (incl + 15) >> 4; ++blk) {
                s_blk[w][blk] = (unsigned char)lane;
            }
        }
        int q0 = 0, base = 0, nbat = 0;
        while (q0 < nq) {
            const unsigned long long fit =
                __ballot(lane >= q0 && lane < nq && incl - base <= kStageMax);
            const int q1 = q0 + __popcll(fit);
            ++nbat;
            if (lane == 0) {
                s_bq[w][nbat] = q1;
            }
            base = __shfl(incl, q1 - 1);
            q0 = q1;
        }
        if (lane == 0) {
            s_off[w][0] = 0;
            s_bq[w][0] = 0;
            s_nbat[w] = nbat;
        }
    }
    __syncthreads();

    // Particle fields as raw buffers: a load past the end returns 0 without
    // a memory request, which idle staging slots use instead of a branch.
    const int pbytes = np * (int)sizeof(double);
    auto rsrc = [&](double *f) {
        return __builtin_amdgcn_make_buffer_rsrc(f, (short)0, pbytes, kBufferDword3);
    };
    const __amdgpu_buffer_rsrc_t r_x = rsrc(P.x), r_y = rsrc(P.y), r_m = rsrc(P.m);
    const __amdgpu_buffer_rsrc_t r_v = rsrc(P.v), r_vx = rsrc(P.vx), r_vy = rsrc(P.vy);
    const __amdgpu_buffer_rsrc_t r_sx = rsrc(P.txx), r_sy = rsrc(P.txy), r_sz = rsrc(P.tyy);
    auto ld = [](__amdgpu_buffer_rsrc_t r, int off) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    };
    auto swz = [](int s) { return s ^ ((s >> 4) & 15); };

    // Staging slot u of a batch holds record s = tid + u * kP2GThreads of the
    // colour's list: entry k of the bin of cell q, where s_off[c][q] <= base
    // + s (the cell comes from the block-owner map).  The bin's next entry is
    // loaded too: bins are filled in lane order and are nearly always sorted
    // by particle index, and the pair shows the exceptions.  Slots past the
    // batch read a slot of an in-grid cell and are discarded, so no branch
    // separates these loads from the particle loads around them.
    struct Slots {
        int p[kStageSlots];     // bin entry (raw load)
        int pn[kStageSlots];    // the bin's next entry (raw load)
        int q[kStageSlots];     // cell, -1 for an idle slot
        int more;               // bit u: the bin has an entry after this one
    };
    auto fetch = [&](int c, int bi, Slots &R) {
        const int ncx = ncx_of(c);
        const int ca = (c & 1) ? 0 : 1, cb = (c >> 1) ? 0 : 1;
        const int base = s_off[c][s_bq[c][bi]];
        const int nb = s_off[c][s_bq[c][bi + 1]] - base;
        R.more = 0;
#pragma unroll
        for (int u = 0; u < kStageSlots; ++u) {
            const int s = tid + u * kP2GThreads;
            const bool valid = s < nb;
            const int gi = base + s;
            int q = s_bq[c][bi];   // idle slots: any cell (the address is clamped)
            if (valid) {
                q = s_blk[c][gi >> 4];
                while (s_off[c][q + 1] <= gi) {
                    ++q;
                }
            }
            const int k = valid ? gi - s_off[c][q] : 0;
            const bool more = valid && gi + 1 < s_off[c][q + 1];
            const int cc = min(max(cx0 + 2 * (q % ncx) + ca, 0), g.Ne - 1);
            const int rr = min(max(cy0 + 2 * (q / ncx) + cb, 0), g.Ne - 1);
            const int *bin = bkt_cur + (size_t)(rr * g.Ne + cc) * CAP;
            R.p[u] = bin[k];
            R.pn[u] = bin[more ? k + 1 : k];
            R.q[u] = valid ? q : -1;
            R.more |= (int)more << u;
        }
    };
    // Raises the wave's flag if a bin entry of the batch exceeds its
    // successor (waits for the batch's bin loads).
    auto flag_order = [&](const Slots &R) {
        bool inv = false;
#pragma unroll
        for (int u = 0; u < kStageSlots; ++u) {
            inv |= ((R.more >> u) & 1) && R.pn[u] < R.p[u];
        }
        const bool winv = __ballot(inv) != 0;
        if ((tid & 63) == 0) {
            s_inv[tid >> 6] = winv;
        }
    };

    const int i = tx0 + (nt % kTileX);
    const int j = ty0 + (nt / kTileX);
    const bool live = i < g.N && j < g.N;
    // body force of this role's force chain
    const double bb = role == ROLES - 2 ? st->bx : st->by;
    // two roles: role 0 m (am), mx (ap), fx (af); role 1 my (ap), fy (af)
    // three roles: role 0 m (am), mx (ap), my (ay); roles 1, 2 fx, fy (af)
    double am = 0.0, ap = 0.0, ay = 0.0, af = 0.0;

    if (any_particles) {
    Slots Rn;
    fetch(0, 0, Rn);
    flag_order(Rn);
    __syncthreads();
    // Colours in the reference's order; colour = 2 * (row % 2) + (col % 2).
    // Unrolled, so each colour's cell geometry is a compile-time constant.
    // Every colour is visited once the tile holds particles (an empty one
    // is a single batch of idle slots).
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int pc = c & 1, pr = c >> 1;
        const int ncx = ncx_of(c);
        const int ca = pc ? 0 : 1, cb = pr ? 0 : 1;

        // This node's cell of colour c, as an index into the colour's cells.
        int myq = -1, dx = 0, dy = 0;
        if (live) {
            const int cc = i - ((i ^ pc) & 1);
            const int rr = j - ((j ^ pr) & 1);
            if (cc >= 0 && rr >= 0 && cc < g.Ne && rr < g.Ne) {
                myq = ((rr - cy0) >> 1) * ncx + ((cc - cx0) >> 1);
                dx = i - cc;
                dy = j - rr;
            }
        }
        // tent(xl) = dx ? xl : 1 - xl as one exact fma (the product with +-1
        // is exact; adding -0 leaves any value unchanged), and grad_tent's
        // sign folded into the 1 / h factor
        const double ax = dx ? 1.0 : -1.0, cx = dx ? -0.0 : 1.0;
        const double ay_ = dy ? 1.0 : -1.0, cy = dy ? -0.0 : 1.0;
        const double hx = dx ? g.inv_h : -g.inv_h, hy = dy ? g.inv_h : -g.inv_h;

        const int nbat = s_nbat[c];
        for (int bi = 0; bi < nbat; ++bi) {
            const Slots Rc = Rn;
            const int qb = s_bq[c][bi], qe = s_bq[c][bi + 1];
            const int base = s_off[c][qb];
            // staged position of each slot's record: its own slot when the
            // bins are in order, else its cell's offset + its rank by
            // particle index within the bin (ranked through LDS)
            int dst[kStageSlots];
#pragma unroll
            for (int u = 0; u < kStageSlots; ++u) {
                dst[u] = tid + u * kP2GThreads;
            }
            int any_inv = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                any_inv |= s_inv[w];
            }
            if (__builtin_amdgcn_readfirstlane(any_inv)) {
#pragma unroll
                for (int u = 0; u < kStageSlots; ++u) {
                    if (Rc.q[u] >= 0) {
                        s_cmp[dst[u]] = Rc.p[u];
                    }
                }
                __syncthreads();
#pragma unroll
                for (int u = 0; u < kStageSlots; ++u) {
                    if (Rc.q[u] >= 0) {
                        const int b0 = s_off[c][Rc.q[u]] - base;
                        const int b1 = s_off[c][Rc.q[u] + 1] - base;
                        int rank = 0;
                        for (int t = b0; t < b1; ++t) {
                            rank += s_cmp[t] < Rc.p[u];
                        }
                        dst[u] = b0 + rank;
                    }
                }
            }
            // the batch's particle loads, then the bins of the next batch
            int off[kStageSlots];
#pragma unroll
            for (int u = 0; u < kStageSlots; ++u) {
                off[u] = Rc.q[u] >= 0 ? Rc.p[u] * (int)sizeof(double) : -(int)sizeof(double);
            }
            double px[kStageSlots], py[kStageSlots], pm[kStageSlots], pv[kStageSlots];
            double pvx[kStageSlots], pvy[kStageSlots];
            double psx[kStageSlots], psy[kStageSlots], psz[kStageSlots];
#pragma unroll
            for (int u = 0; u < kStageSlots; ++u) {
                px[u] = ld(r_x, off[u]);
                py[u] = ld(r_y, off[u]);
                pm[u] = ld(r_m, off[u]);
                pv[u] = ld(r_v, off[u]);
                pvx[u] = ld(r_vx, off[u]);
                pvy[u] = ld(r_vy, off[u]);
                psx[u] = ld(r_sx, off[u]);
                psy[u] = ld(r_sy, off[u]);
                psz[u] = ld(r_sz, off[u]);
            }
            const bool last = bi + 1 == nbat;
            if (!last) {
                fetch(c, bi + 1, Rn);
            } else if (c < 3) {
                fetch(c + 1, 0, Rn);
            }
#pragma unroll
            for (int u = 0; u < kStageSlots; ++u) {
                if (Rc.q[u] >= 0) {
                    const int lo = Rc.q[u];
                    const int cc = cx0 + 2 * (lo % ncx) + ca, rr = cy0 + 2 * (lo / ncx) + cb;
                    const double mp = pm[u], vp = pv[u];
                    const int sl = swz(dst[u]);
                    s_xy[sl] = make_double2(snap_local(div_h(px[u] - (double)cc * g.h, g)),
                                            snap_local(div_h(py[u] - (double)rr * g.h, g)));
                    // pdata = {m, x_t m, y_t m, ..}, stressdata = {-sxx v, -sxy v, -syy v}
                    s_msy[sl] = make_double2(mp, -psy[u] * vp);
                    s_pxs[sl] = make_double2(pvx[u] * mp, -psx[u] * vp);
                    s_pys[sl] = make_double2(pvy[u] * mp, -psz[u] * vp);
                }
            }
            __syncthreads();
            if (myq >= qb && myq < qe) {
                const int k0 = s_off[c][myq] - base;
                const int k1 = s_off[c][myq + 1] - base;
                // with no x body force the fx chain drops sf * (0 * m): af
                // starts at +0 and is only ever a sum, so it is never -0,
                // and adding +-0 (a finite mass) leaves it unchanged.  One
                // loop per role (the role is uniform per wave), so none
                // carries another's selects.
                const bool fx_role = ROLES == 2 ? role == 0 : role == 1;
                const bool bf = !(fx_role && g.m_finite && bb == 0.0);
                auto fold_fx = [&](auto body_force, auto with_m) {
#pragma unroll 2
                    for (int s0 = k0; s0 < k1; ++s0) {
                        const int s = swz(s0);
                        const double2 xy = s_xy[s], msy = s_msy[s], ps = s_pxs[s];
                        const double X = fma(ax, xy.x, cx);
                        const double Y = fma(ay_, xy.y, cy);
                        const double sf = X * Y;
                        const double mp = msy.x;
                        if (decltype(with_m)::value) {
                            am += sf * mp;
                            ap += sf * ps.x;
                        }
                        if (!MUSL2) {
                            if (decltype(body_force)::value) {
                                af += sf * (bb * mp);
                            }
                            af += (Y * hx) * ps.y;     // -sxx v
                            af += (X * hy) * msy.y;    // -sxy v
                        }
                    }
                };
                auto fold_fy = [&](auto with_my) {
#pragma unroll 2
                    for (int s0 = k0; s0 < k1; ++s0) {
                        const int s = swz(s0);
                        const double2 xy = s_xy[s], msy = s_msy[s], ps = s_pys[s];
                        const double X = fma(ax, xy.x, cx);
                        const double Y = fma(ay_, xy.y, cy);
                        const double sf = X * Y;
                        if (decltype(with_my)::value) {
                            ap += sf * ps.x;
                        }
                        if (!MUSL2) {
                            af += sf * (bb * msy.x);
                            af += (Y * hx) * msy.y;    // -sxy v
                            af += (X * hy) * ps.y;     // -syy v
                        }
                    }
                };
                if (ROLES == 2) {
                    if (role == 0) {
                        if (bf) {
                            fold_fx(std::true_type{}, std::true_type{});
                        } else {
                            fold_fx(std::false_type{}, std::true_type{});
                        }
                    } else {
                        fold_fy(std::true_type{});
                    }
                } else if (role == 0) {
#pragma unroll 2
                    for (int s0 = k0; s0 < k1; ++s0) {
                        const int s = swz(s0);
                        const double2 xy = s_xy[s];
                        const double mp = s_msy[s].x, pxm = s_pxs[s].x, pym = s_pys[s].x;
                        const double sf = fma(ax, xy.x, cx) * fma(ay_, xy.y, cy);
                        am += sf * mp;
                        ap += sf * pxm;
                        ay += sf * pym;
                    }
                } else if (!MUSL2) {
                    if (role == 1) {
                        if (bf) {
                            fold_fx(std::true_type{}, std::false_type{});
                        } else {
                            fold_fx(std::false_type{}, std::false_type{});
                        }
                    } else {
                        fold_fy(std::false_type{});
                    }
                }
            }
            // the next batch's order flags (its bin loads had the fold to land)
            if (!last || c < 3) {
                flag_order(Rn);
            }
            __syncthreads();
        }
    }
    }

    // Hand the other roles' chains to the role-0 thread of the node
    // (staging space is free after the last barrier).
    double2 *s_y = s_xy;
    if (role != 0) {
        s_y[(role - 1) * kTileNodes + nt] = make_double2(ap, af);
    }
    __syncthreads();
    if (role != 0 || !live) {
        return;
    }
    double amx = ap, amy, afx, afy;
    if (ROLES == 2) {
        const double2 yv = s_y[nt];
        afx = af;
        amy = yv.x;
        afy = yv.y;
    } else {
        amy = ay;
        afx = s_y[nt].y;
        afy = s_y[kTileNodes + nt].y;
    }
    if (MUSL2) {
        NodeRec r;
        r.m = am;
        r.mx = amx;
        r.my = amy;
        r.fx = 0.0;
        r.fy = 0.0;
        r.vx = am > kTol ? amx / am : 0.0;
        r.vy = am > kTol ? amy / am : 0.0;
        r.pad = 0.0;
        nodes[(size_t)j * g.N + i] = r;
        return;
    }

    // bc_momentum then bc_force, then move_grid_split.
    const unsigned f = bc_fixed(g, i, j, st->t + g.dt, mask);
    if (f & 1u) amx = 0.0;
    if (f & 2u) amy = 0.0;
    if (f & 1u) afx = 0.0;
    if (f & 2u) afy = 0.0;
    double nvx = 0.0, nvy = 0.0;
    if (am > kTol) {
        amx += g.dt * afx;
        amy += g.dt * afy;
        nvx = amx / am;
        nvy = amy / am;
    }
    NodeRec r;
    r.m = am;
    r.mx = amx;
    r.my = amy;
    r.fx = afx;
    r.fy = afy;
    r.vx = nvx;
    r.vy = nvy;
    r.pad = 0.0;
    nodes[(size_t)j * g.N + i] = r;
}

// k_p2g_grid: p2g_tile_body for one tile per workgroup: a single job's from
// p2g_tile, an ensemble's from the host-built dispatch table (BatchTables).
template <int CAP, bool BATCH, bool MUSL2, int ROLES>
__global__ __launch_bounds__(ROLES * kTileNodes) void k_p2g_grid(
    const Particles *__restrict__ Pd, const int *__restrict__ cnt_cur, const int *__restrict__ bkt_cur,
    int *__restrict__ cnt_next, NodeRec *__restrict__ nodes, uint8_t *__restrict__ tile_zero,
    const uint8_t *__restrict__ mask, GridParams g, const DevState *__restrict__ st,
    int tiles_x, int ntiles, int np, TileOrder ord, BatchTables bt)
{
    int tile, mb = 0;
    if constexpr (BATCH) {
        const int2 tm = bt.tile_order[blockIdx.x];
        tile = tm.x;
        mb = tm.y;
    } else {
        tile = p2g_tile(blockIdx.x, tiles_x, ntiles, ord);
    }
    if (tile < 0 || tile >= ntiles) {
        return;
    }
    p2g_tile_body<CAP, BATCH, MUSL2, ROLES>(Pd, cnt_cur, bkt_cur, cnt_next, nodes, tile_zero, mask,
                                            g, st, tiles_x, np, bt, tile, mb);
}

// The ensemble's light tiles (regions without particles at upload),
// kLightGroup per workgroup from BatchTables::light_tiles: the group's bin
// counts are checked in one pass and its own cells' next-step counts
// cleared; a tile that now holds particles, or whose node records are not
// known to be zero, is appended to fix_list for k_p2g_fix.
__global__ __launch_bounds__(256) void k_p2g_light(
    const int *__restrict__ cnt_cur, int *__restrict__ cnt_next,
    const uint8_t *__restrict__ tile_zero, int n_light, BatchTables bt, int *__restrict__ fix_count,
    int2 *__restrict__ fix_list)
{
    __shared__ int4 s_geo[kLightGroup];   // region origin, grid side, cell base
    __shared__ int s_need[kLightGroup];   // the tile needs p2g_tile_body
    __shared__ int2 s_tm[kLightGroup];
    const int e0 = blockIdx.x * kLightGroup;
    const int n = min(kLightGroup, n_light - e0);
    if (n <= 0) {
        return;
    }
    const int tid = threadIdx.x;
    if (tid < n) {
        const int2 tm = bt.light_tiles[e0 + tid];
        const MemberInfo &mi = bt.mem[tm.y];
        const int lt = tm.x - mi.tile_base;
        s_geo[tid] = make_int4((lt % mi.tiles_x) * kTileX, (lt / mi.tiles_x) * kTileY, mi.g.Ne,
                               mi.cell_base);
        s_need[tid] = !tile_zero[tm.x];
        s_tm[tid] = tm;
    }
    __syncthreads();
    for (int it = tid; it < n * kRegCells; it += 256) {
        const int u = it / kRegCells, cl = it % kRegCells;
        const int4 geo = s_geo[u];
        const int ax = cl % kRegX, ay = cl / kRegX;
        const int cc = geo.x - 1 + ax, rr = geo.y - 1 + ay;
        if (cc >= 0 && rr >= 0 && cc < geo.z && rr < geo.z) {
            const int c = geo.w + rr * geo.z + cc;
            if (cnt_cur[c] > 0) {
                s_need[u] = 1;
            }
            if (ax > 0 && ay > 0) {   // the tile's own cell
                cnt_next[c] = 0;
            }
        }
    }
    __syncthreads();
    if (tid < n && s_need[tid]) {
        fix_list[atomicAdd(fix_count, 1)] = s_tm[tid];
    }
}

// p2g_tile_body for the tiles k_p2g_light listed (fix_count of them, in
// list order over a fixed grid; usually none), and the other step parity's
// list emptied for the next step.  (A kernel of its own: the loop around
// the body needs more registers than k_p2g_grid can spare.)
template <int CAP>
__global__ __launch_bounds__(2 * kTileNodes) void k_p2g_fix(
    const Particles *__restrict__ Pd, const int *__restrict__ cnt_cur, const int *__restrict__ bkt_cur,
    int *__restrict__ cnt_next, NodeRec *__restrict__ nodes, uint8_t *__restrict__ tile_zero,
    int np, const int *__restrict__ fix_count, const int2 *__restrict__ fix_list,
    int *__restrict__ next_count, BatchTables bt)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *next_count = 0;
    }
    const int n = *fix_count;
    const GridParams g0{};
    for (int e = blockIdx.x; e < n; e += gridDim.x) {
        const int2 tm = fix_list[e];
        p2g_tile_body<CAP, true, false, 2>(Pd, cnt_cur, bkt_cur, cnt_next, nodes, tile_zero, nullptr,
                                           g0, nullptr, 0, np, bt, tm.x, tm.y);
        __syncthreads();   // LDS is reused by the next tile
    }
}

// FIXED >= 0: the mode is that compile-time constant (the plain USL step's
// instantiation carries no other mode's code or registers); -1: `mode`.
template <bool BATCH, int FIXED>
__global__ __launch_bounds__(kG2PThreads, FIXED >= 0 ? 5 : 1) void k_g2p_stress(
    const Particles *__restrict__ Pd, const NodeRec *__restrict__ nodes, int *__restrict__ cnt_next,
    int *__restrict__ bkt_next, GridParams g, Material mt, DevState *__restrict__ st,
    int np, int mode, BatchTables bt)
{
    const Particles &P = *Pd;   // field pointers, loaded as used
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (FIXED >= 0) {
        mode = FIXED;
    }
    int first = 0;
    if constexpr (BATCH) {
        const int mb = bt.block_member[blockIdx.x];
        const MemberInfo &mi = bt.mem[mb];
        g = mi.g;
        mt = mi.mt;
        first = mi.part_base;
        np = mi.part_base + mi.np;
        nodes += mi.node_base;
        cnt_next += mi.cell_base;
        bkt_next += (size_t)mi.cell_base * g.cap;
        st = bt.st + mb;
    }
    // MUSL runs this kernel twice per step: the particle update (move), then,
    // after the second mapping, strain rate, volume, stress and rebinning
    // (finish); USL does both in one pass.
    const bool move = !(mode & kModeMuslFinish);
    const bool finish = !(mode & kModeMuslMove);
    if (p == first && finish) {
        // t += dt (process_usl.c:522), then time_varying_loads for the next
        // step (src/main.c:975).  No thread of this kernel reads *st.
        const double t = st->t + g.dt;
        st->t = t;
        st->bx = 0.0;
        st->by = gravity(t, st->t0, g.g_mag, g.ramp);
    }
    if (p >= np) {
        return;
    }
    // Loads in three waves, each issued before any of it is used: the
    // particle's kinematics and its nodes' masses, momenta and forces (the
    // move); the nodes' velocities (the strain rate); the mass, volume,
    // stress and state (the stress update).  Each group is dead before the
    // next is live, which keeps the plain step at 96 VGPRs (5 waves per
    // SIMD) without spills.
    const int act = P.active[p];
    const int e = P.elem[p];
    double x = P.x[p], y = P.y[p];
    const double vx0 = P.vx[p], vy0 = P.vy[p];
    const double ux0 = P.ux[p], uy0 = P.uy[p];
    const bool stress_in = !(mode & kModeHostMaterial) || (mode & kModeNonlocal);
    const bool state_in = !(mode & (kModeHostMaterial | kModeNonlocal | kModeIsolin));
    double xn0 = 0.0, yn0 = 0.0;
    if (!move) {
        xn0 = P.xn[p];
        yn0 = P.yn[p];
    }
    if (!act) {
        return;
    }
    if (e < 0) {
        // left the grid during the previous step: deactivated at the start
        // of this one (process_usl.c:660-665)
        P.active[p] = 0;
        return;
    }
    const int cc = e % g.Ne, rr = e / g.Ne;
    const double dt = g.dt;
    // shape_functions, the values now and the gradients after the move
    // (the same operations, fewer registers live)
    Shape sh;
    {
        const double xl = snap_local(div_h(x - (double)cc * g.h, g));
        const double yl = snap_local(div_h(y - (double)rr * g.h, g));
        const double ox = 1.0 - xl, oy = 1.0 - yl;
        sh.s[0] = ox * oy;
        sh.s[1] = xl * oy;
        sh.s[2] = xl * yl;
        sh.s[3] = ox * yl;
        sh.xl = xl;
        sh.yl = yl;
    }

    const size_t n0 = (size_t)rr * g.N + cc;
    const size_t nid[4] = {n0, n0 + 1, n0 + g.N + 1, n0 + g.N};
    struct NodeMove { double m, mx, my, fx, fy; };
    NodeMove nd[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const NodeRec *r = nodes + nid[q];
        nd[q].m = r->m;
        nd[q].mx = r->mx;
        nd[q].my = r->my;
        nd[q].fx = r->fx;
        nd[q].fy = r->fy;
    }

    // move_particles_explicit_usl_split
    double ax = 0.0, ay = 0.0;
    double vx = vx0, vy = vy0;
    double ux, uy;
    if (move) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (nd[q].m < kTol) continue;
            ax += ((sh.s[q] * nd[q].fx) / nd[q].m);
            ay += ((sh.s[q] * nd[q].fy) / nd[q].m);
        }
        vx += dt * ax;
        vy += dt * ay;
        double dux = 0.0, duy = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (nd[q].m < kTol) continue;
            dux += ((sh.s[q] * nd[q].mx) / nd[q].m);
            duy += ((sh.s[q] * nd[q].my) / nd[q].m);
        }
        dux *= dt;
        duy *= dt;
        x += dux;
        y += duy;
        ux = ux0 + dux;
        uy = uy0 + duy;
        if (!finish) {
            // MUSL: the second mapping needs the new velocities at the
            // step's start positions; the new positions wait in xn, yn
            P.vx[p] = vx;
            P.vy[p] = vy;
            P.ux[p] = ux;
            P.uy[p] = uy;
            P.xn[p] = x;
            P.yn[p] = y;
            if (mode & kModeStoreAux) {
                P.ax[p] = ax;
                P.ay[p] = ay;
                P.xl[p] = sh.xl;
                P.yl[p] = sh.yl;
            }
            return;
        }
    } else {
        x = xn0;
        y = yn0;
        ux = ux0;
        uy = uy0;
    }

    // The new kinematics, and the next step's create_particle_to_element_map
    // and bin: none is needed by the strain rate and stress below.
    P.x[p] = x;
    P.y[p] = y;
    P.vx[p] = vx;
    P.vy[p] = vy;
    P.ux[p] = ux;
    P.uy[p] = uy;
    if ((mode & kModeStoreAux) && move) {
        P.ax[p] = ax;
        P.ay[p] = ay;
        P.xl[p] = sh.xl;
        P.yl[p] = sh.yl;
    }
    const int en = which_element(x, y, g);
    P.elem[p] = en;
    const unsigned long long crossed = __ballot(en != e);
    if (crossed && __lane_id() == __builtin_ctzll(__ballot(1))) {
        atomicAdd(&st->crossings, (unsigned long long)__popcll(crossed));
    }
    bin_insert(en >= 0, en, p, cnt_next, bkt_next, g.cap, &st->status);

    {
        const double xl = sh.xl, yl = sh.yl;
        const double ox = 1.0 - xl, oy = 1.0 - yl;
        sh.b1[0] = -oy * g.inv_h;
        sh.b1[1] = oy * g.inv_h;
        sh.b1[2] = yl * g.inv_h;
        sh.b1[3] = -yl * g.inv_h;
        sh.b2[0] = -ox * g.inv_h;
        sh.b2[1] = -xl * g.inv_h;
        sh.b2[2] = xl * g.inv_h;
        sh.b2[3] = ox * g.inv_h;
    }
    // the node velocities, loaded after the move (fewer registers live)
    if constexpr (FIXED >= 0) {
        __builtin_amdgcn_sched_barrier(0);
    }
    struct NodeVel { double vx, vy; };
    NodeVel nv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        nv[q].vx = nodes[nid[q]].vx;
        nv[q].vy = nodes[nid[q]].vy;
    }
    // calculate_strainrate_split (SMEAR4 folds left to right)
    const double exx = 1.0 * (sh.b1[0] * nv[0].vx + sh.b1[1] * nv[1].vx
                              + sh.b1[2] * nv[2].vx + sh.b1[3] * nv[3].vx);
    const double eyy = 1.0 * (sh.b2[0] * nv[0].vy + sh.b2[1] * nv[1].vy
                              + sh.b2[2] * nv[2].vy + sh.b2[3] * nv[3].vy);
    const double dx_tdy = 1.0 * (sh.b2[0] * nv[0].vx + sh.b2[1] * nv[1].vx
                                 + sh.b2[2] * nv[2].vx + sh.b2[3] * nv[3].vx);
    const double dy_tdx = 1.0 * (sh.b1[0] * nv[0].vy + sh.b1[1] * nv[1].vy
                                 + sh.b1[2] * nv[2].vy + sh.b1[3] * nv[3].vy);
    const double exy = 0.5 * (dx_tdy + dy_tdx);
    const double wxy = 0.5 * (dx_tdy - dy_tdx);
    if constexpr (FIXED >= 0) {
        __builtin_amdgcn_sched_barrier(0);
    }
    const double mp = P.m[p], v0 = P.v[p];
    double sxx0 = 0.0, sxy0 = 0.0, syy0 = 0.0, szz0 = 0.0, gp0 = 0.0;
    if (finish && stress_in) {
        sxx0 = P.txx[p];
        sxy0 = P.txy[p];
        syy0 = P.tyy[p];
    }
    if (finish && state_in) {
        szz0 = P.tzz[p];
        gp0 = P.gp[p];
    }

    // update_particle_densities_split
    const double v = v0 * mpm_exp(dt * (exx + eyy));

    if (mode & kModeNonlocal) {
        // the stress update waits for the fluidity solve; keep what it needs
        double gfb, xsi;
        nonlocal_bulk(sxx0, sxy0, syy0, mt, gfb, xsi);
        P.gf_bulk[p] = gfb;
        P.xisq_inv[p] = xsi;
        P.elem_prev[p] = e;
    } else if (mode & kModeIsolin) {
        // isolin calculate_stress_threaded (materialsrc/isolin.c:89-110) on
        // sxx, sxy, syy (held in the Txx, Txy, Tyy arrays)
        const double sxx = sxx0, sxy = sxy0, syy = syy0;
        const double trD = exx + eyy;
        double dsjxx = mt.lambda * trD + mt.two_G * exx;
        double dsjxy = mt.two_G * exy;
        double dsjyy = mt.lambda * trD + mt.two_G * eyy;
        dsjxx += 2 * wxy * sxy;
        dsjxy -= wxy * (sxx - syy);
        dsjyy -= 2 * wxy * sxy;
        P.txx[p] = sxx + dt * dsjxx;
        P.txy[p] = sxy + dt * dsjxy;
        P.tyy[p] = syy + dt * dsjyy;
    } else if (mode & kModeDpRiMt) {
        // dp_ri_mt calculate_stress_threaded (materialsrc/dp_ri_mt.c:133-205)
        // on sxx, sxy, syy (Txx, Txy, Tyy arrays) and szz = state[1] (Tzz)
        const double sxx = sxx0, sxy = sxy0, syy = syy0, szz = szz0;
        const double trD = exx + eyy;
        double dsjxx = mt.lambda * trD + mt.two_G * exx;
        double dsjxy = mt.two_G * exy;
        double dsjyy = mt.lambda * trD + mt.two_G * eyy;
        dsjxx += 2 * wxy * sxy;
        dsjxy -= wxy * (sxx - syy);
        dsjyy -= 2 * wxy * sxy;
        const double sxx_tr = sxx + dt * dsjxx;
        const double sxy_tr = sxy + dt * dsjxy;
        const double syy_tr = syy + dt * dsjyy;
        const double szz_tr = szz + dt * mt.lambda * trD;
        const double p_tr = -(sxx_tr + syy_tr + szz_tr) / 3.0;
        const double t0xx = sxx_tr + p_tr, t0xy = sxy_tr, t0yy = syy_tr + p_tr;
        const double t0zz = szz_tr + p_tr;
        const double tau_tr = sqrt(0.5 * (t0xx * t0xx + 2 * t0xy * t0xy + t0yy * t0yy
                                          + t0zz * t0zz));
        double nup_tau;
        if ((mp / v) < mt.rho_crit || p_tr <= 0.0) {
            nup_tau = tau_tr / (mt.G * dt);
            P.txx[p] = 0.0;
            P.txy[p] = 0.0;
            P.tyy[p] = 0.0;
            P.tzz[p] = 0.0;
        } else if (p_tr > 0.0) {
            const double S0 = mt.mu_s * p_tr;
            double tau_tau, scale;
            if (tau_tr <= S0) {
                tau_tau = tau_tr;
                scale = 1.0;
            } else {
                tau_tau = S0;
                scale = tau_tau / tau_tr;
            }
            nup_tau = ((tau_tr - tau_tau) / mt.G) / dt;
            P.txx[p] = scale * t0xx - p_tr;
            P.txy[p] = scale * t0xy;
            P.tyy[p] = scale * t0yy - p_tr;
            P.tzz[p] = scale * t0zz - p_tr;
        } else {
            nup_tau = 0.0;   // NaN pressure: stress left unchanged
        }
        P.gp[p] = gp0 + nup_tau * dt;
        P.gdp[p] = nup_tau;
    } else if (!(mode & kModeHostMaterial)) {
        double txx = sxx0, txy = sxy0, tyy = syy0, tzz = szz0;
        double gp = gp0, gdp;
        g_local_mu2_update(exx, exy, eyy, wxy, mp, v, dt, mt, txx, txy, tyy, tzz, gp, gdp);
        P.txx[p] = txx;
        P.txy[p] = txy;
        P.tyy[p] = tyy;
        P.tzz[p] = tzz;
        P.gp[p] = gp;
        P.gdp[p] = gdp;
    }
    P.v[p] = v;
    if (mode & kModeStoreAux) {
        P.exx[p] = exx;
        P.exy[p] = exy;
        P.eyy[p] = eyy;
        P.wxy[p] = wxy;
    }
}

// ---------------------------------------------------------------------------
// g_nonlocal: fluidity system, solve and stress update
// ---------------------------------------------------------------------------

struct NlSystem {
    double *K;          // 9-point stencil, [9][N*N], slot (oy+1)*3 + (ox+1)
    double *f, *x, *r, *p, *dinv;
    double *u, *w, *s;  // [2][N*N] each: previous and next iterate
    double *partials;   // [2][4][kNlMaxBlocks]
    unsigned int *barrier;   // grid barrier counter of the solve
    int *stats;         // iterations
    double *res;        // final relative residual
    int N;
};

constexpr int kNlThreads = 256;
constexpr int kCgTile = 32;                     // CG tiles: 32 x 32 nodes
constexpr int kCgThreads = kCgTile * kCgTile;   // one thread per tile node
constexpr int kNlMaxBlocks = 2048;

// Assembly of K g = f (solve_diffusion_part, g_nonlocal.c:356-592) at the
// nodes: K_ij = sum_p v_p (xi_p^-2 N_i N_j + grad N_i . grad N_j),
// f_i = sum_p v_p g_loc,p xi_p^-2 N_i, over the particles of the step just
// taken (their step-start cells and local coordinates).  Unknowns are the
// nodes with m > TOL; other nodes get the row of the identity.  One thread
// per node, 16x16 nodes per workgroup; the cells' particle lists are staged
// and sorted in LDS and folded in colour order, ascending particle index,
// so the sums are deterministic.
template <int CAP>
__global__ __launch_bounds__(kNlThreads) void k_nl_assemble(
    const Particles *__restrict__ Pd, const int *__restrict__ cnt_cur, const int *__restrict__ bkt_cur,
    const NodeRec *__restrict__ nodes, NlSystem S, GridParams g, int tiles_x, int ntiles)
{
    const Particles &P = *Pd;   // field pointers, loaded as used
    constexpr int T = 16, R = T + 1;
    __shared__ int s_cnt[R * R];
    __shared__ int s_lst[R * R * CAP];
    const int tile = blockIdx.x;
    if (tile >= ntiles) {
        return;
    }
    if (tile == 0 && threadIdx.x == 0) {
        *S.barrier = 0u;   // the solve that follows starts a fresh barrier count
    }
    const int tx0 = (tile % tiles_x) * T, ty0 = (tile / tiles_x) * T;
    const int cx0 = tx0 - 1, cy0 = ty0 - 1;
    const int tid = threadIdx.x;
    for (int cl = tid; cl < R * R; cl += kNlThreads) {
        const int cc = cx0 + cl % R, rr = cy0 + cl / R;
        int n = 0;
        if (cc >= 0 && rr >= 0 && cc < g.Ne && rr < g.Ne) {
            n = cnt_cur[rr * g.Ne + cc];
            n = n < CAP ? n : CAP;
        }
        s_cnt[cl] = n;
    }
    __syncthreads();
    for (int idx = tid; idx < R * R * CAP; idx += kNlThreads) {
        const int cl = idx / CAP, k = idx % CAP;
        if (k < s_cnt[cl]) {
            const int cc = cx0 + cl % R, rr = cy0 + cl / R;
            s_lst[idx] = bkt_cur[(size_t)(rr * g.Ne + cc) * CAP + k];
        }
    }
    __syncthreads();
    for (int cl = tid; cl < R * R; cl += kNlThreads) {
        int *a = &s_lst[cl * CAP];
        const int n = s_cnt[cl];
        for (int k = 1; k < n; ++k) {
            const int key = a[k];
            int q = k - 1;
            while (q >= 0 && a[q] > key) {
                a[q + 1] = a[q];
                --q;
            }
            a[q + 1] = key;
        }
    }
    __syncthreads();

    const int i = tx0 + tid % T, j = ty0 + tid / T;
    if (i >= g.N || j >= g.N) {
        return;
    }
    const size_t nn = (size_t)g.N * g.N;
    const size_t me = (size_t)j * g.N + i;
    auto is_dof = [&](int ii, int jj) { return nodes[(size_t)jj * g.N + ii].m > kTol; };
    double Kc[9];
#pragma unroll
    for (int o = 0; o < 9; ++o) {
        Kc[o] = 0.0;
    }
    double fi = 0.0;
    const bool dof = is_dof(i, j);
    if (dof) {
        for (int c = 0; c < 4; ++c) {
            const int cc = i - ((i ^ c) & 1);
            const int rr = j - ((j ^ (c >> 1)) & 1);
            if (cc < 0 || rr < 0 || cc >= g.Ne || rr >= g.Ne) {
                continue;
            }
            const int dx = i - cc, dy = j - rr;
            const int cl = (cc - cx0) + (rr - cy0) * R;
            const int n = s_cnt[cl];
            // corners b = BL, BR, TR, TL of this cell; skip non-unknowns
            const bool db[4] = {is_dof(cc, rr), is_dof(cc + 1, rr), is_dof(cc + 1, rr + 1),
                                is_dof(cc, rr + 1)};
            double kb[4] = {0.0, 0.0, 0.0, 0.0};
            double fc = 0.0;
            for (int k = 0; k < n; ++k) {
                const int p = s_lst[cl * CAP + k];
                const double xl = P.xl[p], yl = P.yl[p];
                const double vp = P.v[p], xsi = P.xisq_inv[p], gfb = P.gf_bulk[p];
                const double ox = 1.0 - xl, oy = 1.0 - yl;
                const double sb[4] = {ox * oy, xl * oy, xl * yl, ox * yl};
                const double gxb[4] = {-oy * g.inv_h, oy * g.inv_h, yl * g.inv_h, -yl * g.inv_h};
                const double gyb[4] = {-ox * g.inv_h, -xl * g.inv_h, xl * g.inv_h, ox * g.inv_h};
                const int a = dy ? (dx ? 2 : 3) : (dx ? 1 : 0);
                const double sa = sb[a], gxa = gxb[a], gya = gyb[a];
                fc += ((vp * gfb) * xsi) * sa;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    if (db[b]) {
                        kb[b] += vp * (xsi * sa * sb[b] + (gxa * gxb[b] + gya * gyb[b]));
                    }
                }
            }
            fi += fc;
            // corner b sits at offset (bx - dx, by - dy) from this node
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int bx = (b == 1 || b == 2) ? 1 : 0, by = (b >= 2) ? 1 : 0;
                const int o = (by - dy + 1) * 3 + (bx - dx + 1);
                Kc[o] += kb[b];
            }
        }
    } else {
        Kc[4] = 1.0;
    }
#pragma unroll
    for (int o = 0; o < 9; ++o) {
        S.K[(size_t)o * nn + me] = Kc[o];
    }
    S.f[me] = fi;
    S.dinv[me] = 1.0 / Kc[4];
    if (!dof) {
        S.x[me] = 0.0;
    }
}

// Sum over a workgroup in a fixed order (per wave: shuffle tree; across
// waves: LDS, wave 0 adds in order).
__device__ __forceinline__ double block_sum(double v, double *s_red)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        v += __shfl_down(v, d);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        s_red[w] = v;
    }
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
        t += s_red[k];
    }
    return t;   // valid in every thread
}

__device__ __forceinline__ double block_max(double v, double *s_red)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        v = fmax(v, __shfl_down(v, d));
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        s_red[w] = v;
    }
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
        t = fmax(t, s_red[k]);
    }
    return t;   // valid in every thread
}

// Sum of the per-workgroup partials, same order in every workgroup.
__device__ __forceinline__ double grid_total(const double *part, int nblocks, double *s_red)
{
    double v = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += blockDim.x) {
        v += part[k];
    }
    return block_sum(v, s_red);
}

// Three workgroup sums at once (s_red holds 3 * waves entries).
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *s_red)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_down(a, d);
        b += __shfl_down(b, d);
        c += __shfl_down(c, d);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) {
        s_red[w] = a;
        s_red[nw + w] = b;
        s_red[2 * nw + w] = c;
    }
    __syncthreads();
    a = b = c = 0.0;
    for (int k = 0; k < nw; ++k) {
        a += s_red[k];
        b += s_red[nw + k];
        c += s_red[2 * nw + k];
    }
}

// grid_total of three partial arrays spaced kNlMaxBlocks apart.
__device__ __forceinline__ void grid_total3(const double *part, int nblocks, double &a,
                                            double &b, double &c, double *s_red)
{
    a = b = c = 0.0;
    for (int k = threadIdx.x; k < nblocks; k += blockDim.x) {
        a += part[k];
        b += part[kNlMaxBlocks + k];
        c += part[2 * kNlMaxBlocks + k];
    }
    block_sum3(a, b, c, s_red);
}

// Grid-wide barrier for a cooperative launch (all workgroups co-resident):
// one agent-scope counter that only grows; the workgroup's barrier episode
// k waits for the count to reach k * gridDim.x.  One release fence before the
// arrival and one acquire fence after the wait; the polling itself is
// relaxed, so it does not invalidate the L2 on every probe.  The counter is
// cleared by the assembly kernel that precedes every solve.
__device__ __forceinline__ void grid_barrier(unsigned int *count, unsigned int &target)
{
    __syncthreads();
    target += gridDim.x;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Sum of the 9-point stencil row i of K against v.
__device__ __forceinline__ double stencil_dot(const double *__restrict__ K, const double *v,
                                              int i, int N, int nn)
{
    const int ix = i % N, iy = i / N;
    double acc = 0.0;
#pragma unroll
    for (int oy = -1; oy <= 1; ++oy) {
#pragma unroll
        for (int ox = -1; ox <= 1; ++ox) {
            const int jx = ix + ox, jy = iy + oy;
            if (jx >= 0 && jx < N && jy >= 0 && jy < N) {
                acc += K[(size_t)((oy + 1) * 3 + (ox + 1)) * nn + i] * v[jy * N + jx];
            }
        }
    }
    return acc;
}

// Jacobi-preconditioned conjugate gradients on the assembled stencil in the
// single-reduction form of Chronopoulos and Gear (J. Comput. Appl. Math. 25,
// 1989): with u = M^-1 r, w = K u, s = K p and q = M^-1 s, one iteration is
//   p = u + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s,
//   u -= alpha M^-1 s,  w = K u,
//   gamma' = (r,u), delta' = (w,u), beta = gamma'/gamma,
//   alpha = gamma' / (delta' - beta gamma' / alpha),
// so the three inner products (with (r,r) for the stopping test) share one
// grid barrier.  The product w = K u needs the new u of the neighbours; each
// thread recomputes it from their previous u, w, s (double-buffered), with
// the same expression the owner uses, so the values agree bit for bit.  One
// persistent cooperative launch; every workgroup reduces the same partials
// in the same order and reaches the same stopping decision.
__global__ __launch_bounds__(kCgThreads) void k_nl_cg(NlSystem S, double rel_tol, int max_iter)
{
    __shared__ double s_red[3 * (kCgThreads / 64)];
    const int N = S.N;
    const int nn = N * N;
    const int nb = gridDim.x;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = nb * blockDim.x;
    const double *__restrict__ K = S.K;
    const double *__restrict__ dinv = S.dinv;
    unsigned int target = 0;

    // r = f - K x, u = M^-1 r, p = s = 0
    double ff = 0.0;
    for (int i = gid; i < nn; i += stride) {
        const double r = S.f[i] - stencil_dot(K, S.x, i, N, nn);
        S.r[i] = r;
        S.u[i] = dinv[i] * r;
        S.p[i] = 0.0;
        S.s[i] = 0.0;
        ff += S.f[i] * S.f[i];
    }
    grid_barrier(S.barrier, target);
    // w = K u and the first inner products
    double g = 0.0, dl = 0.0, rr = 0.0;
    for (int i = gid; i < nn; i += stride) {
        const double w = stencil_dot(K, S.u, i, N, nn);
        S.w[i] = w;
        const double r = S.r[i], u = S.u[i];
        g += r * u;
        dl += w * u;
        rr += r * r;
    }
    double *part = S.partials;   // [2][4][kNlMaxBlocks], by iteration parity
    {
        const double a = block_sum(g, s_red), b = block_sum(dl, s_red);
        const double c = block_sum(rr, s_red), d = block_sum(ff, s_red);
        if (threadIdx.x == 0) {
            part[blockIdx.x] = a;
            part[kNlMaxBlocks + blockIdx.x] = b;
            part[2 * kNlMaxBlocks + blockIdx.x] = c;
            part[3 * kNlMaxBlocks + blockIdx.x] = d;
        }
    }
    grid_barrier(S.barrier, target);
    double gamma = grid_total(part, nb, s_red);
    double delta = grid_total(part + kNlMaxBlocks, nb, s_red);
    rr = grid_total(part + 2 * kNlMaxBlocks, nb, s_red);
    ff = grid_total(part + 3 * kNlMaxBlocks, nb, s_red);
    const double stop = rel_tol * rel_tol * ff;
    double alpha = gamma / delta, beta = 0.0;
    int it = 0;
    // Iterations work on 32 x 32-node tiles: each workgroup forms the new u
    // of its tile and of the one-node halo around it in LDS, then takes the
    // stencil product from there.
    constexpr int E = kCgTile + 2;
    __shared__ double s_u[E * E];
    const int tiles_x = (N + kCgTile - 1) / kCgTile;
    const int ntiles = tiles_x * tiles_x;
    const int lx = threadIdx.x % kCgTile, ly = threadIdx.x / kCgTile;
    int hx = -2, hy = -2;   // this thread's halo node, if any
    {
        const int t = threadIdx.x;
        if (t < E) {
            hx = t - 1;
            hy = -1;
        } else if (t < 2 * E) {
            hx = t - E - 1;
            hy = kCgTile;
        } else if (t < 2 * E + kCgTile) {
            hx = -1;
            hy = t - 2 * E;
        } else if (t < 2 * E + 2 * kCgTile) {
            hx = kCgTile;
            hy = t - 2 * E - kCgTile;
        }
    }
    if (ntiles <= nb) {
        // One tile per workgroup for the whole solve, each thread keeping
        // its node's vectors and stencil row in registers, and a
        // preconditioner that acts inside each tile: M^-1 = a Chebyshev
        // polynomial of odd degree kChebDegree in D^-1 K_tile (K restricted
        // to the tile's nodes) times D^-1, on [lmax / 100, lmax] with lmax
        // the tile's Gershgorin bound of D^-1/2 K_tile D^-1/2.  An odd-degree
        // Chebyshev polynomial keeps M SPD for any positive spectrum, so CG
        // stays valid; it needs no communication between tiles, and cuts
        // the iteration count about threefold against Jacobi.  Standard PCG:
        // two grid barriers per iteration (p.Kp, then r.z with r.r, which
        // also publishes z for the neighbours' halo of p).
        constexpr int kChebDegree = 9;
        __shared__ double s_t[2][E * E];   // preconditioner iterates, zero ring
        const bool have_tile = (int)blockIdx.x < ntiles;
        const int x0 = have_tile ? ((int)blockIdx.x % tiles_x) * kCgTile : 0;
        const int y0 = have_tile ? ((int)blockIdx.x / tiles_x) * kCgTile : 0;
        const int ix = x0 + lx, iy = y0 + ly;
        const bool own = have_tile && ix < N && iy < N;
        const int i = own ? iy * N + ix : 0;
        const int me = (ly + 1) * E + lx + 1;
        double Kr[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            Kr[k] = own ? K[(size_t)k * nn + i] : 0.0;
        }
        double xv = 0.0, rv = 0.0, dv = 0.0;
        if (own) {
            xv = S.x[i];
            rv = S.r[i];
            dv = dinv[i];
        }
        const int hxg = x0 + hx, hyg = y0 + hy;
        const bool hin = have_tile && hx != -2 && hxg >= 0 && hxg < N && hyg >= 0 && hyg < N;
        const int hj = hin ? hyg * N + hxg : 0;
        // the tile's Gershgorin bound, from D^-1 in LDS (zero outside the tile)
        for (int k = threadIdx.x; k < E * E; k += blockDim.x) {
            s_t[0][k] = 0.0;
            s_t[1][k] = 0.0;
        }
        __syncthreads();
        s_t[0][me] = dv;
        __syncthreads();
        double rowg = 0.0;
        if (own) {
#pragma unroll
            for (int oy = -1; oy <= 1; ++oy) {
#pragma unroll
                for (int ox = -1; ox <= 1; ++ox) {
                    rowg += fabs(Kr[(oy + 1) * 3 + (ox + 1)])
                            * sqrt(dv * s_t[0][me + oy * E + ox]);
                }
            }
        }
        __syncthreads();
        s_t[0][me] = 0.0;
        const double lmax = block_max(rowg, s_red);
        const double lmin = lmax / 100.0;
        const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
        const double sigma = theta / delta;
        // z = M^-1 r for this thread's node (every thread of the workgroup
        // takes part: the tile's products go through LDS)
        auto precond = [&](double r) {
            double dd = dv * r / theta;
            double z = dd;
            double rho_old = 1.0 / sigma;
            for (int k = 1; k < kChebDegree; ++k) {
                double *t = s_t[k & 1];
                if (own) {
                    t[me] = z;
                }
                __syncthreads();
                double kz = 0.0;
#pragma unroll
                for (int oy = -1; oy <= 1; ++oy) {
#pragma unroll
                    for (int ox = -1; ox <= 1; ++ox) {
                        kz += Kr[(oy + 1) * 3 + (ox + 1)] * t[me + oy * E + ox];
                    }
                }
                const double rho = 1.0 / (2.0 * sigma - rho_old);
                dd = rho * rho_old * dd + (2.0 * rho / delta) * (dv * (r - kz));
                z += dd;
                rho_old = rho;
            }
            return own ? z : 0.0;
        };
        double * const pz = part + 4 * kNlMaxBlocks;   // [rz0, pq, rz, rr] partials
        double zv = precond(rv);
        if (own) {
            S.u[i] = zv;
        }
        double rz = block_sum(rv * zv, s_red);
        if (threadIdx.x == 0) {
            pz[blockIdx.x] = rz;
        }
        grid_barrier(S.barrier, target);
        double hz = hin ? S.u[hj] : 0.0;
        rz = grid_total(pz, nb, s_red);
        double pv = zv, ph = hz;
        while (it < max_iter && rr > stop) {
            // q = K p, from p on the tile and its halo
            s_u[me] = own ? pv : 0.0;
            if (hx != -2) {
                s_u[(hy + 1) * E + hx + 1] = ph;
            }
            __syncthreads();
            double qv = 0.0;
            if (own) {
#pragma unroll
                for (int oy = -1; oy <= 1; ++oy) {
#pragma unroll
                    for (int ox = -1; ox <= 1; ++ox) {
                        const int jx = ix + ox, jy = iy + oy;
                        if (jx >= 0 && jx < N && jy >= 0 && jy < N) {
                            qv += Kr[(oy + 1) * 3 + (ox + 1)] * s_u[me + oy * E + ox];
                        }
                    }
                }
            }
            const double pq_b = block_sum(pv * qv, s_red);
            if (threadIdx.x == 0) {
                pz[kNlMaxBlocks + blockIdx.x] = pq_b;
            }
            grid_barrier(S.barrier, target);
            const double pq = grid_total(pz + kNlMaxBlocks, nb, s_red);
            const double alpha = rz / pq;
            xv += alpha * pv;
            rv -= alpha * qv;
            zv = precond(rv);
            if (own) {
                S.u[i] = zv;
            }
            double rzl = rv * zv, rrl = rv * rv, unused = 0.0;
            block_sum3(rzl, rrl, unused, s_red);
            if (threadIdx.x == 0) {
                pz[2 * kNlMaxBlocks + blockIdx.x] = rzl;
                pz[3 * kNlMaxBlocks + blockIdx.x] = rrl;
            }
            grid_barrier(S.barrier, target);
            hz = hin ? S.u[hj] : 0.0;
            const double rz_new = grid_total(pz + 2 * kNlMaxBlocks, nb, s_red);
            rr = grid_total(pz + 3 * kNlMaxBlocks, nb, s_red);
            const double beta = rz_new / rz;
            rz = rz_new;
            pv = zv + beta * pv;
            ph = hz + beta * ph;
            ++it;
        }
        if (own) {
            S.x[i] = xv;
        }
        if (gid == 0) {
            S.stats[0] = it;
            S.res[0] = ff > 0.0 ? sqrt(rr / ff) : 0.0;
        }
        return;
    }
    while (it < max_iter && rr > stop) {
        const int cur = it & 1;
        const double *uo = S.u + (size_t)cur * nn, *wo = S.w + (size_t)cur * nn;
        const double *so = S.s + (size_t)cur * nn;
        double *un = S.u + (size_t)(cur ^ 1) * nn, *wn = S.w + (size_t)(cur ^ 1) * nn;
        double *sn = S.s + (size_t)(cur ^ 1) * nn;
        double gl = 0.0, dll = 0.0, rrl = 0.0;
        for (int tile = blockIdx.x; tile < ntiles; tile += nb) {
            const int x0 = (tile % tiles_x) * kCgTile, y0 = (tile / tiles_x) * kCgTile;
            const int ix = x0 + lx, iy = y0 + ly;
            const bool own = ix < N && iy < N;
            const int i = iy * N + ix;
            double si = 0.0, ui = 0.0;
            if (own) {
                si = wo[i] + beta * so[i];
                ui = uo[i] - alpha * (dinv[i] * si);
            }
            double uh = 0.0;
            if (hx != -2) {
                const int jx = x0 + hx, jy = y0 + hy;
                if (jx >= 0 && jx < N && jy >= 0 && jy < N) {
                    const int j = jy * N + jx;
                    const double sj = wo[j] + beta * so[j];
                    uh = uo[j] - alpha * (dinv[j] * sj);
                }
            }
            __syncthreads();   // the previous tile's products are done with s_u
            s_u[(ly + 1) * E + lx + 1] = ui;
            if (hx != -2) {
                s_u[(hy + 1) * E + hx + 1] = uh;
            }
            __syncthreads();
            if (own) {
                double w = 0.0;
#pragma unroll
                for (int oy = -1; oy <= 1; ++oy) {
#pragma unroll
                    for (int ox = -1; ox <= 1; ++ox) {
                        const int jx = ix + ox, jy = iy + oy;
                        if (jx >= 0 && jx < N && jy >= 0 && jy < N) {
                            w += K[(size_t)((oy + 1) * 3 + (ox + 1)) * nn + i]
                                 * s_u[(ly + 1 + oy) * E + lx + 1 + ox];
                        }
                    }
                }
                const double pi = uo[i] + beta * S.p[i];
                S.p[i] = pi;
                S.x[i] += alpha * pi;
                const double ri = S.r[i] - alpha * si;
                S.r[i] = ri;
                sn[i] = si;
                un[i] = ui;
                wn[i] = w;
                gl += ri * ui;
                dll += w * ui;
                rrl += ri * ri;
            }
        }
        double *pt = part + (size_t)((it + 1) & 1) * 4 * kNlMaxBlocks;
        block_sum3(gl, dll, rrl, s_red);
        if (threadIdx.x == 0) {
            pt[blockIdx.x] = gl;
            pt[kNlMaxBlocks + blockIdx.x] = dll;
            pt[2 * kNlMaxBlocks + blockIdx.x] = rrl;
        }
        grid_barrier(S.barrier, target);
        double gamma_new, delta_new;
        grid_total3(pt, nb, gamma_new, delta_new, rr, s_red);
        beta = gamma_new / gamma;
        alpha = gamma_new / (delta_new - beta * gamma_new / alpha);
        gamma = gamma_new;
        ++it;
    }
    if (gid == 0) {
        S.stats[0] = it;
        S.res[0] = ff > 0.0 ? sqrt(rr / ff) : 0.0;
    }
}

// Nonlocal fluidity at the particles and the explicit visco-plastic update
// (g_nonlocal.c:145-301, solve_diffusion_part 538-564), every particle.
__global__ __launch_bounds__(kG2PThreads) void k_nl_stress(
    const Particles *__restrict__ Pd, const double *__restrict__ gnode, GridParams g, Material mt,
    const DevState *__restrict__ st, int np)
{
    const Particles &P = *Pd;   // field pointers, loaded as used
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) {
        return;
    }
    double sxx = P.txx[p], sxy = P.txy[p], syy = P.tyy[p];
    double gf = P.gf[p];
    if (P.active[p]) {
        const int e = P.elem_prev[p];
        const int cc = e % g.Ne, rr = e / g.Ne;
        const double xl = P.xl[p], yl = P.yl[p];
        const double ox = 1.0 - xl, oy = 1.0 - yl;
        const size_t n0 = (size_t)rr * g.N + cc;
        gf = 0.0;
        gf += gnode[n0] * (ox * oy);
        gf += gnode[n0 + 1] * (xl * oy);
        gf += gnode[n0 + g.N + 1] * (xl * yl);
        gf += gnode[n0 + g.N] * (ox * yl);
        P.gf[p] = gf;
    } else {
        double gfb, xsi;
        nonlocal_bulk(sxx, sxy, syy, mt, gfb, xsi);
        P.gf_bulk[p] = gfb;
        P.xisq_inv[p] = xsi;
    }
    const double dt = g.dt;
    const double E = mt.nl_E, nu = mt.nl_nu, c = mt.nl_c;
    const double exx = P.exx[p], exy = P.exy[p], eyy = P.eyy[p], wxy = P.wxy[p];
    double dsjxx = dt * (E / (1 - nu * nu)) * ((exx) + nu * (eyy));
    double dsjyy = dt * (E / (1 - nu * nu)) * ((eyy) + nu * (exx));
    dsjxx -= 2 * dt * wxy * sxy;
    dsjyy += 2 * dt * wxy * sxy;
    const double p_tr = -0.5 * ((sxx + dsjxx) + (syy + dsjyy));
    if (p_tr < c) {
        P.txx[p] = 0.5 * c / mt.nl_mu_s;
        P.txy[p] = 0.0;
        P.tyy[p] = 0.5 * c / mt.nl_mu_s;
        return;
    }
    const double p_t = -0.5 * (sxx + syy);
    const double sxx_t0 = sxx + p_t, sxy_t0 = sxy, syy_t0 = syy + p_t;
    const double tau_t = sqrt(0.5 * (sxx_t0 * sxx_t0 + 2 * sxy_t0 * sxy_t0 + syy_t0 * syy_t0));
    const double r2 = sqrt(0.5);
    const double Npxx = r2 * sxx_t0 / tau_t;
    const double Npxy = r2 * sxy_t0 / tau_t;
    const double Npyy = r2 * syy_t0 / tau_t;
    const double dxx0 = 0.5 * (exx - eyy);
    const double dxy0 = exy;
    const double dyy0 = 0.5 * (eyy - exx);
    const double gammadot = sqrt(2 * (dxx0 * dxx0 + 2 * dxy0 * dxy0 + dyy0 * dyy0));
    double nup = 0.0, dpxx = 0.0, dpxy = 0.0, dpyy = 0.0;
    if (p_t > c && gf > 0) {
        nup = gf * tau_t / p_t;
        if (nup < 0) nup = 0;
        if (nup > gammadot) nup = gammadot;
        dpxx = r2 * (nup) * Npxx;
        dpxy = r2 * (nup) * Npxy;
        dpyy = r2 * (nup) * Npyy;
    }
    if (st->t == 0) {
        nup = 0.0;
        dpxx = dpxy = dpyy = 0.0;
    }
    dsjxx = dt * (E / (1 - nu * nu)) * ((exx - dpxx) + nu * (eyy - dpyy));
    double dsjxy = dt * (E / (2 * (1 + nu))) * (exy - dpxy);
    dsjyy = dt * (E / (1 - nu * nu)) * ((eyy - dpyy) + nu * (exx - dpxx));
    dsjxx -= 2 * dt * wxy * sxy;
    dsjxy += dt * wxy * (sxx - syy);
    dsjyy += 2 * dt * wxy * sxy;
    P.gp[p] += nup * dt;
    P.txx[p] = sxx + dsjxx;
    P.txy[p] = sxy + dsjxy;
    P.tyy[p] = syy + dsjyy;
}

// ---------------------------------------------------------------------------
// Adjoint (reverse mode) of the USL step with g_local_mu2.
//
// The step is differentiated on its recorded branch: particle cells, active
// flags, node-mass thresholds, boundary masks and the material branch
// (density cut-off / elastic / plastic) are held fixed.  One reverse step is
// three kernels after the forward node state of the step has been rebuilt
// (k_p2g_grid on the step's start state):
//   k_adj_g2p    per particle: reverse of the stress update
//                (g_local_mu2.c:197-307), the volume update
//                (process_usl.c:961-970), the strain rate (764-808) and G2P
//                (914-957); leaves in the particle adjoint the terms that do
//                not pass through the grid, and 8 coefficients that carry the
//                rest to the nodes.
//   k_adj_nodes  per node: gathers those coefficients from the particles of
//                the node's four cells (bins sorted by index, so the sums are
//                deterministic), then the reverse of move_grid_split
//                (596-625) and of the boundary zeroing (box.c:183-224).
//   k_adj_p2g    per particle: reverse of the P2G scatter
//                (process_usl.c:812-910, map.c:85-104) by a gather from the
//                particle's four nodes.
// ---------------------------------------------------------------------------

constexpr int kAdjParams = 7;   // adjoints of G, lambda, mu_s, mu_2, I_0, d, rho_s
constexpr int kAdjCoef = 8;     // A, B, Cx, Cy, exx~, dxy~, eyy~, dyx~ per particle

struct AdjState {
    double *lam[MPM_F_EXX];     // particle adjoint, updated in place
    double *coef;               // [kAdjCoef][np]
    double *nadj;               // [5][nn]: m~, px~, py~, fx~, fy~
    double *cellsum;            // [6][4][cells]: per cell and corner, the sums of its
                                // particles' terms (Fx, Fy, Ux, Uy, Vx, Vy) in index order
    double *partials;           // [kAdjParams][blocks]
};

// Reverse of g_local_mu2_update.  In: the update's inputs and the adjoints
// of its outputs (Tb[4] for Txx, Txy, Tyy, Tzz; gpb, gdpb).  Out: adjoints of
// the strain-rate inputs, of the input stress (written over Tb) and of the
// constants (pb[kAdjParams], added).
__device__ __forceinline__ void g_local_mu2_reverse(
    double exx, double exy, double eyy, double wxy, double m, double v_new, double dt,
    const Material &mt, double txx, double txy, double tyy, double tzz, double Tb[4],
    double gpb, double gdpb, double &exxb, double &exyb, double &eyyb, double &wxyb,
    double pb[kAdjParams])
{
    const double w = 0.5 * ((exy + wxy) - (exy - wxy));
    const double dxx = 0.5 * (exx + exx);
    const double dxy = 0.5 * ((exy + wxy) + (exy - wxy));
    const double dyy = 0.5 * (eyy + eyy);
    const double trD = (dxx + dyy) + 0.0;
    const double a = w * txy;
    const double ltr = mt.lambda * trD;
    const double rxx = ((dxx * mt.two_G + ltr) + (a + a)) * dt + txx;
    const double rxy = ((dxy * mt.two_G) + (w * tyy - w * txx)) * dt + txy;
    const double ryy = ((dyy * mt.two_G + ltr) + -(a + a)) * dt + tyy;
    const double rzz = ltr * dt + tzz;
    const double c = ((rxx + ryy) + rzz) / 3.0;
    const double d_xx = rxx - c, d_yy = ryy - c, d_zz = rzz - c;
    const double p = c * -1.0;
    double tau2 = d_xx * d_xx;
    tau2 += rxy * rxy;
    tau2 += rxy * rxy;
    tau2 += d_yy * d_yy;
    tau2 += d_zz * d_zz;
    const double tau = sqrt(0.5 * tau2);

    double rxxb = 0.0, rxyb = 0.0, ryyb = 0.0, rzzb = 0.0;
    exxb = exyb = eyyb = wxyb = 0.0;
    const bool zero = (m / v_new) < mt.rho_crit || p <= 0.0;
    if (zero) {
        Tb[0] = Tb[1] = Tb[2] = Tb[3] = 0.0;
        return;
    }
    if (!(p > 0.0)) {
        return;   // NaN pressure: the stress passes through unchanged
    }
    const double S0 = mt.mu_s * p;
    if (tau <= S0) {
        rxxb = Tb[0];
        rxyb = Tb[1];
        ryyb = Tb[2];
        rzzb = Tb[3];
    } else {
        const double S2 = mt.mu_2 * p;
        const double alpha = mt.G * mt.I_0 * dt * sqrt(p / mt.grains_rho) / mt.grains_d;
        const double B = -(S2 + tau + alpha);
        const double H = S2 * tau + S0 * alpha;
        double taup;
        if (B > 0.0) {
            taup = (-B - sqrt(B * B - 4.0 * H)) / 2.0;
        } else {
            taup = (2.0 * H) / (-B + sqrt(B * B - 4.0 * H));
        }
        const double scale = taup / tau;
        const double nupb = gdpb + gpb * dt;
        const double scaleb = Tb[0] * d_xx + Tb[1] * rxy + Tb[2] * d_yy + Tb[3] * d_zz;
        double dxxb = Tb[0] * scale, dyyb = Tb[2] * scale, dzzb = Tb[3] * scale;
        double rxyb_p = Tb[1] * scale;
        double Pb = -(Tb[0] + Tb[2] + Tb[3]);
        const double Gdt = mt.G * dt;
        double taub = nupb / Gdt;
        double taupb = -nupb / Gdt + scaleb / tau;
        double Gb = -nupb * (tau - taup) / (mt.G * Gdt);
        taub -= scaleb * taup / (tau * tau);
        const double den = 2.0 * taup + B;
        const double Bb = -taupb * taup / den;
        const double Hb = -taupb / den;
        const double S2b = -Bb + Hb * tau;
        taub += -Bb + Hb * S2;
        const double alphab = -Bb + Hb * S0;
        const double S0b = Hb * alpha;
        Gb += alphab * alpha / mt.G;
        pb[4] += alphab * alpha / mt.I_0;
        pb[5] += -alphab * alpha / mt.grains_d;
        pb[6] += -alphab * alpha / (2.0 * mt.grains_rho);
        Pb += alphab * alpha / (2.0 * p);
        pb[2] += S0b * p;
        Pb += S0b * mt.mu_s;
        pb[3] += S2b * p;
        Pb += S2b * mt.mu_2;
        pb[0] += Gb;
        const double tau2b = taub * 0.25 / tau;
        dxxb += 2.0 * d_xx * tau2b;
        rxyb_p += 4.0 * rxy * tau2b;
        dyyb += 2.0 * d_yy * tau2b;
        dzzb += 2.0 * d_zz * tau2b;
        const double cb = -Pb - (dxxb + dyyb + dzzb);
        rxxb = dxxb + cb / 3.0;
        ryyb = dyyb + cb / 3.0;
        rzzb = dzzb + cb / 3.0;
        rxyb = rxyb_p;
    }
    // trial stress
    const double jsxxb = rxxb * dt, jsxyb = rxyb * dt, jsyyb = ryyb * dt;
    const double ltrb = (rxxb + ryyb + rzzb) * dt;
    pb[0] += 2.0 * (rxxb * dxx + rxyb * dxy + ryyb * dyy) * dt;
    pb[1] += ltrb * trD;
    const double trDb = ltrb * mt.lambda;
    exxb = rxxb * dt * mt.two_G + trDb;
    exyb = rxyb * dt * mt.two_G;
    eyyb = ryyb * dt * mt.two_G + trDb;
    wxyb = 2.0 * txy * (jsxxb - jsyyb) + jsxyb * (tyy - txx);
    Tb[0] = rxxb - jsxyb * w;
    Tb[1] = rxyb + 2.0 * w * (jsxxb - jsyyb);
    Tb[2] = ryyb + jsxyb * w;
    Tb[3] = rzzb;
}

__device__ __forceinline__ void block_sum_params(double v[kAdjParams], double *s_red)
{
#pragma unroll
    for (int k = 0; k < kAdjParams; ++k) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            v[k] += __shfl_down(v[k], d);
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kAdjParams; ++k) s_red[k * nw + wv] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kAdjParams; ++k) {
        double t = 0.0;
        for (int q = 0; q < nw; ++q) t += s_red[k * nw + q];
        v[k] = t;
    }
}

__global__ __launch_bounds__(kG2PThreads) void k_adj_g2p(
    const Particles *__restrict__ Pd, const NodeRec *__restrict__ nodes, AdjState A, GridParams g, Material mt,
    int np)
{
    const Particles &P = *Pd;   // field pointers, loaded as used
    __shared__ double s_red[kAdjParams * (kG2PThreads / 64)];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double pb[kAdjParams] = {0, 0, 0, 0, 0, 0, 0};
    const bool upd = p < np && P.active[p] && P.elem[p] >= 0;
    if (p < np) {
        A.lam[MPM_F_M][p] = 0.0;
    }
    if (p < np && !upd) {
#pragma unroll
        for (int k = 0; k < kAdjCoef; ++k) A.coef[(size_t)k * np + p] = 0.0;
    }
    if (upd) {
        const int e = P.elem[p];
        const int cc = e % g.Ne, rr = e / g.Ne;
        const double dt = g.dt;
        Shape sh;
        shape_functions(P.x[p], P.y[p], cc, rr, g, sh);
        const size_t n0 = (size_t)rr * g.N + cc;
        const size_t nid[4] = {n0, n0 + 1, n0 + g.N + 1, n0 + g.N};
        NodeRec nd[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) nd[q] = nodes[nid[q]];
        const double exx = 1.0 * (sh.b1[0] * nd[0].vx + sh.b1[1] * nd[1].vx
                                  + sh.b1[2] * nd[2].vx + sh.b1[3] * nd[3].vx);
        const double eyy = 1.0 * (sh.b2[0] * nd[0].vy + sh.b2[1] * nd[1].vy
                                  + sh.b2[2] * nd[2].vy + sh.b2[3] * nd[3].vy);
        const double dxy = 1.0 * (sh.b2[0] * nd[0].vx + sh.b2[1] * nd[1].vx
                                  + sh.b2[2] * nd[2].vx + sh.b2[3] * nd[3].vx);
        const double dyx = 1.0 * (sh.b1[0] * nd[0].vy + sh.b1[1] * nd[1].vy
                                  + sh.b1[2] * nd[2].vy + sh.b1[3] * nd[3].vy);
        const double exy = 0.5 * (dxy + dyx);
        const double wxy = 0.5 * (dxy - dyx);
        const double ex = mpm_exp(dt * (exx + eyy));
        const double v = P.v[p];
        const double v_new = v * ex;
        const double mp = P.m[p];

        double lb[MPM_F_EXX];
#pragma unroll
        for (int f = 0; f < MPM_F_EXX; ++f) lb[f] = A.lam[f][p];
        double Tb[4] = {lb[MPM_F_TXX], lb[MPM_F_TXY], lb[MPM_F_TYY], lb[MPM_F_TZZ]};
        double exxb, exyb, eyyb, wxyb;
        g_local_mu2_reverse(exx, exy, eyy, wxy, mp, v_new, dt, mt, P.txx[p], P.txy[p],
                            P.tyy[p], P.tzz[p], Tb, lb[MPM_F_GAMMAP], lb[MPM_F_GAMMADOTP],
                            exxb, exyb, eyyb, wxyb, pb);
        // volume
        const double vb = lb[MPM_F_V] * ex;
        exxb += lb[MPM_F_V] * v_new * dt;
        eyyb += lb[MPM_F_V] * v_new * dt;
        const double dxyb = 0.5 * (exyb + wxyb);
        const double dyxb = 0.5 * (exyb - wxyb);
        // G2P
        const double Ab = lb[MPM_F_VX] * dt, Bb = lb[MPM_F_VY] * dt;
        const double Cx = (lb[MPM_F_X] + lb[MPM_F_UX]) * dt;
        const double Cy = (lb[MPM_F_Y] + lb[MPM_F_UY]) * dt;
        const double r2 = g.inv_h * g.inv_h;
        double xb = lb[MPM_F_X], yb = lb[MPM_F_Y];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double Sb = 0.0;
            if (!(nd[q].m < kTol)) {
                Sb = (Ab * nd[q].fx + Bb * nd[q].fy + Cx * nd[q].mx + Cy * nd[q].my) / nd[q].m;
            }
            const double b1b = exxb * nd[q].vx + dyxb * nd[q].vy;
            const double b2b = eyyb * nd[q].vy + dxyb * nd[q].vx;
            const double sg = (q & 1) ? -1.0 : 1.0;
            xb += Sb * sh.b1[q] + b2b * sg * r2;
            yb += Sb * sh.b2[q] + b1b * sg * r2;
        }
        A.lam[MPM_F_X][p] = xb;
        A.lam[MPM_F_Y][p] = yb;
        A.lam[MPM_F_V][p] = vb;
        A.lam[MPM_F_TXX][p] = Tb[0];
        A.lam[MPM_F_TXY][p] = Tb[1];
        A.lam[MPM_F_TYY][p] = Tb[2];
        A.lam[MPM_F_TZZ][p] = Tb[3];
        A.lam[MPM_F_GAMMADOTP][p] = 0.0;
        const double cf[kAdjCoef] = {Ab, Bb, Cx, Cy, exxb, dxyb, eyyb, dyxb};
#pragma unroll
        for (int k = 0; k < kAdjCoef; ++k) A.coef[(size_t)k * np + p] = cf[k];
    }
    block_sum_params(pb, s_red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kAdjParams; ++k) A.partials[(size_t)k * gridDim.x + blockIdx.x] = pb[k];
    }
}

// Sorts the first n of R register values (odd-even transposition network;
// the values past n are replaced by INT_MAX).  Returns whether the first n
// were already sorted (and then leaves v alone).
template <int R>
__device__ __forceinline__ bool sort_in_registers(int (&v)[R], int n)
{
    bool sorted = true;
#pragma unroll
    for (int k = 1; k < R; ++k) {
        sorted &= !(k < n && v[k - 1] > v[k]);
    }
    if (sorted) {
        return true;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        v[k] = k < n ? v[k] : INT_MAX;
    }
#pragma unroll
    for (int phase = 0; phase < R; ++phase) {
#pragma unroll
        for (int k = phase & 1; k + 1 < R; k += 2) {
            const int lo = min(v[k], v[k + 1]), hi = max(v[k], v[k + 1]);
            v[k] = lo;
            v[k + 1] = hi;
        }
    }
    return false;
}

// Per cell, four lanes, one per corner node q: the terms of the cell's
// particles for that node (shape value and gradients times the particle's
// coefficients), summed in ascending particle index.  The four lanes of a
// cell load the same particles (one request) and evaluate the shape
// functions as shape_functions does, each keeping its corner.  Bins are
// filled in lane order and are mostly sorted already; bins of up to 32
// entries are checked and sorted in registers (the first 16 entries are
// loaded with the count), longer ones are sorted by insertion; each lane
// keeps its sorted copy in its own LDS row.
__global__ __launch_bounds__(256) void k_adj_cells(
    const Particles *__restrict__ Pd, const int *__restrict__ cnt, const int *__restrict__ bkt,
    AdjState A, GridParams g, int np)
{
    const Particles &P = *Pd;   // field pointers, loaded as used
    __shared__ int s_bin[256 * 65];   // per thread, stride 65 (no bank conflicts)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int cell = t >> 2, q = t & 3;
    const bool in = cell < g.ncells;
    const int cc = in ? cell % g.Ne : 0, rr = in ? cell / g.Ne : 0;
    const int *l = bkt + (size_t)(in ? cell : 0) * g.cap;
    int *sb = s_bin + threadIdx.x * 65;
    int n = 0;
    if (in) {
        const int4 *l4 = reinterpret_cast<const int4 *>(l);
        int v[32];
#pragma unroll
        for (int k = 0; k < 4; ++k) {   // every capacity holds 16 entries
            const int4 w = l4[k];
            v[4 * k] = w.x;
            v[4 * k + 1] = w.y;
            v[4 * k + 2] = w.z;
            v[4 * k + 3] = w.w;
        }
        n = cnt[cell];
        n = n < g.cap ? n : g.cap;
        if (n <= 16) {
            int (&v16)[16] = *reinterpret_cast<int (*)[16]>(v);
            sort_in_registers<16>(v16, n);
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                sb[k] = v[k];
            }
        } else if (n <= 32) {
#pragma unroll
            for (int k = 4; k < 8; ++k) {
                const int4 w = l4[k];
                v[4 * k] = w.x;
                v[4 * k + 1] = w.y;
                v[4 * k + 2] = w.z;
                v[4 * k + 3] = w.w;
            }
            sort_in_registers<32>(v, n);
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                sb[k] = v[k];
            }
        } else {
            for (int k = 0; k < n; ++k) {
                const int key = l[k];
                int r = k - 1;
                while (r >= 0 && sb[r] > key) {
                    sb[r + 1] = sb[r];
                    --r;
                }
                sb[r + 1] = key;
            }
        }
    }
    if (!in) {
        return;
    }
    const int *bin = sb;
    const double xn = (double)cc * g.h, yn = (double)rr * g.h;
    const bool qx = q == 1 || q == 2, qy = q >= 2;
    double F[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // particles in chunks of 4: the chunk's loads are all in flight
    // together; the sums still run in index order
    for (int k0 = 0; k0 < n; k0 += 4) {
        int pk[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            pk[u] = bin[k0 + u < n ? k0 + u : k0];
        }
        double px[4], py[4], cf[4][8];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            px[u] = P.x[pk[u]];
            py[u] = P.y[pk[u]];
#pragma unroll
            for (int f = 0; f < 8; ++f) {
                cf[u][f] = A.coef[(size_t)f * np + pk[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (k0 + u < n) {
                // shape_functions' values for corner q
                const double xl = snap_local(div_h(px[u] - xn, g));
                const double yl = snap_local(div_h(py[u] - yn, g));
                const double X = qx ? xl : 1.0 - xl;
                const double Y = qy ? yl : 1.0 - yl;
                const double S = X * Y;
                const double b1 = (q == 0 || q == 3 ? -Y : Y) * g.inv_h;
                const double b2 = (q <= 1 ? -X : X) * g.inv_h;
                F[0] += S * cf[u][0];
                F[1] += S * cf[u][1];
                F[2] += S * cf[u][2];
                F[3] += S * cf[u][3];
                F[4] += b1 * cf[u][4] + b2 * cf[u][5];
                F[5] += b2 * cf[u][6] + b1 * cf[u][7];
            }
        }
    }
    const size_t nc = (size_t)g.ncells;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        A.cellsum[(f * 4 + q) * nc + cell] = F[f];
    }
}

// Per node: the reverse of the grid update and of the boundary zeroing,
// from the sums of the node's cells (k_adj_cells) added in colour order.
__global__ __launch_bounds__(256) void k_adj_nodes(
    const NodeRec *__restrict__ nodes, const uint8_t *__restrict__ mask, AdjState A,
    GridParams g, const DevState *__restrict__ st)
{
    const int nidx = blockIdx.x * blockDim.x + threadIdx.x;
    const int nn = g.N * g.N;
    if (nidx >= nn) {
        return;
    }
    const int i = nidx % g.N, j = nidx / g.N;
    const size_t nc = (size_t)g.ncells;
    // the sums of colour c's cell (zero where the node has none), in order
    double v[4][6];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int pc = c & 1, pr = c >> 1;
        const int cc = i - ((i ^ pc) & 1);
        const int rr = j - ((j ^ pr) & 1);
        const bool has = cc >= 0 && rr >= 0 && cc < g.Ne && rr < g.Ne;
        const int dx = i - cc, dy = j - rr;
        const int q = dy ? (dx ? 2 : 3) : (dx ? 1 : 0);
        const int cell = has ? rr * g.Ne + cc : 0;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            v[c][f] = has ? A.cellsum[(f * 4 + q) * nc + cell] : 0.0;
        }
    }
    double Fsum[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        Fsum[f] = ((v[0][f] + v[1][f]) + v[2][f]) + v[3][f];
    }
    const double Fx = Fsum[0], Fy = Fsum[1], Ux = Fsum[2], Uy = Fsum[3];
    const double Vx = Fsum[4], Vy = Fsum[5];
    const NodeRec nd = nodes[nidx];
    const double m = nd.m;
    double mb = 0.0, pxb = 0.0, pyb = 0.0, fxb = 0.0, fyb = 0.0;
    if (!(m < kTol)) {
        fxb = Fx / m;
        fyb = Fy / m;
        pxb = Ux / m;
        pyb = Uy / m;
        mb = -(Fx * nd.fx + Fy * nd.fy + Ux * nd.mx + Uy * nd.my) / (m * m);
    }
    if (m > kTol) {
        // v = p~ / m, p~ = p + dt f
        pxb += Vx / m;
        pyb += Vy / m;
        mb -= (Vx * nd.mx + Vy * nd.my) / (m * m);
        fxb += g.dt * pxb;
        fyb += g.dt * pyb;
    }
    const unsigned f = bc_fixed(g, i, j, st->t + g.dt, mask);
    if (f & 1u) {
        pxb = 0.0;
        fxb = 0.0;
    }
    if (f & 2u) {
        pyb = 0.0;
        fyb = 0.0;
    }
    A.nadj[(size_t)0 * nn + nidx] = mb;
    A.nadj[(size_t)1 * nn + nidx] = pxb;
    A.nadj[(size_t)2 * nn + nidx] = pyb;
    A.nadj[(size_t)3 * nn + nidx] = fxb;
    A.nadj[(size_t)4 * nn + nidx] = fyb;
}

__global__ __launch_bounds__(kG2PThreads) void k_adj_p2g(
    const Particles *__restrict__ Pd, AdjState A, GridParams g, const DevState *__restrict__ st, int np)
{
    const Particles &P = *Pd;   // field pointers, loaded as used
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || !P.active[p] || P.elem[p] < 0) return;
    const int e = P.elem[p];
    const int cc = e % g.Ne, rr = e / g.Ne;
    Shape sh;
    shape_functions(P.x[p], P.y[p], cc, rr, g, sh);
    const int nn = g.N * g.N;
    const size_t n0 = (size_t)rr * g.N + cc;
    const size_t nid[4] = {n0, n0 + 1, n0 + g.N + 1, n0 + g.N};
    const double m = P.m[p], v = P.v[p], vx = P.vx[p], vy = P.vy[p];
    const double sxx = P.txx[p], sxy = P.txy[p], syy = P.tyy[p];
    const double bx = st->bx, by = st->by;
    const double r2 = g.inv_h * g.inv_h;
    double xb = 0, yb = 0, vxb = 0, vyb = 0, vb = 0, txxb = 0, txyb = 0, tyyb = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double mbj = A.nadj[(size_t)0 * nn + nid[q]];
        const double pxbj = A.nadj[(size_t)1 * nn + nid[q]];
        const double pybj = A.nadj[(size_t)2 * nn + nid[q]];
        const double fxbj = A.nadj[(size_t)3 * nn + nid[q]];
        const double fybj = A.nadj[(size_t)4 * nn + nid[q]];
        const double Sb = mbj * m + pxbj * (vx * m) + pybj * (vy * m) + fxbj * (bx * m)
                          + fybj * (by * m);
        const double b1b = fxbj * (-sxx * v) + fybj * (-sxy * v);
        const double b2b = fxbj * (-sxy * v) + fybj * (-syy * v);
        const double sg = (q & 1) ? -1.0 : 1.0;
        xb += Sb * sh.b1[q] + b2b * sg * r2;
        yb += Sb * sh.b2[q] + b1b * sg * r2;
        vxb += sh.s[q] * pxbj;
        vyb += sh.s[q] * pybj;
        txxb += sh.b1[q] * fxbj;
        txyb += sh.b2[q] * fxbj + sh.b1[q] * fybj;
        tyyb += sh.b2[q] * fybj;
        vb += fxbj * (-sxx * sh.b1[q] - sxy * sh.b2[q]) + fybj * (-sxy * sh.b1[q] - syy * sh.b2[q]);
    }
    A.lam[MPM_F_X][p] += xb;
    A.lam[MPM_F_Y][p] += yb;
    A.lam[MPM_F_VX][p] += m * vxb;
    A.lam[MPM_F_VY][p] += m * vyb;
    A.lam[MPM_F_V][p] += vb;
    A.lam[MPM_F_TXX][p] += -v * txxb;
    A.lam[MPM_F_TXY][p] += -v * txyb;
    A.lam[MPM_F_TYY][p] += -v * tyyb;
}

// Copies a particle state (the 14 step fields, active flags, device
// scalars) in one launch, and optionally clears the two bin-count arrays
// for the re-binning that follows a load.
struct StateCopy {
    const double *src[MPM_F_EXX];
    double *dst[MPM_F_EXX];
    const int *asrc;
    int *adst;
    const int *esrc;   // element indices too, when set
    int *edst;
    const DevState *ssrc;
    DevState *sdst;
    int *cnt0, *cnt1;
    int np, ncells;
};

__global__ __launch_bounds__(256) void k_copy_state(StateCopy c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < c.np) {
#pragma unroll
        for (int f = 0; f < MPM_F_EXX; ++f) {
            c.dst[f][i] = c.src[f][i];
        }
        c.adst[i] = c.asrc[i];
        if (c.edst) {
            c.edst[i] = c.esrc[i];
        }
    }
    if (c.cnt0 && i < c.ncells) {
        c.cnt0[i] = 0;
        c.cnt1[i] = 0;
    }
    if (i == 0) {
        *c.sdst = *c.ssrc;
    }
}

// acc[k] += sum over workgroups of the step's parameter partials, in order.
__global__ __launch_bounds__(256) void k_adj_reduce(const double *__restrict__ partials,
                                                    int nblocks, double *__restrict__ acc)
{
    // the seven sums side by side, each in block_sum's order
    __shared__ double s_red[kAdjParams][4];
    double v[kAdjParams];
#pragma unroll
    for (int k = 0; k < kAdjParams; ++k) {
        v[k] = 0.0;
    }
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < kAdjParams; ++k) {
            v[k] += partials[(size_t)k * nblocks + b];
        }
    }
#pragma unroll
    for (int k = 0; k < kAdjParams; ++k) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            v[k] += __shfl_down(v[k], d);
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kAdjParams; ++k) {
            s_red[k][w] = v[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < kAdjParams) {
        double t = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
            t += s_red[threadIdx.x][q];
        }
        acc[threadIdx.x] += t;
    }
}


__global__ void k_eval_ops(int op, const double *a, const double *b, double *out, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (op == 0) {
        out[i] = mpm_exp(a[i]);
    } else if (op == 1) {
        out[i] = a[i] / b[i];
    } else {
        out[i] = sqrt(a[i]);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct mpm_hip_ctx {
    mpm_hip_config cfg;
    int device;
    size_t np;
    GridParams g;
    Material mt;
    hipStream_t stream;
    Particles P;
    Particles *dP;          // device copy of P: kernels load field pointers as used
    double *pfield[MPM_F_COUNT];
    NodeRec *nodes;
    uint8_t *tile_zero;     // per P2G tile: its node records are all zero
    uint8_t *mask;
    int *cnt[2];
    int *bkt[2];
    DevState *st;
    int cur;                // bins holding the current step's lists
    int tiles_x, ntiles, p2g_blocks;
    TileOrder order;        // k_p2g_grid's dispatch order (p2g_blocks follows it)
    int p2g_roles;          // threads per node in k_p2g_grid (2 or 3)
    bool have_particles;
    // g_nonlocal
    bool nonlocal;
    NlSystem nl;
    int nl_tiles_x, nl_ntiles, cg_blocks;
};

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t err_ = (expr);                                                  \
        if (err_ != hipSuccess) {                                                  \
            fprintf(stderr, "mpm_hip: %s failed: %s (%s:%d)\n", #expr,              \
                    hipGetErrorString(err_), __FILE__, __LINE__);                  \
            return MPM_HIP_ERR_DEVICE;                                             \
        }                                                                          \
    } while (0)

// Points a particle-array table at field arrays f[MPM_F_*].
static void bind_fields_to(Particles &P, double *const *f)
{
    P.x = f[MPM_F_X];
    P.y = f[MPM_F_Y];
    P.vx = f[MPM_F_VX];
    P.vy = f[MPM_F_VY];
    P.m = f[MPM_F_M];
    P.v = f[MPM_F_V];
    P.txx = f[MPM_F_TXX];
    P.txy = f[MPM_F_TXY];
    P.tyy = f[MPM_F_TYY];
    P.tzz = f[MPM_F_TZZ];
    P.gp = f[MPM_F_GAMMAP];
    P.gdp = f[MPM_F_GAMMADOTP];
    P.ux = f[MPM_F_UX];
    P.uy = f[MPM_F_UY];
    P.exx = f[MPM_F_EXX];
    P.exy = f[MPM_F_EXY];
    P.eyy = f[MPM_F_EYY];
    P.wxy = f[MPM_F_WXY];
    P.ax = f[MPM_F_AX];
    P.ay = f[MPM_F_AY];
    P.xl = f[MPM_F_XL];
    P.yl = f[MPM_F_YL];
    P.gf = f[MPM_F_GF];
    P.gf_bulk = f[MPM_F_GF_BULK];
    P.xisq_inv = f[MPM_F_XISQ_INV];
}

static void bind_fields(mpm_hip_ctx *ctx)
{
    bind_fields_to(ctx->P, ctx->pfield);
}

// Refreshes the device copy of a particle-array table (after an allocation
// changed it; dst may not exist yet during set-up).
static int upload_particles(Particles *dst, const Particles &src)
{
    if (dst && hipMemcpy(dst, &src, sizeof(Particles), hipMemcpyHostToDevice) != hipSuccess) {
        return MPM_HIP_ERR_DEVICE;
    }
    return MPM_HIP_OK;
}

// Auxiliary per-particle outputs, allocated (zeroed) on first use.
static int alloc_aux(mpm_hip_ctx *ctx)
{
    const size_t bytes = (ctx->np ? ctx->np : 8) * sizeof(double);
    for (int f = MPM_F_EXX; f < MPM_F_COUNT; ++f) {
        if (ctx->pfield[f]) continue;
        if (hipMalloc((void **)&ctx->pfield[f], bytes) != hipSuccess
            || hipMemsetAsync(ctx->pfield[f], 0, bytes, ctx->stream) != hipSuccess) {
            return MPM_HIP_ERR_ALLOC;
        }
    }
    bind_fields(ctx);
    return upload_particles(ctx->dP, ctx->P);
}

// k_p2g_grid's grid: whole rounds of 8 chunks (one per XCD); workgroups
// past the last tile return at once.
static int p2g_grid_blocks(int ntiles, int chunk)
{
    const int chunks = (ntiles + chunk - 1) / chunk;
    return ((chunks + 7) / 8) * 8 * chunk;
}

// Dispatch order of k_p2g_grid (TileOrder): core, rim of the box, the rest.
static void set_tile_order(mpm_hip_ctx *ctx, int cx0, int cy0, int cw, int ch, int bx0, int by0,
                           int bw, int bh)
{
    TileOrder &o = ctx->order;
    o.cx0 = cx0;
    o.cy0 = cy0;
    o.cw = cw;
    o.ch = ch;
    o.bx0 = bx0;
    o.by0 = by0;
    o.bw = bw;
    o.bh = bh;
    const int nc = cw * ch, nb = bw * bh;
    o.core_rows = (ch + 7) / 8;
    o.blocks_a = p2g_grid_blocks(nc, cw * o.core_rows);
    o.blocks_b = nb > nc ? p2g_grid_blocks(nb - nc, 1) : 0;
    ctx->p2g_blocks = o.blocks_a + o.blocks_b
                      + (nb < ctx->ntiles ? p2g_grid_blocks(ctx->ntiles - nb, ctx->tiles_x) : 0);
}

// Row order over the whole grid: no hint.
static void set_row_order(mpm_hip_ctx *ctx)
{
    const int ty = ctx->ntiles / ctx->tiles_x;
    set_tile_order(ctx, 0, 0, ctx->tiles_x, ty, 0, 0, ctx->tiles_x, ty);
}

// The box of tiles whose cell region (17 x 9 cells from one below and left
// of the tile's first node) holds a particle on the grid, and its core: the
// box without an edge column or row that carries less than half the load of
// the box's heaviest column or row (the one-cell halo of a block of
// particles).  Row order when no particle is on the grid.
static void order_tiles_by_particles(mpm_hip_ctx *ctx, const double *x, const double *y)
{
    const int tx = ctx->tiles_x, ty = ctx->ntiles / ctx->tiles_x;
    const GridParams &g = ctx->g;
    long long *col = new (std::nothrow) long long[tx + ty]();
    if (!col) {
        set_row_order(ctx);
        return;
    }
    long long *row = col + tx;
    int c0 = tx, c1 = -1, r0 = ty, r1 = -1;
    for (size_t p = 0; p < ctx->np; ++p) {
        const double fx = std::floor(x[p] / g.h), fy = std::floor(y[p] / g.h);
        if (!(fx >= 0.0 && fy >= 0.0 && fx < (double)g.Ne && fy < (double)g.Ne)) {
            continue;
        }
        // cell c lies in the regions of tile columns c / 16 and (c + 1) / 16
        const int c = (int)fx, r = (int)fy;
        const int ta = c / kTileX, tb = std::min((c + 1) / kTileX, tx - 1);
        const int ra = r / kTileY, rb = std::min((r + 1) / kTileY, ty - 1);
        col[ta] += 1;
        col[tb] += tb != ta;
        row[ra] += 1;
        row[rb] += rb != ra;
        c0 = std::min(c0, ta);
        c1 = std::max(c1, tb);
        r0 = std::min(r0, ra);
        r1 = std::max(r1, rb);
    }
    if (c1 < 0) {
        delete[] col;
        set_row_order(ctx);
        return;
    }
    long long cmax = 0, rmax = 0;
    for (int t = c0; t <= c1; ++t) cmax = std::max(cmax, col[t]);
    for (int t = r0; t <= r1; ++t) rmax = std::max(rmax, row[t]);
    int k0 = c0, k1 = c1, q0 = r0, q1 = r1;
    if (k1 > k0 && 2 * col[k0] < cmax) ++k0;
    if (k1 > k0 && 2 * col[k1] < cmax) --k1;
    if (q1 > q0 && 2 * row[q0] < rmax) ++q0;
    if (q1 > q0 && 2 * row[q1] < rmax) --q1;
    delete[] col;
    set_tile_order(ctx, k0, q0, k1 - k0 + 1, q1 - q0 + 1, c0, r0, c1 - c0 + 1, r1 - r0 + 1);
}

static int all_finite(const double *a, size_t n)
{
    for (size_t k = 0; k < n; ++k) {
        if (!std::isfinite(a[k])) {
            return 0;
        }
    }
    return 1;
}

static int use_device(const mpm_hip_ctx *ctx)
{
    HIP_TRY(hipSetDevice(ctx->device));
    return MPM_HIP_OK;
}

template <bool MUSL2, int ROLES>
static void p2g_kernel_launch_roles(mpm_hip_ctx *ctx)
{
    const int cur = ctx->cur, nxt = 1 - cur;
    const dim3 grid(ctx->p2g_blocks), block(ROLES * kTileNodes);
    switch (ctx->g.cap) {
    case 16:
        k_p2g_grid<16, false, MUSL2, ROLES><<<grid, block, 0, ctx->stream>>>(ctx->dP,
            ctx->cnt[cur], ctx->bkt[cur], ctx->cnt[nxt], ctx->nodes, ctx->tile_zero, ctx->mask, ctx->g, ctx->st,
            ctx->tiles_x, ctx->ntiles, (int)ctx->np, ctx->order, BatchTables{});
        break;
    case 32:
        k_p2g_grid<32, false, MUSL2, ROLES><<<grid, block, 0, ctx->stream>>>(ctx->dP,
            ctx->cnt[cur], ctx->bkt[cur], ctx->cnt[nxt], ctx->nodes, ctx->tile_zero, ctx->mask, ctx->g, ctx->st,
            ctx->tiles_x, ctx->ntiles, (int)ctx->np, ctx->order, BatchTables{});
        break;
    default:
        k_p2g_grid<64, false, MUSL2, ROLES><<<grid, block, 0, ctx->stream>>>(ctx->dP,
            ctx->cnt[cur], ctx->bkt[cur], ctx->cnt[nxt], ctx->nodes, ctx->tile_zero, ctx->mask, ctx->g, ctx->st,
            ctx->tiles_x, ctx->ntiles, (int)ctx->np, ctx->order, BatchTables{});
        break;
    }
}

template <bool MUSL2>
static void p2g_kernel_launch(mpm_hip_ctx *ctx)
{
    if (ctx->p2g_roles == 3) {
        p2g_kernel_launch_roles<MUSL2, 3>(ctx);
    } else {
        p2g_kernel_launch_roles<MUSL2, 2>(ctx);
    }
}

static int launch_p2g(mpm_hip_ctx *ctx)
{
    p2g_kernel_launch<false>(ctx);
    HIP_TRY(hipGetLastError());
    return MPM_HIP_OK;
}

static int launch_nonlocal(mpm_hip_ctx *ctx, int bins)
{
    const size_t nb = ctx->np ? (ctx->np + kG2PThreads - 1) / kG2PThreads : 1;
    const dim3 ag(ctx->nl_ntiles), ab(kNlThreads);
    switch (ctx->g.cap) {
    case 16:
        k_nl_assemble<16><<<ag, ab, 0, ctx->stream>>>(ctx->dP, ctx->cnt[bins], ctx->bkt[bins],
            ctx->nodes, ctx->nl, ctx->g, ctx->nl_tiles_x, ctx->nl_ntiles);
        break;
    case 32:
        k_nl_assemble<32><<<ag, ab, 0, ctx->stream>>>(ctx->dP, ctx->cnt[bins], ctx->bkt[bins],
            ctx->nodes, ctx->nl, ctx->g, ctx->nl_tiles_x, ctx->nl_ntiles);
        break;
    default:
        k_nl_assemble<64><<<ag, ab, 0, ctx->stream>>>(ctx->dP, ctx->cnt[bins], ctx->bkt[bins],
            ctx->nodes, ctx->nl, ctx->g, ctx->nl_tiles_x, ctx->nl_ntiles);
        break;
    }
    HIP_TRY(hipGetLastError());
    double tol = ctx->cfg.cg_rel_tol;
    int max_iter = ctx->cfg.cg_max_iter;
    void *args[] = {&ctx->nl, &tol, &max_iter};
    HIP_TRY(hipLaunchCooperativeKernel((const void *)k_nl_cg, dim3(ctx->cg_blocks),
                                       dim3(kCgThreads), args, 0, ctx->stream));
    k_nl_stress<<<dim3((unsigned)nb), dim3(kG2PThreads), 0, ctx->stream>>>(
        ctx->dP, ctx->nl.x, ctx->g, ctx->mt, ctx->st, (int)ctx->np);
    HIP_TRY(hipGetLastError());
    return MPM_HIP_OK;
}

static int launch_g2p(mpm_hip_ctx *ctx)
{
    const int nxt = 1 - ctx->cur;
    const int bins = ctx->cur;   // the step's own bins (the assembly reads them)
    const size_t nb = ctx->np ? (ctx->np + kG2PThreads - 1) / kG2PThreads : 1;
    int mode = 0;
    if (ctx->cfg.flags & (MPM_HIP_STORE_AUX | MPM_HIP_HOST_MATERIAL)) mode |= kModeStoreAux;
    if (ctx->cfg.flags & MPM_HIP_HOST_MATERIAL) mode |= kModeHostMaterial;
    if (ctx->nonlocal) mode |= kModeNonlocal | kModeStoreAux;
    if (ctx->cfg.material_kind == MPM_MAT_ISOLIN && !(mode & kModeHostMaterial)) {
        mode |= kModeIsolin;
    }
    if (ctx->cfg.material_kind == MPM_MAT_DP_RI_MT && !(mode & kModeHostMaterial)) {
        mode |= kModeDpRiMt;
    }
    if (ctx->cfg.scheme == MPM_SCHEME_MUSL) {
        // particle update; second mapping of the momentum; strain rate,
        // volume, stress and rebinning from the re-mapped velocities
        k_g2p_stress<false, -1><<<dim3((unsigned)nb), dim3(kG2PThreads), 0, ctx->stream>>>(
            ctx->dP, ctx->nodes, ctx->cnt[nxt], ctx->bkt[nxt], ctx->g, ctx->mt, ctx->st,
            (int)ctx->np, mode | kModeMuslMove, BatchTables{});
        p2g_kernel_launch<true>(ctx);
        mode |= kModeMuslFinish;
    }
    if (mode == 0) {
        k_g2p_stress<false, 0><<<dim3((unsigned)nb), dim3(kG2PThreads), 0, ctx->stream>>>(
            ctx->dP, ctx->nodes, ctx->cnt[nxt], ctx->bkt[nxt], ctx->g, ctx->mt, ctx->st,
            (int)ctx->np, 0, BatchTables{});
    } else {
        k_g2p_stress<false, -1><<<dim3((unsigned)nb), dim3(kG2PThreads), 0, ctx->stream>>>(
            ctx->dP, ctx->nodes, ctx->cnt[nxt], ctx->bkt[nxt], ctx->g, ctx->mt, ctx->st,
            (int)ctx->np, mode, BatchTables{});
    }
    HIP_TRY(hipGetLastError());
    ctx->cur = nxt;
    if (ctx->nonlocal) {
        return launch_nonlocal(ctx, bins);
    }
    return MPM_HIP_OK;
}

// ---------------------------------------------------------------------------
// Ensembles: members concatenated in one set of arrays, stepped by shared
// launches (the kernels' BATCH instantiations) replayed from a HIP graph.
// ---------------------------------------------------------------------------

constexpr int kGraphSteps = 8;   // steps per graph replay (even: bins return to
                                 // the parity they had at capture)

struct mpm_hip_batch {
    int device;
    int nmem;
    int cap;
    hipStream_t stream;
    MemberInfo *hmem;          // host copy of the member table
    mpm_hip_config *cfgs;
    unsigned char *have;       // particles set, per member
    MemberInfo *dmem;
    int2 *tile_order;          // P2G workgroup -> (tile, member), -1 tile: spare
    int2 *light_tiles;         // k_p2g_light's tiles: (tile, member)
    int *block_member;
    int *htile_member;         // host: tile -> member
    int *tile_load;            // host: particles in each tile's cell region
    int2 *horder, *hlight;     // host copies of tile_order and light_tiles
    int max_p2g_blocks;        // tile_order capacity
    int n_light;               // light tiles
    int *fix_count;            // [2] per step parity: light tiles listed for k_p2g_fix
    int2 *fix_list;            // [2][light tiles]
    bool order_dirty;          // tile_load changed since tile_order was built
    DevState *st;
    double *pfield[MPM_F_EXX];
    Particles P;
    Particles *dP;             // device copy of P
    NodeRec *nodes;
    uint8_t *tile_zero;        // per P2G tile: its node records are all zero
    int *cnt[2], *bkt[2];
    size_t np_total, nn_total, ncells_total;
    int ntiles, p2g_blocks, g2p_blocks;
    int cur;
    hipGraphExec_t exec;
    int graph_cur;             // bin parity the graph was captured at
};

static BatchTables batch_tables(const mpm_hip_batch *b)
{
    BatchTables t;
    t.mem = b->dmem;
    t.tile_order = b->tile_order;
    t.light_tiles = b->light_tiles;
    t.block_member = b->block_member;
    t.st = b->st;
    return t;
}

template <int CAP>
static void batch_p2g_kernels(mpm_hip_batch *b, int cur)
{
    const int nxt = 1 - cur;
    const BatchTables t = batch_tables(b);
    const GridParams g0{};
    const int np = (int)b->np_total;
    if (b->p2g_blocks) {
        k_p2g_grid<CAP, true, false, 2><<<dim3(b->p2g_blocks), dim3(2 * kTileNodes), 0, b->stream>>>(
            b->dP, b->cnt[cur], b->bkt[cur], b->cnt[nxt], b->nodes, b->tile_zero, nullptr, g0,
            nullptr, 0, b->ntiles, np, TileOrder{}, t);
    }
    if (b->n_light) {
        const int blocks = (b->n_light + kLightGroup - 1) / kLightGroup;
        int *count = b->fix_count + cur;
        int2 *list = b->fix_list + (size_t)cur * b->ntiles;
        k_p2g_light<<<dim3(blocks), dim3(256), 0, b->stream>>>(
            b->cnt[cur], b->cnt[nxt], b->tile_zero, b->n_light, t, count, list);
        const int fix_blocks = std::min(b->n_light, kFixBlocks);
        k_p2g_fix<CAP><<<dim3(fix_blocks), dim3(2 * kTileNodes), 0, b->stream>>>(
            b->dP, b->cnt[cur], b->bkt[cur], b->cnt[nxt], b->nodes, b->tile_zero, np, count, list,
            b->fix_count + nxt, t);
    }
}

static int batch_launch_p2g(mpm_hip_batch *b, int cur)
{
    switch (b->cap) {
    case 16:
        batch_p2g_kernels<16>(b, cur);
        break;
    case 32:
        batch_p2g_kernels<32>(b, cur);
        break;
    default:
        batch_p2g_kernels<64>(b, cur);
        break;
    }
    HIP_TRY(hipGetLastError());
    return MPM_HIP_OK;
}

static int batch_launch_g2p(mpm_hip_batch *b, int cur)
{
    const int nxt = 1 - cur;
    const GridParams g0{};
    const Material m0{};
    k_g2p_stress<true, 0><<<dim3(b->g2p_blocks), dim3(kG2PThreads), 0, b->stream>>>(
        b->dP, b->nodes, b->cnt[nxt], b->bkt[nxt], g0, m0, nullptr, 0, 0, batch_tables(b));
    HIP_TRY(hipGetLastError());
    return MPM_HIP_OK;
}

static int batch_eager_step(mpm_hip_batch *b)
{
    int rc = batch_launch_p2g(b, b->cur);
    if (!rc) rc = batch_launch_g2p(b, b->cur);
    if (!rc) b->cur = 1 - b->cur;
    return rc;
}

static void batch_drop_graph(mpm_hip_batch *b)
{
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
}

// Captures kGraphSteps steps starting at the current bin parity.
static int batch_capture(mpm_hip_batch *b)
{
    batch_drop_graph(b);
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = MPM_HIP_OK;
    int c = b->cur;
    for (int k = 0; k < kGraphSteps && !rc; ++k) {
        rc = batch_launch_p2g(b, c);
        if (!rc) rc = batch_launch_g2p(b, c);
        c = 1 - c;
    }
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(b->stream, &graph);
    if (rc || end != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc ? rc : MPM_HIP_ERR_DEVICE;
    }
    const hipError_t inst = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (inst != hipSuccess) {
        b->exec = nullptr;
        return MPM_HIP_ERR_DEVICE;
    }
    b->graph_cur = b->cur;
    return MPM_HIP_OK;
}

static int batch_rebin(mpm_hip_batch *b, int init)
{
    HIP_TRY(hipMemsetAsync(b->cnt[0], 0, sizeof(int) * b->ncells_total, b->stream));
    HIP_TRY(hipMemsetAsync(b->cnt[1], 0, sizeof(int) * b->ncells_total, b->stream));
    HIP_TRY(hipMemsetAsync(b->fix_count, 0, 2 * sizeof(int), b->stream));
    b->cur = 0;
    const GridParams g0{};
    k_bin_particles<true><<<dim3(b->g2p_blocks), dim3(kG2PThreads), 0, b->stream>>>(
        b->dP, b->cnt[0], b->bkt[0], g0, nullptr, 0, init, batch_tables(b));
    HIP_TRY(hipGetLastError());
    return MPM_HIP_OK;
}

static int rebin(mpm_hip_ctx *ctx, int init)
{
    HIP_TRY(hipMemsetAsync(ctx->cnt[0], 0, sizeof(int) * ctx->g.ncells, ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->cnt[1], 0, sizeof(int) * ctx->g.ncells, ctx->stream));
    ctx->cur = 0;
    if (ctx->np) {
        const size_t nb = (ctx->np + kG2PThreads - 1) / kG2PThreads;
        k_bin_particles<false><<<dim3((unsigned)nb), dim3(kG2PThreads), 0, ctx->stream>>>(
            ctx->dP, ctx->cnt[0], ctx->bkt[0], ctx->g, ctx->st, (int)ctx->np, init,
            BatchTables{});
        HIP_TRY(hipGetLastError());
    }
    return MPM_HIP_OK;
}

// Grid, boundary and material constants of one job (mpm_init,
// process_usl.c:68-280; bc_validate of the boundary plugins; material_init).
static int params_from_config(const mpm_hip_config *cfg, int cap, GridParams &g, Material &mt)
{
    memset(&g, 0, sizeof g);
    memset(&mt, 0, sizeof mt);
    g.N = cfg->N;
    g.Ne = cfg->N - 1;
    g.ncells = g.Ne * g.Ne;
    g.cap = cap;
    g.h = cfg->h;
    g.inv_h = 1.0 / cfg->h;
    {
        int ex = 0;
        g.h_pow2 = frexp(cfg->h, &ex) == 0.5;
    }
    g.dt = cfg->dt;
    g.g_mag = cfg->g_mag;
    g.ramp = cfg->ramp_time;
    g.bc_kind = cfg->bc_kind;
    if (cfg->bc_kind == MPM_BC_BOX) {
        // bc_validate (box.c:42-96)
        const double x0 = cfg->bc_fp64[0], y0 = cfg->bc_fp64[1];
        const double w = cfg->bc_fp64[2], hh = cfg->bc_fp64[3];
        if (x0 < 0 || y0 < 0 || x0 > 1.0 || y0 > 1.0 || w < 0 || hh < 0
            || (x0 + w) > 1.0 || (y0 + hh) > 1.0) {
            return MPM_HIP_ERR_INVALID;
        }
        const size_t nm1 = (size_t)cfg->N - 1;
        g.left_col = (int)(size_t)(x0 * nm1);
        g.right_col = (int)(size_t)((x0 + w) * nm1);
        g.bottom_row = (int)(size_t)(y0 * nm1);
        g.top_row = (int)(size_t)((y0 + hh) * nm1);
        g.left_fric = cfg->bc_int[0];
        g.bottom_fric = cfg->bc_int[1];
        g.right_fric = cfg->bc_int[2];
        g.top_fric = cfg->bc_int[3];
    } else if (cfg->bc_kind == MPM_BC_SQUARE_SILO) {
        // bc_validate (square_silo.c:32-48)
        if (!(cfg->bc_fp64[2] >= 0.0 && cfg->bc_fp64[2] <= 1.0)) {
            return MPM_HIP_ERR_INVALID;
        }
        g.hole_radius = cfg->bc_fp64[0];
        g.open_time = cfg->bc_fp64[1];
        g.wall_col = (int)(((size_t)cfg->N - 1) * cfg->bc_fp64[2]);
    } else if (cfg->bc_kind != MPM_BC_NONE && cfg->bc_kind != MPM_BC_MASK) {
        return MPM_HIP_ERR_INVALID;
    }

    // material_init (g_local_mu2.c:139-142) and calculate_stress_threaded's
    // lambda (g_local_mu2.c:187)
    const double E = cfg->material.E, nu = cfg->material.nu;
    mt.G = E / (2.0 * (1.0 + nu));
    const double K = E / (3.0 * (1.0 - 2 * nu));
    mt.two_G = 2.0 * mt.G;
    mt.lambda = K - 2.0 * mt.G / 3.0;
    mt.mu_s = cfg->material.mu_s;
    mt.mu_2 = cfg->material.mu_2;
    mt.I_0 = cfg->material.I_0;
    mt.grains_d = cfg->material.grains_d;
    mt.grains_rho = cfg->material.grains_rho;
    mt.rho_crit = cfg->material.rho_crit;
    mt.nl_E = cfg->nonlocal.E;
    mt.nl_nu = cfg->nonlocal.nu;
    mt.nl_mu_s = cfg->nonlocal.mu_s;
    mt.nl_A = cfg->nonlocal.A;
    mt.nl_B = cfg->nonlocal.B;
    mt.nl_d = cfg->nonlocal.d;
    mt.nl_rho = cfg->nonlocal.rho;
    mt.nl_p_cap = cfg->nonlocal.p_cap;
    mt.nl_c = cfg->nonlocal.c;
    if (cfg->material_kind != MPM_MAT_G_LOCAL_MU2 && cfg->material_kind != MPM_MAT_G_NONLOCAL
        && cfg->material_kind != MPM_MAT_ISOLIN && cfg->material_kind != MPM_MAT_DP_RI_MT) {
        return MPM_HIP_ERR_INVALID;
    }
    return MPM_HIP_OK;
}

extern "C" {

void mpm_hip_default_config(mpm_hip_config *cfg)
{
    memset(cfg, 0, sizeof *cfg);
    cfg->N = 65;
    cfg->h = 1.0 / 64;
    cfg->dt = 1e-5;
    cfg->t0 = 0.0;
    cfg->loads_t0 = 0.0;
    cfg->g_mag = 9.8;
    cfg->ramp_time = 0.1;
    cfg->material.E = 1e6;
    cfg->material.nu = 0.3;
    cfg->material.mu_s = 0.3819;
    cfg->material.mu_2 = 0.6435;
    cfg->material.I_0 = (0.278 / 1.0);
    cfg->material.grains_d = (0.001 * 5);
    cfg->material.grains_rho = 2450;
    cfg->material.rho_crit = 1485.0f;
    cfg->bc_kind = MPM_BC_NONE;
    cfg->material_kind = MPM_MAT_G_LOCAL_MU2;
    cfg->nonlocal.E = 1e8;
    cfg->nonlocal.nu = 0.3;
    cfg->nonlocal.mu_s = 0.3819;
    cfg->nonlocal.A = 0.48;
    cfg->nonlocal.B = (0.9377 * 20);
    cfg->nonlocal.d = (0.01 * 5);
    cfg->nonlocal.rho = 2450;
    cfg->nonlocal.p_cap = 1e-2;
    cfg->nonlocal.c = 1e-2;
    cfg->cg_rel_tol = 1e-10;
    cfg->cg_max_iter = 20000;
    cfg->cg_max_blocks = 0;
    cfg->scheme = MPM_SCHEME_USL;
    cfg->bin_capacity = 0;
    cfg->flags = 0;
    cfg->device = 0;
}

const char *mpm_hip_status_string(int status)
{
    switch (status) {
    case MPM_HIP_OK: return "ok";
    case MPM_HIP_ERR_INVALID: return "invalid argument";
    case MPM_HIP_ERR_DEVICE: return "HIP runtime error";
    case MPM_HIP_ERR_NO_DEVICE: return "no gfx950 device";
    case MPM_HIP_ERR_BIN_OVERFLOW: return "cell bin capacity exceeded";
    case MPM_HIP_ERR_ALLOC: return "device allocation failed";
    default: return "unknown status";
    }
}

int mpm_hip_create(const mpm_hip_config *cfg, size_t num_particles, mpm_hip_ctx **out)
{
    if (!cfg || !out || cfg->N < 2 || !(cfg->h > 0.0) || !(cfg->dt > 0.0)
        || num_particles > kMaxParticles || cfg->N > 46340) {
        return MPM_HIP_ERR_INVALID;
    }
    int cap = cfg->bin_capacity ? cfg->bin_capacity : 32;
    if (cap != 16 && cap != 32 && cap != 64) {
        return MPM_HIP_ERR_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || cfg->device < 0 || cfg->device >= ndev) {
        return MPM_HIP_ERR_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) != hipSuccess
        || strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        return MPM_HIP_ERR_NO_DEVICE;
    }

    mpm_hip_ctx *ctx = new (std::nothrow) mpm_hip_ctx();
    if (!ctx) {
        return MPM_HIP_ERR_ALLOC;
    }
    ctx->cfg = *cfg;
    ctx->device = cfg->device;
    ctx->np = num_particles;
    *out = nullptr;

    GridParams &g = ctx->g;
    if (params_from_config(cfg, cap, g, ctx->mt) != MPM_HIP_OK) {
        delete ctx;
        return MPM_HIP_ERR_INVALID;
    }
    ctx->nonlocal = cfg->material_kind == MPM_MAT_G_NONLOCAL;
    if ((cfg->scheme != MPM_SCHEME_USL && cfg->scheme != MPM_SCHEME_MUSL)
        || (cfg->scheme == MPM_SCHEME_MUSL && ctx->nonlocal)) {
        delete ctx;
        return MPM_HIP_ERR_INVALID;
    }
    if (ctx->nonlocal && ((cfg->flags & MPM_HIP_HOST_MATERIAL) || !(cfg->cg_rel_tol >= 0.0)
                          || cfg->cg_max_iter < 0)) {
        delete ctx;
        return MPM_HIP_ERR_INVALID;
    }

    ctx->tiles_x = (cfg->N + kTileX - 1) / kTileX;
    ctx->ntiles = ctx->tiles_x * ((cfg->N + kTileY - 1) / kTileY);
    set_row_order(ctx);
    // three threads per node when every tile can be resident at once (two
    // 384-thread workgroups per CU): shorter chains per thread for a grid
    // that is all latency; two otherwise, for three workgroups per CU
    ctx->p2g_roles = ctx->ntiles <= 2 * prop.multiProcessorCount ? 3 : 2;

#define ALLOC_OR_FAIL(ptr, bytes)                                                   \
    do {                                                                            \
        if (hipMalloc((void **)&(ptr), (bytes) ? (bytes) : 64) != hipSuccess) {     \
            mpm_hip_destroy(ctx);                                                   \
            return MPM_HIP_ERR_ALLOC;                                               \
        }                                                                           \
    } while (0)

    if (hipSetDevice(ctx->device) != hipSuccess
        || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return MPM_HIP_ERR_DEVICE;
    }
    const size_t np = num_particles;
    const size_t nn = (size_t)cfg->N * cfg->N;
    for (int f = 0; f < MPM_F_EXX; ++f) {
        ALLOC_OR_FAIL(ctx->pfield[f], np * sizeof(double));
    }
    Particles &P = ctx->P;
    bind_fields(ctx);
    if ((cfg->flags & (MPM_HIP_STORE_AUX | MPM_HIP_HOST_MATERIAL) || ctx->nonlocal)
        && alloc_aux(ctx) != MPM_HIP_OK) {
        mpm_hip_destroy(ctx);
        return MPM_HIP_ERR_ALLOC;
    }
    ALLOC_OR_FAIL(P.elem, np * sizeof(int));
    if (cfg->scheme == MPM_SCHEME_MUSL) {
        ALLOC_OR_FAIL(P.xn, np * sizeof(double));
        ALLOC_OR_FAIL(P.yn, np * sizeof(double));
    }
    if (ctx->nonlocal) {
        ALLOC_OR_FAIL(P.elem_prev, np * sizeof(int));
        NlSystem &S = ctx->nl;
        S.N = cfg->N;
        ALLOC_OR_FAIL(S.K, 9 * nn * sizeof(double));
        ALLOC_OR_FAIL(S.f, nn * sizeof(double));
        ALLOC_OR_FAIL(S.x, nn * sizeof(double));
        ALLOC_OR_FAIL(S.r, nn * sizeof(double));
        ALLOC_OR_FAIL(S.p, nn * sizeof(double));
        ALLOC_OR_FAIL(S.u, 2 * nn * sizeof(double));
        ALLOC_OR_FAIL(S.w, 2 * nn * sizeof(double));
        ALLOC_OR_FAIL(S.s, 2 * nn * sizeof(double));
        ALLOC_OR_FAIL(S.barrier, sizeof(unsigned int));
        ALLOC_OR_FAIL(S.dinv, nn * sizeof(double));
        ALLOC_OR_FAIL(S.partials, 8 * kNlMaxBlocks * sizeof(double));
        ALLOC_OR_FAIL(S.stats, sizeof(int));
        ALLOC_OR_FAIL(S.res, sizeof(double));
        if (hipMemsetAsync(S.x, 0, nn * sizeof(double), ctx->stream) != hipSuccess
            || hipMemsetAsync(S.barrier, 0, sizeof(unsigned int), ctx->stream) != hipSuccess
            || hipMemsetAsync(S.stats, 0, sizeof(int), ctx->stream) != hipSuccess
            || hipMemsetAsync(S.res, 0, sizeof(double), ctx->stream) != hipSuccess) {
            mpm_hip_destroy(ctx);
            return MPM_HIP_ERR_DEVICE;
        }
        ctx->nl_tiles_x = (cfg->N + 15) / 16;
        ctx->nl_ntiles = ctx->nl_tiles_x * ctx->nl_tiles_x;
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)k_nl_cg,
                                                         kCgThreads, 0) != hipSuccess || occ < 1) {
            mpm_hip_destroy(ctx);
            return MPM_HIP_ERR_DEVICE;
        }
        // co-resident workgroups only (the solve synchronises the grid)
        const long tx = (cfg->N + kCgTile - 1) / kCgTile;
        long want = tx * tx;   // one workgroup per tile when they fit
        long have = (long)occ * prop.multiProcessorCount;
        long nbk = want < have ? want : have;
        if (cfg->cg_max_blocks > 0 && nbk > cfg->cg_max_blocks) nbk = cfg->cg_max_blocks;
        ctx->cg_blocks = (int)(nbk < kNlMaxBlocks ? nbk : kNlMaxBlocks);
    }
    ALLOC_OR_FAIL(P.active, np * sizeof(int));
    ALLOC_OR_FAIL(ctx->nodes, nn * sizeof(NodeRec));
    ALLOC_OR_FAIL(ctx->tile_zero, (size_t)ctx->ntiles);
    ALLOC_OR_FAIL(ctx->mask, nn);
    for (int b = 0; b < 2; ++b) {
        ALLOC_OR_FAIL(ctx->cnt[b], (size_t)g.ncells * sizeof(int));
        ALLOC_OR_FAIL(ctx->bkt[b], (size_t)g.ncells * cap * sizeof(int));
    }
    ALLOC_OR_FAIL(ctx->st, sizeof(DevState));
    ALLOC_OR_FAIL(ctx->dP, sizeof(Particles));
#undef ALLOC_OR_FAIL
    if (upload_particles(ctx->dP, P) != MPM_HIP_OK) {
        mpm_hip_destroy(ctx);
        return MPM_HIP_ERR_DEVICE;
    }

    if (hipMemsetAsync(ctx->mask, 0, nn, ctx->stream) != hipSuccess
        || hipMemsetAsync(ctx->nodes, 0, nn * sizeof(NodeRec), ctx->stream) != hipSuccess
        || hipMemsetAsync(ctx->tile_zero, 1, (size_t)ctx->ntiles, ctx->stream) != hipSuccess
        || hipMemsetAsync(P.active, 0, np * sizeof(int), ctx->stream) != hipSuccess
        || hipMemsetAsync(P.elem, 0xff, np * sizeof(int), ctx->stream) != hipSuccess) {
        mpm_hip_destroy(ctx);
        return MPM_HIP_ERR_DEVICE;
    }
    DevState s0;
    memset(&s0, 0, sizeof s0);
    s0.t = cfg->t0;
    s0.t0 = cfg->loads_t0;
    if (hipMemcpyAsync(ctx->st, &s0, sizeof s0, hipMemcpyHostToDevice, ctx->stream) != hipSuccess
        || hipStreamSynchronize(ctx->stream) != hipSuccess) {
        mpm_hip_destroy(ctx);
        return MPM_HIP_ERR_DEVICE;
    }
    *out = ctx;
    return MPM_HIP_OK;
}

int mpm_hip_set_particles(mpm_hip_ctx *ctx, const double *m, const double *v,
                          const double *x, const double *y, const double *vx,
                          const double *vy, const double *sxx, const double *sxy,
                          const double *syy)
{
    if (!ctx || (ctx->np && (!m || !v || !x || !y || !vx || !vy))) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    const size_t bytes = ctx->np * sizeof(double);
    const double *src[] = {x, y, vx, vy, m, v};
    const int dst[] = {MPM_F_X, MPM_F_Y, MPM_F_VX, MPM_F_VY, MPM_F_M, MPM_F_V};
    if (bytes) {
        for (int k = 0; k < 6; ++k) {
            HIP_TRY(hipMemcpyAsync(ctx->pfield[dst[k]], src[k], bytes, hipMemcpyHostToDevice,
                                   ctx->stream));
        }
        // material_init: T from the particle stress, T[ZZ] and state zero
        const double *s3[] = {sxx, sxy, syy};
        const int t3[] = {MPM_F_TXX, MPM_F_TXY, MPM_F_TYY};
        for (int k = 0; k < 3; ++k) {
            if (s3[k]) {
                HIP_TRY(hipMemcpyAsync(ctx->pfield[t3[k]], s3[k], bytes, hipMemcpyHostToDevice,
                                       ctx->stream));
            } else {
                HIP_TRY(hipMemsetAsync(ctx->pfield[t3[k]], 0, bytes, ctx->stream));
            }
        }
        const int zero[] = {MPM_F_TZZ, MPM_F_GAMMAP, MPM_F_GAMMADOTP, MPM_F_UX, MPM_F_UY};
        for (int k = 0; k < 5; ++k) {
            HIP_TRY(hipMemsetAsync(ctx->pfield[zero[k]], 0, bytes, ctx->stream));
        }
        for (int f = MPM_F_EXX; f < MPM_F_COUNT; ++f) {
            if (ctx->pfield[f]) {
                HIP_TRY(hipMemsetAsync(ctx->pfield[f], 0, bytes, ctx->stream));
            }
        }
    }
    ctx->g.m_finite = all_finite(m, ctx->np);
    order_tiles_by_particles(ctx, x, y);
    DevState s0;
    memset(&s0, 0, sizeof s0);
    s0.t = ctx->cfg.t0;
    s0.t0 = ctx->cfg.loads_t0;
    HIP_TRY(hipMemcpyAsync(ctx->st, &s0, sizeof s0, hipMemcpyHostToDevice, ctx->stream));
    rc = rebin(ctx, 1);
    if (rc) return rc;
    ctx->have_particles = true;
    return mpm_hip_sync(ctx);
}

int mpm_hip_set_field(mpm_hip_ctx *ctx, int field, const double *host)
{
    if (!ctx || field < 0 || field >= MPM_F_COUNT || !ctx->pfield[field]
        || (ctx->np && !host)) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    if (ctx->np) {
        HIP_TRY(hipMemcpyAsync(ctx->pfield[field], host, ctx->np * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
    }
    if (field == MPM_F_M) {
        ctx->g.m_finite = all_finite(host, ctx->np);
    }
    if (field == MPM_F_X || field == MPM_F_Y) {
        // positions changed: no dispatch hint
        set_row_order(ctx);
        rc = rebin(ctx, 0);
        if (rc) return rc;
    }
    return mpm_hip_sync(ctx);
}

int mpm_hip_get_field(mpm_hip_ctx *ctx, int field, double *host)
{
    if (!ctx || field < 0 || field >= MPM_F_COUNT || !ctx->pfield[field]
        || (ctx->np && !host)) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    if (ctx->np) {
        HIP_TRY(hipMemcpyAsync(host, ctx->pfield[field], ctx->np * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
    }
    return mpm_hip_sync(ctx);
}

int mpm_hip_get_active(mpm_hip_ctx *ctx, int *host)
{
    if (!ctx || (ctx->np && !host)) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    if (ctx->np) {
        HIP_TRY(hipMemcpyAsync(host, ctx->P.active, ctx->np * sizeof(int),
                               hipMemcpyDeviceToHost, ctx->stream));
    }
    return mpm_hip_sync(ctx);
}

int mpm_hip_get_node_field(mpm_hip_ctx *ctx, int field, double *host)
{
    if (!ctx || field < 0 || field >= MPM_N_COUNT || !host) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    const size_t nn = (size_t)ctx->g.N * ctx->g.N;
    // strided copy of one double out of each 64-byte record
    HIP_TRY(hipMemcpy2DAsync(host, sizeof(double), (const char *)ctx->nodes + field * sizeof(double),
                             sizeof(NodeRec), sizeof(double), nn, hipMemcpyDeviceToHost,
                             ctx->stream));
    return mpm_hip_sync(ctx);
}

int mpm_hip_set_bc_mask(mpm_hip_ctx *ctx, const uint8_t *mask)
{
    if (!ctx || !mask) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    const size_t nn = (size_t)ctx->g.N * ctx->g.N;
    HIP_TRY(hipMemcpyAsync(ctx->mask, mask, nn, hipMemcpyHostToDevice, ctx->stream));
    return mpm_hip_sync(ctx);
}

int mpm_hip_set_body_force(mpm_hip_ctx *ctx, double bx, double by)
{
    if (!ctx) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    const double b[2] = {bx, by};
    HIP_TRY(hipMemcpyAsync((char *)ctx->st + offsetof(DevState, bx), b, sizeof b,
                           hipMemcpyHostToDevice, ctx->stream));
    return mpm_hip_sync(ctx);
}

int mpm_hip_set_store_aux(mpm_hip_ctx *ctx, int on)
{
    if (!ctx) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    if (on) {
        rc = alloc_aux(ctx);
        if (rc) return rc;
        ctx->cfg.flags |= MPM_HIP_STORE_AUX;
    } else {
        ctx->cfg.flags &= ~MPM_HIP_STORE_AUX;
    }
    return mpm_hip_sync(ctx);
}

int mpm_hip_step(mpm_hip_ctx *ctx, int n)
{
    if (!ctx || n < 0 || !ctx->have_particles) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    for (int s = 0; s < n; ++s) {
        if ((rc = launch_p2g(ctx)) || (rc = launch_g2p(ctx))) {
            return rc;
        }
    }
    return MPM_HIP_OK;
}

int mpm_hip_sync(mpm_hip_ctx *ctx)
{
    if (!ctx) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    int status = 0;
    HIP_TRY(hipMemcpyAsync(&status, &ctx->st->status, sizeof(int), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (status & kStatusBinOverflow) {
        return MPM_HIP_ERR_BIN_OVERFLOW;
    }
    return MPM_HIP_OK;
}

int mpm_hip_get_time(mpm_hip_ctx *ctx, double *t, double *by)
{
    if (!ctx) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    DevState s;
    HIP_TRY(hipMemcpyAsync(&s, ctx->st, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (t) *t = s.t;
    if (by) *by = s.by;
    return MPM_HIP_OK;
}

int mpm_hip_get_crossings(mpm_hip_ctx *ctx, long long *count)
{
    if (!ctx || !count) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    DevState s;
    HIP_TRY(hipMemcpyAsync(&s, ctx->st, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *count = (long long)s.crossings;
    return MPM_HIP_OK;
}

int mpm_hip_get_solver_stats(mpm_hip_ctx *ctx, int *iterations, double *rel_residual)
{
    if (!ctx || !ctx->nonlocal) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    int it = 0;
    double res = 0.0;
    HIP_TRY(hipMemcpyAsync(&it, ctx->nl.stats, sizeof it, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(&res, ctx->nl.res, sizeof res, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (iterations) *iterations = it;
    if (rel_residual) *rel_residual = res;
    return MPM_HIP_OK;
}

int mpm_hip_time_steps(mpm_hip_ctx *ctx, int n, double *ms)
{
    if (!ctx || n <= 0 || !ms || !ctx->have_particles) {
        return MPM_HIP_ERR_INVALID;
    }
    int rc = use_device(ctx);
    if (rc) return rc;
    hipEvent_t *ev = (hipEvent_t *)malloc(sizeof(hipEvent_t) * (2 * (size_t)n + 1));
    if (!ev) return MPM_HIP_ERR_ALLOC;
    int created = 0;
    for (; created < 2 * n + 1; ++created) {
        if (hipEventCreate(&ev[created]) != hipSuccess) {
            rc = MPM_HIP_ERR_DEVICE;
            break;
        }
    }
    if (!rc) {
        for (int s = 0; s < n && !rc; ++s) {
            if (hipEventRecord(ev[2 * s], ctx->stream) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
            if (!rc) rc = launch_p2g(ctx);
            if (!rc && hipEventRecord(ev[2 * s + 1], ctx->stream) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
            if (!rc) rc = launch_g2p(ctx);
        }
        if (!rc && hipEventRecord(ev[2 * n], ctx->stream) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
        if (!rc && hipEventSynchronize(ev[2 * n]) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
    }
    if (!rc) {
        double p2g = 0, g2p = 0;
        float a = 0.f;
        for (int s = 0; s < n; ++s) {
            (void)hipEventElapsedTime(&a, ev[2 * s], ev[2 * s + 1]);
            p2g += a;
            (void)hipEventElapsedTime(&a, ev[2 * s + 1], ev[2 * s + 2]);
            g2p += a;
        }
        (void)hipEventElapsedTime(&a, ev[0], ev[2 * n]);
        ms[0] = a;
        ms[1] = p2g;
        ms[2] = g2p;
    }
    for (int k = 0; k < created; ++k) {
        (void)hipEventDestroy(ev[k]);
    }
    free(ev);
    if (rc) return rc;
    return mpm_hip_sync(ctx);
}

int mpm_hip_destroy(mpm_hip_ctx *ctx)
{
    if (!ctx) {
        return MPM_HIP_OK;
    }
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
    }
    for (int f = 0; f < MPM_F_COUNT; ++f) {
        if (ctx->pfield[f]) (void)hipFree(ctx->pfield[f]);
    }
    if (ctx->P.elem) (void)hipFree(ctx->P.elem);
    if (ctx->P.xn) (void)hipFree(ctx->P.xn);
    if (ctx->P.yn) (void)hipFree(ctx->P.yn);
    if (ctx->P.elem_prev) (void)hipFree(ctx->P.elem_prev);
    {
        NlSystem &S = ctx->nl;
        double *arrs[] = {S.K, S.f, S.x, S.r, S.p, S.u, S.w, S.s, S.dinv, S.partials, S.res};
        for (double *a : arrs) {
            if (a) (void)hipFree(a);
        }
        if (S.stats) (void)hipFree(S.stats);
        if (S.barrier) (void)hipFree(S.barrier);
    }
    if (ctx->P.active) (void)hipFree(ctx->P.active);
    if (ctx->dP) (void)hipFree(ctx->dP);
    if (ctx->nodes) (void)hipFree(ctx->nodes);
    if (ctx->tile_zero) (void)hipFree(ctx->tile_zero);
    if (ctx->mask) (void)hipFree(ctx->mask);
    for (int b = 0; b < 2; ++b) {
        if (ctx->cnt[b]) (void)hipFree(ctx->cnt[b]);
        if (ctx->bkt[b]) (void)hipFree(ctx->bkt[b]);
    }
    if (ctx->st) (void)hipFree(ctx->st);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MPM_HIP_OK;
}

int mpm_hip_batch_create(const mpm_hip_config *cfgs, const size_t *num_particles,
                         int num_members, mpm_hip_batch **out)
{
    if (!cfgs || !num_particles || !out || num_members <= 0) {
        return MPM_HIP_ERR_INVALID;
    }
    *out = nullptr;
    const int cap = cfgs[0].bin_capacity ? cfgs[0].bin_capacity : 32;
    if (cap != 16 && cap != 32 && cap != 64) {
        return MPM_HIP_ERR_INVALID;
    }
    const int device = cfgs[0].device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        return MPM_HIP_ERR_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess
        || strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        return MPM_HIP_ERR_NO_DEVICE;
    }
    mpm_hip_batch *b = new (std::nothrow) mpm_hip_batch();
    if (!b) return MPM_HIP_ERR_ALLOC;
    b->device = device;
    b->nmem = num_members;
    b->cap = cap;
    b->hmem = new (std::nothrow) MemberInfo[num_members]();
    b->cfgs = new (std::nothrow) mpm_hip_config[num_members];
    b->have = new (std::nothrow) unsigned char[num_members]();
    if (!b->hmem || !b->cfgs || !b->have) {
        mpm_hip_batch_destroy(b);
        return MPM_HIP_ERR_ALLOC;
    }
    // member table: particle ranges padded to whole workgroups
    size_t pbase = 0, nbase = 0, cbase = 0;
    long tbase = 0;
    for (int k = 0; k < num_members; ++k) {
        const mpm_hip_config *c = &cfgs[k];
        if (c->N < 2 || !(c->h > 0.0) || !(c->dt > 0.0) || c->N > 46340
            || num_particles[k] > (size_t)0x7fffffff || c->bc_kind == MPM_BC_MASK
            || c->material_kind != MPM_MAT_G_LOCAL_MU2 || (c->flags & MPM_HIP_HOST_MATERIAL)
            || c->scheme != MPM_SCHEME_USL
            || (c->bin_capacity && c->bin_capacity != cap)) {
            mpm_hip_batch_destroy(b);
            return MPM_HIP_ERR_INVALID;
        }
        MemberInfo &mi = b->hmem[k];
        if (params_from_config(c, cap, mi.g, mi.mt) != MPM_HIP_OK) {
            mpm_hip_batch_destroy(b);
            return MPM_HIP_ERR_INVALID;
        }
        b->cfgs[k] = *c;
        const size_t npk = num_particles[k];
        mi.np = (int)npk;
        mi.part_base = (int)pbase;
        mi.node_base = (int)nbase;
        mi.cell_base = (int)cbase;
        mi.tiles_x = (c->N + kTileX - 1) / kTileX;
        mi.tile_base = (int)tbase;
        const size_t padded = ((npk ? npk : 1) + kG2PThreads - 1) / kG2PThreads * kG2PThreads;
        pbase += padded;
        nbase += (size_t)c->N * c->N;
        cbase += (size_t)mi.g.ncells;
        tbase += (long)mi.tiles_x * ((c->N + kTileY - 1) / kTileY);
        if (pbase > kMaxParticles || nbase > (size_t)0x7fffffff
            || cbase * (size_t)64 > (size_t)0x7fffffff || tbase > 0x7fffffff / 8) {
            mpm_hip_batch_destroy(b);
            return MPM_HIP_ERR_INVALID;
        }
    }
    b->np_total = pbase;
    b->nn_total = nbase;
    b->ncells_total = cbase;
    b->ntiles = (int)tbase;
    // occupied tiles one per workgroup in whole chunk rounds (batch_build_order)
    b->max_p2g_blocks = p2g_grid_blocks(b->ntiles, kBatchTileChunk);
    b->p2g_blocks = 0;
    b->g2p_blocks = (int)(pbase / kG2PThreads);

    if (hipSetDevice(device) != hipSuccess
        || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        b->stream = nullptr;
        mpm_hip_batch_destroy(b);
        return MPM_HIP_ERR_DEVICE;
    }
#define BALLOC(ptr, bytes)                                                          \
    do {                                                                            \
        if (hipMalloc((void **)&(ptr), (bytes) ? (bytes) : 64) != hipSuccess) {     \
            mpm_hip_batch_destroy(b);                                               \
            return MPM_HIP_ERR_ALLOC;                                               \
        }                                                                           \
    } while (0)
    const size_t np = b->np_total;
    for (int f = 0; f < MPM_F_EXX; ++f) {
        BALLOC(b->pfield[f], np * sizeof(double));
    }
    Particles &P = b->P;
    double **f = b->pfield;
    P.x = f[MPM_F_X];
    P.y = f[MPM_F_Y];
    P.vx = f[MPM_F_VX];
    P.vy = f[MPM_F_VY];
    P.m = f[MPM_F_M];
    P.v = f[MPM_F_V];
    P.txx = f[MPM_F_TXX];
    P.txy = f[MPM_F_TXY];
    P.tyy = f[MPM_F_TYY];
    P.tzz = f[MPM_F_TZZ];
    P.gp = f[MPM_F_GAMMAP];
    P.gdp = f[MPM_F_GAMMADOTP];
    P.ux = f[MPM_F_UX];
    P.uy = f[MPM_F_UY];
    BALLOC(P.elem, np * sizeof(int));
    BALLOC(P.active, np * sizeof(int));
    BALLOC(b->nodes, b->nn_total * sizeof(NodeRec));
    BALLOC(b->tile_zero, (size_t)(b->ntiles ? b->ntiles : 1));
    for (int k = 0; k < 2; ++k) {
        BALLOC(b->cnt[k], b->ncells_total * sizeof(int));
        BALLOC(b->bkt[k], b->ncells_total * (size_t)cap * sizeof(int));
    }
    BALLOC(b->dmem, (size_t)num_members * sizeof(MemberInfo));
    BALLOC(b->st, (size_t)num_members * sizeof(DevState));
    BALLOC(b->tile_order, (size_t)b->max_p2g_blocks * sizeof(int2));
    BALLOC(b->light_tiles, (size_t)b->ntiles * sizeof(int2));
    BALLOC(b->fix_count, 2 * sizeof(int));
    BALLOC(b->fix_list, 2 * (size_t)b->ntiles * sizeof(int2));
    BALLOC(b->block_member, (size_t)b->g2p_blocks * sizeof(int));
    BALLOC(b->dP, sizeof(Particles));
#undef BALLOC
    if (upload_particles(b->dP, P) != MPM_HIP_OK) {
        mpm_hip_batch_destroy(b);
        return MPM_HIP_ERR_DEVICE;
    }

    b->htile_member = new (std::nothrow) int[b->ntiles ? b->ntiles : 1];
    b->tile_load = new (std::nothrow) int[b->ntiles ? b->ntiles : 1]();
    b->horder = new (std::nothrow) int2[b->max_p2g_blocks];
    b->hlight = new (std::nothrow) int2[b->ntiles ? b->ntiles : 1];
    int *tm = b->htile_member;
    int *bm = (int *)malloc(sizeof(int) * (size_t)b->g2p_blocks);
    DevState *s0 = (DevState *)calloc((size_t)num_members, sizeof(DevState));
    if (!tm || !b->tile_load || !b->horder || !b->hlight || !bm || !s0) {
        free(bm);
        free(s0);
        mpm_hip_batch_destroy(b);
        return MPM_HIP_ERR_ALLOC;
    }
    for (int k = 0; k < num_members; ++k) {
        const MemberInfo &mi = b->hmem[k];
        const int nt = mi.tiles_x * ((cfgs[k].N + kTileY - 1) / kTileY);
        for (int t = 0; t < nt; ++t) tm[mi.tile_base + t] = k;
        const int b0 = mi.part_base / kG2PThreads;
        const int nbk = (int)(((mi.np ? mi.np : 1) + kG2PThreads - 1) / kG2PThreads);
        for (int q = 0; q < nbk; ++q) bm[b0 + q] = k;
        s0[k].t = cfgs[k].t0;
        s0[k].t0 = cfgs[k].loads_t0;
    }
    b->order_dirty = true;
    hipError_t e = hipMemcpyAsync(b->block_member, bm, sizeof(int) * (size_t)b->g2p_blocks,
                                  hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->dmem, b->hmem, sizeof(MemberInfo) * (size_t)num_members,
                           hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->st, s0, sizeof(DevState) * (size_t)num_members,
                           hipMemcpyHostToDevice, b->stream);
    for (int fld = 0; fld < MPM_F_EXX && e == hipSuccess; ++fld) {
        e = hipMemsetAsync(b->pfield[fld], 0, np * sizeof(double), b->stream);
    }
    if (e == hipSuccess) e = hipMemsetAsync(P.active, 0, np * sizeof(int), b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(P.elem, 0xff, np * sizeof(int), b->stream);
    if (e == hipSuccess)
        e = hipMemsetAsync(b->nodes, 0, b->nn_total * sizeof(NodeRec), b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->tile_zero, 1, (size_t)b->ntiles, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    free(bm);
    free(s0);
    if (e != hipSuccess) {
        mpm_hip_batch_destroy(b);
        return MPM_HIP_ERR_DEVICE;
    }
    *out = b;
    return MPM_HIP_OK;
}

// The load of member k's P2G tiles (particles in each tile's 17 x 9 cell
// region: cell c lies in the regions of tile columns c / 16 and (c + 1) / 16,
// likewise rows) from its host positions.
static void batch_tile_loads(mpm_hip_batch *b, int k, const double *x, const double *y)
{
    const MemberInfo &mi = b->hmem[k];
    const int tx = mi.tiles_x, ty = (b->cfgs[k].N + kTileY - 1) / kTileY;
    int *load = b->tile_load + mi.tile_base;
    std::fill(load, load + tx * ty, 0);
    for (int p = 0; p < mi.np; ++p) {
        const double fx = std::floor(x[p] / mi.g.h), fy = std::floor(y[p] / mi.g.h);
        if (!(fx >= 0.0 && fy >= 0.0 && fx < (double)mi.g.Ne && fy < (double)mi.g.Ne)) {
            continue;
        }
        const int c = (int)fx, r = (int)fy;
        const int ta = c / kTileX, tb = std::min((c + 1) / kTileX, tx - 1);
        const int ra = r / kTileY, rb = std::min((r + 1) / kTileY, ty - 1);
        load[ra * tx + ta] += 1;
        if (tb != ta) load[ra * tx + tb] += 1;
        if (rb != ra) {
            load[rb * tx + ta] += 1;
            if (tb != ta) load[rb * tx + tb] += 1;
        }
    }
    b->order_dirty = true;
}

// The ensemble's P2G dispatch tables (BatchTables): tiles carrying at least
// half the load of their member's heaviest tile first, then the other tiles
// that hold particles, one per k_p2g_grid workgroup (chunks of
// kBatchTileChunk to the 8 XCDs in turn), so the heavy tiles start first and
// are dealt evenly to the CUs; the tiles whose regions held no particle go
// to k_p2g_light, kLightGroup per workgroup.  The order is a hint: every
// tile is in one table once, and any order gives the same node sums.
static int batch_build_order(mpm_hip_batch *b)
{
    const int nt = b->ntiles;
    int *heavy = new (std::nothrow) int[b->nmem];
    int *seq = new (std::nothrow) int[nt ? nt : 1];
    if (!heavy || !seq) {
        delete[] heavy;
        delete[] seq;
        return MPM_HIP_ERR_ALLOC;
    }
    for (int k = 0; k < b->nmem; ++k) heavy[k] = 0;
    for (int t = 0; t < nt; ++t) {
        int &h = heavy[b->htile_member[t]];
        h = std::max(h, b->tile_load[t]);
    }
    int n_occ = 0, n_light = 0;
    for (int c = 0; c < 2; ++c) {
        for (int t = 0; t < nt; ++t) {
            const int l = b->tile_load[t];
            if (l > 0 && (2 * l >= heavy[b->htile_member[t]]) == (c == 0)) seq[n_occ++] = t;
        }
    }
    for (int t = 0; t < nt; ++t) {
        if (b->tile_load[t] == 0) b->hlight[n_light++] = make_int2(t, b->htile_member[t]);
    }
    constexpr int C = kBatchTileChunk;
    const int blocks = n_occ ? p2g_grid_blocks(n_occ, C) : 0;
    for (int blk = 0; blk < blocks; ++blk) {
        const int xcd = blk & 7, slot = blk >> 3;
        const int v = ((slot / C) * 8 + xcd) * C + slot % C;
        b->horder[blk] = v < n_occ ? make_int2(seq[v], b->htile_member[seq[v]]) : make_int2(-1, -1);
    }
    delete[] heavy;
    delete[] seq;
    batch_drop_graph(b);   // the captured launches have the old grids
    b->p2g_blocks = blocks;
    b->n_light = n_light;
    if (blocks) {
        HIP_TRY(hipMemcpyAsync(b->tile_order, b->horder, sizeof(int2) * (size_t)blocks,
                               hipMemcpyHostToDevice, b->stream));
    }
    if (n_light) {
        HIP_TRY(hipMemcpyAsync(b->light_tiles, b->hlight, sizeof(int2) * (size_t)n_light,
                               hipMemcpyHostToDevice, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));   // the host copies are rewritten by the next build
    b->order_dirty = false;
    return MPM_HIP_OK;
}

int mpm_hip_batch_set_particles(mpm_hip_batch *b, int member, const double *m,
                                const double *v, const double *x, const double *y,
                                const double *vx, const double *vy, const double *sxx,
                                const double *sxy, const double *syy)
{
    if (!b || member < 0 || member >= b->nmem) {
        return MPM_HIP_ERR_INVALID;
    }
    const MemberInfo &mi = b->hmem[member];
    if (mi.np && (!m || !v || !x || !y || !vx || !vy)) {
        return MPM_HIP_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(b->device));
    const size_t bytes = (size_t)mi.np * sizeof(double);
    const size_t off = (size_t)mi.part_base;
    if (bytes) {
        const double *src[] = {x, y, vx, vy, m, v, sxx, sxy, syy};
        const int dst[] = {MPM_F_X, MPM_F_Y, MPM_F_VX, MPM_F_VY, MPM_F_M, MPM_F_V,
                           MPM_F_TXX, MPM_F_TXY, MPM_F_TYY};
        for (int k = 0; k < 9; ++k) {
            if (src[k]) {
                HIP_TRY(hipMemcpyAsync(b->pfield[dst[k]] + off, src[k], bytes,
                                       hipMemcpyHostToDevice, b->stream));
            } else {
                HIP_TRY(hipMemsetAsync(b->pfield[dst[k]] + off, 0, bytes, b->stream));
            }
        }
        const int zero[] = {MPM_F_TZZ, MPM_F_GAMMAP, MPM_F_GAMMADOTP, MPM_F_UX, MPM_F_UY};
        for (int k = 0; k < 5; ++k) {
            HIP_TRY(hipMemsetAsync(b->pfield[zero[k]] + off, 0, bytes, b->stream));
        }
    }
    DevState s0;
    memset(&s0, 0, sizeof s0);
    s0.t = b->cfgs[member].t0;
    s0.t0 = b->cfgs[member].loads_t0;
    HIP_TRY(hipMemcpyAsync(b->st + member, &s0, sizeof s0, hipMemcpyHostToDevice, b->stream));
    batch_tile_loads(b, member, x, y);
    batch_drop_graph(b);
    int rc = batch_rebin(b, member + 1);
    if (rc) return rc;
    b->have[member] = 1;
    return mpm_hip_batch_sync(b);
}

static int batch_check(mpm_hip_batch *b, int member)
{
    if (!b || member < 0 || member >= b->nmem) return MPM_HIP_ERR_INVALID;
    HIP_TRY(hipSetDevice(b->device));
    return MPM_HIP_OK;
}

int mpm_hip_batch_get_field(mpm_hip_batch *b, int member, int field, double *host)
{
    int rc = batch_check(b, member);
    if (rc) return rc;
    if (field < 0 || field >= MPM_F_EXX || (b->hmem[member].np && !host)) {
        return MPM_HIP_ERR_INVALID;
    }
    const MemberInfo &mi = b->hmem[member];
    if (mi.np) {
        HIP_TRY(hipMemcpyAsync(host, b->pfield[field] + mi.part_base,
                               (size_t)mi.np * sizeof(double), hipMemcpyDeviceToHost,
                               b->stream));
    }
    return mpm_hip_batch_sync(b);
}

int mpm_hip_batch_get_active(mpm_hip_batch *b, int member, int *host)
{
    int rc = batch_check(b, member);
    if (rc) return rc;
    const MemberInfo &mi = b->hmem[member];
    if (mi.np && !host) return MPM_HIP_ERR_INVALID;
    if (mi.np) {
        HIP_TRY(hipMemcpyAsync(host, b->P.active + mi.part_base, (size_t)mi.np * sizeof(int),
                               hipMemcpyDeviceToHost, b->stream));
    }
    return mpm_hip_batch_sync(b);
}

int mpm_hip_batch_get_time(mpm_hip_batch *b, int member, double *t, double *by)
{
    int rc = batch_check(b, member);
    if (rc) return rc;
    DevState s;
    HIP_TRY(hipMemcpyAsync(&s, b->st + member, sizeof s, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (t) *t = s.t;
    if (by) *by = s.by;
    return MPM_HIP_OK;
}

static bool batch_ready(const mpm_hip_batch *b)
{
    for (int k = 0; k < b->nmem; ++k) {
        if (!b->have[k]) return false;
    }
    return true;
}

// n steps: eager steps up to the graph's parity, whole graph replays, eager rest.
static int batch_run(mpm_hip_batch *b, int n)
{
    int rc = MPM_HIP_OK;
    if (b->order_dirty && (rc = batch_build_order(b))) {
        return rc;
    }
    if (n >= kGraphSteps) {
        if (!b->exec) {
            rc = batch_capture(b);
            if (rc) return rc;
        }
        if (b->cur != b->graph_cur) {
            rc = batch_eager_step(b);
            if (rc) return rc;
            --n;
        }
        while (n >= kGraphSteps) {
            HIP_TRY(hipGraphLaunch(b->exec, b->stream));
            n -= kGraphSteps;
        }
    }
    for (; n > 0 && !rc; --n) {
        rc = batch_eager_step(b);
    }
    return rc;
}

int mpm_hip_batch_step(mpm_hip_batch *b, int n)
{
    if (!b || n < 0 || !batch_ready(b)) return MPM_HIP_ERR_INVALID;
    HIP_TRY(hipSetDevice(b->device));
    return batch_run(b, n);
}

int mpm_hip_batch_sync(mpm_hip_batch *b)
{
    if (!b) return MPM_HIP_ERR_INVALID;
    HIP_TRY(hipSetDevice(b->device));
    DevState *s = (DevState *)malloc(sizeof(DevState) * (size_t)b->nmem);
    if (!s) return MPM_HIP_ERR_ALLOC;
    hipError_t e = hipMemcpyAsync(s, b->st, sizeof(DevState) * (size_t)b->nmem,
                                  hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    int status = 0;
    for (int k = 0; k < b->nmem && e == hipSuccess; ++k) status |= s[k].status;
    free(s);
    if (e != hipSuccess) {
        fprintf(stderr, "mpm_hip: batch sync failed: %s\n", hipGetErrorString(e));
        return MPM_HIP_ERR_DEVICE;
    }
    return (status & kStatusBinOverflow) ? MPM_HIP_ERR_BIN_OVERFLOW : MPM_HIP_OK;
}

int mpm_hip_batch_time_steps(mpm_hip_batch *b, int n, double *ms)
{
    if (!b || n <= 0 || !ms || !batch_ready(b)) return MPM_HIP_ERR_INVALID;
    HIP_TRY(hipSetDevice(b->device));
    hipEvent_t ev[4] = {};
    int rc = MPM_HIP_OK;
    for (int k = 0; k < 4; ++k) {
        if (hipEventCreate(&ev[k]) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
    }
    if (!rc && hipEventRecord(ev[0], b->stream) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
    if (!rc) rc = batch_run(b, n - 1);
    // the last step eagerly, with events around each kernel
    if (!rc && hipEventRecord(ev[1], b->stream) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
    if (!rc) rc = batch_launch_p2g(b, b->cur);
    if (!rc && hipEventRecord(ev[2], b->stream) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
    if (!rc) rc = batch_launch_g2p(b, b->cur);
    if (!rc) b->cur = 1 - b->cur;
    if (!rc && hipEventRecord(ev[3], b->stream) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
    if (!rc && hipEventSynchronize(ev[3]) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
    if (!rc) {
        float a = 0.f;
        (void)hipEventElapsedTime(&a, ev[0], ev[3]);
        ms[0] = a;
        (void)hipEventElapsedTime(&a, ev[1], ev[2]);
        ms[1] = a;
        (void)hipEventElapsedTime(&a, ev[2], ev[3]);
        ms[2] = a;
    }
    for (int k = 0; k < 4; ++k) {
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    }
    if (rc) return rc;
    return mpm_hip_batch_sync(b);
}

// Replaces the bin arrays with ones of capacity `cap` and re-bins.
static int batch_set_capacity(mpm_hip_batch *b, int cap)
{
    batch_drop_graph(b);
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (int k = 0; k < 2; ++k) {
        if (b->bkt[k]) (void)hipFree(b->bkt[k]);
        b->bkt[k] = nullptr;
        HIP_TRY(hipMalloc((void **)&b->bkt[k], b->ncells_total * (size_t)cap * sizeof(int)));
    }
    b->cap = cap;
    for (int k = 0; k < b->nmem; ++k) b->hmem[k].g.cap = cap;
    HIP_TRY(hipMemcpyAsync(b->dmem, b->hmem, sizeof(MemberInfo) * (size_t)b->nmem,
                           hipMemcpyHostToDevice, b->stream));
    return batch_rebin(b, 0);
}

int mpm_hip_batch_autotune(mpm_hip_batch *b, int trial_steps, int *chosen, double *ms)
{
    if (!b || trial_steps <= 0 || !batch_ready(b)) return MPM_HIP_ERR_INVALID;
    HIP_TRY(hipSetDevice(b->device));
    int rc = mpm_hip_batch_sync(b);
    if (rc) return rc;
    // snapshot of the state a step reads
    const size_t np = b->np_total;
    double *save = nullptr;
    int *save_i = nullptr;
    DevState *save_st = nullptr;
    HIP_TRY(hipMalloc((void **)&save, np * sizeof(double) * MPM_F_EXX));
    if (hipMalloc((void **)&save_i, np * sizeof(int) * 2) != hipSuccess
        || hipMalloc((void **)&save_st, sizeof(DevState) * (size_t)b->nmem) != hipSuccess) {
        (void)hipFree(save);
        if (save_i) (void)hipFree(save_i);
        return MPM_HIP_ERR_ALLOC;
    }
    auto copy_state = [&](bool to_save) -> hipError_t {
        hipError_t e = hipSuccess;
        for (int f = 0; f < MPM_F_EXX && e == hipSuccess; ++f) {
            double *a = save + (size_t)f * np, *p = b->pfield[f];
            e = hipMemcpyAsync(to_save ? a : p, to_save ? p : a, np * sizeof(double),
                               hipMemcpyDeviceToDevice, b->stream);
        }
        int *ia[2] = {b->P.active, b->P.elem};
        for (int k = 0; k < 2 && e == hipSuccess; ++k) {
            e = hipMemcpyAsync(to_save ? save_i + k * np : ia[k], to_save ? ia[k] : save_i + k * np,
                               np * sizeof(int), hipMemcpyDeviceToDevice, b->stream);
        }
        if (e == hipSuccess) {
            e = hipMemcpyAsync(to_save ? save_st : b->st, to_save ? b->st : save_st,
                               sizeof(DevState) * (size_t)b->nmem, hipMemcpyDeviceToDevice,
                               b->stream);
        }
        return e;
    };
    const int caps[3] = {16, 32, 64};
    double t[3] = {-1.0, -1.0, -1.0};
    int best = -1;
    rc = copy_state(true) == hipSuccess ? MPM_HIP_OK : MPM_HIP_ERR_DEVICE;
    for (int k = 0; k < 3 && !rc; ++k) {
        rc = batch_set_capacity(b, caps[k]);
        if (!rc) {
            rc = mpm_hip_batch_sync(b);   // overflow here: the cells do not fit
            if (rc == MPM_HIP_ERR_BIN_OVERFLOW) {
                rc = MPM_HIP_OK;
            } else if (!rc) {
                double tm[3];
                rc = mpm_hip_batch_time_steps(b, trial_steps, tm);
                if (rc == MPM_HIP_ERR_BIN_OVERFLOW) {
                    rc = MPM_HIP_OK;
                } else if (!rc) {
                    t[k] = tm[0] / trial_steps;
                    if (best < 0 || t[k] < t[best]) best = k;
                }
            }
        }
        if (!rc && copy_state(false) != hipSuccess) rc = MPM_HIP_ERR_DEVICE;
    }
    if (!rc) {
        rc = batch_set_capacity(b, best >= 0 ? caps[best] : 64);
    }
    if (!rc) rc = mpm_hip_batch_sync(b);
    (void)hipFree(save);
    (void)hipFree(save_i);
    (void)hipFree(save_st);
    if (rc) return rc;
    if (best < 0) return MPM_HIP_ERR_BIN_OVERFLOW;
    if (chosen) *chosen = caps[best];
    if (ms) {
        for (int k = 0; k < 3; ++k) ms[k] = t[k];
    }
    return MPM_HIP_OK;
}

int mpm_hip_batch_destroy(mpm_hip_batch *b)
{
    if (!b) return MPM_HIP_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    for (int f = 0; f < MPM_F_EXX; ++f) {
        if (b->pfield[f]) (void)hipFree(b->pfield[f]);
    }
    void *arrs[] = {b->P.elem, b->P.active, b->nodes, b->tile_zero, b->cnt[0], b->cnt[1], b->bkt[0],
                    b->bkt[1], b->dmem, b->st, b->tile_order, b->light_tiles, b->fix_count,
                    b->fix_list, b->block_member, b->dP};
    for (void *a : arrs) {
        if (a) (void)hipFree(a);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete[] b->hmem;
    delete[] b->htile_member;
    delete[] b->tile_load;
    delete[] b->horder;
    delete[] b->hlight;
    delete[] b->cfgs;
    delete[] b->have;
    delete b;
    return MPM_HIP_OK;
}

// ---------------------------------------------------------------------------
// Adjoint runs: forward with checkpoints, reverse sweep segment by segment.
// ---------------------------------------------------------------------------

}  // extern "C"

struct mpm_hip_adjoint {
    mpm_hip_ctx *ctx;          // forward state and kernels
    int K;                     // checkpoint interval
    int nsteps;                // steps of the last forward run
    int ncp;                   // checkpoints allocated
    size_t np;
    double *init;              // [MPM_F_EXX][np] state after set_particles
    int *init_active;
    DevState *init_st;
    double *cp;                // [ncp][MPM_F_EXX][np]
    int *cp_active;            // [ncp][np]
    DevState *cp_st;           // [ncp]
    double *seg;               // [K][MPM_F_EXX][np]
    int *seg_active;           // [K][np]
    DevState *seg_st;          // [K]
    // per step of a segment, kept from its forward recompute for the
    // reverse step: element indices, node records (with their tiles' zero
    // flags) and the bins of the step's start state
    int *seg_elem;             // [K][np]
    NodeRec *seg_nodes;        // [K][N * N]
    uint8_t *seg_zero;         // [K][P2G tiles]
    int *seg_cnt;              // [K][cells]
    int *seg_bkt;              // [K][cells][cap]
    Particles *seg_P;          // [K] device tables of the kept start states (the
                               // reverse kernels read them in place)
    // the context's own bins and node records (restored after a reverse sweep)
    int *own_cnt[2], *own_bkt[2];
    NodeRec *own_nodes;
    uint8_t *own_zero;
    AdjState A;
    double *lam_store;         // [MPM_F_EXX][np]
    double *acc;               // [kAdjParams]
    int g2p_blocks;
    bool have_particles, have_forward;
};

// Saves the context's particle state to (f, act, st) or loads it back.
// With `elem`, the element indices are saved or loaded too and a load keeps
// the bins as they are; without, a load re-bins the particles into the
// context's bins 0 (bins cleared in the same launch).
static int adj_copy_state(mpm_hip_adjoint *a, bool save, double *f, int *act, DevState *st,
                          int *elem = nullptr)
{
    mpm_hip_ctx *ctx = a->ctx;
    StateCopy c;
    memset(&c, 0, sizeof c);
    for (int k = 0; k < MPM_F_EXX; ++k) {
        double *dev = ctx->pfield[k], *buf = f + (size_t)k * (a->np ? a->np : 1);
        c.src[k] = save ? dev : buf;
        c.dst[k] = save ? buf : dev;
    }
    c.asrc = save ? ctx->P.active : act;
    c.adst = save ? act : ctx->P.active;
    c.ssrc = save ? ctx->st : st;
    c.sdst = save ? st : ctx->st;
    if (elem) {
        c.esrc = save ? ctx->P.elem : elem;
        c.edst = save ? elem : ctx->P.elem;
    }
    const bool rebin = !save && !elem;
    c.np = (int)a->np;
    if (rebin) {
        c.cnt0 = ctx->cnt[0];
        c.cnt1 = ctx->cnt[1];
        c.ncells = ctx->g.ncells;
    }
    const int n = c.np > c.ncells ? c.np : c.ncells;
    k_copy_state<<<dim3((unsigned)((n > 0 ? n : 1) + 255) / 256), dim3(256), 0, ctx->stream>>>(c);
    HIP_TRY(hipGetLastError());
    if (rebin) {
        ctx->cur = 0;
        if (a->np) {
            k_bin_particles<false><<<dim3((unsigned)a->g2p_blocks), dim3(kG2PThreads), 0,
                                     ctx->stream>>>(ctx->dP, ctx->cnt[0], ctx->bkt[0], ctx->g,
                                                    ctx->st, (int)a->np, 0, BatchTables{});
            HIP_TRY(hipGetLastError());
        }
    }
    return MPM_HIP_OK;
}

// One reverse step for step k of the segment: its start state, node records
// and bins are those kept by the segment's forward recompute.
static int adj_reverse_step(mpm_hip_adjoint *a, int k)
{
    mpm_hip_ctx *ctx = a->ctx;
    const Particles *Pk = a->seg_P + k;
    const DevState *stk = a->seg_st + k;
    const int np = (int)a->np;
    const int nn = ctx->g.N * ctx->g.N;
    const NodeRec *nodes = a->seg_nodes + (size_t)k * nn;
    int *cnt = a->seg_cnt + (size_t)k * ctx->g.ncells;
    int *bkt = a->seg_bkt + (size_t)k * ctx->g.ncells * ctx->g.cap;
    k_adj_g2p<<<dim3(a->g2p_blocks), dim3(kG2PThreads), 0, ctx->stream>>>(
        Pk, nodes, a->A, ctx->g, ctx->mt, np);
    k_adj_cells<<<dim3((4 * ctx->g.ncells + 255) / 256), dim3(256), 0, ctx->stream>>>(
        Pk, cnt, bkt, a->A, ctx->g, np);
    k_adj_nodes<<<dim3((nn + 255) / 256), dim3(256), 0, ctx->stream>>>(
        nodes, ctx->mask, a->A, ctx->g, stk);
    k_adj_p2g<<<dim3(a->g2p_blocks), dim3(kG2PThreads), 0, ctx->stream>>>(
        Pk, a->A, ctx->g, stk, np);
    k_adj_reduce<<<dim3(1), dim3(256), 0, ctx->stream>>>(a->A.partials, a->g2p_blocks, a->acc);
    HIP_TRY(hipGetLastError());
    return MPM_HIP_OK;
}

// The reverse sweep, segment by segment from the last: each segment's
// steps are recomputed from its checkpoint, keeping every step's start
// state, node records and bins (the context's bin and node pointers are
// pointed at the segment's arrays meanwhile), then reversed from the last.
static int adj_reverse_segments(mpm_hip_adjoint *a)
{
    mpm_hip_ctx *ctx = a->ctx;
    const size_t np = a->np ? a->np : 1;
    const size_t nn = (size_t)ctx->g.N * ctx->g.N;
    const size_t nc = (size_t)ctx->g.ncells, cap = (size_t)ctx->g.cap;
    int rc = MPM_HIP_OK;
    const int nseg = (a->nsteps + a->K - 1) / a->K;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int c0 = sg * a->K;
        const int len = (a->nsteps - c0) < a->K ? (a->nsteps - c0) : a->K;
        // the checkpoint, binned into step 0's bins
        ctx->cnt[0] = a->seg_cnt;
        ctx->bkt[0] = a->seg_bkt;
        ctx->cnt[1] = a->own_cnt[1];
        ctx->bkt[1] = a->own_bkt[1];
        rc = adj_copy_state(a, false, a->cp + (size_t)sg * MPM_F_EXX * np,
                            a->cp_active + (size_t)sg * np, a->cp_st + sg);
        if (rc) return rc;
        for (int k = 0; k < len; ++k) {
            rc = adj_copy_state(a, true, a->seg + (size_t)k * MPM_F_EXX * np,
                                a->seg_active + (size_t)k * np, a->seg_st + k,
                                a->seg_elem + (size_t)k * np);
            if (rc) return rc;
            // step k reads bins k and fills bins k + 1 (the context's own
            // for the last step, whose move is not needed)
            const int cur = ctx->cur, nxt = 1 - cur;
            ctx->cnt[cur] = a->seg_cnt + (size_t)k * nc;
            ctx->bkt[cur] = a->seg_bkt + (size_t)k * nc * cap;
            ctx->cnt[nxt] = k + 1 < len ? a->seg_cnt + (size_t)(k + 1) * nc : a->own_cnt[1];
            ctx->bkt[nxt] = k + 1 < len ? a->seg_bkt + (size_t)(k + 1) * nc * cap : a->own_bkt[1];
            ctx->nodes = a->seg_nodes + (size_t)k * nn;
            ctx->tile_zero = a->seg_zero + (size_t)k * ctx->ntiles;
            if ((rc = launch_p2g(ctx))) return rc;
            if (k + 1 < len && (rc = launch_g2p(ctx))) return rc;
        }
        for (int k = len - 1; k >= 0; --k) {
            rc = adj_reverse_step(a, k);
            if (rc) return rc;
        }
    }
    return MPM_HIP_OK;
}

extern "C" {

int mpm_hip_adjoint_create(const mpm_hip_config *cfg, size_t num_particles,
                           int checkpoint_every, mpm_hip_adjoint **out)
{
    if (!cfg || !out || checkpoint_every <= 0 || cfg->material_kind != MPM_MAT_G_LOCAL_MU2
        || cfg->N > 23170   // four lanes per cell stay within int range
        || cfg->scheme != MPM_SCHEME_USL
        || (cfg->flags & MPM_HIP_HOST_MATERIAL)) {
        return MPM_HIP_ERR_INVALID;
    }
    *out = nullptr;
    mpm_hip_adjoint *a = new (std::nothrow) mpm_hip_adjoint();
    if (!a) return MPM_HIP_ERR_ALLOC;
    int rc = mpm_hip_create(cfg, num_particles, &a->ctx);
    if (rc) {
        delete a;
        return rc;
    }
    a->K = checkpoint_every;
    a->np = num_particles;
    for (int k = 0; k < 2; ++k) {
        a->own_cnt[k] = a->ctx->cnt[k];
        a->own_bkt[k] = a->ctx->bkt[k];
    }
    a->own_nodes = a->ctx->nodes;
    a->own_zero = a->ctx->tile_zero;
    const size_t np = num_particles ? num_particles : 1;
    const size_t nn = (size_t)cfg->N * cfg->N;
    a->g2p_blocks = (int)((np + kG2PThreads - 1) / kG2PThreads);
#define AALLOC(ptr, bytes)                                                          \
    do {                                                                            \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) {                    \
            mpm_hip_adjoint_destroy(a);                                             \
            return MPM_HIP_ERR_ALLOC;                                               \
        }                                                                           \
    } while (0)
    AALLOC(a->init, np * sizeof(double) * MPM_F_EXX);
    AALLOC(a->init_active, np * sizeof(int));
    AALLOC(a->init_st, sizeof(DevState));
    AALLOC(a->seg, np * sizeof(double) * MPM_F_EXX * (size_t)a->K);
    AALLOC(a->seg_active, np * sizeof(int) * (size_t)a->K);
    AALLOC(a->seg_st, sizeof(DevState) * (size_t)a->K);
    AALLOC(a->seg_elem, np * sizeof(int) * (size_t)a->K);
    AALLOC(a->seg_nodes, nn * sizeof(NodeRec) * (size_t)a->K);
    AALLOC(a->seg_zero, (size_t)a->ctx->ntiles * a->K);
    AALLOC(a->seg_cnt, (size_t)a->ctx->g.ncells * sizeof(int) * a->K);
    AALLOC(a->seg_bkt, (size_t)a->ctx->g.ncells * a->ctx->g.cap * sizeof(int) * a->K);
    AALLOC(a->lam_store, np * sizeof(double) * MPM_F_EXX);
    for (int f = 0; f < MPM_F_EXX; ++f) a->A.lam[f] = a->lam_store + (size_t)f * np;
    AALLOC(a->A.coef, np * sizeof(double) * kAdjCoef);
    AALLOC(a->A.nadj, nn * sizeof(double) * 5);
    AALLOC(a->A.cellsum, (size_t)a->ctx->g.ncells * sizeof(double) * 24);
    AALLOC(a->seg_P, sizeof(Particles) * (size_t)a->K);
    AALLOC(a->A.partials, sizeof(double) * kAdjParams * (size_t)a->g2p_blocks);
    AALLOC(a->acc, sizeof(double) * kAdjParams);
#undef AALLOC
    // no segment node records are known to be zero yet
    if (hipMemsetAsync(a->seg_zero, 0, (size_t)a->ctx->ntiles * a->K, a->ctx->stream)
        != hipSuccess) {
        mpm_hip_adjoint_destroy(a);
        return MPM_HIP_ERR_DEVICE;
    }
    // tables of the kept states: the saved fields, activity and elements
    for (int k = 0; k < a->K; ++k) {
        Particles Pk = a->ctx->P;
        double *f[MPM_F_COUNT];
        for (int q = 0; q < MPM_F_COUNT; ++q) {
            f[q] = q < MPM_F_EXX ? a->seg + ((size_t)k * MPM_F_EXX + q) * np : a->ctx->pfield[q];
        }
        bind_fields_to(Pk, f);
        Pk.active = a->seg_active + (size_t)k * np;
        Pk.elem = a->seg_elem + (size_t)k * np;
        if (upload_particles(a->seg_P + k, Pk) != MPM_HIP_OK) {
            mpm_hip_adjoint_destroy(a);
            return MPM_HIP_ERR_DEVICE;
        }
    }
    *out = a;
    return MPM_HIP_OK;
}

int mpm_hip_adjoint_set_particles(mpm_hip_adjoint *a, const double *m, const double *v,
                                  const double *x, const double *y, const double *vx,
                                  const double *vy, const double *sxx, const double *sxy,
                                  const double *syy)
{
    if (!a) return MPM_HIP_ERR_INVALID;
    int rc = mpm_hip_set_particles(a->ctx, m, v, x, y, vx, vy, sxx, sxy, syy);
    if (rc) return rc;
    rc = adj_copy_state(a, true, a->init, a->init_active, a->init_st);
    if (rc) return rc;
    a->have_particles = true;
    a->have_forward = false;
    return mpm_hip_sync(a->ctx);
}

int mpm_hip_adjoint_set_material(mpm_hip_adjoint *a, const mpm_hip_material *mat)
{
    if (!a || !mat) return MPM_HIP_ERR_INVALID;
    mpm_hip_config c = a->ctx->cfg;
    c.material = *mat;
    GridParams g;
    Material mt;
    if (params_from_config(&c, a->ctx->g.cap, g, mt) != MPM_HIP_OK) return MPM_HIP_ERR_INVALID;
    a->ctx->cfg.material = *mat;
    a->ctx->mt = mt;
    a->have_forward = false;
    return MPM_HIP_OK;
}

int mpm_hip_adjoint_forward(mpm_hip_adjoint *a, int nsteps)
{
    if (!a || nsteps < 0 || !a->have_particles) return MPM_HIP_ERR_INVALID;
    mpm_hip_ctx *ctx = a->ctx;
    int rc = use_device(ctx);
    if (rc) return rc;
    const int ncp = (nsteps + a->K - 1) / a->K;
    if (ncp > a->ncp) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (a->cp) (void)hipFree(a->cp);
        if (a->cp_active) (void)hipFree(a->cp_active);
        if (a->cp_st) (void)hipFree(a->cp_st);
        a->cp = nullptr;
        a->cp_active = nullptr;
        a->cp_st = nullptr;
        a->ncp = 0;
        const size_t np = a->np ? a->np : 1;
        if (hipMalloc((void **)&a->cp, np * sizeof(double) * MPM_F_EXX * (size_t)ncp) != hipSuccess
            || hipMalloc((void **)&a->cp_active, np * sizeof(int) * (size_t)ncp) != hipSuccess
            || hipMalloc((void **)&a->cp_st, sizeof(DevState) * (size_t)ncp) != hipSuccess) {
            return MPM_HIP_ERR_ALLOC;
        }
        a->ncp = ncp;
    }
    rc = adj_copy_state(a, false, a->init, a->init_active, a->init_st);
    if (rc) return rc;
    const size_t np = a->np ? a->np : 1;
    for (int s = 0; s < nsteps; ++s) {
        if (s % a->K == 0) {
            const int c = s / a->K;
            rc = adj_copy_state(a, true, a->cp + (size_t)c * MPM_F_EXX * np,
                                a->cp_active + (size_t)c * np, a->cp_st + c);
            if (rc) return rc;
        }
        if ((rc = launch_p2g(ctx)) || (rc = launch_g2p(ctx))) return rc;
    }
    a->nsteps = nsteps;
    a->have_forward = true;
    return mpm_hip_sync(ctx);
}

int mpm_hip_adjoint_get_field(mpm_hip_adjoint *a, int field, double *host)
{
    if (!a) return MPM_HIP_ERR_INVALID;
    return mpm_hip_get_field(a->ctx, field, host);
}

int mpm_hip_adjoint_get_active(mpm_hip_adjoint *a, int *host)
{
    if (!a) return MPM_HIP_ERR_INVALID;
    return mpm_hip_get_active(a->ctx, host);
}

int mpm_hip_adjoint_backward(mpm_hip_adjoint *a, const double *const *seed,
                             double *grad_params, double *const *grad_init)
{
    if (!a || !a->have_forward) return MPM_HIP_ERR_INVALID;
    mpm_hip_ctx *ctx = a->ctx;
    int rc = use_device(ctx);
    if (rc) return rc;
    const size_t np = a->np ? a->np : 1;
    const size_t bytes = a->np * sizeof(double);
    for (int f = 0; f < MPM_F_EXX; ++f) {
        if (seed && seed[f] && bytes) {
            HIP_TRY(hipMemcpyAsync(a->A.lam[f], seed[f], bytes, hipMemcpyHostToDevice,
                                   ctx->stream));
        } else {
            HIP_TRY(hipMemsetAsync(a->A.lam[f], 0, np * sizeof(double), ctx->stream));
        }
    }
    HIP_TRY(hipMemsetAsync(a->acc, 0, sizeof(double) * kAdjParams, ctx->stream));
    rc = adj_reverse_segments(a);
    // the context's own bins and node records again, and the initial state
    ctx->cnt[0] = a->own_cnt[0];
    ctx->cnt[1] = a->own_cnt[1];
    ctx->bkt[0] = a->own_bkt[0];
    ctx->bkt[1] = a->own_bkt[1];
    ctx->nodes = a->own_nodes;
    ctx->tile_zero = a->own_zero;
    if (rc) return rc;
    rc = adj_copy_state(a, false, a->init, a->init_active, a->init_st);
    if (rc) return rc;
    double acc[kAdjParams];
    HIP_TRY(hipMemcpyAsync(acc, a->acc, sizeof acc, hipMemcpyDeviceToHost, ctx->stream));
    if (grad_init && bytes) {
        for (int f = 0; f < MPM_F_EXX; ++f) {
            if (grad_init[f]) {
                HIP_TRY(hipMemcpyAsync(grad_init[f], a->A.lam[f], bytes, hipMemcpyDeviceToHost,
                                       ctx->stream));
            }
        }
    }
    rc = mpm_hip_sync(ctx);
    if (rc) return rc;
    if (grad_params) {
        // G = E / (2 (1 + nu)), lambda = E / (3 (1 - 2 nu)) - 2 G / 3
        const double E = ctx->cfg.material.E, nu = ctx->cfg.material.nu;
        const double dG_dE = 1.0 / (2.0 * (1.0 + nu));
        const double dK_dE = 1.0 / (3.0 * (1.0 - 2.0 * nu));
        const double dG_dnu = -E / (2.0 * (1.0 + nu) * (1.0 + nu));
        const double dK_dnu = 2.0 * E / (3.0 * (1.0 - 2.0 * nu) * (1.0 - 2.0 * nu));
        grad_params[0] = acc[0] * dG_dE + acc[1] * (dK_dE - 2.0 / 3.0 * dG_dE);
        grad_params[1] = acc[0] * dG_dnu + acc[1] * (dK_dnu - 2.0 / 3.0 * dG_dnu);
        for (int k = 2; k < kAdjParams; ++k) grad_params[k] = acc[k];
    }
    a->have_forward = false;
    return MPM_HIP_OK;
}

int mpm_hip_adjoint_destroy(mpm_hip_adjoint *a)
{
    if (!a) return MPM_HIP_OK;
    if (a->ctx) {
        (void)hipSetDevice(a->ctx->device);
        (void)hipStreamSynchronize(a->ctx->stream);
    }
    void *arrs[] = {a->init, a->init_active, a->init_st, a->cp, a->cp_active, a->cp_st,
                    a->seg, a->seg_active, a->seg_st, a->seg_elem, a->seg_nodes, a->seg_zero,
                    a->seg_cnt, a->seg_bkt, a->seg_P, a->lam_store, a->A.coef, a->A.nadj,
                    a->A.cellsum,
                    a->A.partials, a->acc};
    for (void *p : arrs) {
        if (p) (void)hipFree(p);
    }
    if (a->ctx) mpm_hip_destroy(a->ctx);
    delete a;
    return MPM_HIP_OK;
}

double mpm_hip_exp_host(double x)
{
    return mpm_exp(x);
}

int mpm_hip_eval_ops(int op, const double *a, const double *b, double *out, size_t n,
                     int device)
{
    if (op < 0 || op > 2 || !a || !out || (op == 1 && !b)) {
        return MPM_HIP_ERR_INVALID;
    }
    if (n == 0) {
        return MPM_HIP_OK;
    }
    HIP_TRY(hipSetDevice(device));
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    const size_t bytes = n * sizeof(double);
    int rc = MPM_HIP_OK;
    if (hipMalloc((void **)&da, bytes) != hipSuccess
        || hipMalloc((void **)&db, bytes) != hipSuccess
        || hipMalloc((void **)&dout, bytes) != hipSuccess) {
        rc = MPM_HIP_ERR_ALLOC;
    }
    if (!rc && (hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) != hipSuccess
                || hipMemcpy(db, b ? b : a, bytes, hipMemcpyHostToDevice) != hipSuccess)) {
        rc = MPM_HIP_ERR_DEVICE;
    }
    if (!rc) {
        k_eval_ops<<<dim3((unsigned)((n + 255) / 256)), dim3(256)>>>(op, da, db, dout, n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess
            || hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = MPM_HIP_ERR_DEVICE;
        }
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dout) (void)hipFree(dout);
    return rc;
}

}  // extern "C"
