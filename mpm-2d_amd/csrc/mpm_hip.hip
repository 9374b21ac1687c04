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
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

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
constexpr int kStageMax = 704;     // staged records per batch (64 B each, incl. one
                                   // padding record per cell)
constexpr int kTileNodes = kTileX * kTileY;
constexpr int kP2GThreads = 2 * kTileNodes;   // two threads per node
constexpr int kG2PThreads = 256;

enum : int { kStatusBinOverflow = 1 };
enum : int { kModeStoreAux = 1, kModeHostMaterial = 2 };

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
};

struct Particles {
    double *x, *y, *vx, *vy, *m, *v;
    double *txx, *txy, *tyy, *tzz, *gp, *gdp, *ux, *uy;
    double *exx, *exy, *eyy, *wxy, *ax, *ay, *xl, *yl;   // auxiliary
    int *elem;          // element of the next step, -1 outside the grid
    int *active;
};

struct alignas(64) NodeRec {
    double m, mx, my, fx, fy, vx, vy, pad;
};

// ---------------------------------------------------------------------------
// Point operations shared by both kernels.
// ---------------------------------------------------------------------------

// WHICH_ELEMENT4 (process_usl.c:43-44).  An index at or beyond the element
// count (an ulp-level rounding case the reference reads out of bounds for)
// is treated as outside.
__device__ __forceinline__ int which_element(double x, double y, const GridParams &g)
{
    if (x < 1.0 && x >= 0.0 && y < 1.0 && y >= 0.0) {
        double e = floor(x / g.h) + floor(y / g.h) * (double)g.Ne;
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
    const double xl = snap_local((x - xn) / g.h);
    const double yl = snap_local((y - yn) / g.h);
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

// Appends particle p to the bin of cell `key` when `valid`.  Lanes of a wave
// that share a cell take their slots with one atomic per distinct cell (the
// slot order inside a bin is irrelevant: readers sort the bin by index).
// Must be reached by every lane of the wave that has not exited.
__device__ __forceinline__ void bin_insert(bool valid, int key, int p, int *cnt, int *bkt,
                                           int cap, int *status)
{
    unsigned long long pending = __ballot(valid);
    const unsigned long long lane_bit = 1ull << __lane_id();
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const int lkey = __shfl(key, leader);
        const unsigned long long group = __ballot(valid && key == lkey) & pending;
        int base = 0;
        if (lane_bit == (1ull << leader)) {
            base = atomicAdd(&cnt[lkey], __popcll(group));
        }
        base = __shfl(base, leader);
        if (group & lane_bit) {
            const int r = base + __popcll(group & (lane_bit - 1));
            if (r < cap) {
                bkt[(size_t)lkey * cap + r] = p;
            } else {
                atomicOr(status, kStatusBinOverflow);
            }
        }
        pending &= ~group;
    }
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
// particles outside the grid (process_usl.c:236-244).
__global__ __launch_bounds__(kG2PThreads) void k_bin_particles(
    Particles P, int *__restrict__ cnt, int *__restrict__ bkt, GridParams g,
    DevState *__restrict__ st, int np, int init)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
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

template <int CAP>
__global__ __launch_bounds__(kP2GThreads) void k_p2g_grid(
    Particles P, const int *__restrict__ cnt_cur, const int *__restrict__ bkt_cur,
    int *__restrict__ cnt_next, NodeRec *__restrict__ nodes,
    const uint8_t *__restrict__ mask, GridParams g, const DevState *__restrict__ st,
    int tiles_x, int ntiles)
{
    // bin entries of one colour's cells held per thread between colours
    constexpr int kListSlots = (kColourCellsMax * CAP + kP2GThreads - 1) / kP2GThreads;
    constexpr int kStageSlots = (kStageMax + kP2GThreads - 1) / kP2GThreads;
    // bins are stored CAP + 1 apart so the per-cell sort threads hit
    // different LDS banks
    constexpr int kLstStride = CAP + 1;

    __shared__ int s_cnt[kRegCells];
    __shared__ int s_lst[kColourCellsMax * kLstStride];
    __shared__ int s_off[kColourCellsMax + 1];
    __shared__ unsigned char s_owner[kColourCellsMax * kLstStride];   // staged slot -> cell
    // staged particles of one colour, paired for 16-byte LDS reads
    __shared__ double2 s_xy[kStageMax];     // local coordinates xl, yl
    __shared__ double2 s_msy[kStageMax];    // m, -sxy * v
    __shared__ double2 s_pxs[kStageMax];    // vx * m, -sxx * v
    __shared__ double2 s_pys[kStageMax];    // vy * m, -syy * v

    // Consecutive tiles go to the same XCD (workgroups are dealt to the 8
    // XCDs round-robin), so neighbouring tiles share that XCD's L2.
    const int per_xcd = gridDim.x >> 3;
    const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (tile >= ntiles) {
        return;
    }
    const int tx0 = (tile % tiles_x) * kTileX;
    const int ty0 = (tile / tiles_x) * kTileY;
    const int cx0 = tx0 - 1, cy0 = ty0 - 1;   // staged cell region origin
    const int tid = threadIdx.x;
    // Two threads per node: role 0 folds the m, mx, fx chains, role 1 the
    // my, fy chains (each chain is one node field, summed in the reference's
    // order by exactly one thread).
    const int role = tid / kTileNodes;
    const int nt = tid % kTileNodes;

    // This tile owns cells (tx0.., ty0..): clear their next-step bin counts.
    if (role == 0) {
        const int cc = tx0 + (nt % kTileX), rr = ty0 + (nt / kTileX);
        if (cc < g.Ne && rr < g.Ne) {
            cnt_next[rr * g.Ne + cc] = 0;
        }
    }
    for (int cl = tid; cl < kRegCells; cl += kP2GThreads) {
        const int cc = cx0 + cl % kRegX, rr = cy0 + cl / kRegX;
        int n = 0;
        if (cc >= 0 && rr >= 0 && cc < g.Ne && rr < g.Ne) {
            n = cnt_cur[rr * g.Ne + cc];
            n = n < CAP ? n : CAP;
        }
        s_cnt[cl] = n;
    }
    __syncthreads();

    // Colour c's cells: region column a = 2*ax + ca holds cell column cx0 + a
    // (cx0 is odd, so colour parity pc <-> ca = 1 - pc); likewise rows.
    auto ncx_of = [](int c) { return (c & 1) ? (kRegX + 1) / 2 : kRegX / 2; };
    auto ncy_of = [](int c) { return (c >> 1) ? (kRegY + 1) / 2 : kRegY / 2; };

    // Issues the loads of colour c's bin entries (all slots at once).
    int lst_reg[kListSlots];
    auto fetch_lists = [&](int c) {
        const int ncx = ncx_of(c), nq = ncx * ncy_of(c);
        const int ca = (c & 1) ? 0 : 1, cb = (c >> 1) ? 0 : 1;
#pragma unroll
        for (int u = 0; u < kListSlots; ++u) {
            const int idx = tid + u * kP2GThreads;
            const int q = idx / CAP, k = idx % CAP;
            int v = -1;
            if (q < nq) {
                const int a = 2 * (q % ncx) + ca, b = 2 * (q / ncx) + cb;
                if (k < s_cnt[b * kRegX + a]) {
                    v = bkt_cur[(size_t)((cy0 + b) * g.Ne + (cx0 + a)) * CAP + k];
                }
            }
            lst_reg[u] = v;
        }
    };
    fetch_lists(0);

    const int i = tx0 + (nt % kTileX);
    const int j = ty0 + (nt / kTileX);
    const bool live = i < g.N && j < g.N;
    const double bb = role == 0 ? st->bx : st->by;   // body force of this role's chain
    double am = 0.0, ap = 0.0, af = 0.0;             // role 0: m, mx, fx; role 1: -, my, fy

    // Colours in the reference's order; colour = 2 * (row % 2) + (col % 2).
    // Unrolled, so each colour's cell geometry is a compile-time constant.
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int pc = c & 1, pr = c >> 1;
        const int ncx = ncx_of(c), ncy = ncy_of(c);
        const int ca = pc ? 0 : 1, cb = pr ? 0 : 1;
        const int nq = ncx * ncy;
#pragma unroll
        for (int u = 0; u < kListSlots; ++u) {
            const int idx = tid + u * kP2GThreads;
            if (idx < kColourCellsMax * CAP && lst_reg[u] >= 0) {
                s_lst[(idx / CAP) * kLstStride + idx % CAP] = lst_reg[u];
            }
        }
        __syncthreads();
        if (tid < 64) {
            // staging offsets: an exclusive scan of (count + 1) over the
            // colour's cells in the first wave (the extra slot per cell
            // staggers the cells' staged records across LDS banks), and the
            // owner cell of every staged slot
            int n = 0;
            if (tid < nq) {
                n = s_cnt[(2 * (tid / ncx) + cb) * kRegX + 2 * (tid % ncx) + ca] + 1;
            }
            int incl = n;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int up = __shfl_up(incl, d);
                if (tid >= d) incl += up;
            }
            if (tid < nq) {
                s_off[tid + 1] = incl;
                for (int t = incl - n; t < incl; ++t) {
                    s_owner[t] = (unsigned char)tid;
                }
            }
            if (tid == 0) {
                s_off[0] = 0;
            }
        }
        if (tid >= kP2GThreads - kColourCellsMax && tid - (kP2GThreads - kColourCellsMax) < nq) {
            // sort one bin by particle index (insertion sort; bins arrive
            // nearly sorted because lanes of a wave insert in lane order)
            const int q = tid - (kP2GThreads - kColourCellsMax);
            const int a = 2 * (q % ncx) + ca, b = 2 * (q / ncx) + cb;
            const int n = s_cnt[b * kRegX + a];
            int *lst = &s_lst[q * kLstStride];
            for (int k = 1; k < n; ++k) {
                const int key = lst[k];
                int r = k - 1;
                while (r >= 0 && lst[r] > key) {
                    lst[r + 1] = lst[r];
                    --r;
                }
                lst[r + 1] = key;
            }
        }
        __syncthreads();
        if (c < 3) {
            fetch_lists(c + 1);   // in flight while this colour is staged and folded
        }

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

        // Stage the colour's particles in batches of whole cells.  Cell q's
        // records sit at [s_off[q], s_off[q+1] - 1); slot s_off[q+1] - 1 is
        // padding.
        int qb = 0;
        while (qb < nq) {
            const int base = s_off[qb];
            int qe = qb + 1;
            while (qe < nq && s_off[qe + 1] - base <= kStageMax) {
                ++qe;
            }
            const int nb = s_off[qe] - base;
            int sp[kStageSlots], sc[kStageSlots];
#pragma unroll
            for (int u = 0; u < kStageSlots; ++u) {
                const int s = tid + u * kP2GThreads;
                sp[u] = 0;
                sc[u] = -1;
                if (s < nb) {
                    const int gi = base + s;
                    const int lo = s_owner[gi];
                    const int k = gi - s_off[lo];
                    if (k < s_off[lo + 1] - s_off[lo] - 1) {   // not the padding slot
                        sp[u] = s_lst[lo * kLstStride + k];
                        sc[u] = lo;
                    }
                }
            }
            double px[kStageSlots], py[kStageSlots], pm[kStageSlots], pv[kStageSlots];
            double pvx[kStageSlots], pvy[kStageSlots];
            double psx[kStageSlots], psy[kStageSlots], psz[kStageSlots];
#pragma unroll
            for (int u = 0; u < kStageSlots; ++u) {
                const int p = sp[u];    // 0 for idle slots: always a valid index
                px[u] = P.x[p];
                py[u] = P.y[p];
                pm[u] = P.m[p];
                pv[u] = P.v[p];
                pvx[u] = P.vx[p];
                pvy[u] = P.vy[p];
                psx[u] = P.txx[p];
                psy[u] = P.txy[p];
                psz[u] = P.tyy[p];
            }
#pragma unroll
            for (int u = 0; u < kStageSlots; ++u) {
                const int s = tid + u * kP2GThreads;
                if (sc[u] >= 0) {
                    const int lo = sc[u];
                    const int cc = cx0 + 2 * (lo % ncx) + ca, rr = cy0 + 2 * (lo / ncx) + cb;
                    const double mp = pm[u], vp = pv[u];
                    s_xy[s] = make_double2(snap_local((px[u] - (double)cc * g.h) / g.h),
                                           snap_local((py[u] - (double)rr * g.h) / g.h));
                    // pdata = {m, x_t m, y_t m, ..}, stressdata = {-sxx v, -sxy v, -syy v}
                    s_msy[s] = make_double2(mp, -psy[u] * vp);
                    s_pxs[s] = make_double2(pvx[u] * mp, -psx[u] * vp);
                    s_pys[s] = make_double2(pvy[u] * mp, -psz[u] * vp);
                }
            }
            __syncthreads();
            if (myq >= qb && myq < qe) {
                const int k1 = s_off[myq + 1] - 1 - base;
                const double2 *s_mom = role == 0 ? s_pxs : s_pys;
                for (int s = s_off[myq] - base; s < k1; ++s) {
                    const double2 xy = s_xy[s], msy = s_msy[s], ps = s_mom[s];
                    const double X = dx ? xy.x : 1.0 - xy.x;
                    const double Y = dy ? xy.y : 1.0 - xy.y;
                    const double sf = X * Y;
                    const double b1 = (dx ? Y : -Y) * g.inv_h;
                    const double b2 = (dy ? X : -X) * g.inv_h;
                    const double mp = msy.x;
                    if (role == 0) {
                        am += sf * mp;
                        ap += sf * ps.x;
                        af += sf * (bb * mp);
                        af += b1 * ps.y;       // -sxx v
                        af += b2 * msy.y;      // -sxy v
                    } else {
                        ap += sf * ps.x;
                        af += sf * (bb * mp);
                        af += b1 * msy.y;      // -sxy v
                        af += b2 * ps.y;       // -syy v
                    }
                }
            }
            __syncthreads();
            qb = qe;
        }
    }

    // Hand the y chains to the role-0 thread of the node (staging space is
    // free after the last barrier).
    double2 *s_y = s_xy;
    if (role == 1) {
        s_y[nt] = make_double2(ap, af);
    }
    __syncthreads();
    if (role != 0 || !live) {
        return;
    }
    double amx = ap, afx = af;
    const double2 yv = s_y[nt];
    double amy = yv.x, afy = yv.y;

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

__global__ __launch_bounds__(kG2PThreads) void k_g2p_stress(
    Particles P, const NodeRec *__restrict__ nodes, int *__restrict__ cnt_next,
    int *__restrict__ bkt_next, GridParams g, Material mt, DevState *__restrict__ st,
    int np, int mode)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        // t += dt (process_usl.c:522), then time_varying_loads for the next
        // step (src/main.c:975).  No thread of this kernel reads *st.
        const double t = st->t + g.dt;
        st->t = t;
        st->bx = 0.0;
        st->by = gravity(t, st->t0, g.g_mag, g.ramp);
    }
    if (p >= np || !P.active[p]) {
        return;
    }
    const int e = P.elem[p];
    if (e < 0) {
        // left the grid during the previous step: deactivated at the start
        // of this one (process_usl.c:660-665)
        P.active[p] = 0;
        return;
    }
    const int cc = e % g.Ne, rr = e / g.Ne;
    const double dt = g.dt;
    double x = P.x[p], y = P.y[p];
    Shape sh;
    shape_functions(x, y, cc, rr, g, sh);

    const size_t n0 = (size_t)rr * g.N + cc;
    const size_t nid[4] = {n0, n0 + 1, n0 + g.N + 1, n0 + g.N};
    NodeRec nd[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        nd[q] = nodes[nid[q]];
    }

    // move_particles_explicit_usl_split
    double ax = 0.0, ay = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (nd[q].m < kTol) continue;
        ax += ((sh.s[q] * nd[q].fx) / nd[q].m);
        ay += ((sh.s[q] * nd[q].fy) / nd[q].m);
    }
    double vx = P.vx[p], vy = P.vy[p];
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
    const double ux = P.ux[p] + dux;
    const double uy = P.uy[p] + duy;

    // calculate_strainrate_split (SMEAR4 folds left to right)
    const double exx = 1.0 * (sh.b1[0] * nd[0].vx + sh.b1[1] * nd[1].vx
                              + sh.b1[2] * nd[2].vx + sh.b1[3] * nd[3].vx);
    const double eyy = 1.0 * (sh.b2[0] * nd[0].vy + sh.b2[1] * nd[1].vy
                              + sh.b2[2] * nd[2].vy + sh.b2[3] * nd[3].vy);
    const double dx_tdy = 1.0 * (sh.b2[0] * nd[0].vx + sh.b2[1] * nd[1].vx
                                 + sh.b2[2] * nd[2].vx + sh.b2[3] * nd[3].vx);
    const double dy_tdx = 1.0 * (sh.b1[0] * nd[0].vy + sh.b1[1] * nd[1].vy
                                 + sh.b1[2] * nd[2].vy + sh.b1[3] * nd[3].vy);
    const double exy = 0.5 * (dx_tdy + dy_tdx);
    const double wxy = 0.5 * (dx_tdy - dy_tdx);

    // update_particle_densities_split
    const double mp = P.m[p];
    const double v = P.v[p] * mpm_exp(dt * (exx + eyy));

    if (!(mode & kModeHostMaterial)) {
        double txx = P.txx[p], txy = P.txy[p], tyy = P.tyy[p], tzz = P.tzz[p];
        double gp = P.gp[p], gdp;
        g_local_mu2_update(exx, exy, eyy, wxy, mp, v, dt, mt, txx, txy, tyy, tzz, gp, gdp);
        P.txx[p] = txx;
        P.txy[p] = txy;
        P.tyy[p] = tyy;
        P.tzz[p] = tzz;
        P.gp[p] = gp;
        P.gdp[p] = gdp;
    }
    P.x[p] = x;
    P.y[p] = y;
    P.vx[p] = vx;
    P.vy[p] = vy;
    P.v[p] = v;
    P.ux[p] = ux;
    P.uy[p] = uy;
    if (mode & kModeStoreAux) {
        P.exx[p] = exx;
        P.exy[p] = exy;
        P.eyy[p] = eyy;
        P.wxy[p] = wxy;
        P.ax[p] = ax;
        P.ay[p] = ay;
        P.xl[p] = sh.xl;
        P.yl[p] = sh.yl;
    }

    // The next step's create_particle_to_element_map, and its bin.
    const int en = which_element(x, y, g);
    P.elem[p] = en;
    const unsigned long long crossed = __ballot(en != e);
    if (crossed && __lane_id() == __builtin_ctzll(__ballot(1))) {
        atomicAdd(&st->crossings, (unsigned long long)__popcll(crossed));
    }
    bin_insert(en >= 0, en, p, cnt_next, bkt_next, g.cap, &st->status);
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
    double *pfield[MPM_F_COUNT];
    NodeRec *nodes;
    uint8_t *mask;
    int *cnt[2];
    int *bkt[2];
    DevState *st;
    int cur;                // bins holding the current step's lists
    int tiles_x, ntiles, p2g_blocks;
    bool have_particles;
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

static void bind_fields(mpm_hip_ctx *ctx)
{
    Particles &P = ctx->P;
    double **f = ctx->pfield;
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
    return MPM_HIP_OK;
}

static int use_device(const mpm_hip_ctx *ctx)
{
    HIP_TRY(hipSetDevice(ctx->device));
    return MPM_HIP_OK;
}

static int launch_p2g(mpm_hip_ctx *ctx)
{
    const int cur = ctx->cur, nxt = 1 - cur;
    const dim3 grid(ctx->p2g_blocks), block(kP2GThreads);
    switch (ctx->g.cap) {
    case 16:
        k_p2g_grid<16><<<grid, block, 0, ctx->stream>>>(ctx->P, ctx->cnt[cur], ctx->bkt[cur],
            ctx->cnt[nxt], ctx->nodes, ctx->mask, ctx->g, ctx->st, ctx->tiles_x, ctx->ntiles);
        break;
    case 32:
        k_p2g_grid<32><<<grid, block, 0, ctx->stream>>>(ctx->P, ctx->cnt[cur], ctx->bkt[cur],
            ctx->cnt[nxt], ctx->nodes, ctx->mask, ctx->g, ctx->st, ctx->tiles_x, ctx->ntiles);
        break;
    default:
        k_p2g_grid<64><<<grid, block, 0, ctx->stream>>>(ctx->P, ctx->cnt[cur], ctx->bkt[cur],
            ctx->cnt[nxt], ctx->nodes, ctx->mask, ctx->g, ctx->st, ctx->tiles_x, ctx->ntiles);
        break;
    }
    HIP_TRY(hipGetLastError());
    return MPM_HIP_OK;
}

static int launch_g2p(mpm_hip_ctx *ctx)
{
    const int nxt = 1 - ctx->cur;
    const size_t nb = ctx->np ? (ctx->np + kG2PThreads - 1) / kG2PThreads : 1;
    int mode = 0;
    if (ctx->cfg.flags & (MPM_HIP_STORE_AUX | MPM_HIP_HOST_MATERIAL)) mode |= kModeStoreAux;
    if (ctx->cfg.flags & MPM_HIP_HOST_MATERIAL) mode |= kModeHostMaterial;
    k_g2p_stress<<<dim3((unsigned)nb), dim3(kG2PThreads), 0, ctx->stream>>>(
        ctx->P, ctx->nodes, ctx->cnt[nxt], ctx->bkt[nxt], ctx->g, ctx->mt, ctx->st,
        (int)ctx->np, mode);
    HIP_TRY(hipGetLastError());
    ctx->cur = nxt;
    return MPM_HIP_OK;
}

static int rebin(mpm_hip_ctx *ctx, int init)
{
    HIP_TRY(hipMemsetAsync(ctx->cnt[0], 0, sizeof(int) * ctx->g.ncells, ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->cnt[1], 0, sizeof(int) * ctx->g.ncells, ctx->stream));
    ctx->cur = 0;
    if (ctx->np) {
        const size_t nb = (ctx->np + kG2PThreads - 1) / kG2PThreads;
        k_bin_particles<<<dim3((unsigned)nb), dim3(kG2PThreads), 0, ctx->stream>>>(
            ctx->P, ctx->cnt[0], ctx->bkt[0], ctx->g, ctx->st, (int)ctx->np, init);
        HIP_TRY(hipGetLastError());
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
        || num_particles > (size_t)0x7fffffff || cfg->N > 46340) {
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
    g.N = cfg->N;
    g.Ne = cfg->N - 1;
    g.ncells = g.Ne * g.Ne;
    g.cap = cap;
    g.h = cfg->h;
    g.inv_h = 1.0 / cfg->h;
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
            delete ctx;
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
            delete ctx;
            return MPM_HIP_ERR_INVALID;
        }
        g.hole_radius = cfg->bc_fp64[0];
        g.open_time = cfg->bc_fp64[1];
        g.wall_col = (int)(((size_t)cfg->N - 1) * cfg->bc_fp64[2]);
    } else if (cfg->bc_kind != MPM_BC_NONE && cfg->bc_kind != MPM_BC_MASK) {
        delete ctx;
        return MPM_HIP_ERR_INVALID;
    }

    // material_init (g_local_mu2.c:139-142) and calculate_stress_threaded's
    // lambda (g_local_mu2.c:187)
    Material &mt = ctx->mt;
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

    ctx->tiles_x = (cfg->N + kTileX - 1) / kTileX;
    ctx->ntiles = ctx->tiles_x * ((cfg->N + kTileY - 1) / kTileY);
    ctx->p2g_blocks = ((ctx->ntiles + 7) / 8) * 8;

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
    if ((cfg->flags & (MPM_HIP_STORE_AUX | MPM_HIP_HOST_MATERIAL)) && alloc_aux(ctx) != MPM_HIP_OK) {
        mpm_hip_destroy(ctx);
        return MPM_HIP_ERR_ALLOC;
    }
    ALLOC_OR_FAIL(P.elem, np * sizeof(int));
    ALLOC_OR_FAIL(P.active, np * sizeof(int));
    ALLOC_OR_FAIL(ctx->nodes, nn * sizeof(NodeRec));
    ALLOC_OR_FAIL(ctx->mask, nn);
    for (int b = 0; b < 2; ++b) {
        ALLOC_OR_FAIL(ctx->cnt[b], (size_t)g.ncells * sizeof(int));
        ALLOC_OR_FAIL(ctx->bkt[b], (size_t)g.ncells * cap * sizeof(int));
    }
    ALLOC_OR_FAIL(ctx->st, sizeof(DevState));
#undef ALLOC_OR_FAIL

    if (hipMemsetAsync(ctx->mask, 0, nn, ctx->stream) != hipSuccess
        || hipMemsetAsync(ctx->nodes, 0, nn * sizeof(NodeRec), ctx->stream) != hipSuccess
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
    if (field == MPM_F_X || field == MPM_F_Y) {
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
            hipEventElapsedTime(&a, ev[2 * s], ev[2 * s + 1]);
            p2g += a;
            hipEventElapsedTime(&a, ev[2 * s + 1], ev[2 * s + 2]);
            g2p += a;
        }
        hipEventElapsedTime(&a, ev[0], ev[2 * n]);
        ms[0] = a;
        ms[1] = p2g;
        ms[2] = g2p;
    }
    for (int k = 0; k < created; ++k) {
        hipEventDestroy(ev[k]);
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
    hipSetDevice(ctx->device);
    if (ctx->stream) {
        hipStreamSynchronize(ctx->stream);
    }
    for (int f = 0; f < MPM_F_COUNT; ++f) {
        if (ctx->pfield[f]) hipFree(ctx->pfield[f]);
    }
    if (ctx->P.elem) hipFree(ctx->P.elem);
    if (ctx->P.active) hipFree(ctx->P.active);
    if (ctx->nodes) hipFree(ctx->nodes);
    if (ctx->mask) hipFree(ctx->mask);
    for (int b = 0; b < 2; ++b) {
        if (ctx->cnt[b]) hipFree(ctx->cnt[b]);
        if (ctx->bkt[b]) hipFree(ctx->bkt[b]);
    }
    if (ctx->st) hipFree(ctx->st);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
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
    if (da) hipFree(da);
    if (db) hipFree(db);
    if (dout) hipFree(dout);
    return rc;
}

}  // extern "C"
