/*
 * usl_oracle.c -- CPU restatement of the reference's explicit USL MPM step.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg load this code, and only as the checker or as the timed
 * CPU baseline; the product (the HIP library under mpm-2d_amd/) never calls
 * it and never falls back to it.
 *
 * What it restates, operation for operation, in the reference's data layout
 * (AoS particle_t/node_t/element_t from include/mpm_abi.h, 12 SoA shape
 * function arrays, per-(thread, colour) particle lists, T pthreads with the
 * reference's barrier structure):
 *   mpm_init                       libmpm/process_usl.c:68-280
 *   WHICH_ELEMENT4                 libmpm/process_usl.c:43-44
 *   explicit_mpm_step_usl_threaded libmpm/process_usl.c:465-592
 *   explicit_mpm_step_musl_threaded libmpm/process_usl.c:284-462
 *   move_grid_split                libmpm/process_usl.c:596-625
 *   create_particle_to_element_map libmpm/process_usl.c:631-674
 *   find_filled_elements           libmpm/process_usl.c:677-723
 *   calculate_shapefunctions_split libmpm/process_usl.c:727-760
 *   calculate_strainrate_split     libmpm/process_usl.c:764-808 (SMEAR4 58-63)
 *   map_to_grid_explicit_split     libmpm/process_usl.c:812-910
 *   accumulate_p_to_n_ds_list4{7,2} libmpm/map.c:85-104
 *   move_particles_explicit_usl    libmpm/process_usl.c:914-957
 *   update_particle_densities      libmpm/process_usl.c:961-970
 *   global_to_local_coords         libmpm/particle.c:14-30
 *   tent, grad_tent                libmpm/interpolate.c:23-65
 *   node_number_to_coords          libmpm/element.c:10-19
 *   tensor_*3                      libmpm/tensor.c:64-437
 *   initial_loads, time_varying    libmpm/loading.c:21-54
 *   g_local_mu2                    materialsrc/g_local_mu2.c:90-311
 *   box                            bcsrc/box.c:42-224
 *   square_silo                    bcsrc/square_silo.c:32-149
 *   mpm_run_until + thread split   src/main.c:747-839, 934-982
 *
 * Deliberate differences from the reference, all documented:
 *   - Constants the reference hard-codes (g_local_mu2's mu_s, mu_2, I_0,
 *     grain diameter, grain density, density cut-off; loading's gravity and
 *     ramp time) are parameters whose defaults are the reference values.
 *   - exp() in the volume update: the parity build (ORACLE_PORTABLE_EXP)
 *     calls mpm_exp() from include/mpm_exp.h, the same IEEE-only function the
 *     GPU kernels use; it is correctly rounded for |x| <= 2^-8 and within
 *     one ulp of glibc's exp elsewhere (tests/test_exp.py).  The timing build
 *     calls glibc's exp like the reference.
 *   - A particle whose element index computes to num_elements or beyond
 *     (x or y within an ulp of 1.0 on a grid whose spacing is not a power of
 *     two) reads past the element array in the reference (undefined); here
 *     it is deactivated like a particle that left the grid.
 *   - The reference's frame/state writers and driver I/O are not part of
 *     the step and are not restated.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpm_abi.h"
#ifdef ORACLE_PORTABLE_EXP
#include "mpm_exp.h"
#define ORACLE_EXP mpm_exp
#else
#define ORACLE_EXP exp
#endif

#define TOL 1e-10

/* ------------------------------------------------------------------------ */
/* Parameters the oracle takes (reference defaults noted).                  */
/* ------------------------------------------------------------------------ */
typedef struct oracle_params {
    int N;                 /* nodes per side */
    double h;              /* node spacing (reference: len / (N-1)) */
    double dt;
    int num_threads;
    /* g_local_mu2 (materialsrc/g_local_mu2.c:30-41, 139-140, 247) */
    double E, nu;
    double mu_s;           /* 0.3819 */
    double mu_2;           /* 0.6435 */
    double I_0;            /* 0.278 */
    double grains_d;       /* 0.001 * 5 */
    double grains_rho;     /* 2450 */
    double rho_crit;       /* 1485 */
    /* loading (libmpm/loading.c:17-18) */
    double g_mag;          /* 9.8 */
    double ramp_time;      /* 0.1 */
    /* boundary conditions: 0 none, 1 box, 2 square_silo */
    int bc_kind;
    double bc_fp64[4];
    int bc_int[4];
    int log_crossings;     /* write the reference's per-crossing log lines */
    const char *log_path;  /* NULL: /dev/null when logging */
    /* material: 0 g_local_mu2, 1 g_nonlocal (materialsrc/g_nonlocal.c),
     * 2 isolin (materialsrc/isolin.c; E and nu from the fields above),
     * 3 dp_ri_mt (materialsrc/dp_ri_mt.c; E, nu, mu_s from the fields above) */
    int material_kind;
    /* g_nonlocal constants (g_nonlocal.c:31-47, 141, 313):
     * E (EMOD 1e8), nu (NUMOD 0.3), mu_s (0.3819), A (0.48), B (0.9377*20),
     * grains_d (0.01*5), grains_rho (2450), p_cap (1e-2), c (1e-2) */
    double nl_E, nl_nu, nl_mu_s, nl_A, nl_B, nl_d, nl_rho, nl_p_cap, nl_c;
    /* step: 0 explicit_mpm_step_usl_threaded, 1 explicit_mpm_step_musl_threaded */
    int scheme;
} oracle_params_t;

typedef struct oracle_job {
    job_t job;
    oracle_params_t prm;
    /* material constants derived in material_init */
    double G, K;
    /* box plugin state (bcsrc/box.c:22-30) */
    size_t left_col, right_col, bottom_row, top_row;
    int left_fric, right_fric, bottom_fric, top_fric;
    /* square_silo plugin state (bcsrc/square_silo.c:19-21) */
    double hole_radius, open_time;
    size_t wall_col;
    threadtask_t *tasks;
    int steps_to_run;
    long long crossings;
    pthread_mutex_t crossings_lock;
} oracle_job_t;

#define OJ(job) ((oracle_job_t *)(job))

/* ------------------------------------------------------------------------ */
/* libmpm/particle.c, interpolate.c, element.c                              */
/* ------------------------------------------------------------------------ */
void oracle_global_to_local_coords(double *x_local, double *y_local,
    double x, double y, double x_ref, double y_ref, double h)
{
    *x_local = (x - x_ref) / h;
    const double xl_sgn = copysign(1.0, *x_local);
    if ((xl_sgn * (*x_local)) > 1.0f && ((xl_sgn * (*x_local)) - 1.0f) < TOL) {
        *x_local = xl_sgn;
    }
    *y_local = (y - y_ref) / h;
    const double yl_sgn = copysign(1.0, *y_local);
    if ((yl_sgn * (*y_local)) > 1.0f && ((yl_sgn * (*y_local)) - 1.0f) < TOL) {
        *y_local = yl_sgn;
    }
}

void oracle_tent(double *h1, double *h2, double *h3, double *h4,
    double xl, double yl)
{
    *h1 = (1 - xl) * (1 - yl);
    *h2 = (xl) * (1 - yl);
    *h3 = (xl) * (yl);
    *h4 = (1 - xl) * (yl);
}

void oracle_grad_tent(double *b11, double *b12, double *b13, double *b14,
    double *b21, double *b22, double *b23, double *b24,
    double xl, double yl, double h)
{
    double dxl_dx = (1.0 / h);
    double dyl_dy = (1.0 / h);
    *b11 = -(1 - yl) * dxl_dx;
    *b12 = (1 - yl) * dxl_dx;
    *b13 = (yl) * dxl_dx;
    *b14 = -(yl) * dxl_dx;
    *b21 = -(1 - xl) * dyl_dy;
    *b22 = -(xl) * dyl_dy;
    *b23 = (xl) * dyl_dy;
    *b24 = (1 - xl) * dyl_dy;
}

static void node_number_to_coords(double *x, double *y, int num, int N, double h)
{
    const int i = num % N;
    const int j = num / N;
    *x = i * h;
    *y = j * h;
}

/* WHICH_ELEMENT4 (libmpm/process_usl.c:43-44) */
static int which_element(double xp, double yp, size_t N, double h)
{
    return (int)((xp < 1.0 && xp >= 0.0 && yp < 1.0 && yp >= 0.0)
        ? ((floor(xp / h) + floor(yp / h) * (N - 1))) : (-1));
}

/* ------------------------------------------------------------------------ */
/* libmpm/tensor.c (3x3 helpers, row-major)                                 */
/* ------------------------------------------------------------------------ */
void oracle_tensor_trace3(double *trA, const double *A)
{
    *trA = 0;
    for (size_t i = 0; i < 3; i++) {
        *trA += A[i * 3 + i];
    }
}

void oracle_tensor_add3(double *C, const double *A, const double *B)
{
    if (C != A) {
        for (size_t i = 0; i < 9; i++) {
            C[i] = A[i];
        }
    }
    for (size_t i = 0; i < 3; i++) {
        for (size_t j = 0; j < 3; j++) {
            C[i * 3 + j] += B[i * 3 + j];
        }
    }
}

void oracle_tensor_copy3(double *C, const double *A)
{
    if (C == A) {
        return;
    }
    for (size_t i = 0; i < 3; i++) {
        for (size_t j = 0; j < 3; j++) {
            C[i * 3 + j] = A[i * 3 + j];
        }
    }
}

void oracle_tensor_decompose3(double *A_0, double *c, const double *A)
{
    double trA;
    oracle_tensor_trace3(&trA, A);
    *c = trA / 3.0;
    oracle_tensor_copy3(A_0, A);
    for (size_t i = 0; i < 3; i++) {
        A_0[i * 3 + i] = A_0[i * 3 + i] - (*c);
    }
}

void oracle_tensor_multiply3(double *C, const double *A, const double *B)
{
    for (size_t i = 0; i < 3; i++) {
        for (size_t j = 0; j < 3; j++) {
            C[i * 3 + j] = A[3 * i + 0] * B[3 * 0 + j]
                + A[3 * i + 1] * B[3 * 1 + j]
                + A[3 * i + 2] * B[3 * 2 + j];
        }
    }
}

void oracle_tensor_scale3(double *A, double c)
{
    for (size_t i = 0; i < 3; i++) {
        for (size_t j = 0; j < 3; j++) {
            A[i * 3 + j] *= c;
        }
    }
}

void oracle_tensor_sym3(double *C, const double *A)
{
    for (size_t i = 0; i < 3; i++) {
        for (size_t j = 0; j < 3; j++) {
            C[i * 3 + j] = 0.5 * (A[i * 3 + j] + A[j * 3 + i]);
        }
    }
}

void oracle_tensor_skw3(double *C, const double *A)
{
    for (size_t i = 0; i < 3; i++) {
        for (size_t j = 0; j < 3; j++) {
            C[i * 3 + j] = 0.5 * (A[i * 3 + j] - A[j * 3 + i]);
        }
    }
}

void oracle_tensor_contraction3(double *c, const double *A, const double *B)
{
    *c = 0;
    for (size_t i = 0; i < 3; i++) {
        for (size_t j = 0; j < 3; j++) {
            (*c) += A[i * 3 + j] * B[i * 3 + j];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* materialsrc/g_local_mu2.c                                                */
/* ------------------------------------------------------------------------ */
double oracle_negative_root(double a, double b, double c)
{
    double x;
    if (b > 0) {
        x = (-b - sqrt(b * b - 4 * a * c)) / (2 * a);
    } else {
        x = (2 * c) / (-b + sqrt(b * b - 4 * a * c));
    }
    return x;
}

static void g_local_mu2_init(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    for (size_t i = 0; i < job->num_particles; i++) {
        particle_t *p = &job->particles[i];
        for (int j = 0; j < DEPVAR; j++) {
            p->state[j] = 0;
        }
        p->T[XX] = p->sxx;
        p->T[XY] = p->sxy;
        p->T[XZ] = 0;
        p->T[YX] = p->sxy;
        p->T[YY] = p->syy;
        p->T[YZ] = 0;
        p->T[ZX] = 0;
        p->T[ZY] = 0;
        p->T[ZZ] = 0;
    }
    oj->G = oj->prm.E / (2.0 * (1.0 + oj->prm.nu));
    oj->K = oj->prm.E / (3.0 * (1.0 - 2 * oj->prm.nu));
}

/* Per-particle body of calculate_stress_threaded (g_local_mu2.c:197-307). */
void oracle_g_local_mu2_point(particle_t *p, double dt, double G, double K,
    double mu_s, double mu_2, double I_0, double grains_d, double grains_rho,
    double rho_crit)
{
    double tau_tau, scale_factor, p_tr, tau_tr, nup_tau;
    double const c = 0;
    int density_flag;
    double trD;
    const double lambda = K - 2.0 * G / 3.0;
    double S0, S2, B, H, alpha;
    double T0[9], Ttr[9], JS[9], W[9], D[9], temp[9];

    p->L[XX] = p->exx_t;
    p->L[XY] = p->exy_t + p->wxy_t;
    p->L[XZ] = 0;
    p->L[YX] = p->exy_t - p->wxy_t;
    p->L[YY] = p->eyy_t;
    p->L[YZ] = 0;
    p->L[ZX] = 0;
    p->L[ZY] = 0;
    p->L[ZZ] = 0;

    oracle_tensor_skw3(W, p->L);
    oracle_tensor_sym3(D, p->L);
    oracle_tensor_copy3(Ttr, D);
    oracle_tensor_trace3(&trD, D);

    oracle_tensor_multiply3(JS, W, p->T);
    oracle_tensor_multiply3(temp, p->T, W);
    oracle_tensor_scale3(temp, -1.0);
    oracle_tensor_add3(JS, JS, temp);

    oracle_tensor_scale3(Ttr, 2.0 * G);
    Ttr[XX] += lambda * trD;
    Ttr[YY] += lambda * trD;
    Ttr[ZZ] += lambda * trD;
    oracle_tensor_add3(Ttr, Ttr, JS);
    oracle_tensor_scale3(Ttr, dt);
    oracle_tensor_add3(Ttr, Ttr, p->T);

    oracle_tensor_decompose3(T0, &p_tr, Ttr);
    oracle_tensor_contraction3(&tau_tr, T0, T0);
    tau_tr = sqrt(0.5 * tau_tr);
    p_tr *= -1.0;

    if ((p->m / p->v) < rho_crit) {
        density_flag = 1;
    } else {
        density_flag = 0;
    }

    if (density_flag || p_tr <= c) {
        nup_tau = 0;
        for (int k = 0; k < 9; k++) {
            p->T[k] = 0;
        }
    } else if (p_tr > c) {
        S0 = mu_s * p_tr;
        if (tau_tr <= S0) {
            tau_tau = tau_tr;
            scale_factor = 1.0;
        } else {
            S2 = mu_2 * p_tr;
            alpha = G * I_0 * dt * sqrt(p_tr / grains_rho) / grains_d;
            B = -(S2 + tau_tr + alpha);
            H = S2 * tau_tr + S0 * alpha;
            tau_tau = oracle_negative_root(1.0, B, H);
            scale_factor = (tau_tau / tau_tr);
        }
        nup_tau = ((tau_tr - tau_tau) / G) / dt;
        oracle_tensor_scale3(T0, scale_factor);
        oracle_tensor_copy3(p->T, T0);
        p->T[XX] -= p_tr;
        p->T[YY] -= p_tr;
        p->T[ZZ] -= p_tr;
    } else {
        nup_tau = 0;
    }

    p->sxx = p->T[XX];
    p->sxy = p->T[XY];
    p->syy = p->T[YY];
    p->state[9] += nup_tau * dt;     /* gammap */
    p->state[10] = nup_tau;          /* gammadotp */
}

static void g_local_mu2_stress_threaded(threadtask_t *task)
{
    oracle_job_t *oj = OJ(task->job);
    job_t *job = task->job;
    const oracle_params_t *q = &oj->prm;
    size_t p_start = task->offset;
    size_t p_stop = task->offset + task->blocksize;
    for (size_t i = p_start; i < p_stop; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        oracle_g_local_mu2_point(&job->particles[i], job->dt, oj->G, oj->K,
            q->mu_s, q->mu_2, q->I_0, q->grains_d, q->grains_rho, q->rho_crit);
    }
}

/* calculate_stress_threaded of materialsrc/isolin.c (72-116): isotropic
 * linear elasticity with the Jaumann rate on sxx, sxy, syy; G and K as
 * material_init (isolin.c:30-58) derives them. */
static void isolin_point(particle_t *p, double dt, double G, double K)
{
    const double lambda = K - 2.0 * G / 3.0;
    const double trD = p->exx_t + p->eyy_t;
    double dsjxx = lambda * trD + 2.0 * G * p->exx_t;
    double dsjxy = 2.0 * G * p->exy_t;
    double dsjyy = lambda * trD + 2.0 * G * p->eyy_t;
    dsjxx += 2 * p->wxy_t * p->sxy;
    dsjxy -= p->wxy_t * (p->sxx - p->syy);
    dsjyy -= 2 * p->wxy_t * p->sxy;
    p->sxx += dt * dsjxx;
    p->sxy += dt * dsjxy;
    p->syy += dt * dsjyy;
}

/* One particle of dp_ri_mt's calculate_stress_threaded
 * (materialsrc/dp_ri_mt.c:133-205): Jaumann trial stress, then the
 * rate-independent Drucker-Prager return tau <= mu_s p; szz is state[1]. */
static void dp_ri_mt_point(particle_t *p, double dt, double G, double K, double mu_s)
{
    const double lambda = K - 2.0 * G / 3.0;
    double *szz = &p->state[1];
    const double trD = p->exx_t + p->eyy_t;
    double dsjxx = lambda * trD + 2.0 * G * p->exx_t;
    double dsjxy = 2.0 * G * p->exy_t;
    double dsjyy = lambda * trD + 2.0 * G * p->eyy_t;
    dsjxx += 2 * p->wxy_t * p->sxy;
    dsjxy -= p->wxy_t * (p->sxx - p->syy);
    dsjyy -= 2 * p->wxy_t * p->sxy;
    const double sxx_tr = p->sxx + dt * dsjxx;
    const double sxy_tr = p->sxy + dt * dsjxy;
    const double syy_tr = p->syy + dt * dsjyy;
    const double szz_tr = *szz + dt * lambda * trD;
    const double p_tr = -(sxx_tr + syy_tr + szz_tr) / 3.0;
    const double t0xx_tr = sxx_tr + p_tr;
    const double t0xy_tr = sxy_tr;
    const double t0yy_tr = syy_tr + p_tr;
    const double t0zz_tr = szz_tr + p_tr;
    const double tau_tr = sqrt(0.5 * (t0xx_tr * t0xx_tr + 2 * t0xy_tr * t0xy_tr
                                      + t0yy_tr * t0yy_tr + t0zz_tr * t0zz_tr));
    double nup_tau;
    if ((p->m / p->v) < 1485.0f || p_tr <= 0) {
        nup_tau = (tau_tr) / (G * dt);
        p->sxx = 0;
        p->sxy = 0;
        p->syy = 0;
        *szz = 0;
    } else if (p_tr > 0) {
        const double S0 = mu_s * p_tr;
        double tau_tau, scale_factor;
        if (tau_tr <= S0) {
            tau_tau = tau_tr;
            scale_factor = 1.0;
        } else {
            tau_tau = S0;
            scale_factor = (tau_tau / tau_tr);
        }
        nup_tau = ((tau_tr - tau_tau) / G) / dt;
        p->sxx = scale_factor * t0xx_tr - p_tr;
        p->sxy = scale_factor * t0xy_tr;
        p->syy = scale_factor * t0yy_tr - p_tr;
        *szz = scale_factor * t0zz_tr - p_tr;
    } else {
        nup_tau = 0;
    }
    p->state[9] += nup_tau * dt;
    p->state[10] = nup_tau;
}

static void dp_ri_mt_stress_threaded(threadtask_t *task)
{
    oracle_job_t *oj = OJ(task->job);
    job_t *job = task->job;
    for (size_t i = task->offset; i < task->offset + task->blocksize; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        dp_ri_mt_point(&job->particles[i], job->dt, oj->G, oj->K, oj->prm.mu_s);
    }
}

static void isolin_stress_threaded(threadtask_t *task)
{
    oracle_job_t *oj = OJ(task->job);
    job_t *job = task->job;
    size_t p_start = task->offset;
    size_t p_stop = task->offset + task->blocksize;
    for (size_t i = p_start; i < p_stop; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        isolin_point(&job->particles[i], job->dt, oj->G, oj->K);
    }
}

/* ------------------------------------------------------------------------ */
/* materialsrc/g_nonlocal.c                                                 */
/* state[3] gf, state[4] eta, state[5] gf_bulk, state[9] gammap,           */
/* state[10] xisq_inv (g_nonlocal.c:49-59)                                  */
/* ------------------------------------------------------------------------ */

/* calculate_bulk_granular_fluidity (g_nonlocal.c:306-352) for one particle:
 * from the in-plane stress at the start of the step, out[0] = eta
 * (state[4]), out[1] = gf_bulk (state[5]), out[2] = xisq_inv (state[10]). */
void oracle_nonlocal_bulk_point(double sxx, double sxy, double syy,
    const oracle_params_t *q, double *out)
{
    const double p_cap = q->nl_p_cap;
    double p_t = -0.5 * (sxx + syy);
    double sxx_tr0 = sxx + p_t;
    double sxy_tr0 = sxy;
    double syy_tr0 = syy + p_t;
    double tau_t = sqrt(0.5 * (sxx_tr0 * sxx_tr0 + 2 * sxy_tr0 * sxy_tr0 + syy_tr0 * syy_tr0));
    if (p_t < p_cap) {
        p_t = p_cap;
    }
    double S0 = p_t * q->nl_mu_s;
    double eta = q->nl_B * q->nl_d * sqrt(p_t * q->nl_rho);
    out[0] = eta;
    out[2] = fabs(tau_t / p_t - q->nl_mu_s) / (q->nl_d * q->nl_d * q->nl_A * q->nl_A);
    if (tau_t > S0) {
        out[1] = (p_t / eta) * (1 - S0 / tau_t);
    } else {
        out[1] = 0;
    }
}

/* The same for every particle. */
static void g_nonlocal_bulk(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    for (size_t i = 0; i < job->num_particles; i++) {
        particle_t *pp = &job->particles[i];
        double out[3];
        oracle_nonlocal_bulk_point(pp->sxx, pp->sxy, pp->syy, &oj->prm, out);
        pp->state[4] = out[0];
        pp->state[5] = out[1];
        pp->state[10] = out[2];
    }
}

/* In-place Cholesky factorisation and solve of a symmetric positive definite
 * band matrix: band[k * n + i] holds K(i, i - k) for k = 0..bw.  Stands in
 * for CXSparse's cs_lusol (absent here) in solve_diffusion_part; pinned
 * against scipy's sparse solver in the tests.  Returns 0 if not SPD. */
int oracle_band_cholesky_solve(double *band, size_t n, size_t bw, double *b)
{
    for (size_t j = 0; j < n; j++) {
        double d = band[j];
        size_t kmax = j < bw ? j : bw;
        for (size_t k = 1; k <= kmax; k++) {
            double l = band[k * n + j];
            d -= l * l;
        }
        if (!(d > 0)) {
            return 0;
        }
        d = sqrt(d);
        band[j] = d;
        for (size_t r = 1; r <= bw && j + r < n; r++) {
            size_t i = j + r;
            double v = band[r * n + i];
            size_t kmax2 = j < (bw - r) ? j : (bw - r);
            for (size_t k = 1; k <= kmax2; k++) {
                v -= band[(r + k) * n + i] * band[k * n + j];
            }
            band[r * n + i] = v / d;
        }
    }
    for (size_t i = 0; i < n; i++) {
        double v = b[i];
        size_t kmax = i < bw ? i : bw;
        for (size_t k = 1; k <= kmax; k++) {
            v -= band[k * n + i] * b[i - k];
        }
        b[i] = v / band[i];
    }
    for (size_t ii = n; ii-- > 0;) {
        double v = b[ii];
        for (size_t k = 1; k <= bw && ii + k < n; k++) {
            v -= band[k * n + ii + k] * b[ii + k];
        }
        b[ii] = v / band[ii];
    }
    return 1;
}

/* Last assembled system, kept for the tests. */
static double *nl_last_band, *nl_last_f;
static size_t nl_last_n, nl_last_bw;

/* solve_diffusion_part (g_nonlocal.c:356-592).  Nodes with m > TOL are the
 * unknowns, numbered in ascending node order; K and f are summed in particle
 * order as cs_entry + cs_dupl sum them.  Entries whose node carries no mass
 * (the reference indexes past its map there) are skipped. */
static int g_nonlocal_solve(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    size_t nn = job->num_nodes;
    long *node_map = malloc(nn * sizeof(long));
    size_t slda = 0;
    for (size_t i = 0; i < nn; i++) {
        node_map[i] = -1;
    }
    for (size_t i = 0; i < nn; i++) {
        size_t i_new = (size_t)(job->node_number_override[NODAL_DOF * i + 0] - 0) / NODAL_DOF;
        if (node_map[i_new] != -1) {
            continue;
        }
        if (job->nodes[i_new].m > TOL) {
            node_map[i_new] = (long)slda++;
        }
    }
    if (slda == 0) {
        free(node_map);
        return -1;
    }
    /* bandwidth of the DOF numbering */
    size_t bw = 0;
    for (size_t i = 0; i < job->num_particles; i++) {
        if (job->active[i] == 0 || job->in_element[i] == -1) continue;
        const int *nd = job->elements[job->in_element[i]].nodes;
        for (int a = 0; a < 4; a++) {
            for (int b2 = 0; b2 < 4; b2++) {
                long sa = node_map[nd[a]], sb = node_map[nd[b2]];
                if (sa >= 0 && sb >= 0) {
                    size_t d = (size_t)labs(sa - sb);
                    if (d > bw) bw = d;
                }
            }
        }
    }
    double *band = calloc((bw + 1) * slda, sizeof(double));
    double *f = calloc(slda, sizeof(double));
    double s[4], grad_s[4][2];
    for (size_t i = 0; i < job->num_particles; i++) {
        if (job->active[i] == 0) continue;
        int p = job->in_element[i];
        if (p == -1) continue;
        const particle_t *pp = &job->particles[i];
        s[0] = job->h1[i]; s[1] = job->h2[i]; s[2] = job->h3[i]; s[3] = job->h4[i];
        const int *nd = job->elements[p].nodes;
        for (int ei = 0; ei < 4; ei++) {
            long sgi = node_map[nd[ei]];
            if (sgi < 0) continue;
            f[sgi] += (pp->v) * pp->state[5] * pp->state[10] * s[ei];
        }
    }
    for (size_t i = 0; i < job->num_particles; i++) {
        if (job->active[i] == 0) continue;
        int p = job->in_element[i];
        if (p == -1) continue;
        const particle_t *pp = &job->particles[i];
        s[0] = job->h1[i]; s[1] = job->h2[i]; s[2] = job->h3[i]; s[3] = job->h4[i];
        grad_s[0][0] = job->b11[i]; grad_s[1][0] = job->b12[i];
        grad_s[2][0] = job->b13[i]; grad_s[3][0] = job->b14[i];
        grad_s[0][1] = job->b21[i]; grad_s[1][1] = job->b22[i];
        grad_s[2][1] = job->b23[i]; grad_s[3][1] = job->b24[i];
        const int *nd = job->elements[p].nodes;
        for (int ei = 0; ei < 4; ei++) {
            for (int ej = 0; ej < 4; ej++) {
                long sgi = node_map[nd[ei]], sgj = node_map[nd[ej]];
                if (sgi < 0 || sgj < 0 || sgj > sgi) continue;   /* lower band */
                band[(size_t)(sgi - sgj) * slda + (size_t)sgi] +=
                    pp->v * (pp->state[10] * s[ei] * s[ej]
                             + (grad_s[ei][0] * grad_s[ej][0] + grad_s[ei][1] * grad_s[ej][1]));
            }
        }
    }
    free(nl_last_band);
    free(nl_last_f);
    nl_last_band = malloc((bw + 1) * slda * sizeof(double));
    nl_last_f = malloc(slda * sizeof(double));
    memcpy(nl_last_band, band, (bw + 1) * slda * sizeof(double));
    memcpy(nl_last_f, f, slda * sizeof(double));
    nl_last_n = slda;
    nl_last_bw = bw;
    int ok = oracle_band_cholesky_solve(band, slda, bw, f);
    if (ok) {
        for (size_t i = 0; i < job->num_particles; i++) {
            if (job->active[i] == 0) continue;
            int p = job->in_element[i];
            if (p == -1) continue;
            particle_t *pp = &job->particles[i];
            s[0] = job->h1[i]; s[1] = job->h2[i]; s[2] = job->h3[i]; s[3] = job->h4[i];
            const int *nd = job->elements[p].nodes;
            double gf = 0;
            for (int ei = 0; ei < 4; ei++) {
                long sgi = node_map[nd[ei]];
                gf += (sgi >= 0 ? f[sgi] : 0.0) * s[ei];
            }
            pp->state[3] = gf;
        }
    }
    free(band);
    free(f);
    free(node_map);
    return ok ? 0 : -1;
}

/* The stress update of calculate_stress (g_nonlocal.c:148-298) for one
 * particle: s[3] = {sxx, sxy, syy} in and out, *gammap (state[9]) advanced;
 * gf is the particle's nonlocal fluidity (state[3]), t the job's time. */
void oracle_nonlocal_stress_point(double *s, double *gammap, double exx_t,
    double exy_t, double eyy_t, double wxy_t, double gf, double t, double dt,
    const oracle_params_t *q)
{
    const double EMOD = q->nl_E, NUMOD = q->nl_nu, MU_S = q->nl_mu_s;
    const double c = q->nl_c;
    double sxx = s[0], sxy = s[1], syy = s[2];
    double dsjxx = dt * (EMOD / (1 - NUMOD * NUMOD)) * ((exx_t) + NUMOD * (eyy_t));
    double dsjxy = dt * (EMOD / (2 * (1 + NUMOD))) * (exy_t);
    double dsjyy = dt * (EMOD / (1 - NUMOD * NUMOD)) * ((eyy_t) + NUMOD * (exx_t));
    dsjxx -= 2 * dt * wxy_t * sxy;
    dsjxy += dt * wxy_t * (sxx - syy);
    dsjyy += 2 * dt * wxy_t * sxy;
    double sxx_tr = sxx + dsjxx;
    double syy_tr = syy + dsjyy;
    double p_tr = -0.5 * (sxx_tr + syy_tr);
    if (p_tr < c) {
        s[0] = 0.5 * c / MU_S;
        s[1] = 0;
        s[2] = 0.5 * c / MU_S;
        return;
    }
    /* explicit constitutive update (g_nonlocal.c:232-296) */
    double p_t = -0.5 * (sxx + syy);
    double sxx_t0 = sxx + p_t;
    double sxy_t0 = sxy;
    double syy_t0 = syy + p_t;
    double tau_t = sqrt(0.5 * (sxx_t0 * sxx_t0 + 2 * sxy_t0 * sxy_t0 + syy_t0 * syy_t0));
    double Npxx = sqrt(0.5) * sxx_t0 / tau_t;
    double Npxy = sqrt(0.5) * sxy_t0 / tau_t;
    double Npyy = sqrt(0.5) * syy_t0 / tau_t;
    double dxx0 = 0.5 * (exx_t - eyy_t);
    double dxy0 = exy_t;
    double dyy0 = 0.5 * (eyy_t - exx_t);
    double gammadot = sqrt(2 * (dxx0 * dxx0 + 2 * dxy0 * dxy0 + dyy0 * dyy0));
    double nup_tau, dpxx, dpxy, dpyy;
    if (p_t > c && gf > 0) {
        nup_tau = gf * tau_t / p_t;
        if (nup_tau < 0) nup_tau = 0;
        if (nup_tau > gammadot) nup_tau = gammadot;
        dpxx = sqrt(0.5) * (nup_tau) * Npxx;
        dpxy = sqrt(0.5) * (nup_tau) * Npxy;
        dpyy = sqrt(0.5) * (nup_tau) * Npyy;
    } else {
        nup_tau = 0;
        dpxx = 0;
        dpxy = 0;
        dpyy = 0;
    }
    if (t == 0) {
        nup_tau = 0;
        dpxx = 0;
        dpxy = 0;
        dpyy = 0;
    }
    dsjxx = dt * (EMOD / (1 - NUMOD * NUMOD)) * ((exx_t - dpxx) + NUMOD * (eyy_t - dpyy));
    dsjxy = dt * (EMOD / (2 * (1 + NUMOD))) * (exy_t - dpxy);
    dsjyy = dt * (EMOD / (1 - NUMOD * NUMOD)) * ((eyy_t - dpyy) + NUMOD * (exx_t - dpxx));
    dsjxx -= 2 * dt * wxy_t * sxy;
    dsjxy += dt * wxy_t * (sxx - syy);
    dsjyy += 2 * dt * wxy_t * sxy;
    *gammap += nup_tau * dt;
    s[0] = sxx + dsjxx;
    s[1] = sxy + dsjxy;
    s[2] = syy + dsjyy;
}

/* calculate_stress (g_nonlocal.c:103-301): every particle. */
static void g_nonlocal_stress(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    g_nonlocal_bulk(oj);
    if (g_nonlocal_solve(oj) != 0) {
        fprintf(stderr, "oracle: nonlocal solve failed\n");
    }
    for (size_t i = 0; i < job->num_particles; i++) {
        particle_t *pp = &job->particles[i];
        double s[3] = {pp->sxx, pp->sxy, pp->syy};
        oracle_nonlocal_stress_point(s, &pp->state[9], pp->exx_t, pp->exy_t,
            pp->eyy_t, pp->wxy_t, pp->state[3], job->t, job->dt, &oj->prm);
        pp->sxx = s[0];
        pp->sxy = s[1];
        pp->syy = s[2];
    }
}

/* The last assembled nonlocal system (lower band of K, and f). */
size_t oracle_nonlocal_system(double *band, double *f, size_t *bw)
{
    if (bw) *bw = nl_last_bw;
    if (band && nl_last_band) memcpy(band, nl_last_band, (nl_last_bw + 1) * nl_last_n * sizeof(double));
    if (f && nl_last_f) memcpy(f, nl_last_f, nl_last_n * sizeof(double));
    return nl_last_n;
}

/* ------------------------------------------------------------------------ */
/* bcsrc/box.c and bcsrc/square_silo.c                                      */
/* ------------------------------------------------------------------------ */
static int box_validate(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    double x0 = oj->prm.bc_fp64[0], y0 = oj->prm.bc_fp64[1];
    double w = oj->prm.bc_fp64[2], h = oj->prm.bc_fp64[3];
    oj->left_fric = oj->prm.bc_int[0];
    oj->bottom_fric = oj->prm.bc_int[1];
    oj->right_fric = oj->prm.bc_int[2];
    oj->top_fric = oj->prm.bc_int[3];
    if (x0 < 0 || y0 < 0 || x0 > 1.0 || y0 > 1.0) {
        return 0;
    }
    if (w < 0 || h < 0) {
        return 0;
    }
    if ((x0 + w) > 1.0 || (y0 + h) > 1.0) {
        return 0;
    }
    oj->left_col = x0 * (job->N - 1);
    oj->right_col = (x0 + w) * (job->N - 1);
    oj->bottom_row = y0 * (job->N - 1);
    oj->top_row = (y0 + h) * (job->N - 1);
    return 1;
}

#define SET_FIXED(node, dof) do { \
        job->u_dirichlet[NODAL_DOF * (node) + (dof)] = 0; \
        job->u_dirichlet_mask[NODAL_DOF * (node) + (dof)] = 1; \
    } while (0)

static void box_time_varying(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    for (size_t i = 0; i < job->num_nodes; i++) {
        for (size_t j = 0; j < NODAL_DOF; j++) {
            job->u_dirichlet_mask[NODAL_DOF * i + j] = 0;
        }
    }
    for (size_t n = oj->left_col; n <= oj->right_col; n++) {
        size_t b = n + oj->bottom_row * job->N;
        size_t t = n + oj->top_row * job->N;
        if (oj->bottom_fric & 0x01) SET_FIXED(b, XDOF_IDX);
        if (oj->bottom_fric & 0x02) SET_FIXED(b, YDOF_IDX);
        if (oj->top_fric & 0x01) SET_FIXED(t, XDOF_IDX);
        if (oj->top_fric & 0x02) SET_FIXED(t, YDOF_IDX);
    }
    for (size_t n = oj->bottom_row; n <= oj->top_row; n++) {
        size_t l = n * job->N + oj->left_col;
        size_t r = n * job->N + oj->right_col;
        if (oj->left_fric & 0x01) SET_FIXED(l, XDOF_IDX);
        if (oj->left_fric & 0x02) SET_FIXED(l, YDOF_IDX);
        if (oj->right_fric & 0x01) SET_FIXED(r, XDOF_IDX);
        if (oj->right_fric & 0x02) SET_FIXED(r, YDOF_IDX);
    }
    for (size_t i = 0; i < job->num_nodes; i++) {
        for (size_t j = 0; j < NODAL_DOF; j++) {
            job->node_number_override[NODAL_DOF * i + j] = (int)(NODAL_DOF * i + j);
        }
    }
}

static int silo_validate(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    oj->hole_radius = oj->prm.bc_fp64[0];
    oj->open_time = oj->prm.bc_fp64[1];
    oj->wall_col = (int)((job->N - 1) * oj->prm.bc_fp64[2]);
    return 1;
}

static void silo_time_varying(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    for (size_t i = 0; i < job->num_nodes; i++) {
        for (size_t j = 0; j < NODAL_DOF; j++) {
            job->u_dirichlet_mask[NODAL_DOF * i + j] = 0;
        }
    }
    for (size_t n = 0; n < job->N; n++) {
        if (job->nodes[n].x <= oj->hole_radius && job->t > oj->open_time) {
            continue;
        }
        SET_FIXED(n, XDOF_IDX);
        SET_FIXED(n, YDOF_IDX);
    }
    for (size_t n = 0; n < job->N; n++) {
        SET_FIXED(0 + n * job->N, XDOF_IDX);
        SET_FIXED(oj->wall_col + n * job->N, XDOF_IDX);
    }
    for (size_t i = 0; i < job->num_nodes; i++) {
        for (size_t j = 0; j < NODAL_DOF; j++) {
            job->node_number_override[NODAL_DOF * i + j] = (int)(NODAL_DOF * i + j);
        }
    }
}

/* bc_momentum / bc_force are identical in both plugins
 * (box.c:183-224, square_silo.c:106-149). */
static void bc_momentum(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    if (oj->prm.bc_kind == 0) {
        return;
    }
    for (size_t i = 0; i < job->num_nodes; i++) {
        for (size_t j = 0; j < NODAL_DOF; j++) {
            size_t n = NODAL_DOF * i + j;
            size_t m = job->node_number_override[n];
            if (job->u_dirichlet_mask[m] != 0) {
                if (job->u_dirichlet[m] == 0) {
                    if (j == XDOF_IDX) {
                        job->nodes[i].mx_t = 0;
                    } else if (j == YDOF_IDX) {
                        job->nodes[i].my_t = 0;
                    }
                }
            }
        }
    }
}

static void bc_force(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    if (oj->prm.bc_kind == 0) {
        return;
    }
    for (size_t i = 0; i < job->num_nodes; i++) {
        for (size_t j = 0; j < NODAL_DOF; j++) {
            size_t n = NODAL_DOF * i + j;
            size_t m = job->node_number_override[n];
            if (job->u_dirichlet_mask[m] != 0) {
                if (job->u_dirichlet[m] == 0) {
                    if (j == XDOF_IDX) {
                        job->nodes[i].fx = 0;
                    } else if (j == YDOF_IDX) {
                        job->nodes[i].fy = 0;
                    }
                }
            }
        }
    }
}

static void bc_time_varying(oracle_job_t *oj)
{
    if (oj->prm.bc_kind == 1) {
        box_time_varying(oj);
    } else if (oj->prm.bc_kind == 2) {
        silo_time_varying(oj);
    } else {
        /* a plugin that fixes nothing: no masks, identity node numbering */
        job_t *job = &oj->job;
        memset(job->u_dirichlet_mask, 0, sizeof(int) * NODAL_DOF * job->num_nodes);
        for (size_t i = 0; i < NODAL_DOF * job->num_nodes; i++) {
            job->node_number_override[i] = (int)i;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* libmpm/loading.c                                                         */
/* ------------------------------------------------------------------------ */
static void initial_loads(job_t *job)
{
    for (size_t i = 0; i < job->num_particles; i++) {
        job->particles[i].bx = 0;
        job->particles[i].by = 0;
        job->particles[i].material = 1;   /* M_DRUCKER_PRAGER */
    }
}

static void time_varying_loads(oracle_job_t *oj)
{
    job_t *job = &oj->job;
    double gravity;
    if ((oj->prm.ramp_time + job->step_start_time) > job->t) {
        gravity = -oj->prm.g_mag * ((job->t - job->step_start_time) / oj->prm.ramp_time);
    } else {
        gravity = -oj->prm.g_mag;
    }
    for (size_t i = 0; i < job->num_particles; i++) {
        job->particles[i].bx = 0;
        job->particles[i].by = gravity;
    }
}

/* ------------------------------------------------------------------------ */
/* libmpm/process_usl.c                                                     */
/* ------------------------------------------------------------------------ */
#define ijton(i, j, N) ((j) * (N) + (i))
#define NODE_OF(j, p, n) (j)->elements[(j)->in_element[p]].nodes[n]
#define SMEAR4(j, tok, i, s) ( \
    j->s ## 1[i] * j->nodes[NODE_OF(j, i, 0)].tok + \
    j->s ## 2[i] * j->nodes[NODE_OF(j, i, 1)].tok + \
    j->s ## 3[i] * j->nodes[NODE_OF(j, i, 2)].tok + \
    j->s ## 4[i] * j->nodes[NODE_OF(j, i, 3)].tok)

static void oracle_mpm_init(oracle_job_t *oj, int N, double h,
    const particle_t *particles, size_t num_particles)
{
    job_t *job = &oj->job;
    job->t = 0;
    job->t_stop = 0;
    job->frame = 0;
    job->N = N;
    job->h = h;
    job->num_nodes = (size_t)N * N;
    job->num_particles = num_particles;
    job->num_elements = (size_t)(N - 1) * (N - 1);

    job->particles = calloc(num_particles, sizeof(particle_t));
    memcpy(job->particles, particles, num_particles * sizeof(particle_t));
    for (size_t i = 0; i < num_particles; i++) {
        particle_t *p = &job->particles[i];
        p->real_sxx = p->sxx;
        p->real_sxy = p->sxy;
        p->real_syy = p->syy;
        p->exx_t = 0;
        p->exy_t = 0;
        p->eyy_t = 0;
        p->wxy_t = 0;
        p->Fxx = 1;
        p->Fxy = 0;
        p->Fyx = 0;
        p->Fyy = 1;
        p->ux = 0;
        p->uy = 0;
        p->material_data = NULL;
        p->id = i;
        p->state[9] = 0;
        p->state[10] = 0;
        memset(p->corners, 0, sizeof p->corners);
        p->color = 0;
    }

    job->nodes = calloc(job->num_nodes, sizeof(node_t));
    for (size_t i = 0; i < job->num_nodes; i++) {
        node_number_to_coords(&job->nodes[i].x, &job->nodes[i].y, (int)i, N, h);
    }

    job->elements = calloc(job->num_elements, sizeof(element_t));
    for (size_t i = 0; i < job->num_elements; i++) {
        int n = ijton(i % (job->N - 1), i / (job->N - 1), job->N);
        job->elements[i].nodes[0] = n + ijton(0, 0, job->N);
        job->elements[i].nodes[1] = n + ijton(1, 0, job->N);
        job->elements[i].nodes[2] = n + ijton(1, 1, job->N);
        job->elements[i].nodes[3] = n + ijton(0, 1, job->N);
    }

    job->num_colors = 4;
    job->color_indices = calloc(job->num_colors, sizeof(size_t));
    for (size_t i = 0; i < job->num_elements; i++) {
        size_t r = i / (job->N - 1);
        size_t c = i % (job->N - 1);
        job->elements[i].color = (int)(2 * (r % 2) + ((c % 2) + 1) - 1);
    }

    job->in_element = malloc(num_particles * sizeof(int));
    job->active = malloc(num_particles * sizeof(int));
    job->h1 = malloc(num_particles * sizeof(double));
    job->h2 = malloc(num_particles * sizeof(double));
    job->h3 = malloc(num_particles * sizeof(double));
    job->h4 = malloc(num_particles * sizeof(double));
    job->b11 = malloc(num_particles * sizeof(double));
    job->b12 = malloc(num_particles * sizeof(double));
    job->b13 = malloc(num_particles * sizeof(double));
    job->b14 = malloc(num_particles * sizeof(double));
    job->b21 = malloc(num_particles * sizeof(double));
    job->b22 = malloc(num_particles * sizeof(double));
    job->b23 = malloc(num_particles * sizeof(double));
    job->b24 = malloc(num_particles * sizeof(double));
    job->vec_len = (int)(NODAL_DOF * job->num_nodes);
    job->u_dirichlet = calloc(NODAL_DOF * job->num_nodes, sizeof(double));
    job->u_dirichlet_mask = calloc(NODAL_DOF * job->num_nodes, sizeof(int));
    job->node_number_override = calloc(job->vec_len, sizeof(int));

    for (size_t i = 0; i < num_particles; i++) {
        job->in_element[i] = which_element(job->particles[i].x,
            job->particles[i].y, job->N, job->h);
        /* reference: `> num_elements` (process_usl.c:239); `>=` here, see header */
        if (job->in_element[i] < 0 || (size_t)job->in_element[i] >= job->num_elements) {
            job->in_element[i] = -1;
            job->active[i] = 0;
        } else {
            job->active[i] = 1;
        }
    }
    for (size_t i = 0; i < num_particles; i++) {
        particle_t *p = &job->particles[i];
        p->v0 = p->v;
        p->r1_initial[0] = 0.5 * sqrt(p->v0);
        p->r1_initial[1] = 0;
        p->r2_initial[0] = 0;
        p->r2_initial[1] = 0.5 * sqrt(p->v0);
    }
    job->use_cpdi = 0;
    job->step_number = 0;
    job->step_start_time = job->t;
}

static void move_grid_split(job_t *job, size_t n_start, size_t n_stop)
{
    for (size_t i = n_start; i < n_stop; i++) {
        node_t *nd = &job->nodes[i];
        double m = nd->m;
        if (m > TOL) {
            nd->x_tt = nd->fx / m;
            nd->y_tt = nd->fy / m;
            nd->mx_t += job->dt * nd->fx;
            nd->my_t += job->dt * nd->fy;
            nd->x_t = nd->mx_t / m;
            nd->y_t = nd->my_t / m;
            nd->ux = job->dt * nd->x_t;
            nd->uy = job->dt * nd->y_t;
        } else {
            nd->x_tt = 0;
            nd->y_tt = 0;
            nd->x_t = 0;
            nd->y_t = 0;
        }
    }
}

static void create_particle_to_element_map_threaded(threadtask_t *task)
{
    job_t *job = task->job;
    oracle_job_t *oj = OJ(job);
    size_t p_start = task->offset;
    size_t p_stop = task->offset + task->blocksize;
    unsigned int changed = 0;
    long long crossed = 0;
    for (size_t i = p_start; i < p_stop; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        int p = which_element(job->particles[i].x, job->particles[i].y, job->N, job->h);
        if (p >= 0 && (size_t)p >= job->num_elements) {
            p = -1;   /* see header: the reference reads past elements[] here */
        }
        if (p != job->in_element[i]) {
            changed = 1;
            crossed++;
            if (oj->prm.log_crossings) {
                fprintf(job->output.log_fd,
                    "[%g] Particle %zu @(%g, %g) left element %d, now in element %d.\n",
                    job->t, i, job->particles[i].x, job->particles[i].y,
                    job->in_element[i], p);
            }
        }
        job->in_element[i] = p;
        if (p == -1) {
            if (oj->prm.log_crossings) {
                fprintf(job->output.log_fd,
                    "[%g] Particle %zu outside of grid (%g, %g), marking as inactive.\n",
                    job->t, i, job->particles[i].x, job->particles[i].y);
            }
            job->active[i] = 0;
            continue;
        }
    }
    job->update_elementlists[task->id] = (int)changed;
    if (crossed) {
        pthread_mutex_lock(&oj->crossings_lock);
        oj->crossings += crossed;
        pthread_mutex_unlock(&oj->crossings_lock);
    }
}

static void find_filled_elements(job_t *job)
{
    size_t i, p, c_idx, t_idx, tc_idx, curr_len;
    for (i = 0; i < job->num_colors; i++) {
        job->color_indices[i] = 0;
    }
    for (i = 0; i < job->num_particles; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        p = job->in_element[i];
        job->elements[p].filled = 1;
        job->elements[p].n++;
        job->elements[p].m += job->particles[i].m;
    }
    for (i = 0; i < job->num_elements; i++) {
        if (job->elements[i].filled) {
            job->elements[i].color_idx = (int)job->color_indices[job->elements[i].color];
            job->color_indices[job->elements[i].color]++;
        }
    }
    for (i = 0; i < (job->num_threads * job->num_colors); i++) {
        job->particle_by_element_color_lengths[i] = 0;
    }
    for (i = 0; i < job->num_particles; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        p = job->in_element[i];
        c_idx = job->elements[p].color;
        t_idx = job->elements[p].color_idx % job->num_threads;
        tc_idx = t_idx * job->num_colors + c_idx;
        curr_len = job->particle_by_element_color_lengths[tc_idx];
        job->particle_by_element_color_lists[tc_idx][curr_len] = i;
        job->particle_by_element_color_lengths[tc_idx]++;
    }
}

static void calculate_shapefunctions_split(job_t *job, size_t p_start, size_t p_stop)
{
    for (size_t i = p_start; i < p_stop; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        size_t p = job->in_element[i];
        size_t n = job->elements[p].nodes[0];
        double xn = job->nodes[n].x;
        double yn = job->nodes[n].y;
        double xl, yl;
        oracle_global_to_local_coords(&xl, &yl, job->particles[i].x,
            job->particles[i].y, xn, yn, job->h);
        oracle_tent(&job->h1[i], &job->h2[i], &job->h3[i], &job->h4[i], xl, yl);
        oracle_grad_tent(&job->b11[i], &job->b12[i], &job->b13[i], &job->b14[i],
            &job->b21[i], &job->b22[i], &job->b23[i], &job->b24[i], xl, yl, job->h);
        job->particles[i].xl = xl;
        job->particles[i].yl = yl;
    }
}

static void calculate_strainrate_split(job_t *job, size_t p_start, size_t p_stop)
{
    for (size_t i = p_start; i < p_stop; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        particle_t *p = &job->particles[i];
        p->exx_t = 0;
        p->exy_t = 0;
        p->eyy_t = 0;
        p->wxy_t = 0;
        p->exx_t = (1.0) * SMEAR4(job, x_t, i, b1);
        p->eyy_t = (1.0) * SMEAR4(job, y_t, i, b2);
        double dx_tdy = (1.0) * SMEAR4(job, x_t, i, b2);
        double dy_tdx = (1.0) * SMEAR4(job, y_t, i, b1);
        p->exy_t = 0.5 * (dx_tdy + dy_tdx);
        p->wxy_t = 0.5 * (dx_tdy - dy_tdx);
    }
}

/* accumulate_p_to_n_ds_list{47,42} (libmpm/map.c:85-104): node-outer,
 * field-inner accumulation of sfvalues[i] * pdata[j]. */
static void accumulate_list(node_t *nodes, const size_t *node_field_offset,
    const int *nodelist, const double *sfvalues, const double *pdata, int Np)
{
    for (int i = 0; i < 4; i++) {
        char *nodeoff = (char *)&nodes[nodelist[i]];
        for (int j = 0; j < Np; j++) {
            double *ndata = (double *)(nodeoff + node_field_offset[j]);
            *ndata += sfvalues[i] * pdata[j];
        }
    }
}

static void map_to_grid_explicit_split(job_t *job, size_t thread_id)
{
    double s[4], ds[4];
    size_t node_field_offsets[7];
    double pdata[7];
    size_t node_d_offsets[3];
    double stressdata[3];

    node_field_offsets[0] = offsetof(node_t, m);
    node_field_offsets[1] = offsetof(node_t, mx_t);
    node_field_offsets[2] = offsetof(node_t, my_t);
    node_field_offsets[3] = offsetof(node_t, mx_tt);
    node_field_offsets[4] = offsetof(node_t, my_tt);
    node_field_offsets[5] = offsetof(node_t, fx);
    node_field_offsets[6] = offsetof(node_t, fy);
    node_d_offsets[0] = offsetof(node_t, fx);
    node_d_offsets[1] = offsetof(node_t, fy);

    for (size_t c = 0; c < job->num_colors; c++) {
        size_t tc_idx = thread_id * job->num_colors + c;
        for (size_t i = 0; i < job->particle_by_element_color_lengths[tc_idx]; i++) {
            size_t p_idx = job->particle_by_element_color_lists[tc_idx][i];
            int p = job->in_element[p_idx];
            const particle_t *pp = &job->particles[p_idx];
            s[0] = job->h1[p_idx];
            s[1] = job->h2[p_idx];
            s[2] = job->h3[p_idx];
            s[3] = job->h4[p_idx];
            pdata[0] = pp->m;
            pdata[1] = pp->x_t * pp->m;
            pdata[2] = pp->y_t * pp->m;
            pdata[3] = pp->x_tt * pp->m;
            pdata[4] = pp->y_tt * pp->m;
            pdata[5] = pp->bx * pp->m;
            pdata[6] = pp->by * pp->m;
            stressdata[0] = -pp->sxx * pp->v;
            stressdata[1] = -pp->sxy * pp->v;
            stressdata[2] = -pp->syy * pp->v;
            accumulate_list(job->nodes, node_field_offsets, job->elements[p].nodes,
                s, pdata, 7);
            ds[0] = job->b11[p_idx];
            ds[1] = job->b12[p_idx];
            ds[2] = job->b13[p_idx];
            ds[3] = job->b14[p_idx];
            accumulate_list(job->nodes, node_d_offsets, job->elements[p].nodes,
                ds, &stressdata[0], 2);
            ds[0] = job->b21[p_idx];
            ds[1] = job->b22[p_idx];
            ds[2] = job->b23[p_idx];
            ds[3] = job->b24[p_idx];
            accumulate_list(job->nodes, node_d_offsets, job->elements[p].nodes,
                ds, &stressdata[1], 2);
        }
        pthread_barrier_wait(job->serialize_barrier);
    }
}

static void move_particles_explicit_usl_split(job_t *job, size_t p_start, size_t p_stop)
{
    for (size_t i = p_start; i < p_stop; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        particle_t *pp = &job->particles[i];
        double s[4];
        s[0] = job->h1[i];
        s[1] = job->h2[i];
        s[2] = job->h3[i];
        s[3] = job->h4[i];
        size_t el = job->in_element[i];
        pp->x_tt = 0;
        pp->y_tt = 0;
        for (size_t j = 0; j < 4; j++) {
            size_t n = job->elements[el].nodes[j];
            double m = job->nodes[n].m;
            if (m < TOL) {
                continue;
            }
            pp->x_tt += ((s[j] * job->nodes[n].fx) / m);
            pp->y_tt += ((s[j] * job->nodes[n].fy) / m);
        }
        pp->x_t += job->dt * pp->x_tt;
        pp->y_t += job->dt * pp->y_tt;
        double dux = 0;
        double duy = 0;
        for (size_t j = 0; j < 4; j++) {
            size_t n = job->elements[el].nodes[j];
            double m = job->nodes[n].m;
            if (m < TOL) {
                continue;
            }
            dux += ((s[j] * job->nodes[n].mx_t) / m);
            duy += ((s[j] * job->nodes[n].my_t) / m);
        }
        dux *= job->dt;
        duy *= job->dt;
        pp->x += dux;
        pp->y += duy;
        pp->ux += dux;
        pp->uy += duy;
    }
}

static void update_particle_densities_split(job_t *job, size_t p_start, size_t p_stop)
{
    for (size_t i = p_start; i < p_stop; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        job->particles[i].v = job->particles[i].v *
            ORACLE_EXP(job->dt * (job->particles[i].exx_t + job->particles[i].eyy_t));
    }
}

static void clear_nodes(job_t *job, size_t n_start, size_t n_stop)
{
    for (size_t i = n_start; i < n_stop; i++) {
        node_t *nd = &job->nodes[i];
        nd->m = 0;
        nd->mx_t = 0;
        nd->my_t = 0;
        nd->mx_tt = 0;
        nd->my_tt = 0;
        nd->x_t = 0;
        nd->y_t = 0;
        nd->x_tt = 0;
        nd->y_tt = 0;
        nd->fx = 0;
        nd->fy = 0;
        nd->ux = 0;
        nd->uy = 0;
    }
}

/* The material's calculate_stress_threaded (process_usl.c:589). */
static void stress_update(threadtask_t *task)
{
    job_t *job = task->job;
    oracle_job_t *oj = OJ(job);
    if (oj->prm.material_kind == 1) {
        /* g_nonlocal runs on one thread (g_nonlocal_mu2.c:113-123); it reads
         * every particle's strain rate, so the others must be done */
        pthread_barrier_wait(job->serialize_barrier);
        if (task->id == 0) {
            g_nonlocal_stress(oj);
        }
    } else if (oj->prm.material_kind == 2) {
        isolin_stress_threaded(task);
    } else if (oj->prm.material_kind == 3) {
        dp_ri_mt_stress_threaded(task);
    } else {
        g_local_mu2_stress_threaded(task);
    }
}

static void explicit_mpm_step_usl_threaded(threadtask_t *task)
{
    job_t *job = task->job;
    oracle_job_t *oj = OJ(job);
    int rc;
    size_t p_start = task->offset;
    size_t p_stop = task->offset + task->blocksize;
    size_t n_start = task->n_offset;
    size_t n_stop = task->n_offset + task->n_blocksize;
    size_t e_start = task->e_offset;
    size_t e_stop = task->e_offset + task->e_blocksize;

    pthread_barrier_wait(job->serialize_barrier);
    clear_nodes(job, n_start, n_stop);
    for (size_t i = e_start; i < e_stop; i++) {
        job->elements[i].filled = 0;
        job->elements[i].n = 0;
        job->elements[i].m = 0;
    }
    create_particle_to_element_map_threaded(task);
    calculate_shapefunctions_split(job, p_start, p_stop);

    rc = pthread_barrier_wait(job->serialize_barrier);
    if (rc == PTHREAD_BARRIER_SERIAL_THREAD) {
        job->t += job->dt;
        int flag = 0;
        for (size_t i = 0; i < job->num_threads; i++) {
            flag += job->update_elementlists[i];
        }
        if (flag != 0) {
            find_filled_elements(job);
        }
        bc_time_varying(oj);
    }
    pthread_barrier_wait(job->serialize_barrier);
    map_to_grid_explicit_split(job, task->id);

    rc = pthread_barrier_wait(job->serialize_barrier);
    if (rc == PTHREAD_BARRIER_SERIAL_THREAD) {
        bc_momentum(oj);
        bc_force(oj);
    }
    pthread_barrier_wait(job->serialize_barrier);
    move_grid_split(job, n_start, n_stop);
    pthread_barrier_wait(job->serialize_barrier);

    move_particles_explicit_usl_split(job, p_start, p_stop);
    calculate_strainrate_split(job, p_start, p_stop);
    update_particle_densities_split(job, p_start, p_stop);
    stress_update(task);
}

/* explicit_mpm_step_musl_threaded (process_usl.c:284-462): the USL step up
 * to the particle update, then the particle momentum is mapped to the grid
 * again with the step's shape functions, the node velocities recomputed
 * (process_usl.c:429-440; bc_momentum then zeroes the node momenta, not the
 * velocities the strain rate reads, 441-445), and the strain rate, volume
 * and stress follow from those velocities. */
static void explicit_mpm_step_musl_threaded(threadtask_t *task)
{
    job_t *job = task->job;
    oracle_job_t *oj = OJ(job);
    int rc;
    size_t p_start = task->offset;
    size_t p_stop = task->offset + task->blocksize;
    size_t n_start = task->n_offset;
    size_t n_stop = task->n_offset + task->n_blocksize;
    size_t e_start = task->e_offset;
    size_t e_stop = task->e_offset + task->e_blocksize;

    pthread_barrier_wait(job->serialize_barrier);
    clear_nodes(job, n_start, n_stop);
    for (size_t i = e_start; i < e_stop; i++) {
        job->elements[i].filled = 0;
        job->elements[i].n = 0;
        job->elements[i].m = 0;
    }
    create_particle_to_element_map_threaded(task);
    calculate_shapefunctions_split(job, p_start, p_stop);

    rc = pthread_barrier_wait(job->serialize_barrier);
    if (rc == PTHREAD_BARRIER_SERIAL_THREAD) {
        job->t += job->dt;
        int flag = 0;
        for (size_t i = 0; i < job->num_threads; i++) {
            flag += job->update_elementlists[i];
        }
        if (flag != 0) {
            find_filled_elements(job);
        }
        bc_time_varying(oj);
    }
    pthread_barrier_wait(job->serialize_barrier);
    map_to_grid_explicit_split(job, task->id);

    rc = pthread_barrier_wait(job->serialize_barrier);
    if (rc == PTHREAD_BARRIER_SERIAL_THREAD) {
        bc_momentum(oj);
        bc_force(oj);
    }
    pthread_barrier_wait(job->serialize_barrier);
    move_grid_split(job, n_start, n_stop);
    pthread_barrier_wait(job->serialize_barrier);
    move_particles_explicit_usl_split(job, p_start, p_stop);

    pthread_barrier_wait(job->serialize_barrier);
    clear_nodes(job, n_start, n_stop);
    pthread_barrier_wait(job->serialize_barrier);
    map_to_grid_explicit_split(job, task->id);
    pthread_barrier_wait(job->serialize_barrier);
    for (size_t i = n_start; i < n_stop; i++) {
        node_t *nd = &job->nodes[i];
        if (nd->m > TOL) {
            nd->x_t = nd->mx_t / nd->m;
            nd->y_t = nd->my_t / nd->m;
        } else {
            nd->x_t = 0;
            nd->y_t = 0;
        }
    }
    rc = pthread_barrier_wait(job->serialize_barrier);
    if (rc == PTHREAD_BARRIER_SERIAL_THREAD) {
        bc_momentum(oj);
    }
    pthread_barrier_wait(job->serialize_barrier);
    calculate_strainrate_split(job, p_start, p_stop);
    update_particle_densities_split(job, p_start, p_stop);
    stress_update(task);
}

/* mpm_run_until (src/main.c:934-982) with an exact step count in place of
 * the t_stop test; the frame writer is not part of the step. */
static void *run_steps(void *arg)
{
    threadtask_t *task = arg;
    job_t *job = task->job;
    oracle_job_t *oj = OJ(job);
    for (int s = 0; s < oj->steps_to_run; s++) {
        if (OJ(job)->prm.scheme == 1) {
            explicit_mpm_step_musl_threaded(task);
        } else {
            explicit_mpm_step_usl_threaded(task);
        }
        int rc = pthread_barrier_wait(job->serialize_barrier);
        if (rc == PTHREAD_BARRIER_SERIAL_THREAD) {
            job->stepcount++;
            time_varying_loads(oj);
        }
    }
    return NULL;
}

/* ------------------------------------------------------------------------ */
/* C API used by the tests and the CPU-baseline leg of the benchmark.        */
/* ------------------------------------------------------------------------ */
void oracle_default_params(oracle_params_t *q)
{
    memset(q, 0, sizeof *q);
    q->num_threads = 1;
    q->E = 1e6;
    q->nu = 0.3;
    q->mu_s = 0.3819;
    q->mu_2 = 0.6435;
    q->I_0 = (0.278 / 1.0);
    q->grains_d = (0.001 * 5);
    q->grains_rho = 2450;
    q->rho_crit = 1485.0f;
    q->g_mag = 9.8;
    q->ramp_time = 0.1;
    q->material_kind = 0;
    q->nl_E = 1e8;
    q->nl_nu = 0.3;
    q->nl_mu_s = 0.3819;
    q->nl_A = 0.48;
    q->nl_B = (0.9377 * 20);
    q->nl_d = (0.01 * 5);
    q->nl_rho = 2450;
    q->nl_p_cap = 1e-2;
    q->nl_c = 1e-2;
}

size_t oracle_params_size(void)
{
    return sizeof(oracle_params_t);
}

/* Builds the job as the driver does (src/main.c:405-839): mpm_init, thread
 * split, colour lists, find_filled_elements, initial_loads, material_init,
 * bc_validate.  Particle fields: m v x y x_t y_t sxx sxy syy (reader.c:40-119). */
void *oracle_create(const oracle_params_t *prm, size_t np,
    const double *m, const double *v, const double *x, const double *y,
    const double *vx, const double *vy,
    const double *sxx, const double *sxy, const double *syy)
{
    oracle_job_t *oj = calloc(1, sizeof *oj);
    oj->prm = *prm;
    pthread_mutex_init(&oj->crossings_lock, NULL);
    particle_t *pdata = calloc(np ? np : 1, sizeof(particle_t));
    for (size_t i = 0; i < np; i++) {
        pdata[i].m = m[i];
        pdata[i].v = v[i];
        pdata[i].x = x[i];
        pdata[i].y = y[i];
        pdata[i].x_t = vx[i];
        pdata[i].y_t = vy[i];
        pdata[i].sxx = sxx ? sxx[i] : 0;
        pdata[i].sxy = sxy ? sxy[i] : 0;
        pdata[i].syy = syy ? syy[i] : 0;
    }
    oracle_mpm_init(oj, prm->N, prm->h, pdata, np);
    free(pdata);

    job_t *job = &oj->job;
    job->dt = prm->dt;
    size_t T = prm->num_threads > 0 ? (size_t)prm->num_threads : 1;
    job->num_threads = T;
    oj->tasks = calloc(T, sizeof(threadtask_t));
    size_t psplit = (job->num_particles / T) + ((job->num_particles % T != 0) ? 1 : 0);
    size_t nsplit = (job->num_nodes / T) + ((job->num_nodes % T != 0) ? 1 : 0);
    size_t esplit = (job->num_elements / T) + ((job->num_elements % T != 0) ? 1 : 0);
#define MINZ(a, b) ((a) < (b) ? (a) : (b))
    for (size_t i = 0; i < T; i++) {
        threadtask_t *tk = &oj->tasks[i];
        tk->id = i;
        tk->num_threads = T;
        tk->job = job;
        tk->offset = MINZ(i * psplit, job->num_particles);
        tk->blocksize = MINZ(psplit, job->num_particles - tk->offset);
        tk->n_offset = MINZ(i * nsplit, job->num_nodes);
        tk->n_blocksize = MINZ(nsplit, job->num_nodes - tk->n_offset);
        tk->e_offset = MINZ(i * esplit, job->num_elements);
        tk->e_blocksize = MINZ(esplit, job->num_elements - tk->e_offset);
        tk->gidx_max = job->num_particles;
        tk->gidx_min = 0;
    }
    job->serialize_barrier = malloc(sizeof(pthread_barrier_t));
    job->step_barrier = malloc(sizeof(pthread_barrier_t));
    pthread_barrier_init(job->serialize_barrier, NULL, (unsigned)T);
    pthread_barrier_init(job->step_barrier, NULL, (unsigned)T);
    job->update_elementlists = malloc(sizeof(int) * T);
    for (size_t i = 0; i < T; i++) {
        job->update_elementlists[i] = 1;
    }
    job->particle_by_element_color_lengths = calloc(job->num_colors * T, sizeof(size_t));
    job->particle_by_element_color_lists = malloc(sizeof(size_t *) * job->num_colors * T);
    for (size_t i = 0; i < job->num_colors * T; i++) {
        job->particle_by_element_color_lists[i] = malloc(sizeof(size_t) * (np ? np : 1));
    }
    for (size_t i = 0; i < job->num_elements; i++) {
        job->elements[i].filled = 0;
        job->elements[i].n = 0;
        job->elements[i].m = 0;
    }
    find_filled_elements(job);
    initial_loads(job);
    g_local_mu2_init(oj);
    if (prm->material_kind == 1) {
        /* g_nonlocal.c material_init (63-93): all state zero */
        for (size_t i = 0; i < job->num_particles; i++) {
            memset(job->particles[i].state, 0, sizeof job->particles[i].state);
        }
    }
    if (prm->log_crossings) {
        job->output.log_fd = fopen(prm->log_path ? prm->log_path : "/dev/null", "w");
    }
    int ok = 1;
    if (prm->bc_kind == 1) {
        ok = box_validate(oj);
    } else if (prm->bc_kind == 2) {
        ok = silo_validate(oj);
    }
    if (!ok) {
        fprintf(stderr, "oracle: boundary condition properties rejected\n");
    }
    job->stepcount = 0;
    return oj;
}

/* Runs `nsteps` USL steps on the configured number of threads. */
int oracle_run(void *handle, int nsteps)
{
    oracle_job_t *oj = handle;
    job_t *job = &oj->job;
    size_t T = job->num_threads;
    oj->steps_to_run = nsteps;
    pthread_t *th = malloc(sizeof(pthread_t) * T);
    for (size_t i = 0; i + 1 < T; i++) {
        pthread_create(&th[i], NULL, run_steps, &oj->tasks[i]);
    }
    run_steps(&oj->tasks[T - 1]);
    for (size_t i = 0; i + 1 < T; i++) {
        pthread_join(th[i], NULL);
    }
    free(th);
    return 0;
}

/* Particle fields, one array of np doubles per field, in this order. */
enum {
    OF_X, OF_Y, OF_VX, OF_VY, OF_M, OF_V, OF_TXX, OF_TXY, OF_TYY, OF_TZZ,
    OF_GAMMAP, OF_GAMMADOTP, OF_UX, OF_UY, OF_EXX, OF_EXY, OF_EYY, OF_WXY,
    OF_AX, OF_AY, OF_XL, OF_YL, OF_SXX, OF_SXY, OF_SYY, OF_TXZ, OF_TYX,
    OF_TYZ, OF_TZX, OF_TZY, OF_BY, OF_GF, OF_GF_BULK, OF_XISQ_INV, OF_STATE1, OF_NFIELDS
};

int oracle_num_particle_fields(void)
{
    return OF_NFIELDS;
}

void oracle_get_particles(void *handle, double *out, int *active, int *in_element)
{
    oracle_job_t *oj = handle;
    job_t *job = &oj->job;
    size_t np = job->num_particles;
    for (size_t i = 0; i < np; i++) {
        const particle_t *p = &job->particles[i];
        double f[OF_NFIELDS] = {
            p->x, p->y, p->x_t, p->y_t, p->m, p->v, p->T[XX], p->T[XY], p->T[YY],
            p->T[ZZ], p->state[9], p->state[10], p->ux, p->uy, p->exx_t, p->exy_t,
            p->eyy_t, p->wxy_t, p->x_tt, p->y_tt, p->xl, p->yl, p->sxx, p->sxy,
            p->syy, p->T[XZ], p->T[YX], p->T[YZ], p->T[ZX], p->T[ZY], p->by,
            p->state[3], p->state[5], p->state[10], p->state[1] };
        for (int k = 0; k < OF_NFIELDS; k++) {
            out[(size_t)k * np + i] = f[k];
        }
        if (active) {
            active[i] = job->active[i];
        }
        if (in_element) {
            in_element[i] = job->in_element[i];
        }
    }
}

/* Node fields after the last step: m mx_t my_t fx fy x_t y_t x_tt y_tt. */
void oracle_get_nodes(void *handle, double *out)
{
    oracle_job_t *oj = handle;
    job_t *job = &oj->job;
    size_t nn = job->num_nodes;
    for (size_t i = 0; i < nn; i++) {
        const node_t *n = &job->nodes[i];
        out[0 * nn + i] = n->m;
        out[1 * nn + i] = n->mx_t;
        out[2 * nn + i] = n->my_t;
        out[3 * nn + i] = n->fx;
        out[4 * nn + i] = n->fy;
        out[5 * nn + i] = n->x_t;
        out[6 * nn + i] = n->y_t;
        out[7 * nn + i] = n->x_tt;
        out[8 * nn + i] = n->y_tt;
    }
}

/* Boundary masks as the BC plugin left them: u_dirichlet_mask[3*Nn]. */
void oracle_get_bc_mask(void *handle, int *mask)
{
    oracle_job_t *oj = handle;
    memcpy(mask, oj->job.u_dirichlet_mask, sizeof(int) * NODAL_DOF * oj->job.num_nodes);
}

double oracle_time(void *handle)
{
    return OJ(handle)->job.t;
}

long long oracle_crossings(void *handle)
{
    return OJ(handle)->crossings;
}

void oracle_destroy(void *handle)
{
    oracle_job_t *oj = handle;
    job_t *job = &oj->job;
    if (job->output.log_fd) {
        fclose(job->output.log_fd);
    }
    for (size_t i = 0; i < job->num_colors * job->num_threads; i++) {
        free(job->particle_by_element_color_lists[i]);
    }
    free(job->particle_by_element_color_lists);
    free(job->particle_by_element_color_lengths);
    free(job->update_elementlists);
    pthread_barrier_destroy(job->serialize_barrier);
    pthread_barrier_destroy(job->step_barrier);
    free(job->serialize_barrier);
    free(job->step_barrier);
    free(oj->tasks);
    free(job->nodes);
    free(job->particles);
    free(job->elements);
    free(job->in_element);
    free(job->active);
    free(job->h1); free(job->h2); free(job->h3); free(job->h4);
    free(job->b11); free(job->b12); free(job->b13); free(job->b14);
    free(job->b21); free(job->b22); free(job->b23); free(job->b24);
    free(job->u_dirichlet);
    free(job->u_dirichlet_mask);
    free(job->node_number_override);
    free(job->color_indices);
    pthread_mutex_destroy(&oj->crossings_lock);
    free(oj);
}

/* Unit hooks for the tests. */
double oracle_exp(double x)
{
    return ORACLE_EXP(x);
}

/* Single-particle driver in the manner of tests/test_material.c:89-195:
 * applies the given velocity gradient (exx, exy, eyy, wxy) for one step to
 * the state (T[4] = xx, xy, yy, zz; m, v; gammap) and returns the result in
 * out[6] = {Txx, Txy, Tyy, Tzz, gammap, gammadotp}. */
/* One isolin update: s = {sxx, sxy, syy} -> out (test hook). */
void oracle_isolin_point(const double *sin, double exx, double exy, double eyy,
    double wxy, double dt, const oracle_params_t *q, double *out)
{
    particle_t p;
    memset(&p, 0, sizeof p);
    p.sxx = sin[0];
    p.sxy = sin[1];
    p.syy = sin[2];
    p.exx_t = exx;
    p.exy_t = exy;
    p.eyy_t = eyy;
    p.wxy_t = wxy;
    double G = q->E / (2.0 * (1.0 + q->nu));
    double K = q->E / (3.0 * (1.0 - 2 * q->nu));
    isolin_point(&p, dt, G, K);
    out[0] = p.sxx;
    out[1] = p.sxy;
    out[2] = p.syy;
}

/* One dp_ri_mt update: s = {sxx, sxy, syy, szz, gammap} -> out = {sxx,
 * sxy, syy, szz, gammap, gammadotp} (test hook). */
void oracle_dp_ri_mt_point(const double *sin, double m, double v, double exx, double exy,
    double eyy, double wxy, double dt, const oracle_params_t *q, double *out)
{
    particle_t p;
    memset(&p, 0, sizeof p);
    p.sxx = sin[0];
    p.sxy = sin[1];
    p.syy = sin[2];
    p.state[1] = sin[3];
    p.state[9] = sin[4];
    p.m = m;
    p.v = v;
    p.exx_t = exx;
    p.exy_t = exy;
    p.eyy_t = eyy;
    p.wxy_t = wxy;
    double G = q->E / (2.0 * (1.0 + q->nu));
    double K = q->E / (3.0 * (1.0 - 2 * q->nu));
    dp_ri_mt_point(&p, dt, G, K, q->mu_s);
    out[0] = p.sxx;
    out[1] = p.sxy;
    out[2] = p.syy;
    out[3] = p.state[1];
    out[4] = p.state[9];
    out[5] = p.state[10];
}

void oracle_stress_point(const double *Tin, double m, double v, double gammap,
    double exx, double exy, double eyy, double wxy, double dt,
    const oracle_params_t *q, double *out)
{
    particle_t p;
    memset(&p, 0, sizeof p);
    p.T[XX] = Tin[0];
    p.T[XY] = Tin[1];
    p.T[YX] = Tin[1];
    p.T[YY] = Tin[2];
    p.T[ZZ] = Tin[3];
    p.m = m;
    p.v = v;
    p.state[9] = gammap;
    p.exx_t = exx;
    p.exy_t = exy;
    p.eyy_t = eyy;
    p.wxy_t = wxy;
    double G = q->E / (2.0 * (1.0 + q->nu));
    double K = q->E / (3.0 * (1.0 - 2 * q->nu));
    oracle_g_local_mu2_point(&p, dt, G, K, q->mu_s, q->mu_2, q->I_0,
        q->grains_d, q->grains_rho, q->rho_crit);
    out[0] = p.T[XX];
    out[1] = p.T[XY];
    out[2] = p.T[YY];
    out[3] = p.T[ZZ];
    out[4] = p.state[9];
    out[5] = p.state[10];
}
