/*
 * g_local_mu2 material plugin: local mu(I) granular plasticity with a
 * Jaumann stress rate and a density cut-off (plugin ABI of
 * materialsrc/g_local_mu2.c: material_init 102-151, calculate_stress
 * 154-163, calculate_stress_threaded 165-311).
 *
 * properties = {E, nu [, mu_s, mu_2, I_0, grains_d, grains_rho, rho_crit]}
 * The optional ones default to the constants the reference hard-codes
 * (g_local_mu2.c:30-41, 247).
 *
 * material_gpu_descriptor() lets the GPU step run this law on the device;
 * the host functions below are the same law for host use (and the fallback
 * a build with -DMPM_NO_GPU_DESCRIPTOR exercises).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "mpm_abi.h"
#include "mpm_dropin.h"
#include "mpm_plugin_gpu.h"

static double E_, nu_, G_, K_;
static double prm[6] = {0.3819, 0.6435, (0.278 / 1.0), (0.001 * 5), 2450, 1485.0f};
enum { MU_S, MU_2, I_0, GRAINS_D, GRAINS_RHO, RHO_CRIT };

void material_init(job_t *job)
{
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
    if (job->material.num_fp64_props < 2) {
        fprintf(stderr, "%s:%s: Need at least 2 properties defined (E, nu).\n",
                __FILE__, __func__);
        exit(EXIT_ERROR_MATERIAL_FILE);
    }
    E_ = job->material.fp64_props[0];
    nu_ = job->material.fp64_props[1];
    G_ = E_ / (2.0 * (1.0 + nu_));
    K_ = E_ / (3.0 * (1.0 - 2 * nu_));
    for (size_t k = 2; k < job->material.num_fp64_props && k < 8; k++) {
        prm[k - 2] = job->material.fp64_props[k];
    }
    printf("%s:%s: properties (E = %g, nu = %g, G = %g, K = %g).\n",
           __FILE__, __func__, E_, nu_, G_, K_);
}

void calculate_stress_threaded(threadtask_t *task);

void calculate_stress(job_t *job)
{
    threadtask_t t;
    t.job = job;
    t.offset = 0;
    t.blocksize = job->num_particles;
    calculate_stress_threaded(&t);
}

static double smaller_root(double b, double c)
{
    /* the smaller root of x^2 + b x + c, without cancellation */
    if (b > 0) {
        return (-b - sqrt(b * b - 4 * c)) / 2;
    }
    return (2 * c) / (-b + sqrt(b * b - 4 * c));
}

void calculate_stress_threaded(threadtask_t *task)
{
    job_t *job = task->job;
    const double dt = job->dt;
    const double lambda = K_ - 2.0 * G_ / 3.0;
    for (size_t i = task->offset; i < task->offset + task->blocksize; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        particle_t *p = &job->particles[i];
        double W[9], D[9], T0[9], Ttr[9], JS[9], tmp[9];
        double trD, p_tr, tau2;
        p->L[XX] = p->exx_t;
        p->L[XY] = p->exy_t + p->wxy_t;
        p->L[XZ] = 0;
        p->L[YX] = p->exy_t - p->wxy_t;
        p->L[YY] = p->eyy_t;
        p->L[YZ] = 0;
        p->L[ZX] = 0;
        p->L[ZY] = 0;
        p->L[ZZ] = 0;
        tensor_skw3(W, p->L);
        tensor_sym3(D, p->L);
        tensor_copy3(Ttr, D);
        tensor_trace3(&trD, D);
        /* objective (Jaumann) rate term W T - T W */
        tensor_multiply3(JS, W, p->T);
        tensor_multiply3(tmp, p->T, W);
        tensor_scale3(tmp, -1.0);
        tensor_add3(JS, JS, tmp);
        /* elastic trial stress T + dt (2G D + lambda tr(D) I + JS) */
        tensor_scale3(Ttr, 2.0 * G_);
        Ttr[XX] += lambda * trD;
        Ttr[YY] += lambda * trD;
        Ttr[ZZ] += lambda * trD;
        tensor_add3(Ttr, Ttr, JS);
        tensor_scale3(Ttr, dt);
        tensor_add3(Ttr, Ttr, p->T);
        tensor_decompose3(T0, &p_tr, Ttr);
        tensor_contraction3(&tau2, T0, T0);
        const double tau_tr = sqrt(0.5 * tau2);
        p_tr *= -1.0;

        double nup = 0;
        if ((p->m / p->v) < prm[RHO_CRIT] || p_tr <= 0) {
            for (int k = 0; k < 9; k++) {
                p->T[k] = 0;
            }
        } else if (p_tr > 0) {
            const double S0 = prm[MU_S] * p_tr;
            double tau, scale;
            if (tau_tr <= S0) {
                tau = tau_tr;
                scale = 1.0;
            } else {
                const double S2 = prm[MU_2] * p_tr;
                const double alpha = G_ * prm[I_0] * dt * sqrt(p_tr / prm[GRAINS_RHO])
                                     / prm[GRAINS_D];
                const double B = -(S2 + tau_tr + alpha);
                const double H = S2 * tau_tr + S0 * alpha;
                tau = smaller_root(B, H);
                scale = tau / tau_tr;
            }
            nup = ((tau_tr - tau) / G_) / dt;
            tensor_scale3(T0, scale);
            tensor_copy3(p->T, T0);
            p->T[XX] -= p_tr;
            p->T[YY] -= p_tr;
            p->T[ZZ] -= p_tr;
        }
        p->sxx = p->T[XX];
        p->sxy = p->T[XY];
        p->syy = p->T[YY];
        p->state[9] += nup * dt;
        p->state[10] = nup;
    }
}

#ifndef MPM_NO_GPU_DESCRIPTOR
int material_gpu_descriptor(const job_t *job, mpm_material_gpu_desc *d)
{
    (void)job;
    d->abi = MPM_PLUGIN_GPU_ABI;
    d->kind = MPM_GPU_MATERIAL_G_LOCAL_MU2;
    d->params[0] = E_;
    d->params[1] = nu_;
    for (int k = 0; k < 6; k++) {
        d->params[2 + k] = prm[k];
    }
    return 1;
}
#endif
