/*
 * dp_ri_mt material plugin: rate-independent Drucker-Prager plasticity
 * (tau <= mu_s p) with the Jaumann stress rate and a density cut-off
 * (plugin ABI of materialsrc/dp_ri_mt.c: material_init 51-88,
 * calculate_stress 91-99, calculate_stress_threaded 101-209).
 *
 * properties = {E, nu, mu_s}
 * State: state[1] szz, state[9] accumulated plastic shear strain,
 * state[10] plastic shear rate.
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

static double E_, nu_, G_, K_, mu_s_;

void material_init(job_t *job)
{
    for (size_t i = 0; i < job->num_particles; i++) {
        for (int j = 0; j < DEPVAR; j++) {
            job->particles[i].state[j] = 0;
        }
    }
    if (job->material.num_fp64_props < 3) {
        fprintf(stderr, "%s:%s: Need at least 3 properties defined (E, nu, mu_s).\n",
                __FILE__, __func__);
        exit(EXIT_ERROR_MATERIAL_FILE);
    }
    E_ = job->material.fp64_props[0];
    nu_ = job->material.fp64_props[1];
    mu_s_ = job->material.fp64_props[2];
    G_ = E_ / (2.0 * (1.0 + nu_));
    K_ = E_ / (3.0 * (1.0 - 2 * nu_));
    printf("%s:%s: properties (E = %g, nu = %g, G = %g, K = %g, mu_s = %g).\n",
           __FILE__, __func__, E_, nu_, G_, K_, mu_s_);
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
        /* elastic trial stress with the Jaumann terms */
        const double trD = p->exx_t + p->eyy_t;
        double rxx = lambda * trD + 2.0 * G_ * p->exx_t;
        double rxy = 2.0 * G_ * p->exy_t;
        double ryy = lambda * trD + 2.0 * G_ * p->eyy_t;
        rxx += 2 * p->wxy_t * p->sxy;
        rxy -= p->wxy_t * (p->sxx - p->syy);
        ryy -= 2 * p->wxy_t * p->sxy;
        const double txx = p->sxx + dt * rxx;
        const double txy = p->sxy + dt * rxy;
        const double tyy = p->syy + dt * ryy;
        const double tzz = p->state[1] + dt * lambda * trD;
        /* pressure, deviator and equivalent shear stress of the trial state */
        const double pr = -(txx + tyy + tzz) / 3.0;
        const double dxx = txx + pr, dxy = txy, dyy = tyy + pr, dzz = tzz + pr;
        const double tau = sqrt(0.5 * (dxx * dxx + 2 * dxy * dxy + dyy * dyy + dzz * dzz));
        double rate;
        if ((p->m / p->v) < 1485.0f || pr <= 0) {
            /* too loose or in tension: stress-free */
            rate = (tau) / (G_ * dt);
            p->sxx = 0;
            p->sxy = 0;
            p->syy = 0;
            p->state[1] = 0;
        } else if (pr > 0) {
            /* return to the cone tau = mu_s p */
            const double limit = mu_s_ * pr;
            double tau_new, s;
            if (tau <= limit) {
                tau_new = tau;
                s = 1.0;
            } else {
                tau_new = limit;
                s = (tau_new / tau);
            }
            rate = ((tau - tau_new) / G_) / dt;
            p->sxx = s * dxx - pr;
            p->sxy = s * dxy;
            p->syy = s * dyy - pr;
            p->state[1] = s * dzz - pr;
        } else {
            rate = 0;
        }
        p->state[9] += rate * dt;
        p->state[10] = rate;
    }
}

#ifndef MPM_NO_GPU_DESCRIPTOR
int material_gpu_descriptor(const job_t *job, mpm_material_gpu_desc *d)
{
    (void)job;
    d->abi = MPM_PLUGIN_GPU_ABI;
    d->kind = MPM_GPU_MATERIAL_DP_RI_MT;
    for (int k = 0; k < 8; k++) {
        d->params[k] = 0.0;
    }
    d->params[0] = E_;
    d->params[1] = nu_;
    d->params[2] = mu_s_;
    return 1;
}
#endif
