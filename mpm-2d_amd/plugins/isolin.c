/*
 * isolin material plugin: isotropic linear elasticity with the Jaumann
 * stress rate, on the in-plane stress (plugin ABI of materialsrc/isolin.c:
 * material_init 30-58, calculate_stress 62-71, calculate_stress_threaded
 * 74-116).
 *
 * properties = {E, nu}
 *
 * material_gpu_descriptor() lets the GPU step run this law on the device;
 * the host functions below are the same law for host use (and the fallback
 * a build with -DMPM_NO_GPU_DESCRIPTOR exercises).
 */
#include <stdio.h>
#include <stdlib.h>

#include "mpm_abi.h"
#include "mpm_dropin.h"
#include "mpm_plugin_gpu.h"

static double E_, nu_, G_, K_, lambda_;

void material_init(job_t *job)
{
    for (size_t i = 0; i < job->num_particles; i++) {
        for (int j = 0; j < DEPVAR; j++) {
            job->particles[i].state[j] = 0;
        }
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
    lambda_ = K_ - 2.0 * G_ / 3.0;
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

void calculate_stress_threaded(threadtask_t *task)
{
    job_t *job = task->job;
    for (size_t i = task->offset; i < task->offset + task->blocksize; i++) {
        if (job->active[i] == 0) {
            continue;
        }
        particle_t *p = &job->particles[i];
        /* Hooke's law on the strain rate plus the Jaumann terms W s - s W */
        const double trD = p->exx_t + p->eyy_t;
        double rxx = lambda_ * trD + 2.0 * G_ * p->exx_t;
        double rxy = 2.0 * G_ * p->exy_t;
        double ryy = lambda_ * trD + 2.0 * G_ * p->eyy_t;
        rxx += 2 * p->wxy_t * p->sxy;
        rxy -= p->wxy_t * (p->sxx - p->syy);
        ryy -= 2 * p->wxy_t * p->sxy;
        p->sxx += job->dt * rxx;
        p->sxy += job->dt * rxy;
        p->syy += job->dt * ryy;
    }
}

#ifndef MPM_NO_GPU_DESCRIPTOR
int material_gpu_descriptor(const job_t *job, mpm_material_gpu_desc *d)
{
    (void)job;
    d->abi = MPM_PLUGIN_GPU_ABI;
    d->kind = MPM_GPU_MATERIAL_ISOLIN;
    for (int k = 0; k < 8; k++) {
        d->params[k] = 0.0;
    }
    d->params[0] = E_;
    d->params[1] = nu_;
    return 1;
}
#endif
