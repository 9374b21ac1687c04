/*
 * square_silo boundary-condition plugin (plugin ABI of bcsrc/square_silo.c):
 * a floor with a trapdoor of half-width hole_radius at x = 0 that opens once
 * t > open_time, a frictionless wall at x = 0 and another at column
 * (N - 1) * wall_frac.
 *
 * properties = {hole_radius, open_time, wall_frac}     (square_silo.c:32-48)
 */
#include <stdio.h>
#include <string.h>

#include "mpm_abi.h"
#include "mpm_plugin_gpu.h"

static double hole_radius, open_time, wall_frac;
static size_t wall_col;

void bc_init(job_t *job)
{
    (void)job;
}

int bc_validate(job_t *job)
{
    if (job->boundary.num_fp64_props != 3) {
        fprintf(stderr, "%s:%s: Wrong number of floating point boundary condition "
                "properties. Expected 3, got %zu.\n", __FILE__, __func__,
                job->boundary.num_fp64_props);
        return 0;
    }
    hole_radius = job->boundary.fp64_props[0];
    open_time = job->boundary.fp64_props[1];
    wall_frac = job->boundary.fp64_props[2];
    wall_col = (int)((job->N - 1) * wall_frac);
    return 1;
}

void bc_time_varying(job_t *job)
{
    memset(job->u_dirichlet_mask, 0, sizeof(int) * NODAL_DOF * job->num_nodes);
    for (size_t n = 0; n < job->N; n++) {
        if (job->nodes[n].x <= hole_radius && job->t > open_time) {
            continue;     /* trapdoor open */
        }
        for (int j = 0; j < 2; j++) {
            job->u_dirichlet[NODAL_DOF * n + j] = 0;
            job->u_dirichlet_mask[NODAL_DOF * n + j] = 1;
        }
    }
    for (size_t n = 0; n < job->N; n++) {
        size_t l = n * job->N, r = wall_col + n * job->N;
        job->u_dirichlet[NODAL_DOF * l + XDOF_IDX] = 0;
        job->u_dirichlet_mask[NODAL_DOF * l + XDOF_IDX] = 1;
        job->u_dirichlet[NODAL_DOF * r + XDOF_IDX] = 0;
        job->u_dirichlet_mask[NODAL_DOF * r + XDOF_IDX] = 1;
    }
    for (size_t i = 0; i < NODAL_DOF * job->num_nodes; i++) {
        job->node_number_override[i] = (int)i;
    }
}

static void zero_fixed(job_t *job, size_t ox, size_t oy)
{
    for (size_t i = 0; i < job->num_nodes; i++) {
        for (size_t j = 0; j < 2; j++) {
            size_t m = (size_t)job->node_number_override[NODAL_DOF * i + j];
            if (job->u_dirichlet_mask[m] != 0 && job->u_dirichlet[m] == 0) {
                *(double *)((char *)&job->nodes[i] + (j == 0 ? ox : oy)) = 0;
            }
        }
    }
}

void bc_momentum(job_t *job)
{
    zero_fixed(job, offsetof(node_t, mx_t), offsetof(node_t, my_t));
}

void bc_force(job_t *job)
{
    zero_fixed(job, offsetof(node_t, fx), offsetof(node_t, fy));
}

#ifndef MPM_NO_GPU_DESCRIPTOR
int bc_gpu_descriptor(const job_t *job, mpm_bc_gpu_desc *d)
{
    (void)job;
    d->abi = MPM_PLUGIN_GPU_ABI;
    d->kind = MPM_GPU_BC_SQUARE_SILO;
    d->fp64[0] = hole_radius;
    d->fp64[1] = open_time;
    d->fp64[2] = wall_frac;
    d->fp64[3] = 0;
    memset(d->i32, 0, sizeof d->i32);
    return 1;
}
#endif
