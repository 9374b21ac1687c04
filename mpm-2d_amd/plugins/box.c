/*
 * box boundary-condition plugin (plugin ABI of bcsrc/box.c): rectangular
 * walls whose nodes have their X and/or Y momentum and force zeroed.
 *
 * properties         = {x0, y0, width, height}            (box.c:42-96)
 * integer-properties = {left, bottom, right, top}; bit 1 fixes X, bit 2 Y
 *
 * bc_gpu_descriptor() lets the GPU step evaluate the same masks on the
 * device; the host functions keep the reference behaviour.
 */
#include <stdio.h>
#include <string.h>

#include "mpm_abi.h"
#include "mpm_plugin_gpu.h"

static size_t cols[2], rows[2];   /* left/right column, bottom/top row */
static int fric[4];               /* left, bottom, right, top */
static double fp[4];

void bc_init(job_t *job)
{
    (void)job;
}

int bc_validate(job_t *job)
{
    if (job->boundary.num_fp64_props != 4) {
        fprintf(stderr, "%s:%s: Wrong number of floating point boundary condition "
                "properties. Expected 4, got %zu.\n", __FILE__, __func__,
                job->boundary.num_fp64_props);
        return 0;
    }
    if (job->boundary.num_int_props != 4) {
        fprintf(stderr, "%s:%s: Wrong number of integer boundary condition properties. "
                "Expected 4, got %zu.\n", __FILE__, __func__, job->boundary.num_int_props);
        return 0;
    }
    memcpy(fp, job->boundary.fp64_props, sizeof fp);
    for (int k = 0; k < 4; k++) {
        fric[k] = job->boundary.int_props[k];
    }
    const double x0 = fp[0], y0 = fp[1], w = fp[2], h = fp[3];
    if (x0 < 0 || y0 < 0 || x0 > 1.0 || y0 > 1.0 || w < 0 || h < 0
        || (x0 + w) > 1.0 || (y0 + h) > 1.0) {
        fprintf(stderr, "%s:%s: Box must lie in [(0,0),(1,1)] with positive size.\n",
                __FILE__, __func__);
        return 0;
    }
    cols[0] = x0 * (job->N - 1);
    cols[1] = (x0 + w) * (job->N - 1);
    rows[0] = y0 * (job->N - 1);
    rows[1] = (y0 + h) * (job->N - 1);
    return 1;
}

static void fix(job_t *job, size_t node, int bits)
{
    for (int j = 0; j < 2; j++) {
        if (bits & (1 << j)) {
            job->u_dirichlet[NODAL_DOF * node + j] = 0;
            job->u_dirichlet_mask[NODAL_DOF * node + j] = 1;
        }
    }
}

void bc_time_varying(job_t *job)
{
    memset(job->u_dirichlet_mask, 0, sizeof(int) * NODAL_DOF * job->num_nodes);
    for (size_t n = cols[0]; n <= cols[1]; n++) {
        fix(job, n + rows[0] * job->N, fric[1]);
        fix(job, n + rows[1] * job->N, fric[3]);
    }
    for (size_t n = rows[0]; n <= rows[1]; n++) {
        fix(job, n * job->N + cols[0], fric[0]);
        fix(job, n * job->N + cols[1], fric[2]);
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
    d->kind = MPM_GPU_BC_BOX;
    memcpy(d->fp64, fp, sizeof fp);
    memcpy(d->i32, fric, sizeof fric);
    return 1;
}
#endif
