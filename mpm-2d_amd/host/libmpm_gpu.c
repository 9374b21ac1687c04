/*
 * libmpm_gpu.c -- drop-in replacement of the reference's libmpm for the
 * explicit USL step, running the step on one MI355X through libmpm_hip.so.
 *
 * The reference driver (src/main.c) calls mpm_init, find_filled_elements,
 * initial_loads, the material's material_init and the BC's bc_validate, then
 * on T threads loops { explicit_mpm_step_usl_threaded; barrier; serial:
 * frame output, time_varying_loads } (src/main.c:742-982).  This library keeps
 * that contract:
 *
 *   - mpm_init builds the host job exactly as libmpm/process_usl.c:68-280.
 *   - On the first step the serial thread uploads the particles, picks the
 *     device material and boundary conditions from the plugins' optional
 *     descriptor symbols (include/mpm_plugin_gpu.h) and checks that the BC
 *     plugin's bc_momentum/bc_force zero exactly the DOFs its masks mark.
 *   - Each step, job->t += job->dt on the host (as process_usl.c:522), then
 *     one device step.  Without a BC descriptor the plugin's bc_time_varying
 *     runs on the host and the masks are uploaded when they change.  Without
 *     a material descriptor the step stops after the volume update, all
 *     threads run the plugin's calculate_stress_threaded on their particle
 *     slices, and the new stresses go back to the device.
 *   - Particle state is copied back into the AoS arrays only when the driver
 *     will read it (frame times and the end of the run).
 *
 * One job per process, as the reference's plugins (static state) assume.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "mpm_abi.h"
#include "mpm_dropin.h"
#include "mpm_hip.h"
#include "mpm_plugin_gpu.h"

#define TOL 1e-10
#define EMOD 1e6          /* libmpm/material.h:12, feeds the automatic dt */
#define G_MAG 9.8         /* libmpm/loading.c:17 */
#define RAMP_TIME 0.1     /* libmpm/loading.c:18 */

typedef struct gpu_job {
    job_t *job;
    mpm_hip_ctx *ctx;
    int ready;
    int host_material;
    int s_law;              /* device law on s: 1 isolin, 2 dp_ri_mt (szz in state[1]) */
    int scheme;             /* MPM_SCHEME_USL or MPM_SCHEME_MUSL, fixed by the first step */
    int host_bc;
    int aux_on;
    uint8_t *mask;        /* effective Dirichlet mask last uploaded */
    int mask_valid;
    double *dbuf;
    int *ibuf;
    double bx, by;        /* body force the next step uses */
    int async_frames;     /* the driver takes frames by snapshot (mpm_gpu_async_frames) */
} gpu_job_t;

static gpu_job_t G;

static void fatal(const char *what, int rc)
{
    fprintf(stderr, "libmpm_gpu: %s failed: %s\n", what, mpm_hip_status_string(rc));
    exit(EXIT_ERROR_THREADING);
}

#define CHECK(call) do { int rc_ = (call); if (rc_ != MPM_HIP_OK) fatal(#call, rc_); } while (0)

/* ------------------------------------------------------------------------ */
/* Host helpers of libmpm (plugins link against these).                      */
/* ------------------------------------------------------------------------ */
void global_to_local_coords(double *x_local, double *y_local, double x, double y,
                            double x_ref, double y_ref, double h)
{
    *x_local = (x - x_ref) / h;
    const double xs = copysign(1.0, *x_local);
    if ((xs * (*x_local)) > 1.0 && ((xs * (*x_local)) - 1.0) < TOL) {
        *x_local = xs;
    }
    *y_local = (y - y_ref) / h;
    const double ys = copysign(1.0, *y_local);
    if ((ys * (*y_local)) > 1.0 && ((ys * (*y_local)) - 1.0) < TOL) {
        *y_local = ys;
    }
}

void tent(double *h1, double *h2, double *h3, double *h4, double xl, double yl)
{
    *h1 = (1 - xl) * (1 - yl);
    *h2 = xl * (1 - yl);
    *h3 = xl * yl;
    *h4 = (1 - xl) * yl;
}

void grad_tent(double *b11, double *b12, double *b13, double *b14,
               double *b21, double *b22, double *b23, double *b24,
               double xl, double yl, double h)
{
    const double r = 1.0 / h;
    *b11 = -(1 - yl) * r;
    *b12 = (1 - yl) * r;
    *b13 = yl * r;
    *b14 = -yl * r;
    *b21 = -(1 - xl) * r;
    *b22 = -xl * r;
    *b23 = xl * r;
    *b24 = (1 - xl) * r;
}

void node_number_to_coords(double *x, double *y, int num, int N, double h)
{
    const int i = num % N;
    const int j = num / N;
    *x = i * h;
    *y = j * h;
}

void zero_array(double *array, size_t numel)
{
    memset(array, 0, numel * sizeof(double));
}

void tensor_trace3(double *trA, const double *A)
{
    *trA = 0;
    for (size_t i = 0; i < 3; i++) {
        *trA += A[i * 3 + i];
    }
}

void tensor_copy3(double *C, const double *A)
{
    if (C != A) {
        memcpy(C, A, 9 * sizeof(double));
    }
}

void tensor_add3(double *C, const double *A, const double *B)
{
    tensor_copy3(C, A);
    for (size_t k = 0; k < 9; k++) {
        C[k] += B[k];
    }
}

void tensor_decompose3(double *A_0, double *c, const double *A)
{
    double trA;
    tensor_trace3(&trA, A);
    *c = trA / 3.0;
    tensor_copy3(A_0, A);
    for (size_t i = 0; i < 3; i++) {
        A_0[i * 3 + i] = A_0[i * 3 + i] - (*c);
    }
}

void tensor_multiply3(double *C, const double *A, const double *B)
{
    for (size_t i = 0; i < 3; i++) {
        for (size_t j = 0; j < 3; j++) {
            C[i * 3 + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
        }
    }
}

void tensor_scale3(double *A, double c)
{
    for (size_t k = 0; k < 9; k++) {
        A[k] *= c;
    }
}

void tensor_sym3(double *C, const double *A)
{
    for (size_t i = 0; i < 3; i++) {
        for (size_t j = 0; j < 3; j++) {
            C[i * 3 + j] = 0.5 * (A[i * 3 + j] + A[j * 3 + i]);
        }
    }
}

void tensor_skw3(double *C, const double *A)
{
    for (size_t i = 0; i < 3; i++) {
        for (size_t j = 0; j < 3; j++) {
            C[i * 3 + j] = 0.5 * (A[i * 3 + j] - A[j * 3 + i]);
        }
    }
}

void tensor_contraction3(double *c, const double *A, const double *B)
{
    *c = 0;
    for (size_t k = 0; k < 9; k++) {
        (*c) += A[k] * B[k];
    }
}

/* ------------------------------------------------------------------------ */
/* mpm_init / mpm_cleanup / element lists / loads                            */
/* ------------------------------------------------------------------------ */
#define WHICH_ELEMENT(xp, yp, N, h) \
    ((int)(((xp) < 1.0 && (xp) >= 0.0 && (yp) < 1.0 && (yp) >= 0.0) \
        ? ((floor((xp) / (h)) + floor((yp) / (h)) * ((N) - 1))) : (-1)))

job_t *mpm_init(int N, double h, particle_t *particles, size_t num_particles, double t)
{
    job_t *job = calloc(1, sizeof(job_t));
    job->t = 0;
    job->t_stop = t;
    job->frame = 0;
    job->N = N;
    job->h = h;
    job->num_nodes = (size_t)N * N;
    job->num_particles = num_particles;
    job->num_elements = (size_t)(N - 1) * (N - 1);

    job->particles = calloc(num_particles ? num_particles : 1, sizeof(particle_t));
    memcpy(job->particles, particles, num_particles * sizeof(particle_t));
    for (size_t i = 0; i < num_particles; i++) {
        particle_t *p = &job->particles[i];
        p->real_sxx = p->sxx;
        p->real_sxy = p->sxy;
        p->real_syy = p->syy;
        p->exx_t = p->exy_t = p->eyy_t = p->wxy_t = 0;
        p->Fxx = 1;
        p->Fxy = 0;
        p->Fyx = 0;
        p->Fyy = 1;
        p->ux = p->uy = 0;
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

    const size_t Ne = job->N - 1;
    job->elements = calloc(job->num_elements ? job->num_elements : 1, sizeof(element_t));
    for (size_t i = 0; i < job->num_elements; i++) {
        int n = (int)((i / Ne) * job->N + (i % Ne));
        element_t *e = &job->elements[i];
        e->nodes[0] = n;
        e->nodes[1] = n + 1;
        e->nodes[2] = n + N + 1;
        e->nodes[3] = n + N;
    }
    job->num_colors = 4;
    job->color_indices = calloc(job->num_colors, sizeof(size_t));
    for (size_t i = 0; i < job->num_elements; i++) {
        size_t r = i / Ne, c = i % Ne;
        element_t *e = &job->elements[i];
        e->color = (int)(2 * (r % 2) + (c % 2));
        /* neighbours, counter-clockwise from the right (process_usl.c:167-198);
         * the index arithmetic is the reference's, including its row/column
         * order in ijton */
        for (int k = 0; k < 8; k++) {
            e->neighbors[k] = -1;
        }
        if (r != Ne) {
            e->neighbors[2] = (int)(c * Ne + (r + 1));
            if (c != Ne) e->neighbors[1] = (int)((c + 1) * Ne + (r + 1));
            if (c != 0) e->neighbors[3] = (int)((c - 1) * Ne + (r + 1));
        }
        if (r != 0) {
            e->neighbors[6] = (int)(c * Ne + (r - 1));
            if (c != Ne) e->neighbors[7] = (int)((c + 1) * Ne + (r - 1));
            if (c != 0) e->neighbors[5] = (int)((c - 1) * Ne + (r - 1));
        }
        if (c != Ne) e->neighbors[0] = (int)((c + 1) * Ne + r);
        if (c != 0) e->neighbors[4] = (int)((c - 1) * Ne + r);
    }

    size_t np = num_particles ? num_particles : 1;
    job->in_element = malloc(np * sizeof(int));
    job->active = malloc(np * sizeof(int));
    double **sf[] = {&job->h1, &job->h2, &job->h3, &job->h4, &job->b11, &job->b12,
                     &job->b13, &job->b14, &job->b21, &job->b22, &job->b23, &job->b24};
    for (size_t k = 0; k < sizeof sf / sizeof sf[0]; k++) {
        *sf[k] = calloc(np, sizeof(double));
    }
    job->vec_len = (int)(NODAL_DOF * job->num_nodes);
    job->u_dirichlet = calloc(NODAL_DOF * job->num_nodes, sizeof(double));
    job->u_dirichlet_mask = calloc(NODAL_DOF * job->num_nodes, sizeof(int));
    job->node_number_override = calloc(job->vec_len, sizeof(int));

    for (size_t i = 0; i < num_particles; i++) {
        job->in_element[i] = WHICH_ELEMENT(job->particles[i].x, job->particles[i].y, job->N, job->h);
        job->active[i] = !(job->in_element[i] < 0
                           || (size_t)job->in_element[i] > job->num_elements);
    }
    for (size_t i = 0; i < num_particles; i++) {
        particle_t *p = &job->particles[i];
        p->v0 = p->v;
        p->r1_initial[0] = 0.5 * sqrt(p->v0);
        p->r1_initial[1] = 0;
        p->r2_initial[0] = 0;
        p->r2_initial[1] = 0.5 * sqrt(p->v0);
    }
    if (num_particles) {
        job->dt = 0.4 * job->h * sqrt(job->particles[0].m / (job->particles[0].v * EMOD));
    }
    job->use_cpdi = 0;
    job->material.num_fp64_props = 0;
    job->material.num_int_props = 0;
    job->step_number = 0;
    job->step_start_time = job->t;
    return job;
}

void find_filled_elements(job_t *job)
{
    for (size_t i = 0; i < job->num_colors; i++) {
        job->color_indices[i] = 0;
    }
    for (size_t i = 0; i < job->num_particles; i++) {
        if (!job->active[i]) continue;
        element_t *e = &job->elements[job->in_element[i]];
        e->filled = 1;
        e->n++;
        e->m += job->particles[i].m;
    }
    for (size_t i = 0; i < job->num_elements; i++) {
        if (job->elements[i].filled) {
            job->elements[i].color_idx = (int)job->color_indices[job->elements[i].color];
            job->color_indices[job->elements[i].color]++;
        }
    }
    if (!job->particle_by_element_color_lengths || !job->particle_by_element_color_lists) {
        return;
    }
    for (size_t i = 0; i < job->num_threads * job->num_colors; i++) {
        job->particle_by_element_color_lengths[i] = 0;
    }
    for (size_t i = 0; i < job->num_particles; i++) {
        if (!job->active[i]) continue;
        const element_t *e = &job->elements[job->in_element[i]];
        size_t tc = (e->color_idx % job->num_threads) * job->num_colors + e->color;
        job->particle_by_element_color_lists[tc][job->particle_by_element_color_lengths[tc]++] = i;
    }
}

void initial_loads(job_t *job)
{
    for (size_t i = 0; i < job->num_particles; i++) {
        job->particles[i].bx = 0;
        job->particles[i].by = 0;
        job->particles[i].material = 1;   /* M_DRUCKER_PRAGER */
    }
    G.bx = 0;
    G.by = 0;
}

static double loading_gravity(const job_t *job)
{
    if ((RAMP_TIME + job->step_start_time) > job->t) {
        return -G_MAG * ((job->t - job->step_start_time) / RAMP_TIME);
    }
    return -G_MAG;
}

void time_varying_loads(job_t *job)
{
    G.bx = 0;
    G.by = loading_gravity(job);
    if (!G.ready) {
        /* before the device owns the state, keep the reference's AoS writes */
        for (size_t i = 0; i < job->num_particles; i++) {
            job->particles[i].bx = 0;
            job->particles[i].by = G.by;
        }
    }
    /* afterwards the device applies the same rule after every step itself;
     * the AoS copy of bx/by is refreshed whenever particles are copied back */
}

void mpm_cleanup(job_t *job)
{
    if (G.ctx && G.job == job) {
        mpm_hip_destroy(G.ctx);
        free(G.mask);
        free(G.dbuf);
        free(G.ibuf);
        memset(&G, 0, sizeof G);
    }
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
}

/* ------------------------------------------------------------------------ */
/* Device set-up                                                            */
/* ------------------------------------------------------------------------ */
static void *symbol_beside(void *fn, const char *name)
{
    Dl_info info;
    if (!fn || !dladdr(fn, &info) || !info.dli_fname) {
        return NULL;
    }
    void *h = dlopen(info.dli_fname, RTLD_LAZY | RTLD_NOLOAD);
    if (!h) {
        return NULL;
    }
    void *sym = dlsym(h, name);
    dlclose(h);
    return sym;
}

/* Effective per-node mask from the Dirichlet arrays, as bc_momentum and
 * bc_force apply them (bcsrc/box.c:183-224): DOF j of node i is zeroed when
 * u_dirichlet_mask[m] != 0 and u_dirichlet[m] == 0, m = override[3i+j]. */
static void effective_mask(const job_t *job, uint8_t *mask)
{
    for (size_t i = 0; i < job->num_nodes; i++) {
        uint8_t f = 0;
        for (int j = 0; j < 2; j++) {
            size_t m = (size_t)job->node_number_override[NODAL_DOF * i + j];
            if (job->u_dirichlet_mask[m] != 0 && job->u_dirichlet[m] == 0) {
                f |= (uint8_t)(1u << j);
            }
        }
        mask[i] = f;
    }
}

/* The GPU applies bc_momentum/bc_force as mask zeroing; check on sentinel
 * node values that this plugin does exactly that. */
static void check_bc_plugin(job_t *job, const uint8_t *mask)
{
    node_t *saved = malloc(job->num_nodes * sizeof(node_t));
    memcpy(saved, job->nodes, job->num_nodes * sizeof(node_t));
    for (size_t i = 0; i < job->num_nodes; i++) {
        job->nodes[i].mx_t = job->nodes[i].my_t = 1.0;
        job->nodes[i].fx = job->nodes[i].fy = 1.0;
    }
    job->boundary.bc_momentum(job);
    job->boundary.bc_force(job);
    size_t bad = 0;
    for (size_t i = 0; i < job->num_nodes; i++) {
        const node_t *n = &job->nodes[i];
        int zx = (mask[i] & 1) != 0, zy = (mask[i] & 2) != 0;
        bad += (n->mx_t == 0.0) != zx || (n->fx == 0.0) != zx;
        bad += (n->my_t == 0.0) != zy || (n->fy == 0.0) != zy;
    }
    memcpy(job->nodes, saved, job->num_nodes * sizeof(node_t));
    free(saved);
    if (bad) {
        fprintf(stderr, "libmpm_gpu: the boundary plugin's bc_momentum/bc_force do not "
                "match its Dirichlet masks at %zu DOFs; it cannot run on the GPU\n", bad);
        exit(EXIT_ERROR_BC_PROPS);
    }
}

static void upload_mask_if_changed(job_t *job)
{
    uint8_t *m = malloc(job->num_nodes);
    effective_mask(job, m);
    if (!G.mask_valid || memcmp(m, G.mask, job->num_nodes) != 0) {
        CHECK(mpm_hip_set_bc_mask(G.ctx, m));
        memcpy(G.mask, m, job->num_nodes);
        G.mask_valid = 1;
    }
    free(m);
}

static void gpu_setup(job_t *job)
{
    mpm_hip_config cfg;
    mpm_hip_default_config(&cfg);
    cfg.N = (int)job->N;
    cfg.h = job->h;
    cfg.dt = job->dt;
    cfg.t0 = job->t;
    cfg.loads_t0 = job->step_start_time;
    cfg.g_mag = G_MAG;
    cfg.ramp_time = RAMP_TIME;
    cfg.scheme = G.scheme;
    const char *dev = getenv("MPM_HIP_DEVICE");
    cfg.device = dev ? atoi(dev) : 0;
    const char *cap = getenv("MPM_HIP_BIN_CAPACITY");
    cfg.bin_capacity = cap ? atoi(cap) : 0;

    mpm_material_gpu_descriptor_fn mdesc = (mpm_material_gpu_descriptor_fn)
        symbol_beside((void *)job->material.calculate_stress_threaded, "material_gpu_descriptor");
    mpm_material_gpu_desc md;
    memset(&md, 0, sizeof md);
    int have = mdesc && mdesc(job, &md) && md.abi == MPM_PLUGIN_GPU_ABI;
    if (have && md.kind == MPM_GPU_MATERIAL_ISOLIN) {
        cfg.material_kind = MPM_MAT_ISOLIN;
        cfg.material.E = md.params[0];
        cfg.material.nu = md.params[1];
        G.host_material = 0;
        G.s_law = 1;
    } else if (have && md.kind == MPM_GPU_MATERIAL_DP_RI_MT) {
        cfg.material_kind = MPM_MAT_DP_RI_MT;
        cfg.material.E = md.params[0];
        cfg.material.nu = md.params[1];
        cfg.material.mu_s = md.params[2];
        cfg.material.rho_crit = 1485.0f;   /* dp_ri_mt.c:165 */
        G.host_material = 0;
        G.s_law = 2;
    } else if (have && md.kind == MPM_GPU_MATERIAL_G_LOCAL_MU2) {
        cfg.material.E = md.params[0];
        cfg.material.nu = md.params[1];
        cfg.material.mu_s = md.params[2];
        cfg.material.mu_2 = md.params[3];
        cfg.material.I_0 = md.params[4];
        cfg.material.grains_d = md.params[5];
        cfg.material.grains_rho = md.params[6];
        cfg.material.rho_crit = md.params[7];
        G.host_material = 0;
    } else {
        if (!job->material.calculate_stress_threaded) {
            fprintf(stderr, "libmpm_gpu: no calculate_stress_threaded\n");
            exit(EXIT_ERROR_MATERIAL_FILE);
        }
        G.host_material = 1;
        cfg.flags |= MPM_HIP_HOST_MATERIAL;
    }

    mpm_bc_gpu_descriptor_fn bdesc = (mpm_bc_gpu_descriptor_fn)
        symbol_beside((void *)job->boundary.bc_time_varying, "bc_gpu_descriptor");
    mpm_bc_gpu_desc bd;
    memset(&bd, 0, sizeof bd);
    if (bdesc && bdesc(job, &bd) && bd.abi == MPM_PLUGIN_GPU_ABI
        && (bd.kind == MPM_GPU_BC_BOX || bd.kind == MPM_GPU_BC_SQUARE_SILO)) {
        cfg.bc_kind = bd.kind == MPM_GPU_BC_BOX ? MPM_BC_BOX : MPM_BC_SQUARE_SILO;
        memcpy(cfg.bc_fp64, bd.fp64, sizeof cfg.bc_fp64);
        memcpy(cfg.bc_int, bd.i32, sizeof cfg.bc_int);
        G.host_bc = 0;
    } else {
        cfg.bc_kind = MPM_BC_MASK;
        G.host_bc = 1;
    }

    const size_t np = job->num_particles;
    CHECK(mpm_hip_create(&cfg, np, &G.ctx));
    G.job = job;
    G.dbuf = malloc((np ? np : 1) * sizeof(double) * 9);
    G.ibuf = malloc((np ? np : 1) * sizeof(int));
    G.mask = calloc(job->num_nodes, 1);

    /* Upload the host state: mpm_init's particles after material_init. */
    double *m = G.dbuf, *v = m + np, *x = v + np, *y = x + np, *vx = y + np, *vy = vx + np;
    double *sxx = vy + np, *sxy = sxx + np, *syy = sxy + np;
    for (size_t i = 0; i < np; i++) {
        const particle_t *p = &job->particles[i];
        m[i] = p->m;
        v[i] = p->v;
        x[i] = p->x;
        y[i] = p->y;
        vx[i] = p->x_t;
        vy[i] = p->y_t;
        /* the device law's stress: isolin and host laws act on s, g_local_mu2
         * on T (which its material_init set from s) */
        const int on_s = G.host_material || G.s_law;
        sxx[i] = on_s ? p->sxx : p->T[XX];
        sxy[i] = on_s ? p->sxy : p->T[XY];
        syy[i] = on_s ? p->syy : p->T[YY];
    }
    CHECK(mpm_hip_set_particles(G.ctx, m, v, x, y, vx, vy, sxx, sxy, syy));
    const struct { int f; size_t off; } extra[] = {
        /* the law's out-of-plane stress: T[ZZ] for g_local_mu2, szz =
         * state[1] for dp_ri_mt (a restarted job carries both) */
        {MPM_F_TZZ, G.s_law == 2 ? offsetof(particle_t, state) + 1 * sizeof(double)
                                 : offsetof(particle_t, T) + ZZ * sizeof(double)},
        {MPM_F_GAMMAP, offsetof(particle_t, state) + 9 * sizeof(double)},
        {MPM_F_GAMMADOTP, offsetof(particle_t, state) + 10 * sizeof(double)},
        {MPM_F_UX, offsetof(particle_t, ux)},
        {MPM_F_UY, offsetof(particle_t, uy)},
    };
    for (size_t k = 0; k < sizeof extra / sizeof extra[0]; k++) {
        for (size_t i = 0; i < np; i++) {
            memcpy(&m[i], (const char *)&job->particles[i] + extra[k].off, sizeof(double));
        }
        CHECK(mpm_hip_set_field(G.ctx, extra[k].f, m));
    }
    CHECK(mpm_hip_set_body_force(G.ctx, np ? job->particles[0].bx : G.bx,
                                 np ? job->particles[0].by : G.by));
    if (G.host_bc) {
        /* the plugin's masks at the first step's time, and the self-check */
        double t_saved = job->t;
        job->t += job->dt;
        job->boundary.bc_time_varying(job);
        uint8_t *mask = malloc(job->num_nodes);
        effective_mask(job, mask);
        check_bc_plugin(job, mask);
        free(mask);
        job->t = t_saved;
    }
    G.ready = 1;
}

/* ------------------------------------------------------------------------ */
/* Copies between device and AoS                                             */
/* ------------------------------------------------------------------------ */
static void get_into(int field, size_t off)
{
    job_t *job = G.job;
    CHECK(mpm_hip_get_field(G.ctx, field, G.dbuf));
    for (size_t i = 0; i < job->num_particles; i++) {
        memcpy((char *)&job->particles[i] + off, &G.dbuf[i], sizeof(double));
    }
}

#define POFF(f) offsetof(particle_t, f)
#define TOFF(k) (offsetof(particle_t, T) + (k) * sizeof(double))
#define SOFF(k) (offsetof(particle_t, state) + (k) * sizeof(double))

static void download_kinematics(void)
{
    job_t *job = G.job;
    get_into(MPM_F_X, POFF(x));
    get_into(MPM_F_Y, POFF(y));
    get_into(MPM_F_VX, POFF(x_t));
    get_into(MPM_F_VY, POFF(y_t));
    get_into(MPM_F_V, POFF(v));
    get_into(MPM_F_UX, POFF(ux));
    get_into(MPM_F_UY, POFF(uy));
    CHECK(mpm_hip_get_active(G.ctx, G.ibuf));
    memcpy(job->active, G.ibuf, job->num_particles * sizeof(int));
    for (size_t i = 0; i < job->num_particles; i++) {
        job->particles[i].bx = G.bx;
        job->particles[i].by = G.by;
    }
    if (G.aux_on) {
        get_into(MPM_F_EXX, POFF(exx_t));
        get_into(MPM_F_EXY, POFF(exy_t));
        get_into(MPM_F_EYY, POFF(eyy_t));
        get_into(MPM_F_WXY, POFF(wxy_t));
        get_into(MPM_F_AX, POFF(x_tt));
        get_into(MPM_F_AY, POFF(y_tt));
        get_into(MPM_F_XL, POFF(xl));
        get_into(MPM_F_YL, POFF(yl));
    }
}

static void download_stress(void)
{
    job_t *job = G.job;
    if (G.s_law) {
        /* isolin and dp_ri_mt keep their stress in s (isolin.c:107-109,
         * dp_ri_mt.c:167-194); dp_ri_mt also szz and the plastic strain in
         * state[1], state[9], state[10] */
        get_into(MPM_F_TXX, POFF(sxx));
        get_into(MPM_F_TXY, POFF(sxy));
        get_into(MPM_F_TYY, POFF(syy));
        if (G.s_law == 2) {
            get_into(MPM_F_TZZ, SOFF(1));
            get_into(MPM_F_GAMMAP, SOFF(9));
            get_into(MPM_F_GAMMADOTP, SOFF(10));
        }
        return;
    }
    get_into(MPM_F_TXX, TOFF(XX));
    get_into(MPM_F_TXY, TOFF(XY));
    get_into(MPM_F_TYY, TOFF(YY));
    get_into(MPM_F_TZZ, TOFF(ZZ));
    get_into(MPM_F_GAMMAP, SOFF(9));
    get_into(MPM_F_GAMMADOTP, SOFF(10));
    for (size_t i = 0; i < job->num_particles; i++) {
        particle_t *p = &job->particles[i];
        p->T[YX] = p->T[XY];
        p->sxx = p->T[XX];
        p->sxy = p->T[XY];
        p->syy = p->T[YY];
    }
}

static void put_from(int field, size_t off)
{
    job_t *job = G.job;
    for (size_t i = 0; i < job->num_particles; i++) {
        memcpy(&G.dbuf[i], (const char *)&job->particles[i] + off, sizeof(double));
    }
    CHECK(mpm_hip_set_field(G.ctx, field, G.dbuf));
}

/* ------------------------------------------------------------------------ */
/* The step                                                                 */
/* ------------------------------------------------------------------------ */
static int driver_reads_after_step(const job_t *job)
{
    return job->t >= (job->frame / job->output.sample_rate_hz) || job->t >= job->t_stop;
}

static void serial_step(job_t *job)
{
    if (!G.ready) {
        gpu_setup(job);
    }
    /* strain rates, accelerations and local coordinates are only needed for
     * the final state dump (src/writer.c:266-380) and by host materials */
    int last = job->t + job->dt >= job->t_stop;
    int want_aux = last || G.host_material;
    if (want_aux != G.aux_on) {
        CHECK(mpm_hip_set_store_aux(G.ctx, want_aux));
        G.aux_on = want_aux;
    }
    job->t += job->dt;
    if (G.host_bc) {
        job->boundary.bc_time_varying(job);
        upload_mask_if_changed(job);
    }
    CHECK(mpm_hip_step(G.ctx, 1));
    /* the loads the next step uses, as time_varying_loads will compute them */
    G.bx = 0;
    G.by = loading_gravity(job);
    /* with snapshot frames only the end of the run is copied back here */
    int reads = driver_reads_after_step(job);
    if (G.async_frames && !G.host_material && job->t < job->t_stop) {
        reads = 0;
    }
    if (G.host_material || reads) {
        CHECK(mpm_hip_sync(G.ctx));
        download_kinematics();
        if (!G.host_material) {
            download_stress();
        }
    }
}

static void step_threaded(void *_task, int scheme)
{
    threadtask_t *task = (threadtask_t *)_task;
    job_t *job = task->job;
    int rc = pthread_barrier_wait(job->serialize_barrier);
    if (rc == PTHREAD_BARRIER_SERIAL_THREAD && !G.ready) {
        G.scheme = scheme;   /* the first step fixes the scheme of the run */
    }
    if (rc == PTHREAD_BARRIER_SERIAL_THREAD) {
        serial_step(job);
    }
    pthread_barrier_wait(job->serialize_barrier);
    if (G.host_material) {
        /* the plugin updates T, state and sxx/sxy/syy of its slice
         * (process_usl.c:589) */
        job->material.calculate_stress_threaded(task);
        rc = pthread_barrier_wait(job->serialize_barrier);
        if (rc == PTHREAD_BARRIER_SERIAL_THREAD) {
            put_from(MPM_F_TXX, POFF(sxx));
            put_from(MPM_F_TXY, POFF(sxy));
            put_from(MPM_F_TYY, POFF(syy));
        }
        pthread_barrier_wait(job->serialize_barrier);
    }
}

void explicit_mpm_step_usl_threaded(void *task)
{
    step_threaded(task, MPM_SCHEME_USL);
}

void explicit_mpm_step_musl_threaded(void *task)
{
    step_threaded(task, MPM_SCHEME_MUSL);
}

/* ------------------------------------------------------------------------ */
/* Frames by snapshot (an extension for this repository's driver)            */
/* ------------------------------------------------------------------------ */
void mpm_gpu_async_frames(job_t *job, int on)
{
    (void)job;
    G.async_frames = on != 0;
}

int mpm_gpu_frame_alloc(job_t *job, mpm_gpu_frame_t *f)
{
    memset(f, 0, sizeof *f);
    const size_t n = job->num_particles ? job->num_particles : 1;
    double **cols[MPM_GPU_FRAME_FIELDS] = {&f->v, &f->x, &f->y, &f->x_t, &f->y_t, &f->sxx,
                                           &f->sxy, &f->syy, &f->ux, &f->uy, &f->state9,
                                           &f->state10};
    for (int k = 0; k < MPM_GPU_FRAME_FIELDS; k++) {
        if (mpm_hip_host_alloc(n * sizeof(double), (void **)cols[k]) != MPM_HIP_OK) {
            return -1;
        }
    }
    if (mpm_hip_host_alloc(n * sizeof(int), (void **)&f->active) != MPM_HIP_OK) {
        return -1;
    }
    return 0;
}

void mpm_gpu_frame_free(mpm_gpu_frame_t *f)
{
    double *cols[MPM_GPU_FRAME_FIELDS] = {f->v, f->x, f->y, f->x_t, f->y_t, f->sxx, f->sxy,
                                          f->syy, f->ux, f->uy, f->state9, f->state10};
    for (int k = 0; k < MPM_GPU_FRAME_FIELDS; k++) {
        mpm_hip_host_free(cols[k]);
    }
    mpm_hip_host_free(f->active);
    memset(f, 0, sizeof *f);
}

int mpm_gpu_frame_begin(job_t *job, mpm_gpu_frame_t *f, size_t frame, double t)
{
    if (!G.async_frames || G.host_material || !G.ready || G.job != job) {
        return 0;   /* the job's AoS is current: write_frame now */
    }
    /* the fields write_frame prints (src/writer.c:169-182) that the device
     * advances; m, color and corners stay in the AoS */
    int fields[MPM_GPU_FRAME_FIELDS] = {MPM_F_V, MPM_F_X, MPM_F_Y, MPM_F_VX, MPM_F_VY,
                                        MPM_F_TXX, MPM_F_TXY, MPM_F_TYY, MPM_F_UX, MPM_F_UY,
                                        MPM_F_GAMMAP, MPM_F_GAMMADOTP};
    double *dst[MPM_GPU_FRAME_FIELDS] = {f->v, f->x, f->y, f->x_t, f->y_t, f->sxx, f->sxy,
                                         f->syy, f->ux, f->uy, f->state9, f->state10};
    /* isolin keeps no plastic state on the device (isolin.c:72-116): its
     * state[9], state[10] are the AoS values */
    const int nf = G.s_law == 1 ? MPM_GPU_FRAME_FIELDS - 2 : MPM_GPU_FRAME_FIELDS;
    f->state_from_job = G.s_law == 1;
    f->frame = frame;
    f->t = t;
    CHECK(mpm_hip_snapshot_begin(G.ctx, fields, nf, dst, f->active));
    return 1;
}

/* One chunk of a frame's particle lines, formatted into memory with the
 * same fprintf calls as write_particle (src/writer.c:169-182). */
typedef struct frame_chunk {
    const job_t *job;
    const mpm_gpu_frame_t *f;
    size_t begin, end;
    char *text;
    size_t len;
} frame_chunk_t;

static void write_lines(FILE *fd, const job_t *job, const mpm_gpu_frame_t *f, size_t begin,
                        size_t end)
{
    for (size_t i = begin; i < end; i++) {
        const particle_t *p = &job->particles[i];
        const double s9 = f->state_from_job ? p->state[9] : f->state9[i];
        const double s10 = f->state_from_job ? p->state[10] : f->state10[i];
        fprintf(fd, "%lg %lg %lg %lg %lg %lg ", p->m, f->v[i], f->x[i], f->y[i], f->x_t[i],
                f->y_t[i]);
        fprintf(fd, "%lg %lg %lg %lg %lg %lg %lg %lg %lg", f->sxx[i], f->sxy[i], f->syy[i],
                f->ux[i], f->uy[i], s9, p->color, s10, (double)f->active[i]);
        for (int k = 0; k < 4; k++) {
            fprintf(fd, " %lg %lg", p->corners[k][0], p->corners[k][1]);
        }
        fprintf(fd, "\n");
    }
}

static void *format_chunk(void *arg)
{
    frame_chunk_t *c = arg;
    FILE *fd = open_memstream(&c->text, &c->len);
    if (!fd) {
        c->text = NULL;
        return NULL;
    }
    write_lines(fd, c->job, c->f, c->begin, c->end);
    if (fclose(fd) != 0) {
        free(c->text);
        c->text = NULL;
    }
    return NULL;
}

int mpm_gpu_frame_write(FILE *fd, const job_t *job, const mpm_gpu_frame_t *f)
{
    CHECK(mpm_hip_snapshot_wait(G.ctx));
    fprintf(fd, "%zu %lg %zu\n", f->frame, f->t, job->num_particles);
    /* the particle lines in up to 16 chunks formatted concurrently, written
     * in order: the same bytes as one sequential pass */
    const size_t np = job->num_particles;
    long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
    size_t nt = ncpu > 16 ? 16 : (ncpu < 1 ? 1 : (size_t)ncpu);
    if (np < 4096 * nt) {
        nt = np / 4096 + 1;
    }
    frame_chunk_t chunk[16];
    pthread_t th[16];
    int started[16] = {0};
    for (size_t k = 0; k < nt; k++) {
        chunk[k] = (frame_chunk_t){job, f, np * k / nt, np * (k + 1) / nt, NULL, 0};
        started[k] = k > 0 && pthread_create(&th[k], NULL, format_chunk, &chunk[k]) == 0;
    }
    format_chunk(&chunk[0]);
    for (size_t k = 1; k < nt; k++) {
        if (started[k]) {
            pthread_join(th[k], NULL);
        } else {
            format_chunk(&chunk[k]);
        }
    }
    int rc = 0;
    for (size_t k = 0; k < nt; k++) {
        if (chunk[k].text) {
            if (fwrite(chunk[k].text, 1, chunk[k].len, fd) != chunk[k].len) {
                rc = -1;
            }
            free(chunk[k].text);
        } else {
            /* no memory for the chunk's text: format it straight out */
            write_lines(fd, job, f, chunk[k].begin, chunk[k].end);
        }
    }
    if (fflush(fd) != 0) {
        rc = -1;
    }
    return rc;
}
