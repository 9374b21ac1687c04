/*
 * mpm_usl_run -- runs the USL loop the way the reference driver does
 * (src/main.c:164-982), against whatever libmpm it is linked with, so the
 * drop-in library can be exercised end to end without libconfuse (absent
 * here; the reference driver needs it for its .cfg parsing).
 *
 *   mpm_usl_run -g GRID -p PARTICLES -u MATERIAL.so -b BC.so [-t THREADS]
 *               --dt DT --steps K [--mprops a,b,..] [--bprops a,b,..]
 *               [--biprops i,j,..] [--sample-rate HZ] [--frames FILE]
 *               [--state FILE]
 *
 * Inputs use the reference's grid and particle file formats
 * (src/reader.c:17-119).  Plugins are loaded with dlopen/dlsym by the same
 * symbol names (src/main.c:427-456, 500-543).  The job is split over T
 * threads as in src/main.c:747-804; the loop is mpm_run_until
 * (src/main.c:934-982) with t_stop = (K - 1/2) dt so it runs exactly K
 * steps.  --frames writes frames in the reference's text format
 * (src/writer.c:169-201); --state writes the final particle state with
 * full-precision hex floats, one particle per line:
 *   x y x_t y_t v sxx sxy syy T[ZZ] state[9] state[10] ux uy active
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mpm_abi.h"
#include "mpm_dropin.h"

#define MINZ(a, b) ((a) < (b) ? (a) : (b))

static FILE *frames_fd;

static int read_grid(const char *path, int *N, double *len)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    int r = fscanf(fp, "%d", N);
    r += fscanf(fp, "%lg", len);
    fclose(fp);
    return r == 2 ? 0 : -1;
}

static int read_particles(const char *path, particle_t **out, size_t *n)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    char s[16384];
    if (!fgets(s, sizeof s, fp) || sscanf(s, "%zu", n) != 1) {
        fclose(fp);
        return -1;
    }
    particle_t *p = calloc(*n ? *n : 1, sizeof(particle_t));
    for (size_t i = 0; i < *n; i++) {
        if (!fgets(s, sizeof s, fp)) {
            fclose(fp);
            free(p);
            return -1;
        }
        double v[9] = {0};
        int field = 0;
        for (char *tok = strtok(s, " ,\n"); tok && field < 9; tok = strtok(NULL, " ,\n")) {
            sscanf(tok, "%lg", &v[field++]);
        }
        if (field < 6) {
            fprintf(stderr, "error reading particle %zu\n", i);
        }
        p[i].m = v[0];
        p[i].v = v[1];
        p[i].x = v[2];
        p[i].y = v[3];
        p[i].x_t = v[4];
        p[i].y_t = v[5];
        p[i].sxx = field >= 9 ? v[6] : 0;
        p[i].sxy = field >= 9 ? v[7] : 0;
        p[i].syy = field >= 9 ? v[8] : 0;
    }
    fclose(fp);
    *out = p;
    return 0;
}

static size_t parse_list(const char *s, double *out, size_t cap)
{
    size_t n = 0;
    char *copy = strdup(s);
    for (char *tok = strtok(copy, ","); tok && n < cap; tok = strtok(NULL, ",")) {
        out[n++] = strtod(tok, NULL);
    }
    free(copy);
    return n;
}

static void write_frame(FILE *fd, size_t frame, double t, job_t *job)
{
    fprintf(fd, "%zu %lg %zu\n", frame, t, job->num_particles);
    for (size_t i = 0; i < job->num_particles; i++) {
        const particle_t *p = &job->particles[i];
        fprintf(fd, "%lg %lg %lg %lg %lg %lg ", p->m, p->v, p->x, p->y, p->x_t, p->y_t);
        fprintf(fd, "%lg %lg %lg %lg %lg %lg %lg %lg %lg", p->sxx, p->sxy, p->syy, p->ux,
                p->uy, p->state[9], p->color, p->state[10], (double)job->active[i]);
        for (int c = 0; c < 4; c++) {
            fprintf(fd, " %lg %lg", p->corners[c][0], p->corners[c][1]);
        }
        fprintf(fd, "\n");
    }
    fflush(fd);
}

static void *run_until(void *arg)
{
    threadtask_t *task = arg;
    job_t *job = task->job;
    while (job->t < job->t_stop) {
        explicit_mpm_step_usl_threaded(task);
        int rc = pthread_barrier_wait(job->serialize_barrier);
        if (rc == PTHREAD_BARRIER_SERIAL_THREAD) {
            job->stepcount++;
            if (frames_fd && job->t >= (job->frame / job->output.sample_rate_hz)) {
                write_frame(frames_fd, job->frame, job->t, job);
                job->frame++;
            }
            time_varying_loads(job);
        }
    }
    return NULL;
}

int main(int argc, char **argv)
{
    const char *grid = NULL, *pfile = NULL, *mat = NULL, *bc = NULL;
    const char *frames = NULL, *state = NULL;
    double dt = 0, rate = 60.0;
    int steps = 0;
    size_t T = 1;
    double mprops[16], bprops[16], biprops_d[16];
    size_t nm = 0, nb = 0, nbi = 0;
    for (int a = 1; a < argc; a++) {
        const char *k = argv[a];
        const char *v = a + 1 < argc ? argv[a + 1] : "";
        if (!strcmp(k, "-g")) grid = v, a++;
        else if (!strcmp(k, "-p")) pfile = v, a++;
        else if (!strcmp(k, "-u")) mat = v, a++;
        else if (!strcmp(k, "-b")) bc = v, a++;
        else if (!strcmp(k, "-t")) T = (size_t)atoi(v), a++;
        else if (!strcmp(k, "--dt")) dt = strtod(v, NULL), a++;
        else if (!strcmp(k, "--steps")) steps = atoi(v), a++;
        else if (!strcmp(k, "--mprops")) nm = parse_list(v, mprops, 16), a++;
        else if (!strcmp(k, "--bprops")) nb = parse_list(v, bprops, 16), a++;
        else if (!strcmp(k, "--biprops")) nbi = parse_list(v, biprops_d, 16), a++;
        else if (!strcmp(k, "--sample-rate")) rate = strtod(v, NULL), a++;
        else if (!strcmp(k, "--frames")) frames = v, a++;
        else if (!strcmp(k, "--state")) state = v, a++;
        else {
            fprintf(stderr, "unknown option %s\n", k);
            return EXIT_ERROR_CFG_PARSE;
        }
    }
    if (!grid || !pfile || !mat || !bc || dt <= 0 || steps <= 0 || T < 1) {
        fprintf(stderr, "usage: mpm_usl_run -g GRID -p PARTICLES -u MAT.so -b BC.so "
                "--dt DT --steps K [-t T] [--mprops ..] [--bprops ..] [--biprops ..]\n");
        return EXIT_ERROR_CFG_PARSE;
    }
    int N;
    double len;
    particle_t *pdata;
    size_t np;
    if (read_grid(grid, &N, &len) || read_particles(pfile, &pdata, &np)) {
        fprintf(stderr, "error reading inputs\n");
        return EXIT_ERROR_CFG_PARSE;
    }
    double h = len / (N - 1);
    job_t *job = mpm_init(N, h, pdata, np, (steps - 0.5) * dt);

    void *mh = dlopen(mat, RTLD_LAZY);
    void *bh = dlopen(bc, RTLD_LAZY);
    if (!mh || !bh) {
        fprintf(stderr, "Can't dlopen() plugin: %s\n", dlerror());
        return EXIT_ERROR_MATERIAL_FILE;
    }
    *(void **)&job->material.material_init = dlsym(mh, "material_init");
    *(void **)&job->material.calculate_stress = dlsym(mh, "calculate_stress");
    *(void **)&job->material.calculate_stress_threaded = dlsym(mh, "calculate_stress_threaded");
    *(void **)&job->boundary.bc_init = dlsym(bh, "bc_init");
    *(void **)&job->boundary.bc_validate = dlsym(bh, "bc_validate");
    *(void **)&job->boundary.bc_time_varying = dlsym(bh, "bc_time_varying");
    *(void **)&job->boundary.bc_momentum = dlsym(bh, "bc_momentum");
    *(void **)&job->boundary.bc_force = dlsym(bh, "bc_force");
    if (!job->material.material_init || !job->material.calculate_stress_threaded
        || !job->boundary.bc_validate || !job->boundary.bc_time_varying
        || !job->boundary.bc_momentum || !job->boundary.bc_force) {
        fprintf(stderr, "missing plugin symbol\n");
        return EXIT_ERROR_MATERIAL_FILE;
    }
    job->material.num_fp64_props = nm;
    job->material.fp64_props = mprops;
    job->boundary.num_fp64_props = nb;
    job->boundary.fp64_props = bprops;
    int biprops[16];
    for (size_t i = 0; i < nbi; i++) biprops[i] = (int)biprops_d[i];
    job->boundary.num_int_props = nbi;
    job->boundary.int_props = biprops;
    job->output.sample_rate_hz = rate;
    job->dt = dt;

    threadtask_t *tasks = calloc(T, sizeof(threadtask_t));
    size_t psplit = job->num_particles / T + (job->num_particles % T != 0);
    size_t nsplit = job->num_nodes / T + (job->num_nodes % T != 0);
    size_t esplit = job->num_elements / T + (job->num_elements % T != 0);
    for (size_t i = 0; i < T; i++) {
        tasks[i].id = i;
        tasks[i].num_threads = T;
        tasks[i].job = job;
        tasks[i].offset = MINZ(i * psplit, job->num_particles);
        tasks[i].blocksize = MINZ(psplit, job->num_particles - tasks[i].offset);
        tasks[i].n_offset = MINZ(i * nsplit, job->num_nodes);
        tasks[i].n_blocksize = MINZ(nsplit, job->num_nodes - tasks[i].n_offset);
        tasks[i].e_offset = MINZ(i * esplit, job->num_elements);
        tasks[i].e_blocksize = MINZ(esplit, job->num_elements - tasks[i].e_offset);
        tasks[i].gidx_max = job->num_particles;
    }
    job->num_threads = T;
    job->serialize_barrier = malloc(sizeof(pthread_barrier_t));
    job->step_barrier = malloc(sizeof(pthread_barrier_t));
    pthread_barrier_init(job->serialize_barrier, NULL, (unsigned)T);
    pthread_barrier_init(job->step_barrier, NULL, (unsigned)T);
    job->update_elementlists = malloc(sizeof(int) * T);
    for (size_t i = 0; i < T; i++) job->update_elementlists[i] = 1;
    job->particle_by_element_color_lengths = calloc(job->num_colors * T, sizeof(size_t));
    job->particle_by_element_color_lists = malloc(sizeof(size_t *) * job->num_colors * T);
    for (size_t i = 0; i < job->num_colors * T; i++) {
        job->particle_by_element_color_lists[i] = malloc(sizeof(size_t) * (np ? np : 1));
    }
    find_filled_elements(job);
    initial_loads(job);
    job->material.material_init(job);
    if (job->boundary.bc_validate(job) == 0) {
        fprintf(stderr, "Error with boundary condition properties.\n");
        return EXIT_ERROR_BC_PROPS;
    }
    job->frame = (size_t)floor(job->t * job->output.sample_rate_hz);
    if (frames) frames_fd = fopen(frames, "w");

    pthread_t *th = malloc(sizeof(pthread_t) * T);
    for (size_t i = 0; i + 1 < T; i++) {
        pthread_create(&th[i], NULL, run_until, &tasks[i]);
    }
    run_until(&tasks[T - 1]);
    for (size_t i = 0; i + 1 < T; i++) {
        pthread_join(th[i], NULL);
    }
    if (frames_fd) fclose(frames_fd);

    if (state) {
        FILE *fd = fopen(state, "w");
        fprintf(fd, "%a %d\n", job->t, job->stepcount);
        for (size_t i = 0; i < job->num_particles; i++) {
            const particle_t *p = &job->particles[i];
            fprintf(fd, "%a %a %a %a %a %a %a %a %a %a %a %a %a %d\n", p->x, p->y, p->x_t,
                    p->y_t, p->v, p->sxx, p->sxy, p->syy, p->T[ZZ], p->state[9],
                    p->state[10], p->ux, p->uy, job->active[i]);
        }
        fclose(fd);
    }
    printf("ran %d steps to t = %g on %zu threads\n", job->stepcount, job->t, T);
    mpm_cleanup(job);
    return 0;
}
