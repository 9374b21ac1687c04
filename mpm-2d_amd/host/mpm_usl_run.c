/*
 * mpm_usl_run -- the reference driver's set-up and loop (src/main.c:164-982)
 * over whatever libmpm it is linked with, so the drop-in library runs end
 * to end without libconfuse (absent here; the reference driver needs it for
 * its .cfg parsing).  Two ways to call it:
 *
 *   mpm_usl_run -c SIM.cfg [-o DIR] [-r STATE] [-p PFILE] [-g GFILE]
 *               [-u MAT.so] [-b BC.so] [-t THREADS] [--musl] [--state FILE]
 *               [--check-cfg] [t_max]
 *     the reference's command line (src/main.c:64, 145-160; t_max defaults
 *     to 1) and configuration file (host/cfg.h reads the libconfuse subset
 *     the reference's .cfg files use; options and defaults as
 *     src/main.c:169-251).  Writes info.txt, the frame file, the element,
 *     state and log files into the output directory as src/main.c:660-715
 *     opens them; frames in the reference's format (src/writer.c:169-201)
 *     and the final state in write_state's format (writer.c:266-380).
 *
 *   mpm_usl_run -g GRID -p PARTICLES -u MATERIAL.so -b BC.so [-t THREADS]
 *               --dt DT --steps K [--mprops a,b,..] [--bprops a,b,..]
 *               [--biprops i,j,..] [--sample-rate HZ] [--frames FILE]
 *               [--musl] [--state FILE]
 *     the same loop with t_stop = (K - 1/2) dt, so it runs exactly K steps.
 *
 * -r STATE restarts from a state file (write_state's format): grid,
 * particles, clock, loads, state variables and active flags come from it
 * (the plane-strain T[ZZ] of g_local_mu2 is not in the format and restarts
 * at zero, as material_init sets it); with the flags form, --steps counts
 * from the saved time.
 * --check-cfg prints the settings a -c run would use and exits.
 * --musl steps with explicit_mpm_step_musl_threaded (process_usl.c:284-462;
 * the reference's loop has that call commented out, src/main.c:948).
 * Frames are taken by snapshot: at a frame time the device copies the
 * printed fields into page-locked host memory while the next steps run, and
 * a writer thread formats and writes the frame (byte-identical to the
 * synchronous write_frame of src/writer.c:190-201, the particle lines
 * formatted in up to 16 chunks at once; with a host material
 * plugin the frame is written synchronously from the AoS).  --sync-frames
 * writes every frame synchronously from the AoS, as the reference does.
 * --state writes the final particle state with full-precision hex floats,
 * one particle per line:
 *   x y x_t y_t v sxx sxy syy T[ZZ] state[9] state[10] ux uy active
 * Inputs use the reference's grid and particle file formats
 * (src/reader.c:17-119).  Plugins are loaded with dlopen/dlsym by the same
 * symbol names (src/main.c:427-456, 500-543).  The job is split over T
 * threads as in src/main.c:747-804.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <sys/stat.h>

#include "cfg.h"
#include "mpm_abi.h"
#include "mpm_dropin.h"

#define MINZ(a, b) ((a) < (b) ? (a) : (b))

static FILE *frames_fd;
static int use_musl;
static int sync_frames;

/* The frame writer thread (snapshot frames): one frame in hand at a time. */
static pthread_mutex_t wlock = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t wcond = PTHREAD_COND_INITIALIZER;
static int writer_on, wfull, wquit;
static mpm_gpu_frame_t wbuf;
static job_t *wjob;
static pthread_t writer;

static void *frame_writer(void *arg)
{
    (void)arg;
    pthread_mutex_lock(&wlock);
    for (;;) {
        while (!wfull && !wquit) {
            pthread_cond_wait(&wcond, &wlock);
        }
        if (!wfull) {
            break;
        }
        pthread_mutex_unlock(&wlock);
        if (mpm_gpu_frame_write(frames_fd, wjob, &wbuf) != 0) {
            fprintf(stderr, "mpm_usl_run: writing frame %zu failed\n", wbuf.frame);
        }
        pthread_mutex_lock(&wlock);
        wfull = 0;
        pthread_cond_broadcast(&wcond);
    }
    pthread_mutex_unlock(&wlock);
    return NULL;
}

/* Waits until the writer has finished the frame in hand. */
static void writer_idle(void)
{
    pthread_mutex_lock(&wlock);
    while (wfull) {
        pthread_cond_wait(&wcond, &wlock);
    }
    pthread_mutex_unlock(&wlock);
}

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
        if (use_musl) {
            explicit_mpm_step_musl_threaded(task);
        } else {
            explicit_mpm_step_usl_threaded(task);
        }
        int rc = pthread_barrier_wait(job->serialize_barrier);
        if (rc == PTHREAD_BARRIER_SERIAL_THREAD) {
            job->stepcount++;
            if (frames_fd && job->t >= (job->frame / job->output.sample_rate_hz)) {
                int begun = 0;
                if (writer_on) {
                    writer_idle();   /* the last frame is out: its buffers are free */
                    begun = mpm_gpu_frame_begin(job, &wbuf, job->frame, job->t);
                }
                if (begun) {
                    pthread_mutex_lock(&wlock);
                    wfull = 1;
                    pthread_cond_broadcast(&wcond);
                    pthread_mutex_unlock(&wlock);
                } else {
                    write_frame(frames_fd, job->frame, job->t, job);
                }
                job->frame++;
            }
            time_varying_loads(job);
        }
    }
    return NULL;
}

/* write_state (src/writer.c:266-380): the format the reference's restart
 * reader expects. */
static void write_state(FILE *fd, const job_t *job)
{
    fprintf(fd, "%lg %lg %lg\n", job->t, job->dt, job->t_stop);
    fprintf(fd, "%zu %zu %zu\n", job->num_particles, job->num_nodes, job->num_elements);
    fprintf(fd, "%zu %lg\n", job->N, job->h);
    for (size_t i = 0; i < job->num_particles; i++) {
        const particle_t *p = &job->particles[i];
        fprintf(fd, "%lg %lg\n", p->x, p->y);
        fprintf(fd, "%lg %lg\n", p->xl, p->yl);
        fprintf(fd, "%lg %lg\n", p->x_t, p->y_t);
        fprintf(fd, "%lg\n", p->m);
        fprintf(fd, "%lg\n", p->v);
        fprintf(fd, "%lg\n", p->v0);
        fprintf(fd, "%lg %lg %lg\n", p->sxx, p->sxy, p->syy);
        fprintf(fd, "%lg %lg %lg %lg\n", p->exx_t, p->exy_t, p->eyy_t, p->wxy_t);
        fprintf(fd, "%lg %lg\n", p->bx, p->by);
        fprintf(fd, "%lg %lg %lg %lg\n", p->Fxx, p->Fxy, p->Fyx, p->Fyy);
        fprintf(fd, "%lg %lg\n", p->ux, p->uy);
        fprintf(fd, "%lg\n", p->color);
        fprintf(fd, "%d\n", DEPVAR);
        for (int j = 0; j < DEPVAR; j++) {
            fprintf(fd, "%lg\n", p->state[j]);
        }
        fprintf(fd, "%d\n", job->active[i]);
    }
    for (size_t i = 0; i < job->num_nodes; i++) {
        fprintf(fd, "%lg %lg\n", job->nodes[i].x, job->nodes[i].y);
    }
    for (size_t i = 0; i < job->num_elements; i++) {
        fprintf(fd, "%d %d %d %d\n", job->elements[i].nodes[0], job->elements[i].nodes[1],
                job->elements[i].nodes[2], job->elements[i].nodes[3]);
    }
}

/* The state file written by write_state, read back for a restart (the
 * format of src/reader.c:123-243; the reference's reader writes through an
 * unallocated active array and returns a job without colours or plugins,
 * src/reader.c:215 -- here the job is rebuilt by mpm_init and the saved
 * per-particle fields are put back after the material's init). */
typedef struct saved_state {
    double t, dt, t_stop, h;
    size_t np, nn, ne, N;
    particle_t *p;
    int *active;
} saved_state;

static int read_state(const char *path, saved_state *st)
{
    FILE *fd = fopen(path, "r");
    if (!fd) {
        fprintf(stderr, "cannot open state file '%s'\n", path);
        return -1;
    }
    memset(st, 0, sizeof *st);
    int r = fscanf(fd, "%lg %lg %lg", &st->t, &st->dt, &st->t_stop);
    r += fscanf(fd, "%zu %zu %zu", &st->np, &st->nn, &st->ne);
    r += fscanf(fd, "%zu %lg", &st->N, &st->h);
    if (r != 8 || st->N < 2 || st->nn != st->N * st->N) {
        fprintf(stderr, "Error reading state header!\n");
        fclose(fd);
        return -1;
    }
    st->p = calloc(st->np ? st->np : 1, sizeof(particle_t));
    st->active = calloc(st->np ? st->np : 1, sizeof(int));
    for (size_t i = 0; i < st->np; i++) {
        particle_t *q = &st->p[i];
        int dep = 0;
        r = fscanf(fd, "%lg %lg", &q->x, &q->y);
        r += fscanf(fd, "%lg %lg", &q->xl, &q->yl);
        r += fscanf(fd, "%lg %lg", &q->x_t, &q->y_t);
        r += fscanf(fd, "%lg", &q->m);
        r += fscanf(fd, "%lg", &q->v);
        r += fscanf(fd, "%lg", &q->v0);
        r += fscanf(fd, "%lg %lg %lg", &q->sxx, &q->sxy, &q->syy);
        r += fscanf(fd, "%lg %lg %lg %lg", &q->exx_t, &q->exy_t, &q->eyy_t, &q->wxy_t);
        r += fscanf(fd, "%lg %lg", &q->bx, &q->by);
        r += fscanf(fd, "%lg %lg %lg %lg", &q->Fxx, &q->Fxy, &q->Fyx, &q->Fyy);
        r += fscanf(fd, "%lg %lg", &q->ux, &q->uy);
        r += fscanf(fd, "%lg", &q->color);
        r += fscanf(fd, "%d", &dep);
        if (dep < 0 || dep > DEPVAR) {
            fprintf(stderr, "error reading state of particle %zu\n", i);
            fclose(fd);
            return -1;
        }
        for (int j = 0; j < dep; j++) {
            r += fscanf(fd, "%lg", &q->state[j]);
        }
        r += fscanf(fd, "%d", &st->active[i]);
        if (r != 26 + dep + 1) {
            fprintf(stderr, "error reading state of particle %zu\n", i);
            fclose(fd);
            return -1;
        }
    }
    fclose(fd);   /* node coordinates and element nodes follow; mpm_init rebuilds them */
    return 0;
}

static void write_hex_state(const char *path, const job_t *job)
{
    FILE *fd = fopen(path, "w");
    if (!fd) {
        fprintf(stderr, "cannot write '%s'\n", path);
        return;
    }
    fprintf(fd, "%a %d\n", job->t, job->stepcount);
    for (size_t i = 0; i < job->num_particles; i++) {
        const particle_t *p = &job->particles[i];
        fprintf(fd, "%a %a %a %a %a %a %a %a %a %a %a %a %a %d\n", p->x, p->y, p->x_t, p->y_t,
                p->v, p->sxx, p->sxy, p->syy, p->T[ZZ], p->state[9], p->state[10], p->ux,
                p->uy, job->active[i]);
    }
    fclose(fd);
}

typedef struct run_opts {
    const char *cfgfile, *outdir, *grid, *pfile, *mat, *bc, *frames, *state, *restart;
    double dt, rate, t_stop;
    int steps, have_t_stop, check_cfg;
    size_t T;
    double mprops[64], bprops[64], biprops_d[64], miprops_d[64];
    size_t nm, nb, nbi, nmi;
} run_opts;

static int load_plugins(job_t *job, const char *mat, const char *bc)
{
    void *mh = dlopen(mat, RTLD_LAZY);
    if (!mh) {
        fprintf(stderr, "Can't dlopen() material file '%s': %s.\n", mat, dlerror());
        return EXIT_ERROR_MATERIAL_FILE;
    }
    void *bh = dlopen(bc, RTLD_LAZY);
    if (!bh) {
        fprintf(stderr, "FATAL -- Can't dlopen() boundary condition file '%s': %s.\n", bc,
                dlerror());
        return EXIT_ERROR_BC_PROPS;
    }
    *(void **)&job->material.material_init = dlsym(mh, "material_init");
    *(void **)&job->material.calculate_stress = dlsym(mh, "calculate_stress");
    *(void **)&job->material.calculate_stress_threaded = dlsym(mh, "calculate_stress_threaded");
    *(void **)&job->boundary.bc_init = dlsym(bh, "bc_init");
    *(void **)&job->boundary.bc_validate = dlsym(bh, "bc_validate");
    *(void **)&job->boundary.bc_time_varying = dlsym(bh, "bc_time_varying");
    *(void **)&job->boundary.bc_momentum = dlsym(bh, "bc_momentum");
    *(void **)&job->boundary.bc_force = dlsym(bh, "bc_force");
    if (!job->material.material_init || !job->material.calculate_stress
        || !job->material.calculate_stress_threaded) {
        fprintf(stderr, "Error loading the material's symbols from '%s'.\n", mat);
        return EXIT_ERROR_MATERIAL_FILE;
    }
    if (!job->boundary.bc_init || !job->boundary.bc_validate || !job->boundary.bc_time_varying
        || !job->boundary.bc_momentum || !job->boundary.bc_force) {
        fprintf(stderr, "FATAL -- No boundary conditions.\n");
        return EXIT_ERROR_BC_PROPS;
    }
    return 0;
}

/* The reference's .cfg options (src/main.c:169-251, 380-730) into `o` and
 * the job's output section; returns 0 or an exit code. */
static int apply_cfg(const cfg_file *cf, run_opts *o, job_t *job)
{
    (void)job;
    if (!o->mat) {
        if (cfg_int(cf, "material|use-builtin", 1) != 0) {
            fprintf(stderr, "The build has no builtin material: set material.use-builtin = 0 "
                    "and material.material-file.\n");
            return EXIT_ERROR_MATERIAL_FILE;
        }
        o->mat = cfg_str(cf, "material|material-file", "builtin.so");
    }
    if (!o->bc) {
        if (cfg_int(cf, "boundary-conditions|use-builtin", 1) != 0) {
            fprintf(stderr, "FATAL -- No boundary conditions.\n");
            return EXIT_ERROR_BC_PROPS;
        }
        o->bc = cfg_str(cf, "boundary-conditions|boundary-conditions-file", "builtin.so");
    }
    o->nm = cfg_size(cf, "material|properties");
    o->nmi = cfg_size(cf, "material|integer-properties");
    o->nb = cfg_size(cf, "boundary-conditions|properties");
    o->nbi = cfg_size(cf, "boundary-conditions|integer-properties");
    if (o->nm > 64 || o->nmi > 64 || o->nb > 64 || o->nbi > 64) {
        fprintf(stderr, "too many properties\n");
        return EXIT_ERROR_CFG_PARSE;
    }
    for (size_t i = 0; i < o->nm; i++) o->mprops[i] = cfg_nfloat(cf, "material|properties", i);
    for (size_t i = 0; i < o->nmi; i++) {
        o->miprops_d[i] = (double)cfg_nint(cf, "material|integer-properties", i);
    }
    for (size_t i = 0; i < o->nb; i++) {
        o->bprops[i] = cfg_nfloat(cf, "boundary-conditions|properties", i);
    }
    for (size_t i = 0; i < o->nbi; i++) {
        o->biprops_d[i] = (double)cfg_nint(cf, "boundary-conditions|integer-properties", i);
    }
    const char *solver = cfg_str(cf, "solver|solver-type", "implicit");
    if (strcmp(solver, "implicit") && strcmp(solver, "explicit-usf")
        && strcmp(solver, "explicit-usl")) {
        fprintf(stderr, "Invalid value for option 'solver-type': %s\n", solver);
        return EXIT_ERROR_CFG_PARSE;
    }
    o->rate = cfg_float(cf, "output|sample-rate", 60.0f);
    if (o->rate < 0) o->rate = 60.0f;
    return 0;
}

int main(int argc, char **argv)
{
    run_opts o;
    memset(&o, 0, sizeof o);
    o.rate = 60.0;
    o.T = 1;
    for (int a = 1; a < argc; a++) {
        const char *k = argv[a];
        const char *v = a + 1 < argc ? argv[a + 1] : "";
        if (!strcmp(k, "-c")) o.cfgfile = v, a++;
        else if (!strcmp(k, "-o")) o.outdir = v, a++;
        else if (!strcmp(k, "-r")) o.restart = v, a++;
        else if (!strcmp(k, "-g")) o.grid = v, a++;
        else if (!strcmp(k, "-p")) o.pfile = v, a++;
        else if (!strcmp(k, "-u")) o.mat = v, a++;
        else if (!strcmp(k, "-b")) o.bc = v, a++;
        else if (!strcmp(k, "-t")) {
            int t = atoi(v);
            o.T = t > 0 ? (size_t)t : 1;
            a++;
        }
        else if (!strcmp(k, "--dt")) o.dt = strtod(v, NULL), a++;
        else if (!strcmp(k, "--steps")) o.steps = atoi(v), a++;
        else if (!strcmp(k, "--mprops")) o.nm = parse_list(v, o.mprops, 64), a++;
        else if (!strcmp(k, "--bprops")) o.nb = parse_list(v, o.bprops, 64), a++;
        else if (!strcmp(k, "--biprops")) o.nbi = parse_list(v, o.biprops_d, 64), a++;
        else if (!strcmp(k, "--sample-rate")) o.rate = strtod(v, NULL), a++;
        else if (!strcmp(k, "--frames")) o.frames = v, a++;
        else if (!strcmp(k, "--state")) o.state = v, a++;
        else if (!strcmp(k, "--musl")) use_musl = 1;
        else if (!strcmp(k, "--sync-frames")) sync_frames = 1;
        else if (!strcmp(k, "--check-cfg")) o.check_cfg = 1;
        else if (k[0] != '-' && !o.have_t_stop) {
            o.t_stop = strtod(k, NULL);
            o.have_t_stop = 1;
        } else {
            fprintf(stderr, "unknown option %s\n", k);
            return EXIT_ERROR_CFG_PARSE;
        }
    }
    cfg_file cf;
    memset(&cf, 0, sizeof cf);
    const int cfg_mode = o.cfgfile != NULL;
    if (cfg_mode) {
        if (cfg_read(&cf, o.cfgfile)) {
            fprintf(stderr, "Fatal -- cannot parse configuration file '%s': %s\n", o.cfgfile,
                    cf.error);
            return EXIT_ERROR_CFG_PARSE;
        }
        if (!o.pfile) o.pfile = cfg_str(&cf, "input|initial-particle-file", "particles.txt");
        if (!o.grid) o.grid = cfg_str(&cf, "input|grid-file", "grid.txt");
    } else if ((!o.restart && (!o.grid || !o.pfile || o.dt <= 0)) || !o.mat || !o.bc
               || o.steps <= 0) {
        fprintf(stderr, "usage: mpm_usl_run -c SIM.cfg [-o DIR] [-p PF] [-g GF] [-u MAT.so] "
                "[-b BC.so] [-t T] [--musl] [t_max]\n"
                "   or: mpm_usl_run -g GRID -p PARTICLES -u MAT.so -b BC.so --dt DT --steps K "
                "[-t T] [--mprops ..] [--bprops ..] [--biprops ..] [--musl]\n");
        return EXIT_ERROR_CFG_PARSE;
    }
    if (cfg_mode && o.check_cfg) {
        /* print the settings the run would use, without running */
        int rc = apply_cfg(&cf, &o, NULL);
        if (rc) {
            cfg_release(&cf);
            return rc;
        }
        printf("particle-file=%s\ngrid-file=%s\nmaterial-file=%s\nbc-file=%s\n", o.pfile,
               o.grid, o.mat, o.bc);
        printf("dt=%.17g\nautomatic-dt=%ld\nsample-rate=%.17g\n",
               cfg_float(&cf, "timestep|dt", 1e-6), cfg_int(&cf, "timestep|automatic-dt", 1),
               o.rate);
        printf("directory=%s\nuser=%s\nsolver=%s\n",
               o.outdir ? o.outdir : cfg_str(&cf, "output|directory", "jobs/"),
               cfg_str(&cf, "output|user", "unknown"), cfg_str(&cf, "solver|solver-type", "implicit"));
        printf("material-properties=");
        for (size_t i = 0; i < o.nm; i++) printf("%s%.17g", i ? "," : "", o.mprops[i]);
        printf("\nbc-properties=");
        for (size_t i = 0; i < o.nb; i++) printf("%s%.17g", i ? "," : "", o.bprops[i]);
        printf("\nbc-integer-properties=");
        for (size_t i = 0; i < o.nbi; i++) printf("%s%d", i ? "," : "", (int)o.biprops_d[i]);
        printf("\nt_stop=%.17g\n", o.have_t_stop ? o.t_stop : 1.0);
        cfg_release(&cf);
        return 0;
    }
    int N;
    double h;
    particle_t *pdata;
    size_t np;
    saved_state saved;
    memset(&saved, 0, sizeof saved);
    if (o.restart) {
        /* -r STATE (src/main.c:326-332): grid and particles from the state */
        if (read_state(o.restart, &saved)) {
            return EXIT_ERROR_CFG_PARSE;
        }
        N = (int)saved.N;
        h = saved.h;
        np = saved.np;
        pdata = calloc(np ? np : 1, sizeof(particle_t));
        memcpy(pdata, saved.p, np * sizeof(particle_t));
        if (!cfg_mode) o.dt = saved.dt;
    } else {
        double len;
        if (read_grid(o.grid, &N, &len)) {
            fprintf(stderr, "Error reading grid file.\n");
            return EXIT_ERROR_CFG_PARSE;
        }
        if (read_particles(o.pfile, &pdata, &np)) {
            fprintf(stderr, "Error reading particle file.\n");
            return EXIT_ERROR_CFG_PARSE;
        }
        h = len / (N - 1);
    }
    const double t_begin = o.restart ? saved.t : 0.0;
    const double t_stop = cfg_mode ? (o.have_t_stop ? o.t_stop : (o.restart ? saved.t_stop : 1.0))
                                   : t_begin + (o.steps - 0.5) * o.dt;
    job_t *job = mpm_init(N, h, pdata, np, t_stop);

    char *dir = NULL;
    FILE *info_fd = NULL, *element_fd = NULL, *state_fd = NULL, *log_fd = NULL;
    if (cfg_mode) {
        int rc = apply_cfg(&cf, &o, job);
        if (rc) return rc;
        /* timestep (src/main.c:615-636): mpm_init's automatic dt unless
         * automatic-dt = 0 */
        const double cfg_dt = cfg_float(&cf, "timestep|dt", 1e-6);
        job->dt = cfg_int(&cf, "timestep|automatic-dt", 1) != 0 ? job->dt : cfg_dt;
        if (o.restart) job->dt = saved.dt;
        /* output files (src/main.c:660-715) */
        const char *d = o.outdir ? o.outdir : cfg_str(&cf, "output|directory", "jobs/");
        struct stat st;
        if (stat(d, &st) != 0 || !S_ISDIR(st.st_mode)) {
            fprintf(stderr, "Unable to open '%s' for option 'output|directory'.\n", d);
            return EXIT_ERROR_CFG_PARSE;
        }
        size_t dl = strlen(d);
        dir = malloc(dl + 2);
        strcpy(dir, d);
        if (dl == 0 || dir[dl - 1] != '/') strcat(dir, "/");
        char path[16384];
        const char *names[5] = {"info.txt",
                                cfg_str(&cf, "output|particle-file", "frame_particle_data.txt"),
                                cfg_str(&cf, "output|element-file", "frame_element_data.txt"),
                                cfg_str(&cf, "output|state-file", "state.txt"),
                                cfg_str(&cf, "output|log-file", "job.log")};
        FILE **fds[5] = {&info_fd, &frames_fd, &element_fd, &state_fd, &log_fd};
        for (int k = 0; k < 5; k++) {
            snprintf(path, sizeof path, "%s%s", dir, names[k]);
            *fds[k] = fopen(path, "w");
            if (!*fds[k]) {
                fprintf(stderr, "Can't open '%s' for output.\n", path);
                return EXIT_ERROR_CFG_PARSE;
            }
        }
        job->output.directory = dir;
        job->output.log_fd = log_fd;
    } else {
        job->dt = o.dt;
        if (o.frames) frames_fd = fopen(o.frames, "w");
    }
    int rc = load_plugins(job, o.mat, o.bc);
    if (rc) return rc;
    job->material.num_fp64_props = o.nm;
    job->material.fp64_props = o.mprops;
    int miprops[64], biprops[64];
    for (size_t i = 0; i < o.nmi; i++) miprops[i] = (int)o.miprops_d[i];
    for (size_t i = 0; i < o.nbi; i++) biprops[i] = (int)o.biprops_d[i];
    job->material.num_int_props = o.nmi;
    job->material.int_props = miprops;
    job->boundary.num_fp64_props = o.nb;
    job->boundary.fp64_props = o.bprops;
    job->boundary.num_int_props = o.nbi;
    job->boundary.int_props = biprops;
    job->output.sample_rate_hz = o.rate;

    const size_t T = o.T;
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
    if (o.restart) {
        /* continue the saved run: its clock, loads, and the per-particle
         * fields mpm_init and material_init reset */
        job->t = saved.t;
        for (size_t i = 0; i < np; i++) {
            particle_t *q = &job->particles[i];
            const particle_t *sp = &saved.p[i];
            q->xl = sp->xl;
            q->yl = sp->yl;
            q->v0 = sp->v0;
            q->exx_t = sp->exx_t;
            q->exy_t = sp->exy_t;
            q->eyy_t = sp->eyy_t;
            q->wxy_t = sp->wxy_t;
            q->bx = sp->bx;
            q->by = sp->by;
            q->Fxx = sp->Fxx;
            q->Fxy = sp->Fxy;
            q->Fyx = sp->Fyx;
            q->Fyy = sp->Fyy;
            q->ux = sp->ux;
            q->uy = sp->uy;
            q->color = sp->color;
            memcpy(q->state, sp->state, sizeof q->state);
            job->active[i] = saved.active[i];
        }
    }
    if (job->boundary.bc_validate(job) == 0) {
        fprintf(stderr, "Error with boundary condition properties.\n");
        return EXIT_ERROR_BC_PROPS;
    }
    job->frame = (size_t)floor(job->t * job->output.sample_rate_hz);

    if (frames_fd && !sync_frames) {
        if (mpm_gpu_frame_alloc(job, &wbuf) == 0) {
            mpm_gpu_async_frames(job, 1);
            wjob = job;
            writer_on = pthread_create(&writer, NULL, frame_writer, NULL) == 0;
            if (!writer_on) {
                mpm_gpu_async_frames(job, 0);
            }
        } else {
            mpm_gpu_frame_free(&wbuf);
        }
    }
    pthread_t *th = malloc(sizeof(pthread_t) * T);
    for (size_t i = 0; i + 1 < T; i++) {
        pthread_create(&th[i], NULL, run_until, &tasks[i]);
    }
    run_until(&tasks[T - 1]);
    for (size_t i = 0; i + 1 < T; i++) {
        pthread_join(th[i], NULL);
    }
    if (writer_on) {
        pthread_mutex_lock(&wlock);
        wquit = 1;
        pthread_cond_broadcast(&wcond);
        pthread_mutex_unlock(&wlock);
        pthread_join(writer, NULL);
        mpm_gpu_frame_free(&wbuf);
        mpm_gpu_async_frames(job, 0);
    }
    if (state_fd) {
        write_state(state_fd, job);
    }
    if (o.state) {
        write_hex_state(o.state, job);
    }
    FILE *close_fds[5] = {info_fd, frames_fd, element_fd, state_fd, log_fd};
    for (int k = 0; k < 5; k++) {
        if (close_fds[k]) fclose(close_fds[k]);
    }
    printf("ran %d steps to t = %g on %zu threads\n", job->stepcount, job->t, T);
    mpm_cleanup(job);
    free(dir);
    cfg_release(&cf);
    return 0;
}
