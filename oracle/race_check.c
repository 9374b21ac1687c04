/*
 * race_check.c -- test infrastructure: drives the oracle's threaded step
 * (usl_oracle.c, included whole) on several threads over small jobs that
 * reach every threaded path -- USL and MUSL, g_local_mu2, g_nonlocal,
 * isolin and dp_ri_mt, box and square_silo walls, particles crossing cells
 * (with the per-crossing log lines written, process_usl.c:651-663) and
 * leaving the grid (660-665), which runs find_filled_elements serially.
 *
 * Built twice by the Makefile: with -fsanitize=thread (data races between
 * the worker threads and the serial sections elected by the barrier, as
 * src/main.c:747-839 and process_usl.c:465-592 arrange them) and with
 * -fsanitize=address,undefined (out-of-bounds accesses, use after free,
 * undefined arithmetic).  Exit status 0 means every case ran; a sanitizer
 * report makes the run fail (halt_on_error in the test's environment).
 *
 *   race_check LOG_PATH
 */
#include "usl_oracle.c"

/* 4 x 4 particles per cell on the cells [cx0, cx0 + ncx) x [cy0, cy0 + ncy)
 * of an n-cell grid, jittered by +-0.1 cell (a fixed LCG: any positions
 * inside the cells serve), density 1500, initial velocity (vx, vy). */
static size_t make_block(int n, int cx0, int cy0, int ncx, int ncy, double vx, double vy,
                         double **f)
{
    const double h = 1.0 / n;
    const size_t np = (size_t)ncx * ncy * 16;
    for (int k = 0; k < 9; k++) {
        f[k] = calloc(np, sizeof(double));
    }
    unsigned long long s = 0x2545F4914F6CDD1DULL;
    size_t i = 0;
    for (int cy = cy0; cy < cy0 + ncy; cy++) {
        for (int cx = cx0; cx < cx0 + ncx; cx++) {
            for (int b = 0; b < 4; b++) {
                for (int a = 0; a < 4; a++, i++) {
                    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
                    const double jx = ((double)(s >> 11) / 9007199254740992.0 - 0.5) * 0.2;
                    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
                    const double jy = ((double)(s >> 11) / 9007199254740992.0 - 0.5) * 0.2;
                    const double v = (h / 4) * (h / 4);
                    f[0][i] = 1500.0 * v;                       /* m */
                    f[1][i] = v;                                /* v */
                    f[2][i] = (cx + (a + 0.5) / 4 + jx) * h;    /* x */
                    f[3][i] = (cy + (b + 0.5) / 4 + jy) * h;    /* y */
                    f[4][i] = vx;
                    f[5][i] = vy;
                    /* a lithostatic-like compression so the stress laws
                     * leave their zero-pressure branches */
                    f[6][i] = -100.0;                           /* sxx */
                    f[8][i] = -150.0;                           /* syy */
                }
            }
        }
    }
    return np;
}

struct check_case {
    const char *name;
    int scheme, material_kind, bc_kind, steps;
    int cx0, cy0, ncx, ncy;
    double vx, vy, dt;
};

int main(int argc, char **argv)
{
    const char *log_path = argc > 1 ? argv[1] : "/dev/null";
    const int n = 32;
    const struct check_case cases[] = {
        /* a block sliding over the box floor: crossings every step */
        {"usl_box_g_local_mu2", 0, 0, 1, 40, 4, 0, 8, 12, 3.0, 0.0, 1e-4},
        {"musl_box_g_local_mu2", 1, 0, 1, 40, 4, 0, 8, 12, 3.0, 0.0, 1e-4},
        /* a block falling through the open silo orifice: particles leave
         * the grid and are deactivated */
        {"usl_silo_falling", 0, 0, 2, 40, 0, 1, 4, 6, 0.0, -40.0, 1e-4},
        {"usl_box_g_nonlocal", 0, 1, 1, 10, 4, 0, 8, 12, 1.0, 0.0, 1e-5},
        {"usl_box_isolin", 0, 2, 1, 40, 4, 0, 8, 12, 3.0, 0.0, 1e-4},
        {"usl_box_dp_ri_mt", 0, 3, 1, 40, 4, 0, 8, 12, 3.0, 0.0, 1e-4},
    };
    for (size_t c = 0; c < sizeof cases / sizeof cases[0]; c++) {
        const struct check_case *k = &cases[c];
        double *f[9];
        const size_t np = make_block(n, k->cx0, k->cy0, k->ncx, k->ncy, k->vx, k->vy, f);
        oracle_params_t q;
        oracle_default_params(&q);
        q.N = n + 1;
        q.h = 1.0 / n;
        q.dt = k->dt;
        q.num_threads = 4;
        q.scheme = k->scheme;
        q.material_kind = k->material_kind;
        q.bc_kind = k->bc_kind;
        if (k->bc_kind == 1) {
            /* box.c: {x0, y0, x1, y1} walls, frictionless sides, no-slip floor */
            const double fp[4] = {0.0, 0.0, 1.0, 1.0};
            const int ip[4] = {1, 3, 1, 0};
            memcpy(q.bc_fp64, fp, sizeof fp);
            memcpy(q.bc_int, ip, sizeof ip);
        } else {
            /* square_silo.c: {hole_radius, open_time, wall_frac} */
            const double fp[4] = {0.2, 0.0, 0.6, 0.0};
            memcpy(q.bc_fp64, fp, sizeof fp);
        }
        q.log_crossings = 1;
        q.log_path = log_path;
        void *oj = oracle_create(&q, np, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7],
                                 f[8]);
        oracle_run(oj, k->steps / 2);
        oracle_run(oj, k->steps - k->steps / 2);
        const long long crossings = oracle_crossings(oj);
        size_t active = 0;
        {
            double *out = malloc(np * OF_NFIELDS * sizeof(double));
            int *act = malloc(np * sizeof(int));
            int *elem = malloc(np * sizeof(int));
            oracle_get_particles(oj, out, act, elem);
            for (size_t i = 0; i < np; i++) {
                active += act[i] != 0;
            }
            free(out);
            free(act);
            free(elem);
        }
        printf("%s: %zu particles, %d steps, %lld crossings, %zu active at the end\n",
               k->name, np, k->steps, crossings, active);
        oracle_destroy(oj);
        for (int i = 0; i < 9; i++) {
            free(f[i]);
        }
    }
    return 0;
}
