/*
 * mpm_abi.h -- plugin-ABI data types of the 2-D MPM USL step.
 *
 * The material and boundary-condition plugins of the reference driver are
 * shared objects that receive job_t* / threadtask_t* and index straight into
 * particle_t / node_t / element_t arrays.  A drop-in replacement of the step
 * therefore has to keep these structures byte-for-byte: same field order,
 * same types, same padding.  The field names are kept too, so plugin sources
 * written against the reference headers compile unchanged against this one.
 *
 * Reference definitions reproduced here:
 *   particle_t   libmpm/particle.h:34-147   (1280 bytes)
 *   node_t       libmpm/node.h:11-54        (152 bytes)
 *   element_t    libmpm/element.h:29-67     (1504 bytes)
 *   job_t        libmpm/process.h:116-213   (800 bytes on x86-64 glibc)
 *   threadtask_t libmpm/process.h:215-249   (96 bytes)
 * Offsets are checked against a probe of the reference headers in the tests.
 */
#ifndef MPM_ABI_H
#define MPM_ABI_H

#include <stddef.h>
#include <stdio.h>
#include <time.h>
#include <pthread.h>

#ifdef __cplusplus
extern "C" {
#endif

/* particle.h constants */
#define DEPVAR 11
#define S_XIDX 0
#define S_YIDX 1
#define S_ZIDX 2
#define NDIM 3
#define XX (NDIM * S_XIDX + S_XIDX)
#define XY (NDIM * S_XIDX + S_YIDX)
#define XZ (NDIM * S_XIDX + S_ZIDX)
#define YX (NDIM * S_YIDX + S_XIDX)
#define YY (NDIM * S_YIDX + S_YIDX)
#define YZ (NDIM * S_YIDX + S_ZIDX)
#define ZX (NDIM * S_ZIDX + S_XIDX)
#define ZY (NDIM * S_ZIDX + S_YIDX)
#define ZZ (NDIM * S_ZIDX + S_ZIDX)

/* element.h constants */
#define NODAL_DOF 3
#define NODES_PER_ELEMENT 4
#define XDOF_IDX 0
#define YDOF_IDX 1

/* exitcodes.h */
enum exitcodes_e {
    EXIT_ERROR_CS_ENTRY = 0x01,
    EXIT_ERROR_CS_SOL,
    EXIT_ERROR_CS_DUP,
    EXIT_ERROR_CFG_PARSE = 0x20,
    EXIT_ERROR_MATERIAL_FILE,
    EXIT_ERROR_BC_PROPS,
    EXIT_ERROR_DT_TOO_SMALL = 0x40,
    EXIT_ERROR_THREADING = 0x80
};

/* One material point.  Only a handful of fields are live in the USL step
 * (x, y, v, m, x_t, y_t, x_tt, y_tt, bx, by, sxx..syy, T, exx_t..wxy_t, L,
 * ux, uy, state); the rest is carried for plugins and output. */
typedef struct particle_s {
    double x, y;                 /* position */
    double xl, yl;               /* local coordinates in the element */
    double v, v0;                /* current / initial volume */
    double s[4];                 /* shape functions (unused by the USL step) */
    double m;                    /* mass */
    double x_t, y_t;             /* velocity */
    double x_tt, y_tt;           /* acceleration */
    double bx, by;               /* body force per unit mass */
    double sxx, sxy, syy;        /* in-plane Cauchy stress */
    double T[NDIM * NDIM];       /* full 3-D stress (row-major) */
    double exx_t, exy_t, eyy_t, wxy_t;  /* strain rate and spin */
    double L[NDIM * NDIM];       /* velocity gradient */
    double Fxx, Fxy, Fyx, Fyy;   /* deformation gradient */
    double ux, uy;               /* accumulated displacement */
    double color;
    double state[DEPVAR];        /* material state variables */
    int material;
    double grad_s[4][2];
    int element;
    double real_sxx, real_sxy, real_syy;
    double real_state[DEPVAR];
    double r1_initial[2], r2_initial[2];
    double r1[2], r2[2];
    double corners[4][2];
    double cornersl[4][2];
    double sc[4][4];
    double grad_sc[4][4][2];
    int corner_elements[4];
    void *material_data;
    size_t id;
} particle_t;

/* One grid node. */
typedef struct node_s {
    int num_filled_element_neighbors;
    double mass_filled_element_neighbors;
    double m;                    /* mass */
    double x, y;                 /* position */
    double ux, uy;               /* displacement increment */
    double x_t, y_t;             /* velocity */
    double x_tt, y_tt;           /* acceleration */
    double mx_t, my_t;           /* momentum */
    double mx_tt, my_tt;         /* "pseudoforce" (dead in USL) */
    double fx, fy;               /* force */
    double rho;
    int velocity_update_flag;
} node_t;

/* One grid cell (element).  nodes[] = {BL, BR, TR, TL}. */
typedef struct element_s {
    int color;
    int color_idx;
    int nodes[NODES_PER_ELEMENT];
    int n;
    double m;
    int filled;
    int neighbors[8];
    double grad_x, grad_y, grad_mag;
    double n_x, n_y, n_theta;
    double v;
    int p_index;
    double p_dist;
    double sxx, sxy, syy;
    double exx_t, exy_t, eyy_t;
    double wxy_t;
    double p, tau;
    double kku_element[NODAL_DOF * NODES_PER_ELEMENT][NODAL_DOF * NODES_PER_ELEMENT];
    double f_element[NODAL_DOF * NODES_PER_ELEMENT];
    double jacobian[2][2];
} element_t;

enum solver_e {
    IMPLICIT_SOLVER = 0,
    EXPLICIT_SOLVER_USF,
    EXPLICIT_SOLVER_USL,
    NUM_SOLVERS
};

typedef struct material_s {
    double E, nu, rho;
} material_t;

typedef struct ts_control_s {
    double dt_max, dt_min, dt;
    int automatic_dt;
    int allow_dt_increase;
    int stable_dt_threshold;
} timestep_control_t;

typedef struct im_control_s {
    double du_norm_ratio, q_norm_ratio, du_norm_converged;
    int unstable_iteration_count;
    int stable_step_count;
} implicit_control_t;

typedef struct op_control_s {
    char *directory, *user;
    char *particle_filename, *element_filename, *state_filename, *log_filename;
    char *particle_filename_fullpath, *element_filename_fullpath;
    char *state_filename_fullpath, *log_filename_fullpath;
    FILE *particle_fd, *element_fd, *state_fd, *log_fd;
    FILE *info_fd;
    char *job_name, *job_description;
    int job_id;
    char *date;
    int override_directory_with_id;
    int prepend_date;
    int modified_directory;
    double sample_rate_hz;
} output_control_t;

struct job_s;

/* Material plugin hooks (resolved by dlsym in the driver). */
typedef struct mat_control_s {
    const char *material_filename;
    int use_builtin;
    void (*material_init)(struct job_s *);
    void (*calculate_stress)(struct job_s *);
    void (*calculate_stress_threaded)(void *);
    double *fp64_props;
    int *int_props;
    size_t num_fp64_props;
    size_t num_int_props;
} material_control_t;

/* Boundary-condition plugin hooks (resolved by dlsym in the driver). */
typedef struct bc_control_s {
    const char *bc_filename;
    int use_builtin;
    void (*bc_init)(void *);
    int (*bc_validate)(void *);
    void (*bc_time_varying)(void *);
    void (*bc_momentum)(void *);
    void (*bc_force)(void *);
    double *fp64_props;
    int *int_props;
    size_t num_fp64_props;
    size_t num_int_props;
} boundary_control_t;

typedef struct job_s {
    double t, dt, t_stop;
    size_t frame;
    int stepcount;

    size_t num_particles, num_nodes, num_elements, num_colors;
    size_t *color_indices;
    size_t *particle_by_element_color_lengths;   /* [thread * num_colors + color] */
    size_t **particle_by_element_color_lists;

    size_t N;                    /* nodes per side */
    double h;                    /* node spacing */

    int use_cpdi;
    struct timespec tic, toc;

    node_t *nodes;
    particle_t *particles;
    element_t *elements;

    int *in_element;
    int *active;

    /* per-particle shape functions and gradients (SoA) */
    double *h1, *h2, *h3, *h4;
    double *b11, *b12, *b13, *b14;
    double *b21, *b22, *b23, *b24;

    int *node_u_map, *inv_node_u_map;

    double *u_dirichlet;         /* [NODAL_DOF * num_nodes] */
    int *u_dirichlet_mask;       /* [NODAL_DOF * num_nodes] */
    int *node_number_override;   /* [NODAL_DOF * num_nodes] */

    int vec_len;
    enum solver_e solver;
    timestep_control_t timestep;
    implicit_control_t implicit;
    output_control_t output;
    material_control_t material;
    boundary_control_t boundary;

    int step_number;
    double step_start_time;

    pthread_barrier_t *step_barrier;
    pthread_barrier_t *serialize_barrier;
    size_t num_threads;

    int *update_elementlists;
    int *update_elementlists_flag;
} job_t;

typedef struct s_threadtask {
    size_t id;
    size_t num_threads;
    size_t gidx_min, gidx_max;
    size_t offset, blocksize;       /* particle slice */
    size_t n_offset, n_blocksize;   /* node slice */
    size_t e_offset, e_blocksize;   /* element slice */
    size_t stride;
    job_t *job;
} threadtask_t;

#ifndef __cplusplus
_Static_assert(sizeof(particle_t) == 1280, "particle_t layout");
_Static_assert(sizeof(node_t) == 152, "node_t layout");
_Static_assert(sizeof(element_t) == 1504, "element_t layout");
_Static_assert(sizeof(threadtask_t) == 96, "threadtask_t layout");
#if defined(__x86_64__) && defined(__GLIBC__)
_Static_assert(sizeof(job_t) == 800, "job_t layout");
#endif
#endif

#ifdef __cplusplus
}
#endif
#endif /* MPM_ABI_H */
