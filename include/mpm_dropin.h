/*
 * mpm_dropin.h -- the reference libmpm entry points that libmpm_gpu.so
 * (mpm-2d_amd/host/libmpm_gpu.c) exports with the reference's names,
 * signatures and semantics, so the reference driver (src/main.c) and its
 * material / boundary-condition plugins link against it unchanged.
 *
 * The step itself runs on the GPU through include/mpm_hip.h; everything else
 * here keeps the reference's host-side behaviour.
 */
#ifndef MPM_DROPIN_H
#define MPM_DROPIN_H

#include <stdio.h>

#include "mpm_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* libmpm/process.h:251, process_usl.c:68-280.  Builds the job exactly as the
 * reference does (AoS particles/nodes/elements, colours, element neighbours,
 * in_element/active, v0, automatic dt); no device work yet. */
job_t *mpm_init(int N, double h, particle_t *particles, size_t num_particles, double t);

/* libmpm/process.h:252, process_usl.c:465-592.  Called by every worker thread
 * once per step with its threadtask_t, like the reference.  The thread that
 * the serialize barrier elects drives the GPU (the first call uploads the
 * job); the others wait at the same barrier.  job->t advances by job->dt on
 * the host exactly as in the reference.  Particle state is copied back into
 * job->particles / job->active whenever the driver is about to read it: at
 * a frame time (job->t >= job->frame / sample_rate_hz, src/main.c:955) and at
 * the end of the run (job->t >= job->t_stop, src/main.c:847).  Device or
 * configuration failures print a message and exit with
 * EXIT_ERROR_THREADING, the reference's code for a failed worker set-up. */
void explicit_mpm_step_usl_threaded(void *task);

/* libmpm/process_usl.h:15, process_usl.c:284-462: the MUSL step, same
 * calling convention (the reference driver's loop has it commented out,
 * src/main.c:948).  A run uses one scheme: the step function called first
 * fixes it. */
void explicit_mpm_step_musl_threaded(void *task);

/* libmpm/process.h:253, process_usl.c:974-1004 (also releases the device). */
void mpm_cleanup(job_t *job);

/* process_usl.c:677-723: per-(thread, colour) particle lists as the
 * reference's driver expects them before the loop (src/main.c:812). */
void find_filled_elements(job_t *job);

/* libmpm/loading.c:21-54 */
void initial_loads(job_t *job);
void time_varying_loads(job_t *job);

/* Host helpers that plugins link against (libmpm/particle.c:14-30,
 * interpolate.c:23-65, element.c:10-19, tensor.c:64-437). */
void global_to_local_coords(double *x_local, double *y_local, double x, double y,
                            double x_ref, double y_ref, double h);
void tent(double *h1, double *h2, double *h3, double *h4, double x_local, double y_local);
void grad_tent(double *b11, double *b12, double *b13, double *b14,
               double *b21, double *b22, double *b23, double *b24,
               double x_local, double y_local, double h);
void node_number_to_coords(double *x, double *y, int num, int N, double h);
void zero_array(double *array, size_t numel);
void tensor_trace3(double *trA, const double *A);
void tensor_add3(double *C, const double *A, const double *B);
void tensor_copy3(double *C, const double *A);
void tensor_decompose3(double *A_0, double *c, const double *A);
void tensor_multiply3(double *C, const double *A, const double *B);
void tensor_scale3(double *A, double c);
void tensor_sym3(double *C, const double *A);
void tensor_skw3(double *C, const double *A);
void tensor_contraction3(double *c, const double *A, const double *B);

/* ------------------------------------------------------------------------
 * Extensions beyond the reference's libmpm, for this repository's driver
 * (mpm-2d_amd/host/mpm_usl_run.c): frames taken by snapshot, so the driver
 * formats and writes frame k on a thread of its own while the steps after
 * it run.  The reference driver writes each frame synchronously from the
 * job's AoS (src/main.c:955-973, src/writer.c:190-201); the file written
 * from a snapshot is byte-identical.
 * ---------------------------------------------------------------------- */
#define MPM_GPU_FRAME_FIELDS 12
typedef struct mpm_gpu_frame {
    /* page-locked host copies of the fields write_frame prints that the
     * device advances (x_t, y_t: the velocity; state9, state10: gammap and
     * gammadotp) */
    double *v, *x, *y, *x_t, *y_t, *sxx, *sxy, *syy, *ux, *uy, *state9, *state10;
    int *active;
    size_t frame;
    double t;
    int state_from_job;     /* state[9], state[10] come from the job's AoS */
} mpm_gpu_frame_t;

/* on = 1: the step no longer copies particles back at frame times (only at
 * the end of the run); the driver takes each frame with mpm_gpu_frame_begin. */
void mpm_gpu_async_frames(job_t *job, int on);
int mpm_gpu_frame_alloc(job_t *job, mpm_gpu_frame_t *f);
void mpm_gpu_frame_free(mpm_gpu_frame_t *f);
/* From the serial section at a frame time, after the step: starts the
 * snapshot of the state after the steps so far and returns 1 at once; 0
 * when the frame must be written from the job's AoS now (a host material
 * plugin keeps the stress there). */
int mpm_gpu_frame_begin(job_t *job, mpm_gpu_frame_t *f, size_t frame, double t);
/* Waits for the snapshot (from any thread) and writes the frame in the
 * format of write_frame; m, color and corners come from the AoS.  The
 * particle lines are formatted in up to 16 chunks on threads of its own and
 * written in order (the same bytes).  Returns 0, or -1 if a write failed. */
int mpm_gpu_frame_write(FILE *fd, const job_t *job, const mpm_gpu_frame_t *f);

#ifdef __cplusplus
}
#endif
#endif /* MPM_DROPIN_H */
