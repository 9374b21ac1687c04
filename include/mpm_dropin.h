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

#ifdef __cplusplus
}
#endif
#endif /* MPM_DROPIN_H */
