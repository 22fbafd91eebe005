/* jcfd_profile.h -- measurement hooks of libjcfd_amd.
 *
 * No counterpart in the reference (jax-cfd has no kernels of its own to
 * time).  For benchmarks: each function launches one kernel class of a plan
 * `reps` times back to back on `stream`, brackets the launches with HIP
 * events recorded on that same stream, and reports the mean duration per
 * launch in milliseconds.  Inputs are read, outputs go to the plan's scratch
 * (the caller's arrays are not modified).
 */
#ifndef JCFD_PROFILE_H_
#define JCFD_PROFILE_H_

#include "jcfd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel classes of one FVM step, in launch order. */
#define JCFD_FVM_KERNEL_ROWS_FORWARD 0 /* stencil + Euler + div + row FFT */
#define JCFD_FVM_KERNEL_POISSON_COLS 1 /* column FFT, Lambda^+, inverse */
#define JCFD_FVM_KERNEL_ROWS_PROJECT 2 /* inverse row FFT + correction */
#define JCFD_FVM_KERNEL_COUNT 3

/* ms_per_launch[k] = mean duration of kernel class k over `reps` launches
 * with the plan's tile shapes; u, v are the plan's (batch, n, n) fields (or
 * (n, n) for one instance) and every instance is processed. */
int jcfd_fvm_time_kernels(jcfd_fvm_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                          int32_t reps, double* ms_per_launch, void* stream);

/* Algorithmic HBM bytes one launch of each kernel class moves over every
 * instance of the plan (distinct logical inputs read once + outputs written
 * once). */
int jcfd_fvm_kernel_bytes(const jcfd_fvm_plan* plan, int64_t* bytes_per_launch);

/* Learned-interpolation step: ms_per_launch[k] for each of the tower's
 * n_layers convolutions (the output layer is the coefficient layer), then
 * the learned explicit-terms kernel, then the combination u + dt k with its
 * projection (three launches, timed together); n_layers + 2 entries. */
int jcfd_li_time_kernels(jcfd_li_plan* plan, const jcfd_tensor* u, const jcfd_tensor* v,
                         int32_t reps, double* ms_per_launch, void* stream);
/* Algorithmic floating-point operations of each convolution of the tower
 * (2 * 9 * cin * cout multiply-adds per cell, cout = the tower's own output
 * channels for the last layer), n_layers entries. */
int jcfd_li_kernel_flops(const jcfd_li_plan* plan, double* flops_per_launch);

/* Kernel classes of one CN-RK4 step of the spectral plan. */
#define JCFD_NS2D_KERNEL_COLS_INVERSE 0       /* stage-0 inverse column FFTs */
#define JCFD_NS2D_KERNEL_ROWS 1               /* row FFTs + nonlinear product */
#define JCFD_NS2D_KERNEL_COLS_UPDATE_INVERSE 2 /* forward cols + update + next inverse */
#define JCFD_NS2D_KERNEL_COLS_UPDATE 3        /* forward cols + final update */
#define JCFD_NS2D_KERNEL_COUNT 4

int jcfd_ns2d_time_kernels(jcfd_ns2d_plan* plan, const jcfd_tensor* w_hat, int32_t reps,
                           double* ms_per_launch, void* stream);
int jcfd_ns2d_kernel_bytes(const jcfd_ns2d_plan* plan, int64_t* bytes_per_launch);

/* Mixed-size groups: `reps` further group steps of the member set of the
 * last jcfd_fvm_group_step (their states advance), with events around each
 * of the three phase launches; ms_per_launch[k] = mean duration of phase k
 * (the JCFD_FVM_KERNEL_* classes). */
int jcfd_fvm_group_time_kernels(jcfd_fvm_group* group, int32_t reps, double* ms_per_launch,
                                void* stream);
/* Shape of the last group step: workgroups per phase launch, work items per
 * phase, and the algorithmic bytes per phase launch (the per-instance
 * jcfd_fvm_kernel_bytes summed over the member instances). */
int jcfd_fvm_group_launch_shape(const jcfd_fvm_group* group, int32_t* grid, int32_t* items,
                                int64_t* bytes_per_launch);

/* Spectral groups: `reps` further steps of the member set of the last
 * jcfd_ns2d_group_step on the plans' internal states (the members' own
 * states are not touched), events around every launch; ms_per_launch[k] =
 * mean duration of class k (the JCFD_NS2D_KERNEL_* classes). */
int jcfd_ns2d_group_time_kernels(jcfd_ns2d_group* group, int32_t reps, double* ms_per_launch,
                                 void* stream);
/* Shape of the last spectral group step: workgroups per class launch, work
 * items per launch, algorithmic bytes per launch (jcfd_ns2d_kernel_bytes per
 * instance, summed over the member instances). */
int jcfd_ns2d_group_launch_shape(const jcfd_ns2d_group* group, int32_t* grid, int32_t* items,
                                 int64_t* bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif /* JCFD_PROFILE_H_ */
