/*
 * elastic3d.h -- C ABI of the 9-field 3D elastic adjoint time step on one
 * GPU (part of librtm_hip.so).
 *
 * Replaces the reference's 3D elastic kernel test interface (K3/gpu.h:39-50:
 * rtm_gpu_init, rtm_gpu_func, rtm_gpu_final; kernel rtm_gpu_kernel,
 * K3/rtmlib-modC-gpu.cu:374-690; CPU loop K3/rtmlib-modC.c:277-589). The
 * reference's names are taken by its 2D library (W/rtmlib-modC-gpu.cu), so
 * these carry a rtm3d_ prefix; arguments and their order are the reference's.
 *
 * Layout (K3/gpu.h n3d_index_ex): every array is padded by 5 cells per face,
 * (nz+10) x (nx+10) x (ny+10) floats, z fastest, then x, y slowest; interior
 * cell (iz, ix, iy) is at (iz+5) + (nz+10)*((ix+5) + (nx+10)*(iy+5)).
 * Field order: Vy, Vx, Vz, syy, sxx, szz, sxy, sxz, syz. Coefficient order:
 * m1_y, m1_x, m1_z, m2, m3, m2m3. One step updates the interior of the 9
 * out fields: out = out + in1 + (adjoint gathers of coefficient * in); the
 * padding of out is not written, the padding of in and of the coefficients
 * is read (the reference keeps it zero).
 *
 * Every function returns 0 or a negative RTM_E_* code (rtmlib_modC.h); the
 * message is in rtm_hip_last_error().
 */
#ifndef ELASTIC3D_H
#define ELASTIC3D_H

#ifdef __cplusplus
extern "C" {
#endif

/* Device-resident form (no counterpart in the reference, which copies host
 * arrays every call). in, in1, out: 9 device pointers each; coef: 6. Runs
 * nsteps steps with the reference's rotation after every step but the last,
 * (out, in, in1) <- (in1, out, in) (K3/rtmlib-modC-gpu.cu:273-372), and
 * leaves the rotated pointers in in / in1 / out. Synchronous. */
int rtm3d_adj_steps_dev(int ny, int nz, int nx, float *in[9], float *in1[9], float *out[9],
                        const float *const coef[6], int nsteps);

/* [0] ms of the last rtm3d_adj_steps_dev / rtm3d_gpu_func compute phase,
 * [1] its launches, [2] interior cells per launch, [3] planes per y-chunk. */
int rtm3d_timings(double *out4);

/* rtm_gpu_init (K3/rtmlib-modC-gpu.cu:65): allocates the device copies. */
int rtm3d_gpu_init(int ny, int nz, int nx);

/* rtm_gpu_func (K3/rtmlib-modC-gpu.cu:698-788): copies the 27 fields and 6
 * coefficients in, runs Steps_write_back = 1 step (K3/gpu.h:11), copies the 9
 * out fields back. debug is unused (as in the reference); gpu_kernel_time,
 * if not null, receives seconds of copy-in, compute and copy-out. */
int rtm3d_gpu_func(int ny, int nz, int nx,
                   float *Vy0_in, float *Vx0_in, float *Vz0_in, float *sigmayy0_in, float *sigmaxx0_in,
                   float *sigmazz0_in, float *sigmaxy0_in, float *sigmaxz0_in, float *sigmayz0_in,
                   float *Vy0_in1, float *Vx0_in1, float *Vz0_in1, float *sigmayy0_in1, float *sigmaxx0_in1,
                   float *sigmazz0_in1, float *sigmaxy0_in1, float *sigmaxz0_in1, float *sigmayz0_in1,
                   float *Vy0_out, float *Vx0_out, float *Vz0_out, float *sigmayy0_out, float *sigmaxx0_out,
                   float *sigmazz0_out, float *sigmaxy0_out, float *sigmaxz0_out, float *sigmayz0_out,
                   float *m1_y, float *m1_x, float *m1_z, float *m2, float *m3, float *m2m3,
                   float *debug, float *gpu_kernel_time);

/* rtm_gpu_final (K3/rtmlib-modC-gpu.cu:218): frees the device copies. */
int rtm3d_gpu_final(void);

#ifdef __cplusplus
}
#endif
#endif /* ELASTIC3D_H */
