/*
 * acoustic3d.h -- C ABI of the 3D acoustic 10th-order propagator and
 * cross-correlation imaging on one GPU's subdomain (part of librtm_hip.so).
 *
 * Replaces, on one GPU, the reference's multi-GPU acoustic RTM functions
 * source_prop (W/gpu_funcs_3d.cu:133-315) and rtm_adjoint (:509-712) with
 * their kernels wave_kernel (W/wave_fkernel.3d8o.cu:31-106), damp_kernel
 * (:183-210), new_src_inject_kernel / new_data_inject_kernel /
 * new_data_extract_kernel (W/gpu_funcs_3d.cu:64-109) and img_kernel
 * (:111-120). The reference is 8th order with sinc time interpolation and
 * P2P halo exchange; this is 10th order (radius 5), samples on the time
 * grid, and one subdomain (halo cells present, not exchanged).
 *
 * Layout: (n1, n2, n3) = (z fastest, x, y), index i1 + n1*(i2 + n2*i3), FP32.
 * Cells within 5 of a face are the halo and stay zero. Receiver data is
 * (nt, n3, n2): level t, plane row i3, column i2. Return 0 or RTM_E_*.
 */
#ifndef ACOUSTIC3D_H
#define ACOUSTIC3D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n1, n2, n3; /* grid including the 5-cell halo */
  int nt;         /* time levels */
  int sz, sx, sy; /* point source cell (interior) */
  int zr;         /* receiver plane z index (interior) */
  int abc;        /* sponge width per face (W/abc-mod.f90 taper per axis; <=0 none) */
  int snap_every; /* imaging / snapshot decimation j */
  int snap_prec;  /* 0 fp32 snapshots, 1 bf16 with a per-snapshot max-abs scale */
} ac3d_cfg_t;

typedef struct ac3d ac3d;

/* c = (v dt / dx)^2 per cell (n1*n2*n3); amp[t-1] = source amplitude added at
 * level t. */
int ac3d_create(ac3d **a, const ac3d_cfg_t *cfg, const float *c, const float *amp);
/* Receiver data on plane zr: data(c_true) - data(c) when c_true is given,
 * else data(c). Device resident afterwards. */
int ac3d_model_data(ac3d *a, const float *c_true);
/* Source pass with snapshots every j levels, then the receiver pass with data
 * injection and img += src(t) * rec(t) every j levels. */
int ac3d_migrate(ac3d *a);
int ac3d_set_data(ac3d *a, const float *rec);
int ac3d_get_data(ac3d *a, float *rec);
int ac3d_get_image(ac3d *a, float *img);
int ac3d_get_snapshot_scales(ac3d *a, float *out, int n);
/* [0] source-pass ms [1] its launches [2] receiver-pass ms [3] its launches
 * [4] last model_data ms */
int ac3d_timings(ac3d *a, double *out6);
int ac3d_destroy(ac3d *a);

#ifdef __cplusplus
}
#endif
#endif /* ACOUSTIC3D_H */
