// elastic2d.hpp -- kernel argument blocks and launchers of the 2D elastic
// velocity-stress time step (forward, adjoint, imaging) for gfx950.
#pragma once

#include "common.hpp"

namespace erm {

// Field order everywhere: 0 Vx, 1 Vz, 2 sigma_xx, 3 sigma_zz, 4 sigma_xz.
// Coefficient order: 0 m1_x, 1 m1_z, 2 aux_m2, 3 aux_m3, 4 aux_m2m3 (each
// already scaled by dt/dx and zeroed outside its operator's index range).
struct StepArgs {
  int nz, nx, ldz;     // grid and column stride of every field below
  float* out[5];       // level being computed
  const float* x2[5];  // level it-2 (forward) / it+2 (adjoint); may alias out
  const float* y1[5];  // level it-1 (forward) / it+1 (adjoint)
  const float* xin[5]; // adjoint only: level-it content added first (drop-in form) or null
  const float* coef[5];
  const float* kz;     // sponge row factors (nz) or null = no sponge
  const float* kx;     // sponge column factors (nx)
  // explosive source (forward): amp added to sxx, szz at (zs, xs)
  int zs, xs;
  float src_amp;
  int has_src;
  const float* bsrc[5];   // Born scattering sources of this level (forward) or null
  // receivers on row zr, columns rx0 .. rx0+nrx-1
  int zr, rx0, nrx;
  float* rec_out[5];      // forward: row sample of this level, contiguous over ix
  const float* rec_in[5]; // adjoint: injected row of this level
  // source-wavefield history written by the forward step (full grid, fp32)
  float* hvx;
  float* hvz;
  // imaging (adjoint): source Vx, Vz at this level and the next one
  const float* svx;
  const float* svz;
  const float* svx1;  // null at the last level (S1 = S2 = 0 there)
  const float* svz1;
  float* img[3];      // drho, dlambda, dmu
  float* imgc[3];     // Neumaier compensation terms (comp mode)
  float inv_dx, inv_dt, img_w;
  int do_image;
  int comp;
  int xcd_remap;
};

// Tile geometry of the step kernels: 64 z-cells (one wave64 along the fast
// axis) by 4*TXR x-cells, 256 threads, radius-5 halo staged in LDS.
constexpr int TILE_Z = 64;
constexpr int TILE_WY = 4;

int launch_fwd_step(const StepArgs& a, int txr, hipStream_t st);
int launch_adj_step(const StepArgs& a, int txr, hipStream_t st);

// Born scattering sources of one level (image_op2 forward branch,
// W/image-mod-rtm.f90:163-179) from the background Vx, Vz at levels it, it+1.
struct BornArgs {
  int nz, nx, ldz;
  const float* svx; const float* svz; const float* svx1; const float* svz1;
  const float* dm[3];
  float* src[5];
  float inv_dx, inv_dt;
};
int launch_born_sources(const BornArgs& a, hipStream_t st);

// out[i] = value over n floats (fill kernel; used instead of hipMemsetAsync so
// the launch sequence stays capturable in a HIP graph).
int launch_fill(float* out, size_t n, float value, hipStream_t st);

// Sponge K applied in place to one field (abc_k_hip).
int launch_apply_k(float* f, int nz, int nx, int ldz, const float* kz, const float* kx,
                   hipStream_t st);

// Final image reduction: img = (double)s + (double)c, rounded to float.
int launch_image_finalize(float* s, const float* c, size_t n, hipStream_t st);

}  // namespace erm
