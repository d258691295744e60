// elastic2d.hpp -- device descriptors and launchers of the 2D elastic
// velocity-stress RTM step kernels (gfx950).
#pragma once

#include "common.hpp"

namespace erm {

// Field order everywhere: 0 Vx, 1 Vz, 2 sigma_xx, 3 sigma_zz, 4 sigma_xz.
// Coefficient order: 0 m1_x, 1 m1_z, 2 aux_m2, 3 aux_m3, 4 aux_m2m3 (each
// already scaled by dt/dx and zeroed outside its operator's index range).

// Snapshot formats of the source wavefield kept for the imaging condition.
enum SnapFmt : int {
  SNAP_HIST = 0,  // fp32 Vx, Vz of every level (the plain form)
  SNAP_F32 = 1,   // every j-th level: Vx, Vz, dVx, dVz in fp32
  SNAP_BF16 = 2,  // same four arrays in bf16 with a max-abs scale per strip
  SNAP_F16 = 3    // same four arrays in fp16 with a max-abs scale per strip
};

// One shot (one computational grid) as the kernels see it. Lives in device
// memory; a kernel reads the descriptor of its tile's shot with scalar loads.
struct ShotDev {
  int nz, nx, ldz, nt;
  long long n2;  // ldz * nx: elements per field
  // Level storage: field f of level L is at
  //   lev + ((L + lev_off) % lev_mod) * lev_stride + f * fstride
  // (ring: lev_mod 2, lev_stride 5*n2, fstride n2; drop-in history: 0-based
  // levels, lev_mod nt, lev_stride n2, fstride nt*n2).
  float* lev;
  long long lev_stride, fstride;
  int lev_mod, lev_off;
  const float* coef;  // 5 * n2, masked
  const float* kz;    // sponge factors (nz), null = none
  const float* kx;    // (nx)
  // explosive source: src_amp[L-1] = dt*source(L-1) is added at (zs, xs) to
  // sxx and szz when level L is computed
  int zs, xs;
  const float* src_amp;
  // receivers: row zr, columns rx0..rx0+nrx-1; rec is [5][nt][nrx], level L
  // at row L-1 (recorded by the forward, injected by the adjoint)
  int zr, rx0, nrx;
  float* rec;
  // source-wavefield snapshots
  int snap_fmt, snap_every, snap_sw, snap_nbz, snap_nstrip;
  void* snap;          // storage (see SnapFmt)
  float* snap_scale;   // [k][4][nstrip] for SNAP_BF16 / SNAP_F16
  // imaging
  float* img;   // 3 * n2: drho, dlambda, dmu
  float* imgc;  // 3 * n2 compensation terms
  float inv_dx, inv_dt, img_w;
  // Born modelling
  float* bsrc;       // 5 * n2 scattering sources of the current level
  const float* dm;   // 3 * n2 model perturbation
};

struct TileRef {
  int shot, z0, x0, tile;
};

// Level L of a shot's storage (ring or drop-in history).
__device__ __forceinline__ float* lev_ptr(float* lev, long long stride, int mod, int off, int L) {
  int slot = (L + off) % mod;
  if (slot < 0) slot += mod;  // levels before the first one map into the ring too
  return lev + (long long)slot * stride;
}

// Forward flags
enum : int {
  F_REC = 1,     // record row zr into rec (store)
  F_RECSUB = 2,  // record by subtracting (observed = true - smooth)
  F_K = 4,       // apply the sponge
  F_SRC = 8,     // inject the explosive source
  F_BORN = 16,   // add the Born sources bsrc
  F_SNAP = 32,   // write the source-wavefield snapshot belonging to this step
  F_NTS = 64,    // non-temporal stores of the new level
  F_SNAP2 = 128  // two-level step: also write the snapshot of level L
};
// Adjoint flags
enum : int {
  A_XIN = 1,   // drop-in form: add the level's own content first
  A_INJ = 2,   // inject the receiver row of this level
  A_K = 4,     // exact-adjoint sponge
  A_IMG = 8,   // accumulate the imaging condition at this level
  A_NEXT = 16, // level L+1 exists (S1, S2 nonzero)
  A_COMP = 32, // compensated image accumulation
  A_NTS = 64,  // non-temporal stores of the new level
  A_IMG2 = 128, // two-level step: also image level L-1
  A_KSRC = 256  // with A_K: the R/Src sponge. The ring holds K-damped levels
                // T(L) = K R(L); R(L) = inj + K T(L+2) + Op^T T(L+1); the
                // level imaged is K T(L) for L >= 3, T(2), R(1)
};

// Workgroups are dispatched round-robin over the 8 XCDs; renumber them so
// each XCD receives a contiguous range of tiles (shot-major, column-major),
// which keeps neighbouring tiles -- and their shared halos -- in one L2.
__device__ __forceinline__ int xcd_tile(int L, int N) {
  const int q = N >> 3, r = N & 7;
  const int x = L & 7, pos = L >> 3;
  return (x < r) ? x * (q + 1) + pos : r * (q + 1) + (x - r) * q + pos;
}

// Tile geometry: 64 z-cells (one wave64 along the fast axis) by 4*TXR
// x-cells, 256 threads, radius-5 halo staged in LDS.
constexpr int TILE_Z = 64;
constexpr int TILE_WY = 4;
inline int tile_x(int txr) { return TILE_WY * txr; }

// Step launches over a tile table covering one or many shots.
int launch_fwd(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int flags,
               int txr, int snap_fmt, int xcd, hipStream_t st);
int launch_adj(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int flags,
               int txr, int snap_fmt, int xcd, hipStream_t st);
// Levels L and L+1 in one pass (temporal blocking; TXR 1, 2, 4). The ring
// must hold four levels. No Born sources.
int launch_fwd2(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int flags,
                int txr, int snap_fmt, int xcd, hipStream_t st);
// Adjoint levels L and L-1 in one pass (TXR 1, 2, 4; four-level ring; not
// the drop-in history form).
int launch_adj2(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int flags,
                int txr, int snap_fmt, int xcd, hipStream_t st);
// Marching forms: one workgroup of nw waves (nw in {1, 2, 4}) per strip of
// 64*nw rows x xw columns (xw even; strip tables start strips on even
// columns and rows on multiples of 64*nw). Bitwise the same results as
// launch_fwd / launch_adj; the 16-bit snapshot strips are 64 x 2 (the layout
// of tile width 2).
int launch_fwdm(const ShotDev* shots, const TileRef* strips, int nstrips, int L, int flags, int xw, int nw,
                int snap_fmt, int xcd, hipStream_t st);
int launch_adjm(const ShotDev* shots, const TileRef* strips, int nstrips, int L, int flags, int xw, int nw,
                int snap_fmt, int xcd, hipStream_t st);
// Workgroups of the marching adjoint (nw waves, snapshot format, compensated
// image sums or not) resident on the current device at once: per-CU
// occupancy times the CU count (0 if unknown).
int adjm_resident_blocks(int nw, int snap_fmt, bool comp);
// The same for the marching forward (without Born sources).
int fwdm_resident_blocks(int nw, int snap_fmt);
// Snapshot of the last level (no later step writes it): dV = 0.
int launch_snap_last(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int txr,
                     int snap_fmt, hipStream_t st);
// Born scattering sources of level L of every shot (image_op2 forward branch).
int launch_born_sources(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int txr,
                        hipStream_t st);

int launch_fill(float* out, size_t n, float value, hipStream_t st);
int launch_apply_k(float* f, int nz, int nx, int ldz, const float* kz, const float* kx,
                   hipStream_t st);
// img = (double)img + (double)imgc, rounded once.
int launch_image_finalize(float* s, const float* c, size_t n, hipStream_t st);
// stack[z0+iz, x0+ix] += img[iz0+iz, ix0+ix] (fp64 stack, 3 images).
int launch_stack(double* stack, int snz, int snx, const float* img, int ldz, long long n2,
                 int iz0, int ix0, int nz, int nx, int z0, int x0, hipStream_t st);

}  // namespace erm
