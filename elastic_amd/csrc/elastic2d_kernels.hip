// elastic2d_kernels.hip -- gfx950 kernels of the 2D elastic velocity-stress
// RTM time step (10th-order staggered grid, radius 5).
//
// One launch advances all five fields (Vx, Vz, sxx, szz, sxz) by one time
// level. A workgroup owns a 64 (z) x TX (x) tile: one wave64 spans the fast z
// axis so every global access of a wave is one contiguous 256-byte column
// segment, four waves split the x extent, and the radius-5 halo of the
// level being differentiated is staged once in LDS. The stencils are not a
// dense contraction, so they run on the VALU; the kernel is bounded by HBM
// bandwidth (80 algorithmic bytes per cell-update).
//
// The reference operators (W/rtmlib-mod.f90:200-633) all reduce to two
// staggered differences along either axis:
//   S+ f(i) = sum_k c_k (f[i+6-k] - f[i-5+k])   (D, E, F, G, Az, Bx)
//   S- f(i) = sum_k c_k (f[i+5-k] - f[i-6+k])   (Ax, Bz, H, J)
// and their transposes satisfy (m S+)^T = -S-(m .), (m S-)^T = -S+(m .), so
// the gather-form adjoint (W/rtmlib-modC.c:120-253) is the same stencil
// applied to coefficient-weighted products, which the adjoint kernel forms
// while staging the halo.

#include "elastic2d.hpp"

namespace erm {

template <int TXR>
struct Tile {
  static constexpr int TZ = TILE_Z;
  static constexpr int WY = TILE_WY;
  static constexpr int TX = WY * TXR;
  static constexpr int H = 5;
  static constexpr int LZ = TZ + 2 * H;  // LDS column length (z incl. halo)
  static constexpr int LX = TX + 2 * H;  // LDS columns (x incl. halo)
  static constexpr int N = LX * LZ;
};

// Staggered differences on an LDS tile; S is the element stride of the axis.
template <int S>
__device__ __forceinline__ float sp(const float* f, int p) {
  float r = C1 * (f[p + 5 * S] - f[p - 4 * S]);
  r += C2 * (f[p + 4 * S] - f[p - 3 * S]);
  r += C3 * (f[p + 3 * S] - f[p - 2 * S]);
  r += C4 * (f[p + 2 * S] - f[p - S]);
  r += C5 * (f[p + S] - f[p]);
  return r;
}
template <int S>
__device__ __forceinline__ float sm(const float* f, int p) {
  float r = C1 * (f[p + 4 * S] - f[p - 5 * S]);
  r += C2 * (f[p + 3 * S] - f[p - 4 * S]);
  r += C3 * (f[p + 2 * S] - f[p - 3 * S]);
  r += C4 * (f[p + S] - f[p - 2 * S]);
  r += C5 * (f[p] - f[p - S]);
  return r;
}

// Workgroups are dispatched round-robin over the 8 XCDs; renumber them so
// each XCD receives a contiguous range of tiles (column-major tile order),
// which keeps neighbouring tiles -- and their shared halos -- in one L2.
__device__ __forceinline__ int xcd_tile(int L, int N) {
  const int q = N >> 3, r = N & 7;
  const int x = L & 7, pos = L >> 3;
  return (x < r) ? x * (q + 1) + pos : r * (q + 1) + (x - r) * q + pos;
}

// Neumaier compensated accumulation of t into (s, c).
__device__ __forceinline__ void neumaier(float& s, float& c, float t) {
  const float u = s + t;
  if (fabsf(s) >= fabsf(t))
    c += (s - u) + t;
  else
    c += (t - u) + s;
  s = u;
}

// ---------------------------------------------------------------------------
// Forward step, W/rtm-mod-adj.f90:225-252 (and the Born propagation :323-348):
//   Vx  = K(Vx(it-2)  + m1_x (S-x sxx + S+z sxz))            Ax, Bx
//   Vz  = K(Vz(it-2)  + m1_z (S+x sxz + S-z szz))            Az, Bz
//   sxx = K(sxx(it-2) + src + m2m3 S+x Vx + m2 S+z Vz)       D, E
//   szz = K(szz(it-2) + src + m2 S+x Vx + m2m3 S+z Vz)       F, G
//   sxz = K(sxz(it-2) + m3 (S-z Vx + S-x Vz))                H, J
// with all differences of level it-1. Coefficients are zero outside each
// operator's range, which reproduces the reference loop bounds exactly.
template <int TXR>
__global__ __launch_bounds__(256) void k_fwd_step(const StepArgs a, int nbz, int ntiles) {
  using T = Tile<TXR>;
  __shared__ float sh[5][T::N];
  const int t = a.xcd_remap ? xcd_tile(blockIdx.x, ntiles) : (int)blockIdx.x;
  const int z0 = (t % nbz) * T::TZ, x0 = (t / nbz) * T::TX;
  const int tz = threadIdx.x, ty = threadIdx.y;

  for (int lx = ty; lx < T::LX; lx += T::WY) {
    const int gx = x0 - T::H + lx;
    const bool xin = gx >= 0 && gx < a.nx;
    for (int lz = tz; lz < T::LZ; lz += T::TZ) {
      const int gz = z0 - T::H + lz;
      const bool in = xin && gz >= 0 && gz < a.nz;
      const size_t gi = (size_t)gz + (size_t)a.ldz * gx;
      const int p = lx * T::LZ + lz;
#pragma unroll
      for (int f = 0; f < 5; ++f) sh[f][p] = in ? a.y1[f][gi] : 0.f;
    }
  }
  __syncthreads();

  const int gz = z0 + tz;
  if (gz >= a.nz) return;
  const float kz = a.kz ? a.kz[gz] : 1.f;
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    if (gx >= a.nx) break;
    const int p = (lxo + T::H) * T::LZ + (tz + T::H);
    const size_t gi = (size_t)gz + (size_t)a.ldz * gx;
    const float m1x = a.coef[0][gi], m1z = a.coef[1][gi], m2 = a.coef[2][gi],
                m3 = a.coef[3][gi], m2m3 = a.coef[4][gi];
    float v[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) v[f] = a.x2[f][gi];
    if (a.bsrc[0]) {
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = a.bsrc[f][gi] + v[f];
    }
    constexpr int SX = T::LZ;
    const float dxx_xm = sm<SX>(sh[2], p);  // Ax
    const float dxz_zp = sp<1>(sh[4], p);   // Bx
    const float dxz_xp = sp<SX>(sh[4], p);  // Az
    const float dzz_zm = sm<1>(sh[3], p);   // Bz
    const float dvx_xp = sp<SX>(sh[0], p);  // D, F
    const float dvz_zp = sp<1>(sh[1], p);   // E, G
    const float dvx_zm = sm<1>(sh[0], p);   // H
    const float dvz_xm = sm<SX>(sh[1], p);  // J
    v[0] = v[0] + m1x * dxx_xm;
    v[0] = v[0] + m1x * dxz_zp;
    v[1] = v[1] + m1z * dxz_xp;
    v[1] = v[1] + m1z * dzz_zm;
    if (a.has_src && gz == a.zs && gx == a.xs) {
      v[2] = v[2] + a.src_amp;
      v[3] = v[3] + a.src_amp;
    }
    v[2] = v[2] + m2m3 * dvx_xp;
    v[2] = v[2] + m2 * dvz_zp;
    v[3] = v[3] + m2 * dvx_xp;
    v[3] = v[3] + m2m3 * dvz_zp;
    v[4] = v[4] + m3 * dvx_zm;
    v[4] = v[4] + m3 * dvz_xm;
    if (a.kz) {
      const float kx = a.kx[gx];
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = (v[f] * kz) * kx;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) a.out[f][gi] = v[f];
    if (a.hvx) {
      a.hvx[gi] = v[0];
      a.hvz[gi] = v[1];
    }
    if (a.rec_out[0] && gz == a.zr && gx >= a.rx0 && gx < a.rx0 + a.nrx) {
#pragma unroll
      for (int f = 0; f < 5; ++f) a.rec_out[f][gx - a.rx0] = v[f];
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint step (gather form of rtm_op3_c_, W/rtmlib-modC.c:120-253, with the
// receiver injection of W/rtmlib-mod.f90:99-146):
//   Vx  = K'(inj + Vx(it+2)  - S-x(m2m3 sxx + m2 szz) - S+z(m3 sxz))
//   Vz  = K'(inj + Vz(it+2)  - S-z(m2 sxx + m2m3 szz) - S+x(m3 sxz))
//   sxx = K'(inj + sxx(it+2) - S+x(m1_x Vx))
//   szz = K'(inj + szz(it+2) - S+z(m1_z Vz))
//   sxz = K'(inj + sxz(it+2) - S-z(m1_x Vx) - S-x(m1_z Vz))
// products of level it+1. K' is the sponge (exact adjoint) or identity.
// With do_image, the Lame imaging condition of this level
// (W/image-mod-rtm.f90:154-162, S1..S6 of :26-67) is accumulated in the same
// pass from the source Vx, Vz history.
template <int TXR>
__global__ __launch_bounds__(256) void k_adj_step(const StepArgs a, int nbz, int ntiles) {
  using T = Tile<TXR>;
  __shared__ float sh[5][T::N];
  const int t = a.xcd_remap ? xcd_tile(blockIdx.x, ntiles) : (int)blockIdx.x;
  const int z0 = (t % nbz) * T::TZ, x0 = (t / nbz) * T::TX;
  const int tz = threadIdx.x, ty = threadIdx.y;

  for (int lx = ty; lx < T::LX; lx += T::WY) {
    const int gx = x0 - T::H + lx;
    const bool xin = gx >= 0 && gx < a.nx;
    for (int lz = tz; lz < T::LZ; lz += T::TZ) {
      const int gz = z0 - T::H + lz;
      const bool in = xin && gz >= 0 && gz < a.nz;
      const int p = lx * T::LZ + lz;
      float p0 = 0.f, p1 = 0.f, p2 = 0.f, p3 = 0.f, p4 = 0.f;
      if (in) {
        const size_t gi = (size_t)gz + (size_t)a.ldz * gx;
        const float vx = a.y1[0][gi], vz = a.y1[1][gi], xx = a.y1[2][gi], zz = a.y1[3][gi],
                    xz = a.y1[4][gi];
        const float m1x = a.coef[0][gi], m1z = a.coef[1][gi], m2 = a.coef[2][gi],
                    m3 = a.coef[3][gi], m2m3 = a.coef[4][gi];
        p0 = m2m3 * xx + m2 * zz;
        p1 = m2 * xx + m2m3 * zz;
        p2 = m3 * xz;
        p3 = m1x * vx;
        p4 = m1z * vz;
      }
      sh[0][p] = p0;
      sh[1][p] = p1;
      sh[2][p] = p2;
      sh[3][p] = p3;
      sh[4][p] = p4;
    }
  }
  __syncthreads();

  const int gz = z0 + tz;
  const bool zok = gz < a.nz;
  const float kz = (a.kz && zok) ? a.kz[gz] : 1.f;
  const bool row = a.rec_in[0] && gz == a.zr;
  float o[TXR][5];
  constexpr int SX = T::LZ;
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    const bool ok = zok && gx < a.nx;
    const int p = (lxo + T::H) * T::LZ + (tz + T::H);
#pragma unroll
    for (int f = 0; f < 5; ++f) o[j][f] = 0.f;
    if (!ok) continue;
    const size_t gi = (size_t)gz + (size_t)a.ldz * gx;
    float b[5];
    if (a.xin[0]) {
#pragma unroll
      for (int f = 0; f < 5; ++f) b[f] = a.xin[f][gi] + a.x2[f][gi];
    } else if (row && gx >= a.rx0 && gx < a.rx0 + a.nrx) {
#pragma unroll
      for (int f = 0; f < 5; ++f) b[f] = a.rec_in[f][gx - a.rx0] + a.x2[f][gi];
    } else {
#pragma unroll
      for (int f = 0; f < 5; ++f) b[f] = a.x2[f][gi];
    }
    float v0 = b[0] - sm<SX>(sh[0], p);
    v0 -= sp<1>(sh[2], p);
    float v1 = b[1] - sm<1>(sh[1], p);
    v1 -= sp<SX>(sh[2], p);
    const float v2 = b[2] - sp<SX>(sh[3], p);
    const float v3 = b[3] - sp<1>(sh[4], p);
    float v4 = b[4] - sm<1>(sh[3], p);
    v4 -= sm<SX>(sh[4], p);
    float v[5] = {v0, v1, v2, v3, v4};
    if (a.kz) {
      const float kx = a.kx[gx];
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = (v[f] * kz) * kx;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      a.out[f][gi] = v[f];
      o[j][f] = v[f];
    }
  }

  if (!a.do_image) return;  // uniform over the grid: no barrier is skipped unevenly
  __syncthreads();
  for (int lx = ty; lx < T::LX; lx += T::WY) {
    const int gx = x0 - T::H + lx;
    const bool xin = gx >= 0 && gx < a.nx;
    for (int lz = tz; lz < T::LZ; lz += T::TZ) {
      const int gz2 = z0 - T::H + lz;
      const bool in = xin && gz2 >= 0 && gz2 < a.nz;
      const size_t gi = (size_t)gz2 + (size_t)a.ldz * gx;
      const int p = lx * T::LZ + lz;
      sh[0][p] = in ? a.svx[gi] : 0.f;
      sh[1][p] = in ? a.svz[gi] : 0.f;
    }
  }
  __syncthreads();
  const bool zin4 = gz >= 4 && gz <= a.nz - 6, zin5 = gz >= 5 && gz <= a.nz - 6;
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    if (!zok || gx >= a.nx) continue;
    const int p = (lxo + T::H) * T::LZ + (tz + T::H);
    const size_t gi = (size_t)gz + (size_t)a.ldz * gx;
    // S1, S2: forward time difference (T, W/image-mod-rtm.f90:288-297)
    float s1 = 0.f, s2 = 0.f;
    if (a.svx1) {
      s1 = a.inv_dt * (a.svx1[gi] - sh[0][p]);
      s2 = a.inv_dt * (a.svz1[gi] - sh[1][p]);
    }
    // S4 = vxdx, S5 = vzdz (range 4..n-6), S6 = vxdz + vzdx (range 5..n-6)
    float s4 = 0.f, s5 = 0.f, s6 = 0.f;
    if (zin4 && gx >= 4 && gx <= a.nx - 6) {
      s4 = sp<SX>(sh[0], p) * a.inv_dx;
      s5 = sp<1>(sh[1], p) * a.inv_dx;
    }
    if (zin5 && gx >= 5 && gx <= a.nx - 6) {
      s6 = sm<1>(sh[0], p) * a.inv_dx;
      s6 += sm<SX>(sh[1], p) * a.inv_dx;
    }
    const float s3 = s4 + s5;
    const float* v = o[j];
    const float w = a.img_w;
    const float t1 = w * (v[0] * s1 + v[1] * s2);
    const float t2 = -w * (v[2] * s3 + v[3] * s3);
    const float t3 = -w * (v[2] * 2.f * s4 + v[3] * 2.f * s5 + v[4] * s6);
    if (a.comp) {
      float s = a.img[0][gi], c = a.imgc[0][gi];
      neumaier(s, c, t1);
      a.img[0][gi] = s; a.imgc[0][gi] = c;
      s = a.img[1][gi]; c = a.imgc[1][gi];
      neumaier(s, c, t2);
      a.img[1][gi] = s; a.imgc[1][gi] = c;
      s = a.img[2][gi]; c = a.imgc[2][gi];
      neumaier(s, c, t3);
      a.img[2][gi] = s; a.imgc[2][gi] = c;
    } else {
      a.img[0][gi] += t1;
      a.img[1][gi] += t2;
      a.img[2][gi] += t3;
    }
  }
}

// ---------------------------------------------------------------------------
// Born scattering sources of one level, image_op2 forward branch
// (W/image-mod-rtm.f90:171-179) with S1..S6 of the background level.
__global__ __launch_bounds__(256) void k_born_sources(const BornArgs a) {
  constexpr int TXR = 4;
  using T = Tile<TXR>;
  __shared__ float sh[2][T::N];
  const int nbz = (a.nz + T::TZ - 1) / T::TZ;
  const int t = blockIdx.x;
  const int z0 = (t % nbz) * T::TZ, x0 = (t / nbz) * T::TX;
  const int tz = threadIdx.x, ty = threadIdx.y;
  for (int lx = ty; lx < T::LX; lx += T::WY) {
    const int gx = x0 - T::H + lx;
    const bool xin = gx >= 0 && gx < a.nx;
    for (int lz = tz; lz < T::LZ; lz += T::TZ) {
      const int gz = z0 - T::H + lz;
      const bool in = xin && gz >= 0 && gz < a.nz;
      const size_t gi = (size_t)gz + (size_t)a.ldz * gx;
      const int p = lx * T::LZ + lz;
      sh[0][p] = in ? a.svx[gi] : 0.f;
      sh[1][p] = in ? a.svz[gi] : 0.f;
    }
  }
  __syncthreads();
  const int gz = z0 + tz;
  if (gz >= a.nz) return;
  constexpr int SX = T::LZ;
  for (int j = 0; j < TXR; ++j) {
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    if (gx >= a.nx) break;
    const int p = (lxo + T::H) * T::LZ + (tz + T::H);
    const size_t gi = (size_t)gz + (size_t)a.ldz * gx;
    float s1 = 0.f, s2 = 0.f, s4 = 0.f, s5 = 0.f, s6 = 0.f;
    if (a.svx1) {
      s1 = a.inv_dt * (a.svx1[gi] - sh[0][p]);
      s2 = a.inv_dt * (a.svz1[gi] - sh[1][p]);
    }
    if (gz >= 4 && gz <= a.nz - 6 && gx >= 4 && gx <= a.nx - 6) {
      s4 = sp<SX>(sh[0], p) * a.inv_dx;
      s5 = sp<1>(sh[1], p) * a.inv_dx;
    }
    if (gz >= 5 && gz <= a.nz - 6 && gx >= 5 && gx <= a.nx - 6) {
      s6 = sm<1>(sh[0], p) * a.inv_dx;
      s6 += sm<SX>(sh[1], p) * a.inv_dx;
    }
    const float s3 = s4 + s5;
    const float d1 = a.dm[0][gi], d2 = a.dm[1][gi], d3 = a.dm[2][gi];
    a.src[0][gi] = d1 * s1;
    a.src[1][gi] = d1 * s2;
    a.src[2][gi] = -(d2 * s3) - (d3 * 2.f * s4);
    a.src[3][gi] = -(d2 * s3) - (d3 * 2.f * s5);
    a.src[4][gi] = -(d3 * s6);
  }
}

__global__ void k_fill(float* out, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = v;
}

__global__ void k_apply_k(float* f, int nz, int nx, int ldz, const float* kz, const float* kx) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y;
  if (iz >= nz || ix >= nx) return;
  const size_t i = (size_t)iz + (size_t)ldz * ix;
  f[i] = (f[i] * kz[iz]) * kx[ix];
}

__global__ void k_image_finalize(float* s, const float* c, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    s[i] = (float)((double)s[i] + (double)c[i]);
}

// ---------------------------------------------------------------------------
template <int TXR>
static int launch_fwd_t(const StepArgs& a, hipStream_t st) {
  using T = Tile<TXR>;
  const int nbz = (a.nz + T::TZ - 1) / T::TZ;
  const int nbx = (a.nx + T::TX - 1) / T::TX;
  const int nt = nbz * nbx;
  hipLaunchKernelGGL(k_fwd_step<TXR>, dim3(nt), dim3(T::TZ, T::WY), 0, st, a, nbz, nt);
  HIPCHK(hipGetLastError());
  return OK;
}
template <int TXR>
static int launch_adj_t(const StepArgs& a, hipStream_t st) {
  using T = Tile<TXR>;
  const int nbz = (a.nz + T::TZ - 1) / T::TZ;
  const int nbx = (a.nx + T::TX - 1) / T::TX;
  const int nt = nbz * nbx;
  hipLaunchKernelGGL(k_adj_step<TXR>, dim3(nt), dim3(T::TZ, T::WY), 0, st, a, nbz, nt);
  HIPCHK(hipGetLastError());
  return OK;
}

int launch_fwd_step(const StepArgs& a, int txr, hipStream_t st) {
  switch (txr) {
    case 1: return launch_fwd_t<1>(a, st);
    case 2: return launch_fwd_t<2>(a, st);
    case 4: return launch_fwd_t<4>(a, st);
    case 8: return launch_fwd_t<8>(a, st);
    default: set_error("unsupported tile width %d", txr); return E_ARG;
  }
}
int launch_adj_step(const StepArgs& a, int txr, hipStream_t st) {
  switch (txr) {
    case 1: return launch_adj_t<1>(a, st);
    case 2: return launch_adj_t<2>(a, st);
    case 4: return launch_adj_t<4>(a, st);
    case 8: return launch_adj_t<8>(a, st);
    default: set_error("unsupported tile width %d", txr); return E_ARG;
  }
}

int launch_born_sources(const BornArgs& a, hipStream_t st) {
  using T = Tile<4>;
  const int nbz = (a.nz + T::TZ - 1) / T::TZ;
  const int nbx = (a.nx + T::TX - 1) / T::TX;
  hipLaunchKernelGGL(k_born_sources, dim3(nbz * nbx), dim3(T::TZ, T::WY), 0, st, a);
  HIPCHK(hipGetLastError());
  return OK;
}

int launch_fill(float* out, size_t n, float value, hipStream_t st) {
  if (n == 0) return OK;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, out, n, value);
  HIPCHK(hipGetLastError());
  return OK;
}

int launch_apply_k(float* f, int nz, int nx, int ldz, const float* kz, const float* kx,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_apply_k, dim3((nz + 255) / 256, nx), dim3(256), 0, st, f, nz, nx, ldz, kz,
                     kx);
  HIPCHK(hipGetLastError());
  return OK;
}

int launch_image_finalize(float* s, const float* c, size_t n, hipStream_t st) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_image_finalize, dim3((unsigned)blocks), dim3(256), 0, st, s, c, n);
  HIPCHK(hipGetLastError());
  return OK;
}

}  // namespace erm
