// elastic2d_kernels.hip -- gfx950 kernels of the 2D elastic velocity-stress
// RTM time step (10th-order staggered grid, radius 5).
//
// One launch advances all five fields (Vx, Vz, sxx, szz, sxz) of every shot
// in a tile table by one time level. A workgroup owns a 64 (z) x 4*TXR (x)
// tile of one shot: one wave64 spans the fast z axis so each global access of
// a wave is one contiguous 256-byte column segment, four waves split the x
// extent, and the radius-5 halo of the level being differentiated is staged
// once in LDS. The stencils are not a dense contraction, so they run on the
// VALU; the step is bounded by HBM bandwidth (80 algorithmic bytes per
// cell-update: 5 fields read at level L-1, 5 at L-2, 5 written, 5
// coefficients).
//
// The reference operators (W/rtmlib-mod.f90:200-633) all reduce to two
// staggered differences along either axis:
//   S+ f(i) = sum_k c_k (f[i+6-k] - f[i-5+k])   (D, E, F, G, Az, Bx)
//   S- f(i) = sum_k c_k (f[i+5-k] - f[i-6+k])   (Ax, Bz, H, J)
// and (m S+)^T = -S-(m .), (m S-)^T = -S+(m .), so the gather-form adjoint
// (W/rtmlib-modC.c:120-253) is the same stencil applied to
// coefficient-weighted products, which the adjoint kernel forms while it
// stages the halo.

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
  r = fmaf(C2, f[p + 4 * S] - f[p - 3 * S], r);
  r = fmaf(C3, f[p + 3 * S] - f[p - 2 * S], r);
  r = fmaf(C4, f[p + 2 * S] - f[p - S], r);
  r = fmaf(C5, f[p + S] - f[p], r);
  return r;
}
template <int S>
__device__ __forceinline__ float sm(const float* f, int p) {
  float r = C1 * (f[p + 4 * S] - f[p - 5 * S]);
  r = fmaf(C2, f[p + 3 * S] - f[p - 4 * S], r);
  r = fmaf(C3, f[p + 2 * S] - f[p - 3 * S], r);
  r = fmaf(C4, f[p + S] - f[p - 2 * S], r);
  r = fmaf(C5, f[p] - f[p - S], r);
  return r;
}


__device__ __forceinline__ void neumaier(float& s, float& c, float t) {
#pragma clang fp contract(off)
  const float u = s + t;
  if (fabsf(s) >= fabsf(t))
    c += (s - u) + t;
  else
    c += (t - u) + s;
  s = u;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// bf16 / fp16 conversions (round to nearest even).
__device__ __forceinline__ unsigned short q_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float dq_bf16(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }
__device__ __forceinline__ unsigned short q_f16(float x) {
  const _Float16 h = (_Float16)x;
  return __builtin_bit_cast(unsigned short, h);
}
__device__ __forceinline__ float dq_f16(unsigned short b) { return (float)__builtin_bit_cast(_Float16, b); }

// Strips whose max-abs is below this are stored as zero: Q / m would
// overflow for subnormal maxima (the precursor ahead of the wavefront).
constexpr float kSnapFloor = 1e-30f;

template <int SNAP>
struct SnapQ {
  static constexpr float Q = SNAP == SNAP_F16 ? 32768.f : 1.f;
  __device__ static unsigned short enc(float x) { return SNAP == SNAP_F16 ? q_f16(x) : q_bf16(x); }
  __device__ static float dec(unsigned short h) { return SNAP == SNAP_F16 ? dq_f16(h) : dq_bf16(h); }
};


// Global-memory accesses through address-space-1 pointers. The descriptor's
// pointers are generic; read through them the compiler emits FLAT
// instructions, which also count against the LDS wait counter, so waiting
// for an LDS read would wait for every global load in flight as well.
template <class T>
__device__ __forceinline__ T gld(const T* p, int i) {
  return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <class T>
__device__ __forceinline__ void gst(T* p, int i, T v) {
  ((__attribute__((address_space(1))) T*)p)[i] = v;
}
template <class T>
__device__ __forceinline__ void gst_nt(T* p, int i, T v) {
  __builtin_nontemporal_store(v, (__attribute__((address_space(1))) T*)p + i);
}
// Load of an element whose index is the same in every lane, of an array no
// kernel writes while it runs: through the constant address space it is
// one scalar load into a scalar register (a divergent index would still be
// served correctly, by vector loads).
template <class T>
__device__ __forceinline__ T gldu(const T* p, int i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}
// Masked load without a branch: lanes outside read element 0 (a valid
// address) and return zero.
template <class T>
__device__ __forceinline__ T gldm(const T* p, int i, bool in) {
  const T v = gld(p, in ? i : 0);
  return in ? v : T(0);
}

// ---------------------------------------------------------------------------
// Forward step, W/rtm-mod-adj.f90:225-252 (Born propagation :323-348):
//   Vx  = K(Vx(L-2)  + m1_x (S-x sxx + S+z sxz))            Ax, Bx
//   Vz  = K(Vz(L-2)  + m1_z (S+x sxz + S-z szz))            Az, Bz
//   sxx = K(sxx(L-2) + src + m2m3 S+x Vx + m2 S+z Vz)       D, E
//   szz = K(szz(L-2) + src + m2 S+x Vx + m2m3 S+z Vz)       F, G
//   sxz = K(sxz(L-2) + m3 (S-z Vx + S-x Vz))                H, J
// all differences of level L-1. Coefficients are zero outside each
// operator's range, which reproduces the reference loop bounds exactly.
// With F_SNAP the step also stores the snapshot of the imaging level L-1
// (Vx, Vz of L-1 from the staged tile and dV = V(L) - V(L-1)), or with
// SNAP_HIST the Vx, Vz of level L.
template <int TXR, int SNAP>
__global__ __launch_bounds__(256) void k_fwd(const ShotDev* __restrict__ shots,
                                             const TileRef* __restrict__ tiles, int ntiles, int L,
                                             int flags, int xcd) {
  using T = Tile<TXR>;
  constexpr int NIT = (T::N + 255) / 256;  // staged cells per thread
  __shared__ float sh[5][T::N];
  const TileRef tr = tiles[xcd ? xcd_tile(blockIdx.x, ntiles) : (int)blockIdx.x];
  const ShotDev& S = shots[tr.shot];
  const int nz = S.nz, nx = S.nx, ldz = S.ldz;
  // offsets inside one field are 32-bit (5 * n2 < 2^31, checked by the
  // hosts); field offsets are 64-bit: a drop-in history field is nt * n2
  const long long fs = S.fstride;
  const int n2 = (int)S.n2;
  float* const out = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L);
  const float* const x2 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L - 2);
  const float* const y1 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L - 1);
  const float* const coef = S.coef;
  const int z0 = tr.z0, x0 = tr.x0;
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int tid = tz + T::TZ * ty;

  // Every global load of this thread is issued before the first LDS store or
  // barrier: the staged tile of level L-1 (flat over tile + halo so all
  // lanes stay busy) and the own-cell operands (level L-2, coefficients).
  float st[NIT][5];
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int c = tid + 256 * k;
    const int lx = c / T::LZ, lz = c - lx * T::LZ;
    const int gx = x0 - T::H + lx, gz = z0 - T::H + lz;
    const bool in = c < T::N && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
    const int gi = gz + ldz * gx;
#pragma unroll
    for (int f = 0; f < 5; ++f) st[k][f] = gldm(y1 + f * fs, gi, in);
  }
  const int gz = z0 + tz;
  const bool zok = gz < nz;
  float ox[TXR][5], oc[TXR][5];
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int gx = x0 + ty * TXR + j;
    const bool ok = zok && gx < nx;
    const int gi = gz + ldz * gx;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      ox[j][f] = gld(x2 + f * fs, ok ? gi : 0);
      oc[j][f] = gld(coef + f * n2, ok ? gi : 0);
    }
  }
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int c = tid + 256 * k;
    if (c < T::N) {
#pragma unroll
      for (int f = 0; f < 5; ++f) sh[f][c] = st[k][f];
    }
  }
  __syncthreads();

  const float kz = ((flags & F_K) && zok) ? gld(S.kz, gz) : 1.f;
  constexpr int SX = T::LZ;
  float sv[4][TXR];
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    const bool ok = zok && gx < nx;
    const int p = (lxo + T::H) * T::LZ + (tz + T::H);
#pragma unroll
    for (int a = 0; a < 4; ++a) sv[a][j] = 0.f;
    if (!ok) continue;
    const int gi = gz + ldz * gx;
    const float m1x = oc[j][0], m1z = oc[j][1], m2 = oc[j][2], m3 = oc[j][3], m2m3 = oc[j][4];
    float v[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) v[f] = ox[j][f];
    if (flags & F_BORN) {
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = gld(S.bsrc + f * n2, gi) + v[f];
    }
    const float dxx_xm = sm<SX>(sh[2], p);  // Ax
    const float dxz_zp = sp<1>(sh[4], p);   // Bx
    const float dxz_xp = sp<SX>(sh[4], p);  // Az
    const float dzz_zm = sm<1>(sh[3], p);   // Bz
    const float dvx_xp = sp<SX>(sh[0], p);  // D, F
    const float dvz_zp = sp<1>(sh[1], p);   // E, G
    const float dvx_zm = sm<1>(sh[0], p);   // H
    const float dvz_xm = sm<SX>(sh[1], p);  // J
    v[0] = fmaf(m1x, dxx_xm, v[0]);
    v[0] = fmaf(m1x, dxz_zp, v[0]);
    v[1] = fmaf(m1z, dxz_xp, v[1]);
    v[1] = fmaf(m1z, dzz_zm, v[1]);
    if ((flags & F_SRC) && gz == S.zs && gx == S.xs) {
      const float amp = gld(S.src_amp, L - 1);
      v[2] = v[2] + amp;
      v[3] = v[3] + amp;
    }
    v[2] = fmaf(m2m3, dvx_xp, v[2]);
    v[2] = fmaf(m2, dvz_zp, v[2]);
    v[3] = fmaf(m2, dvx_xp, v[3]);
    v[3] = fmaf(m2m3, dvz_zp, v[3]);
    v[4] = fmaf(m3, dvx_zm, v[4]);
    v[4] = fmaf(m3, dvz_xm, v[4]);
    if (flags & F_K) {
      const float kx = gldu(S.kx, gx);
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = (v[f] * kz) * kx;
    }
    if (flags & F_NTS) {
#pragma unroll
      for (int f = 0; f < 5; ++f) gst_nt(out + f * fs, gi, v[f]);
    } else {
#pragma unroll
      for (int f = 0; f < 5; ++f) gst(out + f * fs, gi, v[f]);
    }
    if ((flags & (F_REC | F_RECSUB)) && gz == S.zr && gx >= S.rx0 && gx < S.rx0 + S.nrx) {
      float* r = S.rec + (long long)(L - 1) * S.nrx + (gx - S.rx0);
      const long long rs = (long long)S.nt * S.nrx;
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        if (flags & F_RECSUB)
          gst(r + f * rs, 0, gld(r + f * rs, 0) - v[f]);
        else
          gst(r + f * rs, 0, v[f]);
      }
    }
    if (SNAP == SNAP_HIST && (flags & F_SNAP)) {
      float* h = static_cast<float*>(S.snap) + (long long)(L - 1) * 2 * (long long)n2;
      gst(h, gi, v[0]);
      gst(h + n2, gi, v[1]);
    }
    if (SNAP != SNAP_HIST) {
      sv[0][j] = sh[0][p];
      sv[1][j] = sh[1][p];
      sv[2][j] = v[0] - sh[0][p];
      sv[3][j] = v[1] - sh[1][p];
    }
  }

  if (SNAP != SNAP_HIST && (flags & F_SNAP)) {
    // snapshot k of level L-1 = j*k (flag set by the host only then)
    const long long k0 = (L - 1) / S.snap_every - 1;
    if (SNAP == SNAP_F32) {
      float* sp4 = static_cast<float*>(S.snap) + k0 * 4 * n2;
#pragma unroll
      for (int j = 0; j < TXR; ++j) {
        const int gx = x0 + ty * TXR + j;
        if (!zok || gx >= nx) continue;
        const int gi = gz + ldz * gx;
#pragma unroll
        for (int a = 0; a < 4; ++a) gst(sp4 + a * n2, gi, sv[a][j]);
      }
    } else {
      unsigned short* q4 = static_cast<unsigned short*>(S.snap) + k0 * 4 * n2;
      const int strip = (z0 >> 6) + S.snap_nbz * ((x0 >> S.snap_sw) + ty);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        float m = 0.f;
#pragma unroll
        for (int j = 0; j < TXR; ++j) m = fmaxf(m, fabsf(sv[a][j]));
        m = wave_max(m);
        const float inv = m > kSnapFloor ? SnapQ<SNAP>::Q / m : 0.f;
        if (tz == 0) gst(S.snap_scale + (k0 * 4 + a) * S.snap_nstrip, strip, m / SnapQ<SNAP>::Q);
#pragma unroll
        for (int j = 0; j < TXR; ++j) {
          const int gx = x0 + ty * TXR + j;
          if (!zok || gx >= nx) continue;
          const int gi = gz + ldz * gx;
          gst(q4 + a * n2, gi, SnapQ<SNAP>::enc(sv[a][j] * inv));
        }
      }
    }
  }
}

// Store one decimated snapshot k0 of the thread's TXR cells: sv = Vx, Vz,
// dVx, dVz (fp32, or bf16/fp16 with one max-abs scale per 64 x TXR strip of
// each array, reduced over the wave).
template <int TXR, int SNAP>
__device__ __forceinline__ void store_snap(const ShotDev& S, long long k0, int z0, int x0, int tz, int ty,
                                           bool zok, int gz, int nx, int ldz, int n2, const float (&sv)[4][TXR]) {
  if (SNAP == SNAP_F32) {
    float* sp4 = static_cast<float*>(S.snap) + k0 * 4 * n2;
#pragma unroll
    for (int j = 0; j < TXR; ++j) {
      const int gx = x0 + ty * TXR + j;
      if (!zok || gx >= nx) continue;
      const int gi = gz + ldz * gx;
#pragma unroll
      for (int a = 0; a < 4; ++a) sp4[a * n2 + gi] = sv[a][j];
    }
  } else {
    unsigned short* q4 = static_cast<unsigned short*>(S.snap) + k0 * 4 * n2;
    const int strip = (z0 >> 6) + S.snap_nbz * ((x0 >> S.snap_sw) + ty);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      float m = 0.f;
#pragma unroll
      for (int j = 0; j < TXR; ++j) m = fmaxf(m, fabsf(sv[a][j]));
      m = wave_max(m);
      const float inv = m > kSnapFloor ? SnapQ<SNAP>::Q / m : 0.f;
      if (tz == 0) S.snap_scale[(k0 * 4 + a) * S.snap_nstrip + strip] = m / SnapQ<SNAP>::Q;
#pragma unroll
      for (int j = 0; j < TXR; ++j) {
        const int gx = x0 + ty * TXR + j;
        if (!zok || gx >= nx) continue;
        q4[a * n2 + gz + ldz * gx] = SnapQ<SNAP>::enc(sv[a][j] * inv);
      }
    }
  }
}

// Two forward levels per pass (temporal blocking): levels L and L+1 of a
// 64 x 4*TXR tile from levels L-1 and L-2. Level L-1 is staged in LDS with a
// radius-10 halo; level L is computed on the tile plus a radius-5 halo
// (cells outside the grid are zero, as a staged halo would be), kept in LDS
// in place of L-1, and level L+1 is computed on the tile from it. Each cell
// of L and L+1 gets exactly the arithmetic of k_fwd, so results are bitwise
// those of two single-level steps. Per two cell-updates HBM sees 5 fields of
// L-1, 5 of L-2, 5 coefficients and 10 written fields (100 B instead of
// 160 B); halo re-reads come from L2. The ring must hold four levels (L and
// L+1 are written to other slots than L-1 and L-2, which neighbouring tiles
// still read). Flags: F_K, F_SRC, F_REC / F_RECSUB, F_NTS apply to both
// levels; F_SNAP writes the snapshot of level L-1 (SNAP_HIST: the history
// of both levels), F_SNAP2 the snapshot of level L. No Born sources.
template <int TXR>
struct Tile2 {
  static constexpr int TZ = TILE_Z, TX = TILE_WY * TXR, H = 5;
  static constexpr int LZ1 = TZ + 4 * H, LX1 = TX + 4 * H, N1 = LZ1 * LX1;  // level L-1
  static constexpr int LZ0 = TZ + 2 * H, LX0 = TX + 2 * H, N0 = LZ0 * LX0;  // level L
};

template <int TXR, int SNAP>
__global__ __launch_bounds__(256) void k_fwd2(const ShotDev* __restrict__ shots,
                                              const TileRef* __restrict__ tiles, int ntiles, int L,
                                              int flags, int xcd) {
  using T = Tile2<TXR>;
  constexpr int NIT1 = (T::N1 + 255) / 256;
  constexpr int NIT0 = (T::N0 + 255) / 256;
  constexpr int H = T::H;
  __shared__ float sh[5][T::N1];
  const TileRef tr = tiles[xcd ? xcd_tile(blockIdx.x, ntiles) : (int)blockIdx.x];
  const ShotDev& S = shots[tr.shot];
  const int nz = S.nz, nx = S.nx, ldz = S.ldz;
  const long long fs = S.fstride;
  const int n2 = (int)S.n2;
  float* const out0 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L);
  float* const out1 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L + 1);
  const float* const x2 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L - 2);
  const float* const y1 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L - 1);
  const float* const coef = S.coef;
  const int z0 = tr.z0, x0 = tr.x0;
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int tid = tz + T::TZ * ty;

  // all global loads first: level L-1 with a radius-10 halo, then level L-2
  // and the coefficients of the radius-5 extended tile
  float st[NIT1][5];
#pragma unroll
  for (int k = 0; k < NIT1; ++k) {
    const int c = tid + 256 * k;
    const int lx = c / T::LZ1, lz = c - lx * T::LZ1;
    const int gx = x0 - 2 * H + lx, gz = z0 - 2 * H + lz;
    const bool in = c < T::N1 && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
    const int gi = gz + ldz * gx;
#pragma unroll
    for (int f = 0; f < 5; ++f) st[k][f] = in ? (y1 + f * fs)[gi] : 0.f;
  }
  float ex[NIT0][5], ec[NIT0][5];
#pragma unroll
  for (int k = 0; k < NIT0; ++k) {
    const int c = tid + 256 * k;
    const int lx = c / T::LZ0, lz = c - lx * T::LZ0;
    const int gx = x0 - H + lx, gz = z0 - H + lz;
    const bool in = c < T::N0 && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
    const int gi = gz + ldz * gx;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      ex[k][f] = in ? (x2 + f * fs)[gi] : 0.f;
      ec[k][f] = in ? coef[f * n2 + gi] : 0.f;
    }
  }
#pragma unroll
  for (int k = 0; k < NIT1; ++k) {
    const int c = tid + 256 * k;
    if (c < T::N1) {
#pragma unroll
      for (int f = 0; f < 5; ++f) sh[f][c] = st[k][f];
    }
  }
  __syncthreads();

  // level L on the extended tile (registers)
  constexpr int SX1 = T::LZ1;
  float vl[NIT0][5];
#pragma unroll
  for (int k = 0; k < NIT0; ++k) {
    const int c = tid + 256 * k;
    const int lx = c / T::LZ0, lz = c - lx * T::LZ0;
    const int gx = x0 - H + lx, gz = z0 - H + lz;
    const bool in = c < T::N0 && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
#pragma unroll
    for (int f = 0; f < 5; ++f) vl[k][f] = 0.f;
    if (!in) continue;
    const int p = (lx + H) * T::LZ1 + (lz + H);
    const float m1x = ec[k][0], m1z = ec[k][1], m2 = ec[k][2], m3 = ec[k][3], m2m3 = ec[k][4];
    float v[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) v[f] = ex[k][f];
    const float dxx_xm = sm<SX1>(sh[2], p);
    const float dxz_zp = sp<1>(sh[4], p);
    const float dxz_xp = sp<SX1>(sh[4], p);
    const float dzz_zm = sm<1>(sh[3], p);
    const float dvx_xp = sp<SX1>(sh[0], p);
    const float dvz_zp = sp<1>(sh[1], p);
    const float dvx_zm = sm<1>(sh[0], p);
    const float dvz_xm = sm<SX1>(sh[1], p);
    v[0] = fmaf(m1x, dxx_xm, v[0]);
    v[0] = fmaf(m1x, dxz_zp, v[0]);
    v[1] = fmaf(m1z, dxz_xp, v[1]);
    v[1] = fmaf(m1z, dzz_zm, v[1]);
    if ((flags & F_SRC) && gz == S.zs && gx == S.xs) {
      const float amp = S.src_amp[L - 1];
      v[2] = v[2] + amp;
      v[3] = v[3] + amp;
    }
    v[2] = fmaf(m2m3, dvx_xp, v[2]);
    v[2] = fmaf(m2, dvz_zp, v[2]);
    v[3] = fmaf(m2, dvx_xp, v[3]);
    v[3] = fmaf(m2m3, dvz_zp, v[3]);
    v[4] = fmaf(m3, dvx_zm, v[4]);
    v[4] = fmaf(m3, dvz_xm, v[4]);
    if (flags & F_K) {
      const float kz = S.kz[gz], kx = S.kx[gx];
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = (v[f] * kz) * kx;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) vl[k][f] = v[f];
  }
  // own cells of level L-1 (the L-2 operand of level L+1) before the tile
  // is overwritten
  const int gz = z0 + tz;
  const bool zok = gz < nz;
  float om[TXR][5];
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int p = (ty * TXR + j + 2 * H) * T::LZ1 + (tz + 2 * H);
#pragma unroll
    for (int f = 0; f < 5; ++f) om[j][f] = sh[f][p];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NIT0; ++k) {
    const int c = tid + 256 * k;
    if (c < T::N0) {
#pragma unroll
      for (int f = 0; f < 5; ++f) sh[f][c] = vl[k][f];
    }
  }
  __syncthreads();

  // level L+1 on the tile; level L of the tile written from LDS
  constexpr int SX0 = T::LZ0;
  const float kz = ((flags & F_K) && zok) ? S.kz[gz] : 1.f;
  float sa[4][TXR], sb[4][TXR];
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    const bool ok = zok && gx < nx;
    const int p = (lxo + H) * T::LZ0 + (tz + H);
#pragma unroll
    for (int a = 0; a < 4; ++a) sa[a][j] = sb[a][j] = 0.f;
    if (!ok) continue;
    const int gi = gz + ldz * gx;
    float cur[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) cur[f] = sh[f][p];
    const float m1x = coef[gi], m1z = coef[n2 + gi], m2 = coef[2 * n2 + gi], m3 = coef[3 * n2 + gi],
                m2m3 = coef[4 * n2 + gi];
    float v[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) v[f] = om[j][f];
    const float dxx_xm = sm<SX0>(sh[2], p);
    const float dxz_zp = sp<1>(sh[4], p);
    const float dxz_xp = sp<SX0>(sh[4], p);
    const float dzz_zm = sm<1>(sh[3], p);
    const float dvx_xp = sp<SX0>(sh[0], p);
    const float dvz_zp = sp<1>(sh[1], p);
    const float dvx_zm = sm<1>(sh[0], p);
    const float dvz_xm = sm<SX0>(sh[1], p);
    v[0] = fmaf(m1x, dxx_xm, v[0]);
    v[0] = fmaf(m1x, dxz_zp, v[0]);
    v[1] = fmaf(m1z, dxz_xp, v[1]);
    v[1] = fmaf(m1z, dzz_zm, v[1]);
    if ((flags & F_SRC) && gz == S.zs && gx == S.xs) {
      const float amp = S.src_amp[L];
      v[2] = v[2] + amp;
      v[3] = v[3] + amp;
    }
    v[2] = fmaf(m2m3, dvx_xp, v[2]);
    v[2] = fmaf(m2, dvz_zp, v[2]);
    v[3] = fmaf(m2, dvx_xp, v[3]);
    v[3] = fmaf(m2m3, dvz_zp, v[3]);
    v[4] = fmaf(m3, dvx_zm, v[4]);
    v[4] = fmaf(m3, dvz_xm, v[4]);
    if (flags & F_K) {
      const float kx = S.kx[gx];
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = (v[f] * kz) * kx;
    }
    if (flags & F_NTS) {
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        __builtin_nontemporal_store(cur[f], out0 + f * fs + gi);
        __builtin_nontemporal_store(v[f], out1 + f * fs + gi);
      }
    } else {
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        (out0 + f * fs)[gi] = cur[f];
        (out1 + f * fs)[gi] = v[f];
      }
    }
    if ((flags & (F_REC | F_RECSUB)) && gz == S.zr && gx >= S.rx0 && gx < S.rx0 + S.nrx) {
      const long long rs = (long long)S.nt * S.nrx;
      float* r0 = S.rec + (long long)(L - 1) * S.nrx + (gx - S.rx0);
      float* r1 = r0 + S.nrx;
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        if (flags & F_RECSUB) {
          r0[f * rs] = r0[f * rs] - cur[f];
          r1[f * rs] = r1[f * rs] - v[f];
        } else {
          r0[f * rs] = cur[f];
          r1[f * rs] = v[f];
        }
      }
    }
    if (SNAP == SNAP_HIST && (flags & F_SNAP)) {
      float* h = static_cast<float*>(S.snap) + (long long)(L - 1) * 2 * (long long)n2;
      h[gi] = cur[0];
      h[n2 + gi] = cur[1];
      h[2 * n2 + gi] = v[0];
      h[3 * n2 + gi] = v[1];
    }
    if (SNAP != SNAP_HIST) {
      sa[0][j] = om[j][0];
      sa[1][j] = om[j][1];
      sa[2][j] = cur[0] - om[j][0];
      sa[3][j] = cur[1] - om[j][1];
      sb[0][j] = cur[0];
      sb[1][j] = cur[1];
      sb[2][j] = v[0] - cur[0];
      sb[3][j] = v[1] - cur[1];
    }
  }
  if (SNAP != SNAP_HIST) {
    // snapshot of level L-1 (flag set by the host when (L-1) % j == 0) and of
    // level L (F_SNAP2, when L % j == 0)
    if (flags & F_SNAP) store_snap<TXR, SNAP>(S, (L - 1) / S.snap_every - 1, z0, x0, tz, ty, zok, gz, nx, ldz, n2, sa);
    if (flags & F_SNAP2) store_snap<TXR, SNAP>(S, L / S.snap_every - 1, z0, x0, tz, ty, zok, gz, nx, ldz, n2, sb);
  }
}

// Snapshot of the imaging level L that no later forward step writes (the
// last level): Vx, Vz from the ring, dV = 0.
template <int TXR, int SNAP>
__global__ __launch_bounds__(256) void k_snap_last(const ShotDev* __restrict__ shots,
                                                   const TileRef* __restrict__ tiles, int L) {
  const TileRef tr = tiles[blockIdx.x];
  const ShotDev& S = shots[tr.shot];
  const int nz = S.nz, nx = S.nx, ldz = S.ldz;
  const long long fs = S.fstride, n2 = S.n2;
  const float* const cur = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L);
  const int z0 = tr.z0, x0 = tr.x0, tz = threadIdx.x, ty = threadIdx.y;
  const int gz = z0 + tz;
  const bool zok = gz < nz;
  const long long k0 = L / S.snap_every - 1;
  float sv[2][TXR];
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int gx = x0 + ty * TXR + j;
    const bool ok = zok && gx < nx;
    const long long gi = (long long)gz + (long long)ldz * gx;
    sv[0][j] = ok ? cur[gi] : 0.f;
    sv[1][j] = ok ? cur[fs + gi] : 0.f;
  }
  if (SNAP == SNAP_F32) {
    float* sp4 = static_cast<float*>(S.snap) + k0 * 4 * n2;
#pragma unroll
    for (int j = 0; j < TXR; ++j) {
      const int gx = x0 + ty * TXR + j;
      if (!zok || gx >= nx) continue;
      const long long gi = (long long)gz + (long long)ldz * gx;
      sp4[gi] = sv[0][j];
      sp4[n2 + gi] = sv[1][j];
      sp4[2 * n2 + gi] = 0.f;
      sp4[3 * n2 + gi] = 0.f;
    }
  } else {
    unsigned short* q4 = static_cast<unsigned short*>(S.snap) + k0 * 4 * n2;
    const int strip = (z0 >> 6) + S.snap_nbz * ((x0 >> S.snap_sw) + ty);
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      float m = 0.f;
#pragma unroll
      for (int j = 0; j < TXR; ++j) m = fmaxf(m, fabsf(sv[a][j]));
      m = wave_max(m);
      const float inv = m > kSnapFloor ? SnapQ<SNAP>::Q / m : 0.f;
      if (tz == 0) {
        S.snap_scale[(k0 * 4 + a) * S.snap_nstrip + strip] = m / SnapQ<SNAP>::Q;
        S.snap_scale[(k0 * 4 + a + 2) * S.snap_nstrip + strip] = 0.f;
      }
#pragma unroll
      for (int j = 0; j < TXR; ++j) {
        const int gx = x0 + ty * TXR + j;
        if (!zok || gx >= nx) continue;
        const long long gi = (long long)gz + (long long)ldz * gx;
        q4[a * n2 + gi] = SnapQ<SNAP>::enc(sv[a][j] * inv);
        q4[(a + 2) * n2 + gi] = 0;
      }
    }
  }
}

// Source-wavefield view of one imaging level L: per-level base pointers are
// resolved once per block; elements are then addressed with 32-bit offsets.
template <int SNAP>
struct SrcView {
  const float* f[4];           // fp32 arrays (Vx, Vz, dVx, dVz; HIST: Vx, Vz, Vx(L+1), Vz(L+1))
  const unsigned short* q[4];  // bf16 / fp16 arrays
  const float* sc[4];          // per-strip scales
  int nbz, sw;
  __device__ SrcView(const ShotDev& S, int L) {
    const long long n2 = S.n2;
    nbz = S.snap_nbz;
    sw = S.snap_sw;
    if (SNAP == SNAP_HIST) {
      // arrays 2, 3 (level L+1) exist below the last level only; at the last
      // level they alias arrays 0, 1 so that every pointer stays inside the
      // history (element 0 of each array is read by masked loads)
      const float* h = static_cast<const float*>(S.snap) + (long long)(L - 1) * 2 * n2;
      const bool nxt = L < S.nt;
      f[0] = h; f[1] = h + n2; f[2] = nxt ? h + 2 * n2 : h; f[3] = nxt ? h + 3 * n2 : h + n2;
    } else {
      const long long k0 = L / S.snap_every - 1;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        if (SNAP == SNAP_F32) {
          f[a] = static_cast<const float*>(S.snap) + (k0 * 4 + a) * n2;
        } else {
          q[a] = static_cast<const unsigned short*>(S.snap) + (k0 * 4 + a) * n2;
          sc[a] = S.snap_scale + (k0 * 4 + a) * S.snap_nstrip;
        }
      }
    }
  }
  // array a at element gi = gz + ldz*gx
  __device__ __forceinline__ float get(int a, int gi, int gz, int gx) const {
    if (SNAP == SNAP_HIST || SNAP == SNAP_F32) return f[a][gi];
    const int strip = (gz >> 6) + nbz * (gx >> sw);
    return SnapQ<SNAP>::dec(q[a][gi]) * sc[a][strip];
  }
  // an element as loaded (clamped index), decoded and masked by value():
  // loads and their use can be a column apart
  struct Raw {
    float v;
    unsigned short h;
    bool in;
  };
  __device__ __forceinline__ void load(int a, int gi, int gz, int gx, bool in, Raw& r) const {
    r.in = in;
    if (SNAP == SNAP_HIST || SNAP == SNAP_F32) {
      r.v = gld(f[a], in ? gi : 0);
    } else {
      r.h = gld(q[a], in ? gi : 0);
      r.v = gld(sc[a], in ? (gz >> 6) + nbz * (gx >> sw) : 0);
    }
  }
  __device__ __forceinline__ float value(const Raw& r) const {
    if (SNAP == SNAP_HIST || SNAP == SNAP_F32) return r.in ? r.v : 0.f;
    return r.in ? SnapQ<SNAP>::dec(r.h) * r.v : 0.f;
  }
  // The element alone (16-bit formats: the scale comes from scale_u).
  __device__ __forceinline__ void load_q(int a, int gi, bool in, Raw& r) const {
    r.in = in;
    if (SNAP == SNAP_HIST || SNAP == SNAP_F32)
      r.v = gld(f[a], in ? gi : 0);
    else
      r.h = gld(q[a], in ? gi : 0);
  }
  // 16-bit formats: the scale of the strip of 64-row block zb and column gx,
  // read through the constant address space, so that a wave-uniform block and
  // column make it one scalar load into a scalar register.
  __device__ __forceinline__ float scale_u(int a, int zb, int gx) const {
    if (SNAP == SNAP_HIST || SNAP == SNAP_F32) return 1.f;
    return ((const __attribute__((address_space(4))) float*)sc[a])[zb + nbz * (gx >> sw)];
  }
  __device__ __forceinline__ float value_s(const Raw& r, float s) const {
    if (SNAP == SNAP_HIST || SNAP == SNAP_F32) return r.in ? r.v : 0.f;
    return r.in ? SnapQ<SNAP>::dec(r.h) * s : 0.f;
  }
  // the same as a branch-free global load, zero where !in
  __device__ __forceinline__ float getm(int a, int gi, int gz, int gx, bool in) const {
    if (SNAP == SNAP_HIST || SNAP == SNAP_F32) return gldm(f[a], gi, in);
    const int strip = in ? (gz >> 6) + nbz * (gx >> sw) : 0;
    const float v = SnapQ<SNAP>::dec(gld(q[a], in ? gi : 0)) * gld(sc[a], strip);
    return in ? v : 0.f;
  }
};

// ---------------------------------------------------------------------------
// Adjoint step (gather form of rtm_op3_c_, W/rtmlib-modC.c:120-253, with the
// receiver injection of W/rtmlib-mod.f90:99-146):
//   Vx  = K'(b + - S-x(m2m3 sxx + m2 szz) - S+z(m3 sxz))
//   Vz  = K'(b + - S-z(m2 sxx + m2m3 szz) - S+x(m3 sxz))
//   sxx = K'(b - S+x(m1_x Vx))
//   szz = K'(b - S+z(m1_z Vz))
//   sxz = K'(b - S-z(m1_x Vx) - S-x(m1_z Vz))
// products of level L+1, b = (injected row or the level's own content) +
// level L+2. K' is the sponge (exact adjoint) or identity (the drop-in).
// With A_IMG the Lame imaging condition of level L (W/image-mod-rtm.f90:
// 154-162 with S1..S6 of :26-67) is accumulated in the same pass.
template <int TXR, int SNAP>
__global__ __launch_bounds__(256) void k_adj(const ShotDev* __restrict__ shots,
                                             const TileRef* __restrict__ tiles, int ntiles, int L,
                                             int flags, int xcd) {
  using T = Tile<TXR>;
  constexpr int NIT = (T::N + 255) / 256;
  __shared__ float sh[5][T::N];
  const TileRef tr = tiles[xcd ? xcd_tile(blockIdx.x, ntiles) : (int)blockIdx.x];
  const ShotDev& S = shots[tr.shot];
  const int nz = S.nz, nx = S.nx, ldz = S.ldz;
  const long long fs = S.fstride;
  const int n2 = (int)S.n2;
  float* const out = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L);
  const float* const x2 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L + 2);
  const float* const y1 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L + 1);
  const float* const coef = S.coef;
  const int z0 = tr.z0, x0 = tr.x0;
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int tid = tz + T::TZ * ty;

  // Staged tile: coefficient-weighted products of level L+1 (flat over tile +
  // halo); own-cell operands are loaded before the barrier as well.
  float st[NIT][5];
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int c = tid + 256 * k;
    const int lx = c / T::LZ, lz = c - lx * T::LZ;
    const int gx = x0 - T::H + lx, gz = z0 - T::H + lz;
    const bool in = c < T::N && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
    // With the fp32 history out-of-range cells issue no loads; with
    // decimated snapshots the loads are branch-free (clamped index, masked
    // products). Each form measured faster for its format.
    float p0 = 0.f, p1 = 0.f, p2 = 0.f, p3 = 0.f, p4 = 0.f;
    if (SNAP != SNAP_HIST || in) {
      const int gi = in ? gz + ldz * gx : 0;
      const float vx = gld(y1, gi), vz = gld(y1 + fs, gi), xx = gld(y1 + 2 * fs, gi), zz = gld(y1 + 3 * fs, gi),
                  xz = gld(y1 + 4 * fs, gi);
      const float m1x = gld(coef, gi), m1z = gld(coef + n2, gi), m2 = gld(coef + 2 * n2, gi),
                  m3 = gld(coef + 3 * n2, gi), m2m3 = gld(coef + 4 * n2, gi);
      p0 = in ? fmaf(m2m3, xx, m2 * zz) : 0.f;
      p1 = in ? fmaf(m2, xx, m2m3 * zz) : 0.f;
      p2 = in ? m3 * xz : 0.f;
      p3 = in ? m1x * vx : 0.f;
      p4 = in ? m1z * vz : 0.f;
    }
    st[k][0] = p0;
    st[k][1] = p1;
    st[k][2] = p2;
    st[k][3] = p3;
    st[k][4] = p4;
  }
  const int gz = z0 + tz;
  const bool zok = gz < nz;
  const bool row = (flags & A_INJ) && gz == S.zr;
  float b[TXR][5];
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int gx = x0 + ty * TXR + j;
    const bool ok = zok && gx < nx;
    const int gi = gz + ldz * gx;
#pragma unroll
    for (int f = 0; f < 5; ++f) b[j][f] = gld(x2 + f * fs, ok ? gi : 0);  // used where ok only
    if (flags & A_XIN) {
#pragma unroll
      for (int f = 0; f < 5; ++f) b[j][f] = gld(out + f * fs, ok ? gi : 0) + b[j][f];
    } else if (!(flags & A_KSRC) && ok && row && gx >= S.rx0 && gx < S.rx0 + S.nrx) {
      const float* r = S.rec + (long long)(L - 1) * S.nrx + (gx - S.rx0);
      const long long rs = (long long)S.nt * S.nrx;
#pragma unroll
      for (int f = 0; f < 5; ++f) b[j][f] = gld(r + f * rs, 0) + b[j][f];
    }
  }
  // Imaging inputs: the source Vx, Vz of level L with their halo, the
  // own-cell S1/S2 operands and the image (and compensation) values to be
  // updated. For fp32 snapshots they are loaded here, before the barrier,
  // with the step's own operands (measured faster); bf16 / fp16 snapshots
  // need their strip scales too and are loaded after the step (fewer live
  // registers, higher occupancy, measured faster for them).
  const bool imaging = flags & A_IMG;  // uniform over the launch
  constexpr bool kEarly = SNAP == SNAP_HIST || SNAP == SNAP_F32;
  const SrcView<SNAP> sv(S, L);
  float sx[NIT][2], nt2[TXR][2], im[TXR][3], imc[TXR][3];
  auto load_imaging = [&]() {
#pragma unroll
    for (int k = 0; k < NIT; ++k) {
      const int c = tid + 256 * k;
      const int lx = c / T::LZ, lz = c - lx * T::LZ;
      const int gx = x0 - T::H + lx, gz2 = z0 - T::H + lz;
      const bool in = c < T::N && gx >= 0 && gx < nx && gz2 >= 0 && gz2 < nz;
      const int gi = gz2 + ldz * gx;
      sx[k][0] = sv.getm(0, gi, gz2, gx, in);
      sx[k][1] = sv.getm(1, gi, gz2, gx, in);
    }
    const bool next = flags & A_NEXT;
#pragma unroll
    for (int j = 0; j < TXR; ++j) {
      const int gx = x0 + ty * TXR + j;
      const bool ok = zok && gx < nx;
      const int gi = gz + ldz * gx;
      nt2[j][0] = next ? sv.getm(2, gi, gz, gx, ok) : 0.f;  // arrays 2, 3 exist below the last level only
      nt2[j][1] = next ? sv.getm(3, gi, gz, gx, ok) : 0.f;
#pragma unroll
      for (int a = 0; a < 3; ++a) {  // used where ok only
        im[j][a] = gld(S.img + a * n2, ok ? gi : 0);
        imc[j][a] = (flags & A_COMP) ? gld(S.imgc + a * n2, ok ? gi : 0) : 0.f;
      }
    }
  };
  if (kEarly && imaging) load_imaging();
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int c = tid + 256 * k;
    if (c < T::N) {
#pragma unroll
      for (int f = 0; f < 5; ++f) sh[f][c] = st[k][f];
    }
  }
  __syncthreads();

  const float kz = ((flags & A_K) && zok) ? gld(S.kz, gz) : 1.f;
  float o[TXR][5];
  constexpr int SX = T::LZ;
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    const bool ok = zok && gx < nx;
    const int p = (lxo + T::H) * T::LZ + (tz + T::H);
#pragma unroll
    for (int f = 0; f < 5; ++f) o[j][f] = 0.f;
    if (!ok) continue;
    const int gi = gz + ldz * gx;
    if (flags & A_KSRC) {
      // R/Src sponge: the injection is added to the damped level L+2
      const float kxx = gldu(S.kx, gx);
      const bool inj = row && gx >= S.rx0 && gx < S.rx0 + S.nrx;
      const float* r = S.rec + (long long)(L - 1) * S.nrx + (inj ? gx - S.rx0 : 0);
      const long long rs = (long long)S.nt * S.nrx;
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        b[j][f] = (b[j][f] * kz) * kxx;
        if (inj) b[j][f] = gld(r + f * rs, 0) + b[j][f];
      }
    }
    float v0 = b[j][0] - sm<SX>(sh[0], p);
    v0 -= sp<1>(sh[2], p);
    float v1 = b[j][1] - sm<1>(sh[1], p);
    v1 -= sp<SX>(sh[2], p);
    const float v2 = b[j][2] - sp<SX>(sh[3], p);
    const float v3 = b[j][3] - sp<1>(sh[4], p);
    float v4 = b[j][4] - sm<1>(sh[3], p);
    v4 -= sm<SX>(sh[4], p);
    float v[5] = {v0, v1, v2, v3, v4};
    // with the R/Src sponge, level 1 (read by no later step) keeps its
    // undamped value, which is also the one imaged
    const bool ksrc = flags & A_KSRC;
    float kx = 1.f;
    if ((flags & A_K) && (!ksrc || L >= 2)) {
      kx = gldu(S.kx, gx);
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = (v[f] * kz) * kx;
    }
    if (flags & A_NTS) {
#pragma unroll
      for (int f = 0; f < 5; ++f) gst_nt(out + f * fs, gi, v[f]);
    } else {
#pragma unroll
      for (int f = 0; f < 5; ++f) gst(out + f * fs, gi, v[f]);
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) o[j][f] = ksrc && L >= 3 ? (v[f] * kz) * kx : v[f];
  }

  if (!imaging) return;
  if (!kEarly) load_imaging();
  // the source tile replaces the products in LDS
  __syncthreads();
  float* const ss[2] = {sh[0], sh[1]};
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int c = tid + 256 * k;
    if (c < T::N) {
      ss[0][c] = sx[k][0];
      ss[1][c] = sx[k][1];
    }
  }
  __syncthreads();
  const bool next = flags & A_NEXT;
  const bool zin4 = gz >= 4 && gz <= nz - 6, zin5 = gz >= 5 && gz <= nz - 6;
  float* const img = S.img;
  float* const imgc = S.imgc;
  const float w = S.img_w, inv_dx = S.inv_dx, inv_dt = S.inv_dt;
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    // the imaging arithmetic is fused only where written (fmaf), so every
    // kernel variant rounds it the same way
#pragma clang fp contract(off)
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    if (!zok || gx >= nx) continue;
    const int p = (lxo + T::H) * T::LZ + (tz + T::H);
    const int gi = gz + ldz * gx;
    // S1, S2: forward time difference (T, W/image-mod-rtm.f90:288-297)
    float s1 = 0.f, s2 = 0.f;
    if (next) {
      if (SNAP == SNAP_HIST) {  // arrays 2, 3 are Vx, Vz of level L+1
        s1 = inv_dt * (nt2[j][0] - ss[0][p]);
        s2 = inv_dt * (nt2[j][1] - ss[1][p]);
      } else {
        s1 = inv_dt * nt2[j][0];
        s2 = inv_dt * nt2[j][1];
      }
    }
    // S4 = vxdx, S5 = vzdz (range 4..n-6), S6 = vxdz + vzdx (range 5..n-6)
    float s4 = 0.f, s5 = 0.f, s6 = 0.f;
    if (zin4 && gx >= 4 && gx <= nx - 6) {
      s4 = sp<SX>(ss[0], p) * inv_dx;
      s5 = sp<1>(ss[1], p) * inv_dx;
    }
    if (zin5 && gx >= 5 && gx <= nx - 6) {
      s6 = sm<1>(ss[0], p) * inv_dx;
      s6 = fmaf(sm<SX>(ss[1], p), inv_dx, s6);
    }
    const float s3 = s4 + s5;
    const float* v = o[j];
    const float t[3] = {w * fmaf(v[0], s1, v[1] * s2), -w * fmaf(v[2], s3, v[3] * s3),
                        -w * fmaf(v[4], s6, fmaf(v[3] * 2.f, s5, v[2] * 2.f * s4))};
    if (flags & A_COMP) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        float sa = im[j][a], ca = imc[j][a];
        neumaier(sa, ca, t[a]);
        gst(img + a * n2, gi, sa);
        gst(imgc + a * n2, gi, ca);
      }
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) gst(img + a * n2, gi, im[j][a] + t[a]);
    }
  }
}

// Imaging condition of one adjoint level at the thread's TXR cells: o holds
// the adjoint fields of the level, the source Vx, Vz of the level are staged
// in s0 / s1 (LDS, stride LZ, radius-5 halo).
template <int TXR, int SNAP, int LZ>
__device__ __forceinline__ void image_cells(const ShotDev& S, const SrcView<SNAP>& sv, const float* s0, const float* s1,
                                            const float (&o)[TXR][5], bool next, bool comp, int z0, int x0, int tz,
                                            int ty, int nz, int nx, int ldz, int n2) {
  constexpr int H = 5;
  const int gz = z0 + tz;
  if (gz >= nz) return;
  const bool zin4 = gz >= 4 && gz <= nz - 6, zin5 = gz >= 5 && gz <= nz - 6;
  float* const img = S.img;
  float* const imgc = S.imgc;
  const float w = S.img_w, inv_dx = S.inv_dx, inv_dt = S.inv_dt;
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
#pragma clang fp contract(off)
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    if (gx >= nx) continue;
    const int p = (lxo + H) * LZ + (tz + H);
    const int gi = gz + ldz * gx;
    float s1v = 0.f, s2v = 0.f;
    if (next) {
      if (SNAP == SNAP_HIST) {
        s1v = inv_dt * (sv.get(2, gi, gz, gx) - s0[p]);
        s2v = inv_dt * (sv.get(3, gi, gz, gx) - s1[p]);
      } else {
        s1v = inv_dt * sv.get(2, gi, gz, gx);
        s2v = inv_dt * sv.get(3, gi, gz, gx);
      }
    }
    float s4 = 0.f, s5 = 0.f, s6 = 0.f;
    if (zin4 && gx >= 4 && gx <= nx - 6) {
      s4 = sp<LZ>(s0, p) * inv_dx;
      s5 = sp<1>(s1, p) * inv_dx;
    }
    if (zin5 && gx >= 5 && gx <= nx - 6) {
      s6 = sm<1>(s0, p) * inv_dx;
      s6 = fmaf(sm<LZ>(s1, p), inv_dx, s6);
    }
    const float s3 = s4 + s5;
    const float* v = o[j];
    const float t1 = w * fmaf(v[0], s1v, v[1] * s2v);
    const float t2 = -w * fmaf(v[2], s3, v[3] * s3);
    const float t3 = -w * fmaf(v[4], s6, fmaf(v[3] * 2.f, s5, v[2] * 2.f * s4));
    if (comp) {
      float s = img[gi], c = imgc[gi];
      neumaier(s, c, t1);
      img[gi] = s;
      imgc[gi] = c;
      s = img[n2 + gi];
      c = imgc[n2 + gi];
      neumaier(s, c, t2);
      img[n2 + gi] = s;
      imgc[n2 + gi] = c;
      s = img[2 * n2 + gi];
      c = imgc[2 * n2 + gi];
      neumaier(s, c, t3);
      img[2 * n2 + gi] = s;
      imgc[2 * n2 + gi] = c;
    } else {
      img[gi] += t1;
      img[n2 + gi] += t2;
      img[2 * n2 + gi] += t3;
    }
  }
}

// Stage the source Vx, Vz of imaging level L (radius-5 halo) into s0 / s1.
template <int TXR, int SNAP>
__device__ __forceinline__ void stage_source(const SrcView<SNAP>& sv, float* s0, float* s1, int z0, int x0, int tid,
                                             int nz, int nx, int ldz) {
  using T = Tile<TXR>;
  constexpr int NIT = (T::N + 255) / 256;
  float sx[NIT][2];
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int c = tid + 256 * k;
    const int lx = c / T::LZ, lz = c - lx * T::LZ;
    const int gx = x0 - T::H + lx, gz2 = z0 - T::H + lz;
    const bool in = c < T::N && gx >= 0 && gx < nx && gz2 >= 0 && gz2 < nz;
    const int gi = gz2 + ldz * gx;
    sx[k][0] = in ? sv.get(0, gi, gz2, gx) : 0.f;
    sx[k][1] = in ? sv.get(1, gi, gz2, gx) : 0.f;
  }
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int c = tid + 256 * k;
    if (c < T::N) {
      s0[c] = sx[k][0];
      s1[c] = sx[k][1];
    }
  }
}

// Two adjoint levels per pass (temporal blocking): levels L and L-1 from
// levels L+1 and L+2, with the arithmetic of k_adj per cell (bitwise the
// same results). Coefficient-weighted products of level L+1 are staged with
// a radius-10 halo; level L is computed on the tile plus a radius-5 halo
// (receiver injection and sponge included, zero outside the grid), its
// products replace those of L+1 in LDS, and level L-1 is computed on the
// tile. The imaging condition of level L (A_IMG) and of level L-1 (A_IMG2)
// is accumulated in the same pass, in that order. The ring must hold four
// levels. Flags as k_adj (A_NEXT refers to level L; level L-1 always has a
// next level); no drop-in history form (A_XIN).
template <int TXR>
struct Adj2Lds {
  using T = Tile2<TXR>;
  static constexpr int P1 = 5 * T::N1;                 // products of L+1
  static constexpr int P0 = 5 * T::N0 + 5 * T::TZ * T::TX;  // products of L + raw L of the tile
  static constexpr int IMG = 2 * Tile<TXR>::N;          // staged source Vx, Vz
  static constexpr int A = P1 > P0 ? P1 : P0;
  static constexpr int SIZE = A > IMG ? A : IMG;
};

template <int TXR, int SNAP>
__global__ __launch_bounds__(256) void k_adj2(const ShotDev* __restrict__ shots,
                                              const TileRef* __restrict__ tiles, int ntiles, int L,
                                              int flags, int xcd) {
  using T = Tile2<TXR>;
  constexpr int NIT1 = (T::N1 + 255) / 256;
  constexpr int NIT0 = (T::N0 + 255) / 256;
  constexpr int H = T::H;
  constexpr int NT = T::TZ * T::TX;
  __shared__ float shm[Adj2Lds<TXR>::SIZE];
  const TileRef tr = tiles[xcd ? xcd_tile(blockIdx.x, ntiles) : (int)blockIdx.x];
  const ShotDev& S = shots[tr.shot];
  const int nz = S.nz, nx = S.nx, ldz = S.ldz;
  const long long fs = S.fstride;
  const int n2 = (int)S.n2;
  float* const out0 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L);
  float* const out1 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L - 1);
  const float* const x2 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L + 2);
  const float* const y1 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L + 1);
  const float* const coef = S.coef;
  const int z0 = tr.z0, x0 = tr.x0;
  const int tz = threadIdx.x, ty = threadIdx.y;
  const int tid = tz + T::TZ * ty;
  const bool inj = flags & A_INJ;
  const long long rs = (long long)S.nt * S.nrx;

  // products of level L+1 over the radius-10 tile
  float st[NIT1][5];
#pragma unroll
  for (int k = 0; k < NIT1; ++k) {
    const int c = tid + 256 * k;
    const int lx = c / T::LZ1, lz = c - lx * T::LZ1;
    const int gx = x0 - 2 * H + lx, gz = z0 - 2 * H + lz;
    const bool in = c < T::N1 && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
    float p0 = 0.f, p1 = 0.f, p2 = 0.f, p3 = 0.f, p4 = 0.f;
    if (in) {
      const int gi = gz + ldz * gx;
      const float vx = y1[gi], vz = (y1 + fs)[gi], xx = (y1 + 2 * fs)[gi], zz = (y1 + 3 * fs)[gi],
                  xz = (y1 + 4 * fs)[gi];
      const float m1x = coef[gi], m1z = coef[n2 + gi], m2 = coef[2 * n2 + gi],
                  m3 = coef[3 * n2 + gi], m2m3 = coef[4 * n2 + gi];
      p0 = in ? fmaf(m2m3, xx, m2 * zz) : 0.f;
      p1 = in ? fmaf(m2, xx, m2m3 * zz) : 0.f;
      p2 = in ? m3 * xz : 0.f;
      p3 = in ? m1x * vx : 0.f;
      p4 = in ? m1z * vz : 0.f;
    }
    st[k][0] = p0;
    st[k][1] = p1;
    st[k][2] = p2;
    st[k][3] = p3;
    st[k][4] = p4;
  }
  // level L+2 (+ injection of level L) and the coefficients on the radius-5 tile
  float eb[NIT0][5], ec[NIT0][5];
#pragma unroll
  for (int k = 0; k < NIT0; ++k) {
    const int c = tid + 256 * k;
    const int lx = c / T::LZ0, lz = c - lx * T::LZ0;
    const int gx = x0 - H + lx, gz = z0 - H + lz;
    const bool in = c < T::N0 && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
    const int gi = gz + ldz * gx;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      eb[k][f] = in ? (x2 + f * fs)[gi] : 0.f;
      ec[k][f] = in ? coef[f * n2 + gi] : 0.f;
    }
    if (in && inj && gz == S.zr && gx >= S.rx0 && gx < S.rx0 + S.nrx) {
      const float* r = S.rec + (long long)(L - 1) * S.nrx + (gx - S.rx0);
#pragma unroll
      for (int f = 0; f < 5; ++f) eb[k][f] = r[f * rs] + eb[k][f];
    }
  }
  // level L+1 (+ injection of level L-1) at the tile: b of level L-1
  const int gz = z0 + tz;
  const bool zok = gz < nz;
  float b1[TXR][5];
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int gx = x0 + ty * TXR + j;
    const bool ok = zok && gx < nx;
    const int gi = gz + ldz * gx;
#pragma unroll
    for (int f = 0; f < 5; ++f) b1[j][f] = ok ? (y1 + f * fs)[gi] : 0.f;
    if (ok && inj && gz == S.zr && gx >= S.rx0 && gx < S.rx0 + S.nrx) {
      const float* r = S.rec + (long long)(L - 2) * S.nrx + (gx - S.rx0);
#pragma unroll
      for (int f = 0; f < 5; ++f) b1[j][f] = r[f * rs] + b1[j][f];
    }
  }
  float* const sh1[5] = {shm, shm + T::N1, shm + 2 * T::N1, shm + 3 * T::N1, shm + 4 * T::N1};
#pragma unroll
  for (int k = 0; k < NIT1; ++k) {
    const int c = tid + 256 * k;
    if (c < T::N1) {
#pragma unroll
      for (int f = 0; f < 5; ++f) sh1[f][c] = st[k][f];
    }
  }
  __syncthreads();

  // level L on the radius-5 tile
  constexpr int SX1 = T::LZ1;
  float vl[NIT0][5];
#pragma unroll
  for (int k = 0; k < NIT0; ++k) {
    const int c = tid + 256 * k;
    const int lx = c / T::LZ0, lz = c - lx * T::LZ0;
    const int gx = x0 - H + lx, gzz = z0 - H + lz;
    const bool in = c < T::N0 && gx >= 0 && gx < nx && gzz >= 0 && gzz < nz;
#pragma unroll
    for (int f = 0; f < 5; ++f) vl[k][f] = 0.f;
    if (!in) continue;
    const int p = (lx + H) * T::LZ1 + (lz + H);
    float v0 = eb[k][0] - sm<SX1>(sh1[0], p);
    v0 -= sp<1>(sh1[2], p);
    float v1 = eb[k][1] - sm<1>(sh1[1], p);
    v1 -= sp<SX1>(sh1[2], p);
    const float v2 = eb[k][2] - sp<SX1>(sh1[3], p);
    const float v3 = eb[k][3] - sp<1>(sh1[4], p);
    float v4 = eb[k][4] - sm<1>(sh1[3], p);
    v4 -= sm<SX1>(sh1[4], p);
    float v[5] = {v0, v1, v2, v3, v4};
    if (flags & A_K) {
      const float kz = S.kz[gzz], kx = S.kx[gx];
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = (v[f] * kz) * kx;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) vl[k][f] = v[f];
  }
  __syncthreads();
  // products of level L (radius-5 tile) and the raw level L of the tile
  float* const sh0[5] = {shm, shm + T::N0, shm + 2 * T::N0, shm + 3 * T::N0, shm + 4 * T::N0};
  float* const raw = shm + 5 * T::N0;  // [f][NT], tile cell (lxo, tz) at lxo*TZ + tz
#pragma unroll
  for (int k = 0; k < NIT0; ++k) {
    const int c = tid + 256 * k;
    if (c < T::N0) {
      const float m1x = ec[k][0], m1z = ec[k][1], m2 = ec[k][2], m3 = ec[k][3], m2m3 = ec[k][4];
      const float vx = vl[k][0], vz = vl[k][1], xx = vl[k][2], zz = vl[k][3], xz = vl[k][4];
      sh0[0][c] = fmaf(m2m3, xx, m2 * zz);
      sh0[1][c] = fmaf(m2, xx, m2m3 * zz);
      sh0[2][c] = m3 * xz;
      sh0[3][c] = m1x * vx;
      sh0[4][c] = m1z * vz;
      const int lx = c / T::LZ0, lz = c - lx * T::LZ0;
      const int lxo = lx - H, lzo = lz - H;
      if (lxo >= 0 && lxo < T::TX && lzo >= 0 && lzo < T::TZ) {
#pragma unroll
        for (int f = 0; f < 5; ++f) raw[f * NT + lxo * T::TZ + lzo] = vl[k][f];
      }
    }
  }
  __syncthreads();

  // level L-1 on the tile; level L of the tile written from LDS
  constexpr int SX0 = T::LZ0;
  const float kz = ((flags & A_K) && zok) ? S.kz[gz] : 1.f;
  float oa[TXR][5], ob[TXR][5];
#pragma unroll
  for (int j = 0; j < TXR; ++j) {
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    const bool ok = zok && gx < nx;
    const int p = (lxo + H) * T::LZ0 + (tz + H);
#pragma unroll
    for (int f = 0; f < 5; ++f) oa[j][f] = ob[j][f] = 0.f;
    if (!ok) continue;
    const int gi = gz + ldz * gx;
#pragma unroll
    for (int f = 0; f < 5; ++f) oa[j][f] = raw[f * NT + lxo * T::TZ + tz];
    float v0 = b1[j][0] - sm<SX0>(sh0[0], p);
    v0 -= sp<1>(sh0[2], p);
    float v1 = b1[j][1] - sm<1>(sh0[1], p);
    v1 -= sp<SX0>(sh0[2], p);
    const float v2 = b1[j][2] - sp<SX0>(sh0[3], p);
    const float v3 = b1[j][3] - sp<1>(sh0[4], p);
    float v4 = b1[j][4] - sm<1>(sh0[3], p);
    v4 -= sm<SX0>(sh0[4], p);
    float v[5] = {v0, v1, v2, v3, v4};
    if (flags & A_K) {
      const float kx = S.kx[gx];
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = (v[f] * kz) * kx;
    }
    if (flags & A_NTS) {
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        __builtin_nontemporal_store(oa[j][f], out0 + f * fs + gi);
        __builtin_nontemporal_store(v[f], out1 + f * fs + gi);
      }
    } else {
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        (out0 + f * fs)[gi] = oa[j][f];
        (out1 + f * fs)[gi] = v[f];
      }
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) ob[j][f] = v[f];
  }

  // imaging of level L, then of level L-1 (uniform flags over the launch)
  const bool comp = flags & A_COMP;
  float* const s0 = shm;
  float* const s1 = shm + Tile<TXR>::N;
  if (flags & A_IMG) {
    const SrcView<SNAP> sv(S, L);
    __syncthreads();
    stage_source<TXR, SNAP>(sv, s0, s1, z0, x0, tid, nz, nx, ldz);
    __syncthreads();
    image_cells<TXR, SNAP, Tile<TXR>::LZ>(S, sv, s0, s1, oa, (flags & A_NEXT) != 0, comp, z0, x0, tz, ty, nz, nx,
                                         ldz, n2);
  }
  if (flags & A_IMG2) {
    const SrcView<SNAP> sv(S, L - 1);
    __syncthreads();
    stage_source<TXR, SNAP>(sv, s0, s1, z0, x0, tid, nz, nx, ldz);
    __syncthreads();
    image_cells<TXR, SNAP, Tile<TXR>::LZ>(S, sv, s0, s1, ob, true, comp, z0, x0, tz, ty, nz, nx, ldz, n2);
  }
}

// ---------------------------------------------------------------------------
// Born scattering sources of level L, image_op2 forward branch
// (W/image-mod-rtm.f90:171-179) from the background S1..S6 (fp32 history).
template <int TXR>
__global__ __launch_bounds__(256) void k_born_sources(const ShotDev* __restrict__ shots,
                                                      const TileRef* __restrict__ tiles, int L) {
  using T = Tile<TXR>;
  __shared__ float sh[2][T::N];
  const TileRef tr = tiles[blockIdx.x];
  const ShotDev& S = shots[tr.shot];
  const int nz = S.nz, nx = S.nx, ldz = S.ldz;
  const long long n2 = S.n2;
  const float* h = static_cast<const float*>(S.snap) + (long long)(L - 1) * 2 * n2;
  const float* h1 = L < S.nt ? h + 2 * n2 : nullptr;
  const int z0 = tr.z0, x0 = tr.x0, tz = threadIdx.x, ty = threadIdx.y;
  for (int lx = ty; lx < T::LX; lx += T::WY) {
    const int gx = x0 - T::H + lx;
    const bool xin = gx >= 0 && gx < nx;
    for (int lz = tz; lz < T::LZ; lz += T::TZ) {
      const int gz = z0 - T::H + lz;
      const bool in = xin && gz >= 0 && gz < nz;
      const long long gi = (long long)gz + (long long)ldz * gx;
      const int p = lx * T::LZ + lz;
      sh[0][p] = in ? h[gi] : 0.f;
      sh[1][p] = in ? h[n2 + gi] : 0.f;
    }
  }
  __syncthreads();
  const int gz = z0 + tz;
  if (gz >= nz) return;
  constexpr int SX = T::LZ;
  for (int j = 0; j < TXR; ++j) {
#pragma clang fp contract(off)
    const int lxo = ty * TXR + j;
    const int gx = x0 + lxo;
    if (gx >= nx) break;
    const int p = (lxo + T::H) * T::LZ + (tz + T::H);
    const long long gi = (long long)gz + (long long)ldz * gx;
    float s1 = 0.f, s2 = 0.f, s4 = 0.f, s5 = 0.f, s6 = 0.f;
    if (h1) {
      s1 = S.inv_dt * (h1[gi] - sh[0][p]);
      s2 = S.inv_dt * (h1[n2 + gi] - sh[1][p]);
    }
    if (gz >= 4 && gz <= nz - 6 && gx >= 4 && gx <= nx - 6) {
      s4 = sp<SX>(sh[0], p) * S.inv_dx;
      s5 = sp<1>(sh[1], p) * S.inv_dx;
    }
    if (gz >= 5 && gz <= nz - 6 && gx >= 5 && gx <= nx - 6) {
      s6 = sm<1>(sh[0], p) * S.inv_dx;
      s6 = fmaf(sm<SX>(sh[1], p), S.inv_dx, s6);
    }
    const float s3 = s4 + s5;
    const float d1 = S.dm[gi], d2 = S.dm[n2 + gi], d3 = S.dm[2 * n2 + gi];
    float* src = S.bsrc;
    src[gi] = d1 * s1;
    src[n2 + gi] = d1 * s2;
    src[2 * n2 + gi] = fmaf(-(d3 * 2.f), s4, -(d2 * s3));
    src[3 * n2 + gi] = fmaf(-(d3 * 2.f), s5, -(d2 * s3));
    src[4 * n2 + gi] = -(d3 * s6);
  }
}

// ---------------------------------------------------------------------------
// Marching step kernels: one wave (64 lanes = 64 z-rows, the fast axis) walks
// a strip of columns x0 .. x0+xw-1 along x. The x-differences come from
// 11-deep register queues (columns x-5 .. x+5 of each field the x-stencils
// read), so a column is loaded once per strip instead of once per tile it
// is a halo of; only the z-differences go through LDS: the centre column of
// the fields they read plus 5 halo rows on each side, loaded by lanes 0..9.
// The next column's operands are issued before the current column is
// computed. Same arithmetic, in the same order, as k_fwd / k_adj, so the
// results are bitwise the same.

// A strip is NW waves stacked along z (64*NW rows); its LDS column holds
// rows z0-5 .. z0+64*NW+4, so the rows between its waves come from LDS and
// only the 5 rows above and below the strip are loaded as halo.
template <int NW>
struct MarchCol {
  static constexpr int MZ = TILE_Z * NW + 10;
};

// x-differences of a queue. The queue holds columns in the order they were
// loaded, oldest first: q[k] is column x-5+k on a march towards +x (DIR 1)
// and column x+5-k on a march towards -x (DIR -1). qv<DIR>(q, j) is column
// x-5+j either way, so both directions perform the same operations in the
// same order as sp / sm.
template <int DIR>
__device__ __forceinline__ float qv(const float (&q)[11], int j) {
  return q[DIR > 0 ? j : 10 - j];
}
template <int DIR>
__device__ __forceinline__ float spq(const float (&q)[11]) {
  float r = C1 * (qv<DIR>(q, 10) - qv<DIR>(q, 1));
  r = fmaf(C2, qv<DIR>(q, 9) - qv<DIR>(q, 2), r);
  r = fmaf(C3, qv<DIR>(q, 8) - qv<DIR>(q, 3), r);
  r = fmaf(C4, qv<DIR>(q, 7) - qv<DIR>(q, 4), r);
  r = fmaf(C5, qv<DIR>(q, 6) - qv<DIR>(q, 5), r);
  return r;
}
template <int DIR>
__device__ __forceinline__ float smq(const float (&q)[11]) {
  float r = C1 * (qv<DIR>(q, 9) - qv<DIR>(q, 0));
  r = fmaf(C2, qv<DIR>(q, 8) - qv<DIR>(q, 1), r);
  r = fmaf(C3, qv<DIR>(q, 7) - qv<DIR>(q, 2), r);
  r = fmaf(C4, qv<DIR>(q, 6) - qv<DIR>(q, 3), r);
  r = fmaf(C5, qv<DIR>(q, 5) - qv<DIR>(q, 4), r);
  return r;
}
__device__ __forceinline__ void push(float (&q)[11], float v) {
#pragma unroll
  for (int k = 0; k < 10; ++k) q[k] = q[k + 1];
  q[10] = v;
}

// LDS exchange between the lanes of one wave (the marching kernels run one
// wave per workgroup). LDS operations of a wave are performed in order, so
// wavefront-scope fences suffice: they keep the compiler from moving LDS
// accesses across the exchange and emit no wait, unlike a workgroup
// barrier, whose fence would also wait for the global loads of the next
// column that are meant to stay in flight.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// The same for the waves of a multi-wave strip: a workgroup barrier whose
// fences cover LDS only, so the global loads of the next column stay in
// flight across it.
template <int NW>
__device__ __forceinline__ void strip_lds_sync() {
  if (NW == 1) {
    wave_lds_sync();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  }
}

// Lane geometry of a marching strip: own row gz = z0 + lane; lanes 0..9 of
// its first wave also carry one halo row hz (z0-5 .. z0-1, z0+64*NW ..
// z0+64*NW+4) at LDS slot hs.
template <int NW>
struct MarchLane {
  int lane, gz, hz, hs;
  bool zok, hok;
  __device__ MarchLane(int z0, int nz) {
    lane = threadIdx.x;
    gz = z0 + lane;
    zok = gz < nz;
    hz = lane < 5 ? z0 - 5 + lane : z0 + TILE_Z * NW - 5 + lane;
    hs = lane < 5 ? lane : TILE_Z * NW + lane;
    hok = lane < 10 && hz >= 0 && hz < nz;
  }
};

// One strip of the forward step, marching towards +x (DIR 1) or -x (DIR -1).
template <int SNAP, bool BORN, int DIR, int NW>
__device__ __forceinline__ void fwdm_strip(const ShotDev& S, const TileRef& tr, int L, int flags, int xw,
                                           float (*col)[4][MarchCol<NW>::MZ], float (*hring)[3][10]) {
  const int nz = S.nz, nx = S.nx, ldz = S.ldz;
  const long long fs = S.fstride;
  const int n2 = (int)S.n2;
  float* const out = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L);
  const float* const x2 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L - 2);
  const float* const y1 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L - 1);
  const float* const coef = S.coef;
  const float* const bsrc = S.bsrc;
  const int z0 = tr.z0, xb = tr.x0, xe = min(tr.x0 + xw, nx);
  const int ncol = xe - xb, xs = DIR > 0 ? xb : xe - 1;  // columns, first column of the march
  const MarchLane<NW> m(z0, nz);
  const int gz = m.gz;
  const bool born = BORN;  // Born sources (F_BORN) in their own instance: 10 fewer live registers

  // queues: sigma_xx (S-x), sigma_xz (S+x), Vx (S+x), Vz (S-x). The
  // priming columns are loaded in increasing x on both march directions, so
  // the columns a strip shares with its neighbour (which starts at the same
  // end, marching the other way) are requested by both at about the same time.
  float qa[11], qb[11], qc[11], qd[11];
#pragma unroll
  for (int i = 0; i < 11; ++i) {
    const int k = DIR > 0 ? i : 10 - i;
    const int gx = xs + DIR * (k - 5);
    const bool in = m.zok && gx >= 0 && gx < nx;
    qa[k] = gldm(y1 + 2 * fs, gz + ldz * gx, in);
    qb[k] = gldm(y1 + 4 * fs, gz + ldz * gx, in);
    qc[k] = gldm(y1, gz + ldz * gx, in);
    qd[k] = gldm(y1 + fs, gz + ldz * gx, in);
    // Halo rows of sigma_xz, Vx, Vz: lanes 0..9 of the first wave load their
    // halo row of a column when the column enters the queues, the step at
    // which the strips above and below load those rows as their own (so the
    // second read is an L2 hit), and keep it in a 7-slot LDS ring (slot =
    // step % 7) until the column reaches the centre six steps later; only
    // the lane that writes an entry reads it. sigma_zz is not queued: its
    // halo rows come with the centre column's own loads.
    if (k >= 5 && m.lane < 10) {
      const bool hin = m.hok && k - 5 < ncol;
      const int ih = m.hz + ldz * gx;
      hring[k - 5][0][m.lane] = gldm(y1 + 4 * fs, ih, hin);
      hring[k - 5][1][m.lane] = gldm(y1, ih, hin);
      hring[k - 5][2][m.lane] = gldm(y1 + fs, ih, hin);
    }
  }
  // Operands of a column: centre sigma_zz, halo rows of the LDS fields, own
  // cell level L-2 (+ Born sources) and coefficients. Loads are issued with
  // clamped indices and masked where the values are taken, so that nothing
  // waits for them before the next column.
  struct Col {
    float zz, h, o[5], c[5], bs[5];
  };
  auto load_col = [&](int gx, bool more, Col& v) {
    const int ic = m.zok && more ? gz + ldz * gx : 0;
    const int ih = m.hok && more ? m.hz + ldz * gx : 0;
    v.zz = gld(y1 + 3 * fs, ic);
    v.h = gld(y1 + 3 * fs, ih);
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      v.o[f] = gld(x2 + f * fs, ic);
      v.c[f] = gld(coef + f * n2, ic);
      v.bs[f] = born ? gld(bsrc + f * n2, ic) : 0.f;
    }
  };
  Col cur;
  load_col(xs, true, cur);
  bool cur_h = m.hok;  // halo values of the current column are valid
  const float kz = ((flags & F_K) && m.zok) ? gld(S.kz, gz) : 1.f;
  float svp[4] = {0.f, 0.f, 0.f, 0.f};  // previous column's snapshot values (16-bit formats)

  for (int st = 0; st < ncol; ++st) {
    const int x = xs + DIR * st;
    const bool more = st + 1 < ncol;
    const int xq = x + 6 * DIR;  // column entering the queues
    const bool qok = m.zok && more && xq >= 0 && xq < nx;
    const int iq = qok ? gz + ldz * xq : 0;
    const float na = gld(y1 + 2 * fs, iq), nb = gld(y1 + 4 * fs, iq), nc = gld(y1, iq), nd = gld(y1 + fs, iq);
    // halo rows of the column entering the queues (centre at step st + 6)
    const bool hq = m.hok && st + 6 < ncol;
    const int ihq = hq ? m.hz + ldz * xq : 0;
    const float hb = gld(y1 + 4 * fs, ihq), hc = gld(y1, ihq), hd = gld(y1 + fs, ihq);
    Col nxt;
    load_col(x + DIR, more, nxt);

    const int b = st & 1;
    col[b][0][m.lane + 5] = m.zok ? cur.zz : 0.f;
    col[b][1][m.lane + 5] = qb[5];
    col[b][2][m.lane + 5] = qc[5];
    col[b][3][m.lane + 5] = qd[5];
    if (m.lane < 10) {
      col[b][0][m.hs] = cur_h ? cur.h : 0.f;
#pragma unroll
      for (int f = 0; f < 3; ++f) col[b][f + 1][m.hs] = hring[st % 7][f][m.lane];
    }
    strip_lds_sync<NW>();

    float sv[4] = {0.f, 0.f, 0.f, 0.f};
    if (m.zok) {
      const int p = m.lane + 5;
      const int gx = x;
      const int gi = gz + ldz * gx;
      const float m1x = cur.c[0], m1z = cur.c[1], m2 = cur.c[2], m3 = cur.c[3], m2m3 = cur.c[4];
      float v[5];
#pragma unroll
      for (int f = 0; f < 5; ++f) v[f] = born ? cur.bs[f] + cur.o[f] : cur.o[f];
      const float dxx_xm = smq<DIR>(qa);        // Ax
      const float dxz_zp = sp<1>(col[b][1], p);  // Bx
      const float dxz_xp = spq<DIR>(qb);        // Az
      const float dzz_zm = sm<1>(col[b][0], p);  // Bz
      const float dvx_xp = spq<DIR>(qc);        // D, F
      const float dvz_zp = sp<1>(col[b][3], p);  // E, G
      const float dvx_zm = sm<1>(col[b][2], p);  // H
      const float dvz_xm = smq<DIR>(qd);        // J
      v[0] = fmaf(m1x, dxx_xm, v[0]);
      v[0] = fmaf(m1x, dxz_zp, v[0]);
      v[1] = fmaf(m1z, dxz_xp, v[1]);
      v[1] = fmaf(m1z, dzz_zm, v[1]);
      if ((flags & F_SRC) && gz == S.zs && gx == S.xs) {
        const float amp = gld(S.src_amp, L - 1);
        v[2] = v[2] + amp;
        v[3] = v[3] + amp;
      }
      v[2] = fmaf(m2m3, dvx_xp, v[2]);
      v[2] = fmaf(m2, dvz_zp, v[2]);
      v[3] = fmaf(m2, dvx_xp, v[3]);
      v[3] = fmaf(m2m3, dvz_zp, v[3]);
      v[4] = fmaf(m3, dvx_zm, v[4]);
      v[4] = fmaf(m3, dvz_xm, v[4]);
      if (flags & F_K) {
        const float kx = gldu(S.kx, gx);
#pragma unroll
        for (int f = 0; f < 5; ++f) v[f] = (v[f] * kz) * kx;
      }
      if (flags & F_NTS) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst_nt(out + f * fs, gi, v[f]);
      } else {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(out + f * fs, gi, v[f]);
      }
      if ((flags & (F_REC | F_RECSUB)) && gz == S.zr && gx >= S.rx0 && gx < S.rx0 + S.nrx) {
        float* r = S.rec + (long long)(L - 1) * S.nrx + (gx - S.rx0);
        const long long rs = (long long)S.nt * S.nrx;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
          if (flags & F_RECSUB)
            gst(r + f * rs, 0, gld(r + f * rs, 0) - v[f]);
          else
            gst(r + f * rs, 0, v[f]);
        }
      }
      if (SNAP == SNAP_HIST && (flags & F_SNAP)) {
        float* h = static_cast<float*>(S.snap) + (long long)(L - 1) * 2 * (long long)n2;
        gst(h, gi, v[0]);
        gst(h + n2, gi, v[1]);
      }
      if (SNAP != SNAP_HIST) {
        sv[0] = qc[5];
        sv[1] = qd[5];
        sv[2] = v[0] - qc[5];
        sv[3] = v[1] - qd[5];
      }
    }
    if (SNAP != SNAP_HIST && (flags & F_SNAP)) {
      const long long k0 = (L - 1) / S.snap_every - 1;
      if (SNAP == SNAP_F32) {
        float* sp4 = static_cast<float*>(S.snap) + k0 * 4 * n2;
        if (m.zok) {
#pragma unroll
          for (int a = 0; a < 4; ++a) gst(sp4 + a * n2, gz + ldz * x, sv[a]);
        }
      } else if (DIR > 0 ? ((x & 1) || x + 1 == nx) : !(x & 1)) {
        // 16-bit formats: one max-abs scale per 64 x 2 strip (columns 2i,
        // 2i+1; strips start on even columns, and so do strips of a march).
        // A pair is stored at its second column in march order; its first
        // (column x - DIR) is held in svp.
        unsigned short* q4 = static_cast<unsigned short*>(S.snap) + k0 * 4 * n2;
        const int strip = (gz >> 6) + S.snap_nbz * (x >> 1);  // the wave's 64 rows
        const bool pair = DIR > 0 ? (x & 1) : x + 1 < xe;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          float mx = 0.f;
          if (pair) mx = fmaxf(mx, fabsf(svp[a]));
          mx = fmaxf(mx, fabsf(sv[a]));
          mx = wave_max(mx);
          const float inv = mx > kSnapFloor ? SnapQ<SNAP>::Q / mx : 0.f;
          // a wave wholly below the grid (last strip of a short grid) has no strip
          if ((m.lane & 63) == 0 && m.zok)
            gst(S.snap_scale + (k0 * 4 + a) * S.snap_nstrip, strip, mx / SnapQ<SNAP>::Q);
          if (m.zok) {
            if (pair) gst(q4 + a * n2, gz + ldz * (x - DIR), SnapQ<SNAP>::enc(svp[a] * inv));
            gst(q4 + a * n2, gz + ldz * x, SnapQ<SNAP>::enc(sv[a] * inv));
          }
        }
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) svp[a] = sv[a];
    }
    push(qa, qok ? na : 0.f);
    push(qb, qok ? nb : 0.f);
    push(qc, qok ? nc : 0.f);
    push(qd, qok ? nd : 0.f);
    if (m.lane < 10) {
      float* r = &hring[(st + 6) % 7][0][m.lane];
      r[0] = hq ? hb : 0.f;
      r[10] = hq ? hc : 0.f;
      r[20] = hq ? hd : 0.f;
    }
    cur = nxt;
    cur_h = m.hok && more;
  }
}

// held to 128 VGPRs (4 waves per SIMD)
template <int SNAP, bool BORN, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void k_fwdm(
    const ShotDev* __restrict__ shots, const TileRef* __restrict__ strips, int nstrips, int L, int flags, int xw,
    int xcd) {
  __shared__ float col[2][4][MarchCol<NW>::MZ];  // sigma_zz, sigma_xz, Vx, Vz of the centre column
  __shared__ float hring[7][3][10];              // sigma_xz, Vx, Vz of the halo rows of the next six columns
  const TileRef tr = strips[xcd ? xcd_tile(blockIdx.x, nstrips) : (int)blockIdx.x];
  // neighbouring strips march in opposite directions, so the columns they
  // share (a strip's priming columns) are read by both at the same end of
  // their marches, close in time, and come from L2 for the second
  if (((tr.x0 / xw) & 1) == 0)
    fwdm_strip<SNAP, BORN, 1, NW>(shots[tr.shot], tr, L, flags, xw, col, hring);
  else
    fwdm_strip<SNAP, BORN, -1, NW>(shots[tr.shot], tr, L, flags, xw, col, hring);
}

// Adjoint step (gather form, as k_adj) marching along x: the products
// P0 = m2m3 sxx + m2 szz, P1 = m2 sxx + m2m3 szz, P2 = m3 sxz, P3 = m1x Vx,
// P4 = m1z Vz of level L+1 are formed once per column as it enters the
// queues; imaging reads the source Vx, Vz the same way.
// XIN: the drop-in form (the level's own content added first, no imaging
// or injection), a separate instance: the engine's instance carries neither
// its loads nor their registers.
template <int SNAP, int DIR, int NW, bool XIN, bool COMP, bool PROP>
__device__ __forceinline__ void adjm_strip(const ShotDev& S, const TileRef& tr, int L, int flags, int xw,
                                           float (*col)[6][MarchCol<NW>::MZ], float (*hring)[6][10]) {
  const int nz = S.nz, nx = S.nx, ldz = S.ldz;
  const long long fs = S.fstride;
  const int n2 = (int)S.n2;
  float* const out = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L);
  const float* const x2 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L + 2);
  const float* const y1 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L + 1);
  const float* const coef = S.coef;
  const int z0 = tr.z0, xb = tr.x0, xe = min(tr.x0 + xw, nx);
  const int ncol = xe - xb, xs = DIR > 0 ? xb : xe - 1;  // columns, first column of the march
  const MarchLane<NW> m(z0, nz);
  const int gz = m.gz;
  // the drop-in form has no imaging, injection or sponge; PROP: the
  // instance for levels that are not imaged (no imaging code or registers)
  const bool imaging = !XIN && !PROP && (flags & A_IMG);
  const bool next = flags & A_NEXT;
  constexpr bool comp = !XIN && COMP;  // compensated image sums (A_COMP), an instance of its own
  constexpr bool xin = XIN;
  const SrcView<SNAP> sv(S, L);
  // 64-row blocks of the 16-bit snapshot scales: the wave's own rows (clamped
  // for a wave wholly below the grid) and the halo rows above and below the
  // strip; all wave-uniform, so their scales are scalar loads
  const int zbw = min(__builtin_amdgcn_readfirstlane(gz) >> 6, S.snap_nbz - 1);
  const int zbt = max((z0 >> 6) - 1, 0), zbb = min((z0 >> 6) + NW, S.snap_nbz - 1);
  auto cx = [&](int gx) { return min(max(gx, 0), nx - 1); };

  // the ten level-(L+1) operands at element i (clamped index), as loaded
  struct Raw {
    float vx, vz, xx, zz, xz, m1x, m1z, m2, m3, m2m3;
  };
  auto raw = [&](int i, Raw& w) {
    w.vx = gld(y1, i);
    w.vz = gld(y1 + fs, i);
    w.xx = gld(y1 + 2 * fs, i);
    w.zz = gld(y1 + 3 * fs, i);
    w.xz = gld(y1 + 4 * fs, i);
    w.m1x = gld(coef, i);
    w.m1z = gld(coef + n2, i);
    w.m2 = gld(coef + 2 * n2, i);
    w.m3 = gld(coef + 3 * n2, i);
    w.m2m3 = gld(coef + 4 * n2, i);
  };

  // Halo rows: lanes 0..9 of the first wave load their halo row of a column
  // when the column enters the queues, the step at which the strips above
  // and below load those rows as their own (so the second read is an L2
  // hit), and keep its four products and two source values in a 7-slot LDS
  // ring (slot = step % 7) until the column reaches the centre six steps
  // later. Only the lane that writes an entry reads it.
  struct Halo {
    Raw w;
    typename SrcView<SNAP>::Raw s[2];
    float sc[2][2];  // 16-bit scales of the source values: above, below
  };
  auto load_halo = [&](int gx, bool in, Halo& h) {
    const int ih = in ? m.hz + ldz * gx : 0;
    raw(ih, h.w);
    if (imaging) {
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        sv.load_q(a, ih, in, h.s[a]);
        h.sc[a][0] = sv.scale_u(a, zbt, cx(gx));
        h.sc[a][1] = sv.scale_u(a, zbb, cx(gx));
      }
    }
  };
  auto put_halo = [&](int slot, bool in, const Halo& h) {
    if (m.lane < 10) {
      float* r = &hring[slot][0][m.lane];
      r[0] = in ? fmaf(h.w.m2, h.w.xx, h.w.m2m3 * h.w.zz) : 0.f;
      r[10] = in ? h.w.m3 * h.w.xz : 0.f;
      r[20] = in ? h.w.m1x * h.w.vx : 0.f;
      r[30] = in ? h.w.m1z * h.w.vz : 0.f;
      if (imaging) {
        r[40] = sv.value_s(h.s[0], m.lane < 5 ? h.sc[0][0] : h.sc[0][1]);
        r[50] = sv.value_s(h.s[1], m.lane < 5 ? h.sc[1][0] : h.sc[1][1]);
      }
    }
  };

  // queues: P0 (S-x), P2 (S+x), P3 (S+x), P4 (S-x); P1 at columns x .. x+5.
  // Priming columns in increasing x on both march directions (see k_fwdm).
  float q0[11], q2[11], q3[11], q4[11], d1[6];
  float s0q[11], s1q[11];  // source Vx (S+x), Vz (S-x)
#pragma unroll
  for (int i = 0; i < 11; ++i) {
    const int k = DIR > 0 ? i : 10 - i;
    const int gx = xs + DIR * (k - 5);
    const bool in = m.zok && gx >= 0 && gx < nx;
    Raw w;
    raw(in ? gz + ldz * gx : 0, w);
    q0[k] = in ? fmaf(w.m2m3, w.xx, w.m2 * w.zz) : 0.f;
    q2[k] = in ? w.m3 * w.xz : 0.f;
    q3[k] = in ? w.m1x * w.vx : 0.f;
    q4[k] = in ? w.m1z * w.vz : 0.f;
    if (k >= 5) d1[k - 5] = in ? fmaf(w.m2, w.xx, w.m2m3 * w.zz) : 0.f;
    if (imaging) {
      s0q[k] = sv.getm(0, gz + ldz * gx, gz, gx, in);
      s1q[k] = sv.getm(1, gz + ldz * gx, gz, gx, in);
    } else {
      s0q[k] = s1q[k] = 0.f;
    }
    if (k >= 5) {  // the strip's first six columns: halo rows into ring slots 0..5
      const bool hin = m.hok && k - 5 < ncol;
      Halo h;
      load_halo(gx, hin, h);
      put_halo(k - 5, hin, h);
    }
  }
  const bool row = !XIN && (flags & A_INJ) && gz == S.zr;
  // Operands of a column, loaded with clamped indices and masked where they
  // are taken: own cell (level L+2, the level's own content for the drop-in
  // form, the injected receivers), imaging operands.
  struct Col {
    float b[5], add[5];  // level L+2; the level's own content (drop-in form) or the injected receivers
    typename SrcView<SNAP>::Raw nt2[2];
    float ntsc[2];  // their 16-bit scales
    float im[3], imc[3];
  };
  auto load_col = [&](int gx, bool more, Col& c) {
    const int ic = m.zok && more ? gz + ldz * gx : 0;
    const bool inj = row && more && gx >= S.rx0 && gx < S.rx0 + S.nrx;
    const float* r = S.rec + (long long)(L - 1) * S.nrx + (inj ? gx - S.rx0 : 0);
    const long long rs = (long long)S.nt * S.nrx;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      c.b[f] = gld(x2 + f * fs, ic);
      if (xin)
        c.add[f] = gld(out + f * fs, ic);
      else
        c.add[f] = row ? gld(r + f * rs, 0) : 0.f;
    }
    if (imaging) {
      if (next) {  // arrays 2, 3 exist only below the last level
        sv.load_q(2, ic, m.zok && more, c.nt2[0]);
        sv.load_q(3, ic, m.zok && more, c.nt2[1]);
        c.ntsc[0] = sv.scale_u(2, zbw, cx(gx));
        c.ntsc[1] = sv.scale_u(3, zbw, cx(gx));
      } else {
        c.nt2[0].in = c.nt2[1].in = false;
        c.ntsc[0] = c.ntsc[1] = 0.f;
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        c.im[a] = gld(S.img + a * n2, ic);
        c.imc[a] = comp ? gld(S.imgc + a * n2, ic) : 0.f;
      }
    }
  };
  Col cur;
  load_col(xs, true, cur);
  bool cur_inj = row && xs >= S.rx0 && xs < S.rx0 + S.nrx;
  const float kz = ((flags & A_K) && m.zok) ? gld(S.kz, gz) : 1.f;
  const bool zin4 = gz >= 4 && gz <= nz - 6, zin5 = gz >= 5 && gz <= nz - 6;
  const float w = S.img_w, inv_dx = S.inv_dx, inv_dt = S.inv_dt;

  // two columns per trip, their column operands in two buffers that swap
  // roles (no copy of one into the other)
  auto column = [&](int st, Col& cur, Col& nxt) __attribute__((always_inline)) {
    const int x = xs + DIR * st;
    const bool more = st + 1 < ncol;
    const int xq = x + 6 * DIR;  // column entering the queues
    const bool qok = m.zok && more && xq >= 0 && xq < nx;
    const int iq = qok ? gz + ldz * xq : 0;
    Raw nr;
    raw(iq, nr);
    typename SrcView<SNAP>::Raw ns0, ns1;
    float nsc0 = 0.f, nsc1 = 0.f;
    if (imaging) {
      sv.load_q(0, iq, qok, ns0);
      sv.load_q(1, iq, qok, ns1);
      nsc0 = sv.scale_u(0, zbw, cx(xq));
      nsc1 = sv.scale_u(1, zbw, cx(xq));
    }
    // halo rows of the column entering the queues (centre at step st + 6)
    const bool hq = m.hok && st + 6 < ncol;
    Halo nh;
    load_halo(xq, hq, nh);
    load_col(x + DIR, more, nxt);

    const int b = st & 1;
    col[b][0][m.lane + 5] = d1[0];
    col[b][1][m.lane + 5] = q2[5];
    col[b][2][m.lane + 5] = q3[5];
    col[b][3][m.lane + 5] = q4[5];
    if (imaging) {
      col[b][4][m.lane + 5] = s0q[5];
      col[b][5][m.lane + 5] = s1q[5];
    }
    if (m.lane < 10) {
      const float* r = &hring[st % 7][0][m.lane];
#pragma unroll
      for (int f = 0; f < 4; ++f) col[b][f][m.hs] = r[10 * f];
      if (imaging) {
        col[b][4][m.hs] = r[40];
        col[b][5][m.hs] = r[50];
      }
    }
    strip_lds_sync<NW>();

    if (m.zok) {
      const int p = m.lane + 5;
      const int gx = x;
      const int gi = gz + ldz * gx;
      const float kx = (flags & A_K) ? gldu(S.kx, gx) : 1.f;
      float bb[5];
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        bb[f] = cur.b[f];
        if (flags & A_KSRC) bb[f] = (bb[f] * kz) * kx;
        if (xin || cur_inj) bb[f] = cur.add[f] + bb[f];
      }
      float v0 = bb[0] - smq<DIR>(q0);
      v0 -= sp<1>(col[b][1], p);
      float v1 = bb[1] - sm<1>(col[b][0], p);
      v1 -= spq<DIR>(q2);
      const float v2 = bb[2] - spq<DIR>(q3);
      const float v3 = bb[3] - sp<1>(col[b][3], p);
      float v4 = bb[4] - sm<1>(col[b][2], p);
      v4 -= smq<DIR>(q4);
      float v[5] = {v0, v1, v2, v3, v4};
      // with the R/Src sponge, level 1 (read by no later step) keeps its
      // undamped value, which is also the one imaged
      const bool ksrc = flags & A_KSRC;
      if ((flags & A_K) && (!ksrc || L >= 2)) {
#pragma unroll
        for (int f = 0; f < 5; ++f) v[f] = (v[f] * kz) * kx;
      }
      float vi[5];  // the level's imaged values
#pragma unroll
      for (int f = 0; f < 5; ++f) vi[f] = ksrc && L >= 3 ? (v[f] * kz) * kx : v[f];
      if (flags & A_NTS) {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst_nt(out + f * fs, gi, v[f]);
      } else {
#pragma unroll
        for (int f = 0; f < 5; ++f) gst(out + f * fs, gi, v[f]);
      }
      if (imaging) {
        // the imaging arithmetic is fused only where written (fmaf), so
        // every kernel variant rounds it the same way
#pragma clang fp contract(off)
        float s1 = 0.f, s2 = 0.f;
        if (next) {
          if (SNAP == SNAP_HIST) {  // arrays 2, 3 are Vx, Vz of level L+1
            s1 = inv_dt * (sv.value_s(cur.nt2[0], cur.ntsc[0]) - s0q[5]);
            s2 = inv_dt * (sv.value_s(cur.nt2[1], cur.ntsc[1]) - s1q[5]);
          } else {
            s1 = inv_dt * sv.value_s(cur.nt2[0], cur.ntsc[0]);
            s2 = inv_dt * sv.value_s(cur.nt2[1], cur.ntsc[1]);
          }
        }
        float s4 = 0.f, s5 = 0.f, s6 = 0.f;
        if (zin4 && gx >= 4 && gx <= nx - 6) {
          s4 = spq<DIR>(s0q) * inv_dx;
          s5 = sp<1>(col[b][5], p) * inv_dx;
        }
        if (zin5 && gx >= 5 && gx <= nx - 6) {
          s6 = sm<1>(col[b][4], p) * inv_dx;
          s6 = fmaf(smq<DIR>(s1q), inv_dx, s6);
        }
        const float s3 = s4 + s5;
        const float t[3] = {w * fmaf(vi[0], s1, vi[1] * s2), -w * fmaf(vi[2], s3, vi[3] * s3),
                            -w * fmaf(vi[4], s6, fmaf(vi[3] * 2.f, s5, vi[2] * 2.f * s4))};
        if (comp) {
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            float sa = cur.im[a], ca = cur.imc[a];
            neumaier(sa, ca, t[a]);
            gst(S.img + a * n2, gi, sa);
            gst(S.imgc + a * n2, gi, ca);
          }
        } else {
#pragma unroll
          for (int a = 0; a < 3; ++a) gst(S.img + a * n2, gi, cur.im[a] + t[a]);
        }
      }
    }
    push(q0, qok ? fmaf(nr.m2m3, nr.xx, nr.m2 * nr.zz) : 0.f);
    push(q2, qok ? nr.m3 * nr.xz : 0.f);
    push(q3, qok ? nr.m1x * nr.vx : 0.f);
    push(q4, qok ? nr.m1z * nr.vz : 0.f);
#pragma unroll
    for (int k = 0; k < 5; ++k) d1[k] = d1[k + 1];
    d1[5] = qok ? fmaf(nr.m2, nr.xx, nr.m2m3 * nr.zz) : 0.f;
    if (imaging) {
      push(s0q, sv.value_s(ns0, nsc0));
      push(s1q, sv.value_s(ns1, nsc1));
    }
    put_halo((st + 6) % 7, hq, nh);
    cur_inj = row && more && x + DIR >= S.rx0 && x + DIR < S.rx0 + S.nrx;
  };
  Col alt;
  for (int st = 0; st < ncol; st += 2) {
    column(st, cur, alt);
    if (st + 1 < ncol) column(st + 1, alt, cur);
  }
}

template <int SNAP, int NW, bool XIN, bool COMP, bool PROP = false>
__device__ __forceinline__ void adjm_kernel(const ShotDev* __restrict__ shots, const TileRef* __restrict__ strips,
                                            int nstrips, int L, int flags, int xw, int xcd) {
  __shared__ float col[2][6][MarchCol<NW>::MZ];  // P1, P2, P3, P4, source Vx, source Vz of the centre column
  __shared__ float hring[7][6][10];              // the same for the halo rows of the next six columns
  const TileRef tr = strips[xcd ? xcd_tile(blockIdx.x, nstrips) : (int)blockIdx.x];
  // neighbouring strips march in opposite directions (see k_fwdm)
  if (((tr.x0 / xw) & 1) == 0)
    adjm_strip<SNAP, 1, NW, XIN, COMP, PROP>(shots[tr.shot], tr, L, flags, xw, col, hring);
  else
    adjm_strip<SNAP, -1, NW, XIN, COMP, PROP>(shots[tr.shot], tr, L, flags, xw, col, hring);
}
// fp32 snapshot formats: held to 168 VGPRs (3 waves per SIMD; a few spills,
// measured 9% faster than 2 waves without them)
template <int SNAP, int NW, bool COMP>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(3))) void k_adjm(
    const ShotDev* __restrict__ shots, const TileRef* __restrict__ strips, int nstrips, int L, int flags, int xw,
    int xcd) {
  adjm_kernel<SNAP, NW, false, COMP>(shots, strips, nstrips, L, flags, xw, xcd);
}
// the drop-in form (fp32 histories, A_XIN)
template <int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(3))) void k_adjm_xin(
    const ShotDev* __restrict__ shots, const TileRef* __restrict__ strips, int nstrips, int L, int flags, int xw,
    int xcd) {
  adjm_kernel<SNAP_HIST, NW, true, false>(shots, strips, nstrips, L, flags, xw, xcd);
}
// levels that are not imaged (decimated snapshots): propagation, injection
// and sponge only, held to 128 VGPRs (4 waves per SIMD)
template <int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void k_adjm_prop(
    const ShotDev* __restrict__ shots, const TileRef* __restrict__ strips, int nstrips, int L, int flags, int xw,
    int xcd) {
  adjm_kernel<SNAP_HIST, NW, false, false, true>(shots, strips, nstrips, L, flags, xw, xcd);
}
// 16-bit snapshot formats: their decode needs more registers; held to 168
// (3 waves per SIMD, at most 3 spills), measured 7% faster on the BF16
// 64-shot workload than 2 waves per SIMD without spills
template <int SNAP, int NW, bool COMP>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(3))) void k_adjm16(const ShotDev* __restrict__ shots,
                                                   const TileRef* __restrict__ strips, int nstrips, int L, int flags,
                                                   int xw, int xcd) {
  adjm_kernel<SNAP, NW, false, COMP>(shots, strips, nstrips, L, flags, xw, xcd);
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

__global__ void k_stack(double* stack, int snz, int snx, const float* img, int ldz, long long n2,
                        int iz0, int ix0, int nz, int nx, int z0, int x0) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x;
  const int ix = blockIdx.y;
  if (iz >= nz || ix >= nx) return;
  const long long src = (long long)(iz0 + iz) + (long long)ldz * (ix0 + ix);
  const long long dst = (long long)(z0 + iz) + (long long)snz * (x0 + ix);
  const long long sn = (long long)snz * snx;
#pragma unroll
  for (int c = 0; c < 3; ++c) stack[c * sn + dst] += (double)img[c * n2 + src];
}

// ---------------------------------------------------------------------------
template <int TXR>
static void fwd_dispatch(int snap, dim3 g, hipStream_t st, const ShotDev* s, const TileRef* t,
                         int n, int L, int fl, int xcd) {
  const dim3 b(TILE_Z, TILE_WY);
  switch (snap) {
    case SNAP_HIST: hipLaunchKernelGGL((k_fwd<TXR, SNAP_HIST>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    case SNAP_F32: hipLaunchKernelGGL((k_fwd<TXR, SNAP_F32>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    case SNAP_BF16: hipLaunchKernelGGL((k_fwd<TXR, SNAP_BF16>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    default: hipLaunchKernelGGL((k_fwd<TXR, SNAP_F16>), g, b, 0, st, s, t, n, L, fl, xcd); break;
  }
}
template <int TXR>
static void fwd2_dispatch(int snap, dim3 g, hipStream_t st, const ShotDev* s, const TileRef* t,
                          int n, int L, int fl, int xcd) {
  const dim3 b(TILE_Z, TILE_WY);
  switch (snap) {
    case SNAP_HIST: hipLaunchKernelGGL((k_fwd2<TXR, SNAP_HIST>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    case SNAP_F32: hipLaunchKernelGGL((k_fwd2<TXR, SNAP_F32>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    case SNAP_BF16: hipLaunchKernelGGL((k_fwd2<TXR, SNAP_BF16>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    default: hipLaunchKernelGGL((k_fwd2<TXR, SNAP_F16>), g, b, 0, st, s, t, n, L, fl, xcd); break;
  }
}
template <int TXR>
static void adj_dispatch(int snap, dim3 g, hipStream_t st, const ShotDev* s, const TileRef* t,
                         int n, int L, int fl, int xcd) {
  const dim3 b(TILE_Z, TILE_WY);
  switch (snap) {
    case SNAP_HIST: hipLaunchKernelGGL((k_adj<TXR, SNAP_HIST>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    case SNAP_F32: hipLaunchKernelGGL((k_adj<TXR, SNAP_F32>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    case SNAP_BF16: hipLaunchKernelGGL((k_adj<TXR, SNAP_BF16>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    default: hipLaunchKernelGGL((k_adj<TXR, SNAP_F16>), g, b, 0, st, s, t, n, L, fl, xcd); break;
  }
}
template <int TXR>
static void adj2_dispatch(int snap, dim3 g, hipStream_t st, const ShotDev* s, const TileRef* t,
                          int n, int L, int fl, int xcd) {
  const dim3 b(TILE_Z, TILE_WY);
  switch (snap) {
    case SNAP_HIST: hipLaunchKernelGGL((k_adj2<TXR, SNAP_HIST>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    case SNAP_F32: hipLaunchKernelGGL((k_adj2<TXR, SNAP_F32>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    case SNAP_BF16: hipLaunchKernelGGL((k_adj2<TXR, SNAP_BF16>), g, b, 0, st, s, t, n, L, fl, xcd); break;
    default: hipLaunchKernelGGL((k_adj2<TXR, SNAP_F16>), g, b, 0, st, s, t, n, L, fl, xcd); break;
  }
}
template <int TXR>
static void snap_last_dispatch(int snap, dim3 g, hipStream_t st, const ShotDev* s,
                               const TileRef* t, int L) {
  const dim3 b(TILE_Z, TILE_WY);
  switch (snap) {
    case SNAP_F32: hipLaunchKernelGGL((k_snap_last<TXR, SNAP_F32>), g, b, 0, st, s, t, L); break;
    case SNAP_BF16: hipLaunchKernelGGL((k_snap_last<TXR, SNAP_BF16>), g, b, 0, st, s, t, L); break;
    case SNAP_F16: hipLaunchKernelGGL((k_snap_last<TXR, SNAP_F16>), g, b, 0, st, s, t, L); break;
    default: break;
  }
}

int launch_fwd(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int flags, int txr,
               int snap, int xcd, hipStream_t st) {
  if (ntiles <= 0) return OK;
  const dim3 g(ntiles);
  switch (txr) {
    case 1: fwd_dispatch<1>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    case 2: fwd_dispatch<2>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    case 4: fwd_dispatch<4>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    case 8: fwd_dispatch<8>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    default: set_error("unsupported tile width %d", txr); return E_ARG;
  }
  HIPCHK(hipGetLastError());
  return OK;
}

int launch_fwd2(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int flags, int txr,
                int snap, int xcd, hipStream_t st) {
  if (ntiles <= 0) return OK;
  if (flags & F_BORN) { set_error("two-level forward step has no Born sources"); return E_ARG; }
  const dim3 g(ntiles);
  switch (txr) {
    case 1: fwd2_dispatch<1>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    case 2: fwd2_dispatch<2>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    case 4: fwd2_dispatch<4>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    default: set_error("two-level forward step: unsupported tile width %d", txr); return E_ARG;
  }
  HIPCHK(hipGetLastError());
  return OK;
}

int launch_adj(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int flags, int txr,
               int snap, int xcd, hipStream_t st) {
  if (ntiles <= 0) return OK;
  const dim3 g(ntiles);
  switch (txr) {
    case 1: adj_dispatch<1>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    case 2: adj_dispatch<2>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    case 4: adj_dispatch<4>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    case 8: adj_dispatch<8>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    default: set_error("unsupported tile width %d", txr); return E_ARG;
  }
  HIPCHK(hipGetLastError());
  return OK;
}

int launch_adj2(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int flags, int txr,
                int snap, int xcd, hipStream_t st) {
  if (ntiles <= 0) return OK;
  if (flags & A_XIN) { set_error("two-level adjoint step has no drop-in history form"); return E_ARG; }
  const dim3 g(ntiles);
  switch (txr) {
    case 1: adj2_dispatch<1>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    case 2: adj2_dispatch<2>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    case 4: adj2_dispatch<4>(snap, g, st, shots, tiles, ntiles, L, flags, xcd); break;
    default: set_error("two-level adjoint step: unsupported tile width %d", txr); return E_ARG;
  }
  HIPCHK(hipGetLastError());
  return OK;
}

int launch_snap_last(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int txr,
                     int snap, hipStream_t st) {
  if (ntiles <= 0 || snap == SNAP_HIST) return OK;
  const dim3 g(ntiles);
  switch (txr) {
    case 1: snap_last_dispatch<1>(snap, g, st, shots, tiles, L); break;
    case 2: snap_last_dispatch<2>(snap, g, st, shots, tiles, L); break;
    case 4: snap_last_dispatch<4>(snap, g, st, shots, tiles, L); break;
    case 8: snap_last_dispatch<8>(snap, g, st, shots, tiles, L); break;
    default: set_error("unsupported tile width %d", txr); return E_ARG;
  }
  HIPCHK(hipGetLastError());
  return OK;
}

int launch_born_sources(const ShotDev* shots, const TileRef* tiles, int ntiles, int L, int txr,
                        hipStream_t st) {
  if (ntiles <= 0) return OK;
  const dim3 g(ntiles), b(TILE_Z, TILE_WY);
  switch (txr) {
    case 1: hipLaunchKernelGGL(k_born_sources<1>, g, b, 0, st, shots, tiles, L); break;
    case 2: hipLaunchKernelGGL(k_born_sources<2>, g, b, 0, st, shots, tiles, L); break;
    case 4: hipLaunchKernelGGL(k_born_sources<4>, g, b, 0, st, shots, tiles, L); break;
    case 8: hipLaunchKernelGGL(k_born_sources<8>, g, b, 0, st, shots, tiles, L); break;
    default: set_error("Born sources: unsupported tile width %d", txr); return E_ARG;
  }
  HIPCHK(hipGetLastError());
  return OK;
}

template <int NW>
static void fwdm_dispatch(int snap, int flags, dim3 g, hipStream_t st, const ShotDev* s, const TileRef* t, int n,
                          int L, int xw, int xcd) {
  const dim3 b(TILE_Z * NW);
  switch (snap) {
    case SNAP_HIST:
      if (flags & F_BORN)
        hipLaunchKernelGGL((k_fwdm<SNAP_HIST, true, NW>), g, b, 0, st, s, t, n, L, flags, xw, xcd);
      else
        hipLaunchKernelGGL((k_fwdm<SNAP_HIST, false, NW>), g, b, 0, st, s, t, n, L, flags, xw, xcd);
      break;
    case SNAP_F32: hipLaunchKernelGGL((k_fwdm<SNAP_F32, false, NW>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
    case SNAP_BF16: hipLaunchKernelGGL((k_fwdm<SNAP_BF16, false, NW>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
    case SNAP_F16: hipLaunchKernelGGL((k_fwdm<SNAP_F16, false, NW>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
    default: break;
  }
}
template <int NW>
static void adjm_dispatch(int snap, dim3 g, hipStream_t st, const ShotDev* s, const TileRef* t, int n, int L,
                          int flags, int xw, int xcd) {
  const dim3 b(TILE_Z * NW);
  if (flags & A_XIN) {  // drop-in form (fp32 histories, checked by the launcher)
    hipLaunchKernelGGL((k_adjm_xin<NW>), g, b, 0, st, s, t, n, L, flags, xw, xcd);
    return;
  }
  if (!(flags & A_IMG)) {  // a level that is not imaged: any snapshot format
    hipLaunchKernelGGL((k_adjm_prop<NW>), g, b, 0, st, s, t, n, L, flags, xw, xcd);
    return;
  }
  if (flags & A_COMP) {
    switch (snap) {
      case SNAP_HIST: hipLaunchKernelGGL((k_adjm<SNAP_HIST, NW, true>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
      case SNAP_F32: hipLaunchKernelGGL((k_adjm<SNAP_F32, NW, true>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
      case SNAP_BF16: hipLaunchKernelGGL((k_adjm16<SNAP_BF16, NW, true>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
      case SNAP_F16: hipLaunchKernelGGL((k_adjm16<SNAP_F16, NW, true>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
      default: break;
    }
    return;
  }
  switch (snap) {
    case SNAP_HIST: hipLaunchKernelGGL((k_adjm<SNAP_HIST, NW, false>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
    case SNAP_F32: hipLaunchKernelGGL((k_adjm<SNAP_F32, NW, false>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
    case SNAP_BF16: hipLaunchKernelGGL((k_adjm16<SNAP_BF16, NW, false>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
    case SNAP_F16: hipLaunchKernelGGL((k_adjm16<SNAP_F16, NW, false>), g, b, 0, st, s, t, n, L, flags, xw, xcd); break;
    default: break;
  }
}

template <int NW, bool COMP>
static int adjm_occupancy(int snap) {
  int n = 0;
  const int b = TILE_Z * NW;
  hipError_t e = hipSuccess;
  switch (snap) {
    case SNAP_HIST: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjm<SNAP_HIST, NW, COMP>, b, 0); break;
    case SNAP_F32: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjm<SNAP_F32, NW, COMP>, b, 0); break;
    case SNAP_BF16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjm16<SNAP_BF16, NW, COMP>, b, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_adjm16<SNAP_F16, NW, COMP>, b, 0); break;
  }
  return e == hipSuccess ? n : 0;
}

template <int NW>
static int fwdm_occupancy(int snap) {
  int n = 0;
  const int b = TILE_Z * NW;
  hipError_t e = hipSuccess;
  switch (snap) {
    case SNAP_HIST: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fwdm<SNAP_HIST, false, NW>, b, 0); break;
    case SNAP_F32: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fwdm<SNAP_F32, false, NW>, b, 0); break;
    case SNAP_BF16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fwdm<SNAP_BF16, false, NW>, b, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fwdm<SNAP_F16, false, NW>, b, 0); break;
  }
  return e == hipSuccess ? n : 0;
}

static int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return cus;
}

int fwdm_resident_blocks(int nw, int snap) {
  const int per = nw == 1 ? fwdm_occupancy<1>(snap) : nw == 2 ? fwdm_occupancy<2>(snap) : fwdm_occupancy<4>(snap);
  return per * device_cus();
}

int adjm_resident_blocks(int nw, int snap, bool comp) {
  const int per = comp ? (nw == 1 ? adjm_occupancy<1, true>(snap) : nw == 2 ? adjm_occupancy<2, true>(snap)
                                                                          : adjm_occupancy<4, true>(snap))
                       : (nw == 1 ? adjm_occupancy<1, false>(snap) : nw == 2 ? adjm_occupancy<2, false>(snap)
                                                                           : adjm_occupancy<4, false>(snap));
  return per * device_cus();
}

static int march_args(int nstrips, int xw, int nw, int snap) {
  if (xw < 2 || (xw & 1)) { set_error("marching strip width %d is not a positive even number", xw); return E_ARG; }
  if (nw != 1 && nw != 2 && nw != 4) { set_error("marching strip height %d waves not in {1,2,4}", nw); return E_ARG; }
  if (snap < SNAP_HIST || snap > SNAP_F16) { set_error("unknown snapshot format %d", snap); return E_ARG; }
  (void)nstrips;
  return OK;
}

int launch_fwdm(const ShotDev* shots, const TileRef* strips, int nstrips, int L, int flags, int xw, int nw, int snap,
                int xcd, hipStream_t st) {
  if (nstrips <= 0) return OK;
  STCHK(march_args(nstrips, xw, nw, snap));
  if ((flags & F_BORN) && snap != SNAP_HIST) { set_error("Born sources need the fp32 history format"); return E_ARG; }
  const dim3 g(nstrips);
  switch (nw) {
    case 1: fwdm_dispatch<1>(snap, flags, g, st, shots, strips, nstrips, L, xw, xcd); break;
    case 2: fwdm_dispatch<2>(snap, flags, g, st, shots, strips, nstrips, L, xw, xcd); break;
    default: fwdm_dispatch<4>(snap, flags, g, st, shots, strips, nstrips, L, xw, xcd); break;
  }
  HIPCHK(hipGetLastError());
  return OK;
}

int launch_adjm(const ShotDev* shots, const TileRef* strips, int nstrips, int L, int flags, int xw, int nw, int snap,
                int xcd, hipStream_t st) {
  if (nstrips <= 0) return OK;
  STCHK(march_args(nstrips, xw, nw, snap));
  if ((flags & A_XIN) && (snap != SNAP_HIST || (flags & (A_IMG | A_INJ | A_K)))) {
    set_error("drop-in adjoint form: fp32 histories without imaging, injection or sponge");
    return E_ARG;
  }
  const dim3 g(nstrips);
  switch (nw) {
    case 1: adjm_dispatch<1>(snap, g, st, shots, strips, nstrips, L, flags, xw, xcd); break;
    case 2: adjm_dispatch<2>(snap, g, st, shots, strips, nstrips, L, flags, xw, xcd); break;
    default: adjm_dispatch<4>(snap, g, st, shots, strips, nstrips, L, flags, xw, xcd); break;
  }
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

int launch_stack(double* stack, int snz, int snx, const float* img, int ldz, long long n2,
                 int iz0, int ix0, int nz, int nx, int z0, int x0, hipStream_t st) {
  if (nz <= 0 || nx <= 0) return OK;
  hipLaunchKernelGGL(k_stack, dim3((nz + 255) / 256, nx), dim3(256), 0, st, stack, snz, snx, img,
                     ldz, n2, iz0, ix0, nz, nx, z0, x0);
  HIPCHK(hipGetLastError());
  return OK;
}

}  // namespace erm
