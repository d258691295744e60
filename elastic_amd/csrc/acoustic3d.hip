// acoustic3d.hip -- 3D acoustic 10th-order (radius 5) propagator and
// cross-correlation imaging on one GPU's subdomain (gfx950).
//
// The reference's multi-GPU acoustic RTM (W/wave_fkernel.3d8o.cu:31-106,
// :183-210; W/gpu_funcs_3d.cu:64-131) is 8th order with separate damping,
// injection and imaging kernels. Here one kernel does a whole step:
//   p_next = K (2 p - p_prev + c lap p),  c = (v dt/dx)^2,  K = kz kx ky
// for the interior (5 cells from every face), then the point source or the
// receiver-plane injection, the receiver recording, the per-snapshot max-abs
// of the source wavefield and, in the receiver pass, the image update.
// Layout: (n1, n2, n3) = (z fastest, x, y slowest). A workgroup owns a 64 x 16
// tile of the (z, x) plane and marches along y over a chunk of planes; the
// y-neighbours live in an 11-deep register queue, the (z, x) halo of the
// current plane in LDS. Each step moves 16 algorithmic bytes per cell
// (read p, p_prev, c; write p_next): HBM-bound.

#include "../../include/acoustic3d.h"
#include "common.hpp"
#include "engine2d.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

namespace erm {

namespace {

constexpr float W0 = -5269.0f / 1800.0f, W1 = 5.0f / 3.0f, W2 = -5.0f / 21.0f, W3 = 5.0f / 126.0f,
                W4 = -5.0f / 1008.0f, W5 = 1.0f / 3150.0f;
constexpr int AZ = 64;     // tile along n1 (one wave64)
constexpr int AWY = 4;     // waves along n2
constexpr int ATXR = 4;    // n2 cells per thread
constexpr int AX = AWY * ATXR;
constexpr int AH = 5;
constexpr int ALZ = AZ + 2 * AH, ALX = AX + 2 * AH;

struct Ac3Args {
  int n1, n2, n3;
  long long s2, s3;          // strides of axes 2 and 3
  const float* p;            // current level
  float* pp;                 // previous level in, next level out
  const float* c;
  const float* kz; const float* kx; const float* ky;  // null = no sponge
  int mode;                  // 0 source pass, 1 receiver pass
  int sz, sx, sy;
  float amp;                 // source amplitude of this level (mode 0)
  int zr;
  float* rec_out;            // plane record of this level (mode 0), [i3][i2]
  const float* rec_in;       // plane injection of this level (mode 1)
  unsigned* smax;            // per-snapshot max-abs (float bits) (mode 0 snapshot level)
  const void* snap;          // source snapshot of this level (mode 1 image level)
  float snap_scale;
  int snap_bf16;
  float* img;                // image (mode 1 image level)
  int ychunk;
};

__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

// Halo cells of the (z, x) plane tile that are not owned by a thread of the
// block: 10 z-halo rows of the 16 interior columns, then the 10 x-halo
// columns over the full 74 rows.
constexpr int AHALO = 2 * AH * AX + 2 * AH * ALZ;
constexpr int ANH = (AHALO + 255) / 256;

__device__ __forceinline__ void halo_cell(int h, int& lx, int& lz) {
  if (h < 2 * AH * AX) {
    lx = AH + h / (2 * AH);
    const int r = h % (2 * AH);
    lz = r < AH ? r : AZ + r;
  } else {
    const int h2 = h - 2 * AH * AX;
    const int cc = h2 / ALZ;
    lx = cc < AH ? cc : AX + cc;
    lz = h2 % ALZ;
  }
}

// Global loads through address-space-1 pointers (GLOBAL, not FLAT,
// instructions: FLAT ones also count against the LDS wait counter).
template <class T>
__device__ __forceinline__ T g1(const T* p, int i) {
  return ((const __attribute__((address_space(1))) T*)p)[i];
}

// The march along y is software-pipelined: the global loads of plane y+1
// (its new queue entry, and the own-cell p_prev, c, snapshot and image
// operands) are issued before plane y is computed, so their latency hides
// behind the arithmetic of the current plane. The plane tile is
// double-buffered in LDS with one barrier per plane whose fences cover LDS
// only, so those loads stay in flight across it; they are issued with
// clamped indices and masked where the values are taken.
// The halo cells of a plane are loaded with the queue entry of the same
// plane, five planes before it is computed: that is when the neighbouring
// tiles load those cells as their own queue entries, so the second read of
// each line is an L2 hit rather than a second fetch. Each thread keeps its
// halo cells of the five planes ahead in a ring in LDS (slot = plane % 5)
// that only it reads.
template <bool IMG>
__global__ __launch_bounds__(256) void k_ac3_step(const Ac3Args a) {
  __shared__ float sh[2][ALX * ALZ];
  __shared__ float hring[5][ANH][256];
  const int nbz = (a.n1 - 10 + AZ - 1) / AZ;
  const int nbx = (a.n2 - 10 + AX - 1) / AX;
  // z-fastest tile order, renumbered so each XCD gets a contiguous range:
  // the z- and x-neighbours whose halos a tile reads share its L2
  int b = xcd_tile(blockIdx.x, gridDim.x);
  const int bz = b % nbz;
  b /= nbz;
  const int bx = b % nbx;
  const int by = b / nbx;
  const int z0 = AH + bz * AZ, x0 = AH + bx * AX;
  const int y0 = AH + by * a.ychunk;
  const int y1 = min(y0 + a.ychunk, a.n3 - AH);
  // ty is wave-uniform (a wave is one row of 64 threads), so values indexed
  // by it and by the plane come from scalar loads
  const int tz = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int tid = tz + AZ * ty;
  const int gz = z0 + tz;
  const bool zok = gz < a.n1 - AH;
  // element offsets of one field fit in 32 bits (the host checks n < 2^31)
  const int s2 = (int)a.s2, s3 = (int)a.s3;
  constexpr bool imaging = IMG;  // receiver pass at an imaging level
  // per-thread halo cells (plane independent): LDS slot (-1: none) and
  // in-plane offset (-1: outside the grid, staged as zero)
  int hoff[ANH], hlds[ANH];
#pragma unroll
  for (int k = 0; k < ANH; ++k) {
    const int h = tid + 256 * k;
    int lx = 0, lz = 0;
    hlds[k] = -1;
    hoff[k] = -1;
    if (h < AHALO) {
      halo_cell(h, lx, lz);
      hlds[k] = lx * ALZ + lz;
      const int gx = x0 - AH + lx, g1v = z0 - AH + lz;
      if (gx >= 0 && gx < a.n2 && g1v >= 0 && g1v < a.n1) hoff[k] = g1v + gx * s2;
    }
  }
  int coff[ATXR];  // own columns (-1: outside the interior)
#pragma unroll
  for (int j = 0; j < ATXR; ++j) {
    const int gx = x0 + ty * ATXR + j;
    coff[j] = (zok && gx < a.n2 - AH) ? gz + gx * s2 : -1;
  }
  // register queue: q[k][j] = p at plane y-5+k
  float q[11][ATXR];
#pragma unroll
  for (int k = 0; k < 10; ++k)
#pragma unroll
    for (int j = 0; j < ATXR; ++j) {
      const bool ok = coff[j] >= 0;
      const float v = g1(a.p, ok ? coff[j] + (y0 - AH + k) * s3 : 0);
      q[k][j] = ok ? v : 0.f;
    }
  float hv[ANH], ppv[ATXR], cv[ATXR], snv[ATXR], imv[ATXR], qn[ATXR];
  // raw loads of plane y (clamped indices; own-cell values are used only
  // where coff >= 0, queue and halo values are masked where taken)
  auto load_plane = [&](int y, float (&h)[ANH], float (&qv)[ATXR], float (&pv)[ATXR], float (&c)[ATXR],
                        float (&sn)[ATXR], float (&im)[ATXR]) {
    const int py = y * s3;
#pragma unroll
    for (int j = 0; j < ATXR; ++j) {
      const int o = coff[j] >= 0 ? coff[j] + py : 0;
      qv[j] = g1(a.p, coff[j] >= 0 ? o + AH * s3 : 0);
      pv[j] = g1(static_cast<const float*>(a.pp), o);
      c[j] = g1(a.c, o);
      sn[j] = 0.f;
      im[j] = 0.f;
      if (imaging) {
        sn[j] = a.snap_bf16 ? bf16_to_f(g1(static_cast<const unsigned short*>(a.snap), o)) * a.snap_scale
                            : g1(static_cast<const float*>(a.snap), o);
        im[j] = g1(static_cast<const float*>(a.img), o);
      }
    }
    // halo cells of plane y+5, with its queue entry
#pragma unroll
    for (int k = 0; k < ANH; ++k) h[k] = g1(a.p, hoff[k] >= 0 ? hoff[k] + py + AH * s3 : 0);
  };
  // halo cells of the first five planes into the ring
#pragma unroll
  for (int j = 0; j < AH; ++j) {
    const int py = (y0 + j) * s3;
#pragma unroll
    for (int k = 0; k < ANH; ++k) {
      const float v = g1(a.p, hoff[k] >= 0 ? hoff[k] + py : 0);
      hring[(y0 + j) % 5][k][tid] = hoff[k] >= 0 ? v : 0.f;
    }
  }
  if (y0 < y1) load_plane(y0, hv, qn, ppv, cv, snv, imv);
  float mx = 0.f;
  // sponge factors: kz of the thread's row (loop invariant), kx of the
  // wave's columns and ky of the plane (wave-uniform, scalar loads)
  const float kzv = (a.kz && zok) ? g1(a.kz, gz) : 1.f;
  auto sc = [](const float* p, int i) { return ((const __attribute__((address_space(4))) float*)p)[i]; };
  for (int y = y0; y < y1; ++y) {
    const int bb = (y - y0) & 1;
    float* const sp = sh[bb];
#pragma unroll
    for (int j = 0; j < ATXR; ++j) q[10][j] = coff[j] >= 0 ? qn[j] : 0.f;
#pragma unroll
    for (int j = 0; j < ATXR; ++j) sp[(ty * ATXR + j + AH) * ALZ + tz + AH] = q[5][j];
    // this plane's halo cells from the ring; the slot then takes plane y+5's
    const int slot = y % 5;
#pragma unroll
    for (int k = 0; k < ANH; ++k) {
      if (hlds[k] >= 0) sp[hlds[k]] = hring[slot][k][tid];
      hring[slot][k][tid] = hoff[k] >= 0 ? hv[k] : 0.f;
    }
    // the tile of plane y is complete; plane y-1's buffer is free again
    // once every wave has passed this barrier
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    // issue the next plane's loads before this plane's arithmetic
    float hv2[ANH], pp2[ATXR], c2[ATXR], sn2[ATXR], im2[ATXR];
    load_plane(y + 1 < y1 ? y + 1 : y, hv2, qn, pp2, c2, sn2, im2);
    const int py = y * s3;
    if (zok) {
      const float kyv = a.kz ? sc(a.ky, y) : 1.f;
#pragma unroll
      for (int j = 0; j < ATXR; ++j) {
        const int lxo = ty * ATXR + j;
        const int gx = x0 + lxo;
        if (coff[j] < 0) continue;
        const int p0 = (lxo + AH) * ALZ + tz + AH;
        float lap = 3.f * W0 * q[5][j];
        lap += W1 * ((sp[p0 + 1] + sp[p0 - 1]) + (sp[p0 + ALZ] + sp[p0 - ALZ]) + (q[6][j] + q[4][j]));
        lap += W2 * ((sp[p0 + 2] + sp[p0 - 2]) + (sp[p0 + 2 * ALZ] + sp[p0 - 2 * ALZ]) + (q[7][j] + q[3][j]));
        lap += W3 * ((sp[p0 + 3] + sp[p0 - 3]) + (sp[p0 + 3 * ALZ] + sp[p0 - 3 * ALZ]) + (q[8][j] + q[2][j]));
        lap += W4 * ((sp[p0 + 4] + sp[p0 - 4]) + (sp[p0 + 4 * ALZ] + sp[p0 - 4 * ALZ]) + (q[9][j] + q[1][j]));
        lap += W5 * ((sp[p0 + 5] + sp[p0 - 5]) + (sp[p0 + 5 * ALZ] + sp[p0 - 5 * ALZ]) + (q[10][j] + q[0][j]));
        const int gi = coff[j] + py;
        const float kk = a.kz ? kzv * sc(a.kx, gx) * kyv : 1.f;
        float v = kk * (2.f * q[5][j] - ppv[j] + cv[j] * lap);
        if (a.mode == 0) {
          if (gz == a.sz && gx == a.sx && y == a.sy) v += a.amp;
        } else if (a.rec_in && gz == a.zr) {
          v += a.rec_in[(long long)y * a.n2 + gx];
        }
        a.pp[gi] = v;
        if (a.mode == 0) {
          if (a.rec_out && gz == a.zr) a.rec_out[(long long)y * a.n2 + gx] = v;
          mx = fmaxf(mx, fabsf(v));
        } else if (imaging) {
          a.img[gi] = imv[j] + snv[j] * v;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 10; ++k)
#pragma unroll
      for (int j = 0; j < ATXR; ++j) q[k][j] = q[k + 1][j];
#pragma unroll
    for (int k = 0; k < ANH; ++k) hv[k] = hv2[k];
#pragma unroll
    for (int j = 0; j < ATXR; ++j) {
      ppv[j] = pp2[j];
      cv[j] = c2[j];
      snv[j] = sn2[j];
      imv[j] = im2[j];
    }
  }
  if (a.mode == 0 && a.smax) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if (tz == 0) atomicMax(a.smax, __float_as_uint(mx));
  }
}

// Snapshot of one level: bf16 with the level's max-abs scale, or fp32 copy.
__global__ void k_ac3_snap(const float* p, void* out, size_t n, const unsigned* smax, int bf16) {
  const float m = __uint_as_float(*smax);
  const float inv = m > 1e-30f ? 1.f / m : 0.f;  // subnormal maxima would overflow 1/m
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    if (bf16) {
      unsigned u = __float_as_uint(p[i] * inv);
      u += 0x7fffu + ((u >> 16) & 1u);
      static_cast<unsigned short*>(out)[i] = (unsigned short)(u >> 16);
    } else {
      static_cast<float*>(out)[i] = p[i];
    }
  }
}

__global__ void k_fill3(float* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

__global__ void k_sub_rec(float* a, const float* b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] -= b[i];
}

int fill0(float* p, size_t n, hipStream_t st) {
  size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_fill3, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
  HIPCHK(hipGetLastError());
  return OK;
}

// Sponge taper of W/abc-mod.f90:35-50 applied per axis (1 inside).
std::vector<float> taper(int n, int abc) {
  std::vector<float> k(n, 1.f);
  if (abc <= 0) return k;
  std::vector<float> f(abc + 5);
  for (int i = 1; i <= abc; ++i) {
    const float aux = (1.0f * (float)i) / (1.0f * (float)abc);
    f[i - 1] = std::exp(-0.09f * aux * aux);
  }
  for (int i = abc + 1; i <= abc + 5; ++i) f[i - 1] = f[abc - 1];
  const int abc2 = abc + 5;
  for (int i1 = 1; i1 <= abc2; ++i1) {
    const int top = i1 - 1, bot = n - abc2 + i1 - 1;
    if (top < n) k[top] *= f[abc2 - i1];
    if (bot >= 0 && bot < n) k[bot] *= f[i1 - 1];
  }
  return k;
}

}  // namespace

}  // namespace erm

using namespace erm;

struct ac3d {
  ac3d_cfg_t cfg{};
  size_t n = 0;
  long long plane = 0;
  DevBuf c, c_alt, k, f0, f1, rec, rec_tmp, snap, smax, img, amp_h_dummy;
  std::vector<float> amp;
  std::vector<float> scales;  // per snapshot (host copy after the source pass)
  hipStream_t st = nullptr;
  hipEvent_t ev[2] = {};
  double timings[6] = {0, 0, 0, 0, 0, 0};
  ~ac3d() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
};

static int ac3_launch(ac3d* a, const float* p, float* pp, const float* c, int mode, int level, bool record,
                      bool inject, unsigned* smax, const void* snap, float scale, float* img) {
  const ac3d_cfg_t& g = a->cfg;
  Ac3Args x{};
  x.n1 = g.n1; x.n2 = g.n2; x.n3 = g.n3;
  x.s2 = g.n1;
  x.s3 = (long long)g.n1 * g.n2;
  x.p = p; x.pp = pp; x.c = c;
  if (g.abc > 0) {
    x.kz = a->k.as<float>();
    x.kx = x.kz + g.n1;
    x.ky = x.kx + g.n2;
  }
  x.mode = mode;
  x.sz = g.sz; x.sx = g.sx; x.sy = g.sy;
  x.amp = mode == 0 ? a->amp[level - 1] : 0.f;
  x.zr = g.zr;
  x.rec_out = record ? a->rec.as<float>() + (long long)(level - 1) * a->plane : nullptr;
  x.rec_in = inject ? a->rec.as<float>() + (long long)(level - 1) * a->plane : nullptr;
  x.smax = smax;
  x.snap = snap;
  x.snap_scale = scale;
  x.snap_bf16 = g.snap_prec == 1;
  x.img = img;
  // planes per workgroup along y (RTM_AC3_YCHUNK overrides; experiments)
  static const int ychunk = [] {
    const char* e = getenv("RTM_AC3_YCHUNK");
    const int v = e ? atoi(e) : 64;
    return v >= 8 ? v : 64;
  }();
  x.ychunk = ychunk;
  const int nbz = (g.n1 - 10 + AZ - 1) / AZ, nbx = (g.n2 - 10 + AX - 1) / AX;
  const int nby = (g.n3 - 10 + x.ychunk - 1) / x.ychunk;
  if (mode == 1 && img)
    hipLaunchKernelGGL(k_ac3_step<true>, dim3(nbz * nbx * nby), dim3(AZ, AWY), 0, a->st, x);
  else
    hipLaunchKernelGGL(k_ac3_step<false>, dim3(nbz * nbx * nby), dim3(AZ, AWY), 0, a->st, x);
  HIPCHK(hipGetLastError());
  return OK;
}

// Source pass: levels 1..nt with the model c; records the receiver plane
// (subtracting when sub) and, with snapshots, stores every j-th level.
static int ac3_source_pass(ac3d* a, const float* c, int record, int sub, int snaps) {
  const ac3d_cfg_t& g = a->cfg;
  float* p = a->f0.as<float>();
  float* pp = a->f1.as<float>();
  STCHK(fill0(p, a->n, a->st));
  STCHK(fill0(pp, a->n, a->st));
  float* rec_dst = nullptr;
  if (record) {
    if (sub) {
      std::swap(a->rec, a->rec_tmp);  // record into the scratch, subtract afterwards
    }
    STCHK(fill0(a->rec.as<float>(), (size_t)g.nt * a->plane, a->st));
    rec_dst = a->rec.as<float>();
  }
  const int j = g.snap_every;
  const int nsnap = g.nt / j;
  if (snaps) HIPCHK(hipMemsetAsync(a->smax.p, 0, sizeof(unsigned) * std::max(nsnap, 1), a->st));
  for (int t = 1; t <= g.nt; ++t) {
    const bool sl = snaps && t % j == 0;
    unsigned* sm = sl ? a->smax.as<unsigned>() + (t / j - 1) : nullptr;
    STCHK(ac3_launch(a, p, pp, c, 0, t, record != 0, false, sm, nullptr, 0.f, nullptr));
    if (sl) {
      const size_t esz = g.snap_prec == 1 ? 2 : 4;
      char* dst = a->snap.as<char>() + (size_t)(t / j - 1) * a->n * esz;
      size_t blocks = std::min<size_t>((a->n + 255) / 256, 65536);
      hipLaunchKernelGGL(k_ac3_snap, dim3((unsigned)blocks), dim3(256), 0, a->st, pp, (void*)dst, a->n, sm,
                         g.snap_prec == 1);
      HIPCHK(hipGetLastError());
    }
    std::swap(p, pp);
  }
  (void)rec_dst;
  if (record && sub) {
    std::swap(a->rec, a->rec_tmp);  // a->rec is the first record again
    size_t cnt = (size_t)g.nt * a->plane;
    size_t blocks = std::min<size_t>((cnt + 255) / 256, 65536);
    hipLaunchKernelGGL(k_sub_rec, dim3((unsigned)blocks), dim3(256), 0, a->st, a->rec.as<float>(),
                       a->rec_tmp.as<float>(), cnt);
    HIPCHK(hipGetLastError());
  }
  return OK;
}

RTM_EXPORT int ac3d_create(ac3d** out, const ac3d_cfg_t* cfg, const float* c, const float* amp) {
  if (!out || !cfg || !c || !amp) { set_error("null argument"); return E_ARG; }
  const ac3d_cfg_t& g = *cfg;
  if (g.n1 < 11 || g.n2 < 11 || g.n3 < 11 || g.nt < 1 || g.snap_every < 1) {
    set_error("ac3d: grid must exceed the 5-cell halo on every axis; nt, snap_every >= 1");
    return E_ARG;
  }
  if (g.sz < 5 || g.sz >= g.n1 - 5 || g.sx < 5 || g.sx >= g.n2 - 5 || g.sy < 5 || g.sy >= g.n3 - 5 ||
      g.zr < 5 || g.zr >= g.n1 - 5) {
    set_error("ac3d: source or receiver plane outside the interior");
    return E_ARG;
  }
  if (g.abc > 0 && (2 * (g.abc + 5) > g.n1 || 2 * (g.abc + 5) > g.n2 || 2 * (g.abc + 5) > g.n3)) {
    set_error("ac3d: sponge does not fit");
    return E_ARG;
  }
  STCHK(ensure_device());
  std::unique_ptr<ac3d> a(new ac3d());
  a->cfg = g;
  a->n = (size_t)g.n1 * g.n2 * g.n3;
  if (a->n >= (size_t)1 << 31) {  // the step kernel addresses a field with 32-bit offsets
    set_error("3D grid of %zu cells exceeds 2^31 cells per field", a->n);
    return E_ARG;
  }
  a->plane = (long long)g.n2 * g.n3;
  HIPCHK(hipStreamCreateWithFlags(&a->st, hipStreamNonBlocking));
  for (auto& e : a->ev) HIPCHK(hipEventCreate(&e));
  STCHK(a->c.alloc(a->n * 4, "velocity term"));
  HIPCHK(hipMemcpy(a->c.p, c, a->n * 4, hipMemcpyHostToDevice));
  STCHK(a->k.alloc((size_t)(g.n1 + g.n2 + g.n3) * 4, "sponge"));
  std::vector<float> k1 = taper(g.n1, g.abc), k2 = taper(g.n2, g.abc), k3 = taper(g.n3, g.abc);
  HIPCHK(hipMemcpy(a->k.p, k1.data(), g.n1 * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a->k.as<float>() + g.n1, k2.data(), g.n2 * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a->k.as<float>() + g.n1 + g.n2, k3.data(), g.n3 * 4, hipMemcpyHostToDevice));
  STCHK(a->f0.alloc(a->n * 4, "wavefield"));
  STCHK(a->f1.alloc(a->n * 4, "wavefield"));
  STCHK(a->rec.alloc((size_t)g.nt * a->plane * 4, "receiver data"));
  STCHK(fill0(a->rec.as<float>(), (size_t)g.nt * a->plane, a->st));
  STCHK(a->img.alloc(a->n * 4, "image"));
  const int nsnap = std::max(g.nt / g.snap_every, 1);
  STCHK(a->snap.alloc((size_t)nsnap * a->n * (g.snap_prec == 1 ? 2 : 4), "snapshots"));
  STCHK(a->smax.alloc(sizeof(unsigned) * nsnap, "snapshot scales"));
  a->amp.assign(amp, amp + g.nt);
  HIPCHK(hipStreamSynchronize(a->st));
  *out = a.release();
  return OK;
}

RTM_EXPORT int ac3d_model_data(ac3d* a, const float* c_true) {
  if (!a) { set_error("null handle"); return E_ARG; }
  // observed = data(true model) - data(migration model) when c_true is given,
  // else the migration-model data
  HIPCHK(hipEventRecord(a->ev[0], a->st));
  if (c_true) {
    STCHK(a->c_alt.alloc(a->n * 4, "true model"));
    HIPCHK(hipMemcpy(a->c_alt.p, c_true, a->n * 4, hipMemcpyHostToDevice));
    if (!a->rec_tmp.p) STCHK(a->rec_tmp.alloc((size_t)a->cfg.nt * a->plane * 4, "receiver scratch"));
    STCHK(ac3_source_pass(a, a->c_alt.as<float>(), 1, 0, 0));
    STCHK(ac3_source_pass(a, a->c.as<float>(), 1, 1, 0));
    a->c_alt.release();
  } else {
    STCHK(ac3_source_pass(a, a->c.as<float>(), 1, 0, 0));
  }
  HIPCHK(hipEventRecord(a->ev[1], a->st));
  HIPCHK(hipStreamSynchronize(a->st));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, a->ev[0], a->ev[1]));
  a->timings[4] = ms;
  return OK;
}

RTM_EXPORT int ac3d_migrate(ac3d* a) {
  if (!a) { set_error("null handle"); return E_ARG; }
  const ac3d_cfg_t& g = a->cfg;
  HIPCHK(hipEventRecord(a->ev[0], a->st));
  STCHK(ac3_source_pass(a, a->c.as<float>(), 0, 0, 1));
  HIPCHK(hipEventRecord(a->ev[1], a->st));
  HIPCHK(hipStreamSynchronize(a->st));
  float ms_src = 0;
  HIPCHK(hipEventElapsedTime(&ms_src, a->ev[0], a->ev[1]));
  const int j = g.snap_every;
  const int nsnap = g.nt / j;
  std::vector<unsigned> sm(std::max(nsnap, 1));
  HIPCHK(hipMemcpy(sm.data(), a->smax.p, sizeof(unsigned) * sm.size(), hipMemcpyDeviceToHost));
  a->scales.resize(sm.size());
  for (size_t i = 0; i < sm.size(); ++i) {
    float f;
    std::memcpy(&f, &sm[i], 4);
    a->scales[i] = f;
  }
  // receiver pass, levels nt..1
  float* q = a->f0.as<float>();
  float* qq = a->f1.as<float>();
  STCHK(fill0(q, a->n, a->st));
  STCHK(fill0(qq, a->n, a->st));
  STCHK(fill0(a->img.as<float>(), a->n, a->st));
  HIPCHK(hipEventRecord(a->ev[0], a->st));
  const size_t esz = g.snap_prec == 1 ? 2 : 4;
  for (int t = g.nt; t >= 1; --t) {
    const bool il = t % j == 0;
    const void* sp = il ? a->snap.as<char>() + (size_t)(t / j - 1) * a->n * esz : nullptr;
    const float sc = il ? a->scales[t / j - 1] : 0.f;
    STCHK(ac3_launch(a, q, qq, a->c.as<float>(), 1, t, false, true, nullptr, sp, sc,
                     il ? a->img.as<float>() : nullptr));
    std::swap(q, qq);
  }
  HIPCHK(hipEventRecord(a->ev[1], a->st));
  HIPCHK(hipStreamSynchronize(a->st));
  float ms_rcv = 0;
  HIPCHK(hipEventElapsedTime(&ms_rcv, a->ev[0], a->ev[1]));
  a->timings[0] = ms_src;
  a->timings[1] = g.nt;
  a->timings[2] = ms_rcv;
  a->timings[3] = g.nt;
  return OK;
}

RTM_EXPORT int ac3d_set_data(ac3d* a, const float* rec) {
  if (!a || !rec) { set_error("null argument"); return E_ARG; }
  HIPCHK(hipMemcpy(a->rec.p, rec, (size_t)a->cfg.nt * a->plane * 4, hipMemcpyHostToDevice));
  return OK;
}

RTM_EXPORT int ac3d_get_data(ac3d* a, float* rec) {
  if (!a || !rec) { set_error("null argument"); return E_ARG; }
  HIPCHK(hipStreamSynchronize(a->st));
  HIPCHK(hipMemcpy(rec, a->rec.p, (size_t)a->cfg.nt * a->plane * 4, hipMemcpyDeviceToHost));
  return OK;
}

RTM_EXPORT int ac3d_get_image(ac3d* a, float* img) {
  if (!a || !img) { set_error("null argument"); return E_ARG; }
  HIPCHK(hipStreamSynchronize(a->st));
  HIPCHK(hipMemcpy(img, a->img.p, a->n * 4, hipMemcpyDeviceToHost));
  return OK;
}

RTM_EXPORT int ac3d_get_snapshot_scales(ac3d* a, float* out, int n) {
  if (!a || !out) { set_error("null argument"); return E_ARG; }
  for (int i = 0; i < n && i < (int)a->scales.size(); ++i) out[i] = a->scales[i];
  return OK;
}

RTM_EXPORT int ac3d_timings(ac3d* a, double* out6) {
  if (!a || !out6) { set_error("null argument"); return E_ARG; }
  for (int i = 0; i < 6; ++i) out6[i] = a->timings[i];
  return OK;
}

RTM_EXPORT int ac3d_destroy(ac3d* a) {
  delete a;
  return OK;
}
