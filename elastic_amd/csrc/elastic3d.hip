// elastic3d.hip -- 9-field 3D elastic adjoint time step (gfx950).
//
// The reference's 3D elastic kernel test (K3/rtmlib-modC.c:277-589 CPU loop,
// K3/rtmlib-modC-gpu.cu:374-690 CUDA kernel) advances the adjoint of the
// staggered-grid velocity-stress system by one level:
//   out = out + in1 + sum of 10-tap gathers of coefficient * field
// over 9 fields (Vy, Vx, Vz, syy, sxx, szz, sxy, sxz, syz) and 6 coefficient
// arrays (m1_y, m1_x, m1_z, m2, m3, m2m3), every array padded by 5 cells per
// face, (nz+10) x (nx+10) x (ny+10), z fastest. Each output is a sum of
// 1 to 5 gathers along one axis with tap set A (offsets -5..4) or B (-4..5).
//
// Here the gathers of one output along one axis are merged: the products
// coefficient * field that share an axis and a tap set are summed first
// (e.g. m2m3 sxx + m2 syy + m2 szz for Vx along x), which leaves 18 gathers
// over 9 distinct product arrays. A workgroup of 64 x EW threads (EW = 8 or
// 16 waves) owns a 64 (z) x EW (x) tile of the (z, x) plane and marches
// along y over a chunk of planes:
//   * the 6 products gathered along x, over the tile plus its 5-column
//     x-halo, and the 6 gathered along z, over the tile plus its 5-row
//     z-halo, go to LDS;
//   * the 6 products gathered along y live in an 11-deep register queue of
//     the thread's own column (planes y-5 .. y+5), refilled 5 planes ahead.
// Each step moves 168 algorithmic bytes per cell (in, in1, out and the 6
// coefficients read, out written): HBM-bound.

#include "../../include/elastic3d.h"
#include "common.hpp"
#include "elastic2d.hpp"
#include "engine2d.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace erm {

namespace {

enum { VY, VX, VZ, SYY, SXX, SZZ, SXY, SXZ, SYZ };
enum { M1Y, M1X, M1Z, M2, M3, M2M3 };

constexpr int EZ = 64;               // tile rows (z): one wave64
constexpr int EH = 5;                // stencil radius
constexpr int ELZ = EZ + 2 * EH;     // z-band rows (tile + z-halo)
constexpr int NXH = 2 * EH * EZ;     // x-halo cells (10 columns x 64 rows)

struct El3Args {
  int ny, nz, nx;
  int sx, sy;  // element strides of x and y (z is 1)
  const float* in[9];
  const float* in1[9];
  float* out[9];
  const float* m[6];
  int nbz, nbx, ych, nblk;
  int xcd;
};

template <class T>
__device__ __forceinline__ T gl(const T* p, int i) {
  return ((const __attribute__((address_space(1))) T*)p)[i];
}

// The 9 distinct products of one cell that the 18 gathers read:
//   PX = m2m3 sxx + m2 (syy + szz)  (x, set A, Vx)   Pxy = m3 sxy (x: Vy, y: Vx)
//   PY = m2m3 syy + m2 (szz + sxx)  (y, set A, Vy)   Pxz = m3 sxz (x: Vz, z: Vx)
//   PZ = m2m3 szz + m2 (sxx + syy)  (z, set A, Vz)   Pyz = m3 syz (z: Vy, y: Vz)
//   Qx = m1_x Vx (x: sxx, y: sxy, z: sxz)   Qy = m1_y Vy (x: sxy, y: syy, z: syz)
//   Qz = m1_z Vz (x: sxz, y: syz, z: szz); every other gather uses set B.
enum { PX, PY, PZ, PXY, PXZ, PYZ, QX, QY, QZ };

__device__ __forceinline__ void products(const El3Args& a, int gi, float p[9]) {
  const float vy = gl(a.in[VY], gi), vx = gl(a.in[VX], gi), vz = gl(a.in[VZ], gi);
  const float syy = gl(a.in[SYY], gi), sxx = gl(a.in[SXX], gi), szz = gl(a.in[SZZ], gi);
  const float sxy = gl(a.in[SXY], gi), sxz = gl(a.in[SXZ], gi), syz = gl(a.in[SYZ], gi);
  const float m1y = gl(a.m[M1Y], gi), m1x = gl(a.m[M1X], gi), m1z = gl(a.m[M1Z], gi);
  const float m2 = gl(a.m[M2], gi), m3 = gl(a.m[M3], gi), m2m3 = gl(a.m[M2M3], gi);
  p[PX] = fmaf(m2m3, sxx, m2 * (syy + szz));
  p[PY] = fmaf(m2m3, syy, m2 * (szz + sxx));
  p[PZ] = fmaf(m2m3, szz, m2 * (sxx + syy));
  p[PXY] = m3 * sxy;
  p[PXZ] = m3 * sxz;
  p[PYZ] = m3 * syz;
  p[QX] = m1x * vx;
  p[QY] = m1y * vy;
  p[QZ] = m1z * vz;
}

// LDS slots of the x-band (gathers along x) and the z-band (along z).
constexpr int XB[6] = {PX, PXY, PXZ, QX, QY, QZ};
constexpr int ZB[6] = {PXZ, PYZ, PZ, QZ, QX, QY};
// register queue along y: planes y-5 .. y+5 of these products
constexpr int YQ[6] = {PY, PXY, PYZ, QX, QY, QZ};
// the own cell's x/z products that are not in the y-queue wait 5 planes
constexpr int DL[3] = {PX, PZ, PXZ};

// Gathers at cell i of products P along one axis, P(k) = product at i + k:
//   set A: sum_d c_(6-d) (P(-d) - P(d-1))     set B: sum_d c_(6-d) (P(1-d) - P(d))
template <class F>
__device__ __forceinline__ float gather_a(F P) {
  float s = C5 * (P(-1) - P(0));
  s = fmaf(C4, P(-2) - P(1), s);
  s = fmaf(C3, P(-3) - P(2), s);
  s = fmaf(C2, P(-4) - P(3), s);
  return fmaf(C1, P(-5) - P(4), s);
}
template <class F>
__device__ __forceinline__ float gather_b(F P) {
  float s = C5 * (P(0) - P(1));
  s = fmaf(C4, P(-1) - P(2), s);
  s = fmaf(C3, P(-2) - P(3), s);
  s = fmaf(C2, P(-3) - P(4), s);
  return fmaf(C1, P(-4) - P(5), s);
}

// Workgroup barrier whose fences cover LDS only: global loads issued before
// it stay in flight.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// EW waves per workgroup, one tile column (x) each: a 64 x EW tile.
// Every input column is loaded once per plane by its owner, 5 planes ahead
// (plane y+5): its y-products enter the queue and its other x/z products a
// 5-plane delay line, so plane y's own cells reach LDS from registers; only
// the x- and z-halo cells are loaded at plane y.
template <int EW>
__global__ __launch_bounds__(EZ * EW) __attribute__((amdgpu_waves_per_eu(4)))
void k_el3_adj(const El3Args a) {
  constexpr int ET = EZ * EW, ELX = EW + 2 * EH, NZH = 2 * EH * EW;
  __shared__ float xs[6][ELX][EZ];   // x-band: [k][x0-5+lx][z0+lz]
  __shared__ float zs[6][EW][ELZ];   // z-band: [k][x0+lx][z0-5+lz]
  const int b = a.xcd ? xcd_tile(blockIdx.x, a.nblk) : (int)blockIdx.x;
  const int bz = b % a.nbz;
  const int r = b / a.nbz;
  const int bx = r % a.nbx;
  const int yc = r / a.nbx;
  const int z0 = bz * EZ, x0 = bx * EW, y0 = yc * a.ych;
  const int y1 = min(y0 + a.ych, a.ny);
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int gz = z0 + lane, gx = x0 + w;
  const bool own = gz < a.nz && gx < a.nx;
  // the own column's products are needed (as LDS halo of valid cells) up to
  // nz+5 / nx+5; beyond that the padded array ends
  const bool ownp = gz < a.nz + EH && gx < a.nx + EH;
  const int sx = a.sx, sy = a.sy;
  auto at = [&](int iz, int ix, int iy) { return (iz + EH) + sx * (ix + EH) + sy * (iy + EH); };

  float q[6][11];  // YQ products, slot j = plane y-5+j
  float dl[3][6];  // DL products, slot j = plane y+j
  // prime planes y0-5 .. y0+4
#pragma unroll
  for (int j = 0; j < 10; ++j) {
    float p[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (ownp) products(a, at(gz, gx, y0 - 5 + j), p);
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k][j] = p[YQ[k]];
    if (j >= 5) {
#pragma unroll
      for (int k = 0; k < 3; ++k) dl[k][j - 5] = p[DL[k]];
    }
  }

  for (int y = y0; y < y1; ++y) {
    // own column, plane y+5
    {
      float p[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (ownp) products(a, at(gz, gx, y + 5), p);
#pragma unroll
      for (int k = 0; k < 6; ++k) q[k][10] = p[YQ[k]];
#pragma unroll
      for (int k = 0; k < 3; ++k) dl[k][5] = p[DL[k]];
    }
    // own cells of plane y into LDS, from the queue (slot 5) and the delay line (slot 0)
    {
      float c[9];
#pragma unroll
      for (int k = 0; k < 6; ++k) c[YQ[k]] = q[k][5];
#pragma unroll
      for (int k = 0; k < 3; ++k) c[DL[k]] = dl[k][0];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        xs[k][w + EH][lane] = c[XB[k]];
        zs[k][w][lane + EH] = c[ZB[k]];
      }
    }
    // x-halo: columns x0-5..x0-1 and x0+EW..x0+EW+4, 64 rows
#pragma unroll
    for (int h0 = 0; h0 < NXH; h0 += ET) {
      const int h = h0 + t;
      if (h < NXH) {
        const int hc = h >> 6;
        const int lx = hc < EH ? hc : hc + EW;
        const int hx = x0 - EH + lx, hz = z0 + (h & 63);
        float p[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (hz < a.nz + EH && hx < a.nx + EH) products(a, at(hz, hx, y), p);
#pragma unroll
        for (int k = 0; k < 6; ++k) xs[k][lx][h & 63] = p[XB[k]];
      }
    }
    // z-halo: rows z0-5..z0-1 and z0+64..z0+68 of the EW columns
    if (t < NZH) {
      const int rr = t % (2 * EH), cc = t / (2 * EH);
      const int lz = rr < EH ? rr : rr + EZ;
      const int hz = z0 - EH + lz, hx = x0 + cc;
      float p[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (hz < a.nz + EH && hx < a.nx + EH) products(a, at(hz, hx, y), p);
#pragma unroll
      for (int k = 0; k < 6; ++k) zs[k][cc][lz] = p[ZB[k]];
    }
    lds_barrier();

    if (own) {
      const int gi = at(gz, gx, y);
      float o[9];
#pragma unroll
      for (int f = 0; f < 9; ++f) o[f] = gl(a.out[f], gi) + gl(a.in1[f], gi);
      const int cx = w + EH, cz = lane + EH;
      // XB / ZB / YQ slot of each product
      auto X = [&](int k) { return [&, k](int d) { return xs[k][cx + d][lane]; }; };
      auto Z = [&](int k) { return [&, k](int d) { return zs[k][w][cz + d]; }; };
      auto Y = [&](int k) { return [&, k](int d) { return q[k][5 + d]; }; };
      // XB = {PX, PXY, PXZ, QX, QY, QZ}, ZB = {PXZ, PYZ, PZ, QZ, QX, QY},
      // YQ = {PY, PXY, PYZ, QX, QY, QZ}
      o[VX] += gather_a(X(0)) + gather_b(Y(1)) + gather_b(Z(0));
      o[VY] += gather_a(Y(0)) + gather_b(Z(1)) + gather_b(X(1));
      o[VZ] += gather_a(Z(2)) + gather_b(Y(2)) + gather_b(X(2));
      o[SXX] += gather_b(X(3));
      o[SYY] += gather_b(Y(4));
      o[SZZ] += gather_b(Z(3));
      o[SXY] += gather_b(X(4)) + gather_b(Y(3));
      o[SXZ] += gather_b(Z(4)) + gather_b(X(5));
      o[SYZ] += gather_b(Z(5)) + gather_b(Y(5));
#pragma unroll
      for (int f = 0; f < 9; ++f) a.out[f][gi] = o[f];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
      for (int j = 0; j < 10; ++j) q[k][j] = q[k][j + 1];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 5; ++j) dl[k][j] = dl[k][j + 1];
    lds_barrier();
  }
}

// tile width in x (= waves per workgroup): RTM_EL3_W = 4, 8 or 16
int default_el3_width() {
  const char* s = std::getenv("RTM_EL3_W");
  const int v = s ? std::atoi(s) : 8;
  return (v == 4 || v == 16) ? v : 8;
}

int default_el3_ychunk(int ny, int tiles) {
  if (const char* s = std::getenv("RTM_EL3_YCHUNK")) {
    const int v = std::atoi(s);
    if (v > 0) return std::min(v, ny);
  }
  // at least 256 workgroups, 10 to 40 planes per chunk (a chunk primes its
  // queue with 10 planes; 40 measured best on 200^3 and 400^3, 6% and 3%
  // faster than 20)
  const int chunks = std::max(1, (256 + tiles - 1) / tiles);
  return std::min(ny, std::min(40, std::max(10, (ny + chunks - 1) / chunks)));
}

struct El3Timing {
  double ms = 0, launches = 0, cells = 0, ych = 0;
};
thread_local El3Timing g_el3_timing;

}  // namespace

// nsteps with the reference's rotation (K3/rtmlib-modC-gpu.cu:273-372):
// after a step that is not the last, (out, in, in1) <- (in1, out, in).
static int el3_steps(int ny, int nz, int nx, float* in[9], float* in1[9], float* out[9], const float* const m[6],
                     int nsteps, hipStream_t st) {
  if (ny < 1 || nz < 1 || nx < 1 || nsteps < 0) {
    set_error("elastic3d: grid %d x %d x %d or %d steps out of range", ny, nz, nx, nsteps);
    return E_ARG;
  }
  const long long n = (long long)(nz + 10) * (nx + 10) * (ny + 10);
  if (n >= (1LL << 31)) {
    set_error("elastic3d: %lld padded cells per field exceed 2^31", n);
    return E_ARG;
  }
  for (int f = 0; f < 9; ++f)
    if (!in[f] || !in1[f] || !out[f]) { set_error("elastic3d: null field pointer"); return E_ARG; }
  for (int k = 0; k < 6; ++k)
    if (!m[k]) { set_error("elastic3d: null coefficient pointer"); return E_ARG; }
  El3Args a{};
  a.ny = ny; a.nz = nz; a.nx = nx;
  a.sx = nz + 10;
  a.sy = (nz + 10) * (nx + 10);
  for (int k = 0; k < 6; ++k) a.m[k] = m[k];
  const int ew = default_el3_width();
  a.nbz = (nz + EZ - 1) / EZ;
  a.nbx = (nx + ew - 1) / ew;
  a.ych = default_el3_ychunk(ny, a.nbz * a.nbx);
  const int nyc = (ny + a.ych - 1) / a.ych;
  a.nblk = a.nbz * a.nbx * nyc;
  a.xcd = default_xcd();
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
  } eg{e0, e1};
  HIPCHK(hipEventRecord(e0, st));
  for (int s = 0; s < nsteps; ++s) {
    for (int f = 0; f < 9; ++f) {
      a.in[f] = in[f];
      a.in1[f] = in1[f];
      a.out[f] = out[f];
    }
    if (ew == 16)
      hipLaunchKernelGGL(k_el3_adj<16>, dim3(a.nblk), dim3(EZ * 16), 0, st, a);
    else if (ew == 4)
      hipLaunchKernelGGL(k_el3_adj<4>, dim3(a.nblk), dim3(EZ * 4), 0, st, a);
    else
      hipLaunchKernelGGL(k_el3_adj<8>, dim3(a.nblk), dim3(EZ * 8), 0, st, a);
    HIPCHK(hipGetLastError());
    if (s < nsteps - 1) {
      for (int f = 0; f < 9; ++f) {
        float* t = out[f];
        out[f] = in1[f];
        in1[f] = in[f];
        in[f] = t;
      }
    }
  }
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  g_el3_timing.ms = ms;
  g_el3_timing.launches = nsteps;
  g_el3_timing.cells = (double)nz * nx * ny;
  g_el3_timing.ych = a.ych;
  return OK;
}

}  // namespace erm

using namespace erm;

RTM_EXPORT int rtm3d_adj_steps_dev(int ny, int nz, int nx, float* in[9], float* in1[9], float* out[9],
                                   const float* const coef[6], int nsteps) {
  if (!in || !in1 || !out || !coef) { set_error("elastic3d: null pointer array"); return E_ARG; }
  STCHK(ensure_device());
  return el3_steps(ny, nz, nx, in, in1, out, coef, nsteps, nullptr);
}

RTM_EXPORT int rtm3d_timings(double* out4) {
  if (!out4) { set_error("null argument"); return E_ARG; }
  out4[0] = g_el3_timing.ms;
  out4[1] = g_el3_timing.launches;
  out4[2] = g_el3_timing.cells;
  out4[3] = g_el3_timing.ych;
  return OK;
}

// --- the reference's host interface (K3/gpu.h:39-50) ------------------------
namespace {
struct El3Ctx {
  int ny = 0, nz = 0, nx = 0;
  size_t n = 0;
  DevBuf buf;  // 27 fields (in, in1, out) then 6 coefficients
};
std::mutex g_el3_mu;
El3Ctx* g_el3 = nullptr;
}  // namespace

RTM_EXPORT int rtm3d_gpu_init(int ny, int nz, int nx) {
  std::lock_guard<std::mutex> lk(g_el3_mu);
  if (ny < 1 || nz < 1 || nx < 1) { set_error("elastic3d: grid %d x %d x %d out of range", ny, nz, nx); return E_ARG; }
  STCHK(ensure_device());
  delete g_el3;
  g_el3 = new El3Ctx;
  g_el3->ny = ny; g_el3->nz = nz; g_el3->nx = nx;
  g_el3->n = (size_t)(nz + 10) * (nx + 10) * (ny + 10);
  const int s = g_el3->buf.alloc(33 * g_el3->n * sizeof(float), "3D elastic fields");
  if (s != OK) { delete g_el3; g_el3 = nullptr; }
  return s;
}

RTM_EXPORT int rtm3d_gpu_final(void) {
  std::lock_guard<std::mutex> lk(g_el3_mu);
  delete g_el3;
  g_el3 = nullptr;
  return OK;
}

// Copies the 27 fields and 6 coefficients in, runs one step (the reference's
// Steps_write_back = 1, K3/gpu.h:11), copies the 9 out fields back.
// gpu_kernel_time (may be null): seconds of copy-in, compute, copy-out.
RTM_EXPORT int rtm3d_gpu_func(int ny, int nz, int nx, float* Vy0_in, float* Vx0_in, float* Vz0_in, float* syy0_in,
                              float* sxx0_in, float* szz0_in, float* sxy0_in, float* sxz0_in, float* syz0_in,
                              float* Vy0_in1, float* Vx0_in1, float* Vz0_in1, float* syy0_in1, float* sxx0_in1,
                              float* szz0_in1, float* sxy0_in1, float* sxz0_in1, float* syz0_in1, float* Vy0_out,
                              float* Vx0_out, float* Vz0_out, float* syy0_out, float* sxx0_out, float* szz0_out,
                              float* sxy0_out, float* sxz0_out, float* syz0_out, float* m1_y, float* m1_x,
                              float* m1_z, float* m2, float* m3, float* m2m3, float* debug,
                              float* gpu_kernel_time) {
  (void)debug;
  std::lock_guard<std::mutex> lk(g_el3_mu);
  if (!g_el3) { set_error("elastic3d: rtm3d_gpu_init was not called"); return E_ARG; }
  if (g_el3->ny != ny || g_el3->nz != nz || g_el3->nx != nx) {
    set_error("elastic3d: grid %d x %d x %d differs from rtm3d_gpu_init's %d x %d x %d", ny, nz, nx, g_el3->ny,
              g_el3->nz, g_el3->nx);
    return E_ARG;
  }
  float* host[33] = {Vy0_in,  Vx0_in,  Vz0_in,  syy0_in,  sxx0_in,  szz0_in,  sxy0_in,  sxz0_in,  syz0_in,
                     Vy0_in1, Vx0_in1, Vz0_in1, syy0_in1, sxx0_in1, szz0_in1, sxy0_in1, sxz0_in1, syz0_in1,
                     Vy0_out, Vx0_out, Vz0_out, syy0_out, sxx0_out, szz0_out, sxy0_out, sxz0_out, syz0_out,
                     m1_y,    m1_x,    m1_z,    m2,       m3,       m2m3};
  for (float* p : host)
    if (!p) { set_error("elastic3d: null array argument"); return E_ARG; }
  const size_t n = g_el3->n;
  float* d = g_el3->buf.as<float>();
  hipEvent_t ev[4];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() { for (int i = 0; i < 4; ++i) (void)hipEventDestroy(e[i]); }
  } eg{ev};
  HIPCHK(hipEventRecord(ev[0], nullptr));
  for (int k = 0; k < 33; ++k)
    HIPCHK(hipMemcpyAsync(d + k * n, host[k], n * sizeof(float), hipMemcpyHostToDevice, nullptr));
  HIPCHK(hipEventRecord(ev[1], nullptr));
  float *in[9], *in1[9], *out[9];
  const float* m[6];
  for (int f = 0; f < 9; ++f) {
    in[f] = d + f * n;
    in1[f] = d + (9 + f) * n;
    out[f] = d + (18 + f) * n;
  }
  for (int k = 0; k < 6; ++k) m[k] = d + (27 + k) * n;
  STCHK(el3_steps(ny, nz, nx, in, in1, out, m, 1, nullptr));
  HIPCHK(hipEventRecord(ev[2], nullptr));
  for (int f = 0; f < 9; ++f)
    HIPCHK(hipMemcpyAsync(host[18 + f], out[f], n * sizeof(float), hipMemcpyDeviceToHost, nullptr));
  HIPCHK(hipEventRecord(ev[3], nullptr));
  HIPCHK(hipEventSynchronize(ev[3]));
  if (gpu_kernel_time) {
    for (int i = 0; i < 3; ++i) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      gpu_kernel_time[i] = ms / 1000.f;
    }
  }
  return OK;
}
