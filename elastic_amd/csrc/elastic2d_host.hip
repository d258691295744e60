// elastic2d_host.hip -- C ABI of the 2D elastic RTM library (host side).
//
// Entry points are declared in include/rtmlib_modC.h with the reference
// interface each one replaces. All device work is queued on HIP streams;
// nothing here falls back to the CPU.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/rtmlib_modC.h"
#include "engine2d.hpp"

namespace erm {

static thread_local char g_err[1024] = "";
static thread_local double g_timings[8] = {0, 0, 0, 0, 0, 0, 0, 0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* last_error() { return g_err; }

static void store_timing(const Timing& t) {
  g_timings[0] = t.fwd_ms;
  g_timings[1] = (double)t.fwd_launches;
  g_timings[2] = t.adj_ms;
  g_timings[3] = (double)t.adj_launches;
  g_timings[4] = t.total_ms;
  g_timings[5] = t.fwd_cells;
  g_timings[6] = t.adj_cells;
}

static int snap_fmt_of(int prec, int every) {
  if (prec == 1) return SNAP_BF16;
  if (prec == 2) return SNAP_F16;
  return every > 1 ? SNAP_F32 : SNAP_HIST;
}

// ---------------------------------------------------------------------------
// Drop-in adjoint on full (nz,nx,nt) histories, W/rtmlib-modC.c:120-253.
static int op3(int nt, int nz, int nx, float* const* hist, const float* const* coefs) {
  if (nz < 12 || nx < 12) {
    set_error("grid %d x %d too small: the 10th-order operators need nz, nx >= 12", nz, nx);
    return E_ARG;
  }
  for (int f = 0; f < 5; ++f)
    if (!hist[f] || !coefs[f]) { set_error("null array argument"); return E_ARG; }
  if (5LL * nz * nx >= (1LL << 31)) {
    // cells inside a level and the five coefficient arrays are addressed with
    // 32-bit offsets; levels and fields with 64-bit ones (a history field
    // holds nt*nz*nx elements, past 2^31 at 390 x 690 x 2000)
    set_error("grid %d x %d: 5*nz*nx must stay below 2^31", nz, nx);
    return E_ARG;
  }
  if (nt < 3) return OK;  // no level below nt-2 to update
  STCHK(ensure_device());
  // The caller's layout, unpadded: each field's history moves in one
  // contiguous copy (a pitched copy of nx*nt rows from pageable memory runs
  // at a small fraction of the link rate).
  const int ldz = nz;
  const long long n2 = (long long)ldz * nx;
  size_t freeb = 0, totb = 0;
  HIPCHK(hipMemGetInfo(&freeb, &totb));
  const size_t need = (5 * (size_t)nt + 5) * n2 * sizeof(float);
  if (need > freeb) {
    set_error("rtm_op3: %zu bytes of history exceed the %zu bytes of free device memory", need, freeb);
    return E_NOMEM;
  }
  hipStream_t st;
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{st};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  DevBuf h, cf, dd, dt;
  STCHK(h.alloc(5 * (size_t)nt * n2 * sizeof(float), "adjoint histories"));
  STCHK(cf.alloc(5 * n2 * sizeof(float), "coefficients"));
  std::vector<float> c5((size_t)5 * nz * nx);
  for (int c = 0; c < 5; ++c) std::memcpy(c5.data() + (size_t)c * nz * nx, coefs[c], sizeof(float) * nz * nx);
  STCHK(fill_masked_coefs(c5.data(), nz, nx, ldz, cf.as<float>()));
  float* hp = h.as<float>();
  for (int f = 0; f < 5; ++f)
    HIPCHK(hipMemcpyAsync(hp + (size_t)f * nt * n2, hist[f], (size_t)nt * n2 * sizeof(float), hipMemcpyHostToDevice,
                          st));
  // one shot, history layout: field f of 0-based level L at hp + L*n2 + f*nt*n2
  ShotDev d{};
  d.nz = nz; d.nx = nx; d.ldz = ldz; d.nt = nt; d.n2 = n2;
  d.lev = hp; d.lev_stride = n2; d.fstride = (long long)nt * n2; d.lev_mod = nt; d.lev_off = 0;
  d.coef = cf.as<float>();
  const int txr = default_txr();
  // kernel form 2 (RTM_TBLOCK=2): strips of RTM_MARCH_W columns marching
  // along x (form 3 marches the forward only: tiles here)
  const bool march = default_tblock() == 2;
  const int mw = default_march_w(), mnw = default_march_nw();
  if (march && (mw < 2 || (mw & 1))) { set_error("RTM_MARCH_W=%d is not a positive even number", mw); return E_ARG; }
  const int tx = march ? mw : tile_x(txr), tz = march ? TILE_Z * mnw : TILE_Z;
  const int nbz = (nz + tz - 1) / tz, nbx = (nx + tx - 1) / tx;
  std::vector<TileRef> tiles;
  for (int bx = 0; bx < nbx; ++bx)
    for (int bz = 0; bz < nbz; ++bz) tiles.push_back({0, bz * tz, bx * tx, bz + nbz * bx});
  STCHK(dd.alloc(sizeof(ShotDev), "descriptor"));
  STCHK(dt.alloc(tiles.size() * sizeof(TileRef), "tiles"));
  HIPCHK(hipMemcpyAsync(dd.p, &d, sizeof(ShotDev), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dt.p, tiles.data(), tiles.size() * sizeof(TileRef), hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(e0, st));
  for (int it = nt - 3; it >= 0; --it) {
    if (march)
      STCHK(launch_adjm(dd.as<ShotDev>(), dt.as<TileRef>(), (int)tiles.size(), it, A_XIN, mw, mnw, SNAP_HIST,
                        default_xcd(), st));
    else
      STCHK(launch_adj(dd.as<ShotDev>(), dt.as<TileRef>(), (int)tiles.size(), it, A_XIN, txr, SNAP_HIST,
                       default_xcd(), st));
  }
  HIPCHK(hipEventRecord(e1, st));
  for (int f = 0; f < 5; ++f)
    HIPCHK(hipMemcpyAsync(hist[f], hp + (size_t)f * nt * n2, (size_t)nt * n2 * sizeof(float), hipMemcpyDeviceToHost,
                          st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  Timing t;
  t.adj_ms = t.total_ms = ms;
  t.adj_launches = nt - 2;
  t.adj_cells = (double)nz * nx * (nt - 2);
  store_timing(t);
  return OK;
}

// ---------------------------------------------------------------------------
// Allocation-free drop-in adjoint (rtm_op3_hip_noalloc). All device memory it
// uses is the file-scope array `scratch` (each GPU the library runs on holds
// its own copy until the process ends): a ring of three levels of the five
// fields, the masked coefficients, the shot descriptor and the tile table.
// The histories stay in host memory and pass through the ring one level at a
// time. Every call on a GPU shares that GPU's array, so calls that overlap on
// one GPU race; the array keeps the last call's contents, and each call
// overwrites every part of it that it reads.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_dev[kMaxDevices] = {};

// Device address of the current device's scratch. The array's name is a
// host-side stand-in; its address on a device comes from
// hipGetSymbolAddress with that device current, once per device.
static int scratch_base(unsigned char** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device %d beyond the %d devices the scratch table covers", dev, kMaxDevices);
    return E_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_dev[dev];
  return OK;
}

struct MaskOrigin {
  int ix0[5], iz0[5];
};

// Zero each (unpadded, ldz = nz) coefficient array outside its operator's
// range, as fill_masked_coefs does on the host.
__global__ void k_mask_coefs(float* __restrict__ coef, int nz, int nx, MaskOrigin m) {
  const long long n2 = (long long)nz * nx;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < 5 * n2;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i / n2);
    const long long j = i - c * n2;
    const int ix = (int)(j / nz), iz = (int)(j - (long long)ix * nz);
    if (ix < m.ix0[c] || ix > nx - 6 || iz < m.iz0[c] || iz > nz - 6) coef[i] = 0.f;
  }
}

// Tile (or strip) table of one shot: entry bz + nbz*bx covers rows bz*tz,
// columns bx*tx (the order op3 builds on the host).
__global__ void k_tile_table(TileRef* __restrict__ t, int nbz, int nbx, int tz, int tx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbz * nbx) {
    const int bx = i / nbz, bz = i - bx * nbz;
    t[i] = TileRef{0, bz * tz, bx * tx, i};
  }
}

static int op3_noalloc(int nt, int nz, int nx, float* const* hist, const float* const* coefs) {
  if (nz < 12 || nx < 12) {
    set_error("grid %d x %d too small: the 10th-order operators need nz, nx >= 12", nz, nx);
    return E_ARG;
  }
  for (int f = 0; f < 5; ++f)
    if (!hist[f] || !coefs[f]) { set_error("null array argument"); return E_ARG; }
  if (5LL * nz * nx >= (1LL << 31)) {
    set_error("grid %d x %d: 5*nz*nx must stay below 2^31", nz, nx);
    return E_ARG;
  }
  const bool march = default_tblock() == 2;
  const int txr = default_txr(), mw = default_march_w(), mnw = default_march_nw();
  if (march && (mw < 2 || (mw & 1))) { set_error("RTM_MARCH_W=%d is not a positive even number", mw); return E_ARG; }
  if (!march && txr != 1 && txr != 2 && txr != 4 && txr != 8) { set_error("unsupported tile width %d", txr); return E_ARG; }
  const int tx = march ? mw : tile_x(txr), tz = march ? TILE_Z * mnw : TILE_Z;
  const int nbz = (nz + tz - 1) / tz, nbx = (nx + tx - 1) / tx, ntiles = nbz * nbx;
  const size_t n2 = (size_t)nz * nx;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t off_desc = up(20 * n2 * sizeof(float));
  const size_t off_tiles = up(off_desc + sizeof(ShotDev));
  const size_t need = off_tiles + (size_t)ntiles * sizeof(TileRef);
  if (need > kScratchBytes) {
    set_error("rtm_op3_hip_noalloc: a %d x %d grid needs %zu bytes of scratch, above its %zu bytes "
              "(20 * nz * nx floats plus tables: nz * nx up to about 1.67e6)", nz, nx, need, kScratchBytes);
    return E_ARG;
  }
  if (nt < 3) return OK;  // no level below nt-2 to update
  STCHK(ensure_device());
  unsigned char* base = nullptr;
  STCHK(scratch_base(&base));
  float* const ring = reinterpret_cast<float*>(base);
  float* const cf = ring + 15 * n2;
  ShotDev* const dd = reinterpret_cast<ShotDev*>(base + off_desc);
  TileRef* const tt = reinterpret_cast<TileRef*>(base + off_tiles);
  const hipStream_t st = nullptr;  // the null stream: nothing to create
  for (int c = 0; c < 5; ++c)
    HIPCHK(hipMemcpyAsync(cf + c * n2, coefs[c], n2 * sizeof(float), hipMemcpyHostToDevice, st));
  MaskOrigin mo;
  for (int c = 0; c < 5; ++c) { mo.ix0[c] = kMaskIx0[c]; mo.iz0[c] = kMaskIz0[c]; }
  k_mask_coefs<<<1024, 256, 0, st>>>(cf, nz, nx, mo);
  HIPCHK(hipGetLastError());
  k_tile_table<<<(ntiles + 255) / 256, 256, 0, st>>>(tt, nbz, nbx, tz, tx);
  HIPCHK(hipGetLastError());
  // ring of three levels: field f of level L at ring + (L % 3)*5*n2 + f*n2
  ShotDev d{};
  d.nz = nz; d.nx = nx; d.ldz = nz; d.nt = nt; d.n2 = (long long)n2;
  d.lev = ring; d.lev_stride = 5 * (long long)n2; d.fstride = (long long)n2; d.lev_mod = 3; d.lev_off = 0;
  d.coef = cf;
  HIPCHK(hipMemcpyAsync(dd, &d, sizeof(ShotDev), hipMemcpyHostToDevice, st));
  auto slot = [&](int L) { return ring + (size_t)(L % 3) * 5 * n2; };
  auto level_in = [&](int L) -> int {
    for (int f = 0; f < 5; ++f)
      HIPCHK(hipMemcpyAsync(slot(L) + f * n2, hist[f] + (size_t)L * n2, n2 * sizeof(float), hipMemcpyHostToDevice,
                            st));
    return OK;
  };
  STCHK(level_in(nt - 1));
  STCHK(level_in(nt - 2));
  for (int it = nt - 3; it >= 0; --it) {
    // the slot of level it held level it+3, last read by the step of it+1
    STCHK(level_in(it));
    if (march)
      STCHK(launch_adjm(dd, tt, ntiles, it, A_XIN, mw, mnw, SNAP_HIST, default_xcd(), st));
    else
      STCHK(launch_adj(dd, tt, ntiles, it, A_XIN, txr, SNAP_HIST, default_xcd(), st));
    for (int f = 0; f < 5; ++f)
      HIPCHK(hipMemcpyAsync(hist[f] + (size_t)it * n2, slot(it) + f * n2, n2 * sizeof(float), hipMemcpyDeviceToHost,
                            st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return OK;
}

// Single-shot engine from an rtm_shot_t.
static int single_engine(Engine& e, const rtm_shot_t* s, const float* coef5, const float* wavelet,
                         int snap_fmt) {
  if (!s || !coef5 || !wavelet) { set_error("null argument"); return E_ARG; }
  EngineCfg cfg;
  cfg.nt = s->nt; cfg.dx = s->dx; cfg.dt = s->dt; cfg.abc = s->abc; cfg.adj_abc = s->adj_abc;
  cfg.snap_every = s->snap_every < 1 ? 1 : s->snap_every;
  cfg.snap_fmt = snap_fmt;
  cfg.comp = s->comp_image;
  cfg.txr = default_txr();
  cfg.xcd = default_xcd();
  cfg.tblock = default_tblock();
  cfg.march_w = default_march_w();
  cfg.march_nw = default_march_nw();
  cfg.use_graph = getenv("RTM_GRAPH") ? atoi(getenv("RTM_GRAPH")) : 0;
  ShotSpec p{};
  p.nz = s->nz; p.nx = s->nx; p.xs = s->xs; p.zs = s->zs; p.zr = s->zr; p.rx0 = s->rx0; p.nrx = s->nrx;
  return e.setup(cfg, {p}, {coef5}, wavelet);
}

// Model of the Fortran setup (W/model-mod.f90:21-89, W/rtmlib-mod.f90:51-58,
// W/rtm-mod-adj.f90:111-113), FP32 in reference order.
static int model_op(int nz, int nx, float dx, float dt, const float* rho, const float* lam,
                    const float* mu, float* coef5) {
  if (nz < 3 || nx < 3) { set_error("model_op needs nz, nx >= 3"); return E_ARG; }
  const size_t n = (size_t)nz * nx;
  std::vector<float> m2(lam, lam + n), m3(mu, mu + n), m1x(n, 0.f), m1z(n, 0.f);
  auto I = [nz](int iz, int ix) { return (size_t)iz + (size_t)nz * ix; };
  for (int ix = 0; ix < nx; ++ix)
    for (int iz = 0; iz < nz - 1; ++iz) m1x[I(iz, ix)] = 1.0f / (0.5f * (rho[I(iz + 1, ix)] + rho[I(iz, ix)]));
  for (int ix = 0; ix < nx - 1; ++ix)
    for (int iz = 0; iz < nz; ++iz) m1z[I(iz, ix)] = 1.0f / (0.5f * (rho[I(iz, ix + 1)] + rho[I(iz, ix)]));
  for (int iz = 0; iz < nz; ++iz) m1z[I(iz, nx - 1)] = 1.0f / rho[I(iz, nx - 1)];
  for (int ix = 0; ix < nx; ++ix) m1x[I(nz - 1, ix)] = 1.0f / rho[I(nz - 1, ix)];
  // in-place harmonic 2x2 means in the reference's column-major sweep: each
  // update reads only cells not yet overwritten
  for (int ix = 1; ix < nx - 1; ++ix)
    for (int iz = 1; iz < nz - 1; ++iz) {
      const size_t a = I(iz, ix), b = I(iz + 1, ix), c = I(iz, ix + 1), d = I(iz + 1, ix + 1);
      m2[a] = 4.0f / ((1.0f / m2[a]) + (1.0f / m2[b]) + (1.0f / m2[c]) + (1.0f / m2[d]));
      if (m3[a] == 0.f || m3[b] == 0.f || m3[c] == 0.f || m3[d] == 0.f)
        m3[a] = 0.f;
      else
        m3[a] = 4.0f / ((1.0f / m3[a]) + (1.0f / m3[b]) + (1.0f / m3[c]) + (1.0f / m3[d]));
    }
  for (int ix = 0; ix < nx; ++ix) {
    m2[I(0, ix)] = m2[I(1, ix)];
    m2[I(nz - 1, ix)] = m2[I(nz - 2, ix)];
    m3[I(0, ix)] = m3[I(1, ix)];
    m3[I(nz - 1, ix)] = m3[I(nz - 2, ix)];
  }
  for (int iz = 0; iz < nz; ++iz) {
    m2[I(iz, 0)] = m2[I(iz, 1)];
    m2[I(iz, nx - 1)] = m2[I(iz, nx - 2)];
    m3[I(iz, 0)] = m3[I(iz, 1)];
    m3[I(iz, nx - 1)] = m3[I(iz, nx - 2)];
  }
  const float aux = dt / dx;
  for (size_t i = 0; i < n; ++i) {
    coef5[i] = aux * m1x[i];
    coef5[n + i] = aux * m1z[i];
    coef5[2 * n + i] = (dt / dx) * m2[i];
    coef5[3 * n + i] = (dt / dx) * m3[i];
    coef5[4 * n + i] = (dt / dx) * (m2[i] + 2.0f * m3[i]);
  }
  return OK;
}

// Ricker wavelet, W/source-mod.f90:31-44.
static void ricker(int nt, float dt, float fp, int nsource, float* out) {
  const int ft = nsource / 10;
  const float pi = 3.14159265f;
  for (int it = 1; it <= nt; ++it) {
    out[it - 1] = 0.f;
    if (it < ft) continue;
    const float t = ((float)it * dt) - ((float)nsource * dt / 2.0f);
    const float aux = (pi * pi * fp * fp * t * t);
    out[it - 1] = 10.0e7f * (1.0f - 2.0f * aux) * std::exp(-aux);
  }
}

// Non-finite scan of a device buffer.
__global__ void k_nonfinite(const float* p, size_t n, int* flag) {
  int bad = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    bad |= !isfinite(p[i]);
  if (bad) atomicOr(flag, 1);
}


// Streaming copy used to measure the achievable HBM bandwidth and to
// calibrate the FETCH_SIZE / WRITE_SIZE counters (known bytes per dispatch).
typedef float v4f __attribute__((ext_vector_type(4)));

// mode 0: copy; 1: copy with non-temporal loads/stores; 2: read-only
// (reduction into one value per thread); 3: write-only (fill).
__global__ __launch_bounds__(256) void k_copy_probe(const v4f* __restrict__ src,
                                                    v4f* __restrict__ dst, size_t n4, int mode) {
  // four independent 16-byte accesses per thread in flight
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  for (; i + 3 * stride < n4; i += 4 * stride) {
    if (mode == 3) {
      const v4f z = {1.f, 2.f, 3.f, 4.f};
      dst[i] = z;
      dst[i + stride] = z;
      dst[i + 2 * stride] = z;
      dst[i + 3 * stride] = z;
      continue;
    }
    v4f a, b, c, d;
    if (mode == 1) {
      a = __builtin_nontemporal_load(src + i);
      b = __builtin_nontemporal_load(src + i + stride);
      c = __builtin_nontemporal_load(src + i + 2 * stride);
      d = __builtin_nontemporal_load(src + i + 3 * stride);
    } else {
      a = src[i];
      b = src[i + stride];
      c = src[i + 2 * stride];
      d = src[i + 3 * stride];
    }
    if (mode == 2) {
      acc += (a + b) + (c + d);
    } else if (mode == 1) {
      __builtin_nontemporal_store(a, dst + i);
      __builtin_nontemporal_store(b, dst + i + stride);
      __builtin_nontemporal_store(c, dst + i + 2 * stride);
      __builtin_nontemporal_store(d, dst + i + 3 * stride);
    } else {
      dst[i] = a;
      dst[i + stride] = b;
      dst[i + 2 * stride] = c;
      dst[i + 3 * stride] = d;
    }
  }
  for (; i < n4; i += stride) {
    if (mode == 2)
      acc += src[i];
    else
      dst[i] = src[i];
  }
  if (mode == 2 && acc.x == 123.456f) dst[0] = acc;  // keeps the reads live
}

static int host_finite(const float* p, size_t n, const char* what) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) {
      set_error("non-finite value in %s at %zu", what, i);
      return E_NONFINITE;
    }
  return OK;
}

}  // namespace erm

// ===========================================================================
// C ABI
using namespace erm;

struct rtm_ctx {
  int nz, nx, nt;
  float dx, dt;
  int xs, zs, zr;
  int abc;
  int adj_abc = 0, snap_every = 1, snap_prec = 0, comp_image = 0;
  int rx0 = 0, nrx = -1;
  std::vector<float> coef5;
  std::vector<float> wavelet;
};

struct rtm_batch {
  Engine eng;
  int nshots = 0, nt = 0, txr = 0, tblock = 0, march_w = 0;
  std::vector<int> nz, nx, nrx;
};

namespace erm {
Engine* batch_engine(rtm_batch* b) { return b ? &b->eng : nullptr; }
}  // namespace erm

RTM_EXPORT const char* rtm_hip_last_error(void) { return erm::last_error(); }

RTM_EXPORT int rtm_hip_init(int device) {
  if (device >= 0) HIPCHK(hipSetDevice(device));
  return ensure_device();
}

RTM_EXPORT int rtm_hip_finalize(void) {
  HIPCHK(hipDeviceSynchronize());
  return OK;
}

RTM_EXPORT int rtm_last_timings(double* out5) {
  if (!out5) return E_ARG;
  for (int i = 0; i < 5; ++i) out5[i] = g_timings[i];
  return OK;
}

RTM_EXPORT int rtm_op3_hip(int nt, int nz, int nx, int zrec, float* Vx0, float* Vz0,
                           float* sigmaxx0, float* sigmazz0, float* sigmaxz0, const float* m1_x,
                           const float* m1_z, const float* aux_m2_c, const float* aux_m3_c,
                           const float* aux_m2m3_c) {
  (void)zrec;
  float* h[5] = {Vx0, Vz0, sigmaxx0, sigmazz0, sigmaxz0};
  const float* c[5] = {m1_x, m1_z, aux_m2_c, aux_m3_c, aux_m2m3_c};
  return op3(nt, nz, nx, h, c);
}

RTM_EXPORT int rtm_op3_hip_noalloc(int nt, int nz, int nx, int zrec, float* Vx0, float* Vz0,
                                   float* sigmaxx0, float* sigmazz0, float* sigmaxz0, const float* m1_x,
                                   const float* m1_z, const float* aux_m2_c, const float* aux_m3_c,
                                   const float* aux_m2m3_c) {
  (void)zrec;
  float* h[5] = {Vx0, Vz0, sigmaxx0, sigmazz0, sigmaxz0};
  const float* c[5] = {m1_x, m1_z, aux_m2_c, aux_m3_c, aux_m2m3_c};
  return op3_noalloc(nt, nz, nx, h, c);
}

RTM_EXPORT void rtm_op3_c_(int nt, int nz, int nx, int zrec, float* Vx0, float* Vz0,
                           float* sigmaxx0, float* sigmazz0, float* sigmaxz0, float* m1_x,
                           float* m1_z, float* aux_m2_c, float* aux_m3_c, float* aux_m2m3_c) {
  const int st = rtm_op3_hip(nt, nz, nx, zrec, Vx0, Vz0, sigmaxx0, sigmazz0, sigmaxz0, m1_x, m1_z,
                             aux_m2_c, aux_m3_c, aux_m2m3_c);
  if (st != OK) fprintf(stderr, "rtm_op3_c_: error %d: %s\n", st, erm::last_error());
}

RTM_EXPORT int rtm_forward_hip(const rtm_shot_t* s, const float* coef5, const float* wavelet,
                               float* rec, float* last) {
  if (s && s->nrx > 0 && !rec) { set_error("null receiver buffer"); return E_ARG; }
  Engine e;
  STCHK(single_engine(e, s, coef5, wavelet, SNAP_HIST));
  STCHK(e.model_data());
  store_timing(e.timing());
  if (s->nrx > 0) STCHK(e.get_data(0, rec));
  if (last) STCHK(e.get_last_levels(0, last));
  if (s->nrx > 0) STCHK(host_finite(rec, (size_t)5 * s->nt * s->nrx, "receiver data"));
  return OK;
}

RTM_EXPORT int rtm_forward_movie_hip(const rtm_shot_t* s, const float* coef5, const float* wavelet, int nsnap,
                                     int z0, int x0, int wnz, int wnx, float* movie, int* dsnap) {
  if (!s || !coef5 || !wavelet || !movie || nsnap < 1 || wnz < 1 || wnx < 1) { set_error("bad argument"); return E_ARG; }
  const int every = s->nt / nsnap;
  if (every < 1) { set_error("movie: %d frames exceed nt = %d", nsnap, s->nt); return E_ARG; }
  STCHK(ensure_device());
  const size_t frame = (size_t)5 * wnz * wnx;
  DevBuf m;
  STCHK(m.alloc(frame * nsnap * sizeof(float), "movie frames"));
  STCHK(launch_fill(m.as<float>(), frame * nsnap, 0.f, nullptr));
  HIPCHK(hipDeviceSynchronize());
  EngineCfg cfg;
  cfg.nt = s->nt; cfg.dx = s->dx; cfg.dt = s->dt; cfg.abc = s->abc;
  cfg.txr = default_txr();
  cfg.xcd = default_xcd();
  cfg.tblock = default_tblock();
  cfg.march_w = default_march_w();
  cfg.march_nw = default_march_nw();
  cfg.movie_every = every;
  cfg.movie_n = nsnap;
  cfg.movie_z0 = z0; cfg.movie_x0 = x0; cfg.movie_nz = wnz; cfg.movie_nx = wnx;
  cfg.movie = m.as<float>();
  ShotSpec p{};
  p.nz = s->nz; p.nx = s->nx; p.xs = s->xs; p.zs = s->zs; p.zr = s->zr; p.rx0 = s->rx0; p.nrx = s->nrx;
  Engine e;
  STCHK(e.setup(cfg, {p}, {coef5}, wavelet));
  STCHK(e.model_data());
  store_timing(e.timing());
  HIPCHK(hipMemcpy(movie, m.p, frame * nsnap * sizeof(float), hipMemcpyDeviceToHost));
  if (dsnap) *dsnap = every;
  return OK;
}

RTM_EXPORT int rtm_migrate_hip(const rtm_shot_t* s, const float* coef5, const float* wavelet,
                               const float* rec, float* img) {
  if (!s || !img || (s->nrx > 0 && !rec)) { set_error("null argument"); return E_ARG; }
  Engine e;
  STCHK(single_engine(e, s, coef5, wavelet, snap_fmt_of(s->snap_prec, s->snap_every)));
  if (s->nrx > 0) STCHK(e.set_data(0, rec));
  STCHK(e.migrate());
  store_timing(e.timing());
  STCHK(e.get_image(0, img));
  return host_finite(img, (size_t)3 * s->nz * s->nx, "images");
}

RTM_EXPORT int rtm_born_hip(const rtm_shot_t* s, const float* coef5, const float* wavelet,
                            const float* dm, float* rec) {
  if (!s || !dm || (s->nrx > 0 && !rec)) { set_error("null argument"); return E_ARG; }
  Engine e;
  STCHK(single_engine(e, s, coef5, wavelet, SNAP_HIST));
  STCHK(e.born({dm}));
  store_timing(e.timing());
  if (s->nrx > 0) STCHK(e.get_data(0, rec));
  return OK;
}

RTM_EXPORT int rtm_model_op(int nz, int nx, float dx, float dt, const float* rho, const float* lam,
                            const float* mu, float* coef5) {
  if (!rho || !lam || !mu || !coef5) { set_error("null argument"); return E_ARG; }
  return erm::model_op(nz, nx, dx, dt, rho, lam, mu, coef5);
}

RTM_EXPORT int rtm_ricker(int nt, float dt, float fp, int nsource, float* out) {
  if (!out || nt < 0) { set_error("bad argument"); return E_ARG; }
  erm::ricker(nt, dt, fp, nsource, out);
  return OK;
}

RTM_EXPORT int abc_k_hip(int nz, int nx, int abc, float* field) {
  if (!field) { set_error("null argument"); return E_ARG; }
  if (abc <= 0) return OK;
  if (2 * (abc + 5) > nz || 2 * (abc + 5) > nx) { set_error("sponge does not fit the grid"); return E_ARG; }
  STCHK(ensure_device());
  std::vector<float> hz, hx;
  sponge_factors(nz, abc, hz);
  sponge_factors(nx, abc, hx);
  DevBuf d, k;
  STCHK(d.alloc((size_t)nz * nx * sizeof(float), "field"));
  STCHK(k.alloc((size_t)(nz + nx) * sizeof(float), "factors"));
  HIPCHK(hipMemcpy(d.p, field, (size_t)nz * nx * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(k.p, hz.data(), nz * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(k.as<float>() + nz, hx.data(), nx * sizeof(float), hipMemcpyHostToDevice));
  STCHK(launch_apply_k(d.as<float>(), nz, nx, nz, k.as<float>(), k.as<float>() + nz, nullptr));
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(field, d.p, (size_t)nz * nx * sizeof(float), hipMemcpyDeviceToHost));
  return OK;
}

RTM_EXPORT int rtm_cfl_check(float vmax, float dt, float dx, double* courant, double* limit) {
  const double c[5] = {35.0 / 294912.0, -405.0 / 229376.0, 567.0 / 40960.0, -735.0 / 8192.0,
                       19845.0 / 16384.0};
  double s = 0;
  for (double v : c) s += std::fabs(v);
  const double lim = 1.0 / (std::sqrt(2.0) * s);
  const double cr = (double)vmax * (double)dt / (double)dx;
  if (courant) *courant = cr;
  if (limit) *limit = lim;
  if (!(cr < lim)) {
    set_error("Courant number %.4f exceeds the 10th-order staggered limit %.4f", cr, lim);
    return E_CFL;
  }
  return OK;
}


RTM_EXPORT int rtm_bandwidth_probe(size_t nbytes, int reps, int mode, double* gbps, double* ms_per_copy) {
  if (nbytes < 1024 || reps < 1) { set_error("bad probe size"); return E_ARG; }
  STCHK(ensure_device());
  const size_t n4 = nbytes / 16;
  DevBuf a, b;
  STCHK(a.alloc(n4 * 16, "probe source"));
  STCHK(b.alloc(n4 * 16, "probe destination"));
  HIPCHK(hipMemset(a.p, 0, n4 * 16));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  if (mode < 0 || mode > 3) { set_error("probe mode must be 0..3"); return E_ARG; }
  const unsigned blocks = 256 * 32;  // 32 workgroups per CU, grid-stride
  hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, nullptr, a.as<v4f>(), b.as<v4f>(), n4, mode);
  HIPCHK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, nullptr, a.as<v4f>(), b.as<v4f>(), n4, mode);
  HIPCHK(hipEventRecord(e1, nullptr));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double per = ms / reps;
  if (ms_per_copy) *ms_per_copy = per;
  const double moved = (mode >= 2 ? 1.0 : 2.0) * (double)(n4 * 16);  // bytes read + written
  if (gbps) *gbps = moved / (per * 1e-3) / 1e9;
  return OK;
}

RTM_EXPORT int rtm_init_c_(rtm_ctx** ctx, int nz, float dz, int nx, float dx, float dt,
                           int xsource, int zsource, int zrec, float fp, int nt, int abc,
                           int free_surface, const float* m1, const float* m2, const float* m3,
                           int lame) {
  (void)dz;
  (void)free_surface;
  if (!ctx || !m1 || !m2 || !m3) { set_error("null argument"); return E_ARG; }
  if (!lame) { set_error("only the Lame parameterisation (lame=1) is on the RTM path"); return E_ARG; }
  if (nz < 12 || nx < 12) { set_error("grid %d x %d too small", nz, nx); return E_ARG; }
  if (xsource < 1 || xsource > nx || zsource < 1 || zsource > nz || zrec < 1 || zrec > nz) {
    set_error("source (%d,%d) or receiver row %d outside the 1-based %d x %d grid", zsource, xsource,
              zrec, nz, nx);
    return E_ARG;
  }
  std::unique_ptr<rtm_ctx> c(new rtm_ctx());
  c->nz = nz; c->nx = nx; c->nt = nt; c->dx = dx; c->dt = dt;
  c->xs = xsource - 1; c->zs = zsource - 1; c->zr = zrec - 1;
  c->abc = abc;
  c->coef5.resize((size_t)5 * nz * nx);
  STCHK(erm::model_op(nz, nx, dx, dt, m1, m2, m3, c->coef5.data()));
  c->wavelet.resize(nt);
  erm::ricker(nt, dt, fp, 250, c->wavelet.data());  // nsource=250, W/rtm-mod-adj.f90:100
  *ctx = c.release();
  return OK;
}

RTM_EXPORT int rtm_set_option_c_(rtm_ctx* c, int key, int value) {
  if (!c) { set_error("null context"); return E_ARG; }
  switch (key) {
    case RTM_OPT_ADJ_ABC: c->adj_abc = value; break;
    case RTM_OPT_SNAP_EVERY:
      if (value < 1) { set_error("snap_every must be >= 1"); return E_ARG; }
      c->snap_every = value;
      break;
    case RTM_OPT_SNAP_PREC:
      if (value < 0 || value > 2) { set_error("snap_prec must be 0, 1 or 2"); return E_ARG; }
      c->snap_prec = value;
      break;
    case RTM_OPT_COMP_IMAGE: c->comp_image = value; break;
    case RTM_OPT_REC_X0: c->rx0 = value; break;
    case RTM_OPT_REC_NX: c->nrx = value; break;
    default: set_error("unknown option %d", key); return E_ARG;
  }
  return OK;
}

RTM_EXPORT int rtm_op_c_(rtm_ctx* c, int adj, float* dm1, float* dm2, float* dm3, float* Vx,
                         float* Vz, float* sigmaxx, float* sigmazz, float* sigmaxz) {
  if (!c || !dm1 || !dm2 || !dm3 || !Vx || !Vz || !sigmaxx || !sigmazz || !sigmaxz) {
    set_error("null argument");
    return E_ARG;
  }
  rtm_shot_t s{};
  s.nz = c->nz; s.nx = c->nx; s.nt = c->nt; s.dx = c->dx; s.dt = c->dt;
  s.xs = c->xs; s.zs = c->zs; s.zr = c->zr;
  s.rx0 = c->rx0;
  s.nrx = c->nrx < 0 ? c->nx - c->rx0 : c->nrx;
  s.abc = c->abc; s.adj_abc = c->adj_abc; s.snap_every = c->snap_every;
  s.snap_prec = c->snap_prec; s.comp_image = c->comp_image;
  const size_t nr = (size_t)s.nt * s.nrx, nn = (size_t)s.nz * s.nx;
  std::vector<float> rec(5 * nr);
  float* comp[5] = {Vx, Vz, sigmaxx, sigmazz, sigmaxz};
  if (adj) {
    for (int f = 0; f < 5; ++f) std::memcpy(rec.data() + f * nr, comp[f], nr * sizeof(float));
    std::vector<float> img(3 * nn);
    STCHK(rtm_migrate_hip(&s, c->coef5.data(), c->wavelet.data(), rec.data(), img.data()));
    std::memcpy(dm1, img.data(), nn * sizeof(float));
    std::memcpy(dm2, img.data() + nn, nn * sizeof(float));
    std::memcpy(dm3, img.data() + 2 * nn, nn * sizeof(float));
  } else {
    std::vector<float> dm(3 * nn);
    std::memcpy(dm.data(), dm1, nn * sizeof(float));
    std::memcpy(dm.data() + nn, dm2, nn * sizeof(float));
    std::memcpy(dm.data() + 2 * nn, dm3, nn * sizeof(float));
    STCHK(rtm_born_hip(&s, c->coef5.data(), c->wavelet.data(), dm.data(), rec.data()));
    for (int f = 0; f < 5; ++f) std::memcpy(comp[f], rec.data() + f * nr, nr * sizeof(float));
  }
  return OK;
}

RTM_EXPORT int rtm_final_c_(rtm_ctx* c) {
  delete c;
  return OK;
}

// ---------------------------------------------------------------------------
// Batched multi-shot RTM
RTM_EXPORT int rtm_batch_create(rtm_batch** out, const rtm_batch_cfg_t* cfg, int nshots,
                                const rtm_batch_shot_t* shots, const float* const* coef5,
                                const float* wavelet) {
  if (!out || !cfg || !shots || !coef5 || !wavelet || nshots < 1) { set_error("bad argument"); return E_ARG; }
  std::unique_ptr<rtm_batch> b(new rtm_batch());
  EngineCfg ec;
  ec.nt = cfg->nt; ec.dx = cfg->dx; ec.dt = cfg->dt; ec.abc = cfg->abc; ec.adj_abc = cfg->adj_abc;
  ec.snap_every = cfg->snap_every < 1 ? 1 : cfg->snap_every;
  ec.snap_fmt = snap_fmt_of(cfg->snap_prec, ec.snap_every);
  ec.comp = cfg->comp_image;
  ec.txr = cfg->txr > 0 ? cfg->txr : default_txr();
  const bool tune = cfg->txr == 0 && !getenv("RTM_TXR") && !getenv("RTM_TBLOCK");
  ec.xcd = default_xcd();
  ec.tblock = default_tblock();
  ec.march_w = default_march_w();
  ec.march_nw = default_march_nw();
  ec.use_graph = cfg->use_graph;
  ec.max_group = cfg->max_group;
  ec.stack_nz = cfg->stack_nz;
  ec.stack_nx = cfg->stack_nx;
  if (cfg->mem_fraction > 0) ec.mem_fraction = cfg->mem_fraction;
  std::vector<ShotSpec> sp(nshots);
  std::vector<const float*> cf(nshots);
  for (int i = 0; i < nshots; ++i) {
    const rtm_batch_shot_t& s = shots[i];
    sp[i] = ShotSpec{s.nz, s.nx, s.xs, s.zs, s.zr, s.rx0, s.nrx, s.img_z0, s.img_x0, s.img_nz, s.img_nx,
                     s.stack_z0, s.stack_x0};
    cf[i] = coef5[i];
    b->nz.push_back(s.nz);
    b->nx.push_back(s.nx);
    b->nrx.push_back(s.nrx);
  }
  b->nshots = nshots;
  b->nt = cfg->nt;
  if (tune) STCHK(autotune_txr(ec, sp, cf, wavelet, &ec.txr, &ec.tblock, nullptr, true, &ec.march_w));
  b->txr = ec.txr;
  b->tblock = ec.tblock;
  STCHK(b->eng.setup(ec, sp, cf, wavelet));
  b->march_w = b->eng.first_march_w();
  *out = b.release();
  return OK;
}

RTM_EXPORT int rtm_batch_autotune(const rtm_batch_cfg_t* cfg, int nshots, const rtm_batch_shot_t* shots,
                                  const float* const* coef5, const float* wavelet, int* txr, int* tblock,
                                  double* ms8) {
  if (!cfg || !shots || !coef5 || !wavelet || !txr || !tblock || nshots < 1) { set_error("bad argument"); return E_ARG; }
  EngineCfg ec;
  ec.nt = cfg->nt; ec.dx = cfg->dx; ec.dt = cfg->dt; ec.abc = cfg->abc; ec.adj_abc = cfg->adj_abc;
  ec.snap_every = cfg->snap_every < 1 ? 1 : cfg->snap_every;
  ec.snap_fmt = snap_fmt_of(cfg->snap_prec, ec.snap_every);
  ec.comp = cfg->comp_image;
  ec.xcd = default_xcd();
  ec.tblock = default_tblock();
  ec.march_w = default_march_w();
  ec.march_nw = default_march_nw();
  ec.use_graph = cfg->use_graph;
  std::vector<ShotSpec> sp(nshots);
  std::vector<const float*> cf(coef5, coef5 + nshots);
  for (int i = 0; i < nshots; ++i) {
    const rtm_batch_shot_t& s = shots[i];
    sp[i] = ShotSpec{s.nz, s.nx, s.xs, s.zs, s.zr, s.rx0, s.nrx, 0, 0, 0, 0, 0, 0};
  }
  return autotune_txr(ec, sp, cf, wavelet, txr, tblock, ms8, false);
}

RTM_EXPORT int rtm_batch_txr(rtm_batch* b) { return b ? b->txr : 0; }
RTM_EXPORT int rtm_batch_tblock(rtm_batch* b) { return b ? b->tblock : 0; }
RTM_EXPORT int rtm_batch_march_w(rtm_batch* b) { return b ? b->march_w : 0; }

RTM_EXPORT int rtm_batch_set_true_model(rtm_batch* b, const float* const* coef5_true) {
  if (!b || !coef5_true) { set_error("bad argument"); return E_ARG; }
  return b->eng.set_true_model(std::vector<const float*>(coef5_true, coef5_true + b->nshots));
}

RTM_EXPORT int rtm_batch_synthesize(rtm_batch* b) {
  if (!b) { set_error("null batch"); return E_ARG; }
  return b->eng.synthesize();
}

RTM_EXPORT int rtm_batch_set_data(rtm_batch* b, int shot, const float* rec) {
  if (!b) { set_error("null batch"); return E_ARG; }
  return b->eng.set_data(shot, rec);
}

RTM_EXPORT int rtm_batch_get_data(rtm_batch* b, int shot, float* rec) {
  if (!b) { set_error("null batch"); return E_ARG; }
  return b->eng.get_data(shot, rec);
}

RTM_EXPORT int rtm_batch_migrate(rtm_batch* b) {
  if (!b) { set_error("null batch"); return E_ARG; }
  STCHK(b->eng.migrate());
  store_timing(b->eng.timing());
  return OK;
}

RTM_EXPORT int rtm_batch_get_image(rtm_batch* b, int shot, float* img3) {
  if (!b) { set_error("null batch"); return E_ARG; }
  return b->eng.get_image(shot, img3);
}

RTM_EXPORT int rtm_batch_get_stack(rtm_batch* b, float* img3) {
  if (!b) { set_error("null batch"); return E_ARG; }
  return b->eng.get_stack(img3);
}

RTM_EXPORT int rtm_batch_info(rtm_batch* b, double* out8) {
  if (!b || !out8) { set_error("bad argument"); return E_ARG; }
  const Timing& t = b->eng.timing();
  out8[0] = t.fwd_ms;
  out8[1] = (double)t.fwd_launches;
  out8[2] = t.adj_ms;
  out8[3] = (double)t.adj_launches;
  out8[4] = t.total_ms;
  out8[5] = t.fwd_cells;
  out8[6] = t.adj_cells;
  out8[7] = b->eng.group_size();
  return OK;
}

RTM_EXPORT int rtm_batch_destroy(rtm_batch* b) {
  delete b;
  return OK;
}
