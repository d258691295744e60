// elastic2d_host.hip -- C ABI of the 2D elastic RTM library (host side).
//
// Entry points are declared in include/rtmlib_modC.h with the reference
// interface each one replaces. All device work is queued on one HIP stream
// per call; nothing here falls back to the CPU.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/rtmlib_modC.h"
#include "elastic2d.hpp"

namespace erm {

static thread_local char g_err[1024] = "";
static thread_local double g_timings[5] = {0, 0, 0, 0, 0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* last_error() { return g_err; }

// Tile width (x-cells per thread) of the step kernels; 4 gives a 64 x 16
// tile with 38.5 KB of LDS, four workgroups per CU.
static int g_txr = 4;
static int g_xcd = 1;

static int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return v ? atoi(v) : dflt;
}

static int ensure_device() {
  static bool inited = false;
  if (!inited) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
      set_error("no HIP device available (%s)", hipGetErrorString(e));
      return E_HIP;
    }
    g_txr = env_int("RTM_TXR", g_txr);
    g_xcd = env_int("RTM_XCD_REMAP", g_xcd);
    inited = true;
  }
  return OK;
}

// ---------------------------------------------------------------------------
// Host-side setup arithmetic (FP32, reference operation order).

// Sponge taper, W/abc-mod.f90:35-50 (abc+5 values).
static std::vector<float> abc_filter(int abc) {
  std::vector<float> f(abc + 5);
  const float alfa = 0.09f;
  for (int i = 1; i <= abc; ++i) {
    const float aux = (1.0f * (float)i) / (1.0f * (float)abc);
    f[i - 1] = std::exp(-alfa * aux * aux);
  }
  for (int i = abc + 1; i <= abc + 5; ++i) f[i - 1] = f[abc - 1];
  return f;
}

// Per-row and per-column factors of K (W/abc-mod.f90:59-78): top/bottom rows
// then left/right columns; a cell's factor is kz[iz]*kx[ix] applied in that
// order, so corners are damped twice as in the reference.
static void sponge_factors(int n, int abc, std::vector<float>& k) {
  k.assign(n, 1.0f);
  if (abc <= 0) return;
  const int abc2 = abc + 5;
  const std::vector<float> f = abc_filter(abc);
  for (int i1 = 1; i1 <= abc2; ++i1) {
    const int top = i1 - 1;
    if (top < n) k[top] = k[top] * f[abc2 - i1];
    const int bot = n - abc2 + i1 - 1;
    if (bot >= 0 && bot < n) k[bot] = k[bot] * f[i1 - 1];
  }
}

// Coefficient masks: each array is zero outside the (0-based) index range of
// the operators that use it (W/rtmlib-mod.f90: Ax/Bx ix 6..nx-5, iz 5..nz-5;
// Az/Bz ix 5..nx-5, iz 6..nz-5; D/E/F/G 5..n-5; H/J 6..n-5, 1-based).
static const int kMaskIx0[5] = {5, 4, 4, 5, 4};
static const int kMaskIz0[5] = {4, 5, 4, 5, 4};

// Upload the five (nz,nx) coefficient arrays, masked, into ldz-strided
// device arrays.
static int upload_coefs(const float* coef5, int nz, int nx, int ldz, float* dev) {
  const size_t n2 = (size_t)ldz * nx;
  std::vector<float> h(5 * n2, 0.0f);
  for (int c = 0; c < 5; ++c) {
    const float* src = coef5 + (size_t)c * nz * nx;
    float* dst = h.data() + c * n2;
    for (int ix = kMaskIx0[c]; ix <= nx - 6; ++ix)
      for (int iz = kMaskIz0[c]; iz <= nz - 6; ++iz)
        dst[iz + (size_t)ldz * ix] = src[iz + (size_t)nz * ix];
  }
  HIPCHK(hipMemcpy(dev, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  return OK;
}

static int check_dims(int nz, int nx) {
  if (nz < 12 || nx < 12) {
    set_error("grid %d x %d too small: the 10th-order operators need nz, nx >= 12", nz, nx);
    return E_ARG;
  }
  return OK;
}

static double now_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

// RAII stream + events for one call.
struct CallCtx {
  hipStream_t st = nullptr;
  hipEvent_t ev[6] = {};
  int init() {
    STCHK(ensure_device());
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    return OK;
  }
  ~CallCtx() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
};

// ---------------------------------------------------------------------------
// One shot on its computational grid: coefficients, sponge, two-level ring,
// source-wavefield history and receiver buffers, all device resident.
struct Shot {
  rtm_shot_t s{};
  int ldz = 0;
  size_t n2 = 0;
  DevBuf coef, kfac, ring, hist, rec, img, imgc, bsrc, dm;
  float* cp[5] = {};
  float* rp[2][5] = {};
  float *kz = nullptr, *kx = nullptr;

  int setup(const rtm_shot_t& sh, const float* coef5) {
    s = sh;
    STCHK(check_dims(s.nz, s.nx));
    if (s.nt < 1) { set_error("nt must be >= 1"); return E_ARG; }
    if (s.xs < 0 || s.xs >= s.nx || s.zs < 0 || s.zs >= s.nz) {
      set_error("source cell (%d,%d) outside the %d x %d grid", s.zs, s.xs, s.nz, s.nx);
      return E_ARG;
    }
    if (s.zr < 0 || s.zr >= s.nz || s.rx0 < 0 || s.nrx < 0 || s.rx0 + s.nrx > s.nx) {
      set_error("receiver row %d / columns [%d,%d) outside the grid", s.zr, s.rx0, s.rx0 + s.nrx);
      return E_ARG;
    }
    if (s.snap_every < 1) s.snap_every = 1;
    if (s.abc > 0 && (2 * (s.abc + 5) > s.nz || 2 * (s.abc + 5) > s.nx)) {
      set_error("sponge of %d cells per edge does not fit the %d x %d grid", s.abc + 5, s.nz, s.nx);
      return E_ARG;
    }
    ldz = device_ldz(s.nz);
    n2 = (size_t)ldz * s.nx;
    STCHK(coef.alloc(5 * n2 * sizeof(float), "coefficients"));
    for (int c = 0; c < 5; ++c) cp[c] = coef.as<float>() + c * n2;
    STCHK(upload_coefs(coef5, s.nz, s.nx, ldz, coef.as<float>()));
    if (s.abc > 0) {
      std::vector<float> hz, hx;
      sponge_factors(s.nz, s.abc, hz);
      sponge_factors(s.nx, s.abc, hx);
      STCHK(kfac.alloc((s.nz + s.nx) * sizeof(float), "sponge factors"));
      kz = kfac.as<float>();
      kx = kz + s.nz;
      HIPCHK(hipMemcpy(kz, hz.data(), s.nz * sizeof(float), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(kx, hx.data(), s.nx * sizeof(float), hipMemcpyHostToDevice));
    }
    STCHK(ring.alloc(10 * n2 * sizeof(float), "wavefield ring"));
    for (int r = 0; r < 2; ++r)
      for (int f = 0; f < 5; ++f) rp[r][f] = ring.as<float>() + (5 * r + f) * n2;
    return OK;
  }

  StepArgs base_args() const {
    StepArgs a{};
    a.nz = s.nz; a.nx = s.nx; a.ldz = ldz;
    for (int c = 0; c < 5; ++c) a.coef[c] = cp[c];
    a.zr = s.zr; a.rx0 = s.rx0; a.nrx = s.nrx;
    a.zs = s.zs; a.xs = s.xs;
    a.inv_dx = 1.0f / s.dx;
    a.inv_dt = 1.0f / s.dt;
    a.img_w = (float)s.snap_every;
    a.xcd_remap = g_xcd;
    return a;
  }

  float* histp(int which, int level1) const {  // level1: 1-based level
    return hist.as<float>() + ((size_t)which * s.nt + (level1 - 1)) * n2;
  }

  // Forward source propagation, levels 1..nt. wavelet[it-1] = source(it).
  // Records receivers into rec (device [5][nt][nrx]) when record, stores the
  // Vx/Vz history when store_hist, adds Born sources when born.
  int forward(const float* wavelet, bool record, bool store_hist, bool born, hipStream_t st,
              hipEvent_t e0, hipEvent_t e1) {
    const int nt = s.nt;
    STCHK(launch_fill(ring.as<float>(), 10 * n2, 0.f, st));
    if (store_hist) {
      STCHK(launch_fill(histp(0, 1), n2, 0.f, st));
      STCHK(launch_fill(histp(1, 1), n2, 0.f, st));
    }
    if (record) STCHK(launch_fill(rec.as<float>(), (size_t)5 * nt * s.nrx, 0.f, st));
    if (e0) HIPCHK(hipEventRecord(e0, st));
    // level L lives in ring slot L&1
    const int first = born ? 1 : 2;
    for (int it = first; it <= nt; ++it) {
      StepArgs a = base_args();
      float* const* cur = rp[it & 1];
      float* const* prv = rp[(it - 1) & 1];
      for (int f = 0; f < 5; ++f) {
        a.out[f] = cur[f];
        a.x2[f] = cur[f];
        a.y1[f] = prv[f];
      }
      // level 2 of the source propagation is the bare injection, without K
      // (W/rtm-mod-adj.f90:193-198); every other level is sponged.
      const bool k = born || it >= 3;
      a.kz = k ? kz : nullptr;
      a.kx = k ? kx : nullptr;
      if (!born) {
        a.has_src = 1;
        a.src_amp = s.dt * wavelet[it - 2];  // dt*source(it-1), explosive_source
      } else {
        for (int f = 0; f < 5; ++f) a.bsrc[f] = bsrc.as<float>() + f * n2;
        BornArgs b{};
        b.nz = s.nz; b.nx = s.nx; b.ldz = ldz;
        b.svx = histp(0, it); b.svz = histp(1, it);
        b.svx1 = it < nt ? histp(0, it + 1) : nullptr;
        b.svz1 = it < nt ? histp(1, it + 1) : nullptr;
        for (int c = 0; c < 3; ++c) b.dm[c] = dm.as<float>() + c * n2;
        for (int f = 0; f < 5; ++f) b.src[f] = bsrc.as<float>() + f * n2;
        b.inv_dx = 1.0f / s.dx;
        b.inv_dt = 1.0f / s.dt;
        STCHK(launch_born_sources(b, st));
      }
      if (record)
        for (int f = 0; f < 5; ++f)
          a.rec_out[f] = rec.as<float>() + ((size_t)f * nt + (it - 1)) * s.nrx;
      if (store_hist) {
        a.hvx = histp(0, it);
        a.hvz = histp(1, it);
      }
      STCHK(launch_fwd_step(a, g_txr, st));
    }
    if (e1) HIPCHK(hipEventRecord(e1, st));
    return OK;
  }

  // Adjoint of the receiver data with the Lame imaging condition, levels nt..1.
  int adjoint_image(hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    const int nt = s.nt;
    const int j = s.snap_every;
    STCHK(launch_fill(ring.as<float>(), 10 * n2, 0.f, st));
    STCHK(launch_fill(img.as<float>(), 3 * n2, 0.f, st));
    if (s.comp_image) STCHK(launch_fill(imgc.as<float>(), 3 * n2, 0.f, st));
    if (e0) HIPCHK(hipEventRecord(e0, st));
    for (int it = nt; it >= 1; --it) {
      StepArgs a = base_args();
      float* const* cur = rp[it & 1];
      float* const* nxt = rp[(it + 1) & 1];
      for (int f = 0; f < 5; ++f) {
        a.out[f] = cur[f];
        a.x2[f] = cur[f];
        a.y1[f] = nxt[f];
        a.rec_in[f] = rec.as<float>() + ((size_t)f * nt + (it - 1)) * s.nrx;
      }
      if (s.nrx == 0)
        for (int f = 0; f < 5; ++f) a.rec_in[f] = nullptr;
      a.kz = s.adj_abc ? kz : nullptr;
      a.kx = s.adj_abc ? kx : nullptr;
      if (it % j == 0) {
        a.do_image = 1;
        a.svx = histp(0, it);
        a.svz = histp(1, it);
        a.svx1 = it < nt ? histp(0, it + 1) : nullptr;
        a.svz1 = it < nt ? histp(1, it + 1) : nullptr;
        for (int c = 0; c < 3; ++c) {
          a.img[c] = img.as<float>() + c * n2;
          a.imgc[c] = s.comp_image ? imgc.as<float>() + c * n2 : nullptr;
        }
        a.comp = s.comp_image;
      }
      STCHK(launch_adj_step(a, g_txr, st));
    }
    if (s.comp_image)
      for (int c = 0; c < 3; ++c)
        STCHK(launch_image_finalize(img.as<float>() + c * n2, imgc.as<float>() + c * n2, n2, st));
    if (e1) HIPCHK(hipEventRecord(e1, st));
    return OK;
  }

  // Host (nt,nrx) component blocks -> device [5][nt][nrx].
  int upload_rec(const float* h) {
    const int nt = s.nt, nr = s.nrx;
    std::vector<float> t((size_t)5 * nt * nr);
    for (int f = 0; f < 5; ++f)
      for (int ix = 0; ix < nr; ++ix)
        for (int it = 0; it < nt; ++it)
          t[((size_t)f * nt + it) * nr + ix] = h[(size_t)f * nt * nr + it + (size_t)nt * ix];
    if (!t.empty()) HIPCHK(hipMemcpy(rec.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    return OK;
  }
  int download_rec(float* h) {
    const int nt = s.nt, nr = s.nrx;
    std::vector<float> t((size_t)5 * nt * nr);
    if (!t.empty()) HIPCHK(hipMemcpy(t.data(), rec.p, t.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int f = 0; f < 5; ++f)
      for (int ix = 0; ix < nr; ++ix)
        for (int it = 0; it < nt; ++it)
          h[(size_t)f * nt * nr + it + (size_t)nt * ix] = t[((size_t)f * nt + it) * nr + ix];
    return OK;
  }
  // ldz-strided device field <-> (nz,nx) host field.
  int download_field(const float* d, float* h) {
    HIPCHK(hipMemcpy2D(h, s.nz * sizeof(float), d, ldz * sizeof(float), s.nz * sizeof(float), s.nx,
                       hipMemcpyDeviceToHost));
    return OK;
  }
  int upload_field(const float* h, float* d) {
    HIPCHK(hipMemcpy2D(d, ldz * sizeof(float), h, s.nz * sizeof(float), s.nz * sizeof(float), s.nx,
                       hipMemcpyHostToDevice));
    return OK;
  }
};

// Non-finite scan of a device buffer: returns E_NONFINITE when any value is
// NaN or Inf.
__global__ void k_nonfinite(const float* p, size_t n, int* flag) {
  int bad = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    bad |= !isfinite(p[i]);
  if (bad) atomicOr(flag, 1);
}
static int check_finite(const float* p, size_t n, hipStream_t st, const char* what) {
  DevBuf flag;
  STCHK(flag.alloc(sizeof(int), "flag"));
  HIPCHK(hipMemsetAsync(flag.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_nonfinite, dim3(1024), dim3(256), 0, st, p, n, flag.as<int>());
  HIPCHK(hipGetLastError());
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (h) {
    set_error("non-finite values in %s", what);
    return E_NONFINITE;
  }
  return OK;
}

// ---------------------------------------------------------------------------
static int op3(int nt, int nz, int nx, float* const* hist, const float* const* coefs) {
  STCHK(check_dims(nz, nx));
  for (int f = 0; f < 5; ++f)
    if (!hist[f] || !coefs[f]) { set_error("null array argument"); return E_ARG; }
  if (nt < 3) return OK;  // no level below nt-2 to update
  CallCtx cc;
  STCHK(cc.init());
  const int ldz = device_ldz(nz);
  const size_t n2 = (size_t)ldz * nx;
  size_t freeb = 0, totb = 0;
  HIPCHK(hipMemGetInfo(&freeb, &totb));
  const size_t need = 5 * n2 * (size_t)nt * sizeof(float) + 5 * n2 * sizeof(float);
  if (need > freeb) {
    set_error("rtm_op3: %zu bytes of history exceed the %zu bytes of free device memory", need, freeb);
    return E_NOMEM;
  }
  DevBuf h, cf;
  STCHK(h.alloc(5 * n2 * (size_t)nt * sizeof(float), "adjoint histories"));
  STCHK(cf.alloc(5 * n2 * sizeof(float), "coefficients"));
  std::vector<float> c5((size_t)5 * nz * nx);
  for (int c = 0; c < 5; ++c) std::memcpy(c5.data() + (size_t)c * nz * nx, coefs[c], sizeof(float) * nz * nx);
  STCHK(upload_coefs(c5.data(), nz, nx, ldz, cf.as<float>()));
  float* hp[5];
  for (int f = 0; f < 5; ++f) {
    hp[f] = h.as<float>() + (size_t)f * nt * n2;
    HIPCHK(hipMemcpy2DAsync(hp[f], ldz * sizeof(float), hist[f], nz * sizeof(float), nz * sizeof(float),
                            (size_t)nx * nt, hipMemcpyHostToDevice, cc.st));
  }
  HIPCHK(hipEventRecord(cc.ev[0], cc.st));
  for (int it = nt - 3; it >= 0; --it) {
    StepArgs a{};
    a.nz = nz; a.nx = nx; a.ldz = ldz;
    for (int c = 0; c < 5; ++c) a.coef[c] = cf.as<float>() + c * n2;
    for (int f = 0; f < 5; ++f) {
      a.out[f] = hp[f] + (size_t)it * n2;
      a.xin[f] = a.out[f];
      a.x2[f] = hp[f] + (size_t)(it + 2) * n2;
      a.y1[f] = hp[f] + (size_t)(it + 1) * n2;
    }
    a.xcd_remap = g_xcd;
    STCHK(launch_adj_step(a, g_txr, cc.st));
  }
  HIPCHK(hipEventRecord(cc.ev[1], cc.st));
  for (int f = 0; f < 5; ++f)
    HIPCHK(hipMemcpy2DAsync(hist[f], nz * sizeof(float), hp[f], ldz * sizeof(float), nz * sizeof(float),
                            (size_t)nx * nt, hipMemcpyDeviceToHost, cc.st));
  HIPCHK(hipStreamSynchronize(cc.st));
  g_timings[0] = 0; g_timings[1] = 0;
  g_timings[2] = now_ms(cc.ev[0], cc.ev[1]);
  g_timings[3] = nt - 2;
  g_timings[4] = g_timings[2];
  return OK;
}

static int shot_forward(const rtm_shot_t* s, const float* coef5, const float* wavelet, float* rec,
                        float* last) {
  if (!s || !coef5 || !wavelet || (!rec && s->nrx > 0)) { set_error("null argument"); return E_ARG; }
  CallCtx cc;
  STCHK(cc.init());
  Shot sh;
  STCHK(sh.setup(*s, coef5));
  STCHK(sh.rec.alloc((size_t)5 * s->nt * std::max(s->nrx, 1) * sizeof(float), "receivers"));
  STCHK(sh.forward(wavelet, s->nrx > 0, false, false, cc.st, cc.ev[0], cc.ev[1]));
  HIPCHK(hipStreamSynchronize(cc.st));
  g_timings[0] = now_ms(cc.ev[0], cc.ev[1]);
  g_timings[1] = s->nt - 1;
  g_timings[2] = g_timings[3] = 0;
  g_timings[4] = g_timings[0];
  if (s->nrx > 0) STCHK(sh.download_rec(rec));
  if (last) {
    const size_t nn = (size_t)s->nz * s->nx;
    for (int f = 0; f < 5; ++f) {
      if (s->nt >= 2) STCHK(sh.download_field(sh.rp[(s->nt - 1) & 1][f], last + f * nn));
      STCHK(sh.download_field(sh.rp[s->nt & 1][f], last + (5 + f) * nn));
    }
  }
  return OK;
}

static int shot_migrate(const rtm_shot_t* s, const float* coef5, const float* wavelet,
                        const float* rec, float* img) {
  if (!s || !coef5 || !wavelet || !img || (!rec && s->nrx > 0)) { set_error("null argument"); return E_ARG; }
  if (s->snap_prec != 0) { set_error("rtm_migrate_hip: only fp32 snapshots in the single-shot path"); return E_ARG; }
  CallCtx cc;
  STCHK(cc.init());
  Shot sh;
  STCHK(sh.setup(*s, coef5));
  STCHK(sh.rec.alloc((size_t)5 * s->nt * std::max(s->nrx, 1) * sizeof(float), "receivers"));
  STCHK(sh.hist.alloc(2 * (size_t)s->nt * sh.n2 * sizeof(float), "source history"));
  STCHK(sh.img.alloc(3 * sh.n2 * sizeof(float), "images"));
  if (s->comp_image) STCHK(sh.imgc.alloc(3 * sh.n2 * sizeof(float), "image compensation"));
  if (s->nrx > 0) STCHK(sh.upload_rec(rec));
  HIPCHK(hipEventRecord(cc.ev[4], cc.st));
  STCHK(sh.forward(wavelet, false, true, false, cc.st, cc.ev[0], cc.ev[1]));
  STCHK(sh.adjoint_image(cc.st, cc.ev[2], cc.ev[3]));
  HIPCHK(hipEventRecord(cc.ev[5], cc.st));
  HIPCHK(hipStreamSynchronize(cc.st));
  g_timings[0] = now_ms(cc.ev[0], cc.ev[1]);
  g_timings[1] = s->nt - 1;
  g_timings[2] = now_ms(cc.ev[2], cc.ev[3]);
  g_timings[3] = s->nt;
  g_timings[4] = now_ms(cc.ev[4], cc.ev[5]);
  const size_t nn = (size_t)s->nz * s->nx;
  for (int c = 0; c < 3; ++c) STCHK(sh.download_field(sh.img.as<float>() + c * sh.n2, img + c * nn));
  STCHK(check_finite(sh.img.as<float>(), 3 * sh.n2, cc.st, "images"));
  return OK;
}

static int shot_born(const rtm_shot_t* s, const float* coef5, const float* wavelet, const float* dm,
                     float* rec) {
  if (!s || !coef5 || !wavelet || !dm || (!rec && s->nrx > 0)) { set_error("null argument"); return E_ARG; }
  CallCtx cc;
  STCHK(cc.init());
  Shot sh;
  STCHK(sh.setup(*s, coef5));
  STCHK(sh.rec.alloc((size_t)5 * s->nt * std::max(s->nrx, 1) * sizeof(float), "receivers"));
  STCHK(sh.hist.alloc(2 * (size_t)s->nt * sh.n2 * sizeof(float), "source history"));
  STCHK(sh.dm.alloc(3 * sh.n2 * sizeof(float), "model perturbation"));
  STCHK(sh.bsrc.alloc(5 * sh.n2 * sizeof(float), "Born sources"));
  STCHK(launch_fill(sh.dm.as<float>(), 3 * sh.n2, 0.f, cc.st));
  HIPCHK(hipStreamSynchronize(cc.st));
  const size_t nn = (size_t)s->nz * s->nx;
  for (int c = 0; c < 3; ++c) STCHK(sh.upload_field(dm + c * nn, sh.dm.as<float>() + c * sh.n2));
  STCHK(launch_fill(sh.bsrc.as<float>(), 5 * sh.n2, 0.f, cc.st));
  STCHK(sh.forward(wavelet, false, true, false, cc.st, nullptr, nullptr));
  STCHK(sh.forward(wavelet, s->nrx > 0, false, true, cc.st, cc.ev[0], cc.ev[1]));
  HIPCHK(hipStreamSynchronize(cc.st));
  g_timings[0] = now_ms(cc.ev[0], cc.ev[1]);
  g_timings[1] = s->nt;
  if (s->nrx > 0) STCHK(sh.download_rec(rec));
  return OK;
}

// ---------------------------------------------------------------------------
// Host setup of the Fortran operator module (W/model-mod.f90:21-89,
// W/rtmlib-mod.f90:51-58, W/rtm-mod-adj.f90:111-113), FP32 in reference order.
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

RTM_EXPORT void rtm_op3_c_(int nt, int nz, int nx, int zrec, float* Vx0, float* Vz0,
                           float* sigmaxx0, float* sigmazz0, float* sigmaxz0, float* m1_x,
                           float* m1_z, float* aux_m2_c, float* aux_m3_c, float* aux_m2m3_c) {
  const int st = rtm_op3_hip(nt, nz, nx, zrec, Vx0, Vz0, sigmaxx0, sigmazz0, sigmaxz0, m1_x, m1_z,
                             aux_m2_c, aux_m3_c, aux_m2m3_c);
  if (st != OK) fprintf(stderr, "rtm_op3_c_: error %d: %s\n", st, erm::last_error());
}

RTM_EXPORT int rtm_forward_hip(const rtm_shot_t* s, const float* coef5, const float* wavelet,
                               float* rec, float* last) {
  return shot_forward(s, coef5, wavelet, rec, last);
}

RTM_EXPORT int rtm_migrate_hip(const rtm_shot_t* s, const float* coef5, const float* wavelet,
                               const float* rec, float* img) {
  return shot_migrate(s, coef5, wavelet, rec, img);
}

RTM_EXPORT int rtm_born_hip(const rtm_shot_t* s, const float* coef5, const float* wavelet,
                            const float* dm, float* rec) {
  return shot_born(s, coef5, wavelet, dm, rec);
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
  CallCtx cc;
  STCHK(cc.init());
  std::vector<float> hz, hx;
  sponge_factors(nz, abc, hz);
  sponge_factors(nx, abc, hx);
  DevBuf d, k;
  STCHK(d.alloc((size_t)nz * nx * sizeof(float), "field"));
  STCHK(k.alloc((size_t)(nz + nx) * sizeof(float), "factors"));
  HIPCHK(hipMemcpy(d.p, field, (size_t)nz * nx * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(k.p, hz.data(), nz * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(k.as<float>() + nz, hx.data(), nx * sizeof(float), hipMemcpyHostToDevice));
  STCHK(launch_apply_k(d.as<float>(), nz, nx, nz, k.as<float>(), k.as<float>() + nz, cc.st));
  HIPCHK(hipStreamSynchronize(cc.st));
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

RTM_EXPORT int rtm_init_c_(rtm_ctx** ctx, int nz, float dz, int nx, float dx, float dt,
                           int xsource, int zsource, int zrec, float fp, int nt, int abc,
                           int free_surface, const float* m1, const float* m2, const float* m3,
                           int lame) {
  (void)dz;
  (void)free_surface;
  if (!ctx || !m1 || !m2 || !m3) { set_error("null argument"); return E_ARG; }
  if (!lame) { set_error("only the Lame parameterisation (lame=1) is on the RTM path"); return E_ARG; }
  STCHK(check_dims(nz, nx));
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
    case RTM_OPT_SNAP_PREC: c->snap_prec = value; break;
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
    STCHK(shot_migrate(&s, c->coef5.data(), c->wavelet.data(), rec.data(), img.data()));
    std::memcpy(dm1, img.data(), nn * sizeof(float));
    std::memcpy(dm2, img.data() + nn, nn * sizeof(float));
    std::memcpy(dm3, img.data() + 2 * nn, nn * sizeof(float));
  } else {
    std::vector<float> dm(3 * nn);
    std::memcpy(dm.data(), dm1, nn * sizeof(float));
    std::memcpy(dm.data() + nn, dm2, nn * sizeof(float));
    std::memcpy(dm.data() + 2 * nn, dm3, nn * sizeof(float));
    STCHK(shot_born(&s, c->coef5.data(), c->wavelet.data(), dm.data(), rec.data()));
    for (int f = 0; f < 5; ++f) std::memcpy(comp[f], rec.data() + f * nr, nr * sizeof(float));
  }
  return OK;
}

RTM_EXPORT int rtm_final_c_(rtm_ctx* c) {
  delete c;
  return OK;
}
