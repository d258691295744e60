// engine2d.hip -- multi-shot 2D elastic RTM engine (host side).
#include "engine2d.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace erm {

static int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return v ? atoi(v) : dflt;
}

int ensure_device() {
  static int state = 0;  // 0 unknown, 1 ok, -1 none
  if (state == 0) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    state = (e == hipSuccess && n > 0) ? 1 : -1;
    if (state < 0) set_error("no HIP device available (%s)", hipGetErrorString(e));
  }
  return state > 0 ? OK : E_HIP;
}

// Tile width (x-cells per thread) of the step kernels and XCD-aware tile
// numbering; RTM_TXR / RTM_XCD_REMAP override them (experiments). The batch
// API tunes the width per workload unless a width is given.
int default_txr() { return env_int("RTM_TXR", 2); }
int default_xcd() { return env_int("RTM_XCD_REMAP", 1); }
// Non-temporal stores of the new level (RTM_NT=1).
static int nt_stores() { return env_int("RTM_NT", 0); }
// Kernel form (RTM_TBLOCK): 0 one level per launch on tiles, 1 two levels
// per launch, 2 one level per launch marching along x in strips (3: the
// marching forward with the tile adjoint); the batch autotuner chooses the
// form per workload. RTM_MARCH_W fixes the strip width (default for groups
// whose strips fit one round at no width).
int default_tblock() { return env_int("RTM_TBLOCK", 0); }
int default_march_w() { return env_int("RTM_MARCH_W", 32); }
int default_march_nw() { return env_int("RTM_MARCH_NW", 2); }
// Per-group strip widths fitted to one round of resident workgroups, unless
// RTM_MARCH_W fixes the width or RTM_MARCH_FIT=0.
bool march_fit() { return !getenv("RTM_MARCH_W") && env_int("RTM_MARCH_FIT", 1) != 0; }

// Sponge taper, W/abc-mod.f90:35-50, and per-row/column factors of K
// (:59-78): rows first then columns, so corners are damped twice.
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

void sponge_factors(int n, int abc, std::vector<float>& k) {
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

// Each coefficient array is zero outside the 0-based index range of the
// operators that use it (W/rtmlib-mod.f90, 1-based: Ax/Bx ix 6..nx-5, iz
// 5..nz-5; Az/Bz ix 5..nx-5, iz 6..nz-5; D/E/F/G 5..n-5; H/J 6..n-5).
const int kMaskIx0[5] = {5, 4, 4, 5, 4};
const int kMaskIz0[5] = {4, 5, 4, 5, 4};

int fill_masked_coefs(const float* coef5, int nz, int nx, int ldz, float* dev) {
  const size_t n2 = (size_t)ldz * nx;
  std::vector<float> h(5 * n2, 0.0f);
  for (int c = 0; c < 5; ++c) {
    const float* src = coef5 + (size_t)c * nz * nx;
    float* dst = h.data() + c * n2;
    for (int ix = kMaskIx0[c]; ix <= nx - 6; ++ix)
      for (int iz = kMaskIz0[c]; iz <= nz - 6; ++iz) dst[iz + (size_t)ldz * ix] = src[iz + (size_t)nz * ix];
  }
  HIPCHK(hipMemcpy(dev, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  return OK;
}

static int nstrips(int nz, int nx, int txr) {
  const int nbz = (nz + TILE_Z - 1) / TILE_Z;
  const int tx = tile_x(txr);
  const int nbx = (nx + tx - 1) / tx;
  return nbz * nbx * TILE_WY;
}

Engine::~Engine() {
  destroy_graphs();
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (st_) (void)hipStreamDestroy(st_);
}

void Engine::destroy_graphs() {
  for (auto& g : groups_) {
    if (g.gfwd) (void)hipGraphExecDestroy(g.gfwd);
    if (g.gadj) (void)hipGraphExecDestroy(g.gadj);
    if (g.gsyn) (void)hipGraphExecDestroy(g.gsyn);
    g.gfwd = g.gadj = g.gsyn = nullptr;
  }
}

int Engine::setup(const EngineCfg& cfg, const std::vector<ShotSpec>& shots,
                  const std::vector<const float*>& coef5, const float* wavelet) {
  STCHK(ensure_device());
  cfg_ = cfg;
  specs_ = shots;
  if (shots.empty() || coef5.size() != shots.size() || !wavelet) {
    set_error("engine setup: no shots, coefficient list mismatch or null wavelet");
    return E_ARG;
  }
  if (cfg_.nt < 1 || cfg_.snap_every < 1 || !(cfg_.dx > 0) || !(cfg_.dt > 0)) {
    set_error("engine setup: nt, snap_every must be >= 1 and dx, dt > 0");
    return E_ARG;
  }
  if (cfg_.adj_abc < 0 || cfg_.adj_abc > 2) {
    set_error("adjoint sponge mode %d not in {0 none, 1 exact adjoint, 2 R/Src}", cfg_.adj_abc);
    return E_ARG;
  }
  if (cfg_.txr != 1 && cfg_.txr != 2 && cfg_.txr != 4 && cfg_.txr != 8) {
    set_error("tile width %d not in {1,2,4,8}", cfg_.txr);
    return E_ARG;
  }
  if (cfg_.tblock < 0 || cfg_.tblock > 3 ||
      (cfg_.tblock >= 2 && (cfg_.txr != 2 || cfg_.march_w < 2 || (cfg_.march_w & 1) ||
                            (cfg_.march_nw != 1 && cfg_.march_nw != 2 && cfg_.march_nw != 4)))) {
    set_error("kernel form %d with tile width %d / strip %d columns x %d waves: marching steps need tile width 2, "
              "an even strip width and 1, 2 or 4 waves",
              cfg_.tblock, cfg_.txr, cfg_.march_w, cfg_.march_nw);
    return E_ARG;
  }
  if (cfg_.movie) {
    const ShotSpec& p = shots[0];
    if (shots.size() != 1 || cfg_.movie_every < 1 || cfg_.movie_n < 1 || cfg_.movie_z0 < 0 || cfg_.movie_x0 < 0 ||
        cfg_.movie_nz < 1 || cfg_.movie_nx < 1 || cfg_.movie_z0 + cfg_.movie_nz > p.nz ||
        cfg_.movie_x0 + cfg_.movie_nx > p.nx) {
      set_error("movie: one shot, frames >= 1 and a window inside the grid are required");
      return E_ARG;
    }
  }
  if (cfg_.snap_fmt == SNAP_HIST && cfg_.snap_every != 1) cfg_.snap_fmt = SNAP_F32;
  HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  for (auto& e : ev_) HIPCHK(hipEventCreate(&e));
  res_.resize(shots.size());
  for (size_t s = 0; s < shots.size(); ++s) {
    const ShotSpec& p = shots[s];
    if (p.nz < 12 || p.nx < 12) {
      set_error("shot %zu: grid %d x %d too small (the 10th-order operators need >= 12)", s, p.nz, p.nx);
      return E_ARG;
    }
    if (p.xs < 0 || p.xs >= p.nx || p.zs < 0 || p.zs >= p.nz) {
      set_error("shot %zu: source cell (%d,%d) outside the %d x %d grid", s, p.zs, p.xs, p.nz, p.nx);
      return E_ARG;
    }
    if (p.zr < 0 || p.zr >= p.nz || p.rx0 < 0 || p.nrx < 0 || p.rx0 + p.nrx > p.nx) {
      set_error("shot %zu: receiver row %d / columns [%d,%d) outside the grid", s, p.zr, p.rx0, p.rx0 + p.nrx);
      return E_ARG;
    }
    if (cfg_.abc > 0 && (2 * (cfg_.abc + 5) > p.nz || 2 * (cfg_.abc + 5) > p.nx)) {
      set_error("shot %zu: sponge of %d cells per edge does not fit %d x %d", s, cfg_.abc + 5, p.nz, p.nx);
      return E_ARG;
    }
    if (!coef5[s]) { set_error("shot %zu: null coefficients", s); return E_ARG; }
    Res& r = res_[s];
    r.ldz = device_ldz(p.nz);
    r.n2 = (long long)r.ldz * p.nx;
    if (5 * r.n2 >= (1LL << 31)) {
      // the step kernels address cells inside the five coefficient arrays
      // with 32-bit offsets
      set_error("shot %zu: %lld cells per field (5 coefficient arrays) exceed 2^31 / 5", s, r.n2);
      return E_ARG;
    }
    STCHK(r.coef.alloc(5 * r.n2 * sizeof(float), "coefficients"));
    STCHK(fill_masked_coefs(coef5[s], p.nz, p.nx, r.ldz, r.coef.as<float>()));
    if (cfg_.abc > 0) {
      std::vector<float> hz, hx;
      sponge_factors(p.nz, cfg_.abc, hz);
      sponge_factors(p.nx, cfg_.abc, hx);
      STCHK(r.kfac.alloc((size_t)(p.nz + p.nx) * sizeof(float), "sponge factors"));
      HIPCHK(hipMemcpy(r.kfac.p, hz.data(), p.nz * sizeof(float), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(r.kfac.as<float>() + p.nz, hx.data(), p.nx * sizeof(float), hipMemcpyHostToDevice));
    }
    STCHK(r.rec.alloc((size_t)5 * cfg_.nt * std::max(p.nrx, 1) * sizeof(float), "receiver data"));
    STCHK(launch_fill(r.rec.as<float>(), (size_t)5 * cfg_.nt * std::max(p.nrx, 1), 0.f, st_));
    STCHK(r.img.alloc(3 * r.n2 * sizeof(float), "images"));
    STCHK(launch_fill(r.img.as<float>(), 3 * r.n2, 0.f, st_));
    if (cfg_.comp) STCHK(r.imgc.alloc(3 * r.n2 * sizeof(float), "image compensation"));
    n2max_ = std::max(n2max_, r.n2);
    r.nstrip = nstrips(p.nz, p.nx, cfg_.txr);
  }
  // dt*source(L-1) for level L (explosive_source, W/source-mod.f90:47-53)
  std::vector<float> amp(cfg_.nt, 0.f);
  for (int L = 2; L <= cfg_.nt; ++L) amp[L - 1] = cfg_.dt * wavelet[L - 2];
  STCHK(src_amp_.alloc(cfg_.nt * sizeof(float), "source amplitudes"));
  HIPCHK(hipMemcpy(src_amp_.p, amp.data(), cfg_.nt * sizeof(float), hipMemcpyHostToDevice));
  if (cfg_.stack_nz > 0 && cfg_.stack_nx > 0) {
    STCHK(stack_.alloc((size_t)3 * cfg_.stack_nz * cfg_.stack_nx * sizeof(double), "stacked image"));
    HIPCHK(hipMemsetAsync(stack_.p, 0, stack_.bytes, st_));
  }
  HIPCHK(hipStreamSynchronize(st_));
  STCHK(alloc_working());
  return build_groups();
}

// Working set of one shot: the level ring (5 fields per level) followed by
// its source-wavefield snapshots (and their per-strip scales).
static void shot_work(const EngineCfg& c, long long n2, int nstrip, size_t* snap, size_t* scale, size_t* work) {
  const long long K = c.nt / c.snap_every;
  switch (c.snap_fmt) {
    case SNAP_HIST: *snap = (size_t)2 * c.nt * n2 * sizeof(float); *scale = 0; break;
    case SNAP_F32: *snap = (size_t)4 * K * n2 * sizeof(float); *scale = 0; break;
    default:
      *snap = (size_t)4 * K * n2 * sizeof(unsigned short);
      *scale = (size_t)4 * K * nstrip * sizeof(float);
  }
  *snap = (*snap + 255) / 256 * 256;
  *scale = (*scale + 255) / 256 * 256;
  *work = (size_t)5 * kEngineRing * n2 * sizeof(float) + *snap + *scale;
}

int first_group_size(const EngineCfg& c, const std::vector<ShotSpec>& shots) {
  size_t freeb = 0, totb = 0;
  if (hipMemGetInfo(&freeb, &totb) != hipSuccess) return (int)shots.size();
  // per-shot buffers allocated before the working arena
  size_t fixed = 0;
  for (const ShotSpec& p : shots) {
    const long long n2 = (long long)device_ldz(p.nz) * p.nx;
    fixed += (size_t)5 * n2 * sizeof(float) + (size_t)5 * c.nt * std::max(p.nrx, 1) * sizeof(float) +
             (size_t)(c.comp ? 6 : 3) * n2 * sizeof(float) + (size_t)(p.nz + p.nx) * sizeof(float);
  }
  const size_t avail = fixed < freeb ? (size_t)(c.mem_fraction * (double)(freeb - fixed)) : 0;
  size_t used = 0;
  int g = 0;
  for (const ShotSpec& p : shots) {
    size_t snap, scale, work;
    shot_work(c, (long long)device_ldz(p.nz) * p.nx, nstrips(p.nz, p.nx, c.txr), &snap, &scale, &work);
    if (used + work > avail || (c.max_group > 0 && g >= c.max_group)) break;
    used += work;
    ++g;
  }
  return std::max(g, 1);
}

int Engine::alloc_working() {
  size_t largest = 0, total = 0;
  for (auto& r : res_) {
    shot_work(cfg_, r.n2, r.nstrip, &r.snap_bytes, &r.scale_bytes, &r.work_bytes);
    largest = std::max(largest, r.work_bytes);
    total += r.work_bytes;
  }
  size_t freeb = 0, totb = 0;
  HIPCHK(hipMemGetInfo(&freeb, &totb));
  size_t avail = (size_t)(cfg_.mem_fraction * (double)freeb);
  if (largest > avail) {
    set_error("one shot needs %zu bytes of working memory; %zu bytes usable", largest, avail);
    return E_NOMEM;
  }
  arena_bytes_ = std::min(total, avail);
  return work_.alloc(arena_bytes_, "group working set");
}

// Workgroups of the marching adjoint (fwd = false) or forward resident at
// once on the current device; RTM_MARCH_RESIDENT="adj,fwd" caps them (a test
// hook: small grids then get fitted widths above 16, different per kernel).
int march_resident(const EngineCfg& c, bool fwd) {
  int r = fwd ? fwdm_resident_blocks(c.march_nw, c.snap_fmt)
              : adjm_resident_blocks(c.march_nw, c.snap_fmt, c.comp != 0);
  if (const char* e = getenv("RTM_MARCH_RESIDENT")) {
    int ca = 0, cf = 0;
    if (std::sscanf(e, "%d,%d", &ca, &cf) == 2) {
      const int cap = fwd ? cf : ca;
      if (cap > 0 && (r <= 0 || cap < r)) r = cap;
    }
  }
  return r;
}

// The narrowest even strip width from 16 to 64 whose strips of the given
// shots fit one round of `resident` workgroups (0 when none does).
int fitted_march_w(const std::vector<ShotSpec>& shots, int march_nw, int resident) {
  if (resident <= 0) return 0;
  const int sz = TILE_Z * march_nw;
  for (int c = 16; c <= 64; c += 2) {
    long long n = 0;
    for (const ShotSpec& p : shots) n += (long long)((p.nz + sz - 1) / sz) * ((p.nx + c - 1) / c);
    if (n <= resident) return c;
  }
  return 0;
}

// Shots are packed in order into groups whose working sets fit the arena
// (and at most max_group shots); one tile table spans all shots of a group.
int Engine::build_groups() {
  groups_.clear();
  march_fit_ = march_fit();
  resident_ = march_fit_ ? march_resident(cfg_, false) : 0;
  fresident_ = march_fit_ ? march_resident(cfg_, true) : 0;
  const int ns = (int)specs_.size();
  std::vector<TileRef> tiles, strips;
  const int tx = tile_x(cfg_.txr);
  const bool march = cfg_.tblock >= 2;  // forms 2, 3 march (3: the forward only)
  int s = 0;
  G_ = 0;
  while (s < ns) {
    Group g;
    g.desc0 = s;
    g.tile0 = (int)tiles.size();
    g.strip0 = (int)strips.size();
    size_t used = 0;
    while (s < ns && used + res_[s].work_bytes <= arena_bytes_ &&
           (cfg_.max_group <= 0 || (int)g.shots.size() < cfg_.max_group)) {
      res_[s].work_off = used;
      used += res_[s].work_bytes;
      const ShotSpec& p = specs_[s];
      const int slot = (int)g.shots.size();
      const int nbz = (p.nz + TILE_Z - 1) / TILE_Z, nbx = (p.nx + tx - 1) / tx;
      for (int bx = 0; bx < nbx; ++bx)
        for (int bz = 0; bz < nbz; ++bz) tiles.push_back({slot, bz * TILE_Z, bx * tx, bz + nbz * bx});
      g.shots.push_back(s);
      ++s;
    }
    if (march) {
      // strip width of this group: cfg_.march_w, or with fitting the
      // narrowest even width from 16 whose strips all fit one round of
      // resident marching-adjoint workgroups (a launch that spills a few
      // workgroups into a second round costs most of a round)
      const int sz = TILE_Z * cfg_.march_nw;
      std::vector<ShotSpec> gs;
      for (int si : g.shots) gs.push_back(specs_[si]);
      auto fit = [&](int resident) {
        const int w = march_fit_ ? fitted_march_w(gs, cfg_.march_nw, resident) : 0;
        return w > 0 ? w : cfg_.march_w;
      };
      auto add = [&](int w) {
        for (int slot = 0; slot < (int)g.shots.size(); ++slot) {
          const ShotSpec& p = specs_[g.shots[slot]];
          const int nsx = (p.nx + w - 1) / w, nsz = (p.nz + sz - 1) / sz;
          for (int bx = 0; bx < nsx; ++bx)
            for (int bz = 0; bz < nsz; ++bz) strips.push_back({slot, bz * sz, bx * w, bz + nsz * bx});
        }
      };
      // the adjoint's strips, then the forward's (the forward kernel holds
      // more workgroups per CU, so its fitted strips may be narrower)
      g.march_w = fit(resident_);
      add(g.march_w);
      g.nstrips = (int)strips.size() - g.strip0;
      g.fmarch_w = fit(fresident_);
      g.fstrip0 = g.strip0;
      g.fnstrips = g.nstrips;
      if (g.fmarch_w != g.march_w) {
        g.fstrip0 = (int)strips.size();
        add(g.fmarch_w);
        g.fnstrips = (int)strips.size() - g.fstrip0;
      }
    }
    g.ntiles = (int)tiles.size() - g.tile0;
    if (!march) g.nstrips = 0;
    G_ = std::max(G_, (int)g.shots.size());
    groups_.push_back(g);
  }
  // descriptor variants: 0 migration model, 1 true model, 2 Born; descriptor
  // of shot s at [variant * nshots + s]
  h_desc_.assign((size_t)3 * ns, ShotDev{});
  STCHK(d_desc_.alloc(h_desc_.size() * sizeof(ShotDev), "shot descriptors"));
  STCHK(d_tiles_.alloc(tiles.size() * sizeof(TileRef), "tile table"));
  HIPCHK(hipMemcpy(d_tiles_.p, tiles.data(), tiles.size() * sizeof(TileRef), hipMemcpyHostToDevice));
  if (march) {
    STCHK(d_strips_.alloc(strips.size() * sizeof(TileRef), "strip table"));
    HIPCHK(hipMemcpy(d_strips_.p, strips.data(), strips.size() * sizeof(TileRef), hipMemcpyHostToDevice));
  }
  for (int g = 0; g < (int)groups_.size(); ++g) STCHK(fill_desc(g, false, false));
  return build_lcells();
}

// Band and rim cells of every group's shots (learned layer), one table.
int Engine::build_lcells() {
  d_lcells_.release();
  if (!cfg_.labc_w && !cfg_.labc_wa) return OK;
  std::vector<LCell> all;
  for (Group& g : groups_) {
    std::vector<std::pair<int, int>> grids;
    for (int s : g.shots) grids.push_back({specs_[s].nz, specs_[s].nx});
    std::vector<LCell> cells;
    labc_cell_table(grids, cells, &g.nlband, &g.nlrim);
    if (g.nlband == 0) {
      set_error("learned layer: a grid of this group is smaller than 27 cells on a side (rim + band + patch)");
      return E_ARG;
    }
    g.lcell0 = (int)all.size();
    all.insert(all.end(), cells.begin(), cells.end());
  }
  STCHK(d_lcells_.alloc(all.size() * sizeof(LCell), "learned-layer cells"));
  HIPCHK(hipMemcpy(d_lcells_.p, all.data(), all.size() * sizeof(LCell), hipMemcpyHostToDevice));
  return OK;
}

// A failed cell table (grids too small, no memory) drops both networks, so
// that no later run launches the inference without its table.
int Engine::set_learned(const unsigned char* w) {
  destroy_graphs();  // their forward loops change
  cfg_.labc_w = w;
  const int rc = build_lcells();
  if (rc != OK) cfg_.labc_w = cfg_.labc_wa = nullptr;
  return rc;
}

int Engine::set_learned_adjoint(const unsigned char* w) {
  if (w && cfg_.adj_abc == 2) {
    set_error("learned adjoint layer: not with the R/Src adjoint sponge");
    return E_ARG;
  }
  destroy_graphs();  // their adjoint loops change
  cfg_.labc_wa = w;
  const int rc = build_lcells();
  if (rc != OK) cfg_.labc_w = cfg_.labc_wa = nullptr;
  return rc;
}

// Descriptors of group g (variant 0: migration coefficients; 1: true
// coefficients; 2: Born) pointing at the shots' resident buffers and at the
// group's working slots.
int Engine::fill_desc(int g, bool true_model, bool born_mode) {
  const int variant = born_mode ? 2 : (true_model ? 1 : 0);
  const Group& gr = groups_[g];
  const size_t ns = specs_.size();
  for (size_t k = 0; k < gr.shots.size(); ++k) {
    const int s = gr.shots[k];
    const ShotSpec& p = specs_[s];
    Res& r = res_[s];
    ShotDev d{};
    d.nz = p.nz; d.nx = p.nx; d.ldz = r.ldz; d.nt = cfg_.nt; d.n2 = r.n2;
    d.lev = ring_of(s);
    // slot of level L = (L + kRing) % kRing: non-negative for the levels -1
    // and 0 that the first steps read
    d.lev_stride = 5 * r.n2; d.fstride = r.n2; d.lev_mod = kRing; d.lev_off = kRing;
    if (true_model && !r.coef_true.p) { set_error("no true model set"); return E_ARG; }
    d.coef = true_model ? r.coef_true.as<float>() : r.coef.as<float>();
    d.kz = r.kfac.p ? r.kfac.as<float>() : nullptr;
    d.kx = r.kfac.p ? r.kfac.as<float>() + p.nz : nullptr;
    d.zs = p.zs; d.xs = p.xs; d.src_amp = src_amp_.as<float>();
    d.zr = p.zr; d.rx0 = p.rx0; d.nrx = p.nrx; d.rec = r.rec.as<float>();
    d.snap_fmt = born_mode ? SNAP_HIST : cfg_.snap_fmt;
    d.snap_every = born_mode ? 1 : cfg_.snap_every;
    d.snap_sw = cfg_.txr == 1 ? 0 : cfg_.txr == 2 ? 1 : cfg_.txr == 4 ? 2 : 3;
    d.snap_nbz = (p.nz + TILE_Z - 1) / TILE_Z;
    d.snap_nstrip = r.nstrip;
    d.snap = snap_of(s);
    d.snap_scale = r.scale_bytes ? reinterpret_cast<float*>(snap_of(s) + r.snap_bytes) : nullptr;
    d.img = r.img.as<float>();
    d.imgc = r.imgc.p ? r.imgc.as<float>() : nullptr;
    d.inv_dx = 1.0f / cfg_.dx;
    d.inv_dt = 1.0f / cfg_.dt;
    d.img_w = (float)cfg_.snap_every;
    if (born_mode) {
      d.bsrc = born_buf_.as<float>() + k * 8 * n2max_;
      d.dm = born_buf_.as<float>() + k * 8 * n2max_ + 5 * n2max_;
    }
    h_desc_[variant * ns + s] = d;
  }
  const size_t off = variant * ns + gr.desc0;
  HIPCHK(hipMemcpy(d_desc_.as<ShotDev>() + off, h_desc_.data() + off, gr.shots.size() * sizeof(ShotDev),
                   hipMemcpyHostToDevice));
  return OK;
}

// mode 0: imaging forward (migration model, snapshots); 1: record with the
// true model (store); 2: record with the migration model (subtract);
// 3: record with the migration model (store); 4: Born background (fp32
// history of every level); 5: Born scattered field (record, store).
int Engine::enqueue_forward(int g, int mode, hipStream_t st) {
  const Group& gr = groups_[g];
  const int variant = (mode == 1) ? 1 : (mode >= 4 ? 2 : 0);
  const ShotDev* desc = d_desc_.as<ShotDev>() + (size_t)variant * specs_.size() + gr.desc0;
  const TileRef* tiles = d_tiles_.as<TileRef>() + gr.tile0;
  const bool march = cfg_.tblock >= 2;
  const TileRef* strips = march ? d_strips_.as<TileRef>() + gr.fstrip0 : nullptr;
  const int snap = (mode == 0) ? cfg_.snap_fmt : SNAP_HIST;
  for (int s : gr.shots) {
    STCHK(launch_fill(ring_of(s), (size_t)5 * kRing * res_[s].n2, 0.f, st));
    if ((mode == 0 && cfg_.snap_fmt == SNAP_HIST) || mode == 4)
      STCHK(launch_fill(reinterpret_cast<float*>(snap_of(s)), 2 * res_[s].n2, 0.f, st));  // level 1
  }
  const bool born = mode == 5;
  const int first = born ? 1 : 2;
  // the learned layer corrects every level after its plain update (the
  // source runs: imaging, data synthesis, modelling)
  const bool learn = cfg_.labc_w && mode <= 3;
  const LCell* lcells = learn ? d_lcells_.as<LCell>() + gr.lcell0 : nullptr;
  // levels L >= 3 go two per launch (temporal blocking) where the kernel
  // supports the tile width; Born scattering sources and the learned layer
  // are per level
  const bool pairs = cfg_.tblock == 1 && !born && !learn && cfg_.txr <= 4;
  for (int L = first; L <= cfg_.nt;) {
    const bool two = pairs && L >= 3 && L + 1 <= cfg_.nt;
    int fl = 0;
    if (cfg_.abc > 0 && (born || L >= 3)) fl |= F_K;  // level 2 of the source run is the bare injection
    if (!born) fl |= F_SRC;
    if (born) {
      fl |= F_BORN | F_REC;
      STCHK(launch_born_sources(desc, tiles, gr.ntiles, L, cfg_.txr, st));
    }
    if (mode == 1 || mode == 3) fl |= F_REC;
    if (mode == 2) fl |= F_RECSUB;
    if (mode == 4) fl |= F_SNAP;
    if (nt_stores()) fl |= F_NTS;
    if (mode == 0) {
      if (cfg_.snap_fmt == SNAP_HIST) {
        fl |= F_SNAP;
      } else {
        if ((L - 1) % cfg_.snap_every == 0) fl |= F_SNAP;      // snapshot of level L-1
        if (two && L % cfg_.snap_every == 0) fl |= F_SNAP2;    // and of level L
      }
    }
    if (two)
      STCHK(launch_fwd2(desc, tiles, gr.ntiles, L, fl, cfg_.txr, snap, cfg_.xcd, st));
    else if (march)
      STCHK(launch_fwdm(desc, strips, gr.fnstrips, L, fl, gr.fmarch_w, cfg_.march_nw, snap, cfg_.xcd, st));
    else
      STCHK(launch_fwd(desc, tiles, gr.ntiles, L, fl, cfg_.txr, snap, cfg_.xcd, st));
    if (learn && L >= 3)
      STCHK(launch_labc_infer(cfg_.labc_w, desc, lcells, gr.nlband, gr.nlrim, L,
                              mode == 0 && cfg_.snap_fmt == SNAP_HIST, st));
    for (int Lm = L; Lm <= L + (two ? 1 : 0); ++Lm) {
      if (mode == 3 && cfg_.movie && cfg_.movie_every > 0 && Lm % cfg_.movie_every == 0 &&
          Lm / cfg_.movie_every <= cfg_.movie_n) {
        const int s = gr.shots[0];
        const long long k = Lm / cfg_.movie_every - 1;
        const float* lev = ring_of(s) + (size_t)(Lm % kRing) * 5 * res_[s].n2;
        for (int f = 0; f < 5; ++f)
          HIPCHK(hipMemcpy2DAsync(cfg_.movie + ((size_t)k * 5 + f) * cfg_.movie_nz * cfg_.movie_nx,
                                  cfg_.movie_nz * sizeof(float),
                                  lev + f * res_[s].n2 + cfg_.movie_z0 + (size_t)res_[s].ldz * cfg_.movie_x0,
                                  res_[s].ldz * sizeof(float), cfg_.movie_nz * sizeof(float), cfg_.movie_nx,
                                  hipMemcpyDeviceToDevice, st));
      }
    }
    L += two ? 2 : 1;
  }
  if (mode == 0 && cfg_.snap_fmt != SNAP_HIST && cfg_.nt % cfg_.snap_every == 0)
    STCHK(launch_snap_last(desc, tiles, gr.ntiles, cfg_.nt, cfg_.txr, cfg_.snap_fmt, st));
  return OK;
}

int Engine::enqueue_adjoint(int g, hipStream_t st) {
  const Group& gr = groups_[g];
  const ShotDev* desc = d_desc_.as<ShotDev>() + gr.desc0;
  const TileRef* tiles = d_tiles_.as<TileRef>() + gr.tile0;
  const bool march = cfg_.tblock == 2;  // form 3 marches the forward only
  const TileRef* strips = march ? d_strips_.as<TileRef>() + gr.strip0 : nullptr;
  for (size_t k = 0; k < gr.shots.size(); ++k) {
    Res& r = res_[gr.shots[k]];
    STCHK(launch_fill(ring_of(gr.shots[k]), (size_t)5 * kRing * r.n2, 0.f, st));
    STCHK(launch_fill(r.img.as<float>(), 3 * r.n2, 0.f, st));
    if (cfg_.comp) STCHK(launch_fill(r.imgc.as<float>(), 3 * r.n2, 0.f, st));
  }
  // the two-level adjoint has no R/Src sponge form, and the learned layer
  // corrects every level: one level per launch
  const bool pairs = cfg_.tblock == 1 && cfg_.txr <= 4 && cfg_.adj_abc != 2 && !cfg_.labc_wa;
  const LCell* lcells = cfg_.labc_wa ? d_lcells_.as<LCell>() + gr.lcell0 : nullptr;
  for (int L = cfg_.nt; L >= 1;) {
    const bool two = pairs && L - 1 >= 1;
    int fl = A_INJ;
    if (cfg_.adj_abc && cfg_.abc > 0) fl |= A_K;
    if (cfg_.adj_abc == 2 && cfg_.abc > 0) fl |= A_KSRC;
    if (L % cfg_.snap_every == 0) fl |= A_IMG;
    if (two && (L - 1) % cfg_.snap_every == 0) fl |= A_IMG2;
    if (L < cfg_.nt) fl |= A_NEXT;
    if (cfg_.comp) fl |= A_COMP;
    if (nt_stores()) fl |= A_NTS;
    if (two)
      STCHK(launch_adj2(desc, tiles, gr.ntiles, L, fl, cfg_.txr, cfg_.snap_fmt, cfg_.xcd, st));
    else if (march)
      STCHK(launch_adjm(desc, strips, gr.nstrips, L, fl, gr.march_w, cfg_.march_nw, cfg_.snap_fmt, cfg_.xcd, st));
    else
      STCHK(launch_adj(desc, tiles, gr.ntiles, L, fl, cfg_.txr, cfg_.snap_fmt, cfg_.xcd, st));
    if (lcells) STCHK(launch_labc_infer(cfg_.labc_wa, desc, lcells, gr.nlband, gr.nlrim, L, false, st, true));
    L -= two ? 2 : 1;
  }
  if (cfg_.comp)
    for (size_t k = 0; k < gr.shots.size(); ++k) {
      Res& r = res_[gr.shots[k]];
      STCHK(launch_image_finalize(r.img.as<float>(), r.imgc.as<float>(), 3 * r.n2, st));
    }
  return OK;
}

// Step launches of one forward run of a mode / of one adjoint run (the
// loops of enqueue_forward / enqueue_adjoint).
long long Engine::fwd_launch_count(int mode) const {
  const bool born = mode == 5;
  const bool pairs = cfg_.tblock == 1 && !born && !(cfg_.labc_w && mode <= 3) && cfg_.txr <= 4;
  long long n = 0;
  for (int L = born ? 1 : 2; L <= cfg_.nt; ++n) L += (pairs && L >= 3 && L + 1 <= cfg_.nt) ? 2 : 1;
  return n;
}
long long Engine::adj_launch_count() const {
  const bool pairs = cfg_.tblock == 1 && cfg_.txr <= 4 && cfg_.adj_abc != 2 && !cfg_.labc_wa;
  long long n = 0;
  for (int L = cfg_.nt; L >= 1; ++n) L -= (pairs && L - 1 >= 1) ? 2 : 1;
  return n;
}

// which: 0 imaging forward, 1 adjoint+imaging, 2 data synthesis (both runs)
int Engine::run_phase(int g, int which, hipStream_t st) {
  Group& gr = groups_[g];
  hipGraphExec_t* slot = which == 0 ? &gr.gfwd : which == 1 ? &gr.gadj : &gr.gsyn;
  auto enqueue = [&]() -> int {
    if (which == 0) return enqueue_forward(g, 0, st);
    if (which == 1) return enqueue_adjoint(g, st);
    for (size_t k = 0; k < gr.shots.size(); ++k) {
      Res& r = res_[gr.shots[k]];
      STCHK(launch_fill(r.rec.as<float>(), (size_t)5 * cfg_.nt * std::max(specs_[gr.shots[k]].nrx, 1), 0.f, st));
    }
    STCHK(enqueue_forward(g, 1, st));
    return enqueue_forward(g, 2, st);
  };
  if (!cfg_.use_graph) return enqueue();
  if (!*slot) {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc != OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) {
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return E_HIP;
    }
    e = hipGraphInstantiate(slot, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      *slot = nullptr;
      set_error("graph instantiation failed: %s", hipGetErrorString(e));
      return E_HIP;
    }
  }
  HIPCHK(hipGraphLaunch(*slot, st));
  return OK;
}

static double cells_of(const ShotSpec& p) { return (double)p.nz * (double)p.nx; }

int Engine::set_true_model(const std::vector<const float*>& coef5_true) {
  if (coef5_true.size() != specs_.size()) { set_error("true-model list size mismatch"); return E_ARG; }
  for (size_t s = 0; s < specs_.size(); ++s) {
    Res& r = res_[s];
    STCHK(r.coef_true.alloc(5 * r.n2 * sizeof(float), "true-model coefficients"));
    STCHK(fill_masked_coefs(coef5_true[s], specs_[s].nz, specs_[s].nx, r.ldz, r.coef_true.as<float>()));
  }
  for (int g = 0; g < (int)groups_.size(); ++g) STCHK(fill_desc(g, true, false));
  return OK;
}

int Engine::set_data(int shot, const float* h) {
  if (shot < 0 || shot >= (int)specs_.size() || !h) { set_error("bad shot index or null data"); return E_ARG; }
  const int nt = cfg_.nt, nr = specs_[shot].nrx;
  std::vector<float> t((size_t)5 * nt * nr);
  for (int f = 0; f < 5; ++f)
    for (int ix = 0; ix < nr; ++ix)
      for (int it = 0; it < nt; ++it) t[((size_t)f * nt + it) * nr + ix] = h[(size_t)f * nt * nr + it + (size_t)nt * ix];
  if (!t.empty()) HIPCHK(hipMemcpy(res_[shot].rec.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
  return OK;
}

int Engine::get_data(int shot, float* h) {
  if (shot < 0 || shot >= (int)specs_.size() || !h) { set_error("bad shot index or null data"); return E_ARG; }
  HIPCHK(hipStreamSynchronize(st_));
  const int nt = cfg_.nt, nr = specs_[shot].nrx;
  std::vector<float> t((size_t)5 * nt * nr);
  if (!t.empty()) HIPCHK(hipMemcpy(t.data(), res_[shot].rec.p, t.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int f = 0; f < 5; ++f)
    for (int ix = 0; ix < nr; ++ix)
      for (int it = 0; it < nt; ++it) h[(size_t)f * nt * nr + it + (size_t)nt * ix] = t[((size_t)f * nt + it) * nr + ix];
  return OK;
}

int Engine::synthesize() {
  for (auto& r : res_)
    if (!r.coef_true.p) { set_error("synthesize: set the true model first"); return E_ARG; }
  for (int g = 0; g < (int)groups_.size(); ++g) STCHK(run_phase(g, 2, st_));
  HIPCHK(hipStreamSynchronize(st_));
  return OK;
}

int Engine::model_data() {
  tm_ = Timing{};
  HIPCHK(hipEventRecord(ev_[0], st_));
  for (int g = 0; g < (int)groups_.size(); ++g) {
    for (int s : groups_[g].shots)
      STCHK(launch_fill(res_[s].rec.as<float>(), (size_t)5 * cfg_.nt * std::max(specs_[s].nrx, 1), 0.f, st_));
    STCHK(enqueue_forward(g, 3, st_));
    for (int s : groups_[g].shots) tm_.fwd_cells += cells_of(specs_[s]) * (cfg_.nt - 1);
    tm_.fwd_launches += fwd_launch_count(3);
    last_group_ = g;
  }
  HIPCHK(hipEventRecord(ev_[1], st_));
  HIPCHK(hipStreamSynchronize(st_));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
  tm_.fwd_ms = tm_.total_ms = ms;
  return OK;
}

int Engine::migrate() {
  tm_ = Timing{};
  if (stack_.p) HIPCHK(hipMemsetAsync(stack_.p, 0, stack_.bytes, st_));
  for (int g = 0; g < (int)groups_.size(); ++g) {
    HIPCHK(hipEventRecord(ev_[0], st_));
    STCHK(run_phase(g, 0, st_));
    HIPCHK(hipEventRecord(ev_[1], st_));
    STCHK(run_phase(g, 1, st_));
    HIPCHK(hipEventRecord(ev_[2], st_));
    for (int s : groups_[g].shots) {
      const ShotSpec& p = specs_[s];
      if (stack_.p && p.img_nz > 0 && p.img_nx > 0)
        STCHK(launch_stack(stack_.as<double>(), cfg_.stack_nz, cfg_.stack_nx, res_[s].img.as<float>(), res_[s].ldz,
                           res_[s].n2, p.img_z0, p.img_x0, p.img_nz, p.img_nx, p.stack_z0, p.stack_x0, st_));
    }
    HIPCHK(hipEventRecord(ev_[3], st_));
    HIPCHK(hipStreamSynchronize(st_));
    float a = 0, b = 0, c = 0;
    HIPCHK(hipEventElapsedTime(&a, ev_[0], ev_[1]));
    HIPCHK(hipEventElapsedTime(&b, ev_[1], ev_[2]));
    HIPCHK(hipEventElapsedTime(&c, ev_[0], ev_[3]));
    tm_.fwd_ms += a;
    tm_.adj_ms += b;
    tm_.total_ms += c;
    tm_.fwd_launches += fwd_launch_count(0);
    tm_.adj_launches += adj_launch_count();
    for (int s : groups_[g].shots) {
      tm_.fwd_cells += cells_of(specs_[s]) * (cfg_.nt - 1);
      tm_.adj_cells += cells_of(specs_[s]) * cfg_.nt;
    }
    last_group_ = g;
  }
  return OK;
}

int Engine::born(const std::vector<const float*>& dm) {
  if (groups_.size() != 1) { set_error("Born modelling runs single-group engines only"); return E_ARG; }
  if (dm.size() != specs_.size()) { set_error("dm list size mismatch"); return E_ARG; }
  if (cfg_.snap_fmt != SNAP_HIST) { set_error("Born modelling needs the fp32 history format"); return E_ARG; }
  STCHK(born_buf_.alloc((size_t)G_ * 8 * n2max_ * sizeof(float), "Born sources"));
  STCHK(launch_fill(born_buf_.as<float>(), (size_t)G_ * 8 * n2max_, 0.f, st_));
  HIPCHK(hipStreamSynchronize(st_));
  for (size_t k = 0; k < specs_.size(); ++k) {
    const ShotSpec& p = specs_[k];
    float* d = born_buf_.as<float>() + k * 8 * n2max_ + 5 * n2max_;
    for (int c = 0; c < 3; ++c)  // component stride = the shot's own n2 (as the kernels index it)
      HIPCHK(hipMemcpy2D(d + c * res_[k].n2, res_[k].ldz * sizeof(float), dm[k] + (size_t)c * p.nz * p.nx,
                         p.nz * sizeof(float), p.nz * sizeof(float), p.nx, hipMemcpyHostToDevice));
    STCHK(launch_fill(res_[k].rec.as<float>(), (size_t)5 * cfg_.nt * std::max(p.nrx, 1), 0.f, st_));
  }
  STCHK(fill_desc(0, false, true));
  tm_ = Timing{};
  STCHK(enqueue_forward(0, 4, st_));
  HIPCHK(hipEventRecord(ev_[0], st_));
  STCHK(enqueue_forward(0, 5, st_));
  HIPCHK(hipEventRecord(ev_[1], st_));
  HIPCHK(hipStreamSynchronize(st_));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
  tm_.fwd_ms = tm_.total_ms = ms;
  tm_.fwd_launches = fwd_launch_count(5);
  last_group_ = 0;
  return OK;
}

int Engine::get_image(int shot, float* img3) {
  if (shot < 0 || shot >= (int)specs_.size() || !img3) { set_error("bad shot index or null output"); return E_ARG; }
  HIPCHK(hipStreamSynchronize(st_));
  const ShotSpec& p = specs_[shot];
  const Res& r = res_[shot];
  for (int c = 0; c < 3; ++c)
    HIPCHK(hipMemcpy2D(img3 + (size_t)c * p.nz * p.nx, p.nz * sizeof(float), r.img.as<float>() + c * r.n2,
                       r.ldz * sizeof(float), p.nz * sizeof(float), p.nx, hipMemcpyDeviceToHost));
  return OK;
}

int Engine::get_stack(float* img3) {
  if (!stack_.p || !img3) { set_error("no stacked image configured"); return E_ARG; }
  HIPCHK(hipStreamSynchronize(st_));
  const size_t n = (size_t)3 * cfg_.stack_nz * cfg_.stack_nx;
  std::vector<double> h(n);
  HIPCHK(hipMemcpy(h.data(), stack_.p, n * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) img3[i] = (float)h[i];
  return OK;
}

int Engine::get_last_levels(int shot, float* last10) {
  if (last_group_ < 0) { set_error("nothing has run"); return E_ARG; }
  const Group& gr = groups_[last_group_];
  auto it = std::find(gr.shots.begin(), gr.shots.end(), shot);
  if (it == gr.shots.end() || !last10) { set_error("shot %d is not in the last group run", shot); return E_ARG; }
  HIPCHK(hipStreamSynchronize(st_));
  const float* base = ring_of(shot);
  const ShotSpec& p = specs_[shot];
  const Res& r = res_[shot];
  const size_t nn = (size_t)p.nz * p.nx;
  const int nt = cfg_.nt;
  for (int lv = 0; lv < 2; ++lv) {
    const int L = nt - 1 + lv;  // levels nt-1, nt (slot L % kRing)
    for (int f = 0; f < 5; ++f)
      HIPCHK(hipMemcpy2D(last10 + (size_t)(5 * lv + f) * nn, p.nz * sizeof(float),
                         base + (size_t)(L % kRing) * 5 * r.n2 + f * r.n2, r.ldz * sizeof(float), p.nz * sizeof(float),
                         p.nx, hipMemcpyDeviceToHost));
  }
  return OK;
}

}  // namespace erm
