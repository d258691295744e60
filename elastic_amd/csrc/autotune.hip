// autotune.hip -- tile-width autotuner of the 2D step kernels.
//
// The step kernels are compiled for four tile widths (4*TXR x-cells by 64
// z-cells, TXR in {1,2,4,8}). Which one is fastest depends on the grid sizes,
// the snapshot format and the imaging decimation: wider tiles re-read less
// halo but hold more registers per thread. The tuner runs the real engine on
// a short copy of the workload (the shots of the full run's first group, a
// few dozen levels, the same options) once per candidate and keeps the
// fastest. Results are cached for the process, keyed by the workload
// signature, and, when RTM_TUNE_CACHE names a file, in that file (one "key
// code" line per entry, code = txr + 16 * form + 256 * strip width). The
// two-level (temporally blocked) kernels are candidates too: they halve the
// HBM traffic of the fields but recompute a radius-5 ring of the first level
// and run at lower occupancy; so are the marching forms (workgroups per
// strip walking along x, halo columns read once per strip) at strip widths
// 16, 24, 28 and 32. Which form wins is measured, not assumed.
#include "engine2d.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

namespace erm {

static std::mutex g_tune_mu;
static std::map<std::string, int> g_tune_cache;

static std::string tune_key(const EngineCfg& c, const std::vector<ShotSpec>& sp) {
  long long cells = 0;
  int nzmax = 0, nxmax = 0;
  for (const ShotSpec& s : sp) {
    cells += (long long)s.nz * s.nx;
    nzmax = std::max(nzmax, s.nz);
    nxmax = std::max(nxmax, s.nx);
  }
  hipDeviceProp_t prop{};
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipGetDeviceProperties(&prop, dev);
  char buf[320];
  // shapes bucketed to 64 cells, mean cells per shot to 2^16; the marching
  // strip height and width selection (fitting on or off, a fixed width)
  const char* mw_env = getenv("RTM_MARCH_W");
  std::snprintf(buf, sizeof buf, "%s/nz%d/nx%d/c%lld/n%d/f%d/j%d/k%d/a%d/g%d/h%d/fit%d/w%d", prop.gcnArchName,
                (nzmax + 63) / 64, (nxmax + 63) / 64, (cells / (long long)sp.size()) >> 16,
                (int)std::min<size_t>(sp.size(), 8), c.snap_fmt, c.snap_every, c.comp, c.adj_abc, c.use_graph,
                c.march_nw, march_fit() ? 1 : 0, mw_env ? std::atoi(mw_env) : 0);
  for (char* p = buf; *p; ++p)
    if (*p == ' ' || *p == ':') *p = '_';
  return buf;
}

static void load_file_cache() {
  static bool done = false;
  if (done) return;
  done = true;
  const char* path = getenv("RTM_TUNE_CACHE");
  if (!path) return;
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  char key[512];
  int txr = 0;
  while (std::fscanf(f, "%511s %d", key, &txr) == 2) {
    const int w = txr & 15, tb = (txr >> 4) & 15, mw = txr >> 8;
    if ((w == 1 || w == 2 || w == 4 || w == 8) && (tb == 0 || (tb == 1 && w <= 4) || (tb >= 2 && tb <= 3 && w == 2)) &&
        mw >= 0 && (mw & 1) == 0)
      g_tune_cache[key] = txr;
  }
  std::fclose(f);
}

static void append_file_cache(const std::string& key, int txr) {
  const char* path = getenv("RTM_TUNE_CACHE");
  if (!path) return;
  FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%s %d\n", key.c_str(), txr);
  std::fclose(f);
}

int autotune_txr(const EngineCfg& cfg, const std::vector<ShotSpec>& shots, const std::vector<const float*>& coef5,
                 const float* wavelet, int* best, int* best_tblock, double ms_out[8], bool use_cache,
                 int* best_march_w) {
  if (shots.empty() || coef5.size() != shots.size() || !wavelet || !best || !best_tblock) {
    set_error("autotune: bad arguments");
    return E_ARG;
  }
  const std::string key = tune_key(cfg, shots);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    load_file_cache();
    auto it = g_tune_cache.find(key);
    if (use_cache && it != g_tune_cache.end()) {
      *best = it->second & 15;
      *best_tblock = (it->second >> 4) & 15;
      if (best_march_w && (it->second >> 8) > 0) *best_march_w = it->second >> 8;
      if (ms_out)
        for (int i = 0; i < 8; ++i) ms_out[i] = -1.0;
      return OK;
    }
  }
  // A short copy of the workload: the shots of the full run's first group
  // (at most RTM_TUNE_MAXSHOTS, default 24: how many strips or tiles one
  // launch holds decides how the kernel forms fill the GPU; with 8 the
  // 23-shot BF16 groups of the 64-shot survey were ranked wrongly), enough
  // levels for several snapshot periods, the same options; one group, no
  // stack.
  const char* cap_env = getenv("RTM_TUNE_MAXSHOTS");
  const int cap = std::max(1, cap_env ? std::atoi(cap_env) : 24);
  const int ns = std::min<int>({(int)shots.size(), cap, first_group_size(cfg, shots)});
  std::vector<ShotSpec> sp(shots.begin(), shots.begin() + ns);
  std::vector<const float*> cf(coef5.begin(), coef5.begin() + ns);
  for (ShotSpec& s : sp) s.img_nz = s.img_nx = 0;
  EngineCfg c = cfg;
  // 96 levels (16 snapshot periods when decimated): on 48 the marching
  // widths of the 23-shot BF16 groups were ranked differently from run to
  // run, 5% apart on the full run
  c.nt = std::max(96, 16 * c.snap_every);
  c.nt = std::min(c.nt, cfg.nt);
  c.stack_nz = c.stack_nx = 0;
  c.max_group = 0;
  // candidates (width, form): widths 1, 2, 4, 8 with one level per launch
  // (slots 0..3), widths 1, 2, 4 with two levels per launch (slots 4..6) and
  // the marching forms (slot 7: the faster of marching forward and adjoint,
  // form 2, and marching forward with the tile adjoint, form 3)
  int cand[8][2] = {{1, 0}, {2, 0}, {4, 0}, {8, 0}, {1, 1}, {2, 1}, {4, 1}, {2, 2}};
  double t[8];
  int arg = -1;
  // marching strip widths tried (RTM_MARCH_W fixes one). A group whose
  // strips fit one round of resident workgroups at some width gets that
  // width from the engine whatever the candidate, so then only it is timed.
  constexpr int kMW = 4;
  const int mws[kMW] = {16, 24, 28, 32};
  const int wfit = march_fit() ? fitted_march_w(sp, c.march_nw, march_resident(c, false)) : 0;
  const int nmw = getenv("RTM_MARCH_W") || wfit > 0 ? 1 : kMW;
  int mw_run = wfit > 0 ? wfit : cfg.march_w, mw_best = mw_run;
  auto time_form = [&](int txr, int form, double* ms) -> int {
    c.txr = txr;
    c.tblock = form;
    c.march_w = mw_run;
    Engine e;
    STCHK(e.setup(c, sp, cf, wavelet));
    STCHK(e.migrate());  // warm-up (and graph instantiation)
    // the faster of two timed runs
    STCHK(e.migrate());
    const double t1 = e.timing().total_ms;
    STCHK(e.migrate());
    *ms = std::min(t1, e.timing().total_ms);
    return OK;
  };
  for (int i = 0; i < 8; ++i) {
    if (i < 7) {
      STCHK(time_form(cand[i][0], cand[i][1], &t[i]));
    } else {
      // marching: forms 2 and 3 at each strip width; the fastest stands for
      // the slot.
      t[7] = -1;
      double tw[kMW] = {-1, -1, -1, -1};
      int fw[kMW] = {2, 2, 2, 2};
      for (int m = 0; m < nmw; ++m) {
        if (nmw > 1) mw_run = mws[m];
        for (int form = 2; form <= 3; ++form) {
          double tm = 0;
          STCHK(time_form(2, form, &tm));
          if (tw[m] < 0 || tm < tw[m]) {
            tw[m] = tm;
            fw[m] = form;
          }
        }
        if (t[7] < 0 || tw[m] < t[7]) t[7] = tw[m];
      }
      int mi = 0;
      for (int m = 0; m < nmw; ++m)
        if (tw[m] < tw[mi]) mi = m;
      cand[7][1] = fw[mi];
      mw_best = nmw == 1 ? mw_run : mws[mi];
    }
    if (arg < 0 || t[i] < t[arg]) arg = i;
  }
  // The marching forms read 1.05x (forward) and 1.08x (adjoint) their
  // algorithmic bytes, the tile forms 1.01x and 1.02x (priming columns
  // re-read by neighbouring strips); they are taken only when clearly faster
  // on the short run, not on a near-tie that a longer, bandwidth-bound run
  // would lose.
  if (arg == 7) {
    int alt = -1;
    for (int i = 0; i < 7; ++i)
      if (t[i] > 0 && (alt < 0 || t[i] < t[alt])) alt = i;
    if (alt >= 0 && t[7] > 0.97 * t[alt]) arg = alt;
  }
  *best = cand[arg][0];
  *best_tblock = cand[arg][1];
  const int mw_out = cand[arg][1] >= 2 ? mw_best : 0;
  if (best_march_w && mw_out > 0) *best_march_w = mw_out;
  if (ms_out)
    for (int i = 0; i < 8; ++i) ms_out[i] = t[i];
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const int code = *best | (*best_tblock << 4) | (mw_out << 8);
  g_tune_cache[key] = code;
  append_file_cache(key, code);
  return OK;
}

}  // namespace erm
