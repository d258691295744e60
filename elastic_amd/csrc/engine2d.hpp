// engine2d.hpp -- multi-shot 2D elastic RTM engine (host side).
//
// All inputs (coefficients, observed data, the source wavelet) are made
// resident in HBM at setup. Shots are processed in groups sized to the free
// device memory; a group's shots (uneven grids allowed) share one tile table,
// so each time level is ONE kernel launch over all of them. The forward and
// adjoint loops of a group can be captured once in HIP graphs and replayed.
#pragma once

#include <utility>
#include <vector>

#include "elastic2d.hpp"

namespace erm {

// First 0-based column and row of each coefficient's operator range (the
// last ones are nx-6, nz-6); cells outside are zeroed.
extern const int kMaskIx0[5];
extern const int kMaskIz0[5];

struct ShotSpec {
  int nz, nx;                     // computational grid
  int xs, zs, zr, rx0, nrx;       // source cell, receiver row and columns
  int img_z0, img_x0, img_nz, img_nx;  // image region added into the stack
  int stack_z0, stack_x0;         // its position in the stacked image
};

struct EngineCfg {
  int nt = 0;
  float dx = 0, dt = 0;
  int abc = 0;         // sponge width (<=0: none)
  int adj_abc = 0;     // 0: adjoint without sponge, 1: exact adjoint sponge
  int snap_every = 1;  // imaging decimation
  int snap_fmt = SNAP_HIST;
  int comp = 0;        // compensated image sums
  int txr = 4;         // tile width (x cells per thread)
  int xcd = 1;         // XCD-aware tile numbering
  int use_graph = 0;   // capture the time loops in HIP graphs
  int tblock = 0;      // kernel form: 0 one level per launch on tiles, 1 two
                       // levels per launch (temporal blocking), 2 one level
                       // per launch marching along x (tile width 2 only), 3
                       // the marching forward with the tile adjoint
  int march_w = 32;    // strip width (columns, even) of the marching form
  int march_nw = 2;    // strip height of the marching form in waves (64 rows each): 1, 2, 4
  int max_group = 0;   // cap on shots per group (0 = memory bound only)
  int stack_nz = 0, stack_nx = 0;  // stacked image (0 = none)
  double mem_fraction = 0.85;      // share of free memory the working set may use
  // Wavefield movie of a one-shot model_data() run: at every level L that is
  // a multiple of movie_every (up to movie_n frames) the five fields of the
  // window [movie_z0, +movie_nz) x [movie_x0, +movie_nx) are copied to
  // movie (device, [frame][field][x][z]).
  int movie_every = 0, movie_n = 0;
  int movie_z0 = 0, movie_x0 = 0, movie_nz = 0, movie_nx = 0;
  float* movie = nullptr;
  // Learned absorbing layer: the network's packed inference weights on the
  // device (null = none). The forward runs then apply the network's correction to
  // the band cells of every level L >= 3 after its plain update (and clear
  // the rim), one level per launch; the adjoint has its own network
  // (labc_wa).
  const unsigned char* labc_w = nullptr;
  // The adjoint network (null = none): the adjoint runs then correct the
  // band cells of every level after its step, one level per launch. The
  // imaging of a level happens in its step, before the correction: only
  // images of band cells (outside the physical region) miss it.
  const unsigned char* labc_wa = nullptr;
};

struct Timing {
  double fwd_ms = 0, adj_ms = 0, total_ms = 0;
  long long fwd_launches = 0, adj_launches = 0;
  double fwd_cells = 0, adj_cells = 0;  // cell-updates (computational grid)
};

// The level ring of a shot's working set holds four levels, so that a
// two-level forward step writes levels other than the two it reads.
constexpr int kEngineRing = 4;

// Learned absorbing layer (learned_abc.hip): cell table entry of the
// inference launch -- a band cell (ed = edge | depth << 2), a padding entry
// (ed = kLDead) or a rim cell (ed = -1) of grid `shot`.
struct LCell {
  int shot, iz, ix, ed;
};
constexpr int kLDead = 32;
// Band cells, then rim cells, of the grids (nz, nx) listed.
void labc_cell_table(const std::vector<std::pair<int, int>>& grids, std::vector<LCell>& cells, int* nband,
                     int* nrim);
// After the plain update of level L: band cells += learned correction from
// levels L-1 and L-2, rim cells = 0 (w: the network's packed inference
// weights).
// With hist, the correction of Vx, Vz also goes into the fp32 source
// history of level L (SNAP_HIST), which the step wrote before it.
// With adj, after an adjoint step: the correction comes from levels L+1 and
// L+2 of an adjoint state (hist ignored).
int launch_labc_infer(const unsigned char* w, const ShotDev* shots, const LCell* cells, int nband, int nrim, int L,
                      bool hist, hipStream_t st, bool adj = false);

// Strip-width fitting of the marching forms (see Engine::build_groups).
bool march_fit();
int march_resident(const EngineCfg& c, bool fwd);
int fitted_march_w(const std::vector<ShotSpec>& shots, int march_nw, int resident);

class Engine {
 public:
  ~Engine();
  int setup(const EngineCfg& cfg, const std::vector<ShotSpec>& shots,
            const std::vector<const float*>& coef5, const float* wavelet);
  // Second coefficient set per shot (the "true" model) for data synthesis.
  int set_true_model(const std::vector<const float*>& coef5_true);
  int set_data(int shot, const float* rec_host);  // 5*(nt,nrx)
  int get_data(int shot, float* rec_host);
  // observed = forward(true) - forward(migration) for every shot
  int synthesize();
  // forward + adjoint + imaging of every shot, images stacked
  int migrate();
  // forward of every shot with the migration model, recording receivers
  int model_data();
  // Born modelling of one-group runs (single shot API); dm 3*(nz,nx) per shot
  int born(const std::vector<const float*>& dm);
  int get_image(int shot, float* img3);          // 3*(nz,nx)
  int get_stack(float* img3);                    // 3*(stack_nz,stack_nx)
  int get_last_levels(int shot, float* last10);  // levels nt-1, nt of the last run
  // Use (w = the network's packed inference weights on the device) or drop
  // (null) the learned layer in the forward runs.
  int set_learned(const unsigned char* w);
  // The same for the adjoint runs (the network trained on adjoint states);
  // not with the R/Src adjoint sponge.
  int set_learned_adjoint(const unsigned char* w);
  const Timing& timing() const { return tm_; }
  // strip width of the first group's marching launches: the adjoint's for
  // form 2, the forward's for form 3 (cfg march_w when the form does not
  // march)
  int first_march_w() const {
    if (groups_.empty()) return cfg_.march_w;
    const int w = cfg_.tblock == 3 ? groups_[0].fmarch_w : groups_[0].march_w;
    return w > 0 ? w : cfg_.march_w;
  }
  int group_size() const { return G_; }
  int ngroups() const { return (int)groups_.size(); }

 private:
  struct Res {
    int ldz = 0;
    long long n2 = 0;
    int nstrip = 0;
    size_t snap_bytes = 0, scale_bytes = 0, work_bytes = 0;  // working set of one shot
    size_t work_off = 0;  // offset of its working set in the arena (within its group)
    DevBuf coef, coef_true, kfac, rec, img, imgc;
  };
  float* ring_of(int s) const { return reinterpret_cast<float*>(work_.as<char>() + res_[s].work_off); }
  static constexpr int kRing = kEngineRing;
  char* snap_of(int s) const {
    return work_.as<char>() + res_[s].work_off + (size_t)5 * kRing * res_[s].n2 * sizeof(float);
  }
  struct Group {
    std::vector<int> shots;
    int tile0 = 0, ntiles = 0;  // slice of the tile table
    int strip0 = 0, nstrips = 0;  // slice of the strip table (marching form)
    int march_w = 0;              // strip width of this group (marching adjoint)
    int fstrip0 = 0, fnstrips = 0;  // slice of the strip table for the marching forward
    int fmarch_w = 0;               // its strip width
    int desc0 = 0;              // first descriptor
    int lcell0 = 0, nlband = 0, nlrim = 0;  // slice of the learned layer's cell table
    hipGraphExec_t gfwd = nullptr, gadj = nullptr, gsyn = nullptr;
  };
  int build_lcells();
  int alloc_working();
  int build_groups();
  int fill_desc(int g, bool true_model, bool born_mode);
  int enqueue_forward(int g, int mode, hipStream_t st);
  int enqueue_adjoint(int g, hipStream_t st);
  int run_phase(int g, int which, hipStream_t st);
  long long fwd_launch_count(int mode) const;
  long long adj_launch_count() const;
  void destroy_graphs();

  EngineCfg cfg_;
  std::vector<ShotSpec> specs_;
  std::vector<Res> res_;
  std::vector<Group> groups_;
  int G_ = 0;
  bool march_fit_ = false;  // per-group strip widths (march_fit())
  int resident_ = 0;        // resident marching-adjoint workgroups on the device
  int fresident_ = 0;       // resident marching-forward workgroups
  long long n2max_ = 0;
  size_t arena_bytes_ = 0;
  DevBuf src_amp_, work_, d_desc_, d_tiles_, d_strips_, stack_, born_buf_, d_lcells_;
  std::vector<ShotDev> h_desc_;
  hipStream_t st_ = nullptr;
  hipEvent_t ev_[4] = {};
  int last_group_ = -1;
  Timing tm_;
};

// The engine of a batch handle (elastic2d_host.hip).
Engine* batch_engine(struct rtm_batch* b);

int fill_masked_coefs(const float* coef5, int nz, int nx, int ldz, float* dev);
void sponge_factors(int n, int abc, std::vector<float>& k);
int ensure_device();
int default_txr();
int default_xcd();
int default_tblock();
int default_march_w();
int default_march_nw();
bool march_fit();
// Shots in the first group a full run of cfg would form: those, in order,
// whose working sets fit the arena the engine would allocate from the
// memory free now.
int first_group_size(const EngineCfg& cfg, const std::vector<ShotSpec>& shots);
// Tile autotuner (autotune.hip): times the engine on a short copy of the
// workload for TXR 1, 2, 4, 8 with one level per launch, TXR 1, 2, 4 with
// two, and the marching forms (strips of 16, 24, 28 and 32 columns unless
// RTM_MARCH_W fixes one), and returns the fastest in *best / *best_tblock
// and, when a marching form wins, its strip width in *best_march_w (if not
// null). ms_out: the 8 slots' times, -1 when the answer came from the cache.
int autotune_txr(const EngineCfg& cfg, const std::vector<ShotSpec>& shots, const std::vector<const float*>& coef5,
                 const float* wavelet, int* best, int* best_tblock, double ms_out[8], bool use_cache,
                 int* best_march_w = nullptr);

}  // namespace erm
