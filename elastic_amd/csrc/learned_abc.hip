// learned_abc.hip -- learned absorbing layer: band geometry, features,
// training (forward, backward, Adam) and inference kernels (gfx950).
//
// The network is small (9093 parameters) and evaluated on the ~6% of cells
// that form the 8-cell band of a grid. Training and inference both use BF16-
// rounded weights and features with FP32 accumulation; the training step keeps
// FP32 master weights and Adam moments.

#include "../../include/learned_abc.h"
#include "engine2d.hpp"

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

namespace erm {
namespace labc {

constexpr int R = LABC_RIM, W = LABC_WIDTH, NF = LABC_NFEAT, KP = LABC_KPAD, H = LABC_HID,
              NO = LABC_NOUT;
constexpr int OW1 = 0, OB1 = NF * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H * NO, NP = OB3 + NO;
static_assert(NP == LABC_NPARAM, "parameter count");
constexpr float Z0 = LABC_Z0;
constexpr int NRAW = 230;  // wavefield features (scaled by the cell's max-abs)

struct Band {
  int nz, nx, nT, nL, total;
};
__host__ __device__ inline Band band_of(int nz, int nx) {
  Band b;
  b.nz = nz;
  b.nx = nx;
  b.nT = W * (nx - 2 * R);
  b.nL = W * (nz - 2 * R - 2 * W);
  b.total = 2 * b.nT + 2 * b.nL;
  return b;
}
// edge 0 top (inward +z), 1 bottom (-z), 2 left (+x), 3 right (-x); depth 0
// is the band row/column next to the rim.
__host__ __device__ inline void cell_coords(const Band& b, int c, int& iz, int& ix, int& edge,
                                            int& depth) {
  const int cols = b.nx - 2 * R, rows = b.nz - 2 * R - 2 * W;
  if (c < b.nT) {
    depth = c / cols; iz = R + depth; ix = R + c % cols; edge = 0;
    return;
  }
  c -= b.nT;
  if (c < b.nT) {
    depth = c / cols; iz = b.nz - R - 1 - depth; ix = R + c % cols; edge = 1;
    return;
  }
  c -= b.nT;
  if (c < b.nL) {
    depth = c / rows; ix = R + depth; iz = R + W + c % rows; edge = 2;
    return;
  }
  c -= b.nL;
  depth = c / rows; ix = b.nx - R - 1 - depth; iz = R + W + c % rows; edge = 3;
}
__host__ __device__ inline int cell_index(const Band& b, int iz, int ix) {
  const int cols = b.nx - 2 * R, rows = b.nz - 2 * R - 2 * W;
  if (ix < R || ix >= b.nx - R || iz < R || iz >= b.nz - R) return -1;
  if (iz < R + W) return (iz - R) * cols + (ix - R);
  if (iz >= b.nz - R - W) return b.nT + (b.nz - R - 1 - iz) * cols + (ix - R);
  if (ix < R + W) return 2 * b.nT + (ix - R) * rows + (iz - R - W);
  if (ix >= b.nx - R - W) return 2 * b.nT + b.nL + (b.nx - R - 1 - ix) * rows + (iz - R - W);
  return -1;
}

__device__ __forceinline__ float bf16r(float x) {  // round to bf16 (nearest even)
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}
static inline float bf16r_host(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float y;
  std::memcpy(&y, &u, 4);
  return y;
}
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678f)); }
__device__ __forceinline__ float gelu_d(float x) {
  return 0.5f * (1.f + erff(x * 0.70710678f)) + x * 0.39894228f * expf(-0.5f * x * x);
}

// Wavefield value of field f at (iz, ix) scaled to velocity units, 0 outside.
// Fields divided by Z0 in the features, targets and corrections: the
// stresses of a forward state, the velocities of an adjoint state (in each,
// the larger pair, by about the impedance).
__host__ __device__ inline bool z_scaled(int f, int adj) { return adj ? f < 2 : f >= 2; }

struct Lev {
  const float* p;  // field f at p + f*fs
  int fs, ldz, nz, nx;
  int adj;  // 1: an adjoint state
  __device__ __forceinline__ float get(int f, int iz, int ix) const {
    if (iz < 0 || iz >= nz || ix < 0 || ix >= nx) return 0.f;
    const float v = p[f * fs + iz + ldz * ix];
    return z_scaled(f, adj) ? v * (1.f / Z0) : v;
  }
};

// Raw (unnormalised) feature k of a band cell and its scale contribution.
__device__ inline float raw_feature(const Lev& m1, const Lev& m2, int iz, int ix, int edge, int depth,
                                    int k) {
  if (k < 225) {
    const int fm = k / 9, kn = k % 9, f = fm / 5, m = fm % 5 - 2;
    const int dz = edge == 0 ? 1 : edge == 1 ? -1 : 0;
    const int dx = edge == 2 ? 1 : edge == 3 ? -1 : 0;
    const int tz = (edge >= 2) ? 1 : 0, tx = (edge < 2) ? 1 : 0;
    return m1.get(f, iz + kn * dz + m * tz, ix + kn * dx + m * tx);
  }
  if (k < 230) return m2.get(k - 225, iz, ix);
  return 0.f;
}
// Material of a cell from the dt/dx-scaled coefficient arrays (0 m1_x,
// 3 aux_m3, 4 aux_m2m3): P and S Courant numbers vp*dt/dx, vs*dt/dx and the
// P impedance rho*vp / Z0.
struct Mat {
  const float* c;
  int fs, ldz;
  __device__ __forceinline__ void at(int iz, int ix, float& cp, float& cs, float& zp) const {
    const int i = iz + ldz * ix;
    const float m1 = c[i], m3 = c[3 * fs + i], m23 = c[4 * fs + i];
    cp = sqrtf(m1 * m23);
    cs = sqrtf(m1 * m3);
    zp = m1 > 0.f ? sqrtf(m23 / m1) * (1.f / Z0) : 0.f;
  }
};

// Features 230..247: one-hot edge, one-hot depth, material, zero padding.
__device__ inline float static_feature(int edge, int depth, float cp, float cs, float zp, int k) {
  if (k >= 230 && k < 234) return (k - 230) == edge ? 1.f : 0.f;
  if (k >= 234 && k < 242) return (k - 234) == depth ? 1.f : 0.f;
  if (k == 242) return cp;
  if (k == 243) return cs;
  if (k == 244) return zp;
  return 0.f;
}

// Features [ncell][KP] (normalised, bf16-rounded) and scales of every band
// cell of one grid.
__global__ void k_features(Lev m1, Lev m2, Mat mat, Band b, float* feat, float* scale) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.total) return;
  int iz, ix, edge, depth;
  cell_coords(b, c, iz, ix, edge, depth);
  float cp, cs, zp;
  mat.at(iz, ix, cp, cs, zp);
  float s = 0.f;
  for (int k = 0; k < 230; ++k) s = fmaxf(s, fabsf(raw_feature(m1, m2, iz, ix, edge, depth, k)));
  const float inv = s > 1e-30f ? 1.f / s : 0.f;
  float* out = feat + (size_t)c * KP;
  for (int k = 0; k < KP; ++k) {
    const float v = k < 230 ? raw_feature(m1, m2, iz, ix, edge, depth, k) * inv
                            : static_feature(edge, depth, cp, cs, zp, k);
    out[k] = bf16r(v);
  }
  scale[c] = s > 1e-30f ? s : 0.f;
}

// Targets: (wide(L) - plain(L)) / (s * unit) on the band cells.
__global__ void k_targets(const float* wide, int wfs, int wldz, int off, const float* plain, int pfs,
                          int pldz, Band b, const float* scale, float* tgt, int adj) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.total) return;
  int iz, ix, edge, depth;
  cell_coords(b, c, iz, ix, edge, depth);
  const float s = scale[c];
  for (int f = 0; f < NO; ++f) {
    const float w = wide[f * wfs + (iz + off) + wldz * (ix + off)];
    const float p = plain[f * pfs + iz + pldz * ix];
    const float u = z_scaled(f, adj) ? Z0 : 1.f;
    tgt[(size_t)c * NO + f] = s > 0.f ? (w - p) / (s * u) : 0.f;
  }
}

// Window copy of one level (5 fields): small(iz, ix) = wide(iz+off, ix+off),
// rim cells zero.
__global__ void k_window(const float* wide, int wfs, int wldz, int off, float* small, int sfs, int sldz,
                         int nz, int nx) {
  const int iz = blockIdx.x * blockDim.x + threadIdx.x, ix = blockIdx.y;
  if (iz >= nz || ix >= nx) return;
  const bool rim = iz < R || iz >= nz - R || ix < R || ix >= nx - R;
  for (int f = 0; f < 5; ++f)
    small[f * sfs + iz + sldz * ix] = rim ? 0.f : wide[f * wfs + (iz + off) + wldz * (ix + off)];
}

// Bf16-rounded copy of the master weights (the weights the forward uses).
__global__ void k_round_params(const float* master, float* w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NP) w[i] = bf16r(master[i]);
}

// One evaluation of the network for up to 32 samples per block of 256
// threads; weights staged in LDS. With grad != nullptr the block also runs
// the backward pass of 0.5*mean((y - t)^2) (normaliser inv_count) and
// accumulates parameter gradients with atomics. Samples are rows idx[i] of
// feat (or i when idx is null).
constexpr int SB = 32;  // samples per block
__global__ __launch_bounds__(256) void k_mlp(const float* __restrict__ w, const float* __restrict__ feat,
                                             const int* __restrict__ idx, int n, const float* __restrict__ tgt,
                                             float* __restrict__ out, float* __restrict__ grad,
                                             float* __restrict__ loss, float inv_count) {
  __shared__ float sW1[NF * H];
  __shared__ float sx[SB][KP];
  __shared__ float sz1[SB][H], sh1[SB][H], sz2[SB][H], sh2[SB][H], sy[SB][8], sd[SB][H];
  const int t = threadIdx.x;
  for (int i = t; i < NF * H; i += 256) sW1[i] = w[OW1 + i];
  const int s0 = blockIdx.x * SB;
  for (int i = t; i < SB * KP; i += 256) {
    const int s = i / KP, k = i % KP;
    const int gs = s0 + s;
    float v = 0.f;
    if (gs < n) {
      const int row = idx ? idx[gs] : gs;
      v = feat[(size_t)row * KP + k];
    }
    sx[s][k] = v;
  }
  __syncthreads();
  const int s = t / 8, q = t % 8;  // sample, group of 4 hidden units
  const bool valid = s0 + s < n;
  // layer 1
  {
    float acc[4];
    for (int u = 0; u < 4; ++u) acc[u] = w[OB1 + 4 * q + u];
    for (int k = 0; k < NF; ++k) {
      const float x = sx[s][k];
      for (int u = 0; u < 4; ++u) acc[u] += x * sW1[k * H + 4 * q + u];
    }
    for (int u = 0; u < 4; ++u) {
      sz1[s][4 * q + u] = acc[u];
      sh1[s][4 * q + u] = gelu(acc[u]);
    }
  }
  __syncthreads();
  // layer 2
  {
    float acc[4];
    for (int u = 0; u < 4; ++u) acc[u] = w[OB2 + 4 * q + u];
    for (int k = 0; k < H; ++k) {
      const float x = sh1[s][k];
      for (int u = 0; u < 4; ++u) acc[u] += x * w[OW2 + k * H + 4 * q + u];
    }
    for (int u = 0; u < 4; ++u) {
      sz2[s][4 * q + u] = acc[u];
      sh2[s][4 * q + u] = gelu(acc[u]);
    }
  }
  __syncthreads();
  // layer 3 (linear)
  if (q < NO) {
    float acc = w[OB3 + q];
    for (int k = 0; k < H; ++k) acc += sh2[s][k] * w[OW3 + k * NO + q];
    sy[s][q] = acc;
    if (valid && out) out[(size_t)(s0 + s) * NO + q] = acc;
  }
  if (!grad) return;
  __syncthreads();
  // dL/dy = (y - t) * inv_count ; loss = 0.5 * sum (y-t)^2 * inv_count
  if (q < NO) {
    float d = 0.f;
    if (valid) {
      const int row = idx ? idx[s0 + s] : s0 + s;
      d = sy[s][q] - tgt[(size_t)row * NO + q];
      atomicAdd(loss, 0.5f * d * d * inv_count);
    }
    sy[s][q] = d * inv_count;
  }
  __syncthreads();
  // W3, b3 gradients: thread t < H*NO handles one entry, summed over samples
  if (t < H * NO) {
    const int k = t / NO, o = t % NO;
    float g = 0.f;
    for (int ss = 0; ss < SB; ++ss) g += sh2[ss][k] * sy[ss][o];
    atomicAdd(&grad[OW3 + t], g);
  }
  if (t < NO) {
    float g = 0.f;
    for (int ss = 0; ss < SB; ++ss) g += sy[ss][t];
    atomicAdd(&grad[OB3 + t], g);
  }
  // dz2 = (dy W3^T) * gelu'(z2)
  for (int u = 0; u < 4; ++u) {
    const int j = 4 * q + u;
    float g = 0.f;
    for (int o = 0; o < NO; ++o) g += sy[s][o] * w[OW3 + j * NO + o];
    sd[s][j] = g * gelu_d(sz2[s][j]);
  }
  __syncthreads();
  // W2 (1024 entries, 4 per thread), b2
  for (int e = t; e < H * H; e += 256) {
    const int k = e / H, j = e % H;
    float g = 0.f;
    for (int ss = 0; ss < SB; ++ss) g += sh1[ss][k] * sd[ss][j];
    atomicAdd(&grad[OW2 + e], g);
  }
  if (t < H) {
    float g = 0.f;
    for (int ss = 0; ss < SB; ++ss) g += sd[ss][t];
    atomicAdd(&grad[OB2 + t], g);
  }
  // dz1 = (dz2 W2^T) * gelu'(z1) -> reuse sz1 for dz1 after everyone read sd
  float dz1[4];
  for (int u = 0; u < 4; ++u) {
    const int k = 4 * q + u;
    float g = 0.f;
    for (int j = 0; j < H; ++j) g += sd[s][j] * w[OW2 + k * H + j];
    dz1[u] = g * gelu_d(sz1[s][k]);
  }
  __syncthreads();
  for (int u = 0; u < 4; ++u) sz1[s][4 * q + u] = dz1[u];
  __syncthreads();
  // W1 (NF*H entries), b1
  for (int e = t; e < NF * H; e += 256) {
    const int k = e / H, j = e % H;
    float g = 0.f;
    for (int ss = 0; ss < SB; ++ss) g += sx[ss][k] * sz1[ss][j];
    atomicAdd(&grad[OW1 + e], g);
  }
  if (t < H) {
    float g = 0.f;
    for (int ss = 0; ss < SB; ++ss) g += sz1[ss][t];
    atomicAdd(&grad[OB1 + t], g);
  }
}

__global__ void k_adam(float* master, float* m, float* v, float* grad, float lr, float b1c, float b2c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NP) return;
  const float g = grad[i];
  const float mi = 0.9f * m[i] + 0.1f * g;
  const float vi = 0.999f * v[i] + 0.001f * g * g;
  m[i] = mi;
  v[i] = vi;
  master[i] -= lr * (mi / b1c) / (sqrtf(vi / b2c) + 1e-8f);
  grad[i] = 0.f;
}

// Random minibatch indices (LCG per step).
__global__ void k_sample(int* idx, int batch, int n, unsigned seed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch) return;
  unsigned x = seed * 2654435761u + (unsigned)i * 2246822519u + 0x9e3779b9u;
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  idx[i] = (int)(x % (unsigned)n);
}

// ---------------------------------------------------------------------------
// Inference on MFMA, for any number of grids at once (the shots of an engine
// group, or one grid). The cell table lists band cells (ed = edge | depth <<
// 2) and then rim cells (ed = -1) of the grids; level L of shot `shot` is
// read and written through its descriptor. After the plain update of level
// L, each band cell gets s * NN(features(L-1, L-2) / s) added and each rim
// cell is cleared. A wave evaluates 16 cells: its lanes gather the
// features (4 lanes per cell, max-abs over the cell's 230 wavefield values),
// layer 1 (245 -> 32, 90% of the arithmetic) runs as 16 x 16 x 32 BF16
// matrix products with FP32 accumulation (the features and weights are
// BF16-rounded already, so the products are exact and only the summation
// order differs from a scalar loop), layer 2 on FP32 matrix cores and
// layer 3 in FP32 from LDS. A workgroup of 12 waves (the weights copied to
// LDS once, 148 KB of LDS: one workgroup per CU, 3 waves per SIMD) loops
// over batches of 192 cells.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int MC = 16;        // cells per wave (rows of the matrix product)
constexpr int IW = 12;        // waves per inference workgroup (one per CU: 3 per SIMD)
constexpr int IT = 64 * IW;   // its threads
constexpr int KX = 256;       // features padded to 8 K-steps of 32
constexpr int KXP = KX + 8;   // LDS row pitch (bf16): rows 16 B apart in the banks

__device__ __forceinline__ __bf16 to_bf16(float x) {  // x already bf16-representable
  return __builtin_bit_cast(__bf16, (unsigned short)(__float_as_uint(x) >> 16));
}

// Global (not flat) load / store at an element offset.
__device__ __forceinline__ float gldv(const float* p, int i) {
  return ((const __attribute__((address_space(1))) float*)p)[i];
}
__device__ __forceinline__ void gstv(float* p, int i, float v) {
  ((__attribute__((address_space(1))) float*)p)[i] = v;
}

struct GridView {
  const float* l1;  // level L-1, field f at + f*fs
  const float* l2;  // level L-2
  long long fs;
  int ldz, nz, nx;
  __device__ __forceinline__ float get(const float* p, int f, int iz, int ix) const {
    if (iz < 0 || iz >= nz || ix < 0 || ix >= nx) return 0.f;
    const float v = p[f * fs + iz + ldz * ix];
    return f >= 2 ? v * (1.f / Z0) : v;
  }
};

// Inference weights as the kernel keeps them in LDS, packed once per weight
// update: W1 transposed to [hidden][feature] in BF16 (rows KXP long), then
// W2, W3 and the three biases in FP32. The kernel copies the block with
// 16-byte loads.
constexpr int PW1 = H * KXP * 2;             // bytes of the BF16 W1 block
constexpr int PF = H * H + H * NO + 2 * H + NO;  // floats after it
constexpr int PBYTES = (PW1 + PF * 4 + 15) / 16 * 16;

__global__ void k_pack_infer(const float* __restrict__ w, unsigned char* __restrict__ packed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  __bf16* w1 = reinterpret_cast<__bf16*>(packed);
  float* f = reinterpret_cast<float*>(packed + PW1);
  if (i < H * KXP) {
    const int j = i / KXP, k = i % KXP;
    w1[i] = to_bf16(k < NF ? w[OW1 + k * H + j] : 0.f);
  }
  if (i < H * H) f[i] = w[OW2 + i];
  if (i < H * NO) f[H * H + i] = w[OW3 + i];
  if (i < H) {
    f[H * H + H * NO + i] = w[OB1 + i];
    f[H * H + H * NO + H + i] = w[OB2 + i];
  }
  if (i < NO) f[H * H + H * NO + 2 * H + i] = w[OB3 + i];
}

// LDS written by lanes of a wave and read by other lanes of the same wave:
// a wave's LDS operations complete in order, so only the compiler has to be
// kept from moving them across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ADJ: an adjoint step's correction (features from levels L+1, L+2, the
// adjoint's field scaling, no history).
template <bool ADJ>
__global__ __launch_bounds__(IT) void k_infer_mfma(const unsigned char* __restrict__ packed,
                                                    const ShotDev* __restrict__ shots,
                                                    const LCell* __restrict__ cells, int nband, int nrim, int L,
                                                    int hist, int nbw) {
  // workgroups 0 .. nbw-1 loop over the batches of IW * MC band cells (the
  // weights are copied to LDS once per workgroup); the rest clear the rim
  const int nbb = (nband + IW * MC - 1) / (IW * MC);
  if ((int)blockIdx.x >= nbw) {  // rim blocks
    const int i = ((int)blockIdx.x - nbw) * IT + (int)threadIdx.x;
    if (i < nrim) {
      const LCell c = cells[nband + i];
      const ShotDev& S = shots[c.shot];
      float* o = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L);
      for (int f = 0; f < NO; ++f) o[f * S.fstride + c.iz + (long long)S.ldz * c.ix] = 0.f;
    }
    return;
  }
  __shared__ __align__(16) unsigned char sWp[PBYTES];  // the packed weights
  // features and hidden layers: each wave uses only its own [wv] rows, so
  // after the weights are in place the waves of a workgroup run their
  // batches without workgroup barriers
  __shared__ __bf16 sX[IW][MC][KXP];
  __shared__ float sH[IW][MC][H];
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
  for (int i = t; i < PBYTES / 16; i += IT)
    reinterpret_cast<uint4*>(sWp)[i] = reinterpret_cast<const uint4*>(packed)[i];
  const __bf16(*sW1)[KXP] = reinterpret_cast<const __bf16(*)[KXP]>(sWp);
  const float* sW2 = reinterpret_cast<const float*>(sWp + PW1);
  const float* sW3 = sW2 + H * H;
  const float* sB = sW3 + H * NO;
  __syncthreads();
  // features: 4 lanes per cell; each batch's cell entry is loaded during
  // the batch before it
  const int cl = lane >> 2, part = lane & 3;
  auto cell_of = [&](int bb) {
    const int c = bb * IW * MC + wv * MC + cl;
    return c < nband ? cells[c] : LCell{0, R, R, kLDead};
  };
  LCell cnext = cell_of((int)blockIdx.x);
  for (int bb = (int)blockIdx.x; bb < nbb; bb += nbw) {
  const int c = bb * IW * MC + wv * MC + cl;
  const bool in = c < nband;  // uniform over the cell's 4 lanes
  const LCell cc = cnext;
  cnext = cell_of(bb + nbw < nbb ? bb + nbw : bb);
  const bool live = in && cc.ed != kLDead;
  // a wave's 16 cells belong to one grid (the table pads every grid to
  // whole waves), so its descriptor comes from scalar loads
  const ShotDev& S = shots[__builtin_amdgcn_readfirstlane(cc.shot)];
  const int edge = cc.ed & 3, depth = (cc.ed >> 2) & 7;
  const float* l1 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, ADJ ? L + 1 : L - 1);
  const float* l2 = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, ADJ ? L + 2 : L - 2);
  // patch offsets: the cell, one step inward (normal) and one step along
  // the edge (tangent); every patch lies inside the grid (the cell table
  // admits only grids of at least 2 (R + W) + 1 cells a side), and the
  // engine's field offsets fit in 32 bits
  const int fs = (int)S.fstride, ldz = S.ldz;
  const int dz = edge == 0 ? 1 : edge == 1 ? -1 : 0, dx = edge == 2 ? 1 : edge == 3 ? -1 : 0;
  const int nstep = dz + ldz * dx, tstep = edge >= 2 ? 1 : ldz;
  const int base = live ? cc.iz + ldz * cc.ix : 2 * tstep;  // dead lanes read near element 0
  // the cell's coefficients and its outputs (level L, the Vx, Vz history),
  // loaded with the features
  float cp = 0.f, cs = 0.f, zp = 0.f;
  Mat{S.coef, (int)S.n2, S.ldz}.at(cc.iz, cc.ix, cp, cs, zp);
  float* const out = lev_ptr(S.lev, S.lev_stride, S.lev_mod, S.lev_off, L);
  float* const hh = !ADJ && hist ? static_cast<float*>(S.snap) + (long long)(L - 1) * 2 * S.n2 : nullptr;
  const int gi0 = cc.iz + ldz * cc.ix;
  float ov[2], hv[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int o = part + 4 * q;
    ov[q] = o < NO ? gldv(out, o * fs + gi0) : 0.f;
    hv[q] = (hh && o < 2) ? gldv(hh, (int)S.n2 * o + gi0) : 0.f;
  }
  constexpr int NV = (NRAW + 3) / 4;  // values per lane
  float rv[NV];
  float m = 0.f;
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    const int k = part + 4 * u;
    float v = 0.f;
    if (k < 225) {  // field f of level L-1 (L+1) at (kn inward, mm - 2 along the edge)
      const int fm = k / 9, kn = k - 9 * fm, f = fm / 5, mi = fm - 5 * f;
      v = gldv(l1, f * fs + base + kn * nstep + (mi - 2) * tstep);
      if (z_scaled(f, ADJ)) v = v * (1.f / Z0);
    } else if (k < NRAW) {  // the cell's own fields of level L-2 (L+2)
      const int f = k - 225;
      v = gldv(l2, f * fs + base);
      if (z_scaled(f, ADJ)) v = v * (1.f / Z0);
    }
    v = live ? v : 0.f;
    rv[u] = v;
    m = fmaxf(m, fabsf(v));
  }
  m = fmaxf(m, __shfl_xor(m, 1, 64));
  m = fmaxf(m, __shfl_xor(m, 2, 64));
  const float inv = m > 1e-30f ? 1.f / m : 0.f;
#pragma unroll
  for (int u = 0; u < NV; ++u) {
    const int k = part + 4 * u;
    // the hardware conversion (round to nearest even, as bf16r) in one
    // instruction
    if (k < NRAW) sX[wv][cl][k] = (__bf16)(rv[u] * inv);
  }
  for (int k = NRAW + part; k < KX; k += 4)
    sX[wv][cl][k] = (__bf16)(live ? static_feature(edge, depth, cp, cs, zp, k) : 0.f);
  wave_sync();

  // layer 1 on MFMA: Z1[16 x 32] = X[16 x 256] W1[256 x 32], two 16-column tiles
  const int r16 = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < KX / 32; ++kk) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sX[wv][r16][kk * 32 + kq]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sW1[nt * 16 + r16][kk * 32 + kq]);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // this lane holds rows (lane / 16) * 4 + i of column nt * 16 + r16
#pragma unroll
    for (int i = 0; i < 4; ++i) sH[wv][(lane >> 4) * 4 + i][nt * 16 + r16] = gelu(acc[i] + sB[nt * 16 + r16]);
  }
  wave_sync();
  // layer 2 on FP32 MFMA: Z2[16 x 32] = H1[16 x 32] W2[32 x 32] in 8 steps
  // of 4 (lane: A row r16, B column r16, k = lane / 16); the wave's own
  // rows of sH are then overwritten with the hidden layer 2 (its LDS
  // operations run in program order, so every read above precedes them)
  {
    const int kr = lane >> 4;
    f32x4 z[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < H / 4; ++kk)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(sH[wv][r16][kk * 4 + kr], sW2[(kk * 4 + kr) * H + nt * 16 + r16],
                                                   acc, 0, 0, 0);
      z[nt] = acc;
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) sH[wv][kr * 4 + i][nt * 16 + r16] = gelu(z[nt][i] + sB[H + nt * 16 + r16]);
  }
  // layer 3: partial sums over hidden units part*8 .. part*8+7 of cell cl,
  // reduced over the cell's 4 lanes
  float y[NO] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int j = part * 8 + u;
    const float h2 = sH[wv][cl][j];
#pragma unroll
    for (int o = 0; o < NO; ++o) y[o] = fmaf(h2, sW3[j * NO + o], y[o]);
  }
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    y[o] += __shfl_xor(y[o], 1, 64);
    y[o] += __shfl_xor(y[o], 2, 64);
  }
  if (live && m > 1e-30f) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int o = part + 4 * q;
      if (o < NO) {
        const float d = (y[o] + sB[2 * H + o]) * m * (z_scaled(o, ADJ) ? Z0 : 1.f);
        gstv(out, o * fs + gi0, ov[q] + d);
        if (hh && o < 2) gstv(hh, (int)S.n2 * o + gi0, hv[q] + d);
      }
    }
  }
  wave_sync();  // sX and sH are rewritten by the next batch
  }
}

// Multiplicative noise on the band cells of one level (training
// augmentation): v *= 1 + amp * u, u uniform in [-1, 1) from a hash of
// (seed, cell, field).
__global__ void k_noise(float* lev, int fs, int ldz, Band b, float amp, unsigned seed) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.total) return;
  int iz, ix, edge, depth;
  cell_coords(b, c, iz, ix, edge, depth);
  for (int f = 0; f < NO; ++f) {
    unsigned x = seed * 0x9e3779b9u ^ ((unsigned)c * 5u + (unsigned)f) * 0x85ebca6bu;
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    const float u = (float)(x >> 8) * (2.f / 16777216.f) - 1.f;
    float* p = lev + f * fs + iz + ldz * ix;
    *p = *p * (1.f + amp * u);
  }
}

}  // namespace labc

// Band cells, then rim cells, of the given grids (shot = index in grids).
void labc_cell_table(const std::vector<std::pair<int, int>>& grids, std::vector<LCell>& cells, int* nband,
                     int* nrim) {
  using namespace labc;
  cells.clear();
  // The inference kernel gives 16 consecutive cells to a wave, which loads
  // their patches 4 features at a time (one per lane of a cell; consecutive
  // features are consecutive inward steps): top and bottom band cells are
  // listed column by column, the 8 depths of a column in a row (two columns
  // per wave; the 8 cells' 4 inward steps span 11 consecutive rows), left
  // and right band cells row by row at each depth (16 consecutive rows per
  // wave).
  for (int s = 0; s < (int)grids.size(); ++s) {
    const int nz = grids[s].first, nx = grids[s].second;
    // every feature patch of a band cell lies inside the grid from this size
    // on (the inference kernel loads the patches without bounds checks)
    if (nz < 2 * (R + W) + 1 || nx < 2 * (R + W) + 1) {
      cells.clear();
      *nband = *nrim = 0;
      return;
    }
    for (int edge = 0; edge < 2; ++edge)
      for (int ix = R; ix < nx - R; ++ix)
        for (int d = 0; d < W; ++d) cells.push_back(LCell{s, edge == 0 ? R + d : nz - R - 1 - d, ix, edge | (d << 2)});
    for (int edge = 2; edge < 4; ++edge)
      for (int d = 0; d < W; ++d)
        for (int iz = R + W; iz < nz - R - W; ++iz)
          cells.push_back(LCell{s, iz, edge == 2 ? R + d : nx - R - 1 - d, edge | (d << 2)});
    // whole waves per grid, so that a wave's grid is the same in every lane:
    // padding entries (ed = kLDead) are evaluated at a valid cell and discarded
    while (cells.size() % 16) cells.push_back(LCell{s, R, R, kLDead});
  }
  *nband = (int)cells.size();
  for (int s = 0; s < (int)grids.size(); ++s) {
    const int nz = grids[s].first, nx = grids[s].second;
    for (int ix = 0; ix < nx; ++ix)
      for (int iz = 0; iz < nz; ++iz)
        if (iz < R || iz >= nz - R || ix < R || ix >= nx - R) cells.push_back(LCell{s, iz, ix, -1});
  }
  *nrim = (int)cells.size() - *nband;
}

int launch_labc_infer(const unsigned char* w, const ShotDev* shots, const LCell* cells, int nband, int nrim, int L,
                      bool hist, hipStream_t st, bool adj) {
  using namespace labc;
  // band workgroups: one round of resident workgroups, each looping over
  // batches of IW * MC cells
  // (queried once per process, thread-safe; a failed query falls back to
  // one workgroup on each of 256 CUs, which only changes how the batches
  // are spread)
  static const int resident = [] {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_infer_mfma<false>, IT, 0) != hipSuccess ||
        cus * per_cu <= 0)
      return 256;
    return cus * per_cu;
  }();
  const int nbb = (nband + IW * MC - 1) / (IW * MC);
  const int nbw = std::min(nbb, resident);
  const int blocks = nbw + (nrim + IT - 1) / IT;
  if (blocks == 0) return OK;
  if (adj)
    hipLaunchKernelGGL(k_infer_mfma<true>, dim3(blocks), dim3(IT), 0, st, w, shots, cells, nband, nrim, L, 0, nbw);
  else
    hipLaunchKernelGGL(k_infer_mfma<false>, dim3(blocks), dim3(IT), 0, st, w, shots, cells, nband, nrim, L,
                       (int)hist, nbw);
  HIPCHK(hipGetLastError());
  return OK;
}

}  // namespace erm

using namespace erm;
using namespace erm::labc;

static double g_last_ms = 0.0;

RTM_EXPORT double labc_last_forward_ms(void) { return g_last_ms; }

struct labc_net {
  DevBuf master, w, m, v, grad, loss, packed;  // packed: the inference layout of w
  int t = 0;
};

struct labc_data {
  int capacity = 0, size = 0;
  DevBuf feat, tgt;
};

static int launch_mlp(labc_net* net, const float* feat, const int* idx, int n, const float* tgt, float* out,
                      bool train, float inv_count, hipStream_t st) {
  const int blocks = (n + SB - 1) / SB;
  hipLaunchKernelGGL(k_mlp, dim3(blocks), dim3(256), 0, st, net->w.as<float>(), feat, idx, n, tgt, out,
                     train ? net->grad.as<float>() : nullptr, net->loss.as<float>(), inv_count);
  HIPCHK(hipGetLastError());
  return OK;
}

static int refresh_weights(labc_net* n, hipStream_t st) {
  hipLaunchKernelGGL(k_round_params, dim3((NP + 255) / 256), dim3(256), 0, st, n->master.as<float>(),
                     n->w.as<float>());
  hipLaunchKernelGGL(k_pack_infer, dim3((H * KXP + 255) / 256), dim3(256), 0, st, n->w.as<float>(),
                     n->packed.as<unsigned char>());
  HIPCHK(hipGetLastError());
  return OK;
}

RTM_EXPORT int labc_net_create(labc_net** out, unsigned seed) {
  if (!out) { set_error("null argument"); return E_ARG; }
  STCHK(ensure_device());
  std::unique_ptr<labc_net> n(new labc_net());
  std::mt19937 rng(seed);
  std::vector<float> p(NP, 0.f);
  auto fill = [&](int off, int fan_in, int fan_out, int count) {
    const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < count; ++i) p[off + i] = u(rng);
  };
  fill(OW1, NF, H, NF * H);
  fill(OW2, H, H, H * H);
  fill(OW3, H, NO, H * NO);
  for (int i = 0; i < H * NO; ++i) p[OW3 + i] *= 0.01f;  // start near "no correction"
  STCHK(n->master.alloc(NP * 4, "net"));
  STCHK(n->w.alloc(NP * 4, "net"));
  STCHK(n->m.alloc(NP * 4, "net"));
  STCHK(n->v.alloc(NP * 4, "net"));
  STCHK(n->grad.alloc(NP * 4, "net"));
  STCHK(n->loss.alloc(4, "net"));
  STCHK(n->packed.alloc(PBYTES, "net"));
  HIPCHK(hipMemcpy(n->master.p, p.data(), NP * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(n->m.p, 0, NP * 4));
  HIPCHK(hipMemset(n->v.p, 0, NP * 4));
  HIPCHK(hipMemset(n->grad.p, 0, NP * 4));
  STCHK(refresh_weights(n.get(), nullptr));
  HIPCHK(hipDeviceSynchronize());
  *out = n.release();
  return OK;
}

RTM_EXPORT int labc_net_destroy(labc_net* n) {
  delete n;
  return OK;
}

RTM_EXPORT int labc_net_get_params(labc_net* n, float* p) {
  if (!n || !p) { set_error("null argument"); return E_ARG; }
  HIPCHK(hipMemcpy(p, n->master.p, NP * 4, hipMemcpyDeviceToHost));
  return OK;
}

RTM_EXPORT int labc_net_set_params(labc_net* n, const float* p) {
  if (!n || !p) { set_error("null argument"); return E_ARG; }
  HIPCHK(hipMemcpy(n->master.p, p, NP * 4, hipMemcpyHostToDevice));
  STCHK(refresh_weights(n, nullptr));
  HIPCHK(hipDeviceSynchronize());
  return OK;
}

RTM_EXPORT int labc_cell_count(int nz, int nx) {
  if (nz < 2 * (R + W) + 1 || nx < 2 * (R + W) + 1) return 0;
  return band_of(nz, nx).total;
}

RTM_EXPORT int labc_cell_coords(int nz, int nx, int c, int* iz, int* ix, int* edge, int* depth) {
  const Band b = band_of(nz, nx);
  if (c < 0 || c >= b.total || !iz || !ix || !edge || !depth) { set_error("bad cell"); return E_ARG; }
  cell_coords(b, c, *iz, *ix, *edge, *depth);
  return OK;
}

RTM_EXPORT int labc_net_forward(labc_net* n, const float* feat, int cnt, float* out, int mode) {
  if (!n || !feat || !out || cnt < 1) { set_error("bad argument"); return E_ARG; }
  if (mode != 0) { set_error("mode must be 0 (the training forward)"); return E_ARG; }
  DevBuf f, o;
  STCHK(f.alloc((size_t)cnt * KP * 4, "features"));
  STCHK(o.alloc((size_t)cnt * NO * 4, "outputs"));
  HIPCHK(hipMemcpy(f.p, feat, (size_t)cnt * KP * 4, hipMemcpyHostToDevice));
  STCHK(launch_mlp(n, f.as<float>(), nullptr, cnt, nullptr, o.as<float>(), false, 0.f, nullptr));
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(out, o.p, (size_t)cnt * NO * 4, hipMemcpyDeviceToHost));
  return OK;
}

RTM_EXPORT int labc_net_grad(labc_net* n, const float* feat, const float* target, int cnt, float* grad,
                             float* loss) {
  if (!n || !feat || !target || !grad || !loss || cnt < 1) { set_error("bad argument"); return E_ARG; }
  DevBuf f, t;
  STCHK(f.alloc((size_t)cnt * KP * 4, "features"));
  STCHK(t.alloc((size_t)cnt * NO * 4, "targets"));
  HIPCHK(hipMemcpy(f.p, feat, (size_t)cnt * KP * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(t.p, target, (size_t)cnt * NO * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(n->grad.p, 0, NP * 4));
  HIPCHK(hipMemset(n->loss.p, 0, 4));
  STCHK(launch_mlp(n, f.as<float>(), nullptr, cnt, t.as<float>(), nullptr, true, 1.f / (cnt * NO), nullptr));
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(grad, n->grad.p, NP * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(loss, n->loss.p, 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemset(n->grad.p, 0, NP * 4));
  return OK;
}

static int features_impl(int nz, int nx, const float* lev_m1, const float* lev_m2, const float* coef5,
                         float* feat, float* scale, int adj) {
  const int nc = labc_cell_count(nz, nx);
  if (nc == 0 || !lev_m1 || !lev_m2 || !coef5 || !feat || !scale) { set_error("bad argument"); return E_ARG; }
  STCHK(ensure_device());
  const size_t n2 = (size_t)nz * nx;
  DevBuf a, b, cf, f, s;
  STCHK(a.alloc(5 * n2 * 4, "level"));
  STCHK(b.alloc(5 * n2 * 4, "level"));
  STCHK(cf.alloc(5 * n2 * 4, "coefficients"));
  HIPCHK(hipMemcpy(cf.p, coef5, 5 * n2 * 4, hipMemcpyHostToDevice));
  STCHK(f.alloc((size_t)nc * KP * 4, "features"));
  STCHK(s.alloc((size_t)nc * 4, "scales"));
  HIPCHK(hipMemcpy(a.p, lev_m1, 5 * n2 * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b.p, lev_m2, 5 * n2 * 4, hipMemcpyHostToDevice));
  Lev l1{a.as<float>(), (int)n2, nz, nz, nx, adj}, l2{b.as<float>(), (int)n2, nz, nz, nx, adj};
  const Band bd = band_of(nz, nx);
  const Mat mat{cf.as<float>(), (int)n2, nz};
  hipLaunchKernelGGL(k_features, dim3((nc + 255) / 256), dim3(256), 0, nullptr, l1, l2, mat, bd, f.as<float>(),
                     s.as<float>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(feat, f.p, (size_t)nc * KP * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(scale, s.p, (size_t)nc * 4, hipMemcpyDeviceToHost));
  return OK;
}

RTM_EXPORT int labc_features(int nz, int nx, const float* lev_m1, const float* lev_m2, const float* coef5,
                             float* feat, float* scale) {
  return features_impl(nz, nx, lev_m1, lev_m2, coef5, feat, scale, 0);
}

RTM_EXPORT int labc_features_adj(int nz, int nx, const float* lev_p1, const float* lev_p2, const float* coef5,
                                 float* feat, float* scale) {
  return features_impl(nz, nx, lev_p1, lev_p2, coef5, feat, scale, 1);
}

RTM_EXPORT int labc_data_create(labc_data** out, int capacity) {
  if (!out || capacity < 1) { set_error("bad argument"); return E_ARG; }
  STCHK(ensure_device());
  std::unique_ptr<labc_data> d(new labc_data());
  d->capacity = capacity;
  STCHK(d->feat.alloc((size_t)capacity * KP * 4, "training features"));
  STCHK(d->tgt.alloc((size_t)capacity * NO * 4, "training targets"));
  *out = d.release();
  return OK;
}

RTM_EXPORT int labc_data_destroy(labc_data* d) {
  delete d;
  return OK;
}

RTM_EXPORT int labc_data_size(labc_data* d) { return d ? d->size : 0; }

RTM_EXPORT int labc_data_get(labc_data* d, int row0, int n, float* feat, float* tgt) {
  if (!d || row0 < 0 || n < 0 || row0 + n > d->size) { set_error("rows out of range"); return E_ARG; }
  if (n == 0) return OK;
  if (feat) HIPCHK(hipMemcpy(feat, d->feat.as<float>() + (size_t)row0 * KP, (size_t)n * KP * 4, hipMemcpyDeviceToHost));
  if (tgt) HIPCHK(hipMemcpy(tgt, d->tgt.as<float>() + (size_t)row0 * NO, (size_t)n * NO * 4, hipMemcpyDeviceToHost));
  return OK;
}

// One grid (shot) with its own ring/coefficients, driven step by step.
struct LGrid {
  ShotDev h{};
  DevBuf desc, tiles, ring, coef, kfac, amp;
  int ntiles = 0, ldz = 0;
  long long n2 = 0;
  // band_abc / band_off: multiply the sponge by the band grid's sponge of
  // width band_abc on the sub-grid [band_off, band_off + band_n) (clamped
  // outside it), so that a wide grid damps the band as the band grid does.
  int setup(int nz, int nx, int nt, float dt, int zs, int xs, const float* coef5, const float* wavelet, int abc,
            int zr, int rx0, int nrx, float* rec_dev, int band_abc = 0, int band_off = 0, int band_nz = 0,
            int band_nx = 0) {
    ldz = device_ldz(nz);
    n2 = (long long)ldz * nx;
    // four levels, so that level L is written while L-1 and L-2 stay
    // readable for the learned layer's inference after the step
    STCHK(ring.alloc(20 * n2 * 4, "ring"));
    STCHK(coef.alloc(5 * n2 * 4, "coefficients"));
    STCHK(fill_masked_coefs(coef5, nz, nx, ldz, coef.as<float>()));
    STCHK(launch_fill(ring.as<float>(), 20 * n2, 0.f, nullptr));
    h.nz = nz; h.nx = nx; h.ldz = ldz; h.nt = nt; h.n2 = n2;
    h.lev = ring.as<float>(); h.lev_stride = 5 * n2; h.fstride = n2; h.lev_mod = 4; h.lev_off = 4;
    h.coef = coef.as<float>();
    if (abc > 0 || band_abc > 0) {
      std::vector<float> hz(nz, 1.f), hx(nx, 1.f);
      if (abc > 0) {
        sponge_factors(nz, abc, hz);
        sponge_factors(nx, abc, hx);
      }
      if (band_abc > 0) {
        std::vector<float> bz, bx;
        sponge_factors(band_nz, band_abc, bz);
        sponge_factors(band_nx, band_abc, bx);
        for (int i = 0; i < nz; ++i) hz[i] *= bz[std::min(std::max(i - band_off, 0), band_nz - 1)];
        for (int i = 0; i < nx; ++i) hx[i] *= bx[std::min(std::max(i - band_off, 0), band_nx - 1)];
      }
      STCHK(kfac.alloc((size_t)(nz + nx) * 4, "sponge"));
      HIPCHK(hipMemcpy(kfac.p, hz.data(), nz * 4, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(kfac.as<float>() + nz, hx.data(), nx * 4, hipMemcpyHostToDevice));
      h.kz = kfac.as<float>();
      h.kx = h.kz + nz;
    }
    std::vector<float> a(nt, 0.f);
    for (int L = 2; L <= nt; ++L) a[L - 1] = dt * wavelet[L - 2];
    STCHK(amp.alloc(nt * 4, "amplitudes"));
    HIPCHK(hipMemcpy(amp.p, a.data(), nt * 4, hipMemcpyHostToDevice));
    h.zs = zs; h.xs = xs; h.src_amp = amp.as<float>();
    h.zr = zr; h.rx0 = rx0; h.nrx = nrx; h.rec = rec_dev;
    h.snap_fmt = SNAP_HIST; h.snap_every = 1;
    STCHK(desc.alloc(sizeof(ShotDev), "descriptor"));
    HIPCHK(hipMemcpy(desc.p, &h, sizeof(ShotDev), hipMemcpyHostToDevice));
    const int tx = tile_x(4);
    const int nbz = (nz + TILE_Z - 1) / TILE_Z, nbx = (nx + tx - 1) / tx;
    std::vector<TileRef> t;
    for (int bx = 0; bx < nbx; ++bx)
      for (int bz = 0; bz < nbz; ++bz) t.push_back({0, bz * TILE_Z, bx * tx, bz + nbz * bx});
    ntiles = (int)t.size();
    STCHK(tiles.alloc(t.size() * sizeof(TileRef), "tiles"));
    HIPCHK(hipMemcpy(tiles.p, t.data(), t.size() * sizeof(TileRef), hipMemcpyHostToDevice));
    return OK;
  }
  float* level(int L) const { return ring.as<float>() + (size_t)((L + 4) % 4) * 5 * n2; }
  int step(int L, int flags, hipStream_t st) {
    return launch_fwd(desc.as<ShotDev>(), tiles.as<TileRef>(), ntiles, L, flags, 4, SNAP_HIST, 1, st);
  }
  // adjoint step: level L from L+1, L+2 (no imaging)
  int astep(int L, int flags, hipStream_t st) {
    return launch_adj(desc.as<ShotDev>(), tiles.as<TileRef>(), ntiles, L, flags, 4, SNAP_HIST, 1, st);
  }
  Lev lev_view(int L, int adj = 0) const { return Lev{level(L), (int)n2, ldz, h.nz, h.nx, adj}; }
  Mat mat() const { return Mat{coef.as<float>(), (int)n2, ldz}; }
};

// The levels of a collection: each in [lo, hi], none twice (the collectors
// walk the levels in order and take each listed one once).
static int check_levels(const int* levels, int nlev, int lo, int hi) {
  std::vector<int> v(levels, levels + nlev);
  std::sort(v.begin(), v.end());
  for (size_t i = 0; i < v.size(); ++i) {
    if (v[i] < lo || v[i] > hi) {
      set_error("level %d not in [%d, %d]", v[i], lo, hi);
      return E_ARG;
    }
    if (i > 0 && v[i] == v[i - 1]) {
      set_error("level %d listed twice", v[i]);
      return E_ARG;
    }
  }
  return OK;
}

RTM_EXPORT int labc_collect(labc_data* d, int nz_phys, int nx_phys, int pad_wide, int abc_wide, int nt,
                            float dt, int zs_phys, int xs_phys, const float* coef5_wide,
                            const float* coef5_small, const float* wavelet, const int* levels, int nlev,
                            float noise, unsigned seed, int abc_band) {
  if (!d || !coef5_wide || !coef5_small || !wavelet || !levels || nlev < 1) { set_error("bad argument"); return E_ARG; }
  STCHK(check_levels(levels, nlev, 2, nt));
  const int pad_s = R + W;
  if (pad_wide < pad_s + 5) { set_error("wide padding must exceed the band padding"); return E_ARG; }
  const int nzw = nz_phys + 2 * pad_wide, nxw = nx_phys + 2 * pad_wide;
  const int nzs = nz_phys + 2 * pad_s, nxs = nx_phys + 2 * pad_s;
  const int off = pad_wide - pad_s;
  const Band bd = band_of(nzs, nxs);
  if ((long long)d->size + (long long)nlev * bd.total > d->capacity) {
    set_error("training set capacity %d exceeded", d->capacity);
    return E_ARG;
  }
  STCHK(ensure_device());
  LGrid gw, gs;
  if (abc_band < 0 || abc_band > W) { set_error("band sponge width %d not in [0, %d]", abc_band, W); return E_ARG; }
  STCHK(gw.setup(nzw, nxw, nt, dt, pad_wide + zs_phys, pad_wide + xs_phys, coef5_wide, wavelet, abc_wide, 0, 0, 0,
                 nullptr, abc_band, off, nzs, nxs));
  STCHK(gs.setup(nzs, nxs, nt, dt, pad_s + zs_phys, pad_s + xs_phys, coef5_small, wavelet, abc_band, 0, 0, 0,
                 nullptr));
  DevBuf scales;
  STCHK(scales.alloc((size_t)bd.total * 4, "scales"));
  std::vector<int> lv(levels, levels + nlev);
  std::sort(lv.begin(), lv.end());
  size_t next = 0;
  const dim3 wgrid((nzs + 255) / 256, nxs);
  for (int L = 2; L <= nt && next < lv.size(); ++L) {
    const bool take = lv[next] == L;
    if (take && L >= 3) {
      // band grid levels L-1, L-2 = windowed wide levels (rim zero)
      hipLaunchKernelGGL(k_window, wgrid, dim3(256), 0, nullptr, gw.level(L - 1), (int)gw.n2, gw.ldz, off,
                         gs.level(L - 1), (int)gs.n2, gs.ldz, nzs, nxs);
      hipLaunchKernelGGL(k_window, wgrid, dim3(256), 0, nullptr, gw.level(L - 2), (int)gw.n2, gw.ldz, off,
                         gs.level(L - 2), (int)gs.n2, gs.ldz, nzs, nxs);
      if (noise > 0.f) {
        hipLaunchKernelGGL(k_noise, dim3((bd.total + 255) / 256), dim3(256), 0, nullptr, gs.level(L - 1), (int)gs.n2,
                           gs.ldz, bd, noise, seed * 7919u + 2u * (unsigned)L);
        hipLaunchKernelGGL(k_noise, dim3((bd.total + 255) / 256), dim3(256), 0, nullptr, gs.level(L - 2), (int)gs.n2,
                           gs.ldz, bd, noise, seed * 7919u + 2u * (unsigned)L + 1u);
      }
      HIPCHK(hipGetLastError());
      hipLaunchKernelGGL(k_features, dim3((bd.total + 255) / 256), dim3(256), 0, nullptr, gs.lev_view(L - 1),
                         gs.lev_view(L - 2), gs.mat(), bd, d->feat.as<float>() + (size_t)d->size * KP, scales.as<float>());
      HIPCHK(hipGetLastError());
    }
    STCHK(gw.step(L, (L >= 3 && (abc_wide > 0 || abc_band > 0) ? F_K : 0) | F_SRC, nullptr));
    if (take) {
      if (L >= 3) {
        // plain update of the band grid (with its sponge when abc_band > 0)
        STCHK(gs.step(L, F_SRC | (abc_band > 0 ? F_K : 0), nullptr));
        hipLaunchKernelGGL(k_targets, dim3((bd.total + 255) / 256), dim3(256), 0, nullptr, gw.level(L),
                           (int)gw.n2, gw.ldz, off, gs.level(L), (int)gs.n2, gs.ldz, bd,
                           scales.as<float>(), d->tgt.as<float>() + (size_t)d->size * NO, 0);
        HIPCHK(hipGetLastError());
        d->size += bd.total;
      }
      ++next;
    }
  }
  HIPCHK(hipDeviceSynchronize());
  return OK;
}

// Teacher-forced pairs of the adjoint (receiver-side) propagation, the
// counterpart of labc_collect: the wide grid models receiver data at depth
// zr_phys (every physical column; with coef5_wide_obs, the data of that
// model minus those of coef5_wide, as a migration injects them), then
// back-propagates them with the exact adjoint of its sponge (continued by
// the band sponge, as in labc_collect);
// at each listed level L the band grid's levels L+1, L+2 are windows of the
// wide run, the features are those of an adjoint state, and the target is
// the wide run's level L minus the band grid's adjoint update (its band
// sponge, the same injected data).
RTM_EXPORT int labc_collect_adj(labc_data* d, int nz_phys, int nx_phys, int pad_wide, int abc_wide, int nt,
                                float dt, int zs_phys, int xs_phys, int zr_phys, const float* coef5_wide,
                                const float* coef5_small, const float* wavelet, const int* levels, int nlev,
                                float noise, unsigned seed, int abc_band, const float* coef5_wide_obs) {
  if (!d || !coef5_wide || !coef5_small || !wavelet || !levels || nlev < 1 || zr_phys < 0 || zr_phys >= nz_phys) {
    set_error("bad argument");
    return E_ARG;
  }
  STCHK(check_levels(levels, nlev, 1, nt));
  const int pad_s = R + W;
  if (pad_wide < pad_s + 5) { set_error("wide padding must exceed the band padding"); return E_ARG; }
  if (abc_band < 0 || abc_band > W) { set_error("band sponge width %d not in [0, %d]", abc_band, W); return E_ARG; }
  const int nzw = nz_phys + 2 * pad_wide, nxw = nx_phys + 2 * pad_wide;
  const int nzs = nz_phys + 2 * pad_s, nxs = nx_phys + 2 * pad_s;
  const int off = pad_wide - pad_s;
  const Band bd = band_of(nzs, nxs);
  if ((long long)d->size + (long long)nlev * bd.total > d->capacity) {
    set_error("training set capacity %d exceeded", d->capacity);
    return E_ARG;
  }
  STCHK(ensure_device());
  DevBuf rec;
  STCHK(rec.alloc((size_t)5 * nt * nx_phys * 4, "receivers"));
  STCHK(launch_fill(rec.as<float>(), (size_t)5 * nt * nx_phys, 0.f, nullptr));
  LGrid gw, gs;
  STCHK(gw.setup(nzw, nxw, nt, dt, pad_wide + zs_phys, pad_wide + xs_phys, coef5_wide, wavelet, abc_wide,
                 pad_wide + zr_phys, pad_wide, nx_phys, rec.as<float>(), abc_band, off, nzs, nxs));
  STCHK(gs.setup(nzs, nxs, nt, dt, pad_s + zs_phys, pad_s + xs_phys, coef5_small, wavelet, abc_band,
                 pad_s + zr_phys, pad_s, nx_phys, rec.as<float>()));
  const bool kw = abc_wide > 0 || abc_band > 0;
  // the data: one forward run of the wide grid (of the observed model, then
  // subtracting the run of coef5_wide)
  if (coef5_wide_obs) {
    LGrid go;
    STCHK(go.setup(nzw, nxw, nt, dt, pad_wide + zs_phys, pad_wide + xs_phys, coef5_wide_obs, wavelet, abc_wide,
                   pad_wide + zr_phys, pad_wide, nx_phys, rec.as<float>(), abc_band, off, nzs, nxs));
    for (int L = 2; L <= nt; ++L) STCHK(go.step(L, (L >= 3 && kw ? F_K : 0) | F_SRC | F_REC, nullptr));
  }
  const int recf = coef5_wide_obs ? F_RECSUB : F_REC;
  for (int L = 2; L <= nt; ++L) STCHK(gw.step(L, (L >= 3 && kw ? F_K : 0) | F_SRC | recf, nullptr));
  STCHK(launch_fill(gw.ring.as<float>(), 20 * gw.n2, 0.f, nullptr));
  DevBuf scales;
  STCHK(scales.alloc((size_t)bd.total * 4, "scales"));
  std::vector<int> lv(levels, levels + nlev);
  std::sort(lv.begin(), lv.end(), std::greater<int>());
  size_t next = 0;
  const dim3 wgrid((nzs + 255) / 256, nxs);
  for (int L = nt; L >= 1 && next < lv.size(); --L) {
    const bool take = lv[next] == L;
    if (take) {
      // band grid levels L+1, L+2 = windowed wide levels (rim zero)
      hipLaunchKernelGGL(k_window, wgrid, dim3(256), 0, nullptr, gw.level(L + 1), (int)gw.n2, gw.ldz, off,
                         gs.level(L + 1), (int)gs.n2, gs.ldz, nzs, nxs);
      hipLaunchKernelGGL(k_window, wgrid, dim3(256), 0, nullptr, gw.level(L + 2), (int)gw.n2, gw.ldz, off,
                         gs.level(L + 2), (int)gs.n2, gs.ldz, nzs, nxs);
      if (noise > 0.f) {
        hipLaunchKernelGGL(k_noise, dim3((bd.total + 255) / 256), dim3(256), 0, nullptr, gs.level(L + 1), (int)gs.n2,
                           gs.ldz, bd, noise, seed * 7919u + 2u * (unsigned)L);
        hipLaunchKernelGGL(k_noise, dim3((bd.total + 255) / 256), dim3(256), 0, nullptr, gs.level(L + 2), (int)gs.n2,
                           gs.ldz, bd, noise, seed * 7919u + 2u * (unsigned)L + 1u);
      }
      HIPCHK(hipGetLastError());
      hipLaunchKernelGGL(k_features, dim3((bd.total + 255) / 256), dim3(256), 0, nullptr, gs.lev_view(L + 1, 1),
                         gs.lev_view(L + 2, 1), gs.mat(), bd, d->feat.as<float>() + (size_t)d->size * KP,
                         scales.as<float>());
      HIPCHK(hipGetLastError());
    }
    STCHK(gw.astep(L, (kw ? A_K : 0) | A_INJ, nullptr));
    if (take) {
      STCHK(gs.astep(L, (abc_band > 0 ? A_K : 0) | A_INJ, nullptr));
      hipLaunchKernelGGL(k_targets, dim3((bd.total + 255) / 256), dim3(256), 0, nullptr, gw.level(L),
                         (int)gw.n2, gw.ldz, off, gs.level(L), (int)gs.n2, gs.ldz, bd,
                         scales.as<float>(), d->tgt.as<float>() + (size_t)d->size * NO, 1);
      HIPCHK(hipGetLastError());
      d->size += bd.total;
      ++next;
    }
  }
  HIPCHK(hipDeviceSynchronize());
  return OK;
}

RTM_EXPORT int labc_train(labc_net* n, labc_data* d, int steps, int batch, float lr, unsigned seed,
                          float* losses) {
  if (!n || !d || d->size < 1 || steps < 1 || batch < 1) { set_error("bad argument"); return E_ARG; }
  DevBuf idx, lossv;
  STCHK(idx.alloc((size_t)batch * 4, "batch"));
  STCHK(lossv.alloc((size_t)steps * 4, "losses"));
  HIPCHK(hipMemset(lossv.p, 0, (size_t)steps * 4));
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(k_sample, dim3((batch + 255) / 256), dim3(256), 0, nullptr, idx.as<int>(), batch, d->size,
                       seed + 7919u * (unsigned)s);
    const int blocks = (batch + SB - 1) / SB;
    hipLaunchKernelGGL(k_mlp, dim3(blocks), dim3(256), 0, nullptr, n->w.as<float>(), d->feat.as<float>(),
                       idx.as<int>(), batch, d->tgt.as<float>(), nullptr, n->grad.as<float>(),
                       lossv.as<float>() + s, 1.f / (batch * NO));
    n->t += 1;
    const float b1c = 1.f - std::pow(0.9f, (float)n->t), b2c = 1.f - std::pow(0.999f, (float)n->t);
    hipLaunchKernelGGL(k_adam, dim3((NP + 255) / 256), dim3(256), 0, nullptr, n->master.as<float>(),
                       n->m.as<float>(), n->v.as<float>(), n->grad.as<float>(), lr, b1c, b2c);
    STCHK(refresh_weights(n, nullptr));
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipDeviceSynchronize());
  if (losses) HIPCHK(hipMemcpy(losses, lossv.p, (size_t)steps * 4, hipMemcpyDeviceToHost));
  return OK;
}

RTM_EXPORT int labc_eval(labc_net* n, labc_data* d, float* loss) {
  if (!n || !d || !loss || d->size < 1) { set_error("bad argument"); return E_ARG; }
  HIPCHK(hipMemset(n->loss.p, 0, 4));
  DevBuf g;
  STCHK(g.alloc(NP * 4, "scratch gradient"));
  HIPCHK(hipMemset(g.p, 0, NP * 4));
  const int blocks = (d->size + SB - 1) / SB;
  hipLaunchKernelGGL(k_mlp, dim3(blocks), dim3(256), 0, nullptr, n->w.as<float>(), d->feat.as<float>(), nullptr,
                     d->size, d->tgt.as<float>(), nullptr, g.as<float>(), n->loss.as<float>(),
                     1.f / ((float)d->size * NO));
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(loss, n->loss.p, 4, hipMemcpyDeviceToHost));
  return OK;
}

RTM_EXPORT int rtm_batch_set_learned(rtm_batch* b, labc_net* n) {
  Engine* e = batch_engine(b);
  if (!e) { set_error("null batch"); return E_ARG; }
  return e->set_learned(n ? n->packed.as<unsigned char>() : nullptr);
}

RTM_EXPORT int rtm_batch_set_learned_adjoint(rtm_batch* b, labc_net* n) {
  Engine* e = batch_engine(b);
  if (!e) { set_error("null batch"); return E_ARG; }
  return e->set_learned_adjoint(n ? n->packed.as<unsigned char>() : nullptr);
}

RTM_EXPORT int labc_forward(labc_net* n, const rtm_shot_t* s, const float* coef5, const float* wavelet,
                            float* rec, float* last) {
  if (!s || !coef5 || !wavelet || (s->nrx > 0 && !rec)) { set_error("bad argument"); return E_ARG; }
  if (n && labc_cell_count(s->nz, s->nx) == 0) { set_error("grid too small for the band"); return E_ARG; }
  STCHK(ensure_device());
  DevBuf recd;
  const int nr = std::max(s->nrx, 1);
  STCHK(recd.alloc((size_t)5 * s->nt * nr * 4, "receivers"));
  STCHK(launch_fill(recd.as<float>(), (size_t)5 * s->nt * nr, 0.f, nullptr));
  LGrid g;
  STCHK(g.setup(s->nz, s->nx, s->nt, s->dt, s->zs, s->xs, coef5, wavelet, s->abc, s->zr, s->rx0, s->nrx,
                recd.as<float>()));
  DevBuf dcells;
  int nband = 0, nrim = 0;
  if (n) {
    std::vector<LCell> cells;
    labc_cell_table({{s->nz, s->nx}}, cells, &nband, &nrim);
    STCHK(dcells.alloc(cells.size() * sizeof(LCell), "band cells"));
    HIPCHK(hipMemcpy(dcells.p, cells.data(), cells.size() * sizeof(LCell), hipMemcpyHostToDevice));
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, nullptr));
  for (int L = 2; L <= s->nt; ++L) {
    int fl = F_SRC | (s->nrx > 0 ? F_REC : 0);
    if (L >= 3 && s->abc > 0) fl |= F_K;  // with a network: the sponge and the layer together
    STCHK(g.step(L, fl, nullptr));
    if (n && L >= 3)
      STCHK(launch_labc_infer(n->packed.as<unsigned char>(), g.desc.as<ShotDev>(), dcells.as<LCell>(), nband, nrim, L, false,
                              nullptr));
  }
  HIPCHK(hipEventRecord(e1, nullptr));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  g_last_ms = ms;
  if (s->nrx > 0) {
    const int nt = s->nt;
    std::vector<float> t((size_t)5 * nt * s->nrx);
    HIPCHK(hipMemcpy(t.data(), recd.p, t.size() * 4, hipMemcpyDeviceToHost));
    for (int f = 0; f < 5; ++f)
      for (int ix = 0; ix < s->nrx; ++ix)
        for (int it = 0; it < nt; ++it)
          rec[(size_t)f * nt * s->nrx + it + (size_t)nt * ix] = t[((size_t)f * nt + it) * s->nrx + ix];
  }
  if (last) {
    const size_t nn = (size_t)s->nz * s->nx;
    for (int lv = 0; lv < 2; ++lv)
      for (int f = 0; f < 5; ++f)
        HIPCHK(hipMemcpy2D(last + (size_t)(5 * lv + f) * nn, s->nz * 4, g.level(s->nt - 1 + lv) + f * g.n2,
                           g.ldz * 4, s->nz * 4, s->nx, hipMemcpyDeviceToHost));
  }
  return OK;
}
