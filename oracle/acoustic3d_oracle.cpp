// acoustic3d_oracle.cpp -- CPU ORACLE FOR TESTS ONLY (3D acoustic subdomain).
//
// Test infrastructure: restates the 3D acoustic propagator and
// cross-correlation imaging of the reference's multi-GPU code
// (W/wave_fkernel.3d8o.cu:31-106 wave_kernel, :183-210 damp_kernel,
// W/gpu_funcs_3d.cu:64-75 source injection, :77-109 data injection /
// extraction, :111-120 img_kernel) at 10th order (radius 5) on one GPU's
// subdomain. Only tests/ and bench's cpu leg load it.
//
// Grid (n1, n2, n3) = (z fastest, x, y slowest), index i1 + n1*(i2 + n2*i3).
// Cells closer than 5 to any face (the halo, FAT in the reference) are never
// updated. One step, for every interior cell:
//   lap = 3 w0 p + sum_axes sum_k w_k (p[+k] + p[-k])          (index units)
//   p_next = K (2 p - p_prev + c lap),  c = (v dt / dx)^2
// with the separable sponge K = kz*kx*ky of W/abc-mod.f90:35-78 (the 2D
// taper applied per axis; 1 inside). The reference damps p0 and p1 in place
// before every step (damp_kernel); the factor is applied to the new level
// here instead, which keeps the step one pass over memory.
// A point source adds amp[t] to p_next at (sz, sx, sy); receivers on plane
// z = zr record p_next. The receiver wavefield is propagated backwards with
// the same step and the data injected on plane zr; every j-th level the image
// accumulates src(t) * rec(t) (img_kernel without the reference's 1e-6).
//
// pinned by: exact rational 2nd-derivative weights, polynomial exactness of
// the Laplacian up to degree 11, and the homogeneous-medium Green's function
// (tests/test_acoustic3d_pins.py).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace ac3 {

template <typename T>
struct W {
  T w0, w1, w2, w3, w4, w5;
};
template <typename T>
static W<T> weights() {
  return {T(-5269.0) / T(1800.0), T(5.0) / T(3.0), T(-5.0) / T(21.0), T(5.0) / T(126.0),
          T(-5.0) / T(1008.0), T(1.0) / T(3150.0)};
}

struct Grid {
  int n1, n2, n3;
};
static inline int64_t idx(const Grid& g, int i1, int i2, int i3) {
  return (int64_t)i1 + (int64_t)g.n1 * ((int64_t)i2 + (int64_t)g.n2 * i3);
}

template <typename T>
static void step(const Grid& g, const T* c, const T* kz, const T* kx, const T* ky, const T* p,
                 T* prev_next /* p_prev in, p_next out */, int nth) {
  const W<T> w = weights<T>();
  const int64_t s1 = 1, s2 = g.n1, s3 = (int64_t)g.n1 * g.n2;
#pragma omp parallel for num_threads(nth) schedule(static) if (nth > 1)
  for (int i3 = 5; i3 < g.n3 - 5; ++i3)
    for (int i2 = 5; i2 < g.n2 - 5; ++i2)
      for (int i1 = 5; i1 < g.n1 - 5; ++i1) {
        const int64_t i = idx(g, i1, i2, i3);
        T lap = T(3) * w.w0 * p[i];
        const T wk[5] = {w.w1, w.w2, w.w3, w.w4, w.w5};
        for (int k = 1; k <= 5; ++k)
          lap += wk[k - 1] * ((p[i + k * s1] + p[i - k * s1]) + (p[i + k * s2] + p[i - k * s2]) +
                              (p[i + k * s3] + p[i - k * s3]));
        const T kk = kz ? kz[i1] * kx[i2] * ky[i3] : T(1);
        prev_next[i] = kk * (T(2) * p[i] - prev_next[i] + c[i] * lap);
      }
}

}  // namespace ac3

#define AC3_EXPORT extern "C" __attribute__((visibility("default")))

// One propagation of nt steps from zero initial fields.
//   mode 0 (source): amp[t] added at (sz,sx,sy); rec (n2*n3 per step, [t][i3][i2]) records plane zr
//   mode 1 (receiver): data rec[t] injected on plane zr, stepping t = nt-1..0
// snap: optional (n_snap * n) source snapshots p(t) at t % j == 0 (mode 0 output)
// img: optional image (mode 1): img += snap(t) * q(t) at t % j == 0
// The level index t counts computed levels 1..nt: level t is the field after
// t steps; snapshots and images use levels t with t % j == 0.
template <typename T>
static void run(int n1, int n2, int n3, int nt, const T* c, const T* kz, const T* kx, const T* ky,
                int mode, int sz, int sx, int sy, const T* amp, int zr, T* rec, int j, T* snap,
                T* img, int nth) {
  ac3::Grid g{n1, n2, n3};
  const int64_t n = (int64_t)n1 * n2 * n3;
  std::vector<T> a(n, T(0)), b(n, T(0));
  T* p = a.data();     // current level
  T* pp = b.data();    // previous level, overwritten by the next one
  const int64_t plane = (int64_t)n2 * n3;
  for (int s = 1; s <= nt; ++s) {
    const int t = mode == 0 ? s : nt + 1 - s;  // level being produced
    ac3::step<T>(g, c, kz, kx, ky, p, pp, nth);
    if (mode == 0) {
      pp[ac3::idx(g, sz, sx, sy)] += amp[t - 1];
      if (rec)
        for (int i3 = 0; i3 < n3; ++i3)
          for (int i2 = 0; i2 < n2; ++i2) rec[(int64_t)(t - 1) * plane + (int64_t)i3 * n2 + i2] = pp[ac3::idx(g, zr, i2, i3)];
      if (snap && t % j == 0) std::memcpy(snap + (int64_t)(t / j - 1) * n, pp, sizeof(T) * n);
    } else {
      for (int i3 = 5; i3 < n3 - 5; ++i3)
        for (int i2 = 5; i2 < n2 - 5; ++i2)
          pp[ac3::idx(g, zr, i2, i3)] += rec[(int64_t)(t - 1) * plane + (int64_t)i3 * n2 + i2];
      if (img && t % j == 0) {
        const T* sv = snap + (int64_t)(t / j - 1) * n;
        for (int64_t i = 0; i < n; ++i) img[i] += sv[i] * pp[i];
      }
    }
    std::swap(p, pp);
  }
}

AC3_EXPORT void ac3_weights_f64(double* w) {
  auto k = ac3::weights<double>();
  w[0] = k.w0; w[1] = k.w1; w[2] = k.w2; w[3] = k.w3; w[4] = k.w4; w[5] = k.w5;
}

AC3_EXPORT void ac3_laplacian_f64(int n1, int n2, int n3, const double* p, double* out) {
  // lap in index units for interior cells (others 0): one step with c = 1,
  // p_prev = 0, K = 1 gives 2p + lap; subtract 2p.
  ac3::Grid g{n1, n2, n3};
  const int64_t n = (int64_t)n1 * n2 * n3;
  std::vector<double> c(n, 1.0);
  std::memset(out, 0, sizeof(double) * n);
  ac3::step<double>(g, c.data(), nullptr, nullptr, nullptr, p, out, 1);
  for (int i3 = 5; i3 < n3 - 5; ++i3)
    for (int i2 = 5; i2 < n2 - 5; ++i2)
      for (int i1 = 5; i1 < n1 - 5; ++i1) {
        const int64_t i = ac3::idx(g, i1, i2, i3);
        out[i] -= 2.0 * p[i];
      }
}

#define AC3_RUN(T, SFX)                                                                          \
  AC3_EXPORT void ac3_run_##SFX(int n1, int n2, int n3, int nt, const T* c, const T* kz,         \
                                const T* kx, const T* ky, int mode, int sz, int sx, int sy,      \
                                const T* amp, int zr, T* rec, int j, T* snap, T* img, int nth) { \
    run<T>(n1, n2, n3, nt, c, kz, kx, ky, mode, sz, sx, sy, amp, zr, rec, j, snap, img, nth);    \
  }
AC3_RUN(double, f64)
AC3_RUN(float, f32)
