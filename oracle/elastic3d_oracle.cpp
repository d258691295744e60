// elastic3d_oracle.cpp -- CPU ORACLE FOR TESTS ONLY (3D elastic adjoint step).
//
// Test infrastructure: restates the reference's 9-field 3D elastic adjoint
// time step, the CPU loop of K3/rtmlib-modC.c:277-589 (the same operator as
// the CUDA kernel rtm_gpu_kernel, K3/rtmlib-modC-gpu.cu:374-690), and its
// multi-step pointer rotation (K3/rtmlib-modC.c:597-690,
// K3/rtmlib-modC-gpu.cu:273-372). Only tests/ and bench's cpu leg load it.
//
// Layout (K3/gpu.h n3d_index_ex): every array is padded by 5 cells on each
// face, (nz+10) x (nx+10) x (ny+10), z fastest, then x, y slowest; interior
// cell (iz, ix, iy) sits at (iz+5) + (nz+10)*((ix+5) + (nx+10)*(iy+5)).
// Field order (the reference's argument order): Vy, Vx, Vz, syy, sxx, szz,
// sxy, sxz, syz. Coefficient order: m1_y, m1_x, m1_z, m2, m3, m2m3.
//
// One step, for every interior cell:
//   out = out + in1 + sum over the output's terms of a 10-tap gather
//   gather(m, v)[i] = sum_j w_j * m[i+o_j] * v[i+o_j]   along one axis
// with two tap sets (c1..c5 the 10th-order staggered weights):
//   set A: offsets -5..4, weights +c1 +c2 +c3 +c4 +c5 -c5 -c4 -c3 -c2 -c1
//   set B: offsets -4..5, same weights.
// The terms of each output (coefficient, input field, axis, set) are the
// table `kTerms` below, in the reference's order; in FP32 each tap is
// rounded as (m * c) * v and added left to right, as the reference's C does.
//
// The forward partner el3_fwd_apply_f64 is the same operator's transpose
// written as the velocity-stress system it discretises, with
//   D+ f[i] = sum_d c_(6-d) (f[i+d] - f[i-d+1])  (derivative at i + 1/2)
//   D- f[i] = sum_d c_(6-d) (f[i+d-1] - f[i-d])  (derivative at i - 1/2)
//   Vx  += m1_x (D-x sxx + D-y sxy + D-z sxz)         (and Vy, Vz alike)
//   sxx += m2m3 D+x Vx + m2 (D+y Vy + D+z Vz)         (and syy, szz alike)
//   sxy += m3 (D-y Vx + D-x Vy)                       (and sxz, syz alike)
//
// pinned by: polynomial exactness of every term against closed-form
// derivatives at i -/+ 1/2 with distinct constant coefficients, the
// permutation symmetry of the axes, the dot-product test of the gather
// against el3_fwd_apply_f64, and FP32 against FP64
// (tests/test_elastic3d_pins.py). The reference tree has no stored outputs
// and its step test needs its CUDA half to link, so no reference build.

#include <cstddef>
#include <cstring>
#include <vector>

namespace el3 {

enum { VY, VX, VZ, SYY, SXX, SZZ, SXY, SXZ, SYZ };
enum { M1Y, M1X, M1Z, M2, M3, M2M3 };
enum { AZ_, AX_, AY_ };  // axis: z, x, y
enum { SET_A, SET_B };

struct Term {
  int out, coef, field, axis, set;
};

// K3/rtmlib-modC.c: Vx :288-349, Vy :350-413, Vz :414-476, sxx :477-489,
// syy :490-502, szz :503-517, sxy :518-542, sxz :543-565, syz :566-589.
constexpr Term kTerms[] = {
    {VX, M2M3, SXX, AX_, SET_A}, {VX, M2, SYY, AX_, SET_A},   {VX, M2, SZZ, AX_, SET_A},
    {VX, M3, SXY, AY_, SET_B},   {VX, M3, SXZ, AZ_, SET_B},
    {VY, M2M3, SYY, AY_, SET_A}, {VY, M2, SZZ, AY_, SET_A},   {VY, M2, SXX, AY_, SET_A},
    {VY, M3, SYZ, AZ_, SET_B},   {VY, M3, SXY, AX_, SET_B},
    {VZ, M2M3, SZZ, AZ_, SET_A}, {VZ, M2, SXX, AZ_, SET_A},   {VZ, M2, SYY, AZ_, SET_A},
    {VZ, M3, SYZ, AY_, SET_B},   {VZ, M3, SXZ, AX_, SET_B},
    {SXX, M1X, VX, AX_, SET_B},
    {SYY, M1Y, VY, AY_, SET_B},
    {SZZ, M1Z, VZ, AZ_, SET_B},
    {SXY, M1Y, VY, AX_, SET_B},  {SXY, M1X, VX, AY_, SET_B},
    {SXZ, M1X, VX, AZ_, SET_B},  {SXZ, M1Z, VZ, AX_, SET_B},
    {SYZ, M1Y, VY, AZ_, SET_B},  {SYZ, M1Z, VZ, AY_, SET_B},
};
constexpr int kNTerms = sizeof(kTerms) / sizeof(kTerms[0]);

template <typename T>
struct C5 {
  T c[5];
};
// The reference writes float c1 = 35.0/294912.0 etc.: the double quotient
// rounded to the working type.
template <typename T>
C5<T> coeffs() {
  return {{(T)(35.0 / 294912.0), (T)(-405.0 / 229376.0), (T)(567.0 / 40960.0), (T)(-735.0 / 8192.0),
           (T)(19845.0 / 16384.0)}};
}

struct Geo {
  int ny, nz, nx;
  long long sz, sx, sy, n;  // strides of z, x, y and the padded size
  Geo(int ny_, int nz_, int nx_) : ny(ny_), nz(nz_), nx(nx_) {
    sz = 1;
    sx = nz + 10;
    sy = (long long)(nz + 10) * (nx + 10);
    n = sy * (ny + 10);
  }
  long long at(int iz, int ix, int iy) const { return (iz + 5) + sx * (ix + 5) + sy * (iy + 5); }
  long long stride(int axis) const { return axis == AZ_ ? sz : axis == AX_ ? sx : sy; }
};

// One step over every interior cell, reference term order.
template <typename T>
void step(const Geo& g, const T* in, const T* in1, T* out, const T* coef, int nthreads) {
  const C5<T> c = coeffs<T>();
  const long long n = g.n;
#pragma omp parallel for collapse(2) num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
  for (int iy = 0; iy < g.ny; ++iy)
    for (int ix = 0; ix < g.nx; ++ix)
      for (int iz = 0; iz < g.nz; ++iz) {
        const long long i = g.at(iz, ix, iy);
        T acc[9];
        for (int f = 0; f < 9; ++f) acc[f] = out[f * n + i] + in1[f * n + i];
        for (int t = 0; t < kNTerms; ++t) {
          const Term& tm = kTerms[t];
          const T* m = coef + tm.coef * n;
          const T* v = in + tm.field * n;
          const long long s = g.stride(tm.axis);
          const int o0 = tm.set == SET_A ? -5 : -4;
          T a = acc[tm.out];
          for (int j = 0; j < 10; ++j) {
            const long long k = i + s * (o0 + j);
            if (j < 5)
              a = a + (m[k] * c.c[j]) * v[k];
            else
              a = a - (m[k] * c.c[9 - j]) * v[k];
          }
          acc[tm.out] = a;
        }
        for (int f = 0; f < 9; ++f) out[f * n + i] = acc[f];
      }
}

// The reference's rotation after a step that is not the last:
// (out, in, in1) <- (in1, out, in).
template <typename T>
void steps(const Geo& g, T* b0, T* b1, T* b2, const T* coef, int nsteps, int nthreads, int* final_slots) {
  T* in = b0;
  T* in1 = b1;
  T* out = b2;
  int si = 0, si1 = 1, so = 2;
  for (int s = 0; s < nsteps; ++s) {
    step<T>(g, in, in1, out, coef, nthreads);
    if (s < nsteps - 1) {
      T* t = out;
      out = in1;
      in1 = in;
      in = t;
      const int ts = so;
      so = si1;
      si1 = si;
      si = ts;
    }
  }
  final_slots[0] = si;
  final_slots[1] = si1;
  final_slots[2] = so;
}

// D+ / D- along one axis at interior cell i.
inline double dplus(const double* f, long long i, long long s, const C5<double>& c) {
  double a = 0;
  for (int d = 1; d <= 5; ++d) a += c.c[5 - d] * (f[i + s * d] - f[i - s * (d - 1)]);
  return a;
}
inline double dminus(const double* f, long long i, long long s, const C5<double>& c) {
  double a = 0;
  for (int d = 1; d <= 5; ++d) a += c.c[5 - d] * (f[i + s * (d - 1)] - f[i - s * d]);
  return a;
}

}  // namespace el3

using namespace el3;

extern "C" {

// in, in1, out: 9 consecutive padded fields each; coef: 6 padded arrays.
void el3_adj_step_f64(int ny, int nz, int nx, const double* in, const double* in1, double* out,
                      const double* coef, int nthreads) {
  step<double>(Geo(ny, nz, nx), in, in1, out, coef, nthreads);
}

void el3_adj_step_f32(int ny, int nz, int nx, const float* in, const float* in1, float* out, const float* coef,
                      int nthreads) {
  step<float>(Geo(ny, nz, nx), in, in1, out, coef, nthreads);
}

// nsteps with the reference's rotation; the buffers start as (in, in1, out)
// = (b0, b1, b2); slots[3] receives which buffer ends as in, in1, out.
void el3_adj_steps_f64(int ny, int nz, int nx, double* b0, double* b1, double* b2, const double* coef, int nsteps,
                       int nthreads, int* slots) {
  steps<double>(Geo(ny, nz, nx), b0, b1, b2, coef, nsteps, nthreads, slots);
}

void el3_adj_steps_f32(int ny, int nz, int nx, float* b0, float* b1, float* b2, const float* coef, int nsteps,
                       int nthreads, int* slots) {
  steps<float>(Geo(ny, nz, nx), b0, b1, b2, coef, nsteps, nthreads, slots);
}

// y = F x over interior cells (y zero elsewhere), F the transpose of the
// step's gather part, written as the velocity-stress system.
void el3_fwd_apply_f64(int ny, int nz, int nx, const double* x, double* y, const double* coef, int nthreads) {
  const Geo g(ny, nz, nx);
  const C5<double> c = coeffs<double>();
  const long long n = g.n;
  std::memset(y, 0, sizeof(double) * 9 * n);
  const long long SZ = g.sz, SX = g.sx, SY = g.sy;
#pragma omp parallel for collapse(2) num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
  for (int iy = 0; iy < g.ny; ++iy)
    for (int ix = 0; ix < g.nx; ++ix)
      for (int iz = 0; iz < g.nz; ++iz) {
        const long long i = g.at(iz, ix, iy);
        const double* f[9];
        for (int k = 0; k < 9; ++k) f[k] = x + k * n;
        const double m1y = coef[M1Y * n + i], m1x = coef[M1X * n + i], m1z = coef[M1Z * n + i];
        const double m2 = coef[M2 * n + i], m3 = coef[M3 * n + i], m2m3 = coef[M2M3 * n + i];
        y[VX * n + i] = m1x * (dminus(f[SXX], i, SX, c) + dminus(f[SXY], i, SY, c) + dminus(f[SXZ], i, SZ, c));
        y[VY * n + i] = m1y * (dminus(f[SXY], i, SX, c) + dminus(f[SYY], i, SY, c) + dminus(f[SYZ], i, SZ, c));
        y[VZ * n + i] = m1z * (dminus(f[SXZ], i, SX, c) + dminus(f[SYZ], i, SY, c) + dminus(f[SZZ], i, SZ, c));
        const double dxvx = dplus(f[VX], i, SX, c), dyvy = dplus(f[VY], i, SY, c), dzvz = dplus(f[VZ], i, SZ, c);
        y[SXX * n + i] = m2m3 * dxvx + m2 * (dyvy + dzvz);
        y[SYY * n + i] = m2m3 * dyvy + m2 * (dxvx + dzvz);
        y[SZZ * n + i] = m2m3 * dzvz + m2 * (dxvx + dyvy);
        y[SXY * n + i] = m3 * (dminus(f[VX], i, SY, c) + dminus(f[VY], i, SX, c));
        y[SXZ * n + i] = m3 * (dminus(f[VX], i, SZ, c) + dminus(f[VZ], i, SX, c));
        y[SYZ * n + i] = m3 * (dminus(f[VY], i, SZ, c) + dminus(f[VZ], i, SY, c));
      }
}

}  // extern "C"
