// rtm_oracle.cpp -- CPU ORACLE FOR TESTS ONLY.
//
// This file is test infrastructure. It is a plain CPU restatement of the
// reference's 2D elastic RTM hot path, used as the checker for the HIP
// library. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it. The product library never links or calls it.
//
// Every routine is templated on the real type T:
//   T = double : the FP64 accuracy oracle (tolerance gates are taken against it)
//   T = float  : FP32 with the reference's operation order, no FMA
//                (built with -ffp-contract=off), used as the CPU baseline and to
//                compare bit-for-bit with the reference's own C adjoint.
//
// Layout (reference layout, W/gpu.h:8-14): a field (nz,nx) is column-major,
// index iz + nz*ix, 0-based. A history (nz,nx,nt) is iz + nz*ix + nz*nx*it.
// Receiver data (nt,nrx) per component is it + nt*ix (W/rtmlib-mod.f90:76).
//
// Pinning status of each routine is recorded in the header comment of the
// routine ("pinned by: ..."); the tests under tests/ implement those checks.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// a1. 10th-order staggered first-derivative weights, W/rtmlib-mod.f90:15 (also
// W/rtmlib-modC.c:45, W/image-mod-rtm.f90:14). Fortran default-real literals:
// the quotient of two exactly representable singles, rounded once.
// pinned by: exact rational solution of the staggered Taylor system
// (tests/test_oracle_pins.py::test_fd_coefficients_exact).
template <typename T>
struct Coef {
  T c1, c2, c3, c4, c5;
};
template <typename T>
static Coef<T> coefs() {
  return {T(35.0) / T(294912.0), T(-405.0) / T(229376.0), T(567.0) / T(40960.0),
          T(-735.0) / T(8192.0), T(19845.0) / T(16384.0)};
}

static inline int64_t I2(int nz, int iz, int ix) { return (int64_t)iz + (int64_t)nz * ix; }

// Staggered difference on the "+" grid (derivative at i+1/2):
//   S+ f(i) = c1(f[i+5]-f[i-4]) + c2(f[i+4]-f[i-3]) + c3(f[i+3]-f[i-2])
//           + c4(f[i+2]-f[i-1]) + c5(f[i+1]-f[i])
// Summed left to right as Fortran evaluates it. Used by D, E, F, G, Az, Bx
// (W/rtmlib-mod.f90:432-440 etc.) and vxdx / vzdz (W/image-mod-rtm.f90:334).
template <typename T>
static inline T splus(const Coef<T>& c, const T* f, int64_t i, int64_t s) {
  return c.c1 * (f[i + 5 * s] - f[i - 4 * s]) + c.c2 * (f[i + 4 * s] - f[i - 3 * s]) +
         c.c3 * (f[i + 3 * s] - f[i - 2 * s]) + c.c4 * (f[i + 2 * s] - f[i - s]) +
         c.c5 * (f[i + s] - f[i]);
}
// Staggered difference on the "-" grid (derivative at i-1/2):
//   S- f(i) = c1(f[i+4]-f[i-5]) + ... + c5(f[i]-f[i-1])
// Used by Ax, Bz, H, J (W/rtmlib-mod.f90:229-237 etc.) and vxdz / vzdx.
template <typename T>
static inline T sminus(const Coef<T>& c, const T* f, int64_t i, int64_t s) {
  return c.c1 * (f[i + 4 * s] - f[i - 5 * s]) + c.c2 * (f[i + 3 * s] - f[i - 4 * s]) +
         c.c3 * (f[i + 2 * s] - f[i - 3 * s]) + c.c4 * (f[i + s] - f[i - 2 * s]) +
         c.c5 * (f[i] - f[i - s]);
}

// ---------------------------------------------------------------------------
// Operator table. 0-based inclusive ranges from the 1-based Fortran loops.
//   op  name coef      shape axis  ix range      iz range     (Fortran lines)
//   0   Ax   m1_x      S-    x     5..nx-6       4..nz-6      W/rtmlib-mod.f90:229-237
//   1   Bx   m1_x      S+    z     5..nx-6       4..nz-6      :268-276
//   2   Az   m1_z      S+    x     4..nx-6       5..nz-6      :306-314
//   3   Bz   m1_z      S-    z     4..nx-6       5..nz-6      :344-352
//   4   D    aux_m2m3  S+    x     4..nx-6       4..nz-6      :432-440
//   5   E    aux_m2    S+    z     4..nx-6       4..nz-6      :470-478
//   6   F    aux_m2    S+    x     4..nx-6       4..nz-6      :508-516
//   7   G    aux_m2m3  S+    z     4..nx-6       4..nz-6      :546-554
//   8   H    aux_m3    S-    z     5..nx-6       5..nz-6      :584-592
//   9   J    aux_m3    S-    x     5..nx-6       5..nz-6      :622-630
struct OpDesc {
  int plus;  // 1: S+ shape, 0: S- shape
  int xaxis; // 1: derivative along x, 0: along z
  int ix0, iz0; // first index of the range (last is n-6 for both axes)
};
static const OpDesc kOps[10] = {
    {0, 1, 5, 4}, {1, 0, 5, 4}, {1, 1, 4, 5}, {0, 0, 4, 5}, {1, 1, 4, 4},
    {1, 0, 4, 4}, {1, 1, 4, 4}, {1, 0, 4, 4}, {0, 0, 5, 5}, {0, 1, 5, 5}};

// a5. Forward operator (adj=.false., add=.true.): data += coef * S(model).
// pinned by: polynomial exactness and plane-wave symbol of S+/S-, and FP64
// dot-product tests against the scatter adjoint (method of W/dptest_mod.f90:8-37).
template <typename T>
static void fwd_op(int op, int nz, int nx, const T* coef, const T* model, T* data, int nth) {
  const Coef<T> c = coefs<T>();
  const OpDesc& d = kOps[op];
  const int64_t s = d.xaxis ? nz : 1;
#pragma omp parallel for num_threads(nth) schedule(static) if (nth > 1)
  for (int ix = d.ix0; ix <= nx - 6; ++ix)
    for (int iz = d.iz0; iz <= nz - 6; ++iz) {
      const int64_t i = I2(nz, iz, ix);
      const T st = d.plus ? splus(c, model, i, s) : sminus(c, model, i, s);
      data[i] = data[i] + coef[i] * st;
    }
}

// a9/a10 scatter adjoint (adj=.true., add=.true.) in the Fortran loop and
// write order (W/rtmlib-mod.f90:200-633 adj branches; the same order as the
// reference C scatter loops of R/Src_cuda_722-1/rtmlib-modC.c:69-271).
// pinned by: bit-equality (T=float) with the reference's own C adjoint built
// into oracle/_ref, and FP64 dot-product tests against fwd_op.
template <typename T>
static void adj_op_scatter(int op, int nz, int nx, const T* coef, T* model, const T* data) {
  const Coef<T> c = coefs<T>();
  const OpDesc& d = kOps[op];
  const int64_t s = d.xaxis ? nz : 1;
  for (int ix = d.ix0; ix <= nx - 6; ++ix)
    for (int iz = d.iz0; iz <= nz - 6; ++iz) {
      const int64_t i = I2(nz, iz, ix);
      const T a = coef[i];
      const T y = data[i];
      if (d.plus) {  // transpose of S+: targets i+5 .. i-4
        model[i + 5 * s] = model[i + 5 * s] + a * c.c1 * y;
        model[i + 4 * s] = model[i + 4 * s] + a * c.c2 * y;
        model[i + 3 * s] = model[i + 3 * s] + a * c.c3 * y;
        model[i + 2 * s] = model[i + 2 * s] + a * c.c4 * y;
        model[i + 1 * s] = model[i + 1 * s] + a * c.c5 * y;
        model[i] = model[i] - a * c.c5 * y;
        model[i - 1 * s] = model[i - 1 * s] - a * c.c4 * y;
        model[i - 2 * s] = model[i - 2 * s] - a * c.c3 * y;
        model[i - 3 * s] = model[i - 3 * s] - a * c.c2 * y;
        model[i - 4 * s] = model[i - 4 * s] - a * c.c1 * y;
      } else {  // transpose of S-: targets i+4 .. i-5
        model[i + 4 * s] = model[i + 4 * s] + a * c.c1 * y;
        model[i + 3 * s] = model[i + 3 * s] + a * c.c2 * y;
        model[i + 2 * s] = model[i + 2 * s] + a * c.c3 * y;
        model[i + 1 * s] = model[i + 1 * s] + a * c.c4 * y;
        model[i] = model[i] + a * c.c5 * y;
        model[i - 1 * s] = model[i - 1 * s] - a * c.c5 * y;
        model[i - 2 * s] = model[i - 2 * s] - a * c.c4 * y;
        model[i - 3 * s] = model[i - 3 * s] - a * c.c3 * y;
        model[i - 4 * s] = model[i - 4 * s] - a * c.c2 * y;
        model[i - 5 * s] = model[i - 5 * s] - a * c.c1 * y;
      }
    }
}

// Gather form of one adjoint operator for one output cell, in the term order of
// W/rtmlib-modC.c:125-253: the coefficient array is zero outside the operator's
// range (W/rtmlib-modC.c:83-93) and neighbours outside the grid are zero (the
// 5-cell padding of W/rtmlib-modC.c:47-81).
//   S+ transpose gathers i-5 .. i+4 (+c1..+c5, -c5..-c1)
//   S- transpose gathers i-4 .. i+5 (+c1..+c5, -c5..-c1)
template <typename T>
static inline T gather_term(const Coef<T>& c, int op, int nz, int nx, const T* coef, const T* y,
                            int iz, int ix, T acc) {
  const OpDesc& d = kOps[op];
  const T w[10] = {c.c1, c.c2, c.c3, c.c4, c.c5, c.c5, c.c4, c.c3, c.c2, c.c1};
  const int first = d.plus ? -5 : -4;
  for (int k = 0; k < 10; ++k) {
    const int off = first + k;
    const int jz = d.xaxis ? iz : iz + off;
    const int jx = d.xaxis ? ix + off : ix;
    T a = T(0), v = T(0);
    if (jz >= 0 && jz < nz && jx >= 0 && jx < nx) {
      const bool in = jx >= d.ix0 && jx <= nx - 6 && jz >= d.iz0 && jz <= nz - 6;
      a = in ? coef[I2(nz, jz, jx)] : T(0);
      v = y[I2(nz, jz, jx)];
    }
    if (k < 5)
      acc = acc + a * w[k] * v;
    else
      acc = acc - a * w[k] * v;
  }
  return acc;
}

// ---------------------------------------------------------------------------
// a3. model_op + rtmmod_init + aux_*_c. W/model-mod.f90:21-89,
// W/rtmlib-mod.f90:51-58, W/rtm-mod-adj.f90:111-113.
// Inputs rho (m1), lambda (m2), mu (m3) are (nz,nx); outputs are the five
// dt/dx-scaled coefficient arrays handed to rtm_op3_c_.
// pinned by: closed forms (constant model -> constant coefficients, mu=0 ->
// aux_m3=0, harmonic mean of a 2x2 block) in tests/test_oracle_pins.py.
template <typename T>
static void model_op(int nz, int nx, T dx, T dt, const T* rho, const T* lam_in, const T* mu_in,
                     T* m1_x, T* m1_z, T* aux_m2, T* aux_m3, T* aux_m2m3) {
  const int64_t n = (int64_t)nz * nx;
  std::vector<T> m2(lam_in, lam_in + n), m3(mu_in, mu_in + n);
  for (int64_t i = 0; i < n; ++i) m1_x[i] = m1_z[i] = T(0);
  // Rho_vx: average along z (model-mod.f90:44-48)
  for (int ix = 0; ix < nx; ++ix)
    for (int iz = 0; iz < nz - 1; ++iz)
      m1_x[I2(nz, iz, ix)] = T(1.0) / (T(0.5) * (rho[I2(nz, iz + 1, ix)] + rho[I2(nz, iz, ix)]));
  // Rho_vz: average along x (:50-54)
  for (int ix = 0; ix < nx - 1; ++ix)
    for (int iz = 0; iz < nz; ++iz)
      m1_z[I2(nz, iz, ix)] = T(1.0) / (T(0.5) * (rho[I2(nz, iz, ix + 1)] + rho[I2(nz, iz, ix)]));
  for (int iz = 0; iz < nz; ++iz) m1_z[I2(nz, iz, nx - 1)] = T(1.0) / rho[I2(nz, iz, nx - 1)];
  for (int ix = 0; ix < nx; ++ix) m1_x[I2(nz, nz - 1, ix)] = T(1.0) / rho[I2(nz, nz - 1, ix)];
  // Harmonic 4-point means, computed IN PLACE in loop order (:66-75): the
  // (iz,ix) update reads (iz+1,ix), (iz,ix+1), (iz+1,ix+1), which have not been
  // overwritten yet in this column-major sweep, so in-place equals out-of-place
  // except that the reads see original values -- we reproduce the in-place form.
  for (int ix = 1; ix < nx - 1; ++ix)
    for (int iz = 1; iz < nz - 1; ++iz) {
      const int64_t a = I2(nz, iz, ix), b = I2(nz, iz + 1, ix), cc = I2(nz, iz, ix + 1),
                    dd = I2(nz, iz + 1, ix + 1);
      m2[a] = T(4.0) / ((T(1.0) / m2[a]) + (T(1.0) / m2[b]) + (T(1.0) / m2[cc]) + (T(1.0) / m2[dd]));
      if (m3[a] == T(0) || m3[b] == T(0) || m3[cc] == T(0) || m3[dd] == T(0))
        m3[a] = T(0);
      else
        m3[a] = T(4.0) / ((T(1.0) / m3[a]) + (T(1.0) / m3[b]) + (T(1.0) / m3[cc]) + (T(1.0) / m3[dd]));
    }
  for (int ix = 0; ix < nx; ++ix) {  // (:76-81)
    m2[I2(nz, 0, ix)] = m2[I2(nz, 1, ix)];
    m2[I2(nz, nz - 1, ix)] = m2[I2(nz, nz - 2, ix)];
    m3[I2(nz, 0, ix)] = m3[I2(nz, 1, ix)];
    m3[I2(nz, nz - 1, ix)] = m3[I2(nz, nz - 2, ix)];
  }
  for (int iz = 0; iz < nz; ++iz) {  // (:82-87)
    m2[I2(nz, iz, 0)] = m2[I2(nz, iz, 1)];
    m2[I2(nz, iz, nx - 1)] = m2[I2(nz, iz, nx - 2)];
    m3[I2(nz, iz, 0)] = m3[I2(nz, iz, 1)];
    m3[I2(nz, iz, nx - 1)] = m3[I2(nz, iz, nx - 2)];
  }
  const T aux = dt / dx;  // rtmlib-mod.f90:51
  for (int64_t i = 0; i < n; ++i) {
    m1_x[i] = aux * m1_x[i];             // :53
    m1_z[i] = aux * m1_z[i];             // :54
    aux_m2[i] = (dt / dx) * m2[i];       // rtm-mod-adj.f90:111
    aux_m3[i] = (dt / dx) * m3[i];       // :112
    aux_m2m3[i] = (dt / dx) * (m2[i] + T(2.0) * m3[i]);  // :113
  }
}

// a4. Ricker wavelet, W/source-mod.f90:31-44 (nsource=250 is hard-coded at
// W/rtm-mod-adj.f90:100). out[it-1] = source(it), it = 1..nt.
// pinned by: closed form (zero before it=nsource/10, peak 1e8 at it=nsource/2,
// zero crossings at t=+-1/(pi*sqrt(2)*fp)).
template <typename T>
static void ricker(int nt, T dt, T fp, int nsource, T* out) {
  const int ft = nsource / 10;
  const T pi = T(3.14159265);
  for (int it = 1; it <= nt; ++it) {
    out[it - 1] = T(0);
    if (it < ft) continue;
    const T t = (T(it) * dt) - (T(nsource) * dt / T(2.0));
    const T aux = (pi * pi * fp * fp * t * t);
    out[it - 1] = T(10.0e7) * (T(1.0) - T(2.0) * aux) * std::exp(-aux);
  }
}

// a7. Sponge taper, W/abc-mod.f90:35-50: filter(i)=exp(-0.09 (i/abc)^2) for
// i <= abc, filter(abc) for abc < i <= abc+5. Returns abc+5 values.
// pinned by: closed form; known answer n1=45 for abc=40 (R/Dat/filter.H).
template <typename T>
static void abc_filter(int abc, T* filter) {
  if (abc <= 0) return;
  const int abc2 = abc + 5;
  const T alfa = T(0.09);
  for (int i = 1; i <= abc; ++i) {
    const T aux = (T(1.0) * T(i)) / (T(1.0) * T(abc));
    filter[i - 1] = std::exp(-alfa * aux * aux);
  }
  for (int i = abc + 1; i <= abc2; ++i) filter[i - 1] = filter[abc - 1];
}

// a7. Sponge application K, W/abc-mod.f90:59-78, in its loop order: top and
// bottom rows first, then left and right columns, so corner cells are
// multiplied twice. abc <= 0 disables K.
// pinned by: closed form including the doubly damped corners.
template <typename T>
static void abc_K(int nz, int nx, int abc, const T* filter, T* data) {
  if (abc <= 0) return;  // abc=0 would read filter(0) in the reference
  const int abc2 = abc + 5;
  for (int ix = 0; ix < nx; ++ix)
    for (int iz = 1; iz <= abc2; ++iz) {
      T& top = data[I2(nz, iz - 1, ix)];
      top = top * filter[abc2 - iz];  // filter(abc2-iz+1)
      T& bot = data[I2(nz, nz - abc2 + iz - 1, ix)];
      bot = bot * filter[iz - 1];     // filter(iz)
    }
  for (int ix = 1; ix <= abc2; ++ix)
    for (int iz = 0; iz < nz; ++iz) {
      T& l = data[I2(nz, iz, ix - 1)];
      l = l * filter[abc2 - ix];
      T& r = data[I2(nz, iz, nx - abc2 + ix - 1)];
      r = r * filter[ix - 1];
    }
}

// ---------------------------------------------------------------------------
// a10. The reference adjoint rtm_op3_c_ in gather form, W/rtmlib-modC.c:120-253,
// with its exact term order: for it = nt-3 .. 0,
//   X[it] = X[it] + X[it+2] + sum of gathered terms of Y[it+1].
// pinned by: bit-equality (T=float) of the equivalent scatter form with the
// reference C adjoint (oracle/_ref), FP64 dot-product tests.
template <typename T>
static void rtm_op3_c_gather(int nt, int nz, int nx, T* Vx0, T* Vz0, T* sxx0, T* szz0, T* sxz0,
                             const T* m1_x, const T* m1_z, const T* m2, const T* m3, const T* m2m3,
                             int nth) {
  const Coef<T> c = coefs<T>();
  const int64_t n2 = (int64_t)nz * nx;
  for (int it = nt - 3; it >= 0; --it) {
    T *vx = Vx0 + n2 * it, *vz = Vz0 + n2 * it, *xx = sxx0 + n2 * it, *zz = szz0 + n2 * it,
      *xz = sxz0 + n2 * it;
    const T *vx2 = vx + 2 * n2, *vz2 = vz + 2 * n2, *xx2 = xx + 2 * n2, *zz2 = zz + 2 * n2,
            *xz2 = xz + 2 * n2;
    const T *vx1 = vx + n2, *vz1 = vz + n2, *xx1 = xx + n2, *zz1 = zz + n2, *xz1 = xz + n2;
#pragma omp parallel for num_threads(nth) schedule(static) if (nth > 1)
    for (int ix = 0; ix < nx; ++ix)
      for (int iz = 0; iz < nz; ++iz) {
        const int64_t i = I2(nz, iz, ix);
        T a;
        a = vx[i] + vx2[i];
        a = gather_term(c, 4, nz, nx, m2m3, xx1, iz, ix, a);  // D^T sxx
        a = gather_term(c, 6, nz, nx, m2, zz1, iz, ix, a);    // F^T szz
        a = gather_term(c, 8, nz, nx, m3, xz1, iz, ix, a);    // H^T sxz
        const T nvx = a;
        a = vz[i] + vz2[i];
        a = gather_term(c, 5, nz, nx, m2, xx1, iz, ix, a);    // E^T sxx
        a = gather_term(c, 7, nz, nx, m2m3, zz1, iz, ix, a);  // G^T szz
        a = gather_term(c, 9, nz, nx, m3, xz1, iz, ix, a);    // J^T sxz
        const T nvz = a;
        a = xx[i] + xx2[i];
        a = gather_term(c, 0, nz, nx, m1_x, vx1, iz, ix, a);  // Ax^T vx
        const T nxx = a;
        a = zz[i] + zz2[i];
        a = gather_term(c, 3, nz, nx, m1_z, vz1, iz, ix, a);  // Bz^T vz
        const T nzz = a;
        a = xz[i] + xz2[i];
        a = gather_term(c, 1, nz, nx, m1_x, vx1, iz, ix, a);  // Bx^T vx
        a = gather_term(c, 2, nz, nx, m1_z, vz1, iz, ix, a);  // Az^T vz
        const T nxz = a;
        vx[i] = nvx; vz[i] = nvz; xx[i] = nxx; zz[i] = nzz; xz[i] = nxz;
      }
  }
}

// Scatter form of the same loop, R/Src_cuda_722-1/rtmlib-modC.c:57-271 order:
// first X[it] += X[it+2] for every cell, then D_c F_c H_c (into Vx), E_c G_c
// J_c (into Vz), Ax_c (sxx), Bz_c (szz), Bx_c Az_c (sxz).
template <typename T>
static void rtm_op3_c_scatter(int nt, int nz, int nx, T* Vx0, T* Vz0, T* sxx0, T* szz0, T* sxz0,
                              const T* m1_x, const T* m1_z, const T* m2, const T* m3,
                              const T* m2m3) {
  const int64_t n2 = (int64_t)nz * nx;
  for (int it = nt - 3; it >= 0; --it) {
    T *vx = Vx0 + n2 * it, *vz = Vz0 + n2 * it, *xx = sxx0 + n2 * it, *zz = szz0 + n2 * it,
      *xz = sxz0 + n2 * it;
    for (int64_t i = 0; i < n2; ++i) {
      vx[i] = vx[i] + vx[i + 2 * n2];
      vz[i] = vz[i] + vz[i + 2 * n2];
      xx[i] = xx[i] + xx[i + 2 * n2];
      zz[i] = zz[i] + zz[i + 2 * n2];
      xz[i] = xz[i] + xz[i + 2 * n2];
    }
    adj_op_scatter(4, nz, nx, m2m3, vx, xx + n2);  // D_c
    adj_op_scatter(6, nz, nx, m2, vx, zz + n2);    // F_c
    adj_op_scatter(8, nz, nx, m3, vx, xz + n2);    // H_c
    adj_op_scatter(5, nz, nx, m2, vz, xx + n2);    // E_c
    adj_op_scatter(7, nz, nx, m2m3, vz, zz + n2);  // G_c
    adj_op_scatter(9, nz, nx, m3, vz, xz + n2);    // J_c
    adj_op_scatter(0, nz, nx, m1_x, xx, vx + n2);  // Ax_c
    adj_op_scatter(3, nz, nx, m1_z, zz, vz + n2);  // Bz_c
    adj_op_scatter(1, nz, nx, m1_x, xz, vx + n2);  // Bx_c
    adj_op_scatter(2, nz, nx, m1_z, xz, vz + n2);  // Az_c
  }
}

// a9. rtm_op3 pre-step and data injection, W/rtmlib-mod.f90:91-146: zero the
// five histories, overwrite row zrec at level nt with the data, then at level
// nt-1 overwrite the row and scatter the adjoint of level nt into it, then
// overwrite row zrec for levels nt-2..1. Data is (nt,nx) per component.
template <typename T>
static void rtm_op3_prestep(int nt, int nz, int nx, int zrec, T* const* hist, const T* const* data,
                            const T* m1_x, const T* m1_z, const T* m2, const T* m3,
                            const T* m2m3) {
  const int64_t n2 = (int64_t)nz * nx;
  for (int f = 0; f < 5; ++f) std::memset(hist[f], 0, sizeof(T) * n2 * nt);
  auto inject = [&](int f, int lev /*0-based*/) {
    for (int ix = 0; ix < nx; ++ix) hist[f][n2 * lev + I2(nz, zrec, ix)] = data[f][lev + (int64_t)nt * ix];
  };
  for (int f = 0; f < 5; ++f) inject(f, nt - 1);
  T *vxm = hist[0] + n2 * (nt - 2), *vzm = hist[1] + n2 * (nt - 2), *xxm = hist[2] + n2 * (nt - 2),
    *zzm = hist[3] + n2 * (nt - 2), *xzm = hist[4] + n2 * (nt - 2);
  const T *vxn = hist[0] + n2 * (nt - 1), *vzn = hist[1] + n2 * (nt - 1),
          *xxn = hist[2] + n2 * (nt - 1), *zzn = hist[3] + n2 * (nt - 1),
          *xzn = hist[4] + n2 * (nt - 1);
  inject(0, nt - 2);
  adj_op_scatter(4, nz, nx, m2m3, vxm, xxn);
  adj_op_scatter(6, nz, nx, m2, vxm, zzn);
  adj_op_scatter(8, nz, nx, m3, vxm, xzn);
  inject(1, nt - 2);
  adj_op_scatter(5, nz, nx, m2, vzm, xxn);
  adj_op_scatter(7, nz, nx, m2m3, vzm, zzn);
  adj_op_scatter(9, nz, nx, m3, vzm, xzn);
  inject(2, nt - 2);
  adj_op_scatter(0, nz, nx, m1_x, xxm, vxn);
  inject(3, nt - 2);
  adj_op_scatter(3, nz, nx, m1_z, zzm, vzn);
  inject(4, nt - 2);
  adj_op_scatter(1, nz, nx, m1_x, xzm, vxn);
  adj_op_scatter(2, nz, nx, m1_z, xzm, vzn);
  for (int lev = nt - 3; lev >= 0; --lev)
    for (int f = 0; f < 5; ++f) inject(f, lev);
}

// ---------------------------------------------------------------------------
// Parameters of a full single-shot run on a computational grid (all indices
// 0-based computational coordinates).
struct Params {
  int nz, nx, nt;
  double dx, dt;
  int xs, zs;          // explosive source cell
  int zr, rx0, nrx;    // receiver row and column range
  int abc;             // sponge width; K covers abc+5 cells per edge; <0: no K
  int adj_abc;         // adjoint: 0 = no K (W/ drop-in), 1 = exact adjoint K
  int snap_every;      // imaging decimation j: levels it with it % j == 0, scaled by j
  int nthreads;
};

// One forward level: X(it) = K(X(it-2) [+ src] + Op Y(it-1)), W/rtm-mod-adj.f90:225-252.
// Fields f: 0 Vx, 1 Vz, 2 sxx, 3 szz, 4 sxz. cur[f] holds level it-2 on entry
// and level it on exit; prev[f] holds level it-1. srcamp is dt*source(it-1),
// added at (zs,xs) of sxx and szz before the stencils (explosive_source,
// W/source-mod.f90:47-53). The forward operators only read level it-1, so the
// order of the five field updates does not matter.
template <typename T>
static void fwd_level(const Params& p, T* const* cur, const T* const* prev, const T* m1_x,
                      const T* m1_z, const T* m2, const T* m3, const T* m2m3, const T* filter,
                      bool apply_k, bool add_src, T srcamp, const T* const* bsrc) {
  const int nz = p.nz, nx = p.nx, nth = p.nthreads;
  const int64_t n2 = (int64_t)nz * nx;
  if (bsrc)  // Born: X(it) = X(it) [scattering source] + X(it-2), rtm-mod-adj.f90:325
    for (int f = 0; f < 5; ++f)
      for (int64_t i = 0; i < n2; ++i) cur[f][i] = bsrc[f][i] + cur[f][i];
  // Vx
  fwd_op(0, nz, nx, m1_x, prev[2], cur[0], nth);
  fwd_op(1, nz, nx, m1_x, prev[4], cur[0], nth);
  // Vz
  fwd_op(2, nz, nx, m1_z, prev[4], cur[1], nth);
  fwd_op(3, nz, nx, m1_z, prev[3], cur[1], nth);
  // sxx
  if (add_src) cur[2][I2(nz, p.zs, p.xs)] = cur[2][I2(nz, p.zs, p.xs)] + srcamp;
  fwd_op(4, nz, nx, m2m3, prev[0], cur[2], nth);
  fwd_op(5, nz, nx, m2, prev[1], cur[2], nth);
  // szz
  if (add_src) cur[3][I2(nz, p.zs, p.xs)] = cur[3][I2(nz, p.zs, p.xs)] + srcamp;
  fwd_op(6, nz, nx, m2, prev[0], cur[3], nth);
  fwd_op(7, nz, nx, m2m3, prev[1], cur[3], nth);
  // sxz
  fwd_op(8, nz, nx, m3, prev[0], cur[4], nth);
  fwd_op(9, nz, nx, m3, prev[1], cur[4], nth);
  if (apply_k)
    for (int f = 0; f < 5; ++f) abc_K(nz, nx, p.abc, filter, cur[f]);
}

// One adjoint level in gather form (no injection): X(it) = X(it) + X(it+2) +
// Op^T Y(it+1), optional K after (exact adjoint of the K'd forward).
template <typename T>
static void adj_level(const Params& p, T* const* cur, const T* const* nxt, const T* m1_x,
                      const T* m1_z, const T* m2, const T* m3, const T* m2m3, const T* filter,
                      bool apply_k) {
  const int nz = p.nz, nx = p.nx;
  const int64_t n2 = (int64_t)nz * nx;
  const Coef<T> c = coefs<T>();
  std::vector<T> out(5 * n2);
#pragma omp parallel for num_threads(p.nthreads) schedule(static) if (p.nthreads > 1)
  for (int ix = 0; ix < nx; ++ix)
    for (int iz = 0; iz < nz; ++iz) {
      const int64_t i = I2(nz, iz, ix);
      T a = cur[0][i];
      a = gather_term(c, 4, nz, nx, m2m3, nxt[2], iz, ix, a);
      a = gather_term(c, 6, nz, nx, m2, nxt[3], iz, ix, a);
      a = gather_term(c, 8, nz, nx, m3, nxt[4], iz, ix, a);
      out[i] = a;
      a = cur[1][i];
      a = gather_term(c, 5, nz, nx, m2, nxt[2], iz, ix, a);
      a = gather_term(c, 7, nz, nx, m2m3, nxt[3], iz, ix, a);
      a = gather_term(c, 9, nz, nx, m3, nxt[4], iz, ix, a);
      out[n2 + i] = a;
      a = cur[2][i];
      a = gather_term(c, 0, nz, nx, m1_x, nxt[0], iz, ix, a);
      out[2 * n2 + i] = a;
      a = cur[3][i];
      a = gather_term(c, 3, nz, nx, m1_z, nxt[1], iz, ix, a);
      out[3 * n2 + i] = a;
      a = cur[4][i];
      a = gather_term(c, 1, nz, nx, m1_x, nxt[0], iz, ix, a);
      a = gather_term(c, 2, nz, nx, m1_z, nxt[1], iz, ix, a);
      out[4 * n2 + i] = a;
    }
  for (int f = 0; f < 5; ++f) {
    std::memcpy(cur[f], out.data() + f * n2, sizeof(T) * n2);
    if (apply_k) abc_K(nz, nx, p.abc, filter, cur[f]);
  }
}

// a11. Source-wavefield derivatives at one level, W/image-mod-rtm.f90:26-67
// with T (:268-299), vxdx (:301-344), vzdz (:346-390), vxdz (:392-435),
// vzdx (:437-480). All spatial derivatives are divided by dx (also the
// z-derivatives, :351). S1 = (Vx(it+1)-Vx(it))/dt, zero at the last level.
// vx1/vz1 = level it+1 or nullptr at the last level.
template <typename T>
static void image_S(int nz, int nx, T dx, T dt, const T* vx, const T* vz, const T* vx1,
                    const T* vz1, T* S1, T* S2, T* S3, T* S4, T* S5, T* S6) {
  const Coef<T> c = coefs<T>();
  const int64_t n2 = (int64_t)nz * nx;
  const T at = T(1.0) / dt, ax = T(1.0) / dx;
  for (int64_t i = 0; i < n2; ++i) {
    S1[i] = vx1 ? T(0) + at * (vx1[i] - vx[i]) : T(0);
    S2[i] = vz1 ? T(0) + at * (vz1[i] - vz[i]) : T(0);
    S4[i] = S5[i] = S6[i] = T(0);
  }
  for (int ix = 4; ix <= nx - 6; ++ix)
    for (int iz = 4; iz <= nz - 6; ++iz) {
      const int64_t i = I2(nz, iz, ix);
      S4[i] = S4[i] + splus(c, vx, i, nz) * ax;  // vxdx
      S5[i] = S5[i] + splus(c, vz, i, 1) * ax;   // vzdz
    }
  for (int64_t i = 0; i < n2; ++i) S3[i] = S4[i] + S5[i];
  for (int ix = 5; ix <= nx - 6; ++ix)
    for (int iz = 5; iz <= nz - 6; ++iz) {
      const int64_t i = I2(nz, iz, ix);
      S6[i] = S6[i] + sminus(c, vx, i, 1) * ax;   // vxdz
    }
  for (int ix = 5; ix <= nx - 6; ++ix)
    for (int iz = 5; iz <= nz - 6; ++iz) {
      const int64_t i = I2(nz, iz, ix);
      S6[i] = S6[i] + sminus(c, vz, i, nz) * ax;  // vzdx (add=.true.)
    }
}

// a12. Lame imaging condition, adjoint branch, one level,
// W/image-mod-rtm.f90:154-162: drho += Vx S1 + Vz S2; dlambda -= sxx S3 +
// szz S3; dmu -= 2 sxx S4 + 2 szz S5 + sxz S6. 'w' scales a decimated sum.
template <typename T>
static void image_adj_level(int64_t n2, const T* const* a, const T* S1, const T* S2, const T* S3,
                            const T* S4, const T* S5, const T* S6, T w, T* dm1, T* dm2, T* dm3) {
  for (int64_t i = 0; i < n2; ++i) {
    dm1[i] = dm1[i] + w * (a[0][i] * S1[i]) + w * (a[1][i] * S2[i]);
    dm2[i] = dm2[i] - w * (a[2][i] * S3[i]) - w * (a[3][i] * S3[i]);
    dm3[i] = dm3[i] - w * (a[2][i] * T(2.0) * S4[i]) - w * (a[3][i] * T(2.0) * S5[i]) -
             w * (a[4][i] * S6[i]);
  }
}

// Born-direction imaging (image_op2 with adj=.false.), W/image-mod-rtm.f90:171-179:
// scattering sources for one level from dm and S.
template <typename T>
static void image_fwd_level(int64_t n2, T* const* s, const T* S1, const T* S2, const T* S3,
                            const T* S4, const T* S5, const T* S6, const T* dm1, const T* dm2,
                            const T* dm3) {
  for (int64_t i = 0; i < n2; ++i) {
    s[0][i] = T(0) + (dm1[i] * S1[i]);
    s[1][i] = T(0) + (dm1[i] * S2[i]);
    s[2][i] = T(0) - (dm2[i] * S3[i]) - (dm3[i] * T(2.0) * S4[i]);
    s[3][i] = T(0) - (dm2[i] * S3[i]) - (dm3[i] * T(2.0) * S5[i]);
    s[4][i] = T(0) - (dm3[i] * S6[i]);
  }
}

// a6 + a8. Forward source propagation with sponge and explosive source,
// W/rtm-mod-adj.f90:178-260; receivers sampled at row zr (as the Born branch
// does at :361-375). Level 1 is zero; level 2 holds dt*source(1) at the source
// without K (:193-198); levels >= 3 follow the general update (the it=3 block
// :200-222 is the general update with level 1 = 0).
// Outputs (each optional): rec 5*(nt,nrx); hist_vx/hist_vz (nz,nx,nt);
// hist5: all five histories (nz,nx,nt) each.
template <typename T>
static void forward(const Params& p, const T* wavelet, const T* m1_x, const T* m1_z, const T* m2,
                    const T* m3, const T* m2m3, T* rec, T* hist_vx, T* hist_vz, T* const* hist5) {
  const int nz = p.nz, nx = p.nx, nt = p.nt;
  const int64_t n2 = (int64_t)nz * nx;
  std::vector<T> filt(std::max(p.abc + 5, 1));
  if (p.abc >= 0) abc_filter<T>(p.abc, filt.data());
  std::vector<T> buf(10 * n2, T(0));
  T* lev[2][5];
  for (int s = 0; s < 2; ++s)
    for (int f = 0; f < 5; ++f) lev[s][f] = buf.data() + (5 * s + f) * n2;
  const T dt = T(p.dt);
  auto record = [&](int it1, T* const* X) {
    if (rec)
      for (int f = 0; f < 5; ++f)
        for (int ix = 0; ix < p.nrx; ++ix)
          rec[(int64_t)f * nt * p.nrx + (it1 - 1) + (int64_t)nt * ix] = X[f][I2(nz, p.zr, p.rx0 + ix)];
    if (hist_vx) std::memcpy(hist_vx + n2 * (it1 - 1), X[0], sizeof(T) * n2);
    if (hist_vz) std::memcpy(hist_vz + n2 * (it1 - 1), X[1], sizeof(T) * n2);
    if (hist5)
      for (int f = 0; f < 5; ++f) std::memcpy(hist5[f] + n2 * (it1 - 1), X[f], sizeof(T) * n2);
  };
  // level 1 (slot 0) = 0
  record(1, lev[0]);
  if (nt < 2) return;
  // level 2 (slot 1): sources only, no K
  {
    const T amp = dt * wavelet[0];
    lev[1][2][I2(nz, p.zs, p.xs)] = lev[1][2][I2(nz, p.zs, p.xs)] + amp;
    lev[1][3][I2(nz, p.zs, p.xs)] = lev[1][3][I2(nz, p.zs, p.xs)] + amp;
  }
  record(2, lev[1]);
  for (int it = 3; it <= nt; ++it) {
    const int cs = (it - 1) & 1;  // slot holding level it-2 (odd levels slot 0)
    const T amp = dt * wavelet[it - 2];  // dt*source(it-1)
    fwd_level<T>(p, lev[cs], lev[cs ^ 1], m1_x, m1_z, m2, m3, m2m3, filt.data(), true, true, amp,
                 nullptr);
    record(it, lev[cs]);
  }
}

// Full single-shot RTM (rtm_op2 with adj=.true., W/rtm-mod-adj.f90:178-491):
// forward with a stored Vx/Vz history, then the adjoint of the receiver data
// (rtm_op3 + rtm_op3_c_, injection of all five components on row zr over the
// receiver columns) with the Lame imaging condition applied at every level
// (image_op2, W/image-mod-rtm.f90:146-162). The adjoint recursion is
//   A(it) = K'( inj(it) + A(it+2) + Op^T A(it+1) ),  A(nt+1)=A(nt+2)=0,
// K' = identity (adj_abc=0, the drop-in) or K (adj_abc=1, exact adjoint).
// With snap_every = j > 1, only levels it with it % j == 0 are imaged,
// weighted by j.
template <typename T>
static void rtm(const Params& p, const T* wavelet, const T* m1_x, const T* m1_z, const T* m2,
                const T* m3, const T* m2m3, const T* rec, T* dm1, T* dm2, T* dm3) {
  const int nz = p.nz, nx = p.nx, nt = p.nt;
  const int64_t n2 = (int64_t)nz * nx;
  std::vector<T> hvx(n2 * nt), hvz(n2 * nt);
  forward<T>(p, wavelet, m1_x, m1_z, m2, m3, m2m3, nullptr, hvx.data(), hvz.data(), nullptr);
  std::vector<T> filt(std::max(p.abc + 5, 1));
  if (p.abc >= 0) abc_filter<T>(p.abc, filt.data());
  std::vector<T> buf(10 * n2, T(0)), S(6 * n2);
  T* lev[2][5];
  for (int s = 0; s < 2; ++s)
    for (int f = 0; f < 5; ++f) lev[s][f] = buf.data() + (5 * s + f) * n2;
  for (int64_t i = 0; i < n2; ++i) dm1[i] = dm2[i] = dm3[i] = T(0);
  const int j = p.snap_every > 0 ? p.snap_every : 1;
  for (int it = nt; it >= 1; --it) {
    const int cs = it & 1;  // slot of level it (== slot of it+2)
    T* const* cur = lev[cs];
    // cur holds A(it+2) (zero for it >= nt-1); add the injection first
    // (reference order: inj + X(it+2) + terms).
    for (int f = 0; f < 5; ++f)
      for (int ix = 0; ix < p.nrx; ++ix) {
        const int64_t i = I2(nz, p.zr, p.rx0 + ix);
        cur[f][i] = rec[(int64_t)f * nt * p.nrx + (it - 1) + (int64_t)nt * ix] + cur[f][i];
      }
    if (it == nt) {
      if (p.adj_abc)
        for (int f = 0; f < 5; ++f) abc_K(nz, nx, p.abc, filt.data(), cur[f]);
    } else {
      adj_level<T>(p, cur, lev[cs ^ 1], m1_x, m1_z, m2, m3, m2m3, filt.data(), p.adj_abc != 0);
    }
    if (it % j == 0) {
      const T* vx1 = it < nt ? hvx.data() + n2 * it : nullptr;
      const T* vz1 = it < nt ? hvz.data() + n2 * it : nullptr;
      image_S<T>(nz, nx, T(p.dx), T(p.dt), hvx.data() + n2 * (it - 1), hvz.data() + n2 * (it - 1),
                 vx1, vz1, S.data(), S.data() + n2, S.data() + 2 * n2, S.data() + 3 * n2,
                 S.data() + 4 * n2, S.data() + 5 * n2);
      const T* a[5] = {cur[0], cur[1], cur[2], cur[3], cur[4]};
      image_adj_level<T>(n2, a, S.data(), S.data() + n2, S.data() + 2 * n2, S.data() + 3 * n2,
                         S.data() + 4 * n2, S.data() + 5 * n2, T(j), dm1, dm2, dm3);
    }
  }
}

// RTM with the R/Src adjoint (R/Src/rtm-mod-adj.f90:365-475): the sponge K
// is applied, in place, to the stored adjoint levels it+1 and it+2 before
// every step, so each level is damped once before it feeds the stencil and
// again before it is added two levels down; the imaging condition then reads
// the stored (damped) levels. Written literally over full (nz,nx,nt)
// histories in the reference's call order, with the scatter adjoints of
// W/rtmlib-mod.f90 (the operators R/Src calls):
//   it = nt:    A(nt) row zr = data(nt)                           (:380-386)
//   it = nt-1:  K A(nt); A(nt-1) row = data; A(nt-1) += Op^T A(nt) (:388-415)
//   it < nt-1:  K A(it+1); K A(it+2); A(it) row = data;
//               A(it) += A(it+2); A(it) += Op^T A(it+1)            (:418-462)
// followed by image_op2 over it = 1..nt (:472).
// The adjoint part: A holds 5 full histories, field f of level L at
// A + (f*nt + L-1)*n2, zero on entry.
template <typename T>
static void rsrc_adjoint(const Params& p, const T* m1_x, const T* m1_z, const T* m2, const T* m3,
                         const T* m2m3, const T* rec, T* A) {
  const int nz = p.nz, nx = p.nx, nt = p.nt;
  const int64_t n2 = (int64_t)nz * nx;
  std::vector<T> filt(std::max(p.abc + 5, 1));
  if (p.abc >= 0) abc_filter<T>(p.abc, filt.data());
  auto lev = [&](int f, int L) { return A + ((int64_t)f * nt + (L - 1)) * n2; };
  auto inject = [&](int f, int L) {
    for (int ix = 0; ix < p.nrx; ++ix)
      lev(f, L)[I2(nz, p.zr, p.rx0 + ix)] = rec[(int64_t)f * nt * p.nrx + (L - 1) + (int64_t)nt * ix];
  };
  auto damp = [&](int L) {
    for (int f = 0; f < 5; ++f) abc_K(nz, nx, p.abc, filt.data(), lev(f, L));
  };
  auto add2 = [&](int f, int L) {
    T* a = lev(f, L);
    const T* b = lev(f, L + 2);
    for (int64_t i = 0; i < n2; ++i) a[i] = a[i] + b[i];
  };
  // one level L <- data row (+ A(L+2)) + Op^T A(L+1), per field in the
  // reference's order: Vx (D, F, H), Vz (E, G, J), sxx (Ax), szz (Bz),
  // sxz (Bx, Az)
  auto level = [&](int L, bool plus2) {
    inject(0, L);
    if (plus2) add2(0, L);
    adj_op_scatter<T>(4, nz, nx, m2m3, lev(0, L), lev(2, L + 1));
    adj_op_scatter<T>(6, nz, nx, m2, lev(0, L), lev(3, L + 1));
    adj_op_scatter<T>(8, nz, nx, m3, lev(0, L), lev(4, L + 1));
    inject(1, L);
    if (plus2) add2(1, L);
    adj_op_scatter<T>(5, nz, nx, m2, lev(1, L), lev(2, L + 1));
    adj_op_scatter<T>(7, nz, nx, m2m3, lev(1, L), lev(3, L + 1));
    adj_op_scatter<T>(9, nz, nx, m3, lev(1, L), lev(4, L + 1));
    inject(2, L);
    if (plus2) add2(2, L);
    adj_op_scatter<T>(0, nz, nx, m1_x, lev(2, L), lev(0, L + 1));
    inject(3, L);
    if (plus2) add2(3, L);
    adj_op_scatter<T>(3, nz, nx, m1_z, lev(3, L), lev(1, L + 1));
    inject(4, L);
    if (plus2) add2(4, L);
    adj_op_scatter<T>(1, nz, nx, m1_x, lev(4, L), lev(0, L + 1));
    adj_op_scatter<T>(2, nz, nx, m1_z, lev(4, L), lev(1, L + 1));
  };
  for (int f = 0; f < 5; ++f) inject(f, nt);
  if (nt >= 2) {
    damp(nt);
    level(nt - 1, false);
  }
  for (int it = nt - 2; it >= 1; --it) {
    damp(it + 1);
    damp(it + 2);
    level(it, true);
  }
}

template <typename T>
static void rtm_rsrc(const Params& p, const T* wavelet, const T* m1_x, const T* m1_z, const T* m2,
                     const T* m3, const T* m2m3, const T* rec, T* dm1, T* dm2, T* dm3) {
  const int nz = p.nz, nx = p.nx, nt = p.nt;
  const int64_t n2 = (int64_t)nz * nx;
  std::vector<T> hvx(n2 * nt), hvz(n2 * nt);
  forward<T>(p, wavelet, m1_x, m1_z, m2, m3, m2m3, nullptr, hvx.data(), hvz.data(), nullptr);
  std::vector<T> A((size_t)5 * n2 * nt, T(0));
  rsrc_adjoint<T>(p, m1_x, m1_z, m2, m3, m2m3, rec, A.data());
  auto lev = [&](int f, int L) { return A.data() + ((int64_t)f * nt + (L - 1)) * n2; };
  for (int64_t i = 0; i < n2; ++i) dm1[i] = dm2[i] = dm3[i] = T(0);
  std::vector<T> S(6 * n2);
  for (int it = 1; it <= nt; ++it) {
    const T* vx1 = it < nt ? hvx.data() + n2 * it : nullptr;
    const T* vz1 = it < nt ? hvz.data() + n2 * it : nullptr;
    image_S<T>(nz, nx, T(p.dx), T(p.dt), hvx.data() + n2 * (it - 1), hvz.data() + n2 * (it - 1), vx1, vz1,
               S.data(), S.data() + n2, S.data() + 2 * n2, S.data() + 3 * n2, S.data() + 4 * n2,
               S.data() + 5 * n2);
    const T* a[5] = {lev(0, it), lev(1, it), lev(2, it), lev(3, it), lev(4, it)};
    image_adj_level<T>(n2, a, S.data(), S.data() + n2, S.data() + 2 * n2, S.data() + 3 * n2,
                       S.data() + 4 * n2, S.data() + 5 * n2, T(1), dm1, dm2, dm3);
  }
}

// Born modelling (rtm_op2 with adj=.false.), W/rtm-mod-adj.f90:273-375: the
// background forward gives S1..S6 at every level; image_op2 (forward branch)
// turns dm into scattering sources; the scattered wavefield is propagated
//   X(it) = K( s(it) + X(it-2) + Op Y(it-1) ),  X(0)=X(-1)=0,
// and sampled on row zr over the receiver columns.
template <typename T>
static void born(const Params& p, const T* wavelet, const T* m1_x, const T* m1_z, const T* m2,
                 const T* m3, const T* m2m3, const T* dm1, const T* dm2, const T* dm3, T* rec) {
  const int nz = p.nz, nx = p.nx, nt = p.nt;
  const int64_t n2 = (int64_t)nz * nx;
  std::vector<T> hvx(n2 * nt), hvz(n2 * nt);
  forward<T>(p, wavelet, m1_x, m1_z, m2, m3, m2m3, nullptr, hvx.data(), hvz.data(), nullptr);
  std::vector<T> filt(std::max(p.abc + 5, 1));
  if (p.abc >= 0) abc_filter<T>(p.abc, filt.data());
  std::vector<T> buf(10 * n2, T(0)), S(6 * n2), sb(5 * n2);
  T* lev[2][5];
  T* bs[5];
  for (int s = 0; s < 2; ++s)
    for (int f = 0; f < 5; ++f) lev[s][f] = buf.data() + (5 * s + f) * n2;
  for (int f = 0; f < 5; ++f) bs[f] = sb.data() + f * n2;
  for (int it = 1; it <= nt; ++it) {
    const T* vx1 = it < nt ? hvx.data() + n2 * it : nullptr;
    const T* vz1 = it < nt ? hvz.data() + n2 * it : nullptr;
    image_S<T>(nz, nx, T(p.dx), T(p.dt), hvx.data() + n2 * (it - 1), hvz.data() + n2 * (it - 1), vx1,
               vz1, S.data(), S.data() + n2, S.data() + 2 * n2, S.data() + 3 * n2,
               S.data() + 4 * n2, S.data() + 5 * n2);
    image_fwd_level<T>(n2, bs, S.data(), S.data() + n2, S.data() + 2 * n2, S.data() + 3 * n2,
                       S.data() + 4 * n2, S.data() + 5 * n2, dm1, dm2, dm3);
    const int cs = (it - 1) & 1;
    const T* bsc[5] = {bs[0], bs[1], bs[2], bs[3], bs[4]};
    fwd_level<T>(p, lev[cs], lev[cs ^ 1], m1_x, m1_z, m2, m3, m2m3, filt.data(), true, false, T(0),
                 bsc);
    for (int f = 0; f < 5; ++f)
      for (int ix = 0; ix < p.nrx; ++ix)
        rec[(int64_t)f * nt * p.nrx + (it - 1) + (int64_t)nt * ix] = lev[cs][f][I2(nz, p.zr, p.rx0 + ix)];
  }
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI for ctypes (tests/oracle loader). Suffix _f64 / _f32 = T.
#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

#define ORC_INSTANTIATE(T, SFX)                                                                    \
  ORC_EXPORT void orc_fd_coefs_##SFX(T* c) {                                                       \
    auto k = orc::coefs<T>();                                                                      \
    c[0] = k.c1; c[1] = k.c2; c[2] = k.c3; c[3] = k.c4; c[4] = k.c5;                              \
  }                                                                                                \
  ORC_EXPORT void orc_model_op_##SFX(int nz, int nx, T dx, T dt, const T* rho, const T* lam,      \
                                     const T* mu, T* m1x, T* m1z, T* m2, T* m3, T* m2m3) {          \
    orc::model_op<T>(nz, nx, dx, dt, rho, lam, mu, m1x, m1z, m2, m3, m2m3);                        \
  }                                                                                                \
  ORC_EXPORT void orc_ricker_##SFX(int nt, T dt, T fp, int nsource, T* out) {                      \
    orc::ricker<T>(nt, dt, fp, nsource, out);                                                      \
  }                                                                                                \
  ORC_EXPORT void orc_abc_filter_##SFX(int abc, T* f) { orc::abc_filter<T>(abc, f); }              \
  ORC_EXPORT void orc_abc_K_##SFX(int nz, int nx, int abc, T* data) {                              \
    std::vector<T> f(std::max(abc + 5, 1));                                                                     \
    orc::abc_filter<T>(abc, f.data());                                                             \
    orc::abc_K<T>(nz, nx, abc, f.data(), data);                                                    \
  }                                                                                                \
  ORC_EXPORT void orc_fwd_op_##SFX(int op, int nz, int nx, const T* coef, const T* model, T* data) {\
    orc::fwd_op<T>(op, nz, nx, coef, model, data, 1);                                              \
  }                                                                                                \
  ORC_EXPORT void orc_adj_op_scatter_##SFX(int op, int nz, int nx, const T* coef, T* model,        \
                                           const T* data) {                                        \
    orc::adj_op_scatter<T>(op, nz, nx, coef, model, data);                                         \
  }                                                                                                \
  ORC_EXPORT void orc_rtm_op3_c_gather_##SFX(int nt, int nz, int nx, T* vx, T* vz, T* xx, T* zz,   \
                                             T* xz, const T* m1x, const T* m1z, const T* m2,       \
                                             const T* m3, const T* m2m3, int nth) {                \
    orc::rtm_op3_c_gather<T>(nt, nz, nx, vx, vz, xx, zz, xz, m1x, m1z, m2, m3, m2m3, nth);         \
  }                                                                                                \
  ORC_EXPORT void orc_rtm_op3_c_scatter_##SFX(int nt, int nz, int nx, T* vx, T* vz, T* xx, T* zz,  \
                                              T* xz, const T* m1x, const T* m1z, const T* m2,      \
                                              const T* m3, const T* m2m3) {                        \
    orc::rtm_op3_c_scatter<T>(nt, nz, nx, vx, vz, xx, zz, xz, m1x, m1z, m2, m3, m2m3);             \
  }                                                                                                \
  ORC_EXPORT void orc_rtm_op3_prestep_##SFX(int nt, int nz, int nx, int zrec, T* vx, T* vz, T* xx, \
                                            T* zz, T* xz, const T* dvx, const T* dvz,              \
                                            const T* dxx, const T* dzz, const T* dxz,              \
                                            const T* m1x, const T* m1z, const T* m2, const T* m3,  \
                                            const T* m2m3) {                                       \
    T* h[5] = {vx, vz, xx, zz, xz};                                                                \
    const T* d[5] = {dvx, dvz, dxx, dzz, dxz};                                                     \
    orc::rtm_op3_prestep<T>(nt, nz, nx, zrec, h, d, m1x, m1z, m2, m3, m2m3);                       \
  }                                                                                                \
  ORC_EXPORT void orc_forward_##SFX(const orc::Params* p, const T* wav, const T* m1x,              \
                                    const T* m1z, const T* m2, const T* m3, const T* m2m3, T* rec, \
                                    T* hvx, T* hvz, T* h0, T* h1, T* h2, T* h3, T* h4) {           \
    T* h5[5] = {h0, h1, h2, h3, h4};                                                               \
    orc::forward<T>(*p, wav, m1x, m1z, m2, m3, m2m3, rec, hvx, hvz, h0 ? h5 : nullptr);            \
  }                                                                                                \
  ORC_EXPORT void orc_fwd_level_##SFX(const orc::Params* p, T* c0, T* c1, T* c2, T* c3, T* c4,     \
                                      const T* p0, const T* p1, const T* p2, const T* p3,          \
                                      const T* p4, const T* m1x, const T* m1z, const T* m2,        \
                                      const T* m3, const T* m2m3, int apply_k, int add_src,        \
                                      T srcamp) {                                                  \
    T* cur[5] = {c0, c1, c2, c3, c4};                                                              \
    const T* prev[5] = {p0, p1, p2, p3, p4};                                                       \
    std::vector<T> filt(std::max(p->abc + 5, 1));                                                  \
    orc::abc_filter<T>(p->abc, filt.data());                                                       \
    orc::fwd_level<T>(*p, cur, prev, m1x, m1z, m2, m3, m2m3, filt.data(), apply_k != 0,           \
                      add_src != 0, srcamp, nullptr);                                              \
  }                                                                                                \
  ORC_EXPORT void orc_image_S_##SFX(int nz, int nx, T dx, T dt, const T* vx, const T* vz,          \
                                    const T* vx1, const T* vz1, T* S1, T* S2, T* S3, T* S4, T* S5, \
                                    T* S6) {                                                       \
    orc::image_S<T>(nz, nx, dx, dt, vx, vz, vx1, vz1, S1, S2, S3, S4, S5, S6);                     \
  }                                                                                                \
  ORC_EXPORT void orc_image_adj_level_##SFX(int64_t n2, const T* vx, const T* vz, const T* xx,      \
                                            const T* zz, const T* xz, const T* S, T w, T* dm1,     \
                                            T* dm2, T* dm3) {                                      \
    const T* a[5] = {vx, vz, xx, zz, xz};                                                          \
    orc::image_adj_level<T>(n2, a, S, S + n2, S + 2 * n2, S + 3 * n2, S + 4 * n2, S + 5 * n2, w,   \
                            dm1, dm2, dm3);                                                        \
  }                                                                                                \
  ORC_EXPORT void orc_rtm_##SFX(const orc::Params* p, const T* wav, const T* m1x, const T* m1z,    \
                                const T* m2, const T* m3, const T* m2m3, const T* rec, T* dm1,     \
                                T* dm2, T* dm3) {                                                  \
    orc::rtm<T>(*p, wav, m1x, m1z, m2, m3, m2m3, rec, dm1, dm2, dm3);                              \
  }                                                                                                \
  ORC_EXPORT void orc_rsrc_adjoint_##SFX(const orc::Params* p, const T* m1x, const T* m1z,          \
                                         const T* m2, const T* m3, const T* m2m3, const T* rec,    \
                                         T* hist) {                                                \
    orc::rsrc_adjoint<T>(*p, m1x, m1z, m2, m3, m2m3, rec, hist);                                   \
  }                                                                                                \
  ORC_EXPORT void orc_rtm_rsrc_##SFX(const orc::Params* p, const T* wav, const T* m1x,              \
                                     const T* m1z, const T* m2, const T* m3, const T* m2m3,        \
                                     const T* rec, T* dm1, T* dm2, T* dm3) {                       \
    orc::rtm_rsrc<T>(*p, wav, m1x, m1z, m2, m3, m2m3, rec, dm1, dm2, dm3);                         \
  }                                                                                                \
  ORC_EXPORT void orc_born_##SFX(const orc::Params* p, const T* wav, const T* m1x, const T* m1z,   \
                                 const T* m2, const T* m3, const T* m2m3, const T* dm1,            \
                                 const T* dm2, const T* dm3, T* rec) {                             \
    orc::born<T>(*p, wav, m1x, m1z, m2, m3, m2m3, dm1, dm2, dm3, rec);                             \
  }

ORC_INSTANTIATE(double, f64)
ORC_INSTANTIATE(float, f32)
