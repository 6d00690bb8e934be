// ysx_device.hpp -- FP64 device numerics of the spectral path for gfx950.
//
// One lane integrates one radial system of one (l, w) pair.  The deck's
// spline table lives in LDS (field-major, 1-based knot index), everything
// else in registers.  The arithmetic follows the reference routines cited
// at each function; the GPU-specific parts are the data layout, the
// register-lean Cash-Karp step and the kernels that drive these functions
// (ysx_march_kernels.hpp, ysx_k_misc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ysx_common.hpp"

namespace ysx {

// ---------------------------------------------------------------- constants
// src/module_int.f90:80-81 renormalisation thresholds
constexpr double TINY = 1.0e-20;
constexpr double HUGE_ = 1.0 / 1.0e-20;
constexpr double TINY2 = TINY * TINY;
constexpr double HUGE2 = HUGE_ * HUGE_;

// ------------------------------------------------------------ complex FP64
struct cd {
  double x, y;
};
__device__ __host__ inline cd C(double x, double y = 0.0) { return {x, y}; }
__host__ __device__ inline cd operator+(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cd operator-(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cd operator-(cd a) { return {-a.x, -a.y}; }
__host__ __device__ inline cd operator*(cd a, cd b) {
  return {fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x)};
}
__host__ __device__ inline cd operator*(double s, cd a) { return {s * a.x, s * a.y}; }
__host__ __device__ inline cd operator*(cd a, double s) { return {s * a.x, s * a.y}; }
__host__ __device__ inline cd operator+(cd a, double s) { return {a.x + s, a.y}; }
__host__ __device__ inline cd operator+(double s, cd a) { return {a.x + s, a.y}; }
__host__ __device__ inline cd operator-(cd a, double s) { return {a.x - s, a.y}; }
__host__ __device__ inline cd operator-(double s, cd a) { return {s - a.x, -a.y}; }
__host__ __device__ inline cd cinv(cd b) {
  // scaled reciprocal: safe for |b| up to the double range
  double s = fmax(fabs(b.x), fabs(b.y));
  if (s == 0.0) return {1.0 / 0.0, 0.0};
  double bx = b.x / s, by = b.y / s;
  double d = 1.0 / (s * (bx * bx + by * by));
  return {bx * d, -by * d};
}
__host__ __device__ inline cd operator/(cd a, cd b) { return a * cinv(b); }
__host__ __device__ inline cd operator/(cd a, double s) { return {a.x / s, a.y / s}; }
__host__ __device__ inline cd operator/(double s, cd b) { return s * cinv(b); }
__host__ __device__ inline cd cmuli(cd a) { return {-a.y, a.x}; }  // i*a
__host__ __device__ inline double cabs2(cd a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline double cabs_(cd a) { return hypot(a.x, a.y); }
__host__ __device__ inline cd cfma(cd a, cd b, cd c) {  // a*b + c
  return {fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y))};
}
__host__ __device__ inline cd csqrt_(cd z) {
  // principal branch, Re >= 0
  if (z.x == 0.0 && z.y == 0.0) return {0.0, z.y};
  double t = sqrt(0.5 * (fabs(z.x) + hypot(z.x, z.y)));
  if (z.x >= 0.0) return {t, z.y / (2.0 * t)};
  return {fabs(z.y) / (2.0 * t), copysign(t, z.y)};
}
__host__ __device__ inline cd clog_(cd z) { return {log(hypot(z.x, z.y)), atan2(z.y, z.x)}; }

// -------------------------------------------------------------- model table
struct Tab {
  const double* t;  // field-major: t[f * nk + (i - 1)]
  int nk;
  __host__ __device__ inline double operator()(int f, int i) const { return t[f * nk + i - 1]; }
};

// Natural cubic spline evaluation on knot interval klo (splint_dis_r,
// src/module_spline.f90:262-283): weights shared by all fields.
struct SplW {
  double a, b, a3, b3, hh;
  int lo;
  bool deg;
};
__host__ __device__ inline SplW spl_w(const Tab& T, int klo, double x) {
  SplW w;
  w.lo = klo;
  double xl = T(F_R, klo), xh = T(F_R, klo + 1);
  double h = xh - xl;
  w.deg = (h == 0.0);
  double ih = w.deg ? 0.0 : 1.0 / h;
  w.a = (xh - x) * ih;
  w.b = (x - xl) * ih;
  w.a3 = w.a * w.a * w.a - w.a;
  w.b3 = w.b * w.b * w.b - w.b;
  w.hh = h * h * (1.0 / 6.0);
  return w;
}
__host__ __device__ inline double spl(const Tab& T, const SplW& w, int f) {
  double ya = T(f, w.lo);
  if (w.deg) return ya;
  double yb = T(f, w.lo + 1);
  double y2a = T(f + 1, w.lo), y2b = T(f + 1, w.lo + 1);
  return fma(w.a, ya, w.b * yb) + fma(w.a3, y2a, w.b3 * y2b) * w.hh;
}

// ------------------------------------------------------ per-pair constants
// set_l / set_omega, src/module_int.f90:2072-2102
struct PS {
  int ll;
  double lld, llp1, zeta, zeta2, zeta2m2, szeta2m2, nu;
  cd omega, omega2, omegai2, lno;
  double rew;
  double pibigg;
};
__host__ __device__ inline void set_l(PS& s, int l) {
  s.ll = l;
  s.lld = (double)l;
  s.llp1 = (double)(l + 1);
  s.zeta2 = (double)(float)(l * (l + 1));
  s.zeta = sqrt(s.zeta2);
  s.zeta2m2 = s.zeta2 - 2.0;
  s.szeta2m2 = s.zeta2m2 > 0.0 ? sqrt(s.zeta2m2) : 0.0;
  s.nu = sqrt((double)(float)(2 * l + 1) / FOURPI_D);
}
__host__ __device__ inline void set_omega(PS& s, cd w, double tref, double ats) {
  s.omega = w;
  s.omega2 = w * w;
  cd oi = 1.0 / w;
  s.omegai2 = oi * oi;
  cd z = cmuli(w) / tref;  // i*w/tref
  cd lg = clog_(z);
  s.lno = C(ats * (2.0 * lg.x / PI_D), ats * (2.0 * lg.y / PI_D));
  s.rew = w.x;
}

// ----------------------------------------------------- coefficient matrices
// sph_solid_amat, src/module_int.f90:500-560
struct SolidA {
  cd t11, t12, k11, k22, s11, s12, s22;
  double t21, t22, t31, t33, k33, s13, s23;
};
__host__ __device__ inline void solid_amat(const Tab& T, const PS& s, int il, double rr, SolidA& A) {
  SplW w = spl_w(T, il, rr);
  double rrho = spl(T, w, F_RHO), racon = spl(T, w, F_ACON), rccon = spl(T, w, F_CCON),
         rfcon = spl(T, w, F_FCON), rncon = spl(T, w, F_NCON), rlcon = spl(T, w, F_LCON),
         gg = spl(T, w, F_GRAV), rqmu = spl(T, w, F_QMU), rxlam = spl(T, w, F_XLAM),
         rxa2 = spl(T, w, F_XA2);
  cd qaa = 1.0 + rxa2 * s.lno, qff = 1.0 + rxlam * s.lno, qll = 1.0 + rqmu * s.lno;
  cd cacon = racon * qaa, cfcon = rfcon * qff, cncon = rncon * qll;
  cd cccon = cinv(rccon * qaa);
  cd fc = cfcon * cccon;
  cd gamma = cacon - cncon - cfcon * fc;
  double ri = 1.0 / rr, ri2 = ri * ri;
  double rgr = rrho * gg * rr;
  A.t11 = (1.0 - 2.0 * fc) * ri;
  A.t12 = (s.zeta * ri) * fc;
  A.t21 = -s.zeta * ri;
  A.t22 = 2.0 * ri;
  A.t31 = -4.0 * s.pibigg * rrho;
  A.t33 = -s.lld * ri;
  A.k11 = cccon;
  A.k22 = cinv(rlcon * qll);
  A.k33 = 4.0 * s.pibigg;
  cd mw2r = -rrho * s.omega2;
  A.s11 = mw2r + (4.0 * ri2) * (gamma - rgr);
  A.s12 = (s.zeta * ri2) * (rgr - 2.0 * gamma);
  A.s13 = -rrho * s.llp1 * ri;
  A.s22 = mw2r + ri2 * (s.zeta2 * (gamma + cncon) - 2.0 * cncon);
  A.s23 = rrho * s.zeta * ri;
}

// sph_minors_solid_derivs, src/module_int.f90:818-871 (14 minors)
struct SolidMinorRHS {
  static constexpr int N = 14;
  __host__ __device__ inline void operator()(const Tab& T, const PS& s, int il, double rr, const cd* m,
                                    cd* d) const {
    SolidA A;
    solid_amat(T, s, il, rr, A);
    double t21 = A.t21, t31 = A.t31;
    cd a123 = A.t11 + (A.t22 + A.t33), a126 = A.t11 + (A.t22 - A.t33),
       a135 = A.t11 - (A.t22 - A.t33), a234 = (A.t22 + A.t33) - A.t11;
    double k33 = A.k33;
    d[0] = k33 * m[3] - A.k22 * m[5] + A.k11 * m[9] + m[0] * a123;
    d[1] = m[0] * A.s13 - A.k22 * m[6] - m[2] * t21 + m[1] * A.t22 - m[3] * t31;
    d[2] = m[0] * A.s23 - A.k11 * m[10] + m[2] * A.t11 - m[1] * A.t12;
    d[3] = A.k22 * m[8] - A.k11 * m[11] + m[3] * a126;
    d[4] = m[9] * A.t12 - k33 * m[7] - m[0] * A.s12 - m[5] * t21 + m[2] * t31 + m[4] * A.t33;
    d[5] = m[5] * a135 - k33 * m[8] - A.k11 * m[12] - m[0] * A.s22 - 2.0 * (m[4] * A.t12);
    d[6] = m[2] * A.s12 + m[5] * A.s13 - m[1] * A.s22 - m[4] * A.s23 + m[10] * A.t12 -
           m[6] * A.t22 + m[8] * t31;
    d[7] = m[3] * A.s12 - m[2] * A.s13 - m[1] * A.s23 + m[11] * A.t12 - m[8] * t21 -
           m[7] * A.t33;
    d[8] = A.k11 * m[13] + m[3] * A.s22 - 2.0 * (m[2] * A.s23) - 2.0 * (m[7] * A.t12) -
           m[8] * a234;
    d[9] = A.k22 * m[12] - k33 * m[11] + m[0] * A.s11 + 2.0 * (m[4] * t21) -
           2.0 * (m[1] * t31) + m[9] * a234;
    d[10] = m[1] * A.s12 - m[2] * A.s11 - m[4] * A.s13 - m[9] * A.s23 - m[10] * A.t11 +
            m[6] * t21 - m[7] * t31;
    d[11] = 2.0 * (m[1] * A.s13) - A.k22 * m[13] - m[3] * A.s11 + 2.0 * (m[7] * t21) -
            m[11] * a135;
    d[12] = k33 * m[13] - m[5] * A.s11 + 2.0 * (m[4] * A.s12) + m[9] * A.s22 +
            2.0 * (m[6] * t31) - m[12] * a126;
    d[13] = m[8] * A.s11 - 2.0 * (m[7] * A.s12) + 2.0 * (m[6] * A.s13) - m[11] * A.s22 +
            2.0 * (m[10] * A.s23) - m[13] * a123;
  }
};

// sph_bvec_solid_derivs, src/module_int.f90:1016-1057 (14 b-vector terms)
struct BvecRHS {
  static constexpr int N = 14;
  __host__ __device__ inline void operator()(const Tab& T, const PS& s, int il, double rr, const cd* b,
                                    cd* d) const {
    SolidA A;
    solid_amat(T, s, il, rr, A);
    double t21 = A.t21, t31 = A.t31;
    double k33 = A.k33;
    cd t1122 = A.t11 + A.t22, t1133 = A.t11 + A.t33;
    double t2233 = A.t22 + A.t33;
    d[0] = A.s11 * b[6] + A.s12 * (b[7] - b[2]) + A.s13 * b[8] - A.s22 * b[3] -
           A.s23 * b[4] - t1122 * b[0] + t31 * b[5];
    d[1] = A.s11 * b[9] + A.s12 * b[10] - A.s13 * (2.0 * b[2] + b[7]) - A.s23 * b[3] -
           t21 * b[5] - t1133 * b[1];
    d[2] = A.s12 * b[11] + A.s13 * b[12] + A.t12 * b[3] - t21 * b[6] - t31 * b[9];
    d[3] = A.s13 * b[13] - A.k22 * b[0] - A.s11 * b[11] + t21 * (b[2] - b[7]) +
           (A.t22 - A.t11) * b[3] - t31 * b[10];
    d[4] = t31 * (2.0 * b[2] + b[7]) - k33 * b[1] - A.s11 * b[12] - A.s12 * b[13] -
           t21 * b[8] + (A.t33 - A.t11) * b[4];
    d[5] = A.s12 * b[9] - A.s13 * b[6] + A.s22 * b[10] - A.s23 * (b[2] + 2.0 * b[7]) -
           A.t12 * b[1] - t2233 * b[5];
    d[6] = A.k11 * b[0] + A.s22 * b[11] + A.s23 * b[12] - A.t12 * (b[2] - b[7]) +
           (A.t11 - A.t22) * b[6];
    d[7] = A.s23 * b[13] - A.s12 * b[11] - A.t12 * b[3] + t21 * b[6];
    d[8] = t31 * b[6] - k33 * b[5] - A.s12 * b[12] - A.s22 * b[13] - A.t12 * b[4] +
           (A.t33 - A.t22) * b[8];
    d[9] = A.k11 * b[1] + A.s23 * b[11] + A.t12 * b[10] + (A.t11 - A.t33) * b[9];
    d[10] = A.k22 * b[5] - A.s13 * b[11] + t21 * b[9] + (A.t22 - A.t33) * b[10];
    d[11] = A.k22 * b[6] - A.k11 * b[3] + t1122 * b[11];
    d[12] = k33 * b[9] - A.k11 * b[4] + A.t12 * b[13] + t1133 * b[12];
    d[13] = k33 * b[10] - A.k22 * b[8] + t21 * b[12] - t31 * b[11] + t2233 * b[13];
  }
};

// sph_fluid_amat + sph_minors_fluid_derivs, src/module_int.f90:676-715, 941-967
struct FluidMinorRHS {
  static constexpr int N = 5;
  __host__ __device__ inline void operator()(const Tab& T, const PS& s, int il, double rr, const cd* m,
                                    cd* d) const {
    SplW w = spl_w(T, il, rr);
    double rrho = spl(T, w, F_RHO), rccon = spl(T, w, F_CCON), gg = spl(T, w, F_GRAV),
           rxa2 = spl(T, w, F_XA2);
    cd cccon = cinv(rccon * (1.0 + rxa2 * s.lno));
    double ri = 1.0 / rr, ri2 = ri * ri;
    cd zoi = s.zeta2 * s.omegai2;        // zeta2/w^2
    cd g = (gg * ri) * zoi;              // gg*zeta2*ri/w^2
    cd t11 = (g - 1.0) * ri;
    cd t12 = ri2 * zoi;
    double t21 = -4.0 * s.pibigg * rrho;
    double t22 = -s.lld * ri;
    cd k11 = cccon - (ri2 / rrho) * zoi;
    double k22 = 4.0 * s.pibigg;
    cd s11 = (rrho * gg * ri) * (g - 4.0) - rrho * s.omega2;
    cd s12 = (rrho * ri) * (g - s.llp1);
    cd s22 = (rrho * ri2) * zoi;
    d[0] = k22 * m[2] - k11 * m[3] + m[0] * (t11 + t22);
    d[1] = m[0] * s12 + m[3] * t12 - t21 * m[2];
    d[2] = k11 * m[4] + m[0] * s22 - 2.0 * (m[1] * t12) + m[2] * (t11 - t22);
    d[3] = 2.0 * t21 * m[1] - k22 * m[4] - m[0] * s11 + m[3] * (t22 - t11);
    d[4] = m[2] * s11 - 2.0 * (m[1] * s12) - m[3] * s22 - m[4] * (t11 + t22);
  }
};

// tor_amat + tor_derivs, src/module_int.f90:134-184
struct TorRHS {
  static constexpr int N = 2;
  __host__ __device__ inline void operator()(const Tab& T, const PS& s, int il, double rr, const cd* y,
                                    cd* d) const {
    SplW w = spl_w(T, il, rr);
    double rrho = spl(T, w, F_RHO), rlcon = spl(T, w, F_LCON), rncon = spl(T, w, F_NCON),
           rqmu = spl(T, w, F_QMU);
    cd qll = 1.0 + rqmu * s.lno;
    double ri = 1.0 / rr;
    double t = 2.0 * ri;
    cd k = cinv(rlcon * qll);
    cd sv = (ri * ri * s.zeta2m2 * rncon) * qll - rrho * s.omega2;
    d[0] = t * y[0] + k * y[1];
    d[1] = sv * y[0] - t * y[1];
  }
};

// rad_amat + rad_derivs, src/module_int.f90:261-372
// coef: the matrix at one radius (independent of y); apply: its product
// with y.  The radial solves share the six stage evaluations of a step
// between the lanes of a group (rk_step_grp).
struct RadRHS {
  static constexpr int N = 2;
  struct Coef {
    cd t, k, sv;
  };
  __host__ __device__ inline Coef coef(const Tab& T, const PS& s, int il, double rr) const {
    SplW w = spl_w(T, il, rr);
    double rrho = spl(T, w, F_RHO), racon = spl(T, w, F_ACON), rccon = spl(T, w, F_CCON),
           rfcon = spl(T, w, F_FCON), rncon = spl(T, w, F_NCON), gg = spl(T, w, F_GRAV),
           rqmu = spl(T, w, F_QMU), rxlam = spl(T, w, F_XLAM), rxa2 = spl(T, w, F_XA2);
    cd qaa = 1.0 + rxa2 * s.lno, qff = 1.0 + rxlam * s.lno, qll = 1.0 + rqmu * s.lno;
    cd cacon = racon * qaa, cfcon = rfcon * qff, cncon = rncon * qll;
    cd cccon = cinv(rccon * qaa);
    cd fc = cfcon * cccon;
    cd gamma = cacon - cncon - cfcon * fc;
    double ri = 1.0 / rr;
    cd t = (1.0 - 2.0 * fc) * ri;
    cd sv = (4.0 * ri * ri) * (gamma - rrho * gg * rr) - rrho * s.omega2;
    return Coef{t, cccon, sv};
  }
  __host__ __device__ static inline void apply(const Coef& c, const cd* y, cd* d) {
    d[0] = c.t * y[0] + c.k * y[1];
    d[1] = c.sv * y[0] - c.t * y[1];
  }
  __host__ __device__ inline void operator()(const Tab& T, const PS& s, int il, double rr, const cd* y,
                                    cd* d) const {
    apply(coef(T, s, il, rr), y, d);
  }
};

// Lanes of one integration: the default is one lane.  RadGroup: lane q of
// a group of G (1 or 8) lanes that integrate the same solution; sh is the
// group's LDS slot of six stage coefficient sets.
struct NoGroup {};
struct RadGroup {
  int q, G;
  RadRHS::Coef* sh;
};

// ----------------------------------------------------------- Cash-Karp step
// rkck, src/module_int.f90:2016-2064.  Stage vectors are folded into
// partial sums as soon as no later stage needs them separately, so at most
// six N-vectors are live at once.
template <class F>
__host__ __device__ inline void rk_step(const Tab& T, const PS& s, int il, double x, double h, cd* y,
                               const F& f) {
  constexpr int N = F::N;
  constexpr double a2 = 0.2, a3 = 0.3, a4 = 0.6, a5 = 1.0, a6 = 0.875, b21 = 0.2,
                   b31 = 3.0 / 40.0, b32 = 9.0 / 40.0, b41 = 0.3, b42 = -0.9, b43 = 1.2,
                   b51 = -11.0 / 54.0, b52 = 2.5, b53 = -70.0 / 27.0, b54 = 35.0 / 27.0,
                   b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0, b63 = 575.0 / 13824.0,
                   b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0, c1 = 37.0 / 378.0,
                   c3 = 250.0 / 621.0, c4 = 125.0 / 594.0, c6 = 512.0 / 1771.0;
  cd k1[N], k2[N], k3[N], t[N];
  f(T, s, il, x, y, k1);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + (b21 * h) * k1[i];
  f(T, s, il, x + a2 * h, t, k2);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + h * (b31 * k1[i] + b32 * k2[i]);
  f(T, s, il, x + a3 * h, t, k3);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + h * (b41 * k1[i] + b42 * k2[i] + b43 * k3[i]);
  {
    cd k4[N];
    f(T, s, il, x + a4 * h, t, k4);
#pragma unroll
    for (int i = 0; i < N; i++) {
      cd p5 = b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i];
      cd p6 = b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i];
      cd po = c1 * k1[i] + c3 * k3[i] + c4 * k4[i];
      t[i] = y[i] + h * p5;
      k2[i] = p6;
      k1[i] = po;
    }
  }
  f(T, s, il, x + a5 * h, t, k3);  // k5
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + h * (k2[i] + b65 * k3[i]);
  f(T, s, il, x + a6 * h, t, k3);  // k6
#pragma unroll
  for (int i = 0; i < N; i++) y[i] = y[i] + h * (k1[i] + c6 * k3[i]);
}

// rk_step for a lane group of G = 8: lane q < 6 evaluates the coefficients
// of stage q + 1 into the group's LDS slot, then every lane runs the stage
// updates (the same arithmetic as rk_step, so every lane holds the same y).
// LDS accesses of one wave take effect in program order, so the slot needs
// no barrier; groups of a wave may be at different points of their solves.
// G = 1 is the plain rk_step.
template <class F>
__host__ __device__ inline void rk_step_grp(const Tab& T, const PS& s, int il, double x, double h, cd* y,
                                   const F& f, const RadGroup& g) {
  constexpr int N = F::N;
  constexpr double a2 = 0.2, a3 = 0.3, a4 = 0.6, a5 = 1.0, a6 = 0.875, b21 = 0.2,
                   b31 = 3.0 / 40.0, b32 = 9.0 / 40.0, b41 = 0.3, b42 = -0.9, b43 = 1.2,
                   b51 = -11.0 / 54.0, b52 = 2.5, b53 = -70.0 / 27.0, b54 = 35.0 / 27.0,
                   b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0, b63 = 575.0 / 13824.0,
                   b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0, c1 = 37.0 / 378.0,
                   c3 = 250.0 / 621.0, c4 = 125.0 / 594.0, c6 = 512.0 / 1771.0;
  if (g.G == 1) {
    rk_step(T, s, il, x, h, y, f);
    return;
  }
  if (g.q < 6) {
    const int q = g.q;
    double aq = (q == 1) ? a2 : (q == 2) ? a3 : (q == 3) ? a4 : (q == 4) ? a5 : a6;
    g.sh[q] = f.coef(T, s, il, (q == 0) ? x : x + aq * h);
  }
  typename F::Coef c[6];
#pragma unroll
  for (int k = 0; k < 6; k++) c[k] = g.sh[k];
  cd k1[N], k2[N], k3[N], t[N];
  F::apply(c[0], y, k1);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + (b21 * h) * k1[i];
  F::apply(c[1], t, k2);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + h * (b31 * k1[i] + b32 * k2[i]);
  F::apply(c[2], t, k3);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + h * (b41 * k1[i] + b42 * k2[i] + b43 * k3[i]);
  {
    cd k4[N];
    F::apply(c[3], t, k4);
#pragma unroll
    for (int i = 0; i < N; i++) {
      cd p5 = b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i];
      cd p6 = b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i];
      cd po = c1 * k1[i] + c3 * k3[i] + c4 * k4[i];
      t[i] = y[i] + h * p5;
      k2[i] = p6;
      k1[i] = po;
    }
  }
  F::apply(c[4], t, k3);  // k5
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + h * (k2[i] + b65 * k3[i]);
  F::apply(c[5], t, k3);  // k6
#pragma unroll
  for (int i = 0; i < N; i++) y[i] = y[i] + h * (k1[i] + c6 * k3[i]);
}

// renormalisation of yint, src/module_int.f90:1910-1919
template <int N>
__host__ __device__ inline void renorm(cd* y, int& scale) {
  bool small = true, big = false;
#pragma unroll
  for (int i = 0; i < N; i++) small = small && (cabs2(y[i]) < TINY2);
  if (small) {
    scale++;
#pragma unroll
    for (int i = 0; i < N; i++) y[i] = y[i] * HUGE_;
  }
#pragma unroll
  for (int i = 0; i < N; i++) big = big || (cabs2(y[i]) > HUGE2);
  if (big) {
    scale--;
#pragma unroll
    for (int i = 0; i < N; i++) y[i] = y[i] * TINY;
  }
}

// steps, src/module_int.f90:2121-2138 (nw = 4)
__host__ __device__ inline double step_len(const Tab& T, const PS& s, int i, int nknot) {
  double a = T(F_ALPHA, i);
  double tmp = (i < nknot) ? fmax(a, T(F_ALPHA, i + 1)) : a;
  return 1.0 / (4.0 * s.rew / tmp);
}

// a per-step callback of yint that does nothing
struct NoStepHook {
  __host__ __device__ inline void operator()(int, double, double, const cd*) const {}
};

// yint, src/module_int.f90:1847-1968: fixed equal steps per knot interval;
// sh(interval, radius, step, y) is called before each step
template <class F, class G = NoGroup, class SH = NoStepHook>
__host__ __device__ void yint(const Tab& T, const PS& s, int nknot, int i1, int i2, double r1, double r2,
                     cd* y, const F& f, int* scale, int& nsteps, const G& grp = G(),
                     const SH& sh = SH()) {
  if (r2 > r1) {
    for (int i = i1; i <= i2; i++) {
      double rr1 = (i == i1) ? r1 : T(F_R, i);
      double rr2 = (i == i2) ? r2 : T(F_R, i + 1);
      if (rr1 == rr2) continue;
      double drr = step_len(T, s, i, nknot);
      int n = 2 + (int)floor((rr2 - rr1) / drr);
      drr = (rr2 - rr1) / (double)(n - 1);
      nsteps += n - 1;
      for (int j = 1; j <= n - 1; j++) {
        double rr = rr1 + (double)(j - 1) * drr;
        sh(i, rr, drr, y);
        if constexpr (std::is_same<G, RadGroup>::value)
          rk_step_grp(T, s, i, rr, drr, y, f, grp);
        else
          rk_step(T, s, i, rr, drr, y, f);
        if (scale) renorm<F::N>(y, *scale);
      }
    }
  } else if (r2 < r1) {
    for (int i = i1; i >= i2; i--) {
      double rr1 = (i == i1) ? r1 : T(F_R, i + 1);
      double rr2 = (i == i2) ? r2 : T(F_R, i);
      if (rr2 == rr1) continue;
      double drr = step_len(T, s, i, nknot);
      int n = 2 + (int)floor(fabs(rr2 - rr1) / drr);
      drr = (rr2 - rr1) / (double)(n - 1);
      nsteps += n - 1;
      for (int j = 1; j <= n - 1; j++) {
        double rr = rr1 + (double)(j - 1) * drr;
        sh(i, rr, drr, y);
        if constexpr (std::is_same<G, RadGroup>::value)
          rk_step_grp(T, s, i, rr, drr, y, f, grp);
        else
          rk_step(T, s, i, rr, drr, y, f);
        if (scale) renorm<F::N>(y, *scale);
      }
    }
  }
}

// ------------------------------------------------------------ start levels
// start_level, src/module_start.f90:80-178; wt = 1 uses beta, 2 alpha
__host__ __device__ inline int start_level(const Tab& T, int l, double w, int i1, int i2, int wt) {
  const int fv = (wt == 1) ? F_BETA : F_ALPHA;
  int isign = (i2 >= i1) ? 1 : -1;
  double pp = (double)(float)(l * (l + 1)) / (w * w);
  int j = i2;
  for (int i = i1; isign > 0 ? i <= i2 : i >= i2; i += isign) {
    double v = T(fv, i), r = T(F_R, i);
    double q1 = 1.0 / (v * v) - pp / (r * r);
    if (isign == 1) {
      if (q1 > 0.0) { j = max(i1, i - 1); break; }
    } else {
      if (q1 < 0.0) { j = i; break; }
    }
  }
  int is = j;
  int lo = min(i1, i2);
  if (is == lo) return is;
  int j1 = is;
  is = lo;
  double sum = 0.0;
  for (int i = j1; i >= lo + 1; i--) {
    double v1 = T(fv, i), v2 = T(fv, i - 1), r1 = T(F_R, i), r2 = T(F_R, i - 1);
    double q1 = 1.0 / (v1 * v1) - pp / (r1 * r1);
    double q2 = 1.0 / (v2 * v2) - pp / (r2 * r2);
    if (q1 > 0.0 || q2 > 0.0) {
      sum = 0.0;
    } else {
      sum = sum + 0.5 * w * (sqrt(-q1) + sqrt(-q2)) * (r1 - r2);
    }
    if (sum > 11.0) { is = i - 1; break; }
  }
  return is;
}

struct Layers {
  int nknot, nic, noc, nsl, fic, ocean;
};

// tor_start_level / sph_start_level (up = 1), src/module_start.f90:14-75
__host__ __device__ inline int tor_start_level(const Tab& T, const Layers& L, int l, double w, int it) {
  return start_level(T, l, w, L.noc + 1, it, 1);
}
__host__ __device__ inline int sph_start_level(const Tab& T, const Layers& L, int l, double w, int it,
                                      double fre_norm) {
  double wdim = w * fre_norm;
  int is;
  if (wdim <= -2.00e-3 + 1.28e-3 * l) {
    is = start_level(T, l, w, L.noc + 1, it, 1);
    if (is == L.nsl) is--;
    if (is > L.noc + 1) return is;
  }
  if (wdim <= -2.00e-3 + 2.25e-3 * l) {
    is = start_level(T, l, w, L.nic + 1, L.noc, 2);
    if (is == L.noc) is--;
    if (is > L.nic + 1) return is;
  }
  is = start_level(T, l, w, 2, L.nic, L.fic ? 2 : 1);
  if (is == L.nic) is--;
  return is;
}

// ------------------------------------------------------ starting solutions
// bfs + bfs_dum, src/module_function.f90:497-581
__host__ __device__ inline cd bfs_ratio(int l, cd xsq) {
  const double eps = 1.e-9;
  cd fp;
  if (xsq.x > 0.0) {
    cd x = csqrt_(xsq);
    int lp1 = l + 1;
    cd rx = 2.0 / x;
    double nu = lp1 - 0.5;
    cd rj = nu * rx;
    rx = -rx;
    cd denom = (nu + 1.0) * rx;
    cd numer = denom + 1.0 / rj;
    rj = rj * numer / denom;
    int nm1 = 1;
    for (int it = 0; it < 100000; it++) {
      nm1++;
      rx = -rx;
      cd a3 = (nu + (double)nm1) * rx;
      denom = a3 + 1.0 / denom;
      numer = a3 + 1.0 / numer;
      cd ratio = numer / denom;
      rj = rj * ratio;
      if (fabs(cabs_(ratio) - 1.0) <= eps) break;
    }
    fp = rj * x - (double)lp1;
  } else {
    cd f = C(1.0), a = C(1.0);
    double b = 2.0 * l + 1.0, c = 2.0, dd = l + 2.0;
    fp = C((double)l);
    for (int it = 0; it < 100000; it++) {
      a = -(a * xsq) / (c * (b + c));
      f = f + a;
      fp = fp + a * dd;
      if (cabs_(a * dd) < eps) break;
      c += 2.0;
      dd += 2.0;
    }
    fp = fp / f;
  }
  double xr = sqrt(cabs_(xsq));
  return ((double)l - fp) / xr;
}

// sph_solid_start, src/module_int.f90:1192-1270; y[col][row]
__host__ __device__ inline void solid_start(const Tab& T, const PS& s, int is, double rr, cd y[3][6]) {
  SplW w = spl_w(T, is, rr);
  double rrho = spl(T, w, F_RHO), rmu = spl(T, w, F_MU), rqmu = spl(T, w, F_QMU),
         rkappa = spl(T, w, F_KAPPA), rqkappa = spl(T, w, F_QKAPPA);
  double pib = s.pibigg, ll = s.lld, llp1 = s.llp1, zeta = s.zeta, zeta2 = s.zeta2;
  cd cmu = rmu * (1.0 + rqmu * s.lno);
  cd ckappa = rkappa * (1.0 + rqkappa * s.lno);
  cd calpha = (ckappa + 1.3333333333 * cmu) / rrho;
  cd cbeta = cmu / rrho;
  cd tmp1 = s.omega2 / cbeta;
  cd tmp2 = (s.omega2 + 5.3333333 * pib * rrho) / calpha;
  double t3 = 2.666666666 * pib * zeta * rrho;
  cd tmp3 = (t3 * t3) / (calpha * cbeta);
  double sign = 1.0;
  for (int i = 0; i < 2; i++) {
    cd d = tmp1 - tmp2;
    cd g = d * d + tmp3;
    g = 0.5 * (tmp1 + tmp2) + (0.5 * sign) * csqrt_(g);
    cd g2 = g;
    g = csqrt_(g);
    cd zt = (0.75 * cbeta) * (g2 - tmp1) / (pib * rrho);
    cd xi = zt - llp1;
    cd ratio = bfs_ratio(s.ll, g2 * (rr * rr));
    cd gr = (g * rr) * ratio;  // gamma*rr*ratio
    y[i][0] = ll * xi - zt * gr;
    y[i][1] = zeta * (xi + gr);
    y[i][2] = (-4.0 * pib * rrho * rr) * zt;
    y[i][3] = (-rrho * rr) * (calpha * zt * g2) + ((2.0 * ll * (ll - 1.0) / rr) * cmu) * xi +
              (2.0 * cmu) * (2.0 * zt + zeta2) * (g * ratio);
    cd y5 = rr * g2 + (2.0 * (ll - 1.0) / rr) * xi - 2.0 * (zt + 1.0) * (g * ratio);
    y[i][4] = (zeta * cmu) * y5;
    y[i][5] = -rrho * (zeta2 + llp1 * zt);
    sign = -sign;
  }
  y[2][0] = C(ll);
  y[2][1] = C(zeta);
  y[2][2] = (s.omega2 - 4.0 * pib * rrho * ll / 3.0) * rr;
  y[2][3] = (2.0 * ll * (ll - 1.0) / rr) * cmu;
  y[2][4] = (2.0 * zeta * (ll - 1.0) / rr) * cmu;
  y[2][5] = (0.25 / pib) * ((2.0 * ll + 1.0) * s.omega2 - 8.0 * pib * ll * (ll - 1.0) * rrho / 3.0);
}

// sph_fluid_start, src/module_int.f90:1346-1405; y[col][row]
__host__ __device__ inline void fluid_start(const Tab& T, const PS& s, int is, double rr, cd y[2][4]) {
  SplW w = spl_w(T, is, rr);
  double rrho = spl(T, w, F_RHO), rkappa = spl(T, w, F_KAPPA), rqkappa = spl(T, w, F_QKAPPA);
  double pib = s.pibigg, ll = s.lld, llp1 = s.llp1, zeta2 = s.zeta2;
  cd ckappa = rkappa * (1.0 + rqkappa * s.lno);
  cd calpha = ckappa / rrho;
  double gamg = 4.0 * pib * rrho / 3.0;
  cd g = (s.omega2 + 4.0 * gamg - (zeta2 * gamg * gamg) * s.omegai2) / calpha;
  cd g2 = g;
  g = csqrt_(g);
  cd zt = -s.omega2 / gamg;
  cd xi = zt - llp1;
  cd ratio = bfs_ratio(s.ll, g2 * (rr * rr));
  y[0][0] = ll * xi - zt * ((g * rr) * ratio);
  y[0][1] = (-4.0 * pib * rrho * rr) * zt;
  y[0][2] = (-rrho * rr) * (calpha * zt * g2);
  y[0][3] = -rrho * (zeta2 + llp1 * zt);
  y[1][0] = C(ll);
  y[1][1] = (s.omega2 - 4.0 * pib * rrho * ll / 3.0) * rr;
  y[1][2] = C(0.0);
  y[1][3] = (0.25 / pib) * ((2.0 * ll + 1.0) * s.omega2 - 8.0 * pib * ll * (ll - 1.0) * rrho / 3.0);
}

__host__ __device__ inline cd det2(cd a1, cd a2, cd b1, cd b2) { return a1 * b2 - a2 * b1; }

// sph_minor_solid_start, src/module_int.f90:1451-1485 (14 3x3 minors)
__host__ __device__ inline void minor_solid_start(const Tab& T, const PS& s, int is, double rr, cd* m) {
  cd y[3][6];
  solid_start(T, s, is, rr, y);
  const int rows[14][3] = {{0, 1, 2}, {0, 1, 3}, {0, 1, 4}, {0, 1, 5}, {0, 2, 3},
                           {0, 2, 4}, {0, 3, 4}, {0, 3, 5}, {0, 4, 5}, {1, 2, 3},
                           {1, 3, 4}, {1, 3, 5}, {2, 3, 4}, {3, 4, 5}};
#pragma unroll
  for (int k = 0; k < 14; k++) {
    int p = rows[k][0], q = rows[k][1], r = rows[k][2];
    m[k] = y[0][p] * det2(y[1][q], y[1][r], y[2][q], y[2][r]) -
           y[1][p] * det2(y[0][q], y[0][r], y[2][q], y[2][r]) +
           y[2][p] * det2(y[0][q], y[0][r], y[1][q], y[1][r]);
  }
}

// sph_minor_fluid_start, src/module_int.f90:1509-1534 (5 2x2 minors)
__host__ __device__ inline void minor_fluid_start(const Tab& T, const PS& s, int is, double rr, cd* m) {
  cd y[2][4];
  fluid_start(T, s, is, rr, y);
  const int rows[5][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {2, 3}};
#pragma unroll
  for (int k = 0; k < 5; k++) {
    int p = rows[k][0], q = rows[k][1];
    m[k] = det2(y[0][p], y[0][q], y[1][p], y[1][q]);
  }
}

// tor_ystart with the sbessj2 stub, src/module_int.f90:213-250
__host__ __device__ inline void tor_ystart(const Tab& T, const PS& s, int is, double rr, cd* y) {
  SplW w = spl_w(T, is, rr);
  double rrho = spl(T, w, F_RHO), rmu = spl(T, w, F_MU), rqmu = spl(T, w, F_QMU);
  cd cmu = rmu * (1.0 + rqmu * s.lno);
  cd cbeta = csqrt_(cmu / rrho);
  cd arg = (s.omega / cbeta) * rr;
  cd ratio = arg / (double)(2 * (s.ll + 1) + 3);
  y[0] = C(s.zeta * rr);
  y[1] = s.zeta * (cmu * ((s.lld - 1.0) - arg * ratio));
}

// rad_ystart with the sbessj2 stub, src/module_int.f90:417-452
__host__ __device__ inline void rad_ystart(const Tab& T, const PS& s, int is, double rr, cd* y) {
  SplW w = spl_w(T, is, rr);
  double rrho = spl(T, w, F_RHO), rmu = spl(T, w, F_MU), rqmu = spl(T, w, F_QMU),
         rkappa = spl(T, w, F_KAPPA), rqkappa = spl(T, w, F_QKAPPA);
  cd cmu = rmu * (1.0 + rqmu * s.lno);
  cd ckappa = rkappa * (1.0 + rqkappa * s.lno);
  cd calpha = ckappa + (4.0 / 3.0) * cmu;
  cd gamma = s.omega2 + 16.0 * s.pibigg * rrho / 3.0;
  gamma = rrho * gamma / calpha;
  cd ratio = (gamma * rr) / 5.0;
  y[0] = rr * ratio;
  y[1] = rr * (calpha * gamma) - 4.0 * (cmu * ratio);
}

// ----------------------------------------------------------- minor algebra
// sph_minors_solid_2_fluid / fluid_2_solid, src/module_int.f90:1620-1680
__host__ __device__ inline void solid_2_fluid(const cd* a, cd* f) {
  f[0] = -a[5]; f[1] = -a[6]; f[2] = a[8]; f[3] = -a[12]; f[4] = a[13];
}
__host__ __device__ inline void fluid_2_solid(const cd* f, cd* a) {
#pragma unroll
  for (int i = 0; i < 14; i++) a[i] = C(0.0);
  a[0] = f[0]; a[1] = f[1]; a[3] = f[2]; a[9] = -f[3]; a[11] = -f[4];
}

// sph_delta_solid, src/module_int.f90:1705-1725
__host__ __device__ inline cd delta_solid(const cd* a, const cd* b) {
  return (a[0] * b[13] - a[13] * b[0]) + 2.0 * (a[6] * b[1] - a[1] * b[6]) +
         2.0 * (a[10] * b[2] - a[2] * b[10]) + (a[12] * b[3] - a[3] * b[12]) +
         2.0 * (a[7] * b[4] - a[4] * b[7]) + (a[11] * b[5] - a[5] * b[11]) +
         (a[8] * b[9] - a[9] * b[8]);
}

// b_solid_start for the unit source column e_k (k = 0,1,3,4 for the four
// columns of src/yspec.f90:234-238), src/module_int.f90:1563-1592
__host__ __device__ inline void b_start_unit(const cd* m, int k, cd* b) {
#pragma unroll
  for (int i = 0; i < 14; i++) b[i] = C(0.0);
  if (k == 0) {
    b[5] = m[13]; b[6] = -m[6]; b[7] = -m[10]; b[8] = -m[12]; b[9] = -m[7];
    b[10] = -m[11]; b[11] = m[1]; b[12] = m[4]; b[13] = m[9];
  } else if (k == 1) {
    b[1] = -m[13]; b[2] = m[6]; b[3] = m[10]; b[4] = m[12]; b[9] = -m[8];
    b[10] = m[7]; b[11] = m[2]; b[12] = m[5]; b[13] = -m[4];
  } else if (k == 3) {
    b[0] = -m[6]; b[1] = -m[7]; b[3] = -m[1]; b[4] = -m[4]; b[5] = -m[8];
    b[7] = -m[2]; b[8] = -m[5]; b[10] = -m[3]; b[13] = -m[0];
  } else {  // k == 4
    b[0] = -m[10]; b[1] = -m[11]; b[2] = m[1]; b[4] = -m[9]; b[5] = m[7];
    b[6] = m[2]; b[8] = m[4]; b[9] = m[3]; b[12] = m[0];
  }
}

// sph_ysol, src/module_int.f90:1785-1815
__host__ __device__ inline void ysol(const cd* m, const cd* b, cd* y) {
  y[0] = m[2] * b[2] + m[0] * b[5] + m[1] * b[6] + 2.0 * (m[2] * b[7]) + m[3] * b[8] +
         m[4] * b[9] + m[5] * b[10] + m[6] * b[11] + m[7] * b[12] + m[8] * b[13];
  y[1] = m[9] * b[9] + m[10] * b[11] + m[11] * b[12] - m[0] * b[1] -
         m[1] * (2.0 * b[2] + b[7]) - m[2] * b[3] - m[3] * b[4] - m[4] * b[10] - m[7] * b[13];
  y[2] = m[0] * b[0] - m[1] * b[8] + m[2] * b[4] + m[4] * (b[7] - b[2]) - m[5] * b[3] +
         m[6] * b[13] - m[9] * b[6] - m[10] * b[12] + m[12] * b[11];
  y[3] = m[1] * b[0] + m[4] * b[1] - m[6] * b[3] - m[7] * b[4] + m[9] * b[5] -
         m[10] * (b[2] + 2.0 * b[7]) - m[11] * b[8] - m[12] * b[10] + m[13] * b[13];
  y[4] = m[2] * b[0] - m[4] * b[5] + m[5] * b[1] + m[6] * (2.0 * b[2] + b[7]) + m[7] * b[8] -
         m[8] * b[4] + m[10] * b[6] + m[12] * b[9] - m[13] * b[12];
  y[5] = m[1] * b[5] - m[2] * b[1] + m[3] * b[0] + m[6] * b[10] + m[7] * (b[2] - b[7]) +
         m[8] * b[3] - m[10] * b[9] + m[11] * b[6] + m[13] * b[11];
}

__host__ __device__ inline cd pick3(int k, cd a, cd b, cd c) {
  return (k == 0) ? a : ((k == 1) ? b : ((k == 2) ? c : C(0.0)));
}
// coefficients of b in row k = 0, 1, 2 of sph_ysol (src/module_int.f90:1785-1815)
__host__ __device__ inline void ysol_row(const cd* m, int k, cd* x) {
  const cd z = C(0.0);
  x[0] = pick3(k, z, z, m[0]);
  x[1] = pick3(k, z, -m[0], z);
  x[2] = pick3(k, m[2], -2.0 * m[1], -m[4]);
  x[3] = pick3(k, z, -m[2], -m[5]);
  x[4] = pick3(k, z, -m[3], m[2]);
  x[5] = pick3(k, m[0], z, z);
  x[6] = pick3(k, m[1], z, -m[9]);
  x[7] = pick3(k, 2.0 * m[2], -m[1], m[4]);
  x[8] = pick3(k, m[3], z, -m[1]);
  x[9] = pick3(k, m[4], m[9], z);
  x[10] = pick3(k, m[5], -m[4], z);
  x[11] = pick3(k, m[6], m[10], m[12]);
  x[12] = pick3(k, m[7], m[11], -m[10]);
  x[13] = pick3(k, m[8], -m[7], m[6]);
}

__host__ __device__ inline void delta_row(const cd* m2, cd* x) {
  // coefficients of m1 in sph_delta_solid(m1, m2), src/module_int.f90:1705-1725
  x[0] = m2[13];
  x[1] = -2.0 * m2[6];
  x[2] = -2.0 * m2[10];
  x[3] = -m2[12];
  x[4] = -2.0 * m2[7];
  x[5] = -m2[11];
  x[6] = 2.0 * m2[1];
  x[7] = 2.0 * m2[4];
  x[8] = m2[9];
  x[9] = -m2[8];
  x[10] = 2.0 * m2[2];
  x[11] = m2[5];
  x[12] = m2[3];
  x[13] = -m2[0];
}

__host__ __device__ inline double ipow_huge(int n) {
  double r = 1.0;
  if (n > 0)
    for (int i = 0; i < n; i++) r *= HUGE_;
  else
    for (int i = 0; i < -n; i++) r *= TINY;
  return r;
}

}  // namespace ysx
