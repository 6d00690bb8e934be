// ysx_march.hpp -- lockstep radial marches for gfx950.
//
// A wavefront holds pairs that share one frequency w.  The reference's step
// rule (yint, src/module_int.f90:1903-1905) depends only on w and the knot
// interval, so every lane of the wave visits the same stage radii.  For each
// chunk of steps the wave evaluates the splines and the l-independent part
// of the coefficient matrices once per stage radius (one lane per radius)
// into LDS; each lane then adds its own l-dependent terms and runs its
// Cash-Karp step.  Control flow is uniform across the wave.
//
// The systems are linear (dy/dr = A y), so the base coefficients are stored
// multiplied by the step length h: the right-hand sides then give h dy/dr
// directly and the Cash-Karp tableau enters as plain constants.
#pragma once
#include "ysx_device.hpp"

namespace ysx {

// complex multiply-accumulate chains: acc +- a*b in 4 FP64 FMAs
// (complex a) or 2 (real a)
__host__ __device__ inline cd cmad(cd c, cd a, cd b) {
  return {fma(-a.y, b.y, fma(a.x, b.x, c.x)), fma(a.y, b.x, fma(a.x, b.y, c.y))};
}
__host__ __device__ inline cd cmsb(cd c, cd a, cd b) {
  return {fma(a.y, b.y, fma(-a.x, b.x, c.x)), fma(-a.y, b.x, fma(-a.x, b.y, c.y))};
}
__host__ __device__ inline cd rmad(cd c, double a, cd b) { return {fma(a, b.x, c.x), fma(a, b.y, c.y)}; }
__host__ __device__ inline cd rmsb(cd c, double a, cd b) { return {fma(-a, b.x, c.x), fma(-a, b.y, c.y)}; }
__host__ __device__ inline cd rmul(double a, cd b) { return {a * b.x, a * b.y}; }
__host__ __device__ inline cd cmul_(cd a, cd b) {
  return {fma(-a.y, b.y, a.x * b.x), fma(a.y, b.x, a.x * b.y)};
}

constexpr int CH = 10;       // steps per shared-coefficient chunk (6*CH radii)
constexpr int BASE_MAX = 20;  // doubles per stage radius

// w-dependent constants shared by the wave (set_omega, src/module_int.f90:2091-2102)
struct WS {
  cd omega, omega2, omegai2, lno;
  double rew, pibigg;
};
__host__ __device__ inline WS omega_state(cd w, double tref, double ats, double pibigg) {
  PS s;
  set_omega(s, w, tref, ats);
  WS o;
  o.omega = s.omega;
  o.omega2 = s.omega2;
  o.omegai2 = s.omegai2;
  o.lno = s.lno;
  o.rew = s.rew;
  o.pibigg = pibigg;
  return o;
}

// l-dependent constants of one lane (set_l, src/module_int.f90:2072-2089)
struct LS {
  double lld, llp1, zeta, zeta2, zeta2m2;
};
__host__ __device__ inline LS l_state(int l) {
  PS s;
  set_l(s, l);
  return LS{s.lld, s.llp1, s.zeta, s.zeta2, s.zeta2m2};
}

__host__ __device__ inline double step_len_w(const Tab& T, double rew, int i, int nknot) {
  double a = T(F_ALPHA, i);
  double tmp = (i < nknot) ? fmax(a, T(F_ALPHA, i + 1)) : a;
  return 1.0 / (4.0 * rew / tmp);
}

// ------------------------------------------------------------ solid region
// sph_solid_amat (src/module_int.f90:500-560) split into an l-independent
// base per radius (times h) and the l-dependent entries added per lane.
// GRAV = false takes gravity as zero (gravity switch 0, sph_solid_amat_ng,
// src/module_int.f90:564-610); the entries used there and with switch 1
// (sph_solid_amat_cow, :613-672) have the same form as with switch 2.
template <bool GRAV>
__host__ __device__ inline void solid_base(const Tab& T, const WS& w, int il, double rr, double h,
                                double* o) {
  SplW sw = spl_w(T, il, rr);
  double rrho = spl(T, sw, F_RHO), racon = spl(T, sw, F_ACON), rccon = spl(T, sw, F_CCON),
         rfcon = spl(T, sw, F_FCON), rncon = spl(T, sw, F_NCON), rlcon = spl(T, sw, F_LCON),
         gg = GRAV ? spl(T, sw, F_GRAV) : 0.0, rqmu = spl(T, sw, F_QMU),
         rxlam = spl(T, sw, F_XLAM), rxa2 = spl(T, sw, F_XA2);
  cd qaa = 1.0 + rxa2 * w.lno, qff = 1.0 + rxlam * w.lno, qll = 1.0 + rqmu * w.lno;
  cd cacon = racon * qaa, cfcon = rfcon * qff, cncon = rncon * qll;
  cd cccon = cinv(rccon * qaa);
  cd fc = cfcon * cccon;
  cd gamma = cacon - cncon - cfcon * fc;
  double ri = 1.0 / rr, ri2 = ri * ri;
  double rgr = rrho * gg * rr;
  cd mw2r = -rrho * w.omega2;
  cd t11 = (1.0 - 2.0 * fc) * ri;
  cd k22 = cinv(rlcon * qll);
  cd s11 = mw2r + (4.0 * ri2) * (gamma - rgr);
  cd fcri = ri * fc;
  cd q12 = ri2 * (rgr - 2.0 * gamma);
  cd a22 = mw2r - (2.0 * ri2) * cncon;
  cd gc = ri2 * (gamma + cncon);
  o[0] = h * ri; o[1] = rrho;
  o[2] = h * t11.x; o[3] = h * t11.y; o[4] = h * cccon.x; o[5] = h * cccon.y;
  o[6] = h * k22.x; o[7] = h * k22.y; o[8] = h * s11.x; o[9] = h * s11.y;
  o[10] = h * fcri.x; o[11] = h * fcri.y; o[12] = h * q12.x; o[13] = h * q12.y;
  o[14] = h * a22.x; o[15] = h * a22.y; o[16] = h * gc.x; o[17] = h * gc.y;
  o[18] = h * (-4.0 * w.pibigg * rrho); o[19] = h * (4.0 * w.pibigg);
}

struct SolidB {
  static constexpr int BS = 20;
  __host__ __device__ static void base(const Tab& T, const WS& w, int il, double rr, double h, double* o) {
    solid_base<true>(T, w, il, rr, h, o);
  }
  // h A from the base of one stage radius
  __host__ __device__ static inline void amat(const double* b, const LS& L, SolidA& A) {
    double hri = b[0], rrho = b[1];
    A.t11 = C(b[2], b[3]);
    A.k11 = C(b[4], b[5]);
    A.k22 = C(b[6], b[7]);
    A.s11 = C(b[8], b[9]);
    A.t12 = L.zeta * C(b[10], b[11]);
    A.s12 = L.zeta * C(b[12], b[13]);
    A.s22 = C(b[14], b[15]) + L.zeta2 * C(b[16], b[17]);
    A.t21 = -L.zeta * hri;
    A.t22 = 2.0 * hri;
    A.t31 = b[18];
    A.t33 = -L.lld * hri;
    A.k33 = b[19];
    A.s13 = -rrho * L.llp1 * hri;
    A.s23 = rrho * L.zeta * hri;
  }
};

// First term of output i of a derivative chain: the plain product, or the
// product added to an initial value g(i) (a Runge-Kutta combination that
// takes this stage's derivative once, folded into the chain).
struct NoInit {
  __host__ __device__ inline cd r(int, double a, cd m) const { return rmul(a, m); }
  __host__ __device__ inline cd c(int, cd a, cd m) const { return cmul_(a, m); }
};
template <class G>
struct AccInit {
  const G& g;
  __host__ __device__ inline cd r(int i, double a, cd m) const { return rmad(g(i), a, m); }
  __host__ __device__ inline cd c(int i, cd a, cd m) const { return cmad(g(i), a, m); }
};

// sph_minors_solid_derivs, src/module_int.f90:818-871, as FMA chains
template <class Init, class Sink>
__host__ __device__ inline void solid_minor_derivs(const SolidA& A, const cd* m, const Init& I,
                                          Sink&& sink) {
  const double t21 = A.t21, t22 = A.t22, t31 = A.t31, t33 = A.t33, k33 = A.k33, s13 = A.s13,
               s23 = A.s23;
  const cd t11 = A.t11, t12 = A.t12, k11 = A.k11, k22 = A.k22, s11 = A.s11, s12 = A.s12,
           s22 = A.s22;
  const double tp = t22 + t33, tm = t22 - t33;
  const cd a123 = t11 + tp, a126 = t11 + tm, a135 = t11 - tm, a234 = tp - t11;
  const cd t12x2 = 2.0 * t12, s12x2 = 2.0 * s12;
  cd acc;
  acc = I.r(0, k33, m[3]); acc = cmsb(acc, k22, m[5]); acc = cmad(acc, k11, m[9]);
  sink(0, cmad(acc, m[0], a123));
  acc = I.r(1, s13, m[0]); acc = cmsb(acc, k22, m[6]); acc = rmsb(acc, t21, m[2]);
  acc = rmad(acc, t22, m[1]); sink(1, rmsb(acc, t31, m[3]));
  acc = I.r(2, s23, m[0]); acc = cmsb(acc, k11, m[10]); acc = cmad(acc, t11, m[2]);
  sink(2, cmsb(acc, t12, m[1]));
  acc = I.c(3, k22, m[8]); acc = cmsb(acc, k11, m[11]); sink(3, cmad(acc, a126, m[3]));
  acc = I.c(4, t12, m[9]); acc = rmsb(acc, k33, m[7]); acc = cmsb(acc, s12, m[0]);
  acc = rmsb(acc, t21, m[5]); acc = rmad(acc, t31, m[2]); sink(4, rmad(acc, t33, m[4]));
  acc = I.c(5, a135, m[5]); acc = rmsb(acc, k33, m[8]); acc = cmsb(acc, k11, m[12]);
  acc = cmsb(acc, s22, m[0]); sink(5, cmsb(acc, t12x2, m[4]));
  acc = I.c(6, s12, m[2]); acc = rmad(acc, s13, m[5]); acc = cmsb(acc, s22, m[1]);
  acc = rmsb(acc, s23, m[4]); acc = cmad(acc, t12, m[10]); acc = rmsb(acc, t22, m[6]);
  sink(6, rmad(acc, t31, m[8]));
  acc = I.c(7, s12, m[3]); acc = rmsb(acc, s13, m[2]); acc = rmsb(acc, s23, m[1]);
  acc = cmad(acc, t12, m[11]); acc = rmsb(acc, t21, m[8]); sink(7, rmsb(acc, t33, m[7]));
  acc = I.c(8, k11, m[13]); acc = cmad(acc, s22, m[3]); acc = rmsb(acc, 2.0 * s23, m[2]);
  acc = cmsb(acc, t12x2, m[7]); sink(8, cmsb(acc, a234, m[8]));
  acc = I.c(9, k22, m[12]); acc = rmsb(acc, k33, m[11]); acc = cmad(acc, s11, m[0]);
  acc = rmad(acc, 2.0 * t21, m[4]); acc = rmsb(acc, 2.0 * t31, m[1]);
  sink(9, cmad(acc, a234, m[9]));
  acc = I.c(10, s12, m[1]); acc = cmsb(acc, s11, m[2]); acc = rmsb(acc, s13, m[4]);
  acc = rmsb(acc, s23, m[9]); acc = cmsb(acc, t11, m[10]); acc = rmad(acc, t21, m[6]);
  sink(10, rmsb(acc, t31, m[7]));
  acc = I.r(11, 2.0 * s13, m[1]); acc = cmsb(acc, k22, m[13]); acc = cmsb(acc, s11, m[3]);
  acc = rmad(acc, 2.0 * t21, m[7]); sink(11, cmsb(acc, a135, m[11]));
  acc = I.r(12, k33, m[13]); acc = cmsb(acc, s11, m[5]); acc = cmad(acc, s12x2, m[4]);
  acc = cmad(acc, s22, m[9]); acc = rmad(acc, 2.0 * t31, m[6]); sink(12, cmsb(acc, a126, m[12]));
  acc = I.c(13, s11, m[8]); acc = cmsb(acc, s12x2, m[7]); acc = rmad(acc, 2.0 * s13, m[6]);
  acc = cmsb(acc, s22, m[11]); acc = rmad(acc, 2.0 * s23, m[10]);
  sink(13, cmsb(acc, a123, m[13]));
}

// sph_bvec_solid_derivs, src/module_int.f90:1016-1057, as FMA chains
template <class Init, class Sink>
__host__ __device__ inline void bvec_derivs(const SolidA& A, const cd* b, const Init& I, Sink&& sink) {
  const double t21 = A.t21, t22 = A.t22, t31 = A.t31, t33 = A.t33, k33 = A.k33, s13 = A.s13,
               s23 = A.s23;
  const cd t11 = A.t11, t12 = A.t12, k11 = A.k11, k22 = A.k22, s11 = A.s11, s12 = A.s12,
           s22 = A.s22;
  const cd t1122 = t11 + t22, t1133 = t11 + t33;
  const double t2233 = t22 + t33;
  const cd b38m = b[7] - b[2], b2p8 = 2.0 * b[2] + b[7], b3p28 = b[2] + 2.0 * b[7];
  cd acc;
  acc = I.c(0, s11, b[6]); acc = cmad(acc, s12, b38m); acc = rmad(acc, s13, b[8]);
  acc = cmsb(acc, s22, b[3]); acc = rmsb(acc, s23, b[4]); acc = cmsb(acc, t1122, b[0]);
  sink(0, rmad(acc, t31, b[5]));
  acc = I.c(1, s11, b[9]); acc = cmad(acc, s12, b[10]); acc = rmsb(acc, s13, b2p8);
  acc = rmsb(acc, s23, b[3]); acc = rmsb(acc, t21, b[5]); sink(1, cmsb(acc, t1133, b[1]));
  acc = I.c(2, s12, b[11]); acc = rmad(acc, s13, b[12]); acc = cmad(acc, t12, b[3]);
  acc = rmsb(acc, t21, b[6]); sink(2, rmsb(acc, t31, b[9]));
  acc = I.r(3, s13, b[13]); acc = cmsb(acc, k22, b[0]); acc = cmsb(acc, s11, b[11]);
  acc = rmsb(acc, t21, b38m); acc = cmad(acc, t22 - t11, b[3]); sink(3, rmsb(acc, t31, b[10]));
  acc = I.r(4, t31, b2p8); acc = rmsb(acc, k33, b[1]); acc = cmsb(acc, s11, b[12]);
  acc = cmsb(acc, s12, b[13]); acc = rmsb(acc, t21, b[8]); sink(4, cmad(acc, t33 - t11, b[4]));
  acc = I.c(5, s12, b[9]); acc = rmsb(acc, s13, b[6]); acc = cmad(acc, s22, b[10]);
  acc = rmsb(acc, s23, b3p28); acc = cmsb(acc, t12, b[1]); sink(5, rmsb(acc, t2233, b[5]));
  acc = I.c(6, k11, b[0]); acc = cmad(acc, s22, b[11]); acc = rmad(acc, s23, b[12]);
  acc = cmad(acc, t12, b38m); sink(6, cmad(acc, t11 - t22, b[6]));
  acc = I.r(7, s23, b[13]); acc = cmsb(acc, s12, b[11]); acc = cmsb(acc, t12, b[3]);
  sink(7, rmad(acc, t21, b[6]));
  acc = I.r(8, t31, b[6]); acc = rmsb(acc, k33, b[5]); acc = cmsb(acc, s12, b[12]);
  acc = cmsb(acc, s22, b[13]); acc = cmsb(acc, t12, b[4]); sink(8, rmad(acc, t33 - t22, b[8]));
  acc = I.c(9, k11, b[1]); acc = rmad(acc, s23, b[11]); acc = cmad(acc, t12, b[10]);
  sink(9, cmad(acc, t11 - t33, b[9]));
  acc = I.c(10, k22, b[5]); acc = rmsb(acc, s13, b[11]); acc = rmad(acc, t21, b[9]);
  sink(10, rmad(acc, t22 - t33, b[10]));
  acc = I.c(11, k22, b[6]); acc = cmsb(acc, k11, b[3]); sink(11, cmad(acc, t1122, b[11]));
  acc = I.r(12, k33, b[9]); acc = cmsb(acc, k11, b[4]); acc = cmad(acc, t12, b[13]);
  sink(12, cmad(acc, t1133, b[12]));
  acc = I.r(13, k33, b[10]); acc = cmsb(acc, k22, b[8]); acc = rmad(acc, t21, b[12]);
  acc = rmsb(acc, t31, b[11]); sink(13, rmad(acc, t2233, b[13]));
}

// Transpose of sph_bvec_solid_derivs (src/module_int.f90:1016-1057):
// e = A^T x for the b-vector system's matrix A, as FMA chains.  It is the
// derivative of the receiver-side row march (rk_step_T_lds).
template <class Init, class Sink>
__host__ __device__ inline void bvecT_derivs(const SolidA& A, const cd* x, const Init& I, Sink&& sink) {
  const double t21 = A.t21, t22 = A.t22, t31 = A.t31, t33 = A.t33, k33 = A.k33, s13 = A.s13,
               s23 = A.s23;
  const cd t11 = A.t11, t12 = A.t12, k11 = A.k11, k22 = A.k22, s11 = A.s11, s12 = A.s12,
           s22 = A.s22;
  const cd t1122 = t11 + t22, t1133 = t11 + t33;
  const double t2233 = t22 + t33;
  const cd x2m7 = x[2] - x[7];
  cd acc;
  acc = I.c(0, k11, x[6]); acc = cmsb(acc, k22, x[3]); sink(0, cmsb(acc, t1122, x[0]));
  acc = I.c(1, k11, x[9]); acc = rmsb(acc, k33, x[4]); acc = cmsb(acc, t12, x[5]);
  sink(1, cmsb(acc, t1133, x[1]));
  acc = I.r(2, t21, x[3]); acc = cmsb(acc, s12, x[0]); acc = rmsb(acc, 2.0 * s13, x[1]);
  acc = rmad(acc, 2.0 * t31, x[4]); acc = rmsb(acc, s23, x[5]); sink(2, cmsb(acc, t12, x[6]));
  acc = I.c(3, t12, x2m7); acc = cmsb(acc, s22, x[0]); acc = rmsb(acc, s23, x[1]);
  acc = cmad(acc, t22 - t11, x[3]); sink(3, cmsb(acc, k11, x[11]));
  acc = I.c(4, t33 - t11, x[4]); acc = rmsb(acc, s23, x[0]); acc = cmsb(acc, t12, x[8]);
  sink(4, cmsb(acc, k11, x[12]));
  acc = I.c(5, k22, x[10]); acc = rmad(acc, t31, x[0]); acc = rmsb(acc, t21, x[1]);
  acc = rmsb(acc, t2233, x[5]); sink(5, rmsb(acc, k33, x[8]));
  acc = I.c(6, s11, x[0]); acc = rmsb(acc, t21, x2m7); acc = rmsb(acc, s13, x[5]);
  acc = cmad(acc, t11 - t22, x[6]); acc = rmad(acc, t31, x[8]); sink(6, cmad(acc, k22, x[11]));
  acc = I.c(7, s12, x[0]); acc = rmsb(acc, s13, x[1]); acc = rmsb(acc, t21, x[3]);
  acc = rmad(acc, t31, x[4]); acc = rmsb(acc, 2.0 * s23, x[5]); sink(7, cmad(acc, t12, x[6]));
  acc = I.r(8, s13, x[0]); acc = rmsb(acc, t21, x[4]); acc = rmad(acc, t33 - t22, x[8]);
  sink(8, cmsb(acc, k22, x[13]));
  acc = I.c(9, s11, x[1]); acc = rmsb(acc, t31, x[2]); acc = cmad(acc, s12, x[5]);
  acc = cmad(acc, t11 - t33, x[9]); acc = rmad(acc, t21, x[10]); sink(9, rmad(acc, k33, x[12]));
  acc = I.c(10, s12, x[1]); acc = rmsb(acc, t31, x[3]); acc = cmad(acc, s22, x[5]);
  acc = cmad(acc, t12, x[9]); acc = rmad(acc, t22 - t33, x[10]);
  sink(10, rmad(acc, k33, x[13]));
  acc = I.c(11, s12, x2m7); acc = cmsb(acc, s11, x[3]); acc = cmad(acc, s22, x[6]);
  acc = rmad(acc, s23, x[9]); acc = rmsb(acc, s13, x[10]); acc = cmad(acc, t1122, x[11]);
  sink(11, rmsb(acc, t31, x[13]));
  acc = I.r(12, s13, x[2]); acc = cmsb(acc, s11, x[4]); acc = rmad(acc, s23, x[6]);
  acc = cmsb(acc, s12, x[8]); acc = cmad(acc, t1133, x[12]); sink(12, rmad(acc, t21, x[13]));
  acc = I.r(13, s13, x[3]); acc = cmsb(acc, s12, x[4]); acc = rmad(acc, s23, x[7]);
  acc = cmsb(acc, s22, x[8]); acc = cmad(acc, t12, x[12]); sink(13, rmad(acc, t2233, x[13]));
}

struct SolidMinorM : SolidB {
  static constexpr int N = 14;
  template <class Sink>
  __host__ __device__ inline void stream(const double* b, const LS& L, const WS& w, const cd* y,
                                Sink&& sink) const {
    SolidA A;
    amat(b, L, A);
    solid_minor_derivs(A, y, NoInit(), sink);
  }
  // sink(i, g(i) + h A y) for a base scaled by h
  template <class G, class Sink>
  __host__ __device__ inline void stream_acc(const double* b, const LS& L, const WS& w, const cd* y,
                                    const G& g, Sink&& sink) const {
    SolidA A;
    amat(b, L, A);
    solid_minor_derivs(A, y, AccInit<G>{g}, sink);
  }
  __host__ __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* y,
                                    cd* d) const {
    stream(b, L, w, y, [&](int i, cd v) { d[i] = v; });
  }
};
struct BvecM : SolidB {
  static constexpr int N = 14;
  template <class Sink>
  __host__ __device__ inline void stream(const double* b, const LS& L, const WS& w, const cd* y,
                                Sink&& sink) const {
    SolidA A;
    amat(b, L, A);
    bvec_derivs(A, y, NoInit(), sink);
  }
  template <class G, class Sink>
  __host__ __device__ inline void stream_acc(const double* b, const LS& L, const WS& w, const cd* y,
                                    const G& g, Sink&& sink) const {
    SolidA A;
    amat(b, L, A);
    bvec_derivs(A, y, AccInit<G>{g}, sink);
  }
  __host__ __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* y,
                                    cd* d) const {
    stream(b, L, w, y, [&](int i, cd v) { d[i] = v; });
  }
};

// Transpose of sph_minors_solid_derivs (src/module_int.f90:818-871):
// e = A^T x for the solid minor system's matrix A.  Marched from the
// receiver back to the source it carries the coefficients of the secular
// determinant (delta_solid in its first argument) down to the source.
template <class Init, class Sink>
__host__ __device__ inline void minorT_derivs(const SolidA& A, const cd* x, const Init& I, Sink&& sink) {
  const double t21 = A.t21, t22 = A.t22, t31 = A.t31, t33 = A.t33, k33 = A.k33, s13 = A.s13,
               s23 = A.s23;
  const cd t11 = A.t11, t12 = A.t12, k11 = A.k11, k22 = A.k22, s11 = A.s11, s12 = A.s12,
           s22 = A.s22;
  const double tp = t22 + t33, tm = t22 - t33;
  const cd a123 = t11 + tp, a126 = t11 + tm, a135 = t11 - tm, a234 = tp - t11;
  const cd t12x2 = 2.0 * t12, s12x2 = 2.0 * s12;
  cd acc;
  acc = I.c(0, s11, x[9]); acc = cmad(acc, a123, x[0]); acc = rmad(acc, s13, x[1]);
  acc = rmad(acc, s23, x[2]); acc = cmsb(acc, s12, x[4]); sink(0, cmsb(acc, s22, x[5]));
  acc = I.c(1, s12, x[10]); acc = rmad(acc, t22, x[1]); acc = cmsb(acc, t12, x[2]);
  acc = cmsb(acc, s22, x[6]); acc = rmsb(acc, s23, x[7]); acc = rmsb(acc, 2.0 * t31, x[9]);
  sink(1, rmad(acc, 2.0 * s13, x[11]));
  acc = I.c(2, t11, x[2]); acc = rmsb(acc, t21, x[1]); acc = rmad(acc, t31, x[4]);
  acc = cmad(acc, s12, x[6]); acc = rmsb(acc, s13, x[7]); acc = rmsb(acc, 2.0 * s23, x[8]);
  sink(2, cmsb(acc, s11, x[10]));
  acc = I.c(3, a126, x[3]); acc = rmad(acc, k33, x[0]); acc = rmsb(acc, t31, x[1]);
  acc = cmad(acc, s12, x[7]); acc = cmad(acc, s22, x[8]); sink(3, cmsb(acc, s11, x[11]));
  acc = I.c(4, s12x2, x[12]); acc = rmad(acc, t33, x[4]); acc = cmsb(acc, t12x2, x[5]);
  acc = rmsb(acc, s23, x[6]); acc = rmad(acc, 2.0 * t21, x[9]); sink(4, rmsb(acc, s13, x[10]));
  acc = I.c(5, a135, x[5]); acc = cmsb(acc, k22, x[0]); acc = rmsb(acc, t21, x[4]);
  acc = rmad(acc, s13, x[6]); sink(5, cmsb(acc, s11, x[12]));
  acc = I.r(6, t21, x[10]); acc = cmsb(acc, k22, x[1]); acc = rmsb(acc, t22, x[6]);
  acc = rmad(acc, 2.0 * t31, x[12]); sink(6, rmad(acc, 2.0 * s13, x[13]));
  acc = I.r(7, 2.0 * t21, x[11]); acc = rmsb(acc, k33, x[4]); acc = rmsb(acc, t33, x[7]);
  acc = cmsb(acc, t12x2, x[8]); acc = rmsb(acc, t31, x[10]); sink(7, cmsb(acc, s12x2, x[13]));
  acc = I.c(8, k22, x[3]); acc = rmsb(acc, k33, x[5]); acc = rmad(acc, t31, x[6]);
  acc = rmsb(acc, t21, x[7]); acc = cmsb(acc, a234, x[8]); sink(8, cmad(acc, s11, x[13]));
  acc = I.c(9, k11, x[0]); acc = cmad(acc, t12, x[4]); acc = cmad(acc, a234, x[9]);
  acc = rmsb(acc, s23, x[10]); sink(9, cmad(acc, s22, x[12]));
  acc = I.c(10, t12, x[6]); acc = cmsb(acc, k11, x[2]); acc = cmsb(acc, t11, x[10]);
  sink(10, rmad(acc, 2.0 * s23, x[13]));
  acc = I.c(11, t12, x[7]); acc = cmsb(acc, k11, x[3]); acc = rmsb(acc, k33, x[9]);
  acc = cmsb(acc, a135, x[11]); sink(11, cmsb(acc, s22, x[13]));
  acc = I.c(12, k22, x[9]); acc = cmsb(acc, k11, x[5]); sink(12, cmsb(acc, a126, x[12]));
  acc = I.c(13, k11, x[8]); acc = cmsb(acc, k22, x[11]); acc = rmad(acc, k33, x[12]);
  sink(13, cmsb(acc, a123, x[13]));
}

// the transposed solid minor system (the secular determinant's row)
struct MinorTM : SolidB {
  static constexpr int N = 14;
  template <class Sink>
  __host__ __device__ inline void stream(const double* b, const LS& L, const WS& w, const cd* y,
                                Sink&& sink) const {
    SolidA A;
    amat(b, L, A);
    minorT_derivs(A, y, NoInit(), sink);
  }
  template <class G, class Sink>
  __host__ __device__ inline void stream_acc(const double* b, const LS& L, const WS& w, const cd* y,
                                    const G& g, Sink&& sink) const {
    SolidA A;
    amat(b, L, A);
    minorT_derivs(A, y, AccInit<G>{g}, sink);
  }
};

// the transposed b-vector system (receiver-side rows)
struct BvecTM : SolidB {
  static constexpr int N = 14;
  template <class Sink>
  __host__ __device__ inline void stream(const double* b, const LS& L, const WS& w, const cd* y,
                                Sink&& sink) const {
    SolidA A;
    amat(b, L, A);
    bvecT_derivs(A, y, NoInit(), sink);
  }
  template <class G, class Sink>
  __host__ __device__ inline void stream_acc(const double* b, const LS& L, const WS& w, const cd* y,
                                    const G& g, Sink&& sink) const {
    SolidA A;
    amat(b, L, A);
    bvecT_derivs(A, y, AccInit<G>{g}, sink);
  }
};

// ------------------------------------------------------------ fluid region
// sph_fluid_amat + sph_minors_fluid_derivs, src/module_int.f90:676-715,
// 941-967, split into an l-independent base (times h) and l-dependent
// entries.
struct FluidMinorM {
  static constexpr int N = 5, BS = 14;
  __host__ __device__ static void base(const Tab& T, const WS& w, int il, double rr, double h, double* o) {
    SplW sw = spl_w(T, il, rr);
    double rrho = spl(T, sw, F_RHO), rccon = spl(T, sw, F_CCON), gg = spl(T, sw, F_GRAV),
           rxa2 = spl(T, sw, F_XA2);
    cd cccon = cinv(rccon * (1.0 + rxa2 * w.lno));
    double ri = 1.0 / rr, ri2 = ri * ri;
    cd fb = ri2 * w.omegai2;   // ri^2 / w^2
    cd fa = gg * fb;           // gg ri^2 / w^2
    cd s11a = (rrho * gg) * fa;
    cd s11b = (-4.0 * rrho * gg * ri) - rrho * w.omega2;
    o[0] = h * ri; o[1] = rrho;
    o[2] = h * fa.x; o[3] = h * fa.y; o[4] = h * fb.x; o[5] = h * fb.y;
    o[6] = h * cccon.x; o[7] = h * cccon.y; o[8] = h * s11a.x; o[9] = h * s11a.y;
    o[10] = h * s11b.x; o[11] = h * s11b.y;
    o[12] = h * (-4.0 * w.pibigg * rrho); o[13] = h * (4.0 * w.pibigg);
  }
  // h dm/dr
  __host__ __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* m,
                                    cd* d) const {
    double ri = b[0], rrho = b[1];  // ri: h / r
    cd fa = C(b[2], b[3]), fb = C(b[4], b[5]), cccon = C(b[6], b[7]);
    double z2 = L.zeta2;
    cd t11 = z2 * fa - ri;
    cd t12 = z2 * fb;
    double t21 = b[12];
    double t22 = -L.lld * ri;
    cd k11 = cccon - (z2 / rrho) * fb;
    double k22 = b[13];
    cd s11 = z2 * C(b[8], b[9]) + C(b[10], b[11]);
    cd s12 = (z2 * rrho) * fa - L.llp1 * rrho * ri;
    cd s22 = (z2 * rrho) * fb;
    cd tp = t11 + t22, tm = t11 - t22;
    d[0] = cmad(cmsb(rmul(k22, m[2]), k11, m[3]), tp, m[0]);
    d[1] = rmsb(cmad(cmul_(s12, m[0]), t12, m[3]), t21, m[2]);
    d[2] = cmad(cmsb(cmad(cmul_(k11, m[4]), s22, m[0]), 2.0 * t12, m[1]), tm, m[2]);
    d[3] = cmsb(cmsb(rmsb(rmul(2.0 * t21, m[1]), k22, m[4]), s11, m[0]), tm, m[3]);
    d[4] = cmsb(cmsb(cmsb(cmul_(s11, m[2]), 2.0 * s12, m[1]), s22, m[3]), tp, m[4]);
  }
  template <class Sink>
  __host__ __device__ inline void stream(const double* b, const LS& L, const WS& w, const cd* m,
                                         Sink&& sink) const {
    cd d[N];
    (*this)(b, L, w, m, d);
    for (int i = 0; i < N; i++) sink(i, d[i]);
  }
};

// ---------------------------------------------------------------- toroidal
// tor_amat + tor_derivs, src/module_int.f90:134-184; base times h
struct TorM {
  static constexpr int N = 2, BS = 7;
  __host__ __device__ static void base(const Tab& T, const WS& w, int il, double rr, double h, double* o) {
    SplW sw = spl_w(T, il, rr);
    double rrho = spl(T, sw, F_RHO), rlcon = spl(T, sw, F_LCON), rncon = spl(T, sw, F_NCON),
           rqmu = spl(T, sw, F_QMU);
    cd qll = 1.0 + rqmu * w.lno;
    double ri = 1.0 / rr;
    cd k = cinv(rlcon * qll);
    cd sv0 = (ri * ri * rncon) * qll;
    cd mw = -rrho * w.omega2;
    o[0] = h * ri; o[1] = h * k.x; o[2] = h * k.y; o[3] = h * sv0.x; o[4] = h * sv0.y;
    o[5] = h * mw.x; o[6] = h * mw.y;
  }
  __host__ __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* y,
                                    cd* d) const {
    double t = 2.0 * b[0];
    cd k = C(b[1], b[2]);
    cd sv = L.zeta2m2 * C(b[3], b[4]) + C(b[5], b[6]);
    d[0] = cmad(rmul(t, y[0]), k, y[1]);
    d[1] = rmsb(cmul_(sv, y[0]), t, y[1]);
  }
  template <class Sink>
  __host__ __device__ inline void stream(const double* b, const LS& L, const WS& w, const cd* y,
                                         Sink&& sink) const {
    cd d[N];
    (*this)(b, L, w, y, d);
    for (int i = 0; i < N; i++) sink(i, d[i]);
  }
};

// ----------------------------------------------------- Cash-Karp, shared base
// rkck (src/module_int.f90:2016-2064) with the six stage coefficient sets
// read from the wave's LDS chunk (stride F::BS doubles per stage); f gives
// h dy/dr, so the tableau (src/module_int.f90:2039-2049) enters unscaled.
namespace ck {
constexpr double b21 = 0.2, b31 = 3.0 / 40.0, b32 = 9.0 / 40.0, b41 = 0.3, b42 = -0.9,
                 b43 = 1.2, b51 = -11.0 / 54.0, b52 = 2.5, b53 = -70.0 / 27.0,
                 b54 = 35.0 / 27.0, b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0,
                 b63 = 575.0 / 13824.0, b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0,
                 c1 = 37.0 / 378.0, c3 = 250.0 / 621.0, c4 = 125.0 / 594.0,
                 c6 = 512.0 / 1771.0;
}  // namespace ck

// The stages are kept apart by scheduling barriers (the RHS of one stage
// would otherwise be interleaved with the next and spill); mask 0x100 lets
// only LDS reads cross them, so the next stage's base coefficients and slot
// values can be fetched early.
template <class F>
__device__ inline void rk_step_b(const double* sb, const LS& L, const WS& w, cd* y, const F& f) {
  using namespace ck;
  constexpr int N = F::N, BS = F::BS;
  cd k1[N], k2[N], k3[N], t[N];
  f(sb, L, w, y, k1);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = rmad(y[i], b21, k1[i]);
  __builtin_amdgcn_sched_barrier(0x100);
  f(sb + BS, L, w, t, k2);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = rmad(rmad(y[i], b31, k1[i]), b32, k2[i]);
  __builtin_amdgcn_sched_barrier(0x100);
  f(sb + 2 * BS, L, w, t, k3);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = rmad(rmad(rmad(y[i], b41, k1[i]), b42, k2[i]), b43, k3[i]);
  {
    cd k4[N];
    __builtin_amdgcn_sched_barrier(0x100);
    f(sb + 3 * BS, L, w, t, k4);
#pragma unroll
    for (int i = 0; i < N; i++) {
      cd t5 = rmad(rmad(rmad(rmad(y[i], b51, k1[i]), b52, k2[i]), b53, k3[i]), b54, k4[i]);
      cd p6 = rmad(rmad(rmad(rmad(y[i], b61, k1[i]), b62, k2[i]), b63, k3[i]), b64, k4[i]);
      cd po = rmad(rmad(rmad(y[i], c1, k1[i]), c3, k3[i]), c4, k4[i]);
      t[i] = t5;
      k2[i] = p6;
      k1[i] = po;
    }
  }
  __builtin_amdgcn_sched_barrier(0x100);
  f(sb + 4 * BS, L, w, t, k3);  // k5
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = rmad(k2[i], b65, k3[i]);
  __builtin_amdgcn_sched_barrier(0x100);
  f(sb + 5 * BS, L, w, t, k3);  // k6
#pragma unroll
  for (int i = 0; i < N; i++) y[i] = rmad(k1[i], c6, k3[i]);
}

// The default stage hook of rk_step_lds / rk_step_T_lds: nothing.
struct NoHook {
  __device__ void operator()(int, const cd*) const {}
};
// march_interval's hook before every step: pre(y, scale) sees the state
// and renormalisation counter the step starts from
struct NoPre {
  __device__ void operator()(const cd*, int) const {}
};

// Cash-Karp step with stage vectors kept in per-lane LDS slots.
// slot points at this lane's first slot entry; entry (s, i) is at
// slot[(s * N + i) * 64] (lane-contiguous, conflict-free).  Three register
// vectors (y, the stage input t, one accumulator u) and two LDS vectors
// (k1, then y + the stage-6 sum) hold the step; every RHS output is consumed
// as soon as it is produced.  After stage 3, y is dead and its registers
// accumulate the stage-5 input; the stage-6 input is formed in t and the
// result in y, so each component takes 9 LDS accesses per step.  Same
// arithmetic as rk_step_b.  hook(k, z) sees stage k's input z (registers)
// just before its derivative is taken (the adjoint's replay).
template <class F, class Hook = NoHook>
__device__ inline void rk_step_lds(const double* sb, const LS& L, const WS& w, cd* y, const F& f,
                                   cd* slot, Hook&& hook = Hook()) {
  using namespace ck;
  constexpr int N = F::N, BS = F::BS;
  cd* S1 = slot;
  cd* S2 = slot + N * 64;
  cd t[N], u[N];
  // Slot reads are issued one component ahead of their use (p1, p2), so
  // that LDS latency overlaps the arithmetic of the previous component.
  cd p1, p2;  // stages 2 and 3
  // stage 1: k1 -> S1, t = y + b21 k1
  hook(0, y);
  f.stream(sb, L, w, y, [&](int i, cd d) {
    S1[i * 64] = d;
    t[i] = rmad(y[i], b21, d);
  });
  __builtin_amdgcn_sched_barrier(0x100);
  // stage 2: k2 -> u
  hook(1, t);
  f.stream(sb + BS, L, w, t, [&](int i, cd d) { u[i] = d; });
  p1 = S1[0];
#pragma unroll
  for (int i = 0; i < N; i++) {
    cd k1 = p1;
    if (i + 1 < N) p1 = S1[(i + 1) * 64];
    t[i] = rmad(rmad(y[i], b31, k1), b32, u[i]);
  }
  __builtin_amdgcn_sched_barrier(0x100);
  // stage 3: k3 -> S2; then t = t4, y = y + A5, S2 = y + A6, u = y + Aout
  hook(2, t);
  f.stream(sb + 2 * BS, L, w, t, [&](int i, cd d) { S2[i * 64] = d; });
  p1 = S1[0];
  p2 = S2[0];
#pragma unroll
  for (int i = 0; i < N; i++) {
    cd k1 = p1, k2 = u[i], k3 = p2, yi = y[i];
    if (i + 1 < N) {
      p1 = S1[(i + 1) * 64];
      p2 = S2[(i + 1) * 64];
    }
    t[i] = rmad(rmad(rmad(yi, b41, k1), b42, k2), b43, k3);
    y[i] = rmad(rmad(rmad(yi, b51, k1), b52, k2), b53, k3);
    S2[i * 64] = rmad(rmad(rmad(yi, b61, k1), b62, k2), b63, k3);
    u[i] = rmad(rmad(yi, c1, k1), c3, k3);
  }
  __builtin_amdgcn_sched_barrier(0x100);
  // stage 4: fold k4 into the three accumulators; y is now the stage-5 input
  // (reads two components ahead: q[i & 1])
  hook(3, t);
  cd q[2] = {S2[0], S2[64]};
  f.stream(sb + 3 * BS, L, w, t, [&](int i, cd d) {
    cd a6 = q[i & 1];
    if (i + 2 < N) q[i & 1] = S2[(i + 2) * 64];
    y[i] = rmad(y[i], b54, d);
    S2[i * 64] = rmad(a6, b64, d);
    u[i] = rmad(u[i], c4, d);
  });
  __builtin_amdgcn_sched_barrier(0x100);
  // stage 5: t = stage-6 input (the stage-5 base is scaled by b65 h, so the
  // chains started from S2 give y + A6 + b65 k5 directly)
  hook(4, y);
  f.stream_acc(sb + 4 * BS, L, w, y, [&](int i) { return S2[i * 64]; },
               [&](int i, cd d) { t[i] = d; });
  __builtin_amdgcn_sched_barrier(0x100);
  // stage 6: y = y + Aout + c6 k6 (base scaled by c6 h, chains started from u)
  hook(5, t);
  f.stream_acc(sb + 5 * BS, L, w, t, [&](int i) { return u[i]; },
               [&](int i, cd d) { y[i] = d; });
}

// The transpose of rk_step_lds: x <- P^T x for the step matrix P of one
// Cash-Karp step of a linear system (the stages' coefficients as in
// rk_step_lds, stage bases 5 and 6 scaled by b65 and c6).  With stage
// inputs z_i and derivatives k_i = h A_i z_i the step is
// y' = y + sum c_i k_i, z_i = y + sum_{m<i} b_im k_m; its transpose, applied
// to a row x, takes the stages in reverse:
//   w_i = c_i x + sum_{m>i} b_mi zb_m,  zb_i = (h A_i)^T w_i,
//   x' = x + sum_i zb_i.
// Registers y, z and the slots S1, S2 hold the partial sums
// (A3 -> w3, A2 -> w2, A1 -> w1, L -> x'), so the step uses the same
// storage as the forward one.  stage(k, kb) sees the row kb of stage k (the
// input of its transposed derivative, registers) just before it is used.
template <class F, class Hook = NoHook>
__device__ inline void rk_step_T_lds(const double* sb, const LS& L, const WS& w, cd* x,
                                     const F& f, cd* slot, Hook&& stage = Hook()) {
  using namespace ck;
  constexpr int N = F::N, BS = F::BS;
  cd* S1 = slot;
  cd* S2 = slot + N * 64;
  cd y[N], z[N];
  // stage 6 (base scaled by c6 h): y = zb6; stage 5 (by b65 h): z = zb5
  stage(5, x);
  f.stream(sb + 5 * BS, L, w, x, [&](int i, cd d) { y[i] = d; });
  __builtin_amdgcn_sched_barrier(0x100);
  stage(4, y);
  f.stream(sb + 4 * BS, L, w, y, [&](int i, cd d) { z[i] = d; });
  __builtin_amdgcn_sched_barrier(0x100);
  // x = w4, y = A3, z = A1, S1 = A2, S2 = L
#pragma unroll
  for (int i = 0; i < N; i++) {
    cd xi = x[i], z6 = y[i], z5 = z[i];
    S2[i * 64] = xi + z5 + z6;
    S1[i * 64] = rmad(rmul(b52, z5), b62, z6);
    x[i] = rmad(rmad(rmul(c4, xi), b54, z5), b64, z6);
    y[i] = rmad(rmad(rmul(c3, xi), b53, z5), b63, z6);
    z[i] = rmad(rmad(rmul(c1, xi), b51, z5), b61, z6);
  }
  __builtin_amdgcn_sched_barrier(0x100);
  // stage 4: zb4 from w4; y = w3
  stage(3, x);
  f.stream(sb + 3 * BS, L, w, x, [&](int i, cd d) {
    y[i] = rmad(y[i], b43, d);
    z[i] = rmad(z[i], b41, d);
    S1[i * 64] = rmad(S1[i * 64], b42, d);
    S2[i * 64] = S2[i * 64] + d;
  });
  __builtin_amdgcn_sched_barrier(0x100);
  // stage 3: zb3 from w3; x = w2
  stage(2, y);
  f.stream(sb + 2 * BS, L, w, y, [&](int i, cd d) {
    x[i] = rmad(S1[i * 64], b32, d);
    z[i] = rmad(z[i], b31, d);
    S2[i * 64] = S2[i * 64] + d;
  });
  __builtin_amdgcn_sched_barrier(0x100);
  // stage 2: zb2 from w2; z = w1; L moves to y (free once stage 3 has run)
  stage(1, x);
  f.stream(sb + BS, L, w, x, [&](int i, cd d) {
    z[i] = rmad(z[i], b21, d);
    y[i] = S2[i * 64] + d;
  });
  __builtin_amdgcn_sched_barrier(0x100);
  // stage 1: x' = L + zb1 (the chains start from y)
  stage(0, z);
  f.stream_acc(sb, L, w, z, [&](int i) { return y[i]; }, [&](int i, cd d) { x[i] = d; });
}

// One knot interval of a march: the interval's equal steps from ra to rb
// (ra != rb), the step rule of yint (src/module_int.f90:1903-1907,
// 1945-1949), optional renormalisation (:1910-1919).  Every lane of the
// block must call this with identical (i, ra, rb).  Returns the number of
// steps taken.
__host__ __device__ inline double stage_offset(int k) {
  return k == 0 ? 0.0 : (k == 1 ? 0.2 : (k == 2 ? 0.3 : (k == 3 ? 0.6 : (k == 4 ? 1.0 : 0.875))));
}

// Factor on h in the base of stage k: rk_step_lds folds b65 (stage 5) and
// c6 (stage 6) into the derivative, since those stages' derivatives enter
// one combination each.
template <int N>
__host__ __device__ inline double stage_scale(int k) {
  if constexpr (N == 14) return k == 4 ? ck::b65 : (k == 5 ? ck::c6 : 1.0);
  return 1.0;
}

template <class F, class Pre = NoPre>
__device__ int march_interval(const Tab& T, const WS& w, const LS& L, int nknot, int i,
                              double ra, double rb, cd* y, bool scaled, int& scale,
                              double* sbase, const F& f, cd* slot = nullptr,
                              Pre&& pre = Pre()) {
  double drr0 = step_len_w(T, w.rew, i, nknot);
  int n = 2 + (int)floor(fabs(rb - ra) / drr0);
  double drr = (rb - ra) / (double)(n - 1);
  for (int j0 = 1; j0 <= n - 1; j0 += CH) {
    int nst = min(CH, n - j0);
    int q = threadIdx.x;
    if (q < 6 * nst) {
      int jj = j0 + q / 6, k = q % 6;
      double x = ra + (double)(jj - 1) * drr;
      F::base(T, w, i, x + stage_offset(k) * drr, drr * stage_scale<F::N>(k), sbase + q * F::BS);
    }
    __syncthreads();
    for (int jj = 0; jj < nst; jj++) {
      pre(y, scale);
      if constexpr (F::N == 14) {
        rk_step_lds(sbase + jj * 6 * F::BS, L, w, y, f, slot);
      } else {
        rk_step_b(sbase + jj * 6 * F::BS, L, w, y, f);
      }
      if (scaled) renorm<F::N>(y, scale);
    }
    __syncthreads();
  }
  return n - 1;
}

// march_interval's steps in reverse order, each applied as its transpose
// (rk_step_T_lds): the same knot interval, step count, step length and
// stage radii.  Returns the number of steps.
template <class F>
__device__ int march_interval_T(const Tab& T, const WS& w, const LS& L, int nknot, int i,
                                double ra, double rb, cd* y, int& scale, double* sbase,
                                const F& f, cd* slot) {
  static_assert(F::N == 14, "the transposed march is for the 14-component systems");
  double drr0 = step_len_w(T, w.rew, i, nknot);
  int n = 2 + (int)floor(fabs(rb - ra) / drr0);
  double drr = (rb - ra) / (double)(n - 1);
  for (int jhi = n - 1; jhi >= 1; jhi -= CH) {
    int nst = min(CH, jhi);  // steps jhi, jhi - 1, ..., jhi - nst + 1
    int q = threadIdx.x;
    if (q < 6 * nst) {
      int jj = jhi - q / 6, k = q % 6;
      double x = ra + (double)(jj - 1) * drr;
      F::base(T, w, i, x + stage_offset(k) * drr, drr * stage_scale<F::N>(k), sbase + q * F::BS);
    }
    __syncthreads();
    for (int jj = 0; jj < nst; jj++) {
      rk_step_T_lds(sbase + jj * 6 * F::BS, L, w, y, f, slot);
      renorm<F::N>(y, scale);
    }
    __syncthreads();
  }
  return n - 1;
}

__device__ inline int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

}  // namespace ysx
