// ysx_march.hpp -- lockstep radial marches for gfx950.
//
// A wavefront holds pairs that share one frequency w.  The reference's step
// rule (yint, src/module_int.f90:1903-1905) depends only on w and the knot
// interval, so every lane of the wave visits the same stage radii.  For each
// chunk of steps the wave evaluates the splines and the l-independent part
// of the coefficient matrices once per stage radius (one lane per radius)
// into LDS; each lane then adds its own l-dependent terms and runs its
// Cash-Karp step.  Control flow is uniform across the wave.
#pragma once
#include "ysx_device.hpp"

namespace ysx {

constexpr int CH = 10;       // steps per shared-coefficient chunk (6*CH radii)
constexpr int BASE_MAX = 18;  // doubles per stage radius

// w-dependent constants shared by the wave (set_omega, src/module_int.f90:2091-2102)
struct WS {
  cd omega, omega2, omegai2, lno;
  double rew, pibigg;
};
__device__ inline WS omega_state(cd w, double tref, double ats, double pibigg) {
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
__device__ inline LS l_state(int l) {
  PS s;
  set_l(s, l);
  return LS{s.lld, s.llp1, s.zeta, s.zeta2, s.zeta2m2};
}

__device__ inline double step_len_w(const Tab& T, double rew, int i, int nknot) {
  double a = T(F_ALPHA, i);
  double tmp = (i < nknot) ? fmax(a, T(F_ALPHA, i + 1)) : a;
  return 1.0 / (4.0 * rew / tmp);
}

// ------------------------------------------------------------ solid region
// sph_solid_amat (src/module_int.f90:500-560) split into an l-independent
// base per radius and the l-dependent entries added per lane.
struct SolidB {
  static constexpr int BS = 18;
  __device__ static void base(const Tab& T, const WS& w, int il, double rr, double* o) {
    SplW sw = spl_w(T, il, rr);
    double rrho = spl(T, sw, F_RHO), racon = spl(T, sw, F_ACON), rccon = spl(T, sw, F_CCON),
           rfcon = spl(T, sw, F_FCON), rncon = spl(T, sw, F_NCON), rlcon = spl(T, sw, F_LCON),
           gg = spl(T, sw, F_GRAV), rqmu = spl(T, sw, F_QMU), rxlam = spl(T, sw, F_XLAM),
           rxa2 = spl(T, sw, F_XA2);
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
    o[0] = ri; o[1] = rrho;
    o[2] = t11.x; o[3] = t11.y; o[4] = cccon.x; o[5] = cccon.y; o[6] = k22.x; o[7] = k22.y;
    o[8] = s11.x; o[9] = s11.y; o[10] = fcri.x; o[11] = fcri.y; o[12] = q12.x; o[13] = q12.y;
    o[14] = a22.x; o[15] = a22.y; o[16] = gc.x; o[17] = gc.y;
  }
  __device__ static inline void amat(const double* b, const LS& L, double pib, SolidA& A) {
    double ri = b[0], rrho = b[1];
    A.t11 = C(b[2], b[3]);
    A.k11 = C(b[4], b[5]);
    A.k22 = C(b[6], b[7]);
    A.s11 = C(b[8], b[9]);
    A.t12 = L.zeta * C(b[10], b[11]);
    A.s12 = L.zeta * C(b[12], b[13]);
    A.s22 = C(b[14], b[15]) + L.zeta2 * C(b[16], b[17]);
    A.t21 = -L.zeta * ri;
    A.t22 = 2.0 * ri;
    A.t31 = -4.0 * pib * rrho;
    A.t33 = -L.lld * ri;
    A.k33 = 4.0 * pib;
    A.s13 = -rrho * L.llp1 * ri;
    A.s23 = rrho * L.zeta * ri;
  }
};

// sph_minors_solid_derivs, src/module_int.f90:818-871
__device__ inline void solid_minor_derivs(const SolidA& A, const cd* m, cd* d) {
  double t21 = A.t21, t31 = A.t31, k33 = A.k33;
  cd a123 = A.t11 + (A.t22 + A.t33), a126 = A.t11 + (A.t22 - A.t33),
     a135 = A.t11 - (A.t22 - A.t33), a234 = (A.t22 + A.t33) - A.t11;
  d[0] = k33 * m[3] - A.k22 * m[5] + A.k11 * m[9] + m[0] * a123;
  d[1] = m[0] * A.s13 - A.k22 * m[6] - m[2] * t21 + m[1] * A.t22 - m[3] * t31;
  d[2] = m[0] * A.s23 - A.k11 * m[10] + m[2] * A.t11 - m[1] * A.t12;
  d[3] = A.k22 * m[8] - A.k11 * m[11] + m[3] * a126;
  d[4] = m[9] * A.t12 - k33 * m[7] - m[0] * A.s12 - m[5] * t21 + m[2] * t31 + m[4] * A.t33;
  d[5] = m[5] * a135 - k33 * m[8] - A.k11 * m[12] - m[0] * A.s22 - 2.0 * (m[4] * A.t12);
  d[6] = m[2] * A.s12 + m[5] * A.s13 - m[1] * A.s22 - m[4] * A.s23 + m[10] * A.t12 -
         m[6] * A.t22 + m[8] * t31;
  d[7] = m[3] * A.s12 - m[2] * A.s13 - m[1] * A.s23 + m[11] * A.t12 - m[8] * t21 - m[7] * A.t33;
  d[8] = A.k11 * m[13] + m[3] * A.s22 - 2.0 * (m[2] * A.s23) - 2.0 * (m[7] * A.t12) -
         m[8] * a234;
  d[9] = A.k22 * m[12] - k33 * m[11] + m[0] * A.s11 + 2.0 * (m[4] * t21) - 2.0 * (m[1] * t31) +
         m[9] * a234;
  d[10] = m[1] * A.s12 - m[2] * A.s11 - m[4] * A.s13 - m[9] * A.s23 - m[10] * A.t11 +
          m[6] * t21 - m[7] * t31;
  d[11] = 2.0 * (m[1] * A.s13) - A.k22 * m[13] - m[3] * A.s11 + 2.0 * (m[7] * t21) -
          m[11] * a135;
  d[12] = k33 * m[13] - m[5] * A.s11 + 2.0 * (m[4] * A.s12) + m[9] * A.s22 + 2.0 * (m[6] * t31) -
          m[12] * a126;
  d[13] = m[8] * A.s11 - 2.0 * (m[7] * A.s12) + 2.0 * (m[6] * A.s13) - m[11] * A.s22 +
          2.0 * (m[10] * A.s23) - m[13] * a123;
}

// sph_bvec_solid_derivs, src/module_int.f90:1016-1057
__device__ inline void bvec_derivs(const SolidA& A, const cd* b, cd* d) {
  double t21 = A.t21, t31 = A.t31, k33 = A.k33;
  cd t1122 = A.t11 + A.t22, t1133 = A.t11 + A.t33;
  double t2233 = A.t22 + A.t33;
  d[0] = A.s11 * b[6] + A.s12 * (b[7] - b[2]) + A.s13 * b[8] - A.s22 * b[3] - A.s23 * b[4] -
         t1122 * b[0] + t31 * b[5];
  d[1] = A.s11 * b[9] + A.s12 * b[10] - A.s13 * (2.0 * b[2] + b[7]) - A.s23 * b[3] -
         t21 * b[5] - t1133 * b[1];
  d[2] = A.s12 * b[11] + A.s13 * b[12] + A.t12 * b[3] - t21 * b[6] - t31 * b[9];
  d[3] = A.s13 * b[13] - A.k22 * b[0] - A.s11 * b[11] + t21 * (b[2] - b[7]) +
         (A.t22 - A.t11) * b[3] - t31 * b[10];
  d[4] = t31 * (2.0 * b[2] + b[7]) - k33 * b[1] - A.s11 * b[12] - A.s12 * b[13] - t21 * b[8] +
         (A.t33 - A.t11) * b[4];
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

struct SolidMinorM : SolidB {
  static constexpr int N = 14;
  __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* y,
                                    cd* d) const {
    SolidA A;
    amat(b, L, w.pibigg, A);
    solid_minor_derivs(A, y, d);
  }
};
struct BvecM : SolidB {
  static constexpr int N = 14;
  __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* y,
                                    cd* d) const {
    SolidA A;
    amat(b, L, w.pibigg, A);
    bvec_derivs(A, y, d);
  }
};

// ------------------------------------------------------------ fluid region
// sph_fluid_amat + sph_minors_fluid_derivs, src/module_int.f90:676-715,
// 941-967, split into an l-independent base and l-dependent entries.
struct FluidMinorM {
  static constexpr int N = 5, BS = 12;
  __device__ static void base(const Tab& T, const WS& w, int il, double rr, double* o) {
    SplW sw = spl_w(T, il, rr);
    double rrho = spl(T, sw, F_RHO), rccon = spl(T, sw, F_CCON), gg = spl(T, sw, F_GRAV),
           rxa2 = spl(T, sw, F_XA2);
    cd cccon = cinv(rccon * (1.0 + rxa2 * w.lno));
    double ri = 1.0 / rr, ri2 = ri * ri;
    cd fb = ri2 * w.omegai2;   // ri^2 / w^2
    cd fa = gg * fb;           // gg ri^2 / w^2
    cd s11a = (rrho * gg) * fa;
    cd s11b = (-4.0 * rrho * gg * ri) - rrho * w.omega2;
    o[0] = ri; o[1] = rrho;
    o[2] = fa.x; o[3] = fa.y; o[4] = fb.x; o[5] = fb.y; o[6] = cccon.x; o[7] = cccon.y;
    o[8] = s11a.x; o[9] = s11a.y; o[10] = s11b.x; o[11] = s11b.y;
  }
  __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* m,
                                    cd* d) const {
    double ri = b[0], rrho = b[1];
    cd fa = C(b[2], b[3]), fb = C(b[4], b[5]), cccon = C(b[6], b[7]);
    double z2 = L.zeta2;
    cd t11 = z2 * fa - ri;
    cd t12 = z2 * fb;
    double t21 = -4.0 * w.pibigg * rrho;
    double t22 = -L.lld * ri;
    cd k11 = cccon - (z2 / rrho) * fb;
    double k22 = 4.0 * w.pibigg;
    cd s11 = z2 * C(b[8], b[9]) + C(b[10], b[11]);
    cd s12 = (z2 * rrho) * fa - L.llp1 * rrho * ri;
    cd s22 = (z2 * rrho) * fb;
    d[0] = k22 * m[2] - k11 * m[3] + m[0] * (t11 + t22);
    d[1] = m[0] * s12 + m[3] * t12 - t21 * m[2];
    d[2] = k11 * m[4] + m[0] * s22 - 2.0 * (m[1] * t12) + m[2] * (t11 - t22);
    d[3] = 2.0 * t21 * m[1] - k22 * m[4] - m[0] * s11 + m[3] * (t22 - t11);
    d[4] = m[2] * s11 - 2.0 * (m[1] * s12) - m[3] * s22 - m[4] * (t11 + t22);
  }
};

// ---------------------------------------------------------------- toroidal
// tor_amat + tor_derivs, src/module_int.f90:134-184
struct TorM {
  static constexpr int N = 2, BS = 7;
  __device__ static void base(const Tab& T, const WS& w, int il, double rr, double* o) {
    SplW sw = spl_w(T, il, rr);
    double rrho = spl(T, sw, F_RHO), rlcon = spl(T, sw, F_LCON), rncon = spl(T, sw, F_NCON),
           rqmu = spl(T, sw, F_QMU);
    cd qll = 1.0 + rqmu * w.lno;
    double ri = 1.0 / rr;
    cd k = cinv(rlcon * qll);
    cd sv0 = (ri * ri * rncon) * qll;
    cd mw = -rrho * w.omega2;
    o[0] = ri; o[1] = k.x; o[2] = k.y; o[3] = sv0.x; o[4] = sv0.y; o[5] = mw.x; o[6] = mw.y;
  }
  __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* y,
                                    cd* d) const {
    double t = 2.0 * b[0];
    cd k = C(b[1], b[2]);
    cd sv = L.zeta2m2 * C(b[3], b[4]) + C(b[5], b[6]);
    d[0] = t * y[0] + k * y[1];
    d[1] = sv * y[0] - t * y[1];
  }
};

// ----------------------------------------------------- Cash-Karp, shared base
// rkck (src/module_int.f90:2016-2064) with the six stage coefficient sets
// read from the wave's LDS chunk (stride F::BS doubles per stage).
template <class F>
__device__ inline void rk_step_b(const double* sb, const LS& L, const WS& w, double h, cd* y,
                                 const F& f) {
  constexpr int N = F::N, BS = F::BS;
  constexpr double b21 = 0.2, b31 = 3.0 / 40.0, b32 = 9.0 / 40.0, b41 = 0.3, b42 = -0.9,
                   b43 = 1.2, b51 = -11.0 / 54.0, b52 = 2.5, b53 = -70.0 / 27.0,
                   b54 = 35.0 / 27.0, b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0,
                   b63 = 575.0 / 13824.0, b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0,
                   c1 = 37.0 / 378.0, c3 = 250.0 / 621.0, c4 = 125.0 / 594.0,
                   c6 = 512.0 / 1771.0;
  cd k1[N], k2[N], k3[N], t[N];
  f(sb, L, w, y, k1);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + (b21 * h) * k1[i];
  f(sb + BS, L, w, t, k2);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + h * (b31 * k1[i] + b32 * k2[i]);
  f(sb + 2 * BS, L, w, t, k3);
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + h * (b41 * k1[i] + b42 * k2[i] + b43 * k3[i]);
  {
    cd k4[N];
    f(sb + 3 * BS, L, w, t, k4);
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
  f(sb + 4 * BS, L, w, t, k3);  // k5
#pragma unroll
  for (int i = 0; i < N; i++) t[i] = y[i] + h * (k2[i] + b65 * k3[i]);
  f(sb + 5 * BS, L, w, t, k3);  // k6
#pragma unroll
  for (int i = 0; i < N; i++) y[i] = y[i] + h * (k1[i] + c6 * k3[i]);
}

// One knot interval of a march: the interval's equal steps from ra to rb
// (ra != rb), the step rule of yint (src/module_int.f90:1903-1907,
// 1945-1949), optional renormalisation (:1910-1919).  Every lane of the
// block must call this with identical (i, ra, rb).  Returns the number of
// steps taken.
template <class F>
__device__ int march_interval(const Tab& T, const WS& w, const LS& L, int nknot, int i,
                              double ra, double rb, cd* y, bool scaled, int& scale,
                              double* sbase, const F& f) {
  constexpr double a[6] = {0.0, 0.2, 0.3, 0.6, 1.0, 0.875};
  double drr0 = step_len_w(T, w.rew, i, nknot);
  int n = 2 + (int)floor(fabs(rb - ra) / drr0);
  double drr = (rb - ra) / (double)(n - 1);
  for (int j0 = 1; j0 <= n - 1; j0 += CH) {
    int nst = min(CH, n - j0);
    int q = threadIdx.x;
    if (q < 6 * nst) {
      int jj = j0 + q / 6, k = q % 6;
      double x = ra + (double)(jj - 1) * drr;
      F::base(T, w, i, x + a[k] * drr, sbase + q * F::BS);
    }
    __syncthreads();
    for (int jj = 0; jj < nst; jj++) {
      rk_step_b(sbase + jj * 6 * F::BS, L, w, drr, y, f);
      if (scaled) renorm<F::N>(y, scale);
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
