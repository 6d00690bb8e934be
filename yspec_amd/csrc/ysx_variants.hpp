// ysx_variants.hpp -- per-lane solvers of the reference's variant systems:
// gravity switch 0 (no gravity) and 1 (Cowling) with a solid or a fluid
// inner core, the self-gravitating solver with a fluid inner core, and the
// no-gravity radial solver.  One lane integrates one (l, w) pair through
// the per-lane Cash-Karp integrator of ysx_device.hpp; these paths serve
// decks and switches outside the lockstep kernels (grav = 2, solid core).
#pragma once
#include "ysx_device.hpp"

namespace ysx {

// sph_solid_amat_ng / _cow, src/module_int.f90:564-672
struct NgcA {
  cd t11, t12, k11, k22, s11, s12, s22;
  double t21, t22;
};
template <bool COW>
__device__ inline void solid_amat_ngc(const Tab& T, const PS& s, int il, double rr, NgcA& A) {
  SplW w = spl_w(T, il, rr);
  double rrho = spl(T, w, F_RHO), racon = spl(T, w, F_ACON), rccon = spl(T, w, F_CCON),
         rfcon = spl(T, w, F_FCON), rncon = spl(T, w, F_NCON), rlcon = spl(T, w, F_LCON),
         rqmu = spl(T, w, F_QMU), rxlam = spl(T, w, F_XLAM), rxa2 = spl(T, w, F_XA2);
  double gg = COW ? spl(T, w, F_GRAV) : 0.0;
  cd qaa = 1.0 + rxa2 * s.lno, qff = 1.0 + rxlam * s.lno, qll = 1.0 + rqmu * s.lno;
  cd cacon = racon * qaa, cfcon = rfcon * qff, cncon = rncon * qll;
  cd cccon = cinv(rccon * qaa);
  cd fc = cfcon * cccon;
  cd gamma = cacon - cncon - cfcon * fc;
  double ri = 1.0 / rr, ri2 = ri * ri;
  A.t11 = (1.0 - 2.0 * fc) * ri;
  A.t12 = (s.zeta * ri) * fc;
  A.t21 = -s.zeta * ri;
  A.t22 = 2.0 * ri;
  A.k11 = cccon;
  A.k22 = cinv(rlcon * qll);
  cd mw2r = -rrho * s.omega2;
  if (COW) {
    double rgr = rrho * gg * rr;
    A.s11 = mw2r + (4.0 * ri2) * (gamma - rgr);
    A.s12 = (s.zeta * ri2) * (rgr - 2.0 * gamma);
  } else {
    A.s11 = mw2r + (4.0 * ri2) * gamma;
    A.s12 = (-2.0 * s.zeta * ri2) * gamma;
  }
  A.s22 = mw2r + ri2 * (s.zeta2 * (gamma + cncon) - 2.0 * cncon);
}

// sph_minors_solid_derivs_ng / _cow, src/module_int.f90:874-936 (5 minors)
template <bool COW>
struct SolidMinorNgc {
  static constexpr int N = 5;
  __device__ inline void operator()(const Tab& T, const PS& s, int il, double rr, const cd* m,
                                    cd* d) const {
    NgcA A;
    solid_amat_ngc<COW>(T, s, il, rr, A);
    cd a11 = A.t11 + A.t22, a22 = A.t11 - A.t22;
    d[0] = a11 * m[0] + A.k22 * m[2] - A.k11 * m[3];
    d[1] = A.s12 * m[0] - A.t21 * m[2] + A.t12 * m[3];
    d[2] = A.s22 * m[0] - 2.0 * (A.t12 * m[1]) + a22 * m[2] + A.k11 * m[4];
    d[3] = 2.0 * A.t21 * m[1] - A.s11 * m[0] - a22 * m[3] - A.k22 * m[4];
    d[4] = A.s11 * m[2] - 2.0 * (A.s12 * m[1]) - A.s22 * m[3] - a11 * m[4];
  }
};

// sph_fluid_amat_ng / _cow + sph_minors_fluid_derivs_ng / _cow,
// src/module_int.f90:720-786, 971-1012 (a 2-vector)
template <bool COW>
struct FluidNgc {
  static constexpr int N = 2;
  __device__ inline void operator()(const Tab& T, const PS& s, int il, double rr, const cd* m,
                                    cd* d) const {
    SplW w = spl_w(T, il, rr);
    double rrho = spl(T, w, F_RHO), rccon = spl(T, w, F_CCON), rxa2 = spl(T, w, F_XA2);
    cd cccon = cinv(rccon * (1.0 + rxa2 * s.lno));
    double ri = 1.0 / rr;
    cd zoi = s.zeta2 * s.omegai2;
    cd k = cccon - ((ri * ri) / rrho) * zoi;
    cd t, sv;
    if (COW) {
      double gg = spl(T, w, F_GRAV);
      cd g = (gg * ri) * zoi;
      t = (g - 1.0) * ri;
      sv = (rrho * gg * ri) * (g - 4.0) - rrho * s.omega2;
    } else {
      t = C(-ri);
      sv = -rrho * s.omega2;
    }
    d[0] = t * m[0] + k * m[1];
    d[1] = sv * m[0] - t * m[1];
  }
};

// sph_bvec_solid_derivs_ng / _cow, src/module_int.f90:1060-1110
template <bool COW>
struct BvecNgc {
  static constexpr int N = 4;
  __device__ inline void operator()(const Tab& T, const PS& s, int il, double rr, const cd* b,
                                    cd* d) const {
    NgcA A;
    solid_amat_ngc<COW>(T, s, il, rr, A);
    d[0] = -(A.s11 * b[2]) - A.s12 * b[3] - A.t11 * b[0] - A.t21 * b[1];
    d[1] = -(A.s12 * b[2]) - A.s22 * b[3] - A.t12 * b[0] - A.t22 * b[1];
    d[2] = A.t11 * b[2] - A.k11 * b[0] + A.t12 * b[3];
    d[3] = A.t21 * b[2] - A.k22 * b[1] + A.t22 * b[3];
  }
};

// sph_solid_start_ng, src/module_int.f90:1275-1341, with the sbessj2 stub
// ratio = x / real(2 (l+1) + 3) (src/module_function.f90:447-490)
__device__ inline void solid_start_ng(const Tab& T, const PS& s, int is, double rr,
                                      cd y[2][4]) {
  SplW w = spl_w(T, is, rr);
  double rrho = spl(T, w, F_RHO), rmu = spl(T, w, F_MU), rqmu = spl(T, w, F_QMU),
         rkappa = spl(T, w, F_KAPPA), rqkappa = spl(T, w, F_QKAPPA);
  int ll = s.ll, llp1 = ll + 1;
  double den = (double)(float)(2 * llp1 + 3);
  cd cmu = rmu * (1.0 + rqmu * s.lno);
  cd ckappa = rkappa * (1.0 + rqkappa * s.lno);
  cd calpha = (ckappa + 1.3333333333 * cmu) / rrho;
  cd cbeta = cmu / rrho;
  cd arg = s.omega * rr / csqrt_(calpha);
  cd j2 = arg / den;
  y[0][0] = (double)ll - arg * j2;
  y[0][1] = C(s.zeta);
  y[0][2] = (-rrho * calpha * arg * arg +
             2.0 * cmu * ((double)(ll * (ll - 1)) + 2.0 * arg * j2)) / rr;
  y[0][3] = (2.0 * s.zeta * cmu * ((double)(ll - 1) - arg * j2)) / rr;
  arg = s.omega * rr / csqrt_(cbeta);
  j2 = arg / den;
  y[1][0] = C((double)(-ll * llp1));
  y[1][1] = arg * j2 - s.zeta * (double)llp1;
  y[1][2] = (-2.0 * cmu * ((double)(ll * (ll * ll - 1)) - (double)(ll * llp1) * arg * j2)) / rr;
  y[1][3] = (s.zeta * cmu * (arg * arg - 2.0 * (double)(ll * ll - 1) - 2.0 * arg * j2)) / rr;
}

// sph_fluid_start_ng, src/module_int.f90:1407-1447
__device__ inline void fluid_start_ng(const Tab& T, const PS& s, int is, double rr, cd* y) {
  SplW w = spl_w(T, is, rr);
  double rrho = spl(T, w, F_RHO), rkappa = spl(T, w, F_KAPPA), rqkappa = spl(T, w, F_QKAPPA);
  cd calpha = rkappa * (1.0 + rqkappa * s.lno) / rrho;
  cd arg = s.omega * rr / csqrt_(calpha);
  cd ratio = arg / (double)(float)(2 * (s.ll + 1) + 3);
  y[0] = s.lld - arg * ratio;
  y[1] = (-rrho * calpha * arg * arg) / rr;
}

// sph_minor_solid_start_ng, src/module_int.f90:1489-1506
__device__ inline void minor_solid_start_ng(const Tab& T, const PS& s, int is, double rr,
                                            cd* m) {
  cd y[2][4];
  solid_start_ng(T, s, is, rr, y);
  const int pr[5][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {2, 3}};
#pragma unroll
  for (int k = 0; k < 5; k++)
    m[k] = det2(y[0][pr[k][0]], y[0][pr[k][1]], y[1][pr[k][0]], y[1][pr[k][1]]);
}

// b_solid_start_ngcow, src/module_int.f90:1596-1615, for the unit source
// vector e_c
__device__ inline void b_start_ngc(const cd* m, int c, cd* b) {
  // e_c's entries as values, not an array indexed by c (which would live
  // in scratch memory)
  const cd sv0 = C(c == 0 ? 1.0 : 0.0), sv1 = C(c == 1 ? 1.0 : 0.0),
           sv2 = C(c == 2 ? 1.0 : 0.0), sv3 = C(c == 3 ? 1.0 : 0.0);
  b[0] = -(m[4] * sv1) - m[1] * sv2 - m[3] * sv3;
  b[1] = m[4] * sv0 - m[2] * sv2 + m[1] * sv3;
  b[2] = m[1] * sv0 + m[2] * sv1 - m[0] * sv3;
  b[3] = m[3] * sv0 - m[1] * sv1 + m[0] * sv2;
}

// sph_minors_*_2_*_ngcow, sph_delta_solid_ngcow, sph_ysol_ngcow
// (src/module_int.f90:1640-1845)
__device__ inline void solid_2_fluid_ngc(const cd* m, cd* f) {
  f[0] = -m[2];
  f[1] = -m[4];
}
__device__ inline void fluid_2_solid_ngc(const cd* f, cd* m) {
  m[0] = f[0];
  m[3] = -f[1];
  m[1] = m[2] = m[4] = C(0.0);
}
__device__ inline cd delta_ngc(const cd* a, const cd* b) {
  return a[4] * b[0] + 2.0 * (a[1] * b[1]) + a[3] * b[2] + a[2] * b[3] + a[0] * b[4];
}
__device__ inline void ysol_ngc(const cd* m, const cd* b, cd* y) {
  y[0] = m[0] * b[1] + m[1] * b[2] + m[2] * b[3];
  y[1] = m[3] * b[2] - m[0] * b[0] - m[1] * b[3];
  y[2] = m[4] * b[3] - m[1] * b[0] - m[3] * b[1];
  y[3] = m[1] * b[1] - m[2] * b[0] - m[4] * b[2];
}

// Geometry of one solve
struct VarGeo {
  int nknot, nic, noc, nsl, ocean, is, ir;
  double rs, rr, alpha_min, fre_norm;
};

struct VarCount {
  int sol, flu, bvec;
};

// sph_solver_ng / _cow (fic = false) and sph_solver_ng_fic / _cow_fic
// (fic = true) with ss = identity(4), src/module_solver.f90:827-1195,
// 1386-1740: yout (4,4) column-major
template <bool COW>
__device__ void sph_solve_ngc(const Tab& T, const PS& s, const VarGeo& G, bool fic, cd* yout,
                              VarCount& n) {
  SolidMinorNgc<COW> fs;
  FluidNgc<COW> ff;
  BvecNgc<COW> fb;
  Layers L{G.nknot, G.nic, G.noc, G.nsl, fic ? 1 : 0, G.ocean};
  bool up = G.rr >= G.rs;
  int iup = up ? G.is : G.ir, ibot = up ? G.ir : G.is;
  double rup = up ? G.rs : G.rr, rbot = up ? G.rr : G.rs;
  int nk = G.nknot;
  int sc1 = 0, sc2 = 0, sc3 = 0;
  cd mf1[2], mf2[2], ms1[5], ms2[5], ms3[5];
  int i1 = sph_start_level(T, L, s.ll, s.rew, iup, G.fre_norm);
  double rt = T(F_R, i1);
  if ((up || fic) && i1 == 1) {
    double dr = 1.0 / (4.0 * s.rew / G.alpha_min);
    rt = fmin(0.1 * dr, 0.1 * T(F_R, 2));
  }
  if (i1 >= iup) {
    i1 = iup - 1;
    rt = T(F_R, i1);
  }
  for (int q = 0; q < 5; q++) ms1[q] = C(0.0);
  if (i1 <= G.nic && !fic) {
    minor_solid_start_ng(T, s, i1, rt, ms1);
    yint(T, s, nk, i1, G.nic, rt, T(F_R, G.nic), ms1, fs, &sc1, n.sol);
    solid_2_fluid_ngc(ms1, mf1);
    yint(T, s, nk, G.nic + 1, G.noc, T(F_R, G.nic), T(F_R, G.noc), mf1, ff, &sc1, n.flu);
    fluid_2_solid_ngc(mf1, ms1);
    yint(T, s, nk, G.noc + 1, iup, T(F_R, G.noc + 1), rup, ms1, fs, &sc1, n.sol);
  } else if (i1 <= G.noc) {
    fluid_start_ng(T, s, i1, rt, mf1);
    yint(T, s, nk, i1, G.noc, rt, T(F_R, G.noc), mf1, ff, &sc1, n.flu);
    fluid_2_solid_ngc(mf1, ms1);
    yint(T, s, nk, G.noc + 1, iup, T(F_R, G.noc + 1), rup, ms1, fs, &sc1, n.sol);
  } else if (i1 <= G.nsl) {
    minor_solid_start_ng(T, s, i1, rt, ms1);
    yint(T, s, nk, i1, iup, rt, rup, ms1, fs, &sc1, n.sol);
  }
  for (int q = 0; q < 5; q++) ms3[q] = ms1[q];
  if (up && G.rr > G.rs) yint(T, s, nk, G.is, G.ir, G.rs, G.rr, ms3, fs, &sc3, n.sol);
  if (!up && G.rs > G.rr) yint(T, s, nk, G.ir, G.is, G.rr, G.rs, ms3, fs, &sc3, n.sol);
  if (G.ocean) {
    mf2[0] = C(1.0);
    mf2[1] = C(0.0);
    yint(T, s, nk, nk - 1, G.nsl + 1, T(F_R, nk), T(F_R, G.nsl + 1), mf2, ff, &sc2, n.flu);
    fluid_2_solid_ngc(mf2, ms2);
    yint(T, s, nk, G.nsl, ibot, T(F_R, G.nsl), rbot, ms2, fs, &sc2, n.sol);
  } else {
    ms2[0] = C(1.0);
    for (int q = 1; q < 5; q++) ms2[q] = C(0.0);
    yint(T, s, nk, nk, ibot, T(F_R, nk), rbot, ms2, fs, &sc2, n.sol);
  }
  cd delta = delta_ngc(ms3, ms2);
  cd idelta = cinv(delta);
  for (int c = 0; c < 4; c++) {
    int sc4 = 0, nb = 0;
    cd b[4];
    cd* y = yout + 4 * c;
    if (up) {
      b_start_ngc(ms1, c, b);
      yint(T, s, nk, G.is, G.ir, G.rs, G.rr, b, fb, &sc4, nb);
      ysol_ngc(ms2, b, y);
      double f = ipow_huge(sc3 - sc4);
      for (int q = 0; q < 4; q++) y[q] = (y[q] / delta) * f;
    } else {
      b_start_ngc(ms2, c, b);
      for (int q = 0; q < 4; q++) b[q] = -(b[q] * idelta);
      yint(T, s, nk, G.is, G.ir, G.rs, G.rr, b, fb, &sc4, nb);
      ysol_ngc(ms1, b, y);
      double f = ipow_huge(sc3 - sc4);
      for (int q = 0; q < 4; q++) y[q] = y[q] * f;
    }
    if (c == 0) n.bvec = nb;
  }
}

// sph_solver_fic (grav = 2, fluid inner core) with ss = s6tmp,
// src/module_solver.f90:1203-1382: yout (6,4) column-major.  Below a
// source the reference starts a solid minor vector in the (fluid) inner
// core as sph_solver does (:1323-1331); that is reproduced as written.
__device__ void sph_solve_fic(const Tab& T, const PS& s, const VarGeo& G, cd* yout,
                              VarCount& n) {
  SolidMinorRHS fs;
  FluidMinorRHS ff;
  BvecRHS fb;
  Layers L{G.nknot, G.nic, G.noc, G.nsl, 1, G.ocean};
  bool up = G.rr >= G.rs;
  int iup = up ? G.is : G.ir, ibot = up ? G.ir : G.is;
  double rup = up ? G.rs : G.rr, rbot = up ? G.rr : G.rs;
  int nk = G.nknot;
  int sc1 = 0, sc2 = 0, sc3 = 0;
  cd m1[14], m2[14], m3[14], mf[5];
  int i1 = sph_start_level(T, L, s.ll, s.rew, iup, G.fre_norm);
  double rt = T(F_R, i1);
  if (i1 == 1) {
    double dr = 1.0 / (4.0 * s.rew / G.alpha_min);
    rt = fmin(0.1 * dr, 0.1 * T(F_R, 2));
  }
  if (i1 >= iup) {
    i1 = iup - 1;
    rt = T(F_R, i1);
  }
  for (int q = 0; q < 14; q++) m1[q] = C(0.0);
  if (i1 <= G.nic && !up) {
    minor_solid_start(T, s, i1, rt, m1);
    yint(T, s, nk, i1, G.nic, rt, T(F_R, G.nic), m1, fs, &sc1, n.sol);
    solid_2_fluid(m1, mf);
    yint(T, s, nk, G.nic + 1, G.noc, T(F_R, G.nic), T(F_R, G.noc), mf, ff, &sc1, n.flu);
    fluid_2_solid(mf, m1);
    yint(T, s, nk, G.noc + 1, iup, T(F_R, G.noc + 1), rup, m1, fs, &sc1, n.sol);
  } else if (i1 <= G.noc) {
    minor_fluid_start(T, s, i1, rt, m1);
    for (int q = 0; q < 5; q++) mf[q] = m1[q];
    yint(T, s, nk, i1, G.noc, rt, T(F_R, G.noc), mf, ff, &sc1, n.flu);
    fluid_2_solid(mf, m1);
    yint(T, s, nk, G.noc + 1, iup, T(F_R, G.noc + 1), rup, m1, fs, &sc1, n.sol);
  } else if (i1 <= G.nsl) {
    minor_solid_start(T, s, i1, rt, m1);
    yint(T, s, nk, i1, iup, rt, rup, m1, fs, &sc1, n.sol);
  }
  for (int q = 0; q < 14; q++) m3[q] = m1[q];
  if (up && G.rr > G.rs) yint(T, s, nk, G.is, G.ir, G.rs, G.rr, m3, fs, &sc3, n.sol);
  if (!up && G.rs > G.rr) yint(T, s, nk, G.ir, G.is, G.rr, G.rs, m3, fs, &sc3, n.sol);
  if (G.ocean) {
    cd f2[5] = {C(1.0), C(0.0), C(0.0), C(0.0), C(0.0)};
    yint(T, s, nk, nk - 1, G.nsl + 1, T(F_R, nk), T(F_R, G.nsl + 1), f2, ff, &sc2, n.flu);
    fluid_2_solid(f2, m2);
    yint(T, s, nk, G.nsl, ibot, T(F_R, G.nsl), rbot, m2, fs, &sc2, n.sol);
  } else {
    m2[0] = C(1.0);
    for (int q = 1; q < 14; q++) m2[q] = C(0.0);
    yint(T, s, nk, up ? nk - 1 : nk, ibot, T(F_R, nk), rbot, m2, fs, &sc2, n.sol);
  }
  cd delta = delta_solid(m3, m2);
  cd mdi = -cinv(delta);
  const int cols[4] = {0, 1, 3, 4};
  for (int c = 0; c < 4; c++) {
    int sc4 = 0, nb = 0;
    cd b[14];
    cd* y = yout + 6 * c;
    if (up) {
      b_start_unit(m1, cols[c], b);
      yint(T, s, nk, G.is, G.ir, G.rs, G.rr, b, fb, &sc4, nb);
      ysol(m2, b, y);
      double f = ipow_huge(sc3 - sc4);
      for (int q = 0; q < 6; q++) y[q] = (-(y[q] / delta)) * f;
    } else {
      b_start_unit(m2, cols[c], b);
      for (int q = 0; q < 14; q++) b[q] = b[q] * mdi;
      yint(T, s, nk, G.is, G.ir, G.rs, G.rr, b, fb, &sc4, nb);
      ysol(m1, b, y);
      double f = ipow_huge(sc3 - sc4);
      for (int q = 0; q < 6; q++) y[q] = y[q] * f;
    }
    if (c == 0) n.bvec = nb;
  }
}

// rad_ystart_ng, src/module_int.f90:455-489
__device__ inline void rad_ystart_ng(const Tab& T, const PS& s, int is, double rr, cd* y) {
  SplW w = spl_w(T, is, rr);
  double rrho = spl(T, w, F_RHO), rmu = spl(T, w, F_MU), rqmu = spl(T, w, F_QMU),
         rkappa = spl(T, w, F_KAPPA), rqkappa = spl(T, w, F_QKAPPA);
  cd cmu = rmu * (1.0 + rqmu * s.lno);
  cd ckappa = rkappa * (1.0 + rqkappa * s.lno);
  cd calpha = ckappa + 4.0 * cmu / 3.0;
  cd gamma = rrho * s.omega2 / calpha;
  cd arg = gamma * rr;
  cd ratio = arg / 5.0;
  y[0] = rr * ratio;
  y[1] = rr * calpha * gamma - 4.0 * cmu * ratio;
}

// rad_solver_ng with ss = identity(2), src/module_solver.f90:335-441.  Its
// RHS rad_derivs_ng (src/module_int.f90:377-392) evaluates rad_amat, the
// self-gravitating coefficients; reproduced as written.  yout (2,2).
__device__ inline void rad_solve_ng(const Tab& T, const PS& s, const VarGeo& G, cd* yout,
                                    int& nsteps) {
  RadRHS f;
  int nk = G.nknot;
  double dr = 1.0 / (4.0 * s.rew / G.alpha_min);
  double rt = fmin(0.1 * dr, 0.1 * T(F_R, 2));
  int sc1 = 0, sc2 = 0, sc3 = 0;
  cd y1[2], y2[2] = {C(1.0), C(0.0)}, y3[2];
  rad_ystart_ng(T, s, 1, rt, y1);
  bool up = G.rr >= G.rs;
  if (up) {
    yint(T, s, nk, 1, G.is, rt, G.rs, y1, f, &sc1, nsteps);
    if (G.rr != T(F_R, nk)) yint(T, s, nk, nk - 1, G.ir, T(F_R, nk), G.rr, y2, f, &sc2, nsteps);
    y3[0] = y2[0];
    y3[1] = y2[1];
    if (G.rs != G.rr) yint(T, s, nk, G.ir, G.is, G.rr, G.rs, y3, f, &sc3, nsteps);
  } else {
    yint(T, s, nk, 1, G.ir, rt, G.rr, y1, f, &sc1, nsteps);
    double rsl = T(F_R, G.nsl);
    if (G.rs != rsl) yint(T, s, nk, G.nsl - 1, G.is, rsl, G.rs, y2, f, &sc2, nsteps);
    y3[0] = y2[0];
    y3[1] = y2[1];
    if (G.rr != G.rs) yint(T, s, nk, G.is, G.ir, G.rs, G.rr, y3, f, &sc3, nsteps);
  }
  cd delta = y3[1] * y1[0] - y3[0] * y1[1];
  delta = (cabs_(delta) < 1.0e-100) ? C(0.0) : cinv(delta);
  for (int c = 0; c < 2; c++) {
    cd b, ya, yb;
    if (up) {
      b = (c == 0) ? y1[1] : -y1[0];
      ya = y2[0];
      yb = y2[1];
    } else {
      b = (c == 0) ? y2[1] : -y2[0];
      ya = y1[0];
      yb = y1[1];
    }
    yout[2 * c] = -delta * ya * b;
    yout[2 * c + 1] = -delta * yb * b;
  }
  // rr >= rs: times huge**scale3; rr < rs: scale3 multiplications by huge
  // (none for scale3 <= 0), src/module_solver.f90:390-392, 433-437
  if (up) {
    if (sc3 != 0)
      for (int q = 0; q < 4; q++) yout[q] = yout[q] * ipow_huge(sc3);
  } else {
    for (int k = 1; k <= sc3; k++)
      for (int q = 0; q < 4; q++) yout[q] = yout[q] * HUGE_;
  }
}

}  // namespace ysx
