// ysx_variants.hpp -- the reference's variant systems: the matrices, start
// solutions, region transforms and solution algebra of gravity switches 0
// (no gravity) and 1 (Cowling), used by the lockstep march of
// ysx_march_ngc.hpp; and two per-lane solvers (one lane integrates one
// (l, w) pair through the Cash-Karp integrator of ysx_device.hpp): the
// self-gravitating solver with a fluid inner core for a receiver below the
// source, and the no-gravity radial solver.
#pragma once
#include "ysx_device.hpp"

namespace ysx {

// the matrix of sph_solid_amat_ng / _cow, src/module_int.f90:564-672 (built
// per lane from the shared base by NgcMinorM::amat)
struct NgcA {
  cd t11, t12, k11, k22, s11, s12, s22;
  double t21, t22;
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
template <class Grp>
__device__ inline void rad_solve_ng(const Tab& T, const PS& s, const VarGeo& G, cd* yout,
                                    int& nsteps, const Grp& grp) {
  RadRHS f;
  int nk = G.nknot;
  double dr = 1.0 / (4.0 * s.rew / G.alpha_min);
  double rt = fmin(0.1 * dr, 0.1 * T(F_R, 2));
  int sc1 = 0, sc2 = 0, sc3 = 0;
  cd y1[2], y2[2] = {C(1.0), C(0.0)}, y3[2];
  rad_ystart_ng(T, s, 1, rt, y1);
  bool up = G.rr >= G.rs;
  if (up) {
    yint(T, s, nk, 1, G.is, rt, G.rs, y1, f, &sc1, nsteps, grp);
    if (G.rr != T(F_R, nk))
      yint(T, s, nk, nk - 1, G.ir, T(F_R, nk), G.rr, y2, f, &sc2, nsteps, grp);
    y3[0] = y2[0];
    y3[1] = y2[1];
    if (G.rs != G.rr) yint(T, s, nk, G.ir, G.is, G.rr, G.rs, y3, f, &sc3, nsteps, grp);
  } else {
    yint(T, s, nk, 1, G.ir, rt, G.rr, y1, f, &sc1, nsteps, grp);
    double rsl = T(F_R, G.nsl);
    if (G.rs != rsl) yint(T, s, nk, G.nsl - 1, G.is, rsl, G.rs, y2, f, &sc2, nsteps, grp);
    y3[0] = y2[0];
    y3[1] = y2[1];
    if (G.rr != G.rs) yint(T, s, nk, G.is, G.ir, G.rs, G.rr, y3, f, &sc3, nsteps, grp);
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
