// ysx_march_ngc.hpp -- the spheroidal solve of gravity switches 0 and 1
// (sph_solver_ng, _cow, _ng_fic, _cow_fic, src/module_solver.f90:827-1200)
// as a lockstep march: a wave holds 64 consecutive degrees at one w, the
// l-independent coefficients of each stage radius are computed once per
// wave (the base of the self-gravitating solid system, with gravity zero
// for switch 0), and each lane carries the 5-minor, 2-vector fluid and
// 4-vector b systems of its degree through the same radii.
#pragma once
#include "ysx_kern.hpp"

namespace ysx {

// 5 minors of the solid region, sph_minors_solid_derivs_ng / _cow
// (src/module_int.f90:874-936), from the h-scaled base: h dm/dr
template <bool COW>
struct NgcMinorM {
  static constexpr int N = 5, BS = SolidB::BS;
  __device__ static void base(const Tab& T, const WS& w, int il, double rr, double h, double* o) {
    solid_base<COW>(T, w, il, rr, h, o);
  }
  // h A of sph_solid_amat_ng / _cow (src/module_int.f90:564-672)
  __device__ static inline void amat(const double* b, const LS& L, NgcA& A) {
    double hri = b[0];
    A.t11 = C(b[2], b[3]);
    A.k11 = C(b[4], b[5]);
    A.k22 = C(b[6], b[7]);
    A.s11 = C(b[8], b[9]);
    A.t12 = L.zeta * C(b[10], b[11]);
    A.s12 = L.zeta * C(b[12], b[13]);
    A.s22 = C(b[14], b[15]) + L.zeta2 * C(b[16], b[17]);
    A.t21 = -L.zeta * hri;
    A.t22 = 2.0 * hri;
  }
  __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* m,
                                    cd* d) const {
    NgcA A;
    amat(b, L, A);
    cd a11 = A.t11 + A.t22, a22 = A.t11 - A.t22;
    d[0] = a11 * m[0] + A.k22 * m[2] - A.k11 * m[3];
    d[1] = A.s12 * m[0] - A.t21 * m[2] + A.t12 * m[3];
    d[2] = A.s22 * m[0] - 2.0 * (A.t12 * m[1]) + a22 * m[2] + A.k11 * m[4];
    d[3] = 2.0 * A.t21 * m[1] - A.s11 * m[0] - a22 * m[3] - A.k22 * m[4];
    d[4] = A.s11 * m[2] - 2.0 * (A.s12 * m[1]) - A.s22 * m[3] - a11 * m[4];
  }
};

// the b-vector, sph_bvec_solid_derivs_ng / _cow (src/module_int.f90:1060-1110)
template <bool COW>
struct NgcBvecM {
  static constexpr int N = 4, BS = SolidB::BS;
  __device__ static void base(const Tab& T, const WS& w, int il, double rr, double h, double* o) {
    solid_base<COW>(T, w, il, rr, h, o);
  }
  __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* y,
                                    cd* d) const {
    NgcA A;
    NgcMinorM<COW>::amat(b, L, A);
    d[0] = -(A.s11 * y[2]) - A.s12 * y[3] - A.t11 * y[0] - A.t21 * y[1];
    d[1] = -(A.s12 * y[2]) - A.s22 * y[3] - A.t12 * y[0] - A.t22 * y[1];
    d[2] = A.t11 * y[2] - A.k11 * y[0] + A.t12 * y[3];
    d[3] = A.t21 * y[2] - A.k22 * y[1] + A.t22 * y[3];
  }
};

// the fluid 2-vector, sph_fluid_amat_ng / _cow + sph_minors_fluid_derivs_ng
// / _cow (src/module_int.f90:720-786, 971-1012); base times h:
// h/r, h/(r^2 rho), h/C, g/r, h rho g/r, -h rho w^2
template <bool COW>
struct NgcFluidM {
  static constexpr int N = 2, BS = 8;
  __device__ static void base(const Tab& T, const WS& w, int il, double rr, double h, double* o) {
    SplW sw = spl_w(T, il, rr);
    double rrho = spl(T, sw, F_RHO), rccon = spl(T, sw, F_CCON), rxa2 = spl(T, sw, F_XA2);
    double gg = COW ? spl(T, sw, F_GRAV) : 0.0;
    cd cccon = cinv(rccon * (1.0 + rxa2 * w.lno));
    double ri = 1.0 / rr;
    cd mw = -rrho * w.omega2;
    o[0] = h * ri;
    o[1] = h * ((ri * ri) / rrho);
    o[2] = h * cccon.x;
    o[3] = h * cccon.y;
    o[4] = gg * ri;
    o[5] = h * (rrho * gg * ri);
    o[6] = h * mw.x;
    o[7] = h * mw.y;
  }
  __device__ inline void operator()(const double* b, const LS& L, const WS& w, const cd* m,
                                    cd* d) const {
    double hri = b[0];
    cd zoi = L.zeta2 * w.omegai2;
    cd k = C(b[2], b[3]) - b[1] * zoi;
    cd t, sv;
    if (COW) {
      cd g = b[4] * zoi;
      t = (g - 1.0) * hri;
      sv = b[5] * (g - 4.0) + C(b[6], b[7]);
    } else {
      t = C(-hri);
      sv = C(b[6], b[7]);
    }
    d[0] = t * m[0] + k * m[1];
    d[1] = sv * m[0] - t * m[1];
  }
};

namespace {

// One wave: degrees lb*64 .. lb*64+63 at frequency j; sph_solver_ng / _cow
// / _ng_fic / _cow_fic with ss = identity(4) (src/module_solver.f90:
// 827-1195, 1386-1740) with every integration a lockstep march.  Writes
// the 4-vector solutions into the records in the reference driver's
// 6-vector layout (src/yspec.f90:342-374) and the step counts.
template <bool COW>
__device__ __forceinline__ void ngc_item(const Tab& T, const RunK& R, int j, int lb,
                                         cd* __restrict__ rec, int* __restrict__ cnt,
                                         double* sb) {
  if (!R.ngc || (R.grav == 1) != COW) return;
  int l = R.lmin + lb * 64 + (int)threadIdx.x;
  bool valid = (l >= 1 && l <= R.lmax);
  if (wave_min(valid ? 0 : 1)) return;  // no valid lane in this wave
  size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
  WS w = wave_omega(R, j);
  LS L = l_state(valid ? l : 1);
  const int nk = R.nknot, nic = R.nic, noc = R.noc, nsl = R.nsl;
  const bool fic = R.fic != 0, up = R.rr >= R.rs;
  const int iup = up ? R.is : R.ir, ibot = up ? R.ir : R.is;
  const double rup = up ? R.rs : R.rr, rbot = up ? R.rr : R.rs;
  // start level and starting minors of this lane
  // (src/module_solver.f90:849-890 and the _fic forms)
  double dr = 1.0 / (4.0 * w.rew / R.alpha_min);
  double rt1 = (up || fic) ? fmin(0.1 * dr, 0.1 * T(F_R, 2)) : T(F_R, 1);
  int i1 = NOLANE;
  cd st[5];
  for (int q = 0; q < 5; q++) st[q] = C(0.0);
  if (valid) {
    PS s;
    pair_state(R, l, j, s);
    i1 = sph_start_level(T, layers(R), l, s.rew, iup, R.fre_norm);
    double rt = (i1 == 1) ? rt1 : T(F_R, i1);
    if (i1 >= iup) {
      i1 = iup - 1;
      rt = T(F_R, i1);
    }
    if (i1 <= nic && !fic)
      minor_solid_start_ng(T, s, i1, rt, st);
    else if (i1 <= noc)
      fluid_start_ng(T, s, i1, rt, st);
    else if (i1 <= nsl)
      minor_solid_start_ng(T, s, i1, rt, st);
  }
  int ib = wave_min(i1);
  // ms1 from each lane's start level to the upper point
  cd m[5];
  for (int q = 0; q < 5; q++) m[q] = C(0.0);
  int layout = -1, sc1 = 0, nsol = 0, nflu = 0;
  for (int i = ib; i <= iup; i++) {
    bool fluid_i = fic ? (i < noc) : (i > nic && i < noc);
    if (layout == 0 && fluid_i) {
      cd f[2];
      solid_2_fluid_ngc(m, f);
      m[0] = f[0];
      m[1] = f[1];
      m[2] = m[3] = m[4] = C(0.0);
    } else if (layout == 1 && !fluid_i) {
      cd f[2] = {m[0], m[1]};
      fluid_2_solid_ngc(f, m);
    }
    layout = fluid_i ? 1 : 0;
    if (i == i1)
      for (int q = 0; q < 5; q++) m[q] = st[q];
    double a = (i == 1) ? rt1 : T(F_R, i);
    double b = (i == iup) ? rup : T(F_R, i + 1);
    if (a == b) continue;
    int n;
    if (fluid_i)
      n = march_interval(T, w, L, nk, i, a, b, m, true, sc1, sb, NgcFluidM<COW>());
    else
      n = march_interval(T, w, L, nk, i, a, b, m, true, sc1, sb, NgcMinorM<COW>());
    if (i >= i1) (fluid_i ? nflu : nsol) += n;
  }
  // ms3: on from the upper point to the lower one
  cd m3[5];
  for (int q = 0; q < 5; q++) m3[q] = m[q];
  int sc3 = 0;
  nsol += march_yint(T, w, L, nk, up ? R.is : R.ir, up ? R.ir : R.is, rup, rbot, m3, sc3, sb,
                     NgcMinorM<COW>());
  // ms2: from the surface down to the lower point
  cd m2[5];
  int sc2 = 0;
  for (int q = 0; q < 5; q++) m2[q] = C(0.0);
  if (R.ocean) {
    m2[0] = C(1.0);
    nflu += march_yint(T, w, L, nk, nk - 1, nsl + 1, T(F_R, nk), T(F_R, nsl + 1), m2, sc2, sb,
                       NgcFluidM<COW>());
    cd f[2] = {m2[0], m2[1]};
    fluid_2_solid_ngc(f, m2);
    nsol += march_yint(T, w, L, nk, nsl, ibot, T(F_R, nsl), rbot, m2, sc2, sb, NgcMinorM<COW>());
  } else {
    m2[0] = C(1.0);
    nsol += march_yint(T, w, L, nk, nk, ibot, T(F_R, nk), rbot, m2, sc2, sb, NgcMinorM<COW>());
  }
  cd delta = delta_ngc(m3, m2);
  cd idelta = cinv(delta);
  // the four b-vector columns from the source to the receiver
  cd* o = rec + p * REC + O_YSPH;
  int nbv = 0;
  // b starts from ms1 (up) or ms2, the solution from the other (values
  // selected, not pointers, so the arrays stay in registers)
  cd mb[5], my[5];
  for (int q = 0; q < 5; q++) {
    mb[q] = up ? m[q] : m2[q];
    my[q] = up ? m2[q] : m[q];
  }
  for (int c = 0; c < 4; c++) {
    int sc4 = 0;
    cd b[4], y[4];
    b_start_ngc(mb, c, b);
    if (!up)
      for (int q = 0; q < 4; q++) b[q] = -(b[q] * idelta);
    int n = march_yint(T, w, L, nk, R.is, R.ir, R.rs, R.rr, b, sc4, sb, NgcBvecM<COW>());
    if (c == 0) nbv = n;
    ysol_ngc(my, b, y);
    double f = ipow_huge(sc3 - sc4);
    if (up)
      for (int q = 0; q < 4; q++) y[q] = (y[q] / delta) * f;
    else
      for (int q = 0; q < 4; q++) y[q] = y[q] * f;
    if (valid) {
      o[6 * c + 0] = y[0];
      o[6 * c + 1] = y[1];
      o[6 * c + 2] = C(0.0);
      o[6 * c + 3] = y[2];
      o[6 * c + 4] = y[3];
      o[6 * c + 5] = C(0.0);
    }
  }
  if (!valid) return;
  cnt[5 * p + 1] = nsol;
  cnt[5 * p + 2] = nflu;
  cnt[5 * p + 3] = nbv;
}

// items (event, j, l-block of 64) of the events with ngc set and gravity
// switch COW ? 1 : 0
template <bool COW>
__global__ void __launch_bounds__(64) k_ngc_march(const double* __restrict__ gtab,
                                                  const RunK* __restrict__ runs, Queue Q,
                                                  cd* __restrict__ rec, int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  RUN_ITEMS(false, Q, ngc_item<COW>(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, it.z, rec,
                                    cnt, sb));
}

}  // namespace
}  // namespace ysx
