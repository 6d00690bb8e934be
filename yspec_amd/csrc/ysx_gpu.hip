// ysx_gpu.hip -- gfx950 kernels of the spectral path and the C-ABI of
// libysx.so (include/ysx.h).
//
// Per run the GPU work is:
//   k_source      source vectors of every (l, w) pair
//   k_start       start levels and starting solutions
//   k_tor_march   toroidal two-sided shooting, lockstep waves of 64 l at one w
//   k_minor_top   spheroidal minors from the receiver side down to the source
//   k_minor_march spheroidal minor vectors from the start level up to the
//                 source, secular determinant, b-vector starts
//   k_bvec_march  b-vector from the source to the receiver, 4 columns
//   k_rad         radial (l = 0) solves on the reference's grid (2nd stream)
//   k_assemble    output-type scaling and the uu/vv/ww coefficients
//   k_lsum        sum over l and m against the Legendre tables: FP64 MFMA
//                 GEMMs with the azimuthal phases fused into the epilogue
//   k_seis        filtered inverse FFT, exp(eps t), units, rotation
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ysx.h"
#include "ysx_device.hpp"
#include "ysx_march.hpp"
#include "ysx_host.hpp"

using namespace ysx;

namespace {

// ----------------------------------------------------------------- run data
struct RunK {
  int lmin, lmax, nw, nr, is, ir, out, cor, npair;
  int nknot, nic, noc, nsl, fic, ocean;
  double rs, rr, wmin, dw, ep, tref, ats, pibigg, fre_norm, alpha_min, grav_ir;
  double mm[6];
};

__device__ inline Layers layers(const RunK& R) {
  return Layers{R.nknot, R.nic, R.noc, R.nsl, R.fic, R.ocean};
}

__device__ inline void pair_lj(const RunK& R, int p, int& l, int& j) {
  l = R.lmin + p / R.nw;
  j = p % R.nw + 1;
}

__device__ inline void pair_state(const RunK& R, int l, int j, PS& s) {
  set_l(s, l);
  s.pibigg = R.pibigg;
  cd w = C(R.wmin + (double)(j - 1) * R.dw, -R.ep);
  set_omega(s, w, R.tref, R.ats);
}

// copy the model table into LDS; every thread of the block must call it
__device__ inline Tab load_tab(const double* __restrict__ g, int nknot) {
  extern __shared__ double smem[];
  int n = F_COUNT * nknot;
  for (int i = threadIdx.x; i < n; i += blockDim.x) smem[i] = g[i];
  __syncthreads();
  return Tab{smem, nknot};
}

// per-pair output records (AoS, one record of 89 complex numbers per pair)
constexpr int REC = 89;
constexpr int O_SVT = 0, O_YTOR = 10, O_SVS = 14, O_YSPH = 44, O_SR0 = 68, O_YRAD = 70,
              O_UU = 74, O_VV = 79, O_WW = 84;

// -------------------------------------------------------------- k_source
// source_vector_tor / _sph / _rad, src/module_solver.f90:145-215,
// 566-627, 1747-1847.  source_vector_sph keeps only the real part of its
// dispersion factors (they are declared real(dp) there, :1771-1772).
__global__ void __launch_bounds__(256) k_source(const double* __restrict__ gtab, RunK R,
                                                cd* __restrict__ rec, int* __restrict__ cnt) {
  Tab T = load_tab(gtab, R.nknot);
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= R.npair) return;
  int l, j;
  pair_lj(R, p, l, j);
  PS s;
  pair_state(R, l, j, s);
  cd* o = rec + (size_t)p * REC;
  for (int i = 0; i < REC; i++) o[i] = C(0.0);
  for (int i = 0; i < 5; i++) cnt[5 * p + i] = 0;
  SplW w = spl_w(T, R.is, R.rs);
  double rlcon = spl(T, w, F_LCON), rccon = spl(T, w, F_CCON), rfcon = spl(T, w, F_FCON),
         rqmu = spl(T, w, F_QMU), rxlam = spl(T, w, F_XLAM), rxa2 = spl(T, w, F_XA2);
  const double* mm = R.mm;
  double rs = R.rs;
  cd den = cinv(cmuli(s.omega) * (rs * rs));  // 1/(i w rs^2)
  cd nd = s.nu * den;
  if (l == 0) {
    cd qaa = 1.0 + rxa2 * s.lno, qff = 1.0 + rxlam * s.lno;
    cd cc = cinv(rccon * qaa);
    cd cf = rfcon * qff;
    o[O_SR0 + 0] = nd * ((rs * mm[0]) * cc);
    o[O_SR0 + 1] = nd * ((2.0 * mm[0]) * (cf * cc) - (mm[3] + mm[5]));
    return;
  }
  if (l > 1) {
    cd cl = rlcon * (1.0 + rqmu * s.lno);
    cd icl = cinv(cl);
    o[O_SVT + 1] = nd * ((0.5 * rs) * C(-mm[2], mm[1]) * icl);
    o[O_SVT + 3] = nd * ((0.5 * rs) * C(mm[2], mm[1]) * icl);
    o[O_SVT + 5 + 0] = nd * ((0.25 * s.szeta2m2) * C(2.0 * mm[4], mm[5] - mm[3]));
    o[O_SVT + 5 + 4] = nd * ((0.25 * s.szeta2m2) * C(2.0 * mm[4], mm[3] - mm[5]));
  }
  double qaa = 1.0 + rxa2 * s.lno.x, qff = 1.0 + rxlam * s.lno.x, qll = 1.0 + rqmu * s.lno.x;
  double cc = 1.0 / (rccon * qaa), cf = rfcon * qff, cl = rlcon * qll;
  double zeta = s.zeta;
  cd* sv = o + O_SVS;  // svs(5,6) column-major
  sv[1 + 5 * 1] = nd * (C(0.5 * rs * mm[1] / cl, 0.5 * rs * mm[2] / cl));
  sv[2 + 5 * 0] = nd * C(rs * mm[0] * cc);
  sv[2 + 5 * 3] = nd * C(-mm[3] - mm[5] + 2.0 * mm[0] * cf * cc);
  sv[2 + 5 * 4] = nd * C(0.5 * zeta * (mm[3] + mm[5]) - zeta * mm[0] * cf * cc);
  sv[3 + 5 * 1] = nd * (C(-0.5 * rs * mm[1] / cl, 0.5 * rs * mm[2] / cl));
  if (l > 1) {
    sv[0 + 5 * 4] = nd * ((0.25 * s.szeta2m2) * C(mm[5] - mm[3], -2.0 * mm[4]));
    sv[4 + 5 * 4] = nd * ((0.25 * s.szeta2m2) * C(mm[5] - mm[3], 2.0 * mm[4]));
  }
}

// ----------------------------------------------------------------- k_start
// Start levels and starting solutions of every pair (one lane per pair):
// tor_solver's start (src/module_solver.f90:64-75, 103-114) and
// sph_solver's start (:678-722, 758-785).  The marches pick these up when
// they reach each lane's start knot.
struct StartData {
  int* tor_i1;   // [npair] adjusted start knot
  cd* tor_y0;    // [npair][2]
  int* sph_i1;   // [npair] adjusted start knot
  cd* sph_m0;    // [npair][14] solid minors, or 5 fluid minors + zeros
};

__global__ void __launch_bounds__(256) k_start(const double* __restrict__ gtab, RunK R,
                                               StartData SD, int* __restrict__ starts) {
  Tab T{gtab, R.nknot};
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= R.npair) return;
  int l, j;
  pair_lj(R, p, l, j);
  if (l < 1) return;
  PS s;
  pair_state(R, l, j, s);
  Layers L = layers(R);
  bool up = R.rr >= R.rs;
  int itop = up ? R.is : R.ir;
  if (l >= 2) {
    int i1 = tor_start_level(T, L, l, s.rew, itop);
    starts[2 * p] = i1;
    if (i1 == itop && i1 > L.noc + 1) i1--;
    cd y[2];
    if (i1 == L.noc + 1) {
      y[0] = C(1.0);
      y[1] = C(0.0);
    } else {
      tor_ystart(T, s, i1, T(F_R, i1), y);
    }
    SD.tor_i1[p] = i1;
    SD.tor_y0[2 * p] = y[0];
    SD.tor_y0[2 * p + 1] = y[1];
  }
  int i1 = sph_start_level(T, L, l, s.rew, itop, R.fre_norm);
  starts[2 * p + 1] = i1;
  double rt = T(F_R, i1);
  if (up && i1 == 1) {
    double dr = 1.0 / (4.0 * s.rew / R.alpha_min);
    rt = fmin(0.1 * dr, 0.1 * T(F_R, 2));
  }
  if (i1 >= itop) {
    i1 = itop - 1;
    rt = T(F_R, i1);
  }
  cd m[14];
  for (int i = 0; i < 14; i++) m[i] = C(0.0);
  if (i1 > L.nic && i1 <= L.noc)
    minor_fluid_start(T, s, i1, rt, m);
  else
    minor_solid_start(T, s, i1, rt, m);
  SD.sph_i1[p] = i1;
  for (int i = 0; i < 14; i++) SD.sph_m0[(size_t)p * 14 + i] = m[i];
}

// ------------------------------------------------------- march helpers
// yint's interval loop (src/module_int.f90:1888-1963) over knot intervals,
// with wave-uniform arguments
template <class F>
__device__ int march_yint(const Tab& T, const WS& w, const LS& L, int nk, int i1, int i2,
                          double r1, double r2, cd* y, int& sc, double* sb, const F& f,
                          cd* slot = nullptr) {
  int n = 0;
  if (r2 > r1) {
    for (int i = i1; i <= i2; i++) {
      double a = (i == i1) ? r1 : T(F_R, i);
      double b = (i == i2) ? r2 : T(F_R, i + 1);
      if (a == b) continue;
      n += march_interval(T, w, L, nk, i, a, b, y, true, sc, sb, f, slot);
    }
  } else if (r2 < r1) {
    for (int i = i1; i >= i2; i--) {
      double a = (i == i1) ? r1 : T(F_R, i + 1);
      double b = (i == i2) ? r2 : T(F_R, i);
      if (a == b) continue;
      n += march_interval(T, w, L, nk, i, a, b, y, true, sc, sb, f, slot);
    }
  }
  return n;
}

__device__ inline WS wave_omega(const RunK& R, int j) {
  return omega_state(C(R.wmin + (double)(j - 1) * R.dw, -R.ep), R.tref, R.ats, R.pibigg);
}

constexpr int NOLANE = 1 << 30;

// ------------------------------------------------------------- k_tor_march
// tor_solver with ss = identity(2), src/module_solver.f90:25-142; a wave
// holds 64 consecutive l at one w, largest w first.
__global__ void __launch_bounds__(64) k_tor_march(const double* __restrict__ gtab, RunK R,
                                                  StartData SD, cd* __restrict__ rec,
                                                  int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  Tab T{gtab, R.nknot};
  int nl = R.lmax - R.lmin + 1, nlb = (nl + 63) / 64;
  int wv = blockIdx.x;
  int j = R.nw - wv / nlb;
  int l = R.lmin + (wv % nlb) * 64 + (int)threadIdx.x;
  bool valid = (l >= 2 && l <= R.lmax);
  size_t p = (size_t)(l - R.lmin) * R.nw + (j - 1);
  int i1 = valid ? SD.tor_i1[p] : NOLANE;
  int ib = wave_min(i1);
  if (ib == NOLANE) return;
  WS w = wave_omega(R, j);
  LS L = l_state(valid ? l : 2);
  TorM f;
  int nk = R.nknot, is = R.is, ir = R.ir;
  double rs = R.rs, rr = R.rr;
  bool up = rr >= rs;
  int itop = up ? is : ir;
  double rtop = up ? rs : rr;
  cd y1[2] = {C(0.0), C(0.0)};
  int sc1 = 0, sc2 = 0, sc3 = 0, ns = 0;
  for (int i = ib; i <= itop; i++) {
    if (i == i1) {
      y1[0] = SD.tor_y0[2 * p];
      y1[1] = SD.tor_y0[2 * p + 1];
    }
    double a = T(F_R, i), b = (i == itop) ? rtop : T(F_R, i + 1);
    if (a == b) continue;
    int n = march_interval(T, w, L, nk, i, a, b, y1, true, sc1, sb, f);
    if (i >= i1) ns += n;
  }
  double rsl = T(F_R, R.nsl);
  cd y2[2] = {C(1.0), C(0.0)}, y3[2], yo[4];
  if (up) {
    y3[0] = y1[0];
    y3[1] = y1[1];
    if (rr != rs) ns += march_yint(T, w, L, nk, is, ir, rs, rr, y3, sc3, sb, f);
    if (rr != rsl) ns += march_yint(T, w, L, nk, R.nsl - 1, ir, rsl, rr, y2, sc2, sb, f);
    cd delta = cinv(y2[1] * y3[0] - y2[0] * y3[1]);
    cd b1 = y1[1], b2 = -y1[0];
    yo[0] = -delta * y2[0] * b1;
    yo[1] = -delta * y2[1] * b1;
    yo[2] = -delta * y2[0] * b2;
    yo[3] = -delta * y2[1] * b2;
  } else {
    if (rs != rsl) ns += march_yint(T, w, L, nk, R.nsl - 1, is, rsl, rs, y2, sc2, sb, f);
    y3[0] = y2[0];
    y3[1] = y2[1];
    if (rr != rs) ns += march_yint(T, w, L, nk, is, ir, rs, rr, y3, sc3, sb, f);
    cd delta = cinv(y3[1] * y1[0] - y3[0] * y1[1]);
    cd b1 = y2[1], b2 = -y2[0];
    yo[0] = -delta * y1[0] * b1;
    yo[1] = -delta * y1[1] * b1;
    yo[2] = -delta * y1[0] * b2;
    yo[3] = -delta * y1[1] * b2;
  }
  if (!valid) return;
  double sc = ipow_huge(sc3);
  cnt[5 * p + 0] = ns;
  cd* o = rec + p * REC + O_YTOR;
  for (int i = 0; i < 4; i++) o[i] = yo[i] * sc;
}

// ----------------------------------------------------------- k_minor_march
// sph_solver (grav = 2, solid inner core) up to the b-vector integration,
// src/module_solver.f90:640-743, 756-806.  Per pair it leaves the four
// b-vector starts (bst), the minor vector of sph_ysol (mfin), the factor
// applied to sph_ysol (fac) and scale3.
struct SphState {
  cd* bst;   // [npair][4][14]
  cd* mfin;  // [npair][14]
  cd* m2;    // [npair][14] minor vector from the surface
  cd* fac;   // [npair]
  int* sc3;  // [npair]
};

// minor vector from the surface down to radius rb in knot interval ib
// (through the ocean as a fluid when there is one), src/module_solver.f90:730-741
__device__ __forceinline__ int minors_from_top(const Tab& T, const WS& w, const LS& L, const RunK& R,
                                      int itopk, int ib, double rb, cd* m, double* sb,
                                      int& nsol, int& nflu, cd* slot) {
  int nk = R.nknot, sc2 = 0;
  for (int i = 0; i < 14; i++) m[i] = C(0.0);
  m[0] = C(1.0);
  if (R.ocean) {
    nflu += march_yint(T, w, L, nk, nk - 1, R.nsl + 1, T(F_R, nk), T(F_R, R.nsl + 1), m, sc2, sb,
                       FluidMinorM());
    cd mf[5] = {m[0], m[1], m[2], m[3], m[4]};
    fluid_2_solid(mf, m);
    nsol += march_yint(T, w, L, nk, R.nsl, ib, T(F_R, R.nsl), rb, m, sc2, sb, SolidMinorM(), slot);
  } else {
    nsol += march_yint(T, w, L, nk, itopk, ib, T(F_R, nk), rb, m, sc2, sb, SolidMinorM(), slot);
  }
  return sc2;
}

__global__ void __launch_bounds__(64) k_minor_march(const double* __restrict__ gtab, RunK R,
                                                    StartData SD, SphState S,
                                                    int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  __shared__ cd slots[2 * 14 * 64];
  cd* slot = slots + threadIdx.x;
  Tab T{gtab, R.nknot};
  int nl = R.lmax - R.lmin + 1, nlb = (nl + 63) / 64;
  int wv = blockIdx.x;
  int j = R.nw - wv / nlb;
  int l = R.lmin + (wv % nlb) * 64 + (int)threadIdx.x;
  bool valid = (l >= 1 && l <= R.lmax);
  size_t p = (size_t)(l - R.lmin) * R.nw + (j - 1);
  int i1 = valid ? SD.sph_i1[p] : NOLANE;
  int ib = wave_min(i1);
  if (ib == NOLANE) return;
  WS w = wave_omega(R, j);
  LS L = l_state(valid ? l : 1);
  int nk = R.nknot, is = R.is, ir = R.ir, nic = R.nic, noc = R.noc;
  double rs = R.rs, rr = R.rr;
  bool up = rr >= rs;
  int itop = up ? is : ir;
  double rtop = up ? rs : rr;
  int nsol = 0, nflu = 0;
  cd m[14];
  // (2) ms1 from each lane's start level up to the lower point
  double rt1 = up ? fmin(0.1 * (1.0 / (4.0 * w.rew / R.alpha_min)), 0.1 * T(F_R, 2)) : T(F_R, 1);
  for (int i = 0; i < 14; i++) m[i] = C(0.0);
  int layout = -1, sc1 = 0;  // -1 none, 0 solid, 1 fluid
  for (int i = ib; i <= itop; i++) {
    int region = (i < nic) ? 0 : ((i > nic && i < noc) ? 1 : 0);
    bool fluid_i = (i > nic && i < noc);
    if (layout == 0 && fluid_i) {
      cd mf[5];
      solid_2_fluid(m, mf);
      for (int q = 0; q < 14; q++) m[q] = q < 5 ? mf[q] : C(0.0);
    } else if (layout == 1 && !fluid_i) {
      cd mf[5] = {m[0], m[1], m[2], m[3], m[4]};
      fluid_2_solid(mf, m);
    }
    layout = region;
    if (i == i1)
      for (int q = 0; q < 14; q++) m[q] = SD.sph_m0[p * 14 + q];
    double a = (i == 1) ? rt1 : T(F_R, i);
    double b = (i == itop) ? rtop : T(F_R, i + 1);
    if (a == b) continue;
    int n;
    if (fluid_i)
      n = march_interval(T, w, L, nk, i, a, b, m, true, sc1, sb, FluidMinorM());
    else
      n = march_interval(T, w, L, nk, i, a, b, m, true, sc1, sb, SolidMinorM(), slot);
    if (i >= i1) (fluid_i ? nflu : nsol) += n;
  }
  const int cols[4] = {0, 1, 3, 4};
  int sc3 = 0;
  if (up) {
    if (valid)
      for (int c = 0; c < 4; c++) {
        cd bb[14];
        b_start_unit(m, cols[c], bb);
        for (int q = 0; q < 14; q++) S.bst[p * 56 + c * 14 + q] = bb[q];
      }
    if (rr > rs) nsol += march_yint(T, w, L, nk, is, ir, rs, rr, m, sc3, sb, SolidMinorM(), slot);
  } else {
    if (valid)
      for (int q = 0; q < 14; q++) S.mfin[p * 14 + q] = m[q];
    if (rs > rr) nsol += march_yint(T, w, L, nk, ir, is, rr, rs, m, sc3, sb, SolidMinorM(), slot);
  }
  if (!valid) return;
  cd m2[14];
  for (int q = 0; q < 14; q++) m2[q] = S.m2[p * 14 + q];
  cd delta = delta_solid(m, m2);
  if (up) {
    for (int q = 0; q < 14; q++) S.mfin[p * 14 + q] = m2[q];
    S.fac[p] = -cinv(delta);
  } else {
    cd mdi = -cinv(delta);
    for (int c = 0; c < 4; c++) {
      cd bb[14];
      b_start_unit(m2, cols[c], bb);
      for (int q = 0; q < 14; q++) S.bst[p * 56 + c * 14 + q] = bb[q] * mdi;
    }
    S.fac[p] = C(1.0);
  }
  S.sc3[p] = sc3;
  cnt[5 * p + 1] += nsol;
  cnt[5 * p + 2] += nflu;
}

// The surface minor vector ms2 of every pair (src/module_solver.f90:730-741,
// 793-804): from the surface down to the receiver (receiver above the
// source) or to the source.  Zero steps for a receiver at the surface.
__global__ void __launch_bounds__(64) k_minor_top(const double* __restrict__ gtab, RunK R,
                                                  SphState S, int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  __shared__ cd slots[2 * 14 * 64];
  cd* slot = slots + threadIdx.x;
  Tab T{gtab, R.nknot};
  int nl = R.lmax - R.lmin + 1, nlb = (nl + 63) / 64;
  int wv = blockIdx.x;
  int j = R.nw - wv / nlb;
  int l = R.lmin + (wv % nlb) * 64 + (int)threadIdx.x;
  bool valid = (l >= 1 && l <= R.lmax);
  if (wave_min(valid ? 0 : 1)) return;
  size_t p = (size_t)(l - R.lmin) * R.nw + (j - 1);
  WS w = wave_omega(R, j);
  LS L = l_state(valid ? l : 1);
  int nsol = 0, nflu = 0;
  cd m[14];
  if (R.rr >= R.rs)
    minors_from_top(T, w, L, R, R.nknot - 1, R.ir, R.rr, m, sb, nsol, nflu, slot);
  else
    minors_from_top(T, w, L, R, R.nknot, R.is, R.rs, m, sb, nsol, nflu, slot);
  if (!valid) return;
  for (int i = 0; i < 14; i++) S.m2[p * 14 + i] = m[i];
  cnt[5 * p + 1] = nsol;
  cnt[5 * p + 2] = nflu;
}

// ------------------------------------------------------------ k_bvec_march
// b-vector from source to receiver and y = sph_ysol(m, b) for the four
// source columns, src/module_solver.f90:745-755, 809-819.  A wave holds 16
// consecutive l x 4 columns at one w.
__global__ void __launch_bounds__(64) k_bvec_march(const double* __restrict__ gtab, RunK R,
                                                   SphState S, cd* __restrict__ rec,
                                                   int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  __shared__ cd slots[2 * 14 * 64];
  cd* slot = slots + threadIdx.x;
  Tab T{gtab, R.nknot};
  int nl = R.lmax - R.lmin + 1, nlb = (nl + 15) / 16;
  int wv = blockIdx.x;
  int j = R.nw - wv / nlb;
  int c = threadIdx.x & 3;
  int l = R.lmin + (wv % nlb) * 16 + (int)(threadIdx.x >> 2);
  bool valid = (l >= 1 && l <= R.lmax);
  int any = wave_min(valid ? 0 : 1);
  if (any) return;  // no valid lane in this wave
  size_t p = (size_t)(l - R.lmin) * R.nw + (j - 1);
  WS w = wave_omega(R, j);
  LS L = l_state(valid ? l : 1);
  cd b[14];
  for (int q = 0; q < 14; q++) b[q] = valid ? S.bst[p * 56 + c * 14 + q] : C(0.0);
  int sc4 = 0;
  int n = march_yint(T, w, L, R.nknot, R.is, R.ir, R.rs, R.rr, b, sc4, sb, BvecM(), slot);
  if (!valid) return;
  cd m[14];
  for (int q = 0; q < 14; q++) m[q] = S.mfin[p * 14 + q];
  cd y[6];
  ysol(m, b, y);
  cd fac = S.fac[p] * ipow_huge(S.sc3[p] - sc4);
  cd* o = rec + p * REC + O_YSPH + 6 * c;
  for (int q = 0; q < 6; q++) o[q] = y[q] * fac;
  if (c == 0) cnt[5 * p + 3] = n;
}

// ------------------------------------------------------------------- k_rad
// rad_solver_wg with ss = identity(2), src/module_solver.f90:446-561, on
// the grid of make_grid, src/module_model.f90:638-787.  Only yout is
// formed; the caller discards pout (src/yspec.f90:277).
struct Run3 {
  double base[3], step[3];
  int j0[3], j1[3], n;
  int src, rec;  // run whose last point is the source / receiver (-1: none)
};

__device__ inline void grid_runs(const Tab& T, int i, double drr, double rs, double rr, Run3& G) {
  double r1 = T(F_R, i), r2 = T(F_R, i + 1);
  G.n = 0;
  G.src = G.rec = -1;
  if (r1 == r2) return;
  bool ls = (rs > r1 && rs <= r2), lr = (rr > r1 && rr <= r2);
  auto add = [&](double a, double b, int j0) {
    int n = (int)floor((b - a) / drr) + 2;
    G.base[G.n] = a;
    G.step[G.n] = (b - a) / (double)(n - 1);
    G.j0[G.n] = j0;
    G.j1[G.n] = n;
    G.n++;
  };
  if (ls && lr) {
    if (rs < rr) {
      add(r1, rs, 1); G.src = 0;
      add(rs, rr, 2); G.rec = 1;
      add(rr, r2, 2);
    } else if (rr < rs) {
      add(r1, rr, 1); G.rec = 0;
      add(rr, rs, 2); G.src = 1;
      add(rs, r2, 2);
    } else {
      add(r1, rs, 1); G.src = 0; G.rec = 0;
      add(rs, r2, 2);
    }
  } else if (ls) {
    add(r1, rs, 1); G.src = 0;
    add(rs, r2, 2);
  } else if (lr) {
    add(r1, rr, 1); G.rec = 0;
    add(rr, r2, 2);
  } else {
    add(r1, r2, 1);
  }
}

__global__ void __launch_bounds__(64) k_rad(const double* __restrict__ gtab, RunK R,
                                            cd* __restrict__ rec, int* __restrict__ cnt) {
  Tab T = load_tab(gtab, R.nknot);
  int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (j > R.nw || R.lmin != 0) return;
  int p = j - 1;  // l = 0 pairs come first
  PS s;
  pair_state(R, 0, j, s);
  RadRHS f;
  int nk = R.nknot;
  double dr = 1.0 / (4.0 * s.rew / R.alpha_min);
  double rs = R.rs, rr = R.rr;
  int nsr = 0;
  // --- upward sweep over grid points 1 .. source
  cd y1[2], y1r[2] = {C(0.0), C(0.0)};
  bool rec_below = false;  // receiver at or below the source point
  {
    int k = 0;
    double rprev = 0.0;
    int iprev = 1;
    bool done = false;
    for (int i = 1; i <= nk - 1 && !done; i++) {
      Run3 G;
      grid_runs(T, i, dr, rs, rr, G);
      for (int q = 0; q < G.n && !done; q++) {
        for (int jj = G.j0[q]; jj <= G.j1[q]; jj++) {
          double rg = G.base[q] + (double)(jj - 1) * G.step[q];
          k++;
          if (k == 2) {
            double rt = fmin(0.1 * dr, 0.1 * rg);
            rad_ystart(T, s, 1, rt, y1);
            yint(T, s, nk, 1, i, rt, rg, y1, f, (int*)nullptr, nsr);
          } else if (k > 2) {
            yint(T, s, nk, iprev, i, rprev, rg, y1, f, (int*)nullptr, nsr);
          }
          rprev = rg;
          iprev = i;
          bool last = (jj == G.j1[q]);
          if (last && q == G.rec) {
            rec_below = true;
            y1r[0] = y1[0];
            y1r[1] = y1[1];
          }
          if (last && q == G.src) {
            done = true;
            break;
          }
        }
      }
    }
  }
  // --- downward sweep from the top grid point to the source point
  cd y2[2] = {C(1.0), C(0.0)}, y2r[2] = {C(1.0), C(0.0)};
  {
    bool first = true, done = false;
    double rprev = 0.0;
    int iprev = nk - 1;
    for (int i = nk - 1; i >= 1 && !done; i--) {
      Run3 G;
      grid_runs(T, i, dr, rs, rr, G);
      for (int q = G.n - 1; q >= 0 && !done; q--) {
        for (int jj = G.j1[q]; jj >= G.j0[q]; jj--) {
          double rg = G.base[q] + (double)(jj - 1) * G.step[q];
          if (!first) yint(T, s, nk, iprev, i, rprev, rg, y2, f, (int*)nullptr, nsr);
          first = false;
          rprev = rg;
          iprev = i;
          bool last = (jj == G.j1[q]);
          if (last && q == G.src) {
            done = true;
            break;
          }
          if (last && q == G.rec && !rec_below) {
            y2r[0] = y2[0];
            y2r[1] = y2[1];
          }
        }
      }
    }
  }
  cd delta = cinv(y2[1] * y1[0] - y2[0] * y1[1]);
  cd yo[4];
  if (rec_below) {
    cd b1 = y2[1], b2 = -y2[0];
    yo[0] = -delta * y1r[0] * b1;
    yo[1] = -delta * y1r[1] * b1;
    yo[2] = -delta * y1r[0] * b2;
    yo[3] = -delta * y1r[1] * b2;
  } else {
    cd b1 = y1[1], b2 = -y1[0];
    yo[0] = -delta * y2r[0] * b1;
    yo[1] = -delta * y2r[1] * b1;
    yo[2] = -delta * y2r[0] * b2;
    yo[3] = -delta * y2r[1] * b2;
  }
  cd* o = rec + (size_t)p * REC + O_YRAD;
  for (int i = 0; i < 4; i++) o[i] = yo[i];
  cnt[5 * p + 4] = nsr;
}

// ------------------------------------------------------------- k_assemble
// output-type scaling and the uu/vv/ww coefficients, src/yspec.f90:280-288,
// 309-316, 377-400 (zeta there is single precision, :255)
struct GemmK {
  int lpad, jpad, nrpad;
  cd* cf;          // [15][lpad][jpad]: uu m=-2..2, vv m=-2..2, ww m=-2..2
  const double* tb;  // [9][lpad][nrpad]: x, dx/dtheta, x/sin(theta) for |m| = 0,1,2
};

__device__ inline void put_coefs(const RunK& R, const GemmK& G, int l, int j, const cd* o) {
  size_t plane = (size_t)G.lpad * G.jpad;
  size_t at = (size_t)l * G.jpad + (j - 1);
  for (int c = 0; c < 15; c++) G.cf[c * plane + at] = o[O_UU + c];
}

__global__ void __launch_bounds__(256) k_assemble(RunK R, cd* __restrict__ rec, GemmK G) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= R.npair) return;
  int l, j;
  pair_lj(R, p, l, j);
  cd* o = rec + (size_t)p * REC;
  cd w = C(R.wmin + (double)(j - 1) * R.dw, -R.ep);
  cd fo = (R.out == 1) ? cmuli(w) : (R.out == 2 ? -(w * w) : C(1.0));
  double rr = R.rr;
  for (int m = 0; m < 15; m++) o[O_UU + m] = C(0.0);
  if (l == 0) {
    cd u3 = (o[O_SR0] * o[O_YRAD + 0] + o[O_SR0 + 1] * o[O_YRAD + 2]) / rr;
    o[O_UU + 2] = fo * u3;
    put_coefs(R, G, l, j, o);
    return;
  }
  double zeta = (double)sqrtf((float)(l * (l + 1)));
  if (l > 1) {
    cd t1 = fo * o[O_YTOR + 0], t2 = fo * o[O_YTOR + 2];
    for (int m = 0; m < 5; m++)
      o[O_WW + m] = (o[O_SVT + m] * t1 + o[O_SVT + 5 + m] * t2) / (zeta * rr);
  }
  cd y1[4], y2[4];
  for (int c = 0; c < 4; c++) {
    cd a = o[O_YSPH + 6 * c + 0], b = o[O_YSPH + 6 * c + 1], g3 = o[O_YSPH + 6 * c + 2];
    if (R.out == 2 && R.cor != 0) {
      cd c1 = -((double)(l + 1) * g3 + (2.0 * R.grav_ir) * a) / rr;
      cd c2 = zeta * (g3 + R.grav_ir * a) / rr;
      y1[c] = fo * a + c1;
      y2[c] = fo * b + c2;
    } else {
      y1[c] = fo * a;
      y2[c] = fo * b;
    }
  }
  const cd* sv = o + O_SVS;
  for (int m = 0; m < 5; m++) {
    cd su = y1[0] * sv[m] + y1[1] * sv[5 + m] + y1[2] * sv[15 + m] + y1[3] * sv[20 + m];
    cd sw = y2[0] * sv[m] + y2[1] * sv[5 + m] + y2[2] * sv[15 + m] + y2[3] * sv[20 + m];
    o[O_UU + m] = su / rr;
    o[O_VV + m] = sw / (zeta * rr);
  }
  put_coefs(R, G, l, j, o);
}

// ---------------------------------------------------------------- k_lsum
// The receiver sum (src/yspec.f90:290-293, 318-335, 402-419) as FP64 MFMA
// GEMMs: for each Legendre table X in {x, dx/dtheta, x/sin(theta)} and
// |m| = a, C[j][k] = sum_l coef(l, w_j) X_a(l, theta_k) over the coefficient
// sets that table multiplies; the epilogue applies (-1)^a for m < 0, the
// factors i*m and exp(i m phi_k) and accumulates ur, ut, up.  One wave owns a
// 16 (w) x 16 (receiver) tile; v_mfma_f64_16x16x4f64 lane layout: A[i][k] at
// lane i + 16k, B[k][n] at lane n + 16k, D[i][n] at lane n + 16(i%4), reg i/4.
typedef double d4 __attribute__((ext_vector_type(4)));

// one (table, |m|) group: GEMM over l, then the epilogue for this group
template <int X, int A, int NC, int C0, int C1, int C2, int C3>
__device__ __forceinline__ void lsum_group(const GemmK& G, int j0, int k0, int li, int lq,
                                           cd ep, cd en, cd* ur, cd* ut, cd* up) {
  constexpr int cs[4] = {C0, C1, C2, C3};
  const size_t plane = (size_t)G.lpad * G.jpad;
  d4 acc[NC][2];
#pragma unroll
  for (int q = 0; q < NC; q++) acc[q][0] = acc[q][1] = (d4){0.0, 0.0, 0.0, 0.0};
  const double* tb = G.tb + (size_t)(X * 3 + A) * G.lpad * G.nrpad + k0 + li;
  const cd* cf = G.cf + (size_t)lq * G.jpad + j0 + li;
  for (int l0 = 0; l0 < G.lpad; l0 += 4) {
    double b = tb[(size_t)(l0 + lq) * G.nrpad];
    size_t at = (size_t)l0 * G.jpad;
    cd av[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) av[q] = cf[cs[q] * plane + at];
#pragma unroll
    for (int q = 0; q < NC; q++) {
      acc[q][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q].x, b, acc[q][0], 0, 0, 0);
      acc[q][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q].y, b, acc[q][1], 0, 0, 0);
    }
  }
  constexpr double sg = (A & 1) ? -1.0 : 1.0;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    cd P = C(acc[0][0][v], acc[0][1][v]);
    if constexpr (A == 0) {
      if constexpr (X == 0) ur[v] = ur[v] + P;
      else ut[v] = ut[v] + P;  // vv_0 dx/dtheta
    } else {
      cd Mn = C(acc[1][0][v], acc[1][1][v]);
      cd pos = ep * P, neg = sg * (en * Mn);
      if constexpr (X == 0) {
        ur[v] = ur[v] + pos + neg;
        cd Wp = C(acc[NC - 2][0][v], acc[NC - 2][1][v]), Wn = C(acc[NC - 1][0][v], acc[NC - 1][1][v]);
        ut[v] = ut[v] + cmuli((double)A * (ep * Wp - sg * (en * Wn)));
      } else if constexpr (X == 1) {
        ut[v] = ut[v] + pos + neg;
        cd Wp = C(acc[NC - 2][0][v], acc[NC - 2][1][v]), Wn = C(acc[NC - 1][0][v], acc[NC - 1][1][v]);
        up[v] = up[v] - (ep * Wp + sg * (en * Wn));
      } else {
        up[v] = up[v] + cmuli((double)A * (pos - neg));
      }
    }
  }
}

__global__ void __launch_bounds__(64) k_lsum(RunK R, GemmK G, const cd* __restrict__ eia,
                                             cd* __restrict__ spec) {
  int lane = threadIdx.x;
  int njt = G.jpad / 16;
  int j0 = (blockIdx.x % njt) * 16, k0 = (blockIdx.x / njt) * 16;
  int li = lane & 15, lq = lane >> 4;
  int k = k0 + li;  // receiver of this lane's output column
  cd e1 = (k < R.nr) ? eia[k] : C(1.0), e2 = e1 * e1;
  cd em1 = C(e1.x, -e1.y), em2 = C(e2.x, -e2.y);
  cd ur[4], ut[4], up[4];
  for (int v = 0; v < 4; v++) ur[v] = ut[v] = up[v] = C(0.0);
  // coefficient index c: uu m -> m+2, vv m -> 7+m, ww m -> 12+m
  lsum_group<0, 0, 1, 2, 0, 0, 0>(G, j0, k0, li, lq, e1, em1, ur, ut, up);     // uu0 x
  lsum_group<0, 1, 4, 3, 1, 13, 11>(G, j0, k0, li, lq, e1, em1, ur, ut, up);   // uu+-1, ww+-1 x
  lsum_group<0, 2, 4, 4, 0, 14, 10>(G, j0, k0, li, lq, e2, em2, ur, ut, up);   // uu+-2, ww+-2 x
  lsum_group<1, 0, 1, 7, 0, 0, 0>(G, j0, k0, li, lq, e1, em1, ur, ut, up);     // vv0 dx
  lsum_group<1, 1, 4, 8, 6, 13, 11>(G, j0, k0, li, lq, e1, em1, ur, ut, up);   // vv+-1, ww+-1 dx
  lsum_group<1, 2, 4, 9, 5, 14, 10>(G, j0, k0, li, lq, e2, em2, ur, ut, up);   // vv+-2, ww+-2 dx
  lsum_group<2, 1, 2, 8, 6, 0, 0>(G, j0, k0, li, lq, e1, em1, ur, ut, up);     // vv+-1 x/sin
  lsum_group<2, 2, 2, 9, 5, 0, 0>(G, j0, k0, li, lq, e2, em2, ur, ut, up);     // vv+-2 x/sin
  if (k >= R.nr) return;
  size_t nwr = (size_t)R.nw * R.nr;
  for (int v = 0; v < 4; v++) {
    int j = j0 + lq + 4 * v;
    if (j >= R.nw) continue;
    size_t at = (size_t)j + (size_t)R.nw * k;
    spec[at] = ur[v];
    spec[nwr + at] = ut[v];
    spec[2 * nwr + at] = up[v];
  }
}

// ------------------------------------------------------------------ k_seis
// ifft (src/module_fourier.f90:55-145) and src/yspec.f90:439-473 for one
// receiver per block: pack bins i1..i2, cosine filters, Hermitian fill,
// radix-2 DIT inverse FFT with the reference's twiddle angles
// pi_d*m/mmax (single-rounded pi), /(dt*nt), exp(eps t), units, rotation.
struct SeisK {
  int nt, lognt, nw, nr, i1, i2, out;
  double df, dt, ep, f11, f12, f21, f22, norm;
};

__device__ inline cd packed_bin(const SeisK& K, const cd* spec, int k, int i) {
  // 0-based FFT bin i of the packed, filtered, Hermitian-filled row
  bool neg = i > K.nt / 2;
  int b = neg ? K.nt - i : i;
  cd v = C(0.0);
  if (b >= K.i1 - 1 && b <= K.i2 - 1) v = spec[(size_t)(b - (K.i1 - 1)) + (size_t)K.nw * k];
  double w = b * K.df, f = 1.0;
  bool nof = (K.f11 == 0.0 && K.f12 == 0.0 && K.f21 == 0.0 && K.f22 == 0.0);
  if (!nof) {
    if (w < K.f11) f = 0.0;
    else if (w < K.f12) f = 0.5 * (1.0 - cos(PI_D * (w - K.f11) / (K.f12 - K.f11)));
    else if (w < K.f21) f = 1.0;
    else if (w < K.f22) f = 0.5 * (1.0 - cos(PI_D * (K.f22 - w) / (K.f22 - K.f21)));
    else f = 0.0;
  }
  v = v * f;
  if (neg) v.y = -v.y;
  return v;
}

__device__ inline unsigned brev(unsigned x, int bits) { return __brev(x) >> (32 - bits); }

__global__ void __launch_bounds__(1024) k_seis(SeisK K, const cd* __restrict__ spec,
                                               const cd* __restrict__ tw, cd* __restrict__ work,
                                               const double* __restrict__ azep,
                                               double* __restrict__ seis) {
  int k = blockIdx.x;
  int nt = K.nt;
  for (int c = 0; c < 3; c++) {
    const cd* sp = spec + (size_t)c * K.nw * K.nr;
    cd* row = work + ((size_t)k * 3 + c) * nt;
    for (int q = threadIdx.x; q < nt; q += blockDim.x)
      row[q] = packed_bin(K, sp, k, (int)brev((unsigned)q, K.lognt));
    __syncthreads();
    for (int mmax = 1; mmax < nt; mmax <<= 1) {
      for (int b = threadIdx.x; b < nt / 2; b += blockDim.x) {
        int m = b & (mmax - 1);
        int i = ((b - m) << 1) + m;
        int jx = i + mmax;
        cd temp = tw[mmax + m] * row[jx];
        cd a = row[i];
        row[jx] = a - temp;
        row[i] = a + temp;
      }
      __syncthreads();
    }
  }
  double ca = cos(azep[k]), sa = sin(azep[k]);
  const cd* r0 = work + ((size_t)k * 3 + 0) * nt;
  const cd* r1 = work + ((size_t)k * 3 + 1) * nt;
  const cd* r2 = work + ((size_t)k * 3 + 2) * nt;
  double sc = K.norm / (K.dt * nt);
  for (int i = threadIdx.x; i < nt; i += blockDim.x) {
    double g = exp(K.ep * (i * K.dt)) * sc;
    double z = r0[i].x * g, ut = r1[i].x * g, up = r2[i].x * g;
    seis[(size_t)i + (size_t)nt * (0 + 3 * (size_t)k)] = z;
    seis[(size_t)i + (size_t)nt * (1 + 3 * (size_t)k)] = -ut * ca + up * sa;
    seis[(size_t)i + (size_t)nt * (2 + 3 * (size_t)k)] = ut * sa + up * ca;
  }
}

__global__ void k_twiddles(int nt, cd* tw) {
  // tw[mmax + m] = exp(i pi_d m / mmax), mmax = 1, 2, 4, ..., nt/2
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 1 || t >= nt) return;
  int mmax = 1 << (31 - __clz(t));
  int m = t - mmax;
  double th = PI_D * (double)m / (double)mmax;
  double sn, cs;
  sincos(th, &sn, &cs);
  tw[t] = C(cs, sn);
}

}  // namespace

// ====================================================================== API
struct ysx_ctx {
  int device = 0;
  hipStream_t stream = nullptr, stream2 = nullptr;
  std::string err;
  Norms N;
  HostModel M;
  HostRun R;
  bool have_model = false, have_run = false, have_spec = false;
  size_t lds_bytes = 0, lds_max = 0;
  // device buffers
  double* d_tab = nullptr;
  cd* d_rec = nullptr;
  int* d_starts = nullptr;
  int* d_cnt = nullptr;
  cd *d_bst = nullptr, *d_mfin = nullptr, *d_m2 = nullptr, *d_fac = nullptr;
  int *d_tor_i1 = nullptr, *d_sph_i1 = nullptr;
  cd *d_tor_y0 = nullptr, *d_sph_m0 = nullptr;
  int* d_sc3 = nullptr;
  double *d_xa = nullptr, *d_xpa = nullptr, *d_xca = nullptr, *d_azep = nullptr;
  cd* d_eia = nullptr;
  cd* d_spec = nullptr;
  cd* d_cf = nullptr;
  double* d_tb = nullptr;
  int lpad = 0, jpad = 0, nrpad = 0;
  cd* d_tw = nullptr;
  cd* d_work = nullptr;
  double* d_seis = nullptr;
  hipEvent_t ev[16] = {};
  double ms[8] = {0};
  RunK rk;
};

namespace {

int set_err(ysx_ctx* c, int code, const std::string& m) {
  c->err = m;
  return code;
}

#define HIPCHK(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return set_err(ctx, YSX_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void free_run(ysx_ctx* c) {
  dfree(c->d_rec); dfree(c->d_starts); dfree(c->d_cnt); dfree(c->d_bst); dfree(c->d_mfin);
  dfree(c->d_m2); dfree(c->d_tor_i1); dfree(c->d_sph_i1); dfree(c->d_tor_y0); dfree(c->d_sph_m0);
  dfree(c->d_fac); dfree(c->d_sc3); dfree(c->d_xa); dfree(c->d_xpa); dfree(c->d_xca);
  dfree(c->d_azep); dfree(c->d_eia); dfree(c->d_spec); dfree(c->d_cf); dfree(c->d_tb); dfree(c->d_tw); dfree(c->d_work);
  dfree(c->d_seis);
  c->have_run = c->have_spec = false;
}

}  // namespace

extern "C" int ysx_create(ysx_ctx** out, int device) {
  if (!out) return YSX_ERR_ARG;
  ysx_ctx* ctx = new ysx_ctx();
  *out = ctx;
  ctx->device = device;
  ctx->N = make_norms();
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
  for (auto& e : ctx->ev) HIPCHK(hipEventCreate(&e));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  ctx->lds_max = prop.sharedMemPerBlock;
  return YSX_OK;
}

extern "C" void ysx_destroy(ysx_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  free_run(ctx);
  dfree(ctx->d_tab);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  delete ctx;
}

extern "C" const char* ysx_last_error(const ysx_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int ysx_set_model(ysx_ctx* ctx, int nknot, int nic, int noc, int ians,
                             double tref_period, const double* deck) {
  if (!ctx || !deck) return YSX_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  free_run(ctx);
  dfree(ctx->d_tab);
  ctx->have_model = false;
  std::string e = build_model(ctx->N, nknot, nic, noc, ians, tref_period, deck, ctx->M);
  if (!e.empty()) return set_err(ctx, YSX_ERR_INPUT, e);
  ctx->lds_bytes = sizeof(double) * F_COUNT * (size_t)nknot;
  if (ctx->lds_bytes > ctx->lds_max)
    return set_err(ctx, YSX_ERR_UNSUPPORTED, "deck has too many knots for the LDS-resident table");
  HIPCHK(hipMalloc(&ctx->d_tab, ctx->lds_bytes));
  HIPCHK(hipMemcpy(ctx->d_tab, ctx->M.tab.data(), ctx->lds_bytes, hipMemcpyHostToDevice));
  ctx->have_model = true;
  return YSX_OK;
}

extern "C" int ysx_set_run(ysx_ctx* ctx, const int* ip, const double* fp, const double* lat,
                           const double* lon) {
  if (!ctx || !ip || !fp || !lat || !lon) return YSX_ERR_ARG;
  if (!ctx->have_model) return set_err(ctx, YSX_ERR_ARG, "no model");
  HIPCHK(hipSetDevice(ctx->device));
  free_run(ctx);
  std::string e = build_run(ctx->N, ctx->M, ip, fp, lat, lon, ctx->R);
  if (!e.empty()) return set_err(ctx, YSX_ERR_INPUT, e);
  const HostRun& R = ctx->R;
  const HostModel& M = ctx->M;
  RunK& k = ctx->rk;
  k.lmin = R.lmin; k.lmax = R.lmax; k.nw = R.nw; k.nr = R.nr; k.is = R.is; k.ir = R.ir;
  k.out = R.out; k.cor = R.cor;
  k.npair = (R.lmax - R.lmin + 1) * R.nw;
  k.nknot = M.nknot; k.nic = M.nic; k.noc = M.noc; k.nsl = M.nsl; k.fic = M.fic;
  k.ocean = M.ocean;
  k.rs = R.rs; k.rr = R.rr; k.wmin = R.wmin; k.dw = R.dw; k.ep = R.ep; k.tref = M.tref;
  k.ats = (double)R.ats; k.pibigg = ctx->N.pibigg; k.fre_norm = ctx->N.fre_norm;
  double amin = M.at(F_ALPHA, 1);
  for (int i = 2; i <= M.nknot; i++) amin = std::min(amin, M.at(F_ALPHA, i));
  k.alpha_min = amin;
  k.grav_ir = M.at(F_GRAV, R.ir);
  for (int i = 0; i < 6; i++) k.mm[i] = R.mm[i];
  size_t P = (size_t)k.npair, nt3 = R.xa.size();
  HIPCHK(hipMalloc(&ctx->d_rec, P * 89 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_starts, P * 2 * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_cnt, P * 5 * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_bst, P * 56 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_mfin, P * 14 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_m2, P * 14 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_tor_i1, P * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_sph_i1, P * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_tor_y0, P * 2 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_sph_m0, P * 14 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_fac, P * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_sc3, P * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_xa, nt3 * sizeof(double)));
  HIPCHK(hipMalloc(&ctx->d_xpa, nt3 * sizeof(double)));
  HIPCHK(hipMalloc(&ctx->d_xca, nt3 * sizeof(double)));
  HIPCHK(hipMalloc(&ctx->d_azep, R.nr * sizeof(double)));
  HIPCHK(hipMalloc(&ctx->d_eia, R.nr * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_spec, 3 * (size_t)R.nw * R.nr * sizeof(cd)));
  // GEMM operands: coefficients [15][lpad][jpad] (zero where no pair), and
  // the Legendre tables transposed to [X*3 + |m|][lpad][nrpad]
  ctx->lpad = (R.lmax + 1 + 3) / 4 * 4;
  ctx->jpad = (R.nw + 15) / 16 * 16;
  ctx->nrpad = (R.nr + 15) / 16 * 16;
  {
    size_t ncf = 15 * (size_t)ctx->lpad * ctx->jpad, ntb = 9 * (size_t)ctx->lpad * ctx->nrpad;
    HIPCHK(hipMalloc(&ctx->d_cf, ncf * sizeof(cd)));
    HIPCHK(hipMemset(ctx->d_cf, 0, ncf * sizeof(cd)));
    std::vector<double> tb(ntb, 0.0);
    const std::vector<double>* tabs[3] = {&R.xa, &R.xpa, &R.xca};
    int L1 = R.lmax + 1;
    for (int X = 0; X < 3; X++)
      for (int a = 0; a < 3; a++)
        for (int l = 0; l < L1; l++)
          for (int kk = 0; kk < R.nr; kk++)
            tb[((size_t)(X * 3 + a) * ctx->lpad + l) * ctx->nrpad + kk] =
                (*tabs[X])[(size_t)l + (size_t)L1 * (a + 3 * (size_t)kk)];
    HIPCHK(hipMalloc(&ctx->d_tb, ntb * sizeof(double)));
    HIPCHK(hipMemcpy(ctx->d_tb, tb.data(), ntb * sizeof(double), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMalloc(&ctx->d_tw, (size_t)R.nt * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_work, 3 * (size_t)R.nt * R.nr * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_seis, 3 * (size_t)R.nt * R.nr * sizeof(double)));
  HIPCHK(hipMemcpy(ctx->d_xa, R.xa.data(), nt3 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_xpa, R.xpa.data(), nt3 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_xca, R.xca.data(), nt3 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_azep, R.azep.data(), R.nr * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_eia, R.eia.data(), R.nr * sizeof(cd), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_twiddles, dim3((R.nt + 255) / 256), dim3(256), 0, ctx->stream, R.nt,
                     ctx->d_tw);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->have_run = true;
  return YSX_OK;
}

extern "C" int ysx_run_info(const ysx_ctx* ctx, int* iout, double* fout) {
  if (!ctx || !ctx->have_run) return YSX_ERR_ARG;
  const HostRun& R = ctx->R;
  if (iout) {
    int v[8] = {R.nw, R.nt, R.i1, R.i2, R.is, R.ir, R.nout, ctx->rk.npair};
    std::memcpy(iout, v, sizeof v);
  }
  if (fout) {
    double v[8] = {R.rs, R.rr, R.dt, R.tout, R.ep, R.df, R.wmin, R.dw};
    std::memcpy(fout, v, sizeof v);
  }
  return YSX_OK;
}

namespace {

int run_spectra(ysx_ctx* ctx) {
  const RunK& k = ctx->rk;
  hipStream_t st = ctx->stream, st2 = ctx->stream2;
  hipEvent_t* ev = ctx->ev;
  size_t lds = ctx->lds_bytes;
  int P = k.npair;
  dim3 b256(256), b64(64);
  int nl = k.lmax - k.lmin + 1;
  int nw64 = k.nw * ((nl + 63) / 64), nw16 = k.nw * ((nl + 15) / 16);
  SphState S{ctx->d_bst, ctx->d_mfin, ctx->d_m2, ctx->d_fac, ctx->d_sc3};
  StartData SD{ctx->d_tor_i1, ctx->d_tor_y0, ctx->d_sph_i1, ctx->d_sph_m0};
  HIPCHK(hipEventRecord(ev[0], st));
  hipLaunchKernelGGL(k_source, dim3((P + 255) / 256), b256, lds, st, ctx->d_tab, k, ctx->d_rec,
                     ctx->d_cnt);
  HIPCHK(hipEventRecord(ev[1], st));
  // the radial (l = 0) solves run beside the marches on a second stream
  HIPCHK(hipStreamWaitEvent(st2, ev[1], 0));
  HIPCHK(hipEventRecord(ev[10], st2));
  if (k.lmin == 0)
    hipLaunchKernelGGL(k_rad, dim3((k.nw + 63) / 64), dim3(64), lds, st2, ctx->d_tab, k,
                       ctx->d_rec, ctx->d_cnt);
  HIPCHK(hipEventRecord(ev[11], st2));
  hipLaunchKernelGGL(k_start, dim3((P + 255) / 256), b256, 0, st, ctx->d_tab, k, SD,
                     ctx->d_starts);
  HIPCHK(hipEventRecord(ev[2], st));
  hipLaunchKernelGGL(k_tor_march, dim3(nw64), b64, 0, st, ctx->d_tab, k, SD, ctx->d_rec,
                     ctx->d_cnt);
  HIPCHK(hipEventRecord(ev[3], st));
  hipLaunchKernelGGL(k_minor_top, dim3(nw64), b64, 0, st, ctx->d_tab, k, S, ctx->d_cnt);
  hipLaunchKernelGGL(k_minor_march, dim3(nw64), b64, 0, st, ctx->d_tab, k, SD, S, ctx->d_cnt);
  HIPCHK(hipEventRecord(ev[4], st));
  hipLaunchKernelGGL(k_bvec_march, dim3(nw16), b64, 0, st, ctx->d_tab, k, S, ctx->d_rec,
                     ctx->d_cnt);
  HIPCHK(hipEventRecord(ev[5], st));
  HIPCHK(hipStreamWaitEvent(st, ev[11], 0));
  HIPCHK(hipEventRecord(ev[6], st));
  GemmK G{ctx->lpad, ctx->jpad, ctx->nrpad, ctx->d_cf, ctx->d_tb};
  hipLaunchKernelGGL(k_assemble, dim3((P + 255) / 256), b256, 0, st, k, ctx->d_rec, G);
  HIPCHK(hipEventRecord(ev[7], st));
  hipLaunchKernelGGL(k_lsum, dim3((ctx->jpad / 16) * (ctx->nrpad / 16)), b64, 0, st, k, G,
                     ctx->d_eia, ctx->d_spec);
  HIPCHK(hipEventRecord(ev[8], st));
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventSynchronize(ev[8]));
  auto el = [&](int a, int b) {
    float t = 0.f;
    (void)hipEventElapsedTime(&t, ev[a], ev[b]);
    return (double)t;
  };
  ctx->ms[0] = el(0, 1) + el(1, 2);  // source vectors + starts
  ctx->ms[1] = el(2, 3);
  ctx->ms[2] = el(3, 4);
  ctx->ms[3] = el(4, 5);
  ctx->ms[4] = el(10, 11);
  ctx->ms[5] = el(6, 7);
  ctx->ms[6] = el(7, 8);
  ctx->have_spec = true;
  return YSX_OK;
}

}  // namespace

namespace {

int run_seis(ysx_ctx* ctx) {
  const HostRun& R = ctx->R;
  hipStream_t st = ctx->stream;
  SeisK K;
  K.nt = R.nt;
  K.lognt = 0;
  while ((1 << K.lognt) < R.nt) K.lognt++;
  K.nw = R.nw; K.nr = R.nr; K.i1 = R.i1; K.i2 = R.i2; K.out = R.out;
  K.df = R.df; K.dt = R.dt; K.ep = R.ep; K.f11 = R.f11; K.f12 = R.f12; K.f21 = R.f21;
  K.f22 = R.f22;
  K.norm = R.out == 0 ? ctx->N.r_norm : (R.out == 1 ? ctx->N.vel_norm : ctx->N.acl_norm);
  HIPCHK(hipEventRecord(ctx->ev[12], st));
  hipLaunchKernelGGL(k_seis, dim3(R.nr), dim3(1024), 0, st, K, ctx->d_spec, ctx->d_tw,
                     ctx->d_work, ctx->d_azep, ctx->d_seis);
  HIPCHK(hipEventRecord(ctx->ev[13], st));
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventSynchronize(ctx->ev[13]));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, ctx->ev[12], ctx->ev[13]));
  ctx->ms[7] = t;
  return YSX_OK;
}

}  // namespace

extern "C" int ysx_spectra(ysx_ctx* ctx, double* ur, double* ut, double* up) {
  if (!ctx) return YSX_ERR_ARG;
  if (!ctx->have_run) return set_err(ctx, YSX_ERR_ARG, "no run");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = run_spectra(ctx);
  if (rc) return rc;
  size_t n = (size_t)ctx->R.nw * ctx->R.nr;
  double* outs[3] = {ur, ut, up};
  for (int c = 0; c < 3; c++)
    if (outs[c])
      HIPCHK(hipMemcpy(outs[c], ctx->d_spec + c * n, n * sizeof(cd), hipMemcpyDeviceToHost));
  return YSX_OK;
}

extern "C" int ysx_seismograms(ysx_ctx* ctx, double* seis) {
  if (!ctx) return YSX_ERR_ARG;
  if (!ctx->have_spec) return set_err(ctx, YSX_ERR_ARG, "no spectra");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = run_seis(ctx);
  if (rc) return rc;
  if (seis)
    HIPCHK(hipMemcpy(seis, ctx->d_seis, 3 * (size_t)ctx->R.nt * ctx->R.nr * sizeof(double),
                     hipMemcpyDeviceToHost));
  return YSX_OK;
}

extern "C" int ysx_forward(ysx_ctx* ctx, double* seis) {
  int rc = ysx_spectra(ctx, nullptr, nullptr, nullptr);
  if (rc) return rc;
  return ysx_seismograms(ctx, seis);
}

extern "C" int ysx_pair_records(ysx_ctx* ctx, int n, const int* l, const int* j, double* rec) {
  if (!ctx || (n > 0 && (!l || !j || !rec))) return YSX_ERR_ARG;
  if (!ctx->have_spec) return set_err(ctx, YSX_ERR_ARG, "no spectra computed");
  HIPCHK(hipSetDevice(ctx->device));
  const RunK& k = ctx->rk;
  for (int i = 0; i < n; i++) {
    if (l[i] < k.lmin || l[i] > k.lmax || j[i] < 1 || j[i] > k.nw)
      return set_err(ctx, YSX_ERR_ARG, "pair outside the run");
    size_t p = (size_t)(l[i] - k.lmin) * k.nw + (j[i] - 1);
    HIPCHK(hipMemcpy(rec + (size_t)i * 178, ctx->d_rec + p * 89, 89 * sizeof(cd),
                     hipMemcpyDeviceToHost));
  }
  return YSX_OK;
}

extern "C" int ysx_timings(const ysx_ctx* ctx, double* ms8) {
  if (!ctx || !ms8) return YSX_ERR_ARG;
  std::memcpy(ms8, ctx->ms, sizeof ctx->ms);
  return YSX_OK;
}

extern "C" int ysx_step_counts(ysx_ctx* ctx, long long* out5) {
  if (!ctx || !out5) return YSX_ERR_ARG;
  if (!ctx->have_spec) return set_err(ctx, YSX_ERR_ARG, "no spectra computed");
  HIPCHK(hipSetDevice(ctx->device));
  size_t P = (size_t)ctx->rk.npair;
  std::vector<int> c(P * 5);
  HIPCHK(hipMemcpy(c.data(), ctx->d_cnt, c.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (int q = 0; q < 5; q++) out5[q] = 0;
  for (size_t p = 0; p < P; p++)
    for (int q = 0; q < 5; q++) out5[q] += c[5 * p + q];
  return YSX_OK;
}
