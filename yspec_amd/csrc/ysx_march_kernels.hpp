// ysx_march_kernels.hpp -- the lockstep march kernels (toroidal, spheroidal
// minors from the surface and from the start level, b-vector), templated on
// the work-distribution mode; instantiated one per object file by
// ysx_k_march.hip.
#pragma once
#include "ysx_kern.hpp"

namespace ysx {
namespace {

// the per-step states of one item of a march into a StepStore: the state
// (N components) before each step, step s of the item at base + s
template <int N>
struct StorePre {
  const StepStore& FS;
  long long base;
  int cap;
  int& s;
  __device__ void operator()(const cd* y, int sc) const {
    if (s < cap) {
      double* c = reinterpret_cast<double*>(FS.ys + (base + s) * (N * 64) + threadIdx.x);
#pragma unroll
      for (int q = 0; q < N; q++) {
        __builtin_nontemporal_store(y[q].x, c + q * 128);
        __builtin_nontemporal_store(y[q].y, c + q * 128 + 1);
      }
      __builtin_nontemporal_store(sc, FS.yss + (base + s) * 64 + threadIdx.x);
    }
    s++;
  }
};

// the hook of item (j, lb): StorePre when the forward stores, else nothing
template <bool STORE, int N>
__device__ __forceinline__ auto store_pre(const StepStore& FS, int j, int lb, int& s) {
  if constexpr (STORE) {
    const int cap = FS.cap[j - 1];
    return StorePre<N>{FS, FS.off[j - 1] + (long long)lb * cap, cap, s};
  } else {
    return NoPre();
  }
}

// ------------------------------------------------------------- k_tor_march
// tor_solver with ss = identity(2), src/module_solver.f90:25-142; a wave
// holds 64 consecutive l at one w, largest w first.
// STORE: also the state before every step of the marches from the start
// level to the source and on to the receiver into FS, for the gradient
// (receiver at or above the source only)
template <bool STORE = false>
__device__ __forceinline__ void tor_item(const Tab& T, const RunK& R, int j, int lb,
                                         const StartData& SD, cd* __restrict__ rec,
                                         int* __restrict__ cnt, double* sb,
                                         const StepStore& FS) {
  int l = R.lmin + lb * 64 + (int)threadIdx.x;
  bool valid = (l >= 2 && l <= R.lmax);
  size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
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
  int fs = 0, sci1 = 0;
  auto pre = store_pre<STORE, 2>(FS, j, lb, fs);
  for (int i = ib; i <= itop; i++) {
    if (i == i1) {
      y1[0] = SD.tor_y0[2 * p];
      y1[1] = SD.tor_y0[2 * p + 1];
      if constexpr (STORE) sci1 = sc1;
    }
    double a = T(F_R, i), b = (i == itop) ? rtop : T(F_R, i + 1);
    if (a == b) continue;
    int n = march_interval(T, w, L, nk, i, a, b, y1, true, sc1, sb, f, nullptr, pre);
    if (i >= i1) ns += n;
  }
  double rsl = T(F_R, R.nsl);
  cd y2[2] = {C(1.0), C(0.0)}, y3[2], yo[4];
  if (up) {
    y3[0] = y1[0];
    y3[1] = y1[1];
    if (rr != rs) ns += march_yint(T, w, L, nk, is, ir, rs, rr, y3, sc3, sb, f, nullptr, pre);
    if constexpr (STORE) {
      if (threadIdx.x == 0 && fs > FS.cap[j - 1]) atomicOr(FS.err, 1);
      if (valid) {
        FS.sc[3 * p] = sc1;
        FS.sc[3 * p + 1] = sc3;
        FS.sc[3 * p + 2] = sci1;
        FS.yend[2 * p] = y3[0];
        FS.yend[2 * p + 1] = y3[1];
      }
    }
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

// items (event, j, l-block of 64)
template <bool PERSIST, bool STORE = false>
__global__ void __launch_bounds__(64) k_tor_march(const double* __restrict__ gtab,
                                                  const RunK* __restrict__ runs, Queue Q,
                                                  StartData SD, cd* __restrict__ rec,
                                                  int* __restrict__ cnt, StepStore FS) {
  __shared__ double sb[6 * CH * BASE_MAX];
  RUN_ITEMS(PERSIST, Q,
            tor_item<STORE>(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, it.z, SD, rec, cnt,
                            sb, FS));
}

// ------------------------------------------------ k_minor_top, k_minor_march
// sph_solver (grav = 2, solid inner core) up to the b-vector integration,
// src/module_solver.f90:640-743, 756-806, leaving SphState per pair; with a
// fluid inner core and the receiver not below the source, sph_solver_fic
// (:1240-1301), whose minors stay fluid from the inner core to the CMB.

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

// STORE: also the state before every step into FS, for the gradient
// (receiver at or above the source only)
template <bool DROW, bool STORE = false>
__device__ __forceinline__ void minor_item(const Tab& T, const RunK& R, int j, int lb,
                                           const StartData& SD, const SphState& S,
                                           int* __restrict__ cnt, double* sb, cd* slot,
                                           const StepStore& FS) {
  if (R.ngc) return;  // spheroidal solves of this event run in k_ngc_march
  int l = R.lmin + lb * 64 + (int)threadIdx.x;
  bool valid = (l >= 1 && l <= R.lmax);
  size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
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
  int fs = 0, sci1 = 0;
  auto pre = store_pre<STORE, 14>(FS, j, lb, fs);
  cd m[14];
  // (2) ms1 from each lane's start level up to the lower point
  double rt1 = up ? fmin(0.1 * (1.0 / (4.0 * w.rew / R.alpha_min)), 0.1 * T(F_R, 2)) : T(F_R, 1);
  for (int i = 0; i < 14; i++) m[i] = C(0.0);
  int layout = -1, sc1 = 0;  // -1 none, 0 solid, 1 fluid
  for (int i = ib; i <= itop; i++) {
    // fluid intervals: the outer core, and the inner core when it is fluid
    // (only such events with the receiver above the source run here)
    bool fluid_i = R.fic ? (i < noc) : (i > nic && i < noc);
    int region = fluid_i ? 1 : 0;
    if (layout == 0 && fluid_i) {
      cd mf[5];
      solid_2_fluid(m, mf);
      for (int q = 0; q < 14; q++) m[q] = q < 5 ? mf[q] : C(0.0);
    } else if (layout == 1 && !fluid_i) {
      cd mf[5] = {m[0], m[1], m[2], m[3], m[4]};
      fluid_2_solid(mf, m);
    }
    layout = region;
    if (i == i1) {
      for (int q = 0; q < 14; q++) m[q] = SD.sph_m0[p * 14 + q];
      if constexpr (STORE) sci1 = sc1;
    }
    double a = (i == 1) ? rt1 : T(F_R, i);
    double b = (i == itop) ? rtop : T(F_R, i + 1);
    if (a == b) continue;
    int n;
    if (fluid_i)
      n = march_interval(T, w, L, nk, i, a, b, m, true, sc1, sb, FluidMinorM(), nullptr, pre);
    else
      n = march_interval(T, w, L, nk, i, a, b, m, true, sc1, sb, SolidMinorM(), slot, pre);
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
    if (DROW && R.drow) {
      // the determinant comes from k_drow_march's receiver-side row; keep
      // the minors at the source for it
      if (!valid) return;
      for (int q = 0; q < 14; q++) {
        S.ms1[p * 14 + q] = m[q];
        S.mfin[p * 14 + q] = S.m2[p * 14 + q];
      }
      cnt[5 * p + 1] += nsol;
      cnt[5 * p + 2] += nflu;
      return;
    }
  } else {
    if (valid)
      for (int q = 0; q < 14; q++) S.mfin[p * 14 + q] = m[q];
  }
  // on from the lower point to the upper one (one call site for both cases)
  nsol += march_yint(T, w, L, nk, up ? is : ir, up ? ir : is, rtop, up ? rr : rs, m, sc3, sb,
                     SolidMinorM(), slot, pre);
  if constexpr (STORE) {
    if (threadIdx.x == 0 && fs > FS.cap[j - 1]) atomicOr(FS.err, 1);
    if (valid) {
      FS.sc[3 * p] = sc1;
      FS.sc[3 * p + 1] = sc3;
      FS.sc[3 * p + 2] = sci1;
    }
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

// items (event, j, l-block of 64).  DROW: events marked drow stop at the
// source and leave the determinant to k_drow_march (batched path)
template <bool PERSIST, bool DROW, bool STORE = false>
__global__ void __launch_bounds__(64) k_minor_march(const double* __restrict__ gtab,
                                                    const RunK* __restrict__ runs, Queue Q,
                                                    StartData SD, SphState S,
                                                    int* __restrict__ cnt, StepStore FS) {
  __shared__ double sb[6 * CH * BASE_MAX];
  __shared__ cd slots[2 * 14 * 64];
  cd* slot = slots + threadIdx.x;
  RUN_ITEMS(PERSIST, Q,
            (minor_item<DROW, STORE>(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, it.z, SD, S,
                                     cnt, sb, slot, FS)));
}

// The surface minor vector ms2 of every pair (src/module_solver.f90:730-741,
// 793-804): from the surface down to the receiver (receiver above the
// source) or to the source.  Zero steps for a receiver at the surface.
__device__ __forceinline__ void minor_top_item(const Tab& T, const RunK& R, int j, int lb,
                                               const SphState& S, int* __restrict__ cnt,
                                               double* sb, cd* slot) {
  if (R.ngc) return;
  int l = R.lmin + lb * 64 + (int)threadIdx.x;
  bool valid = (l >= 1 && l <= R.lmax);
  if (wave_min(valid ? 0 : 1)) return;
  size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
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

// items (event, j, l-block of 64)
template <bool PERSIST>
__global__ void __launch_bounds__(64) k_minor_top(const double* __restrict__ gtab,
                                                  const RunK* __restrict__ runs, Queue Q,
                                                  SphState S, int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  __shared__ cd slots[2 * 14 * 64];
  cd* slot = slots + threadIdx.x;
  RUN_ITEMS(PERSIST, Q,
            minor_top_item(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, it.z, S, cnt, sb,
                           slot));
}

// ------------------------------------------------------------ k_bvec_march
// b-vector from source to receiver and y = sph_ysol(m, b) for the four
// source columns, src/module_solver.f90:745-755, 809-819.  A wave holds 16
// consecutive l x 4 columns at one w.
__device__ __forceinline__ void bvec_item(const Tab& T, const RunK& R, int j, int lb,
                                          const SphState& S, cd* __restrict__ rec,
                                          int* __restrict__ cnt, double* sb, cd* slot) {
  if (R.ngc) return;
  int c = threadIdx.x & 3;
  int l = R.lmin + lb * 16 + (int)(threadIdx.x >> 2);
  bool valid = (l >= 1 && l <= R.lmax);
  int any = wave_min(valid ? 0 : 1);
  if (any) return;  // no valid lane in this wave
  size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
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

// items (event, j, l-block of 16); lanes 4 l + c integrate column c of l
template <bool PERSIST>
__global__ void __launch_bounds__(64) k_bvec_march(const double* __restrict__ gtab,
                                                   const RunK* __restrict__ runs, Queue Q,
                                                   SphState S, cd* __restrict__ rec,
                                                   int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  __shared__ cd slots[2 * 14 * 64];
  cd* slot = slots + threadIdx.x;
  RUN_ITEMS(PERSIST, Q,
            bvec_item(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, it.z, S, rec, cnt, sb,
                      slot));
}

// ------------------------------------------------------------ k_brow_march
// The same y = sph_ysol(mfin, b(rr)) fac for the four source columns
// (src/module_solver.f90:745-755, 809-819) from the other side: ysol is
// linear in b, so row k of y is x_k . b(rr) with x_k row k of ysol(mfin, .),
// and x_k . (P b(rs)) = (x_k P) . b(rs) for the b-vector march's product of
// step matrices P.  The rows x_k P are marched from the receiver back to
// the source with the transposed steps (march_yint_T), so two rows (three
// with the potential/tilt correction) replace four columns; the results
// differ from the column march by rounding only.  Rows 3..5 of ysph are not
// produced (zero).  A wave holds bl degrees x 64 / bl rows at one w.
// Rows k of pair (l, j) of event R from the marched row x (scale count sl):
// y_c[k] = x . bst_c * fac * huge^(sc3 - sl), c = 0..3, into the record;
// rows the march does not produce are zeroed.
__device__ __forceinline__ void brow_finish(const RunK& R, int l, int j, int k, int rows,
                                            const cd* x, int sl, int nsteps, const SphState& S,
                                            cd* __restrict__ rec, int* __restrict__ cnt) {
  size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
  cd fac = S.fac[p] * ipow_huge(S.sc3[p] - sl);
  cd* o = rec + p * REC + O_YSPH;
  for (int c = 0; c < 4; c++) {
    cd v = C(0.0);
#pragma unroll
    for (int q = 0; q < 14; q++) v = cmad(v, x[q], S.bst[p * 56 + c * 14 + q]);
    if (k < 3) o[6 * c + k] = v * fac;
    if (k == 0)
      for (int r = (rows < 3 ? rows : 3); r < 6; r++) o[6 * c + r] = C(0.0);
  }
  if (k == 0) cnt[5 * p + 3] = nsteps;
}

__device__ __forceinline__ void brow_item(const Tab& T, const RunK& R, int j, int lb,
                                          const SphState& S, cd* __restrict__ rec,
                                          int* __restrict__ cnt, double* sb, cd* slot) {
  if (R.ngc) return;
  const int rows = 64 / R.bl;  // lanes per degree: row k = lane % rows
  int k = threadIdx.x % rows;
  int l = R.lmin + lb * R.bl + (int)(threadIdx.x / rows);
  bool valid = (l >= 1 && l <= R.lmax);
  if (wave_min(valid ? 0 : 1)) return;  // no valid lane in this wave
  size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
  WS w = wave_omega(R, j);
  LS L = l_state(valid ? l : 1);
  cd x[14];
  {
    cd m[14];
    for (int q = 0; q < 14; q++) m[q] = valid ? S.mfin[p * 14 + q] : C(0.0);
    ysol_row(m, k, x);
  }
  int sl = 0;
  int n = march_yint_T(T, w, L, R.nknot, R.is, R.ir, R.rs, R.rr, x, sl, sb, BvecTM(), slot);
  if (!valid) return;
  brow_finish(R, l, j, k, rows, x, sl, n, S, rec, cnt);
}

__device__ __forceinline__ bool lane_has(const RunK& R, int l, int j) {
  return l >= 1 && l >= R.lmin && l <= R.lmax && j <= R.nw;
}

// The receiver-side row march of one (w_j, degree block l0 .. l0+bl-1) of
// the shared events' union grid, from the common receiver down, branching
// at each event's source interval.  Each event's rows see exactly the steps
// of its own march (march_yint_T over is_e..ir): the trunk takes the
// intervals above the source interval whole, the branch the source interval
// from rs_e (trunk state saved in global scratch meanwhile), so the results
// are those of the per-event items.  finish(R, x, scale, steps) is called
// (wave-uniformly) at the end of each event's branch.
template <class F, class Fin>
__device__ __forceinline__ void shared_march(const RunK* runs, const ShareK& SH, const Tab& T,
                                             const WS& w, const LS& L, int j, int l0, int bl,
                                             int u, cd* x, double* sb, cd* slot, const F& f,
                                             Fin&& finish) {
  const RunK& R0 = runs[SH.order[0]];
  cd* sx = SH.sx + (size_t)u * 14 * 64 + threadIdx.x;
  int* ss = SH.ss + (size_t)u * 64 + threadIdx.x;
  const int ir = R0.ir;
  const double rr = R0.rr;
  int sl = 0, ntr = 0, icur = ir, q = 0;
  // one march call site: each pass marches either the next whole trunk
  // interval or an event's source interval
  while (true) {
    // next event with a pair in this wave (wave-uniform)
    while (q < SH.n) {
      const RunK& R = runs[SH.order[q]];
      if (j <= R.nw && l0 <= R.lmax && l0 + bl - 1 >= max(R.lmin, 1)) break;
      q++;
    }
    if (q == SH.n) break;
    const RunK& R = runs[SH.order[q]];
    // branch once the trunk has reached the source interval; a source at
    // the receiver takes no step at all (march_yint returns at once)
    const bool branch = (icur <= R.is) || (R.rs == rr);
    int iv;
    double a, b;
    if (!branch) {
      iv = icur;
      a = T(F_R, icur);
      b = (icur == ir) ? rr : T(F_R, icur + 1);
      icur--;
    } else {
      for (int c = 0; c < 14; c++) sx[c * 64] = x[c];
      *ss = sl;
      iv = R.is;
      a = (R.rs == rr) ? rr : R.rs;
      b = (R.rs == rr || iv == ir) ? rr : T(F_R, iv + 1);
    }
    int n = (a != b) ? march_interval_T(T, w, L, R0.nknot, iv, a, b, x, sl, sb, f, slot) : 0;
    if (!branch) {
      ntr += n;
    } else {
      finish(R, x, sl, (R.rs == rr) ? 0 : ntr + n);
      for (int c = 0; c < 14; c++) x[c] = sx[c * 64];
      sl = *ss;
      q++;
    }
  }
}

// pair index of lane l in event R (lane_has(R, l, j) must hold)
__device__ __forceinline__ size_t pair_of(const RunK& R, int l, int j) {
  return (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
}

// the first shared event holding pair (l, j), or -1
__device__ __forceinline__ int first_holder(const RunK* runs, const ShareK& SH, int l, int j) {
  int pe = -1;
  for (int q = 0; q < SH.n; q++)
    if (pe < 0 && lane_has(runs[SH.order[q]], l, j)) pe = SH.order[q];
  return pe;
}

__device__ __forceinline__ void brow_shared_item(const double* gtab, const RunK* runs,
                                                 const ShareK& SH, int j, int lb, int u,
                                                 const SphState& S, cd* __restrict__ rec,
                                                 int* __restrict__ cnt, double* sb, cd* slot) {
  const RunK& R0 = runs[SH.order[0]];
  const Tab T{gtab, R0.nknot};
  const int rows = 64 / R0.bl;
  const int k = threadIdx.x % rows;
  const int l0 = SH.lmin + lb * R0.bl;
  const int l = l0 + (int)(threadIdx.x / rows);
  WS w = wave_omega(R0, j);
  LS L = l_state((l >= 1 && l <= SH.lmax) ? l : 1);
  // rows from the first event holding this lane's pair (the minors from the
  // receiver, mfin, are the same for all shared events)
  cd x[14];
  {
    int pe = first_holder(runs, SH, l, j);
    cd m[14];
    for (int q = 0; q < 14; q++) m[q] = (pe >= 0) ? S.mfin[pair_of(runs[pe], l, j) * 14 + q] : C(0.0);
    ysol_row(m, k, x);
  }
  shared_march(runs, SH, T, w, L, j, l0, R0.bl, u, x, sb, slot, BvecTM(),
               [&](const RunK& R, const cd* xe, int sl, int nsteps) {
                 if (lane_has(R, l, j)) brow_finish(R, l, j, k, rows, xe, sl, nsteps, S, rec, cnt);
               });
}

// items (event, j, l-block of bl); event -1: a shared union item (.w its index)
template <bool PERSIST>
__global__ void __launch_bounds__(64) k_brow_march(const double* __restrict__ gtab,
                                                   const RunK* __restrict__ runs, Queue Q,
                                                   SphState S, ShareK SH, cd* __restrict__ rec,
                                                   int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  __shared__ cd slots[2 * 14 * 64];
  cd* slot = slots + threadIdx.x;
  RUN_ITEMS(PERSIST, Q, {
    if (it.x >= 0)
      brow_item(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, it.z, S, rec, cnt, sb, slot);
    else
      brow_shared_item(gtab, runs, SH, it.y, it.z, it.w, S, rec, cnt, sb, slot);
  });
}

// ------------------------------------------------------------ k_drow_march
// The secular determinant delta = delta_solid(ms3(rr), ms2(rr)) of
// sph_solver (src/module_solver.f90:730-743) with ms3 = P ms1(rs) the minor
// march from the source to the receiver: delta is linear in ms3, so
// delta = (d P) . ms1(rs) with d its coefficients (given ms2(rr)), and the
// row d P is marched from the receiver back to the source with the
// transposed steps of the same step sequence.  Writes fac = -1/delta and
// the row's scale count as sc3 (so y = ysol(...) fac huge^(sc3 - sc4) keeps
// its form); rounding-level differences from the forward continuation.
__device__ __forceinline__ void drow_finish(const RunK& R, int l, int j, const cd* x, int sd,
                                            int nsteps, const SphState& S,
                                            int* __restrict__ cnt) {
  size_t p = pair_of(R, l, j);
  cd delta = C(0.0);
#pragma unroll
  for (int q = 0; q < 14; q++) delta = cmad(delta, x[q], S.ms1[p * 14 + q]);
  S.fac[p] = -cinv(delta);
  S.sc3[p] = sd;
  cnt[5 * p + 1] += nsteps;
}

__device__ __forceinline__ void drow_shared_item(const double* gtab, const RunK* runs,
                                                 const ShareK& SH, int j, int lb, int u,
                                                 const SphState& S, int* __restrict__ cnt,
                                                 double* sb, cd* slot) {
  const RunK& R0 = runs[SH.order[0]];
  const Tab T{gtab, R0.nknot};
  const int l0 = SH.lmin + lb * 64;
  const int l = l0 + (int)threadIdx.x;
  WS w = wave_omega(R0, j);
  LS L = l_state((l >= 1 && l <= SH.lmax) ? l : 1);
  cd x[14];
  {
    int pe = first_holder(runs, SH, l, j);
    cd m2[14];
    for (int q = 0; q < 14; q++) m2[q] = (pe >= 0) ? S.mfin[pair_of(runs[pe], l, j) * 14 + q] : C(0.0);
    delta_row(m2, x);
  }
  shared_march(runs, SH, T, w, L, j, l0, 64, u, x, sb, slot, MinorTM(),
               [&](const RunK& R, const cd* xe, int sd, int nsteps) {
                 if (lane_has(R, l, j)) drow_finish(R, l, j, xe, sd, nsteps, S, cnt);
               });
}

// items (-1, j, l-block of 64, union index) of the shared events (batched path)
template <bool PERSIST>
__global__ void __launch_bounds__(64) k_drow_march(const double* __restrict__ gtab,
                                                   const RunK* __restrict__ runs, Queue Q,
                                                   SphState S, ShareK SH, int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  __shared__ cd slots[2 * 14 * 64];
  cd* slot = slots + threadIdx.x;
  RUN_ITEMS(PERSIST, Q, drow_shared_item(gtab, runs, SH, it.y, it.z, it.w, S, cnt, sb, slot));
}

}  // namespace
}  // namespace ysx
