// ysx_gpu.hip -- gfx950 kernels of the spectral path and the C-ABI of
// libysx.so (include/ysx.h).
//
// Per forward the GPU work is:
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
//
// A context holds one or more events (source + receivers + frequency grid)
// on one model.  Every kernel reads its event's parameters from a device
// array of RunK and takes its work from a list of items (event, w, l-block):
// one event at a time with one block per item ("plain"), or all events at
// once from cost-sorted lists drained by persistent blocks through an atomic
// queue, the whole forward replayed from a captured HIP graph ("batched").
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ysx.h"
#include "ysx_device.hpp"
#include "ysx_march.hpp"
#include "ysx_variants.hpp"
#include "ysx_model.hpp"
#include "ysx_host.hpp"

using namespace ysx;

namespace {

// ------------------------------------------------------- per-event run data
struct RunK {
  int lmin, lmax, nw, nr, is, ir, out, cor, npair;
  int nknot, nic, noc, nsl, fic, ocean;
  int grav, var;                // gravity switch; var: spheroidal solves per lane
  int lpad, jpad, nrpad;        // GEMM padding of this event
  int nt, lognt, i1, i2, roff;  // FFT length, frequency bins, first receiver
  long long pbase;              // first pair of this event in the per-pair arrays
  long long cf_off, tb_off, spec_off, seis_off;
  double rs, rr, wmin, dw, ep, tref, ats, pibigg, fre_norm, alpha_min, grav_ir;
  double df, dt, f11, f12, f21, f22, norm;
  double mm[6];
};

// A work list: items[0..n) of (event, a, b, 0).  counter == nullptr: block
// b takes items b, b + gridDim.x, ...; otherwise persistent blocks take
// `chunk` consecutive items per atomicAdd on *counter (zeroed before the
// launch by k_reset).  March blocks are one wave: lane 0 takes the items
// and broadcasts them.
struct Queue {
  const int4* items;
  int n;
  int* counter;
  int chunk;
};

__device__ __forceinline__ int4 uniform_item(const int4* items, int it) {
  int4 v = items[it];
  return make_int4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                   __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
}

__device__ __forceinline__ int next_base(const Queue& Q) {
  int base = 0;
  if (threadIdx.x == 0) base = atomicAdd(Q.counter, Q.chunk);
  return __builtin_amdgcn_readfirstlane(base);  // lane 0's value, wave-uniform
}

// Run ITEM_CALL (a statement using `it`, the current int4 item) over the
// queue: PERSIST = false: block b runs item b (the hardware dispatches
// blocks in index order, so a cost-sorted list is drained largest first);
// PERSIST = true: blocks loop over the queue.  Written as a macro, not a
// lambda: captured references keep the body's state out of registers.
#define RUN_ITEMS(PERSIST, Q, ITEM_CALL)                                     \
  do {                                                                       \
    if constexpr (!(PERSIST)) {                                              \
      if ((int)blockIdx.x < (Q).n) {                                         \
        int4 it = uniform_item((Q).items, blockIdx.x);                       \
        ITEM_CALL;                                                           \
      }                                                                      \
    } else {                                                                 \
      for (int base_ = next_base(Q); base_ < (Q).n; base_ = next_base(Q)) {  \
        int end_ = min(base_ + (Q).chunk, (Q).n);                            \
        for (int k_ = base_; k_ < end_; k_++) {                              \
          int4 it = uniform_item((Q).items, k_);                             \
          ITEM_CALL;                                                         \
        }                                                                    \
      }                                                                      \
    }                                                                        \
  } while (0)

__global__ void k_reset(int* counters, int n) {
  if ((int)threadIdx.x < n) counters[threadIdx.x] = 0;
}

// event of global pair g among events [e0, e1) (pbase ascending)
__device__ inline int find_event(const RunK* runs, int e0, int e1, long long g) {
  int lo = e0, hi = e1 - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (runs[mid].pbase <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

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

// the pairs of events [e0, e1): global pair index of thread t
struct PairRange {
  int e0, e1;
  long long g0, n;
};

// per-pair output records (AoS, one record of 89 complex numbers per pair)
constexpr int REC = 89;
constexpr int O_SVT = 0, O_YTOR = 10, O_SVS = 14, O_YSPH = 44, O_SR0 = 68, O_YRAD = 70,
              O_UU = 74, O_VV = 79, O_WW = 84;

// -------------------------------------------------------------- k_source
// source_vector_tor / _sph / _rad, src/module_solver.f90:145-215,
// 566-627, 1747-1847.  source_vector_sph keeps only the real part of its
// dispersion factors (they are declared real(dp) there, :1771-1772).
__global__ void __launch_bounds__(256) k_source(const double* __restrict__ gtab,
                                                const RunK* __restrict__ runs, PairRange PR,
                                                cd* __restrict__ rec, int* __restrict__ cnt) {
  Tab T = load_tab(gtab, runs[PR.e0].nknot);
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= PR.n) return;
  long long g = PR.g0 + t;
  const RunK& R = runs[find_event(runs, PR.e0, PR.e1, g)];
  int p = (int)(g - R.pbase);
  int l, j;
  pair_lj(R, p, l, j);
  PS s;
  pair_state(R, l, j, s);
  // records and step counters are zeroed once when the run is set; every
  // forward writes the same entries of them again
  cd* o = rec + (size_t)g * REC;
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

__global__ void __launch_bounds__(256) k_start(const double* __restrict__ gtab,
                                               const RunK* __restrict__ runs, PairRange PR,
                                               StartData SD, int* __restrict__ starts) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= PR.n) return;
  long long g = PR.g0 + t;
  const RunK& R = runs[find_event(runs, PR.e0, PR.e1, g)];
  Tab T{gtab, R.nknot};
  int l, j;
  pair_lj(R, (int)(g - R.pbase), l, j);
  size_t p = (size_t)g;
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
  if (R.var) return;  // the per-lane variant solvers find their own start
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
__device__ __forceinline__ void tor_item(const Tab& T, const RunK& R, int j, int lb,
                                         const StartData& SD, cd* __restrict__ rec,
                                         int* __restrict__ cnt, double* sb) {
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

// items (event, j, l-block of 64)
template <bool PERSIST>
__global__ void __launch_bounds__(64) k_tor_march(const double* __restrict__ gtab,
                                                  const RunK* __restrict__ runs, Queue Q,
                                                  StartData SD, cd* __restrict__ rec,
                                                  int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  RUN_ITEMS(PERSIST, Q,
            tor_item(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, it.z, SD, rec, cnt, sb));
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

__device__ __forceinline__ void minor_item(const Tab& T, const RunK& R, int j, int lb,
                                           const StartData& SD, const SphState& S,
                                           int* __restrict__ cnt, double* sb, cd* slot) {
  if (R.var) return;  // spheroidal solves of this event run in k_sph_var
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

// items (event, j, l-block of 64)
template <bool PERSIST>
__global__ void __launch_bounds__(64) k_minor_march(const double* __restrict__ gtab,
                                                    const RunK* __restrict__ runs, Queue Q,
                                                    StartData SD, SphState S,
                                                    int* __restrict__ cnt) {
  __shared__ double sb[6 * CH * BASE_MAX];
  __shared__ cd slots[2 * 14 * 64];
  cd* slot = slots + threadIdx.x;
  RUN_ITEMS(PERSIST, Q,
            minor_item(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, it.z, SD, S, cnt, sb,
                       slot));
}

// The surface minor vector ms2 of every pair (src/module_solver.f90:730-741,
// 793-804): from the surface down to the receiver (receiver above the
// source) or to the source.  Zero steps for a receiver at the surface.
__device__ __forceinline__ void minor_top_item(const Tab& T, const RunK& R, int j, int lb,
                                               const SphState& S, int* __restrict__ cnt,
                                               double* sb, cd* slot) {
  if (R.var) return;
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
  if (R.var) return;
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

__device__ __forceinline__ void rad_item(const Tab& T, const RunK& R, int jb,
                                         cd* __restrict__ rec, int* __restrict__ cnt) {
  int j = jb * 64 + threadIdx.x + 1;
  if (j > R.nw || R.lmin != 0) return;
  size_t p = (size_t)R.pbase + (j - 1);  // l = 0 pairs come first
  PS s;
  pair_state(R, 0, j, s);
  if (R.grav == 0) return;  // rad_solver_ng: k_rad_ng
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
  cd* o = rec + p * REC + O_YRAD;
  for (int i = 0; i < 4; i++) o[i] = yo[i];
  cnt[5 * p + 4] = nsr;
}

// items (event, block of 64 frequencies); one lane per l = 0 pair, no
// lockstep (the grid of make_grid depends on w)
__global__ void __launch_bounds__(64) k_rad(const double* __restrict__ gtab,
                                            const RunK* __restrict__ runs, Queue Q,
                                            cd* __restrict__ rec, int* __restrict__ cnt) {
  Tab T = load_tab(gtab, runs[0].nknot);
  RUN_ITEMS(false, Q, rad_item(T, runs[it.x], it.y, rec, cnt));
}

// rad_solver_ng (gravity switch 0, src/yspec.f90:271-273) for the l = 0
// pairs of the events with grav = 0; same items as k_rad
__device__ __forceinline__ void rad_ng_item(const Tab& T, const RunK& R, int jb,
                                            cd* __restrict__ rec, int* __restrict__ cnt) {
  int j = jb * 64 + threadIdx.x + 1;
  if (j > R.nw || R.lmin != 0 || R.grav != 0) return;
  size_t p = (size_t)R.pbase + (j - 1);
  PS s;
  pair_state(R, 0, j, s);
  VarGeo G{R.nknot, R.nic, R.noc, R.nsl, R.ocean, R.is, R.ir, R.rs, R.rr, R.alpha_min,
           R.fre_norm};
  cd yo[4];
  int n = 0;
  rad_solve_ng(T, s, G, yo, n);
  cd* o = rec + p * REC + O_YRAD;
  for (int i = 0; i < 4; i++) o[i] = yo[i];
  cnt[5 * p + 4] = n;
}

__global__ void __launch_bounds__(64) k_rad_ng(const double* __restrict__ gtab,
                                               const RunK* __restrict__ runs, Queue Q,
                                               cd* __restrict__ rec, int* __restrict__ cnt) {
  Tab T = load_tab(gtab, runs[0].nknot);
  RUN_ITEMS(false, Q, rad_ng_item(T, runs[it.x], it.y, rec, cnt));
}

// ------------------------------------------------------------- k_sph_var
// The spheroidal solve of one pair per lane for events outside the lockstep
// kernels: gravity switch 0 / 1 (sph_solver_ng, _cow, _ng_fic, _cow_fic)
// or a fluid inner core with self-gravitation (sph_solver_fic), with the
// reference driver's 4-vector -> 6-vector layout (src/yspec.f90:342-374).
__device__ __forceinline__ void sph_var_item(const Tab& T, const RunK& R, int pb,
                                             cd* __restrict__ rec, int* __restrict__ cnt) {
  int pl = pb * 64 + (int)threadIdx.x;
  if (pl >= R.npair) return;
  int l, j;
  pair_lj(R, pl, l, j);
  if (l < 1) return;
  PS s;
  pair_state(R, l, j, s);
  VarGeo G{R.nknot, R.nic, R.noc, R.nsl, R.ocean, R.is, R.ir, R.rs, R.rr, R.alpha_min,
           R.fre_norm};
  VarCount n{0, 0, 0};
  size_t p = (size_t)R.pbase + pl;
  cd* o = rec + p * REC + O_YSPH;
  if (R.grav == 2) {
    cd y[24];
    sph_solve_fic(T, s, G, y, n);
    for (int q = 0; q < 24; q++) o[q] = y[q];
  } else {
    cd y[16];
    if (R.grav == 1)
      sph_solve_ngc<true>(T, s, G, R.fic != 0, y, n);
    else
      sph_solve_ngc<false>(T, s, G, R.fic != 0, y, n);
    for (int c = 0; c < 4; c++) {
      o[6 * c + 0] = y[4 * c + 0];
      o[6 * c + 1] = y[4 * c + 1];
      o[6 * c + 2] = C(0.0);
      o[6 * c + 3] = y[4 * c + 2];
      o[6 * c + 4] = y[4 * c + 3];
      o[6 * c + 5] = C(0.0);
    }
  }
  cnt[5 * p + 1] = n.sol;
  cnt[5 * p + 2] = n.flu;
  cnt[5 * p + 3] = n.bvec;
}

// items (event, block of 64 pairs) of the events with var set
__global__ void __launch_bounds__(64) k_sph_var(const double* __restrict__ gtab,
                                                const RunK* __restrict__ runs, Queue Q,
                                                cd* __restrict__ rec, int* __restrict__ cnt) {
  RUN_ITEMS(false, Q, sph_var_item(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, rec, cnt));
}

// ------------------------------------------------------------- k_assemble
// output-type scaling and the uu/vv/ww coefficients, src/yspec.f90:280-288,
// 309-316, 377-400 (zeta there is single precision, :255)
// GEMM operands of event e, at offsets cf_off / tb_off of RunK:
//   cf [15][lpad][jpad] complex: uu m=-2..2, vv m=-2..2, ww m=-2..2
//   tb [9][lpad][nrpad] real: x, dx/dtheta, x/sin(theta) for |m| = 0,1,2
struct GemmK {
  int lpad, jpad, nrpad;
  cd* cf;
  const double* tb;
};

__device__ inline GemmK gemm_of(const RunK& R, cd* cf, const double* tb) {
  return GemmK{R.lpad, R.jpad, R.nrpad, cf + R.cf_off, tb + R.tb_off};
}

__device__ inline void put_coefs(const GemmK& G, int l, int j, const cd* o) {
  size_t plane = (size_t)G.lpad * G.jpad;
  size_t at = (size_t)l * G.jpad + (j - 1);
  for (int c = 0; c < 15; c++) G.cf[c * plane + at] = o[O_UU + c];
}

__global__ void __launch_bounds__(256) k_assemble(const RunK* __restrict__ runs, PairRange PR,
                                                  cd* __restrict__ rec, cd* __restrict__ cf,
                                                  const double* __restrict__ tb) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= PR.n) return;
  long long g = PR.g0 + t;
  const RunK& R = runs[find_event(runs, PR.e0, PR.e1, g)];
  GemmK G = gemm_of(R, cf, tb);
  int l, j;
  pair_lj(R, (int)(g - R.pbase), l, j);
  cd* o = rec + (size_t)g * REC;
  cd w = C(R.wmin + (double)(j - 1) * R.dw, -R.ep);
  cd fo = (R.out == 1) ? cmuli(w) : (R.out == 2 ? -(w * w) : C(1.0));
  double rr = R.rr;
  if (l == 0) {
    cd u3 = (o[O_SR0] * o[O_YRAD + 0] + o[O_SR0 + 1] * o[O_YRAD + 2]) / rr;
    o[O_UU + 2] = fo * u3;
    put_coefs(G, l, j, o);
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
  put_coefs(G, l, j, o);
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

__device__ __forceinline__ void lsum_item(const RunK& R, const GemmK& G, int tile,
                                          const cd* __restrict__ eia, cd* __restrict__ spec) {
  int lane = threadIdx.x;
  int njt = G.jpad / 16;
  int j0 = (tile % njt) * 16, k0 = (tile / njt) * 16;
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

// items (event, 16 w x 16 receiver tile); spec of event e at spec_off,
// [3][nr][nw]
__global__ void __launch_bounds__(64) k_lsum(const RunK* __restrict__ runs, Queue Q,
                                             cd* __restrict__ cf, const double* __restrict__ tb,
                                             const cd* __restrict__ eia, cd* __restrict__ spec) {
  RUN_ITEMS(false, Q,
            lsum_item(runs[it.x], gemm_of(runs[it.x], cf, tb), it.y, eia + runs[it.x].roff,
                      spec + runs[it.x].spec_off));
}

// -------------------------------------------------------------- k_legendre
// X_lm, dX_lm/dtheta, X_lm/sin(theta), |m| <= 2, of event e for degree l =
// blockIdx.x * 64 + lane and receiver blockIdx.y, into the receiver-sum
// operand tb[(X * 3 + |m|) * lpad + l][nrpad] (legendre, src/module_function.f90:
// 6-132, through the host's own arithmetic in ysx_model.hpp).  th holds
// theta, cos theta, sin theta per receiver.
__global__ void __launch_bounds__(64) k_legendre(const RunK* __restrict__ runs, int e,
                                                 const double* __restrict__ th,
                                                 double* __restrict__ tb_all) {
  const RunK& R = runs[e];
  int l = blockIdx.x * 64 + (int)threadIdx.x, k = blockIdx.y;
  if (l > R.lmax || k >= R.nr) return;
  const double* t = th + 3 * (size_t)(R.roff + k);
  int m = l < 2 ? l : 2;
  double x[3], xp[3], xc[3];
  legendre_cs(t[0], t[1], t[2], l, m, x, xp, xc);
  double* tb = tb_all + R.tb_off;
  for (int a = 0; a <= m; a++) {
    tb[((size_t)(0 * 3 + a) * R.lpad + l) * R.nrpad + k] = x[a];
    tb[((size_t)(1 * 3 + a) * R.lpad + l) * R.nrpad + k] = xp[a];
    tb[((size_t)(2 * 3 + a) * R.lpad + l) * R.nrpad + k] = xc[a];
  }
}

// ----------------------------------------------------------------- k_model
// The deck -> table transform on the device (one block): moduli per knot,
// natural splines per layer, gravity quadrature with the reference's
// sequential accumulation (src/module_model.f90:253-378), through the same
// functions as the host (ysx_model.hpp).  layer_lo (0-based first knot)
// and layer_n per layer; work holds 3 (nknot + 1) doubles per (layer,
// field) solve; gi holds nknot doubles.
__global__ void __launch_bounds__(256) k_model(const double* __restrict__ deck, int nknot,
                                               int ifanis, double vel_norm,
                                               const int* __restrict__ layer_lo,
                                               const int* __restrict__ layer_n, int nlayer,
                                               double* __restrict__ tab, double* __restrict__ work,
                                               double* __restrict__ intg,
                                               double* __restrict__ qint,
                                               double* __restrict__ gi) {
#pragma clang fp contract(off)
  int tid = threadIdx.x, nt = blockDim.x;
  size_t ws = 3 * ((size_t)nknot + 1);
  for (int i = tid; i < nknot; i += nt) knot_fields(deck, nknot, i, ifanis, vel_norm, tab);
  __syncthreads();
  const double* r = tab + (size_t)F_R * nknot;
  const double* rho = tab + (size_t)F_RHO * nknot;
  for (int i = tid; i < nknot; i += nt) intg[i] = r[i] * r[i] * rho[i];
  __syncthreads();
  for (int t = tid; t < 13 * nlayer; t += nt) {
    int k = t / 13, q = t % 13, lo = layer_lo[k], n = layer_n[k];
    if (q < 12) {
      int f = spline_field(q);
      natural_spline(r + lo, tab + (size_t)f * nknot + lo, tab + (size_t)(f + 1) * nknot + lo, n,
                     work + t * ws);
    } else {
      natural_spline(r + lo, intg + lo, qint + lo, n, work + t * ws);
    }
  }
  __syncthreads();
  for (int i = tid; i < nknot - 1; i += nt) gi[i] = grav_interval(r, intg, qint, i);
  __syncthreads();
  double* grav = tab + (size_t)F_GRAV * nknot;
  if (tid == 0) {
    double fint = 0.0;
    grav[0] = 0.0;
    for (int i = 0; i < nknot - 1; i++) {
      fint += gi[i];
      grav[i + 1] = fint;
    }
    for (int i = 0; i < nknot; i++)
      if (r[i] > 0.0) grav[i] = (4.0 * grav[i]) / (r[i] * r[i]);
  }
  __syncthreads();
  for (int k = tid; k < nlayer; k += nt)
    natural_spline(r + layer_lo[k], grav + layer_lo[k], tab + (size_t)F_GRAV2 * nknot + layer_lo[k],
                   layer_n[k], work + k * ws);
}

// ------------------------------------------------------------------ k_seis
// ifft (src/module_fourier.f90:55-145) and src/yspec.f90:439-473 for one
// receiver per block: pack bins i1..i2, cosine filters, Hermitian fill,
// radix-2 DIT inverse FFT with the reference's twiddle angles
// pi_d*m/mmax (single-rounded pi), /(dt*nt), exp(eps t), units, rotation.
__device__ inline cd packed_bin(const RunK& K, const cd* spec, int k, int i) {
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

// items (event, receiver k of the event), static distribution.  Event e's
// FFT rows are work[seis_off + (3k + c) nt ...] and its seismograms
// seis[seis_off + i + nt (c + 3k)], (nt, 3, nr) column-major.
__global__ void __launch_bounds__(1024) k_seis(const RunK* __restrict__ runs, Queue Q,
                                               const cd* __restrict__ spec_all,
                                               const cd* __restrict__ tw, cd* __restrict__ work_all,
                                               const double* __restrict__ azep_all,
                                               double* __restrict__ seis_all) {
  for (int it = blockIdx.x; it < Q.n; it += gridDim.x) {
    int4 item = Q.items[it];
    const RunK& K = runs[item.x];
    int k = item.y;
    int nt = K.nt;
    const cd* spec = spec_all + K.spec_off;
    cd* work = work_all + K.seis_off;
    double* seis = seis_all + K.seis_off;
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
    double az = azep_all[K.roff + k];
    double ca = cos(az), sa = sin(az);
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
    __syncthreads();
  }
}

// The same transform in two passes for nt <= FFT_B * FFT_G: stages
// mmax = 1 .. FFT_B/2 stay inside contiguous blocks of FFT_B elements of the
// bit-reversed row and run in LDS (k_fft_lds); the remaining stages mix
// only elements i0 + t FFT_B (t = 0 .. nt/FFT_B - 1) and run in registers,
// followed by exp(eps t), units and rotation (k_fft_post).  Butterflies,
// twiddles and their order are those of k_seis, so the results are the same
// bit for bit.
constexpr int FFT_B = 4096;
constexpr int FFT_G = 16;

// items (event, receiver k, c << 16 | block); work rows as in k_seis
__global__ void __launch_bounds__(512) k_fft_lds(const RunK* __restrict__ runs, Queue Q,
                                                 const cd* __restrict__ spec_all,
                                                 const cd* __restrict__ tw,
                                                 cd* __restrict__ work_all) {
  __shared__ cd buf[FFT_B];
  for (int it = blockIdx.x; it < Q.n; it += gridDim.x) {
    int4 item = Q.items[it];
    const RunK& K = runs[item.x];
    int k = item.y, c = item.z >> 16, blk = item.z & 0xffff;
    int nt = K.nt, bk = min(nt, FFT_B);
    const cd* sp = spec_all + K.spec_off + (size_t)c * K.nw * K.nr;
    cd* row = work_all + K.seis_off + ((size_t)k * 3 + c) * nt + (size_t)blk * bk;
    for (int q = threadIdx.x; q < bk; q += blockDim.x)
      buf[q] = packed_bin(K, sp, k, (int)brev((unsigned)(blk * bk + q), K.lognt));
    __syncthreads();
    for (int mmax = 1; mmax < bk; mmax <<= 1) {
      for (int b = threadIdx.x; b < bk / 2; b += blockDim.x) {
        int m = b & (mmax - 1);
        int i = ((b - m) << 1) + m;
        int jx = i + mmax;
        cd temp = tw[mmax + m] * buf[jx];
        cd a = buf[i];
        buf[jx] = a - temp;
        buf[i] = a + temp;
      }
      __syncthreads();
    }
    for (int q = threadIdx.x; q < bk; q += blockDim.x) row[q] = buf[q];
    __syncthreads();
  }
}

template <int G>
__device__ __forceinline__ void fft_post_g(const RunK& K, int k, int i0, const cd* __restrict__ tw,
                                           const cd* __restrict__ work, double ca, double sa,
                                           double* __restrict__ seis) {
  int nt = K.nt, bk = nt / G;
  cd v[3][G];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cd* row = work + ((size_t)k * 3 + c) * nt;
#pragma unroll
    for (int t = 0; t < G; t++) v[c][t] = row[i0 + t * bk];
  }
#pragma unroll
  for (int step = 1; step < G; step <<= 1) {
    int mmax = step * bk;
#pragma unroll
    for (int t = 0; t < G; t++) {
      if (t & step) continue;
      int tj = t + step;
      cd w = tw[mmax + ((i0 + t * bk) & (mmax - 1))];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        cd temp = w * v[c][tj];
        cd a = v[c][t];
        v[c][tj] = a - temp;
        v[c][t] = a + temp;
      }
    }
  }
  double sc = K.norm / (K.dt * nt);
#pragma unroll
  for (int t = 0; t < G; t++) {
    int i = i0 + t * bk;
    double g = exp(K.ep * (i * K.dt)) * sc;
    double z = v[0][t].x * g, ut = v[1][t].x * g, up = v[2][t].x * g;
    seis[(size_t)i + (size_t)nt * (0 + 3 * (size_t)k)] = z;
    seis[(size_t)i + (size_t)nt * (1 + 3 * (size_t)k)] = -ut * ca + up * sa;
    seis[(size_t)i + (size_t)nt * (2 + 3 * (size_t)k)] = ut * sa + up * ca;
  }
}

// items (event, receiver k, block of 256 column indices i0 < min(nt, FFT_B))
__global__ void __launch_bounds__(256) k_fft_post(const RunK* __restrict__ runs, Queue Q,
                                                  const cd* __restrict__ tw,
                                                  const cd* __restrict__ work_all,
                                                  const double* __restrict__ azep_all,
                                                  double* __restrict__ seis_all) {
  for (int it = blockIdx.x; it < Q.n; it += gridDim.x) {
    int4 item = Q.items[it];
    const RunK& K = runs[item.x];
    int k = item.y;
    int bk = min(K.nt, FFT_B), g = K.nt / bk;
    int i0 = item.z * 256 + (int)threadIdx.x;
    if (i0 >= bk) continue;
    double az = azep_all[K.roff + k];
    double ca = cos(az), sa = sin(az);
    const cd* work = work_all + K.seis_off;
    double* seis = seis_all + K.seis_off;
    switch (g) {
      case 1: fft_post_g<1>(K, k, i0, tw, work, ca, sa, seis); break;
      case 2: fft_post_g<2>(K, k, i0, tw, work, ca, sa, seis); break;
      case 4: fft_post_g<4>(K, k, i0, tw, work, ca, sa, seis); break;
      case 8: fft_post_g<8>(K, k, i0, tw, work, ca, sa, seis); break;
      default: fft_post_g<16>(K, k, i0, tw, work, ca, sa, seis); break;
    }
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
namespace {

enum WorkKind { W_TOR = 0, W_BVEC, W_RAD, W_LSUM, W_SEIS, W_VAR, W_FFTA, W_FFTB, W_N };
// queue counters of the batched launches (three marches share the W_TOR list)
enum QueueSlot { Q_TOR = 0, Q_TOP, Q_MINOR, Q_BVEC, Q_RAD, Q_N };

struct WorkList {
  std::vector<int4> plain;    // event-major, each event in its plain order
  std::vector<int> start;     // [nev + 1] offsets of each event in plain
  std::vector<int4> batched;  // all events, cost-sorted
  int4* d_plain = nullptr;
  int4* d_batched = nullptr;
};

}  // namespace

struct ysx_ctx {
  int device = 0, ncu = 256;
  hipStream_t stream = nullptr, stream2 = nullptr;
  std::string err;
  Norms N;
  HostModel M;
  std::vector<HostRun> R;  // one per event
  std::vector<RunK> rk;
  bool have_model = false, have_run = false, have_spec = false, have_seis = false;
  size_t lds_bytes = 0, lds_max = 0;
  long long npair = 0;  // all events
  // device buffers
  double* d_tab = nullptr;
  RunK* d_runs = nullptr;
  cd* d_rec = nullptr;
  int* d_starts = nullptr;
  int* d_cnt = nullptr;
  cd *d_bst = nullptr, *d_mfin = nullptr, *d_m2 = nullptr, *d_fac = nullptr;
  int *d_tor_i1 = nullptr, *d_sph_i1 = nullptr;
  cd *d_tor_y0 = nullptr, *d_sph_m0 = nullptr;
  int* d_sc3 = nullptr;
  double* d_azep = nullptr;
  cd* d_eia = nullptr;
  cd* d_spec = nullptr;
  cd* d_cf = nullptr;
  double* d_tb = nullptr;
  cd* d_tw = nullptr;
  cd* d_work = nullptr;
  double* d_seis = nullptr;
  int* d_queue = nullptr;  // Q_N counters of the batched queues
  WorkList W[W_N];
  // batched mode: persistent blocks or one block per item, items taken per
  // queue access, persistent blocks per CU, and the captured forward
  int persist = 0, chunk = 1, blocks_per_cu = 4;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t ev[16] = {};
  double ms[8] = {0};
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

void drop_graph(ysx_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
}

void free_run(ysx_ctx* c) {
  drop_graph(c);
  dfree(c->d_runs); dfree(c->d_rec); dfree(c->d_starts); dfree(c->d_cnt); dfree(c->d_bst);
  dfree(c->d_mfin); dfree(c->d_m2); dfree(c->d_tor_i1); dfree(c->d_sph_i1); dfree(c->d_tor_y0);
  dfree(c->d_sph_m0); dfree(c->d_fac); dfree(c->d_sc3); dfree(c->d_azep); dfree(c->d_eia);
  dfree(c->d_spec); dfree(c->d_cf); dfree(c->d_tb); dfree(c->d_tw); dfree(c->d_work);
  dfree(c->d_seis); dfree(c->d_queue);
  for (auto& w : c->W) {
    dfree(w.d_plain);
    dfree(w.d_batched);
    w = WorkList();
  }
  c->R.clear();
  c->rk.clear();
  c->npair = 0;
  c->have_run = c->have_spec = c->have_seis = false;
}

// Work lists of every kernel kind.  Plain order per event is the order of
// one launch per event (largest w first for the marches).  The batched
// lists hold every event's items, sorted by an estimate of their cost:
// steps grow with Re w; a wave of degrees starting at l0 integrates from
// about the turning radius (l0 + 1/2) v / w (v ~ 1 in normalised units) up
// to the top point; the b-vector runs between source and receiver.
void build_work(ysx_ctx* c) {
  int nev = (int)c->rk.size();
  std::vector<std::vector<double>> cost(W_N);
  for (int k = 0; k < W_N; k++) c->W[k].start.assign(1, 0);
  for (int e = 0; e < nev; e++) {
    const RunK& R = c->rk[e];
    int nl = R.lmax - R.lmin + 1, nlb = (nl + 63) / 64, nlb16 = (nl + 15) / 16;
    auto rew = [&](int j) { return R.wmin + (double)(j - 1) * R.dw; };
    double rtop = std::max(R.rs, R.rr);
    for (int wv = 0; wv < R.nw * nlb; wv++) {
      int j = R.nw - wv / nlb, lb = wv % nlb;
      c->W[W_TOR].plain.push_back(make_int4(e, j, lb, 0));
      double w = rew(j), rt = std::min(1.0, (R.lmin + 64.0 * lb + 0.5) / std::max(w, 1e-30));
      cost[W_TOR].push_back(w * (rtop - std::min(rt, rtop) + 0.05));
    }
    for (int wv = 0; wv < R.nw * nlb16; wv++) {
      int j = R.nw - wv / nlb16, lb = wv % nlb16;
      c->W[W_BVEC].plain.push_back(make_int4(e, j, lb, 0));
      cost[W_BVEC].push_back(rew(j) * (std::fabs(R.rr - R.rs) + 1e-3));
    }
    if (R.lmin == 0)
      for (int jb = 0; jb < (R.nw + 63) / 64; jb++) {
        c->W[W_RAD].plain.push_back(make_int4(e, jb, 0, 0));
        cost[W_RAD].push_back(rew(std::min(R.nw, 64 * jb + 64)));
      }
    if (R.var)
      for (int pb = 0; pb < (R.npair + 63) / 64; pb++) {
        c->W[W_VAR].plain.push_back(make_int4(e, pb, 0, 0));
        // the block's highest w sets its step count
        cost[W_VAR].push_back(rew(std::min(R.nw, (64 * pb + 63) % R.nw + 1)) * (rtop + 0.05));
      }
    int ntile = (R.jpad / 16) * (R.nrpad / 16);
    for (int t = 0; t < ntile; t++) {
      c->W[W_LSUM].plain.push_back(make_int4(e, t, 0, 0));
      cost[W_LSUM].push_back((double)R.lpad);
    }
    int bk = std::min(R.nt, FFT_B);
    if (R.nt / bk <= FFT_G) {  // two-pass transform
      for (int k = 0; k < R.nr; k++)
        for (int cc = 0; cc < 3; cc++)
          for (int blk = 0; blk < R.nt / bk; blk++) {
            c->W[W_FFTA].plain.push_back(make_int4(e, k, (cc << 16) | blk, 0));
            cost[W_FFTA].push_back(1.0);
          }
      for (int k = 0; k < R.nr; k++)
        for (int ib = 0; ib < (bk + 255) / 256; ib++) {
          c->W[W_FFTB].plain.push_back(make_int4(e, k, ib, 0));
          cost[W_FFTB].push_back(1.0);
        }
    } else {  // long series: radix-2 passes in global memory
      for (int k = 0; k < R.nr; k++) {
        c->W[W_SEIS].plain.push_back(make_int4(e, k, 0, 0));
        cost[W_SEIS].push_back((double)R.nt);
      }
    }
    for (int k = 0; k < W_N; k++) c->W[k].start.push_back((int)c->W[k].plain.size());
  }
  for (int k = 0; k < W_N; k++) {
    WorkList& w = c->W[k];
    std::vector<int> ord(w.plain.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int a, int b) { return cost[k][a] > cost[k][b]; });
    w.batched.resize(ord.size());
    for (size_t i = 0; i < ord.size(); i++) w.batched[i] = w.plain[ord[i]];
  }
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
  ctx->ncu = prop.multiProcessorCount;
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
  // the table the kernels read is built on the device from the raw deck;
  // the host copy (same functions) serves the run set-up
  {
    const HostModel& M = ctx->M;
    int nl = M.nlayer;
    std::vector<int> lo(nl), n(nl);
    for (int k = 0; k < nl; k++) {
      lo[k] = M.layer_lo[k] - 1;
      n[k] = M.layer_hi[k] - M.layer_lo[k] + 1;
    }
    size_t ws = 3 * ((size_t)nknot + 1), nwork = ws * 13 * (size_t)nl;
    double *d_deck = nullptr, *d_work = nullptr;
    int* d_lay = nullptr;
    auto cleanup = [&]() {
      dfree(d_deck);
      dfree(d_work);
      dfree(d_lay);
    };
    hipError_t e1 = hipMalloc(&d_deck, 9 * (size_t)nknot * sizeof(double));
    if (e1 == hipSuccess) e1 = hipMalloc(&d_work, (nwork + 3 * (size_t)nknot) * sizeof(double));
    if (e1 == hipSuccess) e1 = hipMalloc(&d_lay, 2 * (size_t)nl * sizeof(int));
    if (e1 == hipSuccess)
      e1 = hipMemcpy(d_deck, deck, 9 * (size_t)nknot * sizeof(double), hipMemcpyHostToDevice);
    if (e1 == hipSuccess) e1 = hipMemcpy(d_lay, lo.data(), nl * sizeof(int), hipMemcpyHostToDevice);
    if (e1 == hipSuccess)
      e1 = hipMemcpy(d_lay + nl, n.data(), nl * sizeof(int), hipMemcpyHostToDevice);
    if (e1 == hipSuccess) e1 = hipMemset(ctx->d_tab, 0, ctx->lds_bytes);
    if (e1 == hipSuccess) {
      hipLaunchKernelGGL(k_model, dim3(1), dim3(256), 0, ctx->stream, d_deck, nknot, M.ifanis,
                         ctx->N.vel_norm, d_lay, d_lay + nl, nl, ctx->d_tab, d_work,
                         d_work + nwork, d_work + nwork + nknot, d_work + nwork + 2 * (size_t)nknot);
      e1 = hipGetLastError();
    }
    if (e1 == hipSuccess) e1 = hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e1 != hipSuccess)
      return set_err(ctx, YSX_ERR_HIP, std::string("model transform: ") + hipGetErrorString(e1));
  }
  ctx->have_model = true;
  return YSX_OK;
}

extern "C" int ysx_model_table(const ysx_ctx* ctx, double* out) {
  if (!ctx || !out || !ctx->have_model) return YSX_ERR_ARG;
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipMemcpy(out, ctx->d_tab, ctx->lds_bytes, hipMemcpyDeviceToHost);
  return e == hipSuccess ? YSX_OK : YSX_ERR_HIP;
}

extern "C" int ysx_model_table_host(int nknot, int nic, int noc, int ians, double tref_period,
                                    const double* deck, double* out) {
  if (!deck || !out) return YSX_ERR_ARG;
  HostModel M;
  if (!build_model(make_norms(), nknot, nic, noc, ians, tref_period, deck, M).empty())
    return YSX_ERR_INPUT;
  std::memcpy(out, M.tab.data(), M.tab.size() * sizeof(double));
  return YSX_OK;
}

extern "C" int ysx_set_events(ysx_ctx* ctx, int nev, const int* ip, const double* fp,
                              const double* lat, const double* lon) {
  if (!ctx || nev < 1 || !ip || !fp || !lat || !lon) return YSX_ERR_ARG;
  if (!ctx->have_model) return set_err(ctx, YSX_ERR_ARG, "no model");
  HIPCHK(hipSetDevice(ctx->device));
  free_run(ctx);
  const HostModel& M = ctx->M;
  double amin = M.at(F_ALPHA, 1);
  for (int i = 2; i <= M.nknot; i++) amin = std::min(amin, M.at(F_ALPHA, i));
  ctx->R.resize(nev);
  ctx->rk.resize(nev);
  long long pb = 0, cfo = 0, tbo = 0, spo = 0, sio = 0;
  int ro = 0, ntmax = 0;
  for (int e = 0; e < nev; e++) {
    int nr = ip[7 * e + 6];
    if (nr < 1) return set_err(ctx, YSX_ERR_INPUT, "bad lmin/lmax/receiver count");
    std::string err = build_run(ctx->N, M, ip + 7 * e, fp + 18 * e, lat + ro, lon + ro, ctx->R[e]);
    if (!err.empty()) {
      ctx->R.clear();
      ctx->rk.clear();
      return set_err(ctx, YSX_ERR_INPUT, nev > 1 ? "event " + std::to_string(e) + ": " + err : err);
    }
    const HostRun& R = ctx->R[e];
    RunK& k = ctx->rk[e];
    k = RunK{};
    k.lmin = R.lmin; k.lmax = R.lmax; k.nw = R.nw; k.nr = R.nr; k.is = R.is; k.ir = R.ir;
    k.out = R.out; k.cor = R.cor;
    k.npair = (R.lmax - R.lmin + 1) * R.nw;
    k.nknot = M.nknot; k.nic = M.nic; k.noc = M.noc; k.nsl = M.nsl; k.fic = M.fic;
    k.ocean = M.ocean;
    k.grav = R.grav;
    k.var = (R.grav != 2 || M.fic) ? 1 : 0;
    k.lpad = (R.lmax + 1 + 3) / 4 * 4;
    k.jpad = (R.nw + 15) / 16 * 16;
    k.nrpad = (R.nr + 15) / 16 * 16;
    k.nt = R.nt;
    k.lognt = 0;
    while ((1 << k.lognt) < R.nt) k.lognt++;
    k.i1 = R.i1; k.i2 = R.i2; k.roff = ro;
    k.pbase = pb; k.cf_off = cfo; k.tb_off = tbo; k.spec_off = spo; k.seis_off = sio;
    k.rs = R.rs; k.rr = R.rr; k.wmin = R.wmin; k.dw = R.dw; k.ep = R.ep; k.tref = M.tref;
    k.ats = (double)R.ats; k.pibigg = ctx->N.pibigg; k.fre_norm = ctx->N.fre_norm;
    k.alpha_min = amin;
    k.grav_ir = M.at(F_GRAV, R.ir);
    k.df = R.df; k.dt = R.dt; k.f11 = R.f11; k.f12 = R.f12; k.f21 = R.f21; k.f22 = R.f22;
    k.norm = R.out == 0 ? ctx->N.r_norm : (R.out == 1 ? ctx->N.vel_norm : ctx->N.acl_norm);
    for (int i = 0; i < 6; i++) k.mm[i] = R.mm[i];
    pb += k.npair;
    cfo += 15LL * k.lpad * k.jpad;
    tbo += 9LL * k.lpad * k.nrpad;
    spo += 3LL * R.nw * R.nr;
    sio += 3LL * R.nt * R.nr;
    ro += nr;
    ntmax = std::max(ntmax, R.nt);
  }
  ctx->npair = pb;
  size_t P = (size_t)pb;
  HIPCHK(hipMalloc(&ctx->d_runs, nev * sizeof(RunK)));
  HIPCHK(hipMemcpy(ctx->d_runs, ctx->rk.data(), nev * sizeof(RunK), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&ctx->d_rec, P * REC * sizeof(cd)));
  HIPCHK(hipMemset(ctx->d_rec, 0, P * REC * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_starts, P * 2 * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_cnt, P * 5 * sizeof(int)));
  HIPCHK(hipMemset(ctx->d_cnt, 0, P * 5 * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_bst, P * 56 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_mfin, P * 14 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_m2, P * 14 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_tor_i1, P * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_sph_i1, P * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_tor_y0, P * 2 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_sph_m0, P * 14 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_fac, P * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_sc3, P * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_azep, ro * sizeof(double)));
  HIPCHK(hipMalloc(&ctx->d_eia, ro * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_spec, spo * sizeof(cd)));
  // GEMM operands: coefficients (zero where no pair) and the Legendre tables
  // transposed to [X*3 + |m|][lpad][nrpad]
  HIPCHK(hipMalloc(&ctx->d_cf, cfo * sizeof(cd)));
  HIPCHK(hipMemset(ctx->d_cf, 0, cfo * sizeof(cd)));
  {
    // receiver geometry from the host (O(nr)); the Legendre tables from the
    // device, written straight into the GEMM operand
    std::vector<double> az(ro), th(3 * (size_t)ro);
    std::vector<std::complex<double>> eia(ro);
    for (int e = 0; e < nev; e++) {
      const HostRun& R = ctx->R[e];
      const RunK& k = ctx->rk[e];
      for (int kk = 0; kk < R.nr; kk++) {
        az[k.roff + kk] = R.azep[kk];
        eia[k.roff + kk] = R.eia[kk];
        th[3 * (k.roff + kk)] = R.delta[kk];
        th[3 * (k.roff + kk) + 1] = std::cos(R.delta[kk]);
        th[3 * (k.roff + kk) + 2] = std::sin(R.delta[kk]);
      }
    }
    double* d_th = nullptr;
    HIPCHK(hipMalloc(&ctx->d_tb, tbo * sizeof(double)));
    HIPCHK(hipMemset(ctx->d_tb, 0, tbo * sizeof(double)));
    HIPCHK(hipMemcpy(ctx->d_azep, az.data(), ro * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_eia, eia.data(), ro * sizeof(cd), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&d_th, th.size() * sizeof(double)));
    hipError_t et = hipMemcpy(d_th, th.data(), th.size() * sizeof(double), hipMemcpyHostToDevice);
    for (int e = 0; e < nev && et == hipSuccess; e++) {
      const RunK& k = ctx->rk[e];
      hipLaunchKernelGGL(k_legendre, dim3((k.lmax + 64) / 64, k.nr), dim3(64), 0, ctx->stream,
                         ctx->d_runs, e, d_th, ctx->d_tb);
      et = hipGetLastError();
    }
    if (et == hipSuccess) et = hipStreamSynchronize(ctx->stream);
    dfree(d_th);
    if (et != hipSuccess)
      return set_err(ctx, YSX_ERR_HIP, std::string("Legendre tables: ") + hipGetErrorString(et));
  }
  HIPCHK(hipMalloc(&ctx->d_tw, (size_t)ntmax * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_work, sio * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_seis, sio * sizeof(double)));
  HIPCHK(hipMalloc(&ctx->d_queue, Q_N * sizeof(int)));
  build_work(ctx);
  for (auto& w : ctx->W) {
    size_t n = std::max<size_t>(1, w.plain.size());
    HIPCHK(hipMalloc(&w.d_plain, n * sizeof(int4)));
    HIPCHK(hipMalloc(&w.d_batched, n * sizeof(int4)));
    if (!w.plain.empty()) {
      HIPCHK(hipMemcpy(w.d_plain, w.plain.data(), w.plain.size() * sizeof(int4),
                       hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(w.d_batched, w.batched.data(), w.batched.size() * sizeof(int4),
                       hipMemcpyHostToDevice));
    }
  }
  hipLaunchKernelGGL(k_twiddles, dim3((ntmax + 255) / 256), dim3(256), 0, ctx->stream, ntmax,
                     ctx->d_tw);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->have_run = true;
  return YSX_OK;
}

extern "C" int ysx_set_run(ysx_ctx* ctx, const int* ip, const double* fp, const double* lat,
                           const double* lon) {
  return ysx_set_events(ctx, 1, ip, fp, lat, lon);
}

extern "C" int ysx_event_info(const ysx_ctx* ctx, int e, int* iout, double* fout) {
  if (!ctx || !ctx->have_run || e < 0 || e >= (int)ctx->R.size()) return YSX_ERR_ARG;
  const HostRun& R = ctx->R[e];
  if (iout) {
    int v[8] = {R.nw, R.nt, R.i1, R.i2, R.is, R.ir, R.nout, ctx->rk[e].npair};
    std::memcpy(iout, v, sizeof v);
  }
  if (fout) {
    double v[8] = {R.rs, R.rr, R.dt, R.tout, R.ep, R.df, R.wmin, R.dw};
    std::memcpy(fout, v, sizeof v);
  }
  return YSX_OK;
}

extern "C" int ysx_run_info(const ysx_ctx* ctx, int* iout, double* fout) {
  return ysx_event_info(ctx, 0, iout, fout);
}

namespace {

Queue plain_q(const ysx_ctx* c, int k, int e) {
  const WorkList& w = c->W[k];
  return Queue{w.d_plain + w.start[e], w.start[e + 1] - w.start[e], nullptr, 1};
}

// slot < 0: static distribution
Queue batched_q(const ysx_ctx* c, int k, int slot) {
  const WorkList& w = c->W[k];
  return Queue{w.d_batched, (int)w.batched.size(), slot < 0 ? nullptr : c->d_queue + slot,
               c->chunk};
}

int grid_of(const Queue& q) { return std::max(1, q.n); }

PairRange pairs_of(const ysx_ctx* c, int e0, int e1) {
  long long g0 = c->rk[e0].pbase;
  long long g1 = (e1 < (int)c->rk.size()) ? c->rk[e1].pbase : c->npair;
  return PairRange{e0, e1, g0, g1 - g0};
}

// the spectra of events [e0, e1) with one block per item: the plain path.
// With timing, per-phase HIP events are recorded (one event only).
int enqueue_plain(ysx_ctx* ctx, int e0, int e1, bool timing) {
  hipStream_t st = ctx->stream, st2 = ctx->stream2;
  hipEvent_t* ev = ctx->ev;
  size_t lds = ctx->lds_bytes;
  dim3 b256(256), b64(64);
  SphState S{ctx->d_bst, ctx->d_mfin, ctx->d_m2, ctx->d_fac, ctx->d_sc3};
  StartData SD{ctx->d_tor_i1, ctx->d_tor_y0, ctx->d_sph_i1, ctx->d_sph_m0};
  for (int e = e0; e < e1; e++) {
    PairRange PR = pairs_of(ctx, e, e + 1);
    int nb = (int)((PR.n + 255) / 256);
    Queue qt = plain_q(ctx, W_TOR, e), qb = plain_q(ctx, W_BVEC, e), qr = plain_q(ctx, W_RAD, e),
          ql = plain_q(ctx, W_LSUM, e);
    if (timing) HIPCHK(hipEventRecord(ev[0], st));
    hipLaunchKernelGGL(k_source, dim3(nb), b256, lds, st, ctx->d_tab, ctx->d_runs, PR,
                       ctx->d_rec, ctx->d_cnt);
    HIPCHK(hipEventRecord(ev[1], st));
    // the radial (l = 0) solves run beside the marches on a second stream
    HIPCHK(hipStreamWaitEvent(st2, ev[1], 0));
    if (timing) HIPCHK(hipEventRecord(ev[10], st2));
    if (qr.n > 0) {
      if (ctx->rk[e].grav == 0)
        hipLaunchKernelGGL(k_rad_ng, dim3(grid_of(qr)), b64, lds, st2, ctx->d_tab, ctx->d_runs,
                           qr, ctx->d_rec, ctx->d_cnt);
      else
        hipLaunchKernelGGL(k_rad, dim3(grid_of(qr)), b64, lds, st2, ctx->d_tab, ctx->d_runs, qr,
                           ctx->d_rec, ctx->d_cnt);
    }
    HIPCHK(hipEventRecord(ev[11], st2));
    hipLaunchKernelGGL(k_start, dim3(nb), b256, 0, st, ctx->d_tab, ctx->d_runs, PR, SD,
                       ctx->d_starts);
    if (timing) HIPCHK(hipEventRecord(ev[2], st));
    hipLaunchKernelGGL(k_tor_march<false>, dim3(grid_of(qt)), b64, 0, st, ctx->d_tab, ctx->d_runs, qt,
                       SD, ctx->d_rec, ctx->d_cnt);
    if (timing) HIPCHK(hipEventRecord(ev[3], st));
    if (ctx->rk[e].var) {
      Queue qv = plain_q(ctx, W_VAR, e);
      hipLaunchKernelGGL(k_sph_var, dim3(grid_of(qv)), b64, 0, st, ctx->d_tab, ctx->d_runs, qv,
                         ctx->d_rec, ctx->d_cnt);
      if (timing) HIPCHK(hipEventRecord(ev[4], st));
    } else {
      hipLaunchKernelGGL(k_minor_top<false>, dim3(grid_of(qt)), b64, 0, st, ctx->d_tab,
                         ctx->d_runs, qt, S, ctx->d_cnt);
      hipLaunchKernelGGL(k_minor_march<false>, dim3(grid_of(qt)), b64, 0, st, ctx->d_tab,
                         ctx->d_runs, qt, SD, S, ctx->d_cnt);
      if (timing) HIPCHK(hipEventRecord(ev[4], st));
      hipLaunchKernelGGL(k_bvec_march<false>, dim3(grid_of(qb)), b64, 0, st, ctx->d_tab,
                         ctx->d_runs, qb, S, ctx->d_rec, ctx->d_cnt);
    }
    if (timing) HIPCHK(hipEventRecord(ev[5], st));
    HIPCHK(hipStreamWaitEvent(st, ev[11], 0));
    if (timing) HIPCHK(hipEventRecord(ev[6], st));
    hipLaunchKernelGGL(k_assemble, dim3(nb), b256, 0, st, ctx->d_runs, PR, ctx->d_rec,
                       ctx->d_cf, ctx->d_tb);
    if (timing) HIPCHK(hipEventRecord(ev[7], st));
    hipLaunchKernelGGL(k_lsum, dim3(grid_of(ql)), b64, 0, st, ctx->d_runs, ql, ctx->d_cf,
                       ctx->d_tb, ctx->d_eia, ctx->d_spec);
    if (timing) HIPCHK(hipEventRecord(ev[8], st));
    HIPCHK(hipGetLastError());
  }
  return YSX_OK;
}

void enqueue_fft(ysx_ctx* ctx, const Queue& qs, const Queue& qa, const Queue& qb, int maxgrid) {
  hipStream_t st = ctx->stream;
  auto g = [&](const Queue& q) { return dim3(std::max(1, std::min(q.n, maxgrid))); };
  if (qs.n > 0)
    hipLaunchKernelGGL(k_seis, g(qs), dim3(1024), 0, st, ctx->d_runs, qs, ctx->d_spec, ctx->d_tw,
                       ctx->d_work, ctx->d_azep, ctx->d_seis);
  if (qa.n > 0)
    hipLaunchKernelGGL(k_fft_lds, g(qa), dim3(512), 0, st, ctx->d_runs, qa, ctx->d_spec,
                       ctx->d_tw, ctx->d_work);
  if (qb.n > 0)
    hipLaunchKernelGGL(k_fft_post, g(qb), dim3(256), 0, st, ctx->d_runs, qb, ctx->d_tw,
                       ctx->d_work, ctx->d_azep, ctx->d_seis);
}

int enqueue_seis_plain(ysx_ctx* ctx, int e0, int e1) {
  for (int e = e0; e < e1; e++)
    enqueue_fft(ctx, plain_q(ctx, W_SEIS, e), plain_q(ctx, W_FFTA, e), plain_q(ctx, W_FFTB, e),
                1 << 30);
  HIPCHK(hipGetLastError());
  return YSX_OK;
}

// every event at once from the cost-sorted lists: one block per item, or
// persistent blocks taking items from atomic queues
int enqueue_batched(ysx_ctx* ctx) {
  hipStream_t st = ctx->stream, st2 = ctx->stream2;
  hipEvent_t* ev = ctx->ev;
  size_t lds = ctx->lds_bytes;
  dim3 b256(256), b64(64);
  int nev = (int)ctx->rk.size();
  SphState S{ctx->d_bst, ctx->d_mfin, ctx->d_m2, ctx->d_fac, ctx->d_sc3};
  StartData SD{ctx->d_tor_i1, ctx->d_tor_y0, ctx->d_sph_i1, ctx->d_sph_m0};
  PairRange PR = pairs_of(ctx, 0, nev);
  int nb = (int)((PR.n + 255) / 256);
  bool pe = ctx->persist != 0;
  int pers = ctx->blocks_per_cu * ctx->ncu;
  auto grid = [&](const Queue& q) {
    return dim3(std::max(1, pe ? std::min(q.n, pers) : q.n));
  };
  Queue qt = batched_q(ctx, W_TOR, Q_TOR), qtop = batched_q(ctx, W_TOR, Q_TOP),
        qmin = batched_q(ctx, W_TOR, Q_MINOR), qb = batched_q(ctx, W_BVEC, Q_BVEC),
        qr = batched_q(ctx, W_RAD, -1), ql = batched_q(ctx, W_LSUM, -1),
        qs = batched_q(ctx, W_SEIS, -1);
  if (pe) hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, ctx->d_queue, (int)Q_N);
  hipLaunchKernelGGL(k_source, dim3(nb), b256, lds, st, ctx->d_tab, ctx->d_runs, PR, ctx->d_rec,
                     ctx->d_cnt);
  HIPCHK(hipEventRecord(ev[1], st));
  HIPCHK(hipStreamWaitEvent(st2, ev[1], 0));
  if (qr.n > 0) {
    hipLaunchKernelGGL(k_rad, dim3(qr.n), b64, lds, st2, ctx->d_tab, ctx->d_runs, qr,
                       ctx->d_rec, ctx->d_cnt);
    bool any_ng = false;
    for (const RunK& k : ctx->rk) any_ng = any_ng || (k.grav == 0 && k.lmin == 0);
    if (any_ng)
      hipLaunchKernelGGL(k_rad_ng, dim3(qr.n), b64, lds, st2, ctx->d_tab, ctx->d_runs, qr,
                         ctx->d_rec, ctx->d_cnt);
  }
  HIPCHK(hipEventRecord(ev[11], st2));
  hipLaunchKernelGGL(k_start, dim3(nb), b256, 0, st, ctx->d_tab, ctx->d_runs, PR, SD,
                     ctx->d_starts);
  if (pe) {
    hipLaunchKernelGGL(k_tor_march<true>, grid(qt), b64, 0, st, ctx->d_tab, ctx->d_runs, qt, SD,
                       ctx->d_rec, ctx->d_cnt);
    hipLaunchKernelGGL(k_minor_top<true>, grid(qtop), b64, 0, st, ctx->d_tab, ctx->d_runs, qtop,
                       S, ctx->d_cnt);
    hipLaunchKernelGGL(k_minor_march<true>, grid(qmin), b64, 0, st, ctx->d_tab, ctx->d_runs,
                       qmin, SD, S, ctx->d_cnt);
    hipLaunchKernelGGL(k_bvec_march<true>, grid(qb), b64, 0, st, ctx->d_tab, ctx->d_runs, qb, S,
                       ctx->d_rec, ctx->d_cnt);
  } else {
    hipLaunchKernelGGL(k_tor_march<false>, grid(qt), b64, 0, st, ctx->d_tab, ctx->d_runs, qt,
                       SD, ctx->d_rec, ctx->d_cnt);
    hipLaunchKernelGGL(k_minor_top<false>, grid(qt), b64, 0, st, ctx->d_tab, ctx->d_runs, qt, S,
                       ctx->d_cnt);
    hipLaunchKernelGGL(k_minor_march<false>, grid(qt), b64, 0, st, ctx->d_tab, ctx->d_runs, qt,
                       SD, S, ctx->d_cnt);
    hipLaunchKernelGGL(k_bvec_march<false>, grid(qb), b64, 0, st, ctx->d_tab, ctx->d_runs, qb,
                       S, ctx->d_rec, ctx->d_cnt);
  }
  Queue qv = batched_q(ctx, W_VAR, -1);
  if (qv.n > 0)
    hipLaunchKernelGGL(k_sph_var, dim3(qv.n), b64, 0, st, ctx->d_tab, ctx->d_runs, qv,
                       ctx->d_rec, ctx->d_cnt);
  HIPCHK(hipStreamWaitEvent(st, ev[11], 0));
  hipLaunchKernelGGL(k_assemble, dim3(nb), b256, 0, st, ctx->d_runs, PR, ctx->d_rec, ctx->d_cf,
                     ctx->d_tb);
  hipLaunchKernelGGL(k_lsum, dim3(std::max(1, ql.n)), b64, 0, st, ctx->d_runs, ql, ctx->d_cf,
                     ctx->d_tb, ctx->d_eia, ctx->d_spec);
  enqueue_fft(ctx, qs, batched_q(ctx, W_FFTA, -1), batched_q(ctx, W_FFTB, -1), 1 << 30);
  HIPCHK(hipGetLastError());
  return YSX_OK;
}

int elapsed(ysx_ctx* ctx, int a, int b, double& out) {
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, ctx->ev[a], ctx->ev[b]));
  out = t;
  return YSX_OK;
}

int run_spectra(ysx_ctx* ctx) {
  int rc = enqueue_plain(ctx, 0, 1, true);
  if (rc) return rc;
  HIPCHK(hipEventSynchronize(ctx->ev[8]));
  double a, b;
  if ((rc = elapsed(ctx, 0, 1, a)) || (rc = elapsed(ctx, 1, 2, b))) return rc;
  ctx->ms[0] = a + b;  // source vectors + starts
  if ((rc = elapsed(ctx, 2, 3, ctx->ms[1])) || (rc = elapsed(ctx, 3, 4, ctx->ms[2])) ||
      (rc = elapsed(ctx, 4, 5, ctx->ms[3])) || (rc = elapsed(ctx, 10, 11, ctx->ms[4])) ||
      (rc = elapsed(ctx, 6, 7, ctx->ms[5])) || (rc = elapsed(ctx, 7, 8, ctx->ms[6])))
    return rc;
  ctx->have_spec = true;
  ctx->have_seis = false;
  return YSX_OK;
}

int run_seis(ysx_ctx* ctx) {
  hipStream_t st = ctx->stream;
  HIPCHK(hipEventRecord(ctx->ev[12], st));
  int rc = enqueue_seis_plain(ctx, 0, 1);
  if (rc) return rc;
  HIPCHK(hipEventRecord(ctx->ev[13], st));
  HIPCHK(hipEventSynchronize(ctx->ev[13]));
  if ((rc = elapsed(ctx, 12, 13, ctx->ms[7]))) return rc;
  ctx->have_seis = true;
  return YSX_OK;
}

}  // namespace

extern "C" int ysx_spectra(ysx_ctx* ctx, double* ur, double* ut, double* up) {
  if (!ctx) return YSX_ERR_ARG;
  if (!ctx->have_run) return set_err(ctx, YSX_ERR_ARG, "no run");
  if (ctx->rk.size() != 1) return set_err(ctx, YSX_ERR_ARG, "several events: use ysx_forward_events");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = run_spectra(ctx);
  if (rc) return rc;
  size_t n = (size_t)ctx->R[0].nw * ctx->R[0].nr;
  double* outs[3] = {ur, ut, up};
  for (int c = 0; c < 3; c++)
    if (outs[c])
      HIPCHK(hipMemcpy(outs[c], ctx->d_spec + c * n, n * sizeof(cd), hipMemcpyDeviceToHost));
  return YSX_OK;
}

extern "C" int ysx_seismograms(ysx_ctx* ctx, double* seis) {
  if (!ctx) return YSX_ERR_ARG;
  if (!ctx->have_spec) return set_err(ctx, YSX_ERR_ARG, "no spectra");
  if (ctx->rk.size() != 1) return set_err(ctx, YSX_ERR_ARG, "several events: use ysx_forward_events");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = run_seis(ctx);
  if (rc) return rc;
  if (seis)
    HIPCHK(hipMemcpy(seis, ctx->d_seis, 3 * (size_t)ctx->R[0].nt * ctx->R[0].nr * sizeof(double),
                     hipMemcpyDeviceToHost));
  return YSX_OK;
}

extern "C" int ysx_forward(ysx_ctx* ctx, double* seis) {
  int rc = ysx_spectra(ctx, nullptr, nullptr, nullptr);
  if (rc) return rc;
  return ysx_seismograms(ctx, seis);
}

extern "C" int ysx_set_batch_tuning(ysx_ctx* ctx, int persist, int chunk, int blocks_per_cu) {
  if (!ctx || (persist != 0 && persist != 1) || chunk < 1 || chunk > 64 || blocks_per_cu < 1 ||
      blocks_per_cu > 32)
    return YSX_ERR_ARG;
  if (persist != ctx->persist || chunk != ctx->chunk || blocks_per_cu != ctx->blocks_per_cu)
    drop_graph(ctx);
  ctx->persist = persist;
  ctx->chunk = chunk;
  ctx->blocks_per_cu = blocks_per_cu;
  return YSX_OK;
}

extern "C" int ysx_forward_events(ysx_ctx* ctx, int mode, double* seis) {
  if (!ctx || (mode != 0 && mode != 1)) return YSX_ERR_ARG;
  if (!ctx->have_run) return set_err(ctx, YSX_ERR_ARG, "no run");
  HIPCHK(hipSetDevice(ctx->device));
  int nev = (int)ctx->rk.size();
  hipStream_t st = ctx->stream;
  HIPCHK(hipEventRecord(ctx->ev[14], st));
  if (mode == 0) {
    int rc = enqueue_plain(ctx, 0, nev, false);
    if (!rc) rc = enqueue_seis_plain(ctx, 0, nev);
    if (rc) return rc;
  } else {
    if (!ctx->gexec) {
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal));
      int rc = enqueue_batched(ctx);
      hipGraph_t g = nullptr;
      hipError_t ec = hipStreamEndCapture(st, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (ec != hipSuccess)
        return set_err(ctx, YSX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
      ctx->graph = g;
      HIPCHK(hipGraphInstantiate(&ctx->gexec, ctx->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(ctx->gexec, st));
  }
  HIPCHK(hipEventRecord(ctx->ev[15], st));
  HIPCHK(hipEventSynchronize(ctx->ev[15]));
  double t = 0.0;
  int rc = elapsed(ctx, 14, 15, t);
  if (rc) return rc;
  for (double& m : ctx->ms) m = 0.0;
  ctx->ms[7] = t;
  ctx->have_spec = ctx->have_seis = true;
  if (seis) {
    size_t n = 0;
    for (const HostRun& R : ctx->R) n += 3 * (size_t)R.nt * R.nr;
    HIPCHK(hipMemcpy(seis, ctx->d_seis, n * sizeof(double), hipMemcpyDeviceToHost));
  }
  return YSX_OK;
}

extern "C" int ysx_event_spectra(ysx_ctx* ctx, int e, double* ur, double* ut, double* up) {
  if (!ctx || e < 0 || e >= (int)ctx->rk.size()) return YSX_ERR_ARG;
  if (!ctx->have_spec) return set_err(ctx, YSX_ERR_ARG, "no spectra computed");
  HIPCHK(hipSetDevice(ctx->device));
  size_t n = (size_t)ctx->R[e].nw * ctx->R[e].nr;
  double* outs[3] = {ur, ut, up};
  for (int c = 0; c < 3; c++)
    if (outs[c])
      HIPCHK(hipMemcpy(outs[c], ctx->d_spec + ctx->rk[e].spec_off + c * n, n * sizeof(cd),
                       hipMemcpyDeviceToHost));
  return YSX_OK;
}

extern "C" int ysx_event_seismograms(ysx_ctx* ctx, int e, double* seis) {
  if (!ctx || !seis || e < 0 || e >= (int)ctx->rk.size()) return YSX_ERR_ARG;
  if (!ctx->have_seis) return set_err(ctx, YSX_ERR_ARG, "no seismograms computed");
  HIPCHK(hipSetDevice(ctx->device));
  size_t n = 3 * (size_t)ctx->R[e].nt * ctx->R[e].nr;
  HIPCHK(hipMemcpy(seis, ctx->d_seis + ctx->rk[e].seis_off, n * sizeof(double),
                   hipMemcpyDeviceToHost));
  return YSX_OK;
}

extern "C" int ysx_pair_records(ysx_ctx* ctx, int n, const int* l, const int* j, double* rec) {
  if (!ctx || (n > 0 && (!l || !j || !rec))) return YSX_ERR_ARG;
  if (!ctx->have_spec) return set_err(ctx, YSX_ERR_ARG, "no spectra computed");
  HIPCHK(hipSetDevice(ctx->device));
  const RunK& k = ctx->rk[0];
  for (int i = 0; i < n; i++) {
    if (l[i] < k.lmin || l[i] > k.lmax || j[i] < 1 || j[i] > k.nw)
      return set_err(ctx, YSX_ERR_ARG, "pair outside the run");
    size_t p = (size_t)(l[i] - k.lmin) * k.nw + (j[i] - 1);
    HIPCHK(hipMemcpy(rec + (size_t)i * 2 * REC, ctx->d_rec + p * REC, REC * sizeof(cd),
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
  size_t P = (size_t)ctx->npair;
  std::vector<int> c(P * 5);
  HIPCHK(hipMemcpy(c.data(), ctx->d_cnt, c.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (int q = 0; q < 5; q++) out5[q] = 0;
  for (size_t p = 0; p < P; p++)
    for (int q = 0; q < 5; q++) out5[q] += c[5 * p + q];
  return YSX_OK;
}
