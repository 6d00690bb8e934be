// ysx_kern.hpp -- types and device helpers shared by the kernel files of
// libysx.so and the host API: per-event run data (RunK), work lists and
// the item loop, pair indexing, per-pair record layout, the state passed
// between the spheroidal kernels, the receiver-sum operands, and the host
// launch functions of every kernel.
#pragma once
#include <hip/hip_runtime.h>

#include "ysx_device.hpp"
#include "ysx_march.hpp"
#include "ysx_model.hpp"
#include "ysx_variants.hpp"

namespace ysx {

// ------------------------------------------------------- per-event run data
struct RunK {
  int lmin, lmax, nw, nr, is, ir, out, cor, npair;
  int nknot, nic, noc, nsl, fic, ocean;
  int grav, ngc;                // gravity switch; spheroidal solves by the gravity-0/1
                                // lockstep march (ngc)
  int lpad, jpad, nrpad;        // GEMM padding of this event
  int radg;                     // lanes per radial (l = 0) solve: 1 or 8
  int brow, bl;                 // b-vector as receiver-side rows (1) or source-side
                                // columns (0); degrees per b-vector wave
  int drow;                     // shared with other events in the batched path: there the
                                // secular determinant comes from the receiver-side row of
                                // the transposed minor system (k_drow_march)
  int tablds;                   // model table staged in LDS (else read from HBM)
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

// copy the model table into LDS (in_lds; every thread of the block must
// call it), or read it through the cache when it does not fit
__device__ inline Tab load_tab(const double* __restrict__ g, int nknot, int in_lds) {
  if (!in_lds) return Tab{g, nknot};
  extern __shared__ double smem[];
  int n = F_COUNT * nknot;
  for (int i = threadIdx.x; i < n; i += blockDim.x) smem[i] = g[i];
  __syncthreads();
  return Tab{smem, nknot};
}

// the pairs of events [e0, e1): global pair index of thread t
// Several events sharing the receiver depth and the frequency grid: the
// b-vector row march of a (w, degree block) is the same for all of them from
// the receiver down to each event's source interval, so one march serves
// them all (k_brow_march items with event -1).  order: the events by source
// radius, shallowest first; sx/ss: per union item and lane, the march state
// saved while an event's last (partial) interval is marched.
struct ShareK {
  const int* order;
  int n;                 // events in order (0: no shared items)
  int lmin, lmax, nw;    // union grid
  cd* sx;                // [items][14][64]
  int* ss;               // [items][64]
};

struct PairRange {
  int e0, e1;
  long long g0, n;
};

// per-pair output records (AoS, one record of 89 complex numbers per pair)
constexpr int REC = 89;
constexpr int O_SVT = 0, O_YTOR = 10, O_SVS = 14, O_YSPH = 44, O_SR0 = 68, O_YRAD = 70,
              O_UU = 74, O_VV = 79, O_WW = 84;

// k_assemble: pairs per 64-lane block
constexpr int ASM_P = 32;
inline int asm_blocks(const PairRange& PR) { return (int)std::max(1LL, (PR.n + ASM_P - 1) / ASM_P); }
// k_assemble_T: pairs per 64-lane block
constexpr int ASMT_P = 16;

// ------------------------------------------------ state between the kernels
// Start levels and starting solutions of every pair, written by k_start:
// tor_solver's start (src/module_solver.f90:64-75, 103-114) and
// sph_solver's start (:678-722, 758-785).  The marches pick these up when
// they reach each lane's start knot.
struct StartData {
  int* tor_i1;   // [npair] adjusted start knot
  cd* tor_y0;    // [npair][2]
  int* sph_i1;   // [npair] adjusted start knot
  cd* sph_m0;    // [npair][14] solid minors, or 5 fluid minors + zeros
};

// What k_minor_top and k_minor_march leave per pair for k_bvec_march
// (sph_solver, grav = 2, solid inner core, up to the b-vector integration,
// src/module_solver.f90:640-743, 756-806): the four b-vector starts (bst),
// the minor vector of sph_ysol (mfin), the factor applied to sph_ysol (fac)
// and scale3.
struct SphState {
  cd* bst;   // [npair][4][14]
  cd* mfin;  // [npair][14]
  cd* m2;    // [npair][14] minor vector from the surface
  cd* fac;   // [npair]
  int* sc3;  // [npair]
  cd* ms1;   // [npair][14] minors at the source (events with the determinant row, drow)
};

// Per-block scratch of the adjoint sweeps (ysx_k_adjoint.hip) in global
// memory: checkpoints (a state of up to 14 complex numbers and a
// renormalisation counter per knot interval and lane) and the tape of one
// interval's steps (six stage vectors per step and lane, and a counter).
struct AdjWork {
  cd* ys;     // the stored side's state before every step: steps base ..
  int* yss;   // base + smax - 1 of the current item (per block, or per item
  long long base;  // when the forward stored them), ystride complex per step
  cd* scr;    // the step scratch of each block (6 stage vectors + 1)
  int smax;
  // err[0] bit 0: an item with more steps than smax; the first is
  // described in err[1] (interval) and the doubles at err + 2: the step
  // count reached, the interval's two radii, Re w
  int* err;
  int ystride;
};

// Per-step states of a march (the minor march, or the toroidal one) stored
// by a single-event forward for the gradient that follows it: item (j, lb)
// holds cap[j - 1] steps from step off[j - 1] + lb cap[j - 1] on (state
// [step][N][64] complex, counter [step][64]); sc[3 p ..] gets pair p's
// renormalisation counters at the source, at the receiver and at its start
// level, and yend[N p ..] (when given) its state at the receiver; err[0] is
// set when an item has more steps than its capacity.  ys == nullptr: not
// stored.
struct StepStore {
  cd* ys;
  int* yss;
  const long long* off;
  const int* cap;
  int* sc;
  int* err;
  cd* yend;
};

// The minor sweep in two kernels when the forward stored the minor states:
// k_minor_rows carries each item's row through its steps (the sequential
// part), storing the row's state before every step in xs / xss (as the
// StepStore layout, same offsets) and appending one PairUnit per interval
// segment; k_minor_pairs then takes the units in any order (the pairings of
// every step are independent once both states are stored).  count: units
// appended; next: units taken; cap: room in units.
struct PairUnit {
  int e, j, lb, ik;  // event, frequency, degree block, interval | part << 16
  int s0, ns, pad0, pad1;  // first step of the segment in the item's storage, steps
};
struct RowStore {
  cd* xs;
  int* xss;
  PairUnit* units;
  int* count;
  int* next;
  int cap;
};

// ------------------------------------------------------- march helpers
// yint's interval loop (src/module_int.f90:1888-1963) over knot intervals,
// with wave-uniform arguments
template <class F, class Pre = NoPre>
__device__ int march_yint(const Tab& T, const WS& w, const LS& L, int nk, int i1, int i2,
                          double r1, double r2, cd* y, int& sc, double* sb, const F& f,
                          cd* slot = nullptr, Pre&& pre = Pre()) {
  // one call site of march_interval for both directions
  int n = 0;
  if (r2 == r1) return 0;
  const int dir = (r2 > r1) ? 1 : -1, lo = (dir > 0) ? 0 : 1;
  for (int i = i1; (i - i2) * dir <= 0; i += dir) {
    double a = (i == i1) ? r1 : T(F_R, i + lo);
    double b = (i == i2) ? r2 : T(F_R, i + 1 - lo);
    if (a == b) continue;
    n += march_interval(T, w, L, nk, i, a, b, y, true, sc, sb, f, slot, pre);
  }
  return n;
}

// march_yint's intervals and steps in reverse, each step transposed: for a
// row x, x <- x P_N ... P_1 where march_yint over the same arguments gives
// y <- P_N ... P_1 y
template <class F>
__device__ int march_yint_T(const Tab& T, const WS& w, const LS& L, int nk, int i1, int i2,
                            double r1, double r2, cd* x, int& sc, double* sb, const F& f,
                            cd* slot) {
  int n = 0;
  if (r2 == r1) return 0;
  const int dir = (r2 > r1) ? 1 : -1, lo = (dir > 0) ? 0 : 1;
  for (int i = i2; (i - i1) * dir >= 0; i -= dir) {
    double a = (i == i1) ? r1 : T(F_R, i + lo);
    double b = (i == i2) ? r2 : T(F_R, i + 1 - lo);
    if (a == b) continue;
    n += march_interval_T(T, w, L, nk, i, a, b, x, sc, sb, f, slot);
  }
  return n;
}

__device__ inline WS wave_omega(const RunK& R, int j) {
  return omega_state(C(R.wmin + (double)(j - 1) * R.dw, -R.ep), R.tref, R.ats, R.pibigg);
}

constexpr int NOLANE = 1 << 30;

// ------------------------------------------------- receiver-sum operands
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

// inverse FFT: series up to FFT_B points are transformed in LDS, FFT_G
// points per thread in the post-processing pass
constexpr int FFT_B = 4096;
constexpr int FFT_G = 16;

// ---------------------------------------------------------------- launchers
// Host-side launch functions, one per kernel, defined in the kernel files
// (ysx_k_march.hip: the marches, _item with one block per work item, _pers
// with persistent blocks; ysx_k_misc.hip: everything else).  Arguments
// after the stream are the kernel's.
void launch_reset(dim3 g, dim3 b, size_t lds, hipStream_t s, int* counters, int n);
// residual of the first nout samples per trace (block size 256, at most
// as many blocks as part holds)
void launch_residual(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* seis,
                     const double* obs, int nt, int nout, size_t ntrace, double* seisb,
                     double* part);
void launch_source(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                   const RunK* runs, PairRange PR, cd* rec, int* cnt);
// parts: bit 0 the toroidal starts, bit 1 the spheroidal ones
void launch_start(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                  const RunK* runs, PairRange PR, StartData SD, int* starts, int parts = 3);
void launch_rad(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab, const RunK* runs,
                Queue Q, cd* rec, int* cnt);
void launch_rad_ng(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                   const RunK* runs, Queue Q, cd* rec, int* cnt);
void launch_assemble(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, PairRange PR,
                     cd* rec, cd* cf, const double* tb);
void launch_lsum(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, Queue Q, cd* cf,
                 const double* tb, const cd* eia, cd* spec);
void launch_legendre(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                     const double* th, double* tb);
void launch_model(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* deck, int nknot,
                  int ifanis, double vel_norm, const int* layer_lo, const int* layer_n,
                  int nlayer, double* tab, double* work, size_t ws, double* intg, double* qint,
                  double* gi);
void launch_model_vjp(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* deck, int nknot,
                      int ifanis, double vel_norm, const int* layer_lo, const int* layer_n,
                      int nlayer, const double* tab, const double* tbar, double* bar,
                      double* work, size_t ws, double* aux, double* dbar);
void launch_fill0(dim3 g, dim3 b, size_t lds, hipStream_t s, double* p, size_t n);
void launch_seis(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, Queue Q,
                 const cd* spec, const cd* tw, cd* work, const double* azep, double* seis);
void launch_fft_lds(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, Queue Q,
                    const cd* spec, const cd* tw, cd* work);
void launch_fft_post(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, Queue Q,
                     const cd* tw, const cd* work, const double* azep, double* seis);
void launch_twiddles(dim3 g, dim3 b, size_t lds, hipStream_t s, int nt, cd* tw);
// adjoints of the seismogram, receiver-sum and assembly stages (event e)
void launch_seis_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                   const double* seisb, const cd* tw, cd* work, const double* azep, cd* specb);
void launch_post_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                   const double* seisb, const cd* tw, const double* azep, cd* work);
void launch_fft_lds_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                      const cd* tw, cd* work);
void launch_pack_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                   const cd* work, cd* specb);
void launch_lsum_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                   const double* tb, const cd* eia, const cd* specb, cd* cfb);
void launch_assemble_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                       const cd* rec, const cd* cfb, cd* recb);
void launch_source_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                     const RunK* runs, int e, const cd* recb, double* part);
void launch_source_model_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                           const RunK* runs, int e, const cd* recb, double* part);
// adjoint sweeps (ysx_k_adjoint.hip): persistent blocks over the queue Q,
// each adding into its own column of the partial table adjoints tbar
// ((F_COUNT nknot) x gridDim.x, block fastest); launch_part_sum (4 entries
// per 256-thread block) adds the nb columns of n entries into the table
// adjoint
void launch_part_sum(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* tpart, int nb,
                     size_t n, double* tbar);
void launch_brow_adj(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                     const RunK* runs, Queue Q, SphState SS, const cd* rec, const cd* recb,
                     cd* mub, AdjWork W, double* tbar);
// the two kernels of the minor sweep when the forward stored the states
// (FS.ys != nullptr): the rows, then the pairings (grid: persistent blocks)
void launch_minor_rows(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                       const RunK* runs, Queue Q, StartData SD, SphState SS, const cd* rec,
                       const cd* recb, const cd* mub, cd* m0bar, AdjWork W, StepStore FS,
                       RowStore RS, double* tbar);
void launch_minor_pairs(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                        const RunK* runs, AdjWork W, StepStore FS, RowStore RS, double* tbar);
// FS.ys != nullptr: the minor march's per-step states from the forward
// (k_minor_march with STORE) replace the sweep's first pass
void launch_minor_adj(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                      const RunK* runs, Queue Q, StartData SD, SphState SS, const cd* rec,
                      const cd* recb, const cd* mub, cd* m0bar, AdjWork W, StepStore FS,
                      double* tbar);
// FS.ys != nullptr: the toroidal march's per-step states from the forward
// (k_tor_march with STORE) replace the sweep's first pass
void launch_tor_adj(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                    const RunK* runs, Queue Q, StartData SD, const cd* rec, const cd* recb,
                    cd* y0bar, AdjWork W, StepStore FS, double* tbar);
void launch_start_adj(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                      const RunK* runs, PairRange PR, StartData SD, const int* starts,
                      const cd* m0bar, const cd* y0bar, double* tbar);
// radial adjoint of event e, cut into segments (sweep, knot interval):
// isrc upward segments (intervals 1 .. isrc), then downward ones (itop ..
// isrc); nwb blocks of 64 frequencies; cap: the most steps of a segment
// (k_rad_prop's atomic maximum), the tape's length per lane
struct RadSegs {
  int isrc, itop, nseg, nwb, cap;
};
struct RadSt {  // a step of the tape: radius, step, state before it
  double x, h;
  cd y[2];
};
size_t rad_step_bytes();
// grid (nseg, nwb) x 64: propagators prop[(q nw + j - 1) 4 ..], cap
void launch_rad_prop(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                     const RunK* runs, int e, RadSegs G, cd* prop, int* cap);
// grid nwb x 64: states ys and rows xe at the segment ends, start part
void launch_rad_scan(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                     const RunK* runs, int e, RadSegs G, const cd* prop, const cd* recb, cd* ys,
                     cd* xe, double* tbar);
// grid (nseg, nwb) x 64: the segments reversed, tape of nseg nwb cap 64
// steps
void launch_rad_seg_adj(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                        const RunK* runs, int e, RadSegs G, const cd* ys, const cd* xe,
                        void* tape, int* err, double* tbar);
#define YSX_MARCH_LAUNCHERS(SUFFIX)                                                              \
  void launch_tor_march_##SUFFIX(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab, \
                                 const RunK* runs, Queue Q, StartData SD, cd* rec, int* cnt,    \
                                 const StepStore* fs = nullptr);                                \
  void launch_minor_top_##SUFFIX(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab, \
                                 const RunK* runs, Queue Q, SphState S, int* cnt);              \
  void launch_minor_march_##SUFFIX(dim3 g, dim3 b, size_t lds, hipStream_t s,                   \
                                   const double* gtab, const RunK* runs, Queue Q, StartData SD, \
                                   SphState S, int* cnt, int drow_on,                           \
                                   const StepStore* fs = nullptr);                              \
  void launch_bvec_march_##SUFFIX(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,\
                                  const RunK* runs, Queue Q, SphState S, cd* rec, int* cnt);  \
  void launch_brow_march_##SUFFIX(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,\
                                  const RunK* runs, Queue Q, SphState S, ShareK SH, cd* rec,    \
                                  int* cnt);                                                     \
  void launch_drow_march_##SUFFIX(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,\
                                  const RunK* runs, Queue Q, SphState S, ShareK SH, int* cnt);
YSX_MARCH_LAUNCHERS(item)
YSX_MARCH_LAUNCHERS(pers)
#undef YSX_MARCH_LAUNCHERS
// gravity switch 1 (cow) or 0, ysx_k_ngc.hip
void launch_ngc_march(bool cow, dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                      const RunK* runs, Queue Q, cd* rec, int* cnt);

}  // namespace ysx
