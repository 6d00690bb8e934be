// ysx_k_adjoint.hip -- the reverse sweeps of the discrete adjoint: the
// gradient of Re sum(conj(record adjoint) * record) over all (l, w) pairs
// with respect to the model table, through the per-pair solves.
//
// Every march is a frozen list of Cash-Karp steps (the reference's step
// rule, yint, src/module_int.f90:1847-1968), so a solve is a product of
// step matrices and its adjoint pairs, step by step, the column the forward
// carries (the solution) with a row carried the other way (the adjoint).
// A wave keeps the forward's lockstep layout (64 lanes at one w, stage
// radii shared) and for each item:
//   1. marches the stored side again from its start, writing its state
//      before every step to the block's per-step storage;
//   2. goes through the steps in the other direction: per step the stored
//      side's step is replayed from its stored state with every stage
//      input written to the block's step scratch, the other side takes its
//      step with every stage row written beside them, and each stage's
//      row is paired with its input (ysx_adjoint.hpp) into adjoints of the
//      stage's base (the next stage's vectors load while one is paired);
//   3. per chunk of steps, sums those adjoints over the wave (LDS), takes
//      them through the base computation to the spline values and the
//      interval's table entries (one lane per stage radius), and at the end
//      of the interval adds them to the block's column of partial table
//      adjoints.
// The step scratch is read and written through a buffer resource (scalar
// and immediate offsets), so the replay keeps the registers of the
// forward's step; only the stored side's 14-vectors are per step.
// Kernels:
//   k_brow_adj   b-vector system: the receiver-side rows (stored side) and
//                the source-side columns weighted by the record adjoints
//   k_minor_adj  minor vectors: the columns from the start level (stored
//                side) and the row of the secular determinant and b-vector
//                starts; writes the adjoint of each pair's start solution
//   k_minor_rows, k_minor_pairs  the same when the forward stored the minor
//                states: the rows' sequential chains (row states stored, one
//                pairing unit per interval), then every unit's step 2 on all
//                compute units
//   k_tor_adj    toroidal system, the same way
//   k_start_adj  starting solutions: forward-mode tangents with respect to
//                the spline values at the start radius
//   k_rad_prop, k_rad_scan, k_rad_seg_adj  the radial (l = 0) solves, cut
//                into per-interval segments (2 x 2 propagators), each
//                recomputed and reversed by its own lane
#include "ysx_adjoint.hpp"
#include "ysx_kern.hpp"
#include "ysx_radial.hpp"

namespace ysx {
namespace {

// ------------------------------------------------------------ systems
struct SysSolidMinor {
  using F = SolidMinorM;
  using FT = MinorTM;
  using Abar = SolidAbar;
  static constexpr int NOB = SOLID_OB, NV = SOLID_NV;
  __device__ static void bil(const cd* w, const cd* z, Abar& B) { solid_minor_bil(w, z, B); }
  template <class ZL>
  __device__ static void bil_s(const cd* w, ZL&& zl, Abar& B) { solid_minor_bil_s(w, zl, B); }
  __device__ static void ob(const Abar& B, const LS& L, const double* b, double* o) {
    solid_ob(B, L, b, o);
  }
  __device__ static void adj(const Tab& T, const WS& w, int il, double r, double h,
                             const double* ob, double* vb) {
    solid_base_adj(T, w, il, r, h, ob, vb);
  }
  __device__ static int field(int v) { return solid_field(v); }
};
struct SysBvec : SysSolidMinor {
  using F = BvecM;
  using FT = BvecTM;
  __device__ static void bil(const cd* w, const cd* z, Abar& B) { bvec_bil(w, z, B); }
  template <class ZL>
  __device__ static void bil_s(const cd* w, ZL&& zl, Abar& B) { bvec_bil_s(w, zl, B); }
};
struct SysFluidMinor {
  using F = FluidMinorM;
  using FT = FluidMinorT;
  using Abar = FluidAbar;
  static constexpr int NOB = FLUID_OB, NV = FLUID_NV;
  __device__ static void bil(const cd* w, const cd* z, Abar& B) { fluid_minor_bil(w, z, B); }
  template <class ZL>
  __device__ static void bil_s(const cd* w, ZL&& zl, Abar& B) { fluid_minor_bil_s(w, zl, B); }
  __device__ static void ob(const Abar& B, const LS& L, const double* b, double* o) {
    fluid_ob(B, L, b, o);
  }
  __device__ static void adj(const Tab& T, const WS& w, int il, double r, double h,
                             const double* ob, double* vb) {
    fluid_base_adj(T, w, il, r, h, ob, vb);
  }
  __device__ static int field(int v) { return fluid_field(v); }
};
struct SysTor {
  using F = TorM;
  using FT = TorT;
  using Abar = TorAbar;
  static constexpr int NOB = TOR_OB, NV = TOR_NV;
  __device__ static void bil(const cd* w, const cd* z, Abar& B) { tor_bil(w, z, B); }
  template <class ZL>
  __device__ static void bil_s(const cd* w, ZL&& zl, Abar& B) { tor_bil_s(w, zl, B); }
  __device__ static void ob(const Abar& B, const LS& L, const double* b, double* o) {
    tor_ob(B, L, b, o);
  }
  __device__ static void adj(const Tab& T, const WS& w, int il, double r, double h,
                             const double* ob, double* vb) {
    tor_base_adj(T, w, il, r, h, ob, vb);
  }
  __device__ static int field(int v) { return tor_field(v); }
};

constexpr int NOB_MAX = SOLID_OB;
constexpr int NV_MAX = SOLID_NV;

// H^n for the renormalisation counters of a row and a column (the
// factor between the stored product and the true one); beyond 15 powers
// one side is zero (a lane outside its solve, whose counter ran on), and
// the pairing counts nothing
__device__ inline double scale_pow(int n) {
  if (n > 15 || n < -15) return 0.0;
  return ipow_huge(n);
}

// ------------------------------------------------------ LDS of a block
// Steps per chunk of shared stage bases in the sweeps (the forward marches
// use CH): with the two stage slots it keeps a block within 40 KB, four
// blocks per CU.
constexpr int CHA = 6;

struct AdjLds {
  double sb[6 * CHA * BASE_MAX];   // stage bases of a chunk of steps
  double bbar[6 * CHA * NOB_MAX];  // the chunk's wave-summed stage adjoints
  double tacc[NV_MAX * 4];         // the interval's table-entry adjoints
  cd slots[2 * 14 * 64];           // rk_step_lds / rk_step_T_lds stage slots
};
static_assert(sizeof(AdjLds) <= 40 * 1024, "four blocks per CU");

// vectors of a block's step scratch (global memory, 14 x 64 complex each):
// the column's six stage inputs (0..5), the row's six stage rows (6..11),
// a parked 14-vector (12)
constexpr int SCR_ROWS = 6, SCR_PARK = 12, SCR_VECS = 13;
// Layout of one vector: groups of 16 lanes, component q of lane g 16 + t
// at byte g 3584 + q 256 + t 16 (256 contiguous bytes per component and
// group).  Accessed through a buffer resource: the vector's offset is a
// scalar, the component's an instruction immediate, the lane's one VGPR,
// so no address stays live across a step.
constexpr int SCR_VEC_BYTES = 4 * 3584;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
struct Scr {
  __amdgpu_buffer_rsrc_t r;
  int voff;
  __device__ explicit Scr(cd* block)
      : r(__builtin_amdgcn_make_buffer_rsrc(block, 0, SCR_VECS * SCR_VEC_BYTES, 0x00020000)),
        voff((int)(threadIdx.x >> 4) * 3584 + (int)(threadIdx.x & 15) * 16) {}
  __device__ __attribute__((always_inline)) void st(int vec, int q, cd v) const {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, voff + q * 256,
                                           vec * SCR_VEC_BYTES, 0);
  }
  __device__ __attribute__((always_inline)) cd ld(int vec, int q) const {
    return __builtin_bit_cast(
        cd, __builtin_amdgcn_raw_buffer_load_b128(r, voff + q * 256, vec * SCR_VEC_BYTES, 0));
  }
};

// bases of steps j0 .. j0 + nst - 1 (1-based) of the interval's step list,
// one lane per stage radius, as march_interval lays them out
template <class F>
__device__ inline void load_bases(const Tab& T, const WS& w, int i, double ra, double drr,
                                  int j0, int nst, double* sb) {
  int q = threadIdx.x;
  if (q < 6 * nst) {
    int jj = j0 + q / 6, k = q % 6;
    double x = ra + (double)(jj - 1) * drr;
    F::base(T, w, i, x + stage_offset(k) * drr, drr * stage_scale<F::N>(k), sb + q * F::BS);
  }
}

// Orders one wave's LDS accesses around an exchange between its lanes (a
// block of the sweeps is one wave, whose LDS instructions execute in order):
// a compiler barrier, no workgroup barrier.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sums over the wave of the per-lane stage adjoints of one step, batched:
// the values of up to SUM_ROWS / NOB stages go as rows into the stage slots
// (free once the step is taken; row length 65 doubles, so that one lane per
// row reads without bank conflicts), then lane t sums row t with eight
// partial sums.  put(k, ob) files stage k's values; the batch is summed
// into dst + k0 NOB when its last stage is filed.
constexpr int SUM_ROWS = 54, SUM_LD = 65;
static_assert(SUM_ROWS * SUM_LD * sizeof(double) <= 2 * 14 * 64 * sizeof(cd),
              "stage sums fit in the slots");
template <int NOB>
struct StageSums {
  static constexpr int SB = SUM_ROWS / NOB;  // stages per batch
  double* red;
  double* dst;
  __device__ void put(int k, const double* ob) const {
    const int row0 = (k % SB) * NOB, t = threadIdx.x;
#pragma unroll
    for (int v = 0; v < NOB; v++) red[(row0 + v) * SUM_LD + t] = ob[v];
    if (k % SB == SB - 1 || k == 5) {
      const int k0 = k - k % SB, rows = (k - k0 + 1) * NOB;
      wave_sync();
      if (t < rows) {
        const double* r = red + t * SUM_LD;
        double a[8];
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] = r[u];
#pragma unroll
        for (int m = 8; m < 64; m += 8)
#pragma unroll
          for (int u = 0; u < 8; u++) a[u] += r[m + u];
        dst[k0 * NOB + t] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
      }
      wave_sync();
    }
  }
};

// The pairings of one step: stage k's row (scratch vector SCR_ROWS + k,
// loaded) with its column stage input (scratch vector k, streamed), scaled
// by fs, summed over the wave into dst[k NOB .. (k + 1) NOB).
template <class Sys>
__device__ inline void pair_step(const Scr& scr, const LS& L, const double* bs, double fs,
                                 AdjLds& S, double* dst) {
  constexpr int N = Sys::F::N, BS = Sys::F::BS;
  // stage k + 1's row and column are loaded while stage k is paired and
  // summed, so the scratch latency overlaps the arithmetic
  const StageSums<Sys::NOB> sums{reinterpret_cast<double*>(S.slots), dst};
  cd kb[2][N], zs[2][N];
#pragma unroll
  for (int q = 0; q < N; q++) {
    kb[0][q] = scr.ld(SCR_ROWS, q);
    zs[0][q] = scr.ld(0, q);
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int c = k & 1;
    if (k + 1 < 6) {
#pragma unroll
      for (int q = 0; q < N; q++) {
        kb[c ^ 1][q] = scr.ld(SCR_ROWS + k + 1, q);
        zs[c ^ 1][q] = scr.ld(k + 1, q);
      }
    }
    typename Sys::Abar B;
    zero(B);
    Sys::bil_s(kb[c], [&](int q) { return zs[c][q]; }, B);
    double ob[Sys::NOB];
    Sys::ob(B, L, bs + k * BS, ob);
#pragma unroll
    for (int v = 0; v < Sys::NOB; v++) ob[v] *= fs;
    sums.put(k, ob);
  }
}

// scratch writers: hook(k, v) of a step stores v as scratch vector base + k
template <int N>
struct ScrWrite {
  const Scr& scr;
  int base;
  __device__ __attribute__((always_inline)) void operator()(int k, const cd* v) const {
#pragma unroll
    for (int q = 0; q < N; q++) scr.st(base + k, q, v[q]);
  }
};

template <int N>
__device__ inline void park(const Scr& scr, const cd* v) {
#pragma unroll
  for (int q = 0; q < N; q++) scr.st(SCR_PARK, q, v[q]);
  // the compiler reloads the vector from the scratch afterwards instead
  // of keeping it in registers across the replay
  asm volatile("" ::: "memory");
}
template <int N>
__device__ inline void unpark(const Scr& scr, cd* v) {
  asm volatile("" ::: "memory");
#pragma unroll
  for (int q = 0; q < N; q++) v[q] = scr.ld(SCR_PARK, q);
}

// The chunk's summed stage adjoints through the base computation to the
// interval's table entries (lane q: stage radius q), added to S.tacc: each
// lane's contributions to the NV x 4 entries (value and second derivative
// at the interval's two knots, spl_adj's weights) go through the stage
// slots (free between steps) as rows of a table that lane e sums down.
template <class Sys>
__device__ inline void chain_chunk(const Tab& T, const WS& w, int i, double ra, double drr,
                                   int j0, int nst, AdjLds& S) {
  using F = typename Sys::F;
  constexpr int NE = Sys::NV * 4, LD = NE + 1;
  static_assert(6 * CHA * LD * sizeof(double) <= sizeof(S.slots), "chain table in the slots");
  double* ct = reinterpret_cast<double*>(S.slots);
  const int q = threadIdx.x, nq = 6 * nst;
  if (q < nq) {
    int jj = j0 + q / 6, k = q % 6;
    double x = ra + (double)(jj - 1) * drr + stage_offset(k) * drr;
    double vb[Sys::NV];
    Sys::adj(T, w, i, x, drr * stage_scale<F::N>(k), S.bbar + q * Sys::NOB, vb);
    const SplW sw = spl_w(T, i, x);
    double* row = ct + q * LD;
#pragma unroll
    for (int v = 0; v < Sys::NV; v++) {
      row[4 * v + 0] = sw.deg ? vb[v] : vb[v] * sw.a;
      row[4 * v + 1] = sw.deg ? 0.0 : vb[v] * sw.b;
      row[4 * v + 2] = sw.deg ? 0.0 : vb[v] * sw.a3 * sw.hh;
      row[4 * v + 3] = sw.deg ? 0.0 : vb[v] * sw.b3 * sw.hh;
    }
  }
  wave_sync();
  if (q < NE) {
    double a = 0.0;
    for (int r = 0; r < nq; r++) a += ct[r * LD + q];
    S.tacc[q] += a;
  }
  wave_sync();
}

// S.tacc of interval i into the block's own column of the partial table
// adjoints, tpart[(f nk + knot - 1) gridDim.x + blockIdx.x] (plain
// read-modify-write: no other block writes that column; k_part_sum adds
// the columns into the table adjoint afterwards)
template <class Sys>
__device__ inline void flush_interval(int i, int nk, AdjLds& S, double* tpart) {
  __syncthreads();
  int t = threadIdx.x;
  if (t < Sys::NV * 4) {
    int v = t / 4, f = Sys::field(v) + ((t / 2) % 2), knot = i + (t % 2);
    double a = S.tacc[t];
    if (a != 0.0) {
      double* e = tpart + ((size_t)f * nk + knot - 1) * gridDim.x + blockIdx.x;
      *e += a;
    }
    S.tacc[t] = 0.0;
  }
  __syncthreads();
}

// tbar[e] += sum over the nb block columns of tpart[e nb ..], one wave per
// entry e < n
__global__ void __launch_bounds__(256) k_part_sum(const double* __restrict__ tpart, int nb,
                                                  size_t n, double* __restrict__ tbar) {
  const size_t e = (size_t)blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (e >= n) return;
  const double* p = tpart + e * nb;
  double a = 0.0;
  for (int b = lane; b < nb; b += 64) a += p[b];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
  if (lane == 0) tbar[e] += a;
}

// ---------------------------------------------- per-block device storage
// the stored side's state before every step of an item (index s in the
// order the first pass takes them, from step W.base on), and the step
// scratch: the six stage vectors of one step and a parked 14-vector
// (the 2-component toroidal states are packed when the forward stored
// them: W.ystride; the others always take 14 x 64 per step)
template <int N>
__device__ inline cd* ys_of(const AdjWork& W, int s) {
  const long long stride = N == 2 ? W.ystride : 14 * 64;
  return W.ys + (W.base + s) * stride + threadIdx.x;
}
__device__ inline int* yss_of(const AdjWork& W, int s) {
  return W.yss + (W.base + s) * 64 + threadIdx.x;
}
// a block's own per-step storage: smax steps from blockIdx.x smax on
__device__ inline AdjWork block_work(AdjWork W) {
  W.base = (long long)blockIdx.x * W.smax;
  return W;
}
__device__ inline Scr scr_of(const AdjWork& W) {
  return Scr(W.scr + (size_t)blockIdx.x * (SCR_VECS * SCR_VEC_BYTES / sizeof(cd)));
}
// more steps in an item than the per-step storage holds: flag it and
// describe the first one (interval, step index, radii, Re w)
__device__ inline void steps_overflow(const AdjWork& W, int i, int s, double ra, double rb,
                                      double rew) {
  if (threadIdx.x == 0 && !(atomicOr(W.err, 1) & 1)) {
    W.err[1] = i;
    double* d = reinterpret_cast<double*>(W.err + 2);
    d[0] = s;
    d[1] = ra;
    d[2] = rb;
    d[3] = rew;
  }
}

__device__ inline int steps_of(const Tab& T, const WS& w, int i, int nk, double ra, double rb,
                               double& drr) {
  double drr0 = step_len_w(T, w.rew, i, nk);
  int n = 2 + (int)floor(fabs(rb - ra) / drr0);
  drr = (rb - ra) / (double)(n - 1);
  return n - 1;
}

// the stored side's state y (N components, counter sc) before step s; the
// states stream through once each way, so they are written and read
// non-temporally and leave the caches to the step scratch
template <int N>
__device__ inline void put_step(const AdjWork& W, int s, const cd* y, int sc) {
  double* c = reinterpret_cast<double*>(ys_of<N>(W, s));
#pragma unroll
  for (int q = 0; q < N; q++) {
    __builtin_nontemporal_store(y[q].x, c + q * 128);
    __builtin_nontemporal_store(y[q].y, c + q * 128 + 1);
  }
  __builtin_nontemporal_store(sc, yss_of(W, s));
}

template <int N>
__device__ inline void load_step(const AdjWork& W, int s, cd* y, int& sc) {
  const double* c = reinterpret_cast<const double*>(ys_of<N>(W, s));
#pragma unroll
  for (int q = 0; q < N; q++)
    y[q] = C(__builtin_nontemporal_load(c + q * 128), __builtin_nontemporal_load(c + q * 128 + 1));
  sc = __builtin_nontemporal_load(yss_of(W, s));
}

// ------------------------------------------------------ first pass
// Interval i of the stored side's march, ra -> rb, as march_interval
// (forward steps, renormalised), its state stored before every step at
// s, s + 1, ...; s advances by the step count.
template <class F>
__device__ inline void adj_march(const Tab& T, const WS& w, const LS& L, int nk, int i,
                                 double ra, double rb, cd* y, int& sc, AdjLds& S, const F& f,
                                 const AdjWork& W, int& s) {
  constexpr int N = F::N, BS = F::BS;
  double drr;
  const int ns = steps_of(T, w, i, nk, ra, rb, drr);
  if (s + ns > W.smax) {
    steps_overflow(W, i, s + ns, ra, rb, w.rew);
    s += ns;
    return;
  }
  for (int j0 = 1; j0 <= ns; j0 += CHA) {
    const int nst = min(CHA, ns + 1 - j0);
    load_bases<F>(T, w, i, ra, drr, j0, nst, S.sb);
    __syncthreads();
    for (int jj = 0; jj < nst; jj++) {
      put_step<N>(W, s++, y, sc);
      if constexpr (N == 14) {
        rk_step_lds(S.sb + jj * 6 * BS, L, w, y, f, S.slots + threadIdx.x);
      } else {
        rk_step_b(S.sb + jj * 6 * BS, L, w, y, f);
      }
      renorm<N>(y, sc);
    }
    __syncthreads();
  }
}

// The same for a row marched rb -> ra by transposed steps (march_interval_T:
// steps ns, ns - 1, ..., 1), its state stored before every step.
template <class Sys>
__device__ inline void adj_march_T(const Tab& T, const WS& w, const LS& L, int nk, int i,
                                   double ra, double rb, cd* x, int& sl, AdjLds& S,
                                   const AdjWork& W, int& s) {
  using F = typename Sys::F;
  constexpr int BS = F::BS;
  double drr;
  const int ns = steps_of(T, w, i, nk, ra, rb, drr);
  if (s + ns > W.smax) {
    steps_overflow(W, i, s + ns, ra, rb, w.rew);
    s += ns;
    return;
  }
  const int jlast = 1 + ((ns - 1) / CHA) * CHA;
  for (int j0 = jlast; j0 >= 1; j0 -= CHA) {
    const int nst = min(CHA, ns + 1 - j0);
    load_bases<F>(T, w, i, ra, drr, j0, nst, S.sb);
    __syncthreads();
    for (int jj = nst - 1; jj >= 0; jj--) {
      put_step<14>(W, s++, x, sl);
      rk_step_T_lds(S.sb + jj * 6 * BS, L, w, x, typename Sys::FT(), S.slots + threadIdx.x);
      renorm<14>(x, sl);
    }
    __syncthreads();
  }
}

// ------------------------------------------------ one interval, row sweep
// Interval i of a march whose column went ra -> rb, its states before each
// step stored at s_end - ns .. s_end - 1: the row x (counter sx) is taken
// back through the steps, last first.  Per step the column's step is
// replayed from its stored state with every stage input written to the
// step scratch, then the row's transposed step pairs each stage's row with
// the scratch.  Returns s_end - ns.
template <class Sys>
__device__ __forceinline__ int rev_interval(const Tab& T, const WS& w, const LS& L, int nk, int i,
                                            double ra, double rb, cd* x, int& sx,
                                            const AdjWork& W, int s_end, AdjLds& S,
                                            double* tbar) {
  using F = typename Sys::F;
  constexpr int N = F::N, BS = F::BS, NOB = Sys::NOB;
  double drr;
  const int ns = steps_of(T, w, i, nk, ra, rb, drr);
  const int s0 = s_end - ns;
  if (s_end > W.smax) return s0;  // not stored (flagged by the first pass)
  const Scr scr = scr_of(W);
  cd yn[N];  // the stored state of the next step to replay, and its counter
  int ysc;
  load_step<N>(W, s_end - 1, yn, ysc);
  const int jlast = 1 + ((ns - 1) / CHA) * CHA;
  for (int j0 = jlast; j0 >= 1; j0 -= CHA) {
    const int nst = min(CHA, ns + 1 - j0);
    load_bases<F>(T, w, i, ra, drr, j0, nst, S.sb);
    __syncthreads();
    for (int jj = nst - 1; jj >= 0; jj--) {
      const int s = s0 + j0 - 1 + jj;
      const double* bs = S.sb + jj * 6 * BS;
      // the column's step again from its stored state (loaded during the
      // previous step's pairings), stage inputs into the scratch
      const int sc = ysc;
      if constexpr (N == 14) {
        {
          cd y[N];
#pragma unroll
          for (int q = 0; q < N; q++) y[q] = yn[q];
          // the row waits in the scratch: the replay needs the registers
          park<N>(scr, x);
          rk_step_lds(bs, L, w, y, F(), S.slots + threadIdx.x, ScrWrite<N>{scr, 0});
          unpark<N>(scr, x);
        }
        // the row's transposed step, stage rows into the scratch
        rk_step_T_lds(bs, L, w, x, typename Sys::FT(), S.slots + threadIdx.x,
                      ScrWrite<N>{scr, SCR_ROWS});
        const double fs = scale_pow(-(sx + sc));
        renorm<N>(x, sx);
        // the next step's stored state is loaded while this step is paired
        if (s > s0) load_step<N>(W, s - 1, yn, ysc);
        pair_step<Sys>(scr, L, bs, fs, S, S.bbar + jj * 6 * NOB);
      } else {
        // small systems: the six stage vectors of both sides in registers
        cd zst[6][N], kst[6][N];
        {
          cd y[N];
#pragma unroll
          for (int q = 0; q < N; q++) y[q] = yn[q];
          rk_step_hook(bs, L, w, y, F(), [&](int k, const cd* z) {
#pragma unroll
            for (int q = 0; q < N; q++) zst[k][q] = z[q];
          });
        }
        rk_step_T_reg(bs, L, w, x, typename Sys::FT(), [&](int k, const cd* kb) {
#pragma unroll
          for (int q = 0; q < N; q++) kst[k][q] = kb[q];
        });
        const double fs = scale_pow(-(sx + sc));
        renorm<N>(x, sx);
        if (s > s0) load_step<N>(W, s - 1, yn, ysc);
        const StageSums<NOB> sums{reinterpret_cast<double*>(S.slots), S.bbar + jj * 6 * NOB};
#pragma unroll
        for (int k = 0; k < 6; k++) {
          typename Sys::Abar B;
          zero(B);
          Sys::bil_s(kst[k], [&](int q) { return zst[k][q]; }, B);
          double ob[NOB];
          Sys::ob(B, L, bs + k * BS, ob);
#pragma unroll
          for (int v = 0; v < NOB; v++) ob[v] *= fs;
          sums.put(k, ob);
        }
      }
    }
    chain_chunk<Sys>(T, w, i, ra, drr, j0, nst, S);
    __syncthreads();
  }
  flush_interval<Sys>(i, nk, S, tbar);
  return s0;
}

// ------------------------------------ one interval, row chain alone
// rev_interval's sequential part for the two-kernel minor sweep: the row x
// (counter sx) taken back through the interval's steps, its state before
// every step stored in X at the step's index; the interval's pairings are
// left to k_minor_pairs as one PairUnit (event e, frequency j, degree
// block lb, interval i, part: 1 the march below the source, 0 the
// continuation to the receiver).  Returns s_end - ns.
template <class Sys>
__device__ __forceinline__ int rows_interval(const Tab& T, const WS& w, const LS& L, int nk, int i,
                                             double ra, double rb, cd* x, int& sx,
                                             const AdjWork& X, int s_end, AdjLds& S,
                                             const RowStore& RS, int e, int j, int lb, int part) {
  using F = typename Sys::F;
  constexpr int N = F::N, BS = F::BS;
  double drr;
  const int ns = steps_of(T, w, i, nk, ra, rb, drr);
  const int s0 = s_end - ns;
  if (s_end > X.smax) return s0;  // not stored (flagged by the forward)
  const int jlast = 1 + ((ns - 1) / CHA) * CHA;
  for (int j0 = jlast; j0 >= 1; j0 -= CHA) {
    const int nst = min(CHA, ns + 1 - j0);
    load_bases<F>(T, w, i, ra, drr, j0, nst, S.sb);
    __syncthreads();
    for (int jj = nst - 1; jj >= 0; jj--) {
      put_step<N>(X, s0 + j0 - 1 + jj, x, sx);
      if constexpr (N == 14) {
        rk_step_T_lds(S.sb + jj * 6 * BS, L, w, x, typename Sys::FT(), S.slots + threadIdx.x);
      } else {
        rk_step_T_reg(S.sb + jj * 6 * BS, L, w, x, typename Sys::FT(), NoHook());
      }
      renorm<N>(x, sx);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int k = atomicAdd(RS.count, 1);
    if (k < RS.cap)
      RS.units[k] = PairUnit{e, j, lb, i | (part << 16), s0, ns, 0, 0};
    else
      atomicOr(X.err, 4);
  }
  return s0;
}

// ------------------------------- one interval segment, pairings alone
// The other part of rev_interval for the two-kernel minor sweep: steps s0
// .. s0 + ns - 1 of an item, whose column states (W) and row states (X)
// are stored; per step the column's step and the row's transposed step are
// replayed from them and the stages paired, summed over the wave and taken
// to the interval's table entries.  The steps are independent.
template <class Sys>
__device__ __forceinline__ void pairs_interval(const Tab& T, const WS& w, const LS& L, int nk, int i,
                                               double ra, double rb, const AdjWork& W,
                                               const AdjWork& X, int s0, AdjLds& S,
                                               double* tbar) {
  using F = typename Sys::F;
  constexpr int N = F::N, BS = F::BS, NOB = Sys::NOB;
  double drr;
  const int ns = steps_of(T, w, i, nk, ra, rb, drr);
  if (s0 + ns > W.smax) return;
  const Scr scr = scr_of(W);
  for (int j0 = 1; j0 <= ns; j0 += CHA) {
    const int nst = min(CHA, ns + 1 - j0);
    load_bases<F>(T, w, i, ra, drr, j0, nst, S.sb);
    __syncthreads();
    for (int jj = 0; jj < nst; jj++) {
      const int s = s0 + j0 - 1 + jj;
      const double* bs = S.sb + jj * 6 * BS;
      int ysc, xsc;
      if constexpr (N == 14) {
        {
          cd y[N];
          load_step<N>(W, s, y, ysc);
          rk_step_lds(bs, L, w, y, F(), S.slots + threadIdx.x, ScrWrite<N>{scr, 0});
        }
        {
          cd xr[N];
          load_step<N>(X, s, xr, xsc);
          rk_step_T_lds(bs, L, w, xr, typename Sys::FT(), S.slots + threadIdx.x,
                        ScrWrite<N>{scr, SCR_ROWS});
        }
        pair_step<Sys>(scr, L, bs, scale_pow(-(xsc + ysc)), S, S.bbar + jj * 6 * NOB);
      } else {
        cd zst[6][N], kst[6][N];
        {
          cd y[N];
          load_step<N>(W, s, y, ysc);
          rk_step_hook(bs, L, w, y, F(), [&](int k, const cd* z) {
#pragma unroll
            for (int q = 0; q < N; q++) zst[k][q] = z[q];
          });
        }
        {
          cd xr[N];
          load_step<N>(X, s, xr, xsc);
          rk_step_T_reg(bs, L, w, xr, typename Sys::FT(), [&](int k, const cd* kb) {
#pragma unroll
            for (int q = 0; q < N; q++) kst[k][q] = kb[q];
          });
        }
        const double fs = scale_pow(-(xsc + ysc));
        const StageSums<NOB> sums{reinterpret_cast<double*>(S.slots), S.bbar + jj * 6 * NOB};
#pragma unroll
        for (int k = 0; k < 6; k++) {
          typename Sys::Abar B;
          zero(B);
          Sys::bil_s(kst[k], [&](int q) { return zst[k][q]; }, B);
          double ob[NOB];
          Sys::ob(B, L, bs + k * BS, ob);
#pragma unroll
          for (int v = 0; v < NOB; v++) ob[v] *= fs;
          sums.put(k, ob);
        }
      }
    }
    chain_chunk<Sys>(T, w, i, ra, drr, j0, nst, S);
    __syncthreads();
  }
  flush_interval<Sys>(i, nk, S, tbar);
}

// ------------------------------------------- one interval, column sweep
// Interval i of a march whose column goes ra -> rb and whose row came
// rb -> ra by transposed steps, the row's states before each of its steps
// stored at s_end - ns .. s_end - 1 (step ns first): the column c (counter
// sc) is stepped through the interval with every stage input written to
// the step scratch; per step the row's transposed step is then replayed
// from its stored state, each stage's row paired with the scratch.
// Returns s_end - ns.
template <class Sys>
__device__ __forceinline__ int fwd_interval(const Tab& T, const WS& w, const LS& L, int nk, int i,
                                            double ra, double rb, cd* c, int& sc,
                                            const AdjWork& W, int s_end, AdjLds& S,
                                            double* tbar) {
  using F = typename Sys::F;
  constexpr int BS = F::BS, NOB = Sys::NOB;
  static_assert(F::N == 14, "the column sweep is the b-vector system's");
  double drr;
  const int ns = steps_of(T, w, i, nk, ra, rb, drr);
  const int s0 = s_end - ns;
  if (s_end > W.smax) return s0;  // not stored (flagged by the first pass)
  const Scr scr = scr_of(W);
  for (int j0 = 1; j0 <= ns; j0 += CHA) {
    const int nst = min(CHA, ns + 1 - j0);
    load_bases<F>(T, w, i, ra, drr, j0, nst, S.sb);
    __syncthreads();
    for (int jj = 0; jj < nst; jj++) {
      // the row's state before its transposed step j = j0 + jj
      const int s = s0 + (ns - (j0 + jj));
      const double* bs = S.sb + jj * 6 * BS;
      const double fs = scale_pow(-(sc + *yss_of(W, s)));
      // the column's step, stage inputs into the scratch
      rk_step_lds(bs, L, w, c, F(), S.slots + threadIdx.x, ScrWrite<14>{scr, 0});
      renorm<14>(c, sc);
      park<14>(scr, c);
      // the row's transposed step again from its stored state, stage rows
      // into the scratch
      {
        cd xr[14];
        int xsc;
        load_step<14>(W, s, xr, xsc);
        rk_step_T_lds(bs, L, w, xr, typename Sys::FT(), S.slots + threadIdx.x,
                      ScrWrite<14>{scr, SCR_ROWS});
      }
      pair_step<Sys>(scr, L, bs, fs, S, S.bbar + jj * 6 * NOB);
      unpark<14>(scr, c);
    }
    chain_chunk<Sys>(T, w, i, ra, drr, j0, nst, S);
    __syncthreads();
  }
  flush_interval<Sys>(i, nk, S, tbar);
  return s0;
}

// ------------------------------------------------------- record adjoints
__device__ inline cd conj_(cd a) { return C(a.x, -a.y); }

// rows b of b_solid_start's map bst = U_k m for the unit source column e_k
// (b_start_unit): mrow += coef * (x U_k)
__device__ inline void b_start_unit_T(const cd* x, int k, cd coef, cd* mrow) {
  auto add = [&](int q, cd v) { mrow[q] = cmad(mrow[q], coef, v); };
  if (k == 0) {
    add(13, x[5]); add(6, -x[6]); add(10, -x[7]); add(12, -x[8]); add(7, -x[9]);
    add(11, -x[10]); add(1, x[11]); add(4, x[12]); add(9, x[13]);
  } else if (k == 1) {
    add(13, -x[1]); add(6, x[2]); add(10, x[3]); add(12, x[4]); add(8, -x[9]);
    add(7, x[10]); add(2, x[11]); add(5, x[12]); add(4, -x[13]);
  } else if (k == 3) {
    add(6, -x[0]); add(7, -x[1]); add(1, -x[3]); add(4, -x[4]); add(8, -x[5]);
    add(2, -x[7]); add(5, -x[8]); add(3, -x[10]); add(0, -x[13]);
  } else {
    add(10, -x[0]); add(11, -x[1]); add(1, x[2]); add(9, -x[4]); add(7, x[5]);
    add(2, x[6]); add(4, x[8]); add(3, x[9]); add(0, x[12]);
  }
}

// ---------------------------------------------------------- k_brow_adj
// The b-vector part of the gradient (sph_solver's b-vector integration,
// src/module_solver.f90:745-755, as the receiver-side row march of
// k_brow_march): with the rows x_k (k < rows) marched from the receiver to
// the source by Q's transposed steps and y_kc = fac H^(sc3 - sl) x_k Q bst_c,
// the record adjoints a_kc = conj(ybar_kc) give
//   dJ = Re sum_k x_k dQ beta_k,  beta_k = fac H^sc3 sum_c a_kc bst_c,
// a column marched from the source up.  The rows are the stored side
// (checkpoints at each interval's top); the columns sweep up.  Also writes
// mub[p], the adjoint of the minors at the source through the b-vector
// starts: sum_kc a_kc fac H^(sc3 - sl) (x_k Q) U_c.
__device__ __forceinline__ void brow_adj_item(const Tab& T, const RunK& R, int j, int lb,
                                              const SphState& SS, const cd* __restrict__ rec,
                                              const cd* __restrict__ recb, cd* __restrict__ mub,
                                              const AdjWork& W, AdjLds& S, double* tbar) {
  const int rows = 64 / R.bl;
  const int k = threadIdx.x % rows;
  const int l = R.lmin + lb * R.bl + (int)(threadIdx.x / rows);
  const bool valid = (l >= 1 && l <= R.lmax);
  if (wave_min(valid ? 0 : 1)) return;
  const size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
  const WS w = wave_omega(R, j);
  const LS L = l_state(valid ? l : 1);
  const int nk = R.nknot, is = R.is, ir = R.ir;
  const double rs = R.rs, rr = R.rr;
  cd x[14];
  {
    cd m[14];
    for (int q = 0; q < 14; q++) m[q] = valid ? SS.mfin[p * 14 + q] : C(0.0);
    ysol_row(m, k, x);
  }
  // 1. the rows down from the receiver, their state stored before every step
  int sl = 0, s = 0;
  for (int i = ir; i >= is; i--) {
    double a = (i == is) ? rs : T(F_R, i), b = (i == ir) ? rr : T(F_R, i + 1);
    if (a != b) adj_march_T<SysBvec>(T, w, L, nk, i, a, b, x, sl, S, W, s);
  }
  __syncthreads();
  // record adjoints of this lane's row
  cd acoef[4];
  cd fac = valid ? SS.fac[p] : C(0.0);
  const int sc3 = valid ? SS.sc3[p] : 0;
  for (int c = 0; c < 4; c++) {
    cd a = (valid && k < 3) ? conj_(recb[p * REC + O_YSPH + 6 * c + k]) : C(0.0);
    acoef[c] = cmul_(a, fac);
  }
  // the minors' adjoint at the source through the b-vector starts
  {
    const int cols[4] = {0, 1, 3, 4};
    cd mrow[14];
    for (int q = 0; q < 14; q++) mrow[q] = C(0.0);
    const double hs = scale_pow(sc3 - sl);
    for (int c = 0; c < 4; c++) b_start_unit_T(x, cols[c], acoef[c] * hs, mrow);
    for (int o = 1; o < rows; o <<= 1)
      for (int q = 0; q < 14; q++) {
        mrow[q].x += __shfl_xor(mrow[q].x, o, 64);
        mrow[q].y += __shfl_xor(mrow[q].y, o, 64);
      }
    if (valid && k == 0)
      for (int q = 0; q < 14; q++) mub[p * 14 + q] = mrow[q];
  }
  // 2. the column beta_k = fac sum_c a_kc bst_c (true value H^sc3 times
  // this: counter -sc3) up from the source, paired with the rows; the
  // intervals come in the reverse of the first pass's order
  cd bcol[14];
  for (int q = 0; q < 14; q++) {
    cd v = C(0.0);
    if (valid)
      for (int c = 0; c < 4; c++) v = cmad(v, acoef[c], SS.bst[p * 56 + c * 14 + q]);
    bcol[q] = v;
  }
  int scb = -sc3;
  for (int i = is; i <= ir; i++) {
    double a = (i == is) ? rs : T(F_R, i), b = (i == ir) ? rr : T(F_R, i + 1);
    if (a != b) s = fwd_interval<SysBvec>(T, w, L, nk, i, a, b, bcol, scb, W, s, S, tbar);
  }
}

template <bool PERSIST>
__global__ void __launch_bounds__(64) k_brow_adj(const double* __restrict__ gtab,
                                                 const RunK* __restrict__ runs, Queue Q,
                                                 SphState SS, const cd* __restrict__ rec,
                                                 const cd* __restrict__ recb, cd* __restrict__ mub,
                                                 AdjWork W, double* __restrict__ tbar) {
  __shared__ AdjLds S;
  if (threadIdx.x < NV_MAX * 4) S.tacc[threadIdx.x] = 0.0;
  __syncthreads();
  const AdjWork WB = block_work(W);
  RUN_ITEMS(PERSIST, Q,
            brow_adj_item(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, it.z, SS, rec, recb, mub,
                          WB, S, tbar));
}

// --------------------------------------------------------- k_minor_adj
// The minor part (sph_solver, grav = 2, receiver above the source,
// src/module_solver.f90:640-743): the column is the forward's minor march
// (ms1 from the start level to the source, ms3 on to the receiver); the
// records depend on it through the secular determinant (fac = -1/delta,
// delta = d . ms3 with d from the surface minors) and the b-vector starts
// (bst_c = U_c ms1).  The row starts at the receiver as G d (G = fac sum_kc
// a_kc y_kc), is carried back to the source, gets mub (k_brow_adj) added,
// and is carried on down to each lane's start level, where it is the
// adjoint of the lane's starting minors (m0bar, for k_start_adj).
__device__ __forceinline__ bool fluid_at(const RunK& R, int i) {
  return R.fic ? (i < R.noc) : (i > R.nic && i < R.noc);
}

// ROWS (with the forward's stored states only): the row chain alone
// (rows_interval), the pairings left to k_minor_pairs through RS
template <bool ROWS = false>
__device__ __forceinline__ void minor_adj_item(const Tab& T, const RunK& R, int e, int j, int lb,
                                               const StartData& SD, const SphState& SS,
                                               const cd* __restrict__ rec,
                                               const cd* __restrict__ recb,
                                               const cd* __restrict__ mub, cd* __restrict__ m0bar,
                                               AdjWork W, const StepStore& FS, AdjLds& S,
                                               double* tbar, const RowStore& RS = RowStore{}) {
  const int l = R.lmin + lb * 64 + (int)threadIdx.x;
  const bool valid = (l >= 1 && l <= R.lmax);
  const size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
  const int i1 = valid ? SD.sph_i1[p] : NOLANE;
  const int ib = wave_min(i1);
  if (ib == NOLANE) return;
  const WS w = wave_omega(R, j);
  const LS L = l_state(valid ? l : 1);
  const int nk = R.nknot, is = R.is, ir = R.ir;
  const double rs = R.rs, rr = R.rr;
  const double rt1 = fmin(0.1 * (1.0 / (4.0 * w.rew / R.alpha_min)), 0.1 * T(F_R, 2));
  int sc1 = 0, sc3 = 0, s = 0, sci1 = 0;
  if (ROWS || FS.ys) {
    // 1. the forward stored every step's state (k_minor_march, STORE) and
    // the counters: the item's steps are counted, not marched again
    W.ys = FS.ys;
    W.yss = FS.yss;
    W.smax = FS.cap[j - 1];
    W.base = FS.off[j - 1] + (long long)lb * W.smax;
    double drr;
    for (int i = ib; i <= is; i++) {
      const double a = (i == 1) ? rt1 : T(F_R, i), b = (i == is) ? rs : T(F_R, i + 1);
      if (a != b) s += steps_of(T, w, i, nk, a, b, drr);
    }
    for (int i = is; i <= ir; i++) {
      const double a = (i == is) ? rs : T(F_R, i), b = (i == ir) ? rr : T(F_R, i + 1);
      if (a != b) s += steps_of(T, w, i, nk, a, b, drr);
    }
    if (valid) {
      sc1 = FS.sc[3 * p];
      sc3 = FS.sc[3 * p + 1];
      sci1 = FS.sc[3 * p + 2];
    }
  } else if constexpr (!ROWS) {
    // 1. the forward's march (minor_item), every step's state stored, the
    // counter at the lane's start kept for the start adjoint
    cd m[14];
    for (int q = 0; q < 14; q++) m[q] = C(0.0);
    int layout = -1;
    for (int i = ib; i <= is; i++) {
      const bool fl = fluid_at(R, i);
      if (layout == 0 && fl) {
        cd mf[5];
        solid_2_fluid(m, mf);
        for (int q = 0; q < 14; q++) m[q] = q < 5 ? mf[q] : C(0.0);
      } else if (layout == 1 && !fl) {
        cd mf[5] = {m[0], m[1], m[2], m[3], m[4]};
        fluid_2_solid(mf, m);
      }
      layout = fl ? 1 : 0;
      if (i == i1) {
        for (int q = 0; q < 14; q++) m[q] = SD.sph_m0[p * 14 + q];
        sci1 = sc1;
      }
      const double a = (i == 1) ? rt1 : T(F_R, i), b = (i == is) ? rs : T(F_R, i + 1);
      if (a == b) continue;
      if (fl)
        adj_march(T, w, L, nk, i, a, b, m, sc1, S, FluidMinorM(), W, s);
      else
        adj_march(T, w, L, nk, i, a, b, m, sc1, S, SolidMinorM(), W, s);
    }
    for (int i = is; i <= ir; i++) {
      const double a = (i == is) ? rs : T(F_R, i), b = (i == ir) ? rr : T(F_R, i + 1);
      if (a != b) adj_march(T, w, L, nk, i, a, b, m, sc3, S, SolidMinorM(), W, s);
    }
  }
  __syncthreads();
  if (threadIdx.x < NV_MAX * 4) S.tacc[threadIdx.x] = 0.0;
  __syncthreads();
  // the row at the receiver: G d on ms3 (true value H^sc3 times: counter -sc3)
  cd x[14];
  {
    const int rows = (R.out == 2 && R.cor != 0) ? 3 : 2;
    cd g = C(0.0);
    if (valid)
      for (int c = 0; c < 4; c++)
        for (int k = 0; k < rows; k++) {
          const size_t f = p * REC + O_YSPH + 6 * c + k;
          g = cmad(g, conj_(recb[f]), rec[f]);
        }
    // fac H^SS.sc3 is the same for every forward (the batched path's
    // k_drow_march stores fac with its own row's count in sc3); here it
    // multiplies the continuation re-marched above, counter sc3
    g = valid ? cmul_(SS.fac[p] * scale_pow(SS.sc3[p] - sc3), g) : C(0.0);
    cd m2[14];
    for (int q = 0; q < 14; q++) m2[q] = valid ? SS.mfin[p * 14 + q] : C(0.0);
    delta_row(m2, x);
    for (int q = 0; q < 14; q++) x[q] = cmul_(g, x[q]);
  }
  int sx = -sc3;
  // the rows' storage (ROWS): the item's steps as the columns'
  AdjWork X = W;
  if constexpr (ROWS) {
    X.ys = RS.xs;
    X.yss = RS.xss;
  }
  // 2. back from the receiver to the source
  for (int i = ir; i >= is; i--) {
    const double a = (i == is) ? rs : T(F_R, i), b = (i == ir) ? rr : T(F_R, i + 1);
    if (a == b) continue;
    if constexpr (ROWS)
      s = rows_interval<SysSolidMinor>(T, w, L, nk, i, a, b, x, sx, X, s, S, RS, e, j, lb, 0);
    else
      s = rev_interval<SysSolidMinor>(T, w, L, nk, i, a, b, x, sx, W, s, S, tbar);
  }
  // at the source: the row on ms1, plus the b-vector starts' part
  {
    const double h = scale_pow(-sx);
    for (int q = 0; q < 14; q++)
      x[q] = valid ? x[q] * h + mub[p * 14 + q] : C(0.0);
    sx = -sc1;
  }
  // 3. back to the start levels
  for (int i = is; i >= ib; i--) {
    const bool fl = fluid_at(R, i);
    const double a = (i == 1) ? rt1 : T(F_R, i), b = (i == is) ? rs : T(F_R, i + 1);
    if (a != b) {
      if constexpr (ROWS) {
        if (fl)
          s = rows_interval<SysFluidMinor>(T, w, L, nk, i, a, b, x, sx, X, s, S, RS, e, j, lb, 1);
        else
          s = rows_interval<SysSolidMinor>(T, w, L, nk, i, a, b, x, sx, X, s, S, RS, e, j, lb, 1);
      } else {
        if (fl)
          s = rev_interval<SysFluidMinor>(T, w, L, nk, i, a, b, x, sx, W, s, S, tbar);
        else
          s = rev_interval<SysSolidMinor>(T, w, L, nk, i, a, b, x, sx, W, s, S, tbar);
      }
    }
    if (i == i1) {
      const double h = scale_pow(-(sx + sci1));
      for (int q = 0; q < 14; q++) {
        m0bar[p * 14 + q] = x[q] * h;
        x[q] = C(0.0);
      }
    }
    if (i > ib && fluid_at(R, i - 1) != fl) {
      if (fl) {  // the forward went solid -> fluid here
        cd xf[5] = {x[0], x[1], x[2], x[3], x[4]};
        for (int q = 0; q < 14; q++) x[q] = C(0.0);
        x[5] = -xf[0];
        x[6] = -xf[1];
        x[8] = xf[2];
        x[12] = -xf[3];
        x[13] = xf[4];
      } else {  // fluid -> solid
        cd xf[5] = {x[0], x[1], x[3], -x[9], -x[11]};
        for (int q = 0; q < 14; q++) x[q] = q < 5 ? xf[q] : C(0.0);
      }
    }
  }
}

template <bool PERSIST>
__global__ void __launch_bounds__(64) k_minor_adj(const double* __restrict__ gtab,
                                                  const RunK* __restrict__ runs, Queue Q,
                                                  StartData SD, SphState SS,
                                                  const cd* __restrict__ rec,
                                                  const cd* __restrict__ recb,
                                                  const cd* __restrict__ mub,
                                                  cd* __restrict__ m0bar, AdjWork W,
                                                  StepStore FS, double* __restrict__ tbar) {
  __shared__ AdjLds S;
  const AdjWork WB = block_work(W);
  RUN_ITEMS(PERSIST, Q,
            minor_adj_item(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.x, it.y, it.z, SD, SS, rec,
                           recb, mub, m0bar, WB, FS, S, tbar));
}

// The minor sweep's row chains with the forward's stored states (FS): per
// item as k_minor_adj, the pairings of each interval appended to RS
__global__ void __launch_bounds__(64) k_minor_rows(const double* __restrict__ gtab,
                                                   const RunK* __restrict__ runs, Queue Q,
                                                   StartData SD, SphState SS,
                                                   const cd* __restrict__ rec,
                                                   const cd* __restrict__ recb,
                                                   const cd* __restrict__ mub,
                                                   cd* __restrict__ m0bar, AdjWork W,
                                                   StepStore FS, RowStore RS,
                                                   double* __restrict__ tbar) {
  __shared__ AdjLds S;
  const AdjWork WB = block_work(W);
  RUN_ITEMS(true, Q,
            minor_adj_item<true>(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.x, it.y, it.z, SD,
                                 SS, rec, recb, mub, m0bar, WB, FS, S, tbar, RS));
}

// The pairings of every unit k_minor_rows appended (persistent blocks take
// them one at a time from RS.next until RS.count), each into the block's
// own partial table column
__global__ void __launch_bounds__(64) k_minor_pairs(const double* __restrict__ gtab,
                                                    const RunK* __restrict__ runs, AdjWork W,
                                                    StepStore FS, RowStore RS,
                                                    double* __restrict__ tbar) {
  __shared__ AdjLds S;
  if (threadIdx.x < NV_MAX * 4) S.tacc[threadIdx.x] = 0.0;
  __syncthreads();
  const int n = min(*RS.count, RS.cap);
  while (true) {
    int k = 0;
    if (threadIdx.x == 0) k = atomicAdd(RS.next, 1);
    k = __builtin_amdgcn_readfirstlane(__shfl(k, 0, 64));
    if (k >= n) break;
    const PairUnit u = RS.units[k];
    const RunK& R = runs[u.e];
    const Tab T{gtab, R.nknot};
    const int j = u.j, lb = u.lb, i = u.ik & 0xffff, part = u.ik >> 16;
    const int l = R.lmin + lb * 64 + (int)threadIdx.x;
    const bool valid = (l >= 1 && l <= R.lmax);
    const WS w = wave_omega(R, j);
    const LS L = l_state(valid ? l : 1);
    const int nk = R.nknot, is = R.is, ir = R.ir;
    AdjWork CW = W;  // the item's column states, and its row states
    CW.ys = FS.ys;
    CW.yss = FS.yss;
    CW.smax = FS.cap[j - 1];
    CW.base = FS.off[j - 1] + (long long)lb * CW.smax;
    AdjWork XW = CW;
    XW.ys = RS.xs;
    XW.yss = RS.xss;
    double a, b;
    bool fl = false;
    if (part == 1) {
      const double rt1 = fmin(0.1 * (1.0 / (4.0 * w.rew / R.alpha_min)), 0.1 * T(F_R, 2));
      a = (i == 1) ? rt1 : T(F_R, i);
      b = (i == is) ? R.rs : T(F_R, i + 1);
      fl = fluid_at(R, i);
    } else {
      a = (i == is) ? R.rs : T(F_R, i);
      b = (i == ir) ? R.rr : T(F_R, i + 1);
    }
    if (fl)
      pairs_interval<SysFluidMinor>(T, w, L, nk, i, a, b, CW, XW, u.s0, S, tbar);
    else
      pairs_interval<SysSolidMinor>(T, w, L, nk, i, a, b, CW, XW, u.s0, S, tbar);
  }
}

// ----------------------------------------------------------- k_tor_adj
// tor_solver (src/module_solver.f90:25-142), receiver above the source and
// at the solid surface: ytor(q, c) = -H^sc3 y2_q b_c / D with y2 = (1, 0)
// at the surface, b = (y1_2, -y1_1) from the march y1 from the start to
// the source and D = y2_2 y3_1 - y2_1 y3_2 from its continuation y3 to the
// receiver.  The row starts on y3 as (0, T/D), T = sum a_qc ytor_qc, gets
// the b part at the source and ends as each lane's start adjoint (y0bar).
__device__ __forceinline__ void tor_adj_item(const Tab& T, const RunK& R, int j, int lb,
                                             const StartData& SD, const cd* __restrict__ rec,
                                             const cd* __restrict__ recb, cd* __restrict__ y0bar,
                                             AdjWork W, const StepStore& FS, AdjLds& S,
                                             double* tbar) {
  const int l = R.lmin + lb * 64 + (int)threadIdx.x;
  const bool valid = (l >= 2 && l <= R.lmax);
  const size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (j - 1);
  const int i1 = valid ? SD.tor_i1[p] : NOLANE;
  const int ib = wave_min(i1);
  if (ib == NOLANE) return;
  const WS w = wave_omega(R, j);
  const LS L = l_state(valid ? l : 2);
  const int nk = R.nknot, is = R.is, ir = R.ir;
  const double rs = R.rs, rr = R.rr;
  int sc1 = 0, sc3 = 0, s = 0, sci1 = 0;
  cd y3[2] = {C(0.0), C(0.0)};
  if (FS.ys) {
    // 1. the forward stored every step's state (k_tor_march, STORE), the
    // counters and the continuation's end: the steps are counted only
    W.ys = FS.ys;
    W.yss = FS.yss;
    W.smax = FS.cap[j - 1];
    W.base = FS.off[j - 1] + (long long)lb * W.smax;
    W.ystride = 2 * 64;
    double drr;
    for (int i = ib; i <= is; i++) {
      const double a = T(F_R, i), b = (i == is) ? rs : T(F_R, i + 1);
      if (a != b) s += steps_of(T, w, i, nk, a, b, drr);
    }
    for (int i = is; i <= ir; i++) {
      const double a = (i == is) ? rs : T(F_R, i), b = (i == ir) ? rr : T(F_R, i + 1);
      if (a != b) s += steps_of(T, w, i, nk, a, b, drr);
    }
    if (valid) {
      sc1 = FS.sc[3 * p];
      sc3 = FS.sc[3 * p + 1];
      sci1 = FS.sc[3 * p + 2];
      y3[0] = FS.yend[2 * p];
      y3[1] = FS.yend[2 * p + 1];
    }
  } else {
    // 1. the forward's marches (tor_item), every step's state stored, the
    // counter at the lane's start kept for the start adjoint
    cd y1[2] = {C(0.0), C(0.0)};
    for (int i = ib; i <= is; i++) {
      if (i == i1) {
        y1[0] = SD.tor_y0[2 * p];
        y1[1] = SD.tor_y0[2 * p + 1];
        sci1 = sc1;
      }
      const double a = T(F_R, i), b = (i == is) ? rs : T(F_R, i + 1);
      if (a == b) continue;
      adj_march(T, w, L, nk, i, a, b, y1, sc1, S, TorM(), W, s);
    }
    y3[0] = y1[0];
    y3[1] = y1[1];
    for (int i = is; i <= ir; i++) {
      const double a = (i == is) ? rs : T(F_R, i), b = (i == ir) ? rr : T(F_R, i + 1);
      if (a != b) adj_march(T, w, L, nk, i, a, b, y3, sc3, S, TorM(), W, s);
    }
  }
  __syncthreads();
  if (threadIdx.x < NV_MAX * 4) S.tacc[threadIdx.x] = 0.0;
  __syncthreads();
  cd a00 = C(0.0), a01 = C(0.0), tsum = C(0.0);
  if (valid) {
    for (int i = 0; i < 4; i++) {
      const size_t f = p * REC + O_YTOR + i;
      tsum = cmad(tsum, conj_(recb[f]), rec[f]);
    }
    a00 = conj_(recb[p * REC + O_YTOR]);
    a01 = conj_(recb[p * REC + O_YTOR + 2]);
  }
  const cd dinvv = valid ? cinv(-y3[1]) : C(0.0);  // 1 / D
  cd x[2] = {C(0.0), cmul_(tsum, dinvv)};
  int sx = -sc3;
  for (int i = ir; i >= is; i--) {
    const double a = (i == is) ? rs : T(F_R, i), b = (i == ir) ? rr : T(F_R, i + 1);
    if (a != b) s = rev_interval<SysTor>(T, w, L, nk, i, a, b, x, sx, W, s, S, tbar);
  }
  {
    const double h = scale_pow(-sx);
    const cd hd = scale_pow(sc3) * dinvv;
    x[0] = x[0] * h + cmul_(a01, hd);
    x[1] = x[1] * h - cmul_(a00, hd);
    sx = -sc1;
  }
  for (int i = is; i >= ib; i--) {
    const double a = T(F_R, i), b = (i == is) ? rs : T(F_R, i + 1);
    if (a != b) s = rev_interval<SysTor>(T, w, L, nk, i, a, b, x, sx, W, s, S, tbar);
    if (i == i1) {
      const double h = scale_pow(-(sx + sci1));
      y0bar[2 * p] = x[0] * h;
      y0bar[2 * p + 1] = x[1] * h;
      x[0] = x[1] = C(0.0);
    }
  }
}

template <bool PERSIST>
__global__ void __launch_bounds__(64) k_tor_adj(const double* __restrict__ gtab,
                                                const RunK* __restrict__ runs, Queue Q,
                                                StartData SD, const cd* __restrict__ rec,
                                                const cd* __restrict__ recb,
                                                cd* __restrict__ y0bar, AdjWork W,
                                                StepStore FS, double* __restrict__ tbar) {
  __shared__ AdjLds S;
  const AdjWork WB = block_work(W);
  RUN_ITEMS(PERSIST, Q,
            tor_adj_item(Tab{gtab, runs[it.x].nknot}, runs[it.x], it.y, it.z, SD, rec, recb,
                         y0bar, WB, FS, S, tbar));
}

// --------------------------------------------------------- k_start_adj
// The starting solutions' part: for every pair, Re(m0bar . dm0/dv) and
// Re(y0bar . dy0/dv) for the spline values v at the start radius (the
// forward's k_start choices, src/module_solver.f90:64-75, 678-722),
// spread onto the start interval's table entries.  The block sums into
// LDS (fields rho, mu, 1/Qmu, kappa, 1/Qkappa and their second
// derivatives, nknot each), then into the table adjoint.
__global__ void __launch_bounds__(256) k_start_adj(const double* __restrict__ gtab,
                                                   const RunK* __restrict__ runs, PairRange PR,
                                                   StartData SD, const int* __restrict__ starts,
                                                   const cd* __restrict__ m0bar,
                                                   const cd* __restrict__ y0bar,
                                                   double* __restrict__ tbar) {
  extern __shared__ double acc[];  // [SV_N][2][nknot]
  const int nk = runs[PR.e0].nknot;
  for (int q = threadIdx.x; q < SV_N * 2 * nk; q += blockDim.x) acc[q] = 0.0;
  __syncthreads();
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < PR.n) {
    const long long g = PR.g0 + t;
    const RunK& R = runs[find_event(runs, PR.e0, PR.e1, g)];
    const Tab T{gtab, R.nknot};
    int l, j;
    pair_lj(R, (int)(g - R.pbase), l, j);
    const size_t p = (size_t)g;
    auto put = [&](const SplW& sw, const double* vb) {
      for (int v = 0; v < SV_N; v++) {
        const int f0 = start_field(v);
        spl_adj(sw, f0, vb[v], [&](int f, int knot, double val) {
          atomicAdd(&acc[(v * 2 + (f == f0 ? 0 : 1)) * nk + knot - 1], val);
        });
      }
    };
    if (l >= 1) {
      PS s;
      pair_state(R, l, j, s);
      if (l >= 2) {
        const int i1 = SD.tor_i1[p];
        if (i1 != R.noc + 1) {
          const double rt = T(F_R, i1);
          DS y[2];
          tor_ystart_d(T, s, i1, rt, y);
          double vb[SV_N];
          for (int v = 0; v < SV_N; v++)
            vb[v] = cmul_(y0bar[2 * p], y[0].d[v]).x + cmul_(y0bar[2 * p + 1], y[1].d[v]).x;
          put(spl_w(T, i1, rt), vb);
        }
      }
      const int raw = starts[2 * p + 1], itop = R.is;
      int i1 = raw;
      double rt = T(F_R, raw);
      if (raw == 1) rt = fmin(0.1 * (1.0 / (4.0 * s.rew / R.alpha_min)), 0.1 * T(F_R, 2));
      if (i1 >= itop) {
        i1 = itop - 1;
        rt = T(F_R, i1);
      }
      const bool fluid = i1 <= R.noc && (i1 > R.nic || R.fic);
      DS m[14];
      minor_start_d(T, s, i1, rt, fluid, m);
      double vb[SV_N];
      for (int v = 0; v < SV_N; v++) {
        double a = 0.0;
        for (int q = 0; q < 14; q++) a += cmul_(m0bar[p * 14 + q], m[q].d[v]).x;
        vb[v] = a;
      }
      put(spl_w(T, i1, rt), vb);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < SV_N * 2 * nk; q += blockDim.x) {
    const double a = acc[q];
    if (a != 0.0) {
      const int v = q / (2 * nk), which = (q / nk) % 2, knot0 = q % nk;
      atomicAdd(&tbar[(size_t)(start_field(v) + which) * nk + knot0], a);
    }
  }
}

// ------------------------------------------------------- radial adjoint
// rad_solver_wg (src/module_solver.f90:446-561), receiver at the top grid
// point: yrad(q, c) = -y2r_q b_c / D, y2r = (1, 0), b = (y1_2, -y1_1),
// D = y2_2 y1_1 - y2_1 y1_2 with y1 the upward and y2 the downward solution
// at the source point.  The solve is linear and 2 x 2 with no
// renormalisation, so it is cut into its per-interval segments
// (rad_segment), each a 2 x 2 propagator:
//   k_rad_prop     every (segment, frequency): the segment's propagator;
//   k_rad_scan     per frequency: the states at the segment starts, the
//                  rows of the record adjoints at the source point, the
//                  rows at the segment ends, and the start solution's part;
//   k_rad_seg_adj  every (segment, frequency): the segment's steps again
//                  from its start state (states on a tape), the row from
//                  its end back through them with every stage paired; per
//                  wave (64 frequencies, one interval) summed into the
//                  interval's table entries.
// Segment q of a frequency, in sweep order: q < isrc upward in interval
// q + 1, then downward in intervals itop .. isrc.

__device__ inline void seg_of(const RadSegs& G, int q, int& sweep, int& i) {
  if (q < G.isrc) {
    sweep = 0;
    i = q + 1;
  } else {
    sweep = 1;
    i = G.itop - (q - G.isrc);
  }
}

__device__ inline double rad_dr(const RunK& R, const PS& s) {
  return 1.0 / (4.0 * s.rew / R.alpha_min);
}

// propagators, column-major: P[2 c + r] (row r, column c)
__global__ void __launch_bounds__(64) k_rad_prop(const double* __restrict__ gtab,
                                                 const RunK* __restrict__ runs, int e, RadSegs G,
                                                 cd* __restrict__ prop, int* __restrict__ cap) {
  const RunK& R = runs[e];
  const int q = blockIdx.x, j = (int)(blockIdx.y * 64 + threadIdx.x) + 1;
  if (j > R.nw) return;
  const Tab T{gtab, R.nknot};
  PS s;
  pair_state(R, 0, j, s);
  int sweep, i;
  seg_of(G, q, sweep, i);
  cd Y[2][2] = {{C(1.0), C(0.0)}, {C(0.0), C(1.0)}};
  const int n = rad_segment(T, R.nknot, s, rad_dr(R, s), R.rs, R.rr, sweep, i, G.isrc, G.itop,
                            [&](double x, double h) { rad_step_cols<2>(T, s, i, x, h, Y); });
  cd* P = prop + ((size_t)q * R.nw + (j - 1)) * 4;
  P[0] = Y[0][0];
  P[1] = Y[0][1];
  P[2] = Y[1][0];
  P[3] = Y[1][1];
  atomicMax(cap, n);
}

__device__ inline void prop_apply(const cd* P, cd* y) {
  const cd a = P[0] * y[0] + P[2] * y[1], b = P[1] * y[0] + P[3] * y[1];
  y[0] = a;
  y[1] = b;
}
__device__ inline void prop_row(const cd* P, cd* x) {
  const cd a = x[0] * P[0] + x[1] * P[1], b = x[0] * P[2] + x[1] * P[3];
  x[0] = a;
  x[1] = b;
}

__global__ void __launch_bounds__(64) k_rad_scan(const double* __restrict__ gtab,
                                                 const RunK* __restrict__ runs, int e, RadSegs G,
                                                 const cd* __restrict__ prop,
                                                 const cd* __restrict__ recb,
                                                 cd* __restrict__ ys, cd* __restrict__ xe,
                                                 double* __restrict__ tbar) {
  const RunK& R = runs[e];
  const int j = (int)(blockIdx.x * 64 + threadIdx.x) + 1;
  if (j > R.nw || R.lmin != 0) return;
  const Tab T{gtab, R.nknot};
  const int nk = R.nknot;
  const size_t p = (size_t)R.pbase + (j - 1);
  PS s;
  pair_state(R, 0, j, s);
  auto at = [&](int q) { return (size_t)q * R.nw + (j - 1); };
  // the start radius: the first step of segment (upward, interval 1)
  double rt = 0.0;
  bool got = false;
  rad_segment(T, nk, s, rad_dr(R, s), R.rs, R.rr, 0, 1, G.isrc, G.itop, [&](double x, double) {
    if (!got) rt = x;
    got = true;
  });
  cd y[2];
  rad_ystart(T, s, 1, rt, y);
  for (int q = 0; q < G.isrc; q++) {
    ys[2 * at(q)] = y[0];
    ys[2 * at(q) + 1] = y[1];
    prop_apply(prop + 4 * at(q), y);
  }
  const cd y1[2] = {y[0], y[1]};
  y[0] = C(1.0);
  y[1] = C(0.0);
  for (int q = G.isrc; q < G.nseg; q++) {
    ys[2 * at(q)] = y[0];
    ys[2 * at(q) + 1] = y[1];
    prop_apply(prop + 4 * at(q), y);
  }
  const cd y2[2] = {y[0], y[1]}, y2r[2] = {C(1.0), C(0.0)};
  // the rows on y1 and y2 at the source point
  cd a[2][2];
  for (int q = 0; q < 2; q++)
    for (int c = 0; c < 2; c++) a[q][c] = conj_(recb[p * REC + O_YRAD + q + 2 * c]);
  const cd Dv = y2[1] * y1[0] - y2[0] * y1[1], Di = cinv(Dv);
  const cd b[2] = {y1[1], -y1[0]};
  cd Gc[2], E = C(0.0);
  for (int c = 0; c < 2; c++) {
    Gc[c] = C(0.0);
    for (int q = 0; q < 2; q++) {
      Gc[c] = Gc[c] - a[q][c] * y2r[q] * Di;
      E = E + a[q][c] * y2r[q] * b[c] * Di * Di;
    }
  }
  cd x[2] = {-E * y1[1], E * y1[0]};  // on y2
  for (int q = G.nseg - 1; q >= G.isrc; q--) {
    xe[2 * at(q)] = x[0];
    xe[2 * at(q) + 1] = x[1];
    prop_row(prop + 4 * at(q), x);
  }
  x[0] = -Gc[1] + E * y2[1];  // on y1
  x[1] = Gc[0] - E * y2[0];
  for (int q = G.isrc - 1; q >= 0; q--) {
    xe[2 * at(q)] = x[0];
    xe[2 * at(q) + 1] = x[1];
    prop_row(prop + 4 * at(q), x);
  }
  // the starting solution at rt, interval 1
  DS yd[2];
  rad_ystart_d(T, s, 1, rt, yd);
  const SplW sw = spl_w(T, 1, rt);
  for (int v = 0; v < SV_N; v++) {
    const double vb = cmul_(x[0], yd[0].d[v]).x + cmul_(x[1], yd[1].d[v]).x;
    const int f0 = start_field(v);
    spl_adj(sw, f0, vb, [&](int ff, int knot, double val) {
      atomicAdd(&tbar[(size_t)ff * nk + knot - 1], val);
    });
  }
}

__global__ void __launch_bounds__(64) k_rad_seg_adj(const double* __restrict__ gtab,
                                                    const RunK* __restrict__ runs, int e,
                                                    RadSegs G,
                                                    const cd* __restrict__ ys,
                                                    const cd* __restrict__ xe,
                                                    RadSt* __restrict__ tape,
                                                    int* __restrict__ err,
                                                    double* __restrict__ tbar) {
  __shared__ double red[RAD_NV * 4 * 64];
  const RunK& R = runs[e];
  const int q = blockIdx.x, lane = threadIdx.x, j = (int)(blockIdx.y * 64 + lane) + 1;
  const Tab T{gtab, R.nknot};
  int sweep, i;
  seg_of(G, q, sweep, i);
  double acc[RAD_NV * 4];
#pragma unroll
  for (int v = 0; v < RAD_NV * 4; v++) acc[v] = 0.0;
  if (j <= R.nw) {
    PS s;
    pair_state(R, 0, j, s);
    const size_t at = (size_t)q * R.nw + (j - 1);
    cd Y[1][2] = {{ys[2 * at], ys[2 * at + 1]}};
    cd x[2] = {xe[2 * at], xe[2 * at + 1]};
    RadSt* tp = tape + ((size_t)q * G.nwb + blockIdx.y) * (size_t)G.cap * 64 + lane;
    int n = 0;
    rad_segment(T, R.nknot, s, rad_dr(R, s), R.rs, R.rr, sweep, i, G.isrc, G.itop,
                [&](double xr, double h) {
                  if (n < G.cap) {
                    RadSt& t = tp[(size_t)n * 64];
                    t.x = xr;
                    t.h = h;
                    t.y[0] = Y[0][0];
                    t.y[1] = Y[0][1];
                  }
                  n++;
                  rad_step_cols<1>(T, s, i, xr, h, Y);
                });
    if (n > G.cap) {
      atomicOr(err, 2);
    } else {
      for (int m = n - 1; m >= 0; m--) {
        const RadSt t = tp[(size_t)m * 64];
        rad_step_back(T, s, i, t.x, t.h, t.y, x, acc);
      }
    }
  }
#pragma unroll
  for (int v = 0; v < RAD_NV * 4; v++) red[v * 64 + lane] = acc[v];
  __syncthreads();
  if (lane < RAD_NV * 4) {
    double sum = 0.0;
    for (int k = 0; k < 64; k++) sum += red[lane * 64 + k];
    const int f = rad_field(lane / 4) + (lane % 4) / 2;
    if (sum != 0.0) atomicAdd(&tbar[(size_t)f * R.nknot + i - 1 + (lane % 2)], sum);
  }
}

}  // namespace

void launch_brow_adj(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                     const RunK* runs, Queue Q, SphState SS, const cd* rec, const cd* recb,
                     cd* mub, AdjWork W, double* tbar) {
  hipLaunchKernelGGL(k_brow_adj<true>, g, b, lds, s, gtab, runs, Q, SS, rec, recb, mub, W, tbar);
}
void launch_minor_rows(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                       const RunK* runs, Queue Q, StartData SD, SphState SS, const cd* rec,
                       const cd* recb, const cd* mub, cd* m0bar, AdjWork W, StepStore FS,
                       RowStore RS, double* tbar) {
  hipLaunchKernelGGL(k_minor_rows, g, b, lds, s, gtab, runs, Q, SD, SS, rec, recb, mub, m0bar,
                     W, FS, RS, tbar);
}
void launch_minor_pairs(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                        const RunK* runs, AdjWork W, StepStore FS, RowStore RS, double* tbar) {
  hipLaunchKernelGGL(k_minor_pairs, g, b, lds, s, gtab, runs, W, FS, RS, tbar);
}
void launch_minor_adj(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                      const RunK* runs, Queue Q, StartData SD, SphState SS, const cd* rec,
                      const cd* recb, const cd* mub, cd* m0bar, AdjWork W, StepStore FS,
                      double* tbar) {
  hipLaunchKernelGGL(k_minor_adj<true>, g, b, lds, s, gtab, runs, Q, SD, SS, rec, recb, mub,
                     m0bar, W, FS, tbar);
}
void launch_tor_adj(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                    const RunK* runs, Queue Q, StartData SD, const cd* rec, const cd* recb,
                    cd* y0bar, AdjWork W, StepStore FS, double* tbar) {
  hipLaunchKernelGGL(k_tor_adj<true>, g, b, lds, s, gtab, runs, Q, SD, rec, recb, y0bar, W, FS,
                     tbar);
}
void launch_start_adj(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                      const RunK* runs, PairRange PR, StartData SD, const int* starts,
                      const cd* m0bar, const cd* y0bar, double* tbar) {
  hipLaunchKernelGGL(k_start_adj, g, b, lds, s, gtab, runs, PR, SD, starts, m0bar, y0bar, tbar);
}
void launch_part_sum(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* tpart, int nb,
                     size_t n, double* tbar) {
  hipLaunchKernelGGL(k_part_sum, g, b, lds, s, tpart, nb, n, tbar);
}
size_t rad_step_bytes() { return sizeof(RadSt); }
void launch_rad_prop(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                     const RunK* runs, int e, RadSegs G, cd* prop, int* cap) {
  hipLaunchKernelGGL(k_rad_prop, g, b, lds, s, gtab, runs, e, G, prop, cap);
}
void launch_rad_scan(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                     const RunK* runs, int e, RadSegs G, const cd* prop, const cd* recb, cd* ys,
                     cd* xe, double* tbar) {
  hipLaunchKernelGGL(k_rad_scan, g, b, lds, s, gtab, runs, e, G, prop, recb, ys, xe, tbar);
}
void launch_rad_seg_adj(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                        const RunK* runs, int e, RadSegs G, const cd* ys, const cd* xe,
                        void* tape, int* err, double* tbar) {
  hipLaunchKernelGGL(k_rad_seg_adj, g, b, lds, s, gtab, runs, e, G, ys, xe,
                     static_cast<RadSt*>(tape), err, tbar);
}

}  // namespace ysx
