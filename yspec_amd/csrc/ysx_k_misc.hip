// ysx_k_misc.hip -- the kernels of libysx.so other than the lockstep
// marches (source vectors, start levels, radial and variant solvers,
// coefficient assembly, receiver sum, Legendre tables, model transform,
// inverse FFT and post-processing), and their launch functions.
#include <cmath>

#include "ysx_kern.hpp"
#include "ysx_radial.hpp"

namespace ysx {
namespace {

__global__ void k_reset(int* counters, int n) {
  if ((int)threadIdx.x < n) counters[threadIdx.x] = 0;
}

// zero n doubles (a kernel rather than a memset node, so that the call can
// sit in a captured graph)
__global__ void k_fill0(double* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// residual seisb = seis - obs on the first nout samples of every trace and 0
// after them (seismograms column-major, time fastest: nt per trace; obs nout
// per trace), with per-block sums of the squared residual in part
__global__ void __launch_bounds__(256) k_residual(const double* __restrict__ seis,
                                                  const double* __restrict__ obs, int nt,
                                                  int nout, size_t ntrace,
                                                  double* __restrict__ seisb,
                                                  double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  const size_t n = (size_t)nt * ntrace;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const size_t tr = i / (size_t)nt;
    const int t = (int)(i - tr * (size_t)nt);
    double r = 0.0;
    if (t < nout) {
      r = seis[i] - obs[tr * (size_t)nout + t];
      acc += r * r;
    }
    seisb[i] = r;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// -------------------------------------------------------------- k_source
// source_vector_tor / _sph / _rad, src/module_solver.f90:145-215,
// 566-627, 1747-1847.  source_vector_sph keeps only the real part of its
// dispersion factors (they are declared real(dp) there, :1771-1772).
// The source vectors of one pair for moment tensor mm (normalised units),
// linear in mm: put(f, v) receives the value v of record field f (svt,
// svs, sr0 entries; the others stay zero).
template <class Put>
__device__ __forceinline__ void source_vectors(const Tab& T, const RunK& R, const PS& s, int l,
                                               const double* mm, Put put) {
  SplW w = spl_w(T, R.is, R.rs);
  double rlcon = spl(T, w, F_LCON), rccon = spl(T, w, F_CCON), rfcon = spl(T, w, F_FCON),
         rqmu = spl(T, w, F_QMU), rxlam = spl(T, w, F_XLAM), rxa2 = spl(T, w, F_XA2);
  double rs = R.rs;
  cd den = cinv(cmuli(s.omega) * (rs * rs));  // 1/(i w rs^2)
  cd nd = s.nu * den;
  if (l == 0) {
    cd qaa = 1.0 + rxa2 * s.lno, qff = 1.0 + rxlam * s.lno;
    cd cc = cinv(rccon * qaa);
    cd cf = rfcon * qff;
    put(O_SR0 + 0, nd * ((rs * mm[0]) * cc));
    put(O_SR0 + 1, nd * ((2.0 * mm[0]) * (cf * cc) - (mm[3] + mm[5])));
    return;
  }
  if (l > 1) {
    cd cl = rlcon * (1.0 + rqmu * s.lno);
    cd icl = cinv(cl);
    put(O_SVT + 1, nd * ((0.5 * rs) * C(-mm[2], mm[1]) * icl));
    put(O_SVT + 3, nd * ((0.5 * rs) * C(mm[2], mm[1]) * icl));
    put(O_SVT + 5 + 0, nd * ((0.25 * s.szeta2m2) * C(2.0 * mm[4], mm[5] - mm[3])));
    put(O_SVT + 5 + 4, nd * ((0.25 * s.szeta2m2) * C(2.0 * mm[4], mm[3] - mm[5])));
  }
  double qaa = 1.0 + rxa2 * s.lno.x, qff = 1.0 + rxlam * s.lno.x, qll = 1.0 + rqmu * s.lno.x;
  double cc = 1.0 / (rccon * qaa), cf = rfcon * qff, cl = rlcon * qll;
  double zeta = s.zeta;
  put(O_SVS + 1 + 5 * 1, nd * (C(0.5 * rs * mm[1] / cl, 0.5 * rs * mm[2] / cl)));
  put(O_SVS + 2 + 5 * 0, nd * C(rs * mm[0] * cc));
  put(O_SVS + 2 + 5 * 3, nd * C(-mm[3] - mm[5] + 2.0 * mm[0] * cf * cc));
  put(O_SVS + 2 + 5 * 4, nd * C(0.5 * zeta * (mm[3] + mm[5]) - zeta * mm[0] * cf * cc));
  put(O_SVS + 3 + 5 * 1, nd * (C(-0.5 * rs * mm[1] / cl, 0.5 * rs * mm[2] / cl)));
  if (l > 1) {
    put(O_SVS + 0 + 5 * 4, nd * ((0.25 * s.szeta2m2) * C(mm[5] - mm[3], -2.0 * mm[4])));
    put(O_SVS + 4 + 5 * 4, nd * ((0.25 * s.szeta2m2) * C(mm[5] - mm[3], 2.0 * mm[4])));
  }
}

__global__ void __launch_bounds__(256) k_source(const double* __restrict__ gtab,
                                                const RunK* __restrict__ runs, PairRange PR,
                                                cd* __restrict__ rec, int* __restrict__ cnt) {
  Tab T = load_tab(gtab, runs[PR.e0].nknot, runs[PR.e0].tablds);
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
  source_vectors(T, R, s, l, R.mm, [&](int f, cd v) { o[f] = v; });
}

// ----------------------------------------------------------------- k_start
// Start levels and starting solutions of every pair (one lane per pair):
// tor_solver's start (src/module_solver.f90:64-75, 103-114) and
// sph_solver's start (:678-722, 758-785); with a fluid inner core
// (receiver not below the source) sph_solver_fic's (:1240-1265), which
// starts the fluid minors in the inner core.  The marches pick these up
// when they reach each lane's start knot.  parts: bit 0 the toroidal
// starts, bit 1 the spheroidal ones (separate launches let the toroidal
// chain run on another stream beside the spheroidal one).
__global__ void __launch_bounds__(256) k_start(const double* __restrict__ gtab,
                                               const RunK* __restrict__ runs, PairRange PR,
                                               StartData SD, int* __restrict__ starts,
                                               int parts) {
  // the start-level walks read the table knot after knot: from LDS when it
  // fits there (all events share the deck)
  const Tab T = load_tab(gtab, runs[PR.e0].nknot, runs[PR.e0].tablds);
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= PR.n) return;
  long long g = PR.g0 + t;
  const RunK& R = runs[find_event(runs, PR.e0, PR.e1, g)];
  int l, j;
  pair_lj(R, (int)(g - R.pbase), l, j);
  size_t p = (size_t)g;
  if (l < 1) return;
  PS s;
  pair_state(R, l, j, s);
  Layers L = layers(R);
  bool up = R.rr >= R.rs;
  int itop = up ? R.is : R.ir;
  if ((parts & 1) && l >= 2) {
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
  // the gravity-0/1 march finds its own start
  if (!(parts & 2) || R.ngc) return;
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
  if (i1 <= L.noc && (i1 > L.nic || R.fic))
    minor_fluid_start(T, s, i1, rt, m);
  else
    minor_solid_start(T, s, i1, rt, m);
  SD.sph_i1[p] = i1;
  for (int i = 0; i < 14; i++) SD.sph_m0[(size_t)p * 14 + i] = m[i];
}

// ------------------------------------------------------------------- k_rad
// rad_solver_wg with ss = identity(2), src/module_solver.f90:446-561, on
// the grid of make_grid, src/module_model.f90:638-787 (ysx_radial.hpp).
// Only yout is formed; the caller discards pout (src/yspec.f90:277).
// A group of R.radg lanes per frequency (8, or 1 for runs with many
// frequencies, where the radial kernel is hidden beside the spheroidal
// marches and should take few compute units) integrates one solution: the
// control flow depends on w only, so the group follows it together and
// shares the stage evaluations of every step (rk_step_grp); lane 0 writes.
constexpr int RAD_GROUP_MAX = 8;

__device__ __forceinline__ void rad_item(const Tab& T, const RunK& R, int jb,
                                         cd* __restrict__ rec, int* __restrict__ cnt,
                                         RadRHS::Coef (*rsh)[6]) {
  const int G = R.radg;
  int j = jb * (64 / G) + (int)threadIdx.x / G + 1;
  if (j > R.nw || R.lmin != 0) return;
  const RadGroup grp{(int)threadIdx.x % G, G, rsh[(threadIdx.x / G) % RAD_GROUP_MAX]};
  size_t p = (size_t)R.pbase + (j - 1);  // l = 0 pairs come first
  PS s;
  pair_state(R, 0, j, s);
  if (R.grav == 0) return;  // rad_solver_ng: k_rad_ng
  double dr = 1.0 / (4.0 * s.rew / R.alpha_min);
  int nsr = 0;
  cd y1[2], y2[2], y1r[2], y2r[2];
  bool rec_below;
  rad_sweeps(T, R.nknot, s, dr, R.rs, R.rr, grp, NoRadHook(), y1, y2, y1r, y2r, rec_below, nsr);
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
  if (grp.q != 0) return;
  cd* o = rec + p * REC + O_YRAD;
  for (int i = 0; i < 4; i++) o[i] = yo[i];
  cnt[5 * p + 4] = nsr;
}

// items (event, block of 64 / radg frequencies); a lane group per l = 0
// pair, no lockstep between groups (the grid of make_grid depends on w)
__global__ void __launch_bounds__(64) k_rad(const double* __restrict__ gtab,
                                            const RunK* __restrict__ runs, Queue Q,
                                            cd* __restrict__ rec, int* __restrict__ cnt) {
  __shared__ RadRHS::Coef rsh[64 / RAD_GROUP_MAX][6];
  Tab T = load_tab(gtab, runs[0].nknot, runs[0].tablds);
  RUN_ITEMS(false, Q, rad_item(T, runs[it.x], it.y, rec, cnt, rsh));
}

// rad_solver_ng (gravity switch 0, src/yspec.f90:271-273) for the l = 0
// pairs of the events with grav = 0; same items as k_rad
__device__ __forceinline__ void rad_ng_item(const Tab& T, const RunK& R, int jb,
                                            cd* __restrict__ rec, int* __restrict__ cnt,
                                            RadRHS::Coef (*rsh)[6]) {
  const int ng = R.radg;
  int j = jb * (64 / ng) + (int)threadIdx.x / ng + 1;
  if (j > R.nw || R.lmin != 0 || R.grav != 0) return;
  const RadGroup grp{(int)threadIdx.x % ng, ng, rsh[(threadIdx.x / ng) % RAD_GROUP_MAX]};
  size_t p = (size_t)R.pbase + (j - 1);
  PS s;
  pair_state(R, 0, j, s);
  VarGeo G{R.nknot, R.nic, R.noc, R.nsl, R.ocean, R.is, R.ir, R.rs, R.rr, R.alpha_min,
           R.fre_norm};
  cd yo[4];
  int n = 0;
  rad_solve_ng(T, s, G, yo, n, grp);
  if (grp.q != 0) return;
  cd* o = rec + p * REC + O_YRAD;
  for (int i = 0; i < 4; i++) o[i] = yo[i];
  cnt[5 * p + 4] = n;
}

__global__ void __launch_bounds__(64) k_rad_ng(const double* __restrict__ gtab,
                                               const RunK* __restrict__ runs, Queue Q,
                                               cd* __restrict__ rec, int* __restrict__ cnt) {
  __shared__ RadRHS::Coef rsh[64 / RAD_GROUP_MAX][6];
  Tab T = load_tab(gtab, runs[0].nknot, runs[0].tablds);
  RUN_ITEMS(false, Q, rad_ng_item(T, runs[it.x], it.y, rec, cnt, rsh));
}

// ------------------------------------------------------------- k_assemble
// output-type scaling and the uu/vv/ww coefficients, src/yspec.f90:280-288,
// 309-316, 377-400 (zeta there is single precision, :255).  A block takes
// ASM_P consecutive pairs: the input fields [0, O_UU) of their records are
// copied to LDS with coalesced loads, one lane per pair computes, and the
// 15 coefficients go back to the records through LDS the same way (reading
// the 1.4 KB records field by field from one lane per pair fetched each
// cache line several times).
constexpr int ASM_LD = O_UU + 1;  // LDS stride per pair, odd: conflict-free 16-byte reads
static_assert(O_UU == O_YRAD + 4 && O_VV == O_UU + 5 && O_WW == O_VV + 5 && REC == O_WW + 5,
              "record layout: inputs first, then uu, vv, ww");

__global__ void __launch_bounds__(64) k_assemble(const RunK* __restrict__ runs, PairRange PR,
                                                 cd* __restrict__ rec, cd* __restrict__ cf,
                                                 const double* __restrict__ tb) {
  __shared__ cd s[ASM_P * ASM_LD];
  long long b0 = (long long)blockIdx.x * ASM_P;
  if (b0 >= PR.n) return;
  int np = (int)min((long long)ASM_P, PR.n - b0);
  const cd* src = rec + (size_t)(PR.g0 + b0) * REC;
  {
    // All of a lane's loads are issued before the first LDS write.  Indices
    // past the block's last field are clamped to it (the lanes concerned
    // load and store that field's value again), so no load is predicated.
    constexpr int NL = (ASM_P * O_UU + 63) / 64;
    const int kmax = np * O_UU - 1;
    cd v[NL];
#pragma unroll
    for (int i = 0; i < NL; i++) {
      int k = min((int)threadIdx.x + 64 * i, kmax), q = k / O_UU, f = k - q * O_UU;
      v[i] = src[(size_t)q * REC + f];
    }
#pragma unroll
    for (int i = 0; i < NL; i++) {
      int k = min((int)threadIdx.x + 64 * i, kmax), q = k / O_UU, f = k - q * O_UU;
      s[q * ASM_LD + f] = v[i];
    }
  }
  __syncthreads();
  int t = threadIdx.x;
  if (t < np) {
    long long g = PR.g0 + b0 + t;
    const RunK& R = runs[find_event(runs, PR.e0, PR.e1, g)];
    GemmK G = gemm_of(R, cf, tb);
    int l, j;
    pair_lj(R, (int)(g - R.pbase), l, j);
    cd* o = s + t * ASM_LD;
    cd w = C(R.wmin + (double)(j - 1) * R.dw, -R.ep);
    cd fo = (R.out == 1) ? cmuli(w) : (R.out == 2 ? -(w * w) : C(1.0));
    double rr = R.rr;
    cd c[15];
    for (int m = 0; m < 15; m++) c[m] = C(0.0);
    if (l == 0) {
      cd u3 = (o[O_SR0] * o[O_YRAD + 0] + o[O_SR0 + 1] * o[O_YRAD + 2]) / rr;
      c[2] = fo * u3;
    } else {
      double zeta = (double)sqrtf((float)(l * (l + 1)));
      if (l > 1) {
        cd t1 = fo * o[O_YTOR + 0], t2 = fo * o[O_YTOR + 2];
        for (int m = 0; m < 5; m++)
          c[10 + m] = (o[O_SVT + m] * t1 + o[O_SVT + 5 + m] * t2) / (zeta * rr);
      }
      cd y1[4], y2[4];
      for (int k = 0; k < 4; k++) {
        cd a = o[O_YSPH + 6 * k + 0], bb = o[O_YSPH + 6 * k + 1], g3 = o[O_YSPH + 6 * k + 2];
        if (R.out == 2 && R.cor != 0) {
          cd c1 = -((double)(l + 1) * g3 + (2.0 * R.grav_ir) * a) / rr;
          cd c2 = zeta * (g3 + R.grav_ir * a) / rr;
          y1[k] = fo * a + c1;
          y2[k] = fo * bb + c2;
        } else {
          y1[k] = fo * a;
          y2[k] = fo * bb;
        }
      }
      const cd* sv = o + O_SVS;
      for (int m = 0; m < 5; m++) {
        cd su = y1[0] * sv[m] + y1[1] * sv[5 + m] + y1[2] * sv[15 + m] + y1[3] * sv[20 + m];
        cd sw = y2[0] * sv[m] + y2[1] * sv[5 + m] + y2[2] * sv[15 + m] + y2[3] * sv[20 + m];
        c[m] = su / rr;
        c[5 + m] = sw / (zeta * rr);
      }
    }
    size_t plane = (size_t)G.lpad * G.jpad, at = (size_t)l * G.jpad + (j - 1);
    for (int m = 0; m < 15; m++) {
      G.cf[m * plane + at] = c[m];
      o[m] = c[m];  // this lane's own slot; its inputs have all been read
    }
  }
  __syncthreads();
  cd* dst = rec + (size_t)(PR.g0 + b0) * REC + O_UU;
#pragma unroll
  for (int i = 0; i < (ASM_P * 15 + 63) / 64; i++) {
    int k = threadIdx.x + 64 * i, q = k / 15, f = k - q * 15;
    if (k < np * 15) dst[(size_t)q * REC + f] = s[q * ASM_LD + f];
  }
}

// ---------------------------------------------------------------- k_lsum
// The receiver sum (src/yspec.f90:290-293, 318-335, 402-419) as FP64 MFMA
// GEMMs: for each Legendre table X in {x, dx/dtheta, x/sin(theta)} and
// |m| = a, C[j][k] = sum_l coef(l, w_j) X_a(l, theta_k) over the coefficient
// sets that table multiplies; the epilogue applies (-1)^a for m < 0, the
// factors i*m and exp(i m phi_k) and accumulates ur, ut, up.  A block of
// LSUM_WAVES waves owns a 16 (w) x 16 (receiver) tile; wave q sums the
// degrees l0 = 4 (q + LSUM_WAVES i) .. l0 + 3 and applies the (linear)
// epilogue to its partial sums, which one LDS reduction then adds.
// v_mfma_f64_16x16x4f64 lane layout: A[i][k] at lane i + 16k, B[k][n] at
// lane n + 16k, D[i][n] at lane n + 16(i%4), reg i/4.
typedef double d4 __attribute__((ext_vector_type(4)));

// one (table, |m|) group: GEMM over l, then the epilogue for this group
constexpr int LSUM_WAVES = 4;

// An empty asm that claims to modify the running sums: the epilogue of one
// group then stays ahead of the next group's loop instead of being sunk to
// the end, which would keep every group's accumulators alive at once.
__device__ __forceinline__ void pin_sums(cd* o) {
#pragma unroll
  for (int v = 0; v < 12; v++) asm volatile("" : "+v"(o[v].x), "+v"(o[v].y));
}

template <int X, int A, int NC, int C0, int C1, int C2, int C3>
__device__ __forceinline__ void lsum_group(const GemmK& G, int j0, int k0, int li, int lq,
                                           int wave, cd ep, cd en, cd* ur, cd* ut, cd* up) {
  constexpr int cs[4] = {C0, C1, C2, C3};
  const size_t plane = (size_t)G.lpad * G.jpad;
  d4 acc[NC][2];
#pragma unroll
  for (int q = 0; q < NC; q++) acc[q][0] = acc[q][1] = (d4){0.0, 0.0, 0.0, 0.0};
  const double* tb = G.tb + (size_t)(X * 3 + A) * G.lpad * G.nrpad + k0 + li;
  const cd* cf = G.cf + (size_t)lq * G.jpad + j0 + li;
  // this wave's degrees l0 = 4 wave + 16 i; the loads of the next l0 are
  // issued before the MFMAs of the current one (the last one is loaded
  // twice rather than branched around)
  const int step = 4 * LSUM_WAVES;
  int l0 = 4 * wave;
  if (l0 >= G.lpad) return;
  const int last = l0 + (G.lpad - 1 - l0) / step * step;
  double b = tb[(size_t)(l0 + lq) * G.nrpad];
  cd av[NC];
#pragma unroll
  for (int q = 0; q < NC; q++) av[q] = cf[cs[q] * plane + (size_t)l0 * G.jpad];
  for (; l0 <= last; l0 += step) {
    const int ln = min(l0 + step, last);
    double bn = tb[(size_t)(ln + lq) * G.nrpad];
    cd an[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) an[q] = cf[cs[q] * plane + (size_t)ln * G.jpad];
#pragma unroll
    for (int q = 0; q < NC; q++) {
      acc[q][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q].x, b, acc[q][0], 0, 0, 0);
      acc[q][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q].y, b, acc[q][1], 0, 0, 0);
    }
    b = bn;
#pragma unroll
    for (int q = 0; q < NC; q++) av[q] = an[q];
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
                                          const cd* __restrict__ eia, cd* __restrict__ spec,
                                          cd (*red)[12][64]) {
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int njt = G.jpad / 16;
  int j0 = (tile % njt) * 16, k0 = (tile / njt) * 16;
  int li = lane & 15, lq = lane >> 4;
  int k = k0 + li;  // receiver of this lane's output column
  cd e1 = (k < R.nr) ? eia[k] : C(1.0), e2 = e1 * e1;
  cd em1 = C(e1.x, -e1.y), em2 = C(e2.x, -e2.y);
  cd o[12];  // ur[4], ut[4], up[4]
  cd *ur = o, *ut = o + 4, *up = o + 8;
  for (int v = 0; v < 12; v++) o[v] = C(0.0);
  // coefficient index c: uu m -> m+2, vv m -> 7+m, ww m -> 12+m
  lsum_group<0, 0, 1, 2, 0, 0, 0>(G, j0, k0, li, lq, wave, e1, em1, ur, ut, up);    // uu0 x
  pin_sums(o);
  lsum_group<0, 1, 4, 3, 1, 13, 11>(G, j0, k0, li, lq, wave, e1, em1, ur, ut, up);  // uu+-1, ww+-1 x
  pin_sums(o);
  lsum_group<0, 2, 4, 4, 0, 14, 10>(G, j0, k0, li, lq, wave, e2, em2, ur, ut, up);  // uu+-2, ww+-2 x
  pin_sums(o);
  lsum_group<1, 0, 1, 7, 0, 0, 0>(G, j0, k0, li, lq, wave, e1, em1, ur, ut, up);    // vv0 dx
  pin_sums(o);
  lsum_group<1, 1, 4, 8, 6, 13, 11>(G, j0, k0, li, lq, wave, e1, em1, ur, ut, up);  // vv+-1, ww+-1 dx
  pin_sums(o);
  lsum_group<1, 2, 4, 9, 5, 14, 10>(G, j0, k0, li, lq, wave, e2, em2, ur, ut, up);  // vv+-2, ww+-2 dx
  pin_sums(o);
  lsum_group<2, 1, 2, 8, 6, 0, 0>(G, j0, k0, li, lq, wave, e1, em1, ur, ut, up);    // vv+-1 x/sin
  pin_sums(o);
  lsum_group<2, 2, 2, 9, 5, 0, 0>(G, j0, k0, li, lq, wave, e2, em2, ur, ut, up);    // vv+-2 x/sin
  pin_sums(o);
  // tree reduction over the waves: 2,3 -> 0,1, then 1 -> 0
  static_assert(LSUM_WAVES == 4, "reduction written for four waves");
  if (wave >= 2)
    for (int v = 0; v < 12; v++) red[wave - 2][v][lane] = o[v];
  __syncthreads();
  if (wave < 2)
    for (int v = 0; v < 12; v++) o[v] = o[v] + red[wave][v][lane];
  __syncthreads();
  if (wave == 1)
    for (int v = 0; v < 12; v++) red[0][v][lane] = o[v];
  __syncthreads();
  if (wave == 0 && k < R.nr) {
    size_t nwr = (size_t)R.nw * R.nr;
    for (int v = 0; v < 4; v++) {
      int j = j0 + lq + 4 * v;
      if (j >= R.nw) continue;
      size_t at = (size_t)j + (size_t)R.nw * k;
      spec[at] = ur[v] + red[0][v][lane];
      spec[nwr + at] = ut[v] + red[0][4 + v][lane];
      spec[2 * nwr + at] = up[v] + red[0][8 + v][lane];
    }
  }
  __syncthreads();  // red is reused by the block's next item
}

// items (event, 16 w x 16 receiver tile); spec of event e at spec_off,
// [3][nr][nw]; blocks of 64 LSUM_WAVES threads
__global__ void __launch_bounds__(256) k_lsum(const RunK* __restrict__ runs, Queue Q,
                                              cd* __restrict__ cf, const double* __restrict__ tb,
                                              const cd* __restrict__ eia, cd* __restrict__ spec) {
  __shared__ cd red[2][12][64];
  RUN_ITEMS(false, Q,
            lsum_item(runs[it.x], gemm_of(runs[it.x], cf, tb), it.y, eia + runs[it.x].roff,
                      spec + runs[it.x].spec_off, red));
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
// and layer_n per layer; work holds ws >= 3 (largest layer_n + 1) doubles
// per (layer, field) solve, 13 nlayer solves; gi holds nknot doubles.
__global__ void __launch_bounds__(256) k_model(const double* __restrict__ deck, int nknot,
                                               int ifanis, double vel_norm,
                                               const int* __restrict__ layer_lo,
                                               const int* __restrict__ layer_n, int nlayer,
                                               double* __restrict__ tab, double* __restrict__ work,
                                               size_t ws, double* __restrict__ intg,
                                               double* __restrict__ qint,
                                               double* __restrict__ gi) {
#pragma clang fp contract(off)
  int tid = threadIdx.x, nt = blockDim.x;
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

// ------------------------------------------------------------- k_model_vjp
// Adjoint of k_model (one block): the table adjoint tbar (F_COUNT x nknot)
// to the adjoints of the deck's 9 columns per knot, dbar (9 x nknot), by
// the transposes of k_model's steps in reverse order (the sequence of the
// host's model_table_vjp, through the same functions).  bar holds F_COUNT
// x nknot doubles, aux 6 nknot (intg, qint, intgb, qintb, gbar, gib), work
// ws >= 3 (largest layer_n + 2) doubles per (layer, field) solve, 12 nlayer
// solves.  Per-layer solves run one thread each (layers hold disjoint knot
// ranges); the running sums and the quadrature transposes, which add into
// neighbouring knots, run on one thread in the host's order.
__global__ void __launch_bounds__(256) k_model_vjp(const double* __restrict__ deck, int nknot,
                                                   int ifanis, double vel_norm,
                                                   const int* __restrict__ layer_lo,
                                                   const int* __restrict__ layer_n, int nlayer,
                                                   const double* __restrict__ tab,
                                                   const double* __restrict__ tbar,
                                                   double* __restrict__ bar,
                                                   double* __restrict__ work, size_t ws,
                                                   double* __restrict__ aux,
                                                   double* __restrict__ dbar) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x, nt = blockDim.x;
  const size_t nk = nknot;
  const double* r = tab + (size_t)F_R * nk;
  const double* rho = tab + (size_t)F_RHO * nk;
  double *intg = aux, *qint = aux + nk, *intgb = aux + 2 * nk, *qintb = aux + 3 * nk;
  double *gbar = aux + 4 * nk, *gib = aux + 5 * nk;
  for (size_t e = tid; e < (size_t)F_COUNT * nk; e += nt) bar[e] = tbar[e];
  for (int i = tid; i < nknot; i += nt) {
    intgb[i] = qintb[i] = gib[i] = 0.0;
    intg[i] = r[i] * r[i] * rho[i];
  }
  __syncthreads();
  double* bgrav = bar + (size_t)F_GRAV * nk;
  for (int k = tid; k < nlayer; k += nt) {
    const int lo = layer_lo[k], n = layer_n[k];
    natural_spline_T(r + lo, bar + (size_t)F_GRAV2 * nk + lo, bgrav + lo, n, work + k * ws);
    natural_spline(r + lo, intg + lo, qint + lo, n, work + (nlayer + k) * ws);
  }
  __syncthreads();
  for (int i = tid; i < nknot; i += nt)
    gbar[i] = (r[i] > 0.0) ? bgrav[i] * 4.0 / (r[i] * r[i]) : bgrav[i];
  __syncthreads();
  if (tid == 0) {
    double acc = 0.0;
    for (int i = nknot - 1; i >= 1; i--) {
      acc += gbar[i];
      gib[i - 1] = acc;
    }
    for (int i = 0; i < nknot - 1; i++) grav_interval_T(r, i, gib[i], intgb, qintb);
  }
  __syncthreads();
  for (int k = tid; k < nlayer; k += nt)
    natural_spline_T(r + layer_lo[k], qintb + layer_lo[k], intgb + layer_lo[k], layer_n[k],
                     work + k * ws);
  __syncthreads();
  for (int i = tid; i < nknot; i += nt) bar[(size_t)F_RHO * nk + i] += r[i] * r[i] * intgb[i];
  __syncthreads();
  for (int t = tid; t < 12 * nlayer; t += nt) {
    const int k = t / 12, f = spline_field(t % 12), lo = layer_lo[k];
    natural_spline_T(r + lo, bar + (size_t)(f + 1) * nk + lo, bar + (size_t)f * nk + lo,
                     layer_n[k], work + t * ws);
  }
  __syncthreads();
  for (int i = tid; i < nknot; i += nt)
    knot_fields_T(deck, nknot, i, ifanis, vel_norm, tab, bar, dbar + 9 * (size_t)i);
}

// ------------------------------------------------------------------ k_seis
// ifft (src/module_fourier.f90:55-145) and src/yspec.f90:439-473 for one
// receiver per block: pack bins i1..i2, cosine filters, Hermitian fill,
// radix-2 DIT inverse FFT with the reference's twiddle angles
// pi_d*m/mmax (single-rounded pi), /(dt*nt), exp(eps t), units, rotation.
// cosine filter weight of frequency bin b (0-based)
__device__ inline double bin_filter(const RunK& K, int b) {
  double w = b * K.df, f = 1.0;
  bool nof = (K.f11 == 0.0 && K.f12 == 0.0 && K.f21 == 0.0 && K.f22 == 0.0);
  if (!nof) {
    if (w < K.f11) f = 0.0;
    else if (w < K.f12) f = 0.5 * (1.0 - cos(PI_D * (w - K.f11) / (K.f12 - K.f11)));
    else if (w < K.f21) f = 1.0;
    else if (w < K.f22) f = 0.5 * (1.0 - cos(PI_D * (K.f22 - w) / (K.f22 - K.f21)));
    else f = 0.0;
  }
  return f;
}

__device__ inline cd packed_bin(const RunK& K, const cd* spec, int k, int i) {
  // 0-based FFT bin i of the packed, filtered, Hermitian-filled row
  bool neg = i > K.nt / 2;
  int b = neg ? K.nt - i : i;
  cd v = C(0.0);
  if (b >= K.i1 - 1 && b <= K.i2 - 1) v = spec[(size_t)(b - (K.i1 - 1)) + (size_t)K.nw * k];
  v = v * bin_filter(K, b);
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

// ================================================================ adjoints
// Vector-Jacobian products of the linear back end of the forward, for a
// real function of the seismograms: a complex quantity z gets the adjoint
// dF/dRe z + i dF/dIm z, so a complex-linear map y = a z has z_bar =
// conj(a) y_bar and y = Re z has z_bar = y_bar.  Each kernel is the exact
// transpose of the discrete map its forward kernel applies (same twiddles,
// filter weights and factors), not of its continuous counterpart.

__device__ inline cd conj_(cd a) { return C(a.x, -a.y); }

// ---------------------------------------------------------------- k_seis_T
// Adjoint of k_seis / k_fft_lds + k_fft_post for one event, one receiver k
// per block: rotation, units and exp(eps t) transposed; the radix-2
// butterfly network run backwards with each butterfly (a, b) -> (a + w b,
// a - w b) replaced by its conjugate transpose (x, y) -> (x + y,
// conj(w) (x - y)); then the bit-reversal, the Hermitian fill (bin b < nt/2
// also collects the conjugate of bin nt - b) and the filter weights give
// spec_bar of bins i1..i2.  Work rows as in k_seis.
__global__ void __launch_bounds__(1024) k_seis_T(const RunK* __restrict__ runs, int e,
                                                 const double* __restrict__ seisb_all,
                                                 const cd* __restrict__ tw,
                                                 cd* __restrict__ work_all,
                                                 const double* __restrict__ azep_all,
                                                 cd* __restrict__ specb_all) {
  const RunK& K = runs[e];
  const int nt = K.nt;
  const double* seisb = seisb_all + K.seis_off;
  cd* work = work_all + K.seis_off;
  cd* specb = specb_all + K.spec_off;
  for (int k = blockIdx.x; k < K.nr; k += gridDim.x) {
    double az = azep_all[K.roff + k];
    double ca = cos(az), sa = sin(az);
    cd* r0 = work + ((size_t)k * 3 + 0) * nt;
    cd* r1 = work + ((size_t)k * 3 + 1) * nt;
    cd* r2 = work + ((size_t)k * 3 + 2) * nt;
    double sc = K.norm / (K.dt * nt);
    for (int i = threadIdx.x; i < nt; i += blockDim.x) {
      double g = exp(K.ep * (i * K.dt)) * sc;
      double z = seisb[(size_t)i + (size_t)nt * (0 + 3 * (size_t)k)];
      double n = seisb[(size_t)i + (size_t)nt * (1 + 3 * (size_t)k)];
      double ea = seisb[(size_t)i + (size_t)nt * (2 + 3 * (size_t)k)];
      r0[i] = C(z * g);
      r1[i] = C((-ca * n + sa * ea) * g);
      r2[i] = C((sa * n + ca * ea) * g);
    }
    __syncthreads();
    for (int c = 0; c < 3; c++) {
      cd* row = work + ((size_t)k * 3 + c) * nt;
      for (int mmax = nt >> 1; mmax >= 1; mmax >>= 1) {
        for (int b = threadIdx.x; b < nt / 2; b += blockDim.x) {
          int m = b & (mmax - 1);
          int i = ((b - m) << 1) + m;
          int jx = i + mmax;
          cd x = row[i], y = row[jx];
          row[i] = x + y;
          row[jx] = conj_(tw[mmax + m]) * (x - y);
        }
        __syncthreads();
      }
      cd* sb = specb + (size_t)c * K.nw * K.nr + (size_t)K.nw * k;
      for (int jj = threadIdx.x; jj < K.nw; jj += blockDim.x) {
        int b = jj + K.i1 - 1;
        cd v = C(0.0);
        if (b <= nt / 2) {
          v = row[brev((unsigned)b, K.lognt)];
          if (b > 0 && b < nt / 2) v = v + conj_(row[brev((unsigned)(nt - b), K.lognt)]);
          v = v * bin_filter(K, b);
        }
        sb[jj] = v;
      }
      __syncthreads();
    }
  }
}

// The same adjoint in three passes for nt <= FFT_B * FFT_G, the transpose
// of k_fft_lds + k_fft_post: k_post_T transposes the post-processing and the
// register stages (mmax = nt/2 .. FFT_B), k_fft_lds_T the LDS stages
// (mmax = FFT_B/2 .. 1) of each contiguous block, and k_pack_T gathers the
// bit-reversed, Hermitian-folded, filtered bins.  Same butterflies and
// twiddles as k_seis_T, so the results are the same bit for bit.

template <int G>
__device__ __forceinline__ void post_T_g(const RunK& K, int k, int i0, const cd* __restrict__ tw,
                                         const double* __restrict__ seisb, double ca, double sa,
                                         cd* __restrict__ work) {
  const int nt = K.nt, bk = nt / G;
  const double sc = K.norm / (K.dt * nt);
  cd v[3][G];
#pragma unroll
  for (int t = 0; t < G; t++) {
    int i = i0 + t * bk;
    double g = exp(K.ep * (i * K.dt)) * sc;
    double z = seisb[(size_t)i + (size_t)nt * (0 + 3 * (size_t)k)];
    double n = seisb[(size_t)i + (size_t)nt * (1 + 3 * (size_t)k)];
    double ea = seisb[(size_t)i + (size_t)nt * (2 + 3 * (size_t)k)];
    v[0][t] = C(z * g);
    v[1][t] = C((-ca * n + sa * ea) * g);
    v[2][t] = C((sa * n + ca * ea) * g);
  }
#pragma unroll
  for (int step = G >> 1; step >= 1; step >>= 1) {
    int mmax = step * bk;
#pragma unroll
    for (int t = 0; t < G; t++) {
      if (t & step) continue;
      int tj = t + step;
      cd w = conj_(tw[mmax + ((i0 + t * bk) & (mmax - 1))]);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        cd x = v[c][t], y = v[c][tj];
        v[c][t] = x + y;
        v[c][tj] = w * (x - y);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    cd* row = work + ((size_t)k * 3 + c) * nt;
#pragma unroll
    for (int t = 0; t < G; t++) row[i0 + t * bk] = v[c][t];
  }
}

// blocks (receiver k, block of 256 column indices i0 < min(nt, FFT_B))
__global__ void __launch_bounds__(256) k_post_T(const RunK* __restrict__ runs, int e,
                                                const double* __restrict__ seisb_all,
                                                const cd* __restrict__ tw,
                                                const double* __restrict__ azep_all,
                                                cd* __restrict__ work_all) {
  const RunK& K = runs[e];
  const int bk = min(K.nt, FFT_B), g = K.nt / bk, nib = (bk + 255) / 256;
  const int k = blockIdx.x / nib, i0 = (blockIdx.x % nib) * 256 + (int)threadIdx.x;
  if (k >= K.nr || i0 >= bk) return;
  double az = azep_all[K.roff + k];
  double ca = cos(az), sa = sin(az);
  const double* seisb = seisb_all + K.seis_off;
  cd* work = work_all + K.seis_off;
  switch (g) {
    case 1: post_T_g<1>(K, k, i0, tw, seisb, ca, sa, work); break;
    case 2: post_T_g<2>(K, k, i0, tw, seisb, ca, sa, work); break;
    case 4: post_T_g<4>(K, k, i0, tw, seisb, ca, sa, work); break;
    case 8: post_T_g<8>(K, k, i0, tw, seisb, ca, sa, work); break;
    default: post_T_g<16>(K, k, i0, tw, seisb, ca, sa, work); break;
  }
}

// blocks (receiver k, component c, block of min(nt, FFT_B) elements), in place
__global__ void __launch_bounds__(512) k_fft_lds_T(const RunK* __restrict__ runs, int e,
                                                   const cd* __restrict__ tw,
                                                   cd* __restrict__ work_all) {
  __shared__ cd buf[FFT_B];
  const RunK& K = runs[e];
  const int nt = K.nt, bk = min(nt, FFT_B), nblk = nt / bk;
  const int k = blockIdx.x / (3 * nblk), rem = blockIdx.x % (3 * nblk);
  const int c = rem / nblk, blk = rem % nblk;
  if (k >= K.nr) return;
  cd* row = work_all + K.seis_off + ((size_t)k * 3 + c) * nt + (size_t)blk * bk;
  for (int q = threadIdx.x; q < bk; q += blockDim.x) buf[q] = row[q];
  __syncthreads();
  for (int mmax = bk >> 1; mmax >= 1; mmax >>= 1) {
    for (int b = threadIdx.x; b < bk / 2; b += blockDim.x) {
      int m = b & (mmax - 1);
      int i = ((b - m) << 1) + m;
      int jx = i + mmax;
      cd x = buf[i], y = buf[jx];
      buf[i] = x + y;
      buf[jx] = conj_(tw[mmax + m]) * (x - y);
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < bk; q += blockDim.x) row[q] = buf[q];
}

// blocks (receiver k, component c, block of 256 bins)
__global__ void __launch_bounds__(256) k_pack_T(const RunK* __restrict__ runs, int e,
                                                const cd* __restrict__ work_all,
                                                cd* __restrict__ specb_all) {
  const RunK& K = runs[e];
  const int nt = K.nt, nwb = (K.nw + 255) / 256;
  const int k = blockIdx.x / (3 * nwb), rem = blockIdx.x % (3 * nwb);
  const int c = rem / nwb, jj = (rem % nwb) * 256 + (int)threadIdx.x;
  if (k >= K.nr || jj >= K.nw) return;
  const cd* row = work_all + K.seis_off + ((size_t)k * 3 + c) * nt;
  int b = jj + K.i1 - 1;
  cd v = C(0.0);
  if (b <= nt / 2) {
    v = row[brev((unsigned)b, K.lognt)];
    if (b > 0 && b < nt / 2) v = v + conj_(row[brev((unsigned)(nt - b), K.lognt)]);
    v = v * bin_filter(K, b);
  }
  specb_all[K.spec_off + (size_t)c * K.nw * K.nr + (size_t)K.nw * k + jj] = v;
}

// --------------------------------------------------------------- k_lsum_T
// Adjoint of k_lsum for one event: cf_bar[q](l, j) = sum over receivers k
// of X(l, k) conj(F(k)) spec_bar(k, j) for every Legendre table X and
// epilogue factor F through which coefficient plane q reaches the spectra,
// as FP64 MFMA GEMMs over the receivers.  A block owns a 16 (l) x 16 (w)
// tile; each of its four waves owns a fixed set of planes and runs the full
// receiver loop.  Operands of v_mfma_f64_16x16x4f64: A[i][k] = X(l0 + i,
// k0 + k) at lane i + 16k, B[k][n] = W(k0 + k, j0 + n) at lane n + 16k,
// result D[i][n] at lane n + 16 (i % 4), register i / 4.

// s * i^im * (cj ? conj(e^p) : e^p), e = exp(i phi_k)
template <int P, bool CJ, int S, bool IM>
__device__ __forceinline__ cd efac(cd e1, cd e2) {
  cd v = (P == 0) ? C(1.0) : (P == 1 ? e1 : e2);
  if (CJ) v.y = -v.y;
  v = (double)S * v;
  return IM ? cmuli(v) : v;
}

// acc += X_T x (factor * out_bar[O])
template <int T, int O, int P, bool CJ, int S, bool IM>
__device__ __forceinline__ void lsumT_term(d4* acc, const double* a, const cd* ob, cd e1, cd e2) {
  cd w = efac<P, CJ, S, IM>(e1, e2) * ob[O];
  acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[T], w.x, acc[0], 0, 0, 0);
  acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[T], w.y, acc[1], 0, 0, 0);
}

// the conjugated factors of the forward epilogue (lsum_group), per plane:
// uu m -> plane m + 2, vv m -> 7 + m, ww m -> 12 + m; table t = 3 X + |m|
// (X: 0 x, 1 dx/dtheta, 2 x/sin theta); out 0 ur, 1 ut, 2 up
template <int WAVE>
__device__ __forceinline__ void lsumT_terms(d4 (*acc)[2], const double* a, const cd* ob, cd e1,
                                            cd e2) {
  if constexpr (WAVE == 0) {
    lsumT_term<0, 0, 0, false, 1, false>(acc[0], a, ob, e1, e2);   // uu0:  x0, ur, 1
    lsumT_term<1, 0, 1, true, 1, false>(acc[1], a, ob, e1, e2);    // uu+1: x1, ur, conj e
    lsumT_term<1, 0, 1, false, -1, false>(acc[2], a, ob, e1, e2);  // uu-1: x1, ur, -e
    lsumT_term<2, 0, 2, true, 1, false>(acc[3], a, ob, e1, e2);    // uu+2: x2, ur, conj e^2
    lsumT_term<2, 0, 2, false, 1, false>(acc[4], a, ob, e1, e2);   // uu-2: x2, ur, e^2
    lsumT_term<3, 1, 0, false, 1, false>(acc[5], a, ob, e1, e2);   // vv0:  dx0, ut, 1
  } else if constexpr (WAVE == 1) {
    lsumT_term<4, 1, 1, true, 1, false>(acc[0], a, ob, e1, e2);    // vv+1: dx1, ut, conj e
    lsumT_term<7, 2, 1, true, -1, true>(acc[0], a, ob, e1, e2);    //       xs1, up, -i conj e
    lsumT_term<4, 1, 1, false, -1, false>(acc[1], a, ob, e1, e2);  // vv-1: dx1, ut, -e
    lsumT_term<7, 2, 1, false, -1, true>(acc[1], a, ob, e1, e2);   //       xs1, up, -i e
    lsumT_term<5, 1, 2, true, 1, false>(acc[2], a, ob, e1, e2);    // vv+2: dx2, ut, conj e^2
    lsumT_term<8, 2, 2, true, -2, true>(acc[2], a, ob, e1, e2);    //       xs2, up, -2i conj e^2
  } else if constexpr (WAVE == 2) {
    lsumT_term<5, 1, 2, false, 1, false>(acc[0], a, ob, e1, e2);   // vv-2: dx2, ut, e^2
    lsumT_term<8, 2, 2, false, 2, true>(acc[0], a, ob, e1, e2);    //       xs2, up, 2i e^2
    lsumT_term<1, 1, 1, true, -1, true>(acc[1], a, ob, e1, e2);    // ww+1: x1, ut, -i conj e
    lsumT_term<4, 2, 1, true, -1, false>(acc[1], a, ob, e1, e2);   //       dx1, up, -conj e
    lsumT_term<1, 1, 1, false, -1, true>(acc[2], a, ob, e1, e2);   // ww-1: x1, ut, -i e
    lsumT_term<4, 2, 1, false, 1, false>(acc[2], a, ob, e1, e2);   //       dx1, up, e
  } else {
    lsumT_term<2, 1, 2, true, -2, true>(acc[0], a, ob, e1, e2);    // ww+2: x2, ut, -2i conj e^2
    lsumT_term<5, 2, 2, true, -1, false>(acc[0], a, ob, e1, e2);   //       dx2, up, -conj e^2
    lsumT_term<2, 1, 2, false, 2, true>(acc[1], a, ob, e1, e2);    // ww-2: x2, ut, 2i e^2
    lsumT_term<5, 2, 2, false, -1, false>(acc[1], a, ob, e1, e2);  //       dx2, up, -e^2
  }
}

template <int WAVE>
__device__ __forceinline__ void lsumT_wave(const RunK& R, const double* __restrict__ tb,
                                           const cd* __restrict__ eia,
                                           const cd* __restrict__ specb, cd* __restrict__ cfb,
                                           int l0, int j0) {
  // planes of this wave, in the order of lsumT_terms
  constexpr int NP = WAVE == 0 ? 6 : 3;
  constexpr int planes[4][6] = {{2, 3, 1, 4, 0, 7}, {8, 6, 9}, {5, 13, 11}, {14, 10, 12}};
  constexpr int ntab = WAVE == 0 ? 4 : 9;
  const int lane = threadIdx.x & 63, li = lane & 15, lq = lane >> 4;
  const size_t tplane = (size_t)R.lpad * R.nrpad, nwr = (size_t)R.nw * R.nr;
  const int l = l0 + li, j = j0 + li;
  const bool lok = l < R.lpad, jok = j < R.nw;
  d4 acc[NP][2];
#pragma unroll
  for (int s = 0; s < NP; s++) acc[s][0] = acc[s][1] = (d4){0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < R.nrpad; k0 += 4) {
    const int k = k0 + lq;
    const bool kok = k < R.nr;
    double a[ntab];
#pragma unroll
    for (int t = 0; t < ntab; t++) a[t] = 0.0;
    // only the tables this wave's terms read
#pragma unroll
    for (int t = 0; t < ntab; t++) {
      constexpr unsigned used[4] = {0xfu, 0x1b0u, 0x132u, 0x24u};
      if ((used[WAVE] >> t) & 1u) a[t] = lok ? tb[t * tplane + (size_t)l * R.nrpad + k] : 0.0;
    }
    cd ob[3];
#pragma unroll
    for (int o = 0; o < 3; o++)
      ob[o] = (kok && jok) ? specb[o * nwr + (size_t)k * R.nw + j] : C(0.0);
    cd e1 = kok ? eia[k] : C(1.0), e2 = e1 * e1;
    lsumT_terms<WAVE>(acc, a, ob, e1, e2);
  }
  const size_t cplane = (size_t)R.lpad * R.jpad;
#pragma unroll
  for (int s = 0; s < NP; s++) {
#pragma unroll
    for (int v = 0; v < 4; v++) {
      int lo = l0 + lq + 4 * v;
      if (lo < R.lpad) cfb[planes[WAVE][s] * cplane + (size_t)lo * R.jpad + j] = C(acc[s][0][v], acc[s][1][v]);
    }
  }
  if constexpr (WAVE == 3) {
    // ww0 does not reach the spectra
#pragma unroll
    for (int v = 0; v < 4; v++) {
      int lo = l0 + lq + 4 * v;
      if (lo < R.lpad) cfb[12 * cplane + (size_t)lo * R.jpad + j] = C(0.0);
    }
  }
}

// one block of four waves per 16 x 16 tile; grid ceil(lpad / 16) * jpad / 16
__global__ void __launch_bounds__(256) k_lsum_T(const RunK* __restrict__ runs, int e,
                                                const double* __restrict__ tb_all,
                                                const cd* __restrict__ eia_all,
                                                const cd* __restrict__ specb_all,
                                                cd* __restrict__ cfb_all) {
  const RunK& R = runs[e];
  const int njt = R.jpad / 16;
  const int j0 = (blockIdx.x % njt) * 16, l0 = (blockIdx.x / njt) * 16;
  const double* tb = tb_all + R.tb_off;
  const cd* eia = eia_all + R.roff;
  const cd* specb = specb_all + R.spec_off;
  cd* cfb = cfb_all + R.cf_off;
  switch (threadIdx.x >> 6) {
    case 0: lsumT_wave<0>(R, tb, eia, specb, cfb, l0, j0); break;
    case 1: lsumT_wave<1>(R, tb, eia, specb, cfb, l0, j0); break;
    case 2: lsumT_wave<2>(R, tb, eia, specb, cfb, l0, j0); break;
    default: lsumT_wave<3>(R, tb, eia, specb, cfb, l0, j0); break;
  }
}

// ----------------------------------------------------------- k_assemble_T
// Adjoint of k_assemble, for one pair: from the coefficient adjoints cf_bar
// (uu, vv, ww) and the pair's inputs o (the forward's record fields
// [0, O_UU)) to the adjoints of the inputs svt, ytor, svs, ysph (rows
// 0..2), sr0, yrad.  ob gets the record layout: input adjoints in
// [0, O_UU), the coefficient adjoints in the uu, vv, ww slots.
__device__ __forceinline__ void assemble_T_pair(const RunK& R, int l, int j, const cd* o,
                                                const cd* __restrict__ cfb, cd* ob) {
  const size_t plane = (size_t)R.lpad * R.jpad, at = (size_t)l * R.jpad + (j - 1);
  cd c[15];
  for (int m = 0; m < 15; m++) c[m] = cfb[m * plane + at];
  cd w = C(R.wmin + (double)(j - 1) * R.dw, -R.ep);
  cd fo = (R.out == 1) ? cmuli(w) : (R.out == 2 ? -(w * w) : C(1.0));
  cd fc = conj_(fo);
  double rr = R.rr;
  for (int f = 0; f < O_UU; f++) ob[f] = C(0.0);
  if (l == 0) {
    // c2 = fo (sr0_1 yrad_11 + sr0_2 yrad_12) / rr
    cd g = fc * c[2] / rr;
    ob[O_SR0] = conj_(o[O_YRAD + 0]) * g;
    ob[O_SR0 + 1] = conj_(o[O_YRAD + 2]) * g;
    ob[O_YRAD + 0] = conj_(o[O_SR0]) * g;
    ob[O_YRAD + 2] = conj_(o[O_SR0 + 1]) * g;
  } else {
    double zeta = (double)sqrtf((float)(l * (l + 1)));
    if (l > 1) {
      // ww_m = (svt_m1 fo ytor_11 + svt_m2 fo ytor_12) / (zeta rr)
      cd t1 = fo * o[O_YTOR + 0], t2 = fo * o[O_YTOR + 2], t1b = C(0.0), t2b = C(0.0);
      for (int m = 0; m < 5; m++) {
        cd g = c[10 + m] / (zeta * rr);
        ob[O_SVT + m] = conj_(t1) * g;
        ob[O_SVT + 5 + m] = conj_(t2) * g;
        t1b = t1b + conj_(o[O_SVT + m]) * g;
        t2b = t2b + conj_(o[O_SVT + 5 + m]) * g;
      }
      ob[O_YTOR + 0] = fc * t1b;
      ob[O_YTOR + 2] = fc * t2b;
    }
    const bool cor = (R.out == 2 && R.cor != 0);
    cd y1[4], y2[4];
    for (int k = 0; k < 4; k++) {
      cd a = o[O_YSPH + 6 * k + 0], bb = o[O_YSPH + 6 * k + 1], g3 = o[O_YSPH + 6 * k + 2];
      if (cor) {
        y1[k] = fo * a + (-((double)(l + 1) * g3 + (2.0 * R.grav_ir) * a) / rr);
        y2[k] = fo * bb + zeta * (g3 + R.grav_ir * a) / rr;
      } else {
        y1[k] = fo * a;
        y2[k] = fo * bb;
      }
    }
    // uu_m = sum_k y1_k sv_(m, col k) / rr, vv_m = sum_k y2_k sv_(m, col k) / (zeta rr)
    // over the source-vector columns 1, 2, 4, 5
    constexpr int off[4] = {0, 5, 15, 20};
    cd y1b[4], y2b[4];
    for (int k = 0; k < 4; k++) y1b[k] = y2b[k] = C(0.0);
    for (int m = 0; m < 5; m++) {
      cd g1 = c[m] / rr, g2 = c[5 + m] / (zeta * rr);
      for (int k = 0; k < 4; k++) {
        cd sv = o[O_SVS + off[k] + m];
        ob[O_SVS + off[k] + m] = conj_(y1[k]) * g1 + conj_(y2[k]) * g2;
        y1b[k] = y1b[k] + conj_(sv) * g1;
        y2b[k] = y2b[k] + conj_(sv) * g2;
      }
    }
    for (int k = 0; k < 4; k++) {
      cd* yb = ob + O_YSPH + 6 * k;
      if (cor) {
        yb[0] = conj_(fo - 2.0 * R.grav_ir / rr) * y1b[k] + (zeta * R.grav_ir / rr) * y2b[k];
        yb[1] = fc * y2b[k];
        yb[2] = (-(double)(l + 1) / rr) * y1b[k] + (zeta / rr) * y2b[k];
      } else {
        yb[0] = fc * y1b[k];
        yb[1] = fc * y2b[k];
      }
    }
  }
  for (int m = 0; m < 15; m++) ob[O_UU + m] = c[m];
}

// ASMT_P consecutive pairs of one event per 64-lane block: their inputs are
// copied to LDS with coalesced loads, one lane per pair computes into a
// second LDS array, and the ASMT_P contiguous 89-field record adjoints are
// stored linearly.
__global__ void __launch_bounds__(64) k_assemble_T(const RunK* __restrict__ runs, int e,
                                                   const cd* __restrict__ rec,
                                                   const cd* __restrict__ cfb_all,
                                                   cd* __restrict__ recb) {
  __shared__ cd si[ASMT_P * ASM_LD];
  __shared__ cd so[ASMT_P * REC];  // stride 89: the 16 lanes' 16-byte writes hit distinct banks
  const RunK& R = runs[e];
  const long long p0 = (long long)blockIdx.x * ASMT_P;
  if (p0 >= R.npair) return;
  const int np = (int)min((long long)ASMT_P, (long long)R.npair - p0);
  const cd* src = rec + (size_t)(R.pbase + p0) * REC;
  {
    // all loads issued before the first LDS write; indices past the last
    // field are clamped to it, so no load is predicated
    constexpr int NL = (ASMT_P * O_UU + 63) / 64;
    const int kmax = np * O_UU - 1;
    cd v[NL];
#pragma unroll
    for (int i = 0; i < NL; i++) {
      int k = min((int)threadIdx.x + 64 * i, kmax), q = k / O_UU, f = k - q * O_UU;
      v[i] = src[(size_t)q * REC + f];
    }
#pragma unroll
    for (int i = 0; i < NL; i++) {
      int k = min((int)threadIdx.x + 64 * i, kmax), q = k / O_UU, f = k - q * O_UU;
      si[q * ASM_LD + f] = v[i];
    }
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < np) {
    int l, j;
    pair_lj(R, (int)(p0 + t), l, j);
    assemble_T_pair(R, l, j, si + t * ASM_LD, cfb_all + R.cf_off, so + t * REC);
  }
  __syncthreads();
  cd* dst = recb + (size_t)(R.pbase + p0) * REC;
  for (int k = t; k < np * REC; k += 64) dst[k] = so[k];
}

// ------------------------------------------------------------- k_source_T
// Gradient with respect to the normalised moment tensor from the record
// adjoints recb of one event: the source vectors are linear in mm, so
// dF/dmm_i = sum over pairs and source-vector fields f of
// Re(conj(recb_f) sv_f(e_i)), e_i the i-th unit tensor.  One lane per pair;
// part[6 b + i] gets block b's sum (the host adds the blocks in order, so
// the result does not depend on scheduling).
__global__ void __launch_bounds__(256) k_source_T(const double* __restrict__ gtab,
                                                  const RunK* __restrict__ runs, int e,
                                                  const cd* __restrict__ recb,
                                                  double* __restrict__ part) {
  __shared__ double red[4][6];
  const RunK& R = runs[e];
  Tab T{gtab, R.nknot};
  const int p = blockIdx.x * 256 + (int)threadIdx.x;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (p < R.npair) {
    int l, j;
    pair_lj(R, p, l, j);
    PS s;
    pair_state(R, l, j, s);
    const cd* b = recb + (size_t)(R.pbase + p) * REC;
#pragma unroll
    for (int i = 0; i < 6; i++) {
      double mi[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      mi[i] = 1.0;
      double a = 0.0;
      source_vectors(T, R, s, l, mi, [&](int f, cd v) { a += b[f].x * v.x + b[f].y * v.y; });
      acc[i] = a;
    }
  }
#pragma unroll
  for (int i = 0; i < 6; i++)
    for (int off = 32; off > 0; off >>= 1) acc[i] += __shfl_down(acc[i], off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
    for (int i = 0; i < 6; i++) red[wv][i] = acc[i];
  __syncthreads();
  if (threadIdx.x < 6) {
    int i = threadIdx.x;
    part[(size_t)blockIdx.x * 6 + i] = (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
  }
}

// ------------------------------------------------------- k_source_model_T
// Partial derivatives of the source vectors of one pair with respect to the
// six model values they read at the source radius, v = (L, C, F, 1/Qmu,
// xlam, xa2) (source_vectors above): put(f, d) receives for record field f
// the six complex derivatives d[0..5].  Fields that do not depend on v are
// not reported.
template <class Put>
__device__ __forceinline__ void source_vectors_dv(const Tab& T, const RunK& R, const PS& s, int l,
                                                  const double* mm, Put put) {
  SplW w = spl_w(T, R.is, R.rs);
  double L = spl(T, w, F_LCON), Cc = spl(T, w, F_CCON), Fc = spl(T, w, F_FCON),
         Qm = spl(T, w, F_QMU), X = spl(T, w, F_XLAM), A2 = spl(T, w, F_XA2);
  double rs = R.rs;
  cd nd = s.nu * cinv(cmuli(s.omega) * (rs * rs));
  cd lam = s.lno;
  const cd z = C(0.0);
  if (l == 0) {
    cd qaa = 1.0 + A2 * lam, qff = 1.0 + X * lam;
    cd cc = cinv(Cc * qaa), cf = Fc * qff;
    cd dcc_C = -(cc * cc) * qaa, dcc_A2 = -(cc * cc) * (Cc * lam);
    cd u0 = nd * (rs * mm[0]);
    cd d0[6] = {z, u0 * dcc_C, z, z, z, u0 * dcc_A2};
    put(O_SR0 + 0, d0);
    cd u1 = nd * (2.0 * mm[0]);
    cd d1[6] = {z, u1 * cf * dcc_C, u1 * qff * cc, z, u1 * (Fc * lam) * cc, u1 * cf * dcc_A2};
    put(O_SR0 + 1, d1);
    return;
  }
  if (l > 1) {
    cd ql = 1.0 + Qm * lam;
    cd icl = cinv(L * ql);
    cd dicl_L = -(icl * icl) * ql, dicl_Q = -(icl * icl) * (L * lam);
    cd u1 = nd * ((0.5 * rs) * C(-mm[2], mm[1])), u3 = nd * ((0.5 * rs) * C(mm[2], mm[1]));
    cd d1[6] = {u1 * dicl_L, z, z, u1 * dicl_Q, z, z};
    cd d3[6] = {u3 * dicl_L, z, z, u3 * dicl_Q, z, z};
    put(O_SVT + 1, d1);
    put(O_SVT + 3, d3);
  }
  double lx = lam.x;
  double qaa = 1.0 + A2 * lx, qff = 1.0 + X * lx, qll = 1.0 + Qm * lx;
  double cc = 1.0 / (Cc * qaa), cf = Fc * qff, cl = L * qll;
  double dcc_C = -(cc * cc) * qaa, dcc_A2 = -(cc * cc) * (Cc * lx);
  double dicl_L = -qll / (cl * cl), dicl_Q = -(L * lx) / (cl * cl);
  double zeta = s.zeta;
  {
    cd u = nd * C(0.5 * rs * mm[1], 0.5 * rs * mm[2]);
    cd d[6] = {u * dicl_L, z, z, u * dicl_Q, z, z};
    put(O_SVS + 1 + 5 * 1, d);
  }
  {
    cd u = nd * (rs * mm[0]);
    cd d[6] = {z, u * dcc_C, z, z, z, u * dcc_A2};
    put(O_SVS + 2 + 5 * 0, d);
  }
  {
    cd u = nd * (2.0 * mm[0]);
    cd d[6] = {z, u * (cf * dcc_C), u * (qff * cc), z, u * (Fc * lx * cc), u * (cf * dcc_A2)};
    put(O_SVS + 2 + 5 * 3, d);
  }
  {
    cd u = nd * (-zeta * mm[0]);
    cd d[6] = {z, u * (cf * dcc_C), u * (qff * cc), z, u * (Fc * lx * cc), u * (cf * dcc_A2)};
    put(O_SVS + 2 + 5 * 4, d);
  }
  {
    cd u = nd * C(-0.5 * rs * mm[1], 0.5 * rs * mm[2]);
    cd d[6] = {u * dicl_L, z, z, u * dicl_Q, z, z};
    put(O_SVS + 3 + 5 * 1, d);
  }
}

// dF/dv for the six source values v from the record adjoints recb of one
// event: sum over pairs and fields of Re(conj(recb_f) dsv_f/dv).  Layout and
// reduction as k_source_T.
__global__ void __launch_bounds__(256) k_source_model_T(const double* __restrict__ gtab,
                                                        const RunK* __restrict__ runs, int e,
                                                        const cd* __restrict__ recb,
                                                        double* __restrict__ part) {
  __shared__ double red[4][6];
  const RunK& R = runs[e];
  Tab T{gtab, R.nknot};
  const int p = blockIdx.x * 256 + (int)threadIdx.x;
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (p < R.npair) {
    int l, j;
    pair_lj(R, p, l, j);
    PS s;
    pair_state(R, l, j, s);
    const cd* b = recb + (size_t)(R.pbase + p) * REC;
    source_vectors_dv(T, R, s, l, R.mm, [&](int f, const cd* d) {
      cd bf = b[f];
#pragma unroll
      for (int k = 0; k < 6; k++) acc[k] += bf.x * d[k].x + bf.y * d[k].y;
    });
  }
#pragma unroll
  for (int i = 0; i < 6; i++)
    for (int off = 32; off > 0; off >>= 1) acc[i] += __shfl_down(acc[i], off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
    for (int i = 0; i < 6; i++) red[wv][i] = acc[i];
  __syncthreads();
  if (threadIdx.x < 6) {
    int i = threadIdx.x;
    part[(size_t)blockIdx.x * 6 + i] = (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
  }
}

}  // namespace

void launch_source_model_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                           const RunK* runs, int e, const cd* recb, double* part) {
  hipLaunchKernelGGL(k_source_model_T, g, b, lds, s, gtab, runs, e, recb, part);
}

void launch_residual(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* seis,
                     const double* obs, int nt, int nout, size_t ntrace, double* seisb,
                     double* part) {
  hipLaunchKernelGGL(k_residual, g, b, lds, s, seis, obs, nt, nout, ntrace, seisb, part);
}
void launch_reset(dim3 g, dim3 b, size_t lds, hipStream_t s, int* counters, int n) {
  hipLaunchKernelGGL(k_reset, g, b, lds, s, counters, n);
}
void launch_source(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                   const RunK* runs, PairRange PR, cd* rec, int* cnt) {
  hipLaunchKernelGGL(k_source, g, b, lds, s, gtab, runs, PR, rec, cnt);
}
void launch_start(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                  const RunK* runs, PairRange PR, StartData SD, int* starts, int parts) {
  hipLaunchKernelGGL(k_start, g, b, lds, s, gtab, runs, PR, SD, starts, parts);
}
void launch_rad(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab, const RunK* runs,
                Queue Q, cd* rec, int* cnt) {
  hipLaunchKernelGGL(k_rad, g, b, lds, s, gtab, runs, Q, rec, cnt);
}
void launch_rad_ng(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                   const RunK* runs, Queue Q, cd* rec, int* cnt) {
  hipLaunchKernelGGL(k_rad_ng, g, b, lds, s, gtab, runs, Q, rec, cnt);
}
void launch_assemble(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, PairRange PR,
                     cd* rec, cd* cf, const double* tb) {
  hipLaunchKernelGGL(k_assemble, g, b, lds, s, runs, PR, rec, cf, tb);
}
void launch_lsum(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, Queue Q, cd* cf,
                 const double* tb, const cd* eia, cd* spec) {
  hipLaunchKernelGGL(k_lsum, g, b, lds, s, runs, Q, cf, tb, eia, spec);
}
void launch_legendre(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                     const double* th, double* tb) {
  hipLaunchKernelGGL(k_legendre, g, b, lds, s, runs, e, th, tb);
}
void launch_model(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* deck, int nknot,
                  int ifanis, double vel_norm, const int* layer_lo, const int* layer_n,
                  int nlayer, double* tab, double* work, size_t ws, double* intg, double* qint,
                  double* gi) {
  hipLaunchKernelGGL(k_model, g, b, lds, s, deck, nknot, ifanis, vel_norm, layer_lo, layer_n,
                     nlayer, tab, work, ws, intg, qint, gi);
}
void launch_model_vjp(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* deck, int nknot,
                      int ifanis, double vel_norm, const int* layer_lo, const int* layer_n,
                      int nlayer, const double* tab, const double* tbar, double* bar,
                      double* work, size_t ws, double* aux, double* dbar) {
  hipLaunchKernelGGL(k_model_vjp, g, b, lds, s, deck, nknot, ifanis, vel_norm, layer_lo, layer_n,
                     nlayer, tab, tbar, bar, work, ws, aux, dbar);
}
void launch_fill0(dim3 g, dim3 b, size_t lds, hipStream_t s, double* p, size_t n) {
  hipLaunchKernelGGL(k_fill0, g, b, lds, s, p, n);
}
void launch_seis(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, Queue Q,
                 const cd* spec, const cd* tw, cd* work, const double* azep, double* seis) {
  hipLaunchKernelGGL(k_seis, g, b, lds, s, runs, Q, spec, tw, work, azep, seis);
}
void launch_fft_lds(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, Queue Q,
                    const cd* spec, const cd* tw, cd* work) {
  hipLaunchKernelGGL(k_fft_lds, g, b, lds, s, runs, Q, spec, tw, work);
}
void launch_fft_post(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, Queue Q,
                     const cd* tw, const cd* work, const double* azep, double* seis) {
  hipLaunchKernelGGL(k_fft_post, g, b, lds, s, runs, Q, tw, work, azep, seis);
}
void launch_twiddles(dim3 g, dim3 b, size_t lds, hipStream_t s, int nt, cd* tw) {
  hipLaunchKernelGGL(k_twiddles, g, b, lds, s, nt, tw);
}
void launch_seis_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                   const double* seisb, const cd* tw, cd* work, const double* azep, cd* specb) {
  hipLaunchKernelGGL(k_seis_T, g, b, lds, s, runs, e, seisb, tw, work, azep, specb);
}
void launch_post_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                   const double* seisb, const cd* tw, const double* azep, cd* work) {
  hipLaunchKernelGGL(k_post_T, g, b, lds, s, runs, e, seisb, tw, azep, work);
}
void launch_fft_lds_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                      const cd* tw, cd* work) {
  hipLaunchKernelGGL(k_fft_lds_T, g, b, lds, s, runs, e, tw, work);
}
void launch_pack_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                   const cd* work, cd* specb) {
  hipLaunchKernelGGL(k_pack_T, g, b, lds, s, runs, e, work, specb);
}
void launch_lsum_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                   const double* tb, const cd* eia, const cd* specb, cd* cfb) {
  hipLaunchKernelGGL(k_lsum_T, g, b, lds, s, runs, e, tb, eia, specb, cfb);
}
void launch_assemble_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const RunK* runs, int e,
                       const cd* rec, const cd* cfb, cd* recb) {
  hipLaunchKernelGGL(k_assemble_T, g, b, lds, s, runs, e, rec, cfb, recb);
}
void launch_source_T(dim3 g, dim3 b, size_t lds, hipStream_t s, const double* gtab,
                     const RunK* runs, int e, const cd* recb, double* part) {
  hipLaunchKernelGGL(k_source_T, g, b, lds, s, gtab, runs, e, recb, part);
}

}  // namespace ysx
