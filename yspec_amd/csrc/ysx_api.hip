// ysx_api.hip -- the C-ABI of libysx.so (include/ysx.h): contexts, model
// and run set-up, work lists, the plain and batched forwards and their
// captured HIP graph, result copies.  Kernels are launched through the
// launch functions of ysx_kern.hpp.
//
// Per forward the GPU work is:
//   k_source      source vectors of every (l, w) pair
//   k_start       start levels and starting solutions
//   k_tor_march   toroidal two-sided shooting, lockstep waves of 64 l at one w
//   k_minor_top   spheroidal minors from the receiver side down to the source
//   k_minor_march spheroidal minor vectors from the start level up to the
//                 source, secular determinant, b-vector starts
//   k_brow_march  b-vector system: receiver-side rows marched back to the
//                 source with transposed steps (k_bvec_march: the reference's
//                 4 source-side columns, for full ysph records)
//   k_rad         radial (l = 0) solves on the reference's grid (2nd stream)
//   k_ngc_march, k_rad_ng solvers of the gravity switches 0 / 1
//   k_assemble    output-type scaling and the uu/vv/ww coefficients
//   k_lsum        sum over l and m against the Legendre tables: FP64 MFMA
//                 GEMMs with the azimuthal phases fused into the epilogue
//   k_fft_lds, k_fft_post (k_seis for long series)
//                 filtered inverse FFT, exp(eps t), units, rotation
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
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ysx.h"
#include "ysx_host.hpp"
#include "ysx_adjoint.hpp"
#include "ysx_kern.hpp"
#include "ysx_radial.hpp"

using namespace ysx;

// ====================================================================== API
namespace {

enum WorkKind { W_TOR = 0, W_BVEC, W_RAD, W_LSUM, W_SEIS, W_FFTA, W_FFTB, W_DROW, W_N };
// queue counters of the batched launches (three marches share the W_TOR list)
enum QueueSlot { Q_TOR = 0, Q_TOP, Q_MINOR, Q_BVEC, Q_RAD, Q_DROW, Q_N };

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
  hipStream_t gstream[2] = {};  // the minor sweep's further item groups
  std::string err;
  Norms N;
  HostModel M;
  std::vector<HostRun> R;  // one per event
  std::vector<RunK> rk;
  bool have_model = false, have_run = false, have_spec = false, have_seis = false;
  bool refined = false;  // wave lists reordered by measured cost (refine_work)
  bool bcols = false;    // b-vector as source-side columns (full ysph records)
  size_t lds_bytes = 0, lds_max = 0;  // model table bytes, LDS per block
  bool tab_lds = true;                 // the table fits in LDS beside the kernels' own
  long long npair = 0;  // all events
  // device buffers
  double* d_tab = nullptr;
  RunK* d_runs = nullptr;
  cd* d_rec = nullptr;
  int* d_starts = nullptr;
  int* d_cnt = nullptr;
  cd *d_bst = nullptr, *d_mfin = nullptr, *d_m2 = nullptr, *d_fac = nullptr, *d_ms1 = nullptr;
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
  // adjoints (one-event runs): seismograms, spectra, coefficient planes and
  // per-pair records, allocated by the first ysx_seismograms_vjp or
  // ysx_spectra_vjp of a run
  double* d_seisb = nullptr;
  cd *d_specb = nullptr, *d_cfb = nullptr, *d_recb = nullptr;
  double* d_mmpart = nullptr;  // per-block sums of the moment-tensor gradient
  double ms_vjp[3] = {0};      // kernel ms of the last call of each adjoint entry point
  bool have_specb = false, have_recb = false;
  WorkList W[W_N];
  // events whose b-vector rows share one march in the batched path (same
  // receiver depth and frequency grid), by source radius, shallowest first
  std::vector<int> sh_order;
  int sh_lmin = 0, sh_lmax = 0, sh_nw = 0;
  int n_union = 0;   // scratch slots: the larger of the two union lists
  int n_ubrow = 0;   // union items of the b-vector rows
  int* d_shorder = nullptr;
  cd* d_shx = nullptr;
  int* d_shs = nullptr;
  // batched mode: persistent blocks or one block per item, items taken per
  // queue access, persistent blocks per CU, and the captured forward
  int persist = 0, chunk = 1, blocks_per_cu = 4;
  // lanes per radial (l = 0) solve: 0 automatic, 1 or 8 (ysx_set_lanes)
  int rad_lanes = 0;
  // the last plain forward ran its toroidal chain on the second stream
  bool tor_side = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t ev[20] = {};
  double ms[8] = {0};
  double* scratch = nullptr;  // this device's copy of `scratch` (ysx_model_table_noalloc)
  // model gradient (ysx_model_vjp): per-pair adjoints of the minors at the
  // source from the b-vector starts (mub), of the starting solutions (m0bar,
  // y0bar), the table adjoint, the sweeps' per-block checkpoints and tapes
  cd *d_mub = nullptr, *d_m0bar = nullptr, *d_y0bar = nullptr;
  double *d_tbar = nullptr, *d_tpart = nullptr, *d_tpart2 = nullptr;
  double* d_gtpart[2] = {};
  cd *d_ays = nullptr, *d_ascr = nullptr;
  cd* d_gscr[2] = {};
  int *d_ayss = nullptr, *d_aerr = nullptr, *d_aq = nullptr;
  int adj_blocks = 0, adj_smax = 0;
  // the minor and toroidal marches' per-step states stored by the
  // single-event forward of ysx_misfit_grad for the sweeps (StepStore):
  // states, counters, per-pair counters (and the toroidal continuation's
  // end), shared per-w item offsets and capacities and overflow flag;
  // fs_valid while the records of that forward are the context's
  cd *d_fys = nullptr, *d_tys = nullptr, *d_tyend = nullptr;
  int *d_fyss = nullptr, *d_tyss = nullptr;
  // the minor sweep's row states and pairing units (RowStore) beside them
  cd* d_fxs = nullptr;
  int *d_fxss = nullptr, *d_pq = nullptr;
  PairUnit* d_units = nullptr;
  int units_cap = 0;
  long long* d_foff = nullptr;
  int *d_fcap = nullptr, *d_fsc = nullptr, *d_tsc = nullptr, *d_ferr = nullptr;
  size_t fs_steps = 0;
  bool fs_want = false, fs_valid = false;
  // radial adjoint: segment propagators, start states, end rows, tape
  cd *d_rprop = nullptr, *d_rys = nullptr, *d_rxe = nullptr;
  int* d_rcap = nullptr;
  void* d_rtape = nullptr;
  size_t rtape_bytes = 0;
  std::vector<int> rad_caps;  // per event from degree 0: the tape's length
  double ms_grad[7] = {0};  // kernel ms of the last ysx_model_vjp
  // pinned staging of large host copies (copy_d2h / copy_h2d) and the
  // events that order its two halves
  void* pin = nullptr;
  size_t pin_bytes = 0;
  void* opin = nullptr;  // ysx_misfit_grad's observed seismograms, staged
  size_t opin_bytes = 0;
  hipEvent_t pev[2] = {};
  // ysx_misfit_grad: the observed traces and the residual's per-block sums
  double *d_obs = nullptr, *d_rpart = nullptr;
  size_t obs_n = 0;
  // the deck as set (row-major nknot x 9) for the deck -> table adjoint,
  // and on the device: the staged deck and layer lists k_model read
  // (layer_lo, layer_n), and k_model_vjp's buffers (allocated by its first
  // call for the model)
  std::vector<double> deck;
  double* d_mdeck = nullptr;
  int* d_mlay = nullptr;
  double* d_mvjp = nullptr;
  int ians = 0;
  double tref_period = 0.0;
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
  dfree(c->d_mfin); dfree(c->d_ms1); dfree(c->d_m2); dfree(c->d_tor_i1); dfree(c->d_sph_i1); dfree(c->d_tor_y0);
  dfree(c->d_sph_m0); dfree(c->d_fac); dfree(c->d_sc3); dfree(c->d_azep); dfree(c->d_eia);
  dfree(c->d_spec); dfree(c->d_cf); dfree(c->d_tb); dfree(c->d_tw); dfree(c->d_work);
  dfree(c->d_seis); dfree(c->d_queue); dfree(c->d_shorder); dfree(c->d_shx); dfree(c->d_shs);
  dfree(c->d_seisb); dfree(c->d_specb); dfree(c->d_cfb); dfree(c->d_recb); dfree(c->d_mmpart);
  dfree(c->d_mub); dfree(c->d_m0bar); dfree(c->d_y0bar); dfree(c->d_tbar); dfree(c->d_tpart);
  dfree(c->d_tpart2);
  for (int g = 0; g < 2; g++) {
    dfree(c->d_gtpart[g]);
    dfree(c->d_gscr[g]);
  }
  dfree(c->d_ays); dfree(c->d_ayss); dfree(c->d_ascr); dfree(c->d_aerr);
  dfree(c->d_aq);
  dfree(c->d_rtape); dfree(c->d_rprop); dfree(c->d_rys); dfree(c->d_rxe); dfree(c->d_rcap);
  dfree(c->d_obs); dfree(c->d_rpart);
  dfree(c->d_fys); dfree(c->d_fyss); dfree(c->d_foff); dfree(c->d_fcap); dfree(c->d_fsc);
  dfree(c->d_ferr); dfree(c->d_tys); dfree(c->d_tyss); dfree(c->d_tsc); dfree(c->d_tyend);
  dfree(c->d_fxs); dfree(c->d_fxss); dfree(c->d_pq); dfree(c->d_units);
  c->units_cap = 0;
  c->fs_steps = 0;
  c->fs_want = c->fs_valid = false;
  c->obs_n = 0;
  c->rad_caps.clear();
  c->adj_blocks = c->adj_smax = 0;
  c->rtape_bytes = 0;
  c->have_specb = c->have_recb = false;
  c->sh_order.clear();
  c->n_union = 0;
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

// Work lists of every kernel kind.  The plain list holds each event's items
// (one launch per event), the batched lists every event's items, both
// sorted by an estimate of their cost, most expensive first:
// steps grow with Re w; a wave of degrees starting at l0 integrates from
// about the turning radius (l0 + 1/2) v / w (v ~ 1 in normalised units) up
// to the top point; the b-vector runs between source and receiver.
void build_work(ysx_ctx* c) {
  int nev = (int)c->rk.size();
  std::vector<std::vector<double>> cost(W_N);
  for (int k = 0; k < W_N; k++) c->W[k].start.assign(1, 0);
  for (int e = 0; e < nev; e++) {
    const RunK& R = c->rk[e];
    int nl = R.lmax - R.lmin + 1, nlb = (nl + 63) / 64, nlbb = (nl + R.bl - 1) / R.bl;
    auto rew = [&](int j) { return R.wmin + (double)(j - 1) * R.dw; };
    double rtop = std::max(R.rs, R.rr);
    for (int wv = 0; wv < R.nw * nlb; wv++) {
      int j = R.nw - wv / nlb, lb = wv % nlb;
      c->W[W_TOR].plain.push_back(make_int4(e, j, lb, 0));
      double w = rew(j), rt = std::min(1.0, (R.lmin + 64.0 * lb + 0.5) / std::max(w, 1e-30));
      cost[W_TOR].push_back(w * (rtop - std::min(rt, rtop) + 0.05));
    }
    for (int wv = 0; wv < R.nw * nlbb; wv++) {
      int j = R.nw - wv / nlbb, lb = wv % nlbb;
      c->W[W_BVEC].plain.push_back(make_int4(e, j, lb, 0));
      cost[W_BVEC].push_back(rew(j) * (std::fabs(R.rr - R.rs) + 1e-3));
    }
    if (R.lmin == 0) {  // blocks of 64 / radg frequencies, radg lanes each
      const int per = 64 / R.radg;
      for (int jb = 0; jb < (R.nw + per - 1) / per; jb++) {
        c->W[W_RAD].plain.push_back(make_int4(e, jb, 0, 0));
        cost[W_RAD].push_back(rew(std::min(R.nw, per * jb + per)));
      }
    }
    // receiver-sum tiles t = jt + njt kt (16 w x 16 receivers).  Blocks go
    // to the 8 XCDs in turn (block b on XCD b % 8), so the list gives each
    // XCD whole frequency tiles, all their receiver tiles one after another:
    // the coefficient rows of a frequency tile are then read into one L2.
    {
      const int njt = R.jpad / 16, nkt = R.nrpad / 16, nx = 8;
      std::vector<std::vector<int>> seq(nx);
      for (int jt = 0; jt < njt; jt++)
        for (int kt = 0; kt < nkt; kt++) seq[jt % nx].push_back(jt + njt * kt);
      size_t longest = 0;
      for (auto& q : seq) longest = std::max(longest, q.size());
      for (size_t i = 0; i < longest; i++)
        for (int x = 0; x < nx; x++)
          if (i < seq[x].size()) {
            c->W[W_LSUM].plain.push_back(make_int4(e, seq[x][i], 0, 0));
            cost[W_LSUM].push_back((double)R.lpad);
          }
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
  // plain order: each event's items by decreasing estimated cost as well
  // (stable, so lists of equal costs keep their order)
  for (int k = 0; k < W_N; k++) {
    WorkList& w = c->W[k];
    for (int e = 0; e < nev; e++) {
      int a = w.start[e], b = w.start[e + 1];
      std::vector<int> ord(b - a);
      std::iota(ord.begin(), ord.end(), a);
      std::stable_sort(ord.begin(), ord.end(),
                       [&](int x, int y) { return cost[k][x] > cost[k][y]; });
      std::vector<int4> items(b - a);
      std::vector<double> cs(b - a);
      for (int i = 0; i < b - a; i++) {
        items[i] = w.plain[ord[i]];
        cs[i] = cost[k][ord[i]];
      }
      std::copy(items.begin(), items.end(), w.plain.begin() + a);
      std::copy(cs.begin(), cs.end(), cost[k].begin() + a);
    }
  }
  // Batched b-vector rows: events with the same receiver depth, frequency
  // grid and row count (all receivers above their sources, full gravity,
  // no per-lane variant) share one row march per (w, degree block) of
  // their union grid; their own items leave the batched list.
  c->sh_order.clear();
  c->n_union = 0;
  std::vector<char> member(nev, 0);
  {
    int e0 = -1;
    for (int e = 0; e < nev && e0 < 0; e++) {
      const RunK& k = c->rk[e];
      if (k.brow && !k.ngc && k.rr >= k.rs && k.lmax >= 1) e0 = e;
    }
    if (e0 >= 0) {
      const RunK& a = c->rk[e0];
      for (int e = 0; e < nev; e++) {
        const RunK& k = c->rk[e];
        member[e] = k.brow && !k.ngc && k.rr >= k.rs && k.lmax >= 1 && k.rr == a.rr &&
                    k.ir == a.ir && k.wmin == a.wmin && k.dw == a.dw && k.ep == a.ep &&
                    k.tref == a.tref && k.ats == a.ats && k.bl == a.bl;
        if (member[e]) c->sh_order.push_back(e);
      }
    }
    if (c->sh_order.size() < 2) {
      c->sh_order.clear();
      std::fill(member.begin(), member.end(), 0);
    }
    std::stable_sort(c->sh_order.begin(), c->sh_order.end(),
                     [&](int x, int y) { return c->rk[x].rs > c->rk[y].rs; });
    for (int e = 0; e < nev; e++) c->rk[e].drow = member[e] ? 1 : 0;
  }
  std::vector<int4> uitems, ditems;
  std::vector<double> ucost, dcost;
  if (!c->sh_order.empty()) {
    const RunK& a = c->rk[c->sh_order[0]];
    int lmin = 1 << 30, lmax = 0, nw = 0;
    for (int e : c->sh_order) {
      lmin = std::min(lmin, c->rk[e].lmin);
      lmax = std::max(lmax, c->rk[e].lmax);
      nw = std::max(nw, c->rk[e].nw);
    }
    c->sh_lmin = lmin;
    c->sh_lmax = lmax;
    c->sh_nw = nw;
    const int bl = a.bl, nlb = (lmax - lmin + 1 + bl - 1) / bl;
    for (int j = nw; j >= 1; j--)
      for (int lb = 0; lb < nlb; lb++) {
        int l0 = lmin + lb * bl, need = 0;
        double rsmin = a.rr;
        for (int e : c->sh_order) {
          const RunK& k = c->rk[e];
          if (j <= k.nw && l0 <= k.lmax && l0 + bl - 1 >= std::max(k.lmin, 1)) {
            need++;
            rsmin = std::min(rsmin, k.rs);
          }
        }
        if (!need) continue;
        uitems.push_back(make_int4(-1, j, lb, (int)uitems.size()));
        ucost.push_back((a.wmin + (double)(j - 1) * a.dw) * (a.rr - rsmin + 0.02 * need + 1e-3));
      }
    // the determinant rows: blocks of 64 degrees
    const int nlbd = (lmax - lmin + 1 + 63) / 64;
    for (int j = nw; j >= 1; j--)
      for (int lb = 0; lb < nlbd; lb++) {
        int l0 = lmin + lb * 64, need = 0;
        double rsmin = a.rr;
        for (int e : c->sh_order) {
          const RunK& k = c->rk[e];
          if (j <= k.nw && l0 <= k.lmax && l0 + 63 >= std::max(k.lmin, 1)) {
            need++;
            rsmin = std::min(rsmin, k.rs);
          }
        }
        if (!need) continue;
        ditems.push_back(make_int4(-1, j, lb, (int)ditems.size()));
        dcost.push_back((a.wmin + (double)(j - 1) * a.dw) * (a.rr - rsmin + 0.02 * need + 1e-3));
      }
    c->n_union = (int)std::max(uitems.size(), ditems.size());
    c->n_ubrow = (int)uitems.size();
  }
  for (int k = 0; k < W_N; k++) {
    WorkList& w = c->W[k];
    std::vector<int4> items;
    std::vector<double> cs;
    for (size_t i = 0; i < w.plain.size(); i++) {
      if (k == W_BVEC && member[w.plain[i].x]) continue;
      items.push_back(w.plain[i]);
      cs.push_back(cost[k][i]);
    }
    if (k == W_BVEC) {
      items.insert(items.end(), uitems.begin(), uitems.end());
      cs.insert(cs.end(), ucost.begin(), ucost.end());
    }
    if (k == W_DROW) {
      items.insert(items.end(), ditems.begin(), ditems.end());
      cs.insert(cs.end(), dcost.begin(), dcost.end());
    }
    std::vector<int> ord(items.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return cs[a] > cs[b]; });
    w.batched.resize(ord.size());
    for (size_t i = 0; i < ord.size(); i++) w.batched[i] = items[ord[i]];
  }
}

// After the first forward of a run the exact cost of every march wave is
// known from the per-pair step counts: a minor-march wave marches as long
// as its longest lane (5102 FLOP per solid step, 1697 per fluid step), a
// b-vector wave as long as its pairs' b-vector steps.  The wave lists
// (W_TOR, shared by the toroidal and minor marches, and W_BVEC) are
// re-sorted by them, most expensive first, once per run; results do not
// depend on the order.
int refine_work(ysx_ctx* c) {
  if (c->refined) return YSX_OK;
  c->refined = true;
  size_t P = (size_t)c->npair;
  std::vector<int> cnt(P * 5);
  {
    ysx_ctx* ctx = c;
    HIPCHK(hipMemcpy(cnt.data(), c->d_cnt, cnt.size() * sizeof(int), hipMemcpyDeviceToHost));
  }
  // degrees of a wave: 64 (W_TOR) or bl (W_BVEC)
  auto wave_cost = [&](const int4& it, int kind) {
    if (it.x < 0) {  // shared rows: the deepest of the events' marches, plus their branches
      double m = 0.0;
      int n = 0;
      const int bl = c->rk[c->sh_order[0]].bl, l0 = c->sh_lmin + it.z * bl;
      for (int e : c->sh_order) {
        const RunK& R = c->rk[e];
        if (it.y > R.nw || l0 > R.lmax || l0 + bl - 1 < std::max(R.lmin, 1)) continue;
        int l = std::min(l0 + bl - 1, R.lmax);
        size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (it.y - 1);
        m = std::max(m, 16100.0 * cnt[5 * p + 3]);
        n++;
      }
      return m * (1.0 + 0.05 * n);
    }
    const RunK& R = c->rk[it.x];
    const int nl = (kind == W_TOR) ? 64 : R.bl;
    double m = 0.0;
    for (int q = 0; q < nl; q++) {
      int l = R.lmin + it.z * nl + q;
      if (l > R.lmax) break;
      if (l < 1) continue;
      size_t p = (size_t)R.pbase + (size_t)(l - R.lmin) * R.nw + (it.y - 1);
      double v = (kind == W_TOR) ? 5102.0 * cnt[5 * p + 1] + 1697.0 * cnt[5 * p + 2]
                                 : 16100.0 * cnt[5 * p + 3];
      m = std::max(m, v);
    }
    return m;
  };
  ysx_ctx* ctx = c;
  for (int kind : {(int)W_TOR, (int)W_BVEC}) {
    WorkList& w = c->W[kind];
    auto resort = [&](int4* first, int4* last) {
      std::vector<std::pair<double, int4>> v;
      for (int4* i = first; i != last; ++i) v.push_back({wave_cost(*i, kind), *i});
      std::stable_sort(v.begin(), v.end(),
                       [](const auto& a, const auto& b) { return a.first > b.first; });
      for (size_t k = 0; k < v.size(); k++) first[k] = v[k].second;
    };
    for (size_t e = 0; e + 1 < w.start.size(); e++)
      resort(w.plain.data() + w.start[e], w.plain.data() + w.start[e + 1]);
    resort(w.batched.data(), w.batched.data() + w.batched.size());
    if (!w.plain.empty()) {
      HIPCHK(hipMemcpy(w.d_plain, w.plain.data(), w.plain.size() * sizeof(int4),
                       hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(w.d_batched, w.batched.data(), w.batched.size() * sizeof(int4),
                       hipMemcpyHostToDevice));
    }
  }
  return YSX_OK;
}

}  // namespace

// ----------------------------------------------- allocation-free model table
// Device scratch of ysx_model_table_noalloc: the staged deck, the layer
// lists and the spline/gravity work arrays of k_model.  Every GPU holds its
// own copy at its own address until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returns for that device
// (scratch_of), never through the array's name.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)2 << 20;  // 2 MiB
constexpr int NOALLOC_MAX_LAYERS = 64;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// the address of `scratch` on `device`, looked up once per device with that
// device current (the caller has made it current)
hipError_t scratch_of(int device, double** ptr) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *ptr = g_scratch_ptr[device];
  return hipSuccess;
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
  for (auto& g : ctx->gstream) HIPCHK(hipStreamCreateWithFlags(&g, hipStreamNonBlocking));
  for (auto& e : ctx->ev) HIPCHK(hipEventCreate(&e));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  ctx->lds_max = prop.sharedMemPerBlock;
  ctx->ncu = prop.multiProcessorCount;
  HIPCHK(scratch_of(device, &ctx->scratch));
  return YSX_OK;
}

extern "C" void ysx_destroy(ysx_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  free_run(ctx);
  dfree(ctx->d_tab);
  dfree(ctx->d_mdeck);
  dfree(ctx->d_mlay);
  dfree(ctx->d_mvjp);
  if (ctx->pin) (void)hipHostFree(ctx->pin);
  if (ctx->opin) (void)hipHostFree(ctx->opin);
  for (auto& e : ctx->pev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  for (auto& g : ctx->gstream)
    if (g) (void)hipStreamDestroy(g);
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
  dfree(ctx->d_mdeck);
  dfree(ctx->d_mlay);
  dfree(ctx->d_mvjp);
  ctx->have_model = false;
  std::string e = build_model(ctx->N, nknot, nic, noc, ians, tref_period, deck, ctx->M);
  if (!e.empty()) return set_err(ctx, YSX_ERR_INPUT, e);
  ctx->deck.assign(deck, deck + 9 * (size_t)nknot);
  ctx->ians = ians;
  ctx->tref_period = tref_period;
  ctx->lds_bytes = sizeof(double) * F_COUNT * (size_t)nknot;
  // k_source and the radial kernels stage the table in LDS when it fits
  // (with room for their own 2.3 KB); larger decks are read through the cache
  ctx->tab_lds = ctx->lds_bytes + 4096 <= ctx->lds_max;
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
    // the staged deck and layer lists stay for k_model_vjp
    auto keep = [&]() {
      ctx->d_mdeck = d_deck;
      ctx->d_mlay = d_lay;
      d_deck = nullptr;
      d_lay = nullptr;
    };
    hipError_t e1 = hipMalloc(&d_deck, 9 * (size_t)nknot * sizeof(double));
    if (e1 == hipSuccess) e1 = hipMalloc(&d_work, (nwork + 3 * (size_t)nknot) * sizeof(double));
    if (e1 == hipSuccess) e1 = hipMalloc(&d_lay, 2 * (size_t)nl * sizeof(int));
    if (e1 == hipSuccess)
      e1 = hipMemcpy(d_deck, deck, 9 * (size_t)nknot * sizeof(double), hipMemcpyHostToDevice);
    if (e1 == hipSuccess) e1 = hipMemcpy(d_lay, lo.data(), nl * sizeof(int), hipMemcpyHostToDevice);
    if (e1 == hipSuccess)
      e1 = hipMemcpy(d_lay + nl, n.data(), nl * sizeof(int), hipMemcpyHostToDevice);
    // stream-ordered: the context streams do not wait for the null stream
    if (e1 == hipSuccess) e1 = hipMemsetAsync(ctx->d_tab, 0, ctx->lds_bytes, ctx->stream);
    if (e1 == hipSuccess) {
      launch_model(dim3(1), dim3(256), 0, ctx->stream, d_deck, nknot, M.ifanis,
                         ctx->N.vel_norm, d_lay, d_lay + nl, nl, ctx->d_tab, d_work, ws,
                         d_work + nwork, d_work + nwork + nknot, d_work + nwork + 2 * (size_t)nknot);
      e1 = hipGetLastError();
    }
    if (e1 == hipSuccess) e1 = hipStreamSynchronize(ctx->stream);
    if (e1 == hipSuccess) keep();
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

extern "C" int ysx_model_table_noalloc(ysx_ctx* ctx, int nknot, int nic, int noc, int ians,
                                       double tref_period, const double* deck,
                                       double* d_table) {
  if (!ctx || !deck || !d_table) return YSX_ERR_ARG;
  if (!ctx->scratch) return set_err(ctx, YSX_ERR_ARG, "no device scratch for this context");
  if (nknot < 3 || nic < 1 || noc <= nic || noc >= nknot)
    return set_err(ctx, YSX_ERR_INPUT, "bad deck header");
  // layers split at duplicated radii (src/module_model.f90:294-308)
  int lay[2 * NOALLOC_MAX_LAYERS];
  int nl = 0, lo = 0, nmax = 0;
  for (int i = 1; i <= nknot; i++) {
    // compared after normalisation, as build_model compares them
    if (i == nknot || deck[9 * (size_t)i] / R_NORM_M == deck[9 * (size_t)(i - 1)] / R_NORM_M) {
      if (nl == NOALLOC_MAX_LAYERS)
        return set_err(ctx, YSX_ERR_ARG,
                       "ysx_model_table_noalloc: more than 64 layers in the deck");
      lay[nl] = lo;
      lay[NOALLOC_MAX_LAYERS + nl] = i - lo;
      nmax = std::max(nmax, i - lo);
      nl++;
      lo = i;
    }
  }
  // scratch layout (doubles): deck [9 nknot] | work [13 nl ws] | intg, qint,
  // gi [nknot each] | layer_lo, layer_n [nl ints each]
  const size_t ws = 3 * ((size_t)nmax + 1);
  const size_t o_work = 9 * (size_t)nknot, o_intg = o_work + 13 * (size_t)nl * ws,
               o_lay = o_intg + 3 * (size_t)nknot;
  const size_t need = (o_lay * sizeof(double) + 2 * (size_t)nl * sizeof(int));
  if (need > NOALLOC_SCRATCH_BYTES)
    return set_err(ctx, YSX_ERR_ARG,
                   "ysx_model_table_noalloc: this deck needs " + std::to_string(need) +
                       " bytes of device scratch, more than its fixed " +
                       std::to_string(NOALLOC_SCRATCH_BYTES) + " (use ysx_set_model)");
  HIPCHK(hipSetDevice(ctx->device));
  double* sc = ctx->scratch;
  int* slay = reinterpret_cast<int*>(sc + o_lay);
  HIPCHK(hipMemcpyAsync(sc, deck, 9 * (size_t)nknot * sizeof(double), hipMemcpyHostToDevice,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(slay, lay, nl * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(slay + nl, lay + NOALLOC_MAX_LAYERS, nl * sizeof(int),
                        hipMemcpyHostToDevice, ctx->stream));
  // the fields k_model does not write are zero, as in ysx_set_model
  const size_t ntab = (size_t)F_COUNT * nknot;
  launch_fill0(dim3((unsigned)std::min<size_t>(1024, (ntab + 255) / 256)), dim3(256), 0,
               ctx->stream, d_table, ntab);
  HIPCHK(hipGetLastError());
  launch_model(dim3(1), dim3(256), 0, ctx->stream, sc, nknot, ians == 1, ctx->N.vel_norm, slay,
               slay + nl, nl, d_table, sc + o_work, ws, sc + o_intg, sc + o_intg + nknot,
               sc + o_intg + 2 * (size_t)nknot);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return YSX_OK;
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

extern "C" int ysx_model_table_vjp(int nknot, int nic, int noc, int ians, double tref_period,
                                   const double* deck, const double* table_bar,
                                   double* deck_bar) {
  if (!deck || !table_bar || !deck_bar) return YSX_ERR_ARG;
  if (!model_table_vjp(make_norms(), nknot, nic, noc, ians, tref_period, deck, table_bar,
                       deck_bar)
           .empty())
    return YSX_ERR_INPUT;
  return YSX_OK;
}

namespace {
// The work lists and run array of the events set: built on the host
// (build_work, which also marks the events that share marches in
// RunK::drow) and uploaded; replaces lists uploaded before.
int upload_work(ysx_ctx* ctx) {
  for (auto& w : ctx->W) {
    dfree(w.d_plain);
    dfree(w.d_batched);
    w = WorkList();
  }
  build_work(ctx);
  const size_t nev = ctx->rk.size();
  HIPCHK(hipMemcpy(ctx->d_runs, ctx->rk.data(), nev * sizeof(RunK), hipMemcpyHostToDevice));
  for (auto& w : ctx->W) {
    size_t n = std::max<size_t>(1, std::max(w.plain.size(), w.batched.size()));
    HIPCHK(hipMalloc(&w.d_plain, n * sizeof(int4)));
    HIPCHK(hipMalloc(&w.d_batched, n * sizeof(int4)));
    if (!w.plain.empty())
      HIPCHK(hipMemcpy(w.d_plain, w.plain.data(), w.plain.size() * sizeof(int4),
                       hipMemcpyHostToDevice));
    if (!w.batched.empty())
      HIPCHK(hipMemcpy(w.d_batched, w.batched.data(), w.batched.size() * sizeof(int4),
                       hipMemcpyHostToDevice));
  }
  return YSX_OK;
}

// lanes per radial (l = 0) solve: the context's setting, or by default 8
// lanes, one for runs with many frequencies (the radial kernel then runs
// hidden beside the spheroidal marches and one lane each keeps it to few
// compute units)
int rad_lanes_of(const ysx_ctx* ctx, int nw) {
  if (ctx->rad_lanes) return ctx->rad_lanes;
  return (nw <= 640) ? 8 : 1;
}
}  // namespace

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
    // self-gravitating runs go through the lockstep kernels, a fluid inner
    // core too when the receiver is not below the source (sph_solver_fic
    // then differs from sph_solver only in its start, handled by k_start
    // and the marches); gravity switches 0 / 1 through k_ngc_march.  With
    // a fluid inner core and the receiver below the source the reference
    // starts solid minors in the fluid inner core (src/module_solver.f90:
    // 1323-1331) and its program does not complete on such a run, so there
    // is no reference result to match: rejected.
    if (R.grav == 2 && M.fic && R.rr < R.rs && R.lmax >= 1) {
      ctx->R.clear();
      ctx->rk.clear();
      std::string m = "self-gravitation with a fluid inner core and the receiver below the "
                      "source (sph_solver_fic's receiver-below branch) is not supported";
      return set_err(ctx, YSX_ERR_UNSUPPORTED, nev > 1 ? "event " + std::to_string(e) + ": " + m : m);
    }
    k.ngc = (R.grav != 2) ? 1 : 0;
    k.lpad = (R.lmax + 1 + 3) / 4 * 4;
    k.radg = rad_lanes_of(ctx, R.nw);
    k.tablds = ctx->tab_lds ? 1 : 0;
    // b-vector: 2 receiver-side rows per degree (4 lanes, rows 0..2, with the
    // potential/tilt correction), or the 4 source-side columns
    k.brow = ctx->bcols ? 0 : 1;
    k.bl = (k.brow && !(R.out == 2 && R.cor != 0)) ? 32 : 16;
    k.drow = 0;  // set by build_work for the events that share marches
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
  HIPCHK(hipMemsetAsync(ctx->d_rec, 0, P * REC * sizeof(cd), ctx->stream));
  HIPCHK(hipMalloc(&ctx->d_starts, P * 2 * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_cnt, P * 5 * sizeof(int)));
  HIPCHK(hipMemsetAsync(ctx->d_cnt, 0, P * 5 * sizeof(int), ctx->stream));
  HIPCHK(hipMalloc(&ctx->d_bst, P * 56 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_mfin, P * 14 * sizeof(cd)));
  HIPCHK(hipMalloc(&ctx->d_ms1, P * 14 * sizeof(cd)));
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
  HIPCHK(hipMemsetAsync(ctx->d_cf, 0, cfo * sizeof(cd), ctx->stream));
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
    HIPCHK(hipMemsetAsync(ctx->d_tb, 0, tbo * sizeof(double), ctx->stream));
    HIPCHK(hipMemcpy(ctx->d_azep, az.data(), ro * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_eia, eia.data(), ro * sizeof(cd), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&d_th, th.size() * sizeof(double)));
    hipError_t et = hipMemcpy(d_th, th.data(), th.size() * sizeof(double), hipMemcpyHostToDevice);
    for (int e = 0; e < nev && et == hipSuccess; e++) {
      const RunK& k = ctx->rk[e];
      launch_legendre(dim3((k.lmax + 64) / 64, k.nr), dim3(64), 0, ctx->stream,
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
  if (int rc = upload_work(ctx)) return rc;
  if (ctx->n_union > 0) {
    HIPCHK(hipMalloc(&ctx->d_shorder, ctx->sh_order.size() * sizeof(int)));
    HIPCHK(hipMemcpy(ctx->d_shorder, ctx->sh_order.data(), ctx->sh_order.size() * sizeof(int),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&ctx->d_shx, (size_t)ctx->n_union * 14 * 64 * sizeof(cd)));
    HIPCHK(hipMalloc(&ctx->d_shs, (size_t)ctx->n_union * 64 * sizeof(int)));
  }
  launch_twiddles(dim3((ntmax + 255) / 256), dim3(256), 0, ctx->stream, ntmax,
                     ctx->d_tw);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->have_run = true;
  ctx->refined = false;
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
  size_t lds = ctx->tab_lds ? ctx->lds_bytes : 0;
  dim3 b256(256), b64(64);
  SphState S{ctx->d_bst, ctx->d_mfin, ctx->d_m2, ctx->d_fac, ctx->d_sc3, ctx->d_ms1};
  StartData SD{ctx->d_tor_i1, ctx->d_tor_y0, ctx->d_sph_i1, ctx->d_sph_m0};
  // a single event's minor march also stores its per-step states when
  // ysx_misfit_grad asks for them (prepare_step_store)
  const StepStore FS{ctx->d_fys, ctx->d_fyss, ctx->d_foff, ctx->d_fcap,
                     ctx->d_fsc, ctx->d_ferr, nullptr},
      FT{ctx->d_tys, ctx->d_tyss, ctx->d_foff, ctx->d_fcap, ctx->d_tsc, ctx->d_ferr, ctx->d_tyend};
  const bool store = ctx->fs_want && ctx->d_fys && e0 == 0 && e1 == 1 && ctx->rk.size() == 1;
  const StepStore* fsp = store ? &FS : nullptr;
  const StepStore* ftp = store ? &FT : nullptr;
  ctx->fs_valid = store;
  for (int e = e0; e < e1; e++) {
    PairRange PR = pairs_of(ctx, e, e + 1);
    int nb = (int)((PR.n + 255) / 256);
    Queue qt = plain_q(ctx, W_TOR, e), qb = plain_q(ctx, W_BVEC, e), qr = plain_q(ctx, W_RAD, e),
          ql = plain_q(ctx, W_LSUM, e);
    if (timing) HIPCHK(hipEventRecord(ev[0], st));
    launch_source(dim3(nb), b256, lds, st, ctx->d_tab, ctx->d_runs, PR,
                       ctx->d_rec, ctx->d_cnt);
    HIPCHK(hipEventRecord(ev[1], st));
    // the radial (l = 0) solves run beside the marches on a second stream.
    // An event too small to fill the GPU with its spheroidal chain (fewer
    // than four waves per SIMD) runs its toroidal chain (starts and march)
    // there too, ahead of the radial solves: none of them feeds the others
    // before the assembly, and the spheroidal chain is the longer one
    const bool side = qt.n < 16 * ctx->ncu;
    ctx->tor_side = side;
    HIPCHK(hipStreamWaitEvent(st2, ev[1], 0));
    if (side) {
      if (timing) HIPCHK(hipEventRecord(ev[9], st2));
      launch_start(dim3(nb), b256, lds, st2, ctx->d_tab, ctx->d_runs, PR, SD, ctx->d_starts, 1);
      launch_tor_march_item(dim3(grid_of(qt)), b64, 0, st2, ctx->d_tab, ctx->d_runs, qt,
                            SD, ctx->d_rec, ctx->d_cnt, ftp);
      if (timing) HIPCHK(hipEventRecord(ev[3], st2));
    }
    if (timing) HIPCHK(hipEventRecord(ev[10], st2));
    if (qr.n > 0) {
      if (ctx->rk[e].grav == 0)
        launch_rad_ng(dim3(grid_of(qr)), b64, lds, st2, ctx->d_tab, ctx->d_runs,
                           qr, ctx->d_rec, ctx->d_cnt);
      else
        launch_rad(dim3(grid_of(qr)), b64, lds, st2, ctx->d_tab, ctx->d_runs, qr,
                           ctx->d_rec, ctx->d_cnt);
    }
    HIPCHK(hipEventRecord(ev[11], st2));
    launch_start(dim3(nb), b256, lds, st, ctx->d_tab, ctx->d_runs, PR, SD, ctx->d_starts,
                 side ? 2 : 3);
    if (timing) HIPCHK(hipEventRecord(ev[2], st));
    if (!side) {
      launch_tor_march_item(dim3(grid_of(qt)), b64, 0, st, ctx->d_tab, ctx->d_runs, qt,
                            SD, ctx->d_rec, ctx->d_cnt, ftp);
      if (timing) HIPCHK(hipEventRecord(ev[3], st));
    }
    if (ctx->rk[e].ngc) {
      launch_ngc_march(ctx->rk[e].grav == 1, dim3(grid_of(qt)), b64, 0, st, ctx->d_tab,
                       ctx->d_runs, qt, ctx->d_rec, ctx->d_cnt);
      if (timing) HIPCHK(hipEventRecord(ev[4], st));
    } else {
      launch_minor_top_item(dim3(grid_of(qt)), b64, 0, st, ctx->d_tab,
                         ctx->d_runs, qt, S, ctx->d_cnt);
      launch_minor_march_item(dim3(grid_of(qt)), b64, 0, st, ctx->d_tab,
                         ctx->d_runs, qt, SD, S, ctx->d_cnt, 0, fsp);
      if (timing) HIPCHK(hipEventRecord(ev[4], st));
      if (ctx->rk[0].brow)
        launch_brow_march_item(dim3(grid_of(qb)), b64, 0, st, ctx->d_tab, ctx->d_runs, qb, S,
                               ShareK{}, ctx->d_rec, ctx->d_cnt);
      else
        launch_bvec_march_item(dim3(grid_of(qb)), b64, 0, st, ctx->d_tab, ctx->d_runs, qb, S,
                               ctx->d_rec, ctx->d_cnt);
    }
    if (timing) HIPCHK(hipEventRecord(ev[5], st));
    HIPCHK(hipStreamWaitEvent(st, ev[11], 0));
    if (timing) HIPCHK(hipEventRecord(ev[6], st));
    launch_assemble(dim3(asm_blocks(PR)), b64, 0, st, ctx->d_runs, PR, ctx->d_rec,
                       ctx->d_cf, ctx->d_tb);
    if (timing) HIPCHK(hipEventRecord(ev[7], st));
    launch_lsum(dim3(grid_of(ql)), dim3(256), 0, st, ctx->d_runs, ql, ctx->d_cf,
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
    launch_seis(g(qs), dim3(1024), 0, st, ctx->d_runs, qs, ctx->d_spec, ctx->d_tw,
                       ctx->d_work, ctx->d_azep, ctx->d_seis);
  if (qa.n > 0)
    launch_fft_lds(g(qa), dim3(512), 0, st, ctx->d_runs, qa, ctx->d_spec,
                       ctx->d_tw, ctx->d_work);
  if (qb.n > 0)
    launch_fft_post(g(qb), dim3(256), 0, st, ctx->d_runs, qb, ctx->d_tw,
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
  size_t lds = ctx->tab_lds ? ctx->lds_bytes : 0;
  dim3 b256(256), b64(64);
  int nev = (int)ctx->rk.size();
  SphState S{ctx->d_bst, ctx->d_mfin, ctx->d_m2, ctx->d_fac, ctx->d_sc3, ctx->d_ms1};
  StartData SD{ctx->d_tor_i1, ctx->d_tor_y0, ctx->d_sph_i1, ctx->d_sph_m0};
  ShareK SH{ctx->d_shorder, ctx->n_union > 0 ? (int)ctx->sh_order.size() : 0, ctx->sh_lmin,
            ctx->sh_lmax, ctx->sh_nw, ctx->d_shx, ctx->d_shs};
  PairRange PR = pairs_of(ctx, 0, nev);
  int nb = (int)((PR.n + 255) / 256);
  bool pe = ctx->persist != 0;
  int pers = ctx->blocks_per_cu * ctx->ncu;
  auto grid = [&](const Queue& q) {
    return dim3(std::max(1, pe ? std::min(q.n, pers) : q.n));
  };
  Queue qt = batched_q(ctx, W_TOR, Q_TOR), qtop = batched_q(ctx, W_TOR, Q_TOP),
        qmin = batched_q(ctx, W_TOR, Q_MINOR), qb = batched_q(ctx, W_BVEC, Q_BVEC),
        qd = batched_q(ctx, W_DROW, Q_DROW),
        qr = batched_q(ctx, W_RAD, -1), ql = batched_q(ctx, W_LSUM, -1),
        qs = batched_q(ctx, W_SEIS, -1);
  if (pe) launch_reset(dim3(1), dim3(64), 0, st, ctx->d_queue, (int)Q_N);
  launch_source(dim3(nb), b256, lds, st, ctx->d_tab, ctx->d_runs, PR, ctx->d_rec,
                     ctx->d_cnt);
  HIPCHK(hipEventRecord(ev[1], st));
  HIPCHK(hipStreamWaitEvent(st2, ev[1], 0));
  if (qr.n > 0) {
    launch_rad(dim3(qr.n), b64, lds, st2, ctx->d_tab, ctx->d_runs, qr,
                       ctx->d_rec, ctx->d_cnt);
    bool any_ng = false;
    for (const RunK& k : ctx->rk) any_ng = any_ng || (k.grav == 0 && k.lmin == 0);
    if (any_ng)
      launch_rad_ng(dim3(qr.n), b64, lds, st2, ctx->d_tab, ctx->d_runs, qr,
                         ctx->d_rec, ctx->d_cnt);
  }
  HIPCHK(hipEventRecord(ev[11], st2));
  launch_start(dim3(nb), b256, lds, st, ctx->d_tab, ctx->d_runs, PR, SD,
                     ctx->d_starts);
  if (pe) {
    launch_tor_march_pers(grid(qt), b64, 0, st, ctx->d_tab, ctx->d_runs, qt, SD,
                       ctx->d_rec, ctx->d_cnt);
    launch_minor_top_pers(grid(qtop), b64, 0, st, ctx->d_tab, ctx->d_runs, qtop,
                       S, ctx->d_cnt);
    launch_minor_march_pers(grid(qmin), b64, 0, st, ctx->d_tab, ctx->d_runs,
                       qmin, SD, S, ctx->d_cnt, 1);
    if (qd.n > 0)
      launch_drow_march_pers(grid(qd), b64, 0, st, ctx->d_tab, ctx->d_runs, qd, S, SH,
                             ctx->d_cnt);
    if (ctx->rk[0].brow)
      launch_brow_march_pers(grid(qb), b64, 0, st, ctx->d_tab, ctx->d_runs, qb, S, SH,
                             ctx->d_rec, ctx->d_cnt);
    else
      launch_bvec_march_pers(grid(qb), b64, 0, st, ctx->d_tab, ctx->d_runs, qb, S,
                             ctx->d_rec, ctx->d_cnt);
  } else {
    launch_tor_march_item(grid(qt), b64, 0, st, ctx->d_tab, ctx->d_runs, qt,
                       SD, ctx->d_rec, ctx->d_cnt);
    launch_minor_top_item(grid(qt), b64, 0, st, ctx->d_tab, ctx->d_runs, qt, S,
                       ctx->d_cnt);
    launch_minor_march_item(grid(qt), b64, 0, st, ctx->d_tab, ctx->d_runs, qt,
                       SD, S, ctx->d_cnt, 1);
    if (qd.n > 0)
      launch_drow_march_item(grid(qd), b64, 0, st, ctx->d_tab, ctx->d_runs, qd, S, SH,
                             ctx->d_cnt);
    if (ctx->rk[0].brow)
      launch_brow_march_item(grid(qb), b64, 0, st, ctx->d_tab, ctx->d_runs, qb, S, SH,
                             ctx->d_rec, ctx->d_cnt);
    else
      launch_bvec_march_item(grid(qb), b64, 0, st, ctx->d_tab, ctx->d_runs, qb, S,
                             ctx->d_rec, ctx->d_cnt);
  }
  // gravity switches 0 / 1: one launch each over the cost-sorted wave list
  // (the kernel skips the other events' items)
  Queue qn = batched_q(ctx, W_TOR, -1);
  for (int sw = 0; sw <= 1; sw++) {
    bool any = false;
    for (const RunK& k : ctx->rk) any = any || (k.ngc && k.grav == sw);
    if (any)
      launch_ngc_march(sw == 1, dim3(std::max(1, qn.n)), b64, 0, st, ctx->d_tab, ctx->d_runs,
                       qn, ctx->d_rec, ctx->d_cnt);
  }
  HIPCHK(hipStreamWaitEvent(st, ev[11], 0));
  launch_assemble(dim3(asm_blocks(PR)), b64, 0, st, ctx->d_runs, PR, ctx->d_rec, ctx->d_cf,
                     ctx->d_tb);
  launch_lsum(dim3(std::max(1, ql.n)), dim3(256), 0, st, ctx->d_runs, ql, ctx->d_cf,
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
  ctx->ms[0] = a + b;  // source vectors + start levels (the spheroidal ones
                       // alone when the toroidal chain ran beside)
  if ((rc = ctx->tor_side ? elapsed(ctx, 9, 3, ctx->ms[1]) : elapsed(ctx, 2, 3, ctx->ms[1])) ||
      (rc = elapsed(ctx, ctx->tor_side ? 2 : 3, 4, ctx->ms[2])) ||
      (rc = elapsed(ctx, 4, 5, ctx->ms[3])) || (rc = elapsed(ctx, 10, 11, ctx->ms[4])) ||
      (rc = elapsed(ctx, 6, 7, ctx->ms[5])) || (rc = elapsed(ctx, 7, 8, ctx->ms[6])))
    return rc;
  ctx->have_spec = true;
  ctx->have_seis = false;
  ctx->have_recb = false;  // record adjoints belong to the records of a forward
  return refine_work(ctx);
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

namespace {
// Host copies of large arrays through pinned staging.  hipMemcpy to or
// from pageable memory stages through the runtime's own small buffers at a
// few GB/s; here the array moves in chunks of PIN_CHUNK bytes between the
// device and a pinned buffer of two chunks (hipHostMalloc, kept by the
// context) by DMA on the context's stream, while up to PIN_THREADS host
// threads copy the previous chunk between the pinned buffer and the
// caller's memory.  Small arrays take hipMemcpy.
constexpr size_t PIN_CHUNK = 32u << 20, PIN_SMALL = 4u << 20;
constexpr unsigned PIN_THREADS = 8;
// observed seismograms of ysx_misfit_grad staged in pinned memory during its forward
constexpr size_t OBS_PIN = 64u << 20;

void parallel_copy(void* dst, const void* src, size_t bytes) {
  const unsigned nt = std::max(1u, std::min(PIN_THREADS, std::thread::hardware_concurrency()));
  const size_t per = (bytes + nt - 1) / nt;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++) {
    const size_t a = t * per, b = std::min(bytes, a + per);
    if (a >= b) break;
    th.emplace_back([=]() {
      std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
    });
  }
  for (auto& t : th) t.join();
}

int pin_setup(ysx_ctx* ctx) {
  if (!ctx->pin) {
    HIPCHK(hipHostMalloc(&ctx->pin, 2 * PIN_CHUNK, hipHostMallocDefault));
    ctx->pin_bytes = 2 * PIN_CHUNK;
  }
  for (auto& e : ctx->pev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return YSX_OK;
}

int copy_d2h(ysx_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes < PIN_SMALL) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return YSX_OK;
  }
  if (int rc = pin_setup(ctx)) return rc;
  char* pin = static_cast<char*>(ctx->pin);
  const char* s = static_cast<const char*>(src);
  const size_t nch = (bytes + PIN_CHUNK - 1) / PIN_CHUNK;
  auto len = [&](size_t i) { return std::min(PIN_CHUNK, bytes - i * PIN_CHUNK); };
  HIPCHK(hipMemcpyAsync(pin, s, len(0), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipEventRecord(ctx->pev[0], ctx->stream));
  for (size_t i = 0; i < nch; i++) {
    const int h = (int)(i % 2);
    if (i + 1 < nch) {  // the other half's previous chunk was copied out last round
      HIPCHK(hipMemcpyAsync(pin + (1 - h) * PIN_CHUNK, s + (i + 1) * PIN_CHUNK, len(i + 1),
                            hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipEventRecord(ctx->pev[1 - h], ctx->stream));
    }
    HIPCHK(hipEventSynchronize(ctx->pev[h]));
    parallel_copy(static_cast<char*>(dst) + i * PIN_CHUNK, pin + h * PIN_CHUNK, len(i));
  }
  return YSX_OK;
}

int copy_h2d(ysx_ctx* ctx, void* dst, const void* src, size_t bytes) {
  hipStream_t st = ctx->stream;
  if (bytes < PIN_SMALL) {
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return YSX_OK;
  }
  if (int rc = pin_setup(ctx)) return rc;
  char* pin = static_cast<char*>(ctx->pin);
  char* d = static_cast<char*>(dst);
  const size_t nch = (bytes + PIN_CHUNK - 1) / PIN_CHUNK;
  auto len = [&](size_t i) { return std::min(PIN_CHUNK, bytes - i * PIN_CHUNK); };
  for (size_t i = 0; i < nch; i++) {
    const int h = (int)(i % 2);
    if (i >= 2) HIPCHK(hipEventSynchronize(ctx->pev[h]));  // this half's DMA is done
    parallel_copy(pin + h * PIN_CHUNK, static_cast<const char*>(src) + i * PIN_CHUNK, len(i));
    HIPCHK(hipMemcpyAsync(d + i * PIN_CHUNK, pin + h * PIN_CHUNK, len(i), hipMemcpyHostToDevice,
                          st));
    HIPCHK(hipEventRecord(ctx->pev[h], st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return YSX_OK;
}
}  // namespace

extern "C" int ysx_seismograms(ysx_ctx* ctx, double* seis) {
  if (!ctx) return YSX_ERR_ARG;
  if (!ctx->have_spec) return set_err(ctx, YSX_ERR_ARG, "no spectra");
  if (ctx->rk.size() != 1) return set_err(ctx, YSX_ERR_ARG, "several events: use ysx_forward_events");
  HIPCHK(hipSetDevice(ctx->device));
  int rc = run_seis(ctx);
  if (rc) return rc;
  if (seis)
    return copy_d2h(ctx, seis, ctx->d_seis,
                    3 * (size_t)ctx->R[0].nt * ctx->R[0].nr * sizeof(double));
  return YSX_OK;
}

extern "C" int ysx_forward(ysx_ctx* ctx, double* seis) {
  int rc = ysx_spectra(ctx, nullptr, nullptr, nullptr);
  if (rc) return rc;
  return ysx_seismograms(ctx, seis);
}

// takes effect at the next ysx_set_run / ysx_set_events (the b-vector
// work lists are built for one of the two kernels)
extern "C" int ysx_set_records(ysx_ctx* ctx, int full) {
  if (!ctx || (full != 0 && full != 1)) return YSX_ERR_ARG;
  ctx->bcols = (full == 1);
  return YSX_OK;
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

extern "C" int ysx_set_lanes(ysx_ctx* ctx, int rad_lanes) {
  if (!ctx || (rad_lanes != 0 && rad_lanes != 1 && rad_lanes != 8)) return YSX_ERR_ARG;
  if (rad_lanes == ctx->rad_lanes) return YSX_OK;
  ctx->rad_lanes = rad_lanes;
  if (!ctx->have_run) return YSX_OK;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream2));
  for (size_t e = 0; e < ctx->rk.size(); e++) ctx->rk[e].radg = rad_lanes_of(ctx, ctx->R[e].nw);
  drop_graph(ctx);
  if (int rc = upload_work(ctx)) return rc;
  ctx->refined = false;  // the wave lists are re-sorted by measured cost again
  return YSX_OK;
}

extern "C" int ysx_forward_events(ysx_ctx* ctx, int mode, double* seis) {
  if (!ctx || (mode != 0 && mode != 1)) return YSX_ERR_ARG;
  if (!ctx->have_run) return set_err(ctx, YSX_ERR_ARG, "no run");
  HIPCHK(hipSetDevice(ctx->device));
  int nev = (int)ctx->rk.size();
  hipStream_t st = ctx->stream;
  ctx->fs_valid = false;  // the batched forward stores no per-step states
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
  ctx->have_recb = false;  // record adjoints belong to the records of a forward
  if (int rc2 = refine_work(ctx)) return rc2;
  if (seis) {
    size_t n = 0;
    for (const HostRun& R : ctx->R) n += 3 * (size_t)R.nt * R.nr;
    return copy_d2h(ctx, seis, ctx->d_seis, n * sizeof(double));
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
  return copy_d2h(ctx, seis, ctx->d_seis + ctx->rk[e].seis_off, n * sizeof(double));
}

extern "C" int ysx_pair_records(ysx_ctx* ctx, int n, const int* l, const int* j, double* rec) {
  if (!ctx || (n > 0 && (!l || !j || !rec))) return YSX_ERR_ARG;
  if (!ctx->have_spec) return set_err(ctx, YSX_ERR_ARG, "no spectra computed");
  if (ctx->rk.size() != 1) return set_err(ctx, YSX_ERR_ARG, "several events: one-event diagnostic");
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

namespace {

// per-event blocks of the source-adjoint kernels (256 pairs each)
int src_blocks(const RunK& k) { return (int)((k.npair + 255) / 256); }

int alloc_adjoint(ysx_ctx* ctx) {
  size_t ns = 0, nspec = 0, ncf = 0, nb = 0;
  for (size_t e = 0; e < ctx->rk.size(); e++) {
    const HostRun& R = ctx->R[e];
    const RunK& k = ctx->rk[e];
    ns += 3 * (size_t)R.nt * R.nr;
    nspec += 3 * (size_t)R.nw * R.nr;
    ncf += 15 * (size_t)k.lpad * k.jpad;
    nb += (size_t)src_blocks(k);
  }
  if (!ctx->d_seisb) HIPCHK(hipMalloc(&ctx->d_seisb, ns * sizeof(double)));
  if (!ctx->d_specb) HIPCHK(hipMalloc(&ctx->d_specb, nspec * sizeof(cd)));
  if (!ctx->d_cfb) HIPCHK(hipMalloc(&ctx->d_cfb, ncf * sizeof(cd)));
  if (!ctx->d_recb) HIPCHK(hipMalloc(&ctx->d_recb, (size_t)ctx->npair * REC * sizeof(cd)));
  if (!ctx->d_mmpart) HIPCHK(hipMalloc(&ctx->d_mmpart, 6 * nb * sizeof(double)));
  return YSX_OK;
}

int adjoint_run(ysx_ctx* ctx) {
  if (!ctx->have_run) return set_err(ctx, YSX_ERR_ARG, "no run");
  HIPCHK(hipSetDevice(ctx->device));
  return alloc_adjoint(ctx);
}

// The per-block sums of k_source_T / k_source_model_T of every event, added
// per event in block order: out[6 e + i].
// the per-block partial sums of k_source_T / k_source_model_T of every
// event into d_mmpart on stream st, between events ev[a] and ev[b]
int enqueue_source_sums(ysx_ctx* ctx, bool model, hipStream_t st, int a, int b) {
  const int nev = (int)ctx->rk.size();
  size_t b0 = 0;
  HIPCHK(hipEventRecord(ctx->ev[a], st));
  for (int e = 0; e < nev; e++) {
    double* part = ctx->d_mmpart + 6 * b0;
    dim3 g(src_blocks(ctx->rk[e]));
    if (model)
      launch_source_model_T(g, dim3(256), 0, st, ctx->d_tab, ctx->d_runs, e, ctx->d_recb, part);
    else
      launch_source_T(g, dim3(256), 0, st, ctx->d_tab, ctx->d_runs, e, ctx->d_recb, part);
    b0 += src_blocks(ctx->rk[e]);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev[b], st));
  return YSX_OK;
}

// their sums per event (6 each), once the launches are complete
int collect_source_sums(ysx_ctx* ctx, std::vector<double>& out) {
  const int nev = (int)ctx->rk.size();
  std::vector<size_t> b0(nev + 1, 0);
  for (int e = 0; e < nev; e++) b0[e + 1] = b0[e] + src_blocks(ctx->rk[e]);
  std::vector<double> part(6 * b0[nev]);
  HIPCHK(hipMemcpy(part.data(), ctx->d_mmpart, part.size() * sizeof(double), hipMemcpyDeviceToHost));
  out.assign(6 * (size_t)nev, 0.0);
  for (int e = 0; e < nev; e++)
    for (int i = 0; i < 6; i++)
      for (size_t b = b0[e]; b < b0[e + 1]; b++) out[6 * (size_t)e + i] += part[6 * b + i];
  return YSX_OK;
}

int source_sums(ysx_ctx* ctx, bool model, std::vector<double>& out) {
  if (int rc = enqueue_source_sums(ctx, model, ctx->stream, 0, 1)) return rc;
  HIPCHK(hipEventSynchronize(ctx->ev[1]));
  if (int rc = elapsed(ctx, 0, 1, ctx->ms_vjp[2])) return rc;
  return collect_source_sums(ctx, out);
}

// the source-vector term of the table adjoint from the per-event sums G
// (source_model_T's six moduli fields at each event's source radius, on its
// knot interval's spline weights, as spl_w computes them), into table_bar
void source_model_fold(const ysx_ctx* ctx, const std::vector<double>& G, double* table_bar) {
  const int nk = ctx->M.nknot;
  const int fields[6] = {F_LCON, F_CCON, F_FCON, F_QMU, F_XLAM, F_XA2};
  for (size_t e = 0; e < ctx->rk.size(); e++) {
    const RunK& k = ctx->rk[e];
    const double r0 = ctx->M.at(F_R, k.is), r1 = ctx->M.at(F_R, k.is + 1);
    const double h = r1 - r0;
    const bool deg = (h == 0.0);
    const double ih = deg ? 0.0 : 1.0 / h, a = (r1 - k.rs) * ih, bw = (k.rs - r0) * ih;
    const double hh = h * h / 6.0, a3 = (a * a * a - a) * hh, b3 = (bw * bw * bw - bw) * hh;
    for (int i = 0; i < 6; i++) {
      const double g = G[6 * e + i];
      double* y = table_bar + (size_t)fields[i] * nk + (k.is - 1);
      double* y2 = table_bar + (size_t)(fields[i] + 1) * nk + (k.is - 1);
      if (deg) {
        y[0] += g;
        continue;
      }
      y[0] += a * g;
      y[1] += bw * g;
      y2[0] += a3 * g;
      y2[1] += b3 * g;
    }
  }
}

int seis_vjp_device(ysx_ctx* ctx);

}  // namespace

extern "C" int ysx_seismograms_vjp(ysx_ctx* ctx, const double* seis_bar, double* ur_bar,
                                   double* ut_bar, double* up_bar) {
  if (!ctx || !seis_bar) return YSX_ERR_ARG;
  if (int rc = adjoint_run(ctx)) return rc;
  const int nev = (int)ctx->rk.size();
  if (nev > 1 && (ur_bar || ut_bar || up_bar))
    return set_err(ctx, YSX_ERR_ARG, "several events: the spectra adjoints stay on the device");
  size_t ns = 0;
  for (const HostRun& R : ctx->R) ns += 3 * (size_t)R.nt * R.nr;
  if (int rc = copy_h2d(ctx, ctx->d_seisb, seis_bar, ns * sizeof(double))) return rc;
  if (int rc = seis_vjp_device(ctx)) return rc;
  size_t n = (size_t)ctx->R[0].nw * ctx->R[0].nr;
  double* outs[3] = {ur_bar, ut_bar, up_bar};
  for (int c = 0; c < 3; c++)
    if (outs[c])
      HIPCHK(hipMemcpy(outs[c], ctx->d_specb + c * n, n * sizeof(cd), hipMemcpyDeviceToHost));
  return YSX_OK;
}

namespace {

// the seismograms' adjoint from d_seisb to the spectra adjoints d_specb
int seis_vjp_device(ysx_ctx* ctx) {
  const int nev = (int)ctx->rk.size();
  hipStream_t st = ctx->stream;
  HIPCHK(hipEventRecord(ctx->ev[0], st));
  for (int e = 0; e < nev; e++) {
    const HostRun& R = ctx->R[e];
    const int bk = std::min(R.nt, FFT_B), nblk = R.nt / bk;
    if (nblk <= FFT_G) {  // three passes, the transpose of k_fft_lds + k_fft_post
      launch_post_T(dim3(R.nr * ((bk + 255) / 256)), dim3(256), 0, st, ctx->d_runs, e,
                    ctx->d_seisb, ctx->d_tw, ctx->d_azep, ctx->d_work);
      launch_fft_lds_T(dim3(R.nr * 3 * nblk), dim3(512), 0, st, ctx->d_runs, e, ctx->d_tw,
                       ctx->d_work);
      launch_pack_T(dim3(R.nr * 3 * ((R.nw + 255) / 256)), dim3(256), 0, st, ctx->d_runs, e,
                    ctx->d_work, ctx->d_specb);
    } else {
      launch_seis_T(dim3(std::max(1, R.nr)), dim3(1024), 0, st, ctx->d_runs, e, ctx->d_seisb,
                    ctx->d_tw, ctx->d_work, ctx->d_azep, ctx->d_specb);
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev[1], st));
  HIPCHK(hipEventSynchronize(ctx->ev[1]));
  if (int rc = elapsed(ctx, 0, 1, ctx->ms_vjp[0])) return rc;
  ctx->have_specb = true;
  return YSX_OK;
}

}  // namespace

extern "C" int ysx_spectra_vjp(ysx_ctx* ctx, const double* ur_bar, const double* ut_bar,
                               const double* up_bar, int n, const int* l, const int* j,
                               double* rec_bar) {
  if (!ctx || n < 0 || (n > 0 && (!l || !j || !rec_bar))) return YSX_ERR_ARG;
  const bool given = ur_bar || ut_bar || up_bar;
  if (given && !(ur_bar && ut_bar && up_bar))
    return set_err(ctx, YSX_ERR_ARG, "ur_bar, ut_bar, up_bar: give all three or none");
  if (int rc = adjoint_run(ctx)) return rc;
  if (!ctx->have_spec) return set_err(ctx, YSX_ERR_ARG, "no spectra computed");
  const int nev = (int)ctx->rk.size();
  if (nev > 1 && (given || n > 0))
    return set_err(ctx, YSX_ERR_ARG,
                   "several events: spectra adjoints from the device, no per-pair copies");
  const RunK& k0 = ctx->rk[0];
  for (int i = 0; i < n; i++)
    if (l[i] < k0.lmin || l[i] > k0.lmax || j[i] < 1 || j[i] > k0.nw)
      return set_err(ctx, YSX_ERR_ARG, "pair outside the run");
  if (given) {
    size_t ns = (size_t)ctx->R[0].nw * ctx->R[0].nr;
    const double* ins[3] = {ur_bar, ut_bar, up_bar};
    for (int c = 0; c < 3; c++)
      HIPCHK(hipMemcpy(ctx->d_specb + c * ns, ins[c], ns * sizeof(cd), hipMemcpyHostToDevice));
    ctx->have_specb = true;
  } else if (!ctx->have_specb) {
    return set_err(ctx, YSX_ERR_ARG, "no spectra adjoint: call ysx_seismograms_vjp first");
  }
  hipStream_t st = ctx->stream;
  HIPCHK(hipEventRecord(ctx->ev[0], st));
  for (int e = 0; e < nev; e++) {
    const RunK& k = ctx->rk[e];
    int ntile = (k.lpad + 15) / 16 * (k.jpad / 16);
    launch_lsum_T(dim3(ntile), dim3(256), 0, st, ctx->d_runs, e, ctx->d_tb, ctx->d_eia,
                  ctx->d_specb, ctx->d_cfb);
    launch_assemble_T(dim3((unsigned)((k.npair + ASMT_P - 1) / ASMT_P)), dim3(64), 0, st,
                      ctx->d_runs, e, ctx->d_rec, ctx->d_cfb, ctx->d_recb);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ctx->ev[1], st));
  HIPCHK(hipEventSynchronize(ctx->ev[1]));
  if (int rc = elapsed(ctx, 0, 1, ctx->ms_vjp[1])) return rc;
  ctx->have_recb = true;
  for (int i = 0; i < n; i++) {
    size_t p = (size_t)(l[i] - k0.lmin) * k0.nw + (j[i] - 1);
    HIPCHK(hipMemcpy(rec_bar + (size_t)i * 2 * REC, ctx->d_recb + p * REC, REC * sizeof(cd),
                     hipMemcpyDeviceToHost));
  }
  return YSX_OK;
}

extern "C" int ysx_moment_tensor_vjp(ysx_ctx* ctx, double* mm_bar) {
  if (!ctx || !mm_bar) return YSX_ERR_ARG;
  if (int rc = adjoint_run(ctx)) return rc;
  if (!ctx->have_recb) return set_err(ctx, YSX_ERR_ARG, "no record adjoints: call ysx_spectra_vjp first");
  std::vector<double> G;
  if (int rc = source_sums(ctx, false, G)) return rc;
  for (size_t i = 0; i < G.size(); i++)
    mm_bar[i] = G[i] / ctx->N.moment_norm;  // normalised -> parameter-file units
  return YSX_OK;
}

extern "C" int ysx_source_model_vjp(ysx_ctx* ctx, double* table_bar) {
  if (!ctx || !table_bar) return YSX_ERR_ARG;
  if (int rc = adjoint_run(ctx)) return rc;
  if (!ctx->have_recb) return set_err(ctx, YSX_ERR_ARG, "no record adjoints: call ysx_spectra_vjp first");
  std::vector<double> G;
  if (int rc = source_sums(ctx, true, G)) return rc;
  std::fill(table_bar, table_bar + (size_t)F_COUNT * ctx->M.nknot, 0.0);
  source_model_fold(ctx, G, table_bar);
  return YSX_OK;
}

namespace {

// A bound on the steps of one sweep item at Re w = rew (the step rule of
// yint, src/module_int.f90:1903-1905): every knot interval's full step list
// once, the longest twice (a march splits the source interval in two, each
// part taking at most the full list), and two more.
int item_steps_bound(const HostModel& M, double rew) {
  long long sum = 0;
  int mx = 1;
  for (int i = 1; i < M.nknot; i++) {
    double a = M.at(F_ALPHA, i), b = M.at(F_ALPHA, i + 1);
    double drr = 1.0 / (4.0 * rew / std::max(a, b));
    double dr = std::fabs(M.at(F_R, i + 1) - M.at(F_R, i));
    int n = 1 + (int)std::floor(dr / drr);
    sum += n;
    mx = std::max(mx, n);
  }
  return (int)std::min<long long>(sum + mx + 2, 1 << 30);
}

// the most steps of any item of event k (its largest frequency): the
// blocks' per-step storage of the adjoint sweeps
int max_item_steps(const HostModel& M, const RunK& k) {
  return item_steps_bound(M, k.wmin + (double)(k.nw - 1) * k.dw);
}

// Device memory for the minor march's per-step states of a single-event
// forward (StepStore), so that the minor sweep need not march again:
// item (j, lb) gets item_steps_bound at w_j.  Returns whether the forward
// should store them: one event in the solver adjoint's scope, and at most
// 32 GiB or half the free device memory.
bool prepare_step_store(ysx_ctx* ctx) {
  if (ctx->rk.size() != 1) return false;
  const RunK& k = ctx->rk[0];
  const HostModel& M = ctx->M;
  if (k.grav != 2 || k.ngc || M.fic || M.ocean || k.rr < k.rs || k.rr != M.at(F_R, M.nknot) ||
      !k.brow || (k.out == 2 && k.cor != 0) || M.nknot > 800)
    return false;
  const int nlb = (k.lmax - k.lmin + 1 + 63) / 64;
  std::vector<long long> off(k.nw);
  std::vector<int> cap(k.nw);
  long long steps = 0;
  for (int j = 0; j < k.nw; j++) {
    cap[j] = item_steps_bound(M, k.wmin + (double)j * k.dw);
    off[j] = steps;
    steps += (long long)nlb * cap[j];
  }
  // per step of an item: the minor march's 14-component states, the
  // minor sweep's rows of 14 and the toroidal march's 2
  const size_t per_step = 30 * 64 * sizeof(cd) + 3 * 64 * sizeof(int);
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return false;
  const size_t have = ctx->fs_steps * per_step;
  const size_t limit = std::min<size_t>(size_t(32) << 30, (freeb + have) / 2);
  if ((size_t)steps * per_step > limit) return false;
  if ((size_t)steps > ctx->fs_steps) {
    dfree(ctx->d_fys);
    dfree(ctx->d_fyss);
    dfree(ctx->d_tys);
    dfree(ctx->d_tyss);
    dfree(ctx->d_fxs);
    dfree(ctx->d_fxss);
    ctx->fs_steps = 0;
    if (hipMalloc(&ctx->d_fys, (size_t)steps * 14 * 64 * sizeof(cd)) != hipSuccess ||
        hipMalloc(&ctx->d_fyss, (size_t)steps * 64 * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->d_tys, (size_t)steps * 2 * 64 * sizeof(cd)) != hipSuccess ||
        hipMalloc(&ctx->d_tyss, (size_t)steps * 64 * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->d_fxs, (size_t)steps * 14 * 64 * sizeof(cd)) != hipSuccess ||
        hipMalloc(&ctx->d_fxss, (size_t)steps * 64 * sizeof(int)) != hipSuccess) {
      (void)hipGetLastError();
      dfree(ctx->d_fys);
      dfree(ctx->d_fyss);
      dfree(ctx->d_tys);
      dfree(ctx->d_tyss);
      dfree(ctx->d_fxs);
      dfree(ctx->d_fxss);
      return false;
    }
    ctx->fs_steps = (size_t)steps;
  }
  // one pairing unit per item and interval segment: at most two per knot
  // interval (the source's is split)
  const long long ucap = (long long)k.nw * nlb * (2 * (long long)M.nknot + 2);
  if (ucap > ctx->units_cap) {
    dfree(ctx->d_units);
    ctx->units_cap = 0;
    if (ucap > (1ll << 30) ||
        hipMalloc(&ctx->d_units, (size_t)ucap * sizeof(PairUnit)) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    ctx->units_cap = (int)ucap;
  }
  if (!ctx->d_pq && hipMalloc(&ctx->d_pq, 8 * sizeof(int)) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (!ctx->d_foff) {
    if (hipMalloc(&ctx->d_foff, k.nw * sizeof(long long)) != hipSuccess ||
        hipMalloc(&ctx->d_fcap, k.nw * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->d_fsc, (size_t)ctx->npair * 3 * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->d_tsc, (size_t)ctx->npair * 3 * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->d_tyend, (size_t)ctx->npair * 2 * sizeof(cd)) != hipSuccess ||
        hipMalloc(&ctx->d_ferr, sizeof(int)) != hipSuccess) {
      (void)hipGetLastError();
      dfree(ctx->d_foff);
      dfree(ctx->d_fcap);
      dfree(ctx->d_fsc);
      dfree(ctx->d_tsc);
      dfree(ctx->d_tyend);
      dfree(ctx->d_ferr);
      return false;
    }
    if (hipMemcpy(ctx->d_foff, off.data(), k.nw * sizeof(long long), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(ctx->d_fcap, cap.data(), k.nw * sizeof(int), hipMemcpyHostToDevice) !=
            hipSuccess)
      return false;
  }
  launch_reset(dim3(1), dim3(64), 0, ctx->stream, ctx->d_ferr, 1);
  return hipGetLastError() == hipSuccess;
}

}  // namespace

extern "C" int ysx_model_vjp(ysx_ctx* ctx, double* table_bar) {
  if (!ctx || !table_bar) return YSX_ERR_ARG;
  if (int rc = adjoint_run(ctx)) return rc;
  if (!ctx->have_recb)
    return set_err(ctx, YSX_ERR_ARG, "no record adjoints: call ysx_spectra_vjp first");
  const HostModel& M = ctx->M;
  const int nev = (int)ctx->rk.size();
  const double rsurf = M.at(F_R, M.nknot);
  for (int e = 0; e < nev; e++) {
    const RunK& k = ctx->rk[e];
    std::string why;
    if (k.grav != 2) why = "gravity switch 0 or 1";
    else if (M.fic) why = "a fluid inner core";
    else if (M.ocean) why = "an ocean";
    else if (k.rr < k.rs || k.rr != rsurf) why = "a receiver below the surface";
    else if (k.out == 2 && k.cor != 0) why = "the potential/tilt correction";
    else if (!k.brow) why = "the source-side b-vector columns";
    else if (M.nknot > 800) why = "more than 800 knots";
    if (!why.empty())
      return set_err(ctx, YSX_ERR_UNSUPPORTED,
                     "ysx_model_vjp: the solver adjoint is built for self-gravitation, a solid "
                     "inner core, no ocean and receivers at the surface; event " +
                         std::to_string(e) + " has " + why);
  }
  const int nk = M.nknot;
  const size_t P = (size_t)ctx->npair;
  // buffers (kept for the run)
  if (!ctx->d_tbar) {
    ctx->adj_blocks = 4 * ctx->ncu;
    ctx->adj_smax = 1;
    for (const RunK& k : ctx->rk) ctx->adj_smax = std::max(ctx->adj_smax, max_item_steps(M, k));
    const size_t nb = ctx->adj_blocks, sm = ctx->adj_smax;
    HIPCHK(hipMalloc(&ctx->d_mub, P * 14 * sizeof(cd)));
    HIPCHK(hipMalloc(&ctx->d_m0bar, P * 14 * sizeof(cd)));
    HIPCHK(hipMalloc(&ctx->d_y0bar, P * 2 * sizeof(cd)));
    HIPCHK(hipMalloc(&ctx->d_tbar, (size_t)F_COUNT * nk * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->d_tpart, (size_t)F_COUNT * nk * nb * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->d_tpart2, (size_t)F_COUNT * nk * nb * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->d_ays, nb * sm * 14 * 64 * sizeof(cd)));
    HIPCHK(hipMalloc(&ctx->d_ayss, nb * sm * 64 * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->d_ascr, nb * 13 * 14 * 64 * sizeof(cd)));
    HIPCHK(hipMalloc(&ctx->d_aerr, 16 * sizeof(int)));
    HIPCHK(hipMalloc(&ctx->d_aq, 4 * sizeof(int)));
  }
  // the minor sweep's item groups with the forward's states: the first
  // eighth of the cost-sorted items, the rest of the first half, the second
  // half (gcut: the ends of the first two); the groups after the first have
  // their own step scratch and partial table columns
  constexpr int ngroups = 3;
  const double gcut[ngroups - 1] = {0.125, 0.5};
  if (ctx->fs_valid)
    for (int g = 0; g + 1 < ngroups; g++)
      if (!ctx->d_gscr[g]) {
        const size_t nb = ctx->adj_blocks;
        HIPCHK(hipMalloc(&ctx->d_gscr[g], nb * 13 * 14 * 64 * sizeof(cd)));
        HIPCHK(hipMalloc(&ctx->d_gtpart[g], (size_t)F_COUNT * nk * nb * sizeof(double)));
      }
  hipStream_t st = ctx->stream;
  hipEvent_t* ev = ctx->ev;
  const dim3 b64(64), b256(256);
  const size_t ntab = (size_t)F_COUNT * nk;
  launch_fill0(dim3((unsigned)std::min<size_t>(1024, (ntab + 255) / 256)), b256, 0, st,
               ctx->d_tbar, ntab);
  launch_fill0(dim3(2048), b256, 0, st, ctx->d_tpart, ntab * ctx->adj_blocks);
  launch_fill0(dim3(2048), b256, 0, st, ctx->d_tpart2, ntab * ctx->adj_blocks);
  for (double* t : ctx->d_gtpart)
    if (t) launch_fill0(dim3(2048), b256, 0, st, t, ntab * ctx->adj_blocks);
  launch_reset(dim3(1), b64, 0, st, ctx->d_aerr, 16);
  launch_reset(dim3(1), b64, 0, st, ctx->d_aq, 4);
  HIPCHK(hipGetLastError());
  SphState S{ctx->d_bst, ctx->d_mfin, ctx->d_m2, ctx->d_fac, ctx->d_sc3, ctx->d_ms1};
  StartData SD{ctx->d_tor_i1, ctx->d_tor_y0, ctx->d_sph_i1, ctx->d_sph_m0};
  AdjWork AW{ctx->d_ays, ctx->d_ayss, 0, ctx->d_ascr, ctx->adj_smax, ctx->d_aerr, 14 * 64};
  // every event's items in one persistent launch per sweep: the plain
  // lists of all events are contiguous
  auto all_q = [&](int kind) {
    const WorkList& w = ctx->W[kind];
    return Queue{w.d_plain, w.start[nev] - w.start[0], nullptr, 1};
  };
  Queue qb = all_q(W_BVEC), qt = all_q(W_TOR);
  Queue qm = qt;
  qb.counter = ctx->d_aq;
  qm.counter = ctx->d_aq + 1;
  qt.counter = ctx->d_aq + 2;
  qb.chunk = qm.chunk = qt.chunk = 1;
  // the radial solves of the events from degree 0, by segments (sweep,
  // knot interval), one event after another on shared buffers, on stream rs
  auto radial = [&](hipStream_t rs) -> int {
    const Tab HT{M.tab.data(), nk};
    std::vector<int> rev;
    std::vector<RadSegs> segs;
    size_t nsmax = 0;
    for (int e = 0; e < nev; e++) {
      const RunK& k = ctx->rk[e];
      if (k.lmin != 0) continue;
      RadSegs G;
      G.isrc = rad_src_interval(HT, nk, k.rs);
      G.itop = rad_top_interval(HT, nk);
      G.nseg = G.isrc + (G.itop - G.isrc + 1);
      G.nwb = (k.nw + 63) / 64;
      G.cap = 0;
      if (G.isrc < 1 || G.itop < G.isrc || HT(F_R, 1) == HT(F_R, 2))
        return set_err(ctx, YSX_ERR_UNSUPPORTED, "ysx_model_vjp: radial grid without a source "
                                                 "interval or a first interval");
      nsmax = std::max(nsmax, (size_t)G.nseg * k.nw);
      rev.push_back(e);
      segs.push_back(G);
    }
    if (!rev.empty() && !ctx->d_rprop) {
      HIPCHK(hipMalloc(&ctx->d_rprop, nsmax * 4 * sizeof(cd)));
      HIPCHK(hipMalloc(&ctx->d_rys, nsmax * 2 * sizeof(cd)));
      HIPCHK(hipMalloc(&ctx->d_rxe, nsmax * 2 * sizeof(cd)));
      HIPCHK(hipMalloc(&ctx->d_rcap, sizeof(int)));
    }
    if (!rev.empty() && ctx->rad_caps.empty()) {
      // the tape's length per event: the most steps of a segment (once per run)
      size_t bytes = 1;
      for (size_t q = 0; q < rev.size(); q++) {
        const RadSegs& G = segs[q];
        launch_reset(dim3(1), b64, 0, rs, ctx->d_rcap, 1);
        launch_rad_prop(dim3(G.nseg, G.nwb), b64, 0, rs, ctx->d_tab, ctx->d_runs, rev[q], G,
                        ctx->d_rprop, ctx->d_rcap);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(rs));
        int cap = 0;
        HIPCHK(hipMemcpy(&cap, ctx->d_rcap, sizeof(int), hipMemcpyDeviceToHost));
        ctx->rad_caps.push_back(cap);
        bytes = std::max(bytes, (size_t)G.nseg * G.nwb * cap * 64 * rad_step_bytes());
      }
      if (bytes > ctx->rtape_bytes) {
        dfree(ctx->d_rtape);
        HIPCHK(hipMalloc(&ctx->d_rtape, bytes));
        ctx->rtape_bytes = bytes;
      }
    }
    for (size_t q = 0; q < rev.size(); q++) {
      RadSegs G = segs[q];
      G.cap = ctx->rad_caps[q];
      const dim3 gs(G.nseg, G.nwb);
      launch_reset(dim3(1), b64, 0, rs, ctx->d_rcap, 1);
      launch_rad_prop(gs, b64, 0, rs, ctx->d_tab, ctx->d_runs, rev[q], G, ctx->d_rprop,
                      ctx->d_rcap);
      launch_rad_scan(dim3(G.nwb), b64, 0, rs, ctx->d_tab, ctx->d_runs, rev[q], G,
                      ctx->d_rprop, ctx->d_recb, ctx->d_rys, ctx->d_rxe, ctx->d_tbar);
      launch_rad_seg_adj(gs, b64, 0, rs, ctx->d_tab, ctx->d_runs, rev[q], G, ctx->d_rys,
                         ctx->d_rxe, ctx->d_rtape, ctx->d_aerr, ctx->d_tbar);
      HIPCHK(hipGetLastError());
    }
      return YSX_OK;
  };
  const dim3 gb(ctx->adj_blocks);
  const StepStore FS = ctx->fs_valid ? StepStore{ctx->d_fys, ctx->d_fyss, ctx->d_foff,
                                                 ctx->d_fcap, ctx->d_fsc, ctx->d_ferr, nullptr}
                                       : StepStore{};
  const StepStore FT = ctx->fs_valid ? StepStore{ctx->d_tys, ctx->d_tyss, ctx->d_foff,
                                                 ctx->d_fcap, ctx->d_tsc, ctx->d_ferr,
                                                 ctx->d_tyend}
                                       : StepStore{};
  // With the toroidal states from the forward, the toroidal sweep uses none
  // of the blocks' per-step storage or scratch: it runs on the second
  // stream beside the minor sweep, whose persistent blocks hold every
  // compute unit until their queue is empty and then leave them one by one
  // as their last items end, into its own partial table columns.
  // Otherwise the three sweeps share that storage and run one after another.
  const bool side = FT.ys != nullptr;
  hipStream_t st2 = ctx->stream2;
  HIPCHK(hipEventRecord(ev[0], st));
  launch_brow_adj(gb, b64, 0, st, ctx->d_tab, ctx->d_runs, qb, S, ctx->d_rec, ctx->d_recb,
                  ctx->d_mub, AW, ctx->d_tpart);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ev[1], st));
  if (side) {
    HIPCHK(hipStreamWaitEvent(st2, ev[1], 0));
    HIPCHK(hipEventRecord(ev[12], st2));
    launch_tor_adj(gb, b64, 0, st2, ctx->d_tab, ctx->d_runs, qt, SD, ctx->d_rec, ctx->d_recb,
                   ctx->d_y0bar, AW, FT, ctx->d_tpart2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev[14], st2));
    // the radial solves' adjoint after it, also beside the minor sweep
    if (int rc = radial(st2)) return rc;
    HIPCHK(hipEventRecord(ev[15], st2));
  }
  const bool split = FS.ys != nullptr;
  if (split) {
    // the row chains, then every interval's pairings on all compute units.
    // A chain is one wave's sequential march, so the chains alone keep few
    // compute units busy, and an item's pairings wait for its chain.  The
    // items are cost-sorted and cut into groups, each on its own stream with
    // its own step scratch, units, counters and partial columns: the longest
    // chains here, the others beside them, each group's pairings as soon as
    // its own chains end, in the compute units the longer chains leave
    // idle.  (On the lmax 300, 20 mHz gradient: one group 10.1 ms, two
    // 9.4 ms, three 8.9 - 9.1 ms over several cuts.)
    const int per = 2 * nk + 2;  // pairing units of an item, at most
    launch_reset(dim3(1), b64, 0, st, ctx->d_pq, 8);
    HIPCHK(hipEventRecord(ev[16], st));
    int i0 = 0;
    for (int g = 0; g < ngroups; g++) {
      const int i1 = g + 1 < ngroups ? std::max(i0, std::min(qm.n, (int)(gcut[g] * qm.n))) : qm.n;
      Queue q = qm;
      q.items = qm.items + i0;
      q.n = i1 - i0;
      q.counter = g == 0 ? qm.counter : ctx->d_pq + 5 + g;
      const RowStore RG{ctx->d_fxs, ctx->d_fxss, ctx->d_units + (size_t)i0 * per,
                        ctx->d_pq + 2 * g, ctx->d_pq + 2 * g + 1,
                        std::max(0, std::min(ctx->units_cap - i0 * per, q.n * per))};
      AdjWork AG = AW;
      hipStream_t sg = st;
      double* tg = ctx->d_tpart;
      if (g > 0) {
        AG.scr = ctx->d_gscr[g - 1];
        sg = ctx->gstream[g - 1];
        tg = ctx->d_gtpart[g - 1];
        HIPCHK(hipStreamWaitEvent(sg, ev[16], 0));
      }
      if (q.n > 0) {
        launch_minor_rows(dim3((unsigned)std::min(q.n, ctx->adj_blocks)), b64, 0, sg, ctx->d_tab,
                          ctx->d_runs, q, SD, S, ctx->d_rec, ctx->d_recb, ctx->d_mub,
                          ctx->d_m0bar, AG, FS, RG, tg);
        launch_minor_pairs(gb, b64, 0, sg, ctx->d_tab, ctx->d_runs, AG, FS, RG, tg);
      }
      HIPCHK(hipGetLastError());
      if (g > 0) {
        HIPCHK(hipEventRecord(ev[16 + g], sg));
      }
      i0 = i1;
    }
    for (int g = 1; g < ngroups; g++) HIPCHK(hipStreamWaitEvent(st, ev[16 + g], 0));
  } else {
    launch_minor_adj(gb, b64, 0, st, ctx->d_tab, ctx->d_runs, qm, SD, S, ctx->d_rec,
                     ctx->d_recb, ctx->d_mub, ctx->d_m0bar, AW, FS, ctx->d_tpart);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ev[2], st));
  if (side) {
    HIPCHK(hipStreamWaitEvent(st, ev[15], 0));
  } else {
    launch_tor_adj(gb, b64, 0, st, ctx->d_tab, ctx->d_runs, qt, SD, ctx->d_rec, ctx->d_recb,
                   ctx->d_y0bar, AW, FT, ctx->d_tpart);
    HIPCHK(hipGetLastError());
  }
  launch_part_sum(dim3((unsigned)((ntab + 3) / 4)), b256, 0, st, ctx->d_tpart, ctx->adj_blocks,
                  ntab, ctx->d_tbar);
  if (side)
    launch_part_sum(dim3((unsigned)((ntab + 3) / 4)), b256, 0, st, ctx->d_tpart2,
                    ctx->adj_blocks, ntab, ctx->d_tbar);
  if (split)
    for (int g = 0; g + 1 < ngroups; g++)
      launch_part_sum(dim3((unsigned)((ntab + 3) / 4)), b256, 0, st, ctx->d_gtpart[g],
                      ctx->adj_blocks, ntab, ctx->d_tbar);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ev[3], st));
  {
    PairRange PR = pairs_of(ctx, 0, nev);
    launch_start_adj(dim3((unsigned)((PR.n + 255) / 256)), b256,
                     (size_t)SV_N * 2 * nk * sizeof(double), st, ctx->d_tab, ctx->d_runs, PR, SD,
                     ctx->d_starts, ctx->d_m0bar, ctx->d_y0bar, ctx->d_tbar);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(ev[4], st));
  if (!side)
    if (int rc = radial(st)) return rc;
  HIPCHK(hipEventRecord(ev[5], st));
  // the source vectors' part (as ysx_source_model_vjp)
  if (int rc = enqueue_source_sums(ctx, true, st, 6, 7)) return rc;
  HIPCHK(hipEventSynchronize(ev[7]));
  for (int q = 0; q < 5; q++)
    if (int rc = elapsed(ctx, q, q + 1, ctx->ms_grad[q])) return rc;
  // the toroidal sweep and the radial adjoint on the second stream (their
  // own events)
  if (side)
    if (int rc = elapsed(ctx, 12, 14, ctx->ms_grad[2]) || elapsed(ctx, 14, 15, ctx->ms_grad[4]))
      return rc;
  int errw[16];
  HIPCHK(hipMemcpy(errw, ctx->d_aerr, sizeof errw, hipMemcpyDeviceToHost));
  const int err = errw[0];
  if (err & 1) {
    double info[5];
    std::memcpy(info, errw + 2, sizeof info);
    char msg[256];
    std::snprintf(msg, sizeof msg, "ysx_model_vjp: a sweep item with more steps than its "
                  "per-step storage (%d): interval %d reaches %d steps, from %.17g to %.17g at "
                  "w %.17g", ctx->adj_smax, errw[1], (int)info[0], info[1], info[2], info[3]);
    return set_err(ctx, YSX_ERR_HIP, msg);
  }
  if (err & 2)
    return set_err(ctx, YSX_ERR_HIP, "ysx_model_vjp: a radial segment with more steps than "
                                     "its tape");
  if (err & 4)
    return set_err(ctx, YSX_ERR_HIP, "ysx_model_vjp: more minor-sweep pairing units than their "
                                     "queue holds");
  if (FS.ys) {
    int ferr = 0;
    HIPCHK(hipMemcpy(&ferr, ctx->d_ferr, sizeof ferr, hipMemcpyDeviceToHost));
    if (ferr)
      return set_err(ctx, YSX_ERR_HIP, "ysx_model_vjp: a minor or toroidal march item of the "
                                       "forward with more steps than its per-step storage");
  }
  std::vector<double> tb(ntab);
  HIPCHK(hipMemcpy(tb.data(), ctx->d_tbar, ntab * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> G;
  if (int rc = collect_source_sums(ctx, G)) return rc;
  std::fill(table_bar, table_bar + ntab, 0.0);
  source_model_fold(ctx, G, table_bar);
  if (int rc = elapsed(ctx, 6, 7, ctx->ms_grad[5])) return rc;
  ctx->ms_grad[6] = 0.0;
  for (int q = 0; q < 6; q++) ctx->ms_grad[6] += ctx->ms_grad[q];
  for (size_t q = 0; q < ntab; q++) table_bar[q] += tb[q];
  return YSX_OK;
}

namespace {
// k_model_vjp's buffers for the context's model (doubles): table adjoint
// input [F_COUNT nk] | bar [F_COUNT nk] | aux [6 nk] | dbar [9 nk] |
// work [12 nlayer ws]
struct MvjpLayout {
  size_t ntab, ws, o_bar, o_aux, o_dbar, o_work, total;
};
MvjpLayout mvjp_layout(const HostModel& M) {
  MvjpLayout L;
  const size_t nk = M.nknot, nl = M.nlayer;
  int nmax = 0;
  for (size_t k = 0; k < nl; k++) nmax = std::max(nmax, M.layer_hi[k] - M.layer_lo[k] + 1);
  L.ntab = (size_t)F_COUNT * nk;
  L.ws = 3 * ((size_t)nmax + 2);
  L.o_bar = L.ntab;
  L.o_aux = 2 * L.ntab;
  L.o_dbar = L.o_aux + 6 * nk;
  L.o_work = L.o_dbar + 9 * nk;
  L.total = L.o_work + 12 * nl * L.ws;
  return L;
}

// k_model_vjp on the context's model: the table adjoint at d_in (device,
// F_COUNT x nknot; nullptr: the input part of the buffers) to the deck
// adjoint, left at *d_out (device, nknot x 9 row-major)
int model_vjp_deck_device(ysx_ctx* ctx, const double* d_in, double** d_out) {
  const HostModel& M = ctx->M;
  const MvjpLayout L = mvjp_layout(M);
  if (!ctx->d_mvjp) HIPCHK(hipMalloc(&ctx->d_mvjp, L.total * sizeof(double)));
  double* b = ctx->d_mvjp;
  launch_model_vjp(dim3(1), dim3(256), 0, ctx->stream, ctx->d_mdeck, M.nknot, M.ifanis,
                   ctx->N.vel_norm, ctx->d_mlay, ctx->d_mlay + M.nlayer, M.nlayer, ctx->d_tab,
                   d_in ? d_in : b, b + L.o_bar, b + L.o_work, L.ws, b + L.o_aux, b + L.o_dbar);
  HIPCHK(hipGetLastError());
  *d_out = b + L.o_dbar;
  return YSX_OK;
}
}  // namespace

extern "C" int ysx_model_vjp_deck(ysx_ctx* ctx, const double* table_bar, double* deck_bar) {
  if (!ctx || !table_bar || !deck_bar) return YSX_ERR_ARG;
  if (!ctx->have_model || !ctx->d_mdeck) return set_err(ctx, YSX_ERR_ARG, "no model");
  HIPCHK(hipSetDevice(ctx->device));
  const MvjpLayout L = mvjp_layout(ctx->M);
  if (!ctx->d_mvjp) HIPCHK(hipMalloc(&ctx->d_mvjp, L.total * sizeof(double)));
  HIPCHK(hipMemcpyAsync(ctx->d_mvjp, table_bar, L.ntab * sizeof(double), hipMemcpyHostToDevice,
                        ctx->stream));
  double* d_out = nullptr;
  if (int rc = model_vjp_deck_device(ctx, nullptr, &d_out)) return rc;
  HIPCHK(hipMemcpyAsync(deck_bar, d_out, 9 * (size_t)ctx->M.nknot * sizeof(double),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return YSX_OK;
}

extern "C" int ysx_misfit_grad(ysx_ctx* ctx, const double* obs, int nout, double* misfit,
                               double* knot_grad) {
  if (!ctx || !obs || !misfit || !knot_grad) return YSX_ERR_ARG;
  if (!ctx->have_run) return set_err(ctx, YSX_ERR_ARG, "no run");
  const int nev = (int)ctx->rk.size();
  // samples per event: nout, or each event's own output length
  std::vector<int> ne(nev);
  std::vector<size_t> ooff(nev + 1, 0);
  for (int e = 0; e < nev; e++) {
    const HostRun& R = ctx->R[e];
    ne[e] = nout > 0 ? nout : R.nout;
    if (ne[e] < 1 || ne[e] > R.nt)
      return set_err(ctx, YSX_ERR_ARG, "nout outside 1 .. nt (" + std::to_string(R.nt) +
                                           ") for event " + std::to_string(e));
    ooff[e + 1] = ooff[e] + (size_t)ne[e] * 3 * R.nr;
  }
  const size_t nobs = ooff[nev];
  constexpr int RBLK = 1024;
  if (ctx->obs_n < nobs) {
    dfree(ctx->d_obs);
    dfree(ctx->d_rpart);
    HIPCHK(hipMalloc(&ctx->d_obs, nobs * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->d_rpart, (size_t)RBLK * nev * sizeof(double)));
    ctx->obs_n = nobs;
  }
  // up to OBS_PIN bytes of observed seismograms are staged into pinned
  // memory by host threads while the forward runs (no HIP calls there) and
  // go up after it; more go up after the forward through the double buffer
  const size_t obytes = nobs * sizeof(double);
  const bool staged = obytes <= OBS_PIN;
  if (staged && ctx->opin_bytes < obytes) {
    if (ctx->opin) HIPCHK(hipHostFree(ctx->opin));
    ctx->opin = nullptr;
    ctx->opin_bytes = 0;
    HIPCHK(hipHostMalloc(&ctx->opin, obytes, hipHostMallocDefault));
    ctx->opin_bytes = obytes;
  }
  std::thread stage;
  if (staged) stage = std::thread([&]() { parallel_copy(ctx->opin, obs, obytes); });
  // one forward: the plain path for one event, the batched launch for
  // several; a single event's forward also stores the minor march's
  // per-step states for the minor sweep when they fit
  ctx->fs_want = nev == 1 && prepare_step_store(ctx);
  const int frc = nev == 1 ? ysx_forward(ctx, nullptr) : ysx_forward_events(ctx, 1, nullptr);
  ctx->fs_want = false;
  if (staged) stage.join();
  if (frc) return frc;
  if (int rc = adjoint_run(ctx)) return rc;
  if (staged)
    HIPCHK(hipMemcpyAsync(ctx->d_obs, ctx->opin, obytes, hipMemcpyHostToDevice, ctx->stream));
  else if (int rc = copy_h2d(ctx, ctx->d_obs, obs, obytes))
    return rc;
  std::vector<unsigned> nb(nev);
  for (int e = 0; e < nev; e++) {
    const HostRun& R = ctx->R[e];
    const RunK& k = ctx->rk[e];
    const size_t ntrace = 3 * (size_t)R.nr, n = (size_t)R.nt * ntrace;
    nb[e] = (unsigned)std::min<size_t>(RBLK, (n + 255) / 256);
    launch_residual(dim3(nb[e]), dim3(256), 0, ctx->stream, ctx->d_seis + k.seis_off,
                    ctx->d_obs + ooff[e], R.nt, ne[e], ntrace, ctx->d_seisb + k.seis_off,
                    ctx->d_rpart + (size_t)RBLK * e);
  }
  HIPCHK(hipGetLastError());
  // the context's streams do not block on the null stream: wait for the
  // residuals before copying their sums
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::vector<double> part((size_t)RBLK * nev);
  HIPCHK(hipMemcpy(part.data(), ctx->d_rpart, part.size() * sizeof(double),
                   hipMemcpyDeviceToHost));
  double sum = 0.0;
  for (int e = 0; e < nev; e++)
    for (unsigned q = 0; q < nb[e]; q++) sum += part[(size_t)RBLK * e + q];
  *misfit = 0.5 * sum;
  if (int rc = seis_vjp_device(ctx)) return rc;
  if (int rc = ysx_spectra_vjp(ctx, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr))
    return rc;
  const int nk = ctx->M.nknot;
  std::vector<double> tb((size_t)F_COUNT * nk, 0.0), db(9 * (size_t)nk, 0.0);
  if (int rc = ysx_model_vjp(ctx, tb.data())) return rc;
  // the whole table adjoint (solves + source vectors) to the deck: the host
  // function (about 0.1 ms for 200 knots) rather than k_model_vjp (one
  // latency-bound block, 0.76 ms), as the table adjoint is on the host here
  std::string e = model_table_vjp(ctx->N, nk, ctx->M.nic, ctx->M.noc, ctx->ians,
                                  ctx->tref_period, ctx->deck.data(), tb.data(), db.data());
  if (!e.empty()) return set_err(ctx, YSX_ERR_INPUT, e);
  for (int k = 0; k < nk; k++) {
    const double* d = db.data() + 9 * (size_t)k;
    double* g = knot_grad + 4 * (size_t)k;
    g[0] = d[1];         // rho
    g[1] = d[2] + d[6];  // vp: vpv and vph together
    g[2] = d[3] + d[7];  // vs: vsv and vsh together
    g[3] = d[5];         // Qmu
  }
  return YSX_OK;
}

extern "C" int ysx_grad_memory(const ysx_ctx* ctx, double* bytes3) {
  if (!ctx || !bytes3) return YSX_ERR_ARG;
  const double nb = ctx->adj_blocks, sm = ctx->adj_smax;
  const double P = (double)ctx->npair, nk = ctx->M.nknot;
  bytes3[0] = nb * sm * 64 * (14 * sizeof(cd) + sizeof(int)) +        // per-step states, and
              (double)ctx->fs_steps * 64 * (30 * sizeof(cd) + 3 * sizeof(int));  // the forward's
  const double groups = 1.0 + (ctx->d_gscr[0] ? 1 : 0) + (ctx->d_gscr[1] ? 1 : 0);  // minor sweep item groups
  bytes3[1] = groups * nb * 13 * 14 * 64 * sizeof(cd) +                // step scratch,
              (double)ctx->rtape_bytes;                                // radial tape
  bytes3[2] = ctx->d_tbar ? P * (2 * 14 + 2 + REC) * sizeof(cd) +      // per-pair adjoints,
                                (1.0 + groups) * nb * F_COUNT * nk * 8.0  // partial columns
                          : 0.0;
  return YSX_OK;
}

extern "C" int ysx_grad_timings(const ysx_ctx* ctx, double* ms7) {
  if (!ctx || !ms7) return YSX_ERR_ARG;
  std::memcpy(ms7, ctx->ms_grad, sizeof ctx->ms_grad);
  return YSX_OK;
}

extern "C" int ysx_vjp_timings(const ysx_ctx* ctx, double* ms3) {
  if (!ctx || !ms3) return YSX_ERR_ARG;
  std::memcpy(ms3, ctx->ms_vjp, sizeof ctx->ms_vjp);
  return YSX_OK;
}

extern "C" int ysx_pair_steps(ysx_ctx* ctx, int n, const int* l, const int* j, int* out5) {
  if (!ctx || (n > 0 && (!l || !j || !out5))) return YSX_ERR_ARG;
  if (!ctx->have_spec) return set_err(ctx, YSX_ERR_ARG, "no spectra computed");
  if (ctx->rk.size() != 1) return set_err(ctx, YSX_ERR_ARG, "several events: one-event diagnostic");
  HIPCHK(hipSetDevice(ctx->device));
  const RunK& k = ctx->rk[0];
  size_t P = (size_t)k.npair;
  std::vector<int> c(P * 5);
  HIPCHK(hipMemcpy(c.data(), ctx->d_cnt, c.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++) {
    if (l[i] < k.lmin || l[i] > k.lmax || j[i] < 1 || j[i] > k.nw)
      return set_err(ctx, YSX_ERR_ARG, "pair outside the run");
    size_t p = (size_t)(l[i] - k.lmin) * k.nw + (j[i] - 1);
    for (int q = 0; q < 5; q++) out5[5 * (size_t)i + q] = c[5 * p + q];
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
