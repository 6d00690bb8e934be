// Host-side checks of the adjoint numerics in yspec_amd/csrc/ysx_adjoint.hpp
// (compiled and run on the CPU by tests/test_adjoint_math.py).
//
//   adjoint_check <model table file (29 x nknot doubles)> <nknot>
//
// Prints one line per check: "<name> <max relative error>".  Every check is
// an identity the mathematics fixes, not a restatement of the code:
//   * bilinear forms: sum_p w_p d_p(A, z) = sum_a a * abar(a) for random
//     complex entries, w and z (d the system's own right-hand side code);
//   * transposed systems: x . (A z) = (A^T x) . z;
//   * per-lane base adjoints and base adjoints: central differences of
//     Re sum_p w_p d_p with respect to the stored base doubles, and of
//     Re sum ob * base with respect to the model-table entries;
//   * one Cash-Karp step: x . (S y) = (x S) . y, and the stage pairings of
//     the transposed step against central differences of Re x . (S y) with
//     respect to every stored base double of every stage.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "ysx_adjoint.hpp"
#include "ysx_radial.hpp"

using namespace ysx;

static std::mt19937_64 rng(12345);
static double U() { return std::uniform_real_distribution<double>(-1.0, 1.0)(rng); }
static cd CU() { return C(U(), U()); }
static double rel(double a, double b, double scale) { return std::fabs(a - b) / scale; }
static cd dot(const cd* a, const cd* b, int n) {
  cd s = C(0.0);
  for (int i = 0; i < n; i++) s = s + cmul_(a[i], b[i]);
  return s;
}

static void random_solidA(SolidA& A) {
  A.t11 = CU(); A.t12 = CU(); A.k11 = CU(); A.k22 = CU(); A.s11 = CU(); A.s12 = CU();
  A.s22 = CU();
  A.t21 = U(); A.t22 = U(); A.t31 = U(); A.t33 = U(); A.k33 = U(); A.s13 = U(); A.s23 = U();
}

template <class Derivs, class Bil>
static double check_bilinear(Derivs derivs, Bil bil) {
  double worst = 0.0;
  for (int trial = 0; trial < 50; trial++) {
    SolidA A;
    random_solidA(A);
    cd w[14], z[14], d[14];
    for (int i = 0; i < 14; i++) { w[i] = CU(); z[i] = CU(); }
    derivs(A, z, d);
    cd lhs = dot(w, d, 14);
    SolidAbar B;
    zero(B);
    bil(w, z, B);
    cd rhs = cmul_(A.t11, B.t11) + cmul_(A.t12, B.t12) + cmul_(A.k11, B.k11) +
             cmul_(A.k22, B.k22) + cmul_(A.s11, B.s11) + cmul_(A.s12, B.s12) +
             cmul_(A.s22, B.s22) + A.t21 * B.t21 + A.t22 * B.t22 + A.t31 * B.t31 +
             A.t33 * B.t33 + A.k33 * B.k33 + A.s13 * B.s13 + A.s23 * B.s23;
    worst = std::fmax(worst, cabs_(lhs - rhs) / cabs_(lhs));
  }
  return worst;
}

// random stored base of a system: plausible magnitudes, rho > 0
template <int BS>
static void random_base(double* b) {
  for (int i = 0; i < BS; i++) b[i] = U();
  b[0] = 0.3 + 0.1 * U();  // h / r
  b[1] = 1.0 + 0.2 * U();  // rho
}

template <class F, class FT>
static double check_transpose(int n, int bs) {
  double worst = 0.0;
  LS L = l_state(37);
  WS w = omega_state(C(0.7, -0.01), 1.0, 1.0, 1.0);
  for (int trial = 0; trial < 30; trial++) {
    std::vector<double> b(bs);
    for (int i = 0; i < bs; i++) b[i] = U();
    b[1] = 1.0 + 0.2 * U();
    cd x[14], z[14], d[14], e[14];
    for (int i = 0; i < n; i++) { x[i] = CU(); z[i] = CU(); }
    F().stream(b.data(), L, w, z, [&](int i, cd v) { d[i] = v; });
    FT().stream(b.data(), L, w, x, [&](int i, cd v) { e[i] = v; });
    cd a = dot(x, d, n), c = dot(e, z, n);
    worst = std::fmax(worst, cabs_(a - c) / cabs_(a));
  }
  return worst;
}

// the per-lane base adjoints: dF/db for F = Re sum_p w_p d_p(b, z)
template <class F, class Abar, class Bil, class Ob>
static double check_lane_ob(int n, int bs, const int* comp, int ncomp, const int* real_idx,
                            int nreal, Bil bil, Ob ob_of) {
  double worst = 0.0;
  WS w = omega_state(C(0.7, -0.01), 1.0, 1.0, 1.0);
  for (int trial = 0; trial < 20; trial++) {
    LS L = l_state(2 + trial * 7);
    std::vector<double> b(bs);
    for (int i = 0; i < bs; i++) b[i] = U();
    b[0] = 0.3 + 0.1 * U();
    b[1] = 1.0 + 0.2 * U();
    cd wv[14], z[14];
    for (int i = 0; i < n; i++) { wv[i] = CU(); z[i] = CU(); }
    auto Fv = [&](const std::vector<double>& bb) {
      cd d[14];
      F().stream(bb.data(), L, w, z, [&](int i, cd v) { d[i] = v; });
      return dot(wv, d, n).x;
    };
    Abar B;
    zero(B);
    bil(wv, z, B);
    double ob[32];
    ob_of(B, L, b.data(), ob);
    double scale = 0.0;
    std::vector<double> fd(bs, 0.0);
    for (int k = 0; k < bs; k++) {
      std::vector<double> bp = b, bm = b;
      double h = 1e-6 * std::fmax(1.0, std::fabs(b[k]));
      bp[k] += h;
      bm[k] -= h;
      fd[k] = (Fv(bp) - Fv(bm)) / (2 * h);
      scale = std::fmax(scale, std::fabs(fd[k]));
    }
    // complex entries: base (re, im) = (comp[c], comp[c] + 1) <-> ob[2c], ob[2c + 1]
    for (int c = 0; c < ncomp; c++) {
      worst = std::fmax(worst, rel(ob[2 * c], fd[comp[c]], scale));
      worst = std::fmax(worst, rel(-ob[2 * c + 1], fd[comp[c] + 1], scale));
    }
    for (int r = 0; r < nreal; r++)
      worst = std::fmax(worst, rel(ob[2 * ncomp + r], fd[real_idx[r]], scale));
  }
  return worst;
}

struct Model {
  std::vector<double> t;
  int nk;
  Tab T() const { return Tab{t.data(), nk}; }
};

// base adjoints: G(T) = sum over the model-dependent stored base doubles of
// ob-weighted values; its table gradient against the chain
template <class F, class Adj, class Field>
static double check_base_adj(Model& M, int bs, const int* comp, int ncomp, const int* real_idx,
                             int nreal, int nv, Adj adj, Field field, int il) {
  double worst = 0.0;
  Tab T0 = M.T();
  double ra = T0(F_R, il), rb = T0(F_R, il + 1);
  WS w = omega_state(C(0.9, -0.02), 1.0, 1.0, 1.0);
  for (int trial = 0; trial < 6; trial++) {
    double rr = ra + (0.1 + 0.15 * trial) * (rb - ra), h = 0.37 * (rb - ra);
    double ob[32];
    for (int i = 0; i < 2 * ncomp + nreal; i++) ob[i] = U();
    auto G = [&]() {
      std::vector<double> o(bs);
      F::base(M.T(), w, il, rr, h, o.data());
      double g = 0.0;
      for (int c = 0; c < ncomp; c++) g += cmul_(C(ob[2 * c], ob[2 * c + 1]),
                                                 C(o[comp[c]], o[comp[c] + 1])).x;
      for (int r = 0; r < nreal; r++) g += ob[2 * ncomp + r] * o[real_idx[r]];
      return g;
    };
    double vb[16];
    adj(T0, w, il, rr, h, ob, vb);
    std::vector<double> tb(M.t.size(), 0.0);
    SplW sw = spl_w(T0, il, rr);
    for (int v = 0; v < nv; v++)
      spl_adj(sw, field(v), vb[v],
              [&](int f, int knot, double x) { tb[(size_t)f * M.nk + knot - 1] += x; });
    double scale = 0.0;
    std::vector<double> fd(M.t.size(), 0.0);
    for (int v = 0; v < nv; v++)
      for (int f : {field(v), field(v) + 1})
        for (int knot : {il, il + 1}) {
          size_t e = (size_t)f * M.nk + knot - 1;
          double x0 = M.t[e], hh = 1e-6 * std::fmax(std::fabs(x0), 1e-3);
          M.t[e] = x0 + hh;
          double gp = G();
          M.t[e] = x0 - hh;
          double gm = G();
          M.t[e] = x0;
          fd[e] = (gp - gm) / (2 * hh);
          scale = std::fmax(scale, std::fabs(fd[e]));
        }
    for (size_t e = 0; e < fd.size(); e++) worst = std::fmax(worst, rel(tb[e], fd[e], scale));
  }
  return worst;
}

// one step: x.(S y) = (x S).y, and the stage pairings against central
// differences in every stored base double of every stage
template <class F, class FT, class Abar, class Bil, class Ob>
static void check_step(const char* name, int bs, const int* comp, int ncomp, const int* real_idx,
                       int nreal, Bil bil, Ob ob_of) {
  constexpr int N = F::N;
  double wt = 0.0, wg = 0.0;
  WS w = omega_state(C(0.8, -0.015), 1.0, 1.0, 1.0);
  for (int trial = 0; trial < 5; trial++) {
    LS L = l_state(3 + 11 * trial);
    std::vector<double> sb(6 * bs);
    for (int k = 0; k < 6; k++) {
      for (int i = 0; i < bs; i++) sb[k * bs + i] = 0.2 * U();
      sb[k * bs + 0] = 0.05 + 0.01 * U();
      sb[k * bs + 1] = 1.0 + 0.2 * U();
    }
    cd y[N], x[N];
    for (int i = 0; i < N; i++) { y[i] = CU(); x[i] = CU(); }
    auto Fv = [&](const std::vector<double>& b) {
      cd yy[N];
      for (int i = 0; i < N; i++) yy[i] = y[i];
      rk_step_hook(b.data(), L, w, yy, F(), [](int, const cd*) {});
      return dot(x, yy, N);
    };
    cd z[6][N];
    cd yy[N];
    for (int i = 0; i < N; i++) yy[i] = y[i];
    rk_step_hook(sb.data(), L, w, yy, F(), [&](int k, const cd* v) {
      for (int i = 0; i < N; i++) z[k][i] = v[i];
    });
    cd fwd = dot(x, yy, N);
    cd xs[N];
    for (int i = 0; i < N; i++) xs[i] = x[i];
    double ob[6][32];
    rk_step_T_reg(sb.data(), L, w, xs, FT(), [&](int k, const cd* kb) {
      Abar B;
      zero(B);
      bil(kb, z[k], B);
      ob_of(B, L, sb.data() + k * bs, ob[k]);
    });
    wt = std::fmax(wt, cabs_(fwd - dot(xs, y, N)) / cabs_(fwd));
    std::vector<double> fd(6 * bs);
    double scale = 0.0;
    for (size_t e = 0; e < fd.size(); e++) {
      std::vector<double> bp = sb, bm = sb;
      double h = 1e-6 * std::fmax(std::fabs(sb[e]), 1e-2);
      bp[e] += h;
      bm[e] -= h;
      fd[e] = (Fv(bp).x - Fv(bm).x) / (2 * h);
      scale = std::fmax(scale, std::fabs(fd[e]));
    }
    for (int k = 0; k < 6; k++) {
      for (int c = 0; c < ncomp; c++) {
        wg = std::fmax(wg, rel(ob[k][2 * c], fd[k * bs + comp[c]], scale));
        wg = std::fmax(wg, rel(-ob[k][2 * c + 1], fd[k * bs + comp[c] + 1], scale));
      }
      for (int r = 0; r < nreal; r++)
        wg = std::fmax(wg, rel(ob[k][2 * ncomp + r], fd[k * bs + real_idx[r]], scale));
    }
  }
  printf("step_transpose_%s %.3e\n", name, wt);
  printf("step_gradient_%s %.3e\n", name, wg);
}


// radial coefficients: G(T) = Re(tb t + kb k + svb sv) at a radius of
// interval il; its table gradient against rad_coef_adj and the weights
static double check_rad_coef(Model& M, int il) {
  double worst = 0.0;
  PS s;
  set_l(s, 0);
  s.pibigg = 1.0;
  set_omega(s, C(0.9, -0.02), 1.0, 1.0);
  Tab T0 = M.T();
  double ra = T0(F_R, il), rb = T0(F_R, il + 1);
  for (int trial = 0; trial < 5; trial++) {
    double rr = ra + (0.05 + 0.2 * trial) * (rb - ra);
    cd tb = CU(), kb = CU(), svb = CU();
    auto G = [&]() {
      RadRHS::Coef c = RadRHS().coef(M.T(), s, il, rr);
      return (cmul_(tb, c.t) + cmul_(kb, c.k) + cmul_(svb, c.sv)).x;
    };
    double vb[RAD_NV];
    rad_coef_adj(T0, s, il, rr, tb, kb, svb, vb);
    std::vector<double> tbar(M.t.size(), 0.0);
    SplW sw = spl_w(T0, il, rr);
    for (int v = 0; v < RAD_NV; v++)
      spl_adj(sw, rad_field(v), vb[v],
              [&](int f, int knot, double x) { tbar[(size_t)f * M.nk + knot - 1] += x; });
    double scale = 0.0;
    std::vector<double> fd(M.t.size(), 0.0);
    for (int v = 0; v < RAD_NV; v++)
      for (int f : {rad_field(v), rad_field(v) + 1})
        for (int knot : {il, il + 1}) {
          size_t e = (size_t)f * M.nk + knot - 1;
          double x0 = M.t[e], hh = 1e-6 * std::fmax(std::fabs(x0), 1e-3);
          M.t[e] = x0 + hh;
          double gp = G();
          M.t[e] = x0 - hh;
          double gm = G();
          M.t[e] = x0;
          fd[e] = (gp - gm) / (2 * hh);
          scale = std::fmax(scale, std::fabs(fd[e]));
        }
    for (size_t e = 0; e < fd.size(); e++) worst = std::fmax(worst, rel(tbar[e], fd[e], scale));
  }
  return worst;
}

// starting-solution tangents against central differences of the forward
// functions in the table values they read (start at a knot, and inside an
// interval where the spline weights enter)
template <class Fwd, class Tan>
static double check_start_tangent(Model& M, int is, double frac, int l, int n, Fwd fwd, Tan tan) {
  PS s;
  set_l(s, l);
  s.pibigg = 1.0;
  set_omega(s, C(0.6, -0.01), 1.0, 1.0);
  Tab T0 = M.T();
  double rr = T0(F_R, is) + frac * (T0(F_R, is + 1) - T0(F_R, is));
  DS y[14];
  tan(T0, s, is, rr, y);
  SplW sw = spl_w(T0, is, rr);
  double worst = 0.0, scale = 0.0;
  std::vector<std::pair<double, double>> got;
  for (int v = 0; v < SV_N; v++) {
    // four-point central differences with a wide step: the continued
    // fraction of the Bessel ratio stops at a 1e-9 tolerance, noise that a
    // narrow step would amplify
    size_t e = (size_t)start_field(v) * M.nk + is - 1;
    double x0 = M.t[e], h = 1e-3 * std::fmax(std::fabs(x0), 1e-4);
    cd y1p[14], y1m[14], y2p[14], y2m[14];
    M.t[e] = x0 + h;
    fwd(M.T(), s, is, rr, y1p);
    M.t[e] = x0 - h;
    fwd(M.T(), s, is, rr, y1m);
    M.t[e] = x0 + 2 * h;
    fwd(M.T(), s, is, rr, y2p);
    M.t[e] = x0 - 2 * h;
    fwd(M.T(), s, is, rr, y2m);
    M.t[e] = x0;
    for (int q = 0; q < n; q++) {
      cd fd = (8.0 * (y1p[q] - y1m[q]) - (y2p[q] - y2m[q])) / (12 * h);
      cd tg = sw.a * y[q].d[v];  // d value / d T(f, is) = weight a
      got.push_back({cabs_(tg - fd), cabs_(fd)});
      scale = std::fmax(scale, cabs_(fd));
    }
  }
  for (auto& g : got) worst = std::fmax(worst, g.first / scale);
  return worst;
}

// the radial segments: enumerated one by one (rad_segment) and put in sweep
// order, they are rad_sweeps' own steps, bitwise (interval, radius, step);
// returns the number of differences over frequencies and source and
// receiver radii (knot radii and inner points)
static double check_rad_segments(Model& M) {
  const Tab T = M.T();
  const int nk = M.nk;
  double amin = 1e300;
  for (int k = 1; k <= nk; k++) amin = std::fmin(amin, T(F_ALPHA, k));
  const double top = T(F_R, nk);
  int bad = 0, total = 0;
  for (double w : {0.02, 0.3, 1.7, 7.5, 30.0, 100.0})
    for (double rs : {0.3 * top, T(F_R, 31), T(F_R, 90), T(F_R, 150), 0.9922 * top, top})
      for (double rr : {top, 0.95 * top}) {
        PS s;
        set_l(s, 0);
        s.pibigg = 1.0;
        set_omega(s, C(w, -0.01), 1.0, 1.0);
        const double dr = 1.0 / (4.0 * s.rew / amin);
        struct St { int sweep, il; double x, h; };
        std::vector<St> a, b;
        cd y1[2], y2[2], y1r[2], y2r[2];
        bool below;
        int nsr = 0;
        rad_sweeps(T, nk, s, dr, rs, rr, NoGroup(),
                   [&](int sw, int il, double x, double h, const cd*) { a.push_back({sw, il, x, h}); },
                   y1, y2, y1r, y2r, below, nsr);
        const int isrc = rad_src_interval(T, nk, rs), itop = rad_top_interval(T, nk);
        for (int q = 0; q < isrc + (itop - isrc + 1); q++) {
          const int sw = q < isrc ? 0 : 1, i = q < isrc ? q + 1 : itop - (q - isrc);
          rad_segment(T, nk, s, dr, rs, rr, sw, i, isrc, itop,
                      [&](double x, double h) { b.push_back({sw, i, x, h}); });
        }
        total += (int)a.size();
        if (a.size() != b.size()) {
          bad += 1 + (int)std::abs((long)a.size() - (long)b.size());
          continue;
        }
        for (size_t k = 0; k < a.size(); k++)
          if (a[k].sweep != b[k].sweep || a[k].il != b[k].il || a[k].x != b[k].x ||
              a[k].h != b[k].h)
            bad++;
      }
  fprintf(stderr, "rad_segments: %d steps compared\n", total);
  return (double)bad;
}

static int main2(Model& M) {
  printf("rad_segments %.3e\n", check_rad_segments(M));
  printf("rad_coef_adj %.3e\n", std::fmax(check_rad_coef(M, 150), check_rad_coef(M, 60)));
  double ws = 0.0, wf = 0.0, wt = 0.0, wr = 0.0;
  for (double frac : {0.0, 0.37}) {
    // solid start in the mantle and in the inner core, fluid in the outer core
    for (int is : {20, 140}) {
      ws = std::fmax(ws, check_start_tangent(M, is, frac, 25, 14,
          [](const Tab& T, const PS& s, int i, double r, cd* m) { minor_solid_start(T, s, i, r, m); },
          [](const Tab& T, const PS& s, int i, double r, DS* m) { minor_start_d(T, s, i, r, false, m); }));
    }
    wf = std::fmax(wf, check_start_tangent(M, 60, frac, 25, 5,
        [](const Tab& T, const PS& s, int i, double r, cd* m) { minor_fluid_start(T, s, i, r, m); },
        [](const Tab& T, const PS& s, int i, double r, DS* m) { minor_start_d(T, s, i, r, true, m); }));
    wt = std::fmax(wt, check_start_tangent(M, 140, frac, 25, 2,
        [](const Tab& T, const PS& s, int i, double r, cd* y) { tor_ystart(T, s, i, r, y); },
        [](const Tab& T, const PS& s, int i, double r, DS* y) { tor_ystart_d(T, s, i, r, y); }));
    wr = std::fmax(wr, check_start_tangent(M, 1, frac + 0.01, 0, 2,
        [](const Tab& T, const PS& s, int i, double r, cd* y) { rad_ystart(T, s, i, r, y); },
        [](const Tab& T, const PS& s, int i, double r, DS* y) { rad_ystart_d(T, s, i, r, y); }));
  }
  printf("start_tangent_solid %.3e\n", ws);
  printf("start_tangent_fluid %.3e\n", wf);
  printf("start_tangent_tor %.3e\n", wt);
  printf("start_tangent_rad %.3e\n", wr);
  return 0;
}

// The z-streamed bilinear forms (solid_minor_bil_s, ...) give the same
// per-lane base adjoints as the full forms: ob(full) against ob(streamed)
// for random rows, columns, bases and degrees.
template <class Abar, int N, int NOB, class Full, class Stream, class Ob>
static double check_streamed(Full full, Stream stream, Ob ob) {
  double worst = 0.0;
  for (int trial = 0; trial < 50; trial++) {
    cd w[14], z[14];
    double base[20];
    for (int i = 0; i < 14; i++) { w[i] = CU(); z[i] = CU(); }
    for (int i = 0; i < 20; i++) base[i] = U();
    LS L = l_state(2 + trial * 13);
    Abar B1, B2;
    zero(B1);
    zero(B2);
    full(w, z, B1);
    stream(w, [&](int q) { return z[q]; }, B2);
    double o1[NOB], o2[NOB], scale = 0.0;
    ob(B1, L, base, o1);
    ob(B2, L, base, o2);
    for (int v = 0; v < NOB; v++) scale = std::fmax(scale, std::fabs(o1[v]));
    for (int v = 0; v < NOB; v++) worst = std::fmax(worst, rel(o1[v], o2[v], scale));
  }
  return worst;
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  Model M;
  M.nk = atoi(argv[2]);
  M.t.resize((size_t)F_COUNT * M.nk);
  FILE* fp = fopen(argv[1], "rb");
  if (!fp || fread(M.t.data(), sizeof(double), M.t.size(), fp) != M.t.size()) return 3;
  fclose(fp);

  printf("bilinear_solid_minor %.3e\n",
         check_bilinear([](const SolidA& A, const cd* z, cd* d) {
           solid_minor_derivs(A, z, NoInit(), [&](int i, cd v) { d[i] = v; });
         }, [](const cd* w, const cd* z, SolidAbar& B) { solid_minor_bil(w, z, B); }));
  printf("bilinear_bvec %.3e\n",
         check_bilinear([](const SolidA& A, const cd* z, cd* d) {
           bvec_derivs(A, z, NoInit(), [&](int i, cd v) { d[i] = v; });
         }, [](const cd* w, const cd* z, SolidAbar& B) { bvec_bil(w, z, B); }));

  printf("streamed_solid_minor %.3e\n", check_streamed<SolidAbar, 14, SOLID_OB>(
      [](const cd* w, const cd* z, SolidAbar& B) { solid_minor_bil(w, z, B); },
      [](const cd* w, auto zl, SolidAbar& B) { solid_minor_bil_s(w, zl, B); },
      [](const SolidAbar& B, const LS& L, const double* b, double* o) { solid_ob(B, L, b, o); }));
  printf("streamed_bvec %.3e\n", check_streamed<SolidAbar, 14, SOLID_OB>(
      [](const cd* w, const cd* z, SolidAbar& B) { bvec_bil(w, z, B); },
      [](const cd* w, auto zl, SolidAbar& B) { bvec_bil_s(w, zl, B); },
      [](const SolidAbar& B, const LS& L, const double* b, double* o) { solid_ob(B, L, b, o); }));
  printf("streamed_fluid_minor %.3e\n", check_streamed<FluidAbar, 5, FLUID_OB>(
      [](const cd* w, const cd* z, FluidAbar& B) { fluid_minor_bil(w, z, B); },
      [](const cd* w, auto zl, FluidAbar& B) { fluid_minor_bil_s(w, zl, B); },
      [](const FluidAbar& B, const LS& L, const double* b, double* o) { fluid_ob(B, L, b, o); }));
  printf("streamed_tor %.3e\n", check_streamed<TorAbar, 2, TOR_OB>(
      [](const cd* w, const cd* z, TorAbar& B) { tor_bil(w, z, B); },
      [](const cd* w, auto zl, TorAbar& B) { tor_bil_s(w, zl, B); },
      [](const TorAbar& B, const LS& L, const double* b, double* o) { tor_ob(B, L, b, o); }));

  printf("transpose_solid_minor %.3e\n", check_transpose<SolidMinorM, MinorTM>(14, 20));
  printf("transpose_bvec %.3e\n", check_transpose<BvecM, BvecTM>(14, 20));
  printf("transpose_fluid_minor %.3e\n", check_transpose<FluidMinorM, FluidMinorT>(5, 14));
  printf("transpose_tor %.3e\n", check_transpose<TorM, TorT>(2, 7));

  // model-dependent stored base doubles of each system (complex entries by
  // their first double, then the real ones), in the order of the ob layouts
  const int s_comp[8] = {2, 4, 6, 8, 10, 12, 14, 16}, s_real[2] = {1, 18};
  const int f_comp[4] = {2, 6, 8, 10}, f_real[2] = {1, 12};
  const int t_comp[3] = {1, 3, 5};
  auto solid_obf = [](const SolidAbar& B, const LS& L, const double* b, double* o) {
    solid_ob(B, L, b, o);
  };
  auto fluid_obf = [](const FluidAbar& B, const LS& L, const double* b, double* o) {
    fluid_ob(B, L, b, o);
  };
  auto tor_obf = [](const TorAbar& B, const LS& L, const double* b, double* o) {
    tor_ob(B, L, b, o);
  };
  auto sbil = [](const cd* w, const cd* z, SolidAbar& B) { solid_minor_bil(w, z, B); };
  auto bbil = [](const cd* w, const cd* z, SolidAbar& B) { bvec_bil(w, z, B); };
  auto fbil = [](const cd* w, const cd* z, FluidAbar& B) { fluid_minor_bil(w, z, B); };
  auto tbil = [](const cd* w, const cd* z, TorAbar& B) { tor_bil(w, z, B); };
  printf("lane_ob_solid_minor %.3e\n",
         check_lane_ob<SolidMinorM, SolidAbar>(14, 20, s_comp, 8, s_real, 2, sbil, solid_obf));
  printf("lane_ob_bvec %.3e\n",
         check_lane_ob<BvecM, SolidAbar>(14, 20, s_comp, 8, s_real, 2, bbil, solid_obf));
  printf("lane_ob_fluid_minor %.3e\n",
         check_lane_ob<FluidMinorM, FluidAbar>(5, 14, f_comp, 4, f_real, 2, fbil, fluid_obf));
  printf("lane_ob_tor %.3e\n",
         check_lane_ob<TorM, TorAbar>(2, 7, t_comp, 3, nullptr, 0, tbil, tor_obf));

  // base adjoints on a mantle, an outer-core and a crust interval
  const int ils[3] = {150, 60, 196};
  double ws = 0.0, wf = 0.0, wtb = 0.0;
  for (int il : ils) {
    if (il + 1 > M.nk) continue;
    ws = std::fmax(ws, check_base_adj<SolidB>(
                           M, 20, s_comp, 8, s_real, 2, SOLID_NV,
                           [](const Tab& T, const WS& w, int i, double r, double h,
                              const double* ob, double* vb) {
                             solid_base_adj(T, w, i, r, h, ob, vb);
                           },
                           solid_field, il));
    wf = std::fmax(wf, check_base_adj<FluidMinorM>(
                           M, 14, f_comp, 4, f_real, 2, FLUID_NV,
                           [](const Tab& T, const WS& w, int i, double r, double h,
                              const double* ob, double* vb) {
                             fluid_base_adj(T, w, i, r, h, ob, vb);
                           },
                           fluid_field, il));
    wtb = std::fmax(wtb, check_base_adj<TorM>(
                             M, 7, t_comp, 3, nullptr, 0, TOR_NV,
                             [](const Tab& T, const WS& w, int i, double r, double h,
                                const double* ob, double* vb) {
                               tor_base_adj(T, w, i, r, h, ob, vb);
                             },
                             tor_field, il));
  }
  printf("base_adj_solid %.3e\n", ws);
  printf("base_adj_fluid %.3e\n", wf);
  printf("base_adj_tor %.3e\n", wtb);

  check_step<SolidMinorM, MinorTM, SolidAbar>("solid_minor", 20, s_comp, 8, s_real, 2, sbil,
                                               solid_obf);
  check_step<BvecM, BvecTM, SolidAbar>("bvec", 20, s_comp, 8, s_real, 2, bbil, solid_obf);
  check_step<FluidMinorM, FluidMinorT, FluidAbar>("fluid_minor", 14, f_comp, 4, f_real, 2, fbil,
                                                  fluid_obf);
  check_step<TorM, TorT, TorAbar>("tor", 7, t_comp, 3, nullptr, 0, tbil, tor_obf);
  return main2(M);
}
