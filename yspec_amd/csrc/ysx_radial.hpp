// ysx_radial.hpp -- the l = 0 radial solve's grid and sweeps
// (rad_solver_wg, src/module_solver.f90:446-561, on the grid of make_grid,
// src/module_model.f90:638-787), shared by the forward (k_rad) and its
// adjoint (k_rad_adj), which replays the same steps through a hook.
#pragma once
#include "ysx_device.hpp"

namespace ysx {

// The grid points of knot interval i (make_grid): up to three runs of
// equally spaced points split at the source and receiver radii.
struct Run3 {
  double base[3], step[3];
  int j0[3], j1[3], n;
  int src, rec;  // run whose last point is the source / receiver (-1: none)
};

__host__ __device__ inline void grid_runs(const Tab& T, int i, double drr, double rs, double rr, Run3& G) {
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

// The two sweeps of rad_solver_wg for pair (l = 0, w) of event R with
// spacing dr: y1 from the centre up to the source point, y2 (from (1, 0) at
// the top point) down to it; y1r / y2r the solutions at the receiver point
// (rec_below: the receiver is at or below the source point, y1r is set;
// else y2r).  hook(sweep, interval, radius, step, y) is called before every
// Cash-Karp step (sweep 0: upward, 1: downward); nsr counts the steps.
template <class G, class Hook>
__host__ __device__ inline void rad_sweeps(const Tab& T, int nk, const PS& s, double dr, double rs,
                                  double rr, const G& grp, const Hook& hook, cd* y1, cd* y2,
                                  cd* y1r, cd* y2r, bool& rec_below, int& nsr) {
  RadRHS f;
  auto up = [&](int i, double x, double h, const cd* y) { hook(0, i, x, h, y); };
  auto down = [&](int i, double x, double h, const cd* y) { hook(1, i, x, h, y); };
  y1r[0] = y1r[1] = C(0.0);
  rec_below = false;
  // --- upward sweep over grid points 1 .. source
  {
    int k = 0;
    double rprev = 0.0;
    int iprev = 1;
    bool done = false;
    for (int i = 1; i <= nk - 1 && !done; i++) {
      Run3 Gr;
      grid_runs(T, i, dr, rs, rr, Gr);
      for (int q = 0; q < Gr.n && !done; q++) {
        for (int jj = Gr.j0[q]; jj <= Gr.j1[q]; jj++) {
          double rg = Gr.base[q] + (double)(jj - 1) * Gr.step[q];
          k++;
          if (k == 2) {
            double rt = fmin(0.1 * dr, 0.1 * rg);
            rad_ystart(T, s, 1, rt, y1);
            yint(T, s, nk, 1, i, rt, rg, y1, f, (int*)nullptr, nsr, grp, up);
          } else if (k > 2) {
            yint(T, s, nk, iprev, i, rprev, rg, y1, f, (int*)nullptr, nsr, grp, up);
          }
          rprev = rg;
          iprev = i;
          bool last = (jj == Gr.j1[q]);
          if (last && q == Gr.rec) {
            rec_below = true;
            y1r[0] = y1[0];
            y1r[1] = y1[1];
          }
          if (last && q == Gr.src) {
            done = true;
            break;
          }
        }
      }
    }
  }
  // --- downward sweep from the top grid point to the source point
  y2[0] = y2r[0] = C(1.0);
  y2[1] = y2r[1] = C(0.0);
  {
    bool first = true, done = false;
    double rprev = 0.0;
    int iprev = nk - 1;
    for (int i = nk - 1; i >= 1 && !done; i--) {
      Run3 Gr;
      grid_runs(T, i, dr, rs, rr, Gr);
      for (int q = Gr.n - 1; q >= 0 && !done; q--) {
        for (int jj = Gr.j1[q]; jj >= Gr.j0[q]; jj--) {
          double rg = Gr.base[q] + (double)(jj - 1) * Gr.step[q];
          if (!first) yint(T, s, nk, iprev, i, rprev, rg, y2, f, (int*)nullptr, nsr, grp, down);
          first = false;
          rprev = rg;
          iprev = i;
          bool last = (jj == Gr.j1[q]);
          if (last && q == Gr.src) {
            done = true;
            break;
          }
          if (last && q == Gr.rec && !rec_below) {
            y2r[0] = y2[0];
            y2r[1] = y2[1];
          }
        }
      }
    }
  }
}

// ------------------------------------------------ per-interval segments
// The steps of rad_sweeps split into independent segments, one per (sweep,
// knot interval): segment (sweep, i) holds the steps the sweep takes inside
// interval i, in its order, and can be enumerated without the others.
// Consecutive grid points of an interval are joined by yint's equal steps
// in either direction; where a sweep passes into the next interval the
// join runs from the last point of one interval to the knot radius that
// starts the next, and yint takes it only in the sweep's own direction:
// upward as a (rounding-length) step in the interval left, downward in the
// interval entered.

// knot interval of the source point (grid_runs' rule r1 < rs <= r2) and the
// topmost interval of nonzero thickness (0: none)
__host__ __device__ inline int rad_src_interval(const Tab& T, int nk, double rs) {
  for (int i = 1; i <= nk - 1; i++) {
    const double r1 = T(F_R, i), r2 = T(F_R, i + 1);
    if (r1 != r2 && rs > r1 && rs <= r2) return i;
  }
  return 0;
}
__host__ __device__ inline int rad_top_interval(const Tab& T, int nk) {
  for (int i = nk - 1; i >= 1; i--)
    if (T(F_R, i) != T(F_R, i + 1)) return i;
  return 0;
}

// yint's equal steps between radii p0 and p1 of interval i: f(x, h) per step
template <class Fn>
__host__ __device__ inline int rad_cell(const Tab& T, const PS& s, int nk, int i, double p0,
                                        double p1, const Fn& f) {
  if (p1 == p0) return 0;
  double drr = step_len(T, s, i, nk);
  const int n = 2 + (int)floor(fabs(p1 - p0) / drr);
  drr = (p1 - p0) / (double)(n - 1);
  for (int j = 1; j <= n - 1; j++) f(p0 + (double)(j - 1) * drr, drr);
  return n - 1;
}

// The steps of segment (sweep, i) (sweep 0 upward from the centre to the
// source point, 1 downward from the top point), f(x, h) per step; returns
// their number.  isrc = rad_src_interval, itop = rad_top_interval; interval
// 1 must have nonzero thickness (the upward sweep starts in it).
template <class Fn>
__host__ __device__ inline int rad_segment(const Tab& T, int nk, const PS& s, double dr,
                                           double rs, double rr, int sweep, int i, int isrc,
                                           int itop, const Fn& f) {
  Run3 G;
  grid_runs(T, i, dr, rs, rr, G);
  if (G.n == 0) return 0;
  int n = 0;
  if (sweep == 0) {
    if (i > isrc) return 0;
    int k = 0;
    double prev = 0.0;
    for (int q = 0; q < G.n; q++)
      for (int jj = G.j0[q]; jj <= G.j1[q]; jj++) {
        const double rg = G.base[q] + (double)(jj - 1) * G.step[q];
        k++;
        if (k == 2 && i == 1)  // the start radius below the second grid point
          n += rad_cell(T, s, nk, i, fmin(0.1 * dr, 0.1 * rg), rg, f);
        else if (k >= 2)
          n += rad_cell(T, s, nk, i, prev, rg, f);
        prev = rg;
        if (jj == G.j1[q] && q == G.src) return n;
      }
    if (prev < T(F_R, i + 1)) n += rad_cell(T, s, nk, i, prev, T(F_R, i + 1), f);
    return n;
  }
  if (i < isrc) return 0;
  bool first = true;
  double prev = T(F_R, i + 1);
  for (int q = G.n - 1; q >= 0; q--)
    for (int jj = G.j1[q]; jj >= G.j0[q]; jj--) {
      const double rg = G.base[q] + (double)(jj - 1) * G.step[q];
      if (!first)
        n += rad_cell(T, s, nk, i, prev, rg, f);
      else if (i != itop && rg < prev)
        n += rad_cell(T, s, nk, i, prev, rg, f);
      first = false;
      prev = rg;
      if (jj == G.j1[q] && q == G.src) return n;
    }
  return n;
}

// a hook of rad_sweeps that does nothing
struct NoRadHook {
  __host__ __device__ inline void operator()(int, int, double, double, const cd*) const {}
};

}  // namespace ysx
