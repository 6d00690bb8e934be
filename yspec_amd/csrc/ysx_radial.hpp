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
__device__ inline void rad_sweeps(const Tab& T, int nk, const PS& s, double dr, double rs,
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

// a hook of rad_sweeps that does nothing
struct NoRadHook {
  __device__ inline void operator()(int, int, double, double, const cd*) const {}
};

}  // namespace ysx
