// ysx_adjoint.hpp -- numerics of the discrete adjoint of the radial marches.
//
// The systems are linear, dy/dr = A(r) y, and every march is a fixed
// sequence of Cash-Karp steps (the reference's step rule, frozen).  For a
// row vector x carried back through a step y' = S y (x <- x S), the change
// of a scalar J = Re(x_end . y_end) with the coefficients is
//   dJ = Re sum_stages (w_k . (h dA_k) z_k)
// where z_k is the stage's input vector (a column) and w_k the adjoint of
// its derivative (a row): w_6 = c6 x, ..., the partial sums of the
// transposed step.  With dA linear in the entries a of A, w . dA z =
// sum_a da * abar(a), abar(a) = sum of w_p z_q over the terms of a (the
// holomorphic adjoint: J changes by Re(abar * da)).  Here:
//   * the bilinear forms abar(w, z) of each system's right-hand side;
//   * the map from those entry adjoints to adjoints of the l-independent
//     per-radius base (ysx_march.hpp), per lane;
//   * the adjoint of each base with respect to the spline values at the
//     stage radius and, through the spline weights, to the table entries;
//   * the transposed fluid-minor and toroidal systems;
//   * the transposed Cash-Karp step that applies x <- x S and hands each
//     stage's (w_k, z_k) to a callback.
#pragma once
#include "ysx_march.hpp"

namespace ysx {

// ------------------------------------------------------ bilinear forms
// Entry adjoints of the solid coefficient matrix (SolidA).  The real
// entries get complex adjoints too; only their real parts matter.
struct SolidAbar {
  cd t11, t12, t21, t22, t31, t33, k11, k22, k33, s11, s12, s13, s22, s23;
};
__host__ __device__ inline void zero(SolidAbar& B) {
  B.t11 = B.t12 = B.t21 = B.t22 = B.t31 = B.t33 = B.k11 = B.k22 = B.k33 = B.s11 = B.s12 =
      B.s13 = B.s22 = B.s23 = C(0.0);
}

// w . d for sph_minors_solid_derivs (src/module_int.f90:818-871) as written
// in solid_minor_derivs: B += abar(w, z)
__host__ __device__ inline void solid_minor_bil(const cd* w, const cd* z, SolidAbar& B) {
  auto W = [&](int p, int q) { return cmul_(w[p], z[q]); };
  B.t11 = B.t11 + ((W(0, 0) - W(13, 13)) + (W(3, 3) - W(12, 12)) + (W(5, 5) - W(11, 11)) +
                   (W(8, 8) - W(9, 9)) + (W(2, 2) - W(10, 10)));
  B.t22 = B.t22 + ((W(0, 0) - W(13, 13)) + (W(3, 3) - W(12, 12)) - (W(5, 5) - W(11, 11)) -
                   (W(8, 8) - W(9, 9)) + (W(1, 1) - W(6, 6)));
  B.t33 = B.t33 + ((W(0, 0) - W(13, 13)) - (W(3, 3) - W(12, 12)) + (W(5, 5) - W(11, 11)) -
                   (W(8, 8) - W(9, 9)) + (W(4, 4) - W(7, 7)));
  B.t12 = B.t12 + ((W(4, 9) - W(2, 1)) + (W(6, 10) + W(7, 11)) - 2.0 * (W(5, 4) + W(8, 7)));
  B.t21 = B.t21 + ((W(10, 6) - W(1, 2)) - (W(4, 5) + W(7, 8)) + 2.0 * (W(9, 4) + W(11, 7)));
  B.t31 = B.t31 + ((W(4, 2) - W(1, 3)) + (W(6, 8) - W(10, 7)) + 2.0 * (W(12, 6) - W(9, 1)));
  B.k11 = B.k11 + ((W(0, 9) - W(2, 10)) - (W(3, 11) + W(5, 12)) + W(8, 13));
  B.k22 = B.k22 + ((W(3, 8) - W(0, 5)) - (W(1, 6) + W(11, 13)) + W(9, 12));
  B.k33 = B.k33 + ((W(0, 3) - W(4, 7)) - (W(5, 8) + W(9, 11)) + W(12, 13));
  B.s11 = B.s11 + ((W(9, 0) - W(10, 2)) - (W(11, 3) + W(12, 5)) + W(13, 8));
  B.s12 = B.s12 + ((W(6, 2) - W(4, 0)) + (W(7, 3) + W(10, 1)) + 2.0 * (W(12, 4) - W(13, 7)));
  B.s13 = B.s13 + ((W(1, 0) + W(6, 5)) - (W(7, 2) + W(10, 4)) + 2.0 * (W(11, 1) + W(13, 6)));
  B.s22 = B.s22 + ((W(8, 3) - W(5, 0)) - (W(6, 1) + W(13, 11)) + W(12, 9));
  B.s23 = B.s23 + ((W(2, 0) - W(6, 4)) - (W(7, 1) + W(10, 9)) + 2.0 * (W(13, 10) - W(8, 2)));
}

// w . d for sph_bvec_solid_derivs (src/module_int.f90:1016-1057) as written
// in bvec_derivs: B += abar(w, z)
__host__ __device__ inline void bvec_bil(const cd* w, const cd* z, SolidAbar& B) {
  auto W = [&](int p, int q) { return cmul_(w[p], z[q]); };
  const cd z38m = z[7] - z[2], z2p8 = 2.0 * z[2] + z[7], z3p28 = z[2] + 2.0 * z[7];
  B.s11 = B.s11 + ((W(0, 6) + W(1, 9)) - (W(3, 11) + W(4, 12)));
  B.s12 = B.s12 + (cmul_(w[0], z38m) + (W(1, 10) + W(2, 11)) + (W(5, 9) - W(4, 13)) -
                   (W(7, 11) + W(8, 12)));
  B.s13 = B.s13 + ((W(0, 8) - cmul_(w[1], z2p8)) + (W(2, 12) + W(3, 13)) -
                   (W(5, 6) + W(10, 11)));
  B.s22 = B.s22 + ((W(5, 10) - W(0, 3)) + (W(6, 11) - W(8, 13)));
  B.s23 = B.s23 + ((W(6, 12) - W(0, 4)) - (W(1, 3) + cmul_(w[5], z3p28)) +
                   (W(7, 13) + W(9, 11)));
  B.t11 = B.t11 + ((W(6, 6) - W(0, 0)) + (W(9, 9) - W(1, 1)) - (W(3, 3) + W(4, 4)) +
                   (W(11, 11) + W(12, 12)));
  B.t22 = B.t22 + ((W(3, 3) - W(0, 0)) - (W(5, 5) + W(6, 6)) + (W(10, 10) - W(8, 8)) +
                   (W(11, 11) + W(13, 13)));
  B.t33 = B.t33 + ((W(4, 4) - W(1, 1)) - (W(5, 5) + W(9, 9)) + (W(8, 8) - W(10, 10)) +
                   (W(12, 12) + W(13, 13)));
  B.t12 = B.t12 + ((W(2, 3) - W(5, 1)) + cmul_(w[6], z38m) - (W(7, 3) + W(8, 4)) +
                   (W(9, 10) + W(12, 13)));
  B.t21 = B.t21 + ((W(7, 6) - W(1, 5)) - (W(2, 6) + cmul_(w[3], z38m)) - W(4, 8) +
                   (W(10, 9) + W(13, 12)));
  B.t31 = B.t31 + ((W(0, 5) - W(2, 9)) + (cmul_(w[4], z2p8) - W(3, 10)) +
                   (W(8, 6) - W(13, 11)));
  B.k11 = B.k11 + ((W(6, 0) + W(9, 1)) - (W(11, 3) + W(12, 4)));
  B.k22 = B.k22 + ((W(10, 5) - W(3, 0)) + (W(11, 6) - W(13, 8)));
  B.k33 = B.k33 + ((W(12, 9) - W(4, 1)) + (W(13, 10) - W(8, 5)));
}

// The same forms with z streamed: zl(q) returns component q of z (a tape
// read), and only the entries solid_ob reads are formed (t21, t22, t33 and
// k33 do not depend on the base; s13, s23 and t31 enter through their real
// parts).  Each distinct product w_p z_q is taken once.  tests/
// adjoint_check.cpp checks them against the full forms.
template <class ZL>
__host__ __device__ inline void solid_minor_bil_s(const cd* w, ZL&& zl, SolidAbar& B) {
  {
    const cd z = zl(0);
    { const cd P = cmul_(w[0], z); B.t11 = B.t11 + P; }
    { const double P = w[1].x * z.x - w[1].y * z.y; B.s13.x = B.s13.x + P; }
    { const double P = w[2].x * z.x - w[2].y * z.y; B.s23.x = B.s23.x + P; }
    { const cd P = cmul_(w[4], z); B.s12 = B.s12 - P; }
    { const cd P = cmul_(w[5], z); B.s22 = B.s22 - P; }
    { const cd P = cmul_(w[9], z); B.s11 = B.s11 + P; }
  }
  {
    const cd z = zl(1);
    { const cd P = cmul_(w[2], z); B.t12 = B.t12 - P; }
    { const cd P = cmul_(w[6], z); B.s22 = B.s22 - P; }
    { const double P = w[7].x * z.x - w[7].y * z.y; B.s23.x = B.s23.x - P; }
    { const double P = w[9].x * z.x - w[9].y * z.y; B.t31.x = B.t31.x + -2.0 * P; }
    { const cd P = cmul_(w[10], z); B.s12 = B.s12 + P; }
    { const double P = w[11].x * z.x - w[11].y * z.y; B.s13.x = B.s13.x + 2.0 * P; }
  }
  {
    const cd z = zl(2);
    { const cd P = cmul_(w[2], z); B.t11 = B.t11 + P; }
    { const double P = w[4].x * z.x - w[4].y * z.y; B.t31.x = B.t31.x + P; }
    { const cd P = cmul_(w[6], z); B.s12 = B.s12 + P; }
    { const double P = w[7].x * z.x - w[7].y * z.y; B.s13.x = B.s13.x - P; }
    { const double P = w[8].x * z.x - w[8].y * z.y; B.s23.x = B.s23.x + -2.0 * P; }
    { const cd P = cmul_(w[10], z); B.s11 = B.s11 - P; }
  }
  {
    const cd z = zl(3);
    { const double P = w[1].x * z.x - w[1].y * z.y; B.t31.x = B.t31.x - P; }
    { const cd P = cmul_(w[3], z); B.t11 = B.t11 + P; }
    { const cd P = cmul_(w[7], z); B.s12 = B.s12 + P; }
    { const cd P = cmul_(w[8], z); B.s22 = B.s22 + P; }
    { const cd P = cmul_(w[11], z); B.s11 = B.s11 - P; }
  }
  {
    const cd z = zl(4);
    { const cd P = cmul_(w[5], z); B.t12 = B.t12 + -2.0 * P; }
    { const double P = w[6].x * z.x - w[6].y * z.y; B.s23.x = B.s23.x - P; }
    { const double P = w[10].x * z.x - w[10].y * z.y; B.s13.x = B.s13.x - P; }
    { const cd P = cmul_(w[12], z); B.s12 = B.s12 + 2.0 * P; }
  }
  {
    const cd z = zl(5);
    { const cd P = cmul_(w[0], z); B.k22 = B.k22 - P; }
    { const cd P = cmul_(w[5], z); B.t11 = B.t11 + P; }
    { const double P = w[6].x * z.x - w[6].y * z.y; B.s13.x = B.s13.x + P; }
    { const cd P = cmul_(w[12], z); B.s11 = B.s11 - P; }
  }
  {
    const cd z = zl(6);
    { const cd P = cmul_(w[1], z); B.k22 = B.k22 - P; }
    { const double P = w[12].x * z.x - w[12].y * z.y; B.t31.x = B.t31.x + 2.0 * P; }
    { const double P = w[13].x * z.x - w[13].y * z.y; B.s13.x = B.s13.x + 2.0 * P; }
  }
  {
    const cd z = zl(7);
    { const cd P = cmul_(w[8], z); B.t12 = B.t12 + -2.0 * P; }
    { const double P = w[10].x * z.x - w[10].y * z.y; B.t31.x = B.t31.x - P; }
    { const cd P = cmul_(w[13], z); B.s12 = B.s12 + -2.0 * P; }
  }
  {
    const cd z = zl(8);
    { const cd P = cmul_(w[3], z); B.k22 = B.k22 + P; }
    { const double P = w[6].x * z.x - w[6].y * z.y; B.t31.x = B.t31.x + P; }
    { const cd P = cmul_(w[8], z); B.t11 = B.t11 + P; }
    { const cd P = cmul_(w[13], z); B.s11 = B.s11 + P; }
  }
  {
    const cd z = zl(9);
    { const cd P = cmul_(w[0], z); B.k11 = B.k11 + P; }
    { const cd P = cmul_(w[4], z); B.t12 = B.t12 + P; }
    { const cd P = cmul_(w[9], z); B.t11 = B.t11 - P; }
    { const double P = w[10].x * z.x - w[10].y * z.y; B.s23.x = B.s23.x - P; }
    { const cd P = cmul_(w[12], z); B.s22 = B.s22 + P; }
  }
  {
    const cd z = zl(10);
    { const cd P = cmul_(w[2], z); B.k11 = B.k11 - P; }
    { const cd P = cmul_(w[6], z); B.t12 = B.t12 + P; }
    { const cd P = cmul_(w[10], z); B.t11 = B.t11 - P; }
    { const double P = w[13].x * z.x - w[13].y * z.y; B.s23.x = B.s23.x + 2.0 * P; }
  }
  {
    const cd z = zl(11);
    { const cd P = cmul_(w[3], z); B.k11 = B.k11 - P; }
    { const cd P = cmul_(w[7], z); B.t12 = B.t12 + P; }
    { const cd P = cmul_(w[11], z); B.t11 = B.t11 - P; }
    { const cd P = cmul_(w[13], z); B.s22 = B.s22 - P; }
  }
  {
    const cd z = zl(12);
    { const cd P = cmul_(w[5], z); B.k11 = B.k11 - P; }
    { const cd P = cmul_(w[9], z); B.k22 = B.k22 + P; }
    { const cd P = cmul_(w[12], z); B.t11 = B.t11 - P; }
  }
  {
    const cd z = zl(13);
    { const cd P = cmul_(w[8], z); B.k11 = B.k11 + P; }
    { const cd P = cmul_(w[11], z); B.k22 = B.k22 - P; }
    { const cd P = cmul_(w[13], z); B.t11 = B.t11 - P; }
  }
}

template <class ZL>
__host__ __device__ inline void bvec_bil_s(const cd* w, ZL&& zl, SolidAbar& B) {
  {
    const cd z = zl(0);
    { const cd P = cmul_(w[0], z); B.t11 = B.t11 - P; }
    { const cd P = cmul_(w[3], z); B.k22 = B.k22 - P; }
    { const cd P = cmul_(w[6], z); B.k11 = B.k11 + P; }
  }
  {
    const cd z = zl(1);
    { const cd P = cmul_(w[1], z); B.t11 = B.t11 - P; }
    { const cd P = cmul_(w[5], z); B.t12 = B.t12 - P; }
    { const cd P = cmul_(w[9], z); B.k11 = B.k11 + P; }
  }
  {
    const cd z = zl(2);
    { const cd P = cmul_(w[0], z); B.s12 = B.s12 - P; }
    { const double P = w[1].x * z.x - w[1].y * z.y; B.s13.x = B.s13.x + -2.0 * P; }
    { const double P = w[4].x * z.x - w[4].y * z.y; B.t31.x = B.t31.x + 2.0 * P; }
    { const double P = w[5].x * z.x - w[5].y * z.y; B.s23.x = B.s23.x - P; }
    { const cd P = cmul_(w[6], z); B.t12 = B.t12 - P; }
  }
  {
    const cd z = zl(3);
    { const cd P = cmul_(w[0], z); B.s22 = B.s22 - P; }
    { const double P = w[1].x * z.x - w[1].y * z.y; B.s23.x = B.s23.x - P; }
    { const cd P = cmul_(w[2], z); B.t12 = B.t12 + P; }
    { const cd P = cmul_(w[3], z); B.t11 = B.t11 - P; }
    { const cd P = cmul_(w[7], z); B.t12 = B.t12 - P; }
    { const cd P = cmul_(w[11], z); B.k11 = B.k11 - P; }
  }
  {
    const cd z = zl(4);
    { const double P = w[0].x * z.x - w[0].y * z.y; B.s23.x = B.s23.x - P; }
    { const cd P = cmul_(w[4], z); B.t11 = B.t11 - P; }
    { const cd P = cmul_(w[8], z); B.t12 = B.t12 - P; }
    { const cd P = cmul_(w[12], z); B.k11 = B.k11 - P; }
  }
  {
    const cd z = zl(5);
    { const double P = w[0].x * z.x - w[0].y * z.y; B.t31.x = B.t31.x + P; }
    { const cd P = cmul_(w[10], z); B.k22 = B.k22 + P; }
  }
  {
    const cd z = zl(6);
    { const cd P = cmul_(w[0], z); B.s11 = B.s11 + P; }
    { const double P = w[5].x * z.x - w[5].y * z.y; B.s13.x = B.s13.x - P; }
    { const cd P = cmul_(w[6], z); B.t11 = B.t11 + P; }
    { const double P = w[8].x * z.x - w[8].y * z.y; B.t31.x = B.t31.x + P; }
    { const cd P = cmul_(w[11], z); B.k22 = B.k22 + P; }
  }
  {
    const cd z = zl(7);
    { const cd P = cmul_(w[0], z); B.s12 = B.s12 + P; }
    { const double P = w[1].x * z.x - w[1].y * z.y; B.s13.x = B.s13.x - P; }
    { const double P = w[4].x * z.x - w[4].y * z.y; B.t31.x = B.t31.x + P; }
    { const double P = w[5].x * z.x - w[5].y * z.y; B.s23.x = B.s23.x + -2.0 * P; }
    { const cd P = cmul_(w[6], z); B.t12 = B.t12 + P; }
  }
  {
    const cd z = zl(8);
    { const double P = w[0].x * z.x - w[0].y * z.y; B.s13.x = B.s13.x + P; }
    { const cd P = cmul_(w[13], z); B.k22 = B.k22 - P; }
  }
  {
    const cd z = zl(9);
    { const cd P = cmul_(w[1], z); B.s11 = B.s11 + P; }
    { const double P = w[2].x * z.x - w[2].y * z.y; B.t31.x = B.t31.x - P; }
    { const cd P = cmul_(w[5], z); B.s12 = B.s12 + P; }
    { const cd P = cmul_(w[9], z); B.t11 = B.t11 + P; }
  }
  {
    const cd z = zl(10);
    { const cd P = cmul_(w[1], z); B.s12 = B.s12 + P; }
    { const double P = w[3].x * z.x - w[3].y * z.y; B.t31.x = B.t31.x - P; }
    { const cd P = cmul_(w[5], z); B.s22 = B.s22 + P; }
    { const cd P = cmul_(w[9], z); B.t12 = B.t12 + P; }
  }
  {
    const cd z = zl(11);
    { const cd P = cmul_(w[2], z); B.s12 = B.s12 + P; }
    { const cd P = cmul_(w[3], z); B.s11 = B.s11 - P; }
    { const cd P = cmul_(w[6], z); B.s22 = B.s22 + P; }
    { const cd P = cmul_(w[7], z); B.s12 = B.s12 - P; }
    { const double P = w[9].x * z.x - w[9].y * z.y; B.s23.x = B.s23.x + P; }
    { const double P = w[10].x * z.x - w[10].y * z.y; B.s13.x = B.s13.x - P; }
    { const cd P = cmul_(w[11], z); B.t11 = B.t11 + P; }
    { const double P = w[13].x * z.x - w[13].y * z.y; B.t31.x = B.t31.x - P; }
  }
  {
    const cd z = zl(12);
    { const double P = w[2].x * z.x - w[2].y * z.y; B.s13.x = B.s13.x + P; }
    { const cd P = cmul_(w[4], z); B.s11 = B.s11 - P; }
    { const double P = w[6].x * z.x - w[6].y * z.y; B.s23.x = B.s23.x + P; }
    { const cd P = cmul_(w[8], z); B.s12 = B.s12 - P; }
    { const cd P = cmul_(w[12], z); B.t11 = B.t11 + P; }
  }
  {
    const cd z = zl(13);
    { const double P = w[3].x * z.x - w[3].y * z.y; B.s13.x = B.s13.x + P; }
    { const cd P = cmul_(w[4], z); B.s12 = B.s12 - P; }
    { const double P = w[7].x * z.x - w[7].y * z.y; B.s23.x = B.s23.x + P; }
    { const cd P = cmul_(w[8], z); B.s22 = B.s22 - P; }
    { const cd P = cmul_(w[12], z); B.t12 = B.t12 + P; }
  }
}

// Per-lane map from the solid entry adjoints to adjoints of the stored
// base doubles that depend on the model (SolidB::amat builds A from the
// base b and the lane's degree): ob[0..15] the holomorphic adjoints of the
// complex base entries t11, cccon, k22, s11, fc/r, q12, a22, gc (re, im
// pairs, bases 2..17), ob[16] of rho (base 1), ob[17] of t31 (base 18).
constexpr int SOLID_OB = 18;
__host__ __device__ inline void solid_ob(const SolidAbar& B, const LS& L, const double* b,
                                         double* ob) {
  const double hri = b[0];
  ob[0] = B.t11.x; ob[1] = B.t11.y;
  ob[2] = B.k11.x; ob[3] = B.k11.y;
  ob[4] = B.k22.x; ob[5] = B.k22.y;
  ob[6] = B.s11.x; ob[7] = B.s11.y;
  ob[8] = L.zeta * B.t12.x; ob[9] = L.zeta * B.t12.y;
  ob[10] = L.zeta * B.s12.x; ob[11] = L.zeta * B.s12.y;
  ob[12] = B.s22.x; ob[13] = B.s22.y;
  ob[14] = L.zeta2 * B.s22.x; ob[15] = L.zeta2 * B.s22.y;
  ob[16] = hri * (L.zeta * B.s23.x - L.llp1 * B.s13.x);
  ob[17] = B.t31.x;
}

// Entry adjoints of the fluid minor system (FluidMinorM's entries)
struct FluidAbar {
  cd t11, t12, t21, t22, k11, k22, s11, s12, s22;
};
__host__ __device__ inline void zero(FluidAbar& B) {
  B.t11 = B.t12 = B.t21 = B.t22 = B.k11 = B.k22 = B.s11 = B.s12 = B.s22 = C(0.0);
}
// w . d for sph_minors_fluid_derivs (src/module_int.f90:941-967) as written
// in FluidMinorM
__host__ __device__ inline void fluid_minor_bil(const cd* w, const cd* z, FluidAbar& B) {
  auto W = [&](int p, int q) { return cmul_(w[p], z[q]); };
  B.t11 = B.t11 + ((W(0, 0) - W(4, 4)) + (W(2, 2) - W(3, 3)));
  B.t22 = B.t22 + ((W(0, 0) - W(4, 4)) - (W(2, 2) - W(3, 3)));
  B.t12 = B.t12 + (W(1, 3) - 2.0 * W(2, 1));
  B.t21 = B.t21 + (2.0 * W(3, 1) - W(1, 2));
  B.k11 = B.k11 + (W(2, 4) - W(0, 3));
  B.k22 = B.k22 + (W(0, 2) - W(3, 4));
  B.s11 = B.s11 + (W(4, 2) - W(3, 0));
  B.s12 = B.s12 + (W(1, 0) - 2.0 * W(4, 1));
  B.s22 = B.s22 + (W(2, 0) - W(4, 3));
}
// fluid_minor_bil with z streamed, forming the entries fluid_ob reads
// (t21 through its real part)
template <class ZL>
__host__ __device__ inline void fluid_minor_bil_s(const cd* w, ZL&& zl, FluidAbar& B) {
  {
    const cd z = zl(0);
    { const cd P = cmul_(w[0], z); B.t11 = B.t11 + P; }
    { const cd P = cmul_(w[1], z); B.s12 = B.s12 + P; }
    { const cd P = cmul_(w[2], z); B.s22 = B.s22 + P; }
    { const cd P = cmul_(w[3], z); B.s11 = B.s11 - P; }
  }
  {
    const cd z = zl(1);
    { const double P = w[3].x * z.x - w[3].y * z.y; B.t21.x = B.t21.x + 2.0 * P; }
    { const cd P = cmul_(w[4], z); B.s12 = B.s12 + -2.0 * P; }
  }
  {
    const cd z = zl(2);
    { const double P = w[1].x * z.x - w[1].y * z.y; B.t21.x = B.t21.x - P; }
    { const cd P = cmul_(w[2], z); B.t11 = B.t11 + P; }
    { const cd P = cmul_(w[4], z); B.s11 = B.s11 + P; }
  }
  {
    const cd z = zl(3);
    { const cd P = cmul_(w[0], z); B.k11 = B.k11 - P; }
    { const cd P = cmul_(w[3], z); B.t11 = B.t11 - P; }
    { const cd P = cmul_(w[4], z); B.s22 = B.s22 - P; }
  }
  {
    const cd z = zl(4);
    { const cd P = cmul_(w[2], z); B.k11 = B.k11 + P; }
    { const cd P = cmul_(w[4], z); B.t11 = B.t11 - P; }
  }
}

// fluid base adjoints per lane: ob[0..1] fa, [2..3] cccon, [4..5] s11a,
// [6..7] s11b (holomorphic), ob[8] rho, ob[9] t21 (base 12); fb = h/(r w)^2
// does not depend on the model
constexpr int FLUID_OB = 10;
__host__ __device__ inline void fluid_ob(const FluidAbar& B, const LS& L, const double* b,
                                         double* ob) {
  const double ri = b[0], rrho = b[1], z2 = L.zeta2;
  const cd fa = C(b[2], b[3]), fb = C(b[4], b[5]);
  const cd fab = z2 * B.t11 + (z2 * rrho) * B.s12;
  ob[0] = fab.x; ob[1] = fab.y;
  ob[2] = B.k11.x; ob[3] = B.k11.y;
  ob[4] = z2 * B.s11.x; ob[5] = z2 * B.s11.y;
  ob[6] = B.s11.x; ob[7] = B.s11.y;
  cd rb = cmul_(B.k11, (z2 / (rrho * rrho)) * fb) + cmul_(B.s12, z2 * fa - L.llp1 * ri) +
          cmul_(B.s22, z2 * fb);
  ob[8] = rb.x;
  ob[9] = B.t21.x;
}

// toroidal (TorM): entry adjoints of k and sv, and of t for checks
struct TorAbar {
  cd t, k, sv;
};
__host__ __device__ inline void zero(TorAbar& B) { B.t = B.k = B.sv = C(0.0); }
__host__ __device__ inline void tor_bil(const cd* w, const cd* z, TorAbar& B) {
  B.t = B.t + (cmul_(w[0], z[0]) - cmul_(w[1], z[1]));
  B.k = B.k + cmul_(w[0], z[1]);
  B.sv = B.sv + cmul_(w[1], z[0]);
}
// tor_bil with z streamed (the entries tor_ob reads)
template <class ZL>
__host__ __device__ inline void tor_bil_s(const cd* w, ZL&& zl, TorAbar& B) {
  const cd z0 = zl(0), z1 = zl(1);
  B.k = B.k + cmul_(w[0], z1);
  B.sv = B.sv + cmul_(w[1], z0);
}
// toroidal base adjoints per lane: ob[0..1] k, [2..3] sv0, [4..5] mw
constexpr int TOR_OB = 6;
__host__ __device__ inline void tor_ob(const TorAbar& B, const LS& L, const double*, double* ob) {
  ob[0] = B.k.x; ob[1] = B.k.y;
  ob[2] = L.zeta2m2 * B.sv.x; ob[3] = L.zeta2m2 * B.sv.y;
  ob[4] = B.sv.x; ob[5] = B.sv.y;
}

// ------------------------------------------------------ transposed systems
// e = A^T x of sph_minors_fluid_derivs (src/module_int.f90:941-967) with
// the entries of FluidMinorM
struct FluidMinorT {
  static constexpr int N = 5, BS = 14;
  __host__ __device__ static void base(const Tab& T, const WS& w, int il, double rr, double h,
                                       double* o) {
    FluidMinorM::base(T, w, il, rr, h, o);
  }
  template <class Sink>
  __host__ __device__ inline void stream(const double* b, const LS& L, const WS&, const cd* x,
                                         Sink&& sink) const {
    double ri = b[0], rrho = b[1];
    cd fa = C(b[2], b[3]), fb = C(b[4], b[5]), cccon = C(b[6], b[7]);
    double z2 = L.zeta2;
    cd t11 = z2 * fa - ri, t12 = z2 * fb;
    double t21 = b[12], t22 = -L.lld * ri;
    cd k11 = cccon - (z2 / rrho) * fb;
    double k22 = b[13];
    cd s11 = z2 * C(b[8], b[9]) + C(b[10], b[11]);
    cd s12 = (z2 * rrho) * fa - L.llp1 * rrho * ri;
    cd s22 = (z2 * rrho) * fb;
    cd tp = t11 + t22, tm = t11 - t22;
    sink(0, cmsb(cmad(cmad(cmul_(tp, x[0]), s12, x[1]), s22, x[2]), s11, x[3]));
    sink(1, 2.0 * cmsb(cmsb(rmul(t21, x[3]), t12, x[2]), s12, x[4]));
    sink(2, cmad(cmad(rmsb(rmul(k22, x[0]), t21, x[1]), tm, x[2]), s11, x[4]));
    sink(3, cmsb(cmsb(cmad(-k11 * x[0], t12, x[1]), tm, x[3]), s22, x[4]));
    sink(4, cmsb(rmsb(cmul_(k11, x[2]), k22, x[3]), tp, x[4]));
  }
};

// e = A^T x of tor_derivs (src/module_int.f90:168-184) with TorM's entries
struct TorT {
  static constexpr int N = 2, BS = 7;
  __host__ __device__ static void base(const Tab& T, const WS& w, int il, double rr, double h,
                                       double* o) {
    TorM::base(T, w, il, rr, h, o);
  }
  template <class Sink>
  __host__ __device__ inline void stream(const double* b, const LS& L, const WS&, const cd* x,
                                         Sink&& sink) const {
    double t = 2.0 * b[0];
    cd k = C(b[1], b[2]);
    cd sv = L.zeta2m2 * C(b[3], b[4]) + C(b[5], b[6]);
    sink(0, cmad(rmul(t, x[0]), sv, x[1]));
    sink(1, rmsb(cmul_(k, x[0]), t, x[1]));
  }
};

// ------------------------------------------------- base adjoints (per radius)
// Adjoint of solid_base<true> (ysx_march.hpp; sph_solid_amat,
// src/module_int.f90:500-560) at one stage radius: ob holds the wave's
// summed adjoints of the stored base (layout of solid_ob); vb receives the
// adjoints of the spline values rho, A, C, F, N, L, g, 1/Qmu, xlam, xa2.
constexpr int SOLID_NV = 10;
__host__ __device__ inline void solid_base_adj(const Tab& T, const WS& w, int il, double rr,
                                               double h, const double* ob, double* vb) {
  SplW sw = spl_w(T, il, rr);
  double rrho = spl(T, sw, F_RHO), racon = spl(T, sw, F_ACON), rccon = spl(T, sw, F_CCON),
         rfcon = spl(T, sw, F_FCON), rncon = spl(T, sw, F_NCON), rlcon = spl(T, sw, F_LCON),
         gg = spl(T, sw, F_GRAV), rqmu = spl(T, sw, F_QMU), rxlam = spl(T, sw, F_XLAM),
         rxa2 = spl(T, sw, F_XA2);
  cd qaa = 1.0 + rxa2 * w.lno, qff = 1.0 + rxlam * w.lno, qll = 1.0 + rqmu * w.lno;
  cd cfcon = rfcon * qff;
  cd u = rccon * qaa, cccon = cinv(u);
  cd fc = cfcon * cccon;
  cd u2 = rlcon * qll, k22 = cinv(u2);
  double ri = 1.0 / rr, ri2 = ri * ri;
  // adjoints of the unscaled complex entries (the base stores them times h)
  cd t11b = h * C(ob[0], ob[1]), cccb = h * C(ob[2], ob[3]), k22b = h * C(ob[4], ob[5]),
     s11b = h * C(ob[6], ob[7]), fcrib = h * C(ob[8], ob[9]), q12b = h * C(ob[10], ob[11]),
     a22b = h * C(ob[12], ob[13]), gcb = h * C(ob[14], ob[15]);
  double rhob = ob[16] + ob[17] * h * (-4.0 * w.pibigg);
  // t11 = (1 - 2 fc)/r, fcri = fc/r
  cd fcb = ri * (fcrib - 2.0 * t11b);
  // s11 = mw2r + 4/r^2 (gamma - rgr), q12 = (rgr - 2 gamma)/r^2,
  // a22 = mw2r - 2 cncon/r^2, gc = (gamma + cncon)/r^2
  cd mw2rb = s11b + a22b;
  cd gammab = 4.0 * ri2 * s11b - 2.0 * ri2 * q12b + ri2 * gcb;
  double rgrb = ri2 * (q12b.x - 4.0 * s11b.x);
  cd cnconb = ri2 * (gcb - 2.0 * a22b);
  // gamma = cacon - cncon - cfcon fc
  cd caconb = gammab;
  cnconb = cnconb - gammab;
  cd cfconb = -cmul_(fc, gammab);
  fcb = fcb - cmul_(cfcon, gammab);
  // fc = cfcon cccon; cccon = 1/u; k22 = 1/u2
  cfconb = cfconb + cmul_(cccon, fcb);
  cccb = cccb + cmul_(cfcon, fcb);
  cd ub = -cmul_(cmul_(cccon, cccon), cccb);
  cd u2b = -cmul_(cmul_(k22, k22), k22b);
  // moduli with their dispersion factors
  cd qaab = racon * caconb + rccon * ub;
  cd qffb = rfcon * cfconb;
  cd qllb = rncon * cnconb + rlcon * u2b;
  // mw2r = -rho w^2, rgr = rho g r
  vb[0] = rhob + cmul_(mw2rb, -w.omega2).x + rgrb * gg * rr;
  vb[1] = cmul_(caconb, qaa).x;
  vb[2] = cmul_(ub, qaa).x;
  vb[3] = cmul_(cfconb, qff).x;
  vb[4] = cmul_(cnconb, qll).x;
  vb[5] = cmul_(u2b, qll).x;
  vb[6] = rgrb * rrho * rr;
  vb[7] = cmul_(qllb, w.lno).x;
  vb[8] = cmul_(qffb, w.lno).x;
  vb[9] = cmul_(qaab, w.lno).x;
}
__host__ __device__ inline int solid_field(int v) {
  const int f[SOLID_NV] = {F_RHO, F_ACON, F_CCON, F_FCON, F_NCON,
                           F_LCON, F_GRAV, F_QMU, F_XLAM, F_XA2};
  return f[v];
}

// Adjoint of FluidMinorM::base (sph_fluid_amat, src/module_int.f90:676-715):
// vb = adjoints of rho, C, g, xa2 (ob: layout of fluid_ob)
constexpr int FLUID_NV = 4;
__host__ __device__ inline void fluid_base_adj(const Tab& T, const WS& w, int il, double rr,
                                               double h, const double* ob, double* vb) {
  SplW sw = spl_w(T, il, rr);
  double rrho = spl(T, sw, F_RHO), rccon = spl(T, sw, F_CCON), gg = spl(T, sw, F_GRAV),
         rxa2 = spl(T, sw, F_XA2);
  cd qaa = 1.0 + rxa2 * w.lno;
  cd u = rccon * qaa, cccon = cinv(u);
  double ri = 1.0 / rr, ri2 = ri * ri;
  cd fb = ri2 * w.omegai2;
  cd fab = h * C(ob[0], ob[1]), cccb = h * C(ob[2], ob[3]), s11ab = h * C(ob[4], ob[5]),
     s11bb = h * C(ob[6], ob[7]);
  // fa = g fb; s11a = rho g^2 fb; s11b = -4 rho g / r - rho w^2
  double rhob = ob[8] + ob[9] * h * (-4.0 * w.pibigg);
  rhob += cmul_(s11ab, gg * gg * fb).x + cmul_(s11bb, -4.0 * gg * ri - w.omega2).x;
  double gb = cmul_(fab, fb).x + cmul_(s11ab, 2.0 * rrho * gg * fb).x - 4.0 * rrho * ri * s11bb.x;
  cd ub = -cmul_(cmul_(cccon, cccon), cccb);
  vb[0] = rhob;
  vb[1] = cmul_(ub, qaa).x;
  vb[2] = gb;
  vb[3] = cmul_(rccon * ub, w.lno).x;
}
__host__ __device__ inline int fluid_field(int v) {
  const int f[FLUID_NV] = {F_RHO, F_CCON, F_GRAV, F_XA2};
  return f[v];
}

// Adjoint of TorM::base (tor_amat, src/module_int.f90:134-164): vb =
// adjoints of rho, L, N, 1/Qmu (ob: layout of tor_ob)
constexpr int TOR_NV = 4;
__host__ __device__ inline void tor_base_adj(const Tab& T, const WS& w, int il, double rr,
                                             double h, const double* ob, double* vb) {
  SplW sw = spl_w(T, il, rr);
  double rlcon = spl(T, sw, F_LCON), rncon = spl(T, sw, F_NCON), rqmu = spl(T, sw, F_QMU);
  cd qll = 1.0 + rqmu * w.lno;
  double ri = 1.0 / rr;
  cd k = cinv(rlcon * qll);
  cd kb = h * C(ob[0], ob[1]), sv0b = h * C(ob[2], ob[3]), mwb = h * C(ob[4], ob[5]);
  cd ub = -cmul_(cmul_(k, k), kb);
  cd qllb = rlcon * ub + (ri * ri * rncon) * sv0b;
  vb[0] = cmul_(mwb, -w.omega2).x;
  vb[1] = cmul_(ub, qll).x;
  vb[2] = cmul_(sv0b, (ri * ri) * qll).x;
  vb[3] = cmul_(qllb, w.lno).x;
}
__host__ __device__ inline int tor_field(int v) {
  const int f[TOR_NV] = {F_RHO, F_LCON, F_NCON, F_QMU};
  return f[v];
}

// Spread the adjoint vb of the spline value of field f at radius rr on knot
// interval il onto the interval's table entries: value and second
// derivative at knots il and il + 1 (splint_dis_r, src/module_spline.f90:
// 262-283).  add(field, knot, value) takes each contribution.
template <class Add>
__host__ __device__ inline void spl_adj(const SplW& sw, int f, double vb, Add&& add) {
  if (sw.deg) {
    add(f, sw.lo, vb);
    return;
  }
  add(f, sw.lo, vb * sw.a);
  add(f, sw.lo + 1, vb * sw.b);
  add(f + 1, sw.lo, vb * sw.a3 * sw.hh);
  add(f + 1, sw.lo + 1, vb * sw.b3 * sw.hh);
}

// ------------------------------------------------ forward step with hooks
// One Cash-Karp step y <- S y with the stored stage bases of rk_step_lds /
// rk_step_b (same tableau and stage scaling), in registers, calling
// hook(k, z) with each stage's input z before its derivative is taken:
// the replay of a step whose stage inputs the adjoint pairs with its rows.
template <class F, class Hook>
__host__ __device__ inline void rk_step_hook(const double* sb, const LS& L, const WS& w, cd* y,
                                             const F& f, Hook&& hook) {
  using namespace ck;
  constexpr int N = F::N, BS = F::BS;
  constexpr bool SC = (N == 14);
  cd k1[N], k2[N], k3[N], k4[N], t[N];
  hook(0, y);
  f.stream(sb, L, w, y, [&](int i, cd d) { k1[i] = d; });
  for (int i = 0; i < N; i++) t[i] = rmad(y[i], b21, k1[i]);
  hook(1, t);
  f.stream(sb + BS, L, w, t, [&](int i, cd d) { k2[i] = d; });
  for (int i = 0; i < N; i++) t[i] = rmad(rmad(y[i], b31, k1[i]), b32, k2[i]);
  hook(2, t);
  f.stream(sb + 2 * BS, L, w, t, [&](int i, cd d) { k3[i] = d; });
  for (int i = 0; i < N; i++) t[i] = rmad(rmad(rmad(y[i], b41, k1[i]), b42, k2[i]), b43, k3[i]);
  hook(3, t);
  f.stream(sb + 3 * BS, L, w, t, [&](int i, cd d) { k4[i] = d; });
  for (int i = 0; i < N; i++) {
    cd a1 = k1[i], a2 = k2[i], a3 = k3[i], a4 = k4[i], yi = y[i];
    t[i] = rmad(rmad(rmad(rmad(yi, b51, a1), b52, a2), b53, a3), b54, a4);  // z5
    k2[i] = rmad(rmad(rmad(rmad(yi, b61, a1), b62, a2), b63, a3), b64, a4);
    k1[i] = rmad(rmad(rmad(yi, c1, a1), c3, a3), c4, a4);
  }
  hook(4, t);
  f.stream(sb + 4 * BS, L, w, t, [&](int i, cd d) { k3[i] = SC ? k2[i] + d : rmad(k2[i], b65, d); });
  hook(5, k3);
  f.stream(sb + 5 * BS, L, w, k3, [&](int i, cd d) { y[i] = SC ? k1[i] + d : rmad(k1[i], c6, d); });
}

// ---------------------------------------------------- transposed step
// x <- x S for one Cash-Karp step S (rk_step_lds / rk_step_b: stage bases
// at sb, stride FT::BS; with N = 14 the bases of stages 5 and 6 are stored
// times b65 and c6).  FT is the transposed system.  With stage inputs z_k
// and derivatives k_k = h A_k z_k the step is y' = y + sum c_k k_k,
// z_k = y + sum_{m<k} b_km k_m; its transpose takes the stages in reverse:
// kb_k = c_k x + sum_{m>k} b_mk zb_m, zb_k = (h A_k)^T kb_k, x' = x + sum zb_k.
// stage(k - 1, kb) is called with the row kb of stage k = 6..1 (the adjoint
// of the stage's derivative with respect to its stored base), to be paired
// with that stage's input.  Registers: zb6, the partial sums kb4..kb1 and
// x, which accumulates x'.
template <class FT, class Stage>
__host__ __device__ inline void rk_step_T_reg(const double* sb, const LS& L, const WS& w, cd* x,
                                     const FT& f, Stage&& stage) {
  using namespace ck;
  constexpr int N = FT::N, BS = FT::BS;
  constexpr bool SC = (N == 14);
  cd Z[N], w4[N], w3[N], w2[N], w1[N];
  // stage 6: kb6 = c6 x (x itself for the c6-scaled base); Z = zb6
  if constexpr (SC) {
    stage(5, x);
    f.stream(sb + 5 * BS, L, w, x, [&](int i, cd d) { Z[i] = d; });
  } else {
#pragma unroll
    for (int i = 0; i < N; i++) w1[i] = c6 * x[i];
    stage(5, w1);
    f.stream(sb + 5 * BS, L, w, w1, [&](int i, cd d) { Z[i] = d; });
  }
  // stage 5: kb5 = b65 zb6 (zb6 itself for the b65-scaled base); the sink
  // receives zb5 and starts the partial sums
  auto sink5 = [&](int i, cd d) {
    const cd xi = x[i], zi = Z[i];
    w4[i] = rmad(rmad(rmul(c4, xi), b54, d), b64, zi);
    w3[i] = rmad(rmad(rmul(c3, xi), b53, d), b63, zi);
    w2[i] = rmad(rmul(b52, d), b62, zi);
    w1[i] = rmad(rmad(rmul(c1, xi), b51, d), b61, zi);
    x[i] = xi + d + zi;
  };
  if constexpr (SC) {
    stage(4, Z);
    f.stream(sb + 4 * BS, L, w, Z, sink5);
  } else {
    cd A[N];
#pragma unroll
    for (int i = 0; i < N; i++) A[i] = b65 * Z[i];
    stage(4, A);
    f.stream(sb + 4 * BS, L, w, A, sink5);
  }
  stage(3, w4);
  f.stream(sb + 3 * BS, L, w, w4, [&](int i, cd d) {
    w3[i] = rmad(w3[i], b43, d);
    w2[i] = rmad(w2[i], b42, d);
    w1[i] = rmad(w1[i], b41, d);
    x[i] = x[i] + d;
  });
  stage(2, w3);
  f.stream(sb + 2 * BS, L, w, w3, [&](int i, cd d) {
    w2[i] = rmad(w2[i], b32, d);
    w1[i] = rmad(w1[i], b31, d);
    x[i] = x[i] + d;
  });
  stage(1, w2);
  f.stream(sb + BS, L, w, w2, [&](int i, cd d) {
    w1[i] = rmad(w1[i], b21, d);
    x[i] = x[i] + d;
  });
  stage(0, w1);
  f.stream(sb, L, w, w1, [&](int i, cd d) { x[i] = x[i] + d; });
}

// ------------------------------------------------ starting-solution tangents
// Forward-mode tangents of the starting solutions with respect to the K
// spline values they read at the start radius (sph_solid_start,
// sph_fluid_start, tor_ystart, src/module_int.f90:1192-1270, 1346-1405,
// 213-250; bfs/bfs_dum, src/module_function.f90:497-581): each complex
// quantity carries its value and its derivative along each direction.
// The iteration counts of the continued fraction and the series follow the
// values, as in the forward.
template <int K>
struct Dual {
  cd v;
  cd d[K];
};
template <int K>
__host__ __device__ inline Dual<K> dconst(cd v) {
  Dual<K> r;
  r.v = v;
  for (int k = 0; k < K; k++) r.d[k] = C(0.0);
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> dvar(double v, int k0) {
  Dual<K> r = dconst<K>(C(v));
  r.d[k0] = C(1.0);
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v + b.v;
  for (int k = 0; k < K; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v - b.v;
  for (int k = 0; k < K; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r;
  r.v = -a.v;
  for (int k = 0; k < K; k++) r.d[k] = -a.d[k];
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v * b.v;
  for (int k = 0; k < K; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> operator*(double s, const Dual<K>& a) {
  Dual<K> r;
  r.v = s * a.v;
  for (int k = 0; k < K; k++) r.d[k] = s * a.d[k];
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> operator*(const Dual<K>& a, double s) { return s * a; }
template <int K>
__host__ __device__ inline Dual<K> operator*(cd s, const Dual<K>& a) {
  Dual<K> r;
  r.v = s * a.v;
  for (int k = 0; k < K; k++) r.d[k] = s * a.d[k];
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> operator+(const Dual<K>& a, double s) {
  Dual<K> r = a;
  r.v = a.v + s;
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> operator+(double s, const Dual<K>& a) { return a + s; }
template <int K>
__host__ __device__ inline Dual<K> operator+(const Dual<K>& a, cd s) {
  Dual<K> r = a;
  r.v = a.v + s;
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> operator+(cd s, const Dual<K>& a) { return a + s; }
template <int K>
__host__ __device__ inline Dual<K> operator-(const Dual<K>& a, double s) { return a + (-s); }
template <int K>
__host__ __device__ inline Dual<K> operator-(double s, const Dual<K>& a) { return (-a) + s; }
template <int K>
__host__ __device__ inline Dual<K> dinv(const Dual<K>& b) {
  Dual<K> r;
  r.v = cinv(b.v);
  cd r2 = -(r.v * r.v);
  for (int k = 0; k < K; k++) r.d[k] = r2 * b.d[k];
  return r;
}
template <int K>
__host__ __device__ inline Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
  return a * dinv(b);
}
template <int K>
__host__ __device__ inline Dual<K> operator/(double s, const Dual<K>& b) {
  return s * dinv(b);
}
template <int K>
__host__ __device__ inline Dual<K> operator/(const Dual<K>& a, double s) {
  return (1.0 / s) * a;
}
template <int K>
__host__ __device__ inline Dual<K> dsqrt(const Dual<K>& a) {
  Dual<K> r;
  r.v = csqrt_(a.v);
  cd h = (r.v.x == 0.0 && r.v.y == 0.0) ? C(0.0) : cinv(2.0 * r.v);
  for (int k = 0; k < K; k++) r.d[k] = h * a.d[k];
  return r;
}
// |a| as a real-valued dual
template <int K>
__host__ __device__ inline Dual<K> dabs(const Dual<K>& a) {
  Dual<K> r;
  double m = cabs_(a.v);
  r.v = C(m);
  for (int k = 0; k < K; k++)
    r.d[k] = C(m == 0.0 ? 0.0 : (a.v.x * a.d[k].x + a.v.y * a.d[k].y) / m);
  return r;
}

// bfs_ratio with tangents
template <int K>
__host__ __device__ inline Dual<K> bfs_ratio_d(int l, const Dual<K>& xsq) {
  const double eps = 1.e-9;
  Dual<K> fp;
  if (xsq.v.x > 0.0) {
    Dual<K> x = dsqrt(xsq);
    int lp1 = l + 1;
    Dual<K> rx = 2.0 / x;
    double nu = lp1 - 0.5;
    Dual<K> rj = nu * rx;
    rx = -rx;
    Dual<K> denom = (nu + 1.0) * rx;
    Dual<K> numer = denom + 1.0 / rj;
    rj = rj * numer / denom;
    int nm1 = 1;
    for (int it = 0; it < 100000; it++) {
      nm1++;
      rx = -rx;
      Dual<K> a3 = (nu + (double)nm1) * rx;
      denom = a3 + 1.0 / denom;
      numer = a3 + 1.0 / numer;
      Dual<K> ratio = numer / denom;
      rj = rj * ratio;
      if (fabs(cabs_(ratio.v) - 1.0) <= eps) break;
    }
    fp = rj * x - (double)lp1;
  } else {
    Dual<K> f = dconst<K>(C(1.0)), a = dconst<K>(C(1.0));
    double b = 2.0 * l + 1.0, c = 2.0, dd = l + 2.0;
    fp = dconst<K>(C((double)l));
    for (int it = 0; it < 100000; it++) {
      a = -(a * xsq) / (c * (b + c));
      f = f + a;
      fp = fp + a * dd;
      if (cabs_(a.v * dd) < eps) break;
      c += 2.0;
      dd += 2.0;
    }
    fp = fp / f;
  }
  Dual<K> xr = dsqrt(dabs(xsq));
  return ((double)l - fp) / xr;
}

// the spline values a starting solution reads, as directions
enum StartVar { SV_RHO = 0, SV_MU, SV_QMU, SV_KAPPA, SV_QKAPPA, SV_N };
__host__ __device__ inline int start_field(int v) {
  const int f[SV_N] = {F_RHO, F_MU, F_QMU, F_KAPPA, F_QKAPPA};
  return f[v];
}
using DS = Dual<SV_N>;

// sph_solid_start with tangents; y[col][row]
__host__ __device__ inline void solid_start_d(const Tab& T, const PS& s, int is, double rr,
                                              DS y[3][6]) {
  SplW w = spl_w(T, is, rr);
  DS rrho = dvar<SV_N>(spl(T, w, F_RHO), SV_RHO), rmu = dvar<SV_N>(spl(T, w, F_MU), SV_MU),
     rqmu = dvar<SV_N>(spl(T, w, F_QMU), SV_QMU),
     rkappa = dvar<SV_N>(spl(T, w, F_KAPPA), SV_KAPPA),
     rqkappa = dvar<SV_N>(spl(T, w, F_QKAPPA), SV_QKAPPA);
  double pib = s.pibigg, ll = s.lld, llp1 = s.llp1, zeta = s.zeta, zeta2 = s.zeta2;
  DS cmu = rmu * (1.0 + s.lno * rqmu);
  DS ckappa = rkappa * (1.0 + s.lno * rqkappa);
  DS calpha = (ckappa + 1.3333333333 * cmu) / rrho;
  DS cbeta = cmu / rrho;
  DS tmp1 = s.omega2 * dinv(cbeta);
  DS tmp2 = (s.omega2 + 5.3333333 * pib * rrho) / calpha;
  DS t3 = 2.666666666 * pib * zeta * rrho;
  DS tmp3 = (t3 * t3) / (calpha * cbeta);
  double sign = 1.0;
  for (int i = 0; i < 2; i++) {
    DS d = tmp1 - tmp2;
    DS g = d * d + tmp3;
    g = 0.5 * (tmp1 + tmp2) + (0.5 * sign) * dsqrt(g);
    DS g2 = g;
    g = dsqrt(g);
    DS zt = (0.75 * cbeta) * (g2 - tmp1) / (pib * rrho);
    DS xi = zt - llp1;
    DS ratio = bfs_ratio_d(s.ll, g2 * (rr * rr));
    DS gr = (g * rr) * ratio;
    y[i][0] = ll * xi - zt * gr;
    y[i][1] = zeta * (xi + gr);
    y[i][2] = (-4.0 * pib * rr) * (rrho * zt);
    y[i][3] = (-rr) * (rrho * (calpha * zt * g2)) + (2.0 * ll * (ll - 1.0) / rr) * (cmu * xi) +
              2.0 * (cmu * (2.0 * zt + zeta2) * (g * ratio));
    DS y5 = rr * g2 + (2.0 * (ll - 1.0) / rr) * xi - 2.0 * ((zt + 1.0) * (g * ratio));
    y[i][4] = zeta * (cmu * y5);
    y[i][5] = -(rrho * (zeta2 + llp1 * zt));
    sign = -sign;
  }
  y[2][0] = dconst<SV_N>(C(ll));
  y[2][1] = dconst<SV_N>(C(zeta));
  y[2][2] = rr * (s.omega2 + (-4.0 * pib * ll / 3.0) * rrho);
  y[2][3] = (2.0 * ll * (ll - 1.0) / rr) * cmu;
  y[2][4] = (2.0 * zeta * (ll - 1.0) / rr) * cmu;
  y[2][5] = (0.25 / pib) * ((2.0 * ll + 1.0) * s.omega2 + (-8.0 * pib * ll * (ll - 1.0) / 3.0) * rrho);
}

// sph_fluid_start with tangents; y[col][row]
__host__ __device__ inline void fluid_start_d(const Tab& T, const PS& s, int is, double rr,
                                              DS y[2][4]) {
  SplW w = spl_w(T, is, rr);
  DS rrho = dvar<SV_N>(spl(T, w, F_RHO), SV_RHO),
     rkappa = dvar<SV_N>(spl(T, w, F_KAPPA), SV_KAPPA),
     rqkappa = dvar<SV_N>(spl(T, w, F_QKAPPA), SV_QKAPPA);
  double pib = s.pibigg, ll = s.lld, llp1 = s.llp1, zeta2 = s.zeta2;
  DS ckappa = rkappa * (1.0 + s.lno * rqkappa);
  DS calpha = ckappa / rrho;
  DS gamg = (4.0 * pib / 3.0) * rrho;
  DS g = (s.omega2 + 4.0 * gamg - s.omegai2 * (zeta2 * (gamg * gamg))) / calpha;
  DS g2 = g;
  g = dsqrt(g);
  DS zt = -(s.omega2 * dinv(gamg));
  DS xi = zt - llp1;
  DS ratio = bfs_ratio_d(s.ll, g2 * (rr * rr));
  y[0][0] = ll * xi - zt * ((g * rr) * ratio);
  y[0][1] = (-4.0 * pib * rr) * (rrho * zt);
  y[0][2] = (-rr) * (rrho * (calpha * zt * g2));
  y[0][3] = -(rrho * (zeta2 + llp1 * zt));
  y[1][0] = dconst<SV_N>(C(ll));
  y[1][1] = rr * (s.omega2 + (-4.0 * pib * ll / 3.0) * rrho);
  y[1][2] = dconst<SV_N>(C(0.0));
  y[1][3] = (0.25 / pib) * ((2.0 * ll + 1.0) * s.omega2 + (-8.0 * pib * ll * (ll - 1.0) / 3.0) * rrho);
}

__host__ __device__ inline DS det2_d(const DS& a1, const DS& a2, const DS& b1, const DS& b2) {
  return a1 * b2 - a2 * b1;
}

// sph_minor_solid_start / sph_minor_fluid_start with tangents (14 or 5
// minors, zeros after)
__host__ __device__ inline void minor_start_d(const Tab& T, const PS& s, int is, double rr,
                                              bool fluid, DS* m) {
  for (int k = 0; k < 14; k++) m[k] = dconst<SV_N>(C(0.0));
  if (!fluid) {
    DS y[3][6];
    solid_start_d(T, s, is, rr, y);
    const int rows[14][3] = {{0, 1, 2}, {0, 1, 3}, {0, 1, 4}, {0, 1, 5}, {0, 2, 3},
                             {0, 2, 4}, {0, 3, 4}, {0, 3, 5}, {0, 4, 5}, {1, 2, 3},
                             {1, 3, 4}, {1, 3, 5}, {2, 3, 4}, {3, 4, 5}};
    for (int k = 0; k < 14; k++) {
      int p = rows[k][0], q = rows[k][1], r = rows[k][2];
      m[k] = y[0][p] * det2_d(y[1][q], y[1][r], y[2][q], y[2][r]) -
             y[1][p] * det2_d(y[0][q], y[0][r], y[2][q], y[2][r]) +
             y[2][p] * det2_d(y[0][q], y[0][r], y[1][q], y[1][r]);
    }
  } else {
    DS y[2][4];
    fluid_start_d(T, s, is, rr, y);
    const int rows[5][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {2, 3}};
    for (int k = 0; k < 5; k++) {
      int p = rows[k][0], q = rows[k][1];
      m[k] = det2_d(y[0][p], y[0][q], y[1][p], y[1][q]);
    }
  }
}

// tor_ystart with tangents (the sbessj2 stub ratio x / (2l + 3))
__host__ __device__ inline void tor_ystart_d(const Tab& T, const PS& s, int is, double rr, DS* y) {
  SplW w = spl_w(T, is, rr);
  DS rrho = dvar<SV_N>(spl(T, w, F_RHO), SV_RHO), rmu = dvar<SV_N>(spl(T, w, F_MU), SV_MU),
     rqmu = dvar<SV_N>(spl(T, w, F_QMU), SV_QMU);
  DS cmu = rmu * (1.0 + s.lno * rqmu);
  DS cbeta = dsqrt(cmu / rrho);
  DS arg = rr * (s.omega * dinv(cbeta));
  DS ratio = arg / (double)(2 * (s.ll + 1) + 3);
  y[0] = dconst<SV_N>(C(s.zeta * rr));
  y[1] = s.zeta * (cmu * (((s.lld - 1.0) - arg * ratio)));
}

// ------------------------------------------------------- radial (l = 0)
// Adjoint of RadRHS::coef (rad_amat, src/module_int.f90:261-304) at one
// stage radius: tb, kb, svb are the holomorphic adjoints of its entries t,
// k (= 1/C'), sv; vb the adjoints of rho, A, C, F, N, g, 1/Qmu, xlam, xa2.
constexpr int RAD_NV = 9;
__host__ __device__ inline int rad_field(int v) {
  const int f[RAD_NV] = {F_RHO, F_ACON, F_CCON, F_FCON, F_NCON, F_GRAV, F_QMU, F_XLAM, F_XA2};
  return f[v];
}
__host__ __device__ inline void rad_coef_adj(const Tab& T, const PS& s, int il, double rr, cd tb,
                                             cd kb, cd svb, double* vb) {
  SplW w = spl_w(T, il, rr);
  double rrho = spl(T, w, F_RHO), racon = spl(T, w, F_ACON), rccon = spl(T, w, F_CCON),
         rfcon = spl(T, w, F_FCON), rncon = spl(T, w, F_NCON), gg = spl(T, w, F_GRAV),
         rqmu = spl(T, w, F_QMU), rxlam = spl(T, w, F_XLAM), rxa2 = spl(T, w, F_XA2);
  cd qaa = 1.0 + rxa2 * s.lno, qff = 1.0 + rxlam * s.lno, qll = 1.0 + rqmu * s.lno;
  cd cfcon = rfcon * qff;
  cd cccon = cinv(rccon * qaa);
  cd fc = cfcon * cccon;
  double ri = 1.0 / rr, ri2 = ri * ri;
  // t = (1 - 2 fc)/r; sv = 4/r^2 (gamma - rho g r) - rho w^2
  cd fcb = -2.0 * ri * tb;
  cd gammab = 4.0 * ri2 * svb;
  double rgrb = -4.0 * ri2 * svb.x;
  // gamma = cacon - cncon - cfcon fc
  cd caconb = gammab, cnconb = -gammab, cfconb = -cmul_(fc, gammab);
  fcb = fcb - cmul_(cfcon, gammab);
  cfconb = cfconb + cmul_(cccon, fcb);
  cd cccb = kb + cmul_(cfcon, fcb);
  cd ub = -cmul_(cmul_(cccon, cccon), cccb);
  cd qaab = racon * caconb + rccon * ub, qffb = rfcon * cfconb, qllb = rncon * cnconb;
  vb[0] = cmul_(svb, -s.omega2).x + rgrb * gg * rr;
  vb[1] = cmul_(caconb, qaa).x;
  vb[2] = cmul_(ub, qaa).x;
  vb[3] = cmul_(cfconb, qff).x;
  vb[4] = cmul_(cnconb, qll).x;
  vb[5] = rgrb * rrho * rr;
  vb[6] = cmul_(qllb, s.lno).x;
  vb[7] = cmul_(qffb, s.lno).x;
  vb[8] = cmul_(qaab, s.lno).x;
}

// rad_ystart with tangents (the sbessj2 stub ratio), src/module_int.f90:417-452
__host__ __device__ inline void rad_ystart_d(const Tab& T, const PS& s, int is, double rr, DS* y) {
  SplW w = spl_w(T, is, rr);
  DS rrho = dvar<SV_N>(spl(T, w, F_RHO), SV_RHO), rmu = dvar<SV_N>(spl(T, w, F_MU), SV_MU),
     rqmu = dvar<SV_N>(spl(T, w, F_QMU), SV_QMU),
     rkappa = dvar<SV_N>(spl(T, w, F_KAPPA), SV_KAPPA),
     rqkappa = dvar<SV_N>(spl(T, w, F_QKAPPA), SV_QKAPPA);
  DS cmu = rmu * (1.0 + s.lno * rqmu);
  DS ckappa = rkappa * (1.0 + s.lno * rqkappa);
  DS calpha = ckappa + (4.0 / 3.0) * cmu;
  DS gamma = s.omega2 + (16.0 * s.pibigg / 3.0) * rrho;
  gamma = (rrho * gamma) / calpha;
  DS ratio = (rr / 5.0) * gamma;
  y[0] = rr * ratio;
  y[1] = rr * (calpha * gamma) - 4.0 * (cmu * ratio);
}

// ------------------------------------------------------------ radial steps
// One Cash-Karp step of the radial system (rkck's tableau, the stage radii
// of rk_step) applied to the NC columns Y[c], with the six stage
// coefficient sets evaluated once: the step matrix on NC vectors.
template <int NC>
__host__ __device__ inline void rad_step_cols(const Tab& T, const PS& s, int il, double x,
                                              double h, cd (*Y)[2]) {
  using namespace ck;
  const double off[6] = {0.0, 0.2, 0.3, 0.6, 1.0, 0.875};
  RadRHS f;
  RadRHS::Coef c[6];
#pragma unroll
  for (int k = 0; k < 6; k++) c[k] = f.coef(T, s, il, k == 0 ? x : x + off[k] * h);
#pragma unroll
  for (int col = 0; col < NC; col++) {
    cd* y = Y[col];
    cd k1[2], k2[2], k3[2], k4[2], k5[2], k6[2], t[2];
    RadRHS::apply(c[0], y, k1);
    for (int i = 0; i < 2; i++) t[i] = y[i] + (b21 * h) * k1[i];
    RadRHS::apply(c[1], t, k2);
    for (int i = 0; i < 2; i++) t[i] = y[i] + h * (b31 * k1[i] + b32 * k2[i]);
    RadRHS::apply(c[2], t, k3);
    for (int i = 0; i < 2; i++) t[i] = y[i] + h * (b41 * k1[i] + b42 * k2[i] + b43 * k3[i]);
    RadRHS::apply(c[3], t, k4);
    for (int i = 0; i < 2; i++)
      t[i] = y[i] + h * (b51 * k1[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i]);
    RadRHS::apply(c[4], t, k5);
    for (int i = 0; i < 2; i++)
      t[i] = y[i] + h * (b61 * k1[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i]);
    RadRHS::apply(c[5], t, k6);
    for (int i = 0; i < 2; i++) y[i] = y[i] + h * (c1 * k1[i] + c3 * k3[i] + c4 * k4[i] + c6 * k6[i]);
  }
}

// One radial step at (il, xr, h) from state y taken back: the row x <- x S
// (S the step matrix), and every stage's pairing kb_k . (dA_k z_k) through
// rad_coef_adj and the spline weights into acc[v * 4 + e] (field
// rad_field(v); e = 0, 1: the value at knots il, il + 1; e = 2, 3: the
// second derivative at il, il + 1).  Loops unrolled: fixed register
// indices throughout.
__host__ __device__ inline void rad_step_back(const Tab& T, const PS& s, int il, double xr,
                                              double h, const cd* y, cd* x, double* acc) {
  using namespace ck;
  const double off[6] = {0.0, 0.2, 0.3, 0.6, 1.0, 0.875};
  const double cc[6] = {c1, 0.0, c3, c4, 0.0, c6};
  const double bb[6][6] = {{0, 0, 0, 0, 0, 0},
                           {b21, 0, 0, 0, 0, 0},
                           {b31, b32, 0, 0, 0, 0},
                           {b41, b42, b43, 0, 0, 0},
                           {b51, b52, b53, b54, 0, 0},
                           {b61, b62, b63, b64, b65, 0}};
  RadRHS f;
  RadRHS::Coef c[6];
  cd z[6][2], d[6][2];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    c[k] = f.coef(T, s, il, k == 0 ? xr : xr + off[k] * h);
#pragma unroll
    for (int i = 0; i < 2; i++) {
      cd v = C(0.0);
#pragma unroll
      for (int m = 0; m < k; m++) v = v + bb[k][m] * d[m][i];
      z[k][i] = y[i] + h * v;
    }
    RadRHS::apply(c[k], z[k], d[k]);
  }
  cd zb[6][2];
  cd xn[2] = {x[0], x[1]};
#pragma unroll
  for (int k = 5; k >= 0; k--) {
    cd kb[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      cd v = cc[k] * x[i];
#pragma unroll
      for (int m = k + 1; m < 6; m++) v = v + bb[m][k] * zb[m][i];
      kb[i] = h * v;
    }
    // zb = A^T kb for A = [[t, k], [sv, -t]]
    zb[k][0] = c[k].t * kb[0] + c[k].sv * kb[1];
    zb[k][1] = c[k].k * kb[0] - c[k].t * kb[1];
    xn[0] = xn[0] + zb[k][0];
    xn[1] = xn[1] + zb[k][1];
    const cd tb = kb[0] * z[k][0] - kb[1] * z[k][1], kbar = kb[0] * z[k][1],
             svb = kb[1] * z[k][0];
    const double r = k == 0 ? xr : xr + off[k] * h;
    double vb[RAD_NV];
    rad_coef_adj(T, s, il, r, tb, kbar, svb, vb);
    const SplW sw = spl_w(T, il, r);
    const double wa = sw.deg ? 1.0 : sw.a, wb = sw.deg ? 0.0 : sw.b,
                 wa3 = sw.deg ? 0.0 : sw.a3 * sw.hh, wb3 = sw.deg ? 0.0 : sw.b3 * sw.hh;
#pragma unroll
    for (int v = 0; v < RAD_NV; v++) {
      acc[v * 4 + 0] += vb[v] * wa;
      acc[v * 4 + 1] += vb[v] * wb;
      acc[v * 4 + 2] += vb[v] * wa3;
      acc[v * 4 + 3] += vb[v] * wb3;
    }
  }
  x[0] = xn[0];
  x[1] = xn[1];
}

}  // namespace ysx
