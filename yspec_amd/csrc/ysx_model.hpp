// ysx_model.hpp -- numerics compiled for both the host and the device from
// this one source: the deck -> model-table transform (normalisation and
// moduli per knot, natural cubic splines per layer, gravity by
// Gauss-Legendre quadrature of the spline of r^2 rho: read_model numerics,
// src/module_model.f90:253-378; spline_r / tridag_ser_r / splint_dis_r,
// src/module_spline.f90:40-73, 262-283, 426-449; gauslv,
// src/module_model.f90:507-535) and the receiver Legendre functions
// (src/module_function.f90:6-132).  Contraction into FMAs is off so the
// host and the device round every operation alike.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "ysx_common.hpp"

namespace ysx {

constexpr double R_NORM_M = 6.371e+6;    // src/module_model.f90:18
constexpr double RHO_NORM_M = 5.515e+3;  // src/module_model.f90:19

// Fields splined per layer, each followed by its second-derivative table
__host__ __device__ inline int spline_field(int k) {
  const int f[12] = {F_RHO, F_ACON, F_CCON, F_FCON, F_LCON, F_NCON,
                     F_KAPPA, F_MU, F_QKAPPA, F_QMU, F_XLAM, F_XA2};
  return f[k];
}

// Knot i (0-based) of the raw deck (9 columns per knot; ians = 0 decks use
// the first 6) into the table tab[field * nknot + i]; gravity comes later.
__host__ __device__ inline void knot_fields(const double* deck, int nknot, int i, int ifanis,
                                            double vel_norm, double* tab) {
#pragma clang fp contract(off)
  const double* d = deck + 9 * (size_t)i;
  auto at = [&](int f) -> double& { return tab[(size_t)f * nknot + i]; };
  double r = d[0] / R_NORM_M, rho = d[1] / RHO_NORM_M;
  double A = ifanis ? d[6] : d[2], Cc = d[2], L = d[3], Nn = ifanis ? d[7] : d[3];
  double F = ifanis ? d[8] : 1.0, qk = d[4], qm = d[5];
  double vA = A / vel_norm, vC = Cc / vel_norm, vL = L / vel_norm, vN = Nn / vel_norm;
  A = rho * (vA * vA);
  Cc = rho * (vC * vC);
  L = rho * (vL * vL);
  Nn = rho * (vN * vN);
  F = (A - 2.0 * L) * F;
  double kap = (Cc + 4.0 * A - 4.0 * Nn + 4.0 * F) / 9.0;
  double mu = (Cc + A + 6.0 * L + 5.0 * Nn - 2.0 * F) / 15.0;
  double lam = kap - 2.0 * mu / 3.0;
  if (qk > 0.0) qk = 1.0 / qk;
  if (qm > 0.0) qm = 1.0 / qm;
  double rat = 4.0 * mu / (3.0 * (lam + 2.0 * mu));
  at(F_R) = r;
  at(F_RHO) = rho;
  at(F_ACON) = A;
  at(F_CCON) = Cc;
  at(F_LCON) = L;
  at(F_NCON) = Nn;
  at(F_FCON) = F;
  at(F_KAPPA) = kap;
  at(F_MU) = mu;
  at(F_QKAPPA) = qk;
  at(F_QMU) = qm;
  at(F_XLAM) = ((1.0 - rat) * qk - 0.5 * rat * qm) / (1.0 - 1.5 * rat);
  at(F_XA2) = (1.0 - rat) * qk + rat * qm;
  at(F_BETA) = sqrt(mu / rho);
  at(F_ALPHA) = sqrt((kap + 1.3333333333 * mu) / rho);
}

// Natural cubic spline second derivatives y2[0..n) of y over x (one
// layer); work holds 3 (n + 1) doubles.  Returns false on a zero pivot.
__host__ __device__ inline bool natural_spline(const double* x, const double* y, double* y2, int n,
                                               double* work) {
#pragma clang fp contract(off)
  double* sup = work;            // [n + 1]
  double* rhs = work + (n + 1);  // [n + 1]
  double* gam = work + 2 * (n + 1);
  for (int i = 0; i <= n; i++) sup[i] = rhs[i] = gam[i] = 0.0;
  for (int i = 1; i <= n - 1; i++) sup[i] = x[i] - x[i - 1];
  for (int i = 1; i <= n - 1; i++) rhs[i] = 6.0 * ((y[i] - y[i - 1]) / sup[i]);
  for (int i = n - 1; i >= 2; i--) rhs[i] -= rhs[i - 1];
  rhs[1] = rhs[n] = 0.0;
  // tridiagonal: sub[i] = sup[i-1], dia[i] = 2 (sup[i] + sub[i]) inside,
  // 1 at both ends; super-diagonal sup[i] inside, 0 in the first row
  auto sub = [&](int i) { return (i >= 2 && i <= n - 1) ? sup[i - 1] : 0.0; };
  auto dia = [&](int i) { return (i >= 2 && i <= n - 1) ? 2.0 * (sup[i] + sup[i - 1]) : 1.0; };
  auto spr = [&](int i) { return (i >= 2) ? sup[i] : 0.0; };
  double bet = dia(1);
  if (bet == 0.0) return false;
  y2[0] = rhs[1] / bet;
  for (int j = 2; j <= n; j++) {
    gam[j] = spr(j - 1) / bet;
    bet = dia(j) - sub(j) * gam[j];
    if (bet == 0.0) return false;
    y2[j - 1] = (rhs[j] - sub(j) * y2[j - 2]) / bet;
  }
  for (int j = n - 1; j >= 1; j--) y2[j - 1] -= gam[j + 1] * y2[j];
  return true;
}

// splint_dis_r on interval [xa[klo], xa[klo+1]] (0-based klo)
__host__ __device__ inline double splint(const double* xa, const double* ya, const double* y2a,
                                         double x, int klo) {
#pragma clang fp contract(off)
  double h = xa[klo + 1] - xa[klo];
  if (h == 0.0) return ya[klo];
  double a = (xa[klo + 1] - x) / h, b = (x - xa[klo]) / h;
  return a * ya[klo] + b * ya[klo + 1] +
         ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[klo + 1]) * (h * h) / 6.0;
}

// gauslv's 5-point Gauss-Legendre integral of the spline of r^2 rho over
// knot interval i (0-based), without the accumulation
__host__ __device__ inline double grav_interval(const double* r, const double* intg,
                                                const double* qint, int i) {
#pragma clang fp contract(off)
  const double gw0 = .478628670499366, gw1 = .236926885056189;
  const double gx0 = .538469310105683, gx1 = .906179845938664;
  double y1 = 0.5 * (r[i + 1] + r[i]), y2 = 0.5 * (r[i + 1] - r[i]);
  double sum = .568888888888889 * splint(r, intg, qint, y1, i);
  double t1 = gx0 * y2;
  sum += gw0 * (splint(r, intg, qint, y1 + t1, i) + splint(r, intg, qint, y1 - t1, i));
  t1 = gx1 * y2;
  sum += gw1 * (splint(r, intg, qint, y1 + t1, i) + splint(r, intg, qint, y1 - t1, i));
  return y2 * sum;
}

// ------------------------------------------------ adjoint of the transform
// The transposes of the steps above, shared by the host function
// (model_table_vjp) and the device kernel (k_model_vjp) so that both
// round alike.  The radii are not differentiated.

// Transpose of natural_spline on one layer: y2 = T^-1 B y with the
// tridiagonal T and the second-difference operator B of spline_r /
// tridag_ser_r (src/module_spline.f90:40-73, 426-449): ybar += B^T T^-T y2bar.
// work holds 2 (n + 2) doubles.
__host__ __device__ inline void natural_spline_T(const double* x, const double* y2bar,
                                                 double* ybar, int n, double* work) {
#pragma clang fp contract(off)
  if (n < 3) return;  // u_1 = u_n = 0: no dependence on y
  // 1-based rows 1..n of T; h(i) = x[i] - x[i-1] (0-based x)
  auto h = [&](int i) { return x[i] - x[i - 1]; };
  auto sub = [&](int i) { return (i >= 2 && i <= n - 1) ? h(i - 1) : 0.0; };
  auto dia = [&](int i) { return (i >= 2 && i <= n - 1) ? 2.0 * (h(i) + h(i - 1)) : 1.0; };
  auto spr = [&](int i) { return (i >= 2 && i <= n - 1) ? h(i) : 0.0; };
  double* c = work;          // [n + 2]
  double* z = work + n + 2;  // [n + 2]
  for (int j = 0; j < n + 2; j++) c[j] = z[j] = 0.0;
  // T^T z = y2bar: row j of T^T has sub-diagonal spr(j-1) and
  // super-diagonal sub(j+1)
  double bet = dia(1);
  z[1] = y2bar[0] / bet;
  for (int j = 2; j <= n; j++) {
    double sb = spr(j - 1);
    c[j] = sub(j) / bet;
    bet = dia(j) - sb * c[j];
    z[j] = (y2bar[j - 1] - sb * z[j - 1]) / bet;
  }
  for (int j = n - 1; j >= 1; j--) z[j] -= c[j + 1] * z[j + 1];
  // rhs[i] = 6 (D_i - D_{i-1}) for i = 2..n-1, D_i = (y[i] - y[i-1]) / h_i
  for (int i = 2; i <= n - 1; i++) {
    double g = 6.0 * z[i];
    ybar[i] += g / h(i);
    ybar[i - 1] -= g / h(i);
    ybar[i - 1] -= g / h(i - 1);
    ybar[i - 2] += g / h(i - 1);
  }
}

// Transpose of splint on interval i (0-based) at x: the adjoint sbar of
// s = a y[i] + b y[i+1] + ((a^3 - a) y2[i] + (b^3 - b) y2[i+1]) h^2 / 6
// into ybar and y2bar
__host__ __device__ inline void splint_T(const double* xa, double x, int i, double sbar,
                                         double* ybar, double* y2bar) {
#pragma clang fp contract(off)
  double h = xa[i + 1] - xa[i];
  if (h == 0.0) {
    ybar[i] += sbar;
    return;
  }
  double a = (xa[i + 1] - x) / h, b = (x - xa[i]) / h, hh = h * h / 6.0;
  ybar[i] += a * sbar;
  ybar[i + 1] += b * sbar;
  y2bar[i] += (a * a * a - a) * hh * sbar;
  y2bar[i + 1] += (b * b * b - b) * hh * sbar;
}

// Transpose of grav_interval on interval i: gbar (the adjoint of the
// interval's integral) into the adjoints of the spline of r^2 rho
__host__ __device__ inline void grav_interval_T(const double* r, int i, double gbar,
                                                double* intgb, double* qintb) {
#pragma clang fp contract(off)
  const double gw0 = .478628670499366, gw1 = .236926885056189;
  const double gx0 = .538469310105683, gx1 = .906179845938664, gc = .568888888888889;
  double y1 = 0.5 * (r[i + 1] + r[i]), y2 = 0.5 * (r[i + 1] - r[i]), sb = y2 * gbar;
  splint_T(r, y1, i, gc * sb, intgb, qintb);
  splint_T(r, y1 + gx0 * y2, i, gw0 * sb, intgb, qintb);
  splint_T(r, y1 - gx0 * y2, i, gw0 * sb, intgb, qintb);
  splint_T(r, y1 + gx1 * y2, i, gw1 * sb, intgb, qintb);
  splint_T(r, y1 - gx1 * y2, i, gw1 * sb, intgb, qintb);
}

// Transpose of knot_fields for knot i: the table-field adjoints bar
// (field-major, F_COUNT x nknot, gravity and spline terms already folded
// in) into the adjoints of the knot's 9 deck columns db[9]; tab is the
// forward table (for alpha and beta).  A zero shear modulus (fluid) has no
// finite derivative through beta; zero Q stays zero.
__host__ __device__ inline void knot_fields_T(const double* deck, int nknot, int i, int ifanis,
                                              double vel_norm, const double* tab,
                                              const double* bar, double* db) {
#pragma clang fp contract(off)
  const double* d = deck + 9 * (size_t)i;
  auto fb = [&](int f) { return bar[(size_t)f * nknot + i]; };
  for (int c = 0; c < 9; c++) db[c] = 0.0;
  double rho = d[1] / RHO_NORM_M;
  double Av = ifanis ? d[6] : d[2], Cv = d[2], Lv = d[3], Nv = ifanis ? d[7] : d[3];
  double F0 = ifanis ? d[8] : 1.0, qk = d[4], qm = d[5], vn = vel_norm;
  double vA = Av / vn, vC = Cv / vn, vL = Lv / vn, vN = Nv / vn;
  double A = rho * (vA * vA), Cc = rho * (vC * vC), L = rho * (vL * vL), Nn = rho * (vN * vN);
  double F = (A - 2.0 * L) * F0;
  double kap = (Cc + 4.0 * A - 4.0 * Nn + 4.0 * F) / 9.0;
  double mu = (Cc + A + 6.0 * L + 5.0 * Nn - 2.0 * F) / 15.0;
  double lam = kap - 2.0 * mu / 3.0;
  double qk1 = (qk > 0.0) ? 1.0 / qk : qk, qm1 = (qm > 0.0) ? 1.0 / qm : qm;
  double D = lam + 2.0 * mu, rat = 4.0 * mu / (3.0 * D);
  double num = (1.0 - rat) * qk1 - 0.5 * rat * qm1, den = 1.0 - 1.5 * rat;
  double rhob = fb(F_RHO), Ab = fb(F_ACON), Cb = fb(F_CCON), Lb = fb(F_LCON);
  double Nb = fb(F_NCON), Fb = fb(F_FCON), kb = fb(F_KAPPA), mb = fb(F_MU);
  double qkb = fb(F_QKAPPA), qmb = fb(F_QMU), xlb = fb(F_XLAM);
  double xab = fb(F_XA2), bb = fb(F_BETA), ab = fb(F_ALPHA);
  // beta = sqrt(mu / rho), alpha = sqrt((kappa + 4/3 mu) / rho)
  double beta = tab[(size_t)F_BETA * nknot + i], alpha = tab[(size_t)F_ALPHA * nknot + i];
  if (beta > 0.0) {
    mb += bb / (2.0 * beta * rho);
    rhob -= bb * beta / (2.0 * rho);
  }
  if (alpha > 0.0) {
    kb += ab / (2.0 * alpha * rho);
    mb += 1.3333333333 * ab / (2.0 * alpha * rho);
    rhob -= ab * alpha / (2.0 * rho);
  }
  // xa2 = (1 - rat) qk1 + rat qm1; xlam = num / den
  double ratb = (qm1 - qk1) * xab;
  qkb += (1.0 - rat) * xab;
  qmb += rat * xab;
  qkb += (1.0 - rat) / den * xlb;
  qmb += -0.5 * rat / den * xlb;
  ratb += ((-qk1 - 0.5 * qm1) * den + 1.5 * num) / (den * den) * xlb;
  // rat = 4 mu / (3 D), D = lam + 2 mu, lam = kappa - 2 mu / 3
  mb += ratb * 4.0 / (3.0 * D);
  double Db = -ratb * 4.0 * mu / (3.0 * D * D);
  double lb = Db;
  mb += 2.0 * Db;
  kb += lb;
  mb += -2.0 / 3.0 * lb;
  // mu and kappa from the moduli
  Cb += mb / 15.0;
  Ab += mb / 15.0;
  Lb += 6.0 * mb / 15.0;
  Nb += 5.0 * mb / 15.0;
  Fb += -2.0 * mb / 15.0;
  Cb += kb / 9.0;
  Ab += 4.0 * kb / 9.0;
  Nb += -4.0 * kb / 9.0;
  Fb += 4.0 * kb / 9.0;
  // F = (A - 2 L) F0
  Ab += F0 * Fb;
  Lb += -2.0 * F0 * Fb;
  double F0b = (A - 2.0 * L) * Fb;
  // X = rho v^2
  rhob += vA * vA * Ab + vC * vC * Cb + vL * vL * Lb + vN * vN * Nb;
  double vAb = 2.0 * rho * vA * Ab, vCb = 2.0 * rho * vC * Cb;
  double vLb = 2.0 * rho * vL * Lb, vNb = 2.0 * rho * vN * Nb;
  db[1] = rhob / RHO_NORM_M;
  db[2] = vCb / vn;
  db[3] = vLb / vn;
  if (ifanis) {
    db[6] = vAb / vn;
    db[7] = vNb / vn;
    db[8] = F0b;
  } else {
    db[2] += vAb / vn;
    db[3] += vNb / vn;
  }
  db[4] = (qk > 0.0) ? -qkb / (qk * qk) : qkb;
  db[5] = (qm > 0.0) ? -qmb / (qm * qm) : qmb;
}

// legendre, src/module_function.f90:6-132.  X_lm, dX_lm/dtheta and
// X_lm/sin(theta) for orders 0..m at colatitude theta (ct = cos theta,
// st = sin theta); the reference takes sfl3, f1, f2 and the second
// normalising square root in single precision.
__host__ __device__ inline void legendre_cs(double theta, double ct, double st, int l, int m,
                                            double* x, double* xp, double* xc) {
#pragma clang fp contract(off)
  double fct = sqrt((double)(float)(2 * l + 1) / FOURPI_D);
  double sfl3 = (double)sqrtf((float)(l * (l + 1)));
  double compar = (double)(float)(2 * l + 1) / FOURPI_D;
  double small = 1.0e-16 * compar;
  for (int i = 0; i <= m; i++) x[i] = xp[i] = xc[i] = 0.0;
  if (l <= 1 || fabs(theta) <= 1.0e-5) {
    x[0] = fct;
    if (l == 0) return;
    x[0] = ct * fct;
    x[1] = -st * fct / sfl3;
    xp[0] = -st * fct;
    xp[1] = -0.5 * ct * fct * sfl3;
    xc[1] = (fabs(theta) < 1.0e-5) ? xp[1] : x[1] / st;
    return;
  }
  double p0 = 1.0, p1 = ct;  // Legendre recurrence P_{i-2}, P_{i-1}
  for (int i = 2; i <= l; i++) {
    double p2 = ((double)(float)(2 * i - 1) * ct * p1 - (double)(float)(i - 1) * p0) /
                (double)(float)i;
    p0 = p1;
    p1 = p2;
  }
  double cot = ct / st, cosec = 1.0 / st;
  double x0 = p1 * fct;
  double x1 = (double)(float)l * (p0 - ct * p1) * fct / st;
  x[0] = x0;
  xp[0] = -x1;
  double xp1 = (double)(float)(l * (l + 1)) * x0 - cot * x1;
  x[1] = -x1 / sfl3;
  xc[1] = x[1] * cosec;
  xp[1] = -xp1 / sfl3;
  double sum = x0 * x0 + 2.0 * x[1] * x[1];
  if (sum - compar > small) return;
  double xa = x0, xb = -x1 / (double)sqrtf((float)(l * (l + 1)));
  for (int i = 3; i <= m + 1; i++) {
    double f1 = (double)sqrtf((float)(l + i - 1) * (float)(l - i + 2));
    double f2 = (double)sqrtf((float)(l + i - 2) * (float)(l - i + 3));
    double xm = i - 1;
    double x3 = -(2.0 * cot * (xm - 1.0) * xb + f2 * xa) / f1;
    sum += 2.0 * x3 * x3;
    if (sum - compar > small && i != l + 1) return;
    x[i - 1] = x3;
    xc[i - 1] = x3 * cosec;
    xa = xb;
    xp[i - 1] = -(f1 * xb + xm * cot * x3);
    xb = x3;
  }
}

}  // namespace ysx
