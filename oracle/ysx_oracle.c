/*
 * ysx_oracle.c -- CPU restatement of the reference yspec spectral path.
 *
 * TEST INFRASTRUCTURE ONLY (see ysx_oracle.h).  This file restates the
 * reference's Fortran algorithm function by function, in the same order
 * of floating-point operations where that is cheap, so that it can be
 * pinned against FP64 dumps of the reference itself.  Citations are
 * path:line inside the reference tree.
 */
#define _GNU_SOURCE
#include "ysx_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ---- constants (src/nrtype.f90:20-25): the pi family is rounded to
 *      single precision, fourpi_d is truncated. --------------------- */
#define PI_D ((double)3.141592653589793238462643383279502884197f)
#define TWOPI_D ((double)6.283185307179586476925286766559005768394f)
static const double FOURPI_D = 12.56637061;
static const double BIGG = 6.6723e-11;
/* src/module_model.f90:18-20 */
static const double R_NORM = 6.371e+6;
static const double RHO_NORM = 5.515e+3;
static const double BIG = 1.e+40;
/* src/module_int.f90:74-81 */
static const int NWSTEP = 4;
static const double TINY = 1.0e-20;
#define HUGE_ (1.0 / 1.0e-20)
/* src/module_start.f90:6 */
static const double TOL_START = 11.0;

enum {
  A_R, A_RHO, A_RHO_CS, A_ACON, A_ACON_CS, A_CCON, A_CCON_CS, A_FCON,
  A_FCON_CS, A_LCON, A_LCON_CS, A_NCON, A_NCON_CS, A_KAPPA, A_KAPPA_CS, A_MU,
  A_MU_CS, A_QKAPPA, A_QKAPPA_CS, A_QMU, A_QMU_CS, A_RAT, A_RAT_CS, A_XLAM,
  A_XLAM_CS, A_XA2, A_XA2_CS, A_GRAV, A_GRAV_CS, A_LAMBDA, A_BETA, A_ALPHA,
  A_COUNT
};

struct ysxo {
  char err[256];
  /* normalisation (src/module_model.f90:474-492) */
  double vel_norm, acl_norm, fre_norm, t_norm, con_norm, grav_norm,
      moment_norm, pot_norm, pibigg;
  /* model (src/module_model.f90:118-160), 1-based arrays */
  int nknot, nic, noc, nsl, ifanis, ocean, fic, nlayer;
  int *layer_lo, *layer_hi;
  double tref;
  double *a[A_COUNT];
  /* run */
  int have_run;
  int lmin, lmax, ats_in, grav, out, cor, nr;
  int nw, nt, i1, i2, is, ir, nout;
  double rs, rr, dt, tout, ep, df, wmin, dw, f11, f12, f21, f22, mm[6];
  double *delta, *azep, *azst, *xa, *xpa, *xca; /* xa: (lmax+1,3,nr) */
  cplx *eia;
};

#define R(o) ((o)->a[A_R])

static int fail(ysxo *o, const char *msg) {
  snprintf(o->err, sizeof o->err, "%s", msg);
  return 1;
}

/* ================================================================== */
/* splines: src/module_spline.f90                                      */
/* ================================================================== */

/* tridag_ser_r, src/module_spline.f90:426-449.  a(1..n-1) sub-diagonal
 * (a[j] multiplies u(j) in row j+1), b(1..n), c(1..n-1), r(1..n). */
static int tridag(const double *a, const double *b, const double *c,
                  const double *r, double *u, int n) {
  double *gam = malloc(sizeof(double) * (size_t)(n + 1));
  double bet = b[1];
  if (bet == 0.0) { free(gam); return 1; }
  u[1] = r[1] / bet;
  for (int j = 2; j <= n; j++) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j - 1] * gam[j];
    if (bet == 0.0) { free(gam); return 1; }
    u[j] = (r[j] - a[j - 1] * u[j - 1]) / bet;
  }
  for (int j = n - 1; j >= 1; j--) u[j] = u[j] - gam[j + 1] * u[j + 1];
  free(gam);
  return 0;
}

/* spline_r with natural end conditions (yp1 = ypn = big),
 * src/module_spline.f90:40-73.  x, y, y2 are 1-based views of length n. */
static int spline_natural(const double *x, const double *y, double *y2,
                          int n) {
  double *a = calloc((size_t)n + 2, sizeof(double));
  double *b = calloc((size_t)n + 2, sizeof(double));
  double *c = calloc((size_t)n + 2, sizeof(double));
  double *r = calloc((size_t)n + 2, sizeof(double));
  double *asub = calloc((size_t)n + 2, sizeof(double));
  for (int i = 1; i <= n - 1; i++) c[i] = x[i + 1] - x[i];
  for (int i = 1; i <= n - 1; i++) r[i] = 6.0 * ((y[i + 1] - y[i]) / c[i]);
  /* r(2:n-1) = r(2:n-1) - r(1:n-2): whole-array semantics, old values */
  for (int i = n - 1; i >= 2; i--) r[i] = r[i] - r[i - 1];
  for (int i = 2; i <= n - 1; i++) a[i] = c[i - 1];
  for (int i = 2; i <= n - 1; i++) b[i] = 2.0 * (c[i] + a[i]);
  b[1] = 1.0;
  b[n] = 1.0;
  (void)BIG; /* yp1, ypn > 0.99e30: natural */
  r[1] = 0.0;
  c[1] = 0.0;
  r[n] = 0.0;
  a[n] = 0.0;
  /* tridag_ser(a(2:n), b, c(1:n-1), r, y2) */
  for (int i = 1; i <= n - 1; i++) asub[i] = a[i + 1];
  int st = tridag(asub, b, c, r, y2, n);
  free(a); free(b); free(c); free(r); free(asub);
  return st;
}

/* splint_dis_r, src/module_spline.f90:262-283 (1-based, klo given) */
static inline double splint_dis(const double *xa, const double *ya,
                                const double *y2a, double x, int klo) {
  int khi = klo + 1;
  double h = xa[khi] - xa[klo];
  if (h == 0.0) return ya[klo];
  double a = (xa[khi] - x) / h;
  double b = (x - xa[klo]) / h;
  return a * ya[klo] + b * ya[khi] +
         ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]) * (h * h) /
             6.0;
}

#define SPL(o, name, x, k) \
  splint_dis(R(o), (o)->a[A_##name], (o)->a[A_##name##_CS], (x), (k))

/* ================================================================== */
/* model: src/module_model.f90                                         */
/* ================================================================== */

ysxo *ysxo_create(void) {
  ysxo *o = calloc(1, sizeof(ysxo));
  /* set_parameters, src/module_model.f90:474-492 */
  o->vel_norm = R_NORM * sqrt(PI_D * BIGG * RHO_NORM);
  o->acl_norm = PI_D * BIGG * RHO_NORM * R_NORM;
  o->fre_norm = o->vel_norm / R_NORM;
  o->t_norm = 1 / o->fre_norm;
  o->con_norm = o->vel_norm * o->vel_norm * RHO_NORM;
  o->grav_norm = PI_D * BIGG;
  o->moment_norm = R_NORM * R_NORM * R_NORM * R_NORM * R_NORM * RHO_NORM *
                   RHO_NORM * o->grav_norm;
  o->pot_norm = o->acl_norm * R_NORM;
  o->pibigg = PI_D * BIGG / o->grav_norm;
  return o;
}

static void free_run(ysxo *o) {
  free(o->delta); free(o->azep); free(o->azst);
  free(o->xa); free(o->xpa); free(o->xca); free(o->eia);
  o->delta = o->azep = o->azst = o->xa = o->xpa = o->xca = NULL;
  o->eia = NULL;
  o->have_run = 0;
}

static void free_model(ysxo *o) {
  for (int i = 0; i < A_COUNT; i++) { free(o->a[i]); o->a[i] = NULL; }
  free(o->layer_lo); free(o->layer_hi);
  o->layer_lo = o->layer_hi = NULL;
  o->nknot = 0;
}

void ysxo_destroy(ysxo *o) {
  if (!o) return;
  free_run(o);
  free_model(o);
  free(o);
}

const char *ysxo_error(const ysxo *o) { return o->err; }

/* gauslv + intgds, src/module_model.f90:495-535 */
static void gauslv(const double *r, const double *intg, const double *qint,
                   double r1, double r2, int iq, double *fint) {
  static const double w[2] = {.478628670499366, .236926885056189};
  static const double xg[2] = {.538469310105683, .906179845938664};
  double y1 = 0.5 * (r2 + r1);
  double y2 = 0.5 * (r2 - r1);
  double vals = splint_dis(r, intg, qint, y1, iq);
  double sum = .568888888888889 * vals;
  for (int i = 0; i < 2; i++) {
    double t1 = xg[i] * y2;
    double v = splint_dis(r, intg, qint, y1 + t1, iq);
    double v1 = splint_dis(r, intg, qint, y1 - t1, iq);
    sum = sum + w[i] * (v + v1);
  }
  *fint = *fint + y2 * sum;
}

/* read_model numerics, src/module_model.f90:194-378 */
int ysxo_set_model(ysxo *o, int nknot, int nic, int noc, int ians,
                   double tref_period, const double *deck) {
  free_run(o);
  free_model(o);
  if (nknot < 3 || nic < 1 || noc <= nic || noc >= nknot)
    return fail(o, "bad deck header");
  o->nknot = nknot;
  o->nic = nic;
  o->noc = noc;
  for (int i = 0; i < A_COUNT; i++)
    o->a[i] = calloc((size_t)nknot + 2, sizeof(double));
  double *r = o->a[A_R], *rho = o->a[A_RHO], *acon = o->a[A_ACON],
         *ccon = o->a[A_CCON], *fcon = o->a[A_FCON], *lcon = o->a[A_LCON],
         *ncon = o->a[A_NCON], *kappa = o->a[A_KAPPA], *mu = o->a[A_MU],
         *qkappa = o->a[A_QKAPPA], *qmu = o->a[A_QMU], *rat = o->a[A_RAT],
         *xlam = o->a[A_XLAM], *xa2 = o->a[A_XA2], *grav = o->a[A_GRAV],
         *lambda = o->a[A_LAMBDA], *beta = o->a[A_BETA],
         *alpha = o->a[A_ALPHA];
  /* header: tref = twopi_d/tref/fre_norm (:195-197) */
  o->tref = TWOPI_D / tref_period;
  o->tref = o->tref / o->fre_norm;
  o->ifanis = (ians == 1);
  for (int i = 1; i <= nknot; i++) {
    const double *d = deck + 9 * (size_t)(i - 1);
    r[i] = d[0];
    rho[i] = d[1];
    ccon[i] = d[2];
    lcon[i] = d[3];
    qkappa[i] = d[4];
    qmu[i] = d[5];
    if (o->ifanis) {
      acon[i] = d[6];
      ncon[i] = d[7];
      fcon[i] = d[8];
    } else {
      acon[i] = ccon[i];
      ncon[i] = lcon[i];
      fcon[i] = 1.0;
    }
  }
  /* ocean and fluid inner core (:227-248) */
  o->ocean = (lcon[nknot] == 0.0);
  o->nsl = nknot;
  if (o->ocean) {
    for (int i = nknot - 1; i >= 1; i--)
      if (lcon[i] > 0.0) { o->nsl = i; break; }
  }
  o->fic = (lcon[1] == 0.0);
  /* normalise, moduli (:253-277) */
  for (int i = 1; i <= nknot; i++) {
    r[i] = r[i] / R_NORM;
    rho[i] = rho[i] / RHO_NORM;
    acon[i] = acon[i] / o->vel_norm;
    ccon[i] = ccon[i] / o->vel_norm;
    lcon[i] = lcon[i] / o->vel_norm;
    ncon[i] = ncon[i] / o->vel_norm;
    acon[i] = rho[i] * (acon[i] * acon[i]);
    ccon[i] = rho[i] * (ccon[i] * ccon[i]);
    lcon[i] = rho[i] * (lcon[i] * lcon[i]);
    ncon[i] = rho[i] * (ncon[i] * ncon[i]);
    fcon[i] = (acon[i] - 2.0 * lcon[i]) * fcon[i];
    kappa[i] = ccon[i] + 4.0 * acon[i] - 4.0 * ncon[i] + 4.0 * fcon[i];
    kappa[i] = kappa[i] / 9.0;
    mu[i] = ccon[i] + acon[i] + 6.0 * lcon[i] + 5.0 * ncon[i] - 2.0 * fcon[i];
    mu[i] = mu[i] / 15.0;
    lambda[i] = kappa[i] - 2.0 * mu[i] / 3.0;
    beta[i] = sqrt(mu[i] / rho[i]);
    alpha[i] = sqrt((kappa[i] + 1.3333333333 * mu[i]) / rho[i]);
    /* Q -> q (:282-283) */
    if (qkappa[i] > 0.0) qkappa[i] = 1.0 / qkappa[i];
    if (qmu[i] > 0.0) qmu[i] = 1.0 / qmu[i];
    /* attenuation terms as in minos (:287-290) */
    rat[i] = 4.0 * mu[i] / (3.0 * (lambda[i] + 2.0 * mu[i]));
    xlam[i] = (1.0 - rat[i]) * qkappa[i] - 0.5 * rat[i] * qmu[i];
    xlam[i] = xlam[i] / (1.0 - 1.5 * rat[i]);
    xa2[i] = (1.0 - rat[i]) * qkappa[i] + rat[i] * qmu[i];
  }
  /* layers from duplicated radii (:294-308) */
  int nl = 1;
  for (int i = 2; i <= nknot; i++)
    if (r[i] == r[i - 1]) nl++;
  o->nlayer = nl;
  o->layer_lo = calloc((size_t)nl + 1, sizeof(int));
  o->layer_hi = calloc((size_t)nl + 1, sizeof(int));
  o->layer_lo[1] = 1;
  o->layer_hi[nl] = nknot;
  int j = 0;
  for (int i = 2; i <= nknot; i++)
    if (r[i] == r[i - 1]) {
      j++;
      o->layer_hi[j] = i - 1;
      o->layer_lo[j + 1] = i;
    }
  /* splines per layer for 13 parameters (:312-352) */
  static const int val[13] = {A_RHO, A_ACON, A_CCON, A_FCON, A_LCON,
                              A_NCON, A_KAPPA, A_MU, A_QKAPPA, A_QMU,
                              A_RAT, A_XLAM, A_XA2};
  for (int k = 1; k <= nl; k++) {
    int lo = o->layer_lo[k], n = o->layer_hi[k] - lo + 1;
    for (int p = 0; p < 13; p++)
      if (spline_natural(r + lo - 1, o->a[val[p]] + lo - 1,
                         o->a[val[p] + 1] + lo - 1, n))
        return fail(o, "tridag_ser: zero pivot");
  }
  /* gravity by Gauss-Legendre quadrature of r^2 rho (:356-378) */
  double *intg = calloc((size_t)nknot + 2, sizeof(double));
  double *qint = calloc((size_t)nknot + 2, sizeof(double));
  for (int i = 1; i <= nknot; i++) intg[i] = r[i] * r[i] * rho[i];
  for (int k = 1; k <= nl; k++) {
    int lo = o->layer_lo[k], n = o->layer_hi[k] - lo + 1;
    spline_natural(r + lo - 1, intg + lo - 1, qint + lo - 1, n);
  }
  grav[1] = 0.0;
  double fint = 0.0;
  for (int i = 1; i <= nknot - 1; i++) {
    gauslv(r, intg, qint, r[i], r[i + 1], i, &fint);
    grav[i + 1] = fint;
  }
  for (int i = 1; i <= nknot; i++)
    if (r[i] > 0.0) grav[i] = (4.0 * grav[i]) / (r[i] * r[i]);
  free(intg);
  free(qint);
  for (int k = 1; k <= nl; k++) {
    int lo = o->layer_lo[k], n = o->layer_hi[k] - lo + 1;
    spline_natural(r + lo - 1, grav + lo - 1, o->a[A_GRAV_CS] + lo - 1, n);
  }
  return 0;
}

int ysxo_model_arrays(const ysxo *o, double *out) {
  if (!o->nknot) return 1;
  for (int p = 0; p < A_COUNT; p++)
    memcpy(out + (size_t)p * o->nknot, o->a[p] + 1,
           sizeof(double) * (size_t)o->nknot);
  return 0;
}

int ysxo_model_flags(const ysxo *o, int *out) {
  out[0] = o->nknot; out[1] = o->nic; out[2] = o->noc; out[3] = o->nsl;
  out[4] = o->ifanis; out[5] = o->ocean; out[6] = o->fic; out[7] = o->nlayer;
  return 0;
}

int ysxo_norms(const ysxo *o, double *out) {
  out[0] = o->tref; out[1] = R_NORM; out[2] = RHO_NORM; out[3] = o->vel_norm;
  out[4] = o->acl_norm; out[5] = o->fre_norm; out[6] = o->t_norm;
  out[7] = o->moment_norm; out[8] = o->pibigg;
  return 0;
}

/* find_radius_index(rr, -1), src/module_model.f90:537-576 */
static int find_radius_index_down(const ysxo *o, double rr) {
  const double *r = R(o);
  if (rr < 0.0 || rr > 1.0) return 0;
  for (int i = 1; i <= o->nknot - 1; i++) {
    if (rr == r[i] && rr == r[i + 1]) return i;
    if (rr > r[i] && rr <= r[i + 1]) return i;
  }
  return 0;
}

/* ================================================================== */
/* special functions: src/module_function.f90                          */
/* ================================================================== */

/* legendre, src/module_function.f90:6-132.  sfl3, f1, f2 and the second
 * normalising sqrt are single precision in the reference. */
void ysxo_legendre(double theta, int l, int m, double *x, double *xp,
                   double *xcosec) {
  int lp1 = l + 1, mp1 = m + 1;
  double th = theta, ct = cos(th), st = sin(th);
  double fct = sqrt((double)(float)(2 * l + 1) / FOURPI_D);
  double sfl3 = (double)sqrtf((float)(l * (l + 1)));
  double compar = (double)(float)(2 * l + 1) / FOURPI_D;
  double dsfl3 = sfl3;
  double small = 1.0e-16 * compar;
  double sum;
  for (int i = 0; i < mp1; i++) x[i] = xp[i] = xcosec[i] = 0.0;
  if (l <= 1 || fabs(theta) <= 1.0e-5) {
    x[0] = fct;
    if (l == 0) return;
    x[0] = ct * fct;
    x[1] = -st * fct / dsfl3;
    xp[0] = -st * fct;
    xp[1] = -0.5 * ct * fct * dsfl3;
    if (fabs(theta) < 1.0e-5)
      xcosec[1] = xp[1];
    else
      xcosec[1] = x[1] / st;
    return;
  }
  double x1 = 1.0, x2 = ct, x3;
  for (int i = 2; i <= l; i++) {
    x3 = ((double)(float)(2 * i - 1) * ct * x2 - (double)(float)(i - 1) * x1) /
         (double)(float)i;
    x1 = x2;
    x2 = x3;
  }
  double cot = ct / st;
  double cosec = 1.0 / st;
  x3 = x2 * fct;
  x2 = (double)(float)l * (x1 - ct * x2) * fct / st;
  x[0] = x3;
  x[1] = x2;
  sum = x3 * x3;
  xp[0] = -x2;
  xp[1] = (double)(float)(l * (l + 1)) * x3 - cot * x2;
  x[1] = -x[1] / sfl3;
  xcosec[1] = x[1] * cosec;
  xp[1] = -xp[1] / sfl3;
  sum = sum + 2.0 * x[1] * x[1];
  if (sum - compar > small) return;
  x1 = x3;
  x2 = -x2 / (double)sqrtf((float)(l * (l + 1)));
  for (int i = 3; i <= mp1; i++) {
    int k = i - 1;
    double f1 = (double)sqrtf((float)(l + i - 1) * (float)(l - i + 2));
    double f2 = (double)sqrtf((float)(l + i - 2) * (float)(l - i + 3));
    double xm = k;
    x3 = -(2.0 * cot * (xm - 1.0) * x2 + f2 * x1) / f1;
    sum = sum + 2.0 * x3 * x3;
    if (sum - compar > small && i != lp1) return;
    x[i - 1] = x3;
    xcosec[i - 1] = x[i - 1] * cosec;
    x1 = x2;
    xp[i - 1] = -(f1 * x2 + xm * cot * x3);
    x2 = x3;
  }
}

/* delaz, src/module_function.f90:317-390; out = delta, azep, azst */
void ysxo_delaz(double eplati, double eplong, double stlati, double stlong,
                double *out) {
  const double geoco = 0.993277, hpi = 0.5 * PI_D, rad = 0.0174533;
  double eplat = eplati, stlat = stlati;
  if (eplat == -90.0) eplat = -89.98;
  if (stlat == -90.0) stlat = -89.98;
  if (eplat == stlat && eplong == stlong) {
    out[0] = out[1] = out[2] = 0.0;
    return;
  }
  double el = atan(geoco * tan(eplat * rad));
  el = hpi - el;
  double stl = atan(geoco * tan(stlat * rad));
  stl = hpi - stl;
  double elon = eplong * rad, slon = stlong * rad;
  double as = cos(stl), bs = sin(stl), cs = cos(slon), ds = sin(slon);
  double a = cos(el), b = sin(el), c = cos(elon), d = sin(elon);
  double cdel = a * as + b * bs * (c * cs + d * ds);
  if (fabs(cdel) > 1.0) cdel = copysign(1.0, cdel);
  double delt = acos(cdel);
  double sdel = sin(delt);
  double caze = (as - a * cdel) / (sdel * b);
  if (fabs(caze) > 1.0) caze = copysign(1.0, caze);
  double aze = acos(caze);
  double cazs = 0.0;
  if (bs > 0.0) cazs = (a - as * cdel) / (bs * sdel);
  if (bs == 0.0) cazs = copysign(1.0, cazs);
  if (fabs(cazs) > 1.0) cazs = copysign(1.0, cazs);
  double azs = acos(cazs);
  double dif = ds * c - cs * d;
  if (dif < 0.0) aze = TWOPI_D - aze;
  if (dif > 0.0) azs = TWOPI_D - azs;
  out[0] = delt;
  out[1] = aze;
  out[2] = azs;
}

/* bfs, src/module_function.f90:520-581 */
static cplx bfs(int l, cplx xsq, double eps) {
  cplx fp;
  if (creal(xsq) > 0.0) {
    cplx x = csqrt(xsq);
    int lp1 = l + 1;
    cplx rx = 2.0 / x;
    cplx nu = lp1 - 0.5;
    cplx rj = nu * rx;
    rx = -rx;
    cplx denom = (nu + 1.0) * rx;
    cplx numer = denom + 1.0 / rj;
    rj = rj * numer / denom;
    int nm1 = 1;
    for (;;) {
      nm1 = nm1 + 1;
      rx = -rx;
      cplx a3 = (nu + (double)nm1) * rx;
      denom = a3 + 1.0 / denom;
      numer = a3 + 1.0 / numer;
      cplx ratio = numer / denom;
      rj = rj * ratio;
      if (fabs(cabs(ratio) - 1.0) <= eps) break;
    }
    fp = rj * x - (double)lp1;
  } else {
    cplx f = 1.0, a = 1.0, b = (double)(l + l) + 1.0, c = 2.0,
         d = (double)l + 2.0;
    fp = (double)l;
    for (;;) {
      a = -a * xsq / (c * (b + c));
      f = f + a;
      fp = fp + a * d;
      if (cabs(a * d) < eps) break;
      c = c + 2.0;
      d = d + 2.0;
    }
    fp = fp / f;
  }
  return fp;
}

void ysxo_bfs(int l, double re, double im, double eps, double *fp2) {
  cplx fp = bfs(l, re + I * im, eps);
  fp2[0] = creal(fp);
  fp2[1] = cimag(fp);
}

/* bfs_dum, src/module_function.f90:497-516: x is the REAL modulus */
static cplx bfs_dum(int l, cplx xsq) {
  cplx fp = bfs(l, xsq, 1.e-9);
  double x = sqrt(cabs(xsq));
  return ((double)l - fp) / x;
}

/* ================================================================== */
/* integration state and RHS: src/module_int.f90                       */
/* ================================================================== */

typedef struct {
  const ysxo *o;
  int il, ll, llp1;
  double zeta, zeta2, zeta2m2, szeta2m2, nu, dr, ats;
  cplx omega, omega2, omegai, omegai2, lno;
} pst;

/* set_l, src/module_int.f90:2072-2089 */
static void set_l(pst *s, int l) {
  s->ll = l;
  s->llp1 = l + 1;
  s->zeta2 = (double)(float)(l * (l + 1));
  s->zeta = sqrt(s->zeta2);
  s->zeta2m2 = s->zeta2 - 2.0;
  s->szeta2m2 = s->zeta2m2 > 0.0 ? sqrt(s->zeta2m2) : 0.0;
  s->nu = (double)(float)(2 * l + 1) / FOURPI_D;
  s->nu = sqrt(s->nu);
}

/* set_omega, src/module_int.f90:2091-2102 */
static void set_omega(pst *s, cplx w) {
  s->omega = w;
  s->omega2 = w * w;
  s->omegai = 1.0 / w;
  s->omegai2 = s->omegai * s->omegai;
  s->lno = 2.0 * clog(I * w / s->o->tref) / PI_D;
  s->lno = s->ats * s->lno;
}

/* steps, src/module_int.f90:2121-2138 */
static double steps(const pst *s, int i1) {
  const double *alpha = s->o->a[A_ALPHA];
  double tmp = (i1 < s->o->nknot) ? fmax(alpha[i1], alpha[i1 + 1]) : alpha[i1];
  double st = NWSTEP * creal(s->omega) / tmp;
  return 1.0 / st;
}

/* iminloc(alpha): first index of the minimum */
static int iminloc(const double *v, int lo, int hi) {
  int im = lo;
  for (int i = lo + 1; i <= hi; i++)
    if (v[i] < v[im]) im = i;
  return im;
}

/* sph_steps / rad_steps, src/module_int.f90:1145-1166, 394-414 */
static void alpha_steps(pst *s) {
  const double *alpha = s->o->a[A_ALPHA];
  int im = iminloc(alpha, 1, s->o->nknot);
  if (creal(s->omega) == 0.0)
    s->dr = 1.0 / NWSTEP;
  else
    s->dr = 1.0 / (NWSTEP * creal(s->omega) / alpha[im]);
}

typedef void (*derivs_fn)(const pst *, double, const cplx *, cplx *);

/* tor_amat + tor_derivs, src/module_int.f90:134-184 */
static void tor_derivs(const pst *s, double rr, const cplx *y, cplx *dy) {
  const ysxo *o = s->o;
  int il = s->il;
  double rrho = SPL(o, RHO, rr, il), rlcon = SPL(o, LCON, rr, il),
         rncon = SPL(o, NCON, rr, il), rqmu = SPL(o, QMU, rr, il);
  cplx qll = 1.0 + rqmu * s->lno;
  cplx clcon = rlcon * qll, cncon = rncon * qll;
  double ri = 1.0 / rr;
  double t = 2.0 * ri;
  cplx k = 1.0 / clcon;
  cplx sv = -s->omega2 * rrho + ri * ri * s->zeta2m2 * cncon;
  dy[0] = t * y[0] + k * y[1];
  dy[1] = sv * y[0] - t * y[1];
}

/* rad_amat + rad_derivs, src/module_int.f90:261-372 */
static void rad_derivs(const pst *s, double rr, const cplx *y, cplx *dy) {
  const ysxo *o = s->o;
  int il = s->il;
  double rrho = SPL(o, RHO, rr, il), racon = SPL(o, ACON, rr, il),
         rccon = SPL(o, CCON, rr, il), rfcon = SPL(o, FCON, rr, il),
         rncon = SPL(o, NCON, rr, il), gg = SPL(o, GRAV, rr, il),
         rqmu = SPL(o, QMU, rr, il), rxlam = SPL(o, XLAM, rr, il),
         rxa2 = SPL(o, XA2, rr, il);
  cplx qaa = 1.0 + rxa2 * s->lno, qff = 1.0 + rxlam * s->lno,
       qll = 1.0 + rqmu * s->lno;
  cplx cacon = racon * qaa, cccon = rccon * qaa, cfcon = rfcon * qff,
       cncon = rncon * qll;
  cccon = 1.0 / cccon;
  cplx gamma = cacon - cncon - cfcon * cfcon * cccon;
  double ri = 1.0 / rr;
  cplx t = (1.0 - 2.0 * cfcon * cccon) * ri;
  cplx k = cccon;
  cplx sv = -s->omega2 * rrho + 4.0 * (gamma - rrho * gg * rr) * ri * ri;
  dy[0] = t * y[0] + k * y[1];
  dy[1] = sv * y[0] - t * y[1];
}

typedef struct {
  cplx t11, t12, t21, t22, t31, t33, k11, k22, k33, s11, s12, s13, s22, s23;
} solid_amat;

/* sph_solid_amat, src/module_int.f90:500-560 */
static void sph_solid_amat(const pst *s, double rr, solid_amat *A) {
  const ysxo *o = s->o;
  int il = s->il;
  double rrho = SPL(o, RHO, rr, il), racon = SPL(o, ACON, rr, il),
         rccon = SPL(o, CCON, rr, il), rfcon = SPL(o, FCON, rr, il),
         rncon = SPL(o, NCON, rr, il), rlcon = SPL(o, LCON, rr, il),
         gg = SPL(o, GRAV, rr, il), rqmu = SPL(o, QMU, rr, il),
         rxlam = SPL(o, XLAM, rr, il), rxa2 = SPL(o, XA2, rr, il);
  double pibigg = o->pibigg, zeta = s->zeta, zeta2 = s->zeta2;
  cplx qaa = 1.0 + rxa2 * s->lno, qff = 1.0 + rxlam * s->lno,
       qll = 1.0 + rqmu * s->lno;
  cplx cacon = racon * qaa, cccon = rccon * qaa, cfcon = rfcon * qff,
       cncon = rncon * qll, clcon = rlcon * qll;
  cccon = 1.0 / cccon;
  cplx gamma = cacon - cncon - cfcon * cfcon * cccon;
  double ri = 1.0 / rr;
  A->t11 = (1.0 - 2.0 * cfcon * cccon) * ri;
  A->t12 = zeta * ri * cfcon * cccon;
  A->t21 = -zeta * ri;
  A->t22 = 2.0 * ri;
  A->t31 = -4 * pibigg * rrho;
  A->t33 = -s->ll * ri;
  A->k11 = cccon;
  A->k22 = 1.0 / clcon;
  A->k33 = 4.0 * pibigg;
  A->s11 = -s->omega2 * rrho + 4.0 * (gamma - rrho * gg * rr) * ri * ri;
  A->s12 = zeta * (rrho * gg * rr - 2.0 * gamma) * ri * ri;
  A->s13 = -rrho * s->llp1 * ri;
  A->s22 = -s->omega2 * rrho + (zeta2 * (gamma + cncon) - 2.0 * cncon) * ri * ri;
  A->s23 = rrho * zeta * ri;
}

/* sph_minors_solid_derivs, src/module_int.f90:818-871 */
static void sph_minors_solid_derivs(const pst *s, double rr, const cplx *m,
                                    cplx *d) {
  solid_amat A;
  sph_solid_amat(s, rr, &A);
  cplx a123 = A.t11 + A.t22 + A.t33, a126 = A.t11 + A.t22 - A.t33,
       a135 = A.t11 - A.t22 + A.t33, a234 = -A.t11 + A.t22 + A.t33;
  /* m(k) -> m[k-1] */
  d[0] = A.k33 * m[3] - A.k22 * m[5] + A.k11 * m[9] + m[0] * a123;
  d[1] = -A.k22 * m[6] + m[0] * A.s13 - m[2] * A.t21 + m[1] * A.t22 -
         m[3] * A.t31;
  d[2] = -A.k11 * m[10] + m[0] * A.s23 + m[2] * A.t11 - m[1] * A.t12;
  d[3] = A.k22 * m[8] - A.k11 * m[11] + m[3] * a126;
  d[4] = -A.k33 * m[7] - m[0] * A.s12 + m[9] * A.t12 - m[5] * A.t21 +
         m[2] * A.t31 + m[4] * A.t33;
  d[5] = -A.k33 * m[8] - A.k11 * m[12] - m[0] * A.s22 -
         2.0 * m[4] * A.t12 + m[5] * a135;
  d[6] = m[2] * A.s12 + m[5] * A.s13 - m[1] * A.s22 - m[4] * A.s23 +
         m[10] * A.t12 - m[6] * A.t22 + m[8] * A.t31;
  d[7] = m[3] * A.s12 - m[2] * A.s13 - m[1] * A.s23 + m[11] * A.t12 -
         m[8] * A.t21 - m[7] * A.t33;
  d[8] = A.k11 * m[13] + m[3] * A.s22 - 2.0 * m[2] * A.s23 -
         2.0 * m[7] * A.t12 - m[8] * a234;
  d[9] = -A.k33 * m[11] + A.k22 * m[12] + m[0] * A.s11 +
         2.0 * m[4] * A.t21 - 2.0 * m[1] * A.t31 + m[9] * a234;
  d[10] = -m[2] * A.s11 + m[1] * A.s12 - m[4] * A.s13 - m[9] * A.s23 -
          m[10] * A.t11 + m[6] * A.t21 - m[7] * A.t31;
  d[11] = -A.k22 * m[13] - m[3] * A.s11 + 2.0 * m[1] * A.s13 +
          2.0 * m[7] * A.t21 - m[11] * a135;
  d[12] = A.k33 * m[13] - m[5] * A.s11 + 2.0 * m[4] * A.s12 +
          m[9] * A.s22 + 2.0 * m[6] * A.t31 - m[12] * a126;
  d[13] = m[8] * A.s11 - 2.0 * m[7] * A.s12 + 2.0 * m[6] * A.s13 -
          m[11] * A.s22 + 2.0 * m[10] * A.s23 - m[13] * a123;
}

/* sph_bvec_solid_derivs, src/module_int.f90:1016-1057 */
static void sph_bvec_solid_derivs(const pst *s, double rr, const cplx *b,
                                  cplx *d) {
  solid_amat A;
  sph_solid_amat(s, rr, &A);
  d[0] = A.s11 * b[6] + A.s12 * (b[7] - b[2]) + A.s13 * b[8] - A.s22 * b[3] -
         A.s23 * b[4] - (A.t11 + A.t22) * b[0] + A.t31 * b[5];
  d[1] = A.s11 * b[9] + A.s12 * b[10] - A.s13 * (2.0 * b[2] + b[7]) -
         A.s23 * b[3] - A.t21 * b[5] - (A.t11 + A.t33) * b[1];
  d[2] = A.s12 * b[11] + A.s13 * b[12] + A.t12 * b[3] - A.t21 * b[6] -
         A.t31 * b[9];
  d[3] = -A.k22 * b[0] - A.s11 * b[11] + A.s13 * b[13] +
         A.t21 * (b[2] - b[7]) + (A.t22 - A.t11) * b[3] - A.t31 * b[10];
  d[4] = -A.k33 * b[1] - A.s11 * b[12] - A.s12 * b[13] - A.t21 * b[8] +
         A.t31 * (2.0 * b[2] + b[7]) + (A.t33 - A.t11) * b[4];
  d[5] = A.s12 * b[9] - A.s13 * b[6] + A.s22 * b[10] -
         A.s23 * (b[2] + 2.0 * b[7]) - A.t12 * b[1] - (A.t22 + A.t33) * b[5];
  d[6] = A.k11 * b[0] + A.s22 * b[11] + A.s23 * b[12] -
         A.t12 * (b[2] - b[7]) + (A.t11 - A.t22) * b[6];
  d[7] = -A.s12 * b[11] + A.s23 * b[13] - A.t12 * b[3] + A.t21 * b[6];
  d[8] = -A.k33 * b[5] - A.s12 * b[12] - A.s22 * b[13] - A.t12 * b[4] -
         A.t22 * b[8] + A.t31 * b[6] + A.t33 * b[8];
  d[9] = A.k11 * b[1] + A.s23 * b[11] + A.t12 * b[10] + (A.t11 - A.t33) * b[9];
  d[10] = A.k22 * b[5] - A.s13 * b[11] + A.t21 * b[9] + (A.t22 - A.t33) * b[10];
  d[11] = -A.k11 * b[3] + A.k22 * b[6] + (A.t11 + A.t22) * b[11];
  d[12] = -A.k11 * b[4] + A.k33 * b[9] + A.t12 * b[13] + (A.t11 + A.t33) * b[12];
  d[13] = -A.k22 * b[8] + A.k33 * b[10] + A.t21 * b[12] - A.t31 * b[11] +
          (A.t22 + A.t33) * b[13];
}

/* sph_fluid_amat + sph_minors_fluid_derivs, src/module_int.f90:676-715,
 * 941-967 */
static void sph_minors_fluid_derivs(const pst *s, double rr, const cplx *m,
                                    cplx *d) {
  const ysxo *o = s->o;
  int il = s->il;
  double rrho = SPL(o, RHO, rr, il), rccon = SPL(o, CCON, rr, il),
         gg = SPL(o, GRAV, rr, il), rxa2 = SPL(o, XA2, rr, il);
  double pibigg = o->pibigg, zeta2 = s->zeta2;
  cplx qaa = 1.0 + rxa2 * s->lno;
  cplx cccon = rccon * qaa;
  cccon = 1.0 / cccon;
  double ri = 1.0 / rr;
  cplx oi2 = s->omegai2;
  cplx t11 = (gg * zeta2 * oi2 * ri - 1.0) * ri;
  cplx t12 = zeta2 * oi2 * ri * ri;
  double t21 = -4.0 * pibigg * rrho;
  double t22 = -s->ll * ri;
  cplx k11 = cccon - zeta2 * oi2 * ri * ri / rrho;
  double k22 = 4.0 * pibigg;
  cplx s11 = -s->omega2 * rrho + rrho * gg * (gg * zeta2 * ri * oi2 - 4.0) * ri;
  cplx s12 = rrho * (gg * zeta2 * ri * oi2 - s->llp1) * ri;
  cplx s22 = rrho * zeta2 * ri * ri * oi2;
  d[0] = k22 * m[2] - k11 * m[3] + m[0] * (t11 + t22);
  d[1] = m[0] * s12 + m[3] * t12 - m[2] * t21;
  d[2] = k11 * m[4] + m[0] * s22 - 2.0 * m[1] * t12 + m[2] * (t11 - t22);
  d[3] = -k22 * m[4] - m[0] * s11 + 2.0 * m[1] * t21 + m[3] * (-t11 + t22);
  d[4] = m[2] * s11 - 2.0 * m[1] * s12 - m[3] * s22 - m[4] * (t11 + t22);
}

/* rkck, src/module_int.f90:2016-2064 (Cash-Karp, no error estimate used) */
static void rkck(const pst *s, cplx *y, const cplx *dydx, double x, double h,
                 derivs_fn derivs, int n) {
  const double a2 = 0.2, a3 = 0.3, a4 = 0.6, a5 = 1.0, a6 = 0.875,
               b21 = 0.2, b31 = 3.0 / 40.0, b32 = 9.0 / 40.0, b41 = 0.3,
               b42 = -0.9, b43 = 1.2, b51 = -11.0 / 54.0, b52 = 2.5,
               b53 = -70.0 / 27.0, b54 = 35.0 / 27.0,
               b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0,
               b63 = 575.0 / 13824.0, b64 = 44275.0 / 110592.0,
               b65 = 253.0 / 4096.0, c1 = 37.0 / 378.0, c3 = 250.0 / 621.0,
               c4 = 125.0 / 594.0, c6 = 512.0 / 1771.0;
  cplx ak2[14], ak3[14], ak4[14], ak5[14], ak6[14], yt[14];
  for (int i = 0; i < n; i++) yt[i] = y[i] + b21 * h * dydx[i];
  derivs(s, x + a2 * h, yt, ak2);
  for (int i = 0; i < n; i++) yt[i] = y[i] + h * (b31 * dydx[i] + b32 * ak2[i]);
  derivs(s, x + a3 * h, yt, ak3);
  for (int i = 0; i < n; i++)
    yt[i] = y[i] + h * (b41 * dydx[i] + b42 * ak2[i] + b43 * ak3[i]);
  derivs(s, x + a4 * h, yt, ak4);
  for (int i = 0; i < n; i++)
    yt[i] = y[i] +
            h * (b51 * dydx[i] + b52 * ak2[i] + b53 * ak3[i] + b54 * ak4[i]);
  derivs(s, x + a5 * h, yt, ak5);
  for (int i = 0; i < n; i++)
    yt[i] = y[i] + h * (b61 * dydx[i] + b62 * ak2[i] + b63 * ak3[i] +
                        b64 * ak4[i] + b65 * ak5[i]);
  derivs(s, x + a6 * h, yt, ak6);
  for (int i = 0; i < n; i++)
    y[i] = y[i] + h * (c1 * dydx[i] + c3 * ak3[i] + c4 * ak4[i] + c6 * ak6[i]);
}

/* renormalisation inside yint, src/module_int.f90:1910-1919 */
static void renorm(cplx *y, int n, int *scale) {
  int all_small = 1, any_big = 0;
  for (int i = 0; i < n; i++) if (!(cabs(y[i]) < TINY)) { all_small = 0; break; }
  if (all_small) {
    (*scale)++;
    for (int i = 0; i < n; i++) y[i] = y[i] * HUGE_;
  }
  for (int i = 0; i < n; i++) if (cabs(y[i]) > HUGE_) { any_big = 1; break; }
  if (any_big) {
    (*scale)--;
    for (int i = 0; i < n; i++) y[i] = y[i] * TINY;
  }
}

/* yint, src/module_int.f90:1847-1968: fixed steps per knot interval */
static void yint(pst *s, int i1, int i2, double r1, double r2, cplx *y, int n,
                 derivs_fn derivs, int *scale) {
  const double *r = R(s->o);
  cplx dydx[14];
  if (r2 > r1) {
    for (int i = i1; i <= i2; i++) {
      double rr1 = (i == i1) ? r1 : r[i];
      double rr2 = (i == i2) ? r2 : r[i + 1];
      if (rr1 == rr2) continue;
      s->il = i;
      double drr = steps(s, i);
      int ns = 2 + (int)floor((rr2 - rr1) / drr);
      drr = (rr2 - rr1) / (ns - 1);
      for (int j = 1; j <= ns - 1; j++) {
        double rr = rr1 + (j - 1) * drr;
        derivs(s, rr, y, dydx);
        rkck(s, y, dydx, rr, drr, derivs, n);
        if (scale) renorm(y, n, scale);
      }
    }
  } else if (r2 < r1) {
    for (int i = i1; i >= i2; i--) {
      double rr1 = (i == i1) ? r1 : r[i + 1];
      double rr2 = (i == i2) ? r2 : r[i];
      if (rr2 == rr1) continue;
      s->il = i;
      double drr = steps(s, i);
      int ns = 2 + (int)floor(fabs(rr2 - rr1) / drr);
      drr = (rr2 - rr1) / (ns - 1);
      for (int j = 1; j <= ns - 1; j++) {
        double rr = rr1 + (j - 1) * drr;
        derivs(s, rr, y, dydx);
        rkck(s, y, dydx, rr, drr, derivs, n);
        if (scale) renorm(y, n, scale);
      }
    }
  }
}

/* ================================================================== */
/* start levels: src/module_start.f90                                  */
/* ================================================================== */

/* start_level, src/module_start.f90:80-178 (no sum_in) */
static int start_level(const ysxo *o, int l, double w, int i1, int i2,
                       int wt) {
  const double *r = R(o), *beta = o->a[A_BETA], *alpha = o->a[A_ALPHA];
  const double *v = (wt == 1) ? beta : alpha;
  int isign = (i2 >= i1) ? 1 : -1;
  double pp = (double)(float)(l * (l + 1)) / (w * w);
  int j = i2, is;
  for (int i = i1; isign > 0 ? i <= i2 : i >= i2; i += isign) {
    double q1 = 1.0 / (v[i] * v[i]) - pp / (r[i] * r[i]);
    if (isign == 1) {
      if (q1 > 0.0) { j = (i1 > i - 1) ? i1 : i - 1; break; }
    } else {
      if (q1 < 0.0) { j = i; break; }
    }
  }
  is = j;
  int lo = (i1 < i2) ? i1 : i2;
  if (is == lo) return is;
  int j1 = is, j2 = lo;
  is = j2;
  double sum = 0.0;
  for (int i = j1; i >= j2 + 1; i--) {
    double q1 = 1.0 / (v[i] * v[i]) - pp / (r[i] * r[i]);
    double q2 = 1.0 / (v[i - 1] * v[i - 1]) - pp / (r[i - 1] * r[i - 1]);
    if (q1 > 0.0 || q2 > 0.0) {
      sum = 0.0;
    } else {
      q1 = sqrt(-q1);
      q2 = sqrt(-q2);
      sum = sum + 0.5 * w * (q1 + q2) * (r[i] - r[i - 1]);
    }
    if (sum > TOL_START) { is = i - 1; break; }
  }
  return is;
}

/* tor_start_level (up = 1), src/module_start.f90:14-32 */
static int tor_start_level(const ysxo *o, int l, double w, int it) {
  return start_level(o, l, w, o->noc + 1, it, 1);
}

/* sph_start_level (up = 1), src/module_start.f90:35-75 */
static int sph_start_level(const ysxo *o, int l, double w, int it) {
  const double aw = -2.00e-3, bw = 2.25e-3, dw = 1.28e-3;
  double wdim = w * o->fre_norm;
  int is;
  double wsoc = aw + dw * l;
  if (wdim <= wsoc) {
    is = start_level(o, l, w, o->noc + 1, it, 1);
    if (is == o->nsl) is = is - 1;
    if (is > o->noc + 1) return is;
  }
  double wsic = aw + bw * l;
  if (wdim <= wsic) {
    is = start_level(o, l, w, o->nic + 1, o->noc, 2);
    if (is == o->noc) is = is - 1;
    if (is > o->nic + 1) return is;
  }
  if (o->fic)
    is = start_level(o, l, w, 2, o->nic, 2);
  else
    is = start_level(o, l, w, 2, o->nic, 1);
  if (is == o->nic) is = is - 1;
  return is;
}

/* ================================================================== */
/* starting solutions and minor algebra: src/module_int.f90            */
/* ================================================================== */

/* tor_ystart, src/module_int.f90:213-250 (sbessj2 stub: x/(2l+3)) */
static void tor_ystart(const pst *s, int is, double rr, cplx *y) {
  const ysxo *o = s->o;
  double rrho = SPL(o, RHO, rr, is), rmu = SPL(o, MU, rr, is),
         rqmu = SPL(o, QMU, rr, is);
  cplx qll = 1.0 + rqmu * s->lno;
  cplx cmu = rmu * qll;
  cplx cbeta = csqrt(cmu / rrho);
  cplx kbeta = s->omega / cbeta;
  cplx arg = kbeta * rr;
  cplx ratio = arg / (double)(float)(2 * s->llp1 + 3);
  y[0] = rr;
  y[1] = cmu * ((double)(s->ll - 1) - arg * ratio);
  y[0] = s->zeta * y[0];
  y[1] = s->zeta * y[1];
}

/* rad_ystart, src/module_int.f90:417-452 */
static void rad_ystart(const pst *s, int is, double rr, cplx *y) {
  const ysxo *o = s->o;
  double rrho = SPL(o, RHO, rr, is), rmu = SPL(o, MU, rr, is),
         rqmu = SPL(o, QMU, rr, is), rkappa = SPL(o, KAPPA, rr, is),
         rqkappa = SPL(o, QKAPPA, rr, is);
  cplx cmu = rmu * (1.0 + rqmu * s->lno);
  cplx ckappa = rkappa * (1.0 + rqkappa * s->lno);
  cplx calpha = (ckappa + 4.0 * cmu / 3.0);
  cplx gamma = s->omega2 + 16.0 * o->pibigg * rrho / 3.0;
  gamma = rrho * gamma / calpha;
  cplx arg = gamma * rr;
  cplx ratio = arg / 5.0;
  y[0] = rr * ratio;
  y[1] = rr * calpha * gamma - 4.0 * cmu * ratio;
}

/* sph_solid_start, src/module_int.f90:1192-1270; y is (6,3) column-major */
static void sph_solid_start(const pst *s, int is, double rr, cplx y[3][6]) {
  const ysxo *o = s->o;
  double rrho = SPL(o, RHO, rr, is), rmu = SPL(o, MU, rr, is),
         rqmu = SPL(o, QMU, rr, is), rkappa = SPL(o, KAPPA, rr, is),
         rqkappa = SPL(o, QKAPPA, rr, is);
  double pibigg = o->pibigg, zeta = s->zeta, zeta2 = s->zeta2;
  double ll = s->ll, llp1 = s->llp1;
  cplx cmu = rmu * (1.0 + rqmu * s->lno);
  cplx ckappa = rkappa * (1.0 + rqkappa * s->lno);
  cplx calpha = (ckappa + 1.3333333333 * cmu) / rrho;
  cplx cbeta = cmu / rrho;
  double sign = 1.0;
  cplx tmp1 = s->omega2 / cbeta;
  cplx tmp2 = s->omega2 + 5.3333333 * pibigg * rrho;
  tmp2 = tmp2 / calpha;
  cplx tmp3 = 2.666666666 * pibigg * zeta * rrho;
  tmp3 = tmp3 * tmp3 / (calpha * cbeta);
  for (int i = 0; i < 2; i++) {
    cplx gamma = (tmp1 - tmp2) * (tmp1 - tmp2) + tmp3;
    gamma = 0.5 * (tmp1 + tmp2) + 0.5 * sign * csqrt(gamma);
    cplx gamma2 = gamma;
    gamma = csqrt(gamma);
    cplx zt = 0.75 * cbeta * (gamma2 - tmp1) / (pibigg * rrho);
    cplx xi = zt - llp1;
    cplx args = gamma2 * rr * rr;
    cplx ratio = bfs_dum(s->ll, args);
    y[i][0] = ll * xi - zt * gamma * rr * ratio;
    y[i][1] = zeta * (xi + rr * gamma * ratio);
    y[i][2] = -4.0 * pibigg * rrho * rr * zt;
    y[i][3] = -rrho * calpha * zt * rr * gamma2 +
              2.0 * cmu * ll * (ll - 1) * xi / rr +
              2.0 * cmu * (2.0 * zt + zeta2) * gamma * ratio;
    y[i][4] = rr * gamma2 + 2.0 * (ll - 1) * xi / rr -
              2.0 * (zt + 1.0) * gamma * ratio;
    y[i][4] = cmu * zeta * y[i][4];
    y[i][5] = -rrho * (zeta2 + llp1 * zt);
    sign = -sign;
  }
  y[2][0] = ll;
  y[2][1] = zeta;
  y[2][2] = (s->omega2 - 4.0 * pibigg * rrho * ll / 3.0) * rr;
  y[2][3] = 2.0 * ll * (ll - 1) * cmu / rr;
  y[2][4] = 2.0 * zeta * (ll - 1) * cmu / rr;
  y[2][5] = ((double)(2 * s->ll + 1) * s->omega2 -
             8.0 * pibigg * ll * (ll - 1) * rrho / 3.0);
  y[2][5] = 0.25 * y[2][5] / pibigg;
}

/* sph_fluid_start, src/module_int.f90:1346-1405; y is (4,2) */
static void sph_fluid_start(const pst *s, int is, double rr, cplx y[2][4]) {
  const ysxo *o = s->o;
  double rrho = SPL(o, RHO, rr, is), rkappa = SPL(o, KAPPA, rr, is),
         rqkappa = SPL(o, QKAPPA, rr, is);
  double pibigg = o->pibigg, zeta2 = s->zeta2, ll = s->ll, llp1 = s->llp1;
  cplx ckappa = rkappa * (1.0 + rqkappa * s->lno);
  cplx calpha = ckappa / rrho;
  double gamg = 4.0 * pibigg * rrho / 3.0;
  cplx gamma = (s->omega2 + 4.0 * gamg - zeta2 * gamg * gamg * s->omegai2) /
               calpha;
  cplx gamma2 = gamma;
  gamma = csqrt(gamma);
  cplx zt = -s->omega2 / gamg;
  cplx xi = zt - llp1;
  cplx args = gamma2 * rr * rr;
  cplx ratio = bfs_dum(s->ll, args);
  y[0][0] = ll * xi - zt * gamma * rr * ratio;
  y[0][1] = -4.0 * pibigg * rrho * rr * zt;
  y[0][2] = -rrho * calpha * zt * rr * gamma2;
  y[0][3] = -rrho * (zeta2 + llp1 * zt);
  y[1][0] = ll;
  y[1][1] = (s->omega2 - 4.0 * pibigg * rrho * ll / 3.0) * rr;
  y[1][2] = 0.0;
  y[1][3] = ((double)(2 * s->ll + 1) * s->omega2 -
             8.0 * pibigg * ll * (ll - 1) * rrho / 3.0);
  y[1][3] = 0.25 * y[1][3] / pibigg;
}

/* det2 / det3, src/module_int.f90:1537-1560 */
static inline cplx det2(cplx a1, cplx a2, cplx b1, cplx b2) {
  return a1 * b2 - a2 * b1;
}
static cplx det3(cplx y[3][6], int p, int q, int r) {
  /* rows p,q,r (0-based) of the three columns */
  return y[0][p] * det2(y[1][q], y[1][r], y[2][q], y[2][r]) -
         y[1][p] * det2(y[0][q], y[0][r], y[2][q], y[2][r]) +
         y[2][p] * det2(y[0][q], y[0][r], y[1][q], y[1][r]);
}

/* sph_minor_solid_start, src/module_int.f90:1451-1485 */
static void sph_minor_solid_start(const pst *s, int is, double rr, cplx *m) {
  static const int rows[14][3] = {{1, 2, 3}, {1, 2, 4}, {1, 2, 5}, {1, 2, 6},
                                  {1, 3, 4}, {1, 3, 5}, {1, 4, 5}, {1, 4, 6},
                                  {1, 5, 6}, {2, 3, 4}, {2, 4, 5}, {2, 4, 6},
                                  {3, 4, 5}, {4, 5, 6}};
  cplx y[3][6];
  sph_solid_start(s, is, rr, y);
  for (int k = 0; k < 14; k++)
    m[k] = det3(y, rows[k][0] - 1, rows[k][1] - 1, rows[k][2] - 1);
}

/* sph_minor_fluid_start, src/module_int.f90:1509-1534 */
static void sph_minor_fluid_start(const pst *s, int is, double rr, cplx *m) {
  static const int rows[5][2] = {{1, 2}, {1, 3}, {1, 4}, {2, 3}, {3, 4}};
  cplx y[2][4];
  sph_fluid_start(s, is, rr, y);
  for (int k = 0; k < 5; k++) {
    int p = rows[k][0] - 1, q = rows[k][1] - 1;
    m[k] = det2(y[0][p], y[0][q], y[1][p], y[1][q]);
  }
}

/* b_solid_start, src/module_int.f90:1563-1592 (m, s, b 0-based) */
static void b_solid_start(const cplx *m, const cplx *s, cplx *b) {
  b[0] = m[13] * s[2] - m[6] * s[3] - m[10] * s[4] - m[12] * s[5];
  b[1] = -m[13] * s[1] - m[7] * s[3] - m[11] * s[4] + m[10] * s[5];
  b[2] = m[6] * s[1] + m[7] * s[2] + m[1] * s[4] + m[4] * s[5];
  b[3] = m[10] * s[1] + m[11] * s[2] - m[1] * s[3] + m[9] * s[5];
  b[4] = m[12] * s[1] - m[10] * s[2] - m[4] * s[3] - m[9] * s[4];
  b[5] = m[13] * s[0] - m[8] * s[3] + m[7] * s[4] - m[6] * s[5];
  b[6] = -m[6] * s[0] + m[8] * s[2] + m[2] * s[4] + m[5] * s[5];
  b[7] = -m[10] * s[0] - m[7] * s[2] - m[2] * s[3] - m[4] * s[5];
  b[8] = -m[12] * s[0] + m[6] * s[2] - m[5] * s[3] + m[4] * s[4];
  b[9] = -m[7] * s[0] - m[8] * s[1] + m[3] * s[4] - m[2] * s[5];
  b[10] = -m[11] * s[0] + m[7] * s[1] - m[3] * s[3] + m[1] * s[5];
  b[11] = m[1] * s[0] + m[2] * s[1] + m[3] * s[2] - m[0] * s[5];
  b[12] = m[4] * s[0] + m[5] * s[1] - m[2] * s[2] + m[0] * s[4];
  b[13] = m[9] * s[0] - m[4] * s[1] + m[1] * s[2] - m[0] * s[3];
}

/* sph_minors_solid_2_fluid / fluid_2_solid, src/module_int.f90:1620-1680 */
static void solid_2_fluid(const cplx *m1, cplx *m2) {
  m2[0] = -m1[5];
  m2[1] = -m1[6];
  m2[2] = m1[8];
  m2[3] = -m1[12];
  m2[4] = m1[13];
}
static void fluid_2_solid(const cplx *m1, cplx *m2) {
  for (int i = 0; i < 14; i++) m2[i] = 0.0;
  m2[0] = m1[0];
  m2[1] = m1[1];
  m2[3] = m1[2];
  m2[9] = -m1[3];
  m2[11] = -m1[4];
}

/* sph_delta_solid, src/module_int.f90:1705-1725 */
static cplx sph_delta_solid(const cplx *m1, const cplx *m2) {
  return -m1[13] * m2[0] + 2.0 * m1[6] * m2[1] + 2.0 * m1[10] * m2[2] +
         m1[12] * m2[3] + 2.0 * m1[7] * m2[4] + m1[11] * m2[5] -
         2.0 * m1[1] * m2[6] - 2.0 * m1[4] * m2[7] - m1[9] * m2[8] +
         m1[8] * m2[9] - 2.0 * m1[2] * m2[10] - m1[5] * m2[11] -
         m1[3] * m2[12] + m1[0] * m2[13];
}

/* sph_ysol, src/module_int.f90:1785-1815 */
static void sph_ysol(const cplx *m, const cplx *b, cplx *y) {
  y[0] = m[2] * b[2] + m[0] * b[5] + m[1] * b[6] + 2.0 * m[2] * b[7] +
         m[3] * b[8] + m[4] * b[9] + m[5] * b[10] + m[6] * b[11] +
         m[7] * b[12] + m[8] * b[13];
  y[1] = -m[0] * b[1] - m[1] * (2.0 * b[2] + b[7]) - m[2] * b[3] -
         m[3] * b[4] - m[4] * b[10] - m[7] * b[13] + m[9] * b[9] +
         m[10] * b[11] + m[11] * b[12];
  y[2] = m[0] * b[0] - m[1] * b[8] + m[2] * b[4] + m[4] * (b[7] - b[2]) -
         m[5] * b[3] + m[6] * b[13] - m[9] * b[6] - m[10] * b[12] +
         m[12] * b[11];
  y[3] = m[1] * b[0] + m[4] * b[1] - m[6] * b[3] - m[7] * b[4] +
         m[9] * b[5] - m[10] * (b[2] + 2.0 * b[7]) - m[11] * b[8] -
         m[12] * b[10] + m[13] * b[13];
  y[4] = m[2] * b[0] - m[4] * b[5] + m[5] * b[1] +
         m[6] * (2.0 * b[2] + b[7]) + m[7] * b[8] - m[8] * b[4] +
         m[10] * b[6] + m[12] * b[9] - m[13] * b[12];
  y[5] = m[1] * b[5] - m[2] * b[1] + m[3] * b[0] + m[6] * b[10] +
         m[7] * (b[2] - b[7]) + m[8] * b[3] - m[10] * b[9] + m[11] * b[6] +
         m[13] * b[11];
}

/* ================================================================== */
/* solvers: src/module_solver.f90                                      */
/* ================================================================== */

static const double S6[4][6] = {{1, 0, 0, 0, 0, 0}, {0, 1, 0, 0, 0, 0},
                                {0, 0, 0, 1, 0, 0}, {0, 0, 0, 0, 1, 0}};

/* tor_solver with ss = identity(2), src/module_solver.f90:25-142.
 * yout(2,2) column-major. */
static void tor_solver(pst *s, int l, cplx w, double rr, double rs, int ir,
                       int is, cplx *yout, int *start) {
  const ysxo *o = s->o;
  const double *r = R(o);
  set_l(s, l);
  s->ats = o->ats_in;
  set_omega(s, w);
  int sc1 = 0, sc2 = 0, sc3 = 0, i1;
  cplx y1[2], y2[2], y3[2], delta, b;
  if (rr >= rs) {
    i1 = tor_start_level(o, s->ll, creal(s->omega), is);
    *start = i1;
    if (i1 == is && i1 > o->noc + 1) i1 = i1 - 1;
    double rt = r[i1];
    /* the reference's rt > rs branch (:67-70) cannot trigger: i1 <= is */
    if (i1 == o->noc + 1) {
      y1[0] = 1.0;
      y1[1] = 0.0;
    } else {
      tor_ystart(s, i1, rt, y1);
    }
    yint(s, i1, is, rt, rs, y1, 2, tor_derivs, &sc1);
    y3[0] = y1[0];
    y3[1] = y1[1];
    if (rr != rs) yint(s, is, ir, rs, rr, y3, 2, tor_derivs, &sc3);
    y2[0] = 1.0;
    y2[1] = 0.0;
    if (rr != r[o->nsl])
      yint(s, o->nsl - 1, ir, r[o->nsl], rr, y2, 2, tor_derivs, &sc2);
    delta = y2[1] * y3[0] - y2[0] * y3[1];
    delta = 1.0 / delta;
    for (int i = 0; i < 2; i++) {
      b = y1[1] * (i == 0 ? 1.0 : 0.0) - y1[0] * (i == 1 ? 1.0 : 0.0);
      yout[2 * i + 0] = -delta * y2[0] * b;
      yout[2 * i + 1] = -delta * y2[1] * b;
    }
  } else {
    i1 = tor_start_level(o, s->ll, creal(s->omega), ir);
    *start = i1;
    if (i1 == ir && i1 > o->noc + 1) i1 = i1 - 1;
    double rt = r[i1];
    if (i1 == o->noc + 1) {
      y1[0] = 1.0;
      y1[1] = 0.0;
    } else {
      tor_ystart(s, i1, rt, y1);
    }
    yint(s, i1, ir, rt, rr, y1, 2, tor_derivs, &sc1);
    y2[0] = 1.0;
    y2[1] = 0.0;
    if (rs != r[o->nsl])
      yint(s, o->nsl - 1, is, r[o->nsl], rs, y2, 2, tor_derivs, &sc2);
    y3[0] = y2[0];
    y3[1] = y2[1];
    if (rr != rs) yint(s, is, ir, rs, rr, y3, 2, tor_derivs, &sc3);
    delta = y3[1] * y1[0] - y3[0] * y1[1];
    delta = 1.0 / delta;
    for (int i = 0; i < 2; i++) {
      b = y2[1] * (i == 0 ? 1.0 : 0.0) - y2[0] * (i == 1 ? 1.0 : 0.0);
      yout[2 * i + 0] = -delta * y1[0] * b;
      yout[2 * i + 1] = -delta * y1[1] * b;
    }
  }
  if (sc3 != 0)
    for (int i = 0; i < 4; i++) yout[i] = yout[i] * pow(HUGE_, sc3);
}

/* grid_count / make_grid, src/module_model.f90:578-787 */
static int grid_build(const ysxo *o, double drr, double rs, double rr,
                      int *isg, int *irg, int *igrid, double *rgrid) {
  const double *r = R(o);
  int k = 0;
#define PUT(val, lay)                 \
  do {                                \
    k++;                              \
    if (rgrid) {                      \
      rgrid[k] = (val);               \
      igrid[k] = (lay);               \
    }                                 \
  } while (0)
  for (int i = 1; i <= o->nknot - 1; i++) {
    double r1 = r[i], r2 = r[i + 1];
    if (r1 == r2) continue;
    int ls = (rs > r1 && rs <= r2), lr = (rr > r1 && rr <= r2);
    int nr;
    double drt;
    if (ls && lr) {
      if (rs < rr) {
        nr = (int)floor((rs - r1) / drr) + 2;
        drt = (rs - r1) / (nr - 1);
        for (int j = 1; j <= nr; j++) PUT(r1 + (j - 1) * drt, i);
        *isg = k;
        nr = (int)floor((rr - rs) / drr) + 2;
        drt = (rr - rs) / (nr - 1);
        for (int j = 2; j <= nr; j++) PUT(rs + (j - 1) * drt, i);
        *irg = k;
        nr = (int)floor((r2 - rr) / drr) + 2;
        drt = (r2 - rr) / (nr - 1);
        for (int j = 2; j <= nr; j++) PUT(rr + (j - 1) * drt, i);
      } else if (rr < rs) {
        nr = (int)floor((rr - r1) / drr) + 2;
        drt = (rr - r1) / (nr - 1);
        for (int j = 1; j <= nr; j++) PUT(r1 + (j - 1) * drt, i);
        *irg = k;
        nr = (int)floor((rs - rr) / drr) + 2;
        drt = (rs - rr) / (nr - 1);
        for (int j = 2; j <= nr; j++) PUT(rr + (j - 1) * drt, i);
        *isg = k;
        nr = (int)floor((r2 - rs) / drr) + 2;
        drt = (r2 - rs) / (nr - 1);
        for (int j = 2; j <= nr; j++) PUT(rs + (j - 1) * drt, i);
      } else {
        nr = (int)floor((rs - r1) / drr) + 2;
        drt = (rs - r1) / (nr - 1);
        for (int j = 1; j <= nr; j++) PUT(r1 + (j - 1) * drt, i);
        *isg = k;
        *irg = k;
        nr = (int)floor((r2 - rs) / drr) + 2;
        drt = (r2 - rs) / (nr - 1);
        for (int j = 2; j <= nr; j++) PUT(rs + (j - 1) * drt, i);
      }
    } else if (ls) {
      nr = (int)floor((rs - r1) / drr) + 2;
      drt = (rs - r1) / (nr - 1);
      for (int j = 1; j <= nr; j++) PUT(r1 + (j - 1) * drt, i);
      *isg = k;
      nr = (int)floor((r2 - rs) / drr) + 2;
      drt = (r2 - rs) / (nr - 1);
      for (int j = 2; j <= nr; j++) PUT(rs + (j - 1) * drt, i);
    } else if (lr) {
      nr = (int)floor((rr - r1) / drr) + 2;
      drt = (rr - r1) / (nr - 1);
      for (int j = 1; j <= nr; j++) PUT(r1 + (j - 1) * drt, i);
      *irg = k;
      nr = (int)floor((r2 - rr) / drr) + 2;
      drt = (r2 - rr) / (nr - 1);
      for (int j = 2; j <= nr; j++) PUT(rr + (j - 1) * drt, i);
    } else {
      nr = (int)floor((r2 - r1) / drr) + 2;
      drt = (r2 - r1) / (nr - 1);
      for (int j = 1; j <= nr; j++) PUT(r1 + (j - 1) * drt, i);
    }
  }
#undef PUT
  return k;
}

/* rad_solver_wg with ss = identity(2), src/module_solver.f90:446-561.
 * Only yout is kept (the caller discards pout, src/yspec.f90:277). */
static void rad_solver_wg(pst *s, cplx w, double rr, double rs, cplx *yout) {
  const ysxo *o = s->o;
  set_l(s, 0);
  s->ats = o->ats_in;
  set_omega(s, w);
  alpha_steps(s);
  int iss = 0, irr = 0;
  int ngrid = grid_build(o, s->dr, rs, rr, &iss, &irr, NULL, NULL);
  int *igrid = calloc((size_t)ngrid + 1, sizeof(int));
  double *rgrid = calloc((size_t)ngrid + 1, sizeof(double));
  cplx *y1s = calloc(2 * ((size_t)ngrid + 1), sizeof(cplx));
  cplx *y2s = calloc(2 * ((size_t)ngrid + 1), sizeof(cplx));
  grid_build(o, s->dr, rs, rr, &iss, &irr, igrid, rgrid);
  cplx y1[2], y2[2];
  int i1 = 1;
  double rt = fmin(0.1 * s->dr, 0.1 * rgrid[2]);
  rad_ystart(s, i1, rt, y1);
  yint(s, i1, igrid[2], rt, rgrid[2], y1, 2, rad_derivs, NULL);
  y1s[2 * 2] = y1[0];
  y1s[2 * 2 + 1] = y1[1];
  for (int ig = 2; ig <= iss - 1; ig++) {
    yint(s, igrid[ig], igrid[ig + 1], rgrid[ig], rgrid[ig + 1], y1, 2,
         rad_derivs, NULL);
    y1s[2 * (ig + 1)] = y1[0];
    y1s[2 * (ig + 1) + 1] = y1[1];
  }
  y2[0] = 1.0;
  y2[1] = 0.0;
  y2s[2 * ngrid] = y2[0];
  y2s[2 * ngrid + 1] = y2[1];
  for (int ig = ngrid; ig >= iss + 1; ig--) {
    yint(s, igrid[ig], igrid[ig - 1], rgrid[ig], rgrid[ig - 1], y2, 2,
         rad_derivs, NULL);
    y2s[2 * (ig - 1)] = y2[0];
    y2s[2 * (ig - 1) + 1] = y2[1];
  }
  cplx delta = y2[1] * y1[0] - y2[0] * y1[1];
  delta = 1.0 / delta;
  for (int i = 0; i < 2; i++) {
    cplx ss1 = (i == 0) ? 1.0 : 0.0, ss2 = (i == 1) ? 1.0 : 0.0;
    const cplx *src;
    cplx b;
    if (irr <= iss) {
      b = y2[1] * ss1 - y2[0] * ss2;
      src = y1s + 2 * irr;
    } else {
      b = y1[1] * ss1 - y1[0] * ss2;
      src = y2s + 2 * irr;
    }
    yout[2 * i + 0] = -delta * src[0] * b;
    yout[2 * i + 1] = -delta * src[1] * b;
  }
  free(igrid); free(rgrid); free(y1s); free(y2s);
}

/* sph_solver (grav=2, solid inner core) with ss = s6tmp,
 * src/module_solver.f90:640-823.  yout (6,4) column-major. */
static void sph_solver(pst *s, int l, cplx w, double rr, double rs, int ir,
                       int is, cplx *yout, int *start) {
  const ysxo *o = s->o;
  const double *r = R(o);
  int nic = o->nic, noc = o->noc, nsl = o->nsl, nknot = o->nknot;
  set_l(s, l);
  s->ats = o->ats_in;
  set_omega(s, w);
  alpha_steps(s);
  int sc1 = 0, sc2 = 0, sc3 = 0, sc4, i1;
  cplx mf1[5], mf2[5], ms1[14], ms2[14], ms3[14], b[14], delta;
  /* lower branch: from the start level up to the lower of rs/rr */
  int iup = (rr >= rs) ? is : ir;
  double rup = (rr >= rs) ? rs : rr;
  i1 = sph_start_level(o, s->ll, creal(s->omega), iup);
  *start = i1;
  double rt = r[i1];
  if (rr >= rs && i1 == 1) rt = fmin(0.1 * s->dr, 0.1 * r[2]);
  if (i1 >= iup) {
    i1 = iup - 1;
    rt = r[i1];
  }
  if (i1 <= nic) {
    sph_minor_solid_start(s, i1, rt, ms1);
    yint(s, i1, nic, rt, r[nic], ms1, 14, sph_minors_solid_derivs, &sc1);
    solid_2_fluid(ms1, mf1);
    yint(s, nic + 1, noc, r[nic], r[noc], mf1, 5, sph_minors_fluid_derivs,
         &sc1);
    fluid_2_solid(mf1, ms1);
    yint(s, noc + 1, iup, r[noc + 1], rup, ms1, 14, sph_minors_solid_derivs,
         &sc1);
  } else if (i1 > nic && i1 <= noc) {
    sph_minor_fluid_start(s, i1, rt, mf1);
    yint(s, i1, noc, rt, r[noc], mf1, 5, sph_minors_fluid_derivs, &sc1);
    fluid_2_solid(mf1, ms1);
    yint(s, noc + 1, iup, r[noc + 1], rup, ms1, 14, sph_minors_solid_derivs,
         &sc1);
  } else if (i1 > noc && i1 <= nsl) {
    sph_minor_solid_start(s, i1, rt, ms1);
    yint(s, i1, iup, rt, rup, ms1, 14, sph_minors_solid_derivs, &sc1);
  }
  memcpy(ms3, ms1, sizeof ms3);
  if (rr >= rs) {
    if (rr > rs)
      yint(s, is, ir, rs, rr, ms3, 14, sph_minors_solid_derivs, &sc3);
    if (o->ocean) {
      mf2[0] = 1.0;
      for (int i = 1; i < 5; i++) mf2[i] = 0.0;
      yint(s, nknot - 1, nsl + 1, r[nknot], r[nsl + 1], mf2, 5,
           sph_minors_fluid_derivs, &sc2);
      fluid_2_solid(mf2, ms2);
      yint(s, nsl, ir, r[nsl], rr, ms2, 14, sph_minors_solid_derivs, &sc2);
    } else {
      ms2[0] = 1.0;
      for (int i = 1; i < 14; i++) ms2[i] = 0.0;
      yint(s, nknot - 1, ir, r[nknot], rr, ms2, 14, sph_minors_solid_derivs,
           &sc2);
    }
    delta = sph_delta_solid(ms3, ms2);
    for (int c = 0; c < 4; c++) {
      cplx ssv[6];
      for (int i = 0; i < 6; i++) ssv[i] = S6[c][i];
      sc4 = 0;
      b_solid_start(ms1, ssv, b);
      yint(s, is, ir, rs, rr, b, 14, sph_bvec_solid_derivs, &sc4);
      cplx *y = yout + 6 * c;
      sph_ysol(ms2, b, y);
      for (int i = 0; i < 6; i++) {
        y[i] = -y[i] / delta;
        y[i] = y[i] * pow(HUGE_, sc3 - sc4);
      }
    }
  } else {
    if (rs > rr)
      yint(s, ir, is, rr, rs, ms3, 14, sph_minors_solid_derivs, &sc3);
    if (o->ocean) {
      mf2[0] = 1.0;
      for (int i = 1; i < 5; i++) mf2[i] = 0.0;
      yint(s, nknot - 1, nsl + 1, r[nknot], r[nsl + 1], mf2, 5,
           sph_minors_fluid_derivs, &sc2);
      fluid_2_solid(mf2, ms2);
      yint(s, nsl, is, r[nsl], rs, ms2, 14, sph_minors_solid_derivs, &sc2);
    } else {
      ms2[0] = 1.0;
      for (int i = 1; i < 14; i++) ms2[i] = 0.0;
      /* the reference starts this downward sweep at knot nknot
       * (src/module_solver.f90:802), one above the upper branch's
       * nknot-1; the first interval then has zero length. */
      yint(s, nknot, is, r[nknot], rs, ms2, 14, sph_minors_solid_derivs,
           &sc2);
    }
    delta = sph_delta_solid(ms3, ms2);
    for (int c = 0; c < 4; c++) {
      cplx ssv[6];
      for (int i = 0; i < 6; i++) ssv[i] = S6[c][i];
      sc4 = 0;
      b_solid_start(ms2, ssv, b);
      for (int i = 0; i < 14; i++) b[i] = -b[i] / delta;
      yint(s, is, ir, rs, rr, b, 14, sph_bvec_solid_derivs, &sc4);
      cplx *y = yout + 6 * c;
      sph_ysol(ms1, b, y);
      for (int i = 0; i < 6; i++) y[i] = y[i] * pow(HUGE_, sc3 - sc4);
    }
  }
}

/* ================================================================== */
/* variants: gravity 0 (no gravity) / 1 (Cowling), fluid inner core    */
/* ================================================================== */

/* sph_solid_amat_ng / _cow, src/module_int.f90:564-672 */
typedef struct {
  cplx t11, t12, t21, t22, k11, k22, s11, s12, s22;
} ngc_amat;

static void solid_amat_ngc(const pst *s, double rr, int cow, ngc_amat *A) {
  const ysxo *o = s->o;
  int il = s->il;
  double rrho = SPL(o, RHO, rr, il), racon = SPL(o, ACON, rr, il),
         rccon = SPL(o, CCON, rr, il), rfcon = SPL(o, FCON, rr, il),
         rncon = SPL(o, NCON, rr, il), rlcon = SPL(o, LCON, rr, il),
         rqmu = SPL(o, QMU, rr, il), rxlam = SPL(o, XLAM, rr, il),
         rxa2 = SPL(o, XA2, rr, il);
  double gg = cow ? SPL(o, GRAV, rr, il) : 0.0;
  double zeta = s->zeta, zeta2 = s->zeta2;
  cplx qaa = 1.0 + rxa2 * s->lno, qff = 1.0 + rxlam * s->lno,
       qll = 1.0 + rqmu * s->lno;
  cplx cacon = racon * qaa, cccon = rccon * qaa, cfcon = rfcon * qff,
       cncon = rncon * qll, clcon = rlcon * qll;
  cccon = 1.0 / cccon;
  cplx gamma = cacon - cncon - cfcon * cfcon * cccon;
  double ri = 1.0 / rr;
  A->t11 = (1.0 - 2.0 * cfcon * cccon) * ri;
  A->t12 = zeta * ri * cfcon * cccon;
  A->t21 = -zeta * ri;
  A->t22 = 2.0 * ri;
  A->k11 = cccon;
  A->k22 = 1.0 / clcon;
  if (cow) {
    A->s11 = -s->omega2 * rrho + 4.0 * (gamma - rrho * gg * rr) * ri * ri;
    A->s12 = zeta * (rrho * gg * rr - 2.0 * gamma) * ri * ri;
  } else {
    A->s11 = -s->omega2 * rrho + 4.0 * gamma * ri * ri;
    A->s12 = -2.0 * zeta * gamma * ri * ri;
  }
  A->s22 = -s->omega2 * rrho + (zeta2 * (gamma + cncon) - 2.0 * cncon) * ri * ri;
}

/* sph_minors_solid_derivs_ng / _cow, src/module_int.f90:874-936 */
static void minors_solid_ngc(const pst *s, double rr, const cplx *m, cplx *d,
                             int cow) {
  ngc_amat A;
  solid_amat_ngc(s, rr, cow, &A);
  cplx a11 = A.t11 + A.t22, a22 = A.t11 - A.t22;
  d[0] = a11 * m[0] + A.k22 * m[2] - A.k11 * m[3];
  d[1] = A.s12 * m[0] - A.t21 * m[2] + A.t12 * m[3];
  d[2] = A.s22 * m[0] - 2.0 * A.t12 * m[1] + a22 * m[2] + A.k11 * m[4];
  d[3] = -A.s11 * m[0] + 2.0 * A.t21 * m[1] - a22 * m[3] - A.k22 * m[4];
  d[4] = -2.0 * A.s12 * m[1] + A.s11 * m[2] - A.s22 * m[3] - a11 * m[4];
}
static void minors_solid_ng(const pst *s, double rr, const cplx *m, cplx *d) {
  minors_solid_ngc(s, rr, m, d, 0);
}
static void minors_solid_cow(const pst *s, double rr, const cplx *m, cplx *d) {
  minors_solid_ngc(s, rr, m, d, 1);
}

/* sph_fluid_amat_ng / _cow + sph_minors_fluid_derivs_ng / _cow,
 * src/module_int.f90:720-786, 971-1012 */
static void minors_fluid_ngc(const pst *s, double rr, const cplx *m, cplx *d,
                             int cow) {
  const ysxo *o = s->o;
  int il = s->il;
  double rrho = SPL(o, RHO, rr, il), rccon = SPL(o, CCON, rr, il),
         rxa2 = SPL(o, XA2, rr, il);
  double zeta2 = s->zeta2;
  cplx qaa = 1.0 + rxa2 * s->lno;
  cplx cccon = rccon * qaa;
  cccon = 1.0 / cccon;
  double ri = 1.0 / rr;
  cplx t, k, sv;
  if (cow) {
    double gg = SPL(o, GRAV, rr, il);
    t = (gg * zeta2 * s->omegai2 * ri - 1.0) * ri;
    k = cccon - zeta2 * s->omegai2 * ri * ri / rrho;
    sv = -s->omega2 * rrho + rrho * gg * (gg * zeta2 * ri * s->omegai2 - 4.0) * ri;
  } else {
    t = -ri;
    k = cccon - zeta2 * s->omegai2 * ri * ri / rrho;
    sv = -s->omega2 * rrho;
  }
  d[0] = t * m[0] + k * m[1];
  d[1] = sv * m[0] - t * m[1];
}
static void minors_fluid_ng(const pst *s, double rr, const cplx *m, cplx *d) {
  minors_fluid_ngc(s, rr, m, d, 0);
}
static void minors_fluid_cow(const pst *s, double rr, const cplx *m, cplx *d) {
  minors_fluid_ngc(s, rr, m, d, 1);
}

/* sph_bvec_solid_derivs_ng / _cow, src/module_int.f90:1060-1110 */
static void bvec_solid_ngc(const pst *s, double rr, const cplx *b, cplx *d,
                           int cow) {
  ngc_amat A;
  solid_amat_ngc(s, rr, cow, &A);
  d[0] = -A.s11 * b[2] - A.s12 * b[3] - A.t11 * b[0] - A.t21 * b[1];
  d[1] = -A.s12 * b[2] - A.s22 * b[3] - A.t12 * b[0] - A.t22 * b[1];
  d[2] = -A.k11 * b[0] + A.t11 * b[2] + A.t12 * b[3];
  d[3] = -A.k22 * b[1] + A.t21 * b[2] + A.t22 * b[3];
}
static void bvec_solid_ng(const pst *s, double rr, const cplx *b, cplx *d) {
  bvec_solid_ngc(s, rr, b, d, 0);
}
static void bvec_solid_cow(const pst *s, double rr, const cplx *b, cplx *d) {
  bvec_solid_ngc(s, rr, b, d, 1);
}

/* sph_solid_start_ng, src/module_int.f90:1275-1341 (sbessj2 stub
 * ratio = x / real(2 l + 3) with l = llp1, src/module_function.f90:447-490);
 * y is (4,2) column-major */
static void sph_solid_start_ng(const pst *s, int is, double rr, cplx y[2][4]) {
  const ysxo *o = s->o;
  double rrho = SPL(o, RHO, rr, is), rmu = SPL(o, MU, rr, is),
         rqmu = SPL(o, QMU, rr, is), rkappa = SPL(o, KAPPA, rr, is),
         rqkappa = SPL(o, QKAPPA, rr, is);
  int ll = s->ll, llp1 = s->llp1;
  double zeta = s->zeta;
  cplx cmu = rmu * (1.0 + rqmu * s->lno);
  cplx ckappa = rkappa * (1.0 + rqkappa * s->lno);
  cplx calpha = (ckappa + 1.3333333333 * cmu) / rrho;
  cplx cbeta = cmu / rrho;
  cplx arg = s->omega * rr / csqrt(calpha);
  cplx ratio = arg / (double)(float)(2 * llp1 + 3);
  cplx j1 = 1.0, j2 = ratio;
  y[0][0] = (double)ll * j1 - arg * j2;
  y[0][1] = zeta * j1;
  y[0][2] = -rrho * calpha * arg * arg * j1 +
            2.0 * cmu * ((double)(ll * (ll - 1)) * j1 + 2.0 * arg * j2);
  y[0][2] = y[0][2] / rr;
  y[0][3] = 2.0 * zeta * cmu * ((double)(ll - 1) * j1 - arg * j2);
  y[0][3] = y[0][3] / rr;
  arg = s->omega * rr / csqrt(cbeta);
  ratio = arg / (double)(float)(2 * llp1 + 3);
  j1 = 1.0;
  j2 = ratio;
  y[1][0] = (double)(-ll * llp1) * j1;
  y[1][1] = -zeta * llp1 * j1 + arg * j2;
  y[1][2] = -2.0 * cmu * ((double)(ll * (ll * ll - 1)) * j1 -
                          (double)(ll * llp1) * arg * j2);
  y[1][2] = y[1][2] / rr;
  y[1][3] = zeta * cmu * (arg * arg * j1 - 2.0 * (double)(ll * ll - 1) * j1 -
                          2.0 * arg * j2);
  y[1][3] = y[1][3] / rr;
}

/* sph_fluid_start_ng, src/module_int.f90:1407-1447 */
static void sph_fluid_start_ng(const pst *s, int is, double rr, cplx *y) {
  const ysxo *o = s->o;
  double rrho = SPL(o, RHO, rr, is), rkappa = SPL(o, KAPPA, rr, is),
         rqkappa = SPL(o, QKAPPA, rr, is);
  cplx ckappa = rkappa * (1.0 + rqkappa * s->lno);
  cplx calpha = ckappa / rrho;
  cplx arg = s->omega * rr / csqrt(calpha);
  cplx ratio = arg / (double)(float)(2 * s->llp1 + 3);
  y[0] = (double)s->ll - arg * ratio;
  y[1] = -rrho * calpha * arg * arg;
  y[1] = y[1] / rr;
}

/* sph_minor_solid_start_ng, src/module_int.f90:1489-1506 */
static void sph_minor_solid_start_ng(const pst *s, int is, double rr, cplx *m) {
  static const int rows[5][2] = {{1, 2}, {1, 3}, {1, 4}, {2, 3}, {3, 4}};
  cplx y[2][4];
  sph_solid_start_ng(s, is, rr, y);
  for (int k = 0; k < 5; k++) {
    int p = rows[k][0] - 1, q = rows[k][1] - 1;
    m[k] = det2(y[0][p], y[0][q], y[1][p], y[1][q]);
  }
}

/* b_solid_start_ngcow, src/module_int.f90:1596-1615 */
static void b_solid_start_ngc(const cplx *m, const cplx *sv, cplx *b) {
  b[0] = -m[4] * sv[1] - m[1] * sv[2] - m[3] * sv[3];
  b[1] = m[4] * sv[0] - m[2] * sv[2] + m[1] * sv[3];
  b[2] = m[1] * sv[0] + m[2] * sv[1] - m[0] * sv[3];
  b[3] = m[3] * sv[0] - m[1] * sv[1] + m[0] * sv[2];
}

/* sph_minors_solid_2_fluid_ngcow / fluid_2_solid_ngcow, sph_delta_solid_ngcow,
 * sph_ysol_ngcow: src/module_int.f90 (after :1638) */
static void solid_2_fluid_ngc(const cplx *m1, cplx *m2) {
  m2[0] = -m1[2];
  m2[1] = -m1[4];
}
static void fluid_2_solid_ngc(const cplx *m1, cplx *m2) {
  m2[0] = m1[0];
  m2[3] = -m1[1];
  m2[1] = m2[2] = m2[4] = 0.0;
}
static cplx delta_solid_ngc(const cplx *m1, const cplx *m2) {
  return m1[4] * m2[0] + 2.0 * m1[1] * m2[1] + m1[3] * m2[2] + m1[2] * m2[3] +
         m1[0] * m2[4];
}
static void ysol_ngc(const cplx *m, const cplx *b, cplx *y) {
  y[0] = m[0] * b[1] + m[1] * b[2] + m[2] * b[3];
  y[1] = -m[0] * b[0] + m[3] * b[2] - m[1] * b[3];
  y[2] = -m[1] * b[0] - m[3] * b[1] + m[4] * b[3];
  y[3] = -m[2] * b[0] + m[1] * b[1] - m[4] * b[2];
}

/* sph_solver_ng / _cow / _ng_fic / _cow_fic with ss = identity(4),
 * src/module_solver.f90:827-1195, 1386-1740.  yout (4,4) column-major. */
static void sph_solver_ngc(pst *s, int l, cplx w, double rr, double rs, int ir,
                           int is, int cow, int fic, cplx *yout, int *start) {
  const ysxo *o = s->o;
  const double *r = R(o);
  int nic = o->nic, noc = o->noc, nsl = o->nsl, nknot = o->nknot;
  derivs_fn dsol = cow ? minors_solid_cow : minors_solid_ng;
  derivs_fn dflu = cow ? minors_fluid_cow : minors_fluid_ng;
  derivs_fn dbv = cow ? bvec_solid_cow : bvec_solid_ng;
  set_l(s, l);
  s->ats = o->ats_in;
  set_omega(s, w);
  alpha_steps(s);
  int sc1 = 0, sc2 = 0, sc3 = 0, sc4, i1;
  cplx mf1[2], mf2[2], ms1[5], ms2[5], ms3[5], b[4], delta;
  int up = rr >= rs;
  int iup = up ? is : ir;
  double rup = up ? rs : rr;
  i1 = sph_start_level(o, s->ll, creal(s->omega), iup);
  *start = i1;
  double rt = r[i1];
  /* the solid-core solvers take the small first radius only when the
   * receiver is above the source; the fluid-core ones always */
  if ((up || fic) && i1 == 1) rt = fmin(0.1 * s->dr, 0.1 * r[2]);
  if (i1 >= iup) {
    i1 = iup - 1;
    rt = r[i1];
  }
  if (i1 <= nic && !fic) {
    sph_minor_solid_start_ng(s, i1, rt, ms1);
    yint(s, i1, nic, rt, r[nic], ms1, 5, dsol, &sc1);
    solid_2_fluid_ngc(ms1, mf1);
    yint(s, nic + 1, noc, r[nic], r[noc], mf1, 2, dflu, &sc1);
    fluid_2_solid_ngc(mf1, ms1);
    yint(s, noc + 1, iup, r[noc + 1], rup, ms1, 5, dsol, &sc1);
  } else if (i1 <= noc) {
    sph_fluid_start_ng(s, i1, rt, mf1);
    yint(s, i1, noc, rt, r[noc], mf1, 2, dflu, &sc1);
    fluid_2_solid_ngc(mf1, ms1);
    yint(s, noc + 1, iup, r[noc + 1], rup, ms1, 5, dsol, &sc1);
  } else if (i1 <= nsl) {
    sph_minor_solid_start_ng(s, i1, rt, ms1);
    yint(s, i1, iup, rt, rup, ms1, 5, dsol, &sc1);
  }
  memcpy(ms3, ms1, sizeof ms3);
  int ibot = up ? ir : is;
  double rbot = up ? rr : rs;
  if (up && rr > rs) yint(s, is, ir, rs, rr, ms3, 5, dsol, &sc3);
  if (!up && rs > rr) yint(s, ir, is, rr, rs, ms3, 5, dsol, &sc3);
  if (o->ocean) {
    mf2[0] = 1.0;
    mf2[1] = 0.0;
    yint(s, nknot - 1, nsl + 1, r[nknot], r[nsl + 1], mf2, 2, dflu, &sc2);
    fluid_2_solid_ngc(mf2, ms2);
    yint(s, nsl, ibot, r[nsl], rbot, ms2, 5, dsol, &sc2);
  } else {
    ms2[0] = 1.0;
    for (int i = 1; i < 5; i++) ms2[i] = 0.0;
    yint(s, nknot, ibot, r[nknot], rbot, ms2, 5, dsol, &sc2);
  }
  delta = delta_solid_ngc(ms3, ms2);
  for (int c = 0; c < 4; c++) {
    cplx ssv[4] = {0, 0, 0, 0};
    ssv[c] = 1.0;
    sc4 = 0;
    cplx *y = yout + 4 * c;
    if (up) {
      b_solid_start_ngc(ms1, ssv, b);
      yint(s, is, ir, rs, rr, b, 4, dbv, &sc4);
      ysol_ngc(ms2, b, y);
      for (int i = 0; i < 4; i++) {
        y[i] = y[i] / delta;
        y[i] = y[i] * pow(HUGE_, sc3 - sc4);
      }
    } else {
      b_solid_start_ngc(ms2, ssv, b);
      for (int i = 0; i < 4; i++) b[i] = -b[i] / delta;
      yint(s, is, ir, rs, rr, b, 4, dbv, &sc4);
      ysol_ngc(ms1, b, y);
      for (int i = 0; i < 4; i++) y[i] = y[i] * pow(HUGE_, sc3 - sc4);
    }
  }
}

/* sph_solver_fic (grav = 2, fluid inner core), src/module_solver.f90:
 * 1203-1382.  The receiver-below-source branch repeats sph_solver's
 * solid-core start (a solid start in a fluid inner core, :1323-1331);
 * it is reproduced as written. */
static void sph_solver_fic(pst *s, int l, cplx w, double rr, double rs, int ir,
                           int is, cplx *yout, int *start) {
  const ysxo *o = s->o;
  const double *r = R(o);
  int nic = o->nic, noc = o->noc, nsl = o->nsl, nknot = o->nknot;
  set_l(s, l);
  s->ats = o->ats_in;
  set_omega(s, w);
  alpha_steps(s);
  int sc1 = 0, sc2 = 0, sc3 = 0, sc4, i1;
  cplx mf1[5], mf2[5], ms1[14], ms2[14], ms3[14], b[14], delta;
  int up = rr >= rs;
  int iup = up ? is : ir;
  double rup = up ? rs : rr;
  i1 = sph_start_level(o, s->ll, creal(s->omega), iup);
  *start = i1;
  double rt = r[i1];
  if (i1 == 1) rt = fmin(0.1 * s->dr, 0.1 * r[2]);
  if (i1 >= iup) {
    i1 = iup - 1;
    rt = r[i1];
  }
  if (i1 <= nic && !up) {
    sph_minor_solid_start(s, i1, rt, ms1);
    yint(s, i1, nic, rt, r[nic], ms1, 14, sph_minors_solid_derivs, &sc1);
    solid_2_fluid(ms1, mf1);
    yint(s, nic + 1, noc, r[nic], r[noc], mf1, 5, sph_minors_fluid_derivs,
         &sc1);
    fluid_2_solid(mf1, ms1);
    yint(s, noc + 1, iup, r[noc + 1], rup, ms1, 14, sph_minors_solid_derivs,
         &sc1);
  } else if (i1 <= noc) {
    sph_minor_fluid_start(s, i1, rt, mf1);
    yint(s, i1, noc, rt, r[noc], mf1, 5, sph_minors_fluid_derivs, &sc1);
    fluid_2_solid(mf1, ms1);
    yint(s, noc + 1, iup, r[noc + 1], rup, ms1, 14, sph_minors_solid_derivs,
         &sc1);
  } else if (i1 <= nsl) {
    sph_minor_solid_start(s, i1, rt, ms1);
    yint(s, i1, iup, rt, rup, ms1, 14, sph_minors_solid_derivs, &sc1);
  }
  memcpy(ms3, ms1, sizeof ms3);
  int ibot = up ? ir : is;
  double rbot = up ? rr : rs;
  if (up && rr > rs)
    yint(s, is, ir, rs, rr, ms3, 14, sph_minors_solid_derivs, &sc3);
  if (!up && rs > rr)
    yint(s, ir, is, rr, rs, ms3, 14, sph_minors_solid_derivs, &sc3);
  if (o->ocean) {
    mf2[0] = 1.0;
    for (int i = 1; i < 5; i++) mf2[i] = 0.0;
    yint(s, nknot - 1, nsl + 1, r[nknot], r[nsl + 1], mf2, 5,
         sph_minors_fluid_derivs, &sc2);
    fluid_2_solid(mf2, ms2);
    yint(s, nsl, ibot, r[nsl], rbot, ms2, 14, sph_minors_solid_derivs, &sc2);
  } else {
    ms2[0] = 1.0;
    for (int i = 1; i < 14; i++) ms2[i] = 0.0;
    yint(s, up ? nknot - 1 : nknot, ibot, r[nknot], rbot, ms2, 14,
         sph_minors_solid_derivs, &sc2);
  }
  delta = sph_delta_solid(ms3, ms2);
  for (int c = 0; c < 4; c++) {
    cplx ssv[6];
    for (int i = 0; i < 6; i++) ssv[i] = S6[c][i];
    sc4 = 0;
    cplx *y = yout + 6 * c;
    if (up) {
      b_solid_start(ms1, ssv, b);
      yint(s, is, ir, rs, rr, b, 14, sph_bvec_solid_derivs, &sc4);
      sph_ysol(ms2, b, y);
      for (int i = 0; i < 6; i++) {
        y[i] = -y[i] / delta;
        y[i] = y[i] * pow(HUGE_, sc3 - sc4);
      }
    } else {
      b_solid_start(ms2, ssv, b);
      for (int i = 0; i < 14; i++) b[i] = -b[i] / delta;
      yint(s, is, ir, rs, rr, b, 14, sph_bvec_solid_derivs, &sc4);
      sph_ysol(ms1, b, y);
      for (int i = 0; i < 6; i++) y[i] = y[i] * pow(HUGE_, sc3 - sc4);
    }
  }
}

/* rad_ystart_ng, src/module_int.f90:455-489 */
static void rad_ystart_ng(const pst *s, int is, double rr, cplx *y) {
  const ysxo *o = s->o;
  double rrho = SPL(o, RHO, rr, is), rmu = SPL(o, MU, rr, is),
         rqmu = SPL(o, QMU, rr, is), rkappa = SPL(o, KAPPA, rr, is),
         rqkappa = SPL(o, QKAPPA, rr, is);
  cplx cmu = rmu * (1.0 + rqmu * s->lno);
  cplx ckappa = rkappa * (1.0 + rqkappa * s->lno);
  cplx calpha = (ckappa + 4.0 * cmu / 3.0);
  cplx gamma = s->omega2;
  gamma = rrho * gamma / calpha;
  cplx arg = gamma * rr;
  cplx ratio = arg / 5.0;
  y[0] = rr * ratio;
  y[1] = rr * calpha * gamma - 4.0 * cmu * ratio;
}

/* rad_solver_ng with ss = identity(2), src/module_solver.f90:335-441.  Its
 * RHS, rad_derivs_ng (src/module_int.f90:377-392), calls rad_amat -- the
 * self-gravitating coefficients -- not rad_amat_ng; reproduced as written. */
static void rad_solver_ng(pst *s, cplx w, double rr, double rs, int ir, int is,
                          cplx *yout) {
  const ysxo *o = s->o;
  const double *r = R(o);
  int nknot = o->nknot, nsl = o->nsl;
  set_l(s, 0);
  set_omega(s, w);
  s->ats = o->ats_in;
  set_omega(s, w);
  alpha_steps(s);
  int sc1 = 0, sc2 = 0, sc3 = 0;
  cplx y1[2], y2[2], y3[2], delta;
  int i1 = 1;
  double rt = fmin(0.1 * s->dr, 0.1 * r[2]);
  rad_ystart_ng(s, i1, rt, y1);
  if (rr >= rs) {
    yint(s, i1, is, rt, rs, y1, 2, rad_derivs, &sc1);
    y2[0] = 1.0;
    y2[1] = 0.0;
    if (rr != r[nknot]) yint(s, nknot - 1, ir, r[nknot], rr, y2, 2, rad_derivs, &sc2);
    y3[0] = y2[0];
    y3[1] = y2[1];
    if (rs != rr) yint(s, ir, is, rr, rs, y3, 2, rad_derivs, &sc3);
  } else {
    yint(s, i1, ir, rt, rr, y1, 2, rad_derivs, &sc1);
    y2[0] = 1.0;
    y2[1] = 0.0;
    if (rs != r[nsl]) yint(s, nsl - 1, is, r[nsl], rs, y2, 2, rad_derivs, &sc2);
    y3[0] = y2[0];
    y3[1] = y2[1];
    if (rr != rs) yint(s, is, ir, rs, rr, y3, 2, rad_derivs, &sc3);
  }
  delta = y3[1] * y1[0] - y3[0] * y1[1];
  if (cabs(delta) < 1.0e-100)
    delta = 0.0;
  else
    delta = 1.0 / delta;
  for (int i = 0; i < 2; i++) {
    cplx ss1 = (i == 0) ? 1.0 : 0.0, ss2 = (i == 1) ? 1.0 : 0.0;
    if (rr >= rs) {
      cplx b = y1[1] * ss1 - y1[0] * ss2;
      yout[2 * i] = -delta * y2[0] * b;
      yout[2 * i + 1] = -delta * y2[1] * b;
    } else {
      cplx b = y2[1] * ss1 - y2[0] * ss2;
      yout[2 * i] = -delta * y1[0] * b;
      yout[2 * i + 1] = -delta * y1[1] * b;
    }
  }
  if (rr >= rs) {
    if (sc3 != 0)
      for (int i = 0; i < 4; i++) yout[i] = yout[i] * pow(HUGE_, sc3);
  } else {
    /* do i = 1, scale3: nothing when scale3 <= 0 (src/module_solver.f90:433-437) */
    for (int k = 1; k <= sc3; k++)
      for (int i = 0; i < 4; i++) yout[i] = yout[i] * HUGE_;
  }
}

/* source_vector_tor, src/module_solver.f90:145-215; svt (5,2) */
static void source_vector_tor(pst *s, int l, cplx w, int is, double rs,
                              const double *mm, cplx *svt) {
  const ysxo *o = s->o;
  set_l(s, l);
  s->ats = o->ats_in;
  set_omega(s, w);
  double rlcon = SPL(o, LCON, rs, is), rqmu = SPL(o, QMU, rs, is);
  cplx qll = 1.0 + rqmu * s->lno;
  cplx clcon = rlcon * qll;
  for (int i = 0; i < 10; i++) svt[i] = 0.0;
  if (l == 1) {
    svt[1] = 0.5 * rs * (-mm[2] + I * mm[1]) / clcon;
    svt[3] = 0.5 * rs * (mm[2] + I * mm[1]) / clcon;
  } else if (l > 1) {
    svt[1] = 0.5 * rs * (-mm[2] + I * mm[1]) / clcon;
    svt[3] = 0.5 * rs * (mm[2] + I * mm[1]) / clcon;
    svt[5 + 0] = 0.25 * s->szeta2m2 * (-I * mm[3] + I * mm[5] + 2.0 * mm[4]);
    svt[5 + 4] = 0.25 * s->szeta2m2 * (I * mm[3] - I * mm[5] + 2.0 * mm[4]);
  }
  cplx den = I * w * rs * rs;
  for (int i = 0; i < 10; i++) svt[i] = s->nu * svt[i] / den;
}

/* source_vector_sph, src/module_solver.f90:1747-1847; svs (5,6).
 * qaa, qff and qll are declared real(dp) there (:1771-1772), so only the
 * real part of each complex dispersion factor survives. */
static void source_vector_sph(pst *s, int l, cplx w, int is, double rs,
                              const double *mm, cplx *svs) {
  const ysxo *o = s->o;
  set_l(s, l);
  s->ats = o->ats_in;
  set_omega(s, w);
  double rlcon = SPL(o, LCON, rs, is), rccon = SPL(o, CCON, rs, is),
         rfcon = SPL(o, FCON, rs, is), rqmu = SPL(o, QMU, rs, is),
         rxlam = SPL(o, XLAM, rs, is), rxa2 = SPL(o, XA2, rs, is);
  double qaa = creal(1.0 + rxa2 * s->lno);
  double qff = creal(1.0 + rxlam * s->lno);
  double qll = creal(1.0 + rqmu * s->lno);
  cplx cccon = rccon * qaa;
  cccon = 1.0 / cccon;
  cplx cfcon = rfcon * qff;
  cplx clcon = rlcon * qll;
  double zeta = s->zeta;
#define SVS(i, j) svs[((j)-1) * 5 + ((i)-1)]
  for (int i = 0; i < 30; i++) svs[i] = 0.0;
  if (l == 1) {
    SVS(2, 2) = 0.5 * rs * (mm[1] + I * mm[2]) / clcon;
    SVS(3, 1) = rs * mm[0] * cccon;
    SVS(3, 4) = -mm[3] - mm[5] + 2.0 * mm[0] * cfcon * cccon;
    SVS(3, 5) = 0.5 * zeta * (mm[3] + mm[5]) - zeta * mm[0] * cfcon * cccon;
    SVS(4, 2) = 0.5 * rs * (-mm[1] + I * mm[2]) / clcon;
  } else {
    SVS(1, 5) = 0.25 * s->szeta2m2 * (-mm[3] + mm[5] - 2.0 * I * mm[4]);
    SVS(2, 2) = 0.5 * rs * (mm[1] + I * mm[2]) / clcon;
    SVS(3, 1) = rs * mm[0] * cccon;
    SVS(3, 4) = -mm[3] - mm[5] + 2.0 * mm[0] * cfcon * cccon;
    SVS(3, 5) = 0.5 * zeta * (mm[3] + mm[5]) - zeta * mm[0] * cfcon * cccon;
    SVS(4, 2) = 0.5 * rs * (-mm[1] + I * mm[2]) / clcon;
    SVS(5, 5) = 0.25 * s->szeta2m2 * (-mm[3] + mm[5] + 2.0 * I * mm[4]);
  }
#undef SVS
  cplx den = I * w * rs * rs;
  for (int i = 0; i < 30; i++) svs[i] = s->nu * svs[i] / den;
}

/* source_vector_rad, src/module_solver.f90:566-627 */
static void source_vector_rad(pst *s, cplx w, int is, double rs,
                              const double *mm, cplx *s0) {
  const ysxo *o = s->o;
  set_l(s, 0);
  s->ats = o->ats_in;
  set_omega(s, w);
  double rccon = SPL(o, CCON, rs, is), rfcon = SPL(o, FCON, rs, is),
         rxa2 = SPL(o, XA2, rs, is), rxlam = SPL(o, XLAM, rs, is);
  cplx qaa = 1.0 + rxa2 * s->lno, qff = 1.0 + rxlam * s->lno;
  cplx cccon = rccon * qaa;
  cccon = 1.0 / cccon;
  cplx cfcon = rfcon * qff;
  s0[0] = rs * mm[0] * cccon;
  s0[1] = -mm[3] - mm[5] + 2.0 * mm[0] * cfcon * cccon;
  cplx den = I * w * rs * rs;
  s0[0] = s->nu * s0[0] / den;
  s0[1] = s->nu * s0[1] / den;
}

/* ================================================================== */
/* run set-up: src/yspec.f90:141-215, src/module_fourier.f90:362-397    */
/* ================================================================== */

/* fcal, src/module_fourier.f90:362-397 (single-precision log) */
static int fcal(double *f1, double *f2, double dt, double tout, int mex,
                int qex, double *df, double *ep, int *nt, int *i1, int *i2) {
  double fn = 0.5 / dt;
  if (fn < *f2) return 1;
  *ep = mex / tout;
  *df = *ep / (TWOPI_D * qex);
  *nt = (int)(1.0 / (*df * dt));
  int ne = (int)(logf((float)*nt) / logf(2.0f) + 1.0f);
  *nt = 1 << ne;
  *df = 1.0 / (*nt * dt);
  *i1 = (int)floor(*f1 / *df);
  if (*i1 < 2) *i1 = 2;
  *f1 = (*i1 - 1) * *df;
  *i2 = (int)floor(*f2 / *df) + 2;
  *f2 = (*i2 - 1) * *df;
  return 0;
}

int ysxo_fcal(double f1, double f2, double dt, double tout, int *iout,
              double *fout) {
  double df, ep;
  int nt, i1, i2;
  if (fcal(&f1, &f2, dt, tout, 5, 1, &df, &ep, &nt, &i1, &i2)) return 1;
  iout[0] = nt; iout[1] = i1; iout[2] = i2;
  fout[0] = df; fout[1] = ep; fout[2] = f1; fout[3] = f2;
  return 0;
}

int ysxo_set_run(ysxo *o, const int *ip, const double *fp, const double *lat,
                 const double *lon) {
  free_run(o);
  if (!o->nknot) return fail(o, "no model");
  o->lmin = ip[0]; o->lmax = ip[1]; o->ats_in = ip[2]; o->grav = ip[3];
  o->out = ip[4]; o->cor = ip[5]; o->nr = ip[6];
  if (o->grav < 0 || o->grav > 2) return fail(o, "gravity switch must be 0, 1 or 2");
  double f1 = fp[0], f2 = fp[1], tout = fp[2], dt = fp[3];
  double f11 = fp[4], f12 = fp[5], f21 = fp[6], f22 = fp[7];
  double sdep = fp[8], lats = fp[9], lons = fp[10], rdep = fp[17];
  int nr = o->nr, lmax = o->lmax;
  /* geometry (src/yspec.f90:144-174) */
  o->delta = calloc((size_t)nr, sizeof(double));
  o->azep = calloc((size_t)nr, sizeof(double));
  o->azst = calloc((size_t)nr, sizeof(double));
  size_t nt3 = (size_t)(lmax + 1) * 3 * nr;
  o->xa = calloc(nt3, sizeof(double));
  o->xpa = calloc(nt3, sizeof(double));
  o->xca = calloc(nt3, sizeof(double));
  o->eia = calloc((size_t)nr, sizeof(cplx));
  if (lats == 90.0) lats = 89.98;
  for (int k = 0; k < nr; k++) {
    double d3[3];
    ysxo_delaz(lat[k], lon[k], lats, lons, d3);
    o->delta[k] = d3[0];
    o->azep[k] = TWOPI_D - d3[1];
    o->azst[k] = PI_D - d3[2];
  }
  for (int k = 0; k < nr; k++) {
    for (int l = 0; l <= lmax; l++) {
      int m = (l == 0) ? 0 : (l == 1 ? 1 : 2);
      double x[3], xp[3], xc[3];
      ysxo_legendre(o->delta[k], l, m, x, xp, xc);
      for (int mm = 0; mm <= m; mm++) {
        size_t idx = (size_t)l + (size_t)(lmax + 1) * (mm + 3 * (size_t)k);
        o->xa[idx] = x[mm];
        o->xpa[idx] = xp[mm];
        o->xca[idx] = xc[mm];
      }
    }
    o->eia[k] = cexp(I * o->azst[k]);
  }
  /* normalisation (src/yspec.f90:179-189) */
  double *r = R(o);
  f1 = f1 * o->t_norm / 1000.0;
  f2 = f2 * o->t_norm / 1000.0;
  o->dt = dt / o->t_norm;
  o->rs = r[o->nknot] - sdep * 1000.0 / R_NORM;
  o->rr = r[o->nknot] - rdep * 1000.0 / R_NORM;
  o->tout = tout * 60.0 / o->t_norm;
  for (int i = 0; i < 6; i++) o->mm[i] = fp[11 + i] / o->moment_norm;
  o->f11 = f11 / 1000.0 * o->t_norm;
  o->f12 = f12 / 1000.0 * o->t_norm;
  o->f21 = f21 / 1000.0 * o->t_norm;
  o->f22 = f22 / 1000.0 * o->t_norm;
  if (fcal(&f1, &f2, o->dt, o->tout, 5, 1, &o->df, &o->ep, &o->nt, &o->i1,
           &o->i2))
    return fail(o, "fmax above the Nyquist frequency");
  o->dw = TWOPI_D * o->df;
  if (o->i1 == o->i2) return fail(o, "too few frequency steps");
  o->wmin = (o->i1 - 1) * o->dw;
  o->nw = o->i2 - o->i1 + 1;
  o->is = find_radius_index_down(o, o->rs);
  if (o->is == 0) return fail(o, "source not in model!");
  if (o->is > o->nsl) return fail(o, "source in the ocean!");
  o->ir = find_radius_index_down(o, o->rr);
  if (o->ir == 0) return fail(o, "receiver not in model!");
  if (o->ir > o->nsl) return fail(o, "receiver in the ocean!");
  /* nout (src/yspec.f90:479-486) */
  o->nout = o->nt;
  for (int i = 1; i <= o->nt; i++) {
    double t = (i - 1) * o->dt;
    if (t > o->tout) { o->nout = i; break; }
  }
  o->have_run = 1;
  return 0;
}

int ysxo_run_info(const ysxo *o, int *iout, double *fout) {
  if (!o->have_run) return 1;
  iout[0] = o->nw; iout[1] = o->nt; iout[2] = o->i1; iout[3] = o->i2;
  iout[4] = o->is; iout[5] = o->ir; iout[6] = o->nout;
  fout[0] = o->rs; fout[1] = o->rr; fout[2] = o->dt; fout[3] = o->tout;
  fout[4] = o->ep; fout[5] = o->df; fout[6] = o->wmin; fout[7] = o->dw;
  fout[8] = o->f11; fout[9] = o->f12; fout[10] = o->f21; fout[11] = o->f22;
  return 0;
}

int ysxo_geometry(const ysxo *o, double *delta, double *azep, double *azst,
                  double *xa, double *xpa, double *xca) {
  if (!o->have_run) return 1;
  size_t nt3 = (size_t)(o->lmax + 1) * 3 * o->nr;
  memcpy(delta, o->delta, sizeof(double) * o->nr);
  memcpy(azep, o->azep, sizeof(double) * o->nr);
  memcpy(azst, o->azst, sizeof(double) * o->nr);
  memcpy(xa, o->xa, sizeof(double) * nt3);
  memcpy(xpa, o->xpa, sizeof(double) * nt3);
  memcpy(xca, o->xca, sizeof(double) * nt3);
  return 0;
}

/* ================================================================== */
/* one (l, w) pair and the receiver sum: src/yspec.f90:254-421          */
/* ================================================================== */

typedef struct {
  cplx svt[10], ytor[4], svs[30], ysph[24], sr0[2], yrad[4], uu[5], vv[5],
      ww[5];
  int starts[2];
} pair_out;

static void do_pair(const ysxo *o, int l, int j, pair_out *P) {
  pst s;
  memset(&s, 0, sizeof s);
  s.o = o;
  memset(P, 0, sizeof *P);
  P->starts[0] = P->starts[1] = -1;
  double rr = o->rr, rs = o->rs;
  double zeta = (double)sqrtf((float)(l * (l + 1)));
  cplx w = o->wmin + (j - 1) * o->dw - I * o->ep;
  if (l == 0) {
    source_vector_rad(&s, w, o->is, rs, o->mm, P->sr0);
    /* src/yspec.f90:271-276 */
    if (o->grav == 0)
      rad_solver_ng(&s, w, rr, rs, o->ir, o->is, P->yrad);
    else
      rad_solver_wg(&s, w, rr, rs, P->yrad);
    cplx u3 = (P->sr0[0] * P->yrad[0] + P->sr0[1] * P->yrad[2]) / rr;
    if (o->out == 1) u3 = I * w * u3;
    if (o->out == 2) u3 = -w * w * u3;
    P->uu[2] = u3;
    return;
  }
  if (l > 1) {
    cplx yt[4];
    source_vector_tor(&s, l, w, o->is, rs, o->mm, P->svt);
    tor_solver(&s, l, w, rr, rs, o->ir, o->is, P->ytor, &P->starts[0]);
    for (int i = 0; i < 4; i++) yt[i] = P->ytor[i];
    if (o->out == 1)
      for (int i = 0; i < 4; i++) yt[i] = I * w * yt[i];
    if (o->out == 2)
      for (int i = 0; i < 4; i++) yt[i] = -w * w * yt[i];
    for (int m = 0; m < 5; m++)
      P->ww[m] = (P->svt[m] * yt[0] + P->svt[5 + m] * yt[2]) / (zeta * rr);
  }
  source_vector_sph(&s, l, w, o->is, rs, o->mm, P->svs);
  /* solver choice and the 4-vector layout, src/yspec.f90:342-374 */
  if (o->grav == 2) {
    if (o->fic)
      sph_solver_fic(&s, l, w, rr, rs, o->ir, o->is, P->ysph, &P->starts[1]);
    else
      sph_solver(&s, l, w, rr, rs, o->ir, o->is, P->ysph, &P->starts[1]);
  } else {
    cplx yng[16];
    sph_solver_ngc(&s, l, w, rr, rs, o->ir, o->is, o->grav == 1, o->fic, yng,
                   &P->starts[1]);
    for (int c = 0; c < 4; c++) {
      cplx *y6 = P->ysph + 6 * c;
      const cplx *y4 = yng + 4 * c;
      y6[0] = y4[0];
      y6[1] = y4[1];
      y6[2] = 0.0;
      y6[3] = y4[2];
      y6[4] = y4[3];
      y6[5] = 0.0;
    }
  }
  cplx y[24];
  memcpy(y, P->ysph, sizeof y);
#define Y(i, c) y[(c)*6 + (i)-1]
  if (o->out == 1) {
    for (int c = 0; c < 4; c++)
      for (int i = 1; i <= 3; i++) Y(i, c) = I * w * Y(i, c);
  } else if (o->out == 2) {
    if (o->cor == 0) {
      for (int c = 0; c < 4; c++)
        for (int i = 1; i <= 3; i++) Y(i, c) = -w * w * Y(i, c);
    } else {
      double g = o->a[A_GRAV][o->ir];
      for (int c = 0; c < 4; c++) {
        cplx c1 = -((l + 1) * Y(3, c) + 2.0 * g * Y(1, c)) / rr;
        cplx c2 = zeta * (Y(3, c) + g * Y(1, c)) / rr;
        Y(1, c) = -w * w * Y(1, c) + c1;
        Y(2, c) = -w * w * Y(2, c) + c2;
      }
    }
  }
  for (int m = 0; m < 5; m++) {
    const cplx *sv = P->svs;
    P->uu[m] = (Y(1, 0) * sv[0 * 5 + m] + Y(1, 1) * sv[1 * 5 + m] +
                Y(1, 2) * sv[3 * 5 + m] + Y(1, 3) * sv[4 * 5 + m]) /
               rr;
    P->vv[m] = (Y(2, 0) * sv[0 * 5 + m] + Y(2, 1) * sv[1 * 5 + m] +
                Y(2, 2) * sv[3 * 5 + m] + Y(2, 3) * sv[4 * 5 + m]) /
               (zeta * rr);
  }
#undef Y
}

int ysxo_pair(ysxo *o, int l, int j, double *rec, int *starts) {
  if (!o->have_run) return fail(o, "no run");
  pair_out P;
  do_pair(o, l, j, &P);
  cplx *c = (cplx *)rec;
  int n = 0;
#define CP(arr, cnt) \
  for (int i = 0; i < (cnt); i++) c[n++] = P.arr[i];
  CP(svt, 10) CP(ytor, 4) CP(svs, 30) CP(ysph, 24) CP(sr0, 2) CP(yrad, 4)
  CP(uu, 5) CP(vv, 5) CP(ww, 5)
#undef CP
  starts[0] = P.starts[0];
  starts[1] = P.starts[1];
  return 0;
}

/* integer power of a complex number as used for eia(k)**m */
static cplx cipow(cplx z, int m) {
  if (m == 0) return 1.0;
  cplx p = 1.0;
  int a = m < 0 ? -m : m;
  for (int i = 0; i < a; i++) p *= z;
  return m < 0 ? 1.0 / p : p;
}

/* receiver sum for one pair (src/yspec.f90:290-293, 318-335, 402-419) */
static void sum_pair(const ysxo *o, int l, int j, const pair_out *P,
                     cplx *ur, cplx *ut, cplx *up) {
  int nr = o->nr, nw = o->nw, L1 = o->lmax + 1;
  for (int k = 0; k < nr; k++) {
    size_t jk = (size_t)(j - 1) + (size_t)nw * k;
    if (l == 0) {
      ur[jk] += P->uu[2] * o->xa[(size_t)L1 * 3 * k];
      continue;
    }
    for (int m = -2; m <= 2; m++) {
      int am = m < 0 ? -m : m;
      double sg = (m < 0 && (am & 1)) ? -1.0 : 1.0;
      size_t idx = (size_t)l + (size_t)L1 * (am + 3 * (size_t)k);
      double x = sg * o->xa[idx], xp = sg * o->xpa[idx], xc = sg * o->xca[idx];
      cplx ei = cipow(o->eia[k], m);
      if (l > 1 && m != 0) {
        ut[jk] = ut[jk] + I * (double)m * P->ww[m + 2] * x * ei;
        up[jk] = up[jk] - P->ww[m + 2] * xp * ei;
      }
    }
    for (int m = -2; m <= 2; m++) {
      int am = m < 0 ? -m : m;
      double sg = (m < 0 && (am & 1)) ? -1.0 : 1.0;
      size_t idx = (size_t)l + (size_t)L1 * (am + 3 * (size_t)k);
      double x = sg * o->xa[idx], xp = sg * o->xpa[idx], xc = sg * o->xca[idx];
      cplx ei = cipow(o->eia[k], m);
      ur[jk] = ur[jk] + P->uu[m + 2] * x * ei;
      ut[jk] = ut[jk] + P->vv[m + 2] * xp * ei;
      up[jk] = up[jk] + I * (double)m * P->vv[m + 2] * xc * ei;
    }
  }
}

int ysxo_spectra(ysxo *o, int l0, int l1, double *urd, double *utd,
                 double *upd, int nthreads) {
  if (!o->have_run) return fail(o, "no run");
  if (l0 < o->lmin || l1 > o->lmax) return fail(o, "l range outside run");
  cplx *ur = (cplx *)urd, *ut = (cplx *)utd, *up = (cplx *)upd;
  size_t n = (size_t)o->nw * o->nr;
  for (size_t i = 0; i < n; i++) ur[i] = ut[i] = up[i] = 0.0;
  int nw = o->nw;
  (void)nthreads;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int j = 1; j <= nw; j++) {
    pair_out P;
    for (int l = l0; l <= l1; l++) {
      do_pair(o, l, j, &P);
      sum_pair(o, l, j, &P, ur, ut, up);
    }
  }
  return 0;
}

/* ================================================================== */
/* inverse FFT and post-processing: src/module_fourier.f90:7-145,       */
/* src/yspec.f90:432-473                                               */
/* ================================================================== */

/* fourrow_dp for a single row, src/module_fourier.f90:7-50: bit
 * reversal, then Danielson-Lanczos with the trigonometric recurrence and
 * theta = pi_d/(isign*mmax) (single-rounded pi). */
int ysxo_fourrow(double *dd, int n, int isign) {
  cplx *data = (cplx *)dd; /* data[0..n-1] <-> data(1..n) */
  if (n & (n - 1)) return 1;
  int n2 = n / 2, j = n2;
  for (int i = 1; i <= n - 2; i++) {
    if (j > i) {
      cplx t = data[j];
      data[j] = data[i];
      data[i] = t;
    }
    int m = n2;
    while (!(m < 2 || j < m)) {
      j = j - m;
      m = m / 2;
    }
    j = j + m;
  }
  int mmax = 1;
  while (n > mmax) {
    int istep = 2 * mmax;
    double theta = PI_D / (isign * mmax);
    double sh = sin(0.5 * theta);
    cplx wp = -2.0 * sh * sh + I * sin(theta);
    cplx w = 1.0;
    for (int m = 1; m <= mmax; m++) {
      cplx ws = w;
      for (int i = m; i <= n; i += istep) {
        int jj = i + mmax;
        cplx temp = ws * data[jj - 1];
        data[jj - 1] = data[i - 1] - temp;
        data[i - 1] = data[i - 1] + temp;
      }
      w = w * wp + w;
    }
    mmax = istep;
  }
  return 0;
}

int ysxo_seismograms(ysxo *o, const double *urd, const double *utd,
                     const double *upd, double *seis) {
  if (!o->have_run) return fail(o, "no run");
  int nt = o->nt, nr = o->nr, nw = o->nw, i1 = o->i1, i2 = o->i2;
  const cplx *spec[3] = {(const cplx *)urd, (const cplx *)utd,
                         (const cplx *)upd};
  double w1 = o->f11, w2 = o->f12, w3 = o->f21, w4 = o->f22;
  int nofilt = (w1 == 0.0 && w2 == 0.0 && w3 == 0.0 && w4 == 0.0);
  double *filt = malloc(sizeof(double) * (size_t)(nt / 2 + 2));
  for (int i = 1; i <= nt / 2 + 1; i++) {
    double w = (i - 1) * o->df, f = 1.0;
    if (!nofilt) {
      if (w < w1) f = 0.0;
      else if (w >= w1 && w < w2) {
        f = PI_D * (w - w1) / (w2 - w1);
        f = 0.5 * (1.0 - cos(f));
      } else if (w >= w2 && w < w3) f = 1.0;
      else if (w >= w3 && w < w4) {
        f = PI_D * (w4 - w) / (w4 - w3);
        f = 0.5 * (1.0 - cos(f));
      } else if (w >= w4) f = 0.0;
    }
    filt[i] = f;
  }
  cplx *row = malloc(sizeof(cplx) * (size_t)nt);
  double norm = (o->out == 0) ? R_NORM : (o->out == 1 ? o->vel_norm : o->acl_norm);
  cplx *tmp = malloc(sizeof(cplx) * (size_t)nt * 3);
  for (int k = 0; k < nr; k++) {
    for (int c = 0; c < 3; c++) {
      /* pack bins i1..i2, zero elsewhere, filter, Hermitian fill */
      for (int i = 1; i <= nt; i++) row[i - 1] = 0.0;
      for (int i = i1; i <= i2; i++)
        row[i - 1] = spec[c][(size_t)(i - i1) + (size_t)nw * k];
      for (int i = 1; i <= nt / 2 + 1; i++) row[i - 1] = filt[i] * row[i - 1];
      int jj = 0;
      for (int i = nt / 2 + 2; i <= nt; i++) {
        jj++;
        row[i - 1] = conj(row[nt / 2 - jj]);
      }
      ysxo_fourrow((double *)row, nt, 1);
      for (int i = 0; i < nt; i++) row[i] = row[i] / (o->dt * nt);
      for (int i = 0; i < nt; i++) {
        double t = i * o->dt;
        tmp[(size_t)c * nt + i] = row[i] * exp(o->ep * t) * norm;
      }
    }
    double ca = cos(o->azep[k]), sa = sin(o->azep[k]);
    for (int i = 0; i < nt; i++) {
      cplx z = tmp[i], t1 = tmp[(size_t)nt + i], t2 = tmp[2 * (size_t)nt + i];
      cplx n_ = -t1 * ca + t2 * sa;
      cplx e_ = t1 * sa + t2 * ca;
      seis[(size_t)i + (size_t)nt * (0 + 3 * (size_t)k)] = creal(z);
      seis[(size_t)i + (size_t)nt * (1 + 3 * (size_t)k)] = creal(n_);
      seis[(size_t)i + (size_t)nt * (2 + 3 * (size_t)k)] = creal(e_);
    }
  }
  free(row);
  free(tmp);
  free(filt);
  return 0;
}
