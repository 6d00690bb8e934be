// ysx_host.cpp -- host set-up: deck -> normalised model tables, run
// parameters, receiver geometry and Legendre tables.
#include "ysx_host.hpp"
#include "ysx_model.hpp"

#include <cmath>
#include <cstring>

namespace ysx {

namespace {

constexpr double R_NORM = R_NORM_M;
constexpr double RHO_NORM = RHO_NORM_M;

}  // namespace

Norms make_norms() {
  Norms N;
  N.r_norm = R_NORM;
  N.rho_norm = RHO_NORM;
  N.vel_norm = R_NORM * std::sqrt(PI_D * BIGG * RHO_NORM);
  N.acl_norm = PI_D * BIGG * RHO_NORM * R_NORM;
  N.fre_norm = N.vel_norm / R_NORM;
  N.t_norm = 1 / N.fre_norm;
  N.grav_norm = PI_D * BIGG;
  N.moment_norm = R_NORM * R_NORM * R_NORM * R_NORM * R_NORM * RHO_NORM * RHO_NORM * N.grav_norm;
  N.pibigg = PI_D * BIGG / N.grav_norm;
  return N;
}

// read_model numerics, src/module_model.f90:196-378
std::string build_model(const Norms& N, int nknot, int nic, int noc, int ians,
                        double tref_period, const double* deck, HostModel& M) {
  if (nknot < 3 || nic < 1 || noc <= nic || noc >= nknot) return "bad deck header";
  M = HostModel();
  M.nknot = nknot;
  M.nic = nic;
  M.noc = noc;
  M.ifanis = (ians == 1);
  M.tref = TWOPI_D / tref_period / N.fre_norm;
  M.tab.assign((size_t)F_COUNT * nknot, 0.0);
  auto col = [&](int f) { return M.tab.data() + (size_t)f * nknot; };
  const double *r = col(F_R), *rho = col(F_RHO), *L0 = deck + 3;
  M.ocean = (L0[9 * (size_t)(nknot - 1)] == 0.0);
  M.nsl = nknot;
  if (M.ocean)
    for (int i = nknot - 1; i >= 1; i--)
      if (L0[9 * (size_t)(i - 1)] > 0.0) { M.nsl = i; break; }
  M.fic = (L0[0] == 0.0);
  for (int i = 0; i < nknot; i++) knot_fields(deck, nknot, i, M.ifanis, N.vel_norm, M.tab.data());
  // layers split at duplicated radii (src/module_model.f90:294-308)
  M.layer_lo = {1};
  for (int i = 2; i <= nknot; i++)
    if (r[i - 1] == r[i - 2]) {
      M.layer_hi.push_back(i - 1);
      M.layer_lo.push_back(i);
    }
  M.layer_hi.push_back(nknot);
  M.nlayer = (int)M.layer_lo.size();
  std::vector<double> work(3 * ((size_t)nknot + 1));
  for (int k = 0; k < M.nlayer; k++) {
    int lo = M.layer_lo[k] - 1, n = M.layer_hi[k] - M.layer_lo[k] + 1;
    for (int q = 0; q < 12; q++) {
      int f = spline_field(q);
      if (!natural_spline(r + lo, col(f) + lo, col(f + 1) + lo, n, work.data()))
        return "tridag_ser: error at code stage 2";
    }
  }
  // gravity: 5-point Gauss-Legendre on the spline of r^2 rho
  // (gauslv, src/module_model.f90:356-371, 507-535)
  std::vector<double> intg(nknot), qint(nknot);
  for (int i = 0; i < nknot; i++) intg[i] = r[i] * r[i] * rho[i];
  for (int k = 0; k < M.nlayer; k++) {
    int lo = M.layer_lo[k] - 1, n = M.layer_hi[k] - M.layer_lo[k] + 1;
    natural_spline(r + lo, intg.data() + lo, qint.data() + lo, n, work.data());
  }
  double* grav = col(F_GRAV);
  grav[0] = 0.0;
  double fint = 0.0;
  for (int i = 0; i < nknot - 1; i++) {
    fint += grav_interval(r, intg.data(), qint.data(), i);
    grav[i + 1] = fint;
  }
  for (int i = 0; i < nknot; i++)
    if (r[i] > 0.0) grav[i] = (4.0 * grav[i]) / (r[i] * r[i]);
  for (int k = 0; k < M.nlayer; k++) {
    int lo = M.layer_lo[k] - 1, n = M.layer_hi[k] - M.layer_lo[k] + 1;
    natural_spline(r + lo, grav + lo, col(F_GRAV2) + lo, n, work.data());
  }
  return "";
}

// legendre, src/module_function.f90:6-132 (the arithmetic is shared with
// the device in ysx_model.hpp; cos and sin of theta are taken here)
void legendre_m2(double theta, int l, int m, double* x, double* xp, double* xc) {
  legendre_cs(theta, std::cos(theta), std::sin(theta), l, m, x, xp, xc);
}

// delaz, src/module_function.f90:317-390 (geocentric latitude, the
// reference's rounded degree-to-radian factor)
void delaz(double eplati, double eplong, double stlati, double stlong, double* delta,
           double* azep, double* azst) {
  const double geoco = 0.993277, hpi = 0.5 * PI_D, rad = 0.0174533;
  double eplat = (eplati == -90.0) ? -89.98 : eplati;
  double stlat = (stlati == -90.0) ? -89.98 : stlati;
  if (eplat == stlat && eplong == stlong) {
    *delta = *azep = *azst = 0.0;
    return;
  }
  double el = hpi - std::atan(geoco * std::tan(eplat * rad));
  double stl = hpi - std::atan(geoco * std::tan(stlat * rad));
  double elon = eplong * rad, slon = stlong * rad;
  double as = std::cos(stl), bs = std::sin(stl), cs = std::cos(slon), ds = std::sin(slon);
  double a = std::cos(el), b = std::sin(el), c = std::cos(elon), d = std::sin(elon);
  double cdel = a * as + b * bs * (c * cs + d * ds);
  if (std::fabs(cdel) > 1.0) cdel = std::copysign(1.0, cdel);
  double delt = std::acos(cdel), sdel = std::sin(delt);
  double caze = (as - a * cdel) / (sdel * b);
  if (std::fabs(caze) > 1.0) caze = std::copysign(1.0, caze);
  double aze = std::acos(caze);
  double cazs = 0.0;
  if (bs > 0.0) cazs = (a - as * cdel) / (bs * sdel);
  if (bs == 0.0) cazs = std::copysign(1.0, cazs);
  if (std::fabs(cazs) > 1.0) cazs = std::copysign(1.0, cazs);
  double azs = std::acos(cazs);
  double dif = ds * c - cs * d;
  if (dif < 0.0) aze = TWOPI_D - aze;
  if (dif > 0.0) azs = TWOPI_D - azs;
  *delta = delt;
  *azep = aze;
  *azst = azs;
}

// run set-up: src/yspec.f90:141-215 and fcal, src/module_fourier.f90:362-397
std::string build_run(const Norms& N, const HostModel& M, const int* ip, const double* fp,
                      const double* lat, const double* lon, HostRun& R) {
  R = HostRun();
  R.lmin = ip[0];
  R.lmax = ip[1];
  R.ats = ip[2];
  R.grav = ip[3];
  R.out = ip[4];
  R.cor = ip[5];
  R.nr = ip[6];
  if (R.lmin < 0 || R.lmax < R.lmin || R.nr < 1) return "bad lmin/lmax/receiver count";
  if (R.out < 0 || R.out > 2) return "output switch must be 0, 1 or 2";
  if (R.grav < 0 || R.grav > 2) return "gravitation switch must be 0, 1 or 2";
  double f1 = fp[0], f2 = fp[1], tout = fp[2], dt = fp[3];
  double sdep = fp[8], lats = fp[9], lons = fp[10], rdep = fp[17];
  int nr = R.nr;
  R.delta.resize(nr);
  R.azep.resize(nr);
  R.azst.resize(nr);
  R.eia.resize(nr);
  if (lats == 90.0) lats = 89.98;
  for (int k = 0; k < nr; k++) {
    double dl, ae, as;
    delaz(lat[k], lon[k], lats, lons, &dl, &ae, &as);
    R.delta[k] = dl;
    R.azep[k] = TWOPI_D - ae;
    R.azst[k] = PI_D - as;
    R.eia[k] = std::exp(std::complex<double>(0.0, R.azst[k]));
  }
  // the (lmax+1) x 3 x nr Legendre tables are built on the device
  // (legendre_cs, k_legendre) straight into the receiver-sum operand
  // normalisation (src/yspec.f90:179-189)
  double rtop = M.at(F_R, M.nknot);
  f1 = f1 * N.t_norm / 1000.0;
  f2 = f2 * N.t_norm / 1000.0;
  R.dt = dt / N.t_norm;
  R.rs = rtop - sdep * 1000.0 / R_NORM;
  R.rr = rtop - rdep * 1000.0 / R_NORM;
  R.tout = tout * 60.0 / N.t_norm;
  for (int i = 0; i < 6; i++) R.mm[i] = fp[11 + i] / N.moment_norm;
  R.f11 = fp[4] / 1000.0 * N.t_norm;
  R.f12 = fp[5] / 1000.0 * N.t_norm;
  R.f21 = fp[6] / 1000.0 * N.t_norm;
  R.f22 = fp[7] / 1000.0 * N.t_norm;
  // fcal with mex = 5, qex = 1; single-precision log as in the reference
  if (0.5 / R.dt < f2) return "fmax above the Nyquist frequency";
  R.ep = 5.0 / R.tout;
  R.df = R.ep / TWOPI_D;
  int nt = (int)(1.0 / (R.df * R.dt));
  int ne = (int)(std::log((float)nt) / std::log(2.0f) + 1.0f);
  if (ne > 24) return "time series too long";
  R.nt = 1 << ne;
  R.df = 1.0 / (R.nt * R.dt);
  R.i1 = std::max((int)std::floor(f1 / R.df), 2);
  R.i2 = (int)std::floor(f2 / R.df) + 2;
  R.dw = TWOPI_D * R.df;
  if (R.i1 == R.i2) return "too few frequency steps";
  R.wmin = (R.i1 - 1) * R.dw;
  R.nw = R.i2 - R.i1 + 1;
  // find_radius_index(.., -1), src/module_model.f90:537-576
  auto fri = [&](double x) {
    if (x < 0.0 || x > 1.0) return 0;
    for (int i = 1; i <= M.nknot - 1; i++) {
      double a = M.at(F_R, i), b = M.at(F_R, i + 1);
      if (x == a && x == b) return i;
      if (x > a && x <= b) return i;
    }
    return 0;
  };
  R.is = fri(R.rs);
  if (R.is == 0) return " source not in model!";
  if (R.is > M.nsl) return " source in the ocean!";
  R.ir = fri(R.rr);
  if (R.ir == 0) return " receiver not in model!";
  if (R.ir > M.nsl) return " receiver in the ocean!";
  // nout: first sample with t > tout (src/yspec.f90:479-486)
  R.nout = R.nt;
  for (int i = 1; i <= R.nt; i++)
    if ((i - 1) * R.dt > R.tout) { R.nout = i; break; }
  return "";
}


// ----------------------------------------------------- model table adjoint
// The transposes of build_model's steps in reverse order, through the
// shared functions of ysx_model.hpp; k_model_vjp runs the same sequence on
// the device.
std::string model_table_vjp(const Norms& N, int nknot, int nic, int noc, int ians,
                            double tref_period, const double* deck, const double* tbar,
                            double* dbar) {
  HostModel M;
  std::string err = build_model(N, nknot, nic, noc, ians, tref_period, deck, M);
  if (!err.empty()) return err;
  const double* tab = M.tab.data();
  const double* r = tab + (size_t)F_R * nknot;
  const double* rho = tab + (size_t)F_RHO * nknot;
  // adjoints of the knot values of every table field
  std::vector<double> fb(tbar, tbar + (size_t)F_COUNT * nknot);
  auto bar = [&](int f) { return fb.data() + (size_t)f * nknot; };
  std::vector<double> work(3 * ((size_t)nknot + 2));
  auto layer = [&](int k, int& lo, int& n) {
    lo = M.layer_lo[k] - 1;
    n = M.layer_hi[k] - M.layer_lo[k] + 1;
  };
  int lo, n;
  // gravity: grav2 = S(grav) per layer; grav[i] = 4 G[i] / r^2 (r > 0) with
  // G the running sum of the interval integrals gi of the spline of r^2 rho
  for (int k = 0; k < M.nlayer; k++) {
    layer(k, lo, n);
    natural_spline_T(r + lo, bar(F_GRAV2) + lo, bar(F_GRAV) + lo, n, work.data());
  }
  std::vector<double> gbar(nknot, 0.0);  // adjoint of the running sum G
  for (int i = 0; i < nknot; i++)
    gbar[i] = (r[i] > 0.0) ? bar(F_GRAV)[i] * 4.0 / (r[i] * r[i]) : bar(F_GRAV)[i];
  std::vector<double> gib(nknot, 0.0);  // gi[i] enters G[i+1..]
  double acc = 0.0;
  for (int i = nknot - 1; i >= 1; i--) {
    acc += gbar[i];
    gib[i - 1] = acc;
  }
  std::vector<double> intg(nknot), qint(nknot), intgb(nknot, 0.0), qintb(nknot, 0.0);
  for (int i = 0; i < nknot; i++) intg[i] = r[i] * r[i] * rho[i];
  for (int k = 0; k < M.nlayer; k++) {
    layer(k, lo, n);
    natural_spline(r + lo, intg.data() + lo, qint.data() + lo, n, work.data());
  }
  for (int i = 0; i < nknot - 1; i++) grav_interval_T(r, i, gib[i], intgb.data(), qintb.data());
  for (int k = 0; k < M.nlayer; k++) {
    layer(k, lo, n);
    natural_spline_T(r + lo, qintb.data() + lo, intgb.data() + lo, n, work.data());
  }
  for (int i = 0; i < nknot; i++) bar(F_RHO)[i] += r[i] * r[i] * intgb[i];
  // the 12 splined fields: y2 = S(y) per layer
  for (int k = 0; k < M.nlayer; k++) {
    layer(k, lo, n);
    for (int q = 0; q < 12; q++) {
      int f = spline_field(q);
      natural_spline_T(r + lo, bar(f + 1) + lo, bar(f) + lo, n, work.data());
    }
  }
  // knot_fields, per knot
  for (int i = 0; i < nknot; i++)
    knot_fields_T(deck, nknot, i, ians == 1, N.vel_norm, tab, fb.data(), dbar + 9 * (size_t)i);
  return "";
}

}  // namespace ysx
