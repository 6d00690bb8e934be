// ysx_host.hpp -- host-side set-up of the spectral path: deck -> normalised
// model and spline tables, run parameters, receiver geometry and Legendre
// tables.  All O(knots) or O(lmax * receivers) work done once per model or
// run; the per-(l, w) work is on the GPU.
#pragma once
#include <complex>
#include <string>
#include <vector>

#include "ysx_common.hpp"

namespace ysx {

struct Norms {
  // set_parameters, src/module_model.f90:474-492
  double r_norm, rho_norm, vel_norm, acl_norm, fre_norm, t_norm, moment_norm,
      grav_norm, pibigg;
};

struct HostModel {
  int nknot = 0, nic = 0, noc = 0, nsl = 0, ifanis = 0, ocean = 0, fic = 0,
      nlayer = 0;
  double tref = 0.0;
  // field-major table, F_COUNT fields x nknot knots (0-based knot index)
  std::vector<double> tab;
  std::vector<int> layer_lo, layer_hi;  // 1-based knot indices
  double at(int field, int knot1) const { return tab[(size_t)field * nknot + knot1 - 1]; }
};

struct HostRun {
  int lmin = 0, lmax = 0, ats = 0, grav = 0, out = 0, cor = 0, nr = 0;
  int nw = 0, nt = 0, i1 = 0, i2 = 0, is = 0, ir = 0, nout = 0;
  double rs = 0, rr = 0, dt = 0, tout = 0, ep = 0, df = 0, wmin = 0, dw = 0;
  double f11 = 0, f12 = 0, f21 = 0, f22 = 0, mm[6] = {0};
  std::vector<double> delta, azep, azst;
  std::vector<std::complex<double>> eia;
};

Norms make_norms();
// returns "" on success or the reference's stop message
std::string build_model(const Norms& N, int nknot, int nic, int noc, int ians,
                        double tref_period, const double* deck, HostModel& M);
// deck_bar = J^T table_bar for table = build_model(deck) (J: d table /
// d deck); knot radii are not differentiated
std::string model_table_vjp(const Norms& N, int nknot, int nic, int noc, int ians,
                            double tref_period, const double* deck, const double* tbar,
                            double* dbar);
std::string build_run(const Norms& N, const HostModel& M, const int* ip,
                      const double* fp, const double* lat, const double* lon,
                      HostRun& R);

void legendre_m2(double theta, int l, int m, double* x, double* xp, double* xc);
void delaz(double eplat, double eplon, double stlat, double stlon,
           double* delta, double* azep, double* azst);

}  // namespace ysx
