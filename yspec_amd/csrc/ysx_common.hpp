// ysx_common.hpp -- layout constants shared by the host set-up and the
// device kernels.
#pragma once

namespace ysx {

// Per-knot fields of the model table, stored field-major
// (tab[field * nknot + knot - 1]).  Each spline field is followed by its
// second-derivative table (natural cubic spline per layer).
enum Field {
  F_R = 0, F_RHO, F_RHO2, F_ACON, F_ACON2, F_CCON, F_CCON2, F_FCON, F_FCON2,
  F_LCON, F_LCON2, F_NCON, F_NCON2, F_KAPPA, F_KAPPA2, F_MU, F_MU2, F_QKAPPA,
  F_QKAPPA2, F_QMU, F_QMU2, F_XLAM, F_XLAM2, F_XA2, F_XA22, F_GRAV, F_GRAV2,
  F_ALPHA, F_BETA, F_COUNT
};

// src/nrtype.f90:20-25: the pi family is single-precision rounded,
// fourpi_d truncated; bigg as the reference sets it.
constexpr double PI_D = (double)3.141592653589793238462643383279502884197f;
constexpr double TWOPI_D = (double)6.283185307179586476925286766559005768394f;
constexpr double FOURPI_D = 12.56637061;
constexpr double BIGG = 6.6723e-11;

}  // namespace ysx
