/*
 * ysx_oracle -- CPU restatement (plain C11, FP64) of the reference yspec
 * spectral path.  TEST INFRASTRUCTURE ONLY: tests/, the package smoke check
 * and bench.py's cpu_baseline leg may load it, as the checker.  The product
 * (libysx.so) never links, loads or calls it.
 *
 * Every function in ysx_oracle.c names the reference routine (file:line) it
 * restates.  Scope: deck models (ians 0/1, ocean, solid inner core), full
 * self-gravitation (grav=2), any source/receiver depth, attenuation on/off,
 * output types 0/1/2 with the potential/tilt correction, the Legendre
 * receiver sum, the filtered inverse FFT and the post-processing.
 */
#ifndef YSX_ORACLE_H
#define YSX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ysxo ysxo;

/* number of doubles per knot written by ysxo_model_arrays (see .c) */
#define YSXO_MODEL_ARRAYS 32

ysxo *ysxo_create(void);
void ysxo_destroy(ysxo *o);
const char *ysxo_error(const ysxo *o);

/* deck: nknot rows of 9 raw columns (r, rho, vpv, vsv, qkappa, qmu, vph, vsh,
 * eta) in SI units as read from the file; for ians=0 only the first 6 are
 * used.  tref_period is the deck header's reference period in seconds. */
int ysxo_set_model(ysxo *o, int nknot, int nic, int noc, int ians,
                   double tref_period, const double *deck);

/* normalised model arrays, YSXO_MODEL_ARRAYS x nknot, order as ref_dump */
int ysxo_model_arrays(const ysxo *o, double *out);
/* nknot, nic, noc, nsl, ifanis, ocean, fic, nlayer */
int ysxo_model_flags(const ysxo *o, int *out8);
/* tref, r_norm, rho_norm, vel_norm, acl_norm, fre_norm, t_norm,
 * moment_norm, pibigg */
int ysxo_norms(const ysxo *o, double *out9);

/* ip: lmin lmax ats grav out cor nr
 * fp: f1 f2 tout dt f11 f12 f21 f22 (mHz/min/s as in the parameter file),
 *     src_depth_km src_lat src_lon mm[6] (N m) rec_depth_km            */
int ysxo_set_run(ysxo *o, const int *ip, const double *fp,
                 const double *lat, const double *lon);
/* iout: nw nt i1 i2 is ir nout ; fout: rs rr dt tout ep df wmin dw
 *       f11 f12 f21 f22 (normalised) */
int ysxo_run_info(const ysxo *o, int *iout7, double *fout12);

/* delta, azep, azst (nr) and xa/xpa/xca (lmax+1, 3, nr) column-major */
int ysxo_geometry(const ysxo *o, double *delta, double *azep, double *azst,
                  double *xa, double *xpa, double *xca);

/* One (l, w_j) pair, j 1-based.  rec receives 89 complex numbers
 * (interleaved re/im): svt(5,2) ytor(2,2) svs(5,6) ysph(6,4) sr0(2)
 * yrad(2,2) uu(5) vv(5) ww(5); starts[2] = tor/sph start level. */
int ysxo_pair(ysxo *o, int l, int j, double *rec, int *starts);

/* Spectra ur/ut/up (nw, nr) complex interleaved, for l in [l0, l1],
 * accumulated in the reference's order; nthreads>1 uses OpenMP over w. */
int ysxo_spectra(ysxo *o, int l0, int l1, double *ur, double *ut, double *up,
                 int nthreads);

/* Filtered inverse FFT + exp(eps t) + units + rotation of the spectra.
 * seis: (nt, 3, nr) real parts, rows t = 0..nt-1 */
int ysxo_seismograms(ysxo *o, const double *ur, const double *ut,
                     const double *up, double *seis);

/* Reference-structured radix-2 FFT (for unit tests): data (n) complex
 * interleaved, isign = +1 or -1, in place. */
int ysxo_fourrow(double *data, int n, int isign);

/* Scalar helpers for unit tests */
void ysxo_legendre(double theta, int l, int m, double *x, double *xp,
                   double *xc);
void ysxo_delaz(double eplat, double eplon, double stlat, double stlon,
                double *out3);
int ysxo_fcal(double f1, double f2, double dt, double tout, int *iout3,
              double *fout4);
void ysxo_bfs(int l, double xsq_re, double xsq_im, double eps, double *fp2);

#ifdef __cplusplus
}
#endif
#endif
