/*
 * ysx.h -- C-ABI of the MI355X (gfx950) yspec spectral path, libysx.so.
 *
 * The reference has no plugin or FFI layer: the path sits behind Fortran
 * subroutines and module globals called from its driver program.  These
 * entry points replace those call sites one for one; a Fortran host binds
 * them with iso_c_binding (interface in yspec_amd/fortran/ysx_bind.f90).
 *
 * Conventions
 *  - every function returns 0 on success and a nonzero code otherwise; the
 *    message is in ysx_last_error(), and carries the reference's own stop
 *    text where the reference has one;
 *  - all arrays are caller-owned host memory, column-major as the reference
 *    stores them, complex numbers interleaved (re, im) as Fortran
 *    complex(dpc);
 *  - device memory, the HIP stream and every intermediate are owned by the
 *    context; calls are synchronous; a context is not thread-safe;
 *  - units in: those of the parameter file and deck (SI, mHz, minutes, s,
 *    km, degrees, N m); units out: SI (src/yspec.f90:447-460).
 */
#ifndef YSX_H
#define YSX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ysx_ctx ysx_ctx;

#define YSX_OK 0
#define YSX_ERR_ARG 1      /* bad argument or call order */
#define YSX_ERR_INPUT 2    /* input rejected where the reference stops */
#define YSX_ERR_HIP 3      /* HIP runtime error */
#define YSX_ERR_UNSUPPORTED 4

/* Create a context on HIP device `device`: stream, device buffers.
 * Replaces the reference's module-global state (src/module_model.f90:118-160,
 * src/module_int.f90:93-117, src/module_solver.f90:11). */
int ysx_create(ysx_ctx **ctx, int device);
void ysx_destroy(ysx_ctx *ctx);
const char *ysx_last_error(const ysx_ctx *ctx);

/* Load a deck model: nknot rows of 9 raw columns (r, rho, vpv, vsv, qkappa,
 * qmu, vph, vsh, eta) as read from the deck file; ians = 0 uses the first 6
 * (src/module_model.f90:214-221).  tref_period is the header's reference
 * period in s.  Normalises, builds the per-layer splines and gravity and
 * uploads the tables.  Replaces read_model's numerics and set_parameters
 * (src/module_model.f90:175-383, 474-492). */
int ysx_set_model(ysx_ctx *ctx, int nknot, int nic, int noc, int ians,
                  double tref_period, const double *deck);

/* The model table the kernels read, built on the device by ysx_set_model:
 * F_COUNT = 29 fields x nknot knots, field-major -- r, then rho, A, C, F,
 * L, N, kappa, mu, 1/Qkappa, 1/Qmu, xlam, xa2 each followed by its spline
 * second derivatives, then g and its second derivatives, alpha, beta, in
 * the reference's normalised units (the arrays of src/module_model.f90:
 * 118-160 after read_model).  out holds 29 * nknot doubles. */
int ysx_model_table(const ysx_ctx *ctx, double *out);

/* The same table computed on the host without a context or a GPU (the
 * set-up path of the run parameters uses it). */
int ysx_model_table_host(int nknot, int nic, int noc, int ians,
                         double tref_period, const double *deck, double *out);

/* Allocation-free form of ysx_set_model followed by ysx_model_table.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  The
 * deck is host memory as for ysx_set_model; d_table is caller-owned device
 * memory on the context's device for F_COUNT * nknot doubles, which receives
 * the table ysx_model_table would return (the same kernel, the same values).
 * The transform's work arrays live in a fixed 2 MiB device array per GPU,
 * held until the process ends; a deck that needs more (large knot counts,
 * more than 64 layers) is rejected with YSX_ERR_ARG and a message.  Makes no
 * device allocation; does not touch the context's model.  Synchronous on the
 * context's stream.  Calls that overlap on one GPU (several threads or
 * contexts) race on that array. */
int ysx_model_table_noalloc(ysx_ctx *ctx, int nknot, int nic, int noc, int ians,
                            double tref_period, const double *deck, double *d_table);

/* Adjoint of the deck -> model-table transform (no counterpart in the
 * reference; the first stage of a misfit gradient): deck_bar (nknot x 9,
 * the deck's layout) = J^T table_bar for table = ysx_model_table_host(deck)
 * and J its Jacobian with respect to the deck columns rho, vpv, vsv, qkappa,
 * qmu, vph, vsh, eta.  Knot radii are not differentiated; at a zero shear
 * modulus (fluid) beta = sqrt(mu / rho) contributes nothing.  Host function;
 * ysx_model_vjp_deck is the same map on the device. */
int ysx_model_table_vjp(int nknot, int nic, int noc, int ians, double tref_period,
                        const double *deck, const double *table_bar, double *deck_bar);

/* ysx_model_table_vjp on the device for the context's model (the deck of
 * the last ysx_set_model; src/module_model.f90:253-378 transposed): one
 * block of k_model_vjp, the same functions as the host's, so deck_bar
 * equals ysx_model_table_vjp's bit for bit.  table_bar: host, F_COUNT x
 * nknot (the layout of ysx_model_table); deck_bar: host, nknot x 9.  One
 * block, latency-bound (0.76 ms for 200 knots on MI355X, against about
 * 0.1 ms for the host function): for callers whose table adjoint is on the
 * device. */
int ysx_model_vjp_deck(ysx_ctx *ctx, const double *table_bar, double *deck_bar);

/* Set the run: ip = lmin lmax ats grav out cor nr; fp = f1 f2 tout dt f11
 * f12 f21 f22 src_depth src_lat src_lon Mrr Mrt Mrp Mtt Mtp Mpp rec_depth;
 * lat/lon of the nr receivers.  Receiver geometry, Legendre tables, fcal,
 * source/receiver layers; rejects what the reference stops on.
 * Replaces src/yspec.f90:141-215 and src/module_fourier.f90:362-397 (fcal),
 * src/module_solver.f90:1855-1861 (set_ats_switch). */
int ysx_set_run(ysx_ctx *ctx, const int *ip, const double *fp,
                const double *lat, const double *lon);

/* iout = nw nt i1 i2 is ir nout npairs ; fout = rs rr dt tout ep df wmin dw
 * (normalised units) */
int ysx_run_info(const ysx_ctx *ctx, int *iout8, double *fout8);

/* Spectra: every (l, w) pair of the run on the GPU, then the Legendre
 * receiver sum.  ur/ut/up are (nw, nr) complex, or NULL to keep the result
 * on the device only.  Replaces the l x w loop and its solver calls,
 * src/yspec.f90:250-426 (tor_solver, sph_solver, rad_solver_wg,
 * source_vector_*: src/module_solver.f90:25-215, 446-627, 640-823,
 * 1747-1847). */
int ysx_spectra(ysx_ctx *ctx, double *ur, double *ut, double *up);

/* Seismograms from the device-resident spectra: cosine filters, Hermitian
 * fill, inverse FFT, exp(eps t), units, rotation to Z/N/E.  seis is
 * (nt, 3, nr) real (component order Z, N, E), or NULL.  Replaces
 * ifft (src/module_fourier.f90:55-145) and src/yspec.f90:432-473. */
int ysx_seismograms(ysx_ctx *ctx, double *seis);

/* ysx_spectra followed by ysx_seismograms without host copies of the
 * spectra. */
int ysx_forward(ysx_ctx *ctx, double *seis);

/* ---- several events on one model ------------------------------------
 * The reference computes one event per run of its program
 * (src/yspec.f90:1-510 for one parameter file); these entry points take
 * nev events at once: event e has the run parameters ip[7e..7e+6] and
 * fp[18e..18e+17] (as in ysx_set_run) and the receivers lat/lon[r_e ..
 * r_e + nr_e - 1], r_e = sum of nr over the events before e.  Each event
 * is checked as ysx_set_run checks one; an error names the event.
 * ysx_set_run(ctx, ...) is ysx_set_events(ctx, 1, ...). */
int ysx_set_events(ysx_ctx *ctx, int nev, const int *ip, const double *fp,
                   const double *lat, const double *lon);

/* ysx_run_info of event e */
int ysx_event_info(const ysx_ctx *ctx, int e, int *iout8, double *fout8);

/* The full forward (spectra and seismograms) of every event.  mode 0:
 * events one after another, each with its own kernel launches (the plain
 * path).  mode 1: all events in one launch per kernel, their (w, l-block)
 * work items sorted by estimated cost, replayed from a captured HIP graph
 * (the batched path).  seis: the events' (nt_e, 3, nr_e) seismograms one
 * after another, or NULL.  ysx_timings()[7] holds the elapsed ms. */
int ysx_forward_events(ysx_ctx *ctx, int mode, double *seis);

/* Batched-path tuning: persist = 1 runs persistent blocks
 * (blocks_per_cu per compute unit) that take `chunk` items per atomic
 * queue access; persist = 0 launches one block per item. */
int ysx_set_batch_tuning(ysx_ctx *ctx, int persist, int chunk, int blocks_per_cu);

/* Lanes per (l, w) pair where the kernels offer a choice: rad_lanes lanes
 * integrate each l = 0 radial solve (rad_solver_wg / rad_solver_ng,
 * src/module_solver.f90:335-561) together, sharing its stage coefficients:
 * 8, 1, or 0 for the default (8, or 1 for runs with more than 640
 * frequencies).  Applies to the runs set and those set later (their work
 * lists are rebuilt); results agree to rounding.  The tuner of the batched
 * path sets it.  No counterpart in the reference. */
int ysx_set_lanes(ysx_ctx *ctx, int rad_lanes);

/* Diagnostic switch, no counterpart in the reference: full = 1 makes runs
 * set afterwards integrate the spheroidal b-vector system as the
 * reference's four source-side columns (sph_solver,
 * src/module_solver.f90:745-755, 809-819), so that ysx_pair_records
 * returns all six rows of ysph.  full = 0 (the default) marches the two or
 * three receiver-side rows the seismograms need back to the source with the
 * transposed steps of the same step sequence: the same results up to
 * rounding, about half the b-vector work. */
int ysx_set_records(ysx_ctx *ctx, int full);

/* Spectra ((nw_e, nr_e) complex each, or NULL) and seismograms
 * ((nt_e, 3, nr_e)) of event e from the last forward. */
int ysx_event_spectra(ysx_ctx *ctx, int e, double *ur, double *ut, double *up);
int ysx_event_seismograms(ysx_ctx *ctx, int e, double *seis);

/* Per-pair records for n pairs (l[i], j[i]) (j 1-based), 89 complex each:
 * svt(5,2) ytor(2,2) svs(5,6) ysph(6,4) sr0(2) yrad(2,2) uu(5) vv(5) ww(5),
 * the values the reference's loop body holds after its source-vector and
 * solver calls (src/yspec.f90:270-400); the pairs must have been computed
 * by the last ysx_spectra (one-event contexts only).  ysph holds all six
 * rows only for runs set after ysx_set_records(ctx, 1); otherwise rows 0..1
 * (0..2 with the potential/tilt correction), the rows the seismograms use,
 * and zeros. */
int ysx_pair_records(ysx_ctx *ctx, int n, const int *l, const int *j,
                     double *rec);

/* Cash-Karp step counts of n pairs (l[i], j[i]) (j 1-based) of the last
 * ysx_spectra (one-event contexts only), 5 per pair in the order of
 * ysx_step_counts.  The reference
 * keeps no such count; its fixed step rule (yint, src/module_int.f90:
 * 1903-1905) fixes them, so they expose each pair's start level and
 * integration path. */
int ysx_pair_steps(ysx_ctx *ctx, int n, const int *l, const int *j, int *out5);

/* ---- adjoints of the linear back end (no counterpart in the reference;
 * stages of a misfit gradient) -------------------------------------------
 * With several events (ysx_set_events, then ysx_forward_events) seis_bar
 * holds the events' (nt_e, 3, nr_e) arrays one after another, the spectra
 * and record adjoints stay on the device (ur_bar, ut_bar, up_bar, the
 * spectra inputs and rec_bar must be NULL, n = 0), mm_bar gets 6 values per
 * event and table_bar the sum over the events.
 * For a real function F of the seismograms, a complex quantity z gets the
 * adjoint dF/dRe z + i dF/dIm z.  Each entry point applies the exact
 * transpose of the discrete map the forward applies (its twiddles, filter
 * weights and factors), so <x_bar, J x> = <J^T x_bar, x> holds to rounding.
 *
 * Adjoint of ysx_seismograms: seis_bar ((nt, 3, nr) real, the layout of the
 * seismograms) -> ur_bar, ut_bar, up_bar ((nw, nr) complex each, or NULL;
 * the result also stays on the device for ysx_spectra_vjp).  Needs the run
 * only, not a forward. */
int ysx_seismograms_vjp(ysx_ctx *ctx, const double *seis_bar, double *ur_bar,
                        double *ut_bar, double *up_bar);

/* Adjoint of the receiver sum and the coefficient assembly of the last
 * ysx_spectra: from the spectra adjoints (the three (nw, nr) arrays, or all
 * NULL for the device-resident result of ysx_seismograms_vjp) to the
 * adjoints of every pair's record, computed for all pairs on the device;
 * rec_bar gets those of the n pairs (l[i], j[i]) (j 1-based) in the layout
 * of ysx_pair_records: the adjoints of svt, ytor, svs, ysph (rows 0..2),
 * sr0, yrad, and of uu, vv, ww in their slots. */
int ysx_spectra_vjp(ysx_ctx *ctx, const double *ur_bar, const double *ut_bar,
                    const double *up_bar, int n, const int *l, const int *j, double *rec_bar);

/* Gradient with respect to the moment tensor (Mrr, Mrt, Mrp, Mtt, Mtp, Mpp,
 * per N m as in the parameter file; 6 per event) from the record adjoints
 * of the last ysx_spectra_vjp: the source vectors (source_vector_tor / _sph / _rad,
 * src/module_solver.f90:145-215, 566-627, 1747-1847) are linear in the
 * moment tensor, so mm_bar[i] = sum over pairs of <sv_bar, dsv/dM_i>.  After
 * ysx_seismograms_vjp(seis_bar) and ysx_spectra_vjp it is the exact gradient
 * of sum(seis_bar * seismograms). */
int ysx_moment_tensor_vjp(ysx_ctx *ctx, double *mm_bar);

/* The source-vector term of the gradient with respect to the model: the
 * source vectors read L, C, F, 1/Qmu, xlam and xa2 at the source radius
 * (spline values on the source's knot interval), so the record adjoints of
 * the last ysx_spectra_vjp give adjoints of those table entries.
 * table_bar: F_COUNT x nknot in the layout of ysx_model_table (zero except
 * the source interval's entries), the input of ysx_model_table_vjp.  The
 * terms of the per-pair solves are not included. */
int ysx_source_model_vjp(ysx_ctx *ctx, double *table_bar);

/* Gradient with respect to the model table through the per-pair solves
 * (no counterpart in the reference; the core of a misfit gradient): with
 * the record adjoints of the last ysx_spectra_vjp (after
 * ysx_seismograms_vjp(seis_bar): the gradient of sum(seis_bar *
 * seismograms)), table_bar (F_COUNT x nknot, the layout of
 * ysx_model_table) gets the adjoint of every table entry, through
 *   - the toroidal, spheroidal minor and b-vector marches (tor_solver,
 *     sph_solver, src/module_solver.f90:25-142, 640-823; yint and rkck,
 *     src/module_int.f90:1847-1968, 2016-2064) as discrete adjoints of
 *     the same fixed step lists, replayed from knot-interval checkpoints;
 *   - the secular determinant, the b-vector starts and the region
 *     transforms (src/module_int.f90:1563-1592, 1620-1680, 1705-1725);
 *   - the starting solutions (src/module_int.f90:213-250, 1192-1270,
 *     1346-1405, 1451-1534);
 *   - the radial (l = 0) solves (rad_solver_wg, src/module_solver.f90:
 *     446-561), every step of both sweeps replayed from a tape;
 *   - the source vectors (as ysx_source_model_vjp).
 * Start levels, step counts
 * and renormalisations are those of the forward (held fixed).  With
 * several events (ysx_set_events) their gradients are summed on the
 * device: each sweep is one persistent launch over every event's pairs.
 * Built for self-gravitation, a solid inner core, no ocean, receivers at
 * the surface, no potential/tilt correction and the default b-vector rows;
 * other runs return YSX_ERR_UNSUPPORTED.  ysx_model_table_vjp takes
 * table_bar to the deck. */
int ysx_model_vjp(ysx_ctx *ctx, double *table_bar);

/* L2 waveform misfit and its gradient with respect to the deck's knot
 * values (no counterpart in the reference): one forward of the run (the
 * batched launch of ysx_forward_events when several events are set), then
 *   misfit = 0.5 sum (seismogram - obs)^2
 * over the first n samples of every trace of every event, n = nout, or for
 * nout <= 0 each event's own output length (the samples the output files
 * hold), and knot_grad (nknot x 4, row-major) = d misfit / d (rho, vp, vs,
 * Qmu) of each knot in the deck's SI units, summed over the events, where
 * vp (vs) moves vpv and vph (vsv and vsh) together.  obs: per event in
 * order, n x 3 x nr column-major (time fastest), the first n samples of
 * each trace in ysx_seismograms' layout.  The residuals are formed on the
 * device; the chain is ysx_seismograms_vjp, ysx_spectra_vjp, ysx_model_vjp
 * and ysx_model_table_vjp, with ysx_model_vjp's scope; stage times from
 * ysx_timings, ysx_vjp_timings and ysx_grad_timings afterwards.  With one
 * event the forward also stores the minor and toroidal marches' state
 * before every step (when it fits in 32 GiB and half the free device
 * memory; kept with the run, ysx_grad_memory), and their sweeps start from
 * those states instead of marching again.  obs up to 64 MiB is staged into
 * pinned host memory while the forward runs. */
int ysx_misfit_grad(ysx_ctx *ctx, const double *obs, int nout, double *misfit,
                    double *knot_grad);

/* Device bytes the model gradient holds for the run (after a
 * ysx_model_vjp; no counterpart in the reference): [0] the sweeps'
 * per-step states (the stored side's state before every step of the
 * longest item, per persistent block) and those that a single-event
 * ysx_misfit_grad's forward stored, [1] the per-block step scratch (three
 * times with those states: the minor sweep's three item groups run side by
 * side)
 * and the radial tape — the per-block parts do not grow with the number of
 * pairs —
 * and [2] the per-pair adjoints (record adjoints, minors at the source and
 * start) and the per-block partial table adjoints. */
int ysx_grad_memory(const ysx_ctx *ctx, double *bytes3);

/* Kernel milliseconds of the last ysx_model_vjp: [0] b-vector sweep, [1]
 * minor sweep, [2] toroidal sweep, [3] starting solutions, [4] radial
 * solves, [5] source vectors, [6] the sum.  When the forward stored the
 * toroidal states (a single-event ysx_misfit_grad), the toroidal sweep and
 * then the radial solves run on a second stream beside the minor sweep:
 * [2] and [4] are their elapsed times there, overlapped with [1], and [6]
 * exceeds the wall time. */
int ysx_grad_timings(const ysx_ctx *ctx, double *ms7);

/* Kernel milliseconds (HIP events on the context's stream, host copies
 * excluded) of the last call of [0] ysx_seismograms_vjp, [1] ysx_spectra_vjp,
 * [2] ysx_moment_tensor_vjp or ysx_source_model_vjp, whichever ran last. */
int ysx_vjp_timings(const ysx_ctx *ctx, double *ms3);

/* Timings of the last ysx_spectra / ysx_seismograms, milliseconds, from
 * HIP events on the context's streams: [0] source vectors + start levels,
 * [1] toroidal, [2] spheroidal minors, [3] b-vector, [4] radial (second
 * stream), [5] coefficient assembly, [6] Legendre sum, [7] IFFT + post.
 * An event too small to fill the GPU runs its toroidal start levels and
 * march on the second stream ahead of the radial solves; [1] then covers
 * both and [0] holds the spheroidal start levels only. */
int ysx_timings(const ysx_ctx *ctx, double *ms8);

/* Exact Cash-Karp step counts of the last ysx_spectra, summed over pairs:
 * [0] toroidal, [1] solid-region minors, [2] fluid-region minors,
 * [3] b-vector (one count per step of the reference's march; each step
 * advances all 4 source columns there, and the receiver-side rows here take
 * the same steps), [4] radial.
 * The step rule is the reference's fixed-step rule (yint,
 * src/module_int.f90:1903-1905), so these are the reference's counts too;
 * with per-step FLOP constants they give the algorithmic work. */
int ysx_step_counts(ysx_ctx *ctx, long long *out5);

#ifdef __cplusplus
}
#endif
#endif
