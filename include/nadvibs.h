/*
 * nadvibs.h -- C ABI of the MI355X Lanczos engine for the diabatic vibronic
 * Hamiltonian in a harmonic-oscillator direct-product basis.
 *
 * Every entry point returns an int32 status (NV_OK == 0) and never aborts or
 * throws across the ABI; the text of the last error is available from
 * nv_last_error().  Counts and indices are int64 (the reference is built with
 * default 8-byte integers, makefile:14), reals are double.  Host vectors use
 * the reference's Global-Arrays layout q(dimen, nstates) in Fortran order,
 * i.e. C order [nstates][n_1]...[n_M] with the last mode fastest
 * (Source/top.f90:379-386, :400-402), so a Fortran caller passes its arrays
 * unchanged.  Iteration numbers are 1-based as in the reference.
 *
 * The context owns all device memory (operator tables, the Krylov basis,
 * scratch); the caller owns every host array it passes.  Calls on one context
 * are not thread-safe; use one context per host thread.
 */
#ifndef NADVIBS_H
#define NADVIBS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define NV_OK 0
#define NV_ERR_ARG 1     /* invalid argument or inconsistent sizes */
#define NV_ERR_HIP 2     /* a HIP runtime call failed */
#define NV_ERR_STATE 3   /* call out of sequence (e.g. iteration before init) */
#define NV_ERR_IO 4      /* file could not be read or written */
#define NV_ERR_NOMEM 5   /* device or host allocation failed */
#define NV_WARN_NONORTHOGONAL 6  /* nv_krylov_gradient only: the gradient is written,
                                    but the Krylov basis has lost orthogonality
                                    (an overlap |q_j . q_k| above 1e-4, named by
                                    nv_last_error), so the projected gradient is
                                    not the derivative of the computed loss */

/* ---- operator conventions (reference quirks) --------------------------- */
/* <m|q^3|m+1> taken as sqrt((m+1)^3/8), one third of the exact value
 * (Source/function.f90:225). */
#define NV_QUIRK_Q3ELEM 1
/* The vibrationally diagonal part of an inter-state block (constant term and
 * the Delta=0 part of all-even monomials) is applied only for row state >
 * column state (Source/top.f90:545-553, :752-754). */
#define NV_QUIRK_LOWERDIAG 2
/* Reference behaviour: both of the above. */
#define NV_QUIRKS_REF 3
/* Exact, symmetric operator. */
#define NV_QUIRKS_EXACT 0

/* ---- H*v precision ----------------------------------------------------- */
#define NV_PREC_F64 0    /* off-diagonal couplings in FP64 */
#define NV_PREC_F32 1    /* off-diagonal couplings in FP32, diagonal FP64 */
#define NV_PREC_BF16 2   /* off-diagonal couplings applied to a BF16 copy of the
                            vector (8-bit mantissa) in FP32 arithmetic, diagonal
                            FP64: 2^-9 relative input rounding, for bandwidth
                            studies, not for the 1e-3 cm^-1 accuracy target
                            (experimental).  nv_lanczos_init refuses it with
                            reorthog > 0: on a 1.38M-row, 2000-iteration run
                            the rounding outgrew the re-orthogonalisation and
                            the coefficients went non-finite at iteration 1570 */

/* ---- problem description ---------------------------------------------- */
/* The data of basis.in / nadvibs.in after reading, before the reference's
 * term numbering (Source/top.f90:88-246 performs that numbering; the
 * library repeats it from these raw arrays). */
typedef struct nv_problem {
    int64_t nstates;          /* electronic states */
    int64_t nmodes;           /* vibrational modes */
    int64_t ordr;             /* order of the expansion (1..4 supported by hoelem) */
    const int64_t *nfunc;     /* [nmodes] basis functions per mode ("basis") */
    const double *aomega;     /* [nmodes] harmonic frequencies, hartree */
    const double *concoef;    /* [nstblks] constant term of each state block,
                                 block order (1,1),(2,1),(2,2),(3,1),... */
    const double *potterms;   /* POTterms(noterms(ordr), ordr, nstblks), Fortran
                                 order: for block b, order p the noterms(p)
                                 coefficients in nadvibs.in order, the rest of
                                 the noterms(ordr) slots ignored */
    double ztoler;            /* |c| < ztoler dropped (basis.in ztoler) */
    int32_t quirks;           /* NV_QUIRK_* bits */
    int32_t neworigin;        /* basis.in shiftref != 0: Franck-Condon start */
    const double *statew;     /* [nstates] state weights, NULL = all 1 */
    const double *bomega;     /* [nmodes]  (neworigin only) */
    const double *dvec;       /* [nmodes]  (neworigin only) */
    const double *tmat;       /* [nmodes*nmodes] T, column-major (neworigin only) */
    int64_t nirreps;          /* irreducible representations (>=1) */
    const int64_t *npirrep;   /* [nirreps] modes per irrep, NULL = one irrep */
    const int64_t *istate;    /* [nmodes] initial-state quanta, NULL = zeros */
} nv_problem;

typedef struct nv_ctx nv_ctx;

/* ---- lifetime ---------------------------------------------------------- */
/* Builds the operator (replaces read_constants' term numbering
 * Source/top.f90:150-234, hoelem table :501-527, diagonal block list
 * :528-553 and compute_allIndex Source/secondary.f90:3-226) on device
 * `device`.  *out is NULL on failure. */
int32_t nv_create(const nv_problem *prob, int32_t device, nv_ctx **out);
int32_t nv_destroy(nv_ctx *ctx);
/* Text of the last error on ctx (or of the last failed nv_create when ctx is
 * NULL).  Valid until the next call on the same context. */
const char *nv_last_error(const nv_ctx *ctx);

/* dimen = prod nfunc, n = nstates*dimen. */
int32_t nv_dims(const nv_ctx *ctx, int64_t *dimen, int64_t *n);
/* Number of stored terms per number of distinct modes (nztrms,
 * Source/top.f90:150-201): counts[ordr]. */
int32_t nv_term_summary(const nv_ctx *ctx, int64_t *nztrms);
/* The reference's term-count self check (Source/top.f90:438-499): number of
 * non-zero matrix elements ntcom(blk, i) per unique state block and number of
 * distinct modes; out[(i-1)*nstblks + (blk-1)]. */
int32_t nv_term_counts(const nv_ctx *ctx, int64_t *out);

/* ---- matrix-free H*v (replaces Hv, Source/top.f90:672-818) -------------- */
/* y = H x for host arrays x, y of n doubles. */
int32_t nv_hv(nv_ctx *ctx, int32_t prec, const double *x, double *y);
/* y = H x for device arrays of n doubles, enqueued on `stream` (a
 * hipStream_t, NULL = the context's stream).  No host synchronisation. */
int32_t nv_hv_device(nv_ctx *ctx, int32_t prec, const double *x_dev,
                     double *y_dev, void *stream);

/* ---- start vector (replaces generate_initialvec,
 *      Source/secondary.f90:351-809) ------------------------------------ */
/* Builds the normalised start vector on device; copies it to q (n doubles)
 * when q is not NULL. */
int32_t nv_start_vector(nv_ctx *ctx, double *q);

/* ---- Lanczos (replaces the loop of Source/main.f90:63-71) ---------------- */
/* Re-orthogonalisation thresholds of the mixed modes (pro_eps = 0): a vector
 * is re-orthogonalised when an overlap estimate reaches NV_PRO_MIXED_TAU, over
 * the range out to estimates below NV_PRO_MIXED_ETA (FP64 H*v: sqrt(eps) and
 * eps^(3/4), Source/top.f90:922-950). */
#define NV_PRO_MIXED_TAU 9.5367431640625e-07      /* 2^-20 */
#define NV_PRO_MIXED_ETA 1.4551915228366852e-11   /* 2^-36 */

typedef struct nv_lanczos_opts {
    int64_t niter;            /* total iterations (basis.in niter) */
    int32_t reorthog;         /* basis.in reorthog: 0 none, 1 partial with the
                                 omega recurrence, >1 partial with exact overlaps */
    int32_t prec;             /* NV_PREC_* for H*v */
    int64_t chkorthog;        /* exact overlap check period (default 100); with
                                 NV_PREC_F32/BF16 and pro_eps = 0 the period is
                                 min(chkorthog, 10) and the thresholds are
                                 NV_PRO_MIXED_TAU / NV_PRO_MIXED_ETA: the omega
                                 model assumes FP64 rounding of H*v, the checks
                                 keep it honest, and the thresholds sit above
                                 the overlap level that FP32 rounding alone
                                 produces (~1e-8) */
    uint64_t seed;            /* seed of the omega-estimate noise */
    double noise_scale;       /* multiplies every noise draw (1 = reference, 0 = off) */
    double pro_eps;           /* unit roundoff assumed by the omega recurrence and
                                 its thresholds sqrt(eps), eps^(3/4); 0 = 2^-53,
                                 the reference's determine_epsilon */
    int32_t savevecs;         /* != 0: keep every Lanczos vector in device memory
                                 (the reference's savevecs, set there by
                                 reorthog > 0, idroots > 0 or soroots > 0,
                                 Source/secondary.f90:863); needed by
                                 nv_identify_roots.  reorthog > 0 implies it. */
    int64_t checkpoint;       /* K > 0 with neither savevecs nor reorthog: keep
                                 (q_{cK}, r_{cK}) every K iterations so that
                                 nv_lanczos_adjoint can recompute the basis
                                 segment by segment; 0 = none.  Refused with
                                 savevecs or reorthog > 0: those keep every
                                 vector (re-orthogonalisation reads them) */
    double pro_tau;           /* > 0: re-orthogonalisation trigger on |omega|
                                 instead of the default (sqrt(eps), or
                                 NV_PRO_MIXED_TAU in the mixed modes) */
    double pro_eta;           /* > 0: range end on |omega| instead of the
                                 default (eps^(3/4), or NV_PRO_MIXED_ETA);
                                 nv_lanczos_init refuses eta >= tau after the
                                 defaults are applied */
} nv_lanczos_opts;

/* Allocates the Krylov basis for opts->niter iterations, builds the start
 * vector and beta(0) (Source/top.f90:590-613). */
int32_t nv_lanczos_init(nv_ctx *ctx, const nv_lanczos_opts *opts);
/* Per-iteration entry points, in the reference's order (Source/main.f90:64-70):
 * update_lanczos (:646-670), Hv (:672-818), lanczos_step (:820-842),
 * set_omega (:844-908), partial_orthog (:911-1005).  All are asynchronous. */
int32_t nv_update_lanczos(nv_ctx *ctx, int64_t iter);
int32_t nv_hv_iter(nv_ctx *ctx, int64_t iter);
int32_t nv_lanczos_step(nv_ctx *ctx, int64_t iter);
int32_t nv_set_omega(nv_ctx *ctx, int64_t iter);
int32_t nv_partial_orthog(nv_ctx *ctx, int64_t iter);
/* Runs iterations iiter..niter of the loop above with the per-iteration work
 * captured in a HIP graph.  Asynchronous. */
int32_t nv_lanczos_run(nv_ctx *ctx, int64_t iiter, int64_t niter);
/* Copies alpha(1:niter), beta(0:niter) and, when omega is not NULL,
 * omega(2, 0:niter) (Fortran order) to the host.  Synchronises. */
int32_t nv_get_coefficients(nv_ctx *ctx, double *alpha, double *beta,
                            double *omega);
/* Number of partial re-orthogonalisations performed and Krylov vectors
 * touched by them so far. */
int32_t nv_pro_stats(const nv_ctx *ctx, int64_t *northog_events,
                     int64_t *vectors_touched);
/* Copies Lanczos vector q_j (1 <= j <= last iteration + 1; j = last+1 gives
 * the unnormalised residual) to the host. */
int32_t nv_get_lanczos_vector(nv_ctx *ctx, int64_t j, double *q);

/* ---- tridiagonal eigensolve (replaces compute_ritzsystem + GIVENS,
 *      Source/secondary.f90:240-257, Source/yarkony.f90:91-543) ---------- */
/* Eigenvalues (ascending) of the n x n tridiagonal T with diagonal
 * alpha[0..n-1] and off-diagonal beta[1..n-1] (beta[0] = beta(0) unused),
 * intensities inten[i] = v(1,i)^2, convergence bji[i] = |beta[n] * v(n,i)|
 * and sign[i] = sign of v(1,i).  Runs on the context's device. */
int32_t nv_ritz(nv_ctx *ctx, int64_t n, const double *alpha,
                const double *beta, double *eval, double *inten, double *bji,
                int32_t *sign);

/* ---- Ritz vectors and root analysis (identify_roots,
 *      Source/top.f90:1122-1283) ------------------------------------------- */
/* For the lowest nroots roots of T(alpha, beta) (n <= last iteration; the
 * Krylov basis must be kept): Ritz vectors Y_i = sum_k v(k,i) q_k, formed on
 * the device from the archived basis.  Outputs (any may be NULL except as
 * noted):
 *   eval[nroots]       eigenvalues (hartree, absolute);
 *   inten[nroots]      v(1,i)^2 divided by its maximum over the nroots roots;
 *   sign[nroots]       sign of v(1,i);
 *   top_*[nroots*idroots] (required when idroots > 0): the idroots largest
 *                      |Y_i|^2 elements in decreasing order (ties by
 *                      increasing Fortran index), as top_index = 1-based
 *                      vibrational index (get_barray decodes it), top_state =
 *                      1-based state, top_phase = sign of the element,
 *                      top_mag2 = its square;
 *   percent[nroots*nstates]  100 |Y_i[:,k].Y_i[:,k]| / sum_k (state weights);
 *   overlap[((i*nroots + j)*nstates + k)*nstates + l] = Y_i[:,k] . Y_j[:,l]
 *                      (the spin-orbit overlaps of sodata.in for i > j, k != l);
 *   vectors[nroots*N]  the Ritz vectors themselves.
 * nroots * nstates <= 4096.  Synchronises. */
int32_t nv_identify_roots(nv_ctx *ctx, int64_t n, const double *alpha,
                          const double *beta, int64_t nroots, int64_t idroots,
                          double *eval, double *inten, int32_t *sign,
                          int64_t *top_index, int32_t *top_state,
                          int32_t *top_phase, double *top_mag2,
                          double *percent, double *overlap, double *vectors);

/* ---- derivatives (no counterpart in the reference: gradients of a
 *      spectrum loss for fitting the coupling constants) ------------------ */
/* Reverse sweep of the recurrence of the last nv_lanczos_run /
 * per-iteration calls (reorthog = 0, default start vector): the exact
 * derivative of the computed coefficients.  The backward recurrence amplifies
 * rounding along converged Ritz directions (about 2x per iteration), so it is
 * accurate for short runs only (n up to ~30 in FP64); spectral losses use
 * nv_krylov_gradient.  Given
 * abar = dL/dalpha(1:n) and bbar = dL/dbeta(0:n), accumulates
 *   g_omega[nmodes]      dL/daomega (hartree),
 *   g_concoef[nstblks]   dL/dconcoef,
 *   g_potterms[...]      dL/dPOTterms, same layout as nv_problem.potterms
 *                        (every slot, also coefficients below ztoler),
 * in FP64 with H*v in precision prec.  Needs the Krylov basis (savevecs) or
 * checkpoints (opts.checkpoint); with checkpoints each segment of K
 * iterations is recomputed on the device and checked bit for bit against the
 * stored alpha/beta.  Any output may be NULL.  Synchronises. */
int32_t nv_lanczos_adjoint(nv_ctx *ctx, int64_t n, const double *abar,
                           const double *bbar, int32_t prec, double *g_omega,
                           double *g_concoef, double *g_potterms);
/* env[g] = sum_i v(1,i)^2 exp(-((grid[g] - (theta_i - ref) AU2WAVE)/sigma)^2/2)
 * for the eigenpairs (theta_i, v(:,i)) of T(alpha, beta): the broadened
 * stick spectrum (convolve.py:28-48 without the final division by the
 * maximum), grid and sigma in cm^-1, ref = theta_1 (origin = 0, as the
 * eigenvalue table of Source/top.f90:1087-1102) or e0 hartree (origin = 1). */
int32_t nv_spectrum(nv_ctx *ctx, int64_t n, const double *alpha,
                    const double *beta, int64_t ng, const double *grid,
                    double sigma, int32_t origin, double e0, double *env);
/* Given gbar = dL/denv: tbar = dL/dT as a full symmetric n x n matrix
 * (S (D o a a^T) S^T, D the divided differences of the eigenvalues: finite
 * for near-degenerate Ritz values), and its tridiagonal part abar =
 * dL/dalpha(1:n), bbar = dL/dbeta(0:n) (bbar[0] = bbar[n] = 0).  Outputs may
 * be NULL. */
int32_t nv_spectrum_adjoint(nv_ctx *ctx, int64_t n, const double *alpha,
                            const double *beta, int64_t ng,
                            const double *grid, double sigma, int32_t origin,
                            double e0, const double *gbar, double *abar,
                            double *bbar, double *tbar);
/* The gradient of a spectral loss L(T) (tbar = dL/dT from
 * nv_spectrum_adjoint) with respect to the parameters, projected on the
 * kept Krylov basis: dL/dtheta = sum_jk tbar[j][k] q_j . (dH/dtheta) q_k.
 * It is the derivative of v0^T f(H) v0 in the Krylov space; it agrees with
 * central differences of the Lanczos-computed loss once the spectrum in the
 * window has converged, and it is numerically stable for any n (unlike the
 * reverse sweep of nv_lanczos_adjoint).  Needs savevecs, reorthog > 0 or
 * checkpoints (opts.checkpoint = K): then the basis is recomputed segment by
 * segment from the checkpoints, each pair of segments (A, B >= A) once with A
 * resident (nseg (nseg + 1) / 2 segment recomputations of K iterations, each
 * checked bit for bit against the stored alpha/beta; 2 (K + 2) N doubles of
 * device scratch).  Output layouts as nv_lanczos_adjoint.  The basis is
 * checked on the way: the first, second, middle and last vectors (kept basis)
 * or the first and last of every segment (checkpoints) against every later
 * vector; an overlap above 1e-4 returns NV_WARN_NONORTHOGONAL with the
 * gradient written.  Synchronises. */
int32_t nv_krylov_gradient(nv_ctx *ctx, int64_t n, const double *tbar,
                           double *g_omega, double *g_concoef,
                           double *g_potterms);
/* generic != 0: use the generic H*v kernel instead of compiling an
 * operator-specialised one (saves the run-time compile for operators used
 * once, e.g. inside a fit).  The generic kernel walks the terms by step shape
 * where it can (ordr <= 3); generic == 2 keeps the step-table
 * form for every operator.  Call before nv_lanczos_init. */
int32_t nv_select_hv_kernel(nv_ctx *ctx, int32_t generic);

/* ---- batched runs (no counterpart in the reference: many Hamiltonians,
 *      e.g. parameter sets or basis truncations, advance together) --------- */
/* One context per problem (operator tables, start vector, eigensolve via
 * nv_batch_context + nv_ritz); the iteration itself runs as one launch
 * sequence over every problem (tiles of rows, one workgroup each), captured
 * once as a HIP graph.  reorthog = 0, the table-driven H*v kernels. */
typedef struct nv_batch nv_batch;
int32_t nv_batch_create(const nv_problem *probs, int64_t count, int32_t device,
                        nv_batch **out);
int32_t nv_batch_destroy(nv_batch *b);
const char *nv_batch_last_error(const nv_batch *b);
/* Allocates the vectors of niter iterations in precision prec (NV_PREC_F64
 * or NV_PREC_F32) and loads every start vector. */
int32_t nv_batch_init(nv_batch *b, int64_t niter, int32_t prec);
/* Times the batched H*v over its tuning space -- 64, 128 and 256 rows per
 * workgroup, one or two rows per thread (FP32, from 128 rows), the shaped or
 * the step-table kernels -- and keeps the fastest; call between
 * nv_batch_init and the run.  report (may be NULL) receives the timings. */
int32_t nv_batch_tune(nv_batch *b, char *report, int64_t len);
/* Iterations iiter..niter for every problem.  Asynchronous. */
int32_t nv_batch_run(nv_batch *b, int64_t iiter, int64_t niter);
/* alpha(1:n) and beta(0:n) of problem p (n = last iteration).  Synchronises. */
int32_t nv_batch_coefficients(nv_batch *b, int64_t p, double *alpha,
                              double *beta);
/* The context of problem p (owned by the batch). */
nv_ctx *nv_batch_context(nv_batch *b, int64_t p);

/* ---- spectrum broadening (convolve.py:28-48, the Gaussian aperture) ----- */
/* env[g] = sum_j y[j] exp(-((grid[g] - x[j]) / sigma)^2 / 2), divided by its
 * maximum over the grid (sigma is a standard deviation, not a FWHM). */
int32_t nv_convolve(nv_ctx *ctx, int64_t n, const double *x, const double *y,
                    int64_t ng, const double *grid, double sigma, double *env);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): nv_convolve on
 * device arrays x_dev[n], y_dev[n], grid_dev[ng] into env_dev[ng], enqueued
 * on `stream` (a hipStream_t, NULL = the context's stream), with its
 * per-workgroup maxima in a static device array of the library instead of an
 * allocation.  ng <= NV_CONVOLVE_MAX_GRID (else NV_ERR_ARG).  Every call on
 * one GPU shares that array: calls that overlap on one GPU race.  No host
 * synchronisation. */
#define NV_CONVOLVE_MAX_GRID 1048576
int32_t nv_convolve_noalloc(nv_ctx *ctx, int64_t n, const double *x_dev,
                            const double *y_dev, int64_t ng,
                            const double *grid_dev, double sigma,
                            double *env_dev, void *stream);

/* ---- restart files (make_restart / load_restartinfo,
 *      Source/top.f90:1007-1049, Source/secondary.f90:872-892) -------------- */
/* Writes q1, q2, q3 of the last iteration as nstates records each of dimen
 * doubles (the layout of the reference's nadvibs.35.<rank> direct-access
 * file for one process). */
int32_t nv_save_vectors(nv_ctx *ctx, const char *path);
/* Restores the state saved by nv_save_vectors together with alpha(1:iter),
 * beta(0:iter) (and omega when non-NULL), so that nv_update_lanczos(iter+1)
 * continues the run. */
int32_t nv_load_vectors(nv_ctx *ctx, const char *path, int64_t iter,
                        const double *alpha, const double *beta,
                        const double *omega);
/* A run that keeps every Lanczos vector (reorthog > 0 or savevecs) restarts
 * from its archive, as the reference reopens nadvibs.34.<rank>
 * (Source/top.f90:569-572).  nv_save_archive writes q_1..q_n (n = last
 * iteration) as n records of nstates*dimen doubles (that file's layout for
 * one process) and, when pro_path is not NULL, the state of partial
 * re-orthogonalisation (omega rows, pending ranges, noise position) in the
 * library's own format, so that the restarted run continues exactly.
 * nv_load_archive restores them for iteration iter after nv_lanczos_init and
 * before nv_load_vectors; without pro_path the re-orthogonalisation starts
 * afresh from the omega passed to nv_load_vectors, as in the reference. */
int32_t nv_save_archive(nv_ctx *ctx, const char *archive_path,
                        const char *pro_path);
int32_t nv_load_archive(nv_ctx *ctx, const char *archive_path,
                        const char *pro_path, int64_t iter);

/* ---- measurement hooks (no counterpart in the reference: they serve the
 *      benchmark's live roofline) ------------------------------------------ */
/* on != 0: every H*v launch of the Lanczos loop is bracketed by HIP events on
 * the context's stream. */
int32_t nv_profile_enable(nv_ctx *ctx, int32_t on);
/* Milliseconds spent in the timed H*v launches and their number since the
 * previous call; synchronises. */
int32_t nv_profile_read(nv_ctx *ctx, double *hv_ms, int64_t *hv_launches);
/* Floating-point operations of one H*v: twice the multiply-adds of its
 * nonzero structure (every row, step and coupled state pair). */
int32_t nv_hv_flops(const nv_ctx *ctx, int64_t *flops);
/* Average duration in microseconds of `launches` back-to-back launches of
 * the iteration H*v kernel alone (after nv_update_lanczos(1)), timed with one
 * pair of HIP events on the context's stream.  Overwrites alpha(1). */
int32_t nv_profile_hv(nv_ctx *ctx, int64_t launches, double *avg_us);
/* Which H*v kernel serves precision prec: the operator-specialised one
 * (compiled at run time, on first use) or the generic one, as text.  The
 * environment variable NV_HV_KERNEL=generic selects the generic kernel. */
int32_t nv_hv_kernel_info(nv_ctx *ctx, int32_t prec, char *buf, int64_t len);

#ifdef __cplusplus
}
#endif

#endif /* NADVIBS_H */
