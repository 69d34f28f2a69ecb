/*
 * ptbgpu.h - C ABI of the MI355X density-matrix engine (libptbgpu.so).
 *
 * Every integer is int64_t because the reference is compiled with -i8
 * (source/Makefile:37): a Fortran `integer` on the caller's side is 8 bytes.
 * Matrices are the reference's packed upper storage: element (i,j), i <= j,
 * 1-based, at lin(i,j) = i + j(j-1)/2 (source/lin.f:5-11,
 * source/solve.f:149-200), length ndim*(ndim+1)/2, FP64.
 *
 * Entry points without the _dev suffix take host arrays (the true drop-in:
 * the engine copies in and out).  The _dev variants take device pointers on
 * the context's device (inputs already resident in HBM); xyz, aoat and all
 * scalar outputs stay on the host.  Device arrays must hold the element counts
 * stated for each call (packed matrices: ndim*(ndim+1)/2 doubles).
 *
 * Ordering of _dev calls: the context's stream is a non-blocking stream and
 * does not wait on any other stream.  The caller must have completed the work
 * that produces the device inputs (synchronized its stream, or made the
 * context's stream wait on an event recorded after that work, see
 * ptbgpu_stream) before the call.  Every call returns with its own work on
 * the context's stream complete, so outputs are ready for any stream.
 *
 * Threading: one context per host thread; calls on one context are not
 * re-entrant.  Each context owns one HIP stream (default priority) and a
 * device workspace that grows on demand and is kept across calls; caller
 * pointers are never retained after a call returns.
 *
 * Every call returns PTBGPU_OK (0) or one of the status codes below;
 * ptbgpu_last_error() gives a message for the last failure on a context.
 */
#ifndef PTBGPU_H
#define PTBGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ptbgpu_status {
    PTBGPU_OK = 0,
    PTBGPU_E_ARG = 1,      /* invalid argument (size, null pointer, range) */
    PTBGPU_E_MODE = 2,     /* unsupported submatrix_columns / submatrix_mode:
                              the reference stops (source/sm.f90:159-161,249-251,312-314) */
    PTBGPU_E_NOT_SPD = 3,  /* an overlap submatrix is not positive definite
                              (dsygvd INFO>0; the reference ignores it, source/sm.f90:299) */
    PTBGPU_E_NAN = 4,      /* NaN in a result (source/purification.f90:179-181,491-492) */
    PTBGPU_E_KMEANS = 5,   /* k-means++ fault (source/sm.f90:478-481) */
    PTBGPU_E_HIP = 6,      /* HIP runtime error */
    PTBGPU_E_NOMEM = 7,    /* device allocation failed */
    PTBGPU_E_NOCONV = 8    /* eigensolver or iteration did not converge */
};

/* arithmetic of the per-submatrix solve */
enum ptbgpu_precision {
    PTBGPU_FP64 = 0,       /* plain FP64 drop-in */
    PTBGPU_MIXED = 1,      /* the tridiagonalisation on float storage, then one
                              FP64 refinement of the occupied eigenvectors whose
                              correction products run on the BF16 matrix cores
                              as split-BF16x3 with FP32 accumulation (the
                              divide-and-conquer merge products too, for the
                              submatrices up to the "lowp_merge_max_m" knob,
                              0 by default; Cholesky, K, the secular
                              equations, fe, mu and the density stay FP64) */
    PTBGPU_MIXED_BF16 = 2  /* the same with plain BF16 products (8-bit operands);
                              for accuracy studies of the precision ladder */
};

typedef struct ptbgpu_ctx ptbgpu_ctx;

typedef struct {
    int64_t precision;    /* enum ptbgpu_precision */
    int64_t strict;       /* 1: a non-SPD overlap block fails the call with
                             PTBGPU_E_NOT_SPD; 0: count it in the report and go on */
    int64_t verbose;      /* 1: print the reference's progress lines */
    int64_t refine_steps; /* MIXED only: FP64 refinement sweeps, 0 = default (1);
                             negative: none (the float-storage accuracy, for
                             accuracy studies) */
    /* submatrix_columns 4 (k-means): the uniform random numbers k-means++ draws,
     * restarted for every candidate as the reference's initseed does
     * (source/sm.f90:20-32,459-470): stream[0..nstream), nstream >= 64 covers
     * every candidate.  A Fortran host fills it from RANDOM_NUMBER after
     * RANDOM_SEED(PUT = 42*(i-1)) and one discarded draw, which reproduces the
     * reference's groups with the same compiler.  NULL: the engine's own
     * generator (splitmix64 seeded with 42), restarted per candidate. */
    const double *kmeans_stream;
    int64_t nstream;
} ptbgpu_opts;

/* Phase slots of ptbgpu_report.t_phase (labels of source/sm.f90:76-386). */
enum ptbgpu_phase {
    PTBGPU_T_ALLOC = 0,        /* 'alloc' (device upload / workspace) */
    PTBGPU_T_COMBINATION = 1,  /* 'submatrix combination' */
    PTBGPU_T_STATS = 2,        /* 'submatrix stats' (pattern + index sets) */
    PTBGPU_T_CONSTRUCTION = 3, /* 'submatrix construction' (gather) */
    PTBGPU_T_EIGEN = 4,        /* 'submatrix eigendecomp' */
    PTBGPU_T_FE = 5,           /* 'submatrix fe' */
    PTBGPU_T_MUITER = 6,       /* 'mu iter' (occ + Nel bisection) */
    PTBGPU_T_FINAL = 7,        /* 'submatrix final iter' (dmat + scatter) */
    PTBGPU_T_DEALLOC = 8,      /* 'submatrix dealloc' (download) */
    PTBGPU_T_TOTAL = 9,
    PTBGPU_T_NSLOT = 12
};

typedef struct {
    double mu;            /* final chemical potential */
    double nel;           /* sum(fe*occ) at mu */
    int64_t muiter;       /* bisection passes (the printed muiter count) */
    int64_t nsm;          /* number of submatrices */
    int64_t sum_m;        /* sum of submatrix dimensions */
    int64_t max_m;        /* largest submatrix dimension */
    double sum_m3;        /* sum of m^3 (work measure) */
    int64_t not_spd;      /* overlap blocks that failed Cholesky */
    double t_phase[PTBGPU_T_NSLOT]; /* seconds, enum ptbgpu_phase */
} ptbgpu_report;

/* ---- context ------------------------------------------------------------ */

/* Create a context on HIP device `device`. flags: reserved, pass 0. */
int ptbgpu_init(ptbgpu_ctx **ctx, int64_t device, int64_t flags);
int ptbgpu_finalize(ptbgpu_ctx *ctx);
const char *ptbgpu_strerror(int64_t status);
const char *ptbgpu_last_error(const ptbgpu_ctx *ctx);
/* Block until all work on the context's stream is done. */
int ptbgpu_sync(ptbgpu_ctx *ctx);
/* The context's HIP stream (hipStream_t), for callers that time or order work. */
int ptbgpu_stream(ptbgpu_ctx *ctx, void **stream);

/* Tuning knobs of a context (no reference counterpart; set by the autotuner,
 * ptb_amd/tune.py).  "dc_bucket_min": the divide-and-conquer merge nodes of
 * one tree level are launched in power-of-two size classes, each with the LDS
 * of its class; classes below this floor join the smallest launched one, and
 * 0 launches a level at once (default 256).  "mixed_min_m": a mixed-precision
 * build reduces the submatrices below this dimension in FP64 (no float
 * storage, no refinement sweep) and the others on float storage with the FP64
 * refinement; 0 puts every submatrix on float storage (default 288).
 * "gemm_tile": output tile edge of the FP64 products, 64 or 128 (identical
 * results either way); 0 (default) chooses per launch: 128 only for launches
 * of fewer than 8 large products.  "dc_global_min": merge size classes of at
 * least this many rows run their secular roots and vectors over the whole
 * chip from device scratch even when they fit in LDS (identical results;
 * slower on MI355X, so 0, the default, keeps that path for the classes that
 * do not fit).  "lowp_merge_max_m": a mixed build's float-reduced
 * submatrices of up to this many rows take their merge products in the
 * low-precision arithmetic (then two Loewdin steps in the refinement instead
 * of one); 0 (default) keeps FP64 merge products, as fast on MI355X.
 * "graph": 1 makes the construction and eigendecomposition phase of a build
 * (one chunk) replay a captured hipGraph once the same shapes and buffers
 * have been seen before (the first such build runs eagerly, the second
 * captures); 0 (default) turns it off and drops the held graphs.  ptbgpu_get_tuning also reads "graphs",
 * the number of graphs held, and "graph_failures", the captures that failed
 * (those builds ran eagerly). */
int ptbgpu_set_tuning(ptbgpu_ctx *ctx, const char *key, int64_t value);
int ptbgpu_get_tuning(ptbgpu_ctx *ctx, const char *key, int64_t *value);
/* What the eigen phase of the last build did with graphs on: "replayed",
 * "captured", "first sight of the shapes: eager", or why a capture failed. */
const char *ptbgpu_graph_note(const ptbgpu_ctx *ctx);

/* Kernel timing (no reference counterpart; used by the benchmark).  With
 * profiling on, the engine brackets each launch of its hot kernels with HIP
 * events on the context's stream; the query syncs and returns the summed
 * milliseconds and the number of bracketed launch groups of one kernel name
 * (filter, pattern, gather, potrf, trsm, sytrd, dc_leaf, dc_merge, gemm,
 * larft, fe_rows, fe_finish, bisect, dmat_scatter, ...).  Enabling or
 * disabling clears the records. */
int ptbgpu_profile(ptbgpu_ctx *ctx, int64_t enable);
int ptbgpu_profile_query(ptbgpu_ctx *ctx, const char *name, double *total_ms, int64_t *launches);

/* Batched products C_i = alpha op(A_i) op(B_i) + beta C_i on the engine's
 * matrix-core GEMM kernels (no reference counterpart: the product family the
 * eigensolver runs on, exposed for tests and tile tuning).  Device pointers,
 * column-major FP64 storage; op(A_i) is M_i x K_i, op(B_i) K_i x N_i;
 * transa/transb 1 = transposed.  precision: 0 FP64 MFMA; 1 split-BF16x3 and
 * 2 BF16 on the BF16 matrix cores with FP32 accumulation (each operand
 * element x enters as bf16(x) + bf16(x - bf16(x)), or bf16(x) alone).
 * Dimensions <= 32768; each operand below 2^32 elements. */
int ptbgpu_gemm_batch_dev(ptbgpu_ctx *ctx, int64_t precision, int64_t transa, int64_t transb,
                          int64_t count, const int64_t *M, const int64_t *N, const int64_t *K,
                          const double *const *A, const int64_t *lda, const double *const *B,
                          const int64_t *ldb, double *const *C, const int64_t *ldc,
                          double alpha, double beta);

/* ---- filter / analyse ----------------------------------------------------
 * Replaces filter_matrix_sym (source/pgtb.f90:1532-1548): A(k) = 0 where
 * abs(A(k)) < filter (strict), in place over the packed matrix.  counts, if
 * not NULL, receives the analyse_matrix_sym figure (source/pgtb.f90:1549-1579)
 * "nonzero elements" = number of strictly positive entries of the full
 * symmetric matrix (diagonal once, off-diagonal twice) before [0] and after [1]
 * filtering. */
int ptbgpu_filter_sym(ptbgpu_ctx *ctx, int64_t ndim, double *A, double filter,
                      int64_t counts[2]);
int ptbgpu_filter_sym_dev(ptbgpu_ctx *ctx, int64_t ndim, double *dA, double filter,
                          int64_t counts[2]);
/* analyse_matrix_sym alone (source/pgtb.f90:1549-1579): positive-entry count. */
int ptbgpu_analyse_sym(ptbgpu_ctx *ctx, int64_t ndim, const double *A, int64_t *npos);
int ptbgpu_analyse_sym_dev(ptbgpu_ctx *ctx, int64_t ndim, const double *dA, int64_t *npos);

/* ---- submatrix density build ---------------------------------------------
 * Replaces sm_stupid_simple(ndim,nelref,Hvec,Svec,Pvec,n,xyz,aoat,
 * submatrix_columns,submatrix_mode) (source/sm.f90:34-409).
 *   xyz(3,n) bohr, aoat(ndim) 1-based atom of each AO (host arrays);
 *   submatrix_columns: 1 single_columns, 2 atoms_columns, 3 full_matrix,
 *                      4 kmeans_r_heuristic (source/purification_settings.f90:9-12);
 *   submatrix_mode: 2 submatrix_sygv (1 = submatrix_sign -> PTBGPU_E_MODE, as
 *                   the reference stops there).
 * P is fully overwritten.  opts and rep may be NULL (defaults: FP64, compat). */
int ptbgpu_sm_density(ptbgpu_ctx *ctx, int64_t ndim, int64_t nelref, const double *H,
                      const double *S, double *P, int64_t n, const double *xyz,
                      const int64_t *aoat, int64_t submatrix_columns,
                      int64_t submatrix_mode, const ptbgpu_opts *opts,
                      ptbgpu_report *rep);
int ptbgpu_sm_density_dev(ptbgpu_ctx *ctx, int64_t ndim, int64_t nelref, const double *dH,
                          const double *dS, double *dP, int64_t n, const double *xyz,
                          const int64_t *aoat, int64_t submatrix_columns,
                          int64_t submatrix_mode, const ptbgpu_opts *opts,
                          ptbgpu_report *rep);

/* Many independent molecules in one build (no counterpart in the reference,
 * which builds one molecule per call): molecule i has ndim[i], nelref[i],
 * packed H[i], S[i] (filtered) and P[i], natoms[i], xyz[i] (3 x natoms,
 * bohr) and aoat[i] (ndim entries, 1..natoms); every molecule gets the same
 * column grouping and options and its own chemical potential, exactly as
 * sm_stupid_simple would give it alone.  The submatrices of all molecules share
 * one batched eigensolver pass; the O(N^2) phases run per molecule.  reps
 * (nmol entries, may be NULL) carry each molecule's mu, nel, passes and
 * submatrix statistics; their phase times are the batch's.  ptbgpu_sm_spectra
 * afterwards returns the submatrices of all molecules in order (nsm and sum_m
 * summed over the batch).  The _dev variant takes device pointers. */
int ptbgpu_sm_density_batch(ptbgpu_ctx *ctx, int64_t nmol, const int64_t *ndim,
                            const int64_t *nelref, const double *const *H, const double *const *S,
                            double *const *P, const int64_t *natoms, const double *const *xyz,
                            const int64_t *const *aoat, int64_t submatrix_columns,
                            int64_t submatrix_mode, const ptbgpu_opts *opts, ptbgpu_report *reps);
int ptbgpu_sm_density_batch_dev(ptbgpu_ctx *ctx, int64_t nmol, const int64_t *ndim,
                                const int64_t *nelref, const double *const *dH,
                                const double *const *dS, double *const *dP, const int64_t *natoms,
                                const double *const *xyz, const int64_t *const *aoat,
                                int64_t submatrix_columns, int64_t submatrix_mode,
                                const ptbgpu_opts *opts, ptbgpu_report *reps);

/* Per-submatrix dimensions, eigenvalues and fe weights of the last build on
 * this context (the quantities the reference prints per submatrix and sums in
 * the mu bisection, source/sm.f90:298-310,340).  Arrays sized from the report:
 * m[nsm], eig[sum_m], fe[sum_m]; eigenvalues per submatrix are in no particular
 * order but eig[i] and fe[i] belong together.  Any pointer may be NULL. */
int ptbgpu_sm_spectra(ptbgpu_ctx *ctx, int64_t nsm, int64_t sum_m, int64_t *m, double *eig,
                      double *fe);

/* ---- populations -------------------------------------------------------------
 * Replaces mlpop2(n, nao, p, s1, s2, q, qsh) (source/mlpop.f:198-238): atomic
 * populations q(n) and shell populations qsh(10, n) (column-major) from the
 * packed P and s1 = S^3/4, s2 = S^1/4 (REAL(4), nao x nao column-major, as
 * mlpop14 makes them, source/mlpop.f:142-190): diag(S^1/4 x S^3/4) with x = P
 * rounded to float, summed over the AOs of each atom (aoat, 1..n) and each
 * shell (shell2ao, 1..10) in AO order.  The products run in FP64 on the
 * float-rounded inputs; the result differs from the reference's REAL(4)
 * products by their rounding.  nao <= 46340.  The _dev variant takes P, s1,
 * s2 on the device; aoat, shell2ao, q and qsh stay on the host. */
int ptbgpu_mlpop2(ptbgpu_ctx *ctx, int64_t n, int64_t nao, const double *P, const float *s1,
                  const float *s2, const int64_t *aoat, const int64_t *shell2ao, double *q,
                  double *qsh);
int ptbgpu_mlpop2_dev(ptbgpu_ctx *ctx, int64_t n, int64_t nao, const double *dP,
                      const float *ds1, const float *ds2, const int64_t *aoat,
                      const int64_t *shell2ao, double *q, double *qsh);

/* ---- density checks (source/metrics.f90) ---------------------------------
 * get_nel_packed (:202-230) = tr(P S); get_band_structure_E_packed (:253-277)
 * = tr(P H); idempotent_packed (:316-360) deviation = |sum(0.5 P S P - P)|
 * over the full matrix, as the reference writes it (a sum, not a norm). */
int ptbgpu_get_nel(ptbgpu_ctx *ctx, int64_t ndim, const double *P, const double *S,
                   double *nel);
int ptbgpu_band_energy(ptbgpu_ctx *ctx, int64_t ndim, const double *P, const double *H,
                       double *energy);
int ptbgpu_idempotency(ptbgpu_ctx *ctx, int64_t ndim, const double *P, const double *S,
                       double *deviation);

/* The same traces on device-resident packed matrices. */
int ptbgpu_get_nel_dev(ptbgpu_ctx *ctx, int64_t ndim, const double *dP, const double *dS,
                       double *nel);
int ptbgpu_band_energy_dev(ptbgpu_ctx *ctx, int64_t ndim, const double *dP, const double *dH,
                           double *energy);

/* The idempotency deviation without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Device inputs dP, dS (packed, ndim*(ndim+1)/2 doubles); the deviation is
 * written to the device double *d_deviation.  The call only enqueues work on
 * the context's stream and returns without waiting (read the result after
 * ptbgpu_sync or an event on ptbgpu_stream); it calls no hipMalloc, hipFree
 * or memset.  Its scratch is one static 40 MiB device array per GPU, which
 * limits ndim to about 51000 (larger returns PTBGPU_E_ARG); every context on a
 * GPU shares that array, so calls that overlap on one GPU (two contexts or
 * streams at once) race on it. */
int ptbgpu_idempotency_noalloc(ptbgpu_ctx *ctx, int64_t ndim, const double *dP, const double *dS,
                               double *d_deviation);

/* ---- purification (source/purification.f90) -------------------------------
 * The fields of tPurificationSet (source/purification_settings.f90:29-86) that
 * purification_wrapper reads, with the same codes and defaults:
 *   type 1 mcweeny, 2 sign_iter_pade, 3 sign_iter_newton, 4 sign_diagonalization
 *   (default 4); metric 1 inv, 2 inv_sqrt (default), 3 sqrt; metric_iterative
 *   0/1 (default 0); cycles 40, metric_cycles 50, chempot_cycles 60,
 *   chempot_guess -1.5, chempot_increment 0.1; nel = electrons. */
typedef struct ptbgpu_pur_settings {
    int64_t type;
    int64_t metric;
    int64_t metric_iterative;
    int64_t cycles;
    int64_t metric_cycles;
    int64_t chempot_cycles;
    double chempot_guess;
    double chempot_increment;
    int64_t nel;
} ptbgpu_pur_settings;

/* Everything a .PUR file sets (tPurificationSet, source/purification_settings.f90:
 * 29-86), with the reference's defaults: mode 1 full / 2 submatrix (default
 * 1), submatrix_columns 1..4 (default 2 atoms_columns), submatrix_mode 1 sign /
 * 2 sygv (default 2), pr (default 1), the dev (default 1) and check (default 0)
 * flags, and the purification fields. */
typedef struct ptbgpu_pur_file {
    int64_t mode;
    int64_t submatrix_columns;
    int64_t submatrix_mode;
    int64_t prlvl;
    int64_t dev;
    int64_t check;
    ptbgpu_pur_settings pur;
} ptbgpu_pur_file;

/* The defaults of ptbgpu_pur_file (host only, no context). */
int ptbgpu_pur_defaults(ptbgpu_pur_file *f);
/* Replaces initialize_purification (source/purification_settings.f90:91-218):
 * parses the text of a .PUR file (len bytes, lines separated by '\n') into f,
 * starting from the defaults, with the reference's tokenizer
 * (source/readline.f90, source/readl2.f90:52-92) and its quirks (numeric
 * values truncated by int(), so chempot_increment 0.1 reads as 0; unknown
 * keys ignored; '!' comment lines).  An invalid value of a string key, where
 * the reference ends in error stop, returns PTBGPU_E_ARG with the reference's
 * message in errmsg (errlen bytes, may be NULL).  Host only, no context. */
int ptbgpu_pur_parse(const char *text, int64_t len, ptbgpu_pur_file *f, char *errmsg,
                     int64_t errlen);

/* Outcome of the chemical-potential search (source/purification.f90:232-353). */
typedef struct ptbgpu_pur_report {
    double chempot;       /* final chemical potential */
    double nel;           /* tr(P S) of the last purified matrix */
    int64_t cycles;       /* outer search cycles run */
    int64_t metric_cycles;/* Newton-Schulz cycles of the iterative metric (0: direct) */
} ptbgpu_pur_report;

/* Replaces purification_wrapper(pur,ndim,H,S,P) (source/purification.f90:28-80):
 * packed H, S -> packed P.  Where the reference ends in ERROR STOP (a NaN in
 * the iterative metric or in the purified matrix, :179-181,491-492,580-581,
 * 641-643) the call returns PTBGPU_E_NAN.  The direct metric and sign
 * diagonalisation use the engine's symmetric eigensolver (ndim <= 32768; larger
 * returns PTBGPU_E_ARG); the iterative metric with McWeeny, Pade or
 * Newton-Schulz purification is limited only by memory (ndim <= 46340).
 * rep may be NULL. */
int ptbgpu_purify(ptbgpu_ctx *ctx, const ptbgpu_pur_settings *pur, int64_t ndim, const double *H,
                  const double *S, double *P, ptbgpu_pur_report *rep);
int ptbgpu_purify_dev(ptbgpu_ctx *ctx, const ptbgpu_pur_settings *pur, int64_t ndim,
                      const double *dH, const double *dS, double *dP, ptbgpu_pur_report *rep);

/* Replaces adjust_thresholds_dp(ndim, Hmat) (source/metrics.f90:82-113): the
 * size-dependent (high, normal, low) thresholds the purification and the
 * density checks use, into vals[0..2] (the reference's default-REAL(4)
 * literals).  Host only: no context, no device work. */
int ptbgpu_thresholds(int64_t ndim, double *vals);

/* ---- dense spectral routines ----------------------------------------------
 * matrix_sign(ndim,A,S) and matrix_root(ndim,A,sqrtAinv) (source/sm.f90:410-457):
 * S = V sgn(e) V^T (sgn 0 for e == 0) and V e^-1/2 V^T from the eigenpairs of
 * the symmetric A (dense ndim x ndim, column-major; upper triangle read, as
 * dsyevd('U')).  ndim <= 32768. */
int ptbgpu_matrix_sign(ptbgpu_ctx *ctx, int64_t ndim, const double *A, double *S);
int ptbgpu_matrix_root(ptbgpu_ctx *ctx, int64_t ndim, const double *A, double *sqrtAinv);

/* FERMISMEAR(prt,norbs,nel,t,eig,occ,fod,e_fermi,s) (source/fermi.f:5-67),
 * host only: Fermi occupations at temperature t (K) of the ascending eig[norbs]
 * with nel occupied (1 <= nel < norbs); fod, e_fermi, s may be NULL. */
int ptbgpu_fermismear(int64_t norbs, int64_t nel, double t, const double *eig, double *occ,
                      double *fod, double *e_fermi, double *s);

/* solve2(mode,ndim,nel,nopen,homo,et,focc,H,S,P,e,U,fail)
 * (source/pgtb.f90:1066-1170), the -check reference density: full dsygvd of
 * the packed H, S; for mode >= 2 the eigenvalues are shifted to
 * e + shift0 + shift1 e (glob_par(15), glob_par(17) of the caller,
 * :1139-1143); occupations from occu + FERMISMEAR at et (K) for alpha and beta
 * (source/fermi.f); P = U diag(focc) U^T packed.  e[ndim] (ascending),
 * focc[ndim], U[ndim*ndim] (column-major, columns in e's order) and fail
 * (1 when the overlap is not positive definite: dsygvd INFO != 0) may be NULL.
 * ndim <= 32768. */
int ptbgpu_solve2(ptbgpu_ctx *ctx, int64_t mode, int64_t ndim, int64_t nel, int64_t nopen,
                  double et, double shift0, double shift1, const double *H, const double *S,
                  double *P, double *e, double *focc, double *U, int64_t *fail);

#ifdef __cplusplus
}
#endif
#endif /* PTBGPU_H */
