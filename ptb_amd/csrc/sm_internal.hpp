// Internal interfaces of the submatrix density build.
#pragma once

#include "common.hpp"

// One submatrix of the batch.  Dense per-submatrix matrices live in arenas at
// element offset `off`, column-major with leading dimension `ld`.  Per-row
// vectors (map, eigenvalues, fe, d, e, tau) live at offset `mapoff`.
struct SubDesc {
    int64_t off;
    int m;
    int ld;
    int mapoff;
    int gbeg;   // owned columns: gcols[gbeg..gend)
    int gend;
    int flags;  // bit 0: overlap block not positive definite
};

// One product C = alpha op(A) op(B) + beta C of a variable-size batch.
struct GemmDesc {
    const double *A;
    const double *B;
    double *C;
    int M, N, K;
    int lda, ldb, ldc;
    double alpha, beta;
};

// A divide-and-conquer merge node: rows/cols [lo, hi) of submatrix s, split at sp.
struct DcNode {
    int s;
    int lo, sp, hi;
};

namespace ptb {

// pattern.hip
int pattern_bits(ptbgpu_ctx *ctx, const double *dH, const double *dS, int64_t ndim,
                 int64_t words, uint32_t *bits);
int group_union(ptbgpu_ctx *ctx, const uint32_t *bits, int64_t words, const int *dgptr,
                const int *dgcols, int nsm, uint32_t *uni, int *dmcount);
int group_map(ptbgpu_ctx *ctx, const uint32_t *uni, int64_t words, const SubDesc *ddesc,
              int nsm, int *dmap, const int *dgptr, const int *dgcols, int *doloc);
int gather(ptbgpu_ctx *ctx, const double *dH, const double *dS, const SubDesc *ddesc, int nsm,
           int max_m, const int *dmap, double *Hd, double *Sd);

// dense.hip
int gemm_vbatch(ptbgpu_ctx *ctx, bool ta, bool tb, const GemmDesc *ddesc, int count, int max_m,
                int max_n);
int potrf_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                double *A, int *flags);
int trsm_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const double *L,
               double *B, bool trans);
int transpose_sym_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m,
                        const double *in, double *out, bool symmetrize);
int trsm_gemm(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
              const double *L, double *B, double *X, bool trans, const char *tag);
int zero_upper_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, double *A);
int larft_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, const double *V,
                const double *tau, int q, int nb, double *T);

// tridiag.hip
int sytrd_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                double *A, double *W, double *d, double *e, double *tau);

// sm.hip
int eig_sym_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                  double *A, double *Z, double *W, double *X, double *dvec, double *evec,
                  double *tau, double *rot, double *Tb);

// dc.hip
int dc_leaves(ptbgpu_ctx *ctx, const SubDesc *ddesc, const DcNode *dleaves, int nleaves,
              double *d, const double *e, double *Z);
int dc_merge_level(ptbgpu_ctx *ctx, const SubDesc *ddesc, const DcNode *dnodes, int count,
                   int max_n, const double *e, double *D, double *Z, double *W, double *X,
                   double *rot, GemmDesc *dgemm);

// density.hip
int fe_owned_rows(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, int max_own,
                  const int *doloc, const double *L, const double *Y, double *Lown, double *SY);
int fe_finish(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const int *doloc,
              const double *SY, const double *U, double *fe);
int bisect_mu(ptbgpu_ctx *ctx, const double *e, const double *fe, int64_t total, int64_t nelref,
              double *dout);
int dmat_scatter(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, int max_own,
                 const int *dgcols, const int *doloc, const int *dmap, const double *eig,
                 const double *U, const double *dmu, double *P);

}  // namespace ptb
