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
    int flags;  // bit 0: overlap block not positive definite; bit 1 (kSubExact):
                // reduced in FP64 in a mixed-precision build (no refinement)
    int mol;    // molecule of a batch (its chemical potential: mu[mol * kBisStride])
    int pad;
};
constexpr int kSubExact = 2;
// the largest dense matrix or submatrix the eigensolver takes (n^2 < 2^31:
// 32-bit element offsets inside a matrix)
constexpr int kDenseMax = 32768;

// One molecule of a batch on the device: its packed matrices and where its
// pattern bitmap, its submatrices' row unions and its submatrices begin.
struct MolDev {
    const double *H;
    const double *S;
    double *P;
    int64_t ndim;
    int64_t words;   // ceil(ndim / 32)
    int64_t bitoff;  // first word of its ndim x words bitmap
    int64_t unioff;  // first word of its submatrices' unions (words each)
    int sbeg;        // first submatrix (or k-means candidate group)
    int pad;
};

// Per-molecule bisection record: [0] mu, [1] nel, [2] logged passes, [3]
// converged flag, then (mu, nel) of up to 100 passes.
constexpr int kBisStride = 4 + 200;

// One product C = alpha op(A) op(B) + beta C of a variable-size batch.  With
// uplo = 1 only the lower triangle of C (row >= column) is formed: tiles above
// the diagonal are skipped (symmetric updates such as A -= V W^T + W V^T).
// uplo = 2 skips the same tiles but writes the diagonal tiles whole (for a
// product whose reader needs the lower tiles only).
// tri marks op(A) / op(B) as triangular (kTri* bits): each output tile runs
// only over the k range where both factors can be nonzero (triangular products
// at about half the flops).  The other triangle of such a factor must be
// stored as zeros.
struct GemmDesc {
    const double *A;
    const double *B;
    double *C;
    int M, N, K;
    int lda, ldb, ldc;
    double alpha, beta;
    int uplo;
    int tri;
    // optional (k_gemm only): every element written to C also goes, rounded
    // to float, to C32 at the same offset (the mixed build's float copy of K)
    float *C32;
};

constexpr int kTriALower = 1;  // op(A)(i, k) = 0 for k > i
constexpr int kTriAUpper = 2;  // op(A)(i, k) = 0 for k < i
constexpr int kTriBUpper = 4;  // op(B)(k, j) = 0 for k > j
constexpr int kTriBLower = 8;  // op(B)(k, j) = 0 for k < j

// One output tile of a batched GEMM: product g of the launch's descriptors
// (-1: nothing to do), tile row t & 0xffff and tile column t >> 16 (64 x 64).
struct GemmTile {
    int g;
    int t;
};

// A divide-and-conquer merge node: rows/cols [lo, hi) of submatrix s, split at sp.
struct DcNode {
    int s;
    int lo, sp, hi;
};

// largest leaf of the divide-and-conquer tree (rows of the tridiagonal piece
// that the leaf kernel diagonalises directly)
constexpr int kDcLeaf = 16;

// Blocked Householder factor of a batch: reflectors in the columns of V (the
// arena of the reduced matrices), T factors of all 32-reflector blocks.
struct QFactor {
    const double *V = nullptr;
    double *T = nullptr;          // 32 x 32 factors of the 32-reflector blocks
    std::vector<int64_t> toff;  // per submatrix, element offset of its first T block
    int maxblk = 0;
    double *T64 = nullptr;        // 64 x 64 factors of pairs of blocks
    std::vector<int64_t> toff64;
    int maxblk64 = 0;
    double *T128 = nullptr;       // 128 x 128 factors of pairs of those (what apply_q uses)
    std::vector<int64_t> toff128;
    int maxblk128 = 0;
};

// Right-hand sides of a batched triangular solve that are not the submatrices'
// own arena blocks: cols[s] columns at B + boff[s] (input) and X + xoff[s]
// (output), leading dimension of the submatrix.
struct RhsLayout {
    std::vector<int64_t> boff, xoff;
    std::vector<int> cols;
};

// Arithmetic of a batched product: FP64 MFMA (the plain path), or the BF16
// matrix cores with FP32 accumulation on split-BF16x3 or plain BF16 operands
// (lowp.hip; the mixed-precision build).
constexpr int kLowpNone = 0;
constexpr int kLowpBF16x3 = 1;
constexpr int kLowpBF16 = 2;

namespace ptb {

// pattern.hip: one launch per phase for every molecule of a batch.  mols
// (device) has nmol entries; smol (device) the molecule of every submatrix or
// k-means candidate group; bitmaps and unions at the table's offsets.
int pattern_bits(ptbgpu_ctx *ctx, const MolDev *mols, int nmol, int64_t max_ndim,
                 int64_t total_words, uint32_t *bits);
int group_union(ptbgpu_ctx *ctx, const uint32_t *bits, const MolDev *mols, const int *smol,
                int64_t max_words, const int *dgptr, const int *dgcols, int nsm, uint32_t *uni,
                int *dmcount);
int group_map(ptbgpu_ctx *ctx, const uint32_t *uni, const MolDev *mols, const int *smol,
              const SubDesc *ddesc, int nsm, int *dmap, const int *dgptr, const int *dgcols,
              int *doloc);
int gather(ptbgpu_ctx *ctx, const MolDev *mols, const SubDesc *ddesc, int nsm, int max_m,
           const int *dmap, double *Hd, double *Sd);

// lowp.hip: the BF16 matrix-core form of gemm_tiles (same descriptors and tile
// lists); f32: operands stored as float, else double
int gemm_tiles_bf(ptbgpu_ctx *ctx, bool ta, bool tb, const GemmDesc *ddesc, const GemmTile *dwork,
                  int64_t nwork, int lowp, bool f32);

// dense.hip
// f32: A, B, C are float arrays behind the descriptor's pointers (A B^T only).
// gemm_vbatch launches count x tiles(max_m, max_n) workgroups (shapes may be
// written on the device); gemm_tiles launches one workgroup per entry of a
// tile list (TileList).
int gemm_vbatch(ptbgpu_ctx *ctx, bool ta, bool tb, const GemmDesc *ddesc, int count, int max_m,
                int max_n, const char *site = nullptr, bool f32 = false);
int gemm_tiles(ptbgpu_ctx *ctx, bool ta, bool tb, const GemmDesc *ddesc, const GemmTile *dwork,
               int64_t nwork, const char *site = nullptr, bool f32 = false, bool one = false,
               int lowp = kLowpNone, int edge = 64);

// Tile lists of batched GEMMs whose shapes the host knows.  Each launch gets
// exactly the tiles that hold output (lower-triangle tiles for uplo), dealt to
// the 8 XCDs so that a product's tiles share one XCD's L2 and the XCDs get
// equal tile counts.  All launches' lists go to the device as one table.
// The lists are built on the host only when the shapes (or the tile-edge
// knob) differ from those of the table already in the device buffer of that
// name (repeated builds of the same system reuse it).
class TileList {
  public:
    // records the shapes of one launch over products h[0 .. count); returns
    // its index.  fp64: the launch runs in FP64 arithmetic on double storage
    // (never the low-precision or float-storage forms), so it may take
    // 128 x 128 tiles (same results as 64 x 64 ones)
    int add(const GemmDesc *h, int count, bool fp64 = false);
    // builds (if the shapes changed) and uploads the lists of every launch
    int upload(ptbgpu_ctx *ctx, const char *name);
    int launch(ptbgpu_ctx *ctx, int i, bool ta, bool tb, const GemmDesc *ddesc,
               const char *site = nullptr, bool f32 = false, int lowp = kLowpNone) const;

  private:
    void build(std::vector<GemmTile> &work, int64_t tile_mode, int num_cus);
    static constexpr char kOne = 1;       // every product is one k slice (K <= 32, no tri)
    static constexpr char kEdge128 = 2;   // 128 x 128 tiles
    std::vector<int> sig_;        // per launch: count, fp64, then M, N, K, uplo, tri per product
    int nlaunch_ = 0;
    std::vector<int64_t> start_;  // per launch: first entry in the table (+ the total)
    std::vector<char> flags_;     // per launch: kOne, kEdge128
    const GemmTile *dwork_ = nullptr;
};
int potrf_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                double *A, int *flags);
int transpose_sym_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m,
                        const double *in, double *out, bool symmetrize);
// Li = L^-1 of every lower-triangular L (zeros above the diagonal); T scratch
// with the arena layout
int inv_lower_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                    const double *L, double *Li, double *T);
int trsm_gemm(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
              const double *L, double *B, double *X, bool trans, const char *tag,
              const RhsLayout *rhs = nullptr);
int zero_upper_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, double *A);
// A's strict upper triangle from its lower one, in place
int mirror_lower_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, double *A);
int larft_all(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
              const double *V, const double *tau, QFactor &qf);
// float copy of the arena blocks (same offsets and leading dimensions)
int to_f32_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const double *A,
                 float *F);
// B = the lower triangle of A (FP64), per submatrix
int copy_lower_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const double *A,
                     double *B);
// FP64 reflectors (zero above the subdiagonal) and orthogonal tau from the
// float output of sytrd_batch_f32
int promote_reflectors(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const float *F,
                       double *V, double *tau);
int apply_q(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const QFactor &qf, double *Y,
            const std::vector<int64_t> &yoff, const std::vector<int> &ncols, bool trans,
            const char *tag);

// tridiag.hip
int sytrd_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                double *A, double *d, double *e, double *tau);
// the same on float storage (FP64 arithmetic): d, e, tau in FP64, reflectors in A
int sytrd_batch_f32(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                    float *A, double *d, double *e, double *tau);

// sm.hip
// With K32 and V given, the tridiagonalisation runs on float storage (K32) and
// the FP64 reflectors go to V; A (the FP64 matrix) is left intact.
// lowp: arithmetic of the divide-and-conquer merge products (kLowp*)
// k32_ready: the float copy of K's lower triangle is already in K32 (written
// by the product that formed K)
int eig_tri_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                  double *A, double *Z, double *W, double *X, double *dvec, double *evec,
                  double *tau, double *rot, QFactor &qf, float *K32 = nullptr,
                  double *V = nullptr, int lowp = 0, bool k32_ready = false);
int eig_sym_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                  double *A, double *Z, double *W, double *X, double *dvec, double *evec,
                  double *tau, double *rot);

// dense_solve.hip: eigenpairs of one dense symmetric matrix (upper triangle of
// M, leading dimension n), eigenvalues ascending
int dense_eig(ptbgpu_ctx *ctx, const double *M, int64_t n, double *V, std::vector<double> &w,
              const char *pfx);

// dc.hip
int dc_leaves(ptbgpu_ctx *ctx, const SubDesc *ddesc, const DcNode *dleaves, int nleaves,
              double *d, const double *e, double *Z);
// tiles: the level's GEMM tile list, two products per node (rows above and
// below the split: n1 x n and (n - n1) x n bounds, n = hi - lo; the merge
// kernel writes N = the node's non-deflated count <= n); dgemm holds
// 2 x count descriptors
// lowp: arithmetic of the merged-vector products (kLowp*)
// sizes (host, may be NULL): the nodes' row counts, for the size-bucketed
// merge launches.  The products of the first nfp64 nodes run in FP64 (launch 0
// of tiles), the rest in lowp (the next launch); nfp64 >= count: all FP64.
int dc_merge_level(ptbgpu_ctx *ctx, const SubDesc *ddesc, const DcNode *dnodes, int count,
                   int max_n, const double *e, double *D, double *Z, double *W, double *X,
                   double *rot, GemmDesc *dgemm, const TileList &tiles, int lowp = 0,
                   const std::vector<int> *sizes = nullptr, int nfp64 = 1 << 30);

// density.hip
// XA = [L(O,:)^T, Li(:,O)] per submatrix (Li = L^-1)
int xall_build(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_own, const int *doloc,
               const double *L, const double *Li, const int64_t *dxoff, double *XA);
int fe_proj(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const int64_t *dxoff,
            const double *F, double *fe);
// mu bisection of nmol molecules at once: molecule i owns the states
// [rbeg[i], rbeg[i+1]) of e and fe (device arrays); its record goes to
// dout + i * kBisStride
int bisect_batch(ptbgpu_ctx *ctx, const double *e, const double *fe, const int64_t *drbeg,
                 const int64_t *dnelref, int nmol, int64_t max_rows, double *dout);
int occ_compact(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const double *eig,
                const double *dmu, const double *Z, double *Zo, double *wocc, int *nocc,
                int *oidx = nullptr);
// FP64 refinement of the occupied eigenvectors of Q^T K Q (density.hip)
int refine_resid(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_nocc, const int *nocc,
                 const int *oidx, const double *eig, const double *Zo, double *R);
int refine_corr(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_nocc, const int *nocc,
                const int *oidx, const double *eig, const double *dmu, double *C, double *thn);
int refine_shift(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_nocc, const int *nocc,
                 double *G);
int refine_scatter(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_nocc, const int *nocc,
                   const int *oidx, const double *Xo, const double *thn, double *Z, double *eig);
// packed P of every molecule zeroed (when zero), then the owned columns of
// every submatrix of ddesc written into its molecule's P
int dmat_compact(ptbgpu_ctx *ctx, const MolDev *mols, int nmol, int64_t max_packed,
                 const SubDesc *ddesc, int nsm, int max_nocc, int max_own, const int *dgcols,
                 const int *doloc, const int *dmap, const double *wocc, const int *nocc,
                 const double *U, bool zero = true);

}  // namespace ptb
