// Electron-count weights, chemical-potential bisection and the owned-column
// density scatter (source/sm.f90:304-362, :190-198, :365-383, :387).
//
// fe(k) = sum_{owned rows i} (Ssm U)(i,k) U(i,k).  With Ssm = L L^T,
// U = L^-T Q Z (Q the Householder product, Z the tridiagonal eigenvectors) the
// owned rows are (L Q Z)(O,:) and (L^-T Q Z)(O,:): 2c row vectors
// [L(O,:); (L^-1 e_O)^T] (rows of L and columns of L^-1) are pushed through Q
// and Z (c = owned columns), so no full back-transformation is needed for fe;
// owned rows accumulate in the reference's order.
//
// The bisection runs per molecule of a batch, every pass one launch over all
// (eigenvalue, fe) pairs (per-block partials, then a fixed-order sum), so it is
// deterministic; it reproduces the reference's decisions: mu = (mua+mub)/2 from [-10, 10],
// occupations 2/1/0 (tie e == mu -> 1), stop on |mua-mub| < 1e-3 or
// |nel-nelref| < 1e-5 with both literals rounded to single precision, then one
// more pass at the same mu.
//
// The density of owned column j is P(:, j) = U diag(occ) U(loc_j, :)^T over the
// occupied states only, so after the bisection only the occupied eigenvector
// columns are compacted, back-transformed and solved with L^-T; packed P
// (zeroed first) receives rows map(a) <= j, which is the reference's packsym of
// P(map, comb) with no symmetrisation.
#include "sm_internal.hpp"

namespace {

constexpr int kT = 256;
constexpr int kG = 8;  // owned columns per pass

// Right-hand sides of the owned-row projections, per submatrix with c owned
// columns (local rows o_g = oloc[g]):
//   XA(:, g)     = L(o_g, :)^T            (row o_g of L, zero past o_g)
//   XA(:, c + g) = Li(:, o_g) = L^-1 e_{o_g}
// grid: (submatrix, owned column)
__global__ __launch_bounds__(kT) void k_xall_build(const SubDesc *__restrict__ desc,
                                                   const int *__restrict__ oloc,
                                                   const double *__restrict__ La,
                                                   const double *__restrict__ Lia,
                                                   const int64_t *__restrict__ xoff,
                                                   double *__restrict__ XA) {
    const SubDesc d = desc[blockIdx.x];
    const int g = d.gbeg + blockIdx.y;
    if (g >= d.gend) return;
    const int c = d.gend - d.gbeg;
    const int row = oloc[g];
    const double *L = La + d.off;
    const double *li = Lia + d.off + (int64_t)row * d.ld;
    double *xa = XA + xoff[blockIdx.x] + (int64_t)blockIdx.y * d.ld;
    double *xb = XA + xoff[blockIdx.x] + (int64_t)(c + blockIdx.y) * d.ld;
    // (L^-1 is lower triangular: its column `row` is zero above the
    // diagonal, where the inverse's arena is not written)
    for (int t = threadIdx.x; t < d.m; t += kT) {
        xa[t] = t <= row ? L[(int64_t)t * d.ld + row] : 0.0;
        xb[t] = t >= row ? li[t] : 0.0;
    }
}

// fe(k) = sum_g F(k, g) F(k, c + g) over owned columns in order, with
// F = Z^T Q^T [L(O,:)^T, L^-1 e_O]: F(k, g) = (L Y)(o_g, k) = (Ssm U)(o_g, k) and
// F(k, c + g) = (L^-T Y)(o_g, k) = U(o_g, k).  Zero for a submatrix whose
// overlap block failed the Cholesky.
__global__ __launch_bounds__(kT) void k_fe_proj(const SubDesc *__restrict__ desc,
                                                const int64_t *__restrict__ xoff,
                                                const double *__restrict__ F,
                                                double *__restrict__ fe_all) {
    const SubDesc d = desc[blockIdx.x];
    const int k = blockIdx.y * kT + threadIdx.x;
    if (k >= d.m) return;
    const int c = d.gend - d.gbeg;
    const double *f = F + xoff[blockIdx.x] + k;
    double s = 0.0;
    if (!(d.flags & 1))
        for (int g = 0; g < c; ++g) s += f[(int64_t)g * d.ld] * f[(int64_t)(c + g) * d.ld];
    fe_all[d.mapoff + k] = s;
}

__device__ __forceinline__ double occ_of(double e, double mu) {
    return e < mu ? 2.0 : (e == mu ? 1.0 : 0.0);
}

// Bisection state of one molecule.
struct BisState {
    double mua, mub, mu;
    int iter, logged, finaliter, done;
};

__global__ void k_bisect_init(BisState *__restrict__ st, int nmol) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nmol) st[i] = BisState{-10.0, 10.0, 0.0, 0, 0, 0, 0};
}

// One pass, part 1: partial sums of fe(k) occ(e_k, mu) over chunk b of
// molecule i's states.  grid: (nblk, nmol)
__global__ __launch_bounds__(kT) void k_nel_partial(const double *__restrict__ e,
                                                    const double *__restrict__ fe,
                                                    const int64_t *__restrict__ rbeg,
                                                    const BisState *__restrict__ st, int nblk,
                                                    double *__restrict__ partial) {
    const int i = blockIdx.y;
    if (st[i].done) return;
    const double mu = st[i].mu;
    const int64_t lo = rbeg[i], n = rbeg[i + 1] - lo;
    const int64_t chunk = (n + nblk - 1) / nblk;
    const int64_t b0 = lo + (int64_t)blockIdx.x * chunk, b1 = min(lo + n, b0 + chunk);
    double part = 0.0;
    for (int64_t k = b0 + threadIdx.x; k < b1; k += kT) part += fe[k] * occ_of(e[k], mu);
    __shared__ double red[kT / PTB_WAVE];
    part = wave_sum(part);
    if (threadIdx.x % PTB_WAVE == 0) red[threadIdx.x / PTB_WAVE] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kT / PTB_WAVE; ++w) t += red[w];
        partial[(int64_t)i * nblk + blockIdx.x] = t;
    }
}

// One pass, part 2 (one thread per molecule): nel = the partials summed in
// fixed order, then the reference's step (source/sm.f90:365-383): a pass after
// convergence is the final one (its nel is reported); otherwise log (mu, nel),
// test |mua-mub| < 1e-3 or |nel-nelref| < 1e-5 (REAL(4) literals), move the
// bracket and bisect.  At most 100 passes, as the reference's loop.
// out (per molecule, kBisStride): [0] mu, [1] nel, [2] logged passes,
// [3] converged, then (mu, nel) of every logged pass.
__global__ void k_bisect_step(BisState *__restrict__ st, const double *__restrict__ partial,
                              int nblk, const int64_t *__restrict__ nelref, int nmol,
                              double *__restrict__ out_all, int *__restrict__ ndone) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nmol) return;
    BisState s = st[i];
    if (s.done) return;
    const float tol_mu = 1e-3f, tol_nel = 1e-5f;  // REAL(4) literals of the reference
    double *out = out_all + (size_t)i * kBisStride;
    double nel = 0.0;
    for (int b = 0; b < nblk; ++b) nel += partial[(int64_t)i * nblk + b];
    const double ref = (double)nelref[i];
    s.iter += 1;
    bool finish = false;
    if (s.finaliter) {
        finish = true;
    } else {
        out[4 + 2 * s.logged] = s.mu;
        out[5 + 2 * s.logged] = nel;
        s.logged += 1;
        if (fabs(s.mua - s.mub) < (double)tol_mu || fabs(nel - ref) < (double)tol_nel) {
            s.finaliter = 1;
        } else if (nel > ref) {
            s.mub = s.mu;
        } else {
            s.mua = s.mu;
        }
        finish = s.iter >= 100;
    }
    if (finish) {
        out[0] = s.mu;
        out[1] = nel;
        out[2] = (double)s.logged;
        out[3] = s.finaliter ? 1.0 : 0.0;
        s.done = 1;
        atomicAdd(ndone, 1);
    } else {
        s.mu = (s.mua + s.mub) / 2.0;
    }
    st[i] = s;
}

// Occupied states of each submatrix at the final mu, in ascending state
// order: weights occ(e_k) (2, or 1 for e_k == mu) into wocc at the submatrix's
// row offset, their count into nocc, and the eigenvector columns Z(:, k) copied
// to the first nocc columns of Zo.  One workgroup per submatrix; a submatrix
// whose overlap block failed the Cholesky contributes no states.
__global__ __launch_bounds__(kT) void k_occ_compact(const SubDesc *__restrict__ desc,
                                                    const double *__restrict__ eig_all,
                                                    const double *__restrict__ mu_p,
                                                    const double *__restrict__ Za,
                                                    double *__restrict__ Zo,
                                                    double *__restrict__ wocc,
                                                    int *__restrict__ nocc,
                                                    int *__restrict__ oidx) {
    extern __shared__ int occk[];  // m
    __shared__ int wtot[kT / PTB_WAVE];
    __shared__ int base;
    const SubDesc d = desc[blockIdx.x];
    const double mu = mu_p[(size_t)d.mol * kBisStride];
    const double *eig = eig_all + d.mapoff;
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    for (int k0 = 0; k0 < d.m; k0 += kT) {
        const int k = k0 + threadIdx.x;
        const double w = (k < d.m && !(d.flags & 1)) ? occ_of(eig[k], mu) : 0.0;
        const unsigned long long mask = __ballot(w != 0.0);
        const int before = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) wtot[wid] = __popcll(mask);
        __syncthreads();
        int off = base;
        for (int q = 0; q < wid; ++q) off += wtot[q];
        if (w != 0.0) {
            occk[off + before] = k;
            wocc[d.mapoff + off + before] = w;
            if (oidx) oidx[d.mapoff + off + before] = k;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int q = 0; q < kT / PTB_WAVE; ++q) t += wtot[q];
            base += t;
        }
        __syncthreads();
    }
    const int n = base;
    if (threadIdx.x == 0) nocc[blockIdx.x] = n;
    const double *Z = Za + d.off;
    double *Y = Zo + d.off;
    for (int t = 0; t < n; ++t) {
        const double *src = Z + (int64_t)occk[t] * d.ld;
        double *dst = Y + (int64_t)t * d.ld;
        for (int r = threadIdx.x; r < d.m; r += kT) dst[r] = src[r];
    }
}

// ---- FP64 refinement of the occupied eigenvectors (mixed precision) -------
// With float storage in the tridiagonalisation, T and Q satisfy
// Q^T K Q = T + D with D a dense perturbation of float rounding size, and the
// exact eigenpairs (theta, Z) of T are first-order accurate for K.  One
// refinement sweep, in the basis of Z, for the occupied columns z_j (compacted
// in Zo, state index oidx[j]):
//   R = Q^T K Q Zo - Zo diag(theta_o)          (= D Zo, FP64)
//   C = Z^T R,  E(i, j) = C(i, j) / (theta_j - theta_i) for unoccupied i, 0 else
//   Zo <- (Zo + Z E) (3 I - G) / 2,  G = Zo^T Zo   (first-order Loewdin)
//   theta_j <- theta_j + C(oidx_j, j)              (the Rayleigh quotient)
// Rotations among occupied states leave the density unchanged, so only the
// occupied-unoccupied couplings are removed; the error drops from O(|D|) to
// O(|D|^2 / gap).

// R(:, j) -= theta_{oidx_j} Zo(:, j).  grid: (submatrix, occupied column)
__global__ __launch_bounds__(kT) void k_refine_resid(const SubDesc *__restrict__ desc,
                                                     const int *__restrict__ nocc,
                                                     const int *__restrict__ oidx,
                                                     const double *__restrict__ eig_all,
                                                     const double *__restrict__ Zo,
                                                     double *__restrict__ R) {
    const SubDesc d = desc[blockIdx.x];
    const int j = blockIdx.y;
    if (j >= nocc[blockIdx.x]) return;
    const double th = eig_all[d.mapoff + oidx[d.mapoff + j]];
    const double *z = Zo + d.off + (int64_t)j * d.ld;
    double *r = R + d.off + (int64_t)j * d.ld;
    for (int i = threadIdx.x; i < d.m; i += kT) r[i] -= th * z[i];
}

// C -> E in place and the refined eigenvalues thn[mapoff + j].
// grid: (submatrix, occupied column)
__global__ __launch_bounds__(kT) void k_refine_corr(const SubDesc *__restrict__ desc,
                                                    const int *__restrict__ nocc,
                                                    const int *__restrict__ oidx,
                                                    const double *__restrict__ eig_all,
                                                    const double *__restrict__ mu_p,
                                                    double *__restrict__ C,
                                                    double *__restrict__ thn) {
    const SubDesc d = desc[blockIdx.x];
    const int j = blockIdx.y;
    if (j >= nocc[blockIdx.x]) return;
    const double mu = mu_p[(size_t)d.mol * kBisStride];
    const double *eig = eig_all + d.mapoff;
    const int kj = oidx[d.mapoff + j];
    const double th = eig[kj];
    double *c = C + d.off + (int64_t)j * d.ld;
    for (int i = threadIdx.x; i < d.m; i += kT) {
        const double x = c[i];
        if (i == kj) thn[d.mapoff + j] = th + x;
        const double gap = th - eig[i];
        c[i] = (occ_of(eig[i], mu) == 0.0 && fabs(gap) > 1e-12) ? x / gap : 0.0;
    }
}

// G <- (3 I - G) / 2 on the leading nocc x nocc block.
// grid: (submatrix, column)
__global__ __launch_bounds__(kT) void k_refine_shift(const SubDesc *__restrict__ desc,
                                                     const int *__restrict__ nocc,
                                                     double *__restrict__ G) {
    const SubDesc d = desc[blockIdx.x];
    const int j = blockIdx.y, n = nocc[blockIdx.x];
    if (j >= n) return;
    double *g = G + d.off + (int64_t)j * d.ld;
    for (int i = threadIdx.x; i < n; i += kT) g[i] = (i == j ? 1.5 : 0.0) - 0.5 * g[i];
}

// Z(:, oidx_j) = Xo(:, j), eig[oidx_j] = thn[j].  grid: (submatrix, occupied column)
__global__ __launch_bounds__(kT) void k_refine_scatter(const SubDesc *__restrict__ desc,
                                                       const int *__restrict__ nocc,
                                                       const int *__restrict__ oidx,
                                                       const double *__restrict__ Xo,
                                                       const double *__restrict__ thn,
                                                       double *__restrict__ Z,
                                                       double *__restrict__ eig_all) {
    const SubDesc d = desc[blockIdx.x];
    const int j = blockIdx.y;
    if (j >= nocc[blockIdx.x]) return;
    const int kj = oidx[d.mapoff + j];
    const double *x = Xo + d.off + (int64_t)j * d.ld;
    double *z = Z + d.off + (int64_t)kj * d.ld;
    for (int i = threadIdx.x; i < d.m; i += kT) z[i] = x[i];
    if (threadIdx.x == 0) eig_all[d.mapoff + kj] = thn[d.mapoff + j];
}

// Packed P of every molecule zeroed.  grid: (blocks, molecule)
__global__ __launch_bounds__(kT) void k_zero_packed(const MolDev *__restrict__ mols) {
    const MolDev mo = mols[blockIdx.y];
    const int64_t n = mo.ndim * (mo.ndim + 1) / 2;
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
        mo.P[i] = 0.0;
}

// P(:, j) = U_o diag(w) U_o(loc_j, :)^T for the owned columns j, with U_o the
// nocc occupied generalized eigenvectors (first columns of the arena block);
// the packed P of the submatrix's molecule (zeroed first) receives rows
// map(a) <= j, the reference's packsym of P(map, comb) with no symmetrisation.
// grid: (submatrix, chunk of kG owned columns).
__global__ __launch_bounds__(kT) void k_dmat_compact(const MolDev *__restrict__ mols,
                                                     const SubDesc *__restrict__ desc,
                                                     const int *__restrict__ gcols,
                                                     const int *__restrict__ oloc,
                                                     const int *__restrict__ map_all,
                                                     const double *__restrict__ wocc,
                                                     const int *__restrict__ nocc,
                                                     const double *__restrict__ Ua) {
    extern __shared__ double wsh[];  // [kG][n]
    const SubDesc d = desc[blockIdx.x];
    const int g0 = d.gbeg + blockIdx.y * kG;
    const int n = nocc[blockIdx.x];
    if (g0 >= d.gend || n == 0) return;
    double *P = mols[d.mol].P;
    const int ng = min(kG, d.gend - g0);
    const double *U = Ua + d.off;
    const double *w = wocc + d.mapoff;
    const int *map = map_all + d.mapoff;
    for (int l = threadIdx.x; l < kG * n; l += kT) {
        const int q = l / n, t = l % n;
        wsh[(size_t)q * n + t] = q < ng ? w[t] * U[(int64_t)t * d.ld + oloc[g0 + q]] : 0.0;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < d.m; a += kT) {
        double acc[kG];
#pragma unroll
        for (int q = 0; q < kG; ++q) acc[q] = 0.0;
        for (int t = 0; t < n; ++t) {
            const double u = U[(int64_t)t * d.ld + a];
#pragma unroll
            for (int q = 0; q < kG; ++q) acc[q] += u * wsh[(size_t)q * n + t];
        }
        const int64_t row = map[a];
#pragma unroll
        for (int q = 0; q < kG; ++q) {
            if (q < ng) {
                const int64_t j = gcols[g0 + q];
                if (row <= j) P[pk(row, j)] = acc[q];
            }
        }
    }
}

}  // namespace

namespace ptb {

int xall_build(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_own, const int *doloc,
               const double *L, const double *Li, const int64_t *dxoff, double *XA) {
    ProfScope prof_(ctx, "fe_rows");
    if (max_own <= 0) return PTBGPU_OK;
    hipLaunchKernelGGL(k_xall_build, dim3(nsm, max_own), dim3(kT), 0, ctx->stream, ddesc, doloc, L,
                       Li, dxoff, XA);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int fe_proj(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const int64_t *dxoff,
            const double *F, double *fe) {
    ProfScope prof_(ctx, "fe_finish");
    hipLaunchKernelGGL(k_fe_proj, dim3(nsm, (max_m + kT - 1) / kT), dim3(kT), 0, ctx->stream,
                       ddesc, dxoff, F, fe);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int bisect_batch(ptbgpu_ctx *ctx, const double *e, const double *fe, const int64_t *drbeg,
                 const int64_t *dnelref, int nmol, int64_t max_rows, double *dout) {
    ProfScope prof_(ctx, "bisect");
    auto *st = static_cast<BisState *>(ws(ctx, "bis_state", sizeof(BisState) * nmol));
    auto *ndone = static_cast<int *>(ws(ctx, "bis_done", sizeof(int)));
    const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(256, (max_rows + 4095) / 4096));
    auto *partial = static_cast<double *>(ws(ctx, "bis_partial", sizeof(double) * nblk * nmol));
    if (!st || !ndone || !partial) return PTBGPU_E_NOMEM;
    const int tb = (nmol + 63) / 64;
    hipLaunchKernelGGL(k_bisect_init, dim3(tb), dim3(64), 0, ctx->stream, st, nmol);
    PTB_CHECK_LAUNCH(ctx);
    PTB_HIP(ctx, hipMemsetAsync(ndone, 0, sizeof(int), ctx->stream));
    // passes in chunks (a pass after a molecule is done returns at once);
    // 101 passes cover the reference's 100 and the final one
    int passes = 0, chunk = 24, done = 0;
    while (passes < 101 && done < nmol) {
        const int n = std::min(chunk, 101 - passes);
        for (int p = 0; p < n; ++p) {
            hipLaunchKernelGGL(k_nel_partial, dim3(nblk, nmol), dim3(kT), 0, ctx->stream, e, fe,
                               drbeg, st, nblk, partial);
            hipLaunchKernelGGL(k_bisect_step, dim3(tb), dim3(64), 0, ctx->stream, st, partial,
                               nblk, dnelref, nmol, dout, ndone);
        }
        PTB_CHECK_LAUNCH(ctx);
        passes += n;
        PTB_HIP(ctx, hipMemcpyAsync(&done, ndone, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
        chunk = 8;
    }
    return PTBGPU_OK;
}

int occ_compact(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const double *eig,
                const double *dmu, const double *Z, double *Zo, double *wocc, int *nocc,
                int *oidx) {
    ProfScope prof_(ctx, "occ_compact");
    hipLaunchKernelGGL(k_occ_compact, dim3(nsm), dim3(kT), sizeof(int) * (size_t)max_m,
                       ctx->stream, ddesc, eig, dmu, Z, Zo, wocc, nocc, oidx);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int refine_resid(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_nocc, const int *nocc,
                 const int *oidx, const double *eig, const double *Zo, double *R) {
    ProfScope prof_(ctx, "refine");
    if (max_nocc <= 0) return PTBGPU_OK;
    hipLaunchKernelGGL(k_refine_resid, dim3(nsm, max_nocc), dim3(kT), 0, ctx->stream, ddesc, nocc,
                       oidx, eig, Zo, R);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int refine_corr(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_nocc, const int *nocc,
                const int *oidx, const double *eig, const double *dmu, double *C, double *thn) {
    ProfScope prof_(ctx, "refine");
    if (max_nocc <= 0) return PTBGPU_OK;
    hipLaunchKernelGGL(k_refine_corr, dim3(nsm, max_nocc), dim3(kT), 0, ctx->stream, ddesc, nocc,
                       oidx, eig, dmu, C, thn);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int refine_shift(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_nocc, const int *nocc,
                 double *G) {
    ProfScope prof_(ctx, "refine");
    if (max_nocc <= 0) return PTBGPU_OK;
    hipLaunchKernelGGL(k_refine_shift, dim3(nsm, max_nocc), dim3(kT), 0, ctx->stream, ddesc, nocc, G);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int refine_scatter(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_nocc, const int *nocc,
                   const int *oidx, const double *Xo, const double *thn, double *Z, double *eig) {
    ProfScope prof_(ctx, "refine");
    if (max_nocc <= 0) return PTBGPU_OK;
    hipLaunchKernelGGL(k_refine_scatter, dim3(nsm, max_nocc), dim3(kT), 0, ctx->stream, ddesc, nocc,
                       oidx, Xo, thn, Z, eig);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int dmat_compact(ptbgpu_ctx *ctx, const MolDev *mols, int nmol, int64_t max_packed,
                 const SubDesc *ddesc, int nsm, int max_nocc, int max_own, const int *dgcols,
                 const int *doloc, const int *dmap, const double *wocc, const int *nocc,
                 const double *U, bool zero) {
    ProfScope prof_(ctx, "dmat_scatter");
    if (zero) {
        const int zb = (int)std::max<int64_t>(1, std::min<int64_t>((max_packed + kT - 1) / kT,
                                                                  (int64_t)ctx->num_cus * 8 / nmol + 1));
        hipLaunchKernelGGL(k_zero_packed, dim3(zb, nmol), dim3(kT), 0, ctx->stream, mols);
        PTB_CHECK_LAUNCH(ctx);
    }
    if (max_nocc <= 0 || max_own <= 0) return PTBGPU_OK;
    const size_t shmem = sizeof(double) * (size_t)kG * max_nocc;
    if (shmem > 150 * 1024)
        return fail(ctx, PTBGPU_E_ARG, "too many occupied states for dmat (%d)", max_nocc);
    hipLaunchKernelGGL(k_dmat_compact, dim3(nsm, (max_own + kG - 1) / kG), dim3(kT), shmem,
                       ctx->stream, mols, ddesc, dgcols, doloc, dmap, wocc, nocc, U);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

}  // namespace ptb
