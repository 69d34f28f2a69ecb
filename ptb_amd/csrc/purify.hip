// Density-matrix purification on the GPU: purification_wrapper and everything
// it drives (source/purification.f90:28-782, settings of
// source/purification_settings.f90:29-86, norms of source/norms.f90).
//
// All matrices are dense N x N, column-major, leading dimension N.  Every
// product is the engine's MFMA FP64 GEMM; norms, traces and NaN checks are
// device reductions whose few scalars come back to the host, which runs the
// reference's scalar control flow (chemical-potential search, convergence
// tests) in the reference's order.  The GEMM descriptors of one call are
// uploaded once into a small device table, so an iteration launches GEMMs
// back to back and synchronises only where a convergence test needs a norm.
// The metric S^-1 or S^-1/2 is the eigen-route V f(w) V^T through the batched
// symmetric eigensolver (ndim <= kDenseMax) or the reference's Newton-Schulz
// iterations.
#include <cmath>
#include <cstring>

#include "sm_internal.hpp"

namespace {

constexpr int kT = 256;
constexpr int kMaxGemm = 32;   // distinct GEMM descriptors of one call

// dense symmetric matrix from the packed upper triangle (blowsym)
__global__ void k_unpack(const double *__restrict__ p, int64_t n, double *__restrict__ A) {
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x) {
        const double *col = p + j * (j + 1) / 2;
        for (int64_t i = threadIdx.x; i <= j; i += kT) {
            const double v = col[i];
            A[j * n + i] = v;
            A[i * n + j] = v;
        }
    }
}

// packsym: packed element (j,i), j <= i, from column i, upper triangle
__global__ void k_pack(const double *__restrict__ A, int64_t n, double *__restrict__ p) {
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        double *col = p + i * (i + 1) / 2;
        for (int64_t j = threadIdx.x; j <= i; j += kT) col[j] = A[i * n + j];
    }
}

// C = a A + b B + c I   (B may be null; C may alias A or B)
__global__ void k_axpby(double a, const double *A, double b, const double *B, double c,
                        int64_t n, double *C) {
    const int64_t tot = n * n;
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < tot;
         k += (int64_t)gridDim.x * kT) {
        double v = a * A[k];
        if (B) v += b * B[k];
        if (k / n == k % n) v += c;
        C[k] = v;
    }
}

// C(:, j) = A(:, j) * w(j)
__global__ void k_colscale(const double *__restrict__ A, const double *__restrict__ w, int64_t n,
                           double *__restrict__ C) {
    const int64_t tot = n * n;
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < tot;
         k += (int64_t)gridDim.x * kT)
        C[k] = A[k] * w[k / n];
}

// per-block partials: [0] sum a^2, [1] sum (I - a)^2, [2] sum a*b (B optional),
// [3] NaN count
__global__ void k_reduce(const double *__restrict__ A, const double *__restrict__ B, int64_t n,
                         double *__restrict__ part) {
    __shared__ double red[16];
    double s2 = 0.0, si = 0.0, sab = 0.0, nan = 0.0;
    const int64_t tot = n * n;
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < tot;
         k += (int64_t)gridDim.x * kT) {
        const double a = A[k];
        s2 += a * a;
        const double d = ((k / n == k % n) ? 1.0 : 0.0) - a;
        si += d * d;
        if (B) sab += a * B[k];
        if (a != a) nan += 1.0;
    }
    const double t0 = block_sum(s2, red);
    const double t1 = block_sum(si, red);
    const double t2 = block_sum(sab, red);
    const double t3 = block_sum(nan, red);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 0] = t0;
        part[4 * blockIdx.x + 1] = t1;
        part[4 * blockIdx.x + 2] = t2;
        part[4 * blockIdx.x + 3] = t3;
    }
}

// per column j: [j] = sum_i |A(i,j)| (gershgorin, source/norms.f90:5-18),
// [n + j] = sum_i |A(j,i)| (infinity, :32-49), [2n + j] = sum_i A(i,j)^2
__global__ void k_col_sums(const double *__restrict__ A, int64_t n, double *__restrict__ out) {
    __shared__ double red[16];
    const int64_t j = blockIdx.x;
    double c = 0.0, r = 0.0, q = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kT) {
        const double v = A[j * n + i];
        c += fabs(v);
        q += v * v;
        r += fabs(A[i * n + j]);
    }
    c = block_sum(c, red);
    r = block_sum(r, red);
    q = block_sum(q, red);
    if (threadIdx.x == 0) {
        out[j] = c;
        out[n + j] = r;
        out[2 * n + j] = q;
    }
}

// thresholds of adjust_thresholds_dp (source/metrics.f90:82-113); the
// literals there are default REAL(4)
void thresholds(int64_t ndim, double &high, double &normal, double &low) {
    if (ndim < 300) {
        high = 1e-10f; normal = 1e-9f; low = 1e-8f;
    } else if (ndim < 2000) {
        high = 1e-7f; normal = 1e-6f; low = 1e-5f;
    } else if (ndim < 15000) {
        high = 1e-5f; normal = 1e-4f; low = 1e-3f;
    } else {
        high = 1e-3f; normal = 1e-2f; low = 1e-1f;
    }
}

enum { MCWEENY = 1, PADE = 2, NEWTON = 3, SIGN_DIAG = 4 };
enum { INV = 1, INV_SQRT = 2, SQRT_ = 3 };

struct Dense {
    ptbgpu_ctx *ctx;
    int64_t n;
    int blocks;
    double *part;     // 4 * blocks reduction partials
    double *sums;     // 3n column sums
    double *wvec;     // n diagonal weights
    GemmDesc *dg;     // kMaxGemm descriptors
    std::vector<GemmDesc> table;

    // C = alpha op(A) op(B) + beta C, all n x n
    int gemm(const double *A, bool ta, const double *B, bool tb, double *C, double alpha = 1.0,
             double beta = 0.0) {
        const GemmDesc g{A, B, C, (int)n, (int)n, (int)n, (int)n, (int)n, (int)n, alpha, beta};
        int slot = -1;
        for (size_t i = 0; i < table.size(); ++i)
            if (std::memcmp(&table[i], &g, sizeof g) == 0) slot = (int)i;
        if (slot < 0) {
            if ((int)table.size() >= kMaxGemm)
                return ptb::fail(ctx, PTBGPU_E_ARG, "purification GEMM table full");
            slot = (int)table.size();
            table.push_back(g);
            PTB_HIP(ctx, hipMemcpy(dg + slot, &g, sizeof g, hipMemcpyHostToDevice));
        }
        return ptb::gemm_vbatch(ctx, ta, tb, dg + slot, 1, (int)n, (int)n);
    }
    int axpby(double a, const double *A, double b, const double *B, double c, double *C) {
        hipLaunchKernelGGL(k_axpby, dim3(blocks), dim3(kT), 0, ctx->stream, a, A, b, B, c, n, C);
        PTB_CHECK_LAUNCH(ctx);
        return PTBGPU_OK;
    }
    int copy(const double *A, double *C) {
        PTB_HIP(ctx, hipMemcpyAsync(C, A, sizeof(double) * n * n, hipMemcpyDeviceToDevice,
                                    ctx->stream));
        return PTBGPU_OK;
    }
    // r[0] = sum a^2, r[1] = sum (I - a)^2, r[2] = sum a b, r[3] = NaN count
    int reduce(const double *A, const double *B, double r[4]) {
        hipLaunchKernelGGL(k_reduce, dim3(blocks), dim3(kT), 0, ctx->stream, A, B, n, part);
        PTB_CHECK_LAUNCH(ctx);
        auto *h = static_cast<double *>(ptb::hws(ctx, "pur_part_h", sizeof(double) * 4 * blocks));
        if (!h) return PTBGPU_E_NOMEM;
        PTB_HIP(ctx, hipMemcpyAsync(h, part, sizeof(double) * 4 * blocks, hipMemcpyDeviceToHost,
                                    ctx->stream));
        PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int q = 0; q < 4; ++q) r[q] = 0.0;
        for (int b = 0; b < blocks; ++b)
            for (int q = 0; q < 4; ++q) r[q] += h[4 * b + q];
        return PTBGPU_OK;
    }
    // column sums of A into the host vector h (3n: abs col, abs row, squares)
    int col_sums(const double *A, std::vector<double> &h) {
        hipLaunchKernelGGL(k_col_sums, dim3(n), dim3(kT), 0, ctx->stream, A, n, sums);
        PTB_CHECK_LAUNCH(ctx);
        h.resize(3 * n);
        PTB_HIP(ctx, hipMemcpyAsync(h.data(), sums, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                                    ctx->stream));
        PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return PTBGPU_OK;
    }
    // gershgorin = max column abs-sum, infinity = max row abs-sum
    int abs_norms(const double *A, double *gersh, double *inf) {
        std::vector<double> h;
        PTB_TRY(col_sums(A, h));
        double g = 0.0, r = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            g = std::max(g, h[j]);
            r = std::max(r, h[n + j]);
        }
        if (gersh) *gersh = g;
        if (inf) *inf = r;
        return PTBGPU_OK;
    }
    // C = A diag(w) A^T with host weights w
    int scaled_outer(const double *A, const std::vector<double> &w, double *tmp, double *C) {
        PTB_HIP(ctx, hipMemcpyAsync(wvec, w.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                                    ctx->stream));
        hipLaunchKernelGGL(k_colscale, dim3(blocks), dim3(kT), 0, ctx->stream, A, wvec, n, tmp);
        PTB_CHECK_LAUNCH(ctx);
        PTB_TRY(gemm(tmp, false, A, true, C));
        PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));   // w is the caller's host vector
        return PTBGPU_OK;
    }
};

const char *kNaNPur = "Error: NaN is encountered during purification";

}  // namespace

extern "C" {

int ptbgpu_thresholds(int64_t ndim, double *vals) {
    if (ndim < 1 || !vals) return PTBGPU_E_ARG;
    thresholds(ndim, vals[0], vals[1], vals[2]);
    return PTBGPU_OK;
}

int ptbgpu_purify_dev(ptbgpu_ctx *ctx, const ptbgpu_pur_settings *ps, int64_t n, const double *dHp,
                      const double *dSp, double *dPp, ptbgpu_pur_report *rep) {
    if (!ctx) return PTBGPU_E_ARG;
    if (!ps || n < 1 || !dHp || !dSp || !dPp)
        return ptb::fail(ctx, PTBGPU_E_ARG, "bad arguments to purify");
    if (n > 46340)
        return ptb::fail(ctx, PTBGPU_E_ARG, "purification is dense: ndim %lld too large",
                         (long long)n);
    if (ps->type < MCWEENY || ps->type > SIGN_DIAG || ps->metric < INV || ps->metric > SQRT_)
        return ptb::fail(ctx, PTBGPU_E_MODE, "unsupported purification type %lld / metric %lld",
                         (long long)ps->type, (long long)ps->metric);
    const bool iterative = ps->metric_iterative != 0;
    if (iterative && ps->metric == SQRT_)
        return ptb::fail(ctx, PTBGPU_E_MODE, "the iterative metric supports inverse and inverse_sqrt");
    if (ps->type != MCWEENY && ps->metric == SQRT_)
        return ptb::fail(ctx, PTBGPU_E_MODE, "sign purification needs metric inverse or inverse_sqrt");
    const bool need_eig = !iterative || ps->type == SIGN_DIAG;
    if (need_eig && n > kDenseMax)
        return ptb::fail(ctx, PTBGPU_E_ARG,
                         "eigen-route purification needs ndim <= %d (got %lld); use the iterative "
                         "metric with McWeeny, Pade or Newton-Schulz",
                         kDenseMax, (long long)n);
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    ptb::ProfScope prof_(ctx, "purify");
    double thigh, tnorm, tlow;
    thresholds(n, thigh, tnorm, tlow);
    const size_t mb = sizeof(double) * (size_t)n * (size_t)n;
    const char *names[] = {"pur_H", "pur_S", "pur_X", "pur_G", "pur_P", "pur_T1", "pur_T2", "pur_T3"};
    double *m[8];
    for (int i = 0; i < 8; ++i) {
        m[i] = static_cast<double *>(ptb::ws(ctx, names[i], mb));
        if (!m[i]) return PTBGPU_E_NOMEM;
    }
    double *H = m[0], *S = m[1], *X = m[2], *G = m[3], *P = m[4], *T1 = m[5], *T2 = m[6],
           *T3 = m[7];
    Dense D{ctx, n, 0, nullptr, nullptr, nullptr, nullptr, {}};
    D.blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n * n + kT - 1) / kT, 1024));
    D.part = static_cast<double *>(ptb::ws(ctx, "pur_part", sizeof(double) * 4 * D.blocks));
    D.sums = static_cast<double *>(ptb::ws(ctx, "pur_sums", sizeof(double) * 3 * n));
    D.wvec = static_cast<double *>(ptb::ws(ctx, "pur_w", sizeof(double) * n));
    D.dg = static_cast<GemmDesc *>(ptb::ws(ctx, "pur_gemm", sizeof(GemmDesc) * kMaxGemm));
    if (!D.part || !D.sums || !D.wvec || !D.dg) return PTBGPU_E_NOMEM;
    D.table.reserve(kMaxGemm);
    const int ub = (int)std::min<int64_t>(n, 2048);
    hipLaunchKernelGGL(k_unpack, dim3(ub), dim3(kT), 0, ctx->stream, dHp, n, H);
    hipLaunchKernelGGL(k_unpack, dim3(ub), dim3(kT), 0, ctx->stream, dSp, n, S);
    PTB_CHECK_LAUNCH(ctx);
    std::vector<double> w, h;
    double r[4];
    int64_t metric_cycles = 0;

    // ---- get_transformation_matrix (source/purification.f90:83-229)
    if (iterative && ps->metric == INV) {
        double inf;
        PTB_TRY(D.abs_norms(S, nullptr, &inf));
        PTB_TRY(D.axpby(1.0 / (inf * inf), S, 0.0, nullptr, 0.0, X));    // :137
        for (int64_t i = 1; i <= ps->metric_cycles; ++i) {
            metric_cycles = i;
            PTB_TRY(D.gemm(S, false, X, false, T1));                     // atmp = S X
            PTB_TRY(D.gemm(X, false, T1, false, T2));                    // XsX = X atmp
            PTB_TRY(D.axpby(2.0, X, -1.0, T2, 0.0, X));                  // X = 2X - XsX
            PTB_TRY(D.gemm(X, false, S, false, T3));                     // check = X S
            PTB_TRY(D.reduce(T3, nullptr, r));
            if (std::sqrt(r[1]) < tlow) break;
        }
    } else if (iterative) {
        PTB_TRY(D.reduce(S, nullptr, r));
        PTB_TRY(D.axpby(1.0 / r[0], S, 0.0, nullptr, 0.0, X));          // S / |S|_F^2, :158
        for (int64_t i = 1; i <= ps->metric_cycles; ++i) {
            metric_cycles = i;
            PTB_TRY(D.gemm(X, false, X, false, T1));                     // atmp = X X
            PTB_TRY(D.gemm(T1, false, S, false, T2));                    // XsX = atmp S
            PTB_TRY(D.gemm(X, false, T2, false, T3));                    // atmp2 = X XsX
            PTB_TRY(D.axpby(1.5, X, -0.5, T3, 0.0, X));                  // 0.5 (3X - atmp2)
            PTB_TRY(D.reduce(T2, nullptr, r));
            if (std::sqrt(r[1]) < tlow) break;
            if (r[3] > 0.0) return ptb::fail(ctx, PTBGPU_E_NAN, "Error, NaN encountered");
        }
    } else {
        PTB_TRY(ptb::dense_eig(ctx, S, n, T1, w, "pur_e"));               // dsyevd, :191
        std::vector<double> f(n, 0.0);   // 'sqrt' selects no branch: D stays 0
        for (int64_t i = 0; i < n; ++i) {
            if (ps->metric == INV) f[i] = 1.0 / w[i];
            else if (ps->metric == INV_SQRT) f[i] = 1.0 / std::sqrt(w[i]);
        }
        PTB_TRY(D.scaled_outer(T1, f, T2, X));                           // X = V f(w) V^T
    }

    // ---- purification_search (source/purification.f90:232-353)
    const double nel = (double)(float)ps->nel;   // real(nel), default REAL(4)
    double chempot = ps->chempot_guess, incr = ps->chempot_increment, nel_calc = 0.0;
    double lower = 0.0, upper = 0.0;
    bool is_lower = false, is_upper = false, bisection = false;
    int64_t cyc = 0;
    bool have_p = false;
    for (int64_t i = 1; i <= ps->chempot_cycles; ++i) {
        cyc = i;
        if (bisection) {
            chempot = (lower + upper) / 2.0;
            if (std::fabs(upper - lower) < thigh) break;
        }
        // build_guess (:355-436): G
        double ger;
        if (ps->type == MCWEENY) {
            PTB_TRY(D.gemm(H, false, X, false, T1));                     // H X
            PTB_TRY(D.gemm(X, false, T1, false, T2));                    // X H X
            PTB_TRY(D.axpby(chempot, X, -1.0, T2, 0.0, T3));             // mu X - X H X
            PTB_TRY(D.reduce(T3, nullptr, r));
            PTB_TRY(D.abs_norms(T3, &ger, nullptr));
            PTB_TRY(D.axpby(0.5 / std::min(std::sqrt(r[0]), ger), T3, 0.5, X, 0.0, G));
        } else {
            if (ps->metric == INV_SQRT) {
                PTB_TRY(D.gemm(H, false, X, false, T1));
                PTB_TRY(D.gemm(X, false, T1, false, T2));                // X H X
            } else {
                PTB_TRY(D.gemm(X, false, H, false, T2));                 // X H
            }
            PTB_TRY(D.axpby(1.0, T2, 0.0, nullptr, -chempot, T3));       // - mu I
            PTB_TRY(D.reduce(T3, nullptr, r));
            PTB_TRY(D.abs_norms(T3, &ger, nullptr));
            PTB_TRY(D.axpby(1.0 / std::min(std::sqrt(r[0]), ger), T3, 0.0, nullptr, 0.0, G));
        }
        if (ps->type == MCWEENY) {
            // mcweeny_ (:438-510): P <- 3PSP - 2PSPSP until |norm2| settles, then 2P
            PTB_TRY(D.reduce(G, nullptr, r));
            double norm = std::sqrt(r[0]);
            PTB_TRY(D.copy(G, P));
            for (int64_t it = 1; it <= ps->cycles; ++it) {
                PTB_TRY(D.gemm(S, false, G, false, T1));                 // S P
                PTB_TRY(D.gemm(G, false, T1, false, T2));                // P S P
                PTB_TRY(D.gemm(T2, false, T1, false, T3));               // P S P S P
                PTB_TRY(D.axpby(3.0, T2, -2.0, T3, 0.0, P));
                PTB_TRY(D.reduce(P, nullptr, r));
                if (r[3] > 0.0) return ptb::fail(ctx, PTBGPU_E_NAN, kNaNPur);
                if (std::fabs(std::sqrt(r[0]) - norm) < tnorm) break;
                PTB_TRY(D.copy(P, G));
                norm = std::sqrt(r[0]);
            }
            PTB_TRY(D.axpby(2.0, P, 0.0, nullptr, 0.0, P));
        } else {
            double *sign_ = T3;
            if (ps->type == PADE || ps->type == NEWTON) {
                // sign_iterative (:512-601)
                PTB_TRY(D.copy(G, sign_));
                PTB_TRY(D.reduce(sign_, nullptr, r));
                for (int64_t it = 1; it <= ps->cycles; ++it) {
                    const double norm = std::sqrt(r[0]);
                    PTB_TRY(D.gemm(sign_, false, sign_, false, T1));     // X^2
                    if (ps->type == PADE) {
                        PTB_TRY(D.gemm(T1, false, T1, false, T2));       // X^4
                        PTB_TRY(D.axpby(-10.0, T1, 3.0, T2, 15.0, T1));  // 15I - 10X^2 + 3X^4
                        PTB_TRY(D.gemm(sign_, false, T1, false, T2, 0.125));
                    } else {
                        PTB_TRY(D.gemm(sign_, false, T1, false, T2));    // X^3
                        PTB_TRY(D.axpby(1.5, sign_, -0.5, T2, 0.0, T2)); // (3X - X^3) / 2
                    }
                    PTB_TRY(D.copy(T2, sign_));
                    PTB_TRY(D.reduce(sign_, nullptr, r));
                    if (r[3] > 0.0) return ptb::fail(ctx, PTBGPU_E_NAN, kNaNPur);
                    if (std::fabs(std::sqrt(r[0]) - norm) < tnorm) break;
                }
            } else {
                // sign_diagonalization_ (:650-782): the guess is diagonalised on
                // every call (the SAVEd first_time is never cleared, :671,692);
                // chempot is intent(inout) and returns adjusted (:659)
                PTB_TRY(ptb::dense_eig(ctx, G, n, T1, w, "pur_e"));       // Q, lambda
                PTB_TRY(D.col_sums(T1, h));                              // |Q(:,i)|^2 at 2n + i
                std::vector<double> sgn(n, 0.0);
                double inc = ps->chempot_increment, lo = 0.0, up = 0.0;
                bool il = false, iu = false;
                for (int64_t k = 1; k <= ps->chempot_cycles; ++k) {
                    double nc = 0.0;
                    for (int64_t q = 0; q < n; ++q) {
                        const double eg = w[q] - chempot;
                        sgn[q] = std::fabs(eg) < tnorm ? 0.0 : (eg > 0.0 ? 1.0 : -1.0);
                        nc += 1.0 - h[2 * n + q] * sgn[q];
                    }
                    if (std::fabs(nc - nel) < tnorm) break;
                    if (nc < nel) {
                        il = true;
                        lo = chempot;
                        chempot += inc;
                    } else {
                        iu = true;
                        up = chempot;
                        chempot -= inc;
                    }
                    inc *= 2.0;
                    if (il && iu) {
                        chempot = (lo + up) / 2.0;
                        if (std::fabs(up - lo) < thigh) break;
                    }
                }
                PTB_TRY(D.scaled_outer(T1, sgn, T2, sign_));             // Q sgn Q^T
            }
            // sign_to_density_matrix (:603-647), then purified * 2
            PTB_TRY(D.axpby(-1.0, sign_, 0.0, nullptr, 1.0, T1));        // I - sign
            if (ps->metric == INV_SQRT) {
                PTB_TRY(D.gemm(T1, false, X, false, T2));
                PTB_TRY(D.gemm(X, false, T2, false, P, 0.5));
            } else {
                PTB_TRY(D.gemm(T1, false, X, false, P, 0.5));
            }
            PTB_TRY(D.reduce(P, nullptr, r));
            if (r[3] > 0.0) return ptb::fail(ctx, PTBGPU_E_NAN, kNaNPur);
            PTB_TRY(D.axpby(2.0, P, 0.0, nullptr, 0.0, P));
        }
        have_p = true;
        // get_nel (source/metrics.f90:182-199): tr(P S) = sum P .* S (S symmetric)
        PTB_TRY(D.reduce(P, S, r));
        nel_calc = r[2];
        if (std::fabs(nel_calc - nel) < tnorm) break;
        if (nel_calc < (double)ps->nel) {
            is_lower = true;
            lower = chempot;
            chempot += incr;
        } else {
            is_upper = true;
            upper = chempot;
            chempot -= incr;
        }
        bisection = is_lower && is_upper;
        incr *= 2.0;
    }
    if (!have_p) {
        // the search ran no purification (chempot_cycles < 1): the reference
        // packs an undefined matrix; return zeros
        PTB_HIP(ctx, hipMemsetAsync(P, 0, mb, ctx->stream));
    }
    const int pb = (int)std::min<int64_t>(n, 2048);
    hipLaunchKernelGGL(k_pack, dim3(pb), dim3(kT), 0, ctx->stream, P, n, dPp);
    PTB_CHECK_LAUNCH(ctx);
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (rep) {
        rep->chempot = chempot;
        rep->nel = nel_calc;
        rep->cycles = cyc;
        rep->metric_cycles = metric_cycles;
    }
    return PTBGPU_OK;
}

int ptbgpu_purify(ptbgpu_ctx *ctx, const ptbgpu_pur_settings *ps, int64_t n, const double *H,
                  const double *S, double *P, ptbgpu_pur_report *rep) {
    if (!ctx) return PTBGPU_E_ARG;
    if (!ps || n < 1 || !H || !S || !P) return ptb::fail(ctx, PTBGPU_E_ARG, "bad arguments to purify");
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    const size_t bytes = sizeof(double) * (size_t)(n * (n + 1) / 2);
    auto *dH = static_cast<double *>(ptb::ws(ctx, "io_h", bytes));
    auto *dS = static_cast<double *>(ptb::ws(ctx, "io_s", bytes));
    auto *dP = static_cast<double *>(ptb::ws(ctx, "io_p", bytes));
    if (!dH || !dS || !dP) return PTBGPU_E_NOMEM;
    PTB_TRY(ptb::h2d(ctx, dH, H, bytes));
    PTB_TRY(ptb::h2d(ctx, dS, S, bytes));
    PTB_TRY(ptbgpu_purify_dev(ctx, ps, n, dH, dS, dP, rep));
    PTB_TRY(ptb::d2h(ctx, P, dP, bytes));
    return PTBGPU_OK;
}

}  // extern "C"
