// Batched dense FP64 kernels over the per-submatrix arenas.
//
// * k_gemm:     C = alpha op(A) op(B) + beta C for a variable-size batch whose
//               shapes and pointers come from device descriptors (so a kernel
//               earlier in the stream can size the next product).  64x64 output
//               tile per workgroup, 16-deep k slices staged in LDS, 4x4
//               accumulators per thread.
// * k_potrf:    lower Cholesky S = L L^T in place, blocked by 32: the diagonal
//               block is factored in LDS, the panel by forward substitution in
//               registers, the trailing lower triangle by 32x32 LDS tiles.  A
//               non-positive pivot sets flags bit 0 (dsygvd would report INFO>0).
// * k_trsm:     X = L^-1 B or L^-T B in place, one workgroup per 32 right-hand
//               sides of one submatrix.
// * k_transpose, k_zero_upper, k_larft: helpers for K = L^-1 H L^-T and for the
//               blocked (WY) application of the Householder reflectors.
#include "sm_internal.hpp"

namespace {

constexpr int kT = 256;

template <bool TA, bool TB>
__global__ __launch_bounds__(kT) void k_gemm(const GemmDesc *__restrict__ descs) {
    const GemmDesc g = descs[blockIdx.y];
    const int tilesM = (g.M + 63) / 64, tilesN = (g.N + 63) / 64;
    if (g.M <= 0 || g.N <= 0 || (int)blockIdx.x >= tilesM * tilesN) return;
    const int m0 = (blockIdx.x % tilesM) * 64, n0 = (blockIdx.x / tilesM) * 64;
    __shared__ double As[16][65];
    __shared__ double Bs[16][65];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < g.K; k0 += 16) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int l = threadIdx.x + q * kT;
            int mm, kk;
            if (!TA) { mm = l & 63; kk = l >> 6; } else { kk = l & 15; mm = l >> 4; }
            const int gm = m0 + mm, gk = k0 + kk;
            double v = 0.0;
            if (gm < g.M && gk < g.K)
                v = TA ? g.A[(int64_t)gm * g.lda + gk] : g.A[(int64_t)gk * g.lda + gm];
            As[kk][mm] = v;
            int nn;
            if (!TB) { kk = l & 15; nn = l >> 4; } else { nn = l & 63; kk = l >> 6; }
            const int gn = n0 + nn, gk2 = k0 + kk;
            double w = 0.0;
            if (gn < g.N && gk2 < g.K)
                w = TB ? g.B[(int64_t)gk2 * g.ldb + gn] : g.B[(int64_t)gn * g.ldb + gk2];
            Bs[kk][nn] = w;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int gn = n0 + ty + 16 * j;
        if (gn >= g.N) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int gm = m0 + tx + 16 * i;
            if (gm >= g.M) continue;
            double *c = g.C + (int64_t)gn * g.ldc + gm;
            *c = g.beta == 0.0 ? g.alpha * acc[i][j] : g.alpha * acc[i][j] + g.beta * *c;
        }
    }
}

__global__ __launch_bounds__(kT) void k_potrf(const SubDesc *__restrict__ desc,
                                              double *__restrict__ Aa, int *__restrict__ flags) {
    const SubDesc d = desc[blockIdx.x];
    double *a = Aa + d.off;
    const int m = d.m, ld = d.ld;
    __shared__ double Lk[32][33];
    __shared__ double Pi[32][33], Pj[32][33];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    for (int k0 = 0; k0 < m; k0 += 32) {
        const int kb = min(32, m - k0);
        __syncthreads();
        for (int l = threadIdx.x; l < 1024; l += kT) {
            const int r = l & 31, c = l >> 5;
            double v;
            if (r < kb && c < kb) v = a[(int64_t)(k0 + c) * ld + k0 + r];
            else v = (r == c) ? 1.0 : 0.0;  // pad with the identity
            Lk[r][c] = v;
        }
        __syncthreads();
        for (int j = 0; j < kb; ++j) {
            if (threadIdx.x == 0) {
                double p = Lk[j][j];
                if (!(p > 0.0)) {
                    bad = 1;
                    p = 1.0;
                }
                Lk[j][j] = sqrt(p);
            }
            __syncthreads();
            if (threadIdx.x > j && threadIdx.x < kb) Lk[threadIdx.x][j] /= Lk[j][j];
            __syncthreads();
            for (int l = threadIdx.x; l < 1024; l += kT) {
                const int r = l & 31, c = l >> 5;
                if (c > j && r >= c && r < kb) Lk[r][c] -= Lk[r][j] * Lk[c][j];
            }
            __syncthreads();
        }
        for (int l = threadIdx.x; l < 1024; l += kT) {
            const int r = l & 31, c = l >> 5;
            if (r < kb && c < kb && r >= c) a[(int64_t)(k0 + c) * ld + k0 + r] = Lk[r][c];
        }
        const int k1 = k0 + kb;
        // panel rows below: x <- x L_kk^-T  (forward substitution, padded to 32)
        for (int i = k1 + threadIdx.x; i < m; i += kT) {
            double x[32];
#pragma unroll
            for (int c = 0; c < 32; ++c) x[c] = c < kb ? a[(int64_t)(k0 + c) * ld + i] : 0.0;
#pragma unroll
            for (int c = 0; c < 32; ++c) {
                double s = x[c];
#pragma unroll
                for (int t = 0; t < c; ++t) s -= x[t] * Lk[c][t];
                x[c] = s / Lk[c][c];
            }
#pragma unroll
            for (int c = 0; c < 32; ++c)
                if (c < kb) a[(int64_t)(k0 + c) * ld + i] = x[c];
        }
        __syncthreads();
        // trailing lower triangle: A_ij -= sum_t P_it P_jt over 32x32 tiles
        const int rem = m - k1;
        if (rem <= 0) continue;
        const int nt = (rem + 31) / 32;
        const int ntile = nt * (nt + 1) / 2;
        for (int t = 0; t < ntile; ++t) {
            int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
            while (ti * (ti + 1) / 2 > t) --ti;
            while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
            const int tj = t - ti * (ti + 1) / 2;  // tj <= ti
            const int i0 = k1 + ti * 32, j0 = k1 + tj * 32;
            for (int l = threadIdx.x; l < 1024; l += kT) {
                const int r = l & 31, c = l >> 5;
                Pi[r][c] = (i0 + r < m && c < kb) ? a[(int64_t)(k0 + c) * ld + i0 + r] : 0.0;
                Pj[r][c] = (j0 + r < m && c < kb) ? a[(int64_t)(k0 + c) * ld + j0 + r] : 0.0;
            }
            __syncthreads();
            for (int l = threadIdx.x; l < 1024; l += kT) {
                const int r = l & 31, c = l >> 5;
                const int gi = i0 + r, gj = j0 + c;
                if (gi < m && gj < m && gi >= gj) {
                    double s = 0.0;
#pragma unroll 8
                    for (int q = 0; q < 32; ++q) s += Pi[r][q] * Pj[c][q];
                    a[(int64_t)gj * ld + gi] -= s;
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && bad) flags[blockIdx.x] |= 1;
}

// In place B <- L^-1 B (TRANS=false) or L^-T B (TRANS=true), L lower (only the
// lower triangle is read).  grid.x: blocks of 32 right-hand sides, grid.y: submatrix.
template <bool TRANS>
__global__ __launch_bounds__(kT) void k_trsm(const SubDesc *__restrict__ desc,
                                             const double *__restrict__ La,
                                             double *__restrict__ Ba) {
    const SubDesc d = desc[blockIdx.y];
    const int rb = blockIdx.x * 32;
    if (rb >= d.m) return;
    const double *L = La + d.off;
    double *B = Ba + d.off;
    const int m = d.m, ld = d.ld;
    const int nbk = (m + 31) / 32;
    __shared__ double Lt[32][33], Xt[32][33], Bt[32][33];
    for (int step = 0; step < nbk; ++step) {
        const int ib = TRANS ? nbk - 1 - step : step;
        const int i0 = ib * 32;
        double acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int l = threadIdx.x + q * kT, r = l & 31, c = l >> 5;
            acc[q] = (i0 + r < m && rb + c < m) ? B[(int64_t)(rb + c) * ld + i0 + r] : 0.0;
        }
        const int jlo = TRANS ? ib + 1 : 0, jhi = TRANS ? nbk : ib;
        for (int jb = jlo; jb < jhi; ++jb) {
            const int j0 = jb * 32;
            __syncthreads();
            for (int l = threadIdx.x; l < 1024; l += kT) {
                const int r = l & 31, c = l >> 5;
                double lv = 0.0;
                if (!TRANS) {  // L(i0+r, j0+c)
                    if (i0 + r < m && j0 + c < m) lv = L[(int64_t)(j0 + c) * ld + i0 + r];
                } else {       // L^T(i0+r, j0+c) = L(j0+c, i0+r)
                    if (i0 + r < m && j0 + c < m) lv = L[(int64_t)(i0 + r) * ld + j0 + c];
                }
                Lt[r][c] = lv;
                Xt[r][c] = (j0 + r < m && rb + c < m) ? B[(int64_t)(rb + c) * ld + j0 + r] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int l = threadIdx.x + q * kT, r = l & 31, c = l >> 5;
                double s = 0.0;
#pragma unroll 8
                for (int t = 0; t < 32; ++t) s += Lt[r][t] * Xt[t][c];
                acc[q] -= s;
            }
        }
        __syncthreads();
        // diagonal block of L (identity padding) and the right-hand sides
        for (int l = threadIdx.x; l < 1024; l += kT) {
            const int r = l & 31, c = l >> 5;
            double lv;
            if (i0 + r < m && i0 + c < m) {
                lv = !TRANS ? (r >= c ? L[(int64_t)(i0 + c) * ld + i0 + r] : 0.0)
                            : (c >= r ? L[(int64_t)(i0 + r) * ld + i0 + c] : 0.0);
            } else {
                lv = (r == c) ? 1.0 : 0.0;
            }
            Lt[r][c] = lv;  // Lt = diag block of op(L)
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int l = threadIdx.x + q * kT, r = l & 31, c = l >> 5;
            Bt[r][c] = acc[q];
        }
        __syncthreads();
        if (threadIdx.x < 32) {
            const int c = threadIdx.x;
            if (!TRANS) {
                for (int r = 0; r < 32; ++r) {
                    double s = Bt[r][c];
                    for (int t = 0; t < r; ++t) s -= Lt[r][t] * Bt[t][c];
                    Bt[r][c] = s / Lt[r][r];
                }
            } else {
                for (int r = 31; r >= 0; --r) {
                    double s = Bt[r][c];
                    for (int t = r + 1; t < 32; ++t) s -= Lt[r][t] * Bt[t][c];
                    Bt[r][c] = s / Lt[r][r];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int l = threadIdx.x + q * kT, r = l & 31, c = l >> 5;
            if (i0 + r < m && rb + c < m) B[(int64_t)(rb + c) * ld + i0 + r] = Bt[r][c];
        }
        __syncthreads();
    }
}

// out = in^T (or 0.5 (in + in^T)); grid.x: 32x32 tiles, grid.y: submatrix
__global__ __launch_bounds__(kT) void k_transpose(const SubDesc *__restrict__ desc,
                                                  const double *__restrict__ ina,
                                                  double *__restrict__ outa, int symmetrize) {
    const SubDesc d = desc[blockIdx.y];
    const int nt = (d.m + 31) / 32;
    if ((int)blockIdx.x >= nt * nt) return;
    const int r0 = (blockIdx.x % nt) * 32, c0 = (blockIdx.x / nt) * 32;
    const double *in = ina + d.off;
    double *out = outa + d.off;
    __shared__ double t[32][33];
    for (int l = threadIdx.x; l < 1024; l += kT) {
        const int r = l & 31, c = l >> 5;
        if (r0 + r < d.m && c0 + c < d.m) t[r][c] = in[(int64_t)(c0 + c) * d.ld + r0 + r];
    }
    __syncthreads();
    // out(c0+r, r0+c) = in(r0+c, c0+r)
    for (int l = threadIdx.x; l < 1024; l += kT) {
        const int r = l & 31, c = l >> 5;
        if (c0 + r < d.m && r0 + c < d.m) {
            double v = t[c][r];
            if (symmetrize) v = 0.5 * (v + in[(int64_t)(r0 + c) * d.ld + c0 + r]);
            out[(int64_t)(r0 + c) * d.ld + c0 + r] = v;
        }
    }
}

__global__ __launch_bounds__(kT) void k_zero_upper(const SubDesc *__restrict__ desc,
                                                   double *__restrict__ Aa) {
    const SubDesc d = desc[blockIdx.y];
    const int c = blockIdx.x;
    if (c >= d.m) return;
    double *col = Aa + d.off + (int64_t)c * d.ld;
    for (int r = threadIdx.x; r <= c; r += kT) col[r] = 0.0;
}

// T (nb x nb, upper) with H_b0 ... H_{b0+nb-1} = I - V T V^T for the reflectors
// in columns b0.. of V (zero above row k+1, unit at row k+1).  One workgroup per
// submatrix; submatrices without a block b0 skip.  T stored nb x nb per submatrix
// at T + s*nb*nb (column-major).
__global__ __launch_bounds__(kT) void k_larft(const SubDesc *__restrict__ desc,
                                              const double *__restrict__ Va,
                                              const double *__restrict__ tau_all, int q, int nb,
                                              double *__restrict__ Ta) {
    const SubDesc d = desc[blockIdx.x];
    const int nr = d.m - 2;  // number of reflectors
    double *T = Ta + (int64_t)blockIdx.x * nb * nb;
    // q-th block counted from the last one
    const int nblk = nr > 0 ? (nr + nb - 1) / nb : 0;
    const int b = nblk - 1 - q;
    if (b < 0) return;
    const int b0 = b * nb;
    const int kb = min(nb, nr - b0);
    const double *V = Va + d.off + (int64_t)b0 * d.ld;
    const double *tau = tau_all + d.mapoff + b0;
    __shared__ double G[32][33];   // G(i,j) = v_i^T v_j, j < i
    __shared__ double Ts[32][33];
    for (int l = threadIdx.x; l < 1024; l += kT) {
        const int i = l & 31, j = l >> 5;
        double s = 0.0;
        if (j < i && i < kb) {
            const double *vi = V + (int64_t)i * d.ld, *vj = V + (int64_t)j * d.ld;
            for (int r = b0 + i + 1; r < d.m; ++r) s += vj[r] * vi[r];
        }
        G[i][j] = s;
        Ts[i][j] = 0.0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < kb; ++i) {
            Ts[i][i] = tau[i];
            for (int r = 0; r < i; ++r) {
                double s = 0.0;
                for (int c = r; c < i; ++c) s += Ts[r][c] * G[i][c];
                Ts[r][i] = -tau[i] * s;
            }
        }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < nb * nb; l += kT) {
        const int r = l % nb, c = l / nb;
        T[(int64_t)c * nb + r] = (r < 32 && c < 32) ? Ts[r][c] : 0.0;
    }
}

}  // namespace

namespace ptb {

int gemm_vbatch(ptbgpu_ctx *ctx, bool ta, bool tb, const GemmDesc *ddesc, int count, int max_m,
                int max_n) {
    if (count <= 0 || max_m <= 0 || max_n <= 0) return PTBGPU_OK;
    const int tiles = ((max_m + 63) / 64) * ((max_n + 63) / 64);
    for (int c0 = 0; c0 < count; c0 += 65535) {
        const int cnt = std::min(65535, count - c0);
        dim3 grid(tiles, cnt);
        const GemmDesc *dd = ddesc + c0;
        if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(kT), 0, ctx->stream, dd);
        else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(kT), 0, ctx->stream, dd);
        else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(kT), 0, ctx->stream, dd);
        else hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(kT), 0, ctx->stream, dd);
        PTB_CHECK_LAUNCH(ctx);
    }
    return PTBGPU_OK;
}

int potrf_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, double *A, int *flags) {
    hipLaunchKernelGGL(k_potrf, dim3(nsm), dim3(kT), 0, ctx->stream, ddesc, A, flags);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int trsm_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const double *L,
               double *B, bool trans) {
    dim3 grid((max_m + 31) / 32, nsm);
    if (trans) hipLaunchKernelGGL(k_trsm<true>, grid, dim3(kT), 0, ctx->stream, ddesc, L, B);
    else hipLaunchKernelGGL(k_trsm<false>, grid, dim3(kT), 0, ctx->stream, ddesc, L, B);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int transpose_sym_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m,
                        const double *in, double *out, bool symmetrize) {
    const int nt = (max_m + 31) / 32;
    hipLaunchKernelGGL(k_transpose, dim3(nt * nt, nsm), dim3(kT), 0, ctx->stream, ddesc, in, out,
                       symmetrize ? 1 : 0);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int zero_upper_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, double *A) {
    hipLaunchKernelGGL(k_zero_upper, dim3(max_m, nsm), dim3(kT), 0, ctx->stream, ddesc, A);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int larft_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, const double *V,
                const double *tau, int q, int nb, double *T) {
    hipLaunchKernelGGL(k_larft, dim3(nsm), dim3(kT), 0, ctx->stream, ddesc, V, tau, q, nb, T);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

}  // namespace ptb
