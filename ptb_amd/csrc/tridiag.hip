// Householder reduction of each K = L^-1 H L^-T to tridiagonal form, K = Q T Q^T,
// Q = H_0 ... H_{m-3},  H_k = I - tau_k v_k v_k^T  (LAPACK dsytrd, lower).
//
// One 512-thread workgroup per submatrix, blocked by 32 columns (dlatrd scheme):
// inside a panel the trailing matrix stays as it was at the panel start and the
// pending rank-2 terms are applied on the fly as V (W^T v) + W (V^T v); at the
// end of the panel one rank-64 update refreshes the trailing matrix.  The
// per-column product p = A22 v is a GEMV over the full (both triangles)
// trailing matrix: threads own rows, so a wavefront reads 64 consecutive
// elements of one column.  This GEMV is the bandwidth-bound core of the solve:
// sum_k (m-k)^2 * 8 bytes per submatrix.
//
// Output: d (diagonal), e (subdiagonal, e[k] couples k and k+1), tau, and v_k
// stored in column k rows k+1.. of A (unit at row k+1).  W is a work panel
// (m x 32, leading dimension ld) in a separate arena.
#include "sm_internal.hpp"

namespace {

constexpr int kT = 512;
constexpr int kNB = 32;
constexpr int kWaves = kT / PTB_WAVE;

__device__ __forceinline__ double block_reduce(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += red[w];
    return t;
}

__global__ __launch_bounds__(kT) void k_sytrd(const SubDesc *__restrict__ desc,
                                              double *__restrict__ Aa, double *__restrict__ Wa,
                                              double *__restrict__ d_all,
                                              double *__restrict__ e_all,
                                              double *__restrict__ tau_all) {
    extern __shared__ double sm[];
    const SubDesc ds = desc[blockIdx.x];
    const int m = ds.m, ld = ds.ld;
    double *a = Aa + ds.off;
    double *W = Wa + ds.off;  // columns 0..31 used
    double *dd = d_all + ds.mapoff, *ee = e_all + ds.mapoff, *tt = tau_all + ds.mapoff;
    double *vs = sm;           // m
    double *red = vs + m;      // kWaves
    double *c1 = red + kWaves; // kNB  W[k][t]
    double *c2 = c1 + kNB;     // kNB  V[k][k0+t]
    double *q1 = c2 + kNB;     // kNB  W[:,t]^T v
    double *q2 = q1 + kNB;     // kNB  V[:,k0+t]^T v
    __shared__ double sc[4];
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;

    if (m == 1) {
        if (threadIdx.x == 0) dd[0] = a[0];
        return;
    }
    for (int k0 = 0; k0 < m - 2; k0 += kNB) {
        const int kb = min(kNB, m - 2 - k0);
        for (int jj = 0; jj < kb; ++jj) {
            const int k = k0 + jj;
            // (1) pending-update coefficients of row k
            if ((int)threadIdx.x < jj) {
                c1[threadIdx.x] = W[(int64_t)threadIdx.x * ld + k];
                c2[threadIdx.x] = a[(int64_t)(k0 + threadIdx.x) * ld + k];
            }
            __syncthreads();
            for (int i = k + 1 + threadIdx.x; i < m; i += kT) {
                double x = a[(int64_t)k * ld + i];
                for (int t = 0; t < jj; ++t)
                    x -= a[(int64_t)(k0 + t) * ld + i] * c1[t] + W[(int64_t)t * ld + i] * c2[t];
                vs[i] = x;
            }
            if (threadIdx.x == 0) {
                double dkk = a[(int64_t)k * ld + k];
                for (int t = 0; t < jj; ++t) dkk -= 2.0 * c2[t] * c1[t];
                dd[k] = dkk;
            }
            __syncthreads();
            // (2) reflector
            double part = 0.0;
            for (int i = k + 2 + threadIdx.x; i < m; i += kT) part += vs[i] * vs[i];
            const double sigma = block_reduce(part, red);
            const double alpha = vs[k + 1];
            double beta, tk, scale;
            if (sigma == 0.0) {
                beta = alpha;
                tk = 0.0;
                scale = 0.0;
            } else {
                beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
                tk = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            __syncthreads();
            for (int i = k + 1 + threadIdx.x; i < m; i += kT) {
                const double v = (i == k + 1) ? 1.0 : vs[i] * scale;
                vs[i] = v;
                a[(int64_t)k * ld + i] = v;
            }
            if (threadIdx.x == 0) {
                ee[k] = beta;
                tt[k] = tk;
            }
            __syncthreads();
            // pending dots q1[t] = W[:,t]^T v, q2[t] = V[:,k0+t]^T v over rows k+1..
            for (int t = wid; t < jj; t += kWaves) {
                double s1 = 0.0, s2 = 0.0;
                for (int i = k + 1 + lane; i < m; i += PTB_WAVE) {
                    const double v = vs[i];
                    s1 += W[(int64_t)t * ld + i] * v;
                    s2 += a[(int64_t)(k0 + t) * ld + i] * v;
                }
                s1 = wave_sum(s1);
                s2 = wave_sum(s2);
                if (lane == 0) {
                    q1[t] = s1;
                    q2[t] = s2;
                }
            }
            __syncthreads();
            // (3) p = A22 v - V q1 - W q2 ; (4) w = tau p - 0.5 tau^2 (p^T v) v
            double pv = 0.0;
            for (int i = k + 1 + threadIdx.x; i < m; i += kT) {
                double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
                int c = k + 1;
                for (; c + 3 < m; c += 4) {
                    acc0 += a[(int64_t)c * ld + i] * vs[c];
                    acc1 += a[(int64_t)(c + 1) * ld + i] * vs[c + 1];
                    acc2 += a[(int64_t)(c + 2) * ld + i] * vs[c + 2];
                    acc3 += a[(int64_t)(c + 3) * ld + i] * vs[c + 3];
                }
                for (; c < m; ++c) acc0 += a[(int64_t)c * ld + i] * vs[c];
                double p = (acc0 + acc1) + (acc2 + acc3);
                for (int t = 0; t < jj; ++t)
                    p -= a[(int64_t)(k0 + t) * ld + i] * q1[t] + W[(int64_t)t * ld + i] * q2[t];
                p *= tk;
                W[(int64_t)jj * ld + i] = p;  // provisional, finished below
                pv += p * vs[i];
            }
            const double ptv = block_reduce(pv, red);
            const double acoef = -0.5 * tk * ptv;
            for (int i = k + 1 + threadIdx.x; i < m; i += kT)
                W[(int64_t)jj * ld + i] += acoef * vs[i];
            for (int i = threadIdx.x; i <= k; i += kT) W[(int64_t)jj * ld + i] = 0.0;
            __syncthreads();
        }
        // trailing update A[s:, s:] -= Vb Wb^T + Wb Vb^T over both triangles
        const int s = k0 + kb;
        const int rem = m - s;
        if (rem > 0) {
            double *Vi = sm;                 // 32 x kNB each, reusing the vector space
            double *Wi = Vi + 32 * kNB;
            double *Vj = Wi + 32 * kNB;
            double *Wj = Vj + 32 * kNB;
            const int nt = (rem + 31) / 32;
            for (int tile = 0; tile < nt * nt; ++tile) {
                const int i0 = s + (tile % nt) * 32, j0 = s + (tile / nt) * 32;
                __syncthreads();
                for (int l = threadIdx.x; l < 32 * kNB; l += kT) {
                    const int r = l & 31, t = l >> 5;
                    const bool okt = t < kb;
                    Vi[r * kNB + t] = (okt && i0 + r < m) ? a[(int64_t)(k0 + t) * ld + i0 + r] : 0.0;
                    Wi[r * kNB + t] = (okt && i0 + r < m) ? W[(int64_t)t * ld + i0 + r] : 0.0;
                    Vj[r * kNB + t] = (okt && j0 + r < m) ? a[(int64_t)(k0 + t) * ld + j0 + r] : 0.0;
                    Wj[r * kNB + t] = (okt && j0 + r < m) ? W[(int64_t)t * ld + j0 + r] : 0.0;
                }
                __syncthreads();
                for (int l = threadIdx.x; l < 1024; l += kT) {
                    const int r = l & 31, c = l >> 5;
                    if (i0 + r < m && j0 + c < m) {
                        double sacc = 0.0;
#pragma unroll 8
                        for (int t = 0; t < kNB; ++t)
                            sacc += Vi[r * kNB + t] * Wj[c * kNB + t] + Wi[r * kNB + t] * Vj[c * kNB + t];
                        a[(int64_t)(j0 + c) * ld + i0 + r] -= sacc;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        dd[m - 2] = a[(int64_t)(m - 2) * ld + m - 2];
        dd[m - 1] = a[(int64_t)(m - 1) * ld + m - 1];
        ee[m - 2] = a[(int64_t)(m - 2) * ld + m - 1];
        tt[m - 2] = 0.0;
        tt[m - 1] = 0.0;
        ee[m - 1] = 0.0;
    }
    (void)sc;
}

}  // namespace

namespace ptb {

int sytrd_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, double *A, double *W, double *d,
                double *e, double *tau, int max_m) {
    const size_t vec = (size_t)max_m + kWaves + 4 * kNB;
    const size_t tiles = 4 * 32 * kNB;
    const size_t shmem = sizeof(double) * std::max(vec, tiles);
    if (shmem > 160 * 1024) return fail(ctx, PTBGPU_E_ARG, "submatrix too large for sytrd (%d)", max_m);
    hipLaunchKernelGGL(k_sytrd, dim3(nsm), dim3(kT), shmem, ctx->stream, ddesc, A, W, d, e, tau);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

}  // namespace ptb
