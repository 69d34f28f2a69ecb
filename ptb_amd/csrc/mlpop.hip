// Mixed Mulliken/Loewdin populations from the density, the first consumer of
// P (mlpop2, source/mlpop.f:198-238), on the device so that P need not leave
// the GPU.
//
// The reference forms, in REAL(4):
//   x  = P as a full symmetric matrix        (blowsym84: doubles rounded to float)
//   cc = x S^3/4                             (ssymm)
//   x  = S^1/4 cc                            (sgemm)
// and sums diag(x) into atomic populations q(aoat(i)) and shell populations
// qsh(shell2ao(i), aoat(i)) in AO order.  Only the diagonal of the second
// product is read, so here:
//   * x(i,j) = float(P(min(i,j), max(i,j))) as a dense FP64 matrix (the same
//     float-rounded values the reference multiplies);
//   * cc = x S^3/4 on the FP64 matrix cores (the float inputs widened);
//   * d_i = sum_k S^1/4(i,k) cc(k,i), one row-column dot per AO (64x64 tiles
//     through LDS so that both operands are read along their columns);
//   * q and qsh summed on the host in the reference's AO order.
// The arithmetic after the float rounding of the inputs is FP64, so the result
// differs from the reference's by the reference's own float rounding.
#include "sm_internal.hpp"

namespace {

constexpr int kT = 256;

// x = float-rounded full symmetric matrix of packed P; grid (tiles of 32 x 32)
__global__ __launch_bounds__(kT) void k_blow84(const double *__restrict__ p, int64_t nao,
                                               double *__restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * 32 + (threadIdx.x & 31);
    const int64_t j0 = (int64_t)blockIdx.y * 32;
    if (i >= nao) return;
    for (int64_t jj = threadIdx.x >> 5; jj < 32; jj += kT / 32) {
        const int64_t j = j0 + jj;
        if (j >= nao) break;
        const double v = i <= j ? p[pk(i, j)] : p[pk(j, i)];
        x[j * nao + i] = (double)(float)v;
    }
}

__global__ __launch_bounds__(kT) void k_f32_to_f64(const float *__restrict__ a, int64_t len,
                                                   double *__restrict__ b) {
    for (int64_t k = (int64_t)blockIdx.x * kT + threadIdx.x; k < len; k += (int64_t)gridDim.x * kT)
        b[k] = (double)a[k];
}

// d_i = sum_k s2(i,k) cc(k,i): block b owns rows i0..i0+63; k runs in chunks
// of 64 staged through LDS (s2 tile read along i, cc tile along k)
__global__ __launch_bounds__(kT) void k_row_col_dot(const float *__restrict__ s2,
                                                    const double *__restrict__ cc, int64_t nao,
                                                    double *__restrict__ d) {
    __shared__ double ts[64][65];
    __shared__ double tc[64][65];
    __shared__ double part[4][64];
    const int64_t i0 = (int64_t)blockIdx.x * 64;
    const int t = threadIdx.x;
    const int li = t & 63, lq = t >> 6;
    double acc = 0.0;
    for (int64_t k0 = 0; k0 < nao; k0 += 64) {
        __syncthreads();
        for (int e = t; e < 64 * 64; e += kT) {
            const int r = e & 63, c = e >> 6;
            // ts[kk][ii] = s2(i0 + ii, k0 + kk): consecutive threads, consecutive i
            const int64_t ii = i0 + r, kk = k0 + c;
            ts[c][r] = (ii < nao && kk < nao) ? (double)s2[kk * nao + ii] : 0.0;
            // tc[ii][kk] = cc(k0 + kk, i0 + ii): consecutive threads, consecutive k
            const int64_t kr = k0 + r, ic = i0 + c;
            tc[c][r] = (kr < nao && ic < nao) ? cc[ic * nao + kr] : 0.0;
        }
        __syncthreads();
        // thread (li, lq): k in lq, lq + 4, ...
        for (int kk = lq; kk < 64; kk += 4) acc += ts[kk][li] * tc[li][kk];
    }
    part[lq][li] = acc;
    __syncthreads();
    if (t < 64 && i0 + t < nao) {
        // fixed order over the four partial sums
        d[i0 + t] = ((part[0][t] + part[1][t]) + part[2][t]) + part[3][t];
    }
}

int mlpop2_device(ptbgpu_ctx *ctx, int64_t n, int64_t nao, const double *dP, const float *ds1,
                  const float *ds2, const int64_t *aoat, const int64_t *shell2ao, double *q,
                  double *qsh) {
    ptb::ProfScope prof_(ctx, "mlpop2");
    const size_t nn = (size_t)nao * (size_t)nao;
    auto *x = static_cast<double *>(ptb::ws(ctx, "mlpop_x", sizeof(double) * nn));
    auto *s1d = static_cast<double *>(ptb::ws(ctx, "mlpop_s1", sizeof(double) * nn));
    auto *cc = static_cast<double *>(ptb::ws(ctx, "mlpop_cc", sizeof(double) * nn));
    auto *d = static_cast<double *>(ptb::ws(ctx, "mlpop_d", sizeof(double) * (size_t)nao));
    if (!x || !s1d || !cc || !d) return PTBGPU_E_NOMEM;
    const unsigned nt = (unsigned)((nao + 31) / 32);
    hipLaunchKernelGGL(k_blow84, dim3(nt, nt), dim3(kT), 0, ctx->stream, dP, nao, x);
    PTB_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)std::min<int64_t>((nn + kT - 1) / kT, 65536)),
                       dim3(kT), 0, ctx->stream, ds1, (int64_t)nn, s1d);
    PTB_CHECK_LAUNCH(ctx);
    GemmDesc g{x, s1d, cc, (int)nao, (int)nao, (int)nao, (int)nao, (int)nao, (int)nao, 1.0, 0.0, 0, 0};
    auto *dg = static_cast<GemmDesc *>(ptb::upload_table(ctx, "mlpop_g", &g, sizeof(g)));
    if (!dg) return PTBGPU_E_NOMEM;
    ptb::TileList tl;
    tl.add(&g, 1, true);
    PTB_TRY(tl.upload(ctx, "mlpop_w"));
    PTB_TRY(tl.launch(ctx, 0, false, false, dg, "mlpop"));
    hipLaunchKernelGGL(k_row_col_dot, dim3((unsigned)((nao + 63) / 64)), dim3(kT), 0, ctx->stream,
                       ds2, cc, nao, d);
    PTB_CHECK_LAUNCH(ctx);
    std::vector<double> hd((size_t)nao);
    PTB_HIP(ctx, hipMemcpyAsync(hd.data(), d, sizeof(double) * nao, hipMemcpyDeviceToHost, ctx->stream));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t a = 0; a < n; ++a) q[a] = 0.0;
    for (int64_t a = 0; a < 10 * n; ++a) qsh[a] = 0.0;
    for (int64_t i = 0; i < nao; ++i) {
        const int64_t at = aoat[i] - 1, sh = shell2ao[i] - 1;
        q[at] += hd[i];
        qsh[at * 10 + sh] += hd[i];
    }
    return PTBGPU_OK;
}

int mlpop_args(ptbgpu_ctx *ctx, int64_t n, int64_t nao, const void *P, const void *s1,
               const void *s2, const int64_t *aoat, const int64_t *shell2ao, const double *q,
               const double *qsh) {
    if (!ctx) return PTBGPU_E_ARG;
    if (n < 1 || nao < 1 || nao > 46340 || !P || !s1 || !s2 || !aoat || !shell2ao || !q || !qsh)
        return ptb::fail(ctx, PTBGPU_E_ARG, "bad mlpop2 arguments (nao <= 46340: n^2 < 2^31)");
    for (int64_t i = 0; i < nao; ++i)
        if (aoat[i] < 1 || aoat[i] > n || shell2ao[i] < 1 || shell2ao[i] > 10)
            return ptb::fail(ctx, PTBGPU_E_ARG, "AO %lld: atom %lld / shell %lld out of range",
                             (long long)(i + 1), (long long)aoat[i], (long long)shell2ao[i]);
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    return PTBGPU_OK;
}

}  // namespace

extern "C" {

int ptbgpu_mlpop2_dev(ptbgpu_ctx *ctx, int64_t n, int64_t nao, const double *dP,
                      const float *ds1, const float *ds2, const int64_t *aoat,
                      const int64_t *shell2ao, double *q, double *qsh) {
    PTB_TRY(mlpop_args(ctx, n, nao, dP, ds1, ds2, aoat, shell2ao, q, qsh));
    return mlpop2_device(ctx, n, nao, dP, ds1, ds2, aoat, shell2ao, q, qsh);
}

int ptbgpu_mlpop2(ptbgpu_ctx *ctx, int64_t n, int64_t nao, const double *P, const float *s1,
                  const float *s2, const int64_t *aoat, const int64_t *shell2ao, double *q,
                  double *qsh) {
    PTB_TRY(mlpop_args(ctx, n, nao, P, s1, s2, aoat, shell2ao, q, qsh));
    const size_t nn = (size_t)nao * (size_t)nao, npk = (size_t)(nao * (nao + 1) / 2);
    auto *dP = static_cast<double *>(ptb::ws(ctx, "mlpop_in_p", sizeof(double) * npk));
    auto *d1 = static_cast<float *>(ptb::ws(ctx, "mlpop_in_s1", sizeof(float) * nn));
    auto *d2 = static_cast<float *>(ptb::ws(ctx, "mlpop_in_s2", sizeof(float) * nn));
    if (!dP || !d1 || !d2) return PTBGPU_E_NOMEM;
    PTB_HIP(ctx, hipMemcpyAsync(dP, P, sizeof(double) * npk, hipMemcpyHostToDevice, ctx->stream));
    PTB_HIP(ctx, hipMemcpyAsync(d1, s1, sizeof(float) * nn, hipMemcpyHostToDevice, ctx->stream));
    PTB_HIP(ctx, hipMemcpyAsync(d2, s2, sizeof(float) * nn, hipMemcpyHostToDevice, ctx->stream));
    return mlpop2_device(ctx, n, nao, dP, d1, d2, aoat, shell2ao, q, qsh);
}

}  // extern "C"
