// Threshold filter and positive-entry count over a packed-upper FP64 matrix.
//
// filter_matrix_sym (source/pgtb.f90:1543-1547): A(k) = 0 where abs(A(k)) < filter.
// analyse_matrix_sym (source/pgtb.f90:1563-1578): count of A(ij) > 0 over the
// full symmetric matrix, diagonal once and off-diagonal twice.
//
// One streaming pass: each thread reads 16-byte pairs (coalesced dwordx4), counts
// positives before and after, and stores only the pairs whose bits change, so an
// already-filtered matrix costs one read.  Counts are reduced per wavefront, then
// per block, then one 64-bit atomic per block.  The diagonal count (needed to turn
// the packed count into the full-matrix count) comes from a second, ndim-long pass.
#include "common.hpp"

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ bool zap(double v, double filter) {
    // strict <, as the reference; NaN compares false and is kept
    return fabs(v) < filter;
}

__device__ __forceinline__ bool nonzero_bits(double v) {
    return __double_as_longlong(v) != 0;
}

template <bool kFilter>
__global__ __launch_bounds__(kThreads) void k_filter_pairs(double *__restrict__ a, int64_t len,
                                                           double filter,
                                                           unsigned long long *__restrict__ cnt) {
    __shared__ unsigned long long red[16];
    const int64_t npair = len / 2;
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    unsigned long long before = 0, after = 0;
    double2 *a2 = reinterpret_cast<double2 *>(a);
    for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < npair; p += stride) {
        double2 v = a2[p];
        before += (v.x > 0.0) + (v.y > 0.0);
        if (kFilter) {
            bool cx = zap(v.x, filter) && nonzero_bits(v.x);
            bool cy = zap(v.y, filter) && nonzero_bits(v.y);
            if (cx) v.x = 0.0;
            if (cy) v.y = 0.0;
            if (cx || cy) a2[p] = v;
        }
        after += (v.x > 0.0) + (v.y > 0.0);
    }
    // odd tail element
    if (blockIdx.x == 0 && threadIdx.x == 0 && (len & 1)) {
        double v = a[len - 1];
        before += v > 0.0;
        if (kFilter && zap(v, filter) && nonzero_bits(v)) {
            v = 0.0;
            a[len - 1] = v;
        }
        after += v > 0.0;
    }
    unsigned long long b = block_sum(before, red);
    unsigned long long f = block_sum(after, red);
    if (threadIdx.x == 0) {
        atomicAdd(&cnt[0], b);
        atomicAdd(&cnt[1], f);
    }
}

// scalar variant for pointers that are not 16-byte aligned
template <bool kFilter>
__global__ __launch_bounds__(kThreads) void k_filter_scalar(double *__restrict__ a, int64_t len,
                                                            double filter,
                                                            unsigned long long *__restrict__ cnt) {
    __shared__ unsigned long long red[16];
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    unsigned long long before = 0, after = 0;
    for (int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x; k < len; k += stride) {
        double v = a[k];
        before += v > 0.0;
        if (kFilter && zap(v, filter) && nonzero_bits(v)) {
            v = 0.0;
            a[k] = v;
        }
        after += v > 0.0;
    }
    unsigned long long b = block_sum(before, red);
    unsigned long long f = block_sum(after, red);
    if (threadIdx.x == 0) {
        atomicAdd(&cnt[0], b);
        atomicAdd(&cnt[1], f);
    }
}

// positive diagonal entries; slot 0 before / slot 1 after the filter rule
__global__ __launch_bounds__(kThreads) void k_diag_positive(const double *__restrict__ a,
                                                            int64_t ndim, double filter,
                                                            int apply,
                                                            unsigned long long *__restrict__ cnt) {
    __shared__ unsigned long long red[16];
    unsigned long long before = 0, after = 0;
    for (int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x; j < ndim;
         j += (int64_t)gridDim.x * kThreads) {
        double v = a[pk(j, j)];
        before += v > 0.0;
        after += (v > 0.0) && !(apply && zap(v, filter));
    }
    unsigned long long b = block_sum(before, red);
    unsigned long long f = block_sum(after, red);
    if (threadIdx.x == 0) {
        atomicAdd(&cnt[2], b);
        atomicAdd(&cnt[3], f);
    }
}

}  // namespace

namespace ptb {

int filter_count(ptbgpu_ctx *ctx, double *dA, int64_t ndim, double filter, bool do_filter,
                 int64_t counts[2]) {
    ProfScope prof_(ctx, "filter");
    const int64_t len = ndim * (ndim + 1) / 2;
    auto *cnt = static_cast<unsigned long long *>(ws(ctx, "filter_cnt", 4 * sizeof(unsigned long long)));
    if (!cnt) return PTBGPU_E_NOMEM;
    PTB_HIP(ctx, hipMemsetAsync(cnt, 0, 4 * sizeof(unsigned long long), ctx->stream));
    // the diagonal pass runs first: it reads the unfiltered diagonal
    {
        int blocks = (int)std::min<int64_t>((ndim + kThreads - 1) / kThreads, 1024);
        if (blocks < 1) blocks = 1;
        hipLaunchKernelGGL(k_diag_positive, dim3(blocks), dim3(kThreads), 0, ctx->stream, dA,
                           ndim, filter, do_filter ? 1 : 0, cnt);
        PTB_CHECK_LAUNCH(ctx);
    }
    const int64_t work = (len + 1) / 2;
    int blocks = (int)std::min<int64_t>((work + kThreads - 1) / kThreads,
                                        (int64_t)std::max(ctx->num_cus, 1) * 16);
    if (blocks < 1) blocks = 1;
    const bool aligned = (reinterpret_cast<uintptr_t>(dA) % 16) == 0;
    ProfScope kprof_(ctx, "filter_kernel");  // the packed pass alone
    if (aligned) {
        if (do_filter)
            hipLaunchKernelGGL(k_filter_pairs<true>, dim3(blocks), dim3(kThreads), 0, ctx->stream,
                               dA, len, filter, cnt);
        else
            hipLaunchKernelGGL(k_filter_pairs<false>, dim3(blocks), dim3(kThreads), 0,
                               ctx->stream, dA, len, filter, cnt);
    } else {
        if (do_filter)
            hipLaunchKernelGGL(k_filter_scalar<true>, dim3(blocks), dim3(kThreads), 0,
                               ctx->stream, dA, len, filter, cnt);
        else
            hipLaunchKernelGGL(k_filter_scalar<false>, dim3(blocks), dim3(kThreads), 0,
                               ctx->stream, dA, len, filter, cnt);
    }
    PTB_CHECK_LAUNCH(ctx);
    unsigned long long h[4];
    PTB_HIP(ctx, hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (counts) {
        // full symmetric count = 2 * packed count - diagonal count
        counts[0] = (int64_t)(2 * h[0] - h[2]);
        counts[1] = (int64_t)(2 * h[1] - h[3]);
    }
    return PTBGPU_OK;
}

}  // namespace ptb
