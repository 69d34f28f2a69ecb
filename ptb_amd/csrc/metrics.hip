// Density checks over packed-upper matrices (source/metrics.f90).
//
//   get_nel_packed (:202-230)              tr(P S)
//   get_band_structure_E_packed (:253-277) tr(P H)
//   idempotent_packed (:316-360)           |sum_ij (0.5 P S P - P)_ij|
//
// The reference blows both matrices up and runs an N^3 dgemm for each trace.
// For symmetric P, S: tr(P S) = sum_ij P_ij S_ij = sum over the packed upper
// triangle of w_ij P_ij S_ij with w = 1 on the diagonal and 2 off it, one
// streaming pass.  The idempotency sum needs no product either:
// 1^T (P S P) 1 = u^T S u with u = P 1, so the deviation is
// |0.5 u^T S u - 1^T u|, two O(N^2) passes.
//
// All reductions are deterministic: fixed per-thread strides, wavefront then
// block sums into per-block partials, and one ordered final sum.
#include <mutex>

#include "common.hpp"

namespace {

constexpr int kThreads = 256;

// Scratch of ptbgpu_idempotency_noalloc: a file-scope device array, so that
// the call allocates nothing.  Every GPU the library runs on holds its own
// copy from the first context on it until the process ends.  40 MiB covers
// the idempotency check's row-sum partials up to ndim ~ 51000 (2048 waters:
// 47104).  Host code reaches it only through the address hipGetSymbolAddress
// returns for the device (kept per device below); the array's name itself is
// not a device address on the host.
constexpr size_t kScratchDoubles = (size_t)5 << 20;
__device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// packed index k -> column j (largest j with j(j+1)/2 <= k), exact for k < 2^52
__device__ __forceinline__ int64_t col_of(int64_t k) {
    double d = sqrt(8.0 * (double)k + 1.0);
    int64_t j = (int64_t)((d - 1.0) * 0.5);
    while (j * (j + 1) / 2 > k) --j;
    while ((j + 1) * (j + 2) / 2 <= k) ++j;
    return j;
}

// partial[block] = sum_k w_k a_k b_k   (w = 1 diagonal, 2 off-diagonal)
__global__ __launch_bounds__(kThreads) void k_weighted_dot(const double *__restrict__ a,
                                                           const double *__restrict__ b,
                                                           int64_t len,
                                                           double *__restrict__ partial) {
    __shared__ double red[16];
    const int64_t per = (len + gridDim.x - 1) / gridDim.x;
    const int64_t k0 = (int64_t)blockIdx.x * per, k1 = min(len, k0 + per);
    double acc = 0.0;
    if (k0 < k1) {
        // walk the block's range column by column so the diagonal test is cheap
        int64_t j = col_of(k0);
        for (int64_t base = k0; base < k1; base += kThreads) {
            int64_t k = base + threadIdx.x;
            if (k < k1) {
                int64_t jj = j;
                while ((jj + 1) * (jj + 2) / 2 <= k) ++jj;
                const int64_t i = k - jj * (jj + 1) / 2;
                acc += (i == jj ? 1.0 : 2.0) * a[k] * b[k];
            }
            // advance the column hint to the last element handled by this sweep
            int64_t last = min(k1 - 1, base + kThreads - 1);
            while ((j + 1) * (j + 2) / 2 <= last) ++j;
        }
    }
    double t = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Row part of u = P 1: urow[chunk*ndim + i] = sum_{j in chunk, j >= i} P(i,j).
// Threads own rows, so a wavefront reads 64 consecutive rows of one packed
// column (contiguous addresses).
__global__ __launch_bounds__(kThreads) void k_rowsum_upper(const double *__restrict__ p,
                                                           int64_t ndim, int64_t chunk,
                                                           double *__restrict__ urow) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t c = blockIdx.y;
    const int64_t j0 = c * chunk, j1 = min(ndim, j0 + chunk);
    double acc = 0.0;
    if (i < ndim) {
        for (int64_t j = max(j0, i); j < j1; ++j) acc += p[pk(i, j)];
        urow[c * ndim + i] = acc;
    }
}

// Column part (strictly above the diagonal of column i) plus the ordered sum
// of the row chunks: u_i = sum_c urow[c][i] + sum_{j<i} P(j,i).
__global__ __launch_bounds__(kThreads) void k_colsum_finish(const double *__restrict__ p,
                                                            int64_t ndim, int64_t nchunk,
                                                            const double *__restrict__ urow,
                                                            double *__restrict__ u) {
    __shared__ double red[16];
    const int64_t i = blockIdx.x;
    double acc = 0.0;
    const double *col = p + i * (i + 1) / 2;
    for (int64_t r = threadIdx.x; r < i; r += kThreads) acc += col[r];
    double t = block_sum(acc, red);
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int64_t c = 0; c < nchunk; ++c) s += urow[c * ndim + i];
        u[i] = s + t;
    }
}

// partial[block] = sum over packed entries of w_ij u_i S_ij u_j, and the
// block's share of sum_i u_i in partial2
__global__ __launch_bounds__(kThreads) void k_quadform(const double *__restrict__ s,
                                                       const double *__restrict__ u,
                                                       int64_t ndim, double *__restrict__ partial,
                                                       double *__restrict__ partial2) {
    __shared__ double red[16];
    double acc = 0.0, acc2 = 0.0;
    for (int64_t j = blockIdx.x; j < ndim; j += gridDim.x) {
        const double uj = u[j];
        const double *col = s + j * (j + 1) / 2;
        for (int64_t i = threadIdx.x; i <= j; i += kThreads)
            acc += (i == j ? 1.0 : 2.0) * u[i] * col[i] * uj;
        if (threadIdx.x == 0) acc2 += uj;
    }
    double t = block_sum(acc, red);
    double t2 = block_sum(acc2, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = t;
        partial2[blockIdx.x] = t2;
    }
}

__global__ void k_ordered_sum(const double *__restrict__ partial, int n, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += partial[i];
        *out = s;
    }
}

// deviation = |0.5 u^T S u - 1^T u| from the two ordered sums
__global__ void k_idem_finish(const double *__restrict__ fin, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = fabs(0.5 * fin[0] - fin[1]);
}

}  // namespace

namespace ptb {

int weighted_dot(ptbgpu_ctx *ctx, const double *dA, const double *dB, int64_t ndim,
                 double *result) {
    const int64_t len = ndim * (ndim + 1) / 2;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((len + kThreads - 1) / kThreads,
                                                                   (int64_t)ctx->num_cus * 8));
    auto *part = static_cast<double *>(ws(ctx, "metric_part", sizeof(double) * (blocks + 1)));
    if (!part) return PTBGPU_E_NOMEM;
    hipLaunchKernelGGL(k_weighted_dot, dim3(blocks), dim3(kThreads), 0, ctx->stream, dA, dB, len,
                       part);
    PTB_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_ordered_sum, dim3(1), dim3(1), 0, ctx->stream, part, blocks, part + blocks);
    PTB_CHECK_LAUNCH(ctx);
    PTB_HIP(ctx, hipMemcpyAsync(result, part + blocks, sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PTBGPU_OK;
}

// The idempotency check's scratch, in doubles: row-sum partials of 512-column
// chunks, u = P 1, the quadratic form's block partials, two ordered sums.
static constexpr int64_t kIdemChunk = 512;
static int idem_qblocks(const ptbgpu_ctx *ctx, int64_t ndim) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(ndim, (int64_t)ctx->num_cus * 4));
}
static size_t idem_scratch_doubles(const ptbgpu_ctx *ctx, int64_t ndim) {
    const int64_t nchunk = (ndim + kIdemChunk - 1) / kIdemChunk;
    return (size_t)(nchunk * ndim + ndim + 2 * idem_qblocks(ctx, ndim) + 2);
}

// Enqueues the check on the context's stream: buf holds
// idem_scratch_doubles(ndim) doubles, the deviation goes to *dout (device).
static int idempotency_enqueue(ptbgpu_ctx *ctx, const double *dP, const double *dS, int64_t ndim,
                               double *buf, double *dout) {
    const int64_t chunk = kIdemChunk;
    const int64_t nchunk = (ndim + chunk - 1) / chunk;
    const int qblocks = idem_qblocks(ctx, ndim);
    double *urow = buf, *u = urow + nchunk * ndim, *part = u + ndim, *part2 = part + qblocks;
    double *fin = part2 + qblocks;
    hipLaunchKernelGGL(k_rowsum_upper, dim3((ndim + kThreads - 1) / kThreads, nchunk),
                       dim3(kThreads), 0, ctx->stream, dP, ndim, chunk, urow);
    PTB_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_colsum_finish, dim3(ndim), dim3(kThreads), 0, ctx->stream, dP, ndim,
                       nchunk, urow, u);
    PTB_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_quadform, dim3(qblocks), dim3(kThreads), 0, ctx->stream, dS, u, ndim,
                       part, part2);
    PTB_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_ordered_sum, dim3(1), dim3(1), 0, ctx->stream, part, qblocks, fin);
    hipLaunchKernelGGL(k_ordered_sum, dim3(1), dim3(1), 0, ctx->stream, part2, qblocks, fin + 1);
    PTB_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_idem_finish, dim3(1), dim3(1), 0, ctx->stream, fin, dout);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int idempotency(ptbgpu_ctx *ctx, const double *dP, const double *dS, int64_t ndim,
                double *result) {
    const size_t need = idem_scratch_doubles(ctx, ndim) + 1;
    auto *buf = static_cast<double *>(ws(ctx, "metric_idem", sizeof(double) * need));
    if (!buf) return PTBGPU_E_NOMEM;
    PTB_TRY(idempotency_enqueue(ctx, dP, dS, ndim, buf, buf + need - 1));
    PTB_HIP(ctx, hipMemcpyAsync(result, buf + need - 1, sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PTBGPU_OK;
}

// The static scratch's device address on the context's device, looked up once
// per device (the context's device is current when ptbgpu_init calls this).
int noalloc_scratch_init(ptbgpu_ctx *ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return fail(ctx, PTBGPU_E_ARG, "device %d beyond the scratch table", ctx->device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[ctx->device]) return PTBGPU_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(ctx, PTBGPU_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_dev[ctx->device] = static_cast<double *>(p);
    return PTBGPU_OK;
}

// host arrays -> two named device buffers
static int upload2(ptbgpu_ctx *ctx, int64_t ndim, const double *A, const double *B, double **dA,
                   double **dB) {
    const size_t bytes = sizeof(double) * (size_t)(ndim * (ndim + 1) / 2);
    *dA = static_cast<double *>(ws(ctx, "io_a", bytes));
    *dB = static_cast<double *>(ws(ctx, "io_b", bytes));
    if (!*dA || !*dB) return PTBGPU_E_NOMEM;
    PTB_HIP(ctx, hipMemcpyAsync(*dA, A, bytes, hipMemcpyHostToDevice, ctx->stream));
    PTB_HIP(ctx, hipMemcpyAsync(*dB, B, bytes, hipMemcpyHostToDevice, ctx->stream));
    return PTBGPU_OK;
}

}  // namespace ptb

extern "C" {

static int metric_args(ptbgpu_ctx *ctx, int64_t ndim, const void *a, const void *b, void *out) {
    if (!ctx) return PTBGPU_E_ARG;
    if (ndim < 1 || !a || !b || !out) return ptb::fail(ctx, PTBGPU_E_ARG, "bad ndim/pointer");
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    return PTBGPU_OK;
}

int ptbgpu_get_nel(ptbgpu_ctx *ctx, int64_t ndim, const double *P, const double *S, double *nel) {
    PTB_TRY(metric_args(ctx, ndim, P, S, nel));
    double *dP, *dS;
    PTB_TRY(ptb::upload2(ctx, ndim, P, S, &dP, &dS));
    return ptb::weighted_dot(ctx, dP, dS, ndim, nel);
}

int ptbgpu_band_energy(ptbgpu_ctx *ctx, int64_t ndim, const double *P, const double *H,
                       double *energy) {
    PTB_TRY(metric_args(ctx, ndim, P, H, energy));
    double *dP, *dH;
    PTB_TRY(ptb::upload2(ctx, ndim, P, H, &dP, &dH));
    return ptb::weighted_dot(ctx, dP, dH, ndim, energy);
}

int ptbgpu_get_nel_dev(ptbgpu_ctx *ctx, int64_t ndim, const double *dP, const double *dS,
                       double *nel) {
    PTB_TRY(metric_args(ctx, ndim, dP, dS, nel));
    return ptb::weighted_dot(ctx, dP, dS, ndim, nel);
}

int ptbgpu_band_energy_dev(ptbgpu_ctx *ctx, int64_t ndim, const double *dP, const double *dH,
                           double *energy) {
    PTB_TRY(metric_args(ctx, ndim, dP, dH, energy));
    return ptb::weighted_dot(ctx, dP, dH, ndim, energy);
}

int ptbgpu_idempotency(ptbgpu_ctx *ctx, int64_t ndim, const double *P, const double *S,
                       double *deviation) {
    PTB_TRY(metric_args(ctx, ndim, P, S, deviation));
    double *dP, *dS;
    PTB_TRY(ptb::upload2(ctx, ndim, P, S, &dP, &dS));
    return ptb::idempotency(ctx, dP, dS, ndim, deviation);
}

int ptbgpu_idempotency_noalloc(ptbgpu_ctx *ctx, int64_t ndim, const double *dP, const double *dS,
                               double *d_deviation) {
    PTB_TRY(metric_args(ctx, ndim, dP, dS, d_deviation));
    double *scr = g_scratch_dev[ctx->device];
    if (!scr) return ptb::fail(ctx, PTBGPU_E_ARG, "static scratch not set up on device %d", ctx->device);
    const size_t need = ptb::idem_scratch_doubles(ctx, ndim);
    if (need > kScratchDoubles)
        return ptb::fail(ctx, PTBGPU_E_ARG,
                         "ndim %lld needs %zu bytes of scratch, above the static %zu bytes",
                         (long long)ndim, need * sizeof(double), kScratchDoubles * sizeof(double));
    return ptb::idempotency_enqueue(ctx, dP, dS, ndim, scr, d_deviation);
}

}  // extern "C"
