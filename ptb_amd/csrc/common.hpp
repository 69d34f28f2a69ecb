// Shared definitions of the engine: context, device workspace, error macros,
// wavefront reductions.  gfx950 only (64-wide wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/ptbgpu.h"

#define PTB_WAVE 64

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct ProfRecord {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};

struct ptbgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    bool profiling = false;
    std::map<std::string, ProfRecord> prof;  // kernel name -> bracketing events
    std::map<std::string, DevBuf> bufs;      // named device workspace, grows on demand
    std::map<std::string, DevBuf> host_bufs; // pinned host staging
    std::map<std::string, std::vector<char>> shadow;  // last upload_table contents per buffer
    int num_cus = 0;
    // the last completed density build (what ptbgpu_sm_spectra may read back):
    // submatrices and their summed dimensions; 0 while none or during a build
    int64_t last_nsm = 0;
    int64_t last_sum_m = 0;
    // tuning knobs (ptbgpu_set_tuning): the smallest size class of the
    // divide-and-conquer merge launches (0: one launch per tree level)
    int64_t dc_bucket_min = 256;
    // mixed-precision builds reduce submatrices below this dimension in FP64
    // (no float storage, no refinement sweep); 0: every submatrix in float.
    // 288: the tuner's choice on MI355X (below it the float reduction saves
    // less than the refinement costs, and small near-degenerate blocks leave
    // the refined density outside 1e-6 of FP64)
    int64_t mixed_min_m = 288;
    // a mixed build's float-reduced submatrices of up to this many rows take
    // their divide-and-conquer merge products in the low-precision
    // arithmetic (split BF16x3, or BF16); 0: FP64 merge products everywhere
    int64_t lowp_merge_max_m = 0;
    // merge size classes (power-of-two node sizes) at or above this run their
    // secular roots, Loewner vectors and eigenvectors over the whole chip from
    // device scratch even when they fit in LDS (0: only those that do not)
    int64_t dc_global_min = 0;
    // tile edge of the FP64 GEMM launches: 0 chooses per launch (128 x 128
    // tiles for launches with many large products), 64 or 128 forces it
    int64_t gemm_tile = 0;
    // "graph": the eigen phase of repeated one-chunk builds replays a captured
    // hipGraph; the cache of instantiated graphs (keyed by shapes and
    // buffers), the keys seen once, and whether a capture is in progress (no
    // allocation or table upload may happen then)
    int64_t graph_mode = 0;
    std::map<uint64_t, std::shared_ptr<void>> graphs;
    std::set<uint64_t> graph_seen;
    bool capturing = false;
    std::string graph_note;     // what the last eigen phase did (replayed, captured, why not)
    std::string table_prefix;   // prefix of table names while a graph-held phase runs
    int64_t graph_failures = 0;
    // the last chunked build: its precision and submatrix dimensions (the
    // key) and its chunk bounds
    std::vector<int64_t> chunk_key;
    std::vector<int> chunk_cbeg;
    // the pinned staging buffers of large host copies: the event recorded
    // after each buffer's last device copy, and whether one is pending
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_busy[2] = {false, false};
    // PTBGPU_POISON_ARENAS=1 at ptbgpu_init: the floating-point arenas of a
    // density build are filled with NaN before each (eager) use, so that a
    // read of an element the build did not write shows in the result (tests)
    bool poison_arenas = false;
};

namespace ptb {

// Record an error on the context and return its status.
int fail(ptbgpu_ctx *ctx, int status, const char *fmt, ...);

// Named device workspace of at least `bytes` (contents undefined after growth).
void *ws(ptbgpu_ctx *ctx, const char *name, size_t bytes);
// releases the named buffer if it holds more than `bytes` (0: any)
void ws_trim(ptbgpu_ctx *ctx, const char *name, size_t bytes);
// Named device table of `bytes` filled from host memory: the copy is skipped
// when the buffer still holds exactly these bytes from the previous call
// (descriptor tables of repeated builds with the same structure).  Returns the
// device pointer, or nullptr with the error recorded.
void *upload_table(ptbgpu_ctx *ctx, const char *name, const void *src, size_t bytes);
// Named pinned host buffer of at least `bytes`.
void *hws(ptbgpu_ctx *ctx, const char *name, size_t bytes);
// Copies between pageable host memory and the device, on the context's
// stream.  Large copies go in chunks through two pinned staging buffers: the
// host side of each chunk is copied by several threads while the previous
// chunk's DMA runs (the runtime stages pageable memory on one thread).  h2d
// returns once the host memory may be reused (device copies may still be in
// flight); d2h returns with the host memory written.
int h2d(ptbgpu_ctx *ctx, void *dst, const void *src, size_t bytes);
int d2h(ptbgpu_ctx *ctx, void *dst, const void *src, size_t bytes);

// Brackets the launches issued during its lifetime with events on the
// context's stream when profiling is enabled (ptbgpu_profile).
struct ProfScope {
    ptbgpu_ctx *ctx;
    hipEvent_t b = nullptr, e = nullptr;
    ProfRecord *rec = nullptr;
    ProfScope(ptbgpu_ctx *c, const char *name) : ctx(c) {
        if (!ctx->profiling) return;
        if (hipEventCreate(&b) != hipSuccess || hipEventCreate(&e) != hipSuccess) return;
        rec = &ctx->prof[name];
        (void)hipEventRecord(b, ctx->stream);
    }
    ~ProfScope() {
        if (!rec) return;
        (void)hipEventRecord(e, ctx->stream);
        rec->ev.emplace_back(b, e);
    }
};

}  // namespace ptb

#define PTB_HIP(ctx, call)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ptb::fail((ctx), PTBGPU_E_HIP, "%s failed: %s (%s:%d)", #call,         \
                             hipGetErrorString(e_), __FILE__, __LINE__);                   \
    } while (0)

#define PTB_CHECK_LAUNCH(ctx) PTB_HIP(ctx, hipGetLastError())

#define PTB_TRY(expr)                  \
    do {                               \
        int s_ = (expr);               \
        if (s_ != PTBGPU_OK) return s_; \
    } while (0)

// ---- wavefront reductions (64 lanes) -------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = PTB_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, PTB_WAVE);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = PTB_WAVE / 2; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, PTB_WAVE);
        v = o > v ? o : v;
    }
    return v;
}

// Block-wide sum; every thread gets the result. `scratch` holds >= 16 T.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *scratch) {
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
    const int nw = (blockDim.x + PTB_WAVE - 1) / PTB_WAVE;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) scratch[wid] = v;
    __syncthreads();
    T t = 0;
    for (int w = 0; w < nw; ++w) t += scratch[w];
    __syncthreads();
    return t;
}

// packed-upper index of (i,j), 0-based, i <= j
__host__ __device__ __forceinline__ int64_t pk(int64_t i, int64_t j) {
    return i + j * (j + 1) / 2;
}

// ---- kernels exported between translation units ---------------------------
namespace ptb {
int filter_count(ptbgpu_ctx *ctx, double *dA, int64_t ndim, double filter, bool do_filter,
                 int64_t counts[2]);
// metrics.hip: device address of the static scratch of the allocation-free
// entry point, looked up once per device at context creation
int noalloc_scratch_init(ptbgpu_ctx *ctx);
}
