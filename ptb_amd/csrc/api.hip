// C ABI entry points: context management, workspace, filter/analyse.
#include <cstdarg>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "common.hpp"

namespace ptb {

int fail(ptbgpu_ctx *ctx, int status, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->last_error = buf;
    return status;
}

void *ws(ptbgpu_ctx *ctx, const char *name, size_t bytes) {
    DevBuf &b = ctx->bufs[name];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    if (ctx->capturing) {  // no allocation inside a graph capture
        fail(ctx, PTBGPU_E_NOMEM, "buffer '%s' would grow during a graph capture", name);
        return nullptr;
    }
    ctx->shadow.erase(name);  // new memory: no table is known to be in it
    if (b.ptr) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
        (void)hipGetLastError();
        b.ptr = nullptr;
        fail(ctx, PTBGPU_E_NOMEM, "device allocation of %zu bytes for '%s' failed", want, name);
        return nullptr;
    }
    b.bytes = want;
    return b.ptr;
}

void ws_trim(ptbgpu_ctx *ctx, const char *name, size_t bytes) {
    auto it = ctx->bufs.find(name);
    if (it == ctx->bufs.end() || !it->second.ptr || it->second.bytes <= bytes) return;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(it->second.ptr);
    ctx->bufs.erase(it);
    ctx->shadow.erase(name);
}

void *upload_table(ptbgpu_ctx *ctx, const char *name_in, const void *src, size_t bytes) {
    // tables of a graph-held phase live under the graph's own prefix, so that
    // other builds do not overwrite what its replays read
    const std::string full = ctx->table_prefix + name_in;
    const char *name = full.c_str();
    void *d = ws(ctx, name, bytes ? bytes : 1);
    if (!d || bytes == 0) return d;
    auto it = ctx->shadow.find(name);
    if (it != ctx->shadow.end() && it->second.size() == bytes &&
        memcmp(it->second.data(), src, bytes) == 0)
        return d;
    if (ctx->capturing) {  // a graph must not hold a copy from host memory
        fail(ctx, PTBGPU_E_ARG, "table '%s' changed during a graph capture", name);
        return nullptr;
    }
    if (hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        ctx->shadow.erase(name);
        fail(ctx, PTBGPU_E_HIP, "upload of '%s' failed: %s", name, hipGetErrorString(hipGetLastError()));
        return nullptr;
    }
    const char *c = static_cast<const char *>(src);
    ctx->shadow[name].assign(c, c + bytes);
    return d;
}

void *hws(ptbgpu_ctx *ctx, const char *name, size_t bytes) {
    DevBuf &b = ctx->host_bufs[name];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    if (b.ptr) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipHostFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    if (hipHostMalloc(&b.ptr, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        b.ptr = nullptr;
        fail(ctx, PTBGPU_E_NOMEM, "pinned host allocation of %zu bytes for '%s' failed", want,
             name);
        return nullptr;
    }
    b.bytes = want;
    return b.ptr;
}

namespace {

constexpr size_t kStageDirect = (size_t)8 << 20;  // smaller copies: one runtime copy
constexpr size_t kStageChunk = (size_t)32 << 20;

// dst[0, n) = src[0, n) on up to eight host threads (at least 4 MB each)
void par_copy(char *dst, const char *src, size_t n) {
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(8, hw), std::max<size_t>(1, n >> 22));
    if (nt <= 1) {
        memcpy(dst, src, n);
        return;
    }
    const size_t part = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    size_t done = part;  // the calling thread copies [0, part)
    try {
        for (size_t t = 1; t < nt && t * part < n; ++t) {
            const size_t b = t * part, len = std::min(part, n - b);
            th.emplace_back([=] { memcpy(dst + b, src + b, len); });
            done = b + len;
        }
    } catch (...) {
        // no thread could be started: what is not handed out is copied here
    }
    memcpy(dst, src, std::min(part, n));
    if (done < n) memcpy(dst + done, src + done, n - done);
    for (auto &x : th) x.join();
}

// the staging buffers, with their events created
int stage_buffers(ptbgpu_ctx *ctx, char *(&stage)[2]) {
    for (auto &e : ctx->stage_ev)
        if (!e) PTB_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    stage[0] = static_cast<char *>(hws(ctx, "stage0", kStageChunk));
    stage[1] = static_cast<char *>(hws(ctx, "stage1", kStageChunk));
    return stage[0] && stage[1] ? PTBGPU_OK : PTBGPU_E_NOMEM;
}

// waits until staging buffer b is no longer read or written by a device copy
int stage_wait(ptbgpu_ctx *ctx, int b) {
    if (ctx->stage_busy[b]) PTB_HIP(ctx, hipEventSynchronize(ctx->stage_ev[b]));
    ctx->stage_busy[b] = false;
    return PTBGPU_OK;
}

}  // namespace

int h2d(ptbgpu_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes < kStageDirect) {
        PTB_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return PTBGPU_OK;
    }
    char *stage[2];
    PTB_TRY(stage_buffers(ctx, stage));
    size_t i = 0;
    for (size_t off = 0; off < bytes; off += kStageChunk, ++i) {
        const size_t n = std::min(kStageChunk, bytes - off);
        const int b = (int)(i & 1);
        PTB_TRY(stage_wait(ctx, b));
        par_copy(stage[b], static_cast<const char *>(src) + off, n);
        PTB_HIP(ctx, hipMemcpyAsync(static_cast<char *>(dst) + off, stage[b], n, hipMemcpyHostToDevice,
                                    ctx->stream));
        PTB_HIP(ctx, hipEventRecord(ctx->stage_ev[b], ctx->stream));
        ctx->stage_busy[b] = true;
    }
    return PTBGPU_OK;
}

int d2h(ptbgpu_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes < kStageDirect) {
        PTB_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return PTBGPU_OK;
    }
    char *stage[2];
    PTB_TRY(stage_buffers(ctx, stage));
    // chunk held by each buffer, still to be copied out
    size_t pend_off[2] = {0, 0}, pend_n[2] = {0, 0};
    auto drain = [&](int b) -> int {
        PTB_TRY(stage_wait(ctx, b));
        if (pend_n[b]) par_copy(static_cast<char *>(dst) + pend_off[b], stage[b], pend_n[b]);
        pend_n[b] = 0;
        return PTBGPU_OK;
    };
    size_t i = 0;
    for (size_t off = 0; off < bytes; off += kStageChunk, ++i) {
        const size_t n = std::min(kStageChunk, bytes - off);
        const int b = (int)(i & 1);
        PTB_TRY(drain(b));  // the chunk two back (its copy-out overlaps the next DMA)
        PTB_HIP(ctx, hipMemcpyAsync(stage[b], static_cast<const char *>(src) + off, n,
                                    hipMemcpyDeviceToHost, ctx->stream));
        PTB_HIP(ctx, hipEventRecord(ctx->stage_ev[b], ctx->stream));
        ctx->stage_busy[b] = true;
        pend_off[b] = off;
        pend_n[b] = n;
    }
    PTB_TRY(drain((int)(i & 1)));  // the older of the last two chunks first
    PTB_TRY(drain((int)((i + 1) & 1)));
    return PTBGPU_OK;
}

}  // namespace ptb

static void prof_clear(ptbgpu_ctx *ctx);

extern "C" {

const char *ptbgpu_strerror(int64_t status) {
    switch (status) {
        case PTBGPU_OK: return "ok";
        case PTBGPU_E_ARG: return "invalid argument";
        case PTBGPU_E_MODE: return "unsupported submatrix mode";
        case PTBGPU_E_NOT_SPD: return "overlap submatrix is not positive definite";
        case PTBGPU_E_NAN: return "NaN encountered";
        case PTBGPU_E_KMEANS: return "k-means clustering failed";
        case PTBGPU_E_HIP: return "HIP runtime error";
        case PTBGPU_E_NOMEM: return "device allocation failed";
        case PTBGPU_E_NOCONV: return "iteration did not converge";
        default: return "unknown status";
    }
}

const char *ptbgpu_graph_note(const ptbgpu_ctx *ctx) {
    return ctx ? ctx->graph_note.c_str() : "no context";
}

const char *ptbgpu_last_error(const ptbgpu_ctx *ctx) {
    return ctx ? ctx->last_error.c_str() : "no context";
}

int ptbgpu_init(ptbgpu_ctx **out, int64_t device, int64_t flags) {
    (void)flags;
    if (!out) return PTBGPU_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PTBGPU_E_HIP;
    if (device < 0 || device >= ndev) return PTBGPU_E_ARG;
    auto *ctx = new ptbgpu_ctx();
    ctx->device = (int)device;
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    PTB_HIP(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    PTB_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
    ctx->num_cus = prop.multiProcessorCount;
    if (const char *pz = std::getenv("PTBGPU_POISON_ARENAS")) ctx->poison_arenas = pz[0] == '1';
    PTB_TRY(ptb::noalloc_scratch_init(ctx));
    *out = ctx;
    return PTBGPU_OK;
}

int ptbgpu_finalize(ptbgpu_ctx *ctx) {
    if (!ctx) return PTBGPU_E_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->graphs.clear();
    prof_clear(ctx);
    for (auto &kv : ctx->bufs)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    for (auto &kv : ctx->host_bufs)
        if (kv.second.ptr) (void)hipHostFree(kv.second.ptr);
    for (auto &e : ctx->stage_ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PTBGPU_OK;
}

int ptbgpu_set_tuning(ptbgpu_ctx *ctx, const char *key, int64_t value) {
    if (!ctx || !key) return PTBGPU_E_ARG;
    const std::string k(key);
    if (k == "dc_bucket_min") {
        if (value < 0 || value > 4096) return ptb::fail(ctx, PTBGPU_E_ARG, "dc_bucket_min out of range");
        ctx->dc_bucket_min = value;
        return PTBGPU_OK;
    }
    if (k == "mixed_min_m") {
        if (value < 0 || value > 1 << 20) return ptb::fail(ctx, PTBGPU_E_ARG, "mixed_min_m out of range");
        ctx->mixed_min_m = value;
        return PTBGPU_OK;
    }
    if (k == "lowp_merge_max_m") {
        if (value < 0 || value > 4096) return ptb::fail(ctx, PTBGPU_E_ARG, "lowp_merge_max_m out of range");
        ctx->lowp_merge_max_m = value;
        return PTBGPU_OK;
    }
    if (k == "dc_global_min") {
        if (value < 0 || value > (1 << 24)) return ptb::fail(ctx, PTBGPU_E_ARG, "dc_global_min out of range");
        ctx->dc_global_min = value;
        return PTBGPU_OK;
    }
    if (k == "gemm_tile") {
        if (value != 0 && value != 64 && value != 128)
            return ptb::fail(ctx, PTBGPU_E_ARG, "gemm_tile must be 0, 64 or 128");
        ctx->gemm_tile = value;
        return PTBGPU_OK;
    }
    if (k == "graph") {
        ctx->graph_mode = value != 0;
        if (!ctx->graph_mode) {
            ctx->graphs.clear();
            ctx->graph_seen.clear();
        }
        return PTBGPU_OK;
    }
    return ptb::fail(ctx, PTBGPU_E_ARG, "unknown tuning key '%s'", key);
}

int ptbgpu_get_tuning(ptbgpu_ctx *ctx, const char *key, int64_t *value) {
    if (!ctx || !key || !value) return PTBGPU_E_ARG;
    const std::string k(key);
    if (k == "dc_bucket_min") {
        *value = ctx->dc_bucket_min;
        return PTBGPU_OK;
    }
    if (k == "mixed_min_m") {
        *value = ctx->mixed_min_m;
        return PTBGPU_OK;
    }
    if (k == "lowp_merge_max_m") {
        *value = ctx->lowp_merge_max_m;
        return PTBGPU_OK;
    }
    if (k == "dc_global_min") {
        *value = ctx->dc_global_min;
        return PTBGPU_OK;
    }
    if (k == "gemm_tile") {
        *value = ctx->gemm_tile;
        return PTBGPU_OK;
    }
    if (k == "graph") {
        *value = ctx->graph_mode;
        return PTBGPU_OK;
    }
    if (k == "graphs") {  // instantiated graphs held by the context
        *value = (int64_t)ctx->graphs.size();
        return PTBGPU_OK;
    }
    if (k == "poison_arenas") {  // PTBGPU_POISON_ARENAS at ptbgpu_init (read only)
        *value = ctx->poison_arenas ? 1 : 0;
        return PTBGPU_OK;
    }
    if (k == "graph_failures") {  // captures or instantiations that did not succeed
        *value = ctx->graph_failures;
        return PTBGPU_OK;
    }
    return ptb::fail(ctx, PTBGPU_E_ARG, "unknown tuning key '%s'", key);
}

int ptbgpu_sync(ptbgpu_ctx *ctx) {
    if (!ctx) return PTBGPU_E_ARG;
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PTBGPU_OK;
}

int ptbgpu_stream(ptbgpu_ctx *ctx, void **stream) {
    if (!ctx || !stream) return PTBGPU_E_ARG;
    *stream = (void *)ctx->stream;
    return PTBGPU_OK;
}

static void prof_clear(ptbgpu_ctx *ctx) {
    for (auto &kv : ctx->prof)
        for (auto &pe : kv.second.ev) {
            (void)hipEventDestroy(pe.first);
            (void)hipEventDestroy(pe.second);
        }
    ctx->prof.clear();
}

int ptbgpu_profile(ptbgpu_ctx *ctx, int64_t enable) {
    if (!ctx) return PTBGPU_E_ARG;
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    prof_clear(ctx);
    ctx->profiling = enable != 0;
    return PTBGPU_OK;
}

int ptbgpu_profile_query(ptbgpu_ctx *ctx, const char *name, double *total_ms, int64_t *launches) {
    if (!ctx || !name) return PTBGPU_E_ARG;
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double t = 0.0;
    int64_t c = 0;
    auto it = ctx->prof.find(name);
    if (it != ctx->prof.end())
        for (auto &pe : it->second.ev) {
            float ms = 0.f;
            PTB_HIP(ctx, hipEventElapsedTime(&ms, pe.first, pe.second));
            t += ms;
            ++c;
        }
    if (total_ms) *total_ms = t;
    if (launches) *launches = c;
    return PTBGPU_OK;
}

int ptbgpu_filter_sym_dev(ptbgpu_ctx *ctx, int64_t ndim, double *dA, double filter,
                          int64_t counts[2]) {
    if (!ctx) return PTBGPU_E_ARG;
    if (ndim < 0 || (ndim > 0 && !dA)) return ptb::fail(ctx, PTBGPU_E_ARG, "bad ndim/pointer");
    if (ndim == 0) {
        if (counts) counts[0] = counts[1] = 0;
        return PTBGPU_OK;
    }
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    return ptb::filter_count(ctx, dA, ndim, filter, true, counts);
}

int ptbgpu_analyse_sym_dev(ptbgpu_ctx *ctx, int64_t ndim, const double *dA, int64_t *npos) {
    if (!ctx) return PTBGPU_E_ARG;
    if (ndim < 0 || (ndim > 0 && !dA)) return ptb::fail(ctx, PTBGPU_E_ARG, "bad ndim/pointer");
    int64_t c[2] = {0, 0};
    if (ndim > 0) {
        PTB_HIP(ctx, hipSetDevice(ctx->device));
        PTB_TRY(ptb::filter_count(ctx, const_cast<double *>(dA), ndim, 0.0, false, c));
    }
    if (npos) *npos = c[0];
    return PTBGPU_OK;
}

int ptbgpu_filter_sym(ptbgpu_ctx *ctx, int64_t ndim, double *A, double filter,
                      int64_t counts[2]) {
    if (!ctx) return PTBGPU_E_ARG;
    if (ndim < 0 || (ndim > 0 && !A)) return ptb::fail(ctx, PTBGPU_E_ARG, "bad ndim/pointer");
    if (ndim == 0) {
        if (counts) counts[0] = counts[1] = 0;
        return PTBGPU_OK;
    }
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    const size_t bytes = sizeof(double) * (size_t)(ndim * (ndim + 1) / 2);
    auto *d = static_cast<double *>(ptb::ws(ctx, "io_a", bytes));
    if (!d) return PTBGPU_E_NOMEM;
    PTB_TRY(ptb::h2d(ctx, d, A, bytes));
    PTB_TRY(ptb::filter_count(ctx, d, ndim, filter, true, counts));
    return ptb::d2h(ctx, A, d, bytes);
}

int ptbgpu_analyse_sym(ptbgpu_ctx *ctx, int64_t ndim, const double *A, int64_t *npos) {
    if (!ctx) return PTBGPU_E_ARG;
    if (ndim < 0 || (ndim > 0 && !A)) return ptb::fail(ctx, PTBGPU_E_ARG, "bad ndim/pointer");
    if (ndim == 0) {
        if (npos) *npos = 0;
        return PTBGPU_OK;
    }
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    const size_t bytes = sizeof(double) * (size_t)(ndim * (ndim + 1) / 2);
    auto *d = static_cast<double *>(ptb::ws(ctx, "io_a", bytes));
    if (!d) return PTBGPU_E_NOMEM;
    PTB_TRY(ptb::h2d(ctx, d, A, bytes));
    int64_t c[2];
    PTB_TRY(ptb::filter_count(ctx, d, ndim, 0.0, false, c));
    if (npos) *npos = c[0];
    return PTBGPU_OK;
}

}  // extern "C"
