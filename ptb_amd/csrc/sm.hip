// Orchestration of the submatrix density build: the C ABI entry points that
// replace sm_stupid_simple (source/sm.f90:34-409).
//
// Phases (reference timer labels in brackets):
//   [submatrix combination]  column groups from aoat (host, closed forms)
//   [submatrix stats]        pattern bitmap, per-group row union, map (device)
//   [submatrix construction] gather of dense Hsm, Ssm
//   [submatrix eigendecomp]  Cholesky Ssm = L L^T, K = L^-1 Hsm L^-T,
//                            Householder tridiagonalisation K = Q T Q^T,
//                            divide & conquer T = Z diag(e) Z^T
//   [submatrix fe]           fe weights from the owned rows of L Q Z and
//                            L^-T Q Z (2c row vectors through Q and Z)
//   [mu iter]                bisection of the chemical potential
//   [submatrix final iter]   occupied columns only: U_o = L^-T Q Z_o, owned-
//                            column density, scatter into packed P
// Every phase is batched over all submatrices: one launch (or one per tree level
// / reflector block) covers the whole batch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "sm_internal.hpp"


namespace {

// nocc = 0 for the submatrices a mixed build reduced in FP64
__global__ __launch_bounds__(256) void k_skip_exact(const SubDesc *__restrict__ desc, int nsm,
                                                    int *__restrict__ nocc) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s < nsm && (desc[s].flags & kSubExact)) nocc[s] = 0;
}

struct Timer {
    hipEvent_t ev[PTBGPU_T_NSLOT + 1];
    bool ok = false;
    void init(hipStream_t s) {
        ok = true;
        for (auto &e : ev) ok = ok && hipEventCreate(&e) == hipSuccess;
        (void)s;
    }
    void mark(int slot, hipStream_t s) {
        if (ok) (void)hipEventRecord(ev[slot], s);
    }
    // 0 for a pair that was not recorded; the failed query's error is
    // cleared so that the next launch check does not report it
    float between(int a, int b) {
        float ms = 0.f;
        if (ok && hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess) {
            (void)hipGetLastError();
            ms = 0.f;
        }
        return ms;
    }
    ~Timer() {
        if (ok)
            for (auto &e : ev) (void)hipEventDestroy(e);
    }
};

// column groups, reference modes 1-3 (source/sm.f90:92-128)
int column_groups(int64_t mode, int64_t ndim, const int64_t *aoat, std::vector<int> &gptr,
                  std::vector<int> &gcols) {
    gptr.clear();
    gcols.resize(ndim);
    for (int64_t i = 0; i < ndim; ++i) gcols[i] = (int)i;
    gptr.push_back(0);
    if (mode == 1) {
        for (int64_t i = 1; i <= ndim; ++i) gptr.push_back((int)i);
    } else if (mode == 2) {
        for (int64_t i = 1; i < ndim; ++i)
            if (aoat[i] != aoat[i - 1]) gptr.push_back((int)i);
        gptr.push_back((int)ndim);
    } else if (mode == 3) {
        gptr.push_back((int)ndim);
    } else {
        return PTBGPU_E_MODE;  // 4 (k-means) is handled by kmeans_groups
    }
    return PTBGPU_OK;
}

// The uniform numbers k-means++ draws: the caller's stream (the compiler's
// RANDOM_NUMBER values after the reference's initseed), or splitmix64 seeded
// with 42.  Restarted for every candidate, as initseed reseeds
// (source/sm.f90:20-32,469).
struct Uniform {
    const double *stream;
    int64_t len, pos = 0;
    uint64_t state = 42;
    Uniform(const double *s, int64_t n) : stream(s), len(n) {}
    bool next(double &u) {
        if (stream) {
            if (pos >= len) return false;
            u = stream[pos++];
            return true;
        }
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        u = (double)(z >> 11) * 0x1.0p-53;
        return true;
    }
};

// k-means++ seeding and Lloyd iterations (KMPP, source/kmpp.f:24-179) on the
// atom positions x(3, n) (column-major): z receives 1-based cluster labels.
// Returns the reference's IFAULT: 0 converged, 1 empty cluster, 2 no
// convergence in 1000 sweeps, 3 bad k; -1 when the stream runs out.
int kmpp(const double *x, int n, int k, Uniform &rng, std::vector<int> &z) {
    const double big = (double)1e33f;  // REAL(4) literal
    z.assign(n, 0);
    if (k < 1 || k > n) return 3;
    std::vector<double> c(3 * (size_t)k, 0.0), work(n, big), cnt(k);
    double w;
    if (!rng.next(w)) return -1;
    int i1 = std::min((int)(w * (double)(float)n) + 1, n);
    for (int j = 0; j < 3; ++j) c[j] = x[3 * (i1 - 1) + j];
    for (int l = 2; l <= k; ++l) {
        double tot = 0.0;
        for (int i = 0; i < n; ++i) {
            double best = work[i], d2 = 0.0;
            bool skip = false;
            for (int j = 0; j < 3; ++j) {
                const double d = x[3 * i + j] - c[3 * (l - 2) + j];
                d2 += d * d;
                if (d2 >= best) {
                    skip = true;
                    break;
                }
            }
            if (!skip && d2 < best) {
                best = d2;
                work[i] = best;
            }
            tot += best;
        }
        if (!rng.next(w)) return -1;
        w *= tot;
        tot = 0.0;
        i1 = n;
        for (int i = 0; i < n; ++i) {
            i1 = i + 1;
            tot += work[i];
            if (tot > w) break;
        }
        for (int j = 0; j < 3; ++j) c[3 * (l - 1) + j] = x[3 * (i1 - 1) + j];
    }
    for (int sweep = 0; sweep < 1000; ++sweep) {
        bool change = false;
        for (int i = 0; i < n; ++i) {
            int l1 = 0;
            double best = big;
            for (int l = 1; l <= k; ++l) {
                double d2 = 0.0;
                for (int j = 0; j < 3; ++j) {
                    const double d = x[3 * i + j] - c[3 * (l - 1) + j];
                    d2 += d * d;
                    if (d2 >= best) break;
                }
                if (d2 < best) {
                    best = d2;
                    l1 = l;
                }
            }
            if (z[i] != l1) {
                z[i] = l1;
                change = true;
            }
        }
        if (!change) return 0;
        std::fill(cnt.begin(), cnt.end(), 0.0);
        std::fill(c.begin(), c.end(), 0.0);
        for (int i = 0; i < n; ++i) {
            cnt[z[i] - 1] += 1.0;
            for (int j = 0; j < 3; ++j) c[3 * (z[i] - 1) + j] += x[3 * i + j];
        }
        for (int l = 0; l < k; ++l)
            if (cnt[l] < 0.5) return 1;
        for (int l = 0; l < k; ++l)
            for (int j = 0; j < 3; ++j) c[3 * l + j] *= 1.0 / cnt[l];
    }
    return 2;
}

// Columns of every cluster: atom a's columns go to group z(a) - 1, ascending
// (source/sm.f90:476-490).  Appends nsm groups to gptr / gcols.
void cluster_columns(const std::vector<int> &z, int nsm, int64_t ndim, const int64_t *aoat,
                     std::vector<int> &gptr, std::vector<int> &gcols) {
    std::vector<std::vector<int>> groups(nsm);
    for (int64_t j = 0; j < ndim; ++j) groups[z[aoat[j] - 1] - 1].push_back((int)j);
    for (int g = 0; g < nsm; ++g) {
        gcols.insert(gcols.end(), groups[g].begin(), groups[g].end());
        gptr.push_back((int)gcols.size());
    }
}

void build_tree(int s, int lo, int hi, std::vector<DcNode> &leaves,
                std::vector<std::vector<DcNode>> &levels, int &height) {
    if (hi - lo <= kDcLeaf) {
        leaves.push_back({s, lo, hi, hi});
        height = 0;
        return;
    }
    const int sp = lo + (hi - lo) / 2;
    int h1, h2;
    build_tree(s, lo, sp, leaves, levels, h1);
    build_tree(s, sp, hi, leaves, levels, h2);
    height = 1 + std::max(h1, h2);
    if ((int)levels.size() < height) levels.resize(height);
    levels[height - 1].push_back({s, lo, sp, hi});
}

template <typename T>
int upload(ptbgpu_ctx *ctx, const char *name, const std::vector<T> &v, T **out) {
    *out = static_cast<T *>(ptb::upload_table(ctx, name, v.data(), sizeof(T) * v.size()));
    return *out ? PTBGPU_OK : PTBGPU_E_NOMEM;
}

// k-means column groups (submatrix_columns 4, source/sm.f90:129-157): for
// nsm = 1 .. min(64, n) cluster the atoms, take every cluster's columns and
// cost the candidate by sum m^2.5 over its submatrix dimensions; the first
// minimum wins.  All candidates' row unions are counted in one launch on the
// pattern bitmap.
int kmeans_groups(ptbgpu_ctx *ctx, int64_t ndim, int64_t n, const double *xyz,
                  const int64_t *aoat, const uint32_t *bits, int64_t words,
                  const ptbgpu_opts &opts, std::vector<int> &gptr, std::vector<int> &gcols) {
    if (!xyz || n < 1) return ptb::fail(ctx, PTBGPU_E_ARG, "k-means needs the atom positions");
    for (int64_t j = 0; j < ndim; ++j)
        if (aoat[j] < 1 || aoat[j] > n) return ptb::fail(ctx, PTBGPU_E_ARG, "aoat out of range");
    const int kmax = (int)std::min<int64_t>(64, n);
    std::vector<int> cptr{0}, ccols, z;
    std::vector<int> first(kmax + 1, 0);
    for (int nsm = 1; nsm <= kmax; ++nsm) {
        Uniform rng(opts.kmeans_stream, opts.nstream);
        const int ifault = kmpp(xyz, (int)n, nsm, rng, z);
        if (ifault == -1)
            return ptb::fail(ctx, PTBGPU_E_ARG, "k-means stream of %lld numbers too short",
                             (long long)opts.nstream);
        if (ifault != 0)
            return ptb::fail(ctx, PTBGPU_E_KMEANS, "kmeans failed %d (nsm %d)", ifault, nsm);
        first[nsm - 1] = (int)cptr.size() - 1;
        cluster_columns(z, nsm, ndim, aoat, cptr, ccols);
    }
    const int total = (int)cptr.size() - 1;
    int *dcptr, *dccols;
    PTB_TRY(upload(ctx, "km_gptr", cptr, &dcptr));
    PTB_TRY(upload(ctx, "km_gcols", ccols, &dccols));
    auto *uni = static_cast<uint32_t *>(ptb::ws(ctx, "km_union", sizeof(uint32_t) * total * words));
    auto *dcount = static_cast<int *>(ptb::ws(ctx, "km_count", sizeof(int) * total));
    if (!uni || !dcount) return PTBGPU_E_NOMEM;
    // every candidate group belongs to the one molecule whose bitmap is `bits`
    std::vector<MolDev> km{{nullptr, nullptr, nullptr, ndim, words, 0, 0, 0, 0}};
    std::vector<int> kmol(total, 0);
    MolDev *dkm;
    int *dkmol;
    PTB_TRY(upload(ctx, "km_mols", km, &dkm));
    PTB_TRY(upload(ctx, "km_smol", kmol, &dkmol));
    PTB_TRY(ptb::group_union(ctx, bits, dkm, dkmol, words, dcptr, dccols, total, uni, dcount));
    std::vector<int> m(total);
    PTB_HIP(ctx, hipMemcpyAsync(m.data(), dcount, sizeof(int) * total, hipMemcpyDeviceToHost,
                                ctx->stream));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double best = HUGE_VAL;
    int best_nsm = 1;
    for (int nsm = 1; nsm <= kmax; ++nsm) {
        double effort = 0.0;
        for (int g = 0; g < nsm; ++g) effort += std::pow((double)m[first[nsm - 1] + g], 2.5);
        if (opts.verbose) printf(" kmeans %d %.17g\n", nsm, effort);
        if (effort < best) {
            best = effort;
            best_nsm = nsm;
        }
    }
    gptr.assign(1, 0);
    gcols.clear();
    const int g0 = first[best_nsm - 1];
    for (int g = 0; g < best_nsm; ++g) {
        gcols.insert(gcols.end(), ccols.begin() + cptr[g0 + g], ccols.begin() + cptr[g0 + g + 1]);
        gptr.push_back((int)gcols.size());
    }
    return PTBGPU_OK;
}

// One FP64 refinement sweep of the eigenvectors occupied at mu = bis[0] (the
// mixed-precision solve, whose Q and T come from a float-storage
// tridiagonalisation of K; A holds K in FP64, both triangles):
//   Zo = occupied columns of Z;  R = Q^T K Q Zo - Zo diag(theta_o);  C = Z^T R;
//   E = C / (theta_j - theta_i) on unoccupied rows;  Zo += Z E;
//   Zo <- Zo (3 I - Zo^T Zo) / 2, twice;  theta_j += C(j, j)
// and the refined columns and eigenvalues are written back into Z and eig.
// Uses W, X and R as scratch.
int refine_occupied(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                    const QFactor &qf, const double *K, double *Z, double *W, double *X, double *R,
                    double *eig, const double *bis, double *wocc, int *nocc, int *oidx,
                    double *thn, int lowp, int lowdin) {
    hipStream_t st = ctx->stream;
    const int nsm = (int)desc.size();
    int max_m = 0;
    for (const auto &d : desc) max_m = std::max(max_m, d.m);
    PTB_TRY(ptb::occ_compact(ctx, ddesc, nsm, max_m, eig, bis, Z, W, wocc, nocc, oidx));  // W = Zo
    PTB_TRY(ptb::occ_compact(ctx, ddesc, nsm, max_m, eig, bis, Z, X, wocc, nocc, nullptr));
    // FP64-reduced submatrices are not refined: no occupied columns to correct
    hipLaunchKernelGGL(k_skip_exact, dim3((nsm + 255) / 256), dim3(256), 0, st, ddesc, nsm, nocc);
    PTB_CHECK_LAUNCH(ctx);
    std::vector<int> hn(nsm);
    PTB_HIP(ctx, hipMemcpyAsync(hn.data(), nocc, sizeof(int) * nsm, hipMemcpyDeviceToHost, st));
    PTB_HIP(ctx, hipStreamSynchronize(st));
    int max_n = 0;
    std::vector<int64_t> off(nsm);
    for (int s = 0; s < nsm; ++s) {
        max_n = std::max(max_n, hn[s]);
        off[s] = desc[s].off;
    }
    if (max_n == 0) return PTBGPU_OK;
    ptb::ProfScope prof_(ctx, "refine_all");
    PTB_TRY(ptb::apply_q(ctx, desc, qf, X, off, hn, false, "r1"));  // X = Q Zo
    std::vector<GemmDesc> g(nsm);
    // prec: kLowpNone for the products whose rounding the result keeps (the
    // residual, the Gram matrix of the Loewdin step, the final rotation), the
    // low-precision arithmetic for the correction terms (their rounding
    // scales with the correction itself)
    auto run = [&](const char *name, bool ta, int prec, auto make) -> int {
        for (int s = 0; s < nsm; ++s) g[s] = make(desc[s], hn[s]);
        GemmDesc *dg;
        PTB_TRY(upload(ctx, name, g, &dg));
        ptb::TileList tl;
        tl.add(g.data(), nsm, prec == kLowpNone);
        PTB_TRY(tl.upload(ctx, (std::string(name) + "_w").c_str()));
        return tl.launch(ctx, 0, ta, false, dg, "refine", false, prec);
    };
    // R = K (Q Zo)
    PTB_TRY(run("rf_g1", false, kLowpNone, [&](const SubDesc &d, int n) {
        return GemmDesc{K + d.off, X + d.off, R + d.off, d.m, n, d.m, d.ld, d.ld, d.ld, 1.0, 0.0, 0, 0};
    }));
    PTB_TRY(ptb::apply_q(ctx, desc, qf, R, off, hn, true, "r2"));  // R = Q^T K Q Zo
    PTB_TRY(ptb::refine_resid(ctx, ddesc, nsm, max_n, nocc, oidx, eig, W, R));
    // X = Z^T R, then E in place and the Rayleigh quotients
    PTB_TRY(run("rf_g2", true, lowp, [&](const SubDesc &d, int n) {
        return GemmDesc{Z + d.off, R + d.off, X + d.off, d.m, n, d.m, d.ld, d.ld, d.ld, 1.0, 0.0, 0, 0};
    }));
    PTB_TRY(ptb::refine_corr(ctx, ddesc, nsm, max_n, nocc, oidx, eig, bis, X, thn));
    // W = Zo + Z E
    PTB_TRY(run("rf_g3", false, lowp, [&](const SubDesc &d, int n) {
        return GemmDesc{Z + d.off, X + d.off, W + d.off, d.m, n, d.m, d.ld, d.ld, d.ld, 1.0, 1.0, 0, 0};
    }));
    // Loewdin orthonormalisation by Newton-Schulz steps of the inverse
    // square root, R = (3 I - W^T W) / 2, X = W R (then the same on X).  Each
    // step squares the departure from orthonormality: one step after FP64
    // merge products (W departs by the correction's square), two after
    // low-precision ones (they leave the columns of Z orthonormal only to
    // about 1e-5).
    PTB_TRY(run("rf_g4", true, kLowpNone, [&](const SubDesc &d, int n) {
        return GemmDesc{W + d.off, W + d.off, R + d.off, n, n, d.m, d.ld, d.ld, d.ld, 1.0, 0.0, 0, 0};
    }));
    PTB_TRY(ptb::refine_shift(ctx, ddesc, nsm, max_n, nocc, R));
    PTB_TRY(run("rf_g5", false, kLowpNone, [&](const SubDesc &d, int n) {
        return GemmDesc{W + d.off, R + d.off, X + d.off, d.m, n, n, d.ld, d.ld, d.ld, 1.0, 0.0, 0, 0};
    }));
    if (lowdin < 2) return ptb::refine_scatter(ctx, ddesc, nsm, max_n, nocc, oidx, X, thn, Z, eig);
    PTB_TRY(run("rf_g6", true, kLowpNone, [&](const SubDesc &d, int n) {
        return GemmDesc{X + d.off, X + d.off, R + d.off, n, n, d.m, d.ld, d.ld, d.ld, 1.0, 0.0, 0, 0};
    }));
    PTB_TRY(ptb::refine_shift(ctx, ddesc, nsm, max_n, nocc, R));
    PTB_TRY(run("rf_g7", false, kLowpNone, [&](const SubDesc &d, int n) {
        return GemmDesc{X + d.off, R + d.off, W + d.off, d.m, n, n, d.ld, d.ld, d.ld, 1.0, 0.0, 0, 0};
    }));
    return ptb::refine_scatter(ctx, ddesc, nsm, max_n, nocc, oidx, W, thn, Z, eig);
}

// The buffers whose sizes follow the chunk plan: the per-submatrix arenas,
// the owned-row projections, the tridiagonalisation's panel buffers and the
// reflector T factors.
const char *const kChunkBufs[] = {"sm_A",      "sm_B",       "sm_Z",  "sm_W",   "sm_X",
                                  "sm_Li",     "sm_K32",     "sm_V",  "sm_R",   "sm_xa",
                                  "sm_f",      "sytrd_pw",   "sytrd_pw32", "q_T", "q_T64",
                                  "q_T128",    "q_t128_gm"};

// Device bytes the per-submatrix arenas of one chunk may take: the free
// device memory plus what this context holds in the chunk-sized buffers
// (released when a new chunked plan starts), with a reserve for everything
// else.  PTBGPU_ARENA_BUDGET_MB, when set, caps it (tests use it to force the
// chunked passes on small systems).
size_t arena_budget(ptbgpu_ctx *ctx) {
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) freeb = (size_t)64 << 30;
    size_t held = 0;
    for (const char *nm : kChunkBufs) {
        auto it = ctx->bufs.find(nm);
        if (it != ctx->bufs.end()) held += it->second.bytes;
    }
    const size_t reserve = ((size_t)4 << 30) + (size_t)(0.03 * (double)total);
    size_t budget = freeb + held > reserve ? (size_t)(0.85 * (double)(freeb + held - reserve)) : 0;
    if (const char *env = std::getenv("PTBGPU_ARENA_BUDGET_MB")) {
        const long long mb = std::atoll(env);
        if (mb > 0) budget = std::min(budget, (size_t)mb << 20);
    }
    return std::max<size_t>(budget, (size_t)64 << 20);
}

// A captured eigen phase: the graph and the host-side state its launches
// leave behind (the blocked reflector factors and the fe product's table).
struct EigenGraph {
    hipGraphExec_t exec = nullptr;
    QFactor qf;
    ptb::TileList tf;
    GemmDesc *dgf = nullptr;
    ~EigenGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

// Runs the eigen phase of a one-chunk build through a hipGraph keyed by its
// shapes and buffers: the first build with a key runs eagerly (and uploads
// every table), the second captures the launches (every table and buffer is
// then already in place, so the capture holds kernels only) and replays the
// graph, later builds replay it.  replayed = false tells the caller to run the
// launches eagerly (first sight of the key, or a capture that could not be
// made).
template <typename F>
int eigen_graph_run(ptbgpu_ctx *ctx, uint64_t key, F launches, QFactor &qf, ptb::TileList &tf,
                    GemmDesc *&dgf, bool &replayed) {
    replayed = false;
    auto it = ctx->graphs.find(key);
    if (it != ctx->graphs.end()) {
        auto *g = static_cast<EigenGraph *>(it->second.get());
        PTB_HIP(ctx, hipGraphLaunch(g->exec, ctx->stream));
        qf = g->qf;
        tf = g->tf;
        dgf = g->dgf;
        replayed = true;
        ctx->graph_note = "replayed";
        return PTBGPU_OK;
    }
    if (!ctx->graph_seen.count(key)) {
        ctx->graph_seen.insert(key);
        ctx->graph_note = "first sight of the shapes: eager";
        return PTBGPU_OK;
    }
    PTB_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    const int rc = launches();
    ctx->capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != PTBGPU_OK || ec != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        ctx->graph_seen.erase(key);
        ctx->graph_note = std::string("capture failed: ") +
                          (rc != PTBGPU_OK ? ctx->last_error : std::string(hipGetErrorString(ec)));
        ++ctx->graph_failures;
        return PTBGPU_OK;  // not capturable now: run eagerly
    }
    auto g = std::make_shared<EigenGraph>();
    const hipError_t ei = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        g->exec = nullptr;
        (void)hipGetLastError();
        ctx->graph_note = std::string("instantiation failed: ") + hipGetErrorString(ei);
        ++ctx->graph_failures;
        return PTBGPU_OK;
    }
    ctx->graph_note = "captured";
    g->qf = qf;
    g->tf = tf;
    g->dgf = dgf;
    ctx->graphs[key] = g;
    PTB_HIP(ctx, hipGraphLaunch(g->exec, ctx->stream));
    replayed = true;
    return PTBGPU_OK;
}

// One molecule of a batch: its packed H, S (filtered) and P on the device and
// the host arrays the column grouping reads.
struct MolIn {
    int64_t ndim, nelref, natoms;
    const double *dH, *dS;
    double *dP;
    const double *xyz;
    const int64_t *aoat;
};

// The density build for a batch of independent molecules.  The O(N^2) phases
// (pattern, row unions, index maps, gather, bisection, scatter) run per
// molecule on the molecule's own packed matrices; the eigendecomposition, fe
// and back-transformation phases run once over the submatrices of every
// molecule together.  Each submatrix carries its molecule index (SubDesc::mol),
// which selects the chemical potential its occupations use.  A single
// molecule is a batch of one.
int sm_density_batch_device(ptbgpu_ctx *ctx, const std::vector<MolIn> &mols, int64_t cols,
                            int64_t mode, const ptbgpu_opts &opts, ptbgpu_report *reps) {
    hipStream_t st = ctx->stream;
    Timer tm;
    tm.init(st);
    tm.mark(0, st);
    ctx->last_nsm = ctx->last_sum_m = 0;
    if (mode != 2)
        return ptb::fail(ctx, PTBGPU_E_MODE, "submatrix_mode %lld not implemented "
                         "(the reference stops, source/sm.f90:249-251,312-314)", (long long)mode);
    if (opts.precision != PTBGPU_FP64 && opts.precision != PTBGPU_MIXED &&
        opts.precision != PTBGPU_MIXED_BF16)
        return ptb::fail(ctx, PTBGPU_E_ARG, "unknown precision %lld", (long long)opts.precision);
    const int nmol = (int)mols.size();
    if (nmol < 1) return ptb::fail(ctx, PTBGPU_E_ARG, "empty batch");
    // ---- coupling pattern of every molecule (the k-means candidates are costed on it)
    std::vector<int64_t> words(nmol), bitoff(nmol + 1, 0);
    for (int i = 0; i < nmol; ++i) {
        words[i] = (mols[i].ndim + 31) / 32;
        bitoff[i + 1] = bitoff[i] + mols[i].ndim * words[i];
    }
    auto *bits = static_cast<uint32_t *>(ptb::ws(ctx, "sm_bits", sizeof(uint32_t) * bitoff[nmol]));
    if (!bits) return PTBGPU_E_NOMEM;
    // device table of the molecules (union offsets and first submatrices are
    // filled in once the groups are known)
    std::vector<MolDev> mtab(nmol);
    int64_t max_ndim = 0, max_words = 0, max_packed = 0;
    for (int i = 0; i < nmol; ++i) {
        mtab[i] = {mols[i].dH, mols[i].dS, mols[i].dP, mols[i].ndim, words[i], bitoff[i], 0, 0, 0};
        max_ndim = std::max(max_ndim, mols[i].ndim);
        max_words = std::max(max_words, words[i]);
        max_packed = std::max(max_packed, mols[i].ndim * (mols[i].ndim + 1) / 2);
    }
    MolDev *dmols;
    PTB_TRY(upload(ctx, "sm_mols", mtab, &dmols));
    PTB_TRY(ptb::pattern_bits(ctx, dmols, nmol, max_ndim, bitoff[nmol], bits));
    // ---- combination: groups of every molecule; gcols holds molecule-local
    // column indices, gptr positions in the concatenated gcols
    std::vector<int> gptr{0}, gcols, sbeg(nmol + 1, 0);
    for (int i = 0; i < nmol; ++i) {
        const MolIn &mo = mols[i];
        std::vector<int> gp, gc;
        if (cols == 4) {
            PTB_TRY(kmeans_groups(ctx, mo.ndim, mo.natoms, mo.xyz, mo.aoat, bits + bitoff[i],
                                  words[i], opts, gp, gc));
        } else if (column_groups(cols, mo.ndim, mo.aoat, gp, gc) != PTBGPU_OK) {
            return ptb::fail(ctx, PTBGPU_E_MODE, "submatrix_columns %lld not implemented",
                             (long long)cols);
        }
        const int base = (int)gcols.size();
        for (size_t g = 1; g < gp.size(); ++g) gptr.push_back(base + gp[g]);
        gcols.insert(gcols.end(), gc.begin(), gc.end());
        sbeg[i + 1] = (int)gptr.size() - 1;
    }
    const int nsm = (int)gptr.size() - 1;
    int *dgptr, *dgcols;
    PTB_TRY(upload(ctx, "sm_gptr", gptr, &dgptr));
    PTB_TRY(upload(ctx, "sm_gcols", gcols, &dgcols));
    tm.mark(1, st);
    // ---- stats: unions, dimensions
    std::vector<int64_t> unioff(nmol + 1, 0);
    std::vector<int> smol(nsm);
    for (int i = 0; i < nmol; ++i) {
        unioff[i + 1] = unioff[i] + (int64_t)(sbeg[i + 1] - sbeg[i]) * words[i];
        mtab[i].unioff = unioff[i];
        mtab[i].sbeg = sbeg[i];
        for (int s = sbeg[i]; s < sbeg[i + 1]; ++s) smol[s] = i;
    }
    int *dsmol;
    PTB_TRY(upload(ctx, "sm_mols", mtab, &dmols));
    PTB_TRY(upload(ctx, "sm_smol", smol, &dsmol));
    auto *uni = static_cast<uint32_t *>(ptb::ws(ctx, "sm_union", sizeof(uint32_t) * std::max<int64_t>(unioff[nmol], 1)));
    auto *dcount = static_cast<int *>(ptb::ws(ctx, "sm_count", sizeof(int) * nsm));
    if (!uni || !dcount) return PTBGPU_E_NOMEM;
    PTB_TRY(ptb::group_union(ctx, bits, dmols, dsmol, max_words, dgptr, dgcols, nsm, uni, dcount));
    std::vector<int> mcount(nsm);
    PTB_HIP(ctx, hipMemcpyAsync(mcount.data(), dcount, sizeof(int) * nsm, hipMemcpyDeviceToHost, st));
    PTB_HIP(ctx, hipStreamSynchronize(st));
    std::vector<SubDesc> desc(nsm);
    // a mixed build reduces the submatrices below mixed_min_m in FP64; when
    // that is all of them it is the FP64 build (same kernels, same results,
    // none of the float copies)
    bool mixed = false;
    if (opts.precision != PTBGPU_FP64)
        for (int s = 0; s < nsm && !mixed; ++s) mixed = mcount[s] >= ctx->mixed_min_m;
    std::vector<int64_t> rbeg(nmol + 1, 0);
    std::vector<int> mol_max_m(nmol, 0);
    int64_t arena = 0, rows = 0;
    int max_m = 0, max_own = 0;
    for (int i = 0; i < nmol; ++i) {
        rbeg[i] = rows;
        for (int s = sbeg[i]; s < sbeg[i + 1]; ++s) {
            SubDesc &d = desc[s];
            d.m = mcount[s];
            d.ld = (d.m + 15) & ~15;
            d.off = arena;
            d.mapoff = (int)rows;
            d.gbeg = gptr[s];
            d.gend = gptr[s + 1];
            d.flags = mixed && d.m < ctx->mixed_min_m ? kSubExact : 0;
            d.mol = i;
            arena += (int64_t)d.ld * d.m;
            rows += d.m;
            max_m = std::max(max_m, d.m);
            mol_max_m[i] = std::max(mol_max_m[i], d.m);
            max_own = std::max(max_own, d.gend - d.gbeg);
        }
    }
    rbeg[nmol] = rows;
    if (max_m > kDenseMax)
        return ptb::fail(ctx, PTBGPU_E_ARG, "submatrix dimension %d exceeds %d", max_m, kDenseMax);
    if (rows > INT32_MAX)
        return ptb::fail(ctx, PTBGPU_E_ARG, "batch too large (%lld submatrix rows)", (long long)rows);
    SubDesc *ddesc;
    PTB_TRY(upload(ctx, "sm_desc", desc, &ddesc));
    auto *dmap = static_cast<int *>(ptb::ws(ctx, "sm_map", sizeof(int) * rows));
    auto *doloc = static_cast<int *>(ptb::ws(ctx, "sm_oloc", sizeof(int) * std::max<size_t>(gcols.size(), 1)));
    if (!dmap || !doloc) return PTBGPU_E_NOMEM;
    PTB_TRY(ptb::group_map(ctx, uni, dmols, dsmol, ddesc, nsm, dmap, dgptr, dgcols, doloc));
    tm.mark(2, st);
    // ---- the numeric phases, in chunks of submatrices that fit the device
    // arenas.  One chunk (every workload up to the 512-water box) runs each
    // phase once.  With several chunks the build is done in passes, each
    // chunk recomputed in every pass (the arithmetic is deterministic, so a
    // recomputed chunk is bit-identical): pass 1 gives every submatrix's
    // eigenvalues and fe, the bisection runs over all of them; for the mixed
    // solve a second pass refines the states occupied at that mu, writes the
    // density at that mu, and the bisection runs again; the last pass writes
    // the density (skipped when the second bisection returns the same mu).
    // the BF16 matrix cores for the products whose rounding the FP64
    // refinement removes: split-BF16x3 by default, plain BF16 on request
    const int lowp = !mixed ? kLowpNone
                            : (opts.precision == PTBGPU_MIXED_BF16 ? kLowpBF16 : kLowpBF16x3);
    // bytes per arena element: A, B, Z, W, X, Li in FP64; the mixed solve's
    // float copy of K, FP64 reflectors and refinement residual
    const size_t per_elem = sizeof(double) * 6 + (mixed ? sizeof(float) + 2 * sizeof(double) : 0);
    auto plan_chunks = [&](size_t budget) {
        std::vector<int> cb{0};
        size_t used = 0;
        for (int s = 0; s < nsm; ++s) {
            // the arenas, the owned-row projections, the tridiagonalisation's
            // panel buffer (ld x 96) and the 32/64/128-reflector T factors
            const size_t c = per_elem * (size_t)desc[s].ld * desc[s].m +
                             2 * sizeof(double) * (size_t)desc[s].ld * 2 * (desc[s].gend - desc[s].gbeg) +
                             sizeof(double) * (size_t)desc[s].ld * (96 + 32 + 64 + 128);
            if (used > 0 && used + c > budget) {
                cb.push_back(s);
                used = 0;
            }
            used += c;
        }
        cb.push_back(nsm);
        return cb;
    };
    // a build of the same submatrices and precision as the last chunked one
    // keeps its plan (and the arenas it holds): the free memory the budget
    // reads moves a little from build to build, and a new plan would
    // reallocate every arena
    const char *cap = std::getenv("PTBGPU_ARENA_BUDGET_MB");
    std::vector<int64_t> pkey{mixed ? 1 : 0, cap ? (int64_t)std::atoll(cap) : -1};
    for (const auto &d : desc) pkey.push_back((int64_t)d.m);
    std::vector<int> cbeg;
    if (!ctx->chunk_cbeg.empty() && pkey == ctx->chunk_key) cbeg = ctx->chunk_cbeg;
    size_t budget = cbeg.empty() ? arena_budget(ctx) : 0;
    if (cbeg.empty()) cbeg = plan_chunks(budget);
    for (int attempt = 0; cbeg.size() > 2; ++attempt) {
        // Several chunks: the arenas then take most of the device, and the
        // budget counts the chunk-sized buffers already held as free.  When
        // the plan differs from the last chunked one (other chunk bounds or
        // precision) they are all released first; a repeated build of the
        // same plan reuses them.  The largest chunk's arenas are allocated
        // here; when the device cannot hold them (memory the budget could not
        // see), the budget shrinks by a fifth and the plan is made again.
        if (pkey != ctx->chunk_key || cbeg != ctx->chunk_cbeg) {
            for (const char *nm : kChunkBufs) ptb::ws_trim(ctx, nm, 0);
            ctx->chunk_key = pkey;
            ctx->chunk_cbeg = cbeg;
        }
        int64_t elems = 1;
        for (size_t k = 0; k + 1 < cbeg.size(); ++k) {
            int64_t e = 0;
            for (int s = cbeg[k]; s < cbeg[k + 1]; ++s) e += (int64_t)desc[s].ld * desc[s].m;
            elems = std::max(elems, e);
        }
        const size_t dbl = sizeof(double) * (size_t)elems;
        bool ok = true;
        for (const char *nm : {"sm_A", "sm_B", "sm_Z", "sm_W", "sm_X", "sm_Li"})
            ok = ok && ptb::ws(ctx, nm, dbl) != nullptr;
        if (mixed)
            ok = ok && ptb::ws(ctx, "sm_K32", sizeof(float) * (size_t)elems) != nullptr &&
                 ptb::ws(ctx, "sm_V", dbl) != nullptr && ptb::ws(ctx, "sm_R", dbl) != nullptr;
        if (ok) break;
        if (attempt == 4) return PTBGPU_E_NOMEM;
        for (const char *nm : kChunkBufs) ptb::ws_trim(ctx, nm, 0);
        ctx->chunk_key.clear();
        ctx->chunk_cbeg.clear();
        if (budget == 0) budget = arena_budget(ctx);
        budget = budget / 5 * 4;
        cbeg = plan_chunks(budget);
    }
    const int nchunk = (int)cbeg.size() - 1;
    // per-row vectors are global (indexed by mapoff): d, e, tau, fe, rotations,
    // occupation weights
    auto *vec = static_cast<double *>(ptb::ws(ctx, "sm_vec", sizeof(double) * rows * 9));
    auto *bis = static_cast<double *>(ptb::ws(ctx, "sm_bis", sizeof(double) * kBisStride * nmol));
    auto *bis1 = static_cast<double *>(ptb::ws(ctx, "sm_bis1", sizeof(double) * kBisStride * nmol));
    auto *flags = static_cast<int *>(ptb::ws(ctx, "sm_flags", sizeof(int) * nsm));
    auto *nocc = static_cast<int *>(ptb::ws(ctx, "sm_nocc", sizeof(int) * nsm));
    double *thn = nullptr;
    int *oidx = nullptr;
    if (mixed) {
        thn = static_cast<double *>(ptb::ws(ctx, "sm_thn", sizeof(double) * rows));
        oidx = static_cast<int *>(ptb::ws(ctx, "sm_oidx", sizeof(int) * rows));
        if (!thn || !oidx) return PTBGPU_E_NOMEM;
    }
    if (!vec || !bis || !bis1 || !flags || !nocc) return PTBGPU_E_NOMEM;
    double *dvec = vec, *evec = vec + rows, *tau = vec + 2 * rows, *fe = vec + 3 * rows;
    double *rot = vec + 4 * rows;   // 4 * rows
    double *wocc = vec + 8 * rows;  // rows
    PTB_HIP(ctx, hipMemsetAsync(flags, 0, sizeof(int) * nsm, st));
    std::vector<int64_t> not_spd(nmol, 0);
    int64_t not_spd_all = 0;
    // mu bisection, per molecule over its own states
    int64_t max_rows = 0;
    std::vector<int64_t> nelrefs(nmol);
    for (int i = 0; i < nmol; ++i) {
        max_rows = std::max(max_rows, rbeg[i + 1] - rbeg[i]);
        nelrefs[i] = mols[i].nelref;
    }
    int64_t *drbeg, *dnelref;
    PTB_TRY(upload(ctx, "sm_rbeg", rbeg, &drbeg));
    PTB_TRY(upload(ctx, "sm_nelref", nelrefs, &dnelref));
    auto bisect_all = [&]() -> int {
        return ptb::bisect_batch(ctx, dvec, fe, drbeg, dnelref, nmol, max_rows, bis);
    };
    // Loewdin steps of a refinement sweep: two when some float-reduced
    // submatrix took low-precision merge products, else one
    int lowdin = 1;
    for (const auto &d : desc)
        if (mixed && !(d.flags & kSubExact) && d.m <= ctx->lowp_merge_max_m) lowdin = 2;
    // the refinement sweeps correct the float-reduced submatrices
    const int sweeps = !mixed || opts.refine_steps < 0
                           ? 0
                           : (int)std::max<int64_t>(1, opts.refine_steps);

    // The arenas and tables of one chunk: submatrices [c0, c1) at offsets from
    // the chunk's arena base.
    struct Chunk {
        int c0 = 0, c1 = 0, n = 0, max_m = 0, max_own = 0;
        std::string sfx;  // suffix of the chunk's table names
        std::vector<SubDesc> desc;
        SubDesc *ddesc = nullptr;
        std::vector<int64_t> xoff;
        std::vector<int> c2;
        int64_t *dxoff = nullptr;
        double *A, *B, *Z, *W, *X, *Li, *XA, *F;
        float *K32 = nullptr;
        double *V = nullptr, *R = nullptr;
        QFactor qf;
        GemmDesc *dgf = nullptr;
        ptb::TileList tf;
    };
    auto setup_chunk = [&](int k, Chunk &c) -> int {
        c = Chunk{};
        c.c0 = cbeg[k];
        c.c1 = cbeg[k + 1];
        c.n = c.c1 - c.c0;
        c.desc.assign(desc.begin() + c.c0, desc.begin() + c.c1);
        int64_t off = 0, xtot = 0;
        c.xoff.resize(c.n);
        c.c2.resize(c.n);
        for (int i = 0; i < c.n; ++i) {
            SubDesc &d = c.desc[i];
            d.off = off;
            off += (int64_t)d.ld * d.m;
            c.max_m = std::max(c.max_m, d.m);
            const int own = d.gend - d.gbeg;
            c.max_own = std::max(c.max_own, own);
            c.xoff[i] = xtot;
            c.c2[i] = 2 * own;
            xtot += (int64_t)d.ld * 2 * own;
        }
        c.sfx = nchunk == 1 ? std::string() : "_c" + std::to_string(k);
        PTB_TRY(upload(ctx, ("sm_cdesc" + c.sfx).c_str(), c.desc, &c.ddesc));
        PTB_TRY(upload(ctx, ("sm_xoff" + c.sfx).c_str(), c.xoff, &c.dxoff));
        const size_t abytes = sizeof(double) * (size_t)std::max<int64_t>(off, 1);
        c.A = static_cast<double *>(ptb::ws(ctx, "sm_A", abytes));
        c.B = static_cast<double *>(ptb::ws(ctx, "sm_B", abytes));
        c.Z = static_cast<double *>(ptb::ws(ctx, "sm_Z", abytes));
        c.W = static_cast<double *>(ptb::ws(ctx, "sm_W", abytes));
        c.X = static_cast<double *>(ptb::ws(ctx, "sm_X", abytes));
        c.Li = static_cast<double *>(ptb::ws(ctx, "sm_Li", abytes));
        // owned-row projection buffers: XA = [L(O,:)^T, L^-1 e_O] (m x 2c),
        // F = Z^T Q^T XA (m x 2c)
        c.XA = static_cast<double *>(ptb::ws(ctx, "sm_xa", sizeof(double) * std::max<int64_t>(xtot, 1)));
        c.F = static_cast<double *>(ptb::ws(ctx, "sm_f", sizeof(double) * std::max<int64_t>(xtot, 1)));
        if (!c.A || !c.B || !c.Z || !c.W || !c.X || !c.Li || !c.XA || !c.F) return PTBGPU_E_NOMEM;
        if (mixed) {
            c.K32 = static_cast<float *>(ptb::ws(ctx, "sm_K32", sizeof(float) * (size_t)std::max<int64_t>(off, 1)));
            c.V = static_cast<double *>(ptb::ws(ctx, "sm_V", abytes));
            c.R = static_cast<double *>(ptb::ws(ctx, "sm_R", abytes));
            if (!c.K32 || !c.V || !c.R) return PTBGPU_E_NOMEM;
        }
        if (ctx->poison_arenas && !ctx->capturing) {  // all bytes 0xff: NaN as double and float
            const size_t xbytes = sizeof(double) * (size_t)std::max<int64_t>(xtot, 1);
            for (double *p : {c.A, c.B, c.Z, c.W, c.X, c.Li})
                PTB_HIP(ctx, hipMemsetAsync(p, 0xff, abytes, st));
            PTB_HIP(ctx, hipMemsetAsync(c.XA, 0xff, xbytes, st));
            PTB_HIP(ctx, hipMemsetAsync(c.F, 0xff, xbytes, st));
            if (mixed) {
                PTB_HIP(ctx, hipMemsetAsync(c.K32, 0xff, sizeof(float) * (size_t)std::max<int64_t>(off, 1), st));
                PTB_HIP(ctx, hipMemsetAsync(c.V, 0xff, abytes, st));
                PTB_HIP(ctx, hipMemsetAsync(c.R, 0xff, abytes, st));
            }
        }
        return PTBGPU_OK;
    };
    // construction, eigendecomposition and fe of a chunk: eigenvalues into
    // dvec, fe into fe (rows mapoff of each submatrix)
    // the launches of the construction, eigendecomposition and fe products
    // (no host synchronisation inside: the phase a graph can hold)
    auto eigen_launches = [&](Chunk &c, bool timed) -> int {
        const int n = c.n;
        const std::vector<SubDesc> &cd = c.desc;
        PTB_TRY(ptb::gather(ctx, dmols, c.ddesc, n, c.max_m, dmap, c.A, c.B));
        if (timed) tm.mark(3, st);
        PTB_TRY(ptb::potrf_batch(ctx, cd, c.ddesc, c.B, flags + c.c0));
        // K = L^-1 H L^-T from Li = L^-1: Z = Li H, A = Z Li^T (the lower
        // triangle, which the tridiagonalisation reads; the mixed solve, whose
        // refinement multiplies by K, mirrors it into the upper one).  A(i, j),
        // i >= j, needs Z(i, k) for k <= j only, so Z is formed on the tiles
        // on and below the diagonal (diagonal tiles whole: the second product
        // reads them whole)
        PTB_TRY(ptb::inv_lower_batch(ctx, cd, c.ddesc, c.B, c.Li, c.X));
        {
            std::vector<GemmDesc> g1(n), g2(n);
            for (int s = 0; s < n; ++s) {
                const SubDesc &d = cd[s];
                g1[s] = {c.Li + d.off, c.A + d.off, c.Z + d.off, d.m, d.m, d.m, d.ld, d.ld, d.ld, 1.0,
                         0.0, 2, kTriALower};
                // (a mixed build's float-reduced submatrices: the float copy
                // of K's lower triangle in the same epilogue)
                float *k32 = mixed && !(d.flags & kSubExact) ? c.K32 + d.off : nullptr;
                g2[s] = {c.Z + d.off, c.Li + d.off, c.A + d.off, d.m, d.m, d.m, d.ld, d.ld, d.ld, 1.0,
                         0.0, 1, kTriBUpper, k32};
            }
            GemmDesc *dg1, *dg2;
            PTB_TRY(upload(ctx, "sm_gk1", g1, &dg1));
            PTB_TRY(upload(ctx, "sm_gk2", g2, &dg2));
            ptb::TileList t1, t2;
            t1.add(g1.data(), n, true);
            t2.add(g2.data(), n, true);
            PTB_TRY(t1.upload(ctx, "sm_wk1"));
            PTB_TRY(t2.upload(ctx, "sm_wk2"));
            PTB_TRY(t1.launch(ctx, 0, false, false, dg1, "kform"));
            PTB_TRY(t2.launch(ctx, 0, false, true, dg2, "kform"));
            if (mixed) PTB_TRY(ptb::mirror_lower_batch(ctx, c.ddesc, n, c.max_m, c.A));
        }
        // mixed precision: the tridiagonalisation on float storage, its
        // reflectors in V, and K (FP64) kept in A for the refinement
        c.qf = QFactor{};
        PTB_TRY(ptb::eig_tri_batch(ctx, cd, c.ddesc, c.A, c.Z, c.W, c.X, dvec, evec, tau, rot, c.qf,
                                   c.K32, c.V, lowp, true));
        if (timed) tm.mark(4, st);
        // ---- fe from the owned rows: F = Z^T Q^T [L(O,:)^T, L^-1 e_O]
        PTB_TRY(ptb::xall_build(ctx, c.ddesc, n, c.max_own, doloc, c.B, c.Li, c.dxoff, c.XA));
        PTB_TRY(ptb::apply_q(ctx, cd, c.qf, c.XA, c.xoff, c.c2, true, "x"));    // Q^T XA
        {
            std::vector<GemmDesc> gf(n);
            for (int s = 0; s < n; ++s) {
                const SubDesc &d = cd[s];
                gf[s] = {c.Z + d.off, c.XA + c.xoff[s], c.F + c.xoff[s], d.m, c.c2[s], d.m, d.ld, d.ld,
                         d.ld, 1.0, 0.0};
            }
            PTB_TRY(upload(ctx, "sm_gf", gf, &c.dgf));
            c.tf = ptb::TileList{};
            c.tf.add(gf.data(), n, true);
            PTB_TRY(c.tf.upload(ctx, "sm_wf"));
            PTB_TRY(c.tf.launch(ctx, 0, true, false, c.dgf, "fe"));
        }
        return PTBGPU_OK;
    };
    // construction, eigendecomposition and fe of a chunk: eigenvalues into
    // dvec, fe into fe (rows mapoff of each submatrix).  With graphs on
    // (ptbgpu_set_tuning "graph"), a one-chunk build whose shapes and buffers
    // were seen before replays the launches as one captured hipGraph.
    auto eigen_chunk = [&](Chunk &c, bool timed) -> int {
        const int n = c.n;
        bool replayed = false;
        if (ctx->graph_mode && nchunk == 1 && !ctx->profiling) {
            uint64_t key = 1469598103934665603ull;
            auto mix = [&](const void *p, size_t b) {
                const unsigned char *q = static_cast<const unsigned char *>(p);
                for (size_t i = 0; i < b; ++i) key = (key ^ q[i]) * 1099511628211ull;
            };
            // the shapes name the graph's private tables; the shapes and
            // the addresses of every shared buffer key the graph itself
            mix(c.desc.data(), sizeof(SubDesc) * c.desc.size());
            const int64_t knobs[] = {(int64_t)lowp, (int64_t)mixed, ctx->dc_bucket_min,
                                     ctx->mixed_min_m, ctx->gemm_tile, ctx->dc_global_min,
                                     ctx->lowp_merge_max_m};
            mix(knobs, sizeof(knobs));
            char pfx[32];
            snprintf(pfx, sizeof(pfx), "g%016llx:", (unsigned long long)key);
            for (const auto &kv : ctx->bufs) {
                if (kv.first.compare(0, 1, "g") == 0 && kv.first.find(':') != std::string::npos)
                    continue;  // the graphs' own tables
                mix(kv.first.data(), kv.first.size());
                mix(&kv.second.ptr, sizeof(void *));
                mix(&kv.second.bytes, sizeof(size_t));
            }
            ctx->table_prefix = pfx;
            // a replay records the construction and eigen marks around the
            // graph (the eager launches record them themselves)
            if (timed) tm.mark(3, st);
            const int rc = eigen_graph_run(ctx, key, [&]() { return eigen_launches(c, false); }, c.qf,
                                           c.tf, c.dgf, replayed);
            if (rc != PTBGPU_OK || replayed) ctx->table_prefix.clear();
            PTB_TRY(rc);
            if (replayed && timed) tm.mark(4, st);
        }
        // eager launches (with graphs on, the first sight of the shapes runs
        // under the graph's table prefix, so that the capture finds them)
        const int erc = replayed ? PTBGPU_OK : eigen_launches(c, timed);
        ctx->table_prefix.clear();
        PTB_TRY(erc);
        std::vector<int> hflags(n);
        PTB_HIP(ctx, hipMemcpyAsync(hflags.data(), flags + c.c0, sizeof(int) * n,
                                    hipMemcpyDeviceToHost, st));
        PTB_HIP(ctx, hipStreamSynchronize(st));
        int bad = 0;
        for (int s = 0; s < n; ++s)
            if (hflags[s] & 1) {
                ++bad;
                c.desc[s].flags |= 1;
                desc[c.c0 + s].flags |= 1;
            }
        if (bad) {
            if (opts.strict)
                return ptb::fail(ctx, PTBGPU_E_NOT_SPD, "%d overlap submatrices are not positive definite",
                                 bad);
            PTB_TRY(upload(ctx, ("sm_cdesc" + c.sfx).c_str(), c.desc, &c.ddesc));
        }
        PTB_TRY(ptb::fe_proj(ctx, c.ddesc, n, c.max_m, c.dxoff, c.F, fe));
        if (timed) tm.mark(5, st);
        return PTBGPU_OK;
    };
    // FP64 refinement of the states occupied at mu = mu_ref[mol], then fe again
    auto refine_chunk = [&](Chunk &c, const double *mu_ref) -> int {
        for (int sweep = 0; sweep < sweeps; ++sweep)
            PTB_TRY(refine_occupied(ctx, c.desc, c.ddesc, c.qf, c.A, c.Z, c.W, c.X, c.R, dvec, mu_ref,
                                    wocc, nocc + c.c0, oidx, thn, lowp, lowdin));
        PTB_TRY(c.tf.launch(ctx, 0, true, false, c.dgf, "fe"));
        return ptb::fe_proj(ctx, c.ddesc, c.n, c.max_m, c.dxoff, c.F, fe);
    };
    // final density of a chunk: occupied columns only, Y_o = Q Z_o,
    // U_o = L^-T Y_o, owned columns scattered into P (zeroed first when zero)
    auto density_chunk = [&](Chunk &c, bool zero) -> int {
        const int n = c.n;
        PTB_TRY(ptb::occ_compact(ctx, c.ddesc, n, c.max_m, dvec, bis, c.Z, c.W, wocc, nocc + c.c0));
        std::vector<int> hnocc(n);
        PTB_HIP(ctx, hipMemcpyAsync(hnocc.data(), nocc + c.c0, sizeof(int) * n, hipMemcpyDeviceToHost, st));
        PTB_HIP(ctx, hipStreamSynchronize(st));
        int max_nocc = 0;
        std::vector<int64_t> aoff(n);
        for (int s = 0; s < n; ++s) {
            max_nocc = std::max(max_nocc, hnocc[s]);
            aoff[s] = c.desc[s].off;
        }
        PTB_TRY(ptb::apply_q(ctx, c.desc, c.qf, c.W, aoff, hnocc, false, "o"));     // W = Q Z_o
        {
            // X = L^-T Q Z_o = Li^T W
            std::vector<GemmDesc> gu(n);
            for (int s = 0; s < n; ++s) {
                const SubDesc &d = c.desc[s];
                gu[s] = {c.Li + d.off, c.W + d.off, c.X + d.off, d.m, hnocc[s], d.m, d.ld, d.ld, d.ld,
                         1.0, 0.0, 0, kTriAUpper};
            }
            GemmDesc *dgu;
            PTB_TRY(upload(ctx, "sm_gu", gu, &dgu));
            ptb::TileList tu;
            tu.add(gu.data(), n, true);
            PTB_TRY(tu.upload(ctx, "sm_wu"));
            PTB_TRY(tu.launch(ctx, 0, true, false, dgu, "u"));
        }
        return ptb::dmat_compact(ctx, dmols, nmol, max_packed, c.ddesc, n, max_nocc, c.max_own, dgcols,
                                 doloc, dmap, wocc, nocc + c.c0, c.X, zero);
    };
    auto save_mu = [&]() -> int {
        PTB_HIP(ctx, hipMemcpyAsync(bis1, bis, sizeof(double) * kBisStride * nmol,
                                    hipMemcpyDeviceToDevice, st));
        return PTBGPU_OK;
    };
    Chunk ch;
    if (nchunk == 1) {
        PTB_TRY(setup_chunk(0, ch));
        PTB_TRY(eigen_chunk(ch, true));
        PTB_TRY(bisect_all());
        if (sweeps > 0) {
            // FP64 refinement of the states occupied at this mu, then fe and
            // the bisection again with the refined eigenpairs
            PTB_TRY(save_mu());
            PTB_TRY(refine_chunk(ch, bis1));
            PTB_TRY(bisect_all());
        }
        tm.mark(6, st);
        PTB_TRY(density_chunk(ch, true));
    } else {
        tm.mark(3, st);
        for (int k = 0; k < nchunk; ++k) {
            PTB_TRY(setup_chunk(k, ch));
            PTB_TRY(eigen_chunk(ch, false));
        }
        tm.mark(4, st);
        tm.mark(5, st);
        PTB_TRY(bisect_all());
        bool final_done = false;
        if (sweeps > 0) {
            // The refinement pass also writes the density at the first mu.
            // The density reads only mu (and the refined pairs, which the
            // last pass would recompute bit for bit), so when the bisection
            // over the refined eigenvalues returns the same mu for every
            // molecule that density is final and the last pass is skipped.
            PTB_TRY(save_mu());
            for (int k = 0; k < nchunk; ++k) {
                PTB_TRY(setup_chunk(k, ch));
                PTB_TRY(eigen_chunk(ch, false));
                PTB_TRY(refine_chunk(ch, bis1));
                PTB_TRY(density_chunk(ch, k == 0));
            }
            PTB_TRY(bisect_all());
            std::vector<double> m1((size_t)kBisStride * nmol), m2((size_t)kBisStride * nmol);
            PTB_HIP(ctx, hipMemcpyAsync(m1.data(), bis1, sizeof(double) * m1.size(), hipMemcpyDeviceToHost, st));
            PTB_HIP(ctx, hipMemcpyAsync(m2.data(), bis, sizeof(double) * m2.size(), hipMemcpyDeviceToHost, st));
            PTB_HIP(ctx, hipStreamSynchronize(st));
            final_done = true;
            for (int i = 0; i < nmol; ++i)
                final_done = final_done && m1[(size_t)i * kBisStride] == m2[(size_t)i * kBisStride];
        }
        tm.mark(6, st);
        for (int k = 0; k < nchunk && !final_done; ++k) {
            PTB_TRY(setup_chunk(k, ch));
            PTB_TRY(eigen_chunk(ch, false));
            if (sweeps > 0) PTB_TRY(refine_chunk(ch, bis1));
            PTB_TRY(density_chunk(ch, k == 0));
        }
    }
    for (int s = 0; s < nsm; ++s)
        if (desc[s].flags & 1) {
            ++not_spd[desc[s].mol];
            ++not_spd_all;
        }
    tm.mark(7, st);
    std::vector<double> hb((size_t)kBisStride * nmol);
    PTB_HIP(ctx, hipMemcpyAsync(hb.data(), bis, sizeof(double) * hb.size(), hipMemcpyDeviceToHost, st));
    PTB_HIP(ctx, hipStreamSynchronize(st));
    tm.mark(8, st);
    PTB_HIP(ctx, hipStreamSynchronize(st));
    ctx->last_nsm = nsm;
    ctx->last_sum_m = rows;
    const int slots[] = {PTBGPU_T_COMBINATION, PTBGPU_T_STATS, PTBGPU_T_CONSTRUCTION, PTBGPU_T_EIGEN,
                         PTBGPU_T_FE, PTBGPU_T_MUITER, PTBGPU_T_FINAL};
    for (int i = 0; i < nmol; ++i) {
        const double *b = hb.data() + (size_t)i * kBisStride;
        ptbgpu_report &rep = reps[i];
        rep.mu = b[0];
        rep.nel = b[1];
        rep.muiter = (int64_t)b[2];
        rep.nsm = sbeg[i + 1] - sbeg[i];
        rep.sum_m = rbeg[i + 1] - rbeg[i];
        rep.max_m = mol_max_m[i];
        rep.sum_m3 = 0.0;
        for (int s = sbeg[i]; s < sbeg[i + 1]; ++s) rep.sum_m3 += (double)desc[s].m * desc[s].m * desc[s].m;
        rep.not_spd = not_spd[i];
        // the phases are shared by the batch: every report carries the batch's times
        for (int t = 0; t < 7; ++t) rep.t_phase[slots[t]] += 1e-3 * tm.between(t, t + 1);
        if (opts.verbose) {
            printf(" Stupid simple implementation of the submatrix method: %lld %lld\n", (long long)cols,
                   (long long)mode);
            printf(" number of submatrices= %lld\n average sm dim= %lld\n", (long long)rep.nsm,
                   (long long)(rep.sum_m / std::max<int64_t>(rep.nsm, 1)));
            for (int k = 0; k < rep.muiter; ++k)
                printf(" muiter %d mu %.17g %.17g of %lld\n", k + 1, b[4 + 2 * k], b[5 + 2 * k],
                       (long long)mols[i].nelref);
            printf(" iteration of chemical potential converged\n");
            fflush(stdout);
        }
    }
    return PTBGPU_OK;
}

int check_sm_args(ptbgpu_ctx *ctx, int64_t ndim, int64_t nelref, const void *H, const void *S,
                  const void *P, int64_t n, const int64_t *aoat) {
    if (!ctx) return PTBGPU_E_ARG;
    if (ndim < 1 || ndim > (int64_t)1 << 30 || !H || !S || !P || !aoat || n < 1 || nelref < 0)
        return ptb::fail(ctx, PTBGPU_E_ARG, "bad arguments to sm_density");
    for (int64_t i = 0; i < ndim; ++i)
        if (aoat[i] < 1 || aoat[i] > n)
            return ptb::fail(ctx, PTBGPU_E_ARG, "aoat(%lld) = %lld outside 1..n", (long long)(i + 1),
                             (long long)aoat[i]);
    return PTBGPU_OK;
}

}  // namespace

namespace ptb {

// Tridiagonal eigendecomposition of a batch of dense symmetric matrices (full
// storage in A, both triangles): Householder tridiagonalisation A = Q T Q^T and
// divide and conquer T = Z diag(d) Z^T.  On return Z holds the eigenvectors of
// T (column order of the eigenvalues in d, not sorted), A the Householder
// vectors and qf the blocked form of Q (apply_q multiplies by Q or Q^T).
int eig_tri_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                  double *A, double *Z, double *W, double *X, double *dvec, double *evec,
                  double *tau, double *rot, QFactor &qf, float *K32, double *V, int lowp,
                  bool k32_ready) {
    hipStream_t st = ctx->stream;
    const int nsm = (int)desc.size();
    int max_m = 0;
    size_t arena = 0;
    for (const auto &x : desc) {
        max_m = std::max(max_m, x.m);
        arena = std::max(arena, (size_t)(x.off + (int64_t)x.ld * x.m));
    }
    const size_t abytes = sizeof(double) * arena;
    if (K32 && V) {
        // mixed precision, A (K) untouched: float storage and FP64
        // reflectors promoted into V for the submatrices the refinement
        // corrects; the kSubExact ones are reduced in FP64 on a copy of K in V
        std::vector<SubDesc> lo, ex;
        int lo_m = 0, ex_m = 0;
        for (const auto &x : desc) {
            if (x.flags & kSubExact) {
                ex.push_back(x);
                ex_m = std::max(ex_m, x.m);
            } else {
                lo.push_back(x);
                lo_m = std::max(lo_m, x.m);
            }
        }
        SubDesc *dlo = const_cast<SubDesc *>(ddesc), *dex = const_cast<SubDesc *>(ddesc);
        if (!lo.empty() && !ex.empty()) {
            PTB_TRY(upload(ctx, "sm_desc_lo", lo, &dlo));
            PTB_TRY(upload(ctx, "sm_desc_ex", ex, &dex));
        }
        if (!lo.empty()) {
            const int n = (int)lo.size();
            if (!k32_ready) PTB_TRY(to_f32_batch(ctx, dlo, n, lo_m, A, K32));
            PTB_TRY(sytrd_batch_f32(ctx, lo, dlo, K32, dvec, evec, tau));
            PTB_TRY(promote_reflectors(ctx, dlo, n, lo_m, K32, V, tau));
        }
        if (!ex.empty()) {
            const int n = (int)ex.size();
            PTB_TRY(copy_lower_batch(ctx, dex, n, ex_m, A, V));
            PTB_TRY(sytrd_batch(ctx, ex, dex, V, dvec, evec, tau));
            PTB_TRY(zero_upper_batch(ctx, dex, n, ex_m, V));
        }
        PTB_TRY(larft_all(ctx, desc, ddesc, V, tau, qf));
    } else {
        PTB_TRY(sytrd_batch(ctx, desc, ddesc, A, dvec, evec, tau));
        PTB_TRY(zero_upper_batch(ctx, ddesc, nsm, max_m, A));              // A = V
        PTB_TRY(larft_all(ctx, desc, ddesc, A, tau, qf));
    }
    // divide and conquer
    std::vector<DcNode> leaves;
    std::vector<std::vector<DcNode>> levels;
    for (int s = 0; s < nsm; ++s) {
        int h;
        build_tree(s, 0, desc[s].m, leaves, levels, h);
    }
    DcNode *dleaves;
    PTB_TRY(upload(ctx, "sm_leaves", leaves, &dleaves));
    PTB_TRY(dc_leaves(ctx, ddesc, dleaves, (int)leaves.size(), dvec, evec, Z));
    for (size_t lv = 0; lv < levels.size(); ++lv) {
        // Merge products in the low-precision arithmetic only for the
        // float-reduced submatrices of up to lowp_merge_max_m rows (0 by
        // default: on MI355X the FP64 products are as fast, and they leave
        // the eigenvectors orthonormal to FP64 rounding).  The low-precision
        // products carry an error of about sqrt(n) 2^-17 per vector element
        // at a node of n rows, and the tree's levels add up; up to 1024 rows
        // that stays within what one FP64 refinement sweep removes.  The
        // FP64-reduced submatrices (kSubExact), which no refinement follows,
        // always keep FP64 products.  Per level: the nodes of large submatrices first,
        // then the others, each group largest first (so that the level's last
        // workgroups are small); the two groups' products are two launches.
        auto &nodes = levels[lv];
        auto big = [&](const DcNode &nd) {
            return lowp != kLowpNone &&
                   (desc[nd.s].m > ctx->lowp_merge_max_m || (desc[nd.s].flags & kSubExact));
        };
        std::stable_sort(nodes.begin(), nodes.end(), [&](const DcNode &a, const DcNode &b) {
            if (big(a) != big(b)) return big(a);
            return a.hi - a.lo > b.hi - b.lo;
        });
        int nbig = 0;
        for (const auto &nd : nodes) nbig += big(nd) ? 1 : 0;
        if (lowp == kLowpNone) nbig = (int)nodes.size();
        int max_n = 0;
        for (const auto &nd : nodes) max_n = std::max(max_n, nd.hi - nd.lo);
        std::string nm = "sm_level" + std::to_string(lv);
        DcNode *dnodes;
        PTB_TRY(upload(ctx, nm.c_str(), nodes, &dnodes));
        std::string gm = "sm_lgemm" + std::to_string(lv);
        // two products per node: the rows above and below its split
        auto *dg = static_cast<GemmDesc *>(ws(ctx, gm.c_str(), sizeof(GemmDesc) * 2 * nodes.size()));
        if (!dg) return PTBGPU_E_NOMEM;
        std::vector<GemmDesc> shape(2 * nodes.size(), GemmDesc{});
        for (size_t i = 0; i < nodes.size(); ++i) {
            const int n = nodes[i].hi - nodes[i].lo, n1 = nodes[i].sp - nodes[i].lo;
            shape[2 * i].M = n1;
            shape[2 * i + 1].M = n - n1;
            shape[2 * i].N = shape[2 * i + 1].N = n;
        }
        TileList tl;
        if (nbig > 0) tl.add(shape.data(), 2 * nbig, true);  // launched in FP64
        if (nbig < (int)nodes.size()) tl.add(shape.data() + 2 * nbig, 2 * ((int)nodes.size() - nbig));
        PTB_TRY(tl.upload(ctx, ("sm_lwork" + std::to_string(lv)).c_str()));
        std::vector<int> sizes(nodes.size());
        for (size_t i = 0; i < nodes.size(); ++i) sizes[i] = nodes[i].hi - nodes[i].lo;
        PTB_TRY(dc_merge_level(ctx, ddesc, dnodes, (int)nodes.size(), max_n, evec, dvec, Z,
                                    W, X, rot, dg, tl, lowp, &sizes, nbig));
    }
    return PTBGPU_OK;
}

// Full eigenvectors: eig_tri_batch, then Z <- Q Z for all columns.
int eig_sym_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                  double *A, double *Z, double *W, double *X, double *dvec, double *evec,
                  double *tau, double *rot) {
    QFactor qf;
    PTB_TRY(eig_tri_batch(ctx, desc, ddesc, A, Z, W, X, dvec, evec, tau, rot, qf));
    const int nsm = (int)desc.size();
    std::vector<int64_t> zoff(nsm);
    std::vector<int> cols(nsm);
    for (int s = 0; s < nsm; ++s) {
        zoff[s] = desc[s].off;
        cols[s] = desc[s].m;
    }
    return apply_q(ctx, desc, qf, Z, zoff, cols, false, "z");
}

}  // namespace ptb


extern "C" {

int ptbgpu_sm_density_dev(ptbgpu_ctx *ctx, int64_t ndim, int64_t nelref, const double *dH,
                          const double *dS, double *dP, int64_t n, const double *xyz,
                          const int64_t *aoat, int64_t submatrix_columns,
                          int64_t submatrix_mode, const ptbgpu_opts *opts, ptbgpu_report *rep) {
    PTB_TRY(check_sm_args(ctx, ndim, nelref, dH, dS, dP, n, aoat));
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    ptbgpu_opts o = opts ? *opts : ptbgpu_opts{PTBGPU_FP64, 0, 0, 0, nullptr, 0};
    ptbgpu_report r;
    memset(&r, 0, sizeof(r));
    auto t0 = std::chrono::steady_clock::now();
    std::vector<MolIn> mols{{ndim, nelref, n, dH, dS, dP, xyz, aoat}};
    int st = sm_density_batch_device(ctx, mols, submatrix_columns, submatrix_mode, o, &r);
    r.t_phase[PTBGPU_T_TOTAL] =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rep) *rep = r;
    return st;
}

int ptbgpu_sm_density_batch_dev(ptbgpu_ctx *ctx, int64_t nmol, const int64_t *ndim,
                                const int64_t *nelref, const double *const *dH,
                                const double *const *dS, double *const *dP, const int64_t *natoms,
                                const double *const *xyz, const int64_t *const *aoat,
                                int64_t submatrix_columns, int64_t submatrix_mode,
                                const ptbgpu_opts *opts, ptbgpu_report *reps) {
    if (!ctx) return PTBGPU_E_ARG;
    if (nmol < 1 || !ndim || !nelref || !dH || !dS || !dP || !natoms || !xyz || !aoat)
        return ptb::fail(ctx, PTBGPU_E_ARG, "bad arguments to sm_density_batch");
    std::vector<MolIn> mols(nmol);
    for (int64_t i = 0; i < nmol; ++i) {
        PTB_TRY(check_sm_args(ctx, ndim[i], nelref[i], dH[i], dS[i], dP[i], natoms[i], aoat[i]));
        mols[i] = {ndim[i], nelref[i], natoms[i], dH[i], dS[i], dP[i], xyz[i], aoat[i]};
    }
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    ptbgpu_opts o = opts ? *opts : ptbgpu_opts{PTBGPU_FP64, 0, 0, 0, nullptr, 0};
    std::vector<ptbgpu_report> r(nmol);
    memset(r.data(), 0, sizeof(ptbgpu_report) * nmol);
    auto t0 = std::chrono::steady_clock::now();
    int st = sm_density_batch_device(ctx, mols, submatrix_columns, submatrix_mode, o, r.data());
    const double tot = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (auto &x : r) x.t_phase[PTBGPU_T_TOTAL] = tot;
    if (reps) memcpy(reps, r.data(), sizeof(ptbgpu_report) * nmol);
    return st;
}

int ptbgpu_sm_density_batch(ptbgpu_ctx *ctx, int64_t nmol, const int64_t *ndim,
                            const int64_t *nelref, const double *const *H, const double *const *S,
                            double *const *P, const int64_t *natoms, const double *const *xyz,
                            const int64_t *const *aoat, int64_t submatrix_columns,
                            int64_t submatrix_mode, const ptbgpu_opts *opts, ptbgpu_report *reps) {
    if (!ctx) return PTBGPU_E_ARG;
    if (nmol < 1 || !ndim || !H || !S || !P)
        return ptb::fail(ctx, PTBGPU_E_ARG, "bad arguments to sm_density_batch");
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    std::vector<size_t> off(nmol + 1, 0);
    for (int64_t i = 0; i < nmol; ++i) {
        if (ndim[i] < 1 || ndim[i] > (int64_t)1 << 30 || !H[i] || !S[i] || !P[i])
            return ptb::fail(ctx, PTBGPU_E_ARG, "bad arguments to sm_density_batch (molecule %lld)",
                             (long long)i);
        off[i + 1] = off[i] + (size_t)(ndim[i] * (ndim[i] + 1) / 2);
    }
    const size_t bytes = sizeof(double) * off[nmol];
    auto *dH = static_cast<double *>(ptb::ws(ctx, "io_h", bytes));
    auto *dS = static_cast<double *>(ptb::ws(ctx, "io_s", bytes));
    auto *dP = static_cast<double *>(ptb::ws(ctx, "io_p", bytes));
    if (!dH || !dS || !dP) return PTBGPU_E_NOMEM;
    std::vector<const double *> ph(nmol), ps(nmol);
    std::vector<double *> pp(nmol);
    for (int64_t i = 0; i < nmol; ++i) {
        const size_t b = sizeof(double) * (off[i + 1] - off[i]);
        PTB_TRY(ptb::h2d(ctx, dH + off[i], H[i], b));
        PTB_TRY(ptb::h2d(ctx, dS + off[i], S[i], b));
        ph[i] = dH + off[i];
        ps[i] = dS + off[i];
        pp[i] = dP + off[i];
    }
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double t_up = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<ptbgpu_report> r(nmol);
    int st = ptbgpu_sm_density_batch_dev(ctx, nmol, ndim, nelref, ph.data(), ps.data(), pp.data(),
                                         natoms, xyz, aoat, submatrix_columns, submatrix_mode, opts,
                                         r.data());
    if (st == PTBGPU_OK) {
        auto t1 = std::chrono::steady_clock::now();
        for (int64_t i = 0; i < nmol; ++i)
            PTB_TRY(ptb::d2h(ctx, P[i], dP + off[i], sizeof(double) * (off[i + 1] - off[i])));
        PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const double t_down = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        const double tot = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (auto &x : r) {
            x.t_phase[PTBGPU_T_ALLOC] = t_up;
            x.t_phase[PTBGPU_T_DEALLOC] = t_down;
            x.t_phase[PTBGPU_T_TOTAL] = tot;
        }
    }
    if (reps) memcpy(reps, r.data(), sizeof(ptbgpu_report) * nmol);
    return st;
}

int ptbgpu_sm_spectra(ptbgpu_ctx *ctx, int64_t nsm, int64_t sum_m, int64_t *m, double *eig,
                      double *fe) {
    if (!ctx || nsm < 1 || sum_m < 1) return PTBGPU_E_ARG;
    auto it = ctx->bufs.find("sm_desc");
    auto iv = ctx->bufs.find("sm_vec");
    if (it == ctx->bufs.end() || iv == ctx->bufs.end() || ctx->last_nsm == 0)
        return ptb::fail(ctx, PTBGPU_E_ARG, "no completed density build on this context");
    // the descriptor table holds last_nsm entries: a different count would read
    // past it (or stop short)
    if (nsm != ctx->last_nsm || sum_m != ctx->last_sum_m)
        return ptb::fail(ctx, PTBGPU_E_ARG,
                         "nsm %lld / sum_m %lld do not match the last build (%lld / %lld)",
                         (long long)nsm, (long long)sum_m, (long long)ctx->last_nsm,
                         (long long)ctx->last_sum_m);
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<SubDesc> desc(nsm);
    PTB_HIP(ctx, hipMemcpyAsync(desc.data(), it->second.ptr, sizeof(SubDesc) * nsm,
                                hipMemcpyDeviceToHost, ctx->stream));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int64_t rows = 0;
    for (int64_t s = 0; s < nsm; ++s) rows += desc[s].m;
    if (rows != sum_m) return ptb::fail(ctx, PTBGPU_E_ARG, "sum_m does not match the last build");
    const double *vec = static_cast<const double *>(iv->second.ptr);
    if (m)
        for (int64_t s = 0; s < nsm; ++s) m[s] = desc[s].m;
    if (eig)
        PTB_HIP(ctx, hipMemcpyAsync(eig, vec, sizeof(double) * rows, hipMemcpyDeviceToHost, ctx->stream));
    if (fe)
        PTB_HIP(ctx, hipMemcpyAsync(fe, vec + 3 * rows, sizeof(double) * rows, hipMemcpyDeviceToHost,
                                    ctx->stream));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PTBGPU_OK;
}

int ptbgpu_sm_density(ptbgpu_ctx *ctx, int64_t ndim, int64_t nelref, const double *H,
                      const double *S, double *P, int64_t n, const double *xyz,
                      const int64_t *aoat, int64_t submatrix_columns, int64_t submatrix_mode,
                      const ptbgpu_opts *opts, ptbgpu_report *rep) {
    PTB_TRY(check_sm_args(ctx, ndim, nelref, H, S, P, n, aoat));
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    const size_t bytes = sizeof(double) * (size_t)(ndim * (ndim + 1) / 2);
    auto *dH = static_cast<double *>(ptb::ws(ctx, "io_h", bytes));
    auto *dS = static_cast<double *>(ptb::ws(ctx, "io_s", bytes));
    auto *dP = static_cast<double *>(ptb::ws(ctx, "io_p", bytes));
    if (!dH || !dS || !dP) return PTBGPU_E_NOMEM;
    PTB_TRY(ptb::h2d(ctx, dH, H, bytes));
    PTB_TRY(ptb::h2d(ctx, dS, S, bytes));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double t_up = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ptbgpu_report r;
    int st = ptbgpu_sm_density_dev(ctx, ndim, nelref, dH, dS, dP, n, xyz, aoat, submatrix_columns,
                                   submatrix_mode, opts, &r);
    if (st == PTBGPU_OK) {
        auto t1 = std::chrono::steady_clock::now();
        PTB_TRY(ptb::d2h(ctx, P, dP, bytes));
        r.t_phase[PTBGPU_T_DEALLOC] =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    }
    r.t_phase[PTBGPU_T_ALLOC] = t_up;
    r.t_phase[PTBGPU_T_TOTAL] =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rep) *rep = r;
    return st;
}

}  // extern "C"
