// Coupling pattern, submatrix index sets and the gather of dense sub-pencils.
//
// Reference (source/sm.f90:207-246): for submatrix s with owned columns comb(s,:),
//   map = rows i (ascending) with  H(i,j)>0 .or. S(i,j)>0 .or. i==j  for some owned j,
//   Hsm(a,b) = H(map(a),map(b)),  Ssm(a,b) = S(map(a),map(b)).
// Note the strictly positive test (not "nonzero").
//
// Layout: the pattern is a bitmap of the full symmetric matrix, one row of
// ceil(N/32) 32-bit words per column (17 MB at N = 11776).  Built in one pass
// over the packed-upper H and S: an entry (i,j), i<=j, that passes the test sets
// bit i of column j and bit j of column i (atomicOr).  A submatrix's row set is
// the OR of its owned columns' bitmaps; its map follows from a popcount prefix.
// Dense sub-pencils are column-major with leading dimension ld (a multiple of 4).
#include "sm_internal.hpp"

namespace {

constexpr int kThreads = 256;

// grid: (column blocks, molecule)
__global__ __launch_bounds__(kThreads) void k_pattern_bits(const MolDev *__restrict__ mols,
                                                           uint32_t *__restrict__ bits_all) {
    const MolDev mo = mols[blockIdx.y];
    const int64_t ndim = mo.ndim, words = mo.words;
    uint32_t *bits = bits_all + mo.bitoff;
    for (int64_t j = blockIdx.x; j < ndim; j += gridDim.x) {
        const int64_t c0 = j * (j + 1) / 2;
        for (int64_t i = threadIdx.x; i <= j; i += kThreads) {
            const bool on = (i == j) || (mo.H[c0 + i] > 0.0) || (mo.S[c0 + i] > 0.0);
            if (on) {
                atomicOr(&bits[j * words + (i >> 5)], 1u << (i & 31));
                if (i != j) atomicOr(&bits[i * words + (j >> 5)], 1u << (j & 31));
            }
        }
    }
}

// grid: (group, word blocks); the group's molecule gives its bitmap
__global__ __launch_bounds__(kThreads) void k_group_union(const uint32_t *__restrict__ bits_all,
                                                          const MolDev *__restrict__ mols,
                                                          const int *__restrict__ smol,
                                                          const int *__restrict__ gptr,
                                                          const int *__restrict__ gcols,
                                                          uint32_t *__restrict__ uni_all,
                                                          int *__restrict__ mcount) {
    __shared__ int red[16];
    const int s = blockIdx.x;
    const MolDev mo = mols[smol[s]];
    const int64_t words = mo.words;
    const uint32_t *bits = bits_all + mo.bitoff;
    uint32_t *uni = uni_all + mo.unioff + (int64_t)(s - mo.sbeg) * words;
    const int g0 = gptr[s], g1 = gptr[s + 1];
    int pc = 0;
    for (int64_t w = (int64_t)blockIdx.y * kThreads + threadIdx.x; w < words;
         w += (int64_t)gridDim.y * kThreads) {
        uint32_t acc = 0;
        for (int g = g0; g < g1; ++g) acc |= bits[(int64_t)gcols[g] * words + w];
        uni[w] = acc;
        pc += __popc(acc);
    }
    int t = block_sum(pc, red);
    if (threadIdx.x == 0 && t) atomicAdd(&mcount[s], t);
}

__device__ __forceinline__ const uint32_t *union_of(const uint32_t *uni_all, const MolDev &mo,
                                                    int s) {
    return uni_all + mo.unioff + (int64_t)(s - mo.sbeg) * mo.words;
}

// grid: (submatrix)
__global__ __launch_bounds__(kThreads) void k_group_map(const uint32_t *__restrict__ uni_all,
                                                        const MolDev *__restrict__ mols,
                                                        const int *__restrict__ smol,
                                                        const SubDesc *__restrict__ desc,
                                                        int *__restrict__ map) {
    __shared__ int pref[kThreads + 1];
    const int s = blockIdx.x;
    const MolDev mo = mols[smol[s]];
    const int64_t words = mo.words;
    const uint32_t *u = union_of(uni_all, mo, s);
    int *out = map + desc[s].mapoff;
    // each thread owns a contiguous chunk of words
    const int64_t per = (words + kThreads - 1) / kThreads;
    const int64_t w0 = threadIdx.x * per, w1 = min(words, w0 + per);
    int cnt = 0;
    for (int64_t w = w0; w < w1; ++w) cnt += __popc(u[w]);
    pref[threadIdx.x + 1] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        pref[0] = 0;
        for (int t = 1; t <= kThreads; ++t) pref[t] += pref[t - 1];
    }
    __syncthreads();
    int pos = pref[threadIdx.x];
    for (int64_t w = w0; w < w1; ++w) {
        uint32_t b = u[w];
        while (b) {
            int bit = __ffs(b) - 1;
            out[pos++] = (int)(w * 32 + bit);
            b &= b - 1;
        }
    }
}

// grid: (submatrix)
__global__ __launch_bounds__(kThreads) void k_owned_local(const uint32_t *__restrict__ uni_all,
                                                          const MolDev *__restrict__ mols,
                                                          const int *__restrict__ smol,
                                                          const int *__restrict__ gptr,
                                                          const int *__restrict__ gcols,
                                                          int nsm, int *__restrict__ oloc) {
    const int s = blockIdx.x;
    if (s >= nsm) return;
    const MolDev mo = mols[smol[s]];
    const uint32_t *u = union_of(uni_all, mo, s);
    for (int g = gptr[s] + threadIdx.x; g < gptr[s + 1]; g += kThreads) {
        const int c = gcols[g];
        const int64_t wc = c >> 5;
        int r = 0;
        for (int64_t w = 0; w < wc; ++w) r += __popc(u[w]);
        r += __popc(u[wc] & ((1u << (c & 31)) - 1u));
        oloc[g] = r;
    }
}

// grid: (upper 32x32 tile pairs, submatrix)
__global__ __launch_bounds__(kThreads) void k_gather(const MolDev *__restrict__ mols,
                                                     const SubDesc *__restrict__ desc,
                                                     const int *__restrict__ map,
                                                     double *__restrict__ Hd,
                                                     double *__restrict__ Sd) {
    __shared__ double th[32][33], ts[32][33];
    __shared__ int mr[32], mc[32];
    const SubDesc d = desc[blockIdx.y];
    const double *H = mols[d.mol].H, *S = mols[d.mol].S;
    const int nt = (d.m + 31) / 32;
    const int ntri = nt * (nt + 1) / 2;
    for (int t = blockIdx.x; t < ntri; t += gridDim.x) {
        // t -> (ta <= tb) in column-major order of the upper tile triangle
        int tb = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while (tb * (tb + 1) / 2 > t) --tb;
        while ((tb + 1) * (tb + 2) / 2 <= t) ++tb;
        const int ta = t - tb * (tb + 1) / 2;
        const int a0 = ta * 32, b0 = tb * 32;
        __syncthreads();
        if (threadIdx.x < 32) {
            int a = a0 + threadIdx.x;
            mr[threadIdx.x] = a < d.m ? map[d.mapoff + a] : -1;
        } else if (threadIdx.x < 64) {
            int b = b0 + threadIdx.x - 32;
            mc[threadIdx.x - 32] = b < d.m ? map[d.mapoff + b] : -1;
        }
        __syncthreads();
        const int lr = threadIdx.x & 31;
        for (int lc = threadIdx.x >> 5; lc < 32; lc += kThreads / 32) {
            const int gi = mr[lr], gj = mc[lc];
            double h = 0.0, s = 0.0;
            if (gi >= 0 && gj >= 0) {
                const int64_t lo = min(gi, gj), hi = max(gi, gj);
                const int64_t k = pk(lo, hi);
                h = H[k];
                s = S[k];
            }
            th[lr][lc] = h;
            ts[lr][lc] = s;
            const int a = a0 + lr, b = b0 + lc;
            if (a < d.m && b < d.m) {
                Hd[d.off + (int64_t)b * d.ld + a] = h;
                // S: the lower triangle only (and the diagonal tiles whole):
                // it is factored in place and read as a lower-triangular L
                if (ta == tb) Sd[d.off + (int64_t)b * d.ld + a] = s;
            }
        }
        if (ta != tb) {
            __syncthreads();
            // transposed tile: rows b0.., cols a0..
            for (int lc = threadIdx.x >> 5; lc < 32; lc += kThreads / 32) {
                const int b = b0 + lr, a = a0 + lc;
                if (a < d.m && b < d.m) {
                    Hd[d.off + (int64_t)a * d.ld + b] = th[lc][lr];
                    Sd[d.off + (int64_t)a * d.ld + b] = ts[lc][lr];
                }
            }
        }
    }
}

}  // namespace

namespace ptb {

int pattern_bits(ptbgpu_ctx *ctx, const MolDev *mols, int nmol, int64_t max_ndim,
                 int64_t total_words, uint32_t *bits) {
    ProfScope prof_(ctx, "pattern");
    PTB_HIP(ctx, hipMemsetAsync(bits, 0, sizeof(uint32_t) * (size_t)total_words, ctx->stream));
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(max_ndim, (int64_t)ctx->num_cus * 16 / nmol + 1));
    hipLaunchKernelGGL(k_pattern_bits, dim3(blocks, nmol), dim3(kThreads), 0, ctx->stream, mols,
                       bits);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int group_union(ptbgpu_ctx *ctx, const uint32_t *bits, const MolDev *mols, const int *smol,
                int64_t max_words, const int *dgptr, const int *dgcols, int nsm, uint32_t *uni,
                int *dmcount) {
    ProfScope prof_(ctx, "group_union");
    PTB_HIP(ctx, hipMemsetAsync(dmcount, 0, sizeof(int) * nsm, ctx->stream));
    int bx = (int)std::min<int64_t>((max_words + kThreads - 1) / kThreads, 64);
    hipLaunchKernelGGL(k_group_union, dim3(nsm, std::max(bx, 1)), dim3(kThreads), 0, ctx->stream,
                       bits, mols, smol, dgptr, dgcols, uni, dmcount);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int group_map(ptbgpu_ctx *ctx, const uint32_t *uni, const MolDev *mols, const int *smol,
              const SubDesc *ddesc, int nsm, int *dmap, const int *dgptr, const int *dgcols,
              int *doloc) {
    ProfScope prof_(ctx, "group_map");
    hipLaunchKernelGGL(k_group_map, dim3(nsm), dim3(kThreads), 0, ctx->stream, uni, mols, smol,
                       ddesc, dmap);
    PTB_CHECK_LAUNCH(ctx);
    hipLaunchKernelGGL(k_owned_local, dim3(nsm), dim3(kThreads), 0, ctx->stream, uni, mols, smol,
                       dgptr, dgcols, nsm, doloc);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int gather(ptbgpu_ctx *ctx, const MolDev *mols, const SubDesc *ddesc, int nsm, int max_m,
           const int *dmap, double *Hd, double *Sd) {
    ProfScope prof_(ctx, "gather");
    const int nt = (max_m + 31) / 32;
    const int ntri = nt * (nt + 1) / 2;
    hipLaunchKernelGGL(k_gather, dim3(std::max(ntri, 1), nsm), dim3(kThreads), 0, ctx->stream, mols,
                       ddesc, dmap, Hd, Sd);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

}  // namespace ptb
