// Batched GEMM on the BF16 matrix cores with FP32 accumulation, for the
// mixed-precision density build.
//
// C = alpha op(A) op(B) + beta C over the same descriptors and tile lists as
// the FP64 kernel (GemmDesc, GemmTile: 64x64 output tiles, uplo and
// triangular-factor k ranges), with the operands stored as double or float.
// Each operand element x is split when it is staged into LDS:
//
//   x_hi = bf16(x),  x_lo = bf16(x - x_hi)          (round to nearest even)
//
// and the product is formed from NT terms per k step:
//
//   NT = 3 (split BF16x3):  a_hi b_hi + a_hi b_lo + a_lo b_hi
//   NT = 1 (BF16):          a_hi b_hi
//
// The dropped terms are below 2^-16 |a||b| for BF16x3 (x_hi + x_lo carries
// 16 significant bits) and 2^-8 for BF16.  Sums run in FP32 on the matrix
// core (v_mfma_f32_16x16x32_bf16); alpha, beta and the stored result in
// FP64 / T.
//
// Tiling: one 256-thread workgroup per 64x64 tile, four wavefronts of 32x32
// (2x2 MFMA tiles of 16x16), k slices of 32 (one MFMA depth).  MFMA operand
// layout of v_mfma_f32_16x16x32_bf16: lane l supplies A(l%16, 8(l/16)..+7)
// and B(8(l/16)..+7, l%16); result register r of lane l is D(4(l/16)+r, l%16).
// LDS holds both operands k-contiguous (As[m][k], Bs[n][k]), so a lane reads
// its eight k values as one 16-byte load; rows are 40 bf16 apart, which puts
// the 16 rows of one read on distinct banks.  The next slice is loaded into
// registers while the current one runs, as in the FP64 kernel.
#include "sm_internal.hpp"

namespace {

constexpr int kT = 256;
constexpr int kKS = 32;   // k slice (one MFMA depth)
constexpr int kLd = 40;   // LDS row stride in bf16

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
using gmem = __attribute__((address_space(1))) T;

template <bool TA, bool TB, typename T, int NT>
__global__ __launch_bounds__(kT, 2) void k_gemm_bf(const GemmDesc *__restrict__ descs,
                                                  const GemmTile *__restrict__ work) {
    const GemmTile w = work[blockIdx.x];
    if (w.g < 0) return;
    const GemmDesc g = descs[w.g];
    const int m0 = (w.t & 0xffff) * 64, n0 = (w.t >> 16) * 64;
    if (m0 >= g.M || n0 >= g.N) return;
    const gmem<T> *__restrict__ gA = (const gmem<T> *)(const void *)g.A;
    const gmem<T> *__restrict__ gB = (const gmem<T> *)(const void *)g.B;
    gmem<T> *gC = (gmem<T> *)(void *)g.C;
    int kbeg = 0, kend = g.K;
    if (g.tri & kTriALower) kend = min(kend, m0 + 64);
    if (g.tri & kTriAUpper) kbeg = max(kbeg, m0);
    if (g.tri & kTriBUpper) kend = min(kend, n0 + 64);
    if (g.tri & kTriBLower) kbeg = max(kbeg, n0);
    kbeg &= ~(kKS - 1);
    // [hi|lo] x [A|B] slices of 64 x kLd bf16; the epilogue reuses the space
    // as the 64 x 65 float result tile
    constexpr int kSlice = 64 * kLd;
    __shared__ __attribute__((aligned(16))) __bf16 smem[4 * kSlice + 64];
    __bf16 *Ah = smem, *Al = smem + kSlice, *Bh = smem + 2 * kSlice, *Bl = smem + 3 * kSlice;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int lr = lane & 15, lk = (lane >> 4) * 8;
    f4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
    T xa[8], xb[8];
    // Staging of each operand: where it is contiguous along m (n) in global
    // memory, a wavefront instruction reads 64 consecutive rows and each
    // thread takes eight consecutive k of its row (row = thread % 64, k =
    // 8 (thread / 64) + q), so the eight split values go to LDS as one
    // 16-byte store per half; where it is contiguous along k, an instruction
    // reads 32 consecutive k of two rows (k = e % 32, row = e / 32, e =
    // thread + 256 q) and each value is stored on its own (eight consecutive
    // k per thread there would spread one instruction over 16 rows).
    constexpr bool kPackA = !TA, kPackB = TB;
    auto a_at = [&](int q, int &mm, int &kk) {
        const int e = threadIdx.x + q * kT;
        if (kPackA) { mm = threadIdx.x & 63; kk = (threadIdx.x >> 6) * 8 + q; }
        else { kk = e & 31; mm = e >> 5; }
    };
    auto b_at = [&](int q, int &kk, int &nn) {
        const int e = threadIdx.x + q * kT;
        if (kPackB) { nn = threadIdx.x & 63; kk = (threadIdx.x >> 6) * 8 + q; }
        else { kk = e & 31; nn = e >> 5; }
    };
    const unsigned ulda = (unsigned)g.lda, uldb = (unsigned)g.ldb;
    // loads from clamped addresses, masked when staged (no load under a branch)
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int mm, kk;
            a_at(q, mm, kk);
            const unsigned cm = (unsigned)min(m0 + mm, g.M - 1), ck = (unsigned)min(k0 + kk, kend - 1);
            xa[q] = gA[TA ? cm * ulda + ck : ck * ulda + cm];
            int kb, nn;
            b_at(q, kb, nn);
            const unsigned cn = (unsigned)min(n0 + nn, g.N - 1), ckb = (unsigned)min(k0 + kb, kend - 1);
            xb[q] = gB[TB ? ckb * uldb + cn : cn * uldb + ckb];
        }
    };
    auto store = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) asm volatile("" : "+v"(xa[q]), "+v"(xb[q]));
        bf16x8 pah, pal, pbh, pbl;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int mm, kk, kb, nn;
            a_at(q, mm, kk);
            b_at(q, kb, nn);
            // split in double first: x - hi is exact there, and its rounding
            // to float before the second bf16 rounding loses nothing
            const double va = (m0 + mm < g.M && k0 + kk < kend) ? (double)xa[q] : 0.0;
            const double vb = (n0 + nn < g.N && k0 + kb < kend) ? (double)xb[q] : 0.0;
            const __bf16 ah = (__bf16)(float)va, bh = (__bf16)(float)vb;
            const __bf16 al = (__bf16)(float)(va - (double)(float)ah);
            const __bf16 bl = (__bf16)(float)(vb - (double)(float)bh);
            if (kPackA) {
                pah[q] = ah;
                pal[q] = al;
            } else {
                Ah[mm * kLd + kk] = ah;
                if (NT > 1) Al[mm * kLd + kk] = al;
            }
            if (kPackB) {
                pbh[q] = bh;
                pbl[q] = bl;
            } else {
                Bh[nn * kLd + kb] = bh;
                if (NT > 1) Bl[nn * kLd + kb] = bl;
            }
        }
        const int prow = threadIdx.x & 63, pkb = (threadIdx.x >> 6) * 8;
        if (kPackA) {
            *reinterpret_cast<bf16x8 *>(Ah + prow * kLd + pkb) = pah;
            if (NT > 1) *reinterpret_cast<bf16x8 *>(Al + prow * kLd + pkb) = pal;
        }
        if (kPackB) {
            *reinterpret_cast<bf16x8 *>(Bh + prow * kLd + pkb) = pbh;
            if (NT > 1) *reinterpret_cast<bf16x8 *>(Bl + prow * kLd + pkb) = pbl;
        }
    };
    if (kbeg < kend) {
        load(kbeg);
        store(kbeg);
        __syncthreads();
    }
    for (int k0 = kbeg; k0 < kend; k0 += kKS) {
        const bool more = k0 + kKS < kend;
        if (more) load(k0 + kKS);
        bf16x8 ah[2], bh[2], al[2], bl[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            ah[i] = *reinterpret_cast<const bf16x8 *>(Ah + (wm + 16 * i + lr) * kLd + lk);
            bh[i] = *reinterpret_cast<const bf16x8 *>(Bh + (wn + 16 * i + lr) * kLd + lk);
            if (NT > 1) {
                al[i] = *reinterpret_cast<const bf16x8 *>(Al + (wm + 16 * i + lr) * kLd + lk);
                bl[i] = *reinterpret_cast<const bf16x8 *>(Bl + (wn + 16 * i + lr) * kLd + lk);
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (NT > 1) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                }
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        __syncthreads();
        if (more) {
            store(k0 + kKS);
            __syncthreads();
        }
    }
    // epilogue through LDS: C read and written one column per wavefront
    float *Ct = reinterpret_cast<float *>(smem);  // Ct[col * 65 + row]
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                Ct[(wn + 16 * j + lr) * 65 + wm + 16 * i + (lane >> 4) * 4 + r] = acc[i][j][r];
    __syncthreads();
    const int gm = m0 + lane;
    for (int cc = wave; cc < 64; cc += kT / 64) {
        const int gn = n0 + cc;
        if (gm >= g.M || gn >= g.N || (g.uplo == 1 && gm < gn)) continue;
        auto *c = gC + (int64_t)gn * g.ldc + gm;
        const double v = (double)Ct[cc * 65 + lane];
        *c = (T)(g.beta == 0.0 ? g.alpha * v : g.alpha * v + g.beta * (double)*c);
    }
}

template <typename T, int NT>
void launch_t(bool ta, bool tb, dim3 grid, hipStream_t st, const GemmDesc *d, const GemmTile *w) {
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm_bf<false, false, T, NT>), grid, dim3(kT), 0, st, d, w);
    else if (ta && !tb) hipLaunchKernelGGL((k_gemm_bf<true, false, T, NT>), grid, dim3(kT), 0, st, d, w);
    else if (!ta && tb) hipLaunchKernelGGL((k_gemm_bf<false, true, T, NT>), grid, dim3(kT), 0, st, d, w);
    else hipLaunchKernelGGL((k_gemm_bf<true, true, T, NT>), grid, dim3(kT), 0, st, d, w);
}

}  // namespace

namespace ptb {

int gemm_tiles_bf(ptbgpu_ctx *ctx, bool ta, bool tb, const GemmDesc *ddesc, const GemmTile *dwork,
                  int64_t nwork, int lowp, bool f32) {
    if (nwork <= 0) return PTBGPU_OK;
    if (nwork > 0x7fffffff) return fail(ctx, PTBGPU_E_ARG, "GEMM tile list too long (%lld)", (long long)nwork);
    const dim3 grid((unsigned)nwork);
    if (lowp == kLowpBF16x3) {
        if (f32) launch_t<float, 3>(ta, tb, grid, ctx->stream, ddesc, dwork);
        else launch_t<double, 3>(ta, tb, grid, ctx->stream, ddesc, dwork);
    } else if (lowp == kLowpBF16) {
        if (f32) launch_t<float, 1>(ta, tb, grid, ctx->stream, ddesc, dwork);
        else launch_t<double, 1>(ta, tb, grid, ctx->stream, ddesc, dwork);
    } else {
        return fail(ctx, PTBGPU_E_ARG, "unknown low-precision GEMM mode %d", lowp);
    }
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

}  // namespace ptb

extern "C" int ptbgpu_gemm_batch_dev(ptbgpu_ctx *ctx, int64_t precision, int64_t transa,
                                     int64_t transb, int64_t count, const int64_t *M,
                                     const int64_t *N, const int64_t *K, const double *const *A,
                                     const int64_t *lda, const double *const *B,
                                     const int64_t *ldb, double *const *C, const int64_t *ldc,
                                     double alpha, double beta) {
    if (!ctx) return PTBGPU_E_ARG;
    if (count < 1 || count > (1 << 20) || !M || !N || !K || !A || !B || !C || !lda || !ldb || !ldc)
        return ptb::fail(ctx, PTBGPU_E_ARG, "bad batch description");
    if (precision < kLowpNone || precision > kLowpBF16)
        return ptb::fail(ctx, PTBGPU_E_ARG, "unknown GEMM precision %lld", (long long)precision);
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    const bool ta = transa != 0, tb = transb != 0;
    std::vector<GemmDesc> g((size_t)count);
    for (int64_t i = 0; i < count; ++i) {
        // op(A) is M x K, op(B) K x N; stored A is (ta ? K x M : M x K)
        const int64_t ar = ta ? K[i] : M[i], ac = ta ? M[i] : K[i];
        const int64_t br = tb ? N[i] : K[i], bc = tb ? K[i] : N[i];
        if (M[i] < 1 || N[i] < 1 || K[i] < 1 || M[i] > 32768 || N[i] > 32768 || K[i] > 32768 ||
            lda[i] < ar || ldb[i] < br || ldc[i] < M[i] || !A[i] || !B[i] || !C[i] ||
            lda[i] * ac >= ((int64_t)1 << 32) || ldb[i] * bc >= ((int64_t)1 << 32) ||
            ldc[i] * N[i] >= ((int64_t)1 << 32))
            return ptb::fail(ctx, PTBGPU_E_ARG, "product %lld: bad shape or leading dimension",
                             (long long)i);
        g[i] = GemmDesc{A[i], B[i], C[i], (int)M[i], (int)N[i], (int)K[i], (int)lda[i], (int)ldb[i],
                        (int)ldc[i], alpha, beta, 0, 0};
    }
    auto *dg = static_cast<GemmDesc *>(ptb::upload_table(ctx, "gemm_api_desc", g.data(),
                                                          sizeof(GemmDesc) * g.size()));
    if (!dg) return PTBGPU_E_NOMEM;
    ptb::TileList tl;
    tl.add(g.data(), (int)count, precision == kLowpNone);
    PTB_TRY(tl.upload(ctx, "gemm_api_tiles"));
    PTB_TRY(tl.launch(ctx, 0, ta, tb, dg, "api", false, (int)precision));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PTBGPU_OK;
}
