// Batched dense FP64 kernels over the per-submatrix arenas.
//
// * k_gemm:     C = alpha op(A) op(B) + beta C for a variable-size batch whose
//               shapes and pointers come from device descriptors (so a kernel
//               earlier in the stream can size the next product).  64x64 (or
//               128x128) output tile per workgroup, 16-deep k slices staged in
//               LDS with the next slice prefetched into registers.
// * k_potrf_panel: one 32-column panel of the lower Cholesky S = L L^T: the
//               diagonal block factored and inverted in LDS, the panel below as
//               X = P L_kk^-T; the trailing update is a batched GEMM.  A
//               non-positive pivot sets flags bit 0 (dsygvd would report INFO>0).
// * k_diag_inv: inverses of the 64x64 diagonal blocks of L, for the blocked
//               triangular solves (one GEMM update + one GEMM with the
//               inverted block per 64 rows).
// * k_transpose, k_zero_upper, k_larft_all: helpers for K = L^-1 H L^-T and for
//               the blocked (WY) application of the Householder reflectors.
#include <algorithm>
#include <cstring>

#include "sm_internal.hpp"

namespace {

constexpr int kT = 256;
// reflectors per block of the WY application of Q.  Every reader of the
// reflector arena (T factors, Q application) reads column c only from row
// (c rounded down to a block) + 1 on, so only those rows above the unit need
// to hold zeros.
constexpr int kQBlk = 128;

typedef double d4_t __attribute__((ext_vector_type(4)));

// Operand pointers read from a descriptor in memory are generic (flat) to the
// compiler.  Flat loads also count against the LDS wait counter, so the
// k loop's wait for its LDS operands would wait for the next slice's global
// prefetch as well.  Pointers into device memory are global: say so.
template <typename T>
using gmem = __attribute__((address_space(1))) T;

// TE x TE output tile per 256-thread workgroup (TE = 64 or 128); each
// wavefront owns a TE/2 x TE/2 quarter as (TE/32)^2 blocks of
// v_mfma_f64_16x16x4.  MFMA operand layout: lane l supplies A(l%16, l/16) of
// a 16x4 slice and B(l/16, l%16) of a 4x16 slice; result value r of lane l is
// D(4r + l/16, l%16).  K advances 16 at a time (TE/16 elements of A and of B
// per thread): the next slices are loaded into registers (coalesced along the
// contiguous dimension of each operand) while the MFMAs of the current slices
// run from LDS.  128 x 128 tiles read half the operand bytes per flop of
// 64 x 64 ones (large products); both add every output element's k terms in
// the same order (4 at a time, ascending, zeros outside the k range), so the
// two tile sizes give identical results.
//
// The 64 x 64 instance is latency-bound: its speed follows the wavefronts
// resident per SIMD.  16-deep slices and 32-column epilogue passes (17 KB of
// LDS, 88 VGPRs) hold five per SIMD; on the 512-water build that is 4 % less
// GEMM time than 32-deep slices at four per SIMD, and 64-deep slices at two
// per SIMD cost 20 % more.
constexpr int kKS = 32;  // k depth up to which a product counts as one slice (k_gemm_nt1)

template <int TE>
struct GemmTileShape {
    static constexpr int KS = 16;         // k slice
    static constexpr int NL = TE * KS / 256;  // elements of A (and of B) per thread per slice
    static constexpr int BW = TE / 32;    // 16 x 16 blocks along a wavefront's quarter
    static constexpr int LD = TE + 4;     // row stride of the LDS slices
    static constexpr int CLD = TE + 1;    // column stride of the epilogue's LDS columns
    static constexpr int CW = 32;         // columns of the tile per epilogue pass
    static constexpr int SMEM = 2 * KS * LD > CW * CLD ? 2 * KS * LD : CW * CLD;
};

// Workgroups are dealt to the 8 XCDs round-robin by linear id, so the 1-D grid
// is read as (xcd, slot): XCD x runs the products x, x+8, x+16, ... tile after
// tile.  Every product's tiles share one XCD's L2, and products of any shape
// spread evenly over the XCDs (a (tile, product) grid would put tile t of every
// product on XCD t % 8, idling most of the chip for narrow products).
//
// T is the storage type of A, B and C (the descriptor's pointers are read as
// T*, leading dimensions in elements of T); products and sums are FP64 either
// way, so float storage halves the traffic of the FP32 tridiagonalisation's
// updates without changing the arithmetic.
//
// With a tile list (work != nullptr) workgroup b computes the tile work[b],
// already dealt to XCDs by the host; the grid is the list's length.
template <bool TA, bool TB, typename T, int TE>
__global__ __attribute__((amdgpu_flat_work_group_size(1, kT), amdgpu_waves_per_eu(TE == 64 ? 5 : 2)))
void k_gemm(const GemmDesc *__restrict__ descs, int count, int tiles,
            const GemmTile *__restrict__ work) {
    using S = GemmTileShape<TE>;
    constexpr int KS = S::KS, BW = S::BW, LD = S::LD, CLD = S::CLD, NL = S::NL, CW = S::CW;
    int gi, m0, n0;
    if (work) {
        const GemmTile w = work[blockIdx.x];
        if (w.g < 0) return;
        gi = w.g;
        m0 = (w.t & 0xffff) * TE;
        n0 = (w.t >> 16) * TE;
    } else {
        const int64_t lin = blockIdx.x;
        int tile;
        if (tiles < 0) {
            // fewer products than XCDs: consecutive workgroups (dealt
            // round-robin to the XCDs) take consecutive tiles of a product,
            // so that one large product runs on the whole chip
            gi = (int)(lin / -tiles);
            if (gi >= count) return;
            tile = (int)(lin % -tiles);
        } else {
            const int64_t slot = lin >> 3;
            gi = (int)((slot / tiles) * 8 + (lin & 7));
            if (gi >= count) return;
            tile = (int)(slot % tiles);
        }
        const int tM = (descs[gi].M + TE - 1) / TE, tN = (descs[gi].N + TE - 1) / TE;
        if (tile >= tM * tN) return;
        m0 = (tile % max(tM, 1)) * TE;
        n0 = (tile / max(tM, 1)) * TE;
        if (descs[gi].uplo && n0 > m0 + TE - 1) return;  // tile strictly above the diagonal
    }
    const GemmDesc g = descs[gi];
    // (a list may hold tiles of a bound on a shape the device fills in)
    if (m0 >= g.M || n0 >= g.N) return;
    const gmem<T> *__restrict__ gA = (const gmem<T> *)(const void *)g.A;
    const gmem<T> *__restrict__ gB = (const gmem<T> *)(const void *)g.B;
    gmem<T> *gC = (gmem<T> *)(void *)g.C;
    // k range of this tile where triangular factors can be nonzero, aligned
    // down to the slice depth: the other triangle of a triangular factor must
    // hold zeros (the few of its elements inside the range add nothing)
    int kbeg = 0, kend = g.K;
    if (g.tri & kTriALower) kend = min(kend, m0 + TE);
    if (g.tri & kTriAUpper) kbeg = max(kbeg, m0);
    if (g.tri & kTriBUpper) kend = min(kend, n0 + TE);
    if (g.tri & kTriBLower) kbeg = max(kbeg, n0);
    kbeg &= ~(KS - 1);
    // A and B slices; after the k loop the same LDS holds 64 columns of the
    // result tile at a time (column-major, leading dimension TE + 1) for the
    // epilogue
    __shared__ double smem[S::SMEM];
    double (*As)[LD] = reinterpret_cast<double (*)[LD]>(smem);
    double (*Bs)[LD] = reinterpret_cast<double (*)[LD]>(smem + KS * LD);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = (wave & 1) * (TE / 2), wn = (wave >> 1) * (TE / 2);
    const int lr = lane & 15, lk = lane >> 4;
    d4_t acc[BW][BW];
#pragma unroll
    for (int i = 0; i < BW; ++i)
#pragma unroll
        for (int j = 0; j < BW; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
    T xa[NL], xb[NL];  // the next slices' raw elements (clamped addresses)
    // element q of this thread's share: A_op(mm, kk) and B_op(kk, nn)
    auto a_index = [&](int q, int &mm, int &kk) {
        const int e = threadIdx.x + q * kT;
        if (!TA) { mm = e & (TE - 1); kk = e / TE; } else { kk = e & (KS - 1); mm = e / KS; }
    };
    auto b_index = [&](int q, int &kk, int &nn) {
        const int e = threadIdx.x + q * kT;
        if (!TB) { kk = e & (KS - 1); nn = e / KS; } else { nn = e & (TE - 1); kk = e / TE; }
    };
    // element offsets are 32-bit (every operand block of the batch is below
    // 2^32 elements from its descriptor pointer), which keeps the 16 addresses
    // of a slice in half the registers of 64-bit pointers
    // Every load is issued, from an address clamped into the operand block;
    // elements outside it are zeroed when the slice is stored to LDS.  The
    // empty asm statement there consumes the raw values unconditionally, so
    // that no load can be moved under the range test: a guarded load waits
    // for memory inside its branch, one element at a time.  Callers ensure
    // kend > kbeg >= 0.
    const unsigned ulda = (unsigned)g.lda, uldb = (unsigned)g.ldb;
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            int mm, kk;
            a_index(q, mm, kk);
            const unsigned cm = (unsigned)min(m0 + mm, g.M - 1), ck = (unsigned)min(k0 + kk, kend - 1);
            xa[q] = gA[TA ? cm * ulda + ck : ck * ulda + cm];
            int kb, nn;
            b_index(q, kb, nn);
            const unsigned cn = (unsigned)min(n0 + nn, g.N - 1), ckb = (unsigned)min(k0 + kb, kend - 1);
            xb[q] = gB[TB ? ckb * uldb + cn : cn * uldb + ckb];
        }
    };
    auto store = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NL; ++q) asm volatile("" : "+v"(xa[q]), "+v"(xb[q]));
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            int mm, kk, kb, nn;
            a_index(q, mm, kk);
            b_index(q, kb, nn);
            As[kk][mm] = (m0 + mm < g.M && k0 + kk < kend) ? (double)xa[q] : 0.0;
            Bs[kb][nn] = (n0 + nn < g.N && k0 + kb < kend) ? (double)xb[q] : 0.0;
        }
    };
    if (kbeg < kend) {
        load(kbeg);
        store(kbeg);
        __syncthreads();
    }
    for (int k0 = kbeg; k0 < kend; k0 += KS) {
        const bool more = k0 + KS < kend;
        if (more) load(k0 + KS);
#pragma unroll
        for (int ks = 0; ks < KS / 4; ++ks) {
            const int kk = ks * 4 + lk;
            double a[BW], b[BW];
#pragma unroll
            for (int i = 0; i < BW; ++i) {
                a[i] = As[kk][wm + 16 * i + lr];
                b[i] = Bs[kk][wn + 16 * i + lr];
            }
#pragma unroll
            for (int i = 0; i < BW; ++i)
#pragma unroll
                for (int j = 0; j < BW; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
        if (more) {
            store(k0 + KS);
            __syncthreads();
        }
    }
    // epilogue, CW columns of the tile at a time: they go through LDS so that
    // C is read and written 64 consecutive rows per wavefront instruction; the
    // k loop ended with a barrier, so the slices' LDS is free.  uplo = 2
    // writes the lower 64 x 64 blocks whole at either tile size.
    double *Ct = smem;  // Ct[col * CLD + row]
#pragma unroll
    for (int h = 0; h < TE / CW; ++h) {
        if (h > 0) __syncthreads();  // the previous columns have been read
#pragma unroll
        for (int j = 0; j < BW; ++j) {
            if ((wn + 16 * j) / CW != h) continue;  // (uniform over the wavefront)
#pragma unroll
            for (int i = 0; i < BW; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Ct[(wn - CW * h + 16 * j + lr) * CLD + wm + 16 * i + 4 * r + lk] = acc[i][j][r];
        }
        __syncthreads();
        for (int cc = wave; cc < CW; cc += kT / 64) {
            const int gn = n0 + CW * h + cc;
#pragma unroll
            for (int rr = 0; rr < TE; rr += 64) {
                const int gm = m0 + rr + lane;
                if (gm >= g.M || gn >= g.N || (g.uplo == 1 && gm < gn) ||
                    (g.uplo == 2 && (gm >> 6) < (gn >> 6)))
                    continue;
                auto *c = gC + (int64_t)gn * g.ldc + gm;
                const double v = Ct[cc * CLD + rr + lane];
                const T r = (T)(g.beta == 0.0 ? g.alpha * v : g.alpha * v + g.beta * (double)*c);
                *c = r;
                if (g.C32) ((gmem<float> *)(void *)g.C32)[(int64_t)gn * g.ldc + gm] = (float)r;
            }
        }
    }
}

// Products of one k slice in the A B^T form (K <= 32, no triangular factor),
// such as the rank-32 symmetric updates C -= [W V] [V W]^T: the C tile is
// read into registers together with the A and B slices, so that a tile waits
// on memory once before its MFMAs instead of a second time in the epilogue.
// Tiles come from a tile list; layouts as in k_gemm.
//
// The slices stay in LDS in the storage type (float rows padded to 80
// elements: the four k rows a wavefront reads at once fall on disjoint
// banks), and the result goes out in two passes of 32 columns: 20 KB of LDS
// with float storage, so that up to eight workgroups share a CU (these
// updates are bound by the trailing matrix's traffic).  KH < 32 stages the
// slices in parts of KH k rows (less LDS); with FP64 storage that keeps the
// later parts' operands live during the first part's products, which spills
// registers at five wavefronts per SIMD, so FP64 slices go whole (four per
// SIMD, LDS-bound).
template <typename T>
struct Nt1Shape {
    static constexpr int LD = sizeof(T) == 4 ? 80 : 68;  // slice row stride (elements of T)
    static constexpr int KH = kKS;  // k rows staged at a time
    static constexpr int SLICE = 2 * KH * LD * (int)sizeof(T);
    static constexpr int EPI = 32 * 65 * (int)sizeof(double);
    static constexpr int BYTES = SLICE > EPI ? SLICE : EPI;
};

template <typename T>
__global__ __launch_bounds__(kT, sizeof(T) == 4 ? 6 : 4) void k_gemm_nt1(
    const GemmDesc *__restrict__ descs, const GemmTile *__restrict__ work) {
    using S = Nt1Shape<T>;
    constexpr int LD = S::LD, KH = S::KH;
    constexpr int NQ = 8 * KH / kKS;  // elements per thread of each staged part
    const GemmTile wt = work[blockIdx.x];
    if (wt.g < 0) return;
    const GemmDesc g = descs[wt.g];
    const int m0 = (wt.t & 0xffff) * 64, n0 = (wt.t >> 16) * 64;
    if (m0 >= g.M || n0 >= g.N) return;
    const gmem<T> *__restrict__ gA = (const gmem<T> *)(const void *)g.A;
    const gmem<T> *__restrict__ gB = (const gmem<T> *)(const void *)g.B;
    gmem<T> *gC = (gmem<T> *)(void *)g.C;
    __shared__ double smem[S::BYTES / sizeof(double)];
    T (*As)[LD] = reinterpret_cast<T (*)[LD]>(smem);
    T (*Bs)[LD] = reinterpret_cast<T (*)[LD]>(reinterpret_cast<T *>(smem) + KH * LD);
    // wave-uniform wavefront index: the C column offsets stay scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    const int gm = m0 + lane;
    const unsigned uldc = (unsigned)g.ldc;
    // All loads are issued from addresses clamped into their blocks, with
    // out-of-range elements zeroed afterwards (no load waits under a branch).
    // C: wavefront w holds columns n0 + w + 4 t, lane = row; read only when
    // beta != 0 (a uniform branch).
    const unsigned cmr = (unsigned)min(gm, g.M - 1);
    T craw[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) craw[t] = T(0);
    if (g.beta != 0.0) {
#pragma unroll
        for (int t = 0; t < 16; ++t)
            craw[t] = gC[(unsigned)min(n0 + wave + 4 * t, g.N - 1) * uldc + cmr];
    }
    // A (64 x K) and B^T (K x 64) slices: element q of this thread is row
    // mm = e & 63, depth kk = e >> 6 (e = thread + 256 q) of both
    T xa[8], xb[8];
    const unsigned ulda = (unsigned)g.lda, uldb = (unsigned)g.ldb;
    const int kmax = max(g.K - 1, 0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = threadIdx.x + q * kT, mm = e & 63;
        const unsigned kk = (unsigned)min(e >> 6, kmax);
        xa[q] = gA[kk * ulda + (unsigned)min(m0 + mm, g.M - 1)];
        xb[q] = gB[kk * uldb + (unsigned)min(n0 + mm, g.N - 1)];
    }
    // consume the raw values unconditionally (see k_gemm), then mask
#pragma unroll
    for (int q = 0; q < 8; ++q) asm volatile("" : "+v"(xa[q]), "+v"(xb[q]));
#pragma unroll
    for (int t = 0; t < 16; ++t) asm volatile("" : "+v"(craw[t]));
    d4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
    // k rows [KH h, KH h + KH) in part h (elements q of this thread have
    // kk = thread / 64 + 4 q, so part h holds q in [NQ h, NQ h + NQ))
#pragma unroll
    for (int h = 0; h < kKS / KH; ++h) {
        if (h > 0) __syncthreads();  // the previous part is read
#pragma unroll
        for (int q = NQ * h; q < NQ * h + NQ; ++q) {
            const int e = threadIdx.x + q * kT, mm = e & 63, kk = e >> 6;
            const bool kin = kk < g.K;
            As[kk - KH * h][mm] = (kin && m0 + mm < g.M) ? xa[q] : T(0);
            Bs[kk - KH * h][mm] = (kin && n0 + mm < g.N) ? xb[q] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KH / 4; ++ks) {
            const int kk = ks * 4 + lk;
            const double a0 = As[kk][wm + lr], a1 = As[kk][wm + 16 + lr];
            const double b0 = Bs[kk][wn + lr], b1 = Bs[kk][wn + 16 + lr];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
    }
    // columns 32 h .. 32 h + 31 per pass, from the two wavefronts that own them
    double *Ct = smem;  // Ct[(col - 32 h) * 65 + row]
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        __syncthreads();  // the slices (h = 0) or the previous columns (h = 1) are read
        if (wn == 32 * h) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        Ct[(16 * j + lr) * 65 + wm + 16 * i + 4 * r + lk] = acc[i][j][r];
        }
        __syncthreads();
#pragma unroll
        for (int t = 8 * h; t < 8 * h + 8; ++t) {
            const int cc = wave + 4 * t, gn = n0 + cc;
            if (gm >= g.M || gn >= g.N || (g.uplo == 1 && gm < gn)) continue;
            const double v = g.alpha * Ct[(cc - 32 * h) * 65 + lane];
            gC[(unsigned)gn * uldc + gm] = (T)(g.beta == 0.0 ? v : v + g.beta * (double)craw[t]);
        }
    }
}

// One 32-column panel of the lower Cholesky factor of every submatrix: factor
// the diagonal block in LDS, invert it, and form the panel below as
// X = P L_kk^-T (rows independent).  The trailing update that follows is a
// batched GEMM.  A non-positive pivot sets flags bit 0 (dsygvd: INFO > 0).
//
// mode 0: the whole panel (grid (submatrix, 1)).  For few submatrices the
// panel is split: mode 1 forms the rows below the diagonal block, sliced over
// grid.y (each workgroup factors the block in LDS for itself), then mode 2
// writes the factored diagonal block and the flag (the rows read the block
// unfactored, so they go first).
__global__ __launch_bounds__(kT) void k_potrf_panel(const SubDesc *__restrict__ desc,
                                                    double *__restrict__ Aa,
                                                    int *__restrict__ flags, int panel, int mode) {
    const SubDesc d = desc[blockIdx.x];
    const int m = d.m, ld = d.ld;
    const int k0 = panel * 32;
    if (k0 >= m) return;
    double *a = Aa + d.off;
    const int kb = min(32, m - k0);
    __shared__ double Lk[32][33], Li[32][33];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    for (int l = threadIdx.x; l < 1024; l += kT) {
        const int r = l & 31, c = l >> 5;
        double v;
        if (r < kb && c < kb) v = r >= c ? a[(int64_t)(k0 + c) * ld + k0 + r] : 0.0;
        else v = (r == c) ? 1.0 : 0.0;  // identity padding
        Lk[r][c] = v;
        Li[r][c] = 0.0;
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
        if ((int)threadIdx.x > j && (int)threadIdx.x < kb) Lk[threadIdx.x][j] /= Lk[j][j];
        __syncthreads();
        for (int l = threadIdx.x; l < 1024; l += kT) {
            const int r = l & 31, c = l >> 5;
            if (c > j && r >= c && r < kb) Lk[r][c] -= Lk[r][j] * Lk[c][j];
        }
        __syncthreads();
    }
    if (mode != 1)
        for (int l = threadIdx.x; l < 1024; l += kT) {
            const int r = l & 31, c = l >> 5;
            if (r < kb && c < kb && r >= c) a[(int64_t)(k0 + c) * ld + k0 + r] = Lk[r][c];
        }
    if (mode == 2) {
        if (threadIdx.x == 0 && bad) flags[blockIdx.x] |= 1;
        return;
    }
    // Li = Lk^-1, column c by thread c (forward substitution)
    if (threadIdx.x < 32) {
        const int c = threadIdx.x;
        for (int r = c; r < 32; ++r) {
            double sacc = (r == c) ? 1.0 : 0.0;
            for (int t = c; t < r; ++t) sacc -= Lk[r][t] * Li[t][c];
            Li[r][c] = sacc / Lk[r][r];
        }
    }
    __syncthreads();
    // X(i, c) = sum_{t<=c} P(i, t) Li(c, t)
    for (int i = k0 + kb + blockIdx.y * kT + threadIdx.x; i < m; i += gridDim.y * kT) {
        double p[32];
#pragma unroll
        for (int t = 0; t < 32; ++t) p[t] = t < kb ? a[(int64_t)(k0 + t) * ld + i] : 0.0;
        for (int c = 0; c < kb; ++c) {
            double sacc = 0.0;
#pragma unroll
            for (int t = 0; t < 32; ++t) sacc += p[t] * Li[c][t];
            a[(int64_t)(k0 + c) * ld + i] = sacc;
        }
    }
    if (mode == 0 && threadIdx.x == 0 && bad) flags[blockIdx.x] |= 1;
}

// L = I (lower triangle) for the submatrices whose Cholesky met a
// non-positive pivot: the factorisation continued past it and its values can
// overflow, so the products downstream (whose results the caller discards for
// these submatrices) get finite operands.  One workgroup per submatrix (the
// unflagged ones return at once).
__global__ __launch_bounds__(kT) void k_identity_if_bad(const SubDesc *__restrict__ desc,
                                                        const int *__restrict__ flags,
                                                        double *__restrict__ Aa) {
    if (!(flags[blockIdx.x] & 1)) return;
    const SubDesc d = desc[blockIdx.x];
    for (int c = 0; c < d.m; ++c) {
        double *col = Aa + d.off + (int64_t)c * d.ld;
        for (int r = c + threadIdx.x; r < d.m; r += kT) col[r] = r == c ? 1.0 : 0.0;
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

// A(r, c) = A(c, r) for r < c, in place (the lower triangle is read, the
// strict upper one written).  grid: (32x32 tile, submatrix)
__global__ __launch_bounds__(kT) void k_mirror_lower(const SubDesc *__restrict__ desc,
                                                     double *__restrict__ Aa) {
    const SubDesc d = desc[blockIdx.y];
    const int nt = (d.m + 31) / 32;
    if ((int)blockIdx.x >= nt * nt) return;
    const int r0 = (blockIdx.x % nt) * 32, c0 = (blockIdx.x / nt) * 32;
    if (r0 > c0) return;  // tiles on or above the diagonal receive
    double *a = Aa + d.off;
    __shared__ double t[32][33];
    for (int l = threadIdx.x; l < 1024; l += kT) {  // t[r][c] = A(c0 + r, r0 + c)
        const int r = l & 31, c = l >> 5;
        if (c0 + r < d.m && r0 + c < d.m) t[r][c] = a[(int64_t)(r0 + c) * d.ld + c0 + r];
    }
    __syncthreads();
    for (int l = threadIdx.x; l < 1024; l += kT) {  // A(r0 + r, c0 + c) = A(c0 + c, r0 + r)
        const int r = l & 31, c = l >> 5;
        if (r0 + r < c0 + c && r0 + r < d.m && c0 + c < d.m)
            a[(int64_t)(c0 + c) * d.ld + r0 + r] = t[c][r];
    }
}

__global__ __launch_bounds__(kT) void k_zero_upper(const SubDesc *__restrict__ desc,
                                                   double *__restrict__ Aa) {
    const SubDesc d = desc[blockIdx.x];
    const int c = blockIdx.y;
    if (c >= d.m) return;
    double *col = Aa + d.off + (int64_t)c * d.ld;
    // rows above the unit that the block readers see (see kQBlk)
    for (int r = (c & ~(kQBlk - 1)) + 1 + threadIdx.x; r <= c; r += kT) col[r] = 0.0;
}

// Float copy of the lower triangle of every arena block (column c of
// submatrix s per workgroup row): the input of the FP32 tridiagonalisation.
// grid: (submatrix, column)
__global__ __launch_bounds__(kT) void k_to_f32(const SubDesc *__restrict__ desc,
                                               const double *__restrict__ Aa,
                                               float *__restrict__ Fa) {
    const SubDesc d = desc[blockIdx.x];
    const int c = blockIdx.y;
    if (c >= d.m) return;
    const double *src = Aa + d.off + (int64_t)c * d.ld;
    float *dst = Fa + d.off + (int64_t)c * d.ld;
    // the lower triangle only: the float tridiagonalisation never reads above it
    for (int r = c + threadIdx.x; r < d.m; r += kT) dst[r] = (float)src[r];
}

// FP64 copy of the lower triangle (the mixed solve's FP64-reduced
// submatrices: K stays in A for the refinement, the reduction runs on the copy)
__global__ __launch_bounds__(kT) void k_copy_lower(const SubDesc *__restrict__ desc,
                                                   const double *__restrict__ Aa,
                                                   double *__restrict__ Ba) {
    const SubDesc d = desc[blockIdx.x];
    const int c = blockIdx.y;
    if (c >= d.m) return;
    const double *src = Aa + d.off + (int64_t)c * d.ld;
    double *dst = Ba + d.off + (int64_t)c * d.ld;
    for (int r = c + threadIdx.x; r < d.m; r += kT) dst[r] = src[r];
}

// Reflectors of the FP32 tridiagonalisation in FP64: V(:, c) = the stored
// v_c (rows > c, unit at c+1), zero above, and tau_c = 2 / (v_c^T v_c) so that
// every H_c = I - tau_c v_c v_c^T is orthogonal to FP64 rounding (tau_c = 0,
// no reflection, stays 0).  One wavefront per column.
// grid: (submatrix, ceil(max_m / 4)), 256 threads
__global__ __launch_bounds__(kT) void k_promote_reflectors(const SubDesc *__restrict__ desc,
                                                           const float *__restrict__ Fa,
                                                           double *__restrict__ Va,
                                                           double *__restrict__ tau_all) {
    const SubDesc d = desc[blockIdx.x];
    const int c = blockIdx.y * (kT / PTB_WAVE) + threadIdx.x / PTB_WAVE;
    const int lane = threadIdx.x % PTB_WAVE;
    if (c >= d.m) return;
    const float *src = Fa + d.off + (int64_t)c * d.ld;
    double *dst = Va + d.off + (int64_t)c * d.ld;
    double ss = 0.0;
    // zeros only in the rows above the unit that the block readers see (see kQBlk)
    for (int r = (c & ~(kQBlk - 1)) + 1 + lane; r < d.m; r += PTB_WAVE) {
        const double v = r > c ? (double)src[r] : 0.0;
        dst[r] = v;
        ss += v * v;
    }
    ss = wave_sum(ss);
    if (lane == 0 && c < d.m - 2) {
        double *t = tau_all + d.mapoff + c;
        if (*t != 0.0) *t = 2.0 / ss;
    }
}

// T factor of every reflector block of every submatrix in one launch (grid:
// block b counted from the first, submatrix).  Q_b = H_b0 ... H_b0+kb-1 =
// I - V_b T_b V_b^T with T_b upper triangular (LAPACK dlarft, forward,
// columnwise).  G = V_b^T V_b is accumulated over 64-row chunks of the block
// staged in LDS (each column read once); the recurrence
// T(r, i) = -tau_i sum_{c=r}^{i-1} T(r, c) G(i, c) runs one column i at a time
// with the rows r < i in parallel.  T_b is stored 32 x 32 column-major at
// Ta + toff[s] + b * 1024.
__global__ __launch_bounds__(kT) void k_larft_all(const SubDesc *__restrict__ desc,
                                                  const double *__restrict__ Va,
                                                  const double *__restrict__ tau_all,
                                                  const int64_t *__restrict__ toff,
                                                  double *__restrict__ Ta) {
    const SubDesc d = desc[blockIdx.x];
    const int nr = d.m - 2;
    const int nblk = nr > 0 ? (nr + 31) / 32 : 0;
    const int b = blockIdx.y;
    if (b >= nblk) return;
    const int b0 = b * 32;
    const int kb = min(32, nr - b0);
    const double *V = Va + d.off + (int64_t)b0 * d.ld;
    const double *tau = tau_all + d.mapoff + b0;
    double *T = Ta + toff[blockIdx.x] + (int64_t)b * 1024;
    __shared__ double Vs[64][33];
    __shared__ double G[32][33];
    __shared__ double Ts[32][33];
    // G = V^T V on the matrix cores: wavefront w owns the 16 x 16 tile
    // (rows 16 (w & 1), columns 16 (w >> 1)) as v_mfma_f64_16x16x4 over the
    // block's rows, 64 at a time through LDS (lane l supplies V(k0 + l/16,
    // i0 + l%16) and V(k0 + l/16, j0 + l%16); result value r of lane l is
    // G(i0 + 4r + l/16, j0 + l%16)).  Column c is zero above its unit at row
    // b0 + c + 1, so every element runs over rows b0 + 1 .. m - 1.  The next
    // 64 rows are loaded into registers while the current ones run.
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = (wave & 1) * 16, j0 = (wave >> 1) * 16;
    const int lr = lane & 15, lk = lane >> 4;
    d4_t acc = d4_t{0.0, 0.0, 0.0, 0.0};
    double x[64 * 32 / kT];
    // every load issued from a clamped address, then masked when staged (the
    // empty asm keeps the loads out of the range tests)
    auto load = [&](int r0) {
#pragma unroll
        for (int u = 0; u < 64 * 32 / kT; ++u) {
            const int l = threadIdx.x + u * kT, rr = l & 63, cc = l >> 6;
            x[u] = V[(int64_t)min(cc, kb - 1) * d.ld + min(r0 + rr, d.m - 1)];
        }
    };
    auto stage = [&](int r0) {
#pragma unroll
        for (int u = 0; u < 64 * 32 / kT; ++u) asm volatile("" : "+v"(x[u]));
#pragma unroll
        for (int u = 0; u < 64 * 32 / kT; ++u) {
            const int l = threadIdx.x + u * kT, rr = l & 63, cc = l >> 6;
            Vs[rr][cc] = (r0 + rr < d.m && cc < kb) ? x[u] : 0.0;
        }
    };
    if (b0 + 1 < d.m) {
        load(b0 + 1);
        stage(b0 + 1);
    }
    __syncthreads();
    for (int r0 = b0 + 1; r0 < d.m; r0 += 64) {
        const bool more = r0 + 64 < d.m;
        if (more) load(r0 + 64);
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) {
            const int kk = 4 * ks + lk;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[kk][i0 + lr], Vs[kk][j0 + lr], acc, 0, 0, 0);
        }
        __syncthreads();
        if (more) {
            stage(r0 + 64);
            __syncthreads();
        }
    }
    for (int l = threadIdx.x; l < 1024; l += kT) Ts[l & 31][l >> 5] = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) G[i0 + 4 * r + lk][j0 + lr] = acc[r];
    __syncthreads();
    for (int i = 0; i < kb; ++i) {
        const int r = threadIdx.x;
        if (r == i) Ts[i][i] = tau[i];
        if (r < i) {
            double sacc = 0.0;
            for (int c = r; c < i; ++c) sacc += Ts[r][c] * G[i][c];
            Ts[r][i] = -tau[i] * sacc;
        }
        __syncthreads();
    }
    for (int l = threadIdx.x; l < 1024; l += kT) {
        const int r = l & 31, c = l >> 5;
        T[(int64_t)c * 32 + r] = Ts[r][c];
    }
}

// T factor of every 64-reflector block from the factors of its two
// 32-reflector halves (the second may be absent):
//   (I - V1 T1 V1^T)(I - V2 T2 V2^T) = I - [V1 V2] [T1 T12; 0 T2] [V1 V2]^T,
//   T12 = -T1 (V1^T V2) T2.
// V1^T V2 runs over the rows where V2 is nonzero (below its first unit),
// staged through LDS 64 rows at a time.  T64 is 64 x 64 column-major at
// T64a + toff64[s] + p * 4096.  grid: (submatrix, pair)
__global__ __launch_bounds__(kT) void k_larft_merge(const SubDesc *__restrict__ desc,
                                                    const double *__restrict__ Va,
                                                    const double *__restrict__ Ta,
                                                    const int64_t *__restrict__ toff,
                                                    const int64_t *__restrict__ toff64,
                                                    double *__restrict__ T64a) {
    const SubDesc d = desc[blockIdx.x];
    const int nr = d.m - 2;
    const int nblk = nr > 0 ? (nr + 31) / 32 : 0;
    const int p = blockIdx.y, b = 2 * p;
    if (b >= nblk) return;
    const bool two = b + 1 < nblk;
    __shared__ double V1s[64][33], V2s[64][33];
    __shared__ double T1[32][33], T2[32][33], G[32][33];
    const double *T1g = Ta + toff[blockIdx.x] + (int64_t)b * 1024;
    for (int l = threadIdx.x; l < 1024; l += kT) {
        const int r = l & 31, c = l >> 5;
        T1[r][c] = T1g[c * 32 + r];
        T2[r][c] = two ? Ta[toff[blockIdx.x] + (int64_t)(b + 1) * 1024 + c * 32 + r] : 0.0;
    }
    // G = V1^T V2 on the matrix cores, 64 rows at a time through LDS with the
    // next rows' loads in flight (tiles and lane layout as in k_larft_all)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = (wave & 1) * 16, j0 = (wave >> 1) * 16;
    const int lr = lane & 15, lk = lane >> 4;
    d4_t acc = d4_t{0.0, 0.0, 0.0, 0.0};
    if (two) {
        // the second half holds kb2 reflectors (its columns past them are not
        // read: they may lie past the submatrix)
        const int kb2 = min(32, nr - (b + 1) * 32);
        const double *V1 = Va + d.off + (int64_t)(b * 32) * d.ld;
        const double *V2 = Va + d.off + (int64_t)((b + 1) * 32) * d.ld;
        // 2 x 8 elements per thread, loads issued together (see k_larft_all)
        double x1[64 * 32 / kT], x2[64 * 32 / kT];
        auto load = [&](int r0) {
#pragma unroll
            for (int u = 0; u < 64 * 32 / kT; ++u) {
                const int l = threadIdx.x + u * kT, rr = l & 63, cc = l >> 6;
                const int64_t r = min(r0 + rr, d.m - 1);
                x1[u] = V1[(int64_t)cc * d.ld + r];
                x2[u] = V2[(int64_t)min(cc, kb2 - 1) * d.ld + r];
            }
        };
        auto stage = [&](int r0) {
#pragma unroll
            for (int u = 0; u < 64 * 32 / kT; ++u) asm volatile("" : "+v"(x1[u]), "+v"(x2[u]));
#pragma unroll
            for (int u = 0; u < 64 * 32 / kT; ++u) {
                const int l = threadIdx.x + u * kT, rr = l & 63, cc = l >> 6;
                const bool in = r0 + rr < d.m;
                V1s[rr][cc] = in ? x1[u] : 0.0;
                V2s[rr][cc] = (in && cc < kb2) ? x2[u] : 0.0;
            }
        };
        const int rb = (b + 1) * 32 + 1;
        if (rb < d.m) {
            load(rb);
            stage(rb);
        }
        __syncthreads();
        for (int r0 = rb; r0 < d.m; r0 += 64) {
            const bool more = r0 + 64 < d.m;
            if (more) load(r0 + 64);
#pragma unroll
            for (int ks = 0; ks < 16; ++ks) {
                const int kk = 4 * ks + lk;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(V1s[kk][i0 + lr], V2s[kk][j0 + lr], acc, 0, 0, 0);
            }
            __syncthreads();
            if (more) {
                stage(r0 + 64);
                __syncthreads();
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) G[i0 + 4 * r + lk][j0 + lr] = acc[r];
    __syncthreads();
    // M = G T2 into V1s (32 x 32), then T12 = -T1 M
    double *M = &V1s[0][0];  // M(i, j) at M[i * 33 + j]
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const int l = threadIdx.x + h * kT, i = l & 31, j = l >> 5;
        double sacc = 0.0;
        for (int k = 0; k < 32; ++k) sacc += G[i][k] * T2[k][j];
        M[i * 33 + j] = sacc;
    }
    __syncthreads();
    double *T64 = T64a + toff64[blockIdx.x] + (int64_t)p * 4096;
    for (int l = threadIdx.x; l < 4096; l += kT) {
        const int r = l & 63, c = l >> 6;
        double v;
        if (r < 32 && c < 32) {
            v = T1[r][c];
        } else if (r >= 32 && c >= 32) {
            v = T2[r - 32][c - 32];
        } else if (r < 32) {  // T12(r, c - 32) = -sum_k T1(r, k) M(k, c - 32)
            double sacc = 0.0;
            for (int k = 0; k < 32; ++k) sacc += T1[r][k] * M[k * 33 + (c - 32)];
            v = -sacc;
        } else {
            v = 0.0;
        }
        T64[c * 64 + r] = v;
    }
}

// Diagonal blocks of every 128-reflector factor T = [T1 T12; 0 T2] from the
// 64-reflector factors of its halves; T12 = -T1 (V1^T V2) T2 is written by
// three batched products afterwards (zero when the second half is absent).
// T128 is 128 x 128 column-major at T128a + toff128[s] + p * 16384.
// grid: (submatrix, block)
__global__ __launch_bounds__(kT) void k_t128_assemble(const SubDesc *__restrict__ desc,
                                                      const double *__restrict__ T64a,
                                                      const int64_t *__restrict__ toff64,
                                                      const int64_t *__restrict__ toff128,
                                                      double *__restrict__ T128a) {
    const SubDesc d = desc[blockIdx.x];
    const int nr = d.m - 2;
    const int n64 = nr > 0 ? (nr + 63) / 64 : 0;
    const int p = blockIdx.y;
    if (2 * p >= n64) return;
    const bool two = 2 * p + 1 < n64;
    const double *T1 = T64a + toff64[blockIdx.x] + (int64_t)(2 * p) * 4096;
    const double *T2 = T1 + 4096;
    double *T = T128a + toff128[blockIdx.x] + (int64_t)p * 16384;
    for (int l = threadIdx.x; l < 16384; l += kT) {  // l = c * 128 + r
        const int r = l & 127, c = l >> 7;
        if (r < 64 && c >= 64 && two) continue;  // T12
        double v = 0.0;
        if (r < 64 && c < 64) v = T1[c * 64 + r];
        else if (r >= 64 && c >= 64 && two) v = T2[(c - 64) * 64 + (r - 64)];
        T[l] = v;
    }
}

// Inverse of every 64x64 diagonal block of each lower-triangular L (identity
// padding past m).  grid: (block, submatrix); Dinv block b of submatrix s at
// dinv + s_doff + b*4096, column-major, leading dimension 64.
__global__ __launch_bounds__(64) void k_diag_inv(const SubDesc *__restrict__ desc,
                                                 const int64_t *__restrict__ doff,
                                                 const double *__restrict__ La,
                                                 double *__restrict__ dinv) {
    const SubDesc d = desc[blockIdx.x];
    const int b = blockIdx.y;
    const int i0 = b * 64;
    if (i0 >= d.m) return;
    __shared__ double T[64][65];
    const double *L = La + d.off;
    const int c = threadIdx.x;
    for (int r = 0; r < 64; ++r) {
        double v;
        if (i0 + r < d.m && i0 + c < d.m) v = r >= c ? L[(int64_t)(i0 + c) * d.ld + i0 + r] : 0.0;
        else v = (r == c) ? 1.0 : 0.0;
        T[r][c] = v;
    }
    __syncthreads();
    // column c of T^-1 by forward substitution
    double x[64];
#pragma unroll
    for (int r = 0; r < 64; ++r) x[r] = 0.0;
#pragma unroll
    for (int r = 0; r < 64; ++r) {
        if (r < c) continue;
        double sacc = (r == c) ? 1.0 : 0.0;
#pragma unroll
        for (int t = 0; t < r; ++t) sacc -= T[r][t] * x[t];
        x[r] = sacc / T[r][r];
    }
    double *out = dinv + doff[blockIdx.x] + (int64_t)b * 4096 + (int64_t)c * 64;
#pragma unroll
    for (int r = 0; r < 64; ++r) out[r] = x[r];
}

// The first level of the batched triangular inverse: the 64x64 diagonal
// blocks of Li are the inverses from k_diag_inv (zero above their diagonal),
// and the 64x64 block above each diagonal block is zero.  Nothing else is
// written: the level products write every block below the diagonal, and the
// readers of Li take it as a triangular factor (kTri* k ranges), which reach
// at most one 64-column block past the diagonal block with 128 x 128 tiles.
// grid: (submatrix, 64-column block)
__global__ __launch_bounds__(kT) void k_linv_init(const SubDesc *__restrict__ desc,
                                                  const int64_t *__restrict__ doff,
                                                  const double *__restrict__ dinv,
                                                  double *__restrict__ Li) {
    const SubDesc d = desc[blockIdx.x];
    const int b = blockIdx.y, c0 = 64 * b;
    if (c0 >= d.m) return;
    const double *blk = dinv + doff[blockIdx.x] + (int64_t)b * 4096;
    const int nc = min(64, d.m - c0);
    const int r0 = max(c0 - 64, 0), nr = min(c0 + 64, d.m) - r0;  // rows written
    for (int l = threadIdx.x; l < nc * nr; l += kT) {
        const int c = l / nr, r = r0 + l % nr;
        const int rr = r - c0;
        const double v = (rr >= c && rr < 64) ? blk[c * 64 + rr] : 0.0;
        Li[d.off + (int64_t)(c0 + c) * d.ld + r] = v;
    }
}

}  // namespace

namespace ptb {

// Li = L^-1 for every lower-triangular L of the batch (the upper triangle of L
// is not read).  The 64x64 diagonal blocks are inverted in LDS; then, with the
// block size h doubling, every pair of diagonal blocks [L11 0; L21 L22] (the
// second one possibly partial) gets Li21 = -Li22 (L21 Li11): two batched
// triangular GEMMs per level, T the scratch for L21 Li11.  Li is lower
// triangular with zeros above the diagonal.
int inv_lower_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                    const double *L, double *Li, double *T) {
    ProfScope prof_(ctx, "inv_lower");
    const int nsm = (int)desc.size();
    std::vector<int64_t> doff(nsm);
    int64_t tot = 0;
    int max_nb = 0, max_m = 0;
    for (int s = 0; s < nsm; ++s) {
        doff[s] = tot;
        const int nb = (desc[s].m + 63) / 64;
        tot += (int64_t)nb * 4096;
        max_nb = std::max(max_nb, nb);
        max_m = std::max(max_m, desc[s].m);
    }
    auto *dinv = static_cast<double *>(ws(ctx, "linv_dinv", sizeof(double) * std::max<int64_t>(tot, 1)));
    auto *ddoff = static_cast<int64_t *>(upload_table(ctx, "linv_doff", doff.data(), sizeof(int64_t) * nsm));
    if (!dinv || !ddoff) return PTBGPU_E_NOMEM;
    {
        ProfScope kprof_(ctx, "diag_inv");
        hipLaunchKernelGGL(k_diag_inv, dim3(nsm, max_nb), dim3(64), 0, ctx->stream, ddesc, ddoff, L,
                           dinv);
        PTB_CHECK_LAUNCH(ctx);
        hipLaunchKernelGGL(k_linv_init, dim3(nsm, max_nb), dim3(kT), 0, ctx->stream, ddesc, ddoff,
                           dinv, Li);
        PTB_CHECK_LAUNCH(ctx);
    }
    int level = 0;
    for (int h = 64; h < max_m; h *= 2, ++level) {
        std::vector<GemmDesc> g1, g2;
        for (int s = 0; s < nsm; ++s) {
            const SubDesc &x = desc[s];
            for (int o = 0; o + h < x.m; o += 2 * h) {
                const int o2 = o + h, n2 = std::min(h, x.m - o2);
                const int64_t b11 = x.off + (int64_t)o * x.ld + o;
                const int64_t b21 = x.off + (int64_t)o * x.ld + o2;
                const int64_t b22 = x.off + (int64_t)o2 * x.ld + o2;
                // T21 = L21 Li11 (Li11 lower)
                g1.push_back({L + b21, Li + b11, T + b21, n2, h, h, x.ld, x.ld, x.ld, 1.0, 0.0, 0,
                              kTriBLower});
                // Li21 = -Li22 T21 (Li22 lower)
                g2.push_back({Li + b22, T + b21, Li + b21, n2, h, n2, x.ld, x.ld, x.ld, -1.0, 0.0, 0,
                              kTriALower});
            }
        }
        if (g1.empty()) continue;
        const std::string n1 = "linv_g1_" + std::to_string(level), n2 = "linv_g2_" + std::to_string(level);
        auto *dg1 = static_cast<GemmDesc *>(upload_table(ctx, n1.c_str(), g1.data(), sizeof(GemmDesc) * g1.size()));
        auto *dg2 = static_cast<GemmDesc *>(upload_table(ctx, n2.c_str(), g2.data(), sizeof(GemmDesc) * g2.size()));
        if (!dg1 || !dg2) return PTBGPU_E_NOMEM;
        TileList t1, t2;
        t1.add(g1.data(), (int)g1.size(), true);
        t2.add(g2.data(), (int)g2.size(), true);
        PTB_TRY(t1.upload(ctx, ("linv_w1_" + std::to_string(level)).c_str()));
        PTB_TRY(t2.upload(ctx, ("linv_w2_" + std::to_string(level)).c_str()));
        PTB_TRY(t1.launch(ctx, 0, false, false, dg1, "linv"));
        PTB_TRY(t2.launch(ctx, 0, false, false, dg2, "linv"));
    }
    return PTBGPU_OK;
}

template <int TE>
static void launch_gemm_te(ptbgpu_ctx *ctx, bool ta, bool tb, dim3 grid, const GemmDesc *ddesc,
                           int count, int tiles, const GemmTile *work) {
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false, double, TE>), grid, dim3(kT), 0, ctx->stream, ddesc, count, tiles, work);
    else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false, double, TE>), grid, dim3(kT), 0, ctx->stream, ddesc, count, tiles, work);
    else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true, double, TE>), grid, dim3(kT), 0, ctx->stream, ddesc, count, tiles, work);
    else hipLaunchKernelGGL((k_gemm<true, true, double, TE>), grid, dim3(kT), 0, ctx->stream, ddesc, count, tiles, work);
}

// edge: 64 or 128 (double storage only)
static int launch_gemm(ptbgpu_ctx *ctx, bool ta, bool tb, bool f32, dim3 grid, const GemmDesc *ddesc,
                       int count, int tiles, const GemmTile *work, int edge) {
    ProfScope eprof_(ctx, edge == 128 ? "gemm_tile128" : "gemm_tile64");
    if (f32) {  // float storage: only the tridiagonal update's form is used
        if (ta || !tb || edge != 64)
            return fail(ctx, PTBGPU_E_ARG, "float-storage GEMM supports A B^T on 64 x 64 tiles only");
        hipLaunchKernelGGL((k_gemm<false, true, float, 64>), grid, dim3(kT), 0, ctx->stream, ddesc, count, tiles, work);
    } else if (edge == 128) {
        launch_gemm_te<128>(ctx, ta, tb, grid, ddesc, count, tiles, work);
    } else {
        launch_gemm_te<64>(ctx, ta, tb, grid, ddesc, count, tiles, work);
    }
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

// Tile edge of an FP64 launch of `count` products from its 128 x 128 tile
// count and the multiply-adds its tiles issue at both edges (padded output
// times the k range of each tile): the gemm_tile knob forces 64 or 128; else
// 128 x 128 tiles only for launches of fewer products than XCDs (single large
// products, spread over the chip) with at least 4 such tiles per CU and at
// most 3 % more work than 64 x 64 tiles.  On MI355X they reach 58 TFLOP/s on
// one 4096^3 product against 50; on the batches of ragged, partly triangular
// submatrix products (m about 100 to 1000) their coarser padding costs more
// than the halved operand traffic saves, so those stay at 64.
static int gemm_edge(int64_t mode, int num_cus, int count, int64_t tiles128, double work64,
                     double work128) {
    if (mode == 64 || mode == 128) return (int)mode;
    return count < 8 && tiles128 >= 4 * (int64_t)std::max(num_cus, 1) && work128 <= 1.03 * work64
               ? 128
               : 64;
}

// Multiply-adds issued by the tiles of edge te of one product: per tile,
// te^2 times its k range (the triangular factors' bounds, the start aligned
// down to the slice depth 2048 / te), as k_gemm runs it.  K of a shape the
// device fills in (0 on the host) counts as 1.
static double tile_work(int M, int N, int K, int uplo, int tri, int te) {
    if (M <= 0 || N <= 0) return 0.0;
    const int ks = 2048 / te, tM = (M + te - 1) / te, tN = (N + te - 1) / te;
    K = std::max(K, 1);
    double w = 0.0;
    for (int tn = 0; tn < tN; ++tn)
        for (int tm = uplo ? tn : 0; tm < tM; ++tm) {
            const int m0 = tm * te, n0 = tn * te;
            int kbeg = 0, kend = K;
            if (tri & kTriALower) kend = std::min(kend, m0 + te);
            if (tri & kTriAUpper) kbeg = std::max(kbeg, m0);
            if (tri & kTriBUpper) kend = std::min(kend, n0 + te);
            if (tri & kTriBLower) kbeg = std::max(kbeg, n0);
            kbeg &= ~(ks - 1);
            if (kend > kbeg) w += (double)te * te * (((kend - kbeg) + ks - 1) / ks * ks);
        }
    return w;
}

int gemm_vbatch(ptbgpu_ctx *ctx, bool ta, bool tb, const GemmDesc *ddesc, int count, int max_m,
                int max_n, const char *site, bool f32) {
    ProfScope prof_(ctx, "gemm");
    ProfScope sprof_(ctx, site ? (std::string("gemm_") + site).c_str() : "gemm_other");
    if (count <= 0 || max_m <= 0 || max_n <= 0) return PTBGPU_OK;
    auto tiles_at = [&](int te) { return (int64_t)((max_m + te - 1) / te) * ((max_n + te - 1) / te); };
    // (the k extent and triangles are on the device: padded outputs only)
    const int edge = f32 ? 64
                         : gemm_edge(ctx->gemm_tile, ctx->num_cus, count, count * tiles_at(128),
                                     4096.0 * tiles_at(64), 16384.0 * tiles_at(128));
    const int64_t tiles = tiles_at(edge);
    // fewer products than XCDs: each product's tiles over the whole chip
    // (a negative tile count selects that mapping in k_gemm)
    const bool spread = count < 8;
    const int64_t groups8 = (count + 7) / 8;
    const int64_t total = spread ? (int64_t)count * tiles : groups8 * 8 * tiles;
    if (total > 0x7fffffff) return fail(ctx, PTBGPU_E_ARG, "GEMM batch too large (%lld workgroups)",
                                        (long long)total);
    return launch_gemm(ctx, ta, tb, f32, dim3((unsigned)total), ddesc, count,
                       (int)(spread ? -tiles : tiles), nullptr, edge);
}

// one: every product of the launch is a single k slice without triangular
// factors (the A B^T form then runs k_gemm_nt1); edge: the list's tile edge
int gemm_tiles(ptbgpu_ctx *ctx, bool ta, bool tb, const GemmDesc *ddesc, const GemmTile *dwork,
               int64_t nwork, const char *site, bool f32, bool one, int lowp, int edge) {
    ProfScope prof_(ctx, "gemm");
    ProfScope sprof_(ctx, site ? (std::string("gemm_") + site).c_str() : "gemm_other");
    if (nwork <= 0) return PTBGPU_OK;
    if (edge != 64 && (edge != 128 || f32 || lowp != kLowpNone))
        return fail(ctx, PTBGPU_E_ARG, "GEMM tile edge %d is for FP64 products on double storage", edge);
    if (lowp != kLowpNone) {
        ProfScope lprof_(ctx, lowp == kLowpBF16 ? "gemm_bf16" : "gemm_bf16x3");
        return gemm_tiles_bf(ctx, ta, tb, ddesc, dwork, nwork, lowp, f32);
    }
    if (nwork > 0x7fffffff) return fail(ctx, PTBGPU_E_ARG, "GEMM tile list too long (%lld)", (long long)nwork);
    const dim3 grid((unsigned)nwork);
    if (one && !ta && tb && edge == 64) {
        if (f32) hipLaunchKernelGGL(k_gemm_nt1<float>, grid, dim3(kT), 0, ctx->stream, ddesc, dwork);
        else hipLaunchKernelGGL(k_gemm_nt1<double>, grid, dim3(kT), 0, ctx->stream, ddesc, dwork);
        PTB_CHECK_LAUNCH(ctx);
        return PTBGPU_OK;
    }
    return launch_gemm(ctx, ta, tb, f32, grid, ddesc, 0, 1, dwork, edge);
}

int TileList::add(const GemmDesc *h, int count, bool fp64) {
    sig_.push_back(count);
    sig_.push_back(fp64 ? 1 : 0);
    for (int i = 0; i < count; ++i) {
        const GemmDesc &g = h[i];
        sig_.insert(sig_.end(), {g.M, g.N, g.K, g.uplo, g.tri});
    }
    return nlaunch_++;
}

// Every launch: the tile edge (128 for an FP64 launch where gemm_edge says
// so, else 64); products largest first (longest-processing-time order), each
// dealt whole to the XCD with the fewest tiles so far (in a launch of fewer
// products than XCDs, a product larger than an XCD's even share in runs of
// tiles instead); entry 8 k + x of the launch's block is XCD x's k-th tile
// (-1 where its queue has ended).  Tiles of a product in column-major tile
// order, lower-triangle tiles only for uplo.
void TileList::build(std::vector<GemmTile> &work, int64_t tile_mode, int num_cus) {
    start_.assign(1, 0);
    flags_.clear();
    size_t p = 0;
    for (int l = 0; l < nlaunch_; ++l) {
        const int count = sig_[p++];
        const bool fp64 = sig_[p++] != 0;
        const int *sh = sig_.data() + p;  // 5 ints per product
        p += 5 * (size_t)count;
        auto tiles_of = [&](int i, int te) -> int64_t {
            const int M = sh[5 * i], N = sh[5 * i + 1];
            if (M <= 0 || N <= 0) return 0;
            const int64_t tM = (M + te - 1) / te, tN = (N + te - 1) / te;
            if (!sh[5 * i + 3]) return tM * tN;
            const int64_t c = std::min(tM, tN);  // columns tn < c hold tM - tn tiles
            return c * tM - c * (c - 1) / 2;
        };
        bool one = true;
        int64_t t128 = 0;
        for (int i = 0; i < count; ++i) {
            if (tiles_of(i, 64) == 0) continue;
            one = one && sh[5 * i + 2] <= kKS && sh[5 * i + 4] == 0;
            t128 += tiles_of(i, 128);
        }
        int te = 64;
        if (fp64 && !one && t128 > 0) {
            double w64 = 0.0, w128 = 0.0;
            if (tile_mode == 0 && count < 8)  // (gemm_edge needs the work only then)
                for (int i = 0; i < count; ++i) {
                    const int *d = sh + 5 * i;
                    w64 += tile_work(d[0], d[1], d[2], d[3], d[4], 64);
                    w128 += tile_work(d[0], d[1], d[2], d[3], d[4], 128);
                }
            te = gemm_edge(tile_mode, num_cus, count, t128, w64, w128);
        }
        std::vector<int> ord(count);
        std::vector<int64_t> area(count);
        for (int i = 0; i < count; ++i) {
            ord[i] = i;
            area[i] = (int64_t)std::max(sh[5 * i], 0) * std::max(sh[5 * i + 1], 0) * std::max(sh[5 * i + 2], 1);
        }
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return area[a] > area[b]; });
        int64_t total = 0;
        for (int i = 0; i < count; ++i) total += tiles_of(i, te);
        // With fewer products than XCDs, a product with more tiles than an
        // XCD's even share of the launch is dealt in runs of consecutive tiles
        // (column-major), each run to the XCD with the fewest tiles so far;
        // every other product goes whole (its tiles share one L2).
        const int64_t share = (total + 7) / 8;
        std::vector<GemmTile> q[8];
        for (int i : ord) {
            const int64_t nt = tiles_of(i, te);
            if (nt == 0) continue;
            const int tM = (sh[5 * i] + te - 1) / te, tN = (sh[5 * i + 1] + te - 1) / te;
            const int64_t run = count < 8 && nt > share ? std::max<int64_t>(8, nt / 32) : nt;
            int64_t k = 0;
            int x = 0;
            for (int tn = 0; tn < tN; ++tn)
                for (int tm = sh[5 * i + 3] ? tn : 0; tm < tM; ++tm, ++k) {
                    if (k % run == 0) {
                        x = 0;
                        for (int y = 1; y < 8; ++y)
                            if (q[y].size() < q[x].size()) x = y;
                    }
                    q[x].push_back(GemmTile{i, tm | (tn << 16)});
                }
        }
        size_t len = 0;
        for (int x = 0; x < 8; ++x) len = std::max(len, q[x].size());
        const size_t base = work.size();
        work.resize(base + 8 * len, GemmTile{-1, 0});
        for (int x = 0; x < 8; ++x)
            for (size_t k = 0; k < q[x].size(); ++k) work[base + 8 * k + x] = q[x][k];
        start_.push_back((int64_t)work.size());
        flags_.push_back((char)((one ? kOne : 0) | (te == 128 ? kEdge128 : 0)));
    }
}

int TileList::upload(ptbgpu_ctx *ctx, const char *name_in) {
    const std::string full = ctx->table_prefix + name_in;  // see upload_table
    const char *name = full.c_str();
    // the shadow entry of `name` holds the shapes of the table in its buffer
    // and the tile-edge knob they were dealt under (ws drops it when the
    // buffer is reallocated); "<name>#meta" holds the launch starts and flags
    // built from them
    const std::string mkey = std::string(name) + "#meta";
    std::vector<int> sig(sig_);
    sig.push_back((int)ctx->gemm_tile);
    const char *sb = reinterpret_cast<const char *>(sig.data());
    const size_t sn = sizeof(int) * sig.size();
    auto sh = ctx->shadow.find(name);
    auto mh = ctx->shadow.find(mkey);
    const size_t nl = (size_t)nlaunch_;
    if (sh != ctx->shadow.end() && mh != ctx->shadow.end() && sh->second.size() == sn &&
        memcmp(sh->second.data(), sb, sn) == 0 &&
        mh->second.size() == sizeof(int64_t) * (nl + 1) + nl) {
        start_.resize(nl + 1);
        flags_.resize(nl);
        memcpy(start_.data(), mh->second.data(), sizeof(int64_t) * (nl + 1));
        memcpy(flags_.data(), mh->second.data() + sizeof(int64_t) * (nl + 1), nl);
        if (start_.back() == 0) return PTBGPU_OK;
        dwork_ = static_cast<const GemmTile *>(ws(ctx, name, sizeof(GemmTile) * start_.back()));
        return dwork_ ? PTBGPU_OK : PTBGPU_E_NOMEM;
    }
    std::vector<GemmTile> work;
    build(work, ctx->gemm_tile, ctx->num_cus);
    std::vector<char> meta(sizeof(int64_t) * (nl + 1) + nl);
    memcpy(meta.data(), start_.data(), sizeof(int64_t) * (nl + 1));
    memcpy(meta.data() + sizeof(int64_t) * (nl + 1), flags_.data(), nl);
    if (work.empty()) {
        ctx->shadow[name].assign(sb, sb + sn);
        ctx->shadow[mkey] = meta;
        return PTBGPU_OK;
    }
    if (ctx->capturing)  // a graph must not hold a copy from host memory
        return fail(ctx, PTBGPU_E_ARG, "tile list '%s' changed during a graph capture", name);
    const size_t bytes = sizeof(GemmTile) * work.size();
    void *d = ws(ctx, name, bytes);
    if (!d) return PTBGPU_E_NOMEM;
    if (hipMemcpyAsync(d, work.data(), bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        ctx->shadow.erase(name);
        return fail(ctx, PTBGPU_E_HIP, "upload of tile list '%s' failed: %s", name,
                    hipGetErrorString(hipGetLastError()));
    }
    ctx->shadow[name].assign(sb, sb + sn);
    ctx->shadow[mkey] = meta;
    dwork_ = static_cast<const GemmTile *>(d);
    return PTBGPU_OK;
}

int TileList::launch(ptbgpu_ctx *ctx, int i, bool ta, bool tb, const GemmDesc *ddesc,
                     const char *site, bool f32, int lowp) const {
    const int64_t n = start_[i + 1] - start_[i];
    if (n == 0) return PTBGPU_OK;
    if (!dwork_) return fail(ctx, PTBGPU_E_ARG, "GEMM tile list launched before upload");
    return gemm_tiles(ctx, ta, tb, ddesc, dwork_ + start_[i], n, site, f32, (flags_[i] & kOne) != 0,
                      lowp, (flags_[i] & kEdge128) ? 128 : 64);
}

// X = L^-1 B (trans=false) or L^-T B (trans=true), all m columns, blocked by
// 64 rows: per block row one update GEMM (in place on B) and one product with
// the inverted diagonal block (into X).  B is overwritten.
int trsm_gemm(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
              const double *L, double *B, double *X, bool trans, const char *tag,
              const RhsLayout *rhs) {
    ProfScope prof_(ctx, "trsm");
    ProfScope tprof_(ctx, (std::string("trsm_") + tag).c_str());
    const int nsm = (int)desc.size();
    std::vector<int64_t> doff(nsm);
    int64_t tot = 0;
    int max_nb = 0, max_m = 0;
    for (int s = 0; s < nsm; ++s) {
        doff[s] = tot;
        const int nb = (desc[s].m + 63) / 64;
        tot += (int64_t)nb * 4096;
        max_nb = std::max(max_nb, nb);
        max_m = std::max(max_m, desc[s].m);
    }
    auto *dinv = static_cast<double *>(ws(ctx, "trsm_dinv", sizeof(double) * tot));
    auto *ddoff = static_cast<int64_t *>(upload_table(ctx, "trsm_doff", doff.data(), sizeof(int64_t) * nsm));
    if (!dinv || !ddoff) return PTBGPU_E_NOMEM;
    {
        ProfScope kprof_(ctx, "diag_inv");
        hipLaunchKernelGGL(k_diag_inv, dim3(nsm, max_nb), dim3(64), 0, ctx->stream, ddesc, ddoff, L,
                           dinv);
        PTB_CHECK_LAUNCH(ctx);
    }
    std::vector<GemmDesc> g1, g2;
    std::vector<int> s1(max_nb + 1), s2(max_nb + 1);
    for (int q = 0; q < max_nb; ++q) {
        s1[q] = (int)g1.size();
        s2[q] = (int)g2.size();
        for (int s = 0; s < nsm; ++s) {
            const SubDesc &d = desc[s];
            const int nb = (d.m + 63) / 64;
            if (q >= nb) continue;
            const int ib = trans ? nb - 1 - q : q;
            const int i0 = ib * 64, ibs = std::min(64, d.m - i0);
            const int nc = rhs ? rhs->cols[s] : d.m;
            if (nc <= 0) continue;
            const double *Lb = L + d.off;
            double *Bb = B + (rhs ? rhs->boff[s] : d.off), *Xb = X + (rhs ? rhs->xoff[s] : d.off);
            const double *Di = dinv + doff[s] + (int64_t)ib * 4096;
            if (!trans) {
                if (i0 > 0)
                    g1.push_back({Lb + i0, Xb, Bb + i0, ibs, nc, i0, d.ld, d.ld, d.ld, -1.0, 1.0});
                g2.push_back({Di, Bb + i0, Xb + i0, ibs, nc, ibs, 64, d.ld, d.ld, 1.0, 0.0});
            } else {
                const int rest = d.m - i0 - ibs;
                if (rest > 0)
                    g1.push_back({Lb + (int64_t)i0 * d.ld + i0 + ibs, Xb + i0 + ibs, Bb + i0, ibs, nc,
                                  rest, d.ld, d.ld, d.ld, -1.0, 1.0});
                g2.push_back({Di, Bb + i0, Xb + i0, ibs, nc, ibs, 64, d.ld, d.ld, 1.0, 0.0});
            }
        }
    }
    s1[max_nb] = (int)g1.size();
    s2[max_nb] = (int)g2.size();
    std::string n1 = std::string("trsm_g1_") + tag, n2 = std::string("trsm_g2_") + tag;
    auto *dg1 = static_cast<GemmDesc *>(upload_table(ctx, n1.c_str(), g1.data(), sizeof(GemmDesc) * g1.size()));
    auto *dg2 = static_cast<GemmDesc *>(upload_table(ctx, n2.c_str(), g2.data(), sizeof(GemmDesc) * g2.size()));
    if (!dg1 || !dg2) return PTBGPU_E_NOMEM;
    for (int q = 0; q < max_nb; ++q) {
        const int c1 = s1[q + 1] - s1[q], c2 = s2[q + 1] - s2[q];
        if (c1 > 0) PTB_TRY(gemm_vbatch(ctx, trans, false, dg1 + s1[q], c1, 64, max_m, "trsm"));
        if (c2 > 0) PTB_TRY(gemm_vbatch(ctx, trans, false, dg2 + s2[q], c2, 64, max_m, "trsm"));
    }
    return PTBGPU_OK;
}

// Blocks of kPB = 4 panels of 32 columns (right-looking within a block):
// after panel j of a block, one rank-32 update of the block's remaining
// columns; after the block, one rank-128 update of the trailing matrix (lower
// triangle), so the trailing matrix is read and written once per 128 columns.
int potrf_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                double *A, int *flags) {
    ProfScope prof_(ctx, "potrf");
    constexpr int kPB = 4, kBW = 32 * kPB;
    const int nsm = (int)desc.size();
    int max_m = 0;
    for (const auto &x : desc) max_m = std::max(max_m, x.m);
    const int npanel = (max_m + 31) / 32;
    const int nblk = (max_m + kBW - 1) / kBW;
    std::vector<GemmDesc> gs, gb;
    // gs launches (block b, panel j < kPB - 1) in order; gb launches per block
    std::vector<int> ss(nblk * (kPB - 1) + 1, 0), sb(nblk + 1, 0);
    for (int b = 0; b < nblk; ++b) {
        const int k0 = kBW * b;
        for (int j = 0; j < kPB - 1; ++j) {
            ss[b * (kPB - 1) + j] = (int)gs.size();
            for (int si = 0; si < nsm; ++si) {
                const SubDesc &x = desc[si];
                double *a = A + x.off;
                const int pc = k0 + 32 * j, s1 = pc + 32, se = std::min(k0 + kBW, x.m);
                if (s1 < se) {  // block columns s1..se-1 -= P P(s1:se, :)^T
                    double *P = a + (int64_t)pc * x.ld + s1;
                    gs.push_back({P, P, a + (int64_t)s1 * x.ld + s1, x.m - s1, se - s1, 32, x.ld, x.ld,
                                  x.ld, -1.0, 1.0, 1, 0});
                }
            }
        }
        sb[b] = (int)gb.size();
        for (int si = 0; si < nsm; ++si) {
            const SubDesc &x = desc[si];
            double *a = A + x.off;
            const int s2 = k0 + kBW;
            if (s2 < x.m) {  // trailing matrix -= P P^T, P = the block's columns
                const int n = x.m - s2;
                double *P = a + (int64_t)k0 * x.ld + s2;
                gb.push_back({P, P, a + (int64_t)s2 * x.ld + s2, n, n, kBW, x.ld, x.ld, x.ld, -1.0, 1.0,
                              1, 0});
            }
        }
    }
    ss[nblk * (kPB - 1)] = (int)gs.size();
    sb[nblk] = (int)gb.size();
    auto *dgs = static_cast<GemmDesc *>(upload_table(ctx, "potrf_gs", gs.data(), sizeof(GemmDesc) * gs.size()));
    auto *dgb = static_cast<GemmDesc *>(upload_table(ctx, "potrf_gb", gb.data(), sizeof(GemmDesc) * gb.size()));
    if (!dgs || !dgb) return PTBGPU_E_NOMEM;
    TileList ts, tb;
    for (int l = 0; l < nblk * (kPB - 1); ++l) ts.add(gs.data() + ss[l], ss[l + 1] - ss[l], true);
    for (int b = 0; b < nblk; ++b) tb.add(gb.data() + sb[b], sb[b + 1] - sb[b], true);
    PTB_TRY(ts.upload(ctx, "potrf_ws"));
    PTB_TRY(tb.upload(ctx, "potrf_wb"));
    // rows per panel over the device: one workgroup per submatrix when the
    // batch fills the chip, else the rows split over several
    const int ny = (int)std::min<int64_t>(std::max<int64_t>(1, ((int64_t)max_m + kT - 1) / kT),
                                          std::max<int64_t>(1, (int64_t)ctx->num_cus / nsm));
    auto panel = [&](int p) -> int {
        ProfScope kprof_(ctx, "potrf_panel");
        if (ny == 1) {
            hipLaunchKernelGGL(k_potrf_panel, dim3(nsm), dim3(kT), 0, ctx->stream, ddesc, A, flags, p, 0);
        } else {
            hipLaunchKernelGGL(k_potrf_panel, dim3(nsm, ny), dim3(kT), 0, ctx->stream, ddesc, A, flags, p, 1);
            hipLaunchKernelGGL(k_potrf_panel, dim3(nsm), dim3(kT), 0, ctx->stream, ddesc, A, flags, p, 2);
        }
        PTB_CHECK_LAUNCH(ctx);
        return PTBGPU_OK;
    };
    for (int b = 0; b < nblk; ++b) {
        for (int j = 0; j < kPB; ++j) {
            const int p = kPB * b + j;
            if (p < npanel) PTB_TRY(panel(p));
            if (j < kPB - 1) {
                const int l = b * (kPB - 1) + j;
                PTB_TRY(ts.launch(ctx, l, false, true, dgs + ss[l], "potrf"));
            }
        }
        PTB_TRY(tb.launch(ctx, b, false, true, dgb + sb[b], "potrf"));
    }
    // a failed factorisation (flags bit 0) leaves L = I
    hipLaunchKernelGGL(k_identity_if_bad, dim3(nsm), dim3(kT), 0, ctx->stream, ddesc, flags, A);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int transpose_sym_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m,
                        const double *in, double *out, bool symmetrize) {
    ProfScope prof_(ctx, "transpose");
    const int nt = (max_m + 31) / 32;
    hipLaunchKernelGGL(k_transpose, dim3(nt * nt, nsm), dim3(kT), 0, ctx->stream, ddesc, in, out,
                       symmetrize ? 1 : 0);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int mirror_lower_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, double *A) {
    ProfScope prof_(ctx, "mirror");
    const int nt = (max_m + 31) / 32;
    hipLaunchKernelGGL(k_mirror_lower, dim3(nt * nt, nsm), dim3(kT), 0, ctx->stream, ddesc, A);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int zero_upper_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, double *A) {
    ProfScope prof_(ctx, "zero_upper");
    hipLaunchKernelGGL(k_zero_upper, dim3(nsm, max_m), dim3(kT), 0, ctx->stream, ddesc, A);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int to_f32_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const double *A,
                 float *F) {
    ProfScope prof_(ctx, "to_f32");
    hipLaunchKernelGGL(k_to_f32, dim3(nsm, max_m), dim3(kT), 0, ctx->stream, ddesc, A, F);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int copy_lower_batch(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const double *A,
                     double *B) {
    ProfScope prof_(ctx, "copy_lower");
    hipLaunchKernelGGL(k_copy_lower, dim3(nsm, max_m), dim3(kT), 0, ctx->stream, ddesc, A, B);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int promote_reflectors(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const float *F,
                       double *V, double *tau) {
    ProfScope prof_(ctx, "promote");
    const int per = kT / PTB_WAVE;
    hipLaunchKernelGGL(k_promote_reflectors, dim3(nsm, (max_m + per - 1) / per), dim3(kT), 0,
                       ctx->stream, ddesc, F, V, tau);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int larft_all(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
              const double *V, const double *tau, QFactor &qf) {
    ProfScope prof_(ctx, "larft");
    const int nsm = (int)desc.size();
    qf.toff.assign(nsm, 0);
    int64_t tot = 0;
    int maxblk = 0;
    for (int s = 0; s < nsm; ++s) {
        const int nr = desc[s].m - 2;
        const int nblk = nr > 0 ? (nr + 31) / 32 : 0;
        qf.toff[s] = tot;
        tot += (int64_t)nblk * 1024;
        maxblk = std::max(maxblk, nblk);
    }
    qf.V = V;
    qf.maxblk = maxblk;
    qf.T = static_cast<double *>(ws(ctx, "q_T", sizeof(double) * std::max<int64_t>(tot, 1)));
    auto *dtoff = static_cast<int64_t *>(upload_table(ctx, "q_toff", qf.toff.data(), sizeof(int64_t) * nsm));
    if (!qf.T || !dtoff) return PTBGPU_E_NOMEM;
    qf.toff64.assign(nsm, 0);
    int64_t tot64 = 0;
    for (int s = 0; s < nsm; ++s) {
        const int nr = desc[s].m - 2;
        const int npair = nr > 0 ? (nr + 63) / 64 : 0;
        qf.toff64[s] = tot64;
        tot64 += (int64_t)npair * 4096;
    }
    qf.maxblk64 = (maxblk + 1) / 2;
    qf.T64 = static_cast<double *>(ws(ctx, "q_T64", sizeof(double) * std::max<int64_t>(tot64, 1)));
    auto *dtoff64 = static_cast<int64_t *>(upload_table(ctx, "q_toff64", qf.toff64.data(), sizeof(int64_t) * nsm));
    if (!qf.T64 || !dtoff64) return PTBGPU_E_NOMEM;
    if (maxblk > 0) {
        hipLaunchKernelGGL(k_larft_all, dim3(nsm, maxblk), dim3(kT), 0, ctx->stream, ddesc, V, tau,
                           dtoff, qf.T);
        PTB_CHECK_LAUNCH(ctx);
        hipLaunchKernelGGL(k_larft_merge, dim3(nsm, qf.maxblk64), dim3(kT), 0, ctx->stream, ddesc, V,
                           qf.T, dtoff, dtoff64, qf.T64);
        PTB_CHECK_LAUNCH(ctx);
    }
    // 128-reflector factors from pairs of 64-reflector ones:
    //   T = [T1 T12; 0 T2],  T12 = -T1 (V1^T V2) T2,
    // V1^T V2 over the rows where V2 is nonzero (from its first unit on)
    qf.toff128.assign(nsm, 0);
    int64_t tot128 = 0, npairs = 0;
    for (int s = 0; s < nsm; ++s) {
        const int nr = desc[s].m - 2;
        const int n64 = nr > 0 ? (nr + 63) / 64 : 0;
        qf.toff128[s] = tot128;
        tot128 += (int64_t)((n64 + 1) / 2) * 16384;
        npairs += n64 / 2;
    }
    qf.maxblk128 = (qf.maxblk64 + 1) / 2;
    qf.T128 = static_cast<double *>(ws(ctx, "q_T128", sizeof(double) * std::max<int64_t>(tot128, 1)));
    auto *dtoff128 = static_cast<int64_t *>(upload_table(ctx, "q_toff128", qf.toff128.data(), sizeof(int64_t) * nsm));
    auto *GM = static_cast<double *>(ws(ctx, "q_t128_gm", sizeof(double) * std::max<int64_t>(8192 * npairs, 1)));
    if (!qf.T128 || !dtoff128 || !GM) return PTBGPU_E_NOMEM;
    if (qf.maxblk128 == 0) return PTBGPU_OK;
    hipLaunchKernelGGL(k_t128_assemble, dim3(nsm, qf.maxblk128), dim3(kT), 0, ctx->stream, ddesc,
                       qf.T64, dtoff64, dtoff128, qf.T128);
    PTB_CHECK_LAUNCH(ctx);
    std::vector<GemmDesc> ga, gm, gt;
    int64_t sc = 0;
    for (int s = 0; s < nsm; ++s) {
        const SubDesc &d = desc[s];
        const int nr = d.m - 2;
        const int n64 = nr > 0 ? (nr + 63) / 64 : 0;
        for (int p = 0; 2 * p + 1 < n64; ++p) {
            const int b0 = 128 * p, kb2 = std::min(64, nr - (b0 + 64)), r0 = b0 + 65;
            const double *V1 = V + d.off + (int64_t)b0 * d.ld + r0;
            const double *V2 = V + d.off + (int64_t)(b0 + 64) * d.ld + r0;
            const double *T1 = qf.T64 + qf.toff64[s] + (int64_t)(2 * p) * 4096, *T2 = T1 + 4096;
            double *G = GM + sc, *Mm = G + 4096;
            sc += 8192;
            double *T12 = qf.T128 + qf.toff128[s] + (int64_t)p * 16384 + 64 * 128;
            ga.push_back({V1, V2, G, 64, kb2, d.m - r0, d.ld, d.ld, 64, 1.0, 0.0, 0, 0});  // V1^T V2
            gm.push_back({G, T2, Mm, 64, kb2, kb2, 64, 64, 64, 1.0, 0.0, 0, kTriBUpper});  // G T2
            gt.push_back({T1, Mm, T12, 64, kb2, 64, 64, 64, 128, -1.0, 0.0, 0, kTriAUpper});  // -T1 M
        }
    }
    if (ga.empty()) return PTBGPU_OK;
    const size_t gbytes = sizeof(GemmDesc) * ga.size();
    auto *dga = static_cast<GemmDesc *>(upload_table(ctx, "q_t128_ga", ga.data(), gbytes));
    auto *dgm = static_cast<GemmDesc *>(upload_table(ctx, "q_t128_gm_d", gm.data(), gbytes));
    auto *dgt = static_cast<GemmDesc *>(upload_table(ctx, "q_t128_gt", gt.data(), gbytes));
    if (!dga || !dgm || !dgt) return PTBGPU_E_NOMEM;
    TileList ta, tm, tt;
    ta.add(ga.data(), (int)ga.size(), true);
    tm.add(gm.data(), (int)gm.size(), true);
    tt.add(gt.data(), (int)gt.size(), true);
    PTB_TRY(ta.upload(ctx, "q_t128_wa"));
    PTB_TRY(tm.upload(ctx, "q_t128_wm"));
    PTB_TRY(tt.upload(ctx, "q_t128_wt"));
    PTB_TRY(ta.launch(ctx, 0, true, false, dga, "larft"));
    PTB_TRY(tm.launch(ctx, 0, false, false, dgm, "larft"));
    return tt.launch(ctx, 0, false, false, dgt, "larft");
}

// Y <- Q Y (trans = false; blocks from the last to the first) or Y <- Q^T Y
// (trans = true; blocks from the first, Q_b^T = I - V_b T_b^T V_b^T) for
// ncols[s] columns of submatrix s at Y + yoff[s] (leading dimension ld), in
// 128-reflector blocks (each block reads and rewrites Y once).  Per block:
// W1 = V_b^T Y, W2 = T_b W1 (or T_b^T W1), Y -= V_b W2, each a batched GEMM;
// W1 and W2 are 128 x ncols panels in engine workspace.  V_b spans the
// block's columns from the row after its first reflector's pivot; every
// column is zero above its own unit.
int apply_q(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const QFactor &qf, double *Y,
            const std::vector<int64_t> &yoff, const std::vector<int> &ncols, bool trans,
            const char *tag) {
    ProfScope prof_(ctx, "apply_q");
    ProfScope tprof_(ctx, (std::string("apply_q_") + tag).c_str());
    const int nsm = (int)desc.size();
    if (qf.maxblk128 == 0) return PTBGPU_OK;
    constexpr int kQB = kQBlk;
    std::vector<int64_t> woff(nsm);
    int64_t wtot = 0;
    int max_m = 0, max_c = 0;
    for (int s = 0; s < nsm; ++s) {
        woff[s] = wtot;
        wtot += kQB * (int64_t)std::max(ncols[s], 0);
        max_m = std::max(max_m, desc[s].m);
        max_c = std::max(max_c, ncols[s]);
    }
    if (max_c == 0) return PTBGPU_OK;
    auto *W1 = static_cast<double *>(ws(ctx, "q_w1", sizeof(double) * std::max<int64_t>(wtot, 1)));
    auto *W2 = static_cast<double *>(ws(ctx, "q_w2", sizeof(double) * std::max<int64_t>(wtot, 1)));
    if (!W1 || !W2) return PTBGPU_E_NOMEM;
    std::vector<GemmDesc> g1, g2, g3;
    std::vector<int> start(qf.maxblk128 + 1, 0);
    for (int q = 0; q < qf.maxblk128; ++q) {
        start[q] = (int)g1.size();
        for (int s = 0; s < nsm; ++s) {
            const SubDesc &d = desc[s];
            const int nc = ncols[s];
            const int nr = d.m - 2;
            const int nblk = nr > 0 ? (nr + kQB - 1) / kQB : 0;
            if (nc <= 0 || q >= nblk) continue;
            const int p = trans ? q : nblk - 1 - q;
            const int b0 = p * kQB, kb = std::min(kQB, nr - b0);
            const int rowsv = d.m - b0 - 1;
            const double *V = qf.V + d.off + (int64_t)b0 * d.ld + b0 + 1;
            double *Yb = Y + yoff[s] + b0 + 1;
            const double *T = qf.T128 + qf.toff128[s] + (int64_t)p * kQB * kQB;
            double *w1 = W1 + woff[s], *w2 = W2 + woff[s];
            g1.push_back({V, Yb, w1, kb, nc, rowsv, d.ld, d.ld, kQB, 1.0, 0.0});
            g2.push_back({T, w1, w2, kb, nc, kb, kQB, kQB, kQB, 1.0, 0.0});
            g3.push_back({V, w2, Yb, rowsv, nc, kb, d.ld, kQB, d.ld, -1.0, 1.0});
        }
    }
    start[qf.maxblk128] = (int)g1.size();
    const std::string t(tag);
    if (g1.empty()) return PTBGPU_OK;
    const size_t gb = sizeof(GemmDesc) * g1.size();
    GemmDesc *dg1 = static_cast<GemmDesc *>(upload_table(ctx, ("q_g1_" + t).c_str(), g1.data(), gb));
    GemmDesc *dg2 = static_cast<GemmDesc *>(upload_table(ctx, ("q_g2_" + t).c_str(), g2.data(), gb));
    GemmDesc *dg3 = static_cast<GemmDesc *>(upload_table(ctx, ("q_g3_" + t).c_str(), g3.data(), gb));
    if (!dg1 || !dg2 || !dg3) return PTBGPU_E_NOMEM;
    TileList t1, t2, t3;
    for (int q = 0; q < qf.maxblk128; ++q) {
        const int c = start[q + 1] - start[q];
        t1.add(g1.data() + start[q], c, true);
        t2.add(g2.data() + start[q], c, true);
        t3.add(g3.data() + start[q], c, true);
    }
    PTB_TRY(t1.upload(ctx, ("q_w1_" + t).c_str()));
    PTB_TRY(t2.upload(ctx, ("q_w2_" + t).c_str()));
    PTB_TRY(t3.upload(ctx, ("q_w3_" + t).c_str()));
    for (int q = 0; q < qf.maxblk128; ++q) {
        if (start[q + 1] == start[q]) continue;
        PTB_TRY(t1.launch(ctx, q, true, false, dg1 + start[q], "q"));
        PTB_TRY(t2.launch(ctx, q, trans, false, dg2 + start[q], "q"));
        PTB_TRY(t3.launch(ctx, q, false, false, dg3 + start[q], "q"));
    }
    return PTBGPU_OK;
}

}  // namespace ptb
