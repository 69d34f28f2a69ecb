// Householder reduction of each K = L^-1 H L^-T to tridiagonal form, K = Q T Q^T,
// Q = H_0 ... H_{m-3},  H_k = I - tau_k v_k v_k^T  (LAPACK dsytrd, lower).
//
// One 256-thread workgroup per submatrix, blocked by kNB = 16 columns (dlatrd
// scheme): inside a panel the trailing matrix stays as it was at the panel start
// and the pending rank-2 terms are applied on the fly as V (W^T v) + W (V^T v);
// at the end of the panel one rank-32 update refreshes the lower triangle of the
// trailing matrix (the upper triangle is never read again).  The per-column
// product p = A22 v is a symmetric GEMV over the lower triangle: lanes own
// rows, so a wavefront reads 64 consecutive elements of one column, and the
// transposed half of each element's work is reduced across the wavefront.
// This GEMV is the bandwidth-bound core of the solve:
// sum_k 8 (m-k)(m-k+1)/2 bytes per submatrix.
//
// Output: d (diagonal), e (subdiagonal, e[k] couples k and k+1), tau, and v_k
// stored in column k rows k+1.. of A (unit at row k+1).  W is a work panel
// (m x kNB, leading dimension ld) in a separate panel buffer.
//
// The matrix and the panel buffer are stored as T: double for the plain
// solve, float for the mixed-precision one (half the bytes of the GEMV).
// Sums, the reflectors and the LDS vectors are FP64 in both cases; only the
// stored values are rounded to T.
#include <algorithm>

#include "sm_internal.hpp"

namespace {

constexpr int kT = 256;
constexpr int kNB = 16;
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

// Sum over the wavefront of 16 values per lane, returned in the lanes with
// (lane >> 2) & 15 == q for value q: a reduce-scatter butterfly (partners
// 32, 16, 8, 4 halve the values held, partners 1 and 2 finish the sums) in
// 17 exchanges instead of 16 full reductions.
__device__ __forceinline__ double butterfly16(const double (&p)[16], int lane) {
    double h8[8], h4[4], h2[2];
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const double send = b5 ? p[j] : p[j + 8];
        const double keep = b5 ? p[j + 8] : p[j];
        h8[j] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double send = b4 ? h8[j] : h8[j + 4];
        const double keep = b4 ? h8[j + 4] : h8[j];
        h4[j] = keep + __shfl_xor(send, 16);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const double send = b3 ? h4[j] : h4[j + 2];
        const double keep = b3 ? h4[j + 2] : h4[j];
        h2[j] = keep + __shfl_xor(send, 8);
    }
    const double send = b2 ? h2[0] : h2[1];
    double v = (b2 ? h2[1] : h2[0]) + __shfl_xor(send, 4);
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    return v;
}

// The same for 8 values per lane: value q ends in the lanes with
// (lane >> 3) & 7 == q (partners 32, 16, 8 halve, partners 1, 2, 4 finish).
__device__ __forceinline__ double butterfly8(const double (&p)[8], int lane) {
    double h4[4], h2[2];
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const double send = b5 ? p[j] : p[j + 4];
        const double keep = b5 ? p[j + 4] : p[j];
        h4[j] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const double send = b4 ? h4[j] : h4[j + 2];
        const double keep = b4 ? h4[j + 2] : h4[j];
        h2[j] = keep + __shfl_xor(send, 16);
    }
    const double send = b3 ? h2[0] : h2[1];
    double v = (b3 ? h2[1] : h2[0]) + __shfl_xor(send, 8);
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    return v;
}

// Float storage runs three workgroups per CU (the FP64 instance would spill
// at that register budget and runs two).
template <typename T>
__global__ __launch_bounds__(kT, sizeof(T) == 4 ? 3 : 2) void k_sytrd_panel(const SubDesc *__restrict__ desc,
                                                    T *__restrict__ Aa,
                                                    T *__restrict__ Wa,
                                                    double *__restrict__ d_all,
                                                    double *__restrict__ e_all,
                                                    double *__restrict__ tau_all,
                                                    const int64_t *__restrict__ pwoff,
                                                    const int *__restrict__ order, int panel) {
    extern __shared__ double sm[];
    // submatrices largest first, so that the last workgroups to start are small
    const int si = order[blockIdx.x];
    const SubDesc ds = desc[si];
    const int m = ds.m, ld = ds.ld;
    T *a = Aa + ds.off;
    // panel buffer [W | V | W], ld x 96: W columns 0..31, a copy of the panel's
    // reflectors in 32..63 and of W in 64..95, so that the trailing update
    // A -= W V^T + V W^T is one product [W V] [V W]^T
    T *W = Wa + pwoff[si];
    double *dd = d_all + ds.mapoff, *ee = e_all + ds.mapoff, *tt = tau_all + ds.mapoff;
    double *vs = sm;              // m   current Householder vector (rows > k)
    double *cp = vs + m;          // kWaves x m  per-wavefront partials of A22 v
    double *praw = cp;            // m   A22 v before the pending corrections
    double *red = cp + kWaves * m; // kWaves
    double *c1 = red + kWaves;    // kNB  W[k][t]
    double *c2 = c1 + kNB;        // kNB  V[k][k0+t]
    double *q1 = c2 + kNB;        // kNB  W[:,t]^T v
    double *q2 = q1 + kNB;        // kNB  V[:,k0+t]^T v
    double *qred = q2 + kNB;      // kWaves x 8 (float storage: wavefront parts of the dots)
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;

    const int k0 = panel * kNB;
    if (k0 >= m - 2) return;
    {
        const int kb = min(kNB, m - 2 - k0);
        if (kb < kNB) {  // the product reads all 64 columns: zero the unused ones
            for (int l = threadIdx.x; l < (kNB - kb) * m; l += kT) {
                const int jj = kb + l / m, i = l % m;
                W[(int64_t)jj * ld + i] = T(0);
                W[(int64_t)(kNB + jj) * ld + i] = T(0);
                W[(int64_t)(2 * kNB + jj) * ld + i] = T(0);
            }
            __syncthreads();
        }
        for (int jj = 0; jj < kb; ++jj) {
            const int k = k0 + jj;
            const int n = m - k - 1;  // trailing size
            // (a) pending-update coefficients of row k
            if ((int)threadIdx.x < jj) {
                c1[threadIdx.x] = W[(int64_t)threadIdx.x * ld + k];
                c2[threadIdx.x] = a[(int64_t)(k0 + threadIdx.x) * ld + k];
            }
            __syncthreads();
            // (b) column k with the pending terms applied; sigma = |x(k+2:)|^2
            double part_s = 0.0;
            for (int i = k + 1 + threadIdx.x; i < m; i += kT) {
                double x = a[(int64_t)k * ld + i];
#pragma unroll 4
                for (int t = 0; t < jj; ++t)
                    x -= a[(int64_t)(k0 + t) * ld + i] * c1[t] + W[(int64_t)t * ld + i] * c2[t];
                vs[i] = x;
                if (i > k + 1) part_s += x * x;
            }
            if (threadIdx.x == 0) {
                double dkk = a[(int64_t)k * ld + k];
                for (int t = 0; t < jj; ++t) dkk -= 2.0 * c2[t] * c1[t];
                dd[k] = dkk;
            }
            const double sigma = block_reduce(part_s, red);  // includes a barrier
            // (c) reflector
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
                a[(int64_t)k * ld + i] = (T)v;
                W[(int64_t)(kNB + jj) * ld + i] = (T)v;
            }
            if (threadIdx.x == 0) {
                ee[k] = beta;
                tt[k] = tk;
            }
            __syncthreads();
            // (d) pending dots q1[t] = W[:,t]^T v, q2[t] = V[:,k0+t]^T v,
            // eight sums per butterfly (lanes with lane % 8 == 0 end with
            // sum (lane >> 3) & 7).  FP64 storage: wavefront wid takes the
            // groups of four t from 4 wid, 4 (wid + kWaves), ... over all rows,
            // with no barrier (q1 and q2 are read in (f), after the barriers
            // of (e)).  Float storage (three workgroups per CU, fewer
            // registers): the workgroup takes one group at a time and sums
            // the wavefronts' parts through LDS.
            if constexpr (sizeof(T) == 8) {
                for (int t0 = 4 * wid; t0 < jj; t0 += 4 * kWaves) {
                    double acc[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) acc[q] = 0.0;
#pragma unroll 2
                    for (int i = k + 1 + lane; i < m; i += PTB_WAVE) {
                        const double v = vs[i];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int t = min(t0 + q, jj - 1);
                            acc[q] += W[(int64_t)t * ld + i] * v;
                            acc[4 + q] += a[(int64_t)(k0 + t) * ld + i] * v;
                        }
                    }
                    const double sq = butterfly8(acc, lane);
                    if ((lane & 7) == 0) {
                        const int q = (lane >> 3) & 7, t = t0 + (q & 3);
                        if (t < jj) {
                            if (q < 4) q1[t] = sq;
                            else q2[t] = sq;
                        }
                    }
                }
            } else {
                for (int t0 = 0; t0 < jj; t0 += 4) {
                    double acc[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) acc[q] = 0.0;
                    for (int i = k + 1 + threadIdx.x; i < m; i += kT) {
                        const double v = vs[i];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int t = min(t0 + q, jj - 1);
                            acc[q] += W[(int64_t)t * ld + i] * v;
                            acc[4 + q] += a[(int64_t)(k0 + t) * ld + i] * v;
                        }
                    }
                    const double sq = butterfly8(acc, lane);
                    if ((lane & 7) == 0) qred[wid * 8 + ((lane >> 3) & 7)] = sq;
                    __syncthreads();
                    if ((int)threadIdx.x < 8) {
                        const int q = threadIdx.x & 3;
                        double t8 = 0.0;
                        for (int w = 0; w < kWaves; ++w) t8 += qred[w * 8 + threadIdx.x];
                        if (t0 + q < jj) {
                            if (threadIdx.x < 4) q1[t0 + q] = t8;
                            else q2[t0 + q] = t8;
                        }
                    }
                    __syncthreads();
                }
            }
            // (e) praw = A22 v from the lower triangle only (A22 is symmetric
            // and only its lower triangle is kept up to date).  Vertical
            // strips of 16 columns c0 .. c0+15 (c0 = base + 16 s, base = k+1
            // rounded down to 16, so that column segments start on 64-byte
            // lines) over the rows c0 .. m-1 in blocks of 64 (one row per
            // lane); wavefront wid takes the strips wid, wid + kWaves, ...
            // Per element A(i, c) one load and two FMAs: the row sum
            // y_i += A(i, c) v_c and the column partial z_c += A(i, c) v_i,
            // kept in registers down the strip.  The row sums of a block go
            // to the wavefront's partial vector cp[wid][:]; the 16 column
            // partials are summed over the wavefront once per strip by a
            // butterfly.  praw = sum of the partial vectors in fixed order.
            for (int i = k + 1 + threadIdx.x; i < m; i += kT)
#pragma unroll
                for (int w = 0; w < kWaves; ++w) cp[w * m + i] = 0.0;
            __syncthreads();
            {
                // strip width: 16 columns for float storage, 8 for FP64
                // storage (fewer registers per wavefront: two workgroups per
                // CU keep more loads in flight)
                constexpr int kSW = sizeof(T) == 8 ? 8 : 16;
                const int base = (k + 1) & ~(kSW - 1);
                const int nstrip = (m - base + kSW - 1) / kSW;
                double *cpw = cp + wid * m;
                const unsigned uld = (unsigned)ld;
                // loads of one 64 x kSW block through 32-bit offsets from the
                // submatrix block, all in flight together (no branch per
                // element); the last row block and the last strip clamp to
                // in-bounds addresses (their excess is masked below)
                auto fetch = [&](int r, int c0, T (&raw)[kSW]) {
                    const unsigned ic = (unsigned)min(r + lane, m - 1);
#pragma unroll
                    for (int q = 0; q < kSW; ++q) raw[q] = a[(unsigned)min(c0 + q, m - 1) * uld + ic];
                };
                // with float storage the next block's loads are issued before
                // this block's arithmetic
                constexpr bool kAhead = sizeof(T) == 4;
                for (int sidx = wid; sidx < nstrip; sidx += kWaves) {
                    const int c0 = base + sidx * kSW;
                    // v of the strip's columns: in registers for FP64 storage;
                    // for float storage (three workgroups per CU, fewer
                    // registers) read from LDS per block (broadcast reads)
                    constexpr bool kVcReg = sizeof(T) == 8;
                    double vc[kSW], zc[kSW];
#pragma unroll
                    for (int q = 0; q < kSW; ++q) {
                        const int c = c0 + q;
                        if (kVcReg) vc[q] = (c > k && c < m) ? vs[c] : 0.0;
                        zc[q] = 0.0;
                    }
                    const bool cols_in = c0 > k && c0 + kSW <= m;  // no column of the strip masked
                    T raw[kSW];
                    fetch(c0, c0, raw);
                    for (int r = c0; r < m; r += 64) {
                        T rawn[kSW];
                        if (kAhead) fetch(min(r + 64, m - 1), c0, rawn);  // past the end: clamped, unused
                        const int i = r + lane;
                        double yrow = 0.0;
                        if (cols_in && r > c0 && r + 64 <= m) {
                            // interior block: strictly below the diagonal, all rows
                            // and columns inside the trailing matrix
                            const double vi = vs[i];
#pragma unroll
                            for (int q = 0; q < kSW; ++q) {
                                const double x = (double)raw[q];
                                yrow += x * (kVcReg ? vc[q] : vs[c0 + q]);
                                zc[q] += x * vi;
                            }
                            cpw[i] += yrow;
                        } else {
                            const bool ok = i < m && i > k;
                            const double vi = ok ? vs[i] : 0.0;
#pragma unroll
                            for (int q = 0; q < kSW; ++q) {
                                const int c = c0 + q;
                                const double x = (ok && c <= i && c > k) ? (double)raw[q] : 0.0;
                                yrow += x * (kVcReg ? vc[q] : vs[min(max(c, k + 1), m - 1)]);
                                if (c < i) zc[q] += x * vi;
                            }
                            if (ok) cpw[i] += yrow;
                        }
                        if (kAhead) {
#pragma unroll
                            for (int q = 0; q < kSW; ++q) raw[q] = rawn[q];
                        } else if (r + 64 < m) {
                            fetch(r + 64, c0, raw);
                        }
                    }
                    // column partials summed over the wavefront: sum q ends in
                    // the lanes with (lane >> 2) & 15 == q (16 wide) or
                    // (lane >> 3) & 7 == q (8 wide)
                    if constexpr (kSW == 16) {
                        const double colsum = butterfly16(zc, lane);
                        const int cq = c0 + ((lane >> 2) & 15);
                        if ((lane & 3) == 0 && cq < m && cq > k) cpw[cq] += colsum;
                    } else {
                        const double colsum = butterfly8(zc, lane);
                        const int cq = c0 + ((lane >> 3) & 7);
                        if ((lane & 7) == 0 && cq < m && cq > k) cpw[cq] += colsum;
                    }
                }
            }
            __syncthreads();
            for (int i = k + 1 + threadIdx.x; i < m; i += kT) {
                double sacc = 0.0;
#pragma unroll
                for (int w = 0; w < kWaves; ++w) sacc += cp[w * m + i];
                praw[i] = sacc;
            }
            __syncthreads();
            // (f) p = tau (praw - V q1 - W q2);  w = p - 0.5 tau (p^T v) v
            double pv = 0.0;
            for (int i = k + 1 + threadIdx.x; i < m; i += kT) {
                double p = praw[i];
#pragma unroll 4
                for (int t = 0; t < jj; ++t)
                    p -= a[(int64_t)(k0 + t) * ld + i] * q1[t] + W[(int64_t)t * ld + i] * q2[t];
                p *= tk;
                W[(int64_t)jj * ld + i] = (T)p;
                pv += p * vs[i];
            }
            const double ptv = block_reduce(pv, red);
            const double acoef = -0.5 * tk * ptv;
            for (int i = k + 1 + threadIdx.x; i < m; i += kT) {
                const double w = W[(int64_t)jj * ld + i] + acoef * vs[i];
                W[(int64_t)jj * ld + i] = (T)w;
                W[(int64_t)(2 * kNB + jj) * ld + i] = (T)w;
            }
            for (int i = threadIdx.x; i <= k; i += kT) W[(int64_t)jj * ld + i] = T(0);
            __syncthreads();
        }
    }
}

// d, e of the last two rows once every panel and trailing update is done
template <typename T>
__global__ void k_sytrd_finish(const SubDesc *__restrict__ desc, const T *__restrict__ Aa,
                               double *__restrict__ d_all, double *__restrict__ e_all,
                               double *__restrict__ tau_all, int nsm) {
    const int si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si >= nsm) return;
    const SubDesc ds = desc[si];
    const int m = ds.m, ld = ds.ld;
    const T *a = Aa + ds.off;
    double *dd = d_all + ds.mapoff, *ee = e_all + ds.mapoff, *tt = tau_all + ds.mapoff;
    if (m == 1) {
        dd[0] = a[0];
        ee[0] = 0.0;
        tt[0] = 0.0;
        return;
    }
    dd[m - 2] = a[(int64_t)(m - 2) * ld + m - 2];
    dd[m - 1] = a[(int64_t)(m - 1) * ld + m - 1];
    ee[m - 2] = a[(int64_t)(m - 2) * ld + m - 1];
    tt[m - 2] = 0.0;
    tt[m - 1] = 0.0;
    ee[m - 1] = 0.0;
}


// ---- Large matrices: the same panel scheme, one launch per step of each
// column over the whole GPU (the batched kernel gives one workgroup to a
// submatrix and holds its vectors in LDS, which bounds m and, for a single
// large matrix, leaves all but one CU idle).  Per column k of a panel:
//   k_big_col      x = column k with the pending terms applied (rows > k),
//                  d_k, and per-workgroup parts of sigma = |x(k+2:)|^2;
//   k_big_reflect  the reflector from the summed parts (every workgroup
//                  forms the same scalars), v stored in A and the panel
//                  buffer, and per-workgroup parts of W^T v and V^T v;
//   k_big_symv     A22 v from the lower triangle in 64 x 64 tiles: tile
//                  (R, S) adds its row sums to PR[S][rows of R] and, for its
//                  elements strictly below the diagonal, its column sums to
//                  PQ[R][columns of S];
//   k_big_p        p = tau (A22 v - V (W^T v) - W (V^T v)) from the partials
//                  in fixed order, and parts of p^T v;
//   k_big_w        w = p - tau/2 (p^T v) v into the panel buffer.
// Every sum is formed in a fixed order, so the result is reproducible.  The
// trailing rank-2 kNB update after each panel is the batched GEMM.
constexpr int kBR = 256;  // rows per workgroup of the row kernels
// the largest m whose vectors fit the batched kernel's 160 KB of LDS
constexpr int kSmallMaxM = (160 * 1024 / 8 - kWaves - 4 * kNB - kWaves * 8) / (1 + kWaves);
// a lone matrix above this goes through the multi-workgroup path
constexpr int kLoneBigM = 2048;
constexpr int kBT = 64;   // tile edge of k_big_symv

template <typename T>
__global__ __launch_bounds__(kBR) void k_big_zero_w(T *__restrict__ W, int m, int ld, int kb) {
    for (int64_t l = (int64_t)blockIdx.x * kBR + threadIdx.x; l < (int64_t)(kNB - kb) * m;
         l += (int64_t)gridDim.x * kBR) {
        const int jj = kb + (int)(l / m), i = (int)(l % m);
        W[(int64_t)jj * ld + i] = T(0);
        W[(int64_t)(kNB + jj) * ld + i] = T(0);
        W[(int64_t)(2 * kNB + jj) * ld + i] = T(0);
    }
}

template <typename T>
__global__ __launch_bounds__(kBR) void k_big_col(const T *__restrict__ a, int m, int ld,
                                                 const T *__restrict__ W, int k0, int jj,
                                                 double *__restrict__ vs, double *__restrict__ psig,
                                                 double *__restrict__ scal, double *__restrict__ dd) {
    __shared__ double c1[kNB], c2[kNB], red[kWaves];
    const int k = k0 + jj;
    if ((int)threadIdx.x < jj) {
        c1[threadIdx.x] = W[(int64_t)threadIdx.x * ld + k];
        c2[threadIdx.x] = a[(int64_t)(k0 + threadIdx.x) * ld + k];
    }
    __syncthreads();
    const int i = k + 1 + blockIdx.x * kBR + threadIdx.x;
    double part = 0.0;
    if (i < m) {
        double x = a[(int64_t)k * ld + i];
        for (int t = 0; t < jj; ++t)
            x -= a[(int64_t)(k0 + t) * ld + i] * c1[t] + W[(int64_t)t * ld + i] * c2[t];
        vs[i] = x;
        if (i > k + 1) part = x * x;
        else scal[0] = x;  // alpha, read by every workgroup of the next kernel
    }
    const double tot = block_reduce(part, red);
    if (threadIdx.x == 0) psig[blockIdx.x] = tot;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double dkk = a[(int64_t)k * ld + k];
        for (int t = 0; t < jj; ++t) dkk -= 2.0 * c2[t] * c1[t];
        dd[k] = dkk;
    }
}

template <typename T>
__global__ __launch_bounds__(kBR) void k_big_reflect(T *__restrict__ a, int m, int ld,
                                                     T *__restrict__ W, int k0, int jj,
                                                     double *__restrict__ vs,
                                                     const double *__restrict__ psig, int nblk,
                                                     double *__restrict__ scal,
                                                     double *__restrict__ ee,
                                                     double *__restrict__ tt,
                                                     double *__restrict__ pq) {
    __shared__ double sig, wred[kWaves][2 * kNB];
    const int k = k0 + jj;
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += psig[b];
        sig = s;
    }
    __syncthreads();
    const double sigma = sig, alpha = scal[0];
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
    const int i = k + 1 + blockIdx.x * kBR + threadIdx.x;
    double v = 0.0;
    if (i < m) {
        v = (i == k + 1) ? 1.0 : vs[i] * scale;
        vs[i] = v;
        a[(int64_t)k * ld + i] = (T)v;
        W[(int64_t)(kNB + jj) * ld + i] = (T)v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ee[k] = beta;
        tt[k] = tk;
        scal[1] = tk;
    }
    // parts of q1[t] = W(:, t)^T v and q2[t] = V(:, k0 + t)^T v over this
    // workgroup's rows: sum t of a wavefront ends in the lanes with
    // (lane >> 2) & 15 == t
    double p1[kNB], p2[kNB];
#pragma unroll
    for (int t = 0; t < kNB; ++t) {
        const bool on = i < m && t < jj;
        p1[t] = on ? (double)W[(int64_t)t * ld + i] * v : 0.0;
        p2[t] = on ? (double)a[(int64_t)(k0 + t) * ld + i] * v : 0.0;
    }
    const double s1 = butterfly16(p1, lane);
    const double s2 = butterfly16(p2, lane);
    if ((lane & 3) == 0) {
        wred[wid][(lane >> 2) & 15] = s1;
        wred[wid][kNB + ((lane >> 2) & 15)] = s2;
    }
    __syncthreads();
    if ((int)threadIdx.x < 2 * kNB) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s += wred[w][threadIdx.x];
        pq[(int64_t)blockIdx.x * 2 * kNB + threadIdx.x] = s;
    }
}

// tile t of the lower triangle of block rows/columns [R0, R0 + nb): (R, S),
// R >= S, in row-major order of the pairs
template <typename T>
__global__ __launch_bounds__(kBR) void k_big_symv(const T *__restrict__ a, int m, int ld, int k,
                                                  const double *__restrict__ vs, int R0,
                                                  double *__restrict__ PR, double *__restrict__ PQ) {
    __shared__ double ysh[kWaves][kBT];
    const int t = blockIdx.x;
    int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    const int R = R0 + r, S = R0 + (t - r * (r + 1) / 2);
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
    const int row = R * kBT + lane;
    const int c0 = S * kBT + wid * kNB;
    const unsigned uld = (unsigned)ld, ir = (unsigned)min(row, m - 1);
    T raw[kNB];
#pragma unroll
    for (int q = 0; q < kNB; ++q) raw[q] = a[(unsigned)min(c0 + q, m - 1) * uld + ir];
    const bool rok = row < m && row > k;
    const double vi = rok ? vs[row] : 0.0;
    double y = 0.0, z[kNB];
#pragma unroll
    for (int q = 0; q < kNB; ++q) {
        const int c = c0 + q;
        const bool on = rok && c > k && c <= row;
        const double x = on ? (double)raw[q] : 0.0;
        y += x * vs[min(max(c, k + 1), m - 1)];
        z[q] = c < row ? x * vi : 0.0;
    }
    ysh[wid][lane] = y;
    const double zs = butterfly16(z, lane);
    const int cq = c0 + ((lane >> 2) & 15);
    if ((lane & 3) == 0 && cq < m && cq > k) PQ[(int64_t)R * m + cq] = zs;
    __syncthreads();
    if (wid == 0 && rok)
        PR[(int64_t)S * m + row] = ((ysh[0][lane] + ysh[1][lane]) + ysh[2][lane]) + ysh[3][lane];
}

template <typename T>
__global__ __launch_bounds__(kBR) void k_big_p(const T *__restrict__ a, int m, int ld,
                                               T *__restrict__ W, int k0, int jj,
                                               const double *__restrict__ vs,
                                               const double *__restrict__ PR,
                                               const double *__restrict__ PQ, int R0, int nbr,
                                               const double *__restrict__ pq, int nblk,
                                               const double *__restrict__ scal,
                                               double *__restrict__ ppv) {
    __shared__ double q[2 * kNB], red[kWaves];
    const int k = k0 + jj;
    if ((int)threadIdx.x < 2 * kNB) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += pq[(int64_t)b * 2 * kNB + threadIdx.x];
        q[threadIdx.x] = s;
    }
    __syncthreads();
    const double tk = scal[1];
    const int i = k + 1 + blockIdx.x * kBR + threadIdx.x;
    double part = 0.0;
    if (i < m) {
        const int Ri = i / kBT;
        // the partials in a fixed order, eight loads in flight at a time
        double s = 0.0;
        auto sum_rows = [&](const double *P, int b0, int b1) {
            int b = b0;
            for (; b + 8 <= b1; b += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = P[(int64_t)(b + u) * m + i];
#pragma unroll
                for (int u = 0; u < 8; ++u) s += v[u];
            }
            for (; b < b1; ++b) s += P[(int64_t)b * m + i];
        };
        sum_rows(PR, R0, Ri + 1);
        sum_rows(PQ, Ri, R0 + nbr);
        double p = s;
        for (int t = 0; t < jj; ++t)
            p -= a[(int64_t)(k0 + t) * ld + i] * q[t] + W[(int64_t)t * ld + i] * q[kNB + t];
        p *= tk;
        W[(int64_t)jj * ld + i] = (T)p;
        part = p * vs[i];
    }
    const double tot = block_reduce(part, red);
    if (threadIdx.x == 0) ppv[blockIdx.x] = tot;
}

template <typename T>
__global__ __launch_bounds__(kBR) void k_big_w(T *__restrict__ W, int m, int ld, int k0, int jj,
                                               const double *__restrict__ vs,
                                               const double *__restrict__ ppv, int nblk,
                                               const double *__restrict__ scal) {
    __shared__ double sh;
    const int k = k0 + jj;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += ppv[b];
        sh = s;
    }
    __syncthreads();
    const double acoef = -0.5 * scal[1] * sh;
    for (int i = blockIdx.x * kBR + threadIdx.x; i < m; i += gridDim.x * kBR) {
        if (i > k) {
            const double w = (double)W[(int64_t)jj * ld + i] + acoef * vs[i];
            W[(int64_t)jj * ld + i] = (T)w;
            W[(int64_t)(2 * kNB + jj) * ld + i] = (T)w;
        } else {
            W[(int64_t)jj * ld + i] = T(0);
        }
    }
}


}  // namespace

namespace ptb {

// Panel by panel: the panel kernel reduces kNB = 16 columns of every submatrix,
// then one batched GEMM applies the rank-32 update A(s:,s:) -= [W V] [V W]^T to
// the lower triangle.
namespace {
// One large submatrix (descriptor x, its device descriptor dx) by the kernels
// above; output as the batched path's.
template <typename T>
int sytrd_big(ptbgpu_ctx *ctx, const SubDesc &x, const SubDesc *dx, T *A, double *d, double *e,
              double *tau) {
    constexpr bool f32 = sizeof(T) == 4;
    ProfScope kprof_(ctx, "sytrd_big");
    const int m = x.m, ld = x.ld;
    T *a = A + x.off;
    double *dd = d + x.mapoff, *ee = e + x.mapoff, *tt = tau + x.mapoff;
    const int nbt = (m + kBT - 1) / kBT;         // 64-row blocks
    const int gmax = (m + kBR - 1) / kBR;        // row workgroups at most
    const size_t nsc = (size_t)m + 2 * (size_t)nbt * m + (size_t)gmax * (2 + 2 * kNB) + 4;
    auto *sc = static_cast<double *>(ws(ctx, "sybig_scr", sizeof(double) * nsc));
    auto *W = static_cast<T *>(ws(ctx, f32 ? "sybig_pw32" : "sybig_pw", sizeof(T) * (size_t)ld * 3 * kNB));
    if (!sc || !W) return PTBGPU_E_NOMEM;
    double *vs = sc, *PR = vs + m, *PQ = PR + (size_t)nbt * m, *psig = PQ + (size_t)nbt * m;
    double *ppv = psig + gmax, *pq = ppv + gmax, *scal = pq + (size_t)gmax * 2 * kNB;
    const int npanel = m > 2 ? (m - 2 + kNB - 1) / kNB : 0;
    std::vector<GemmDesc> g;
    for (int p = 0; p < npanel; ++p) {
        const int k0 = p * kNB, kb = std::min(kNB, m - 2 - k0), s0 = k0 + kb, n = m - s0;
        T *C = a + (int64_t)s0 * ld + s0;
        g.push_back({reinterpret_cast<const double *>(W + s0),
                     reinterpret_cast<const double *>(W + (int64_t)kNB * ld + s0),
                     reinterpret_cast<double *>(C), n, n, 2 * kNB, ld, ld, ld, -1.0, 1.0, 1, 0});
    }
    GemmDesc *dg = nullptr;
    TileList tl;
    if (npanel > 0) {
        dg = static_cast<GemmDesc *>(upload_table(ctx, f32 ? "sybig_g32" : "sybig_g", g.data(),
                                                  sizeof(GemmDesc) * g.size()));
        if (!dg) return PTBGPU_E_NOMEM;
        for (int p = 0; p < npanel; ++p) tl.add(g.data() + p, 1);
        PTB_TRY(tl.upload(ctx, f32 ? "sybig_w32" : "sybig_w"));
    }
    hipStream_t st = ctx->stream;
    for (int p = 0; p < npanel; ++p) {
        const int k0 = p * kNB, kb = std::min(kNB, m - 2 - k0);
        if (kb < kNB) {
            hipLaunchKernelGGL(k_big_zero_w<T>, dim3(64), dim3(kBR), 0, st, W, m, ld, kb);
            PTB_CHECK_LAUNCH(ctx);
        }
        for (int jj = 0; jj < kb; ++jj) {
            const int k = k0 + jj;
            const int nrow = m - k - 1;
            const int G = (nrow + kBR - 1) / kBR;
            const int R0 = (k + 1) / kBT, nbr = nbt - R0;
            const unsigned ntile = (unsigned)(nbr * (nbr + 1) / 2);
            hipLaunchKernelGGL(k_big_col<T>, dim3(G), dim3(kBR), 0, st, a, m, ld, W, k0, jj, vs, psig,
                               scal, dd);
            hipLaunchKernelGGL(k_big_reflect<T>, dim3(G), dim3(kBR), 0, st, a, m, ld, W, k0, jj, vs,
                               psig, G, scal, ee, tt, pq);
            hipLaunchKernelGGL(k_big_symv<T>, dim3(ntile), dim3(kBR), 0, st, a, m, ld, k, vs, R0, PR,
                               PQ);
            hipLaunchKernelGGL(k_big_p<T>, dim3(G), dim3(kBR), 0, st, a, m, ld, W, k0, jj, vs, PR, PQ,
                               R0, nbr, pq, G, scal, ppv);
            hipLaunchKernelGGL(k_big_w<T>, dim3(std::min(G + 1, 256)), dim3(kBR), 0, st, W, m, ld, k0,
                               jj, vs, ppv, G, scal);
            PTB_CHECK_LAUNCH(ctx);
        }
        PTB_TRY(tl.launch(ctx, p, false, true, dg + p, "sytrd", f32));
    }
    hipLaunchKernelGGL(k_sytrd_finish<T>, dim3(1), dim3(256), 0, st, dx, A, d, e, tau, 1);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

template <typename T>
int sytrd_impl(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc, T *A,
               double *d, double *e, double *tau) {
    constexpr bool f32 = sizeof(T) == 4;
    ProfScope prof_(ctx, "sytrd");
    const int nsm = (int)desc.size();
    // Submatrices whose vectors do not fit the batched kernel's LDS, and a
    // lone matrix above kLoneBigM (one workgroup would leave the GPU idle),
    // take the multi-workgroup path; the rest stay batched.
    {
        std::vector<SubDesc> rest;
        int nbig = 0;
        for (int si = 0; si < nsm; ++si) {
            const int m = desc[si].m;
            if (m > kSmallMaxM || (nsm == 1 && m > kLoneBigM)) {
                PTB_TRY(sytrd_big<T>(ctx, desc[si], ddesc + si, A, d, e, tau));
                ++nbig;
            } else {
                rest.push_back(desc[si]);
            }
        }
        if (nbig == nsm) return PTBGPU_OK;
        if (nbig > 0) {
            auto *drest = static_cast<SubDesc *>(upload_table(ctx, f32 ? "sytrd_rest_f" : "sytrd_rest",
                                                              rest.data(), sizeof(SubDesc) * rest.size()));
            if (!drest) return PTBGPU_E_NOMEM;
            return sytrd_impl(ctx, rest, drest, A, d, e, tau);
        }
    }
    int max_m = 0;
    for (const auto &x : desc) max_m = std::max(max_m, x.m);
    const size_t shmem = sizeof(double) * ((1 + kWaves) * (size_t)max_m + kWaves + 4 * kNB + kWaves * 8);
    if (shmem > 160 * 1024) return fail(ctx, PTBGPU_E_ARG, "submatrix too large for sytrd (%d)", max_m);
    const int npanel = max_m > 2 ? (max_m - 2 + kNB - 1) / kNB : 0;
    std::vector<int64_t> pwoff(nsm);
    int64_t pwtot = 0;
    for (int si = 0; si < nsm; ++si) {
        pwoff[si] = pwtot;
        pwtot += (int64_t)desc[si].ld * 3 * kNB;
    }
    auto *PW = static_cast<T *>(ws(ctx, f32 ? "sytrd_pw32" : "sytrd_pw", sizeof(T) * std::max<int64_t>(pwtot, 1)));
    auto *dpwoff = static_cast<int64_t *>(upload_table(ctx, f32 ? "sytrd_pwoff_f" : "sytrd_pwoff", pwoff.data(), sizeof(int64_t) * nsm));
    if (!PW || !dpwoff) return PTBGPU_E_NOMEM;
    std::vector<int> order(nsm);
    for (int si = 0; si < nsm; ++si) order[si] = si;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return desc[a].m > desc[b].m; });
    auto *dorder = static_cast<int *>(upload_table(ctx, f32 ? "sytrd_order_f" : "sytrd_order", order.data(), sizeof(int) * nsm));
    if (!dorder) return PTBGPU_E_NOMEM;
    std::vector<GemmDesc> g1;
    std::vector<int> start(npanel + 1, 0);
    for (int p = 0; p < npanel; ++p) {
        start[p] = (int)g1.size();
        for (int si = 0; si < nsm; ++si) {
            const SubDesc &x = desc[si];
            const int k0 = p * kNB;
            if (k0 >= x.m - 2) continue;
            const int kb = std::min(kNB, x.m - 2 - k0);
            const int s0 = k0 + kb, n = x.m - s0;
            if (n <= 0) continue;
            T *pw = PW + pwoff[si] + s0;
            T *C = A + x.off + (int64_t)s0 * x.ld + s0;
            // C -= [W V] [V W]^T (lower triangle); float storage goes through
            // the descriptor's pointers unchanged (the kernel reads them as T)
            g1.push_back({reinterpret_cast<const double *>(pw),
                          reinterpret_cast<const double *>(pw + (int64_t)kNB * x.ld),
                          reinterpret_cast<double *>(C), n, n, 2 * kNB, x.ld, x.ld, x.ld, -1.0, 1.0,
                          1, 0});
        }
    }
    start[npanel] = (int)g1.size();
    auto *dg1 = static_cast<GemmDesc *>(upload_table(ctx, f32 ? "sytrd_g1f" : "sytrd_g1", g1.data(),
                                                     sizeof(GemmDesc) * g1.size()));
    if (!dg1) return PTBGPU_E_NOMEM;
    TileList tl;
    for (int p = 0; p < npanel; ++p) tl.add(g1.data() + start[p], start[p + 1] - start[p]);
    PTB_TRY(tl.upload(ctx, f32 ? "sytrd_w1f" : "sytrd_w1"));
    for (int p = 0; p < npanel; ++p) {
        {
            ProfScope kprof_(ctx, f32 ? "sytrd_panel_f32" : "sytrd_panel");
            hipLaunchKernelGGL(k_sytrd_panel<T>, dim3(nsm), dim3(kT), shmem, ctx->stream, ddesc, A, PW,
                               d, e, tau, dpwoff, dorder, p);
            PTB_CHECK_LAUNCH(ctx);
        }
        PTB_TRY(tl.launch(ctx, p, false, true, dg1 + start[p], "sytrd", f32));
    }
    hipLaunchKernelGGL(k_sytrd_finish<T>, dim3((nsm + 255) / 256), dim3(256), 0, ctx->stream, ddesc, A,
                       d, e, tau, nsm);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}
}  // namespace

int sytrd_batch(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                double *A, double *d, double *e, double *tau) {
    return sytrd_impl(ctx, desc, ddesc, A, d, e, tau);
}

int sytrd_batch_f32(ptbgpu_ctx *ctx, const std::vector<SubDesc> &desc, const SubDesc *ddesc,
                    float *A, double *d, double *e, double *tau) {
    return sytrd_impl(ctx, desc, ddesc, A, d, e, tau);
}

}  // namespace ptb
