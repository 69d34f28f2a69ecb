// Divide-and-conquer eigensolver for the symmetric tridiagonal T of every
// submatrix (Cuppen's method with Gu-Eisenstat / Loewner eigenvectors; the
// algorithm of LAPACK dstedc, which dsygvd uses).
//
// Tree: each submatrix's index range is halved recursively until a piece has at
// most kDcLeaf rows.  A split at sp couples rows sp-1 and sp through beta = e[sp-1]:
//   T = diag(T1 - |beta| e_last e_last^T, T2 - |beta| e_1 e_1^T) + |beta| u u^T,
//   u = [e_last ; sgn(beta) e_1].
// Leaves: implicit QL with Wilkinson-type shifts (tql2), four leaves per
// wavefront (kDcLeaf lanes each), lane r owning row r of its leaf's eigenvector
// block in LDS.
// Merge of (Q1,D1),(Q2,D2): z = [Q1(last,:); sgn Q2(first,:)] / sqrt 2,
// rho = 2|beta|; sort D; deflate (small rho|z_i| or close pairs rotated together);
// solve the secular equation 1 + rho sum w_i^2/(d_i - lam) = 0 for
// lam - d_origin (origin = nearer pole, so that d_i - lam is formed accurately)
// with rational two-pole model steps inside a shrinking bracket;
// Loewner vector zhat; U(i,j) = zhat_i/(d_i - lam_j) normalised; the merged
// eigenvectors are Qnd U (a GEMM launched for all nodes of a tree level; this
// kernel writes its descriptor).  Eigenvalues are kept in the same column order
// as the eigenvectors (not sorted).
#include <cfloat>

#include "sm_internal.hpp"

namespace {

constexpr int kMT = 256;
constexpr double kEps = DBL_EPSILON;

// Leaves per wavefront: kDcLeaf lanes each.  Every lane of a group runs the
// leaf's scalar recurrence (same values, same branches), so one wavefront
// issues the work of four leaves; the groups diverge, so the group's LDS
// hand-offs are ordered by wavefront-scope fences instead of barriers (the
// lanes of a wavefront are one scope, and the group's lanes are converged).
constexpr int kLeafPack = PTB_WAVE / kDcLeaf;

__device__ __forceinline__ void group_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(PTB_WAVE) void k_dc_leaf(const SubDesc *__restrict__ desc,
                                                      const DcNode *__restrict__ leaves, int nleaves,
                                                      double *__restrict__ d_all,
                                                      const double *__restrict__ e_all,
                                                      double *__restrict__ Za) {
    __shared__ double zs[kLeafPack][kDcLeaf][kDcLeaf + 1];
    __shared__ double dls[kLeafPack][kDcLeaf], els[kLeafPack][kDcLeaf];
    const int grp = threadIdx.x / kDcLeaf, lane = threadIdx.x % kDcLeaf;
    const int li = blockIdx.x * kLeafPack + grp;
    if (li >= nleaves) return;
    double(*z)[kDcLeaf + 1] = zs[grp];
    double *dl = dls[grp], *el = els[grp];
    const DcNode nd = leaves[li];
    const SubDesc ds = desc[nd.s];
    const int n = nd.hi - nd.lo;
    const double *e = e_all + ds.mapoff;
    double *d = d_all + ds.mapoff;
    if (lane < n) {
        double dv = d[nd.lo + lane];
        if (lane == 0 && nd.lo > 0) dv -= fabs(e[nd.lo - 1]);
        if (lane == n - 1 && nd.hi < ds.m) dv -= fabs(e[nd.hi - 1]);
        dl[lane] = dv;
        el[lane] = (lane < n - 1) ? e[nd.lo + lane] : 0.0;
    }
    for (int c = 0; c < kDcLeaf; ++c) z[lane][c] = (lane == c) ? 1.0 : 0.0;
    group_lds_sync();
    // every lane of the group runs the scalar recurrence on the shared d/e
    // (same values); lane r applies each rotation to row r of z
    for (int l = 0; l < n; ++l) {
        for (int iter = 0; iter < 60; ++iter) {
            int mm = l;
            for (; mm < n - 1; ++mm) {
                const double dd = fabs(dl[mm]) + fabs(dl[mm + 1]);
                if (fabs(el[mm]) <= kEps * dd) break;
            }
            if (mm == l) break;
            double g = (dl[l + 1] - dl[l]) / (2.0 * el[l]);
            double r = hypot(g, 1.0);
            g = dl[mm] - dl[l] + el[l] / (g + copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool under = false;
            for (int i = mm - 1; i >= l; --i) {
                const double f = s * el[i], b = c * el[i];
                r = hypot(f, g);
                group_lds_sync();
                if (lane == 0) el[i + 1] = r;
                if (r == 0.0) {
                    if (lane == 0) {
                        dl[i + 1] -= p;
                        el[mm] = 0.0;
                    }
                    group_lds_sync();
                    under = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = dl[i + 1] - p;
                r = (dl[i] - g) * s + 2.0 * c * b;
                p = s * r;
                if (lane == 0) dl[i + 1] = g + p;
                g = c * r - b;
                if (lane < n) {
                    const double zi = z[lane][i], zi1 = z[lane][i + 1];
                    z[lane][i + 1] = s * zi + c * zi1;
                    z[lane][i] = c * zi - s * zi1;
                }
                group_lds_sync();
            }
            if (under) continue;
            group_lds_sync();
            if (lane == 0) {
                dl[l] -= p;
                el[l] = g;
                el[mm] = 0.0;
            }
            group_lds_sync();
        }
    }
    group_lds_sync();
    if (lane < n) d[nd.lo + lane] = dl[lane];
    double *Z = Za + ds.off;
    for (int c = 0; c < n; ++c)
        if (lane < n) Z[(int64_t)(nd.lo + c) * ds.ld + nd.lo + lane] = z[lane][c];
}

// bitonic sort of (key, idx) pairs in LDS, ascending, ties by idx, by a
// workgroup of NT threads
template <int NT>
__device__ void bitonic_sort(double *key, int *idx, int npow) {
    for (int size = 2; size <= npow; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = threadIdx.x; t < npow / 2; t += NT) {
                const int lo = 2 * stride * (t / stride) + (t % stride);
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const double ka = key[lo], kb = key[hi];
                const int ia = idx[lo], ib = idx[hi];
                const bool gt = (ka > kb) || (ka == kb && ia > ib);
                if (gt == up) {
                    key[lo] = kb;
                    key[hi] = ka;
                    idx[lo] = ib;
                    idx[hi] = ia;
                }
            }
        }
    }
    __syncthreads();
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double *red) {
    v = wave_max(v);
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = red[0];
    for (int w = 1; w < NT / PTB_WAVE; ++w) t = fmax(t, red[w]);
    __syncthreads();
    return t;
}

// 1/x: the hardware reciprocal estimate refined by two Newton-Raphson steps
// r <- r + r (1 - x r), each of which squares the relative error, so the
// result is within about an ulp of the rounded quotient for normal finite x
// (the divide sequence costs twice as many instructions in these O(K^2) loops)
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// s = sum_{i0 <= i < i1} q_i, ds = sum q_i r_i with r_i = 1 / ((d_i - dorg) - t),
// q_i = w2_i r_i: the secular function's pole terms and their derivatives.
// Four interleaved partial sums, so that the terms' dependent chains
// (reciprocal, Newton steps, accumulation) overlap.
__device__ __forceinline__ void secular_sums(const double *dl, const double *w2, int i0, int i1,
                                             double dorg, double t, double &s, double &ds) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
    int i = i0;
    for (; i + 4 <= i1; i += 4) {
        const double r0 = rcp64((dl[i] - dorg) - t), r1 = rcp64((dl[i + 1] - dorg) - t);
        const double r2 = rcp64((dl[i + 2] - dorg) - t), r3 = rcp64((dl[i + 3] - dorg) - t);
        const double q0 = w2[i] * r0, q1 = w2[i + 1] * r1, q2 = w2[i + 2] * r2, q3 = w2[i + 3] * r3;
        a0 += q0;
        a1 += q1;
        a2 += q2;
        a3 += q3;
        b0 += q0 * r0;
        b1 += q1 * r1;
        b2 += q2 * r2;
        b3 += q3 * r3;
    }
    for (; i < i1; ++i) {
        const double r = rcp64((dl[i] - dorg) - t), q = w2[i] * r;
        a0 += q;
        b0 += q * r;
    }
    s = (a0 + a1) + (a2 + a3);
    ds = (b0 + b1) + (b2 + b3);
}

// dst(wb:we) = src(wb:we) for one column by one wavefront, 8 loads in flight
// per lane; rows outside [rb, re) are known to be zero and are written, not read
__device__ __forceinline__ void copy_column(double *__restrict__ dst,
                                            const double *__restrict__ src, int wb, int we,
                                            int lane, int rb = 0, int re = 1 << 30) {
    constexpr int kU = 8;
    for (int r0 = wb; r0 < we; r0 += kU * PTB_WAVE) {
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int r = r0 + u * PTB_WAVE + lane;
            v[u] = (r < we && r >= rb && r < re) ? src[r] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int r = r0 + u * PTB_WAVE + lane;
            if (r < we) dst[r] = v[u];
        }
    }
}

// Secular root j of a merge with K poles dl (ascending), weights w and
// w2 = rho w^2: the root's origin pole org[j] and offset tau[j].
__device__ void dc_root(int j, int K, const double *dl, const double *w2, const double *w,
                        double rho, int *org, double *tau) {
    int o;
    double lo, hi;
    if (j < K - 1) {
        const double mid = 0.5 * (dl[j + 1] - dl[j]);
        double fs, fds;
        secular_sums(dl, w2, 0, K, dl[j], mid, fs, fds);
        const double f = 1.0 + fs;
        if (f >= 0.0) {
            o = j;
            lo = 0.0;
            hi = mid;
        } else {
            o = j + 1;
            lo = -mid;
            hi = 0.0;
        }
    } else {
        o = j;
        double s2 = 0.0;
        for (int i = 0; i < K; ++i) s2 += w[i] * w[i];
        lo = 0.0;
        hi = rho * s2;
    }
    // f(t) = 1 + rho sum w_i^2 / (delta_i - t) (delta_i = d_i - d_origin) is
    // increasing between its poles.  Each step fits the two poles that
    // bracket the root exactly, a + b1/(D1 - eta) + b2/(D2 - eta) matching
    // f and the derivatives of the parts left (psi) and right (phi) of the
    // root ("middle way" rational model, as LAPACK dlaed4), and takes the
    // model's root when it lies inside the bracket, else a Newton step or
    // the midpoint.  Every evaluation also shrinks the bracket; stop when
    // |f| is at rounding level or the bracket reaches adjacent doubles.
    const double dorg = dl[o];
    const bool last = (j == K - 1);
    const double da = dl[j] - dorg, db = last ? 0.0 : dl[j + 1] - dorg;
    double t = 0.5 * (lo + hi);
    for (int it = 0; it < 200; ++it) {
        double psi, dpsi, phi, dphi;
        secular_sums(dl, w2, 0, j + 1, dorg, t, psi, dpsi);
        secular_sums(dl, w2, j + 1, K, dorg, t, phi, dphi);
        const double f = 1.0 + psi + phi;
        if (f == 0.0 || fabs(f) <= 8.0 * kEps * (double)K * (1.0 + fabs(psi) + fabs(phi))) {
            lo = hi = t;
            break;
        }
        if (f < 0.0) lo = t;
        else hi = t;
        const double d1 = da - t;  // distance to the left pole
        double tn;
        if (last) {
            // one-pole model a + b1/(d1 - eta): eta = d1 + b1/a
            const double b1 = d1 * d1 * dpsi;
            const double a = f - b1 / d1;
            tn = (a != 0.0) ? t + d1 + b1 / a : 0.5 * (lo + hi);
        } else {
            const double d2 = db - t;  // distance to the right pole
            const double b1 = d1 * d1 * dpsi, b2 = d2 * d2 * dphi;
            const double a = f - b1 / d1 - b2 / d2;
            // a eta^2 - B eta + C = 0
            const double B = a * (d1 + d2) + b1 + b2;
            const double C = a * d1 * d2 + b1 * d2 + b2 * d1;
            // the root between the poles is the small one, 2C / (B + sgn(B) sqrt),
            // formed without cancellation; the other one is tried if it is not
            const double sq = sqrt(fmax(B * B - 4.0 * a * C, 0.0));
            const double den = B >= 0.0 ? B + sq : B - sq;
            tn = (den != 0.0) ? t + 2.0 * C / den : t;
            if (!(tn > lo && tn < hi) && a != 0.0) tn = t + den / (2.0 * a);
        }
        if (!(tn > lo && tn < hi)) {
            tn = t - f / (dpsi + dphi);
            if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        }
        if (tn == t || tn == lo || tn == hi) break;
        t = tn;
    }
    org[j] = o;
    tau[j] = (lo == hi) ? lo : t;
}

// Loewner vector entry i: rho zhat_i^2 = (lam_i - d_i) prod_{j != i} (lam_j - d_i)/(d_j - d_i)
__device__ double dc_loewner(int i, int K, const double *dl, const int *org, const double *tau,
                             const double *w, double rho) {
    const double di = dl[i];
    double prod = -((di - dl[org[i]]) - tau[i]);
    for (int j = 0; j < K; ++j) {
        if (j == i) continue;
        const double dlt = (di - dl[org[j]]) - tau[j];  // d_i - lam_j
        prod *= (-dlt) * rcp64(dl[j] - di);
    }
    return copysign(sqrt(fmax(prod, 0.0) / rho), w[i]);
}

// Merged eigenvector column j by one wavefront: U(:, j) = zhat ./ (d - lam_j)
// normalised, written to xc at the rows prow; the eigenvalue lam_j to Dn[j].
__device__ void dc_vector(int j, int lane, int K, const double *dl, const int *org,
                          const double *tau, const double *zh, const int *prow, double *xc,
                          double *Dn) {
    const double dorg = dl[org[j]], tj = tau[j];
    double nrm = 0.0;
    for (int i = lane; i < K; i += PTB_WAVE) {
        const double u = zh[i] * rcp64((dl[i] - dorg) - tj);
        nrm += u * u;
    }
    const double ci = 1.0 / sqrt(wave_sum(nrm));
    for (int i = lane; i < K; i += PTB_WAVE) xc[prow[i]] = ci * zh[i] * rcp64((dl[i] - dorg) - tj);
    if (lane == 0) Dn[j] = dorg + tj;
}

// exclusive prefix sum of one int per thread over the workgroup of NT
// threads; *total gets the sum of all (wsum: NT / PTB_WAVE ints of LDS)
template <int NT>
__device__ int block_scan_excl(int v, int *wsum, int *total) {
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
    int inc = v;
    for (int off = 1; off < PTB_WAVE; off <<= 1) {
        const int o = __shfl_up(inc, off, PTB_WAVE);
        if (lane >= off) inc += o;
    }
    __syncthreads();
    if (lane == PTB_WAVE - 1) wsum[wid] = inc;
    __syncthreads();
    int base = 0, tot = 0;
    for (int w = 0; w < NT / PTB_WAVE; ++w) {
        if (w < wid) base += wsum[w];
        tot += wsum[w];
    }
    *total = tot;
    return base + inc - v;
}

// One merge node per workgroup.  Scratch: W node block receives the sorted,
// rotated, compacted eigenvector columns [nondeflated | deflated]; X node block
// receives U (K x K); rot holds the deflation rotations (4 doubles each).
//
// A left child's column is zero below the split and a right child's above it;
// a deflation rotation between a left and a right column makes both dense.
// The non-deflated columns are placed [left-only | dense | right-only] (pole
// order within each group), with U's rows in the same order, so that the
// merged vectors are two products of half the height:
//   rows above the split:  W(0:n1, left+dense) U(left+dense, :),
//   rows below it:         W(n1:n, dense+right) U(dense+right, :)
// (about half the flops of W(:, 0:K) U; the structure dlaed3 exploits).  The
// kernel writes both descriptors, gd[2 node] and gd[2 node + 1].
//
// The per-node vectors (6 doubles and 4 ints per row, padded to the power of
// two npow) live in LDS up to npow = 2048; larger nodes (kGlobal) keep them
// in a device scratch of 8 npow doubles per node at gscr, which only this
// workgroup touches (its barriers order the accesses as they do in LDS).
template <bool kGlobal, int NT>
__global__ __launch_bounds__(NT) void k_dc_merge(const SubDesc *__restrict__ desc,
                                                  const DcNode *__restrict__ nodes, int npow,
                                                  const double *__restrict__ e_all,
                                                  double *__restrict__ D_all,
                                                  double *__restrict__ Za, double *__restrict__ Wa,
                                                  double *__restrict__ Xa,
                                                  double *__restrict__ rot_all,
                                                  GemmDesc *__restrict__ gd, double *gscr,
                                                  int *gK) {
    extern __shared__ double shm_[];
    double *shm;
    if constexpr (kGlobal) shm = gscr + (size_t)blockIdx.x * 8 * (size_t)npow;
    else shm = shm_;
    const DcNode nd = nodes[blockIdx.x];
    const SubDesc ds = desc[nd.s];
    const int n = nd.hi - nd.lo, n1 = nd.sp - nd.lo, ld = ds.ld;
    double *Dk = shm;               // npow sorted keys
    double *z = Dk + npow;          // n (sorted order after step 3)
    double *dl = z + npow;          // K non-deflated poles
    double *w = dl + npow;          // K weights
    double *tau = w + npow;         // K root offsets
    double *zh = tau + npow;        // K Loewner vector (rho w_i^2 before that)
    int *ix = reinterpret_cast<int *>(zh + npow);  // npow sort permutation
    int *slot = ix + npow;          // n output slot of sorted position
    int *org = slot + npow;         // K sorted position of each pole, then origin pole of each root
    int *prow = org + npow;         // K W slot (and U row) of each pole
    __shared__ double red[NT / PTB_WAVE];
    __shared__ int wsum[NT / PTB_WAVE];
    __shared__ int sK, sNr, sClose, sKL, sKD;

    const double beta = e_all[ds.mapoff + nd.sp - 1];
    const double rho = 2.0 * fabs(beta);
    const double sgn = beta >= 0.0 ? 1.0 : -1.0;
    double *Z = Za + ds.off, *W = Wa + ds.off, *X = Xa + ds.off;
    double *D = D_all + ds.mapoff;
    double *rot = rot_all + 4 * (int64_t)(ds.mapoff + nd.lo);
    GemmDesc *g = gd + 2 * blockIdx.x;
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
    constexpr int kWaves = NT / PTB_WAVE;

    if (rho == 0.0) {  // decoupled: children are already eigenpairs of T
        if (kGlobal && threadIdx.x == 0) gK[blockIdx.x] = 0;
        if (threadIdx.x < 2) {
            g[threadIdx.x].M = 0;
            g[threadIdx.x].N = 0;
            g[threadIdx.x].K = 0;
            g[threadIdx.x].C32 = nullptr;
        }
        // the node block is diag(Z1, Z2): zero its off-diagonal blocks
        for (int c = wid; c < n; c += kWaves) {
            double *zc = Z + (int64_t)(nd.lo + c) * ld + nd.lo;
            const int rb = c < n1 ? n1 : 0, re = c < n1 ? n : n1;
            for (int r = rb + lane; r < re; r += PTB_WAVE) zc[r] = 0.0;
        }
        return;
    }
    const double rs2 = 1.0 / sqrt(2.0);
    double *zraw = tau;  // temporary: raw z by column before sorting
    // Keys: the eigenvalues, a NaN as +inf; padding +inf.  Ties go by index,
    // so the n real entries always sort ahead of the padding and the sorted
    // positions 0..n-1 name only columns of this node, whatever the data (a
    // NaN key would otherwise leave the order undefined).
    for (int i = threadIdx.x; i < npow; i += NT) {
        if (i < n) {
            const double dv = D[nd.lo + i];
            Dk[i] = dv == dv ? dv : HUGE_VAL;
            ix[i] = i;
            const int col = nd.lo + i;
            const double zv = (i < n1) ? Z[(int64_t)col * ld + nd.sp - 1]
                                       : sgn * Z[(int64_t)col * ld + nd.sp];
            zraw[i] = zv * rs2;
        } else {
            Dk[i] = HUGE_VAL;
            ix[i] = i;
        }
    }
    bitonic_sort<NT>(Dk, ix, npow);
    for (int j = threadIdx.x; j < n; j += NT) z[j] = zraw[ix[j]];
    __syncthreads();
    double mx = 0.0;
    for (int j = threadIdx.x; j < n; j += NT) mx = fmax(mx, fmax(fabs(Dk[j]), fabs(z[j])));
    const double tol = 8.0 * kEps * block_max<NT>(mx, red);

    // Deflation.  Type 1: rho |z_j| <= tol.  Type 2: neighbouring surviving
    // poles so close that a Givens rotation zeroes one weight (|t c s| <= tol);
    // a rotation changes the weight carried to the next test, so that scan is
    // sequential.  When no pair of neighbouring survivors passes the test on
    // the unrotated values the scan rotates nothing and its outcome is the
    // type-1 split, formed here in parallel from a prefix count of survivors:
    // survivor j gets slot = survivors before j, deflated j gets
    // slot = #survivors + deflated before j (the scan's order).
    {
        const int chunk = (n + NT - 1) / NT;
        const int jb = min(n, (int)threadIdx.x * chunk), je = min(n, jb + chunk);
        int cnt = 0;
        for (int j = jb; j < je; ++j) cnt += (rho * fabs(z[j]) > tol);
        int nc;
        int pos = block_scan_excl<NT>(cnt, wsum, &nc);
        if (threadIdx.x == 0) sClose = 0;
        for (int j = jb; j < je; ++j) {
            if (rho * fabs(z[j]) > tol) {
                org[pos] = j;  // survivors in order (org is free until the roots)
                slot[j] = pos++;
            } else {
                slot[j] = nc + (j - pos);
            }
        }
        __syncthreads();
        for (int q = threadIdx.x + 1; q < nc; q += NT) {
            const int pj = org[q - 1], j = org[q];
            double s = z[pj], c = z[j];
            const double tv = hypot(c, s);
            const double t = Dk[j] - Dk[pj];
            c /= tv;
            s = -s / tv;
            if (fabs(t * c * s) <= tol) sClose = 1;
        }
        __syncthreads();
        if (!sClose) {
            for (int q = threadIdx.x; q < nc; q += NT) {
                dl[q] = Dk[org[q]];
                w[q] = z[org[q]];
            }
            if (threadIdx.x == 0) {
                sK = nc;
                sNr = 0;
            }
        }
        __syncthreads();
    }
    // sequential scan (rotations recorded, applied to the columns below)
    if (sClose && threadIdx.x == 0) {
        int K = 0, nd_ = 0, nr = 0, pj = -1;
        // slot: >= 0 final slot for non-deflated (assigned after counting)
        // first pass marks: slot[j] = -1 deflated, -2 non-deflated
        for (int j = 0; j < n; ++j) {
            if (rho * fabs(z[j]) <= tol) {
                slot[j] = -1;
                continue;
            }
            if (pj >= 0) {
                double s = z[pj], c = z[j];
                const double tv = hypot(c, s);
                const double t = Dk[j] - Dk[pj];
                c /= tv;
                s = -s / tv;
                if (fabs(t * c * s) <= tol) {
                    z[j] = tv;
                    z[pj] = 0.0;
                    rot[4 * nr + 0] = (double)pj;
                    rot[4 * nr + 1] = (double)j;
                    rot[4 * nr + 2] = c;
                    rot[4 * nr + 3] = s;
                    ++nr;
                    const double t1 = Dk[pj] * c * c + Dk[j] * s * s;
                    Dk[j] = Dk[pj] * s * s + Dk[j] * c * c;
                    Dk[pj] = t1;
                    slot[pj] = -1;
                    pj = j;
                    continue;
                }
                slot[pj] = -2;
            }
            pj = j;
        }
        if (pj >= 0) slot[pj] = -2;
        for (int j = 0; j < n; ++j)
            if (slot[j] == -2) {
                dl[K] = Dk[j];
                w[K] = z[j];
                org[K] = j;
                slot[j] = K++;
            }
        nd_ = K;
        for (int j = 0; j < n; ++j)
            if (slot[j] == -1) slot[j] = nd_++;
        sK = K;
        sNr = nr;
    }
    __syncthreads();
    const int K = sK, nr = sNr;
    // Column types: bit 0 = rows above the split, bit 1 = rows below, from the
    // child of origin and the rotations in order; bit 2 = touched by a
    // rotation (zh is free until the roots).
    int *cmask = reinterpret_cast<int *>(zh);
    for (int j = threadIdx.x; j < n; j += NT) cmask[j] = ix[j] < n1 ? 1 : 2;
    __syncthreads();
    if (nr > 0 && threadIdx.x == 0)
        for (int q = 0; q < nr; ++q) {
            const int a = (int)rot[4 * q], b = (int)rot[4 * q + 1];
            const int mm = cmask[a] | cmask[b] | 4;
            cmask[a] = mm;
            cmask[b] = mm;
        }
    __syncthreads();
    // slots [left-only | dense | right-only] by prefix counts over the poles
    {
        const int chunk = (K + NT - 1) / NT;
        const int ib = min(K, (int)threadIdx.x * chunk), ie = min(K, ib + chunk);
        int cl = 0, cd = 0;
        for (int i = ib; i < ie; ++i) {
            const int t = cmask[org[i]] & 3;
            cl += t == 1;
            cd += t == 3;
        }
        int KL, KD;
        int pl = block_scan_excl<NT>(cl, wsum, &KL);
        int pd = block_scan_excl<NT>(cd, wsum, &KD);
        int pr = ib - pl - pd;  // right-only poles before ib
        for (int i = ib; i < ie; ++i) {
            const int t = cmask[org[i]] & 3;
            prow[i] = t == 1 ? pl++ : (t == 3 ? KL + pd++ : KL + KD + pr++);
        }
        if (threadIdx.x == 0) {
            sKL = KL;
            sKD = KD;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < K; i += NT) slot[org[i]] = prow[i];
        __syncthreads();
    }
    const int KL = sKL, KD = sKD;
    // sorted, compacted eigenvector columns into W: W(:, slot[j]) = Z(:, lo+ix[j]),
    // one column per wavefront.  A child's column is nonzero only on the
    // child's rows, which are the only rows read.  Left-only and right-only
    // survivors that no rotation touches feed one of the two products, so only
    // their child's rows are written; every other column (rotated row by row
    // below, or copied back whole) gets the other rows as zeros, so the arena
    // needs no clearing beforehand.
    for (int j = wid; j < n; j += kWaves) {
        const int c = ix[j], sl = slot[j];
        const int cb = c < n1 ? 0 : n1, ce = c < n1 ? n1 : n;
        const bool half = !(cmask[j] & 4) && (sl < KL || (sl >= KL + KD && sl < K));
        copy_column(W + (int64_t)(nd.lo + sl) * ld + nd.lo, Z + (int64_t)(nd.lo + c) * ld + nd.lo,
                    half ? cb : 0, half ? ce : n, lane, cb, ce);
    }
    __syncthreads();
    // deflation rotations, in order, row-parallel
    for (int r = threadIdx.x; r < n; r += NT) {
        for (int q = 0; q < nr; ++q) {
            const int cp = nd.lo + slot[(int)rot[4 * q]], cj = nd.lo + slot[(int)rot[4 * q + 1]];
            const double c = rot[4 * q + 2], s = rot[4 * q + 3];
            double *pp = W + (int64_t)cp * ld + nd.lo + r, *pjj = W + (int64_t)cj * ld + nd.lo + r;
            const double qp = *pp, qj = *pjj;
            *pp = c * qp + s * qj;
            *pjj = c * qj - s * qp;
        }
    }
    __syncthreads();
    // deflated columns and eigenvalues go straight to their final place
    for (int c = K + wid; c < n; c += kWaves)
        copy_column(Z + (int64_t)(nd.lo + c) * ld + nd.lo, W + (int64_t)(nd.lo + c) * ld + nd.lo,
                    0, n, lane);
    for (int j = threadIdx.x; j < n; j += NT)
        if (slot[j] >= K) D[nd.lo + slot[j]] = Dk[j];
    __syncthreads();
    double *w2 = zh;
    for (int i = threadIdx.x; i < K; i += NT) w2[i] = rho * w[i] * w[i];
    if (threadIdx.x < 2) {
        // 0: rows above the split from slots [0, KL + KD); 1: rows below it
        // from slots [KL, K)
        const bool top = threadIdx.x == 0;
        const int r0 = top ? 0 : n1, s0 = top ? 0 : KL;
        GemmDesc &gg = g[threadIdx.x];
        gg.A = W + (int64_t)(nd.lo + s0) * ld + nd.lo + r0;
        gg.B = X + (int64_t)nd.lo * ld + nd.lo + s0;
        gg.C = Z + (int64_t)nd.lo * ld + nd.lo + r0;
        gg.M = top ? n1 : n - n1;
        gg.N = K;
        gg.K = top ? KL + KD : K - KL;
        gg.lda = ld;
        gg.ldb = ld;
        gg.ldc = ld;
        gg.alpha = 1.0;
        gg.beta = 0.0;
        gg.uplo = 0;
        gg.tri = 0;
        gg.C32 = nullptr;
    }
    // a node from device scratch leaves the O(K^2) phases to three kernels
    // over the whole chip (k_dc_roots_big, k_dc_loewner_big, k_dc_vectors_big)
    if constexpr (kGlobal) {
        if (threadIdx.x == 0) gK[blockIdx.x] = K;
        return;
    }
    __syncthreads();
    // secular roots, one thread per root
    for (int j = threadIdx.x; j < K; j += NT) dc_root(j, K, dl, w2, w, rho, org, tau);
    __syncthreads();
    // Loewner vector: rho zhat_i^2 = (lam_i - d_i) prod_{j != i} (lam_j - d_i)/(d_j - d_i)
    for (int i = threadIdx.x; i < K; i += NT) zh[i] = dc_loewner(i, K, dl, org, tau, w, rho);
    __syncthreads();
    // U(:, j) = zhat ./ (d - lam_j), normalised; eigenvalues lam_j.  One
    // column per wavefront: its norm, then the column (lanes on rows).
    for (int j = wid; j < K; j += kWaves)
        dc_vector(j, lane, K, dl, org, tau, zh, prow, X + (int64_t)(nd.lo + j) * ld + nd.lo, D + nd.lo);
    __syncthreads();
}


// The O(K^2) phases of the merges run from device scratch, over the whole
// chip: grid (blocks, node of the run); the node's vectors at gscr + 8 npow
// node, laid out as in k_dc_merge, and its K in gK.
struct BigNode {
    const double *dl, *w, *tau;
    double *zh;
    int *org, *prow;
    double rho;
    int K;
};

__device__ BigNode big_node(const SubDesc *desc, const DcNode *nodes, int npow,
                            const double *e_all, double *gscr, const int *gK, int node) {
    const DcNode nd = nodes[node];
    const SubDesc ds = desc[nd.s];
    double *b = gscr + (size_t)node * 8 * (size_t)npow;
    BigNode g;
    g.dl = b + 2 * (size_t)npow;
    g.w = b + 3 * (size_t)npow;
    g.tau = b + 4 * (size_t)npow;
    g.zh = b + 5 * (size_t)npow;
    int *ib = reinterpret_cast<int *>(b + 6 * (size_t)npow);
    g.org = ib + 2 * (size_t)npow;
    g.prow = ib + 3 * (size_t)npow;
    g.rho = 2.0 * fabs(e_all[ds.mapoff + nd.sp - 1]);
    g.K = gK[node];
    return g;
}

__global__ __launch_bounds__(kMT) void k_dc_roots_big(const SubDesc *__restrict__ desc,
                                                      const DcNode *__restrict__ nodes, int npow,
                                                      const double *__restrict__ e_all,
                                                      double *gscr, const int *__restrict__ gK) {
    BigNode g = big_node(desc, nodes, npow, e_all, gscr, gK, blockIdx.y);
    const int j = blockIdx.x * kMT + threadIdx.x;
    if (j < g.K) dc_root(j, g.K, g.dl, g.zh, g.w, g.rho, g.org, const_cast<double *>(g.tau));
}

__global__ __launch_bounds__(kMT) void k_dc_loewner_big(const SubDesc *__restrict__ desc,
                                                        const DcNode *__restrict__ nodes, int npow,
                                                        const double *__restrict__ e_all,
                                                        double *gscr, const int *__restrict__ gK) {
    BigNode g = big_node(desc, nodes, npow, e_all, gscr, gK, blockIdx.y);
    const int i = blockIdx.x * kMT + threadIdx.x;
    if (i < g.K) g.zh[i] = dc_loewner(i, g.K, g.dl, g.org, g.tau, g.w, g.rho);
}

__global__ __launch_bounds__(kMT) void k_dc_vectors_big(const SubDesc *__restrict__ desc,
                                                        const DcNode *__restrict__ nodes, int npow,
                                                        const double *__restrict__ e_all,
                                                        double *__restrict__ D_all,
                                                        double *__restrict__ Xa, double *gscr,
                                                        const int *__restrict__ gK) {
    BigNode g = big_node(desc, nodes, npow, e_all, gscr, gK, blockIdx.y);
    const DcNode nd = nodes[blockIdx.y];
    const SubDesc ds = desc[nd.s];
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
    const int j = blockIdx.x * (kMT / PTB_WAVE) + wid;
    if (j < g.K)
        dc_vector(j, lane, g.K, g.dl, g.org, g.tau, g.zh, g.prow,
                  Xa + ds.off + (int64_t)(nd.lo + j) * ds.ld + nd.lo, D_all + ds.mapoff + nd.lo);
}

}  // namespace

namespace ptb {

int dc_leaves(ptbgpu_ctx *ctx, const SubDesc *ddesc, const DcNode *dleaves, int nleaves,
              double *d, const double *e, double *Z) {
    ProfScope prof_(ctx, "dc_leaf");
    if (nleaves <= 0) return PTBGPU_OK;
    hipLaunchKernelGGL(k_dc_leaf, dim3((nleaves + kLeafPack - 1) / kLeafPack), dim3(PTB_WAVE), 0, ctx->stream,
                       ddesc, dleaves, nleaves, d, e, Z);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int dc_merge_level(ptbgpu_ctx *ctx, const SubDesc *ddesc, const DcNode *dnodes, int count,
                   int max_n, const double *e, double *D, double *Z, double *W, double *X,
                   double *rot, GemmDesc *dgemm, const TileList &tiles, int lowp,
                   const std::vector<int> *sizes, int nfp64) {
    if (count <= 0) return PTBGPU_OK;
    auto pow2 = [](int n) {
        int p = 1;
        while (p < n) p <<= 1;
        return p;
    };
    const int npow_max = pow2(max_n);
    // classes whose vectors exceed the LDS budget run from device scratch
    auto in_lds = [](int np) { return (size_t)np * (6 * sizeof(double) + 4 * sizeof(int)) <= 150 * 1024; };
    if (max_n > (1 << 24)) return fail(ctx, PTBGPU_E_ARG, "merge of %d rows is too large", max_n);
    // Size buckets: the nodes come largest first, and each run of nodes whose
    // power-of-two size class is the same is one launch with LDS for that
    // class, so that small nodes are not held to the largest node's LDS (and
    // occupancy).  Classes below the tuned floor (ctx->dc_bucket_min) join the
    // smallest launched class; a floor of 0 launches the level at once.
    std::vector<std::pair<int, int>> runs;  // (first node, npow)
    const int floor = (int)ctx->dc_bucket_min;
    if (!sizes || floor <= 0) {
        runs.push_back({0, npow_max});
    } else {
        for (int i = 0; i < count; ++i) {
            const int np = std::max(pow2((*sizes)[i]), std::min(pow2(floor), npow_max));
            if (runs.empty() || runs.back().second != np) runs.push_back({i, np});
        }
    }
    {
        ProfScope prof_(ctx, "dc_merge");
        for (size_t r = 0; r < runs.size(); ++r) {
            const int i0 = runs[r].first;
            const int i1 = r + 1 < runs.size() ? runs[r + 1].first : count;
            const int np = runs[r].second;
            const bool global = !in_lds(np) || (ctx->dc_global_min > 0 && np >= ctx->dc_global_min);
            if (!global) {
                // a workgroup of min(max(np / 2, 64), 512) threads per node:
                // the small nodes of the low levels leave fewer threads of a
                // larger group idle at every barrier, and more of them fit
                // per CU (np / 2 measured faster than np and np / 4); nodes
                // of 1024 rows and more, whose LDS allows two or one per CU,
                // fill the CU's wavefront slots with 512 threads
                const size_t shmem = (size_t)np * (6 * sizeof(double) + 4 * sizeof(int));
                const int half = np / 2;
                auto k = half <= 64    ? k_dc_merge<false, 64>
                         : half <= 128 ? k_dc_merge<false, 128>
                         : half <= 256 ? k_dc_merge<false, 256>
                                       : k_dc_merge<false, 512>;
                const int nt = half <= 64 ? 64 : (half <= 128 ? 128 : (half <= 256 ? 256 : 512));
                hipLaunchKernelGGL(k, dim3(i1 - i0), dim3(nt), shmem, ctx->stream, ddesc, dnodes + i0,
                                   np, e, D, Z, W, X, rot, dgemm + 2 * i0, nullptr, nullptr);
            } else {
                const int nn = i1 - i0;
                auto *scr = static_cast<double *>(
                    ws(ctx, "dc_gscr", sizeof(double) * 8 * (size_t)np * (size_t)nn));
                auto *gK = static_cast<int *>(ws(ctx, "dc_gK", sizeof(int) * (size_t)nn));
                if (!scr || !gK) return PTBGPU_E_NOMEM;
                hipLaunchKernelGGL((k_dc_merge<true, kMT>), dim3(nn), dim3(kMT), 0, ctx->stream, ddesc,
                                   dnodes + i0, np, e, D, Z, W, X, rot, dgemm + 2 * i0, scr, gK);
                const unsigned nb = (unsigned)((np + kMT - 1) / kMT);
                hipLaunchKernelGGL(k_dc_roots_big, dim3(nb, nn), dim3(kMT), 0, ctx->stream, ddesc,
                                   dnodes + i0, np, e, scr, gK);
                hipLaunchKernelGGL(k_dc_loewner_big, dim3(nb, nn), dim3(kMT), 0, ctx->stream, ddesc,
                                   dnodes + i0, np, e, scr, gK);
                hipLaunchKernelGGL(k_dc_vectors_big, dim3((unsigned)((np + 3) / 4), nn), dim3(kMT), 0,
                                   ctx->stream, ddesc, dnodes + i0, np, e, D, X, scr, gK);
            }
            PTB_CHECK_LAUNCH(ctx);
        }
    }
    // the first nfp64 nodes' products in FP64 (launch 0 of the tile list),
    // the rest in the low-precision arithmetic (the next launch)
    nfp64 = std::min(std::max(nfp64, 0), count);
    int launch = 0;
    if (nfp64 > 0) PTB_TRY(tiles.launch(ctx, launch++, false, false, dgemm, "dc", false, kLowpNone));
    if (nfp64 < count)
        PTB_TRY(tiles.launch(ctx, launch, false, false, dgemm + 2 * nfp64, "dc", false, lowp));
    return PTBGPU_OK;
}

}  // namespace ptb
