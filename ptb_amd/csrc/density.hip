// Electron-count weights, chemical-potential bisection and the owned-column
// density scatter (source/sm.f90:304-362, :190-198, :365-383, :387).
//
// fe(k) = sum_{owned rows i} (Ssm U)(i,k) U(i,k).  With Ssm = L L^T and
// U = L^-T Y this is sum_i (L Y)(i,k) U(i,k), so the overlap block is not
// needed after the Cholesky: fe_rows computes (L Y)(owned, :) from L and Y
// before the back-substitution, fe_finish multiplies with U(owned, :) after it,
// accumulating owned rows in the reference's order.
//
// The bisection runs in one 1024-thread workgroup over every (eigenvalue, fe)
// pair of the batch with fixed-order reductions, so it is deterministic; it
// reproduces the reference's decisions: mu = (mua+mub)/2 from [-10, 10],
// occupations 2/1/0 (tie e == mu -> 1), stop on |mua-mub| < 1e-3 or
// |nel-nelref| < 1e-5 with both literals rounded to single precision, then one
// more pass at the same mu.
//
// The density of owned column j is P(:, j) = U diag(occ) U(loc_j, :)^T over the
// occupied states only; packed P (zeroed first) receives rows map(a) <= j, which
// is the reference's packsym of P(map, comb) with no symmetrisation.
#include "sm_internal.hpp"

namespace {

constexpr int kT = 256;
constexpr int kG = 8;  // owned columns per pass

// Lown[g][t] = L(oloc_g, t) for t <= oloc_g (0 above): contiguous owned rows of L
__global__ __launch_bounds__(kT) void k_own_rows(const SubDesc *__restrict__ desc,
                                                 const int *__restrict__ oloc,
                                                 const double *__restrict__ La, int rowld,
                                                 double *__restrict__ Lown) {
    const SubDesc d = desc[blockIdx.y];
    const int g = d.gbeg + blockIdx.x;
    if (g >= d.gend) return;
    const int row = oloc[g];
    const double *L = La + d.off;
    double *out = Lown + (int64_t)g * rowld;
    for (int t = threadIdx.x; t < d.m; t += kT) out[t] = t <= row ? L[(int64_t)t * d.ld + row] : 0.0;
}

// SY[g][k] = sum_t Lown[g][t] Y(t, k): one wavefront per column k, kG owned rows
// at a time
__global__ __launch_bounds__(kT) void k_fe_rows(const SubDesc *__restrict__ desc,
                                                const double *__restrict__ Lown, int rowld,
                                                const double *__restrict__ Ya,
                                                double *__restrict__ SY) {
    const SubDesc d = desc[blockIdx.y];
    const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
    const int k = blockIdx.x * (kT / PTB_WAVE) + wid;
    if (k >= d.m) return;
    const double *Y = Ya + d.off + (int64_t)k * d.ld;
    for (int g0 = d.gbeg; g0 < d.gend; g0 += kG) {
        double acc[kG];
#pragma unroll
        for (int q = 0; q < kG; ++q) acc[q] = 0.0;
        for (int t = lane; t < d.m; t += PTB_WAVE) {
            const double y = Y[t];
#pragma unroll
            for (int q = 0; q < kG; ++q)
                if (g0 + q < d.gend) acc[q] += Lown[(int64_t)(g0 + q) * rowld + t] * y;
        }
#pragma unroll
        for (int q = 0; q < kG; ++q) {
            const double s = wave_sum(acc[q]);
            if (lane == 0 && g0 + q < d.gend) SY[(int64_t)(g0 + q) * rowld + k] = s;
        }
    }
}

// fe(k) = sum_g SY[g][k] U(oloc_g, k) in owned-column order; zero for a
// submatrix whose overlap block failed the Cholesky
__global__ __launch_bounds__(kT) void k_fe_finish(const SubDesc *__restrict__ desc,
                                                  const int *__restrict__ oloc,
                                                  const double *__restrict__ SY, int rowld,
                                                  const double *__restrict__ Ua,
                                                  double *__restrict__ fe_all) {
    const SubDesc d = desc[blockIdx.y];
    const int k = blockIdx.x * kT + threadIdx.x;
    if (k >= d.m) return;
    const double *U = Ua + d.off + (int64_t)k * d.ld;
    double s = 0.0;
    if (!(d.flags & 1))
        for (int g = d.gbeg; g < d.gend; ++g) s += SY[(int64_t)g * rowld + k] * U[oloc[g]];
    fe_all[d.mapoff + k] = s;
}

__device__ __forceinline__ double occ_of(double e, double mu) {
    return e < mu ? 2.0 : (e == mu ? 1.0 : 0.0);
}

constexpr int kBT = 1024;

// out: [0] mu, [1] nel at mu, [2] number of logged passes, then the trace
// (mu, nel) of every logged pass from out[4]
__global__ __launch_bounds__(kBT) void k_bisect(const double *__restrict__ e,
                                                const double *__restrict__ fe, int64_t total,
                                                double nelref, double *__restrict__ out) {
    __shared__ double red[kBT / PTB_WAVE];
    const float tol_mu = 1e-3f, tol_nel = 1e-5f;  // REAL(4) literals of the reference
    double mua = -10.0, mub = 10.0, mu = 0.0, nel = 0.0;
    bool finaliter = false;
    int logged = 0;
    for (int muiter = 1; muiter <= 100; ++muiter) {
        mu = (mua + mub) / 2.0;
        double part = 0.0;
        for (int64_t i = threadIdx.x; i < total; i += kBT) part += fe[i] * occ_of(e[i], mu);
        part = wave_sum(part);
        const int lane = threadIdx.x % PTB_WAVE, wid = threadIdx.x / PTB_WAVE;
        __syncthreads();
        if (lane == 0) red[wid] = part;
        __syncthreads();
        nel = 0.0;
        for (int w = 0; w < kBT / PTB_WAVE; ++w) nel += red[w];
        if (finaliter) break;
        if (threadIdx.x == 0) {
            out[4 + 2 * logged] = mu;
            out[5 + 2 * logged] = nel;
        }
        ++logged;
        if (fabs(mua - mub) < (double)tol_mu || fabs(nel - nelref) < (double)tol_nel) finaliter = true;
        if (!finaliter) {
            if (nel > nelref) mub = mu;
            else mua = mu;
        }
    }
    if (threadIdx.x == 0) {
        out[0] = mu;
        out[1] = nel;
        out[2] = (double)logged;
        out[3] = finaliter ? 1.0 : 0.0;
    }
}

// grid: (submatrix, chunk of kG owned columns)
__global__ __launch_bounds__(kT) void k_dmat_scatter(const SubDesc *__restrict__ desc,
                                                     const int *__restrict__ gcols,
                                                     const int *__restrict__ oloc,
                                                     const int *__restrict__ map_all,
                                                     const double *__restrict__ eig_all,
                                                     const double *__restrict__ Ua,
                                                     const double *__restrict__ mu_p,
                                                     double *__restrict__ P) {
    extern __shared__ double wsh[];  // [kG][m] weights then int list of occupied k
    const SubDesc d = desc[blockIdx.x];
    const int g0 = d.gbeg + blockIdx.y * kG;
    if (g0 >= d.gend || (d.flags & 1)) return;
    const int ng = min(kG, d.gend - g0);
    const int m = d.m;
    const double mu = mu_p[0];
    const double *eig = eig_all + d.mapoff;
    const double *U = Ua + d.off;
    const int *map = map_all + d.mapoff;
    int *occk = reinterpret_cast<int *>(wsh + (size_t)kG * m);
    __shared__ int nocc;
    if (threadIdx.x == 0) {
        int c = 0;
        for (int k = 0; k < m; ++k)
            if (occ_of(eig[k], mu) != 0.0) occk[c++] = k;
        nocc = c;
    }
    __syncthreads();
    const int no = nocc;
    for (int l = threadIdx.x; l < kG * no; l += kT) {
        const int q = l / no, kk = l % no;
        const int k = occk[kk];
        wsh[(size_t)q * m + kk] =
            q < ng ? occ_of(eig[k], mu) * U[(int64_t)k * d.ld + oloc[g0 + q]] : 0.0;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < m; a += kT) {
        double acc[kG];
#pragma unroll
        for (int q = 0; q < kG; ++q) acc[q] = 0.0;
        for (int kk = 0; kk < no; ++kk) {
            const double u = U[(int64_t)occk[kk] * d.ld + a];
#pragma unroll
            for (int q = 0; q < kG; ++q) acc[q] += u * wsh[(size_t)q * m + kk];
        }
        const int64_t row = map[a];
#pragma unroll
        for (int q = 0; q < kG; ++q) {
            if (q < ng) {
                const int64_t j = gcols[g0 + q];
                if (row <= j) P[pk(row, j)] = acc[q];
            }
        }
    }
}

}  // namespace

namespace ptb {

int fe_owned_rows(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, int max_own,
                  const int *doloc, const double *L, const double *Y, double *Lown, double *SY) {
    ProfScope prof_(ctx, "fe_rows");
    hipLaunchKernelGGL(k_own_rows, dim3(max_own, nsm), dim3(kT), 0, ctx->stream, ddesc, doloc, L,
                       max_m, Lown);
    PTB_CHECK_LAUNCH(ctx);
    const int wpb = kT / PTB_WAVE;
    hipLaunchKernelGGL(k_fe_rows, dim3((max_m + wpb - 1) / wpb, nsm), dim3(kT), 0, ctx->stream,
                       ddesc, Lown, max_m, Y, SY);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int fe_finish(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, const int *doloc,
              const double *SY, const double *U, double *fe) {
    ProfScope prof_(ctx, "fe_finish");
    hipLaunchKernelGGL(k_fe_finish, dim3((max_m + kT - 1) / kT, nsm), dim3(kT), 0, ctx->stream,
                       ddesc, doloc, SY, max_m, U, fe);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int bisect_mu(ptbgpu_ctx *ctx, const double *e, const double *fe, int64_t total, int64_t nelref,
              double *dout) {
    ProfScope prof_(ctx, "bisect");
    hipLaunchKernelGGL(k_bisect, dim3(1), dim3(kBT), 0, ctx->stream, e, fe, total, (double)nelref,
                       dout);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

int dmat_scatter(ptbgpu_ctx *ctx, const SubDesc *ddesc, int nsm, int max_m, int max_own,
                 const int *dgcols, const int *doloc, const int *dmap, const double *eig,
                 const double *U, const double *dmu, double *P) {
    ProfScope prof_(ctx, "dmat_scatter");
    const size_t shmem = sizeof(double) * (size_t)kG * max_m + sizeof(int) * (size_t)max_m;
    if (shmem > 150 * 1024) return fail(ctx, PTBGPU_E_ARG, "submatrix too large for dmat (%d)", max_m);
    hipLaunchKernelGGL(k_dmat_scatter, dim3(nsm, (max_own + kG - 1) / kG), dim3(kT), shmem,
                       ctx->stream, ddesc, dgcols, doloc, dmap, eig, U, dmu, P);
    PTB_CHECK_LAUNCH(ctx);
    return PTBGPU_OK;
}

}  // namespace ptb
