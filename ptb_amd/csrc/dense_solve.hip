// Dense (one-matrix) eigen routes of the reference:
//  * matrix_sign / matrix_root (source/sm.f90:410-457): V f(e) V^T with
//    f = signum or 1/sqrt from dsyevd of A;
//  * solve2 (source/pgtb.f90:1066-1170): the -check reference density, full
//    dsygvd of (H, S), the iteration-2 eigenvalue shift, Fermi smearing of the
//    alpha and beta occupations (FERMISMEAR, occu: source/fermi.f:5-124) and
//    P = U diag(focc) U^T.
// They run the batched eigensolver of the submatrix build on a batch of one
// (above 2048 rows its multi-workgroup tridiagonalisation and device-scratch
// merges; ndim <= kDenseMax, so that n^2 stays below 2^31).  Smearing is O(ndim) scalar work on the host, as in the
// reference.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "sm_internal.hpp"

namespace {

constexpr int kT = 256;

__global__ void k_unpack_ld(const double *__restrict__ p, int64_t n, double *__restrict__ A,
                            int64_t ld) {
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x) {
        const double *col = p + j * (j + 1) / 2;
        for (int64_t i = threadIdx.x; i <= j; i += kT) {
            const double v = col[i];
            A[j * ld + i] = v;
            A[i * ld + j] = v;
        }
    }
}

// the symmetric matrix dsyevd('U') sees: B(i,j) = M(min, max), ld n -> ld ldb
__global__ void k_load_upper(const double *__restrict__ M, int64_t n, double *__restrict__ B,
                             int64_t ldb) {
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x)
        for (int64_t i = threadIdx.x; i < n; i += kT)
            B[j * ldb + i] = i <= j ? M[j * n + i] : M[i * n + j];
}

// C(:, j) = A(:, perm[j]) * w[j]  (A with leading dimension lda, C with n)
__global__ void k_cols_scaled(const double *__restrict__ A, int64_t lda, const int *__restrict__ perm,
                              const double *__restrict__ w, int64_t n, double *__restrict__ C) {
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x) {
        const double *src = A + (int64_t)perm[j] * lda;
        const double s = w ? w[j] : 1.0;
        for (int64_t i = threadIdx.x; i < n; i += kT) C[j * n + i] = src[i] * s;
    }
}

// packsym of a dense symmetric result: packed (j, i), j <= i, from column i
__global__ void k_pack(const double *__restrict__ A, int64_t n, double *__restrict__ p) {
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        double *col = p + i * (i + 1) / 2;
        for (int64_t j = threadIdx.x; j <= i; j += kT) col[j] = A[i * n + j];
    }
}

int grid_of(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>(n, 1), 2048); }

// one submatrix covering the whole matrix
std::vector<SubDesc> one_desc(int64_t n) {
    SubDesc d{};
    d.off = 0;
    d.m = (int)n;
    d.ld = (int)((n + 15) & ~15);
    d.mapoff = 0;
    d.gbeg = 0;
    d.gend = (int)n;
    d.flags = 0;
    return {d};
}

struct Arena {
    std::vector<SubDesc> desc;
    SubDesc *ddesc = nullptr;
    double *A = nullptr, *B = nullptr, *Z = nullptr, *W = nullptr, *X = nullptr, *vec = nullptr;
    int *flags = nullptr;
    int ld = 0;
};

int arena_setup(ptbgpu_ctx *ctx, int64_t n, const char *pfx, Arena &a) {
    a.desc = one_desc(n);
    a.ld = a.desc[0].ld;
    const size_t ab = sizeof(double) * (size_t)a.ld * (size_t)n;
    const std::string p(pfx);
    a.A = static_cast<double *>(ptb::ws(ctx, (p + "A").c_str(), ab));
    a.B = static_cast<double *>(ptb::ws(ctx, (p + "B").c_str(), ab));
    a.Z = static_cast<double *>(ptb::ws(ctx, (p + "Z").c_str(), ab));
    a.W = static_cast<double *>(ptb::ws(ctx, (p + "W").c_str(), ab));
    a.X = static_cast<double *>(ptb::ws(ctx, (p + "X").c_str(), ab));
    a.vec = static_cast<double *>(ptb::ws(ctx, (p + "vec").c_str(), sizeof(double) * 8 * n + 4096));
    a.flags = static_cast<int *>(ptb::ws(ctx, (p + "flags").c_str(), sizeof(int)));
    a.ddesc = static_cast<SubDesc *>(ptb::ws(ctx, (p + "desc").c_str(), sizeof(SubDesc)));
    if (!a.A || !a.B || !a.Z || !a.W || !a.X || !a.vec || !a.flags || !a.ddesc) return PTBGPU_E_NOMEM;
    PTB_HIP(ctx, hipMemcpyAsync(a.ddesc, a.desc.data(), sizeof(SubDesc), hipMemcpyHostToDevice,
                                ctx->stream));
    return PTBGPU_OK;
}

// ascending order of the eigenvalues (dsyevd / dsygvd order); a NaN (from a
// failed factorisation) sorts as +inf, so that the order is always defined
std::vector<int> ascending(const std::vector<double> &w) {
    std::vector<int> p(w.size());
    std::iota(p.begin(), p.end(), 0);
    auto key = [&](int i) { return std::isnan(w[i]) ? HUGE_VAL : w[i]; };
    std::stable_sort(p.begin(), p.end(), [&](int a, int b) { return key(a) < key(b); });
    return p;
}

// FERMISMEAR (source/fermi.f:5-67).  BOLTZ and the entropy factor are REAL(4)
// literals in the reference; nel is the index of the highest occupied orbital.
void fermismear(int64_t norbs, int64_t nel, double t, const double *eig, double *occ, double *fod,
                double *e_fermi, double *s) {
    const double boltz = (double)3.166808578545117E-06f;
    const double bkt = boltz * t;
    double ef = 0.5 * (eig[nel - 1] + eig[nel]);
    const double occt = (double)nel;
    int ncycle = 0;
    double total_number = 0.0;
    do {
        total_number = 0.0;
        double total_dfermi = 0.0;
        ++ncycle;
        for (int64_t i = 0; i < norbs; ++i) {
            double f = 0.0, df = 0.0;
            const double x = (eig[i] - ef) / bkt;
            if (x < 50.0) {
                const double ex = std::exp(x);
                f = 1.0 / (ex + 1.0);
                df = ex / (bkt * (ex + 1.0) * (ex + 1.0));
            }
            occ[i] = f;
            total_number += f;
            total_dfermi += df;
        }
        ef += (occt - total_number) / (total_dfermi + 1e-14);
    } while (std::fabs(occt - total_number) > 1e-9 && ncycle < 200);
    double fd = 0.0, ss = 0.0;
    for (int64_t i = 0; i < norbs; ++i) {
        if (occ[i] > 1e-9 && 1.0 - occ[i] > 1e-9)
            ss += occ[i] * std::log(occ[i]) + (1.0 - occ[i]) * std::log(1.0 - occ[i]);
        fd += eig[i] < ef ? 1.0 - occ[i] : occ[i];
    }
    if (fod) *fod = fd;
    if (e_fermi) *e_fermi = ef;
    if (s) *s = ss * boltz * t;
}

// occu (source/fermi.f:69-124): alpha/beta occupations and their highest
// occupied indices (1-based; 0 if none)
int occu(int64_t ndim, int64_t nel, int64_t nopen, std::vector<double> &fa,
         std::vector<double> &fb, int64_t &ia, int64_t &ib) {
    std::vector<int> focc(ndim, 0);
    if (nel % 2 == 0) {
        const int64_t ihomo = nel / 2;
        for (int64_t i = 0; i < ihomo && i < ndim; ++i) focc[i] = 2;
        if (nopen > 1)
            for (int64_t i = 1; i <= nopen / 2; ++i) {
                if (ihomo - i >= 0 && ihomo - i < ndim) focc[ihomo - i] -= 1;
                if (ihomo + i - 1 < ndim) focc[ihomo + i - 1] += 1;
            }
    } else {
        const int64_t na = nel / 2 + (nopen - 1) / 2 + 1, nb = nel / 2 - (nopen - 1) / 2;
        for (int64_t i = 0; i < na && i < ndim; ++i) focc[i] += 1;
        for (int64_t i = 0; i < nb && i < ndim; ++i) focc[i] += 1;
    }
    fa.assign(ndim, 0.0);
    fb.assign(ndim, 0.0);
    ia = ib = 0;
    for (int64_t i = 0; i < ndim; ++i) {
        if (focc[i] == 2) fa[i] = fb[i] = 1.0;
        if (focc[i] == 1) fa[i] = 1.0;
        if (fa[i] > 0.99) ia = i + 1;
        if (fb[i] > 0.99) ib = i + 1;
    }
    return ia < 1 ? PTBGPU_E_ARG : PTBGPU_OK;  // the reference stops ('internal error in occu')
}

}  // namespace

namespace ptb {

// Eigenpairs of the symmetric matrix given by the upper triangle of M (n x n,
// leading dimension n, device): eigenvalues ascending into w (host),
// eigenvectors into V (n x n, leading dimension n, device) in the same order.
int dense_eig(ptbgpu_ctx *ctx, const double *M, int64_t n, double *V, std::vector<double> &w,
              const char *pfx) {
    if (n < 1 || n > kDenseMax)
        return fail(ctx, PTBGPU_E_ARG, "dense eigensolver needs 1 <= n <= %d", kDenseMax);
    Arena a;
    PTB_TRY(arena_setup(ctx, n, pfx, a));
    hipLaunchKernelGGL(k_load_upper, dim3(grid_of(n)), dim3(kT), 0, ctx->stream, M, n, a.A,
                       (int64_t)a.ld);
    PTB_CHECK_LAUNCH(ctx);
    double *dvec = a.vec, *evec = a.vec + n, *tau = a.vec + 2 * n, *rot = a.vec + 3 * n;
    PTB_TRY(eig_sym_batch(ctx, a.desc, a.ddesc, a.A, a.Z, a.W, a.X, dvec, evec, tau, rot));
    std::vector<double> raw(n);
    PTB_HIP(ctx, hipMemcpyAsync(raw.data(), dvec, sizeof(double) * n, hipMemcpyDeviceToHost,
                                ctx->stream));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const std::vector<int> perm = ascending(raw);
    w.resize(n);
    for (int64_t i = 0; i < n; ++i) w[i] = raw[perm[i]];
    int *dperm = static_cast<int *>(ws(ctx, (std::string(pfx) + "perm").c_str(), sizeof(int) * n));
    if (!dperm) return PTBGPU_E_NOMEM;
    PTB_HIP(ctx, hipMemcpy(dperm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_cols_scaled, dim3(grid_of(n)), dim3(kT), 0, ctx->stream, a.Z,
                       (int64_t)a.ld, dperm, nullptr, n, V);
    PTB_CHECK_LAUNCH(ctx);
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PTBGPU_OK;
}

}  // namespace ptb

namespace {

// V diag(f) V^T for the eigenpairs of the upper triangle of the dense A
int spectral_function(ptbgpu_ctx *ctx, int64_t n, const double *dA, double *dOut, int which) {
    const size_t mb = sizeof(double) * (size_t)n * (size_t)n;
    double *V = static_cast<double *>(ptb::ws(ctx, "ms_V", mb));
    double *T = static_cast<double *>(ptb::ws(ctx, "ms_T", mb));
    double *fw = static_cast<double *>(ptb::ws(ctx, "ms_f", sizeof(double) * n));
    int *id = static_cast<int *>(ptb::ws(ctx, "ms_id", sizeof(int) * n));
    GemmDesc *dg = static_cast<GemmDesc *>(ptb::ws(ctx, "ms_g", sizeof(GemmDesc)));
    if (!V || !T || !fw || !id || !dg) return PTBGPU_E_NOMEM;
    std::vector<double> w;
    PTB_TRY(ptb::dense_eig(ctx, dA, n, V, w, "ms_"));
    std::vector<double> f(n);
    for (int64_t i = 0; i < n; ++i) {
        if (which == 0) f[i] = w[i] > 0.0 ? 1.0 : (w[i] < 0.0 ? -1.0 : 0.0);
        else f[i] = 1.0 / std::sqrt(w[i]);
    }
    std::vector<int> ident(n);
    std::iota(ident.begin(), ident.end(), 0);
    PTB_HIP(ctx, hipMemcpy(fw, f.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    PTB_HIP(ctx, hipMemcpy(id, ident.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_cols_scaled, dim3(grid_of(n)), dim3(kT), 0, ctx->stream, V, n, id, fw, n, T);
    PTB_CHECK_LAUNCH(ctx);
    const GemmDesc g{T, V, dOut, (int)n, (int)n, (int)n, (int)n, (int)n, (int)n, 1.0, 0.0, 0, 0};
    PTB_HIP(ctx, hipMemcpy(dg, &g, sizeof g, hipMemcpyHostToDevice));
    PTB_TRY(ptb::gemm_vbatch(ctx, false, true, dg, 1, (int)n, (int)n, "dense"));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PTBGPU_OK;
}

int spectral_host(ptbgpu_ctx *ctx, int64_t n, const double *A, double *out, int which) {
    if (!ctx) return PTBGPU_E_ARG;
    if (n < 1 || n > kDenseMax || !A || !out)
        return ptb::fail(ctx, PTBGPU_E_ARG, "matrix_sign/matrix_root need 1 <= ndim <= %d", kDenseMax);
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    const size_t mb = sizeof(double) * (size_t)n * (size_t)n;
    double *dA = static_cast<double *>(ptb::ws(ctx, "ms_in", mb));
    double *dO = static_cast<double *>(ptb::ws(ctx, "ms_out", mb));
    if (!dA || !dO) return PTBGPU_E_NOMEM;
    PTB_TRY(ptb::h2d(ctx, dA, A, mb));
    PTB_TRY(spectral_function(ctx, n, dA, dO, which));
    PTB_TRY(ptb::d2h(ctx, out, dO, mb));
    return PTBGPU_OK;
}

}  // namespace

extern "C" {

int ptbgpu_matrix_sign(ptbgpu_ctx *ctx, int64_t ndim, const double *A, double *S) {
    return spectral_host(ctx, ndim, A, S, 0);
}

int ptbgpu_matrix_root(ptbgpu_ctx *ctx, int64_t ndim, const double *A, double *sqrtAinv) {
    return spectral_host(ctx, ndim, A, sqrtAinv, 1);
}

int ptbgpu_fermismear(int64_t norbs, int64_t nel, double t, const double *eig, double *occ,
                      double *fod, double *e_fermi, double *s) {
    if (norbs < 2 || nel < 1 || nel >= norbs || !eig || !occ) return PTBGPU_E_ARG;
    fermismear(norbs, nel, t, eig, occ, fod, e_fermi, s);
    return PTBGPU_OK;
}

int ptbgpu_solve2(ptbgpu_ctx *ctx, int64_t mode, int64_t ndim, int64_t nel, int64_t nopen,
                  double et, double shift0, double shift1, const double *H, const double *S,
                  double *P, double *e, double *focc, double *U, int64_t *fail) {
    if (!ctx) return PTBGPU_E_ARG;
    if (ndim < 1 || ndim > kDenseMax || nel < 0 || !H || !S || !P)
        return ptb::fail(ctx, PTBGPU_E_ARG, "solve2 needs 1 <= ndim <= %d and H, S, P", kDenseMax);
    PTB_HIP(ctx, hipSetDevice(ctx->device));
    ptb::ProfScope prof_(ctx, "solve2");
    const int64_t n = ndim;
    const size_t pb = sizeof(double) * (size_t)(n * (n + 1) / 2), mb = sizeof(double) * (size_t)n * n;
    Arena a;
    PTB_TRY(arena_setup(ctx, n, "s2_", a));
    double *dHp = static_cast<double *>(ptb::ws(ctx, "io_h", pb));
    double *dSp = static_cast<double *>(ptb::ws(ctx, "io_s", pb));
    double *dPp = static_cast<double *>(ptb::ws(ctx, "io_p", pb));
    double *Uo = static_cast<double *>(ptb::ws(ctx, "s2_U", mb));
    double *T = static_cast<double *>(ptb::ws(ctx, "s2_T", mb));
    double *Pd = static_cast<double *>(ptb::ws(ctx, "s2_P", mb));
    double *fw = static_cast<double *>(ptb::ws(ctx, "s2_f", sizeof(double) * n));
    int *dperm = static_cast<int *>(ptb::ws(ctx, "s2_perm", sizeof(int) * n));
    GemmDesc *dg = static_cast<GemmDesc *>(ptb::ws(ctx, "s2_g", sizeof(GemmDesc)));
    if (!dHp || !dSp || !dPp || !Uo || !T || !Pd || !fw || !dperm || !dg) return PTBGPU_E_NOMEM;
    PTB_TRY(ptb::h2d(ctx, dHp, H, pb));
    PTB_TRY(ptb::h2d(ctx, dSp, S, pb));
    hipLaunchKernelGGL(k_unpack_ld, dim3(grid_of(n)), dim3(kT), 0, ctx->stream, dHp, n, a.A,
                       (int64_t)a.ld);
    hipLaunchKernelGGL(k_unpack_ld, dim3(grid_of(n)), dim3(kT), 0, ctx->stream, dSp, n, a.B,
                       (int64_t)a.ld);
    PTB_CHECK_LAUNCH(ctx);
    // dsygvd: S = L L^T, K = L^-1 H L^-T = Y diag(e) Y^T, U = L^-T Y
    PTB_HIP(ctx, hipMemsetAsync(a.flags, 0, sizeof(int), ctx->stream));
    PTB_TRY(ptb::potrf_batch(ctx, a.desc, a.ddesc, a.B, a.flags));
    PTB_TRY(ptb::trsm_gemm(ctx, a.desc, a.ddesc, a.B, a.A, a.Z, false, "s2h"));
    PTB_TRY(ptb::transpose_sym_batch(ctx, a.ddesc, 1, (int)n, a.Z, a.A, false));
    PTB_TRY(ptb::trsm_gemm(ctx, a.desc, a.ddesc, a.B, a.A, a.Z, false, "s2k"));
    PTB_TRY(ptb::transpose_sym_batch(ctx, a.ddesc, 1, (int)n, a.Z, a.A, true));
    double *dvec = a.vec, *evec = a.vec + n, *tau = a.vec + 2 * n, *rot = a.vec + 3 * n;
    PTB_TRY(ptb::eig_sym_batch(ctx, a.desc, a.ddesc, a.A, a.Z, a.W, a.X, dvec, evec, tau, rot));
    PTB_TRY(ptb::trsm_gemm(ctx, a.desc, a.ddesc, a.B, a.Z, a.X, true, "s2u"));
    std::vector<double> raw(n);
    int hflag = 0;
    PTB_HIP(ctx, hipMemcpyAsync(raw.data(), dvec, sizeof(double) * n, hipMemcpyDeviceToHost,
                                ctx->stream));
    PTB_HIP(ctx, hipMemcpyAsync(&hflag, a.flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const std::vector<int> perm = ascending(raw);
    std::vector<double> ev(n);
    for (int64_t i = 0; i < n; ++i) ev[i] = raw[perm[i]];
    // iteration-2 shift to absolute orbital energies (source/pgtb.f90:1139-1143)
    if (mode >= 2)
        for (auto &x : ev) x = x + shift0 + shift1 * x;
    // Fermi smearing of the alpha and beta occupations
    std::vector<double> fa, fbeta, oc(n, 0.0);
    int64_t ia = 0, ib = 0;
    if (nel > 0) {
        if (occu(n, nel, nopen, fa, fbeta, ia, ib) != PTBGPU_OK)
            return ptb::fail(ctx, PTBGPU_E_ARG, "internal error in occu");
    } else {
        fa.assign(n, 0.0);
        fbeta.assign(n, 0.0);
    }
    if (ia + 1 <= n && ia >= 1) fermismear(n, ia, et, ev.data(), fa.data(), nullptr, nullptr, nullptr);
    if (ib + 1 <= n && ib >= 1 && nel > 1)
        fermismear(n, ib, et, ev.data(), fbeta.data(), nullptr, nullptr, nullptr);
    for (int64_t i = 0; i < n; ++i) oc[i] = fa[i] + fbeta[i];
    // U with ascending columns; P = U diag(focc) U^T
    PTB_HIP(ctx, hipMemcpyAsync(dperm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                                ctx->stream));
    PTB_HIP(ctx, hipMemcpyAsync(fw, oc.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                                ctx->stream));
    hipLaunchKernelGGL(k_cols_scaled, dim3(grid_of(n)), dim3(kT), 0, ctx->stream, a.X,
                       (int64_t)a.ld, dperm, nullptr, n, Uo);
    hipLaunchKernelGGL(k_cols_scaled, dim3(grid_of(n)), dim3(kT), 0, ctx->stream, a.X,
                       (int64_t)a.ld, dperm, fw, n, T);
    PTB_CHECK_LAUNCH(ctx);
    const GemmDesc g{T, Uo, Pd, (int)n, (int)n, (int)n, (int)n, (int)n, (int)n, 1.0, 0.0, 0, 0};
    PTB_HIP(ctx, hipMemcpyAsync(dg, &g, sizeof g, hipMemcpyHostToDevice, ctx->stream));
    PTB_TRY(ptb::gemm_vbatch(ctx, false, true, dg, 1, (int)n, (int)n, "dense"));
    hipLaunchKernelGGL(k_pack, dim3(grid_of(n)), dim3(kT), 0, ctx->stream, Pd, n, dPp);
    PTB_CHECK_LAUNCH(ctx);
    PTB_TRY(ptb::d2h(ctx, P, dPp, pb));
    if (U) PTB_TRY(ptb::d2h(ctx, U, Uo, mb));
    PTB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (e) std::copy(ev.begin(), ev.end(), e);
    if (focc) std::copy(oc.begin(), oc.end(), focc);
    if (fail) *fail = hflag ? 1 : 0;  // dsygvd INFO != 0
    return PTBGPU_OK;
}

}  // extern "C"
