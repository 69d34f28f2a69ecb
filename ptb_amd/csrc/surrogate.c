/*
 * Surrogate generator of the (H, S) pencil that the density build consumes.
 *
 * PTB's own Hamiltonian/integral builder is not part of this engine, so the
 * workloads use this stand-in model with the same shapes:
 *   - basis per element in the vDZP shell layout (H: s,s,p; C/O: s,s,p,p,d
 *     with spherical d), one normalised primitive Gaussian per shell;
 *   - S = analytic overlap of those functions (Obara-Saika recursion), so S is
 *     a Gram matrix (symmetric positive definite);
 *   - H = extended-Hueckel: H_ii = eps(shell), H_ij = K/2 (eps_i + eps_j) S_ij
 *     with the Wolfsberg-Helmholz constant K = 1.75.
 * Output is packed upper storage, element (i,j), i<=j (0-based) at
 * i + j(j+1)/2, the layout of the reference's Hvec/Svec (source/solve.f:149).
 *
 * Plain C, deterministic (no -ffast-math), OpenMP over atom pairs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int l; double alpha; double eps; } shell_t;

/* exponents (bohr^-2) and on-site energies (Hartree) */
static const shell_t SH_H[] = {{0, 0.90, -0.500}, {0, 0.068, -0.050}, {1, 0.21, 0.020}};
static const shell_t SH_C[] = {{0, 1.60, -0.786}, {0, 0.105, -0.060}, {1, 0.90, -0.419},
                               {1, 0.068, -0.020}, {2, 0.34, 0.050}};
static const shell_t SH_O[] = {{0, 2.40, -1.190}, {0, 0.13, -0.080}, {1, 1.20, -0.544},
                               {1, 0.088, -0.030}, {2, 0.38, 0.050}};

static const double WH_K = 1.75;

static int shells_of(int64_t z, const shell_t **sh) {
    switch (z) {
        case 1: *sh = SH_H; return 3;
        case 6: *sh = SH_C; return 5;
        case 8: *sh = SH_O; return 5;
        default: *sh = NULL; return 0;
    }
}

static int nfunc(int l) { return l == 0 ? 1 : (l == 1 ? 3 : 5); }

/* cartesian exponent triples: s; p x,y,z; d xx,yy,zz,xy,xz,yz */
static const int CART[10][3] = {{0,0,0}, {1,0,0},{0,1,0},{0,0,1},
                                {2,0,0},{0,2,0},{0,0,2},{1,1,0},{1,0,1},{0,1,1}};

/* 1D Obara-Saika overlap table s[i][j], i,j <= 2 (without the common prefactor) */
static void os1d(double pa, double pb, double inv2p, double s[3][3]) {
    s[0][0] = 1.0;
    s[1][0] = pa;
    s[0][1] = pb;
    s[1][1] = pb * s[1][0] + inv2p * s[0][0];
    s[2][0] = pa * s[1][0] + inv2p * s[0][0];
    s[0][2] = pb * s[0][1] + inv2p * s[0][0];
    s[2][1] = pb * s[2][0] + inv2p * 2.0 * s[1][0];
    s[1][2] = pb * s[1][1] + inv2p * (s[0][1] + s[1][0]);
    s[2][2] = pb * s[2][1] + inv2p * (2.0 * s[1][1] + 2.0 * s[2][0]);
}

/* cartesian -> spherical rows for d: xy, yz, 2zz-xx-yy, xz, xx-yy */
static void cart_to_sph(int l, const double *cart, double *out) {
    if (l < 2) {
        for (int i = 0; i < nfunc(l); i++) out[i] = cart[i];
        return;
    }
    /* cart order: xx yy zz xy xz yz */
    out[0] = cart[3];
    out[1] = cart[5];
    out[2] = 2.0 * cart[2] - cart[0] - cart[1];
    out[3] = cart[4];
    out[4] = cart[0] - cart[1];
}

/* raw overlap block between shell a at A and shell b at B: out[na][nb] (spherical) */
static void shell_overlap(const shell_t *a, const double *A, const shell_t *b, const double *B,
                          double out[5][5]) {
    double p = a->alpha + b->alpha, mu = a->alpha * b->alpha / p;
    double P[3], ab2 = 0.0;
    for (int k = 0; k < 3; k++) {
        P[k] = (a->alpha * A[k] + b->alpha * B[k]) / p;
        ab2 += (A[k] - B[k]) * (A[k] - B[k]);
    }
    double pref = pow(M_PI / p, 1.5) * exp(-mu * ab2);
    double sx[3][3][3];
    for (int k = 0; k < 3; k++) os1d(P[k] - A[k], P[k] - B[k], 0.5 / p, sx[k]);
    int ca0 = a->l == 0 ? 0 : (a->l == 1 ? 1 : 4), nca = a->l == 0 ? 1 : (a->l == 1 ? 3 : 6);
    int cb0 = b->l == 0 ? 0 : (b->l == 1 ? 1 : 4), ncb = b->l == 0 ? 1 : (b->l == 1 ? 3 : 6);
    double cart[6][6];
    for (int i = 0; i < nca; i++)
        for (int j = 0; j < ncb; j++) {
            const int *ei = CART[ca0 + i], *ej = CART[cb0 + j];
            cart[i][j] = pref * sx[0][ei[0]][ej[0]] * sx[1][ei[1]][ej[1]] * sx[2][ei[2]][ej[2]];
        }
    /* transform rows then columns */
    double half[5][6];
    for (int j = 0; j < ncb; j++) {
        double col[6], tmp[5];
        for (int i = 0; i < nca; i++) col[i] = cart[i][j];
        cart_to_sph(a->l, col, tmp);
        for (int i = 0; i < nfunc(a->l); i++) half[i][j] = tmp[i];
    }
    for (int i = 0; i < nfunc(a->l); i++) {
        double tmp[5];
        cart_to_sph(b->l, half[i], tmp);
        for (int j = 0; j < nfunc(b->l); j++) out[i][j] = tmp[j];
    }
}

static double self_norm(const shell_t *s, int comp) {
    double o[3] = {0, 0, 0}, blk[5][5];
    shell_overlap(s, o, s, o, blk);
    return 1.0 / sqrt(blk[comp][comp]);
}

/* per-element table of function norms: [shell][component] */
static void norm_table(int64_t z, double tab[5][5]) {
    const shell_t *sh;
    int ns = shells_of(z, &sh);
    for (int s = 0; s < ns; s++)
        for (int c = 0; c < nfunc(sh[s].l); c++) tab[s][c] = self_norm(&sh[s], c);
}

int64_t ptbsurr_ndim(int64_t n, const int64_t *znum) {
    int64_t nd = 0;
    for (int64_t i = 0; i < n; i++) {
        const shell_t *sh;
        int ns = shells_of(znum[i], &sh);
        if (!ns) return -1;
        for (int s = 0; s < ns; s++) nd += nfunc(sh[s].l);
    }
    return nd;
}

/* Fill packed-U H and S (length ndim(ndim+1)/2, zeroed here).  xyz is (n,3)
 * row-major in bohr.  Atom pairs farther than rcut contribute nothing (all
 * such overlaps are below 1e-12 for the exponents above).  Returns ndim. */
int64_t ptbsurr_build(int64_t n, const int64_t *znum, const double *xyz, double rcut,
                      double *Hp, double *Sp) {
    int64_t nd = ptbsurr_ndim(n, znum);
    if (nd < 0) return -1;
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
    off[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        const shell_t *sh;
        int ns = shells_of(znum[i], &sh), c = 0;
        for (int s = 0; s < ns; s++) c += nfunc(sh[s].l);
        off[i + 1] = off[i] + c;
    }
    int64_t npk = nd * (nd + 1) / 2;
    memset(Hp, 0, sizeof(double) * npk);
    memset(Sp, 0, sizeof(double) * npk);
    double rc2 = rcut * rcut;
    double ntab[9][5][5];
    norm_table(1, ntab[1]);
    norm_table(6, ntab[6]);
    norm_table(8, ntab[8]);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t ib = 0; ib < n; ib++) {
        const shell_t *shb;
        int nsb = shells_of(znum[ib], &shb);
        for (int64_t ia = 0; ia <= ib; ia++) {
            double d2 = 0.0;
            for (int k = 0; k < 3; k++) {
                double d = xyz[3 * ia + k] - xyz[3 * ib + k];
                d2 += d * d;
            }
            if (d2 > rc2) continue;
            const shell_t *sha;
            int nsa = shells_of(znum[ia], &sha);
            int64_t mu0 = off[ia];
            for (int sa = 0; sa < nsa; sa++) {
                int64_t nu0 = off[ib];
                for (int sb = 0; sb < nsb; sb++) {
                    double blk[5][5];
                    shell_overlap(&sha[sa], xyz + 3 * ia, &shb[sb], xyz + 3 * ib, blk);
                    for (int i = 0; i < nfunc(sha[sa].l); i++) {
                        double ni = ntab[znum[ia]][sa][i];
                        for (int j = 0; j < nfunc(shb[sb].l); j++) {
                            int64_t mu = mu0 + i, nu = nu0 + j;
                            if (mu > nu) continue; /* same-atom lower part */
                            double s = blk[i][j] * ni * ntab[znum[ib]][sb][j];
                            int64_t idx = mu + nu * (nu + 1) / 2;
                            if (mu == nu) {
                                Sp[idx] = 1.0;
                                Hp[idx] = sha[sa].eps;
                            } else {
                                Sp[idx] = s;
                                Hp[idx] = 0.5 * WH_K * (sha[sa].eps + shb[sb].eps) * s;
                            }
                        }
                    }
                    nu0 += nfunc(shb[sb].l);
                }
                mu0 += nfunc(sha[sa].l);
            }
        }
    }
    free(off);
    return nd;
}
