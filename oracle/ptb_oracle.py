"""CPU restatement of the reference's density-build hot path (numpy/scipy).

TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.

Every function cites the reference lines it restates (paths under the
reference checkout).  Dense LAPACK calls go to scipy (OpenBLAS ILP64 in this
image): dsygvd via ``eigh(..., driver="gvd")``, dsyevd via ``driver="evd"``.

Parity pins (what checks each function; tests/test_oracle.py,
tests/test_oracle_configs.py and tests/test_thresholds.py):
  filter_matrix_sym, analyse_matrix_sym   elementwise definition + compiled
                                          reference (oracle/_ref) on fixtures
  lin / blowsym / packsym                 closed-form index map, round trip
  column_groups modes 1-3                 closed forms + compiled reference
  column_groups mode 4 (k-means)          compiled reference: fed the compiler's
                                          RANDOM_NUMBER stream (extracted through
                                          the reference's own initseed), KMPP
                                          reproduces its groups and densities
  index sets / gather / sm_stupid_simple  compiled reference (oracle/_ref) on
                                          fixtures; brute force on tiny inputs;
                                          full_matrix mode == dense projector
  matrix_sign / matrix_root               textbook closed forms (V f(L) V^T)
  metrics (get_nel, band_structure,
    idempotent)                           compiled reference + closed forms
  adjust_thresholds_dp                    compiled reference on both sides of
                                          every tier boundary (ndim 299/300,
                                          1999/2000, 14999/15000)
  purification (metric, guess, McWeeny,
    sign Pade/Newton, sign-diag, search)  compiled reference on the 2-water
                                          pencil and, with the search, on the
                                          8- and 18-water pencils (ndim 184 and
                                          414, thresholds tier 2); converged
                                          P == dense projector
  fermismear / occu / solve2              compiled reference FERMISMEAR/occu;
                                          zero-temperature limit
  mlpop2                                  compiled reference mlpop14 + mlpop2 on
                                          two pencils (float rounding bound);
                                          sum q = tr(P S) for S^1/4 S^3/4 = S
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

# default-REAL(4) literals of the reference, widened to double as Fortran does
F32_1EM3 = float(np.float32(1e-3))   # source/sm.f90:370 abs(mua-mub).lt.1e-3
F32_1EM5 = float(np.float32(1e-5))   # source/sm.f90:370 abs(nel-nelref).lt.1e-5
BOLTZ = float(np.float32(3.166808578545117e-06))  # source/fermi.f:17 REAL(4) literal
KMPP_BIG = float(np.float32(1e33))                 # source/kmpp.f:34


# --------------------------------------------------------------------------
# packed storage (source/lin.f:5-11, source/solve.f:149-200)
# --------------------------------------------------------------------------
def lin(i: int, j: int) -> int:
    """1-based packed index of (i,j): min + max(max-1)/2 (source/lin.f:5-11)."""
    a, b = max(i, j), min(i, j)
    return b + a * (a - 1) // 2


def packed_len(n: int) -> int:
    return n * (n + 1) // 2


def _upper_index(n: int):
    """(rows, cols) of the packed-upper order: column by column, rows 0..j."""
    cols = np.repeat(np.arange(n), np.arange(1, n + 1))
    starts = np.repeat(np.cumsum(np.r_[0, np.arange(1, n)]), np.arange(1, n + 1))
    rows = np.arange(packed_len(n)) - starts
    return rows, cols


def blowsym(v: np.ndarray, n: int) -> np.ndarray:
    """Packed upper -> dense symmetric (source/solve.f:149-165)."""
    r, c = _upper_index(n)
    A = np.zeros((n, n))
    A[r, c] = v
    A[c, r] = v
    return A


def packsym(A: np.ndarray) -> np.ndarray:
    """Dense -> packed upper, element (j,i), j<=i, taken from column i
    (source/solve.f:185-200)."""
    n = A.shape[0]
    r, c = _upper_index(n)
    return A[r, c].copy()


# --------------------------------------------------------------------------
# filter / analyse (source/pgtb.f90:1532-1579)
# --------------------------------------------------------------------------
def filter_matrix_sym(A: np.ndarray, n: int, filter_: float) -> None:
    """In place: A(i)=0 where abs(A(i)) < filter (source/pgtb.f90:1543-1547)."""
    m = np.abs(A[: packed_len(n)]) < filter_
    A[: packed_len(n)][m] = 0.0


def analyse_matrix_sym(A: np.ndarray, n: int) -> dict:
    """Positive-entry statistics (source/pgtb.f90:1549-1579): colcnt counts
    A(ij)>0, once on the diagonal and twice off it; printed figures are
    sum(colcnt), sum/n**2 and the integer sum/n."""
    r, c = _upper_index(n)
    pos = A[: packed_len(n)] > 0
    colcnt = np.bincount(c[pos], minlength=n) + np.bincount(r[pos & (r != c)], minlength=n)
    tot = int(colcnt.sum())
    return {"dimension": n, "nonzero": tot, "sparsity": tot / float(n) ** 2,
            "average": tot // n if n else 0, "colcnt": colcnt}


# --------------------------------------------------------------------------
# column grouping (source/sm.f90:91-162, :459-514; source/kmpp.f:24-179)
# --------------------------------------------------------------------------
def kmpp(x: np.ndarray, k: int, uniform) -> tuple:
    """K-means++ (source/kmpp.f:24-179).  x is (P, N) like the Fortran X(P,N);
    ``uniform()`` yields the RANDOM_NUMBER stream.  Returns (C, Z 1-based,
    ifault)."""
    p, n = x.shape
    z = np.zeros(n, dtype=np.int64)
    c = np.zeros((p, k))
    if k < 1 or k > n:
        return c, z, 3
    work = np.full(n, KMPP_BIG)
    w = uniform()
    i1 = min(int(w * float(np.float32(n))) + 1, n)
    c[:, 0] = x[:, i1 - 1]
    for l in range(2, k + 1):
        tot = 0.0
        for i in range(n):
            best = work[i]
            d2 = 0.0
            skip = False
            for j in range(p):
                d2 += (x[j, i] - c[j, l - 2]) ** 2
                if d2 >= best:
                    skip = True
                    break
            if not skip:
                if d2 < best:
                    best = d2
                work[i] = best
            tot += best
        w = uniform() * tot
        tot = 0.0
        i1 = n
        for i in range(n):
            i1 = i + 1
            tot += work[i]
            if tot > w:
                break
        c[:, l - 1] = x[:, i1 - 1]
    for _ in range(1000):
        change = False
        for i in range(n):
            l0, l1, best = z[i], 0, KMPP_BIG
            for l in range(1, k + 1):
                d2 = 0.0
                for j in range(p):
                    d2 += (x[j, i] - c[j, l - 1]) ** 2
                    if d2 >= best:
                        break
                if d2 < best:
                    best, l1 = d2, l
            if l0 != l1:
                z[i] = l1
                change = True
        if not change:
            return c, z, 0
        cnt = np.zeros(k)
        c[:] = 0.0
        for i in range(n):
            cnt[z[i] - 1] += 1.0
            c[:, z[i] - 1] += x[:, i]
        if np.any(cnt < 0.5):
            return c, z, 1
        c *= 1.0 / cnt
    return c, z, 2


def pattern_dense(Hd: np.ndarray, Sd: np.ndarray) -> np.ndarray:
    """Coupling test H(i,j)>0 .or. S(i,j)>0 .or. i==j (source/sm.f90:173,214,504)."""
    pat = (Hd > 0) | (Sd > 0)
    np.fill_diagonal(pat, True)
    return pat


def smdim_of(pat: np.ndarray, cols: np.ndarray) -> int:
    """Rows coupled to any owned column (source/sm.f90:169-179)."""
    return int(pat[:, cols].any(axis=1).sum())


def column_groups(mode: int, ndim: int, aoat: np.ndarray, n: int = 0, xyz=None,
                  pat=None, uniform=None) -> list:
    """comb/ncomb as a list of 0-based column arrays (source/sm.f90:91-162).

    mode 1 single_columns, 2 atoms_columns (contiguous runs of equal aoat),
    3 full_matrix, 4 kmeans_r_heuristic (min sum m**2.5 over nsm=1..min(64,n),
    k-means++ on the atom positions with the stream ``uniform``)."""
    aoat = np.asarray(aoat)
    if mode == 1:
        return [np.array([i]) for i in range(ndim)]
    if mode == 2:
        starts = np.r_[0, np.nonzero(aoat[1:] != aoat[:-1])[0] + 1, ndim]
        return [np.arange(starts[k], starts[k + 1]) for k in range(len(starts) - 1)]
    if mode == 3:
        return [np.arange(ndim)]
    if mode == 4:
        if uniform is None:
            raise ValueError("mode 4 needs a RANDOM_NUMBER stream")
        best, best_groups = math.inf, None
        for nsm in range(1, min(64, n) + 1):
            groups, effort = clustering_kmeans(n, ndim, xyz, aoat, nsm, pat, uniform(nsm))
            if effort < best:
                best, best_groups = effort, groups
        return best_groups
    raise ValueError(f"submatrix_columns {mode} not implemented")


def clustering_kmeans(n, ndim, xyz, aoat, nsm, pat, uniform):
    """One k-means candidate and its effort sum(m**2.5) (source/sm.f90:459-514)."""
    _, z, ifault = kmpp(np.asarray(xyz).reshape(3, n) if np.shape(xyz) == (3, n)
                        else np.asarray(xyz).T, nsm, uniform)
    if ifault != 0:
        raise RuntimeError(f"kmeans failed {ifault}")
    groups = []
    for ism in range(1, nsm + 1):
        atoms = np.nonzero(z == ism)[0] + 1
        cols = np.nonzero(np.isin(aoat, atoms))[0]
        groups.append(cols)
    effort = 0.0
    for cols in groups:
        effort += float(smdim_of(pat, cols)) ** 2.5
    return groups, effort


# --------------------------------------------------------------------------
# submatrix method (source/sm.f90:34-409)
# --------------------------------------------------------------------------
@dataclass
class SubmatrixResult:
    P: np.ndarray                      # packed output
    mu: float
    nel: float
    muiter: int                        # passes printed as "muiter k mu ..."
    trace: list                        # [(muiter, mu, nel)]
    maps: list                         # per submatrix sorted 0-based row map
    groups: list                       # per submatrix owned columns (0-based)
    eigvals: list = field(default_factory=list)
    fe: list = field(default_factory=list)
    info: list = field(default_factory=list)  # eigh failures (reference ignores INFO)
    stats: dict = field(default_factory=dict)


def sm_eigendecomp(Hsm: np.ndarray, Ssm: np.ndarray):
    """la_sygvd -> dsygvd(1,'V','U') (source/sm.f90:290-302,
    source/lapack_eig.f90:318-386).  Returns (e, U, info)."""
    try:
        e, U = sla.eigh(Hsm, Ssm, lower=False, driver="gvd", check_finite=False)
        return e, U, 0
    except np.linalg.LinAlgError:
        m = Hsm.shape[0]
        return np.zeros(m), np.zeros((m, m)), 1


def sm_fe(Ssm: np.ndarray, U: np.ndarray, owned_local: np.ndarray) -> np.ndarray:
    """fe(k) = sum_{owned rows i} (Ssm U)(i,k) U(i,k) (source/sm.f90:304-310)."""
    Psm = Ssm @ U
    fe = np.zeros(U.shape[1])
    for i in owned_local:
        fe = fe + Psm[i, :] * U[i, :]
    return fe


def occupations(e: np.ndarray, mu: float) -> np.ndarray:
    """2 if e<mu, 1 if e==mu, else 0 (source/sm.f90:322-330)."""
    occ = np.zeros_like(e)
    occ[e < mu] = 2.0
    occ[e == mu] = 1.0
    return occ


def dmat(occ: np.ndarray, C: np.ndarray) -> np.ndarray:
    """P = C diag(occ) C^T (source/solve.f:263-283)."""
    return C @ (C * occ).T


def sm_stupid_simple(ndim: int, nelref: int, Hvec: np.ndarray, Svec: np.ndarray, n: int,
                     xyz, aoat, submatrix_columns: int, submatrix_mode: int,
                     uniform=None, groups=None, verbose=False) -> SubmatrixResult:
    """The submatrix density build (source/sm.f90:34-409), returning packed P.

    ``groups`` overrides the column grouping (used to pin mode 4 from the
    compiled reference's own k-means groups)."""
    if submatrix_mode != 2:
        raise NotImplementedError("submatrix_mode 1 is not implemented in the reference "
                                  "(source/sm.f90:249-251)")
    H = blowsym(Hvec, ndim)
    S = blowsym(Svec, ndim)
    pat = pattern_dense(H, S)
    if groups is None:
        groups = column_groups(submatrix_columns, ndim, aoat, n, xyz, pat, uniform)
    nsm = len(groups)
    # submatrix stats (source/sm.f90:165-187)
    dims = [smdim_of(pat, g) for g in groups]
    stats = {"nsm": nsm, "average": sum(dims) // nsm if nsm else 0,
             "effort_full": float(ndim) ** 2.5,
             "effort_sms": float(sum(float(d) ** 2.5 for d in dims))}
    stats["ratio"] = stats["effort_sms"] / stats["effort_full"]

    maps, eig, fes, Us, infos = [], [], [], [], []
    # small dense problems run faster on one BLAS thread than on many (the
    # arithmetic is the same LAPACK routine either way)
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1 if max(dims, default=0) < 400 else None):
        for g in groups:
            rows = np.nonzero(pat[:, g].any(axis=1))[0]          # :210-228
            imap = np.full(ndim, -1, dtype=np.int64)
            imap[rows] = np.arange(len(rows))
            Hsm = H[np.ix_(rows, rows)]                           # :240-246
            Ssm = S[np.ix_(rows, rows)]
            e, U, info = sm_eigendecomp(Hsm, Ssm)
            maps.append(rows)
            eig.append(e)
            Us.append(U)
            infos.append(info)
            fes.append(sm_fe(Ssm, U, imap[g]))

    mua, mub = -10.0, 10.0                                    # :191-192
    finaliter = False
    trace = []
    P = None
    mu = nel = 0.0
    muiter = 0
    for muiter in range(1, 101):                               # :197
        mu = (mua + mub) / 2.0
        nel = 0.0
        if finaliter:
            P = np.zeros((ndim, ndim))
        for ism in range(nsm):
            occ = occupations(eig[ism], mu)
            nel = nel + float(np.sum(fes[ism] * occ))          # :340
            if finaliter:                                      # :343-361
                Psm = dmat(occ, Us[ism])
                rows = maps[ism]
                imap = np.full(ndim, -1, dtype=np.int64)
                imap[rows] = np.arange(len(rows))
                for j in groups[ism]:
                    P[rows, j] = Psm[:, imap[j]]
        if finaliter:
            break
        trace.append((muiter, mu, nel))
        if verbose:
            print("muiter", muiter, "mu", mu, nel, "of", nelref)
        if abs(mua - mub) < F32_1EM3 or abs(nel - nelref) < F32_1EM5:   # :370
            finaliter = True
        if not finaliter:
            if nel > nelref:
                mub = mu
            else:
                mua = mu
    return SubmatrixResult(P=packsym(P), mu=mu, nel=nel, muiter=len(trace), trace=trace,
                           maps=maps, groups=list(groups), eigvals=eig, fe=fes,
                           info=infos, stats=stats)


def matrix_sign(A: np.ndarray) -> np.ndarray:
    """sign(A) = V sgn(L) V^T via dsyevd (source/sm.f90:410-436)."""
    w, V = sla.eigh(A, lower=False, driver="evd")
    return V @ np.diag(np.sign(w)) @ V.T


def matrix_root(A: np.ndarray) -> np.ndarray:
    """A^-1/2 = V L^-1/2 V^T via dsyevd (source/sm.f90:437-457)."""
    w, V = sla.eigh(A, lower=False, driver="evd")
    return V @ np.diag(1.0 / np.sqrt(w)) @ V.T


# --------------------------------------------------------------------------
# norms (source/norms.f90) and metrics (source/metrics.f90)
# --------------------------------------------------------------------------
def gershgorin(A):   # source/norms.f90:5-18, max column abs-sum
    return float(np.max(np.sum(np.abs(A), axis=0)))


def frobenius(A):    # source/norms.f90:21-28
    return float(np.sqrt(np.sum(A * A)))


def infinity(A):     # source/norms.f90:32-49, max row abs-sum
    return float(np.max(np.sum(np.abs(A), axis=1)))


@dataclass
class Thresholds:
    high: float
    normal: float
    low: float


def adjust_thresholds_dp(ndim: int) -> Thresholds:
    """Size-dependent thresholds (source/metrics.f90:82-113); the literals are
    default REAL(4)."""
    f = lambda x: float(np.float32(x))
    if ndim < 300:
        return Thresholds(f(1e-10), f(1e-9), f(1e-8))
    if ndim < 2000:
        return Thresholds(f(1e-7), f(1e-6), f(1e-5))
    if ndim < 15000:
        return Thresholds(f(1e-5), f(1e-4), f(1e-3))
    return Thresholds(f(1e-3), f(1e-2), f(1e-1))


def get_nel(ndim, Pvec, Svec) -> float:
    """tr(P S) (source/metrics.f90:202-230)."""
    return float(np.trace(blowsym(Pvec, ndim) @ blowsym(Svec, ndim)))


def band_structure(ndim, Pvec, Hvec) -> float:
    """tr(P H) (source/metrics.f90:253-277)."""
    return float(np.trace(blowsym(Pvec, ndim) @ blowsym(Hvec, ndim)))


def idempotency_deviation(ndim, Pvec, Svec) -> float:
    """sqrt(sum(mm2 - P)**2) = |sum(0.5 P S P - P)| (source/metrics.f90:346-358)."""
    P = blowsym(Pvec, ndim)
    S = blowsym(Svec, ndim)
    mm2 = 0.5 * ((P @ S) @ P)
    return float(math.sqrt(np.sum(mm2 - P) ** 2))


def check_density(ndim, Pvec, Svec, nel, thrs: Thresholds) -> dict:
    """check_density_packed (source/metrics.f90:144-179): N_el against
    thrs%normal, idempotency deviation against thrs%low."""
    nelp = get_nel(ndim, Pvec, Svec)
    dev = idempotency_deviation(ndim, Pvec, Svec)
    return {"nel": nelp, "nel_ok": not (abs(nelp - nel) > thrs.normal),
            "idempotency": dev, "idempotent": dev < thrs.low}


# --------------------------------------------------------------------------
# .PUR settings (source/purification_settings.f90, source/readline.f90,
# source/readl2.f90:52-92)
# --------------------------------------------------------------------------
FULL, SUBMATRIX = 1, 2
MCWEENY, SIGN_PADE, SIGN_NEWTON, SIGN_DIAG = 1, 2, 3, 4
INV, INV_SQRT, SQRT_ = 1, 2, 3


@dataclass
class PurSettings:
    """tPurificationSet with its defaults (source/purification_settings.f90:29-86)."""
    mode: int = FULL
    type: int = SIGN_DIAG
    cycles: int = 40
    prlvl: int = 1
    dev: bool = True
    check: bool = False
    nel: int = 0
    submatrix_mode: int = 2
    submatrix_columns: int = 2
    metric_type: int = INV_SQRT
    metric_cycles: int = 50
    metric_iterative: bool = False
    chempot_cycles: int = 60
    chempot_guess: float = -1.5
    chempot_increment: float = 0.1


class PurError(ValueError):
    pass


def _is_number(tok: str) -> bool:
    """checktype's F10.5 read succeeds (source/readl2.f90:62-64)."""
    try:
        float(tok[:10].replace("d", "e").replace("D", "e"))
        return True
    except ValueError:
        return False


def parse_pur(text: str) -> PurSettings:
    """initialize_purification (source/purification_settings.f90:91-218).

    Tokens are split on blanks/tabs; numeric tokens go to floats, others to
    strings (source/readline.f90).  ``key = value`` needs blanks around '=';
    int() truncates numeric values (so chempot_increment 0.1 -> 0)."""
    s = PurSettings()
    for line in text.splitlines():
        line = line[:80]
        toks = line.split()
        strs = [t for t in toks if not _is_number(t)]
        nums = [float(t.replace("d", "e").replace("D", "e")) for t in toks if _is_number(t)]
        arg1 = strs[0] if strs else ""
        if arg1[:1] == "!":
            continue
        arg2 = strs[1] if len(strs) > 1 else ""
        if arg2 == "=":
            arg3 = strs[2] if len(strs) > 2 else ""
            if arg1 == "mode":
                table = {"full": FULL, "submatrix": SUBMATRIX}
                if arg3 not in table:
                    raise PurError("Error: .PUR contains invalid type definition")
                s.mode = table[arg3]
            elif arg1 == "submatrix_columns":
                table = {"single_columns": 1, "atoms_columns": 2, "full_matrix": 3,
                         "kmeans_r_heuristic": 4}
                if arg3 not in table:
                    raise PurError("Error: .PUR contains invalid type definition")
                s.submatrix_columns = table[arg3]
            elif arg1 == "submatrix_mode":
                table = {"submatrix_sign": 1, "submatrix_sygv": 2}
                if arg3 not in table:
                    raise PurError("Error: .PUR contains invalid type definition")
                s.submatrix_mode = table[arg3]
            elif arg1 == "type":
                table = {"mcweeny": MCWEENY, "pade": SIGN_PADE, "newton": SIGN_NEWTON,
                         "diag": SIGN_DIAG}
                if arg3 not in table:
                    raise PurError("Error: .PUR contains invalid type definition")
                s.type = table[arg3]
            elif arg1 == "metric":
                table = {"inverse": INV, "inverse_sqrt": INV_SQRT, "sqrt": SQRT_}
                if arg3 not in table:
                    raise PurError("Error: .PUR contains invalid metric definition.")
                s.metric_type = table[arg3]
            elif arg1 == "cycles" and nums:
                s.cycles = int(nums[0])
            elif arg1 == "chempot_cycles" and nums:
                s.chempot_cycles = int(nums[0])
            elif arg1 == "chempot_increment" and nums:
                s.chempot_increment = float(int(nums[0]))
            elif arg1 == "pr" and nums:
                s.prlvl = int(nums[0])
        else:
            if arg1 == "dev":
                s.dev = True
            elif arg1 == "check":
                s.check = True
            elif arg1 == "iterative_inversion":
                s.metric_iterative = True
    return s


# --------------------------------------------------------------------------
# populations (source/mlpop.f)
# --------------------------------------------------------------------------
def mlpop2(n: int, nao: int, P: np.ndarray, s1: np.ndarray, s2: np.ndarray, aoat, shell2ao):
    """mlpop2 (source/mlpop.f:198-238), in REAL(4) like the reference:
    x = blowsym84(P) (P rounded to float), cc = x s1 (s1 = S^3/4), x = s2 cc
    (s2 = S^1/4), then q(aoat(i)) += x(i,i) and qsh(shell2ao(i), aoat(i)) +=
    x(i,i) in AO order, in double."""
    x = blowsym(np.asarray(P, dtype=np.float64), nao).astype(np.float32)
    cc = x @ np.asarray(s1, dtype=np.float32)
    d = np.einsum("ik,ki->i", np.asarray(s2, dtype=np.float32), cc).astype(np.float64)
    q = np.zeros(n)
    qsh = np.zeros((10, n))
    for i in range(nao):
        q[aoat[i] - 1] += d[i]
        qsh[shell2ao[i] - 1, aoat[i] - 1] += d[i]
    return q, qsh


# --------------------------------------------------------------------------
# purification (source/purification.f90)
# --------------------------------------------------------------------------
class NaNError(RuntimeError):
    pass


def get_transformation_matrix(pur: PurSettings, S: np.ndarray, thrs: Thresholds):
    """X = S^-1 or S^-1/2 (source/purification.f90:83-229).  Returns (X, cycles)."""
    n = S.shape[0]
    I = np.eye(n)
    if pur.metric_iterative:
        if pur.metric_type == INV:
            X = S / infinity(S) ** 2                          # :137
            for i in range(1, pur.metric_cycles + 1):
                atmp = S @ X
                XsX = X @ atmp
                X = 2.0 * X - XsX
                check = X @ S
                if np.sqrt(np.sum((I - check) ** 2)) < thrs.low:
                    return X, i
            return X, pur.metric_cycles
        if pur.metric_type == INV_SQRT:
            X = S / frobenius(S) ** 2                         # :158
            for i in range(1, pur.metric_cycles + 1):
                atmp = X @ X
                XsX = atmp @ S
                atmp2 = X @ XsX
                X = 0.5 * (3 * X - atmp2)
                if frobenius(I - XsX) < thrs.low:
                    return X, i
                if np.any(np.isnan(XsX)):
                    raise NaNError("Error, NaN encountered")
            return X, pur.metric_cycles
        raise ValueError("iterative metric supports inverse and inverse_sqrt only")
    w, V = sla.eigh(S, lower=False, driver="evd")             # :191
    if pur.metric_type == INV:
        D = np.diag(1.0 / w)
    elif pur.metric_type == INV_SQRT:
        D = np.diag(1.0 / np.sqrt(w))
    else:
        D = np.zeros((n, n))   # 'sqrt' selects no branch in the reference: D stays 0
    return (V @ D) @ V.T, 0


def build_guess(pur: PurSettings, H: np.ndarray, X: np.ndarray, chempot: float):
    """Initial guess (source/purification.f90:355-436)."""
    n = H.shape[0]
    I = np.eye(n)
    if pur.type == MCWEENY:
        term1 = chempot * X
        term3 = X @ (H @ X)
        term4 = term1 - term3
        term5 = (1.0 / min(frobenius(term4), gershgorin(term4))) * term4
        return 0.5 * (term5 + X)
    if pur.metric_type == INV_SQRT:
        term4 = X @ (H @ X) - chempot * I
        return (1.0 / min(frobenius(term4), gershgorin(term4))) * term4
    if pur.metric_type == INV:
        term3 = X @ H - chempot * I
        return (1.0 / min(frobenius(term3), gershgorin(term3))) * term3
    raise ValueError("sign guess needs metric inverse or inverse_sqrt")


def mcweeny(pur: PurSettings, guess: np.ndarray, S: np.ndarray, thrs: Thresholds):
    """P <- 3PSP - 2PSPSP, stop on |norm2(P)-norm| < thrs%normal, then 2P
    (source/purification.f90:438-510).  Returns (P, cycles)."""
    g = guess
    norm = frobenius(guess)
    P = g
    it = 0
    for it in range(1, pur.cycles + 1):
        t1 = S @ g
        t2 = g @ t1
        t3 = t2 @ t1
        P = 3 * t2 - 2 * t3
        if np.any(np.isnan(P)):
            raise NaNError("Error: NaN is encountered during purification")
        if abs(frobenius(P) - norm) < thrs.normal:
            break
        g = P
        norm = frobenius(P)
    return P * 2, it


def sign_to_density_matrix(pur: PurSettings, sign_: np.ndarray, X: np.ndarray):
    """P = 0.5 X (I - sign) X (inv_sqrt) or 0.5 (I - sign) X (inv)
    (source/purification.f90:603-647)."""
    I = np.eye(sign_.shape[0])
    if pur.metric_type == INV_SQRT:
        P = 0.5 * (X @ ((I - sign_) @ X))
    elif pur.metric_type == INV:
        P = 0.5 * ((I - sign_) @ X)
    else:
        raise ValueError("sign_to_density_matrix needs inverse or inverse_sqrt")
    if np.any(np.isnan(P)):
        raise NaNError("Error: NaN is encountered during purification")
    return P


def sign_iterative(pur: PurSettings, guess: np.ndarray, X: np.ndarray, thrs: Thresholds):
    """Pade X(15I-10X^2+3X^4)/8 or Newton-Schulz (3X-X^3)/2 iterations, then
    2 * sign_to_density_matrix (source/purification.f90:512-601)."""
    n = guess.shape[0]
    I = np.eye(n)
    Xs = guess
    it = 0
    for it in range(1, pur.cycles + 1):
        norm = frobenius(Xs)
        if pur.type == SIGN_PADE:
            t1 = Xs @ Xs
            t2 = t1 @ t1
            t3 = 15.0 * I - 10.0 * t1 + 3.0 * t2
            Xs = 0.125 * (Xs @ t3)
        else:
            t1 = Xs @ Xs
            t2 = Xs @ t1
            Xs = 0.5 * (3 * Xs - t2)
        if np.any(np.isnan(Xs)):
            raise NaNError("Error: NaN is encountered during purification")
        if abs(frobenius(Xs) - norm) < thrs.normal:
            break
    return 2.0 * sign_to_density_matrix(pur, Xs, X), it


def sign_diagonalization(pur: PurSettings, chempot: float, guess: np.ndarray, X: np.ndarray,
                         thrs: Thresholds, state=None):
    """Internal chemical-potential bisection on the guess's eigenvalues, then
    sign = Q sgn Q^T (source/purification.f90:650-782).  The SAVEd flag
    first_time is initialised .true. and never cleared (:671,692), so the guess
    is diagonalised on every call.  chempot is intent(inout) (:659): the
    internally adjusted value goes back to the outer search.  Returns
    (P, chempot)."""
    n = guess.shape[0]
    lam, Q = sla.eigh(guess, lower=False, driver="evd")
    nel = float(np.float32(pur.nel))       # real(pur%nel), default REAL(4) (:683)
    incr = pur.chempot_increment
    lower = upper = 0.0
    is_lower = is_upper = False
    sgn = np.zeros(n)
    colnorm = np.sum(Q * Q, axis=0)
    for _k in range(1, pur.chempot_cycles + 1):
        eg = lam - chempot
        sgn = np.where(np.abs(eg) < thrs.normal, 0.0, np.where(eg > 0.0, 1.0, -1.0))
        nelcalc = float(np.sum(1.0 - colnorm * sgn))
        if abs(nelcalc - nel) < thrs.normal:
            break
        if nelcalc < nel:
            is_lower, lower = True, chempot
            chempot = chempot + incr
        else:
            is_upper, upper = True, chempot
            chempot = chempot - incr
        incr = incr * 2.0
        if is_lower and is_upper:
            chempot = (lower + upper) / 2.0
            if abs(upper - lower) < thrs.high:
                break
    sign_ = Q @ (np.diag(sgn) @ Q.T)
    return 2.0 * sign_to_density_matrix(pur, sign_, X), chempot


def purification_search(pur: PurSettings, H, S, X, thrs: Thresholds, state=None):
    """Outer chemical-potential search (source/purification.f90:232-353).
    Returns (P dense, chempot, nel_calc, cycles)."""
    chempot = pur.chempot_guess
    incr = pur.chempot_increment
    nel = pur.nel
    lower = upper = 0.0
    is_lower = is_upper = bisection = False
    P = None
    nel_calc = 0.0
    i = 0
    for i in range(1, pur.chempot_cycles + 1):
        if bisection:
            chempot = (lower + upper) / 2.0
            if abs(upper - lower) < thrs.high:
                break
        guess = build_guess(pur, H, X, chempot)
        if pur.type == MCWEENY:
            P, _ = mcweeny(pur, guess, S, thrs)
        elif pur.type in (SIGN_PADE, SIGN_NEWTON):
            P, _ = sign_iterative(pur, guess, X, thrs)
        else:
            P, chempot = sign_diagonalization(pur, chempot, guess, X, thrs, state)
        nel_calc = float(np.trace(P @ S))
        if abs(nel_calc - float(np.float32(nel))) < thrs.normal:   # real(nel) (:327)
            break
        if nel_calc < nel:
            is_lower, lower = True, chempot
            chempot = chempot + incr
        else:
            is_upper, upper = True, chempot
            chempot = chempot - incr
        bisection = is_lower and is_upper
        incr = incr * 2.0
    return P, chempot, nel_calc, i


def purification_wrapper(pur: PurSettings, ndim: int, Hvec, Svec, thrs: Thresholds = None,
                         state=None):
    """Packed H, S -> packed P (source/purification.f90:28-80)."""
    thrs = thrs or adjust_thresholds_dp(ndim)
    H = blowsym(Hvec, ndim)
    S = blowsym(Svec, ndim)
    X, _ = get_transformation_matrix(pur, S, thrs)
    P, chempot, nel_calc, cycles = purification_search(pur, H, S, X, thrs, state)
    return packsym(P), {"chempot": chempot, "nel": nel_calc, "cycles": cycles}


# --------------------------------------------------------------------------
# dense reference solve with Fermi smearing (source/pgtb.f90:1066-1170,
# source/fermi.f:5-124)
# --------------------------------------------------------------------------
def fermismear(norbs: int, nel: int, t: float, eig: np.ndarray):
    """FERMISMEAR (source/fermi.f:5-67); returns (occ, e_fermi, fod, s)."""
    bkt = BOLTZ * t
    e_fermi = 0.5 * (eig[nel - 1] + eig[nel])
    occt = float(nel)
    occ = np.zeros(norbs)
    ncycle = 0
    while True:
        total_number = 0.0
        total_dfermi = 0.0
        ncycle += 1
        for i in range(norbs):
            x = (eig[i] - e_fermi) / bkt
            if x < 50:
                ex = math.exp(x)
                f = 1.0 / (ex + 1.0)
                df = ex / (bkt * (ex + 1.0) ** 2)
            else:
                f = 0.0
                df = 0.0
            occ[i] = f
            total_number += f
            total_dfermi += df
        change = (occt - total_number) / (total_dfermi + 1e-14)
        e_fermi = e_fermi + change
        if not (abs(occt - total_number) > 1e-9 and ncycle < 200):
            break
    fod = 0.0
    s = 0.0
    for i in range(norbs):
        if occ[i] > 1e-9 and 1.0 - occ[i] > 1e-9:
            s += occ[i] * math.log(occ[i]) + (1.0 - occ[i]) * math.log(1.0 - occ[i])
        fod += (1.0 - occ[i]) if eig[i] < e_fermi else occ[i]
    s = s * BOLTZ * t
    return occ, e_fermi, fod, s


def occu(ndim: int, nel: int, nopen: int = 0):
    """Closed/open-shell alpha/beta occupations (source/fermi.f:69-124)."""
    focc = np.zeros(ndim, dtype=np.int64)
    if nel % 2 == 0:
        ihomo = nel // 2
        focc[:ihomo] = 2
        if nopen > 1:
            for i in range(1, nopen // 2 + 1):
                focc[ihomo - i] -= 1
                focc[ihomo + i - 1] += 1
    else:
        na = nel // 2 + (nopen - 1) // 2 + 1
        nb = nel // 2 - (nopen - 1) // 2
        focc[:na] += 1
        focc[:nb] += 1
    focca = np.where(focc >= 1, 1.0, 0.0)
    foccb = np.where(focc == 2, 1.0, 0.0)
    ihomoa = int(np.max(np.nonzero(focca > 0.99)[0]) + 1) if np.any(focca > 0.99) else 0
    ihomob = int(np.max(np.nonzero(foccb > 0.99)[0]) + 1) if np.any(foccb > 0.99) else 0
    return focca, foccb, ihomoa, ihomob


def solve2(ndim: int, nel: int, Hvec, Svec, et: float = 300.0, nopen: int = 0, mode: int = 1,
           shift=(0.0, 0.0)):
    """Dense reference P with Fermi smearing (source/pgtb.f90:1066-1170); for
    mode >= 2 the eigenvalues are shifted to e + shift[0] + shift[1] e
    (glob_par(15), glob_par(17), :1139-1143) before the smearing.  Returns
    (packed P, e, focc)."""
    H = blowsym(Hvec, ndim)
    S = blowsym(Svec, ndim)
    e, U = sla.eigh(H, S, lower=False, driver="gvd")
    if mode >= 2:
        e = e + shift[0] + shift[1] * e
    focca, foccb, ia, ib = occu(ndim, nel, nopen)
    if ia + 1 <= ndim:
        focca, _, _, _ = fermismear(ndim, ia, et, e)
    if ib + 1 <= ndim and nel > 1:
        foccb, _, _, _ = fermismear(ndim, ib, et, e)
    focc = focca + foccb
    return packsym(dmat(focc, U)), e, focc
