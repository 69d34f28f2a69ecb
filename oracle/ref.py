"""ctypes binding of oracle/_ref/libptbref.so (the compiled reference units).

TEST INFRASTRUCTURE ONLY.  The library exists only where the reference
sources were present when ``make -C oracle`` ran; callers must check
``available()`` and skip otherwise.  The reference prints to stdout;
``quiet()`` redirects file descriptor 1 around a call.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libptbref.so")

_lib = None


def available() -> bool:
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"{LIB} is not built (make -C oracle)")
        _lib = ctypes.CDLL(LIB, mode=ctypes.RTLD_GLOBAL)
        _lib.ref_get_nel.restype = ctypes.c_double
        _lib.ref_band_structure.restype = ctypes.c_double
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


I64 = ctypes.c_int64
F64 = ctypes.c_double


@contextlib.contextmanager
def quiet(capture: list | None = None):
    """Redirect fd 1 (Fortran's stdout) into a temp file; append text to capture."""
    import sys
    sys.stdout.flush()
    saved = os.dup(1)
    with tempfile.TemporaryFile(mode="w+b") as tmp:
        os.dup2(tmp.fileno(), 1)
        try:
            yield
        finally:
            try:
                libc = ctypes.CDLL(None)
                libc.fflush(None)
            except OSError:
                pass
            os.dup2(saved, 1)
            os.close(saved)
            if capture is not None:
                tmp.seek(0)
                capture.append(tmp.read().decode(errors="replace"))


def sm_stupid_simple(ndim, nelref, H, S, n, xyz, aoat, cols, mode, log=None):
    """Packed P from the reference's sm_stupid_simple (source/sm.f90:34-409)."""
    P = np.zeros(ndim * (ndim + 1) // 2)
    H = np.ascontiguousarray(H, dtype=np.float64)
    S = np.ascontiguousarray(S, dtype=np.float64)
    x = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(n, 3))  # == xyz(3,n)
    a = np.ascontiguousarray(aoat, dtype=np.int64)
    with quiet(log):
        lib().ref_sm_stupid_simple(I64(ndim), I64(nelref), _p(H), _p(S), _p(P), I64(n), _p(x),
                                   _p(a), I64(cols), I64(mode))
    return P


def random_stream(count):
    """The compiler's RANDOM_NUMBER values after initseed, in the order KMPP
    draws them."""
    out = np.zeros(count)
    lib().ref_random_stream(I64(count), _p(out))
    return out


def clustering_kmeans(n, ndim, xyz, H, S, aoat, nsm):
    """Column groups (0-based arrays) and effort of one k-means candidate."""
    x = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(n, 3))
    ncomb = np.zeros(n, dtype=np.int64)
    comb = np.zeros((ndim, n), dtype=np.int64)      # Fortran comb(n, ndim)
    eff = F64(0.0)
    with quiet():
        lib().ref_clustering_kmeans(I64(n), I64(ndim), _p(x), _p(np.ascontiguousarray(H)),
                                    _p(np.ascontiguousarray(S)),
                                    _p(np.ascontiguousarray(aoat, dtype=np.int64)), I64(nsm),
                                    _p(ncomb), _p(comb), ctypes.byref(eff))
    groups = [comb[: ncomb[k], k] - 1 for k in range(nsm)]
    return groups, eff.value


def get_nel(ndim, P, S):
    return lib().ref_get_nel(I64(ndim), _p(np.ascontiguousarray(P)), _p(np.ascontiguousarray(S)))


def band_structure(ndim, P, H):
    return lib().ref_band_structure(I64(ndim), _p(np.ascontiguousarray(P)),
                                    _p(np.ascontiguousarray(H)))


def thresholds(ndim):
    """(high, normal, low) of adjust_thresholds_dp (source/metrics.f90:82-113)."""
    out = np.zeros(3)
    lib().ref_thresholds(I64(ndim), _p(out))
    return tuple(out)


def check_density(ndim, P, S, nel) -> str:
    log = []
    with quiet(log):
        lib().ref_check_density(I64(ndim), _p(np.ascontiguousarray(P)),
                                _p(np.ascontiguousarray(S)), I64(nel))
    return log[0]


def purification(ndim, nel, ptype, metric, iterative, cycles, metric_cycles, chempot_cycles,
                 chempot_guess, chempot_increment, H, S, prlvl=0, log=None):
    P = np.zeros(ndim * (ndim + 1) // 2)
    with quiet(log):
        lib().ref_purification(I64(ndim), I64(nel), I64(ptype), I64(metric), I64(iterative),
                               I64(cycles), I64(metric_cycles), I64(chempot_cycles),
                               F64(chempot_guess), F64(chempot_increment), I64(prlvl),
                               _p(np.ascontiguousarray(H)), _p(np.ascontiguousarray(S)), _p(P))
    return P


def fermismear(norbs, nel, t, eig):
    occ = np.zeros(norbs)
    fod, ef, s = F64(), F64(), F64()
    with quiet():
        lib().ref_fermismear(I64(norbs), I64(nel), F64(t),
                             _p(np.ascontiguousarray(eig, dtype=np.float64)), _p(occ),
                             ctypes.byref(fod), ctypes.byref(ef), ctypes.byref(s))
    return occ, ef.value, fod.value, s.value


def occu(ndim, nel, nopen=0):
    no = I64(nopen)
    ia, ib = I64(), I64()
    fa, fb = np.zeros(ndim), np.zeros(ndim)
    lib().ref_occu(I64(ndim), I64(nel), ctypes.byref(no), ctypes.byref(ia), ctypes.byref(ib),
                   _p(fa), _p(fb))
    return fa, fb, ia.value, ib.value


def sygvd(A, B):
    """la_sygvd on copies (column-major views of the symmetric inputs)."""
    n = A.shape[0]
    a = np.asfortranarray(A.copy())
    b = np.asfortranarray(B.copy())
    w = np.zeros(n)
    info = I64()
    lib().ref_sygvd(I64(n), _p(a), _p(b), _p(w), ctypes.byref(info))
    return w, a, info.value


def mlpop14(nao, S):
    """(s34, s14) float32 (nao x nao, column-major as Fortran arrays) from
    packed S (source/mlpop.f:142-190)."""
    s34 = np.zeros((nao, nao), dtype=np.float32, order="F")
    s14 = np.zeros((nao, nao), dtype=np.float32, order="F")
    lib().ref_mlpop14(I64(nao), _p(np.ascontiguousarray(S, dtype=np.float64)), _p(s34), _p(s14))
    return s34, s14


def mlpop2(n, nao, P, s1, s2, aoat, shell2ao):
    """(q(n), qsh(10, n)) of source/mlpop.f:198-238."""
    q = np.zeros(n)
    qsh = np.zeros((10, n), order="F")
    lib().ref_mlpop2(I64(n), I64(nao), _p(np.ascontiguousarray(P, dtype=np.float64)),
                     _p(np.asfortranarray(s1, dtype=np.float32)),
                     _p(np.asfortranarray(s2, dtype=np.float32)),
                     _p(np.ascontiguousarray(aoat, dtype=np.int64)),
                     _p(np.ascontiguousarray(shell2ao, dtype=np.int64)), _p(q), _p(qsh))
    return q, qsh
