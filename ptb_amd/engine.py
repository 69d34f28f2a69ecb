"""Python mirror of the reference's hot-path interface over the C ABI.

Names, argument meaning and error behaviour follow the reference routines:

* ``filter_matrix_sym(A, n, filter)``      source/pgtb.f90:1532-1548 (in place)
* ``analyse_matrix_sym(A, n)``             source/pgtb.f90:1549-1579
* ``sm_stupid_simple(ndim, nelref, Hvec, Svec, Pvec, n, xyz, aoat,
  submatrix_columns, submatrix_mode)``     source/sm.f90:34-409 (Pvec overwritten)
* ``get_nel``, ``band_structure``, ``idempotent``, ``check_density``
                                           source/metrics.f90
* ``purification_wrapper(pur, ndim, H, S, P)``
                                           source/purification.f90:28-80

Where the reference stops the process (unsupported mode, k-means fault, NaN),
the engine returns a status code and these wrappers raise ``PtbGpuError``.
Arrays are numpy (host, the drop-in path) or torch CUDA tensors for the
``*_dev`` variants (inputs resident in HBM).  There is no CPU fallback: a
missing libptbgpu.so raises at load time.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native

I64 = ctypes.c_int64
F64 = ctypes.c_double
VP = ctypes.c_void_p

PRECISION = {"fp64": 0, "mixed": 1, "bf16x3": 1, "mixed_bf16": 2, "bf16": 2}
GEMM_PRECISION = {"fp64": 0, "bf16x3": 1, "bf16": 2}
COLUMNS = {"single_columns": 1, "atoms_columns": 2, "full_matrix": 3, "kmeans_r_heuristic": 4}

PHASES = ["alloc", "submatrix combination", "submatrix stats", "submatrix construction",
          "submatrix eigendecomp", "submatrix fe", "mu iter", "submatrix final iter",
          "submatrix dealloc", "total"]


class PtbGpuError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"ptbgpu status {status}: {message}")
        self.status = status


class Opts(ctypes.Structure):
    _fields_ = [("precision", I64), ("strict", I64), ("verbose", I64), ("refine_steps", I64),
                ("kmeans_stream", ctypes.c_void_p), ("nstream", I64)]


class Report(ctypes.Structure):
    _fields_ = [("mu", F64), ("nel", F64), ("muiter", I64), ("nsm", I64), ("sum_m", I64),
                ("max_m", I64), ("sum_m3", F64), ("not_spd", I64), ("t_phase", F64 * 12)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "t_phase"}
        d["t_phase"] = {PHASES[i]: self.t_phase[i] for i in range(len(PHASES))}
        return d


class PurSettings(ctypes.Structure):
    """tPurificationSet fields read by purification_wrapper, with the reference's
    defaults (source/purification_settings.f90:29-86)."""
    _fields_ = [("type", I64), ("metric", I64), ("metric_iterative", I64), ("cycles", I64),
                ("metric_cycles", I64), ("chempot_cycles", I64), ("chempot_guess", F64),
                ("chempot_increment", F64), ("nel", I64)]

    def __init__(self, type=4, metric=2, metric_iterative=False, cycles=40, metric_cycles=50,
                 chempot_cycles=60, chempot_guess=-1.5, chempot_increment=0.1, nel=0):
        super().__init__(int(type), int(metric), 1 if metric_iterative else 0, int(cycles),
                         int(metric_cycles), int(chempot_cycles), float(chempot_guess),
                         float(chempot_increment), int(nel))


class PurReport(ctypes.Structure):
    _fields_ = [("chempot", F64), ("nel", F64), ("cycles", I64), ("metric_cycles", I64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


PUR_TYPES = {"mcweeny": 1, "sign_iter_pade": 2, "sign_iter_newton": 3, "sign_diagonalization": 4}
PUR_METRICS = {"inv": 1, "inv_sqrt": 2, "sqrt": 3}


def _lib():
    lib = _native.engine_lib()
    if getattr(lib, "_ptb_typed", False):
        return lib
    lib.ptbgpu_strerror.restype = ctypes.c_char_p
    lib.ptbgpu_strerror.argtypes = [I64]
    lib.ptbgpu_last_error.restype = ctypes.c_char_p
    lib.ptbgpu_last_error.argtypes = [VP]
    lib.ptbgpu_init.argtypes = [ctypes.POINTER(VP), I64, I64]
    for name in ("ptbgpu_finalize", "ptbgpu_sync"):
        getattr(lib, name).argtypes = [VP]
    lib.ptbgpu_stream.argtypes = [VP, ctypes.POINTER(VP)]
    lib.ptbgpu_profile.argtypes = [VP, I64]
    lib.ptbgpu_profile_query.argtypes = [VP, ctypes.c_char_p, ctypes.POINTER(F64),
                                         ctypes.POINTER(I64)]
    lib.ptbgpu_filter_sym.argtypes = [VP, I64, VP, F64, VP]
    lib.ptbgpu_filter_sym_dev.argtypes = [VP, I64, VP, F64, VP]
    lib.ptbgpu_analyse_sym.argtypes = [VP, I64, VP, VP]
    lib.ptbgpu_analyse_sym_dev.argtypes = [VP, I64, VP, VP]
    sm_args = [VP, I64, I64, VP, VP, VP, I64, VP, VP, I64, I64, ctypes.POINTER(Opts),
               ctypes.POINTER(Report)]
    for name in ("ptbgpu_sm_density", "ptbgpu_sm_density_dev"):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = sm_args
    batch_args = [VP, I64, VP, VP, VP, VP, VP, VP, VP, VP, I64, I64, ctypes.POINTER(Opts),
                  VP]
    for name in ("ptbgpu_sm_density_batch", "ptbgpu_sm_density_batch_dev"):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = batch_args
    if hasattr(lib, "ptbgpu_sm_spectra"):
        lib.ptbgpu_sm_spectra.argtypes = [VP, I64, I64, VP, VP, VP]
    for name in ("ptbgpu_get_nel", "ptbgpu_band_energy", "ptbgpu_idempotency",
                 "ptbgpu_get_nel_dev", "ptbgpu_band_energy_dev"):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = [VP, I64, VP, VP, ctypes.POINTER(F64)]
    for name in ("ptbgpu_matrix_sign", "ptbgpu_matrix_root"):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = [VP, I64, VP, VP]
    if hasattr(lib, "ptbgpu_thresholds"):
        lib.ptbgpu_thresholds.argtypes = [I64, VP]
    if hasattr(lib, "ptbgpu_fermismear"):
        lib.ptbgpu_fermismear.argtypes = [I64, I64, F64, VP, VP, ctypes.POINTER(F64),
                                          ctypes.POINTER(F64), ctypes.POINTER(F64)]
    if hasattr(lib, "ptbgpu_solve2"):
        lib.ptbgpu_solve2.argtypes = [VP, I64, I64, I64, I64, F64, F64, F64, VP, VP, VP, VP, VP,
                                      VP, ctypes.POINTER(I64)]
    for name in ("ptbgpu_purify", "ptbgpu_purify_dev"):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = [VP, ctypes.POINTER(PurSettings), I64, VP, VP, VP,
                                           ctypes.POINTER(PurReport)]
    for name in ("ptbgpu_mlpop2", "ptbgpu_mlpop2_dev"):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = [VP, I64, I64, VP, VP, VP, VP, VP, VP, VP]
    for name in ("ptbgpu_set_tuning",):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = [VP, ctypes.c_char_p, I64]
    if hasattr(lib, "ptbgpu_graph_note"):
        lib.ptbgpu_graph_note.restype = ctypes.c_char_p
        lib.ptbgpu_graph_note.argtypes = [VP]
    if hasattr(lib, "ptbgpu_get_tuning"):
        lib.ptbgpu_get_tuning.argtypes = [VP, ctypes.c_char_p, ctypes.POINTER(I64)]
    if hasattr(lib, "ptbgpu_gemm_batch_dev"):
        lib.ptbgpu_gemm_batch_dev.argtypes = [VP, I64, I64, I64, I64, VP, VP, VP, VP, VP, VP, VP,
                                              VP, VP, F64, F64]
    lib._ptb_typed = True
    return lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(VP)
    return VP(a.data_ptr())   # torch tensor


def _f64(a, n, name):
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError(f"{name} must be a contiguous float64 numpy array")
    if a.size < n:
        raise ValueError(f"{name} has {a.size} elements, needs {n}")
    return a


def _dev_f64(t, n, name, device):
    """A device operand of a ``*_dev`` call: a contiguous float64 torch tensor on
    the engine's device with at least ``n`` elements.  The kernels index it up
    to ``n``; anything else would make them read or write outside it."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on cuda:{device}")
    if not t.is_cuda or t.device.index != device:
        raise ValueError(f"{name} is on {t.device}, the engine runs on cuda:{device}")
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be float64, not {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() < n:
        raise ValueError(f"{name} has {t.numel()} elements, needs {n}")
    return t


def _producers_done(device):
    """The engine's stream does not wait on other streams: work queued on
    torch's current stream (the tensors' producers) must be complete first."""
    import torch
    torch.cuda.current_stream(device).synchronize()


def fermismear(norbs, nel, t, eig):
    """FERMISMEAR (source/fermi.f:5-67) on the host through the C ABI (no
    device context): returns (occ, e_fermi, fod, s)."""
    lib = _lib()
    e = np.ascontiguousarray(eig, dtype=np.float64)
    occ = np.zeros(norbs)
    fod, ef, s = F64(), F64(), F64()
    st = lib.ptbgpu_fermismear(norbs, nel, t, _ptr(e), _ptr(occ), ctypes.byref(fod),
                               ctypes.byref(ef), ctypes.byref(s))
    if st != 0:
        raise PtbGpuError(st, lib.ptbgpu_strerror(st).decode())
    return occ, ef.value, fod.value, s.value


def thresholds(ndim):
    """adjust_thresholds_dp (source/metrics.f90:82-113) through the C ABI
    (host only): (high, normal, low) for a problem of ndim basis functions."""
    lib = _lib()
    out = np.zeros(3)
    st = lib.ptbgpu_thresholds(ndim, _ptr(out))
    if st != 0:
        raise PtbGpuError(st, lib.ptbgpu_strerror(st).decode())
    return tuple(out)


class Engine:
    """One engine context (one HIP stream) on one device."""

    def __init__(self, device: int = 0):
        self.lib = _lib()
        self.ctx = VP()
        st = self.lib.ptbgpu_init(ctypes.byref(self.ctx), device, 0)
        if st != 0:
            raise PtbGpuError(st, self.lib.ptbgpu_strerror(st).decode())
        self.device = device
        self.apply_tuning()

    def set_tuning(self, key: str, value: int):
        """A tuning knob of this context (ptbgpu_set_tuning)."""
        self._check(self.lib.ptbgpu_set_tuning(self.ctx, key.encode(), int(value)))

    def get_tuning(self, key: str) -> int:
        v = I64(0)
        self._check(self.lib.ptbgpu_get_tuning(self.ctx, key.encode(), ctypes.byref(v)))
        return v.value

    def graph_note(self) -> str:
        return self.lib.ptbgpu_graph_note(self.ctx).decode()

    def apply_tuning(self, path: str | None = None):
        """Apply the autotuner's choices (ptb_amd/tuning.json by default, written
        by ``python -m ptb_amd.tune``); a missing file leaves the defaults."""
        import json
        import os
        path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
        if not os.path.exists(path):
            return {}
        with open(path) as fh:
            chosen = json.load(fh).get("chosen", {})
        for k, v in chosen.items():
            self.set_tuning(k, v)
        return chosen

    def close(self):
        if self.ctx:
            self.lib.ptbgpu_finalize(self.ctx)
            self.ctx = VP()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int):
        if st != 0:
            msg = self.lib.ptbgpu_last_error(self.ctx).decode() or \
                self.lib.ptbgpu_strerror(st).decode()
            raise PtbGpuError(st, msg)

    def stream_handle(self) -> int:
        s = VP()
        self._check(self.lib.ptbgpu_stream(self.ctx, ctypes.byref(s)))
        return s.value or 0

    def sync(self):
        self._check(self.lib.ptbgpu_sync(self.ctx))

    def profile(self, enable: bool):
        """Start (clearing old records) or stop per-kernel event timing."""
        self._check(self.lib.ptbgpu_profile(self.ctx, 1 if enable else 0))

    def profile_query(self, name: str):
        """(total milliseconds, launch groups) recorded for one kernel name."""
        t = F64(0.0)
        c = I64(0)
        self._check(self.lib.ptbgpu_profile_query(self.ctx, name.encode(), ctypes.byref(t),
                                                  ctypes.byref(c)))
        return t.value, c.value

    # ---- filter / analyse -------------------------------------------------
    def filter_matrix_sym(self, A: np.ndarray, n: int, filter_: float):
        """In-place filter of a packed matrix; returns positive counts
        (before, after) as analyse_matrix_sym prints them."""
        _f64(A, n * (n + 1) // 2, "A")
        c = np.zeros(2, dtype=np.int64)
        self._check(self.lib.ptbgpu_filter_sym(self.ctx, n, _ptr(A), filter_, _ptr(c)))
        return int(c[0]), int(c[1])

    def filter_matrix_sym_dev(self, A, n: int, filter_: float):
        _dev_f64(A, n * (n + 1) // 2, "A", self.device)
        _producers_done(self.device)
        c = np.zeros(2, dtype=np.int64)
        self._check(self.lib.ptbgpu_filter_sym_dev(self.ctx, n, _ptr(A), filter_, _ptr(c)))
        return int(c[0]), int(c[1])

    def analyse_matrix_sym(self, A, n: int) -> dict:
        c = np.zeros(1, dtype=np.int64)
        if isinstance(A, np.ndarray):
            _f64(A, n * (n + 1) // 2, "A")
            self._check(self.lib.ptbgpu_analyse_sym(self.ctx, n, _ptr(A), _ptr(c)))
        else:
            _dev_f64(A, n * (n + 1) // 2, "A", self.device)
            _producers_done(self.device)
            self._check(self.lib.ptbgpu_analyse_sym_dev(self.ctx, n, _ptr(A), _ptr(c)))
        tot = int(c[0])
        return {"dimension": n, "nonzero": tot, "sparsity": tot / float(n) ** 2 if n else 0.0,
                "average": tot // n if n else 0}

    # ---- submatrix density build -----------------------------------------
    def _opts(self, precision, strict, verbose, refine_steps, kmeans_stream=None):
        o = Opts(PRECISION[precision] if isinstance(precision, str) else int(precision),
                 1 if strict else 0, 1 if verbose else 0, int(refine_steps), None, 0)
        if kmeans_stream is not None:
            self._stream = np.ascontiguousarray(kmeans_stream, dtype=np.float64)
            o.kmeans_stream = self._stream.ctypes.data
            o.nstream = self._stream.size
        return o

    def sm_stupid_simple(self, ndim, nelref, Hvec, Svec, Pvec, n, xyz, aoat,
                         submatrix_columns, submatrix_mode, precision="fp64", strict=False,
                         verbose=False, refine_steps=0, kmeans_stream=None) -> dict:
        """Drop-in for sm_stupid_simple on host arrays; Pvec is overwritten."""
        npk = ndim * (ndim + 1) // 2
        _f64(Hvec, npk, "Hvec")
        _f64(Svec, npk, "Svec")
        _f64(Pvec, npk, "Pvec")
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(n, 3)
        a = np.ascontiguousarray(aoat, dtype=np.int64)
        if a.size != ndim:
            raise ValueError("aoat must have ndim entries")
        rep = Report()
        opts = self._opts(precision, strict, verbose, refine_steps, kmeans_stream)
        self._check(self.lib.ptbgpu_sm_density(
            self.ctx, ndim, nelref, _ptr(Hvec), _ptr(Svec), _ptr(Pvec), n, _ptr(x), _ptr(a),
            submatrix_columns, submatrix_mode, ctypes.byref(opts), ctypes.byref(rep)))
        return rep.as_dict()

    def sm_stupid_simple_dev(self, ndim, nelref, dH, dS, dP, n, xyz, aoat,
                             submatrix_columns, submatrix_mode, precision="fp64",
                             strict=False, verbose=False, refine_steps=0,
                             kmeans_stream=None) -> dict:
        """Same with device-resident packed H, S, P (torch CUDA tensors)."""
        npk = ndim * (ndim + 1) // 2
        for t, name in ((dH, "dH"), (dS, "dS"), (dP, "dP")):
            _dev_f64(t, npk, name, self.device)
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(n, 3)
        a = np.ascontiguousarray(aoat, dtype=np.int64)
        if a.size != ndim:
            raise ValueError("aoat must have ndim entries")
        _producers_done(self.device)
        rep = Report()
        opts = self._opts(precision, strict, verbose, refine_steps, kmeans_stream)
        self._check(self.lib.ptbgpu_sm_density_dev(
            self.ctx, ndim, nelref, _ptr(dH), _ptr(dS), _ptr(dP), n, _ptr(x), _ptr(a),
            submatrix_columns, submatrix_mode, ctypes.byref(opts), ctypes.byref(rep)))
        return rep.as_dict()

    def sm_stupid_simple_batch(self, mols, submatrix_columns, submatrix_mode, precision="fp64",
                               strict=False, verbose=False, refine_steps=0, kmeans_stream=None,
                               device=False) -> list:
        """Many independent molecules in one build (ptbgpu_sm_density_batch):
        ``mols`` is a list of (ndim, nelref, H, S, P, n, xyz, aoat) with packed
        numpy arrays (``device=False``; P is overwritten) or torch CUDA tensors
        (``device=True``).  Returns one report dict per molecule."""
        k = len(mols)
        ndim = np.array([m[0] for m in mols], dtype=np.int64)
        nel = np.array([m[1] for m in mols], dtype=np.int64)
        nat = np.array([m[5] for m in mols], dtype=np.int64)
        keep = []
        ptrs = {name: (VP * k)() for name in ("H", "S", "P", "xyz", "aoat")}
        for i, (nd, _, H, S, P, n, xyz, aoat) in enumerate(mols):
            npk = nd * (nd + 1) // 2
            check = _f64 if not device else \
                (lambda t, k, name: _dev_f64(t, k, f"{name}[{i}]", self.device))
            check(H, npk, "H")
            check(S, npk, "S")
            check(P, npk, "P")
            x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(n, 3)
            a = np.ascontiguousarray(aoat, dtype=np.int64)
            if a.size != nd:
                raise ValueError("aoat must have ndim entries")
            keep += [x, a]
            ptrs["H"][i] = _ptr(H).value
            ptrs["S"][i] = _ptr(S).value
            ptrs["P"][i] = _ptr(P).value
            ptrs["xyz"][i] = x.ctypes.data
            ptrs["aoat"][i] = a.ctypes.data
        if device:
            _producers_done(self.device)
        reps = (Report * k)()
        opts = self._opts(precision, strict, verbose, refine_steps, kmeans_stream)
        fn = self.lib.ptbgpu_sm_density_batch_dev if device else self.lib.ptbgpu_sm_density_batch
        self._check(fn(self.ctx, k, _ptr(ndim), _ptr(nel), ptrs["H"], ptrs["S"], ptrs["P"],
                       _ptr(nat), ptrs["xyz"], ptrs["aoat"], submatrix_columns, submatrix_mode,
                       ctypes.byref(opts), reps))
        return [r.as_dict() for r in reps]

    def sm_spectra(self, nsm: int, sum_m: int):
        """(m per submatrix, eigenvalues, fe) of the last build, split per submatrix."""
        m = np.zeros(nsm, dtype=np.int64)
        e = np.zeros(sum_m)
        fe = np.zeros(sum_m)
        self._check(self.lib.ptbgpu_sm_spectra(self.ctx, nsm, sum_m, _ptr(m), _ptr(e), _ptr(fe)))
        cuts = np.cumsum(m)[:-1]
        return m, np.split(e, cuts), np.split(fe, cuts)

    # ---- purification --------------------------------------------------------
    def purification_wrapper(self, pur: PurSettings, ndim: int, H, S, P) -> dict:
        """Drop-in for purification_wrapper(pur, ndim, H, S, P)
        (source/purification.f90:28-80) on packed host arrays; P is overwritten.
        Where the reference ends in ERROR STOP (NaN) this raises PtbGpuError
        with status 4."""
        npk = ndim * (ndim + 1) // 2
        _f64(H, npk, "H")
        _f64(S, npk, "S")
        _f64(P, npk, "P")
        rep = PurReport()
        self._check(self.lib.ptbgpu_purify(self.ctx, ctypes.byref(pur), ndim, _ptr(H), _ptr(S),
                                           _ptr(P), ctypes.byref(rep)))
        return rep.as_dict()

    def purification_wrapper_dev(self, pur: PurSettings, ndim: int, dH, dS, dP) -> dict:
        """Same with device-resident packed H, S, P (torch CUDA tensors)."""
        npk = ndim * (ndim + 1) // 2
        for t, name in ((dH, "dH"), (dS, "dS"), (dP, "dP")):
            _dev_f64(t, npk, name, self.device)
        _producers_done(self.device)
        rep = PurReport()
        self._check(self.lib.ptbgpu_purify_dev(self.ctx, ctypes.byref(pur), ndim, _ptr(dH),
                                               _ptr(dS), _ptr(dP), ctypes.byref(rep)))
        return rep.as_dict()

    # ---- dense spectral routines ---------------------------------------------
    def _spectral(self, fn, A):
        A = np.ascontiguousarray(A, dtype=np.float64)
        n = A.shape[0]
        out = np.zeros((n, n))
        # column-major in the ABI: pass A^T's buffer, read out^T back
        a = np.ascontiguousarray(A.T)
        self._check(fn(self.ctx, n, _ptr(a), _ptr(out)))
        return np.ascontiguousarray(out.T)

    def matrix_sign(self, A):
        """matrix_sign (source/sm.f90:410-435) of the dense symmetric A."""
        return self._spectral(self.lib.ptbgpu_matrix_sign, A)

    def matrix_root(self, A):
        """matrix_root (source/sm.f90:436-457): A^-1/2 of the dense SPD A."""
        return self._spectral(self.lib.ptbgpu_matrix_root, A)

    def solve2(self, mode, ndim, nel, nopen, et, H, S, shift=(0.0, 0.0)):
        """solve2 (source/pgtb.f90:1066-1170) on packed H, S: returns
        (P packed, e ascending, focc, U (ndim x ndim), fail)."""
        npk = ndim * (ndim + 1) // 2
        _f64(H, npk, "H")
        _f64(S, npk, "S")
        P = np.zeros(npk)
        e = np.zeros(ndim)
        focc = np.zeros(ndim)
        U = np.zeros((ndim, ndim))
        fail = I64(0)
        self._check(self.lib.ptbgpu_solve2(self.ctx, mode, ndim, nel, nopen, et, shift[0],
                                           shift[1], _ptr(H), _ptr(S), _ptr(P), _ptr(e),
                                           _ptr(focc), _ptr(U), ctypes.byref(fail)))
        return P, e, focc, np.ascontiguousarray(U.T), bool(fail.value)

    # ---- populations ---------------------------------------------------------
    def mlpop2(self, n, nao, P, s1, s2, aoat, shell2ao):
        """mlpop2 (source/mlpop.f:198-238): (q(n), qsh(10, n)) from packed P
        and the float32 S^3/4 (s1), S^1/4 (s2), nao x nao.  numpy arrays (host)
        or torch CUDA tensors for P, s1, s2 (device); s1/s2 are read
        column-major (pass a symmetric matrix or its transpose)."""
        a = np.ascontiguousarray(aoat, dtype=np.int64)
        sh = np.ascontiguousarray(shell2ao, dtype=np.int64)
        if a.size != nao or sh.size != nao:
            raise ValueError("aoat and shell2ao must have nao entries")
        q = np.zeros(n)
        qsh = np.zeros((n, 10))      # qsh(10, n) column-major
        if isinstance(P, np.ndarray):
            _f64(P, nao * (nao + 1) // 2, "P")
            s1 = np.asfortranarray(s1, dtype=np.float32)
            s2 = np.asfortranarray(s2, dtype=np.float32)
            fn = self.lib.ptbgpu_mlpop2
        else:
            import torch
            _dev_f64(P, nao * (nao + 1) // 2, "P", self.device)
            for t, name in ((s1, "s1"), (s2, "s2")):
                if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32
                        and t.is_contiguous() and t.numel() >= nao * nao):
                    raise TypeError(f"{name} must be a contiguous float32 CUDA tensor of nao^2")
            _producers_done(self.device)
            fn = self.lib.ptbgpu_mlpop2_dev
        self._check(fn(self.ctx, n, nao, _ptr(P), _ptr(s1), _ptr(s2), _ptr(a), _ptr(sh),
                       _ptr(q), _ptr(qsh)))
        return q, qsh.T.copy()

    # ---- matrix-core products --------------------------------------------------
    def gemm_batch_dev(self, precision, transa, transb, As, Bs, Cs, alpha=1.0, beta=0.0):
        """C_i = alpha op(A_i) op(B_i) + beta C_i for lists of 2-D float64 torch
        CUDA tensors (column-major views: pass A.t().contiguous() layouts as
        torch row-major arrays of the transposed shape; see tests).  Each tensor
        is read as a column-major matrix of shape (cols, rows) of its torch
        shape, i.e. a torch tensor X of shape (c, r) holds the r x c matrix X^T."""
        k = len(As)
        if not (len(Bs) == len(Cs) == k) or k < 1:
            raise ValueError("A, B, C lists must have the same nonzero length")
        M = np.zeros(k, dtype=np.int64)
        N = np.zeros(k, dtype=np.int64)
        K = np.zeros(k, dtype=np.int64)
        lda, ldb, ldc = (np.zeros(k, dtype=np.int64) for _ in range(3))
        pa, pb, pc = (VP * k)(), (VP * k)(), (VP * k)()
        for i, (a, b, c) in enumerate(zip(As, Bs, Cs)):
            for t, name in ((a, "A"), (b, "B"), (c, "C")):
                _dev_f64(t, t.numel(), f"{name}[{i}]", self.device)
                if t.dim() != 2:
                    raise ValueError(f"{name}[{i}] must be 2-D")
            ar, ac = a.shape[1], a.shape[0]      # stored (column-major) rows, cols
            br, bc = b.shape[1], b.shape[0]
            M[i], K[i] = (ac, ar) if transa else (ar, ac)
            kb, N[i] = (bc, br) if transb else (br, bc)
            if kb != K[i] or c.shape[1] != M[i] or c.shape[0] != N[i]:
                raise ValueError(f"product {i}: shapes do not conform")
            lda[i], ldb[i], ldc[i] = ar, br, M[i]
            pa[i], pb[i], pc[i] = a.data_ptr(), b.data_ptr(), c.data_ptr()
        _producers_done(self.device)
        prec = GEMM_PRECISION[precision] if isinstance(precision, str) else int(precision)
        self._check(self.lib.ptbgpu_gemm_batch_dev(
            self.ctx, prec, 1 if transa else 0, 1 if transb else 0, k, _ptr(M), _ptr(N), _ptr(K),
            pa, _ptr(lda), pb, _ptr(ldb), pc, _ptr(ldc), alpha, beta))

    # ---- metrics -----------------------------------------------------------
    def _scalar(self, fn, ndim, A, B):
        out = F64(0.0)
        self._check(fn(self.ctx, ndim, _ptr(A), _ptr(B), ctypes.byref(out)))
        return out.value

    def get_nel(self, ndim, P, S) -> float:
        return self._scalar(self.lib.ptbgpu_get_nel, ndim, P, S)

    def get_nel_dev(self, ndim, dP, dS) -> float:
        for t, name in ((dP, "dP"), (dS, "dS")):
            _dev_f64(t, ndim * (ndim + 1) // 2, name, self.device)
        _producers_done(self.device)
        return self._scalar(self.lib.ptbgpu_get_nel_dev, ndim, dP, dS)

    def band_structure_dev(self, ndim, dP, dH) -> float:
        for t, name in ((dP, "dP"), (dH, "dH")):
            _dev_f64(t, ndim * (ndim + 1) // 2, name, self.device)
        _producers_done(self.device)
        return self._scalar(self.lib.ptbgpu_band_energy_dev, ndim, dP, dH)

    def band_structure(self, ndim, P, H) -> float:
        return self._scalar(self.lib.ptbgpu_band_energy, ndim, P, H)

    def idempotency_deviation(self, ndim, P, S) -> float:
        return self._scalar(self.lib.ptbgpu_idempotency, ndim, P, S)
