"""On-disk formats around the density path.

* ``.PUR`` settings (source/purification_settings.f90:91-218), parsed by the
  engine library's ``ptbgpu_pur_parse`` (host code, no device needed).
* Fortran sequential unformatted records, the framing of PTB's ``ptb_dump``
  (``open(..., form='unformatted')`` + one ``write`` per array,
  source/pgtb.f90:275-290): each record is its payload between two 4-byte
  length markers; a payload of 2 GiB or more is split into subrecords whose
  leading markers are negated except on the last one.
* ``ptb_dump`` itself: pa(n), psh(10,n), wbo(n,n), P (packed upper,
  ndim(ndim+1)/2), qm(n), dipm(3,n), qp(6,n), eps(ndim), Scv(cnsao,ndim),
  norm(ndim) and, for open shells, one more record - all FP64.
* A compact pencil file for filtered packed matrices (the fixture format the
  tests, the engine and the oracle share): an ``.npz`` with the metadata
  (ndim, nel, natoms, xyz, aoat, filter) and each packed-upper matrix stored as
  its nonzero entries (CSC over the upper triangle: column pointers, row
  indices, values).  Loading restores the packed arrays bit for bit, zeros
  included.
"""
from __future__ import annotations

import ctypes
import os
import struct

import numpy as np

from . import _native

I64 = ctypes.c_int64
F64 = ctypes.c_double


class PurSettingsC(ctypes.Structure):
    _fields_ = [("type", I64), ("metric", I64), ("metric_iterative", I64), ("cycles", I64),
                ("metric_cycles", I64), ("chempot_cycles", I64), ("chempot_guess", F64),
                ("chempot_increment", F64), ("nel", I64)]


class PurFile(ctypes.Structure):
    """ptbgpu_pur_file: everything a .PUR sets, with the reference's defaults."""
    _fields_ = [("mode", I64), ("submatrix_columns", I64), ("submatrix_mode", I64),
                ("prlvl", I64), ("dev", I64), ("check", I64), ("pur", PurSettingsC)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "pur"}
        d.update({k: getattr(self.pur, k) for k, _ in PurSettingsC._fields_})
        return d


class PurFormatError(ValueError):
    """An invalid .PUR value (the reference ends in error stop)."""


def parse_pur(text: str) -> dict:
    """initialize_purification on the text of a .PUR file."""
    lib = _native.engine_lib()
    lib.ptbgpu_pur_parse.argtypes = [ctypes.c_char_p, I64, ctypes.POINTER(PurFile),
                                     ctypes.c_char_p, I64]
    raw = text.encode()
    out = PurFile()
    msg = ctypes.create_string_buffer(256)
    st = lib.ptbgpu_pur_parse(raw, len(raw), ctypes.byref(out), msg, 256)
    if st != 0:
        raise PurFormatError(msg.value.decode() or f"status {st}")
    return out.as_dict()


def read_pur(path: str = ".PUR") -> dict:
    """The settings of a .PUR file; the defaults when it does not exist (the
    reference reads the file only when present)."""
    if not os.path.exists(path):
        return parse_pur("")
    with open(path, "r", errors="replace") as fh:
        return parse_pur(fh.read())


# ---- Fortran sequential unformatted records --------------------------------
_SUBREC_MAX = 2**31 - 9   # largest payload of one subrecord


def write_records(path: str, arrays) -> None:
    """One record per array (its bytes in Fortran order)."""
    with open(path, "wb") as fh:
        for a in arrays:
            buf = np.asfortranarray(a).tobytes(order="F")
            n = len(buf)
            if n == 0:
                fh.write(struct.pack("<ii", 0, 0))
                continue
            pos = 0
            while pos < n:
                chunk = min(_SUBREC_MAX, n - pos)
                last = pos + chunk == n
                first = pos == 0
                lead = chunk if last else -chunk
                tail = chunk if first else -chunk
                fh.write(struct.pack("<i", lead))
                fh.write(buf[pos:pos + chunk])
                fh.write(struct.pack("<i", tail))
                pos += chunk


def read_records(path: str) -> list:
    """Every record's payload as bytes."""
    out = []
    with open(path, "rb") as fh:
        data = fh.read()
    pos, n = 0, len(data)
    while pos < n:
        parts = []
        while True:
            if pos + 4 > n:
                raise ValueError(f"{path}: truncated record marker at byte {pos}")
            lead = struct.unpack_from("<i", data, pos)[0]
            size = abs(lead)
            if pos + 8 + size > n:
                raise ValueError(f"{path}: truncated record at byte {pos}")
            parts.append(data[pos + 4: pos + 4 + size])
            tail = abs(struct.unpack_from("<i", data, pos + 4 + size)[0])
            if tail != size:
                raise ValueError(f"{path}: record markers {lead} / {tail} disagree at byte {pos}")
            pos += 8 + size
            if lead >= 0:
                break
        out.append(b"".join(parts))
    return out


PTB_DUMP_FIELDS = ["pa", "psh", "wbo", "P", "qm", "dipm", "qp", "eps", "Scv", "norm", "scal"]


def read_ptb_dump(path: str, n: int, ndim: int) -> dict:
    """The records of a ptb_dump file (source/pgtb.f90:275-290) as float64
    arrays: pa(n), psh(10,n), wbo(n,n), P packed, qm(n), dipm(3,n), qp(6,n),
    eps(ndim), Scv(cnsao,ndim), norm(ndim) [, scal].  Shapes are checked
    against n and ndim."""
    recs = read_records(path)
    if len(recs) < 10:
        raise ValueError(f"{path}: {len(recs)} records, a ptb_dump has at least 10")
    out = {}
    for name, rec in zip(PTB_DUMP_FIELDS, recs):
        out[name] = np.frombuffer(rec, dtype="<f8").copy()
    want = {"pa": n, "psh": 10 * n, "wbo": n * n, "P": ndim * (ndim + 1) // 2, "qm": n,
            "dipm": 3 * n, "qp": 6 * n, "eps": ndim, "norm": ndim}
    for name, size in want.items():
        if out[name].size != size:
            raise ValueError(f"{path}: record {name} has {out[name].size} values, expected {size}")
    if out["Scv"].size % ndim:
        raise ValueError(f"{path}: record Scv is not a multiple of ndim")
    out["psh"] = out["psh"].reshape(n, 10).T
    out["wbo"] = out["wbo"].reshape(n, n).T
    out["dipm"] = out["dipm"].reshape(n, 3).T
    out["qp"] = out["qp"].reshape(n, 6).T
    out["Scv"] = out["Scv"].reshape(ndim, -1).T
    return out


def write_ptb_dump(path: str, fields: dict) -> None:
    """A ptb_dump file from named arrays (the record order of the reference;
    ``scal`` only when given)."""
    names = PTB_DUMP_FIELDS[:10] + (["scal"] if "scal" in fields else [])
    write_records(path, [np.asarray(fields[k], dtype=np.float64) for k in names])


# ---- pencil files (filtered packed matrices) --------------------------------
def _packed_csc(v: np.ndarray, ndim: int):
    nz = np.flatnonzero(v)
    starts = np.arange(ndim + 1, dtype=np.int64) * (np.arange(ndim + 1, dtype=np.int64) + 1) // 2
    colptr = np.searchsorted(nz, starts).astype(np.int64)
    cols = np.repeat(np.arange(ndim, dtype=np.int64), np.diff(colptr))
    rows = (nz - starts[cols]).astype(np.int32)
    return colptr, rows, v[nz]


def save_pencil(path: str, ndim: int, matrices: dict, **meta) -> None:
    """matrices: name -> packed-upper float64 array (e.g. H, S, P); meta:
    nel, natoms, xyz, aoat, filter, ... (stored as given)."""
    out = {"format": np.array("ptb-pencil-1"), "ndim": np.int64(ndim)}
    for k, v in meta.items():
        out[f"meta_{k}"] = np.asarray(v)
    for name, v in matrices.items():
        v = np.ascontiguousarray(v, dtype=np.float64)
        if v.size != ndim * (ndim + 1) // 2:
            raise ValueError(f"{name}: {v.size} values, packed ndim {ndim} needs "
                             f"{ndim * (ndim + 1) // 2}")
        colptr, rows, vals = _packed_csc(v, ndim)
        out[f"{name}_colptr"], out[f"{name}_rows"], out[f"{name}_vals"] = colptr, rows, vals
    np.savez_compressed(path, **out)


def load_pencil(path: str):
    """(ndim, {name: packed array}, {meta name: value})."""
    d = np.load(path, allow_pickle=False)
    if str(d["format"]) != "ptb-pencil-1":
        raise ValueError(f"{path}: not a pencil file")
    ndim = int(d["ndim"])
    npk = ndim * (ndim + 1) // 2
    mats, meta = {}, {}
    for key in d.files:
        if key.startswith("meta_"):
            meta[key[5:]] = d[key]
        elif key.endswith("_colptr"):
            name = key[:-7]
            colptr, rows, vals = d[key], d[f"{name}_rows"], d[f"{name}_vals"]
            cols = np.repeat(np.arange(ndim, dtype=np.int64), np.diff(colptr))
            v = np.zeros(npk)
            v[cols * (cols + 1) // 2 + rows.astype(np.int64)] = vals
            mats[name] = v
    return ndim, mats, meta
