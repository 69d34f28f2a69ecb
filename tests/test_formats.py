"""On-disk formats (ptb_amd/formats.py): the .PUR parser of the engine library
against the oracle's restatement of initialize_purification
(source/purification_settings.f90:91-218), Fortran sequential records (the
ptb_dump framing, source/pgtb.f90:275-290) and pencil files."""
import numpy as np
import pytest

from conftest import small_system

# the reference checkout's own .PUR (its only input of this kind), as data
REFERENCE_PUR = """submatrix_columns = kmeans_r_heuristic
!submatrix_columns = full_matrix
!submatrix_columns = atoms_columns
submatrix_mode = submatrix_sygv
mode = submatrix
check
"""

CORPUS = [
    "",
    REFERENCE_PUR,
    "type = mcweeny\nmetric = inverse\ncycles = 25\n",
    "type = pade\nmetric = sqrt\niterative_inversion\nchempot_cycles = 12.9\n",
    "chempot_increment = 0.1\npr = 2\n",          # int() truncation: 0.1 -> 0
    "chempot_increment = 3.7\n",
    "\ttype\t=\tnewton  \n  submatrix_columns   =   single_columns\n",
    "mode = full\ndev\nunknown_key = 5\n! type = pade\n",
    "cycles = -3\ncycles = 1e2\n",
    "submatrix_columns = atoms_columns   trailing words\n",
    "x" * 90 + "\n",                              # beyond 80 characters
    "type = diag\r\nmetric = inverse_sqrt\r\n",
]


def _oracle_dict(text):
    from oracle import ptb_oracle as O
    s = O.parse_pur(text)
    return {"mode": s.mode, "submatrix_columns": s.submatrix_columns,
            "submatrix_mode": s.submatrix_mode, "prlvl": s.prlvl, "dev": int(s.dev),
            "check": int(s.check), "type": s.type, "metric": s.metric_type,
            "metric_iterative": int(s.metric_iterative), "cycles": s.cycles,
            "metric_cycles": s.metric_cycles, "chempot_cycles": s.chempot_cycles,
            "chempot_guess": s.chempot_guess, "chempot_increment": s.chempot_increment}


@pytest.mark.parametrize("text", CORPUS)
def test_pur_parser_matches_oracle(text):
    from ptb_amd import formats
    got = formats.parse_pur(text)
    want = _oracle_dict(text.replace("\r", ""))
    for k, v in want.items():
        assert got[k] == v, (k, got[k], v)


def test_pur_reference_file_and_quirks():
    from ptb_amd import formats
    d = formats.parse_pur(REFERENCE_PUR)
    assert (d["mode"], d["submatrix_columns"], d["submatrix_mode"], d["check"]) == (2, 4, 2, 1)
    assert formats.parse_pur("chempot_increment = 0.1\n")["chempot_increment"] == 0.0
    assert formats.parse_pur("")["chempot_increment"] == 0.1      # the default survives


@pytest.mark.parametrize("text,msg", [("type = tc2\n", "invalid type"),
                                      ("metric = cholesky\n", "invalid metric"),
                                      ("mode = \n", "invalid type")])
def test_pur_invalid_values_are_errors(text, msg):
    from ptb_amd import formats
    from oracle import ptb_oracle as O
    with pytest.raises(formats.PurFormatError, match=msg):
        formats.parse_pur(text)
    with pytest.raises(O.PurError):
        O.parse_pur(text)


def test_records_round_trip_and_subrecords(tmp_path, monkeypatch):
    from ptb_amd import formats
    gen = np.random.default_rng(0)
    arrays = [gen.standard_normal(17), np.arange(6.0).reshape(2, 3), np.zeros(0),
              gen.standard_normal(1000)]
    p = tmp_path / "r.bin"
    formats.write_records(str(p), arrays)
    raw = p.read_bytes()
    # framing: 4-byte length, payload, 4-byte length
    assert int.from_bytes(raw[:4], "little") == 17 * 8
    assert int.from_bytes(raw[4 + 136:8 + 136], "little") == 17 * 8
    recs = formats.read_records(str(p))
    for a, r in zip(arrays, recs):
        assert r == np.asfortranarray(a).tobytes(order="F")
    # payloads above the subrecord limit are split with negated markers
    monkeypatch.setattr(formats, "_SUBREC_MAX", 40)
    formats.write_records(str(p), [np.arange(13.0)])
    raw = p.read_bytes()
    assert int.from_bytes(raw[:4], "little", signed=True) == -40
    assert formats.read_records(str(p))[0] == np.arange(13.0).tobytes()


def test_ptb_dump_round_trip(tmp_path):
    from ptb_amd import formats
    n, ndim, cnsao = 3, 7, 4
    gen = np.random.default_rng(1)
    f = {"pa": gen.random(n), "psh": gen.random((10, n)), "wbo": gen.random((n, n)),
         "P": gen.random(ndim * (ndim + 1) // 2), "qm": gen.random(n),
         "dipm": gen.random((3, n)), "qp": gen.random((6, n)), "eps": gen.random(ndim),
         "Scv": gen.random((cnsao, ndim)), "norm": gen.random(ndim)}
    p = tmp_path / "ptb_dump"
    formats.write_ptb_dump(str(p), f)
    back = formats.read_ptb_dump(str(p), n, ndim)
    for k, v in f.items():
        assert np.array_equal(back[k], v), k
    with pytest.raises(ValueError):
        formats.read_ptb_dump(str(p), n + 1, ndim)


def test_pencil_round_trip_is_bit_exact(tmp_path):
    from ptb_amd import formats
    mol, H, S = small_system("water", (2, 2, 1), 0)
    P = np.where(np.abs(H) > 1e-2, H * 0.5, 0.0)
    p = tmp_path / "pencil.npz"
    formats.save_pencil(str(p), mol.ndim, {"H": H, "S": S, "P": P}, nel=mol.nel, natoms=mol.n,
                        xyz=mol.xyz, aoat=mol.aoat, filter=1e-4)
    ndim, mats, meta = formats.load_pencil(str(p))
    assert ndim == mol.ndim and int(meta["nel"]) == mol.nel
    for name, v in (("H", H), ("S", S), ("P", P)):
        assert mats[name].tobytes() == v.tobytes(), name
    assert np.array_equal(meta["aoat"], mol.aoat)
