"""The Fortran drop-in (module sm over the C ABI) driven as twoscf drives the
reference: filter S and H in place, sm_stupid_simple, N_el and Tr(HP)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "ptb_amd", "lib", "sm_gpu_driver")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
@pytest.mark.parametrize("cols", [2, 3])
def test_fortran_drop_in(tmp_path, cols):
    from ptb_amd import geometry, surrogate
    from oracle import ptb_oracle as O
    mol = geometry.water_box((2, 2, 2), 3.1, 0.2, 0)
    H, S = surrogate.build_hs(mol)
    n = mol.ndim
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([n, mol.nel, mol.n, cols, 2], dtype=np.int64).tofile(f)
        np.array([1e-4]).tofile(f)
        np.ascontiguousarray(mol.xyz).tofile(f)   # (n,3) C order == xyz(3,n)
        np.ascontiguousarray(mol.aoat, dtype=np.int64).tofile(f)
        H.tofile(f)
        S.tofile(f)
    out = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True,
                         timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    raw = np.fromfile(fout, dtype=np.float64)
    P, nelp, eband = raw[:-2], raw[-2], raw[-1]
    Hf, Sf = H.copy(), S.copy()
    O.filter_matrix_sym(Sf, n, 1e-4)
    O.filter_matrix_sym(Hf, n, 1e-4)
    res = O.sm_stupid_simple(n, mol.nel, Hf, Sf, mol.n, mol.xyz, mol.aoat, cols, 2)
    assert np.abs(P - res.P).max() <= 1e-8
    assert abs(eband - O.band_structure(n, res.P, Hf)) <= 1e-7
    assert abs(nelp - O.get_nel(n, res.P, Sf)) <= 1e-9


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
@pytest.mark.parametrize("fixture", ["sm_w16_kmeans.npz", "sm_c40_kmeans.npz"])
def test_fortran_kmeans_matches_reference(tmp_path, fixture):
    """The drop-in draws the k-means numbers from the compiler's generator with
    the reference's seeding, so mode 4 reproduces the reference's density
    (tests/golden/sm_*_kmeans.npz, already filtered H and S)."""
    d = np.load(os.path.join(ROOT, "tests", "golden", fixture))
    n = int(d["ndim"])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([n, int(d["nel"]), int(d["n"]), 4, 2], dtype=np.int64).tofile(f)
        np.array([float(d["filter"])]).tofile(f)
        np.ascontiguousarray(d["xyz"]).tofile(f)
        np.ascontiguousarray(d["aoat"], dtype=np.int64).tofile(f)
        d["H"].tofile(f)
        d["S"].tofile(f)
    out = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True,
                         timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    P = np.fromfile(fout, dtype=np.float64)[:-2]
    assert np.abs(P - d["P"]).max() <= 1e-8
