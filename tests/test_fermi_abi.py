"""FERMISMEAR of the C ABI (host code, no device) against the reference's
outputs in tests/golden/fermi.npz (source/fermi.f:5-67)."""
import os

import numpy as np

from conftest import ROOT


def test_fermismear_matches_reference():
    from ptb_amd.engine import fermismear
    f = np.load(os.path.join(ROOT, "tests", "golden", "fermi.npz"))
    e, nel = f["eig"], int(f["nel"])
    for t, key in ((300.0, ""), (20000.0, "_hot")):
        occ, ef, fod, s = fermismear(len(e), nel // 2, t, e)
        assert np.abs(occ - f["occ" + key]).max() < 1e-12
        assert abs(ef - float(f["ef" + key])) < 1e-12
        assert abs(fod - float(f["fod" + key])) < 1e-10
        assert abs(s - float(f["s" + key])) < 1e-12
