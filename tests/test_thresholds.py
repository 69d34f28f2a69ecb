"""adjust_thresholds_dp (source/metrics.f90:82-113), the size-dependent
thresholds of the purification and the density checks: the oracle and the
engine's C ABI (host code, no device) against the compiled reference's values
on both sides of every tier boundary (tests/golden/thresholds.npz: ndim 1, 46,
299 | 300, 1999 | 2000, 14999 | 15000, 20000)."""
import os

import numpy as np

from conftest import ROOT


def _fixture():
    return np.load(os.path.join(ROOT, "tests", "golden", "thresholds.npz"))


def test_oracle_thresholds_match_reference():
    from oracle import ptb_oracle as O
    f = _fixture()
    for n, vals in zip(f["ndim"], f["vals"]):
        t = O.adjust_thresholds_dp(int(n))
        assert (t.high, t.normal, t.low) == tuple(vals), n


def test_engine_thresholds_match_reference():
    from ptb_amd.engine import thresholds
    f = _fixture()
    for n, vals in zip(f["ndim"], f["vals"]):
        assert thresholds(int(n)) == tuple(vals), n
    # every tier appears: the boundaries are where the reference puts them
    assert len({tuple(v) for v in f["vals"]}) == 4
