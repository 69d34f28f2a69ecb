"""Shared test setup: the `gpu` marker and small deterministic systems."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def small_system(kind="water", size=(2, 2, 2), seed=0, filt=1e-4):
    """(mol, H, S) with H and S filtered in place like twoscf does
    (source/pgtb.f90:566-578)."""
    from ptb_amd import geometry, surrogate
    from oracle import ptb_oracle as O
    if kind == "water":
        mol = geometry.water_box(size, 3.1, 0.2, seed)
    else:
        mol = geometry.alkane(size)
    H, S = surrogate.build_hs(mol)
    O.filter_matrix_sym(S, mol.ndim, filt)
    O.filter_matrix_sym(H, mol.ndim, filt)
    return mol, H, S


@pytest.fixture(scope="session")
def engine():
    from ptb_amd.engine import Engine
    eng = Engine(0)
    yield eng
    eng.close()
