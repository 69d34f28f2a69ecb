"""Density checks (source/metrics.f90) against the oracle."""
import numpy as np
import pytest

from conftest import small_system

pytestmark = pytest.mark.gpu


def _density():
    from oracle import ptb_oracle as O
    mol, H, S = small_system(size=(2, 2, 2))
    res = O.sm_stupid_simple(mol.ndim, mol.nel, H, S, mol.n, mol.xyz, mol.aoat, 2, 2)
    return mol, H, S, res.P


def test_traces(engine):
    from oracle import ptb_oracle as O
    mol, H, S, P = _density()
    n = mol.ndim
    assert abs(engine.get_nel(n, P, S) - O.get_nel(n, P, S)) < 1e-11
    assert abs(engine.band_structure(n, P, H) - O.band_structure(n, P, H)) < 1e-10


def test_idempotency_deviation(engine):
    from oracle import ptb_oracle as O
    mol, H, S, P = _density()
    n = mol.ndim
    got = engine.idempotency_deviation(n, P, S)
    ref = O.idempotency_deviation(n, P, S)
    assert abs(got - ref) <= 1e-9 * max(1.0, abs(ref))


def test_idempotency_of_exact_projector(engine):
    """Full-matrix P = 2 C_occ C_occ^T with S-orthonormal C: 0.5 PSP == P."""
    import scipy.linalg as sla
    from oracle import ptb_oracle as O
    mol, H, S = small_system(size=(2, 2, 1))
    n = mol.ndim
    e, C = sla.eigh(O.blowsym(H, n), O.blowsym(S, n))
    occ = np.zeros(n)
    occ[: mol.nel // 2] = 2.0
    P = O.packsym(O.dmat(occ, C))
    assert engine.idempotency_deviation(n, P, S) < 1e-9
    assert abs(engine.get_nel(n, P, S) - mol.nel) < 1e-10


def test_device_traces_equal_host_traces(engine):
    """get_nel / band_structure on device-resident matrices give exactly the
    host entry points' values (same kernels, same reduction order)."""
    import torch
    mol, H, S, P = _density()
    n = mol.ndim
    dP, dS, dH = (torch.from_numpy(x.copy()).cuda() for x in (P, S, H))
    assert engine.get_nel_dev(n, dP, dS) == engine.get_nel(n, P, S)
    assert engine.band_structure_dev(n, dP, dH) == engine.band_structure(n, P, H)
