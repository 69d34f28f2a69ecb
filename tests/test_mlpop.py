"""mlpop2 (source/mlpop.f:198-238), the first consumer of P.

CPU: the oracle against the compiled reference's mlpop14 + mlpop2 on the
reference's own densities of two systems (tests/golden/mlpop.npz; float
rounding of the REAL(4) products, bar 2e-6 per population), and the closed
form sum_i q_i = tr(P S) for exact S^1/4, S^3/4.
GPU: the engine's device mlpop2 (host and device entry points) against the
same fixtures; its products run in FP64 on the float-rounded inputs, so the
bar is the reference's own float rounding, 1e-5 per population.
"""
import os

import numpy as np
import pytest

from conftest import ROOT

CASES = ["h2o8", "c8h18"]


def _case(tag):
    d = np.load(os.path.join(ROOT, "tests", "golden", "mlpop.npz"))
    return {k[len(tag) + 1:]: d[k] for k in d.files if k.startswith(tag + "_")}


@pytest.mark.parametrize("tag", CASES)
def test_oracle_matches_reference(tag):
    from oracle import ptb_oracle as O
    c = _case(tag)
    q, qsh = O.mlpop2(int(c["n"]), int(c["nao"]), c["P"], c["s34"], c["s14"], c["aoat"],
                      c["shell2ao"])
    assert np.abs(q - c["q"]).max() <= 2e-6
    assert np.abs(qsh - c["qsh"]).max() <= 2e-6


def test_oracle_sum_is_trace_ps():
    """S^1/4 P S^3/4 has the trace of P S: with exact fractional powers (in
    double) the populations sum to get_nel."""
    import scipy.linalg as sla
    from conftest import small_system
    from oracle import ptb_oracle as O
    mol, H, S = small_system("water", (2, 2, 1), 0)
    n = mol.ndim
    P = O.sm_stupid_simple(n, mol.nel, H, S, mol.n, mol.xyz, mol.aoat, 2, 2).P
    w, V = sla.eigh(O.blowsym(S, n))
    s34 = (V * w ** 0.75) @ V.T
    s14 = (V * w ** 0.25) @ V.T
    q, qsh = O.mlpop2(mol.n, n, P, s34, s14, mol.aoat, mol.shell2ao)
    assert abs(q.sum() - O.get_nel(n, P, S)) < 1e-3 * mol.nel * 1e-2
    assert np.allclose(qsh.sum(axis=0), q, atol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("tag", CASES)
def test_gpu_mlpop2_matches_reference(engine, tag):
    import torch
    c = _case(tag)
    n, nao = int(c["n"]), int(c["nao"])
    q, qsh = engine.mlpop2(n, nao, c["P"].copy(), c["s34"], c["s14"], c["aoat"], c["shell2ao"])
    print(f"\n{tag}: max|q - q_ref| {np.abs(q - c['q']).max():.2e}, "
          f"max|qsh - qsh_ref| {np.abs(qsh - c['qsh']).max():.2e}")
    assert np.abs(q - c["q"]).max() <= 1e-5
    assert np.abs(qsh - c["qsh"]).max() <= 1e-5
    dP = torch.from_numpy(c["P"].copy()).cuda()
    # column-major nao x nao == the row-major buffer of the transpose
    d1 = torch.from_numpy(np.ascontiguousarray(c["s34"].T)).cuda()
    d2 = torch.from_numpy(np.ascontiguousarray(c["s14"].T)).cuda()
    qd, qshd = engine.mlpop2(n, nao, dP, d1, d2, c["aoat"], c["shell2ao"])
    assert np.array_equal(qd, q) and np.array_equal(qshd, qsh)
