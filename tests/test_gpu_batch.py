"""Many independent molecules in one build (ptbgpu_sm_density_batch): every
molecule's density equals what the single-molecule drop-in gives it alone
(same kernels, same arithmetic order), and matches the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(count, seed=2):
    """Half water clusters (filter 1e-4), half alkanes (filter 1e-5)."""
    from ptb_amd import geometry, surrogate
    from oracle import ptb_oracle as O
    out = []
    for mol in geometry.mixed_batch(count, seed):
        H, S = surrogate.build_hs(mol)
        filt = 1e-4 if mol.name.startswith("h2o") else 1e-5
        O.filter_matrix_sym(S, mol.ndim, filt)
        O.filter_matrix_sym(H, mol.ndim, filt)
        out.append((mol, H, S))
    return out


def _args(batch):
    return [(m.ndim, m.nel, H, S, np.full_like(H, np.nan), m.n, m.xyz, m.aoat)
            for m, H, S in batch]


@pytest.mark.parametrize("cols", [2, 1])
def test_batch_equals_single_molecule_builds(engine, cols):
    batch = _batch(10)
    args = _args(batch)
    reps = engine.sm_stupid_simple_batch(args, cols, 2)
    worst = 0.0
    for (mol, H, S), a, rep in zip(batch, args, reps):
        P1 = np.full_like(H, np.nan)
        r1 = engine.sm_stupid_simple(mol.ndim, mol.nel, H, S, P1, mol.n, mol.xyz, mol.aoat, cols, 2)
        assert rep["mu"] == r1["mu"] and rep["muiter"] == r1["muiter"]
        assert rep["nsm"] == r1["nsm"] and rep["sum_m"] == r1["sum_m"]
        worst = max(worst, float(np.abs(a[4] - P1).max()))
    print(f"\n{len(batch)} molecules, cols={cols}: max |P_batch - P_single| = {worst:.3e}")
    assert worst == 0.0


def test_batch_matches_oracle(engine):
    from oracle import ptb_oracle as O
    batch = _batch(6, seed=5)
    args = _args(batch)
    reps = engine.sm_stupid_simple_batch(args, 2, 2)
    for (mol, H, S), a, rep in zip(batch, args, reps):
        res = O.sm_stupid_simple(mol.ndim, mol.nel, H, S, mol.n, mol.xyz, mol.aoat, 2, 2)
        err = float(np.abs(a[4] - res.P).max())
        print(f"{mol.name:10s} ndim {mol.ndim:4d} nsm {rep['nsm']:3d} mu {rep['mu']:+.6f} "
              f"max|P - P_oracle| {err:.2e}")
        assert rep["mu"] == res.mu and rep["muiter"] == res.muiter
        assert err <= 1e-8
        assert abs(O.band_structure(mol.ndim, a[4], H) - O.band_structure(mol.ndim, res.P, H)) <= 1e-7


def test_batch_device_and_mixed(engine):
    import torch
    from oracle import ptb_oracle as O
    batch = _batch(8, seed=7)
    ref = _args(batch)
    engine.sm_stupid_simple_batch(ref, 2, 2)
    dev = [(m.ndim, m.nel, torch.from_numpy(H).cuda(), torch.from_numpy(S).cuda(),
            torch.zeros(H.size, dtype=torch.float64, device="cuda"), m.n, m.xyz, m.aoat)
           for m, H, S in batch]
    engine.sm_stupid_simple_batch(dev, 2, 2, device=True)
    torch.cuda.synchronize()
    for r, d in zip(ref, dev):
        assert np.array_equal(r[4], d[4].cpu().numpy())
    mixed = _args(batch)
    engine.sm_stupid_simple_batch(mixed, 2, 2, precision="mixed")
    for (mol, H, S), r, x in zip(batch, ref, mixed):
        assert np.abs(r[4] - x[4]).max() <= 1e-6
        assert abs(O.band_structure(mol.ndim, r[4], H) - O.band_structure(mol.ndim, x[4], H)) <= 1e-6


def test_batch256_seed2_equals_single_and_oracle(engine):
    """The benchmark's own batch: 256 molecules from seed 2 (128 water clusters
    of 4-40 H2O at filter 1e-4, 128 alkanes C3..C40 at filter 1e-5) in one
    build.  Every molecule's density is bit-identical to its single-molecule
    build; 16 molecules spread over the batch match the oracle to 1e-8 with
    identical mu and pass count."""
    from oracle import ptb_oracle as O
    batch = _batch(256, seed=2)
    args = _args(batch)
    reps = engine.sm_stupid_simple_batch(args, 2, 2)
    worst = 0.0
    for (mol, H, S), a, rep in zip(batch, args, reps):
        P1 = np.full_like(H, np.nan)
        r1 = engine.sm_stupid_simple(mol.ndim, mol.nel, H, S, P1, mol.n, mol.xyz, mol.aoat, 2, 2)
        assert rep["mu"] == r1["mu"] and rep["muiter"] == r1["muiter"]
        worst = max(worst, float(np.abs(a[4] - P1).max()))
    print(f"\n256 molecules ({sum(r['nsm'] for r in reps)} submatrices): "
          f"max |P_batch - P_single| = {worst:.3e}")
    assert worst == 0.0
    for i in np.linspace(0, 255, 16).round().astype(int):
        mol, H, S = batch[i]
        res = O.sm_stupid_simple(mol.ndim, mol.nel, H, S, mol.n, mol.xyz, mol.aoat, 2, 2)
        err = float(np.abs(args[i][4] - res.P).max())
        print(f"{i:3d} {mol.name:10s} ndim {mol.ndim:4d} mu {reps[i]['mu']:+.6f} "
              f"max|P - P_oracle| {err:.2e}")
        assert reps[i]["mu"] == res.mu and reps[i]["muiter"] == res.muiter
        assert err <= 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_is_bit_identical(engine, precision):
    """With graphs on, the eigen phase of a repeated build is captured into a
    hipGraph (second sight of the shapes) and replayed afterwards.  Every
    replay, with other eager builds and synchronisations in between, gives
    the eager density bit for bit."""
    import torch
    batch = _batch(10, seed=4)
    items = [(m.ndim, m.nel, torch.from_numpy(H).cuda(), torch.from_numpy(S).cuda(),
              torch.zeros(H.size, dtype=torch.float64, device="cuda"), m.n, m.xyz, m.aoat)
             for m, H, S in batch]
    other = _args(_batch(3, seed=9))
    engine.sm_stupid_simple_batch(items, 2, 2, precision=precision, device=True)
    torch.cuda.synchronize()
    ref = [it[4].cpu().numpy().copy() for it in items]
    engine.set_tuning("graph", 1)
    try:
        for rep in range(4):
            for it in items:
                it[4].fill_(float("nan"))
            torch.cuda.synchronize()
            engine.sm_stupid_simple_batch(items, 2, 2, precision=precision, device=True)
            torch.cuda.synchronize()
            print(f"build {rep}: {engine.graph_note()}")
            for r, it in zip(ref, items):
                assert np.array_equal(r, it[4].cpu().numpy()), f"build {rep}"
            engine.sm_stupid_simple_batch(other, 2, 2)   # eager work in between
            engine.sync()
        held = engine.get_tuning("graphs")
        print(f"\n{precision}: graphs held {held}")
        assert held >= 1
    finally:
        engine.set_tuning("graph", 0)
