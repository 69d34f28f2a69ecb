"""The benchmark's own workloads on the GPU against the compiled reference.

For the 64-water box (filter 1e-4), the 512-water box the throughput is quoted
on (filter 1e-4, 1536 submatrices with m = 203..999) and C400H802 (filter
1e-5), all atoms_columns / submatrix_sygv, the fixtures under tests/golden/
hold what the reference's own sm_stupid_simple (source/sm.f90:34-409) and
filter_matrix_sym / analyse_matrix_sym (source/pgtb.f90:1532-1579) give on the
same surrogate pencils: positive counts before and after filtering, a SHA-256
digest of the filtered packed S and H, the mu trace, tr(HP), tr(PS), and the
packed columns of sampled column groups spread over the range of m.

Bars (FP64 drop-in): filter bytes and counts bit-exact; identical mu and pass
count; the sampled columns within 1e-8 per element with the same exact-zero
rows; |dE_band| <= 1e-7 Eh; |dN_el| <= 1e-10.  Mixed precision against the FP64
build: max |dP| <= 1e-6 and |dE_band| <= 1e-6 Eh.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden")
CONFIGS = ["c0_h2o64", "c1_h2o512", "c2_c400h802"]
_cache = {}


def _fixture(name):
    path = os.path.join(GOLD, f"sm_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{path} not generated")
    return np.load(path)


def _system(engine, name):
    """The workload's pencil, filtered on the GPU (S, then H, as twoscf does,
    source/pgtb.f90:566-578), checked byte for byte against the reference's."""
    if name in _cache:
        return _cache[name]
    from ptb_amd import geometry, surrogate
    d = _fixture(name)
    mol = geometry.config_molecule(str(d["workload"]))
    H, S = surrogate.build_hs(mol)
    n = mol.ndim
    assert n == int(d["ndim"]) and mol.nel == int(d["nel"])
    filt = float(d["filter"])
    cs = engine.filter_matrix_sym(S, n, filt)
    ch = engine.filter_matrix_sym(H, n, filt)
    print(f"\n{name}: ndim {n}, positive S {cs}, H {ch}")
    assert list(cs) + list(ch) == [int(x) for x in d["counts"][[0, 1, 2, 3]]]
    assert hashlib.sha256(S.tobytes()).hexdigest() == str(d["sha_S"])
    assert hashlib.sha256(H.tobytes()).hexdigest() == str(d["sha_H"])
    _cache.clear()
    _cache[name] = (mol, H, S)
    return _cache[name]


def _build(engine, mol, H, S, **kw):
    P = np.full_like(H, np.nan)
    rep = engine.sm_stupid_simple(mol.ndim, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, 2, 2,
                                  **kw)
    return P, rep


def _check_columns(d, P):
    """The sampled packed columns: same nonzero rows, values within 1e-8."""
    cols, nnz = d["sampled_cols"], d["col_nnz"]
    rows, vals = d["col_rows"], d["col_vals"]
    at, worst = 0, 0.0
    for j, k in zip(cols, nnz):
        j = int(j)
        col = P[j * (j + 1) // 2: j * (j + 1) // 2 + j + 1]
        nz = np.flatnonzero(col)
        assert np.array_equal(nz, rows[at:at + k]), f"zero pattern of column {j}"
        worst = max(worst, float(np.abs(col[nz] - vals[at:at + k]).max()))
        at += k
    return worst


@pytest.mark.parametrize("name", CONFIGS)
def test_fp64_matches_reference(engine, name):
    d = _fixture(name)
    mol, H, S = _system(engine, name)
    n = mol.ndim
    P, rep = _build(engine, mol, H, S)
    trace = d["mu_trace"]
    m = engine.sm_spectra(rep["nsm"], rep["sum_m"])[0]
    print(f"nsm {rep['nsm']} m {m.min()}..{m.max()} mu {rep['mu']} ({trace[-1, 1]}) "
          f"muiter {rep['muiter']} ({len(trace)})")
    assert rep["nsm"] == int(d["nsm"])
    assert np.array_equal(m, d["group_m"])
    assert rep["muiter"] == len(trace)
    assert rep["mu"] == trace[-1, 1]
    assert rep["nel"] == pytest.approx(trace[-1, 2], abs=1e-9)
    err = _check_columns(d, P)
    eb = engine.band_structure(n, P, H)
    ne = engine.get_nel(n, P, S)
    print(f"{len(d['sampled_cols'])} sampled columns: max|P - P_ref| {err:.3e}; "
          f"dE_band {eb - float(d['eband']):.3e}, dN_el {ne - float(d['nelp']):.3e}")
    assert err <= 1e-8
    assert abs(eb - float(d["eband"])) <= 1e-7
    assert abs(ne - float(d["nelp"])) <= 1e-10
    _cache[name + "_P"] = P


@pytest.mark.parametrize("name", CONFIGS)
def test_mixed_matches_fp64(engine, name):
    mol, H, S = _system(engine, name)
    n = mol.ndim
    P64 = _cache.get(name + "_P")
    if P64 is None:
        P64, _ = _build(engine, mol, H, S)
    Pm, rm = _build(engine, mol, H, S, precision="mixed")
    dP = float(np.abs(Pm - P64).max())
    dE = abs(engine.band_structure(n, Pm, H) - engine.band_structure(n, P64, H))
    print(f"mixed: max|P - P_fp64| {dP:.3e}, |dE_band| {dE:.3e}, mu {rm['mu']}")
    assert dP <= 1e-6 and dE <= 1e-6
    assert np.array_equal(Pm == 0.0, P64 == 0.0)


def test_h2o64_full_density_matches_oracle(engine):
    """Every element of the 64-water density (ndim 1472, 192 submatrices up to
    m = 811) against the oracle, not a sample."""
    from oracle import ptb_oracle as O
    mol, H, S = _system(engine, "c0_h2o64")
    P, rep = _build(engine, mol, H, S)
    res = O.sm_stupid_simple(mol.ndim, mol.nel, H, S, mol.n, mol.xyz, mol.aoat, 2, 2)
    err = float(np.abs(P - res.P).max())
    print(f"max|P - P_oracle| {err:.3e} over {P.size} packed entries")
    assert rep["mu"] == res.mu and rep["muiter"] == res.muiter
    assert err <= 1e-8
    assert np.array_equal(P == 0.0, res.P == 0.0)


@pytest.mark.parametrize("name", ["c2_c400h802", "c1_h2o512"])
def test_precision_ladder(engine, name):
    """The mixed build's accuracy per matrix-core precision of its
    divide-and-conquer merge products, against the FP64 build: split-BF16x3
    merge products with one FP64 refinement sweep meet max |dP| <= 1e-6 and
    |dE_band| <= 1e-6 Eh.  Plain
    BF16 products (8-bit operands) in the divide-and-conquer merges leave the
    eigenvectors too far from orthogonal for a first-order refinement to
    recover: reported with one and three sweeps, and required to be far worse
    than BF16x3 (which is why the mixed build uses BF16x3)."""
    mol, H, S = _system(engine, name)
    n = mol.ndim
    P64 = _cache.get(name + "_P")
    if P64 is None:
        P64, _ = _build(engine, mol, H, S)
    e64 = engine.band_structure(n, P64, H)
    rows = {}
    # every submatrix on float storage (mixed_min_m 0) and its merge products
    # in the low-precision arithmetic (lowp_merge_max_m 1024): the precision
    # of the low-precision path itself, not of the size policy
    saved = engine.get_tuning("mixed_min_m")
    saved_merge = engine.get_tuning("lowp_merge_max_m")
    engine.set_tuning("mixed_min_m", 0)
    engine.set_tuning("lowp_merge_max_m", 1024)
    try:
        for label, prec, sweeps in (("bf16x3", "mixed", 1), ("bf16", "mixed_bf16", 1),
                                    ("bf16+3", "mixed_bf16", 3), ("none", "mixed", -1)):
            P, rep = _build(engine, mol, H, S, precision=prec, refine_steps=sweeps)
            rows[label] = (float(np.abs(P - P64).max()), abs(engine.band_structure(n, P, H) - e64))
            print(f"{name} {label:7s} sweeps {sweeps:2d}: max|dP| {rows[label][0]:.3e} "
                  f"|dE_band| {rows[label][1]:.3e} mu {rep['mu']}")
    finally:
        engine.set_tuning("mixed_min_m", saved)
        engine.set_tuning("lowp_merge_max_m", saved_merge)
    assert rows["bf16x3"][0] <= 1e-6 and rows["bf16x3"][1] <= 1e-6
    assert rows["bf16"][0] > 100 * rows["bf16x3"][0]
    assert rows["none"][0] > rows["bf16x3"][0]    # the refinement does the work


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chunked_passes_are_bit_identical(engine, precision):
    """When the submatrix arenas do not fit the device (the 2048-water box),
    the build runs in chunks and passes: eigenvalues and fe of every chunk,
    the bisection over all, (for the mixed solve) the refinement pass and a
    second bisection, then the density chunk by chunk, each chunk recomputed
    in every pass.  Forced here on the 64-water box with a small arena budget
    (several chunks), the density, mu and pass count are bit-identical to the
    one-chunk build."""
    import os
    mol, H, S = _system(engine, "c0_h2o64")
    P1, r1 = _build(engine, mol, H, S, precision=precision)
    os.environ["PTBGPU_ARENA_BUDGET_MB"] = "700"
    try:
        Pc, rc = _build(engine, mol, H, S, precision=precision)
    finally:
        del os.environ["PTBGPU_ARENA_BUDGET_MB"]
    print(f"{precision}: mu {rc['mu']} / {r1['mu']}, muiter {rc['muiter']} / {r1['muiter']}, "
          f"max|dP| {np.abs(Pc - P1).max():.3e}")
    assert rc["mu"] == r1["mu"] and rc["muiter"] == r1["muiter"]
    assert np.array_equal(Pc, P1)


def test_h2o2048_full_size(engine):
    """The 2048-water box (6144 atoms, ndim 47104, filter 1e-5): its submatrix
    arenas exceed the GPU's memory, so the build runs in chunks and passes.
    At full size: every atom gives a submatrix, the FP64 and the mixed build
    converge to the same mu, the mixed density is within 1e-6 of the FP64 one
    (max per element) and 1e-6 Eh in band energy, and tr(P S) is the electron
    count to the accuracy of the submatrix method (the reference's own
    density misses it by 0.7 to 1.1 % on the smaller workloads)."""
    import torch
    from ptb_amd import geometry, surrogate
    mol = geometry.config_molecule("h2o_2048")
    H, S = surrogate.build_hs(mol)
    n = mol.ndim
    dH, dS = torch.from_numpy(H).cuda(), torch.from_numpy(S).cuda()
    del H, S
    torch.cuda.synchronize()
    engine.filter_matrix_sym_dev(dS, n, 1e-5)
    engine.filter_matrix_sym_dev(dH, n, 1e-5)
    import time
    print(f"\nh2o_2048: H, S filtered on the GPU (ndim {n})", flush=True)
    dP64 = torch.zeros_like(dH)
    t0 = time.perf_counter()
    r64 = engine.sm_stupid_simple_dev(n, mol.nel, dH, dS, dP64, mol.n, mol.xyz, mol.aoat, 2, 2)
    t64 = time.perf_counter() - t0
    print(f"h2o_2048: FP64 build {t64:.2f} s, nsm {r64['nsm']}, max m {r64['max_m']}", flush=True)
    dPm = torch.zeros_like(dH)
    t0 = time.perf_counter()
    rm = engine.sm_stupid_simple_dev(n, mol.nel, dH, dS, dPm, mol.n, mol.xyz, mol.aoat, 2, 2,
                                     precision="mixed")
    tm = time.perf_counter() - t0
    torch.cuda.synchronize()
    dp = float((dP64 - dPm).abs().max().item())
    e64 = engine.band_structure_dev(n, dP64, dH)
    em = engine.band_structure_dev(n, dPm, dH)
    nel = engine.get_nel_dev(n, dP64, dS)
    print(f"\nh2o_2048: nsm {r64['nsm']} max m {r64['max_m']} mu {r64['mu']} / {rm['mu']}, "
          f"muiter {r64['muiter']}; fp64 {t64:.2f} s, mixed {tm:.2f} s; "
          f"max|dP| {dp:.3e} |dE| {abs(em - e64):.3e}; tr(PS) {nel:.6f} of {mol.nel}")
    assert r64["nsm"] == mol.n
    assert rm["mu"] == r64["mu"] and rm["muiter"] == r64["muiter"]
    assert dp <= 1e-6 and abs(em - e64) <= 1e-6
    # the submatrix method's own electron-count error: the compiled
    # reference's P gives tr(P S) = 508.36 of 512 (64 waters), 4064.57 of
    # 4096 (512 waters) and 2428.27 of 2402 (C400H802), 0.7 to 1.1 %
    assert abs(nel - mol.nel) <= 1.5e-2 * mol.nel
