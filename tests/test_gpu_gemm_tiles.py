"""Knobs that change how work is laid out but not its arithmetic.

The FP64 product kernel at both tile edges (64 x 64 and 128 x 128 output
tiles per workgroup).  Both add every output element's k terms in the same
order, 4 at a time ascending, so forcing either edge (tuning knob
"gemm_tile") must give bit-identical products, density matrices, dense
solves and purifications; the products also match torch's FP64 ones to the
FP64 bound.  The automatic choice (knob 0) is checked to pick 128 x 128
tiles for the large batched products and 64 x 64 for small ones."""
import contextlib

import numpy as np
import pytest

from conftest import small_system

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def _tile(engine, edge):
    saved = engine.get_tuning("gemm_tile")
    engine.set_tuning("gemm_tile", edge)
    try:
        yield
    finally:
        engine.set_tuning("gemm_tile", saved)


def _launches(engine, fn):
    """fn() under profiling: launch groups of 64- and 128-edge FP64 products."""
    engine.profile(True)
    try:
        fn()
    finally:
        counts = (engine.profile_query("gemm_tile64")[1], engine.profile_query("gemm_tile128")[1])
        engine.profile(False)
    return counts


def _mat(rows, cols, gen):
    import torch
    x = gen.standard_normal((cols, rows))   # torch (cols, rows) = rows x cols column-major
    x[:, ::5] *= 1e3
    return torch.from_numpy(x).cuda()


# ragged shapes around both tile edges and both k-slice depths (32 and 16),
# one product much larger than the rest
SHAPES = [(1000, 700, 300), (129, 257, 1000), (128, 128, 16), (127, 129, 33), (64, 64, 32),
          (7, 200, 513), (300, 33, 1), (1, 1, 1), (513, 130, 17)]


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_edges_identical_and_fp64_accurate(engine, ta, tb):
    import torch
    gen = np.random.default_rng(21)
    As, Bs, C0, refs = [], [], [], []
    alpha, beta = -1.25, 0.5
    for (m, n, k) in SHAPES:
        A = _mat(k, m, gen) if ta else _mat(m, k, gen)
        B = _mat(n, k, gen) if tb else _mat(k, n, gen)
        C = _mat(m, n, gen)
        opA = A if ta else A.t()     # A.t() is the stored column-major matrix
        opB = B if tb else B.t()
        refs.append((alpha * (opA @ opB) + beta * C.t(),
                     abs(alpha) * (opA.abs() @ opB.abs()) + abs(beta) * C.t().abs()))
        As.append(A)
        Bs.append(B)
        C0.append(C)
    out = {}
    for edge in (64, 128):
        Cs = [c.clone() for c in C0]
        with _tile(engine, edge):
            n64, n128 = _launches(engine, lambda: engine.gemm_batch_dev("fp64", ta, tb, As, Bs, Cs,
                                                                        alpha, beta))
        assert (n64, n128) == ((1, 0) if edge == 64 else (0, 1))
        torch.cuda.synchronize()
        out[edge] = Cs
    worst = 0.0
    for c64, c128, (ref, scale) in zip(out[64], out[128], refs):
        assert torch.equal(c64, c128)
        worst = max(worst, ((c128.t() - ref).abs() / scale.clamp_min(1e-300)).max().item())
    print(f"\nta={ta} tb={tb}: identical at both edges, max scaled error {worst:.2e}")
    assert worst <= 1e-13


def test_single_large_product_over_the_chip(engine):
    """One 2500 x 2300 x 1900 product (fewer products than XCDs: tiles dealt
    in runs over the whole chip) at both edges."""
    import torch
    gen = np.random.default_rng(5)
    A, B = _mat(2500, 1900, gen), _mat(1900, 2300, gen)
    ref = A.t() @ B.t()
    out = {}
    for edge in (64, 128):
        C = torch.zeros(2300, 2500, dtype=torch.float64, device="cuda")
        with _tile(engine, edge):
            engine.gemm_batch_dev("fp64", False, False, [A], [B], [C])
        torch.cuda.synchronize()
        out[edge] = C
    assert torch.equal(out[64], out[128])
    err = ((out[128].t() - ref).abs().max() / ref.abs().max()).item()
    print(f"\nrelative error {err:.2e}")
    assert err < 1e-14


def test_automatic_edge(engine):
    """Knob 0: one 4096 x 4096 product (1024 big tiles, no extra padding)
    takes 128 x 128 tiles; one of 4000 x 4000 (3 % more padded work at 128)
    and batches (8 or more products, however large) stay at 64 x 64."""
    import torch
    gen = np.random.default_rng(9)
    assert engine.get_tuning("gemm_tile") == 0

    def run(n, count):
        As = [_mat(n, n, gen) for _ in range(count)]
        Cs = [torch.zeros_like(a) for a in As]
        return _launches(engine, lambda: engine.gemm_batch_dev("fp64", False, False, As, As, Cs))

    assert run(4096, 1) == (0, 1)
    assert run(4000, 1) == (1, 0)
    assert run(1024, 8) == (1, 0)
    assert run(300, 8) == (1, 0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_density_identical_at_both_edges(engine, precision):
    """The submatrix build of 64 waters (Cholesky, L^-1, K, Q application,
    refinement, projections): forcing 64 x 64 or 128 x 128 tiles on every FP64
    product gives the same P bit for bit."""
    mol, H, S = small_system("water", (4, 4, 4), 3)
    n = mol.ndim
    Ps = {}
    for edge in (64, 128):
        P = np.zeros_like(H)
        with _tile(engine, edge):
            n64, n128 = _launches(engine, lambda: engine.sm_stupid_simple(
                n, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, 2, 2, precision=precision))
        print(f"\n{precision} edge {edge}: launch groups 64: {n64}, 128: {n128}")
        assert n128 > 0 if edge == 128 else n128 == 0
        Ps[edge] = P
    assert np.array_equal(Ps[64], Ps[128])


def test_dense_solve_and_purification_identical_at_both_edges(engine):
    """solve2 (ndim 1104) and a Pade sign-iteration purification with its
    chemical-potential search (ndim 414) at both edges: identical results."""
    from ptb_amd.engine import PurSettings
    mol, H, S = small_system("water", (4, 4, 3), 1, filt=1e-5)
    res = {}
    for edge in (64, 128):
        with _tile(engine, edge):
            res[edge] = engine.solve2(1, mol.ndim, mol.nel, 0, 300.0, H, S)
    for a, b in zip(res[64][:4], res[128][:4]):
        assert np.array_equal(a, b)
    mol, H, S = small_system("water", (3, 3, 2), 0)
    pur = PurSettings(type=2, metric=2, cycles=40, metric_cycles=50, chempot_cycles=60,
                      chempot_guess=-1.5, chempot_increment=0.1, nel=mol.nel)
    Ps = {}
    for edge in (64, 128):
        P = np.zeros_like(H)
        with _tile(engine, edge):
            engine.purification_wrapper(pur, mol.ndim, H.copy(), S.copy(), P)
        Ps[edge] = P
    assert np.array_equal(Ps[64], Ps[128])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chip_wide_merges_identical(engine, precision):
    """Tuning knob dc_global_min = 64: every divide-and-conquer merge class of
    64 rows and up runs its roots, Loewner vectors and eigenvectors over the
    whole chip from device scratch (the path of nodes too large for LDS)
    instead of one workgroup per node in LDS.  Same device functions, same
    order: the density is the same bit for bit."""
    mol, H, S = small_system("water", (4, 4, 4), 3)
    n = mol.ndim
    Ps = {}
    saved = engine.get_tuning("dc_global_min")
    try:
        for v in (0, 64):
            engine.set_tuning("dc_global_min", v)
            P = np.zeros_like(H)
            engine.sm_stupid_simple(n, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, 2, 2,
                                    precision=precision)
            Ps[v] = P
    finally:
        engine.set_tuning("dc_global_min", saved)
    assert np.array_equal(Ps[0], Ps[64])
