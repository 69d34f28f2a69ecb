"""The matrix-core product family of the mixed-precision build
(ptbgpu_gemm_batch_dev): FP64 MFMA, split-BF16x3 and BF16 on the BF16 matrix
cores with FP32 accumulation, against an FP64 torch reference of the same
products.

Bars, per element of C = A B with the error scaled by (|A| |B|)_ij:
FP64 1e-13; BF16x3 2^-15 (each operand carries 16 significant bits, the
dropped lo*lo term is 2^-16, FP32 sums add k 2^-24); BF16 2^-7.  Ragged
shapes (not multiples of the 64x64 tile or the 32-deep k slice), transposed
operands, alpha/beta and a batch of different sizes are covered.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOUND = {"fp64": 1e-13, "bf16x3": 2.0 ** -15, "bf16": 2.0 ** -7}


def _mat(rows, cols, gen):
    import torch
    # torch (cols, rows) row-major == a rows x cols column-major matrix
    x = gen.standard_normal((cols, rows))
    x[:, ::7] *= 1e3           # mixed magnitudes within a column
    return torch.from_numpy(x).cuda()


@pytest.mark.parametrize("prec", ["fp64", "bf16x3", "bf16"])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_products_within_precision_bound(engine, prec, ta, tb):
    import torch
    gen = np.random.default_rng(11)
    shapes = [(64, 64, 32), (130, 77, 45), (7, 200, 513), (300, 33, 1), (1, 1, 1)]
    As, Bs, Cs, refs = [], [], [], []
    alpha, beta = 0.75, -0.5
    for (m, n, k) in shapes:
        A = _mat(k, m, gen) if ta else _mat(m, k, gen)
        B = _mat(n, k, gen) if tb else _mat(k, n, gen)
        C = _mat(m, n, gen)
        Am = A.t() if True else A                 # the stored column-major matrix
        Bm = B.t()
        opA = Am.t() if ta else Am
        opB = Bm.t() if tb else Bm
        ref = alpha * (opA @ opB) + beta * C.t()
        scale = abs(alpha) * (opA.abs() @ opB.abs()) + abs(beta) * C.t().abs()
        As.append(A)
        Bs.append(B)
        Cs.append(C)
        refs.append((ref, scale))
    engine.gemm_batch_dev(prec, ta, tb, As, Bs, Cs, alpha, beta)
    torch.cuda.synchronize()
    worst = 0.0
    for C, (ref, scale) in zip(Cs, refs):
        err = ((C.t() - ref).abs() / scale.clamp_min(1e-300)).max().item()
        worst = max(worst, err)
    print(f"\n{prec} ta={ta} tb={tb}: max scaled error {worst:.3e} (bound {BOUND[prec]:.1e})")
    assert worst <= BOUND[prec]
    if prec == "bf16x3":
        assert worst > 1e-12          # really the low-precision units, not FP64


def test_bf16x3_is_far_more_accurate_than_bf16(engine):
    import torch
    gen = np.random.default_rng(3)
    A, B = _mat(256, 256, gen), _mat(256, 256, gen)
    ref = A.t() @ B.t()
    errs = {}
    for prec in ("bf16x3", "bf16"):
        C = torch.zeros_like(A)
        engine.gemm_batch_dev(prec, False, False, [A], [B], [C])
        torch.cuda.synchronize()
        errs[prec] = ((C.t() - ref).abs().max() / ref.abs().max()).item()
    print(f"\nrelative max error: bf16x3 {errs['bf16x3']:.2e}, bf16 {errs['bf16']:.2e}")
    assert errs["bf16x3"] * 100 < errs["bf16"]
