"""Algorithmic flops per GEMM site as the benchmark line prices them: the split
of each site into FP64 and split-BF16x3 flops against closed forms of the
products each build runs (no GPU)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

M = np.array([120, 250, 288, 300, 640, 999])
C = np.array([3, 5, 3, 9, 13, 3])
NO = np.array([60, 110, 140, 150, 320, 500])
THR = 288


@pytest.mark.parametrize("site", ["gemm_potrf", "gemm_linv", "gemm_kform", "gemm_sytrd",
                                  "gemm_dc", "gemm_u", "gemm_fe", "gemm_q", "gemm_refine"])
def test_fp64_build_has_no_bf16x3_flops(site):
    fp, bf = bench.gemm_site_flops(site, M, C, NO, False, THR, 1024)
    assert bf == 0.0 and fp >= 0.0


def test_unknown_site():
    assert bench.gemm_site_flops("gemm_other", M, C, NO, True, THR, 0) is None


def test_merges_fp64_unless_lowp_limit():
    full = float(np.sum(2.0 * M.astype(float) ** 3 / 3.0))
    # lowp_merge_max_m 0: every merge product in FP64
    assert bench.gemm_site_flops("gemm_dc", M, C, NO, True, THR, 0) == (full, 0.0)
    # limit 640: the refined submatrices (m >= 288) up to 640 rows in BF16x3
    fp, bf = bench.gemm_site_flops("gemm_dc", M, C, NO, True, THR, 640)
    low = [288, 300, 640]
    assert bf == pytest.approx(sum(2.0 * m ** 3 / 3.0 for m in low))
    assert fp + bf == pytest.approx(full)
    # an FP64 build never takes low-precision merges
    assert bench.gemm_site_flops("gemm_dc", M, C, NO, False, THR, 1024) == (full, 0.0)


def test_refinement_products():
    # refined submatrices only: residual K X (2 m^2 n) in FP64, the two
    # correction products Z^T R and Z E (2 m^2 n each) in BF16x3, and per
    # Loewdin step W^T W and W R (2 m n^2 each) in FP64
    ref = M >= THR
    m, n = M[ref].astype(float), NO[ref].astype(float)
    one = bench.gemm_site_flops("gemm_refine", M, C, NO, True, THR, 0)
    assert one[0] == pytest.approx(np.sum(2 * m * m * n + 4 * m * n * n))
    assert one[1] == pytest.approx(np.sum(4 * m * m * n))
    # low-precision merges on some refined submatrix: a second Loewdin step
    two = bench.gemm_site_flops("gemm_refine", M, C, NO, True, THR, 300)
    assert two[0] == pytest.approx(np.sum(2 * m * m * n + 8 * m * n * n))
    assert two[1] == one[1]
    # nothing refined below the threshold
    assert bench.gemm_site_flops("gemm_refine", M, C, NO, True, 10 ** 6, 0) == (0.0, 0.0)


def test_q_application_adds_the_refinement_passes():
    base = float(np.sum(4.0 * M * M * (2.0 * C + NO)))
    assert bench.gemm_site_flops("gemm_q", M, C, NO, False, THR, 0) == (pytest.approx(base), 0.0)
    ref = M >= THR
    extra = float(np.sum(8.0 * M[ref].astype(float) ** 2 * NO[ref]))
    fp, bf = bench.gemm_site_flops("gemm_q", M, C, NO, True, THR, 0)
    assert fp == pytest.approx(base + extra) and bf == 0.0
