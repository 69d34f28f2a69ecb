"""ptb_amd: an MI355X-native density-matrix engine for PTB's hot path.

The engine replaces the reference's submatrix density build
(``sm_stupid_simple``, source/sm.f90:34-409), its in-place threshold filter
(``filter_matrix_sym``, source/pgtb.f90:1532-1548), the density checks of
source/metrics.f90 and the purification driver (source/purification.f90)
with hand-written HIP kernels for gfx950 behind a C ABI (include/ptbgpu.h).
``ptb_amd.engine`` mirrors the reference interface in Python over that ABI.
"""
