"""oracle: CPU restatement of the reference's density-build hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it, and
never as the thing measured or shipped.  The engine (ptb_amd, libptbgpu.so)
does not import, link or call anything here.

Modules
-------
ptb_oracle   numpy/scipy restatement of source/sm.f90, the filter/analyse
             routines of source/pgtb.f90, source/metrics.f90, source/norms.f90,
             source/purification.f90, source/purification_settings.f90,
             source/fermi.f and the solve2 dense reference (source/pgtb.f90).
ref          ctypes binding of oracle/_ref/libptbref.so, the reference's own
             Fortran hot-path units compiled here by oracle/Makefile (built only
             where the reference sources are present; used to pin ptb_oracle).

Parity status per function is listed in the header of ptb_oracle.
"""
