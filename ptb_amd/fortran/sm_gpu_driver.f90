! A host program that uses the drop-in `module sm` exactly as twoscf does
! (source/pgtb.f90:566-618): filter S and H in place, call sm_stupid_simple,
! then evaluate N_el and Tr(HP).  Inputs come from a raw binary file written by
! the test-suite: ndim, nel, n, columns, mode (int64), filter (double), xyz(3,n),
! aoat(ndim) (int64), then packed H and S (double).  Packed P and the two
! traces are written to the output file (stream access) and, as two Fortran
! sequential records, to <output>.seq.
program sm_gpu_driver
   use iso_c_binding
   use sm, only: sm_stupid_simple
   use ptbgpu_fortran, only: gpu_filter_matrix_sym, gpu_get_nel, gpu_band_structure
   implicit none
   integer(c_int64_t) :: ndim, nel, n, cols, mode, npk
   real(c_double) :: filter, nelp, eband
   real(c_double), allocatable :: xyz(:, :), H(:), S(:), P(:)
   integer(c_int64_t), allocatable :: aoat(:)
   character(len=1024) :: fin, fout
   integer :: u

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=u, file=trim(fin), access="stream", form="unformatted", status="old")
   read (u) ndim, nel, n, cols, mode
   read (u) filter
   allocate (xyz(3, n), aoat(ndim))
   read (u) xyz
   read (u) aoat
   npk = ndim*(ndim + 1)/2
   allocate (H(npk), S(npk), P(npk))
   read (u) H
   read (u) S
   close (u)

   if (filter > 0) then
      call gpu_filter_matrix_sym(S, ndim, filter)
      call gpu_filter_matrix_sym(H, ndim, filter)
   end if
   P = 0
   call sm_stupid_simple(ndim, nel, H, S, P, n, xyz, aoat, cols, mode)
   nelp = gpu_get_nel(ndim, P, S)
   eband = gpu_band_structure(ndim, P, H)
   print *, "nel in P=", nelp, "required nel=", nel
   print *, "Tr(H P)=", eband

   open (newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace")
   write (u) P
   write (u) nelp, eband
   close (u)
   ! the same two records with the framing of PTB's ptb_dump (sequential
   ! unformatted, source/pgtb.f90:275-290), for the format readers' tests
   open (newunit=u, file=trim(fout)//".seq", form="unformatted", status="replace")
   write (u) P
   write (u) nelp, eband
   close (u)
end program sm_gpu_driver
