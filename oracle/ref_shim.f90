! C-callable wrappers around the reference's own hot-path routines, compiled
! together with them into oracle/_ref/libptbref.so (see oracle/Makefile).
! TEST INFRASTRUCTURE ONLY: used to pin the numpy oracle; never shipped.
!
! Every wrapper takes explicit-shape arrays and 8-byte integers (the reference
! is built with -i8) and forwards to the module procedure it names.
module ptbref_shim
   use iso_c_binding
   use iso_fortran_env, only: output_unit
   implicit none
contains

   ! sm_stupid_simple (source/sm.f90:34-409)
   subroutine ref_sm_stupid_simple(ndim, nelref, H, S, P, n, xyz, aoat, cols, mode) &
         bind(C, name="ref_sm_stupid_simple")
      use sm, only: sm_stupid_simple
      integer(c_int64_t), value :: ndim, nelref, n, cols, mode
      real(c_double), intent(in) :: H(ndim*(ndim+1)/2), S(ndim*(ndim+1)/2), xyz(3, n)
      real(c_double), intent(inout) :: P(ndim*(ndim+1)/2)
      integer(c_int64_t), intent(in) :: aoat(ndim)
      call sm_stupid_simple(ndim, nelref, H, S, P, n, xyz, aoat, cols, mode)
      flush (output_unit)
   end subroutine

   ! clustering_kmeans (source/sm.f90:459-514) for one nsm: groups and effort
   ! The RANDOM_NUMBER stream KMPP consumes after initseed
   ! (source/sm.f90:20-32 reseeds before every k-means candidate).
   subroutine ref_random_stream(count, out) bind(C, name="ref_random_stream")
      use sm, only: initseed
      integer(c_int64_t), value :: count
      real(c_double), intent(out) :: out(count)
      integer(c_int64_t) :: i
      call initseed()
      do i = 1, count
         call random_number(out(i))
      end do
   end subroutine

   subroutine ref_clustering_kmeans(n, ndim, xyz, Hp, Sp, aoat, nsm, ncomb, comb, effort) &
         bind(C, name="ref_clustering_kmeans")
      use sm, only: clustering_kmeans
      integer(c_int64_t), value :: n, ndim, nsm
      real(c_double), intent(in) :: xyz(3, n), Hp(ndim*(ndim+1)/2), Sp(ndim*(ndim+1)/2)
      integer(c_int64_t), intent(in) :: aoat(ndim)
      integer(c_int64_t), intent(out) :: ncomb(n), comb(n, ndim)
      real(c_double), intent(out) :: effort
      real(c_double), allocatable :: H(:, :), S(:, :), kc(:, :), kw(:)
      integer(c_int64_t), allocatable :: kz(:)
      allocate (H(ndim, ndim), S(ndim, ndim), kc(3, n), kw(n), kz(n))
      call blowsym(ndim, Hp, H)
      call blowsym(ndim, Sp, S)
      ncomb = 0
      comb = 0
      call clustering_kmeans(n, ndim, xyz, kc, kz, kw, H, S, aoat, nsm, ncomb, comb, effort)
   end subroutine

   ! get_nel_packed / get_band_structure_E_packed (source/metrics.f90:202-277)
   function ref_get_nel(ndim, P, S) result(r) bind(C, name="ref_get_nel")
      use metrics, only: get_nel
      integer(c_int64_t), value :: ndim
      real(c_double), intent(in) :: P(ndim*(ndim+1)/2), S(ndim*(ndim+1)/2)
      real(c_double) :: r
      r = get_nel(ndim, P, S)
   end function

   function ref_band_structure(ndim, P, H) result(r) bind(C, name="ref_band_structure")
      use metrics, only: band_structure
      integer(c_int64_t), value :: ndim
      real(c_double), intent(in) :: P(ndim*(ndim+1)/2), H(ndim*(ndim+1)/2)
      real(c_double) :: r
      r = band_structure(ndim, P, H)
   end function

   ! adjust_thresholds_dp (source/metrics.f90:82-113) for ndim: thrs is set
   ! only when unallocated, so it is released first (each call takes its own
   ! size's tier)
   subroutine ref_thresholds(ndim, vals) bind(C, name="ref_thresholds")
      use metrics, only: adjust_thresholds, thrs
      integer(c_int64_t), value :: ndim
      real(c_double), intent(out) :: vals(3)
      real(c_double), allocatable :: dummy(:)
      allocate (dummy(ndim*(ndim+1)/2))
      dummy = 0.0_c_double
      if (allocated(thrs)) deallocate (thrs)
      call adjust_thresholds(ndim, dummy)
      vals(1) = thrs%high
      vals(2) = thrs%normal
      vals(3) = thrs%low
      deallocate (dummy)
   end subroutine

   ! check_density_packed (source/metrics.f90:144-179); thresholds set by
   ! adjust_thresholds_dp (:82-113) for this ndim. Prints the reference lines.
   subroutine ref_check_density(ndim, P, S, nel) bind(C, name="ref_check_density")
      use metrics, only: check_density, adjust_thresholds, thrs
      integer(c_int64_t), value :: ndim, nel
      real(c_double), intent(in) :: P(ndim*(ndim+1)/2), S(ndim*(ndim+1)/2)
      if (allocated(thrs)) deallocate (thrs)
      call adjust_thresholds(ndim, P)
      call check_density(ndim, P, S, nel)
      flush (output_unit)
   end subroutine

   ! purification_wrapper (source/purification.f90:28-80) with the settings of
   ! tPurificationSet (source/purification_settings.f90:55-86)
   subroutine ref_purification(ndim, nel, ptype, metric, iterative, cycles, metric_cycles, &
         chempot_cycles, chempot_guess, chempot_increment, prlvl, H, S, P) &
         bind(C, name="ref_purification")
      use purification_, only: purification
      use purification_settings, only: tPurificationSet
      use metrics, only: adjust_thresholds, thrs
      integer(c_int64_t), value :: ndim, nel, ptype, metric, iterative, cycles, &
         metric_cycles, chempot_cycles, prlvl
      real(c_double), value :: chempot_guess, chempot_increment
      real(c_double), intent(in) :: H(ndim*(ndim+1)/2), S(ndim*(ndim+1)/2)
      real(c_double), intent(out) :: P(ndim*(ndim+1)/2)
      type(tPurificationSet) :: pur
      pur%nel = nel
      pur%type = ptype
      pur%metric%type = metric
      pur%metric%iterative = iterative /= 0
      pur%cycles = cycles
      pur%metric%cycles = metric_cycles
      pur%chempot%cycles = chempot_cycles
      pur%chempot%guess = chempot_guess
      pur%chempot%increment = chempot_increment
      pur%prlvl = prlvl
      if (allocated(thrs)) deallocate (thrs)
      call adjust_thresholds(ndim, H)
      call purification(pur, ndim, H, S, P)
      flush (output_unit)
   end subroutine

   ! FERMISMEAR (source/fermi.f:5-67) and occu (source/fermi.f:69-124)
   subroutine ref_fermismear(norbs, nel, t, eig, occ, fod, efermi, s) &
         bind(C, name="ref_fermismear")
      integer(c_int64_t), value :: norbs, nel
      real(c_double), value :: t
      real(c_double), intent(in) :: eig(norbs)
      real(c_double), intent(out) :: occ(norbs), fod, efermi, s
      call fermismear(.false., norbs, nel, t, eig, occ, fod, efermi, s)
   end subroutine

   subroutine ref_occu(ndim, nel, nopen, ihomoa, ihomob, focca, foccb) bind(C, name="ref_occu")
      integer(c_int64_t), value :: ndim, nel
      integer(c_int64_t), intent(inout) :: nopen
      integer(c_int64_t), intent(out) :: ihomoa, ihomob
      real(c_double), intent(out) :: focca(ndim), foccb(ndim)
      call occu(ndim, nel, nopen, ihomoa, ihomob, focca, foccb)
   end subroutine

   ! dmat (source/solve.f:263-283) and the LAPACK call convention of
   ! la_sygvd_rdp (source/lapack_eig.f90:318-386), as used by solve2
   subroutine ref_dmat(ndim, focc, C, P) bind(C, name="ref_dmat")
      integer(c_int64_t), value :: ndim
      real(c_double), intent(in) :: focc(ndim), C(ndim, ndim)
      real(c_double), intent(out) :: P(ndim, ndim)
      call dmat(ndim, focc, C, P)
   end subroutine

   subroutine ref_sygvd(ndim, A, B, w, info) bind(C, name="ref_sygvd")
      use gtb_lapack_eig, only: la_sygvd
      integer(c_int64_t), value :: ndim
      real(c_double), intent(inout) :: A(ndim, ndim), B(ndim, ndim)
      real(c_double), intent(out) :: w(ndim)
      integer(c_int64_t), intent(out) :: info
      call la_sygvd(A, B, w, info)
   end subroutine

   ! mlpop14 (source/mlpop.f:142-190): S^3/4 and S^1/4 in REAL(4) from packed S
   subroutine ref_mlpop14(nao, S, s34, s14) bind(C, name="ref_mlpop14")
      integer(c_int64_t), value :: nao
      real(c_double), intent(in) :: S(nao*(nao+1)/2)
      real(c_float), intent(out) :: s34(nao, nao), s14(nao, nao)
      call mlpop14(nao, S, s34, s14)
   end subroutine

   ! mlpop2 (source/mlpop.f:198-238): atomic and shell populations from the
   ! packed P and the REAL(4) S^3/4 (s1), S^1/4 (s2); aoat and shell2ao are the
   ! module arrays of bascom the routine reads
   subroutine ref_mlpop2(n, nao, P, s1, s2, aoat_in, shell2ao_in, q, qsh) &
         bind(C, name="ref_mlpop2")
      use bascom, only: aoat, shell2ao
      integer(c_int64_t), value :: n, nao
      real(c_double), intent(in) :: P(nao*(nao+1)/2)
      real(c_float), intent(in) :: s1(nao, nao), s2(nao, nao)
      integer(c_int64_t), intent(in) :: aoat_in(nao), shell2ao_in(nao)
      real(c_double), intent(out) :: q(n), qsh(10, n)
      if (allocated(aoat)) deallocate (aoat)
      if (allocated(shell2ao)) deallocate (shell2ao)
      allocate (aoat(nao), shell2ao(nao))
      aoat = aoat_in
      shell2ao = shell2ao_in
      call mlpop2(n, nao, P, s1, s2, q, qsh)
   end subroutine

end module ptbref_shim
