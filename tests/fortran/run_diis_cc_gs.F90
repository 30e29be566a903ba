!
!  eT's ground-state loop (run_diis_cc_gs, src/solvers/cc/diis_cc_gs_class.F90:209-315)
!  written against the ISO_C_BINDING module et_amd/fortran/etgpu_interface.F90, one
!  library call per eT procedure, every vector kept on the device.
!
!  Usage: run_diis_cc_gs DIR
!    DIR/dims  (stream): n_o, n_v, n_J (int64), E_HF, residual threshold (float64)
!    DIR/L     (stream): L_mo(n_J, n_mo, n_mo), DIR/h: h_mo(n_mo, n_mo), DIR/eps: eps(n_mo)
!  Prints one line per iteration ("iteration k E |Omega| |dE|") and "E_CCSD E".
!
program run_diis_cc_gs
!
   use iso_c_binding, only: c_ptr, c_int, c_int64_t, c_double
   use etgpu_interface
!
   implicit none
!
   type(c_ptr) :: wf
   integer(c_int64_t) :: n_o, n_v, n_J, n_mo, n_t1, n_t2, iteration
   real(c_double) :: e_hf, threshold, e_corr, energy, previous, omega_norm
   real(c_double), dimension(:), allocatable :: L, h, eps, t, fock
   character(len=4096) :: dir
   integer :: unit
   logical :: converged
!
   call get_command_argument(1, dir)
   open(newunit=unit, file=trim(dir)//'/dims', access='stream', form='unformatted', status='old')
   read(unit) n_o, n_v, n_J, e_hf, threshold
   close(unit)
!
   n_mo = n_o + n_v
   n_t1 = n_o*n_v
   n_t2 = n_t1*(n_t1 + 1)/2
   allocate(L(n_J*n_mo*n_mo), h(n_mo*n_mo), eps(n_mo), t(n_t1 + n_t2), fock(n_mo*n_mo))
   call read_doubles(trim(dir)//'/L', L)
   call read_doubles(trim(dir)//'/h', h)
   call read_doubles(trim(dir)//'/eps', eps)
!
!  initialize_ccsd, the Cholesky hand-over and the one-electron integrals
   call check(etg_create(n_o, n_v, n_J, wf))
   call check(etg_set_cholesky(wf, L))
   call check(etg_set_oei(wf, h))
   call check(etg_set_orbital_energies(wf, eps))
   deallocate(L)
!
!  set_initial_amplitudes_guess: t1 = 0, L_t1 = L_mo, t2 from the MP2 guess
   t = 0.0_c_double
   call check(etg_set_amplitudes(wf, t))
   call check(etg_construct_t1_cholesky(wf))
   call check(etg_set_t2_to_mp2_guess(wf))
!
   call check(etg_diis_initialize(wf, 8_c_int64_t))
   previous = 0.0_c_double
   converged = .false.
   iteration = 0
   do while (.not. converged .and. iteration < 100)
      iteration = iteration + 1
      call check(etg_construct_fock(wf, fock))
      call check(etg_calculate_energy(wf, e_corr))
      energy = e_hf + e_corr
      call check(etg_construct_omega_resident(wf, etg_fp64))
      call check(etg_omega_norm(wf, omega_norm))
      write(*, '(a, i4, f26.15, 2es14.6)') 'iteration', iteration, energy, omega_norm, abs(energy - previous)
      converged = omega_norm < threshold
      if (.not. converged) then
         call check(etg_quasi_newton_update(wf))
         call check(etg_diis_update(wf))
         call check(etg_construct_t1_cholesky(wf))
      end if
      previous = energy
   end do
!
   call check(etg_get_amplitudes(wf, t))
   call check(etg_destroy(wf))
   if (.not. converged) then
      write(*, '(a)') 'not converged'
      stop 2
   end if
   write(*, '(a, f26.15)') 'E_CCSD', energy
!
contains
!
   subroutine check(status)
      integer(c_int), intent(in) :: status
      if (status /= 0) then
         write(*, '(a, a)') 'etgpu error: ', etg_error_message()
         stop 1
      end if
   end subroutine check
!
   subroutine read_doubles(path, x)
      character(len=*), intent(in) :: path
      real(c_double), dimension(:), intent(out) :: x
      integer :: u
      open(newunit=u, file=path, access='stream', form='unformatted', status='old')
      read(u) x
      close(u)
   end subroutine read_doubles
!
end program run_diis_cc_gs
