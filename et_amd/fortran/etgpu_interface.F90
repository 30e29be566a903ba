!
!  ISO_C_BINDING interface to libetgpu.so (include/etgpu.h).
!
!  This is the binding a `ccsd_gpu` wavefunction (extends eT's `ccsd`) uses
!  to forward its type-bound procedures to the GPU: every routine is a
!  bind(C) function returning 0 on success; etg_error_message converts the
!  C error string for output%error_msg.  Sizes are integer(c_int64_t) (eT's
!  default 8-byte integers), arrays are passed by address in eT's layouts.
!
module etgpu_interface
!
   use iso_c_binding, only: c_int, c_int64_t, c_double, c_ptr, c_char, c_null_char, &
                            c_f_pointer, c_associated
!
   implicit none
!
   integer(c_int), parameter :: etg_fp64 = 0
   integer(c_int), parameter :: etg_fp32 = 1
   integer(c_int), parameter :: etg_bf16x3 = 2
   integer(c_int), parameter :: etg_bf16x3_any_shape = 3   ! etg_test_gemm / etg_bench_gemm only
!
   interface
!
      integer(c_int) function etg_create(n_o, n_v, n_J, handle) bind(C, name='etg_create')
         import :: c_int, c_int64_t, c_ptr
         integer(c_int64_t), value :: n_o, n_v, n_J
         type(c_ptr), intent(out) :: handle
      end function etg_create
!
      integer(c_int) function etg_destroy(handle) bind(C, name='etg_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_destroy
!
      type(c_ptr) function etg_last_error() bind(C, name='etg_last_error')
         import :: c_ptr
      end function etg_last_error
!
      integer(c_int) function etg_set_cholesky(handle, L) bind(C, name='etg_set_cholesky')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(in) :: L
      end function etg_set_cholesky
!
      integer(c_int) function etg_set_oei(handle, h) bind(C, name='etg_set_oei')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(in) :: h
      end function etg_set_oei
!
      integer(c_int) function etg_set_orbital_energies(handle, eps) &
                                                   bind(C, name='etg_set_orbital_energies')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(in) :: eps
      end function etg_set_orbital_energies
!
      integer(c_int) function etg_set_amplitudes(handle, t) bind(C, name='etg_set_amplitudes')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(in) :: t
      end function etg_set_amplitudes
!
      integer(c_int) function etg_get_amplitudes(handle, t) bind(C, name='etg_get_amplitudes')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(out) :: t
      end function etg_get_amplitudes
!
      integer(c_int) function etg_set_t2_to_mp2_guess(handle) bind(C, name='etg_set_t2_to_mp2_guess')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_set_t2_to_mp2_guess
!
      integer(c_int) function etg_construct_t1_cholesky(handle) &
                                                   bind(C, name='etg_construct_t1_cholesky')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_construct_t1_cholesky
!
      integer(c_int) function etg_construct_fock(handle, F) bind(C, name='etg_construct_fock')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(out) :: F
      end function etg_construct_fock
!
      integer(c_int) function etg_calculate_energy(handle, e_corr) bind(C, name='etg_calculate_energy')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), intent(out) :: e_corr
      end function etg_calculate_energy
!
      integer(c_int) function etg_construct_omega(handle, omega, precision) &
                                                   bind(C, name='etg_construct_omega')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(out) :: omega
         integer(c_int), value :: precision
      end function etg_construct_omega
!
      integer(c_int) function etg_omega_norm(handle, norm) bind(C, name='etg_omega_norm')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), intent(out) :: norm
      end function etg_omega_norm
!
      integer(c_int) function etg_get_orbital_differences(handle, x, n) &
                                                   bind(C, name='etg_get_orbital_differences')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(out) :: x
         integer(c_int64_t), value :: n
      end function etg_get_orbital_differences
!
      integer(c_int) function etg_prepare_for_jacobian(handle) bind(C, name='etg_prepare_for_jacobian')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_prepare_for_jacobian
!
      integer(c_int) function etg_jacobian_transformation(handle, c, rho) &
                                                   bind(C, name='etg_jacobian_transformation')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(in) :: c
         real(c_double), dimension(*), intent(out) :: rho
      end function etg_jacobian_transformation
!
      integer(c_int) function etg_jacobian_transpose_transformation(handle, b, sigma) &
                                                   bind(C, name='etg_jacobian_transpose_transformation')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(in) :: b
         real(c_double), dimension(*), intent(out) :: sigma
      end function etg_jacobian_transpose_transformation
!
!     diis%initialize (diis_tool_class.F90:249-277)
      integer(c_int) function etg_linear_davidson(handle, side, rhs, start_vector, residual_threshold, &
                                                  max_iterations, max_dim, precision, solution, &
                                                  residual_norms, iterations) bind(C, name='etg_linear_davidson')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int), value :: side, precision
         real(c_double), dimension(*), intent(in) :: rhs
         type(c_ptr), value :: start_vector, residual_norms
         real(c_double), value :: residual_threshold
         integer(c_int64_t), value :: max_iterations, max_dim
         real(c_double), dimension(*), intent(out) :: solution
         integer(c_int64_t), intent(out) :: iterations
      end function etg_linear_davidson
!
      integer(c_int) function etg_set_cvs(handle, n_core, core_mos) bind(C, name='etg_set_cvs')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int64_t), value :: n_core
         type(c_ptr), value :: core_mos
      end function etg_set_cvs
!
      integer(c_int) function etg_set_remove_core(handle, n_core, core_mos) bind(C, name='etg_set_remove_core')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int64_t), value :: n_core
         type(c_ptr), value :: core_mos
      end function etg_set_remove_core
!
      integer(c_int) function etg_set_ionization(handle, on) bind(C, name='etg_set_ionization')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: on
      end function etg_set_ionization
!
      integer(c_int) function etg_newton_raphson_update(handle, rel_threshold, max_iterations, precision, &
                                                        micro_norms, iterations) &
                                                        bind(C, name='etg_newton_raphson_update')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: handle
         real(c_double), value :: rel_threshold
         integer(c_int64_t), value :: max_iterations
         integer(c_int), value :: precision
         type(c_ptr), value :: micro_norms
         integer(c_int64_t), intent(out) :: iterations
      end function etg_newton_raphson_update
!
      integer(c_int) function etg_diis_excited_states(handle, side, n_states, residual_threshold, &
                                                      energy_threshold, max_iterations, diis_dimension, &
                                                      precision, start_vectors, energies, vectors, &
                                                      residual_norms, energy_history, iterations) &
                                                      bind(C, name='etg_diis_excited_states')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int), value :: side, precision
         integer(c_int64_t), value :: n_states, max_iterations, diis_dimension
         real(c_double), value :: residual_threshold, energy_threshold
         type(c_ptr), value :: start_vectors, vectors, residual_norms, energy_history
         real(c_double), dimension(*), intent(out) :: energies
         integer(c_int64_t), intent(out) :: iterations
      end function etg_diis_excited_states
!
!     diis%finalize (diis_tool_class.F90:280-299): the records are released
      integer(c_int) function etg_diis_finalize(handle) bind(C, name='etg_diis_finalize')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_diis_finalize
!
      integer(c_int) function etg_diis_initialize(handle, dimension) bind(C, name='etg_diis_initialize')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int64_t), value :: dimension
      end function etg_diis_initialize
!
      integer(c_int) function etg_diis_initialize_crop(handle, dimension) bind(C, name='etg_diis_initialize_crop')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int64_t), value :: dimension
      end function etg_diis_initialize_crop
!
!     t_updater%update (quasi_newton_updater_class.F90:81-118), on the device
      integer(c_int) function etg_quasi_newton_update(handle) bind(C, name='etg_quasi_newton_update')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_quasi_newton_update
!
!     diis%update(omega, amplitudes) (diis_tool_class.F90:302-347), on the device
      integer(c_int) function etg_diis_update(handle) bind(C, name='etg_diis_update')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_diis_update
!
!     prepare_for_multiplier_equation (multiplier_equation_ccsd.F90:41-53)
      integer(c_int) function etg_prepare_for_multiplier_equation(handle) bind(C, name='etg_prepare_for_multiplier_equation')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_prepare_for_multiplier_equation
!
!     set_initial_multipliers_guess (ccsd_class.F90:423-475)
      integer(c_int) function etg_set_initial_multipliers_guess(handle) bind(C, name='etg_set_initial_multipliers_guess')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_set_initial_multipliers_guess
!
!     tbar_updater%update (diis_cc_multipliers_class.F90:269)
      integer(c_int) function etg_multiplier_quasi_newton_update(handle) bind(C, name='etg_multiplier_quasi_newton_update')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_multiplier_quasi_newton_update
!
!     diis%update(residual, multipliers) (diis_cc_multipliers_class.F90:271)
      integer(c_int) function etg_multiplier_diis_update(handle) bind(C, name='etg_multiplier_diis_update')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_multiplier_diis_update
!
!     set_multipliers (set_get_ccsd.F90:68-82)
      integer(c_int) function etg_set_multipliers(handle, x) bind(C, name='etg_set_multipliers')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(in) :: x
      end function etg_set_multipliers
!
!     get_multipliers (set_get_ccsd.F90:85-99)
      integer(c_int) function etg_get_multipliers(handle, x) bind(C, name='etg_get_multipliers')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(out) :: x
      end function etg_get_multipliers
!
!     construct_eta (multiplier_equation_ccsd.F90:56-110)
      integer(c_int) function etg_construct_eta(handle, x) bind(C, name='etg_construct_eta')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(out) :: x
      end function etg_construct_eta
!
!     construct_multiplier_equation (multiplier_equation_ccs.F90:58-94)
      integer(c_int) function etg_construct_multiplier_equation(handle, x, precision) &
                                                   bind(C, name='etg_construct_multiplier_equation')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(out) :: x
         integer(c_int), value :: precision
      end function etg_construct_multiplier_equation
!
!     get_l2_norm of the multiplier residual (diis_cc_multipliers_class.F90:249)
      integer(c_int) function etg_multiplier_residual_norm(handle, norm) &
                                                   bind(C, name='etg_multiplier_residual_norm')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), intent(out) :: norm
      end function etg_multiplier_residual_norm
!
!     construct_Jacobian_transform over n_vec trial vectors (ccs_class.F90:925-978)
      integer(c_int) function etg_jacobian_transformations(handle, side, n_vec, X, R, precision) &
                                                   bind(C, name='etg_jacobian_transformations')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int), value :: side
         integer(c_int64_t), value :: n_vec
         real(c_double), dimension(*), intent(in) :: X
         real(c_double), dimension(*), intent(out) :: R
         integer(c_int), value :: precision
      end function etg_jacobian_transformations
!
!     run_eigen_davidson_solver as one device call (eigen_davidson_solver_class.F90:122-318);
!     start_vectors may be c_null_ptr (unit vectors on the lowest orbital differences)
      integer(c_int) function etg_davidson(handle, side, n_states, residual_threshold, max_iterations, &
                                           max_dim, precision, start_vectors, energies, vectors, &
                                           residual_norms, iterations) bind(C, name='etg_davidson')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int), value :: side
         integer(c_int64_t), value :: n_states
         real(c_double), value :: residual_threshold
         integer(c_int64_t), value :: max_iterations, max_dim
         integer(c_int), value :: precision
         type(c_ptr), value :: start_vectors, vectors, residual_norms
         real(c_double), dimension(*), intent(out) :: energies
         integer(c_int64_t), intent(out) :: iterations
      end function etg_davidson
!
!     tuning: trial vectors per transformation pass (0: automatic)
      integer(c_int) function etg_set_transform_batch(handle, max_vectors) &
                                                   bind(C, name='etg_set_transform_batch')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int64_t), value :: max_vectors
      end function etg_set_transform_batch
!
!     diagnostic: bytes of the Jacobian record and per trial vector of a pass
      integer(c_int) function etg_jacobian_memory(handle, record_bytes, bytes_per_vector, &
                                                  unshared_bytes_per_vector) &
                                                  bind(C, name='etg_jacobian_memory')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int64_t), intent(out) :: record_bytes, bytes_per_vector, unshared_bytes_per_vector
      end function etg_jacobian_memory
!
!     construct_omega with the residual left in HBM (omega_ccsd.F90:39-110)
      integer(c_int) function etg_construct_omega_resident(handle, precision) &
                                                   bind(C, name='etg_construct_omega_resident')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: precision
      end function etg_construct_omega_resident
!
!     CC mean value of a one-electron operator with the Lagrangian density
!     (mean_value_ccs.F90:94-122, cc_mean_values_task_class.F90:88-92)
      integer(c_int) function etg_one_electron_expectation(handle, X_mo, value) &
                                                           bind(C, name='etg_one_electron_expectation')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(in) :: X_mo
         real(c_double), intent(out) :: value
      end function etg_one_electron_expectation
!
!     CC one-electron density, <X> = sum D_pq X_pq (mean_value_ccs.F90:94-122)
      integer(c_int) function etg_one_electron_density(handle, D) bind(C, name='etg_one_electron_density')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(out) :: D
      end function etg_one_electron_density
!
!     molecular producer: integrals -> Cholesky -> RHF -> MO vectors
!     (src/libint/eri.cpp:33-89, eri_cd_class.F90:241,3132, scf_solver_class.F90)
      integer(c_int) function etg_molecule_create(n_atoms, charges, xyz_bohr, n_shells, shell_atom, shell_l, &
                                                  shell_nprim, exponents, coefficients, n_electrons, molecule) &
                                                  bind(C, name='etg_molecule_create')
         import :: c_int, c_ptr, c_double, c_int64_t
         integer(c_int64_t), value :: n_atoms, n_shells, n_electrons
         real(c_double), dimension(*), intent(in) :: charges, xyz_bohr, exponents, coefficients
         integer(c_int64_t), dimension(*), intent(in) :: shell_atom, shell_l, shell_nprim
         type(c_ptr), intent(out) :: molecule
      end function etg_molecule_create
!
      integer(c_int) function etg_molecule_run(molecule, cholesky_threshold, scf_threshold, max_scf_iterations) &
                                               bind(C, name='etg_molecule_run')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: molecule
         real(c_double), value :: cholesky_threshold, scf_threshold
         integer(c_int64_t), value :: max_scf_iterations
      end function etg_molecule_run
!
      integer(c_int) function etg_molecule_set_orbitals(molecule, n_ao, n_mo, C, eps) &
                                                        bind(C, name='etg_molecule_set_orbitals')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: molecule
         integer(c_int64_t), value :: n_ao, n_mo
         real(c_double), dimension(*), intent(in) :: C, eps
      end function etg_molecule_set_orbitals
!
      integer(c_int) function etg_molecule_set_frozen_core(molecule, n_frozen) &
                                                           bind(C, name='etg_molecule_set_frozen_core')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: molecule
         integer(c_int64_t), value :: n_frozen
      end function etg_molecule_set_frozen_core
!
      integer(c_int) function etg_molecule_set_cartesian(molecule, cartesian) &
                                                         bind(C, name='etg_molecule_set_cartesian')
         import :: c_int, c_ptr
         type(c_ptr), value :: molecule
         integer(c_int), value :: cartesian
      end function etg_molecule_set_cartesian
!
      integer(c_int) function etg_molecule_dims(molecule, n_ao, n_occ, n_J, scf_iterations) &
                                                bind(C, name='etg_molecule_dims')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: molecule
         integer(c_int64_t), intent(out) :: n_ao, n_occ, n_J, scf_iterations
      end function etg_molecule_dims
!
      integer(c_int) function etg_molecule_results(molecule, e_nuc, e_hf, eps, h_mo, L_mo, C) &
                                                   bind(C, name='etg_molecule_results')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: molecule
         real(c_double), intent(out) :: e_nuc, e_hf
         real(c_double), dimension(*), intent(out) :: eps, h_mo, L_mo, C
      end function etg_molecule_results
!
      integer(c_int) function etg_molecule_destroy(molecule) bind(C, name='etg_molecule_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: molecule
      end function etg_molecule_destroy
!
!     AO dipole integrals <mu| r_c - origin_c |nu>, c = x, y, z (three n x n matrices)
      integer(c_int) function etg_molecule_dipole_integrals(molecule, origin, D) &
                                                   bind(C, name='etg_molecule_dipole_integrals')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: molecule
         real(c_double), dimension(3), intent(in) :: origin
         real(c_double), dimension(*), intent(out) :: D
      end function etg_molecule_dipole_integrals
!
!     checks: S, T, V and the dense AO integral matrix (small molecules)
      integer(c_int) function etg_molecule_one_electron(molecule, S, T, V) &
                                                   bind(C, name='etg_molecule_one_electron')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: molecule
         real(c_double), dimension(*), intent(out) :: S, T, V
      end function etg_molecule_one_electron
!
      integer(c_int) function etg_molecule_ao_eri(molecule, G) bind(C, name='etg_molecule_ao_eri')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: molecule
         real(c_double), dimension(*), intent(out) :: G
      end function etg_molecule_ao_eri
!
!     the last residual, copied to the host
      integer(c_int) function etg_get_omega(handle, omega) bind(C, name='etg_get_omega')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), dimension(*), intent(out) :: omega
      end function etg_get_omega
!
!     precision of the start block's transformation in etg_davidson (-1: the solve's)
      integer(c_int) function etg_set_davidson_start_precision(handle, precision) &
                                                   bind(C, name='etg_set_davidson_start_precision')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: precision
      end function etg_set_davidson_start_precision
!
!     placement of the Davidson space: 0 automatic, 1 HBM, 2 pinned host memory
      integer(c_int) function etg_set_davidson_space(handle, mode) bind(C, name='etg_set_davidson_space')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: mode
      end function etg_set_davidson_space
!
!     diagnostic: seconds in the transformations and in total, vectors transformed
      integer(c_int) function etg_davidson_timing(handle, reset, transform_seconds, total_seconds, &
                                                  transforms) bind(C, name='etg_davidson_timing')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int), value :: reset
         real(c_double), intent(out) :: transform_seconds, total_seconds
         integer(c_int64_t), intent(out) :: transforms
      end function etg_davidson_timing
!
!     virtual batch length of the ladder term's g_acbd blocks
      integer(c_int) function etg_set_ladder_batch(handle, batch) bind(C, name='etg_set_ladder_batch')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: handle
         integer(c_int64_t), value :: batch
      end function etg_set_ladder_batch
!
      integer(c_int) function etg_synchronize(handle) bind(C, name='etg_synchronize')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function etg_synchronize
!
!     diagnostics and tuning: the product ledger, graphs on or off, the measured tile choice
      integer(c_int) function etg_last_omega_flops(handle, flops) bind(C, name='etg_last_omega_flops')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: handle
         real(c_double), intent(out) :: flops
      end function etg_last_omega_flops
!
      integer(c_int) function etg_set_kernel_ledger(handle, on) bind(C, name='etg_set_kernel_ledger')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: on
      end function etg_set_kernel_ledger
!
      integer(c_int) function etg_kernel_ledger(handle, index, name, name_len, launches, seconds, flops, &
                                                bytes, count) bind(C, name='etg_kernel_ledger')
         import :: c_int, c_ptr, c_int64_t, c_double, c_char
         type(c_ptr), value :: handle
         integer(c_int64_t), value :: index, name_len
         character(kind=c_char), dimension(*), intent(out) :: name
         integer(c_int64_t), intent(out) :: launches, count
         real(c_double), intent(out) :: seconds, flops, bytes
      end function etg_kernel_ledger
!
      integer(c_int) function etg_set_transform_graphs(handle, on) bind(C, name='etg_set_transform_graphs')
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
         integer(c_int), value :: on
      end function etg_set_transform_graphs
!
      integer(c_int) function etg_set_gemm_autotune(on, tuned) bind(C, name='etg_set_gemm_autotune')
         import :: c_int, c_int64_t
         integer(c_int), value :: on
         integer(c_int64_t), intent(out) :: tuned
      end function etg_set_gemm_autotune
!
      integer(c_int) function etg_test_gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, precision) &
                                            bind(C, name='etg_test_gemm')
         import :: c_int, c_int64_t, c_double, c_char
         character(kind=c_char), value :: ta, tb
         integer(c_int64_t), value :: m, n, k, lda, ldb, ldc
         real(c_double), value :: alpha, beta
         real(c_double), dimension(*), intent(in) :: A, B
         real(c_double), dimension(*), intent(inout) :: C
         integer(c_int), value :: precision
      end function etg_test_gemm
!
      integer(c_int) function etg_bench_gemm(ta, tb, m, n, k, precision, reps, seconds) &
                                             bind(C, name='etg_bench_gemm')
         import :: c_int, c_int64_t, c_double, c_char
         character(kind=c_char), value :: ta, tb
         integer(c_int64_t), value :: m, n, k
         integer(c_int), value :: precision, reps
         real(c_double), intent(out) :: seconds
      end function etg_bench_gemm
!
      integer(c_int) function etg_test_reduced_eigen(k, G, n_want, wr, wi, X) &
                                                     bind(C, name='etg_test_reduced_eigen')
         import :: c_int, c_int64_t, c_double
         integer(c_int64_t), value :: k, n_want
         real(c_double), dimension(*), intent(in) :: G
         real(c_double), dimension(*), intent(out) :: wr, wi, X
      end function etg_test_reduced_eigen
!
!     The same solve on device memory with a static workspace: allocates nothing
!     (for callers that cannot allocate; k <= 128).  G and X are device addresses.
      integer(c_int) function etg_reduced_eigen_noalloc(k, G, ldg, n_want, wr, wi, X, ldx) &
                                                        bind(C, name='etg_reduced_eigen_noalloc')
         import :: c_int, c_int64_t, c_double, c_ptr
         integer(c_int64_t), value :: k, ldg, n_want, ldx
         type(c_ptr), value :: G, X
         real(c_double), dimension(*), intent(out) :: wr, wi
      end function etg_reduced_eigen_noalloc
!
   end interface
!
contains
!
   function etg_error_message() result(message)
!!
!!    Copies the library's last error string into a Fortran character variable.
!!
      implicit none
!
      character(len=:), allocatable :: message
      character(kind=c_char), dimension(:), pointer :: chars
      integer :: n
!
      message = ''
      if (.not. c_associated(etg_last_error())) return
      call c_f_pointer(etg_last_error(), chars, [4096])
      n = 0
      do while (n < 4096)
         if (chars(n + 1) == c_null_char) exit
         n = n + 1
      enddo
      allocate(character(len=n) :: message)
      message = transfer(chars(1:n), message)
!
   end function etg_error_message
!
end module etgpu_interface
