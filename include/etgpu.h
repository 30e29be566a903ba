/*
 * etgpu.h - C ABI of the MI355X CCSD residual library (libetgpu.so).
 *
 * The entry points replace the type-bound procedures that eT's ground-state
 * solvers call on a `ccsd` wavefunction.  Each one is a thin, synchronous
 * call: host arrays are borrowed for the duration of the call only, all
 * tensors (Cholesky vectors, amplitudes, Fock matrix, residual) live in GPU
 * memory owned by the handle.  Array layouts are Fortran (column-major) and
 * exactly those of eT:
 *
 *   L(n_J, n_mo, n_mo)   Cholesky vectors, J fastest, occupied orbitals first
 *   t1(n_v, n_o)         singles,   compound index ai = a + n_v*i (0-based)
 *   t2(n_t2)             doubles packed as idx(ai,bj) = max*(max+1)/2 + min,
 *                        n_t2 = n_o*n_v*(n_o*n_v+1)/2
 *   amplitude / residual vectors are [t1 | t2] of length n_t1 + n_t2.
 *
 * Sizes are int64_t (eT is built with 8-byte default integers,
 * cmake/Fortran_flags.cmake:92).  Every function returns 0 on success and a
 * non-zero status otherwise; etg_last_error() then returns a message
 * (thread-local, valid until the next failing call on the same thread).
 * No exception crosses the ABI.  A handle is not re-entrant: call it from one
 * thread at a time (eT's solver loops are single threaded).
 */
#ifndef ETGPU_H
#define ETGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct etg_handle etg_handle;

/* Arithmetic of the bulk contractions in etg_construct_omega. */
enum etg_precision {
  ETG_FP64 = 0,   /* FP64 matrix cores throughout */
  ETG_FP32 = 1,   /* FP32 matrix-core products, FP32 accumulation */
  ETG_BF16X3 = 2, /* operands split into bf16 hi + lo, three bf16 matrix-core
                     products per term (hi*hi + hi*lo + lo*hi), FP32 accumulation;
                     a product with few outputs per operand element (m n below
                     64 (m + n): bound by HBM even in FP64) and operands of 2^24
                     elements or more runs in FP64 */
  ETG_BF16X3_ANY_SHAPE = 3 /* etg_test_gemm and etg_bench_gemm only: the split-
                     bf16 kernel for every shape */
};

/* Creates the device state for a CCSD wavefunction with n_o occupied and
 * n_v virtual orbitals and n_J Cholesky vectors, on the current HIP device.
 * Replaces: ccsd%initialize (src/wavefunctions/ccsd/ccsd_class.F90:323-354). */
int etg_create(int64_t n_o, int64_t n_v, int64_t n_J, etg_handle** handle);

/* Releases all device memory of the handle.  Replaces: ccsd%cleanup. */
int etg_destroy(etg_handle* handle);

/* Message of the last failing call on this thread ("" if none). */
const char* etg_last_error(void);

/* Uploads the MO Cholesky vectors L(n_J, n_mo, n_mo) and T1-dresses them with
 * the current singles (zero unless amplitudes were set before).
 * Replaces: the L_mo producer hand-over (src/engines/cc/tasks/
 * cholesky_decomposition_task_class.F90:127-134, L_t1 <- L_mo :133). */
int etg_set_cholesky(etg_handle* handle, const double* L);

/* MO one-electron Hamiltonian h(n_mo, n_mo).
 * Replaces: the MO-transformed integrals of get_t1_oei('hamiltonian')
 * (src/wavefunctions/ccs/oei_ccs.F90:35-97) before the T1 transformation. */
int etg_set_oei(etg_handle* handle, const double* h);

/* Orbital energies eps(n_mo).  Replaces: wf%orbital_energies
 * (src/wavefunctions/wavefunction_class.F90:35-60). */
int etg_set_orbital_energies(etg_handle* handle, const double* eps);

/* Amplitudes [t1 | t2].  Replaces: set_amplitudes / get_amplitudes
 * (src/wavefunctions/ccsd/set_get_ccsd.F90). */
int etg_set_amplitudes(etg_handle* handle, const double* t);
int etg_get_amplitudes(etg_handle* handle, double* t);

/* t1 = 0 is not implied: sets t2 to the MP2 guess from the current
 * T1-dressed integrals and orbital energies.
 * Replaces: set_t2_to_cc2_guess (src/wavefunctions/ccsd/ccsd_class.F90:478-533). */
int etg_set_t2_to_mp2_guess(etg_handle* handle);

/* L_t1 = X L_mo Y^T from the current singles.
 * Replaces: construct_t1_cholesky (src/wavefunctions/ccs/t1_ccs.F90:208-246). */
int etg_construct_t1_cholesky(etg_handle* handle);

/* T1-transformed Fock matrix, all four blocks; F (n_mo, n_mo) is copied out
 * when non-NULL.  Replaces: construct_fock (src/wavefunctions/ccsd/fock_ccsd.F90:34-79). */
int etg_construct_fock(etg_handle* handle, double* F);

/* Correlation energy sum (t_aibj + t_ai t_bj)(2 g_iajb - g_ibja).
 * Replaces: calculate_energy (src/wavefunctions/ccsd/mean_value_ccsd.F90:36-99);
 * the caller adds E_HF as eT does. */
int etg_calculate_energy(etg_handle* handle, double* e_corr);

/* Residual Omega(n_t1 + n_t2) for the current amplitudes, in eT's packed
 * convention (doubles diagonal carries the factor 1/2).  Copied to `omega`
 * when non-NULL, otherwise left in device memory (see etg_get_omega).
 * Replaces: construct_omega (src/wavefunctions/ccsd/omega_ccsd.F90:39-110). */
int etg_construct_omega(etg_handle* handle, double* omega, int precision);

/* Copies the last residual to the host. */
int etg_get_omega(etg_handle* handle, double* omega);

/* ||Omega||_2 of the last residual (FP64 reduction on the device).
 * Replaces: get_l2_norm(omega) in run_diis_cc_gs (src/solvers/cc/diis_cc_gs_class.F90:260). */
int etg_omega_norm(etg_handle* handle, double* norm);

/* eps_a - eps_i (+ eps_b - eps_j) for n = n_t1 or n_t1 + n_t2 entries.
 * Replaces: get_orbital_differences (src/wavefunctions/doubles/doubles_class.F90:421-470). */
int etg_get_orbital_differences(etg_handle* handle, double* x, int64_t n);

/* Records the residual program (T1 dressing, Fock, Omega) at the current
 * amplitudes with every intermediate kept in HBM, for the Jacobian
 * transformations below.  Invalidated by any change of amplitudes or inputs.
 * Replaces: prepare_for_jacobian / prepare_for_jacobian_transpose
 * (src/wavefunctions/ccsd/jacobian_ccsd.F90:46-71, jacobian_transpose_ccsd.F90:44-127). */
int etg_prepare_for_jacobian(etg_handle* handle);

/* rho = A c with A_mu,nu = <mu~| exp(-T) [H, tau_nu] exp(T) |R> in eT's
 * biorthonormal packed basis (n_t1 + n_t2 entries each).
 * Replaces: jacobian_transformation (src/wavefunctions/ccsd/jacobian_ccsd.F90:74-178). */
int etg_jacobian_transformation(etg_handle* handle, const double* c, double* rho);

/* sigma^T = b^T A (same A).
 * Replaces: jacobian_transpose_transformation
 * (src/wavefunctions/ccsd/jacobian_transpose_ccsd.F90:129-236). */
int etg_jacobian_transpose_transformation(etg_handle* handle, const double* b, double* sigma);

/* etg_construct_omega with the residual kept in HBM (for the device solvers;
 * same as passing omega = NULL, for callers that cannot pass a null array).
 * Replaces: construct_omega (omega_ccsd.F90:39-110) inside run_diis_cc_gs. */
int etg_construct_omega_resident(etg_handle* handle, int precision);

/* Allocates HBM records for DIIS of the given dimension (1..32) and clears
 * the history.
 * Replaces: diis%initialize (src/solver_tools/diis_tool_class.F90:249-277)
 * as run_diis_cc_gs calls it (src/solvers/cc/diis_cc_gs_class.F90:230). */
int etg_diis_initialize(etg_handle* handle, int64_t dimension);

/* Releases the DIIS records (their HBM returns to the device) at the end of a
 * solve; etg_diis_initialize allocates them again.
 * Replaces: diis%finalize (finalize_diis_tool, diis_tool_class.F90:280-299),
 * called by eT's solvers when they finish (diis_cc_gs_class.F90:313). */
int etg_diis_finalize(etg_handle* handle);

/* etg_diis_initialize with eT's CROP variant: after each extrapolation the newest
 * records become the extrapolated parameters <x> and error <e> = sum_i w_i e_i.
 * Replaces: diis_tool(..., crop=.true.) (diis_tool_class.F90:150-210, 763-805) as
 * the keyword crop of the solver cc gs / solver cc multipliers sections sets it
 * (diis_cc_gs_class.F90:393, diis_cc_multipliers_class.F90:370). */
int etg_diis_initialize_crop(etg_handle* handle, int64_t dimension);

/* Quasi-Newton step on the device, FP64: e = -Omega / eps (orbital
 * differences; doubles diagonal doubled), amplitudes += e.  Needs a residual
 * constructed at the current amplitudes; e is kept for etg_diis_update.
 * Replaces: t_updater%update(wf, amplitudes, omega)
 * (src/solver_tools/amplitude_updaters/amplitude_updater_class.F90:86-121,
 * quasi_newton_updater_class.F90:81-118) at diis_cc_gs_class.F90:284-285. */
int etg_quasi_newton_update(etg_handle* handle);

/* DIIS extrapolation on the device, FP64: stores (e, amplitudes) as the newest
 * record (the oldest is dropped once the space is full), adds the new row of
 * G_ij = e_i . e_j, solves the padded DIIS equation and sets the amplitudes to
 * the extrapolated vector.  Fails (amplitudes unchanged) on a singular system.
 * Replaces: diis%update(omega, amplitudes)
 * (src/solver_tools/diis_tool_class.F90:302-347) at diis_cc_gs_class.F90:287. */
int etg_diis_update(etg_handle* handle);

/* Davidson solution of the n_states lowest roots of A (side 0, right) or A^T
 * (side 1, left) with trial vectors, transforms and all reduced-space algebra
 * (Gram-Schmidt, G = V^T A V, its eigenproblem, residuals) on the device in
 * FP64; the transformations themselves in `precision`.  Start vectors: the
 * columns of start_vectors ((n_t1 + n_t2) x n_states, host or device,
 * orthonormalised), or with NULL unit vectors on the lowest orbital
 * differences (eT's default); diagonal preconditioner, convergence when every
 * ||A x - w x|| / ||x|| < residual_threshold, restart on the Ritz vectors at
 * max_dim (<= 0: max(100, 10 n_states); at most 128).  Needs
 * etg_prepare_for_jacobian.  energies: n_states; vectors: (n_t1 + n_t2) x
 * n_states, host or device, may be NULL; residual_norms: max_iterations x
 * n_states per-iteration norms, may be NULL; iterations: may be NULL.  With
 * precision FP32 or BF16X3 the solve also ends, with *iterations = -(count),
 * when the largest residual has not improved by 10 % in 6 iterations (the
 * transforms' rounding floor); vectors then hold the current Ritz vectors, the start of an
 * FP64 solve.
 * Replaces: run_eigen_davidson_solver
 * (src/solvers/eigen_davidson_solver_class.F90:122-318) with eT's eigen
 * Davidson tool (src/solver_tools/davidson/eigen_davidson_tool_class.F90). */
int etg_davidson(etg_handle* handle, int side, int64_t n_states, double residual_threshold,
                 int64_t max_iterations, int64_t max_dim, int precision,
                 const double* start_vectors, double* energies, double* vectors,
                 double* residual_norms, int64_t* iterations);

/* Precision of the first transformation pass of etg_davidson (the start
 * vectors); the expansion vectors of later iterations use the solve's own
 * precision.  -1 (the default): the solve's precision.  With ETG_FP64 here and
 * a reduced-precision solve started from converged reduced-precision Ritz
 * vectors, every Ritz vector is an FP64-transformed start block plus
 * reduced-precision corrections whose weights are of the size of the start
 * residuals, so the solve reaches residuals far below the reduced precision's
 * rounding floor while transforming only the start block in FP64.
 * Replaces: no eT procedure (eT's Davidson transforms in FP64 throughout); the
 * precision schedule of the excited-state solve, set before etg_davidson. */
int etg_set_davidson_start_precision(etg_handle* handle, int precision);

/* Where etg_davidson keeps its trial space V and transforms A V (n x max_dim
 * each): 0 (the default) in HBM when both fit in half of the free device
 * memory, otherwise A V and, if needed, V as well in pinned host memory that the
 * kernels read and write over PCIe; 1 always HBM; 2 always host memory.
 * Results do not depend on the placement.
 * Replaces: the records_in_memory choice of eT's Davidson tool
 * (davidson_tool_class.F90:150, record_storer_class.F90), which keeps the
 * trial and transformed vectors in memory or on disk. */
int etg_set_davidson_space(etg_handle* handle, int mode);

/* Diagnostic: wall seconds etg_davidson spent in its transformation passes and
 * in total, and the number of vectors it transformed, summed over the solves
 * since the last reset (reset != 0 zeroes them after reading). */
int etg_davidson_timing(etg_handle* handle, int reset, double* transform_seconds, double* total_seconds,
                        int64_t* transforms);

/* A X = F (side 0) or A^T X = F (side 1) by eT's linear Davidson solver, on
 * the device: start from start_vector (may be NULL: the preconditioner guess
 * F / eps), one new trial per iteration (-R / eps of the residual R = A X - F,
 * orthonormalised against the space), the reduced equation solved exactly,
 * restart on the current solution at max_dim (<= 0: 50; at most 128),
 * converged when ||R|| <= residual_threshold.  rhs, start_vector, solution:
 * n_t1 + n_t2, host or device; residual_norms (max_iterations, may be NULL);
 * fails after max_iterations.  Needs etg_prepare_for_jacobian.  With side 1 and
 * rhs = -eta this is eT's default CCSD multiplier solver.
 * Replaces: run_linear_davidson_solver (src/solvers/linear_davidson_solver_class.F90:122-270)
 * with the linear Davidson tool (src/solver_tools/davidson/linear_davidson_tool_class.F90),
 * as cc_multipliers_solver_factory_class.F90:137-230 builds it. */
int etg_linear_davidson(etg_handle* handle, int side, const double* rhs, const double* start_vector,
                        double residual_threshold, int64_t max_iterations, int64_t max_dim, int precision,
                        double* solution, double* residual_norms, int64_t* iterations);

/* Core-valence separation for etg_davidson and etg_diis_excited_states (eT's
 * `core excitation` keyword): each transformed vector is projected onto
 * excitations with at least one occupied index among the n_core core_mos (0-based
 * occupied MOs): singles (a, i) with i a core MO, doubles (ai, bj) with i or j one.
 * n_core = 0 switches it off.  Start vectors are the caller's: eT starts on the
 * unit singles (a, i) with a = 1, 2, ... and, for each a, the core MOs in order.
 * Replaces: es_cvs_projection_tool%project (es_cvs_projection_tool_class.F90:73-85;
 * cvs_projection_ccs, ccs_class.F90:1219-1250; cvs_projection_doubles,
 * doubles_class.F90:287-328) in eigen_davidson_solver_class.F90:187 and
 * diis_cc_es_class.F90:359; start vectors set_cvs_start_indices_ccs (ccs_class.F90:1393-1438). */
int etg_set_cvs(etg_handle* handle, int64_t n_core, const int64_t* core_mos);

/* The opposite projection (eT's `remove core`): onto excitations with no occupied
 * index among the core_mos, for etg_davidson and etg_diis_excited_states (start
 * vectors: eT's usual ones, the lowest orbital differences).  n_core = 0 switches
 * it off, as does etg_set_cvs(handle, 0, NULL).
 * Replaces: es_rm_core_projection_tool (rm_core_projection_ccs, ccs_class.F90:1251-1280;
 * rm_core_projection_doubles, doubles_class.F90:330-370; cc_es_projector_factory_class.F90:100-102). */
int etg_set_remove_core(etg_handle* handle, int64_t n_core, const int64_t* core_mos);

/* Ionisation (eT's `ionization` with `bath orbital`): the caller appends one virtual
 * orbital with zero Cholesky vectors, zero h and orbital energy 0 (eT's
 * make_bath_orbital_ccs, ccs_class.F90:1494-1582); with on != 0 etg_davidson and
 * etg_diis_excited_states keep only excitations into it (singles (bath, i),
 * doubles with a or b the bath), and their roots are ionisation energies.  eT
 * starts on the singles (bath, i) from the highest occupied down
 * (set_ip_start_indices_ccs, ccs_class.F90:1585-1608).  Switches off the core
 * projections; etg_set_cvs / etg_set_remove_core switch it off.
 * Replaces: es_ip_projection_tool (ip_projection_ccs, ccs_class.F90:1611-1640;
 * ip_projection_doubles, doubles_class.F90:376-418). */
int etg_set_ionization(etg_handle* handle, int on);

/* eT's Newton-Raphson amplitude update: A u = -Omega by the linear Davidson
 * solver (as etg_linear_davidson, start -Omega / eps, reduced space 50) to
 * ||A u + Omega|| < rel_threshold ||Omega||, the packed doubles diagonal of u
 * doubled, t += u; the following etg_diis_update records (Omega, t).  Needs the
 * residual of the current amplitudes (etg_construct_omega) and
 * etg_prepare_for_jacobian at them.  micro_norms (max_iterations, may be NULL):
 * the linear iterations' residual norms; iterations: their number.
 * Replaces: newton_raphson_updater%update
 * (src/solver_tools/amplitude_updaters/newton_raphson_updater_class.F90:118-210,
 * amplitude_updater_class.F90:86-121) in run_diis_cc_gs. */
int etg_newton_raphson_update(etg_handle* handle, double rel_threshold, int64_t max_iterations, int precision,
                              double* micro_norms, int64_t* iterations);

/* The n_states eigenpairs of A (side 0) or A^T (side 1) by eT's DIIS
 * excited-state solver, on the device: each state's vector X follows
 * X += -(A X - E X) / (eps - E), E = X . A X, with DIIS over that state's own
 * records (dimension diis_dimension, eT's default 20) and X normalised; a state
 * is converged when ||A X - E X|| < residual_threshold and |dE| <
 * energy_threshold (energy_threshold <= 0: residual only; the first iteration
 * tests the residual only), and is not transformed again.  The unconverged
 * states share one transformation pass per iteration.  Start vectors as in
 * etg_davidson (the roots reached are those the start vectors lead to, as in
 * eT).  On return energies (n_states) are sorted ascending and vectors
 * ((n_t1 + n_t2) x n_states, host or device, may be NULL) follow them;
 * residual_norms and energy_history (max_iterations x n_states, in the start
 * order, may be NULL) hold each iteration's norms and energies; fails after
 * max_iterations.
 * Replaces: run_diis_cc_es (src/solvers/cc/diis_cc_es_class.F90:260-492). */
int etg_diis_excited_states(etg_handle* handle, int side, int64_t n_states, double residual_threshold,
                            double energy_threshold, int64_t max_iterations, int64_t diis_dimension,
                            int precision, const double* start_vectors, double* energies, double* vectors,
                            double* residual_norms, double* energy_history, int64_t* iterations);

/* ---- multiplier (Lambda) equations  t-bar^T A + eta = 0 ---------------- */

/* Records the residual program at the current amplitudes for the left
 * transformation (same record as etg_prepare_for_jacobian).
 * Replaces: prepare_for_multiplier_equation
 * (src/wavefunctions/ccsd/multiplier_equation_ccsd.F90:41-53). */
int etg_prepare_for_multiplier_equation(etg_handle* handle);

/* t-bar1 = t1, t-bar2 = 2 t_aibj - t_ajbi (packed), on the device.
 * Replaces: set_initial_multipliers_guess (ccsd_class.F90:423-475, no restart). */
int etg_set_initial_multipliers_guess(etg_handle* handle);

/* Multipliers t-bar (n_t1 + n_t2, amplitude layout).
 * Replaces: set_multipliers / get_multipliers (set_get_ccsd.F90:68-99). */
int etg_set_multipliers(etg_handle* handle, const double* tbar);
int etg_get_multipliers(etg_handle* handle, double* tbar);

/* eta_ai = 2 F_ia, eta_aibj = 4 g_iajb - 2 g_ibja (packed), from the T1 Fock
 * matrix of the last Fock construction or record.  eta may be NULL (kept in HBM).
 * Replaces: construct_eta (multiplier_equation_ccsd.F90:56-110). */
int etg_construct_eta(etg_handle* handle, double* eta);

/* R = A^T t-bar + eta on the device (needs etg_prepare_for_multiplier_equation
 * at the current amplitudes), the left transformation's bulk products in
 * `precision` (etg_fp64 for eT's equation; the reduced forms for the early
 * iterations of a mixed-precision solve).  residual may be NULL (kept in HBM).
 * Replaces: construct_multiplier_equation (multiplier_equation_ccs.F90:58-94). */
int etg_construct_multiplier_equation(etg_handle* handle, double* residual, int precision);

/* ||R||_2 (FP64 reduction on the device) of the last multiplier residual.
 * Replaces: get_l2_norm(residual) (diis_cc_multipliers_class.F90:249). */
int etg_multiplier_residual_norm(etg_handle* handle, double* norm);

/* e = -R / eps (no diagonal scaling), t-bar += e, on the device.
 * Replaces: tbar_updater%update (diis_cc_multipliers_class.F90:269) with the
 * quasi-Newton updater eT builds for it (cc_multipliers_solver_factory_class.F90:249-251). */
int etg_multiplier_quasi_newton_update(etg_handle* handle);

/* DIIS extrapolation of the multipliers over the records of etg_diis_initialize.
 * Replaces: diis%update(residual, multipliers) (diis_cc_multipliers_class.F90:271). */
int etg_multiplier_diis_update(etg_handle* handle);

/* n_vec transformations at once: R_k = A X_k (side 0) or A^T X_k (side 1)
 * with the bulk products in `precision` (etg_precision), X and R column-major
 * (n_t1 + n_t2) x n_vec, host or device (HIP) pointers.  Same preparation and
 * conventions as etg_jacobian_transformation.
 * Replaces: the per-vector loop of construct_Jacobian_transform(r_or_l, X, R)
 * (src/wavefunctions/ccs/ccs_class.F90:925-978) over a Davidson iteration's
 * new trial vectors. */
int etg_jacobian_transformations(etg_handle* handle, int side, int64_t n_vec, const double* X,
                                 double* R, int precision);

/* Virtual batch length of the ladder term's g_acbd blocks (memory/launch
 * trade-off; results do not depend on it beyond rounding). */
int etg_set_ladder_batch(etg_handle* handle, int64_t batch);

/* Waits for all work queued on the handle's stream. */
int etg_synchronize(etg_handle* handle);

/* Floating-point operations issued by the last etg_construct_omega (sum of
 * 2 m n k over the products it ran). */
int etg_last_omega_flops(etg_handle* handle, double* flops);

/* ---- Diagnostics: no counterpart in the reference interface; used by the
 * tests and the roofline measurement of the matrix-core product kernel. ---- */

/* Turns the per-kernel ledger on (1) or off (0) and clears it.  While on,
 * every matrix-core product is bracketed by HIP events on the handle's
 * stream; entries aggregate launches, seconds and 2mnk flops per kernel
 * instantiation. */
int etg_set_kernel_ledger(etg_handle* handle, int on);

/* Entry `index` of the ledger (name copied into `name`, NUL-terminated):
 * launches, seconds, 2mnk flops and algorithmic bytes (each operand read once,
 * C written once and read when beta != 0; FP64 operands) summed over the
 * kernel's launches; `count` receives the number of entries.  index < 0 only
 * returns count. */
int etg_kernel_ledger(etg_handle* handle, int64_t index, char* name, int64_t name_len,
                      int64_t* launches, double* seconds, double* flops, double* bytes,
                      int64_t* count);

/* Diagnostic: replay captured HIP graphs of the transformation passes (on, the
 * default) or run every pass eagerly (off).  Results are identical. */
int etg_set_transform_graphs(etg_handle* handle, int on);

/* Tuning: at most max_vectors trial vectors per transformation pass (0, the
 * default: as many as 35 % of the free HBM holds, at most 16).  The vectors of
 * one pass are the batch dimension of every product, so the ladder's g+/- and
 * every t-dependent operand are read once per group.  Results agree with
 * one-vector passes to the products' rounding. */
int etg_set_transform_batch(etg_handle* handle, int64_t max_vectors);

/* Diagnostic, after etg_prepare_for_jacobian (each may be NULL): bytes of the
 * recorded intermediates; bytes one trial vector of a transformation pass
 * needs (derivatives of intermediates with disjoint lifetimes share memory);
 * the same with one derivative buffer per intermediate. */
int etg_jacobian_memory(etg_handle* handle, int64_t* record_bytes, int64_t* bytes_per_vector,
                        int64_t* unshared_bytes_per_vector);

/* C = alpha op(A) op(B) + beta C on host arrays (copied to the device and
 * back), column-major, op = 'N' or 'T', with the given precision. */
int etg_test_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                  const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                  double* C, int64_t ldc, int precision);

/* Diagnostic: lowest n_want eigenpairs of a k x k real matrix (host, column-
 * major) with the device reduced-space solver: wr, wi (k, ascending real
 * part), X (k x n_want, unit vectors).  Fails on complex wanted roots. */
int etg_test_reduced_eigen(int64_t k, const double* G, int64_t n_want, double* wr, double* wi,
                           double* X);

/* The reduced-space eigensolver of etg_test_reduced_eigen on caller-owned
 * device memory, allocating nothing: G (device, k x k, ld ldg) -> wr, wi (host,
 * k values by ascending real part), X (device, k x n_want, ld ldx).  Its
 * workspace (eigenvalue staging, the complex-pair systems, the status flag;
 * 526 KB) is a static device array, one per GPU, held until the process ends;
 * k above 128 (the array's size) is rejected with an error.  Calls that overlap
 * on one GPU share that array and race.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight solver or serving loops).  Runs on the null stream and synchronises it. */
int etg_reduced_eigen_noalloc(int64_t k, const double* G_device, int64_t ldg, int64_t n_want,
                              double* wr, double* wi, double* X_device, int64_t ldx);

/* Seconds per product for `reps` back-to-back m x n x k products on
 * device-resident operands. */
int etg_bench_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, int precision, int reps,
                   double* seconds);

/* CC mean value of a one-electron operator X (n_mo x n_mo, MO basis, host)
 * with the CC Lagrangian (Lambda) density at the current amplitudes and
 * multipliers: <X> = <HF|(1 + Lambda) e^-T X e^T|HF>, evaluated as dL/de for
 * h -> h + e X (one forward replay of the recorded residual program plus the
 * reference and t1 terms of the energy).  Needs etg_prepare_for_jacobian (or
 * etg_prepare_for_multiplier_equation) at the converged amplitudes and the
 * multipliers.  The electronic dipole is -<r_c> (Debye: times 2.541746473).
 * Replaces: construct_gs_density (src/wavefunctions/ccs/mean_value_ccs.F90:94-122,
 * src/wavefunctions/doubles/mean_value_doubles.F90) with the CC mean values
 * (src/engines/cc/tasks/cc_mean_values_task_class.F90:88-92). */
int etg_one_electron_expectation(etg_handle* handle, const double* X_mo, double* value);

/* The CC one-electron (Lambda) density D (n_mo x n_mo, MO basis, host) with
 * <X> = sum_pq D_pq X_pq for every one-electron operator X: D = dL/dh, from
 * one backward replay of the recorded residual program seeded with the
 * multipliers, plus the reference and t1 terms of the energy.  One replay serves
 * every operator (the three dipole components at once).  Same preconditions as
 * etg_one_electron_expectation.
 * Replaces: construct_gs_density (src/wavefunctions/ccs/mean_value_ccs.F90:94-122,
 * src/wavefunctions/doubles/mean_value_doubles.F90:78), back-transformed from
 * the T1 basis to the MO basis. */
int etg_one_electron_density(etg_handle* handle, double* D);

/* ---- molecular producer of the inputs: AO integrals -> RHF -> Cholesky -> MO -- */

typedef struct etg_molecule etg_molecule;

/* A closed-shell molecule: n_atoms nuclei (charges, xyz in bohr, 3 per atom),
 * n_shells contracted Gaussian shells in AO order (shell_atom, angular momentum
 * shell_l <= 2 with spherical d, shell_nprim primitives whose exponents and
 * unnormalised contraction coefficients are consecutive in the flat arrays),
 * n_electrons even.
 * Replaces: eT's molecular system and AO basis set-up feeding Libint
 * (src/libint/eri.cpp:33-89). */
int etg_molecule_create(int64_t n_atoms, const double* charges, const double* xyz_bohr, int64_t n_shells,
                        const int64_t* shell_atom, const int64_t* shell_l, const int64_t* shell_nprim,
                        const double* exponents, const double* coefficients, int64_t n_electrons,
                        etg_molecule** molecule);

/* On the device: the AO integrals (McMurchie-Davidson; the dense (pq|rs)
 * matrix in HBM), its pivoted Cholesky decomposition to cholesky_threshold on
 * the largest remaining diagonal, RHF with DIIS (core-Hamiltonian guess, J and
 * K from the Cholesky vectors; converged when max|FDS - SDF| < scf_threshold
 * and |dE| < 1e-13) and the AO->MO transformation of the Cholesky vectors.
 * Replaces: eri_cd (src/solvers/cholesky/eri_cd_class.F90:241, AO->MO :3132)
 * and the RHF solver (src/solvers/hf/scf_solver_class.F90). */
int etg_molecule_run(etg_molecule* molecule, double cholesky_threshold, double scf_threshold,
                     int64_t max_scf_iterations);

/* Restart orbitals for etg_molecule_run: C (n_ao x n_mo, column-major) and the
 * orbital energies, as eT's orbital_coefficients file holds them; n_ao and n_mo
 * must equal the molecule's AO count.  The run then takes these orbitals as they
 * are (no RHF iterations): one Fock build gives E_HF, and the run fails unless
 * max|FDS - SDF| < scf_threshold.
 * Replaces: read_for_scf_restart_hf with the SCF skipped (src/wavefunctions/hf/
 * hf_class.F90:326-338, 1525-1537; read_orbitals_hf, file_handling_hf.F90:55-81). */
int etg_molecule_set_orbitals(etg_molecule* molecule, int64_t n_ao, int64_t n_mo, const double* C,
                              const double* eps);

/* Cartesian Gaussians (eT's `cartesian gaussians` in `system`): every d shell
 * keeps its six Cartesian components (each normalised) instead of the five real
 * solid harmonics; the AO count grows accordingly (etg_molecule_dims).  Call before
 * any integrals are built.
 * Replaces: the Cartesian basis of eT's AO tool (ao_tool_class.F90, Libint's
 * Cartesian shells, src/libint/eri.cpp). */
int etg_molecule_set_cartesian(etg_molecule* molecule, int cartesian);

/* Frozen core for the next etg_molecule_run: h_mo then includes the Coulomb and
 * exchange of the n_frozen lowest MOs, F^fc_pq = sum_c 2 (pq|cc) - (pc|cq), and the
 * caller passes the MOs above them to the CCSD handle (n_o - n_frozen occupied;
 * E_HF is unchanged).  eT freezes one orbital per atom with 5 <= Z <= 12, five for
 * 13..30, nine for 31..38 (ao_tool_class.F90:2425-2507).
 * Replaces: remove_core_orbitals_hf and construct_mo_fock_fc_term_hf
 * (src/wavefunctions/hf/frozen_orbital_hf.F90:81-205, 547-600). */
int etg_molecule_set_frozen_core(etg_molecule* molecule, int64_t n_frozen);

/* Sizes: AO (= MO) count, occupied orbitals, Cholesky vectors, RHF iterations. */
int etg_molecule_dims(etg_molecule* molecule, int64_t* n_ao, int64_t* n_occ, int64_t* n_J,
                      int64_t* scf_iterations);

/* After etg_molecule_run (each may be NULL): E_nuc, E_HF, orbital energies
 * (n), h_mo (n x n), L_mo (n_J x n x n, J fastest: eT's layout, the input of
 * etg_set_cholesky; host or device pointer), MO coefficients C (n x n). */
int etg_molecule_results(etg_molecule* molecule, double* e_nuc, double* e_hf, double* eps, double* h_mo,
                         double* L_mo, double* C);

/* Checks: S, T, V (n x n) and the dense (pq|rs) matrix (n^2 x n^2; small
 * molecules) in the spherical AO basis. */
int etg_molecule_one_electron(etg_molecule* molecule, double* S, double* T, double* V);

/* Dipole integrals <mu| r_c - origin_c |nu>, c = x, y, z, in the AO basis: D
 * holds three n x n matrices.  Replaces: the AO dipole integrals eT takes from
 * Libint for its CC mean values (src/libint/eri.cpp). */
int etg_molecule_dipole_integrals(etg_molecule* molecule, const double* origin, double* D);
int etg_molecule_ao_eri(etg_molecule* molecule, double* G);

int etg_molecule_destroy(etg_molecule* molecule);

/* Tuning, process-wide: measured tile choice for the FP64/FP32 products (on by
 * default).  The first product of a shape outside graph capture times the
 * cost model's tile size and split-K count against its neighbours (the other
 * tile size's best, half and twice the split count) on the real operands and
 * keeps the fastest; off keeps the model's pick.  `tuned` (may be NULL)
 * receives the number of shapes measured so far. */
int etg_set_gemm_autotune(int on, int64_t* tuned);

#ifdef __cplusplus
}
#endif

#endif /* ETGPU_H */
