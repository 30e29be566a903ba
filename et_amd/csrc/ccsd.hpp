// Device-resident state of one CCSD wavefunction and the operations eT's
// solvers call on it (reference: src/wavefunctions/ccsd/ccsd_class.F90).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "ops.hpp"

namespace etg {

// Lowest n_want eigenpairs of a small non-symmetric matrix on the device (one
// wavefront; dimension <= 128): G (device, ld ldg) -> wr, wi (host, k values by
// ascending real part), X (device, k x n_want, ld ldx).  Throws on complex
// wanted roots or QR non-convergence.
void reduced_eigen_solve(hipStream_t stream, DeviceBuffer& workspace, const double* G_dev, int64_t k,
                         int64_t ldg, int64_t n_want, double* wr, double* wi, double* X_dev,
                         int64_t ldx);
// The same solve with its workspace in a static device array (no allocation;
// dimension <= 128).  Calls that overlap on one GPU share that array and race.
void reduced_eigen_solve_noalloc(hipStream_t stream, const double* G_dev, int64_t k, int64_t ldg,
                                 int64_t n_want, double* wr, double* wi, double* X_dev, int64_t ldx);

// One DIIS space in HBM (diis_tool_class.F90): up to `dim` records of an error
// vector and a parameter vector of length n, the matrix of error overlaps and
// the weights of the last solve.
struct DiisSpace {
  int dim = 0;
  bool crop = false;          // eT's CROP variant (conjugate residual with optimal trial vectors)
  int64_t count = 0, n = 0;
  DeviceBuffer e, x, G, w, info;
};

class CCSD {
 public:
  CCSD(int64_t n_o, int64_t n_v, int64_t n_J);
  ~CCSD();

  int64_t n_o, n_v, n_J, n_mo, n_t1, n_t2;

  // inputs
  void set_cholesky(const double* L_host);        // L(J, p, q), full MO range
  void set_oei(const double* h_host);             // h(p, q)
  void set_orbital_energies(const double* eps);   // eps(p)
  void set_amplitudes(const double* t_host);      // t1 (v,o) then packed t2
  void get_amplitudes(double* t_host);
  void set_ladder_batch(int64_t nb);              // virtual batch length of the ladder term

  // reference-interface operations
  void construct_t1_cholesky();                   // L_t1 from L_mo and t1
  void construct_fock();                          // T1-transformed Fock matrix
  double calculate_energy();                      // correlation energy
  void construct_omega(Prec prec);                // residual into device buffer
  void get_omega(double* host);
  void get_fock(double* host);
  void set_t2_to_mp2_guess();
  void get_orbital_differences(double* host, int64_t n);
  double omega_norm();                            // ||omega||_2 (FP64 reduction)

  // ground-state amplitude update on the device (FP64): quasi-Newton step and
  // DIIS extrapolation over records kept in HBM
  static constexpr int kMaxDiis = 32;
  void diis_initialize(int64_t dimension, bool crop = false);
  void diis_finalize();
  void quasi_newton_update();                     // e = -omega/eps (doubles diagonal x2); t += e
  void diis_update();                             // records (e, t); t = DIIS-extrapolated t

  // multiplier (Lambda) equations t-bar^T A + eta = 0 on the device
  // (multiplier_equation_ccs.F90:58-94, multiplier_equation_ccsd.F90:41-110)
  void set_initial_multipliers_guess();           // t-bar1 = t1, t-bar2 = 2 t_aibj - t_ajbi
  void set_multipliers(const double* host);
  void get_multipliers(double* host);
  void construct_eta(double* host);               // eta (host may be null: kept in HBM)
  // R = A^T t-bar + eta (host may be null), the left transformation in `prec`
  void construct_multiplier_equation(double* host, Prec prec = Prec::F64);
  double multiplier_residual_norm();
  void multiplier_quasi_newton_update();          // e = -R/eps; t-bar += e
  void multiplier_diis_update();                  // records (e, t-bar); t-bar extrapolated
  // CC mean value <X> of a one-electron operator X (host, MO basis) with the
  // Lagrangian density at the current amplitudes, multipliers and record
  // one-electron density D (n_mo x n_mo, host) with <X> = sum D_pq X_pq
  void one_electron_density(double* D);
  double one_electron_expectation(const double* X_host);

  // Jacobian transformations at the current amplitudes (A = dOmega/dt D2)
  void prepare_for_jacobian();                    // records the residual program
  void jacobian_transform(const double* c_host, double* rho_host);            // rho = A c
  void jacobian_transpose_transform(const double* b_host, double* sigma_host);  // sigma = A^T b
  // n_vec transformations, side 0 (A x) or 1 (A^T x); host or device pointers
  void transform_batch(int side, int64_t n_vec, const double* X, double* R, Prec prec);
  // replay captured HIP graphs of the transformation passes (on by default)
  void set_transform_graphs(bool on) { use_graphs_ = on; }
  // at most `nb` vectors per transformation pass (0: as many as HBM allows, <= 16)
  void set_transform_batch(int64_t nb);
  // core-valence separation for the excited-state solvers (eT's `core excitation`):
  // the Davidson and DIIS excited-state solvers project every transformed vector onto
  // excitations out of these occupied MOs (0-based); n_core = 0 switches it off.
  // remove_core: the opposite projection (eT's `remove core`), onto excitations with
  // no occupied index among them
  void set_cvs(int64_t n_core, const int64_t* core_mos, bool remove_core = false);
  // ionisation (eT's `ionization` with a bath orbital, the last virtual): the
  // excited-state solvers keep only excitations into the bath orbital
  void set_ionization(bool on);
  void project_cvs(double* x, int64_t n_vec);
  int64_t transform_batch() const { return nb_cap_; }
  // bytes of the record's intermediates, and of one trial vector's derivatives
  // (shared slots, and as one buffer per intermediate would need)
  void jacobian_memory(int64_t* record, int64_t* per_vector, int64_t* per_vector_unshared);

  // Davidson eigensolver on the device (eigen_davidson_solver_class.F90:122-318):
  // n_states lowest roots of A (side 0) or A^T (side 1), transforms in `prec`,
  // reduced-space algebra in FP64.  start (n x n_states, host or device) or null
  // for unit vectors on the lowest orbital differences; energies (n_states), vectors (n x n_states,
  // host or device, may be null), residual_norms (max_iterations x n_states, may
  // be null).  Returns the number of iterations.
  int64_t davidson(int side, int64_t n_states, double threshold, int64_t max_iterations,
                   int64_t max_dim, Prec prec, const double* start, double* energies,
                   double* vectors, double* residual_norms);
  // eT's linear Davidson solver (run_linear_davidson_solver,
  // linear_davidson_solver_class.F90:122-270) for A X = F (side 0) or A^T X = F
  // (side 1): start from `start` or F / eps, restart on the solution at max_dim
  // (<= 0: 50), new trials -R / eps, converged when ||A X - F|| <= threshold.  F,
  // start, X: host or device (n_t1 + n_t2); residual_norms (max_iterations, may be
  // null).  Returns the number of iterations.
  int64_t linear_davidson(int side, const double* F, const double* start, double residual_threshold,
                          int64_t max_iterations, int64_t max_dim, Prec prec, double* X_out,
                          double* residual_norms, double lindep = 0.0);
  // eT's Newton-Raphson amplitude update (newton_raphson_updater_class.F90:118-210,
  // as cc_amplitudes_solver_factory_class.F90 sets it up for CCSD): solves A u = -Omega
  // with the linear Davidson solver to ||A u + Omega|| < rel_threshold ||Omega||
  // (start -Omega / eps, reduced space 50), doubles the packed doubles diagonal of u
  // (scale_amplitudes), t += u, and leaves Omega as the error of the next DIIS step
  // (scale_residual off).  Needs the residual of the current amplitudes and
  // prepare_for_jacobian at them.  micro_norms: max_iterations, may be null.
  // Returns the number of linear iterations.
  int64_t newton_raphson_update(double rel_threshold, int64_t max_iterations, Prec prec, double* micro_norms);
  // eT's DIIS excited-state solver (run_diis_cc_es, diis_cc_es_class.F90:260-492):
  // each state iterates X += -(A X - E X)/(eps - E), DIIS over its own records,
  // X normalised, until ||A X - E X|| < residual_threshold and |dE| <
  // energy_threshold (energy_threshold <= 0: residual only).  Start vectors as
  // davidson's.  energies sorted ascending on return, vectors in the same order;
  // residual_norms and energy_history (max_iterations x n_states, iteration-major,
  // may be null) per iteration.  Returns the number of iterations.
  int64_t diis_excited_states(int side, int64_t n_states, double residual_threshold,
                              double energy_threshold, int64_t max_iterations, int64_t diis_dimension,
                              Prec prec, const double* start, double* energies, double* vectors,
                              double* residual_norms, double* energy_history);
  // precision of the first transformation of a Davidson solve (its start vectors);
  // the expansion vectors use the solve's own precision.  With FP64 here and a
  // reduced precision for the solve, every Ritz vector is its FP64-transformed start
  // block plus reduced-precision corrections of the size of the start residuals.
  // start_precision < 0: the solve's precision.
  void set_davidson_start_precision(int start_precision) { dav_start_prec_ = start_precision; }
  // Where the Davidson space V, A V lives: 0 automatic (HBM when both fit in
  // half of the free HBM, else pinned host memory), 1 HBM, 2 pinned host memory.
  void set_davidson_space(int mode) {
    if (mode < 0 || mode > 2) fail("davidson space: mode must be 0 (automatic), 1 (HBM) or 2 (host)");
    dav_space_ = mode;
  }
  // lowest n_want eigenpairs of the k x k matrix G (device, ld ldg): wr, wi (host,
  // k, ascending real part), X (device, k x n_want, ld ldx)
  void reduced_eigen(const double* G_dev, int64_t k, int64_t ldg, int64_t n_want, double* wr,
                     double* wi, double* X_dev, int64_t ldx);

  // device pointers (for callers that keep data in HBM)
  double* d_omega() { return omega_.as<double>(); }
  double* d_amplitudes() { return t_.as<double>(); }
  hipStream_t stream() const { return stream_; }
  void synchronize();

  // flop ledger of the last residual evaluation (sum of 2 m n k over products)
  double last_omega_flops = 0.0;

  // Per-kernel ledger: when on, every product is bracketed by HIP events on
  // the handle's stream; entries are aggregated by kernel instantiation.
  struct LedgerEntry {
    std::string kernel;
    int64_t launches = 0;
    double seconds = 0.0, flops = 0.0, bytes = 0.0;   // bytes: algorithmic operand traffic
  };
  void set_ledger(bool on);
  std::vector<LedgerEntry> ledger();

  // used by the ops layer
  void run_gemm(const Gemm& g, Prec prec);
  // `elems` doubles in the current product workspace (moves count as a workspace move)
  double* gemm_partials(int64_t elems);
  double* scratch(int slot, int64_t elems, const char* name);
  void ladder_primal(const double* t2, const double* LvvT, double* om2, Prec prec);
  // derivatives for nb vectors, vector k at pointer + k * stride
  void ladder_tangent(const double* t2, const double* dt2, int64_t s_dt2, const double* LvvT,
                      const double* dLvvT, int64_t s_dL, double* dom2, int64_t s_dom2, int64_t nb,
                      Prec prec);
  void ladder_adjoint(const double* t2, double* t2b, int64_t s_t2b, const double* LvvT, double* LvvTb,
                      int64_t s_Lb, const double* om2b, int64_t s_om2b, int64_t nb, Prec prec);

 private:
  void require(bool ok, const char* what) const;
  void build_ladder_plan();
  void diis_step(DiisSpace& d, const double* e, double* x);
  // the residual pipeline, written once against Ops
  void program_t1_cholesky(Ops& X);
  void program_fock(Ops& X);
  void program_omega(Ops& X, Prec prec);
  void omega_singles(Ops& X, Prec prec, const double* U);
  void omega_e2_b2(Ops& X, Prec prec, const double* T, const double* U);
  void omega_c2_d2(Ops& X, Prec prec, const double* T, const double* U);

  hipStream_t stream_ = nullptr;
  GemmWorkspace gws_;

  DeviceBuffer Lmo_oo_, Lmo_ov_, Lmo_vo_, Lmo_vv_;
  DeviceBuffer Lt_oo_, Lt_vo_, Lt_vv_, Lt_vvT_;   // L_t1 ov block equals L_mo ov
  DeviceBuffer h_, eps_, F_, t_, omega_, diff_;
  DeviceBuffer g_ovov_;                           // independent of t1 (ov block is not dressed)
  DeviceBuffer red_partial_, red_result_;
  DeviceBuffer small_[4];
  std::vector<DeviceBuffer> work_;                // (ov)^2 scratch slots

  // ladder plan
  int64_t ladder_nb_ = 0;
  uint64_t ladder_gen_ = 0;                       // pack-cache tag generation of the ladder operands
  struct Block { int64_t a0, la, b0, lb, col0, npairs; };
  std::vector<Block> blocks_;
  DeviceBuffer pair_a_, pair_b_, pair_col_, pcol_a_, pcol_b_;
  DeviceBuffer tp_, tm_, wp_, wm_, gblk_, gp_, gm_;
  DeviceBuffer dtp_, dtm_, dwp_, dwm_, lz1_, lz2_;   // per derivative vector, x batch
  int64_t lsz_tp_ = 0, lsz_tm_ = 0, lsz_wp_ = 0, lsz_wm_ = 0;
  int64_t ladder_maxpairs_ = 0;
  DeviceBuffer gpk_, gmk_;                          // split-bf16 packs of g+- (one block)
  int64_t ladder_vector_elems() const;
  bool reserve_ladder_vectors(int64_t nb);
  bool reserve_ladder_packs();
  void ladder_G(const Block& b, const double* LT, Prec prec, bool fp32, uint64_t tag);
  // W_kbij = sum_cd (sum_J L_kc L_t1,bd) t_cidj of the current record (T1 part of
  // the ladder derivative)
  DeviceBuffer lw_;
  bool have_lw_ = false;
  void build_ladder_t1_intermediate();

  // DIIS records: errors and parameters, dim x (n_t1 + n_t2) each
  DiisSpace diis_;                                // ground state and multipliers
  std::vector<DiisSpace> es_diis_;                // one per excited state (DIIS excited-state solver)
  DeviceBuffer update_, es_X_, es_R_, nr_rhs_;
  bool have_update_ = false, have_omega_ = false;

  // Davidson space (V, AV), Ritz block, reduced matrices
  DeviceBuffer dav_V_, dav_AV_, dav_Y_, dav_small_, red_eig_;
  int dav_start_prec_ = -1;                       // < 0: the solve's precision
  int dav_space_ = 0;                             // V, AV placement (set_davidson_space)
 public:
  // seconds in the transformations, seconds in davidson() altogether, vectors
  // transformed, summed over the solves since the last reset
  double dav_timing_[3] = {0.0, 0.0, 0.0};
 private:
  DeviceBuffer cvs_core_;                         // n_o flags: core occupied MOs
  bool cvs_ = false, remove_core_ = false, ionization_ = false;

  // multipliers: t-bar, eta, the multiplier residual and its update vector
  DeviceBuffer tbar_, eta_, mres_, mupdate_;
  bool have_tbar_ = false, have_mres_ = false, have_mupdate_ = false;

  // Jacobian record
  std::unique_ptr<Tape> tape_;
  // captured transformation passes (valid for the record graph_tape_)
  struct TransformGraph {
    int side = 0;
    Prec prec = Prec::F64;
    int64_t nb = 1;                 // vectors per pass
    uint64_t ws_gen = 0;            // gws_pass_ generation the capture saw
    hipGraphExec_t exec = nullptr;
  };
  std::vector<TransformGraph> tgraphs_;
  const Tape* graph_tape_ = nullptr;
  bool use_graphs_ = true, pack_cache_ = true;
  DeviceBuffer jin_, jout_;
  // vectors per transformation pass: bounded by HBM (derivative buffers of the
  // record and the ladder's per-vector work) and by transform_batch_limit_
  static constexpr int64_t kMaxTransformBatch = 16;
  int64_t transform_batch_limit_ = 0;   // 0: automatic
  int64_t nb_cap_ = 0;                  // for the current record (0: not yet sized)
  int64_t reserved_nb_ = 0;             // derivative vectors reserved so far
  int64_t transform_batch_capacity();
  void transform_body(int side, Prec prec, int64_t nb);
  void drop_transform_graphs();
  void reset_record();
  GemmWorkspace gws_pass_;                        // products inside transformation passes
  GemmWorkspace* gws_cur_ = &gws_;
  DeviceBuffer jvec_;

  struct LedgerRec {
    const char* kernel;
    double flops, bytes;
    hipEvent_t start, stop;
  };
  bool ledger_on_ = false;
  std::vector<LedgerRec> ledger_recs_;
  std::vector<hipEvent_t> event_pool_;
  hipEvent_t take_event();

  bool have_L_ = false, have_h_ = false, have_eps_ = false, have_t_ = false, have_F_ = false;
};

// Empties a DIIS space and sizes it for `dimension` records of length n.
void diis_space_initialize(hipStream_t s, DiisSpace& d, int64_t n, int64_t dimension, bool crop = false);

}  // namespace etg
