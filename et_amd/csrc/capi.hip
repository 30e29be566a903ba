// extern "C" entry points of libetgpu.so (declared in include/etgpu.h).
// Each entry catches every C++ exception and turns it into a status code
// plus a thread-local message: nothing but plain C types crosses the ABI.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>

#include "../../include/etgpu.h"
#include "ccsd.hpp"
#include "frontend.hpp"

struct etg_handle {
  etg::CCSD* wf;
};

struct etg_molecule {
  etg::Molecule* m;
};

namespace {
thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
  } catch (const std::exception& e) {
    g_last_error = e.what();
  } catch (...) {
    g_last_error = "unknown error";
  }
  return 1;
}

etg::CCSD& wf_of(etg_handle* h) {
  if (h == nullptr || h->wf == nullptr) etg::fail("null etg handle");
  return *h->wf;
}

etg::Molecule& mol_of(etg_molecule* h) {
  if (h == nullptr || h->m == nullptr) etg::fail("null molecule handle");
  return *h->m;
}

etg::Prec prec_of(int p) {
  switch (p) {
    case ETG_FP64: return etg::Prec::F64;
    case ETG_FP32: return etg::Prec::F32;
    case ETG_BF16X3: return etg::Prec::BF16X3;
    default: etg::fail("unknown precision " + std::to_string(p));
  }
}
}  // namespace

extern "C" {

int etg_create(int64_t n_o, int64_t n_v, int64_t n_J, etg_handle** handle) {
  return guarded([&] {
    if (handle == nullptr) etg::fail("etg_create: null handle pointer");
    *handle = nullptr;
    etg::CCSD* wf = new etg::CCSD(n_o, n_v, n_J);
    *handle = new etg_handle{wf};
  });
}

int etg_destroy(etg_handle* handle) {
  return guarded([&] {
    if (handle == nullptr) return;
    delete handle->wf;
    delete handle;
  });
}

const char* etg_last_error(void) { return g_last_error.c_str(); }

int etg_set_cholesky(etg_handle* h, const double* L) {
  return guarded([&] { wf_of(h).set_cholesky(L); });
}

int etg_set_oei(etg_handle* h, const double* oei) {
  return guarded([&] { wf_of(h).set_oei(oei); });
}

int etg_set_orbital_energies(etg_handle* h, const double* eps) {
  return guarded([&] { wf_of(h).set_orbital_energies(eps); });
}

int etg_set_amplitudes(etg_handle* h, const double* t) {
  return guarded([&] { wf_of(h).set_amplitudes(t); });
}

int etg_get_amplitudes(etg_handle* h, double* t) {
  return guarded([&] { wf_of(h).get_amplitudes(t); });
}

int etg_set_t2_to_mp2_guess(etg_handle* h) {
  return guarded([&] { wf_of(h).set_t2_to_mp2_guess(); });
}

int etg_construct_t1_cholesky(etg_handle* h) {
  return guarded([&] {
    wf_of(h).construct_t1_cholesky();
    wf_of(h).synchronize();
  });
}

int etg_construct_fock(etg_handle* h, double* F) {
  return guarded([&] {
    etg::CCSD& wf = wf_of(h);
    wf.construct_fock();
    if (F) wf.get_fock(F);
    wf.synchronize();
  });
}

int etg_calculate_energy(etg_handle* h, double* e) {
  return guarded([&] {
    if (e == nullptr) etg::fail("etg_calculate_energy: null output");
    *e = wf_of(h).calculate_energy();
  });
}

int etg_construct_omega(etg_handle* h, double* omega, int precision) {
  return guarded([&] {
    etg::CCSD& wf = wf_of(h);
    wf.construct_omega(prec_of(precision));
    if (omega) wf.get_omega(omega);
    else wf.synchronize();
  });
}

int etg_get_omega(etg_handle* h, double* omega) {
  return guarded([&] { wf_of(h).get_omega(omega); });
}

int etg_omega_norm(etg_handle* h, double* norm) {
  return guarded([&] {
    if (norm == nullptr) etg::fail("etg_omega_norm: null output");
    *norm = wf_of(h).omega_norm();
  });
}

int etg_get_orbital_differences(etg_handle* h, double* x, int64_t n) {
  return guarded([&] {
    if (x == nullptr) etg::fail("etg_get_orbital_differences: null output");
    wf_of(h).get_orbital_differences(x, n);
  });
}

int etg_set_ladder_batch(etg_handle* h, int64_t batch) {
  return guarded([&] { wf_of(h).set_ladder_batch(batch); });
}

int etg_synchronize(etg_handle* h) {
  return guarded([&] { wf_of(h).synchronize(); });
}

int etg_last_omega_flops(etg_handle* h, double* flops) {
  return guarded([&] {
    if (flops == nullptr) etg::fail("etg_last_omega_flops: null output");
    *flops = wf_of(h).last_omega_flops;
  });
}

int etg_prepare_for_jacobian(etg_handle* h) {
  return guarded([&] { wf_of(h).prepare_for_jacobian(); });
}

int etg_jacobian_transformation(etg_handle* h, const double* c, double* rho) {
  return guarded([&] { wf_of(h).jacobian_transform(c, rho); });
}

int etg_jacobian_transpose_transformation(etg_handle* h, const double* b, double* sigma) {
  return guarded([&] { wf_of(h).jacobian_transpose_transform(b, sigma); });
}

int etg_construct_omega_resident(etg_handle* h, int precision) {
  return guarded([&] { wf_of(h).construct_omega(prec_of(precision)); });
}

int etg_jacobian_transformations(etg_handle* h, int side, int64_t n_vec, const double* X, double* R,
                                 int precision) {
  return guarded([&] { wf_of(h).transform_batch(side, n_vec, X, R, prec_of(precision)); });
}

int etg_davidson(etg_handle* h, int side, int64_t n_states, double residual_threshold,
                 int64_t max_iterations, int64_t max_dim, int precision,
                 const double* start_vectors, double* energies, double* vectors,
                 double* residual_norms, int64_t* iterations) {
  return guarded([&] {
    if (energies == nullptr) etg::fail("etg_davidson: null energies");
    const int64_t it = wf_of(h).davidson(side, n_states, residual_threshold, max_iterations, max_dim,
                                         prec_of(precision), start_vectors, energies, vectors,
                                         residual_norms);
    if (iterations) *iterations = it;
  });
}

int etg_davidson_timing(etg_handle* h, int reset, double* transform_seconds, double* total_seconds,
                        int64_t* transforms) {
  return guarded([&] {
    double* t = wf_of(h).dav_timing_;
    if (transform_seconds) *transform_seconds = t[0];
    if (total_seconds) *total_seconds = t[1];
    if (transforms) *transforms = int64_t(t[2]);
    if (reset) t[0] = t[1] = t[2] = 0.0;
  });
}

int etg_set_davidson_space(etg_handle* h, int mode) {
  return guarded([&] { wf_of(h).set_davidson_space(mode); });
}

int etg_set_davidson_start_precision(etg_handle* h, int precision) {
  return guarded([&] {
    if (precision >= 0) (void)prec_of(precision);
    wf_of(h).set_davidson_start_precision(precision < 0 ? -1 : precision);
  });
}

int etg_linear_davidson(etg_handle* h, int side, const double* rhs, const double* start_vector,
                        double residual_threshold, int64_t max_iterations, int64_t max_dim, int precision,
                        double* solution, double* residual_norms, int64_t* iterations) {
  return guarded([&] {
    const int64_t it = wf_of(h).linear_davidson(side, rhs, start_vector, residual_threshold, max_iterations,
                                                max_dim, prec_of(precision), solution, residual_norms);
    if (iterations) *iterations = it;
  });
}

int etg_set_cvs(etg_handle* h, int64_t n_core, const int64_t* core_mos) {
  return guarded([&] { wf_of(h).set_cvs(n_core, core_mos); });
}

int etg_set_remove_core(etg_handle* h, int64_t n_core, const int64_t* core_mos) {
  return guarded([&] { wf_of(h).set_cvs(n_core, core_mos, true); });
}

int etg_set_ionization(etg_handle* h, int on) {
  return guarded([&] { wf_of(h).set_ionization(on != 0); });
}

int etg_newton_raphson_update(etg_handle* h, double rel_threshold, int64_t max_iterations, int precision,
                              double* micro_norms, int64_t* iterations) {
  return guarded([&] {
    const int64_t it = wf_of(h).newton_raphson_update(rel_threshold, max_iterations, prec_of(precision), micro_norms);
    if (iterations) *iterations = it;
  });
}

int etg_diis_excited_states(etg_handle* h, int side, int64_t n_states, double residual_threshold,
                            double energy_threshold, int64_t max_iterations, int64_t diis_dimension,
                            int precision, const double* start_vectors, double* energies, double* vectors,
                            double* residual_norms, double* energy_history, int64_t* iterations) {
  return guarded([&] {
    const int64_t it = wf_of(h).diis_excited_states(side, n_states, residual_threshold, energy_threshold,
                                                    max_iterations, diis_dimension, prec_of(precision),
                                                    start_vectors, energies, vectors, residual_norms,
                                                    energy_history);
    if (iterations) *iterations = it;
  });
}

int etg_diis_finalize(etg_handle* h) {
  return guarded([&] { wf_of(h).diis_finalize(); });
}

int etg_diis_initialize(etg_handle* h, int64_t dimension) {
  return guarded([&] { wf_of(h).diis_initialize(dimension); });
}

int etg_quasi_newton_update(etg_handle* h) {
  return guarded([&] { wf_of(h).quasi_newton_update(); });
}

int etg_diis_initialize_crop(etg_handle* h, int64_t dimension) {
  return guarded([&] { wf_of(h).diis_initialize(dimension, true); });
}

int etg_diis_update(etg_handle* h) {
  return guarded([&] { wf_of(h).diis_update(); });
}

int etg_prepare_for_multiplier_equation(etg_handle* h) {
  return guarded([&] { wf_of(h).prepare_for_jacobian(); });
}

int etg_set_initial_multipliers_guess(etg_handle* h) {
  return guarded([&] { wf_of(h).set_initial_multipliers_guess(); });
}

int etg_set_multipliers(etg_handle* h, const double* tbar) {
  return guarded([&] { wf_of(h).set_multipliers(tbar); });
}

int etg_get_multipliers(etg_handle* h, double* tbar) {
  return guarded([&] { wf_of(h).get_multipliers(tbar); });
}

int etg_construct_eta(etg_handle* h, double* eta) {
  return guarded([&] { wf_of(h).construct_eta(eta); });
}

int etg_construct_multiplier_equation(etg_handle* h, double* residual, int precision) {
  return guarded([&] { wf_of(h).construct_multiplier_equation(residual, prec_of(precision)); });
}

int etg_multiplier_residual_norm(etg_handle* h, double* norm) {
  return guarded([&] {
    if (norm == nullptr) etg::fail("etg_multiplier_residual_norm: null output");
    *norm = wf_of(h).multiplier_residual_norm();
  });
}

int etg_multiplier_quasi_newton_update(etg_handle* h) {
  return guarded([&] { wf_of(h).multiplier_quasi_newton_update(); });
}

int etg_multiplier_diis_update(etg_handle* h) {
  return guarded([&] { wf_of(h).multiplier_diis_update(); });
}

int etg_one_electron_density(etg_handle* h, double* D) {
  return guarded([&] { wf_of(h).one_electron_density(D); });
}

int etg_one_electron_expectation(etg_handle* h, const double* X_mo, double* value) {
  return guarded([&] {
    if (value == nullptr) etg::fail("etg_one_electron_expectation: null result");
    *value = wf_of(h).one_electron_expectation(X_mo);
  });
}

int etg_set_transform_graphs(etg_handle* h, int on) {
  return guarded([&] { wf_of(h).set_transform_graphs(on != 0); });
}

int etg_set_transform_batch(etg_handle* h, int64_t max_vectors) {
  return guarded([&] { wf_of(h).set_transform_batch(max_vectors); });
}

int etg_jacobian_memory(etg_handle* h, int64_t* record_bytes, int64_t* bytes_per_vector,
                        int64_t* unshared_bytes_per_vector) {
  return guarded([&] { wf_of(h).jacobian_memory(record_bytes, bytes_per_vector, unshared_bytes_per_vector); });
}

int etg_set_kernel_ledger(etg_handle* h, int on) {
  return guarded([&] { wf_of(h).set_ledger(on != 0); });
}

int etg_kernel_ledger(etg_handle* h, int64_t index, char* name, int64_t name_len,
                      int64_t* launches, double* seconds, double* flops, double* bytes,
                      int64_t* count) {
  return guarded([&] {
    const auto entries = wf_of(h).ledger();
    if (count) *count = int64_t(entries.size());
    if (index < 0) return;
    if (index >= int64_t(entries.size())) etg::fail("etg_kernel_ledger: index out of range");
    const auto& e = entries[size_t(index)];
    if (name && name_len > 0) {
      std::strncpy(name, e.kernel.c_str(), size_t(name_len - 1));
      name[name_len - 1] = '\0';
    }
    if (launches) *launches = e.launches;
    if (seconds) *seconds = e.seconds;
    if (flops) *flops = e.flops;
    if (bytes) *bytes = e.bytes;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Diagnostics (not part of the reference interface): direct access to the
// matrix-core product for tests and roofline measurements.
// ---------------------------------------------------------------------------
#include "gemm.hpp"

namespace {
struct DeviceArrays {
  etg::DeviceBuffer a, b, c;
};
}  // namespace

extern "C" {

int etg_test_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                  int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                  int precision) {
  return guarded([&] {
    const int64_t colsA = (ta == 'T' || ta == 't') ? m : k;
    const int64_t colsB = (tb == 'T' || tb == 't') ? k : n;
    DeviceArrays d;
    d.a.resize(sizeof(double) * std::max<int64_t>(1, lda * colsA));
    d.b.resize(sizeof(double) * std::max<int64_t>(1, ldb * colsB));
    d.c.resize(sizeof(double) * std::max<int64_t>(1, ldc * n));
    ETG_HIP(hipMemcpy(d.a.as<double>(), A, sizeof(double) * lda * colsA, hipMemcpyHostToDevice));
    ETG_HIP(hipMemcpy(d.b.as<double>(), B, sizeof(double) * ldb * colsB, hipMemcpyHostToDevice));
    ETG_HIP(hipMemcpy(d.c.as<double>(), C, sizeof(double) * ldc * n, hipMemcpyHostToDevice));
    etg::GemmWorkspace ws;
    etg::Gemm g = etg::make_gemm(ta, tb, m, n, k, alpha, d.a.as<double>(), lda, d.b.as<double>(), ldb, beta,
                                 d.c.as<double>(), ldc);
    g.split_any_shape = precision == ETG_BF16X3_ANY_SHAPE;
    etg::gemm(nullptr, g, g.split_any_shape ? etg::Prec::BF16X3 : prec_of(precision), &ws);
    ETG_HIP(hipDeviceSynchronize());
    ETG_HIP(hipMemcpy(C, d.c.as<double>(), sizeof(double) * ldc * n, hipMemcpyDeviceToHost));
  });
}

int etg_test_reduced_eigen(int64_t k, const double* G, int64_t n_want, double* wr, double* wi,
                           double* X) {
  return guarded([&] {
    if (G == nullptr || wr == nullptr || wi == nullptr || (n_want > 0 && X == nullptr))
      etg::fail("etg_test_reduced_eigen: null pointer");
    if (k < 1 || k > 128) etg::fail("etg_test_reduced_eigen: dimension must be in [1, 128]");
    DeviceArrays d;
    d.a.resize(sizeof(double) * k * k);
    d.b.resize(sizeof(double) * k * std::max<int64_t>(1, n_want));
    ETG_HIP(hipMemcpy(d.a.as<double>(), G, sizeof(double) * k * k, hipMemcpyHostToDevice));
    etg::reduced_eigen_solve(nullptr, d.c, d.a.as<double>(), k, k, n_want, wr, wi, d.b.as<double>(), k);
    if (n_want > 0)
      ETG_HIP(hipMemcpy(X, d.b.as<double>(), sizeof(double) * k * n_want, hipMemcpyDeviceToHost));
  });
}

int etg_reduced_eigen_noalloc(int64_t k, const double* G_device, int64_t ldg, int64_t n_want, double* wr,
                              double* wi, double* X_device, int64_t ldx) {
  return guarded([&] {
    if (G_device == nullptr || wr == nullptr || wi == nullptr || (n_want > 0 && X_device == nullptr))
      etg::fail("etg_reduced_eigen_noalloc: null pointer");
    etg::reduced_eigen_solve_noalloc(nullptr, G_device, k, ldg, n_want, wr, wi, X_device, ldx);
  });
}

// Times `reps` products C = A B (op N,N unless ta/tb say otherwise) on
// device-resident operands filled with a fixed pattern; returns seconds per product.
int etg_bench_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, int precision, int reps,
                   double* seconds) {
  return guarded([&] {
    if (seconds == nullptr || reps < 1) etg::fail("etg_bench_gemm: bad arguments");
    const bool tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
    const int64_t lda = tA ? k : m, ldb = tB ? n : k;
    DeviceArrays d;
    d.a.resize(sizeof(double) * m * k);
    d.b.resize(sizeof(double) * k * n);
    d.c.resize(sizeof(double) * m * n);
    ETG_HIP(hipMemset(d.a.as<double>(), 0, sizeof(double) * m * k));
    ETG_HIP(hipMemset(d.b.as<double>(), 0, sizeof(double) * k * n));
    etg::GemmWorkspace ws;
    etg::Gemm g = etg::make_gemm(ta, tb, m, n, k, 1.0, d.a.as<double>(), lda, d.b.as<double>(), ldb, 0.0,
                                 d.c.as<double>(), m);
    g.split_any_shape = precision == ETG_BF16X3_ANY_SHAPE;
    const etg::Prec prec = g.split_any_shape ? etg::Prec::BF16X3 : prec_of(precision);
    auto run = [&] { etg::gemm(nullptr, g, prec, &ws); };
    run();
    ETG_HIP(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    ETG_HIP(hipEventCreate(&e0));
    ETG_HIP(hipEventCreate(&e1));
    ETG_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) run();
    ETG_HIP(hipEventRecord(e1, nullptr));
    ETG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    ETG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *seconds = double(ms) * 1e-3 / reps;
  });
}

int etg_molecule_create(int64_t n_atoms, const double* charges, const double* xyz_bohr, int64_t n_shells,
                        const int64_t* shell_atom, const int64_t* shell_l, const int64_t* shell_nprim,
                        const double* exponents, const double* coefficients, int64_t n_electrons,
                        etg_molecule** molecule) {
  return guarded([&] {
    if (molecule == nullptr) etg::fail("etg_molecule_create: null handle pointer");
    *molecule = nullptr;
    if (!charges || !xyz_bohr || !shell_atom || !shell_l || !shell_nprim || !exponents || !coefficients)
      etg::fail("etg_molecule_create: null array");
    etg::Molecule* m = new etg::Molecule(n_atoms, charges, xyz_bohr, n_shells, shell_atom, shell_l, shell_nprim,
                                         exponents, coefficients, n_electrons);
    *molecule = new etg_molecule{m};
  });
}

int etg_molecule_run(etg_molecule* molecule, double cholesky_threshold, double scf_threshold,
                     int64_t max_scf_iterations) {
  return guarded([&] { mol_of(molecule).run(cholesky_threshold, scf_threshold, max_scf_iterations); });
}

int etg_molecule_set_orbitals(etg_molecule* molecule, int64_t n_ao, int64_t n_mo, const double* C,
                              const double* eps) {
  return guarded([&] {
    etg::Molecule& m = mol_of(molecule);
    if (n_ao != m.n_ao() || n_mo != m.n_ao())
      etg::fail(n_ao != m.n_ao() ? "Number of AOs does not match" : "Number of MOs does not match");
    m.set_orbitals(C, eps);
  });
}

int etg_molecule_set_cartesian(etg_molecule* molecule, int cartesian) {
  return guarded([&] { mol_of(molecule).set_cartesian(cartesian != 0); });
}

int etg_molecule_set_frozen_core(etg_molecule* molecule, int64_t n_frozen) {
  return guarded([&] { mol_of(molecule).set_frozen_core(n_frozen); });
}

int etg_molecule_dims(etg_molecule* molecule, int64_t* n_ao, int64_t* n_occ, int64_t* n_J,
                      int64_t* scf_iterations) {
  return guarded([&] {
    etg::Molecule& m = mol_of(molecule);
    if (n_ao) *n_ao = m.n_ao();
    if (n_occ) *n_occ = m.n_occ();
    if (n_J) *n_J = m.n_J();
    if (scf_iterations) *scf_iterations = m.scf_iterations();
  });
}

int etg_molecule_results(etg_molecule* molecule, double* e_nuc, double* e_hf, double* eps, double* h_mo,
                         double* L_mo, double* C) {
  return guarded([&] { mol_of(molecule).results(e_nuc, e_hf, eps, h_mo, L_mo, C); });
}

int etg_molecule_one_electron(etg_molecule* molecule, double* S, double* T, double* V) {
  return guarded([&] {
    if (!S || !T || !V) etg::fail("etg_molecule_one_electron: null array");
    mol_of(molecule).one_electron_integrals(S, T, V);
  });
}

int etg_molecule_dipole_integrals(etg_molecule* molecule, const double* origin, double* D) {
  return guarded([&] {
    if (!origin || !D) etg::fail("etg_molecule_dipole_integrals: null array");
    mol_of(molecule).dipole_integrals(origin, D);
  });
}

int etg_molecule_ao_eri(etg_molecule* molecule, double* G) {
  return guarded([&] {
    if (!G) etg::fail("etg_molecule_ao_eri: null array");
    etg::Molecule& m = mol_of(molecule);
    m.build_eri();
    m.copy_eri(G);
  });
}

int etg_molecule_destroy(etg_molecule* molecule) {
  return guarded([&] {
    if (molecule == nullptr) return;
    delete molecule->m;
    delete molecule;
  });
}

int etg_set_gemm_autotune(int on, int64_t* tuned) {
  return guarded([&] {
    etg::set_gemm_autotune(on != 0);
    if (tuned) *tuned = etg::gemm_tuned_shapes();
  });
}

}  // extern "C"
