// Molecular producer of the CCSD inputs (AO integrals -> RHF -> pivoted
// Cholesky -> MO Cholesky vectors), replacing eT's Libint interface
// (src/libint/eri.cpp:33-89), its RHF solver (src/solvers/hf/scf_solver_class.F90)
// and its AO Cholesky decomposition and AO->MO transformation
// (src/solvers/cholesky/eri_cd_class.F90:241, :3132).
#pragma once

#include <vector>

#include "common.hpp"

namespace etg {

namespace fe {
struct HostShell {
  double c[3];
  int l;
  std::vector<double> exps, coefs;    // coefs normalised for the x^l component
  int f0;                             // first AO function
  bool cart = false;                  // Cartesian functions (else spherical d)
};
}  // namespace fe

class Molecule {
 public:
  // charges and xyz (bohr) per atom; shells in AO order, each on an atom with
  // angular momentum l <= 2 (spherical d) and its primitives (exponents,
  // unnormalised contraction coefficients) consecutive in the flat arrays.
  Molecule(int64_t n_atoms, const double* charges, const double* xyz, int64_t n_shells,
           const int64_t* shell_atom, const int64_t* shell_l, const int64_t* shell_nprim,
           const double* exponents, const double* coefficients, int64_t n_electrons);
  ~Molecule();
  Molecule(const Molecule&) = delete;
  Molecule& operator=(const Molecule&) = delete;

  int64_t n_ao() const { return n_; }
  int64_t n_occ() const { return n_occ_; }
  int64_t n_J() const { return n_J_; }
  int64_t scf_iterations() const { return scf_iterations_; }

  // the whole producer: integrals, decomposition, RHF, MO vectors
  void run(double cholesky_threshold, double scf_threshold, int64_t max_scf_iterations);
  // restart orbitals (n x n column-major C, n orbital energies) for the next run:
  // the RHF iterations are skipped, one Fock build gives E_HF and checks that
  // max|FDS - SDF| < scf_threshold (eT's `skip` of a restarted SCF)
  void set_orbitals(const double* C, const double* eps);
  // frozen core for the next run: h_mo then carries the Coulomb and exchange of the
  // n_frozen lowest (doubly occupied) MOs, F^fc_pq = sum_c 2 (pq|cc) - (pc|cq), so that
  // the MOs above them describe the correlated problem with E_HF unchanged
  void set_frozen_core(int64_t n_frozen);
  // Cartesian Gaussians (eT's `cartesian gaussians`): every d shell keeps its six
  // Cartesian components instead of the five spherical ones; before any integrals
  void set_cartesian(bool cartesian);
  // after run: E_nuc, E_HF, eps (n), h_mo (n x n), L_mo (n_J x n x n, J fastest), C (n x n); any may be null
  void results(double* e_nuc, double* e_hf, double* eps, double* h_mo, double* L_mo, double* C) const;

  // parts, for checks on small molecules
  void one_electron_integrals(double* S, double* T, double* V) const;
  // <mu| r_c - origin_c |nu>, c = x, y, z: 3 consecutive n x n matrices
  void dipole_integrals(const double* origin, double* D) const;
  void build_eri();
  void copy_eri(double* host) const;     // (pq|rs) as an n^2 x n^2 matrix

 private:
  void decompose(double threshold);
  void scf(const double* S, const double* H, double conv, int64_t max_iterations);
  void mo_transform();

  std::vector<fe::HostShell> shells_;
  std::vector<double> Z_, xyz_;
  int64_t n_ = 0, n_occ_ = 0, n_J_ = 0, scf_iterations_ = 0, n_frozen_core_ = 0;
  double e_nuc_ = 0.0, e_hf_ = 0.0;
  bool done_ = false, given_orbitals_ = false;
  hipStream_t stream_ = nullptr;
  DeviceBuffer G_, Lao_, Lmo_;
  std::vector<double> C_, eps_, h_ao_, h_mo_;
};

}  // namespace etg
