"""CC mean-value properties over the device wavefunction.

dipole_moment follows eT's CC mean-value dipole (cc_mean_values_task_class.F90:88-92,
calculate_and_print_dipole): electronic part -<r_c> with the CC Lagrangian
(Lambda) density, nuclear part sum_A Z_A (R_A - origin), in atomic units and
Debye (eT's factor, CODATA 2018, src/various/parameters.F90:110).
"""
import numpy as np

AU_TO_DEBYE = 2.541746473


def dipole_moment(wf, molecule, C, origin=(0.0, 0.0, 0.0), C_frozen=None):
    """wf: converged ground state with its multipliers and the record of the residual
    program at those amplitudes (prepare_for_jacobian or prepare_for_multiplier_equation);
    molecule: the et_amd.molecule.Molecule it came from; C: its (active) MO coefficients;
    C_frozen: the frozen-core MOs, doubly occupied (their density 2 C_f C_f^T joins the CC
    density, as eT's frozen_CCT does, frozen_orbital_hf.F90:547-600).
    The three components come from one CC density (one backward replay)."""
    o = np.asarray(origin, dtype=float)
    D = molecule.dipole_integrals(o)
    rho = wf.one_electron_density()
    electronic = np.array([-np.sum(rho * (C.T @ D[c] @ C)) for c in range(3)])
    if C_frozen is not None and C_frozen.shape[1]:
        electronic -= np.array([2.0 * np.trace(C_frozen.T @ D[c] @ C_frozen) for c in range(3)])
    nuclear = (molecule.charges[:, None] * (molecule.xyz_bohr - o[None, :])).sum(axis=0)
    total = electronic + nuclear
    return dict(electronic=electronic, nuclear=nuclear, total=total, electronic_debye=electronic * AU_TO_DEBYE,
                nuclear_debye=nuclear * AU_TO_DEBYE, total_debye=total * AU_TO_DEBYE)
