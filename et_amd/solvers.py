"""Ground-state solver driving the GPU wavefunction, mirroring eT's DIIS CC solver.

Follows run_diis_cc_gs (src/solvers/cc/diis_cc_gs_class.F90:209-315): each
iteration builds the T1 Fock matrix, the energy and the residual on the GPU;
the quasi-Newton step with orbital-energy denominators
(quasi_newton_updater_class.F90:81-118, scaled amplitudes and residual as
amplitude_updater_class.F90:86-120 does) and the DIIS extrapolation
(diis_tool_class.F90:302-760, dimension 8, no CROP) follow, then the
Cholesky vectors are T1-dressed again.  Convergence on ||Omega|| like
convergence_tool%has_converged with energy convergence off
(convergence_tool_class.F90:79-120).
"""
import numpy as np

from . import ccsd as _ccsd


class DIIS:
    """Pulay DIIS with records in host memory (eT's diis_tool with records_in_memory)."""

    def __init__(self, dimension=8):
        self.dimension = dimension
        self.e, self.x = [], []
        self.G = np.zeros((0, 0))

    def update(self, e, x):
        if len(self.e) == self.dimension:
            self.e.pop(0)
            self.x.pop(0)
            self.G = self.G[1:, 1:]
        self.e.append(e.copy())
        self.x.append(x.copy())
        n = len(self.e)
        G = np.zeros((n, n))
        G[:n - 1, :n - 1] = self.G
        for k in range(n):
            G[n - 1, k] = G[k, n - 1] = float(self.e[k] @ e)
        self.G = G
        A = -np.ones((n + 1, n + 1))
        A[:n, :n] = G
        A[n, n] = 0.0
        rhs = np.zeros(n + 1)
        rhs[n] = -1.0
        w = np.linalg.solve(A, rhs)[:n]
        out = np.zeros_like(x)
        for wk, xk in zip(w, self.x):
            out += wk * xk
        return out


class DiisCCGroundState:
    """run_diis_cc_gs over a et_amd.ccsd.CCSD wavefunction."""

    def __init__(self, wf, residual_threshold=1e-5, max_iterations=100, diis_dimension=8,
                 precision=_ccsd.FP64, log=None):
        self.wf = wf
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.diis = DIIS(diis_dimension)
        self.precision = precision
        self.log = log
        self.history = []

    def run(self, restart=False):
        wf = self.wf
        if not restart:
            wf.set_initial_amplitudes_guess()
        denom = wf.get_orbital_differences()
        diag = wf.n_t1 + np.arange(wf.n_t1) * (np.arange(wf.n_t1) + 1) // 2 + np.arange(wf.n_t1)
        previous = 0.0
        for it in range(1, self.max_iterations + 1):
            wf.construct_fock('gs')
            wf.calculate_energy()
            omega = wf.construct_omega(self.precision)
            norm = float(np.linalg.norm(omega))
            self.history.append((wf.energy, norm, abs(wf.energy - previous)))
            if self.log:
                self.log('%3d %25.12f %11.4e %11.4e' % (it, wf.energy, norm, abs(wf.energy - previous)))
            if norm < self.residual_threshold:
                return wf.energy
            amplitudes = wf.get_amplitudes()
            update = -omega / denom
            update[diag] *= 2.0            # scale_amplitudes_ccsd on the update and the residual
            amplitudes += update
            amplitudes = self.diis.update(update, amplitudes)
            wf.set_amplitudes(amplitudes)
            wf.construct_t1_cholesky()
            previous = wf.energy
        raise _ccsd.ETGError('Did not converge in the max number of iterations.')
