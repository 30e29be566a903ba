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
    """run_diis_cc_gs over a et_amd.ccsd.CCSD wavefunction, every vector in HBM.

    Per iteration (diis_cc_gs_class.F90:253-293): T1 Fock, energy, residual
    (in `precision`), its FP64 norm, then the device quasi-Newton step and DIIS
    extrapolation and the T1 dressing of the Cholesky vectors.  Only scalars
    (energy, norm) reach the host.
    """

    def __init__(self, wf, residual_threshold=1e-5, max_iterations=100, diis_dimension=8,
                 precision=_ccsd.FP64, log=None):
        self.wf = wf
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.diis_dimension = diis_dimension
        self.precision = precision
        self.log = log
        self.history = []

    def run(self, restart=False):
        wf = self.wf
        if not restart:
            wf.set_initial_amplitudes_guess()
        wf.diis_initialize(self.diis_dimension)
        previous = 0.0
        for it in range(1, self.max_iterations + 1):
            wf.construct_fock('gs')
            wf.calculate_energy()
            wf.construct_omega(self.precision, copy_out=False)
            norm = wf.omega_norm()
            self.history.append((wf.energy, norm, abs(wf.energy - previous)))
            if self.log:
                self.log('%3d %25.12f %11.4e %11.4e' % (it, wf.energy, norm, abs(wf.energy - previous)))
            if norm < self.residual_threshold:
                return wf.energy
            wf.quasi_newton_update()
            wf.diis_update()
            wf.construct_t1_cholesky()
            previous = wf.energy
        raise _ccsd.ETGError('Did not converge in the max number of iterations.')


class MixedPrecisionGroundState:
    """Amplitude iterations on reduced-precision residuals, finished on FP64 ones.

    The residual is evaluated with the bulk contractions in the precisions of
    `schedule` (e.g. split-bf16, then FP32) and finally in FP64.  Each reduced
    form has a noise floor, ||Omega_p(t) - Omega_FP64(t)||, measured once at the
    starting amplitudes; a phase ends when its residual norm falls below
    `margin` times its floor (below that, its steps follow rounding noise) or
    stops decreasing.  The step is taken with the last residual of a phase and
    the DIIS space carries over (or restarts, `restart_diis`).  The last phase
    is run_diis_cc_gs unchanged, so convergence is decided on the FP64 residual
    norm against `residual_threshold`.  Quasi-Newton steps and DIIS
    extrapolation are FP64 on the device throughout.
    """

    def __init__(self, wf, residual_threshold=1e-10, max_iterations=100, diis_dimension=8,
                 schedule=(_ccsd.BF16X3, _ccsd.FP32), margin=30.0, restart_diis=False, log=None):
        self.wf = wf
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.diis_dimension = diis_dimension
        self.schedule = tuple(schedule)
        self.margin = margin
        self.restart_diis = restart_diis
        self.log = log
        self.history = []          # (precision, energy, norm)
        self.noise = {}

    def _measure_noise(self):
        wf = self.wf
        wf.construct_fock('gs')
        ref = wf.construct_omega(_ccsd.FP64)
        for p in self.schedule:
            self.noise[p] = float(np.linalg.norm(wf.construct_omega(p) - ref))

    def run(self):
        wf = self.wf
        wf.set_initial_amplitudes_guess()
        self._measure_noise()
        phases = [(p, self.margin * self.noise[p]) for p in self.schedule]
        phases.append((_ccsd.FP64, self.residual_threshold))
        it, phase = 0, 0
        best, since_best = np.inf, 0
        wf.diis_initialize(self.diis_dimension)
        while it < self.max_iterations:
            it += 1
            precision, threshold = phases[phase]
            wf.construct_fock('gs')
            wf.calculate_energy()
            wf.construct_omega(precision, copy_out=False)
            norm = wf.omega_norm()
            self.history.append((precision, wf.energy, norm))
            if self.log:
                self.log('%3d  %-6s %25.12f %11.4e' % (it, _PREC_NAMES[precision], wf.energy, norm))
            if norm < threshold and precision == _ccsd.FP64:
                return wf.energy
            if norm < 0.5 * best:
                best, since_best = norm, 0
            else:
                since_best += 1
            # the step is taken with this residual; the next one is evaluated in the
            # next phase's precision once this one is at its noise margin or stalls
            advance = precision != _ccsd.FP64 and (norm < threshold or since_best >= 4)
            wf.quasi_newton_update()
            wf.diis_update()
            wf.construct_t1_cholesky()
            if advance:
                phase += 1
                best, since_best = np.inf, 0
                if self.restart_diis:
                    wf.diis_initialize(self.diis_dimension)
        raise _ccsd.ETGError('Did not converge in the max number of iterations.')


class DiisCCMultipliers:
    """run_diis_cc_multipliers (diis_cc_multipliers_class.F90:195-299), vectors in HBM.

    t-bar starts at (t1, 2 t_aibj - t_ajbi); each iteration forms R = A^T t-bar + eta
    with the left transformation, tests ||R||, then takes the quasi-Newton step
    t-bar += -R/eps (no diagonal scaling) and the DIIS extrapolation, all FP64 on
    the device.
    """

    def __init__(self, wf, residual_threshold=1e-5, max_iterations=100, diis_dimension=8, log=None):
        self.wf = wf
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.diis_dimension = diis_dimension
        self.log = log
        self.history = []

    def run(self):
        wf = self.wf
        wf.prepare_for_multiplier_equation()
        wf.diis_initialize(self.diis_dimension)
        wf.set_initial_multipliers_guess()
        for it in range(1, self.max_iterations + 1):
            wf.construct_multiplier_equation(copy_out=False)
            norm = wf.multiplier_residual_norm()
            self.history.append(norm)
            if self.log:
                self.log('%3d %11.4e' % (it, norm))
            if norm < self.residual_threshold:
                return wf.get_multipliers()
            wf.multiplier_quasi_newton_update()
            wf.multiplier_diis_update()
        raise _ccsd.ETGError('was not able to converge the equations in the given number of '
                             'maximum iterations.')


_PREC_NAMES = {_ccsd.FP64: 'fp64', _ccsd.FP32: 'fp32', _ccsd.BF16X3: 'bf16x3'}


class EigenDavidson:
    """Lowest eigenpairs of the CC Jacobian, after eT's eigen Davidson solver.

    Follows run_eigen_davidson_solver (src/solvers/eigen_davidson_solver_class.F90:122-318):
    start vectors on the lowest orbital differences, each new trial vector
    transformed by the Jacobian (right) or its transpose (left), reduced
    matrix X^T (A X) and its non-symmetric eigenproblem
    (eigen_davidson_tool_class.F90:217-379), residuals preconditioned by the
    orbital differences and Gram-Schmidt orthonormalised against the space
    (davidson_tool_class.F90:354).  Reduced-space linear algebra in FP64.
    """

    def __init__(self, wf, n_states, side='right', residual_threshold=1e-6, max_iterations=100,
                 log=None):
        self.wf = wf
        self.n_states = n_states
        self.side = side
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.log = log

    def run(self):
        wf = self.wf
        wf.prepare_for_jacobian(self.side)
        transform = (wf.jacobian_transformation if self.side == 'right'
                     else wf.jacobian_transpose_transformation)
        diag = wf.get_orbital_differences()
        n = diag.size
        V = np.zeros((n, 0))
        AV = np.zeros((n, 0))
        new = []
        for k in np.argsort(diag, kind='stable')[:self.n_states]:
            e = np.zeros(n)
            e[k] = 1.0
            new.append(e)
        for it in range(1, self.max_iterations + 1):
            for x in new:
                for _ in range(2):
                    x = x - V @ (V.T @ x)
                nx = np.linalg.norm(x)
                if nx < 1e-12:
                    continue
                x = x / nx
                V = np.column_stack([V, x])
                AV = np.column_stack([AV, transform(x)])
            w, S = np.linalg.eig(V.T @ AV)
            order = np.argsort(w.real)[:self.n_states]
            w, S = w[order].real, S[:, order].real
            new, norms = [], []
            for r in range(self.n_states):
                x = V @ S[:, r]
                res = AV @ S[:, r] - w[r] * x
                rn = float(np.linalg.norm(res) / np.linalg.norm(x))
                norms.append(rn)
                if rn > self.residual_threshold:
                    new.append(res / (w[r] - diag + 1e-12))
            if self.log:
                self.log('%3d  %s  max|res| %.3e' % (it, ' '.join('%.10f' % e for e in w), max(norms)))
            if not new:
                self.eigenvectors = V @ S
                return w
        raise _ccsd.ETGError('Davidson did not converge in the max number of iterations.')
