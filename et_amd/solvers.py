"""Solvers mirroring eT's CC solvers over the device-resident CCSD wavefunction.

DiisCCGroundState: run_diis_cc_gs (src/solvers/cc/diis_cc_gs_class.F90:209-315);
quasi-Newton step and DIIS extrapolation on the device in FP64.
MixedPrecisionGroundState: the same iteration on split-bf16, FP32, then FP64
residuals.
NewtonRaphsonCCGroundState: the same loop with eT's Newton-Raphson amplitude updater
(A u = -Omega by the linear Davidson solver).
DiisCCMultipliers: run_diis_cc_multipliers (diis_cc_multipliers_class.F90:195-299).
MixedPrecisionMultipliers: the same iteration on split-bf16, FP32, then FP64 left
transformations.
DavidsonCCMultipliers: eT's default CCSD multiplier solver, the linear Davidson solver
on A^T t-bar = -eta (cc_multipliers_solver_factory_class.F90:137-230).
EigenDavidson: run_eigen_davidson_solver (eigen_davidson_solver_class.F90:122-318),
with the whole Davidson space on the device.
DiisCCExcitedStates: run_diis_cc_es (diis_cc_es_class.F90:260-492), one DIIS space
per state on the device.
MixedPrecisionDavidson: the same eigensolver on split-bf16 transformations,
finished by an FP64 solve started from the reduced-precision eigenvectors.
Only scalars (energies, norms) reach the host during the iterations.
"""
import numpy as np

from . import ccsd as _ccsd


class DiisCCGroundState:
    """run_diis_cc_gs over a et_amd.ccsd.CCSD wavefunction, every vector in HBM.

    Per iteration (diis_cc_gs_class.F90:253-293): T1 Fock, energy, residual
    (in `precision`), its FP64 norm, then the device quasi-Newton step and DIIS
    extrapolation and the T1 dressing of the Cholesky vectors.  Only scalars
    (energy, norm) reach the host.
    """

    def __init__(self, wf, residual_threshold=1e-5, max_iterations=100, diis_dimension=8,
                 precision=_ccsd.FP64, log=None, crop=False):
        self.wf = wf
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.diis_dimension = diis_dimension
        self.precision = precision
        self.log = log
        self.crop = crop
        self.history = []

    def run(self, restart=False):
        """restart: start from the wavefunction's t file (initialize_files) where it has one."""
        wf = self.wf
        wf.set_initial_amplitudes_guess(restart)
        wf.diis_initialize(self.diis_dimension, self.crop)
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
                wf.diis_finalize()
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
                 schedule=(_ccsd.BF16X3, _ccsd.FP32), margin=10.0, restart_diis=False, log=None):
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
                wf.diis_finalize()
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


class NewtonRaphsonCCGroundState:
    """run_diis_cc_gs with eT's Newton-Raphson amplitude updater (algorithm: newton-raphson,
    cc_amplitudes_solver_factory_class.F90): each iteration forms the T1 Fock matrix, the
    energy and the FP64 residual, tests ||Omega|| (and |dE| when energy_threshold is set;
    the first iteration tests the residual only), records the residual program at the
    current amplitudes, solves A u = -Omega by the linear Davidson solver to
    rel_threshold ||Omega|| (`precision` for its transformations), t += u, then the DIIS
    extrapolation over (Omega, t) and the T1 dressing, all on the device.
    """

    def __init__(self, wf, residual_threshold=1e-5, energy_threshold=None, max_iterations=100,
                 diis_dimension=8, rel_threshold=1e-2, max_micro_iterations=100, precision=_ccsd.FP64, log=None,
                 crop=False):
        self.wf = wf
        self.residual_threshold = residual_threshold
        self.energy_threshold = energy_threshold
        self.max_iterations = max_iterations
        self.diis_dimension = diis_dimension
        self.rel_threshold = rel_threshold
        self.max_micro_iterations = max_micro_iterations
        self.precision = precision
        self.log = log
        self.crop = crop
        self.history = []          # (energy, ||Omega||, |dE|)
        self.micro = []            # residual norms of each update's linear iterations

    def run(self, restart=False):
        wf = self.wf
        wf.set_initial_amplitudes_guess(restart)
        wf.diis_initialize(self.diis_dimension, self.crop)
        previous = 0.0
        for it in range(1, self.max_iterations + 1):
            wf.construct_fock('gs')
            wf.calculate_energy()
            wf.construct_omega(_ccsd.FP64, copy_out=False)
            norm = wf.omega_norm()
            de = abs(wf.energy - previous)
            self.history.append((wf.energy, norm, de))
            if self.log:
                self.log('%3d %25.12f %11.4e %11.4e' % (it, wf.energy, norm, de))
            energy_ok = self.energy_threshold is None or de < self.energy_threshold or it == 1
            if norm < self.residual_threshold and energy_ok:
                wf.diis_finalize()
                return wf.energy
            wf.prepare_for_jacobian()
            self.micro.append(list(wf.newton_raphson_update(self.rel_threshold, self.max_micro_iterations,
                                                             self.precision)))
            wf.diis_update()
            wf.construct_t1_cholesky()
            previous = wf.energy
        raise _ccsd.ETGError('Did not converge in the max number of iterations.')


class DiisCCMultipliers:
    """run_diis_cc_multipliers (diis_cc_multipliers_class.F90:195-299), vectors in HBM.

    t-bar starts at (t1, 2 t_aibj - t_ajbi); each iteration forms R = A^T t-bar + eta
    with the left transformation, tests ||R||, then takes the quasi-Newton step
    t-bar += -R/eps (no diagonal scaling) and the DIIS extrapolation, all FP64 on
    the device.
    """

    def __init__(self, wf, residual_threshold=1e-5, max_iterations=100, diis_dimension=8, log=None, crop=False):
        self.wf = wf
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.diis_dimension = diis_dimension
        self.log = log
        self.crop = crop
        self.history = []

    def run(self, restart=False):
        """restart: start from the wavefunction's tbar file (initialize_files) where it has one."""
        wf = self.wf
        wf.prepare_for_multiplier_equation()
        wf.diis_initialize(self.diis_dimension, self.crop)
        wf.set_initial_multipliers_guess(restart)
        for it in range(1, self.max_iterations + 1):
            wf.construct_multiplier_equation(copy_out=False)
            norm = wf.multiplier_residual_norm()
            self.history.append(norm)
            if self.log:
                self.log('%3d %11.4e' % (it, norm))
            if norm < self.residual_threshold:
                tbar = wf.get_multipliers()
                wf.diis_finalize()
                return tbar
            wf.multiplier_quasi_newton_update()
            wf.multiplier_diis_update()
        raise _ccsd.ETGError('was not able to converge the equations in the given number of '
                             'maximum iterations.')


class DavidsonCCMultipliers:
    """eT's default CCSD multiplier solver (cc_multipliers_solver_factory_class.F90:137-230):
    the linear Davidson solver (linear_davidson_solver_class.F90:122-270) on
    A^T t-bar = -eta, started from the initial multipliers (t1, 2 t_aibj - t_ajbi, or the
    tbar file with `restart`), reduced space at most `max_dim` (eT's default 50); the
    trial space, its transforms and the solution stay on the device."""

    def __init__(self, wf, residual_threshold=1e-5, max_iterations=100, max_dim=50, precision=_ccsd.FP64,
                 log=None):
        self.wf = wf
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.max_dim = max_dim
        self.precision = precision
        self.log = log
        self.history = []

    def run(self, restart=False):
        wf = self.wf
        wf.prepare_for_multiplier_equation()
        wf.set_initial_multipliers_guess(restart)
        start = wf.get_multipliers()
        rhs = -wf.construct_eta()
        tbar, hist = wf.linear_davidson(rhs, 'left', start, self.residual_threshold, self.max_iterations,
                                        self.max_dim, self.precision)
        self.history = list(hist)
        if self.log:
            for it, r in enumerate(hist, 1):
                self.log('%3d %11.4e' % (it, r))
        wf.set_multipliers(tbar)
        return tbar


class MixedPrecisionMultipliers:
    """Multiplier iterations on reduced-precision left transformations, finished on FP64 ones.

    The iteration of DiisCCMultipliers with R = A^T t-bar + eta formed with the left
    transformation's bulk products in the precisions of `schedule`, then in FP64.
    Each reduced form's noise floor ||R_p - R_FP64|| is measured once at the
    starting multipliers; a phase ends when ||R_p|| falls below `margin` times its
    floor or stops halving within four iterations.  The DIIS space carries over,
    and the last phase is run_diis_cc_multipliers unchanged, so convergence is
    decided on the FP64 residual against `residual_threshold`.
    """

    def __init__(self, wf, residual_threshold=1e-5, max_iterations=100, diis_dimension=8,
                 schedule=(_ccsd.BF16X3, _ccsd.FP32), margin=10.0, log=None):
        self.wf = wf
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.diis_dimension = diis_dimension
        self.schedule = tuple(schedule)
        self.margin = margin
        self.log = log
        self.history = []          # (precision, norm)
        self.noise = {}

    def run(self, restart=False):
        wf = self.wf
        wf.prepare_for_multiplier_equation()
        wf.diis_initialize(self.diis_dimension)
        wf.set_initial_multipliers_guess(restart)
        ref = wf.construct_multiplier_equation(precision=_ccsd.FP64)
        for p in self.schedule:
            self.noise[p] = float(np.linalg.norm(wf.construct_multiplier_equation(precision=p) - ref))
        phases = [(p, self.margin * self.noise[p]) for p in self.schedule]
        phases.append((_ccsd.FP64, self.residual_threshold))
        it, phase = 0, 0
        best, since_best = np.inf, 0
        while it < self.max_iterations:
            it += 1
            precision, threshold = phases[phase]
            wf.construct_multiplier_equation(copy_out=False, precision=precision)
            norm = wf.multiplier_residual_norm()
            self.history.append((precision, norm))
            if self.log:
                self.log('%3d  %-6s %11.4e' % (it, _PREC_NAMES[precision], norm))
            if norm < threshold and precision == _ccsd.FP64:
                tbar = wf.get_multipliers()
                wf.diis_finalize()
                return tbar
            if norm < 0.5 * best:
                best, since_best = norm, 0
            else:
                since_best += 1
            advance = precision != _ccsd.FP64 and (norm < threshold or since_best >= 4)
            wf.multiplier_quasi_newton_update()
            wf.multiplier_diis_update()
            if advance:
                phase += 1
                best, since_best = np.inf, 0
        raise _ccsd.ETGError('was not able to converge the equations in the given number of '
                             'maximum iterations.')


_PREC_NAMES = {_ccsd.FP64: 'fp64', _ccsd.FP32: 'fp32', _ccsd.BF16X3: 'bf16x3'}


class EigenDavidson:
    """Lowest eigenpairs of the CC Jacobian, eT's eigen Davidson solver on the device.

    run_eigen_davidson_solver (src/solvers/eigen_davidson_solver_class.F90:122-318):
    start vectors on the lowest orbital differences (with `restart`, the states the
    wavefunction's restart files hold), each new trial vector
    transformed by the Jacobian (right) or its transpose (left), the reduced
    matrix X^T (A X) and its non-symmetric eigenproblem, residuals preconditioned
    by the orbital differences and orthonormalised against the space.  Trial
    vectors, transforms and the reduced-space algebra stay on the device (FP64);
    the transformations run in `precision`.  core_mos (0-based occupied MOs): eT's
    `core excitation`, core-valence separation with eT's CVS start vectors; remove_core:
    eT's `remove core`, the excitations out of those MOs projected away; ionization: eT's
    `ionization` (the last virtual is the bath orbital, et_amd.molecule.with_bath_orbital),
    with eT's start vectors.
    """

    def __init__(self, wf, n_states, side='right', residual_threshold=1e-6, max_iterations=100,
                 max_dim=0, precision=_ccsd.FP64, log=None, start=None, restart=False, core_mos=None,
                 remove_core=None, ionization=False):
        self.wf = wf
        self.ionization = ionization
        self.core_mos = core_mos
        self.remove_core = remove_core
        self.restart = restart
        self.n_states = n_states
        self.side = side
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.max_dim = max_dim
        self.precision = precision
        self.log = log
        self.start = start

    def run(self):
        wf = self.wf
        wf.prepare_for_jacobian(self.side)
        start = self.start
        if start is None and self.restart:
            start = wf.restart_start_vectors(self.n_states, self.side)[0]
        if self.core_mos is not None and start is None:
            start = wf.cvs_start_vectors(self.n_states, self.core_mos)
        wf.set_cvs(self.core_mos if self.core_mos is not None else ())
        if self.remove_core is not None:
            wf.set_remove_core(self.remove_core)
        if self.ionization:
            if start is None:
                start = wf.ip_start_vectors(self.n_states)
            wf.set_ionization(True)
        try:
            w, hist, X = wf.davidson(self.n_states, self.side, self.residual_threshold, self.max_iterations,
                                     self.max_dim, self.precision, vectors=True, start=start)
        finally:
            wf.set_cvs(())
            wf.set_ionization(False)
        self.history = hist
        self.eigenvectors = X
        if self.log:
            for it, row in enumerate(hist, 1):
                self.log('%3d  max|res| %.3e' % (it, row.max()))
        return w


class DiisCCExcitedStates:
    """eT's DIIS excited-state solver (run_diis_cc_es, diis_cc_es_class.F90:260-492).

    Each state's vector X follows X += -(A X - E X)/(eps - E) with E = X . A X and
    DIIS over its own records, normalised every iteration, until its residual (and,
    with energy_threshold, its energy change) is converged; unconverged states share
    a transformation pass.  Start vectors on the lowest orbital differences (or the
    restart files, `restart`); the roots found are those the start vectors lead to.
    Energies are returned sorted.  Vectors, records and the DIIS algebra stay in HBM.
    core_mos, remove_core: core projections as in EigenDavidson.
    """

    def __init__(self, wf, n_states, side='right', residual_threshold=1e-6, energy_threshold=None,
                 max_iterations=100, diis_dimension=20, precision=_ccsd.FP64, start=None, restart=False,
                 log=None, core_mos=None, remove_core=None, ionization=False):
        self.wf = wf
        self.ionization = ionization
        self.core_mos = core_mos
        self.remove_core = remove_core
        self.n_states = n_states
        self.side = side
        self.residual_threshold = residual_threshold
        self.energy_threshold = energy_threshold
        self.max_iterations = max_iterations
        self.diis_dimension = diis_dimension
        self.precision = precision
        self.start = start
        self.restart = restart
        self.log = log

    def run(self):
        wf = self.wf
        wf.prepare_for_jacobian(self.side)
        start = self.start
        if start is None and self.restart:
            start = wf.restart_start_vectors(self.n_states, self.side)[0]
        if self.core_mos is not None and start is None:
            start = wf.cvs_start_vectors(self.n_states, self.core_mos)
        wf.set_cvs(self.core_mos if self.core_mos is not None else ())
        if self.remove_core is not None:
            wf.set_remove_core(self.remove_core)
        if self.ionization:
            if start is None:
                start = wf.ip_start_vectors(self.n_states)
            wf.set_ionization(True)
        try:
            w, hist, X = wf.diis_excited_states(self.n_states, self.side, self.residual_threshold,
                                                self.energy_threshold or 0.0, self.max_iterations,
                                                self.diis_dimension, self.precision, start)
        finally:
            wf.set_cvs(())
            wf.set_ionization(False)
        self.history = hist
        self.energy_history = wf.diis_energy_history
        self.eigenvectors = X
        if self.log:
            for it, row in enumerate(hist, 1):
                self.log('%3d  max|res| %.3e' % (it, row.max()))
        return w


class MixedPrecisionDavidson:
    """EOM roots from reduced-precision transformations, finished against FP64.

    Each precision of `schedule` (split-bf16, optionally then FP32) runs the device
    Davidson with the transformations in that precision down to `margin` (one factor,
    or a dict of factors per precision) times its noise floor ||A_p x - A_FP64 x||
    (measured once on a normalised random probe), or until its residuals stop
    decreasing; its Ritz vectors start the next phase.

    refine=True (the default) then runs the refinement phase: the device Davidson
    started from those Ritz vectors with the start block transformed in FP64 and
    every expansion vector in the last reduced precision
    (etg_set_davidson_start_precision).  Each Ritz vector is then the FP64 image of
    the start block plus reduced-precision corrections whose weights are of the size
    of the start residuals, so the residual falls far below the reduced precision's
    floor.  The result is accepted when an FP64 transformation of every returned
    vector, outside the solver, gives ||A x - w x|| / ||x|| below
    `residual_threshold`; otherwise the phase repeats from the new vectors (at most
    `refine_rounds` times) and an FP64 Davidson solve finishes as without refinement.
    refine=False: the last phase is the FP64 eigen Davidson solver.  Either way
    convergence is decided on FP64 transformations; the reduced-space algebra is
    FP64 throughout.
    """

    def __init__(self, wf, n_states, side='right', residual_threshold=1e-6, max_iterations=100,
                 max_dim=0, schedule=(_ccsd.BF16X3,), margin=10.0, log=None, seed=7, start=None,
                 restart=False, refine=True, refine_rounds=2):
        self.wf = wf
        self.restart = restart
        self.n_states = n_states
        self.side = side
        self.residual_threshold = residual_threshold
        self.max_iterations = max_iterations
        self.max_dim = max_dim
        self.schedule = tuple(schedule) if refine or len(schedule) != 1 else tuple(schedule)
        self.margin = margin
        self.log = log
        self.seed = seed
        self.start = start
        self.refine = refine
        self.refine_rounds = refine_rounds
        self.noise = {}
        self.phases = []          # (precision, iterations, stalled, final max residual)
        self.checks = []          # FP64 residual norms of the refined vectors, per round

    def _measure_noise(self):
        wf = self.wf
        x = np.random.default_rng(self.seed).standard_normal((wf.n_gs_amplitudes, 1))
        x /= np.linalg.norm(x)
        ref = wf.jacobian_transformations(x, self.side, _ccsd.FP64)
        for p in self.schedule:
            self.noise[p] = float(np.linalg.norm(wf.jacobian_transformations(x, self.side, p) - ref))

    def _say(self, p, hist, stalled, tag=''):
        if self.log:
            self.log('%-6s%s %3d iterations%s, max|res| %.3e' % (_PREC_NAMES[p], tag, len(hist),
                                                                 ' (stalled)' if stalled else '', hist[-1].max()))

    def run(self):
        wf = self.wf
        wf.prepare_for_jacobian(self.side)
        self._measure_noise()
        start = self.start
        if start is None and self.restart:
            start = wf.restart_start_vectors(self.n_states, self.side)[0]
        margin = self.margin if isinstance(self.margin, dict) else {p: self.margin for p in self.schedule}
        phases = [(p, max(self.residual_threshold, margin[p] * self.noise[p])) for p in self.schedule]
        w = None
        for p, threshold in phases:
            w, hist, start = wf.davidson(self.n_states, self.side, threshold, self.max_iterations,
                                         self.max_dim, p, vectors=True, start=start)
            stalled = wf.davidson_stalled
            self.phases.append((p, len(hist), stalled, float(hist[-1].max())))
            self._say(p, hist, stalled)
        if self.refine:
            p = self.schedule[-1]
            for _ in range(self.refine_rounds):
                wf.set_davidson_start_precision(_ccsd.FP64)
                try:
                    w, hist, X = wf.davidson(self.n_states, self.side, self.residual_threshold,
                                             self.max_iterations, self.max_dim, p, vectors=True, start=start)
                finally:
                    wf.set_davidson_start_precision(None)
                stalled = wf.davidson_stalled
                self.phases.append(('refined-' + _PREC_NAMES[p], len(hist), stalled, float(hist[-1].max())))
                self._say(p, hist, stalled, ' refined')
                AX = wf.jacobian_transformations(X, self.side, _ccsd.FP64)
                res = np.linalg.norm(AX - X * np.asarray(w)[None, :], axis=0) / np.linalg.norm(X, axis=0)
                self.checks.append(res)
                start = X
                if self.log:
                    self.log('FP64 check of the refined vectors: max|A x - w x| %.3e' % res.max())
                if res.max() < self.residual_threshold:
                    self.eigenvectors = X
                    return w
        w, hist, start = wf.davidson(self.n_states, self.side, self.residual_threshold, self.max_iterations,
                                     self.max_dim, _ccsd.FP64, vectors=True, start=start)
        self.phases.append((_ccsd.FP64, len(hist), False, float(hist[-1].max())))
        self._say(_ccsd.FP64, hist, False)
        self.eigenvectors = start
        return w
