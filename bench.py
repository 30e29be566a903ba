#!/usr/bin/env python3
"""CCSD residual throughput on one MI355X.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--precision fp64] [--no-cpu]

Workload: the synthetic omega-contraction case o = 48, v = 400, n_J = 1600,
seed 0 (Cholesky vectors 0.5 exp(-J/400) N(0,1), symmetric in p,q; packed
t2 ~ 1e-3 N(0,1); t1 = 0; canonical orbital energies).  A step is one full
residual evaluation (T1 Fock, singles and doubles terms, the O^2V^4 ladder
with g_acbd generated on the fly) with all inputs resident in HBM and the
result left in HBM.  The same residual is also timed with FP32 and split-bf16
matrix-core products and compared with the FP64 one.

Time to solution (skip with --no-tts): (H2O)_8 / cc-pVDZ (o = 40, v = 152) built
by the device producer (seeded water lattice, integrals, Cholesky at 1e-8, RHF);
ground state from the MP2 guess to ||Omega_FP64|| <= 1e-10 with FP64 residuals
and with the mixed-precision schedule (split-bf16, FP32, then FP64 residuals);
then 12 EOM roots, right and left Davidson to residuals <= 1e-6, with FP64
transformations and with split-bf16 transformations finished by FP64 Davidson
from their vectors; differences against the FP64 GPU results in Eh and eV.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from et_amd import ccsd  # noqa: E402
from et_amd import flops as flopcount  # noqa: E402
from et_amd import synthetic as syn  # noqa: E402

METRIC = 'CCSD omega evals/s & time-to-solution, mixed vs FP64 GPU at |ΔE|≤1e-8 Eh'
# dense matrix-core peaks; split-bf16 counts its three bf16 products per term
PEAKS = {'fp64': 78.6e12, 'fp32': 157.3e12, 'bf16x3': 2.5e15 / 3}
WORKLOAD = dict(n_o=48, n_v=400, n_J=1600, seed=0)
# the 1-thread CPU residual: the dimensions of (H2O)_8 / cc-pVDZ at Cholesky 1e-8
# (o = 40, v = 152, n_J = 2541), synthetic inputs of that shape
CPU_SINGLE_THREAD = dict(n_o=40, n_v=152, n_J=2541, seed=0)
HARTREE_TO_EV = 27.21138602          # CODATA 2014, the factor eT prints
_PN = {ccsd.BF16X3: 'bf16x3', ccsd.FP32: 'fp32', ccsd.FP64: 'fp64'}
WORKLOAD_NAME = 'omega_o48_v400_J1600_seed0'
# HBM bytes per launch of the residual's dominant kernel on this workload: rocprofv3
# --pmc FETCH_SIZE and --pmc WRITE_SIZE (KiB per dispatch, averaged over the kernel's
# launches in the last of three residual evaluations), each scaled by what the counter reads for a
# plain 1 GiB device copy (bench.py --copy-calibration: FETCH_SIZE x 2.000,
# WRITE_SIZE x 1.000).  Summary and commands: profiles/omega_fp64_o48_v400_J1600_hbm_traffic.md
MEASURED_TRAFFIC = {(WORKLOAD_NAME, 'gemm_mfma16<f64,NN,128x128>'): 1.067e11}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_system(n_o, n_v, n_J, seed):
    t0 = time.time()
    L = syn.cholesky_vectors(seed, n_o + n_v, n_J)
    eps = syn.orbital_energies(seed, n_o, n_v)
    h = syn.canonical_core_hamiltonian(L, eps, n_o)
    t2 = syn.doubles_amplitudes(seed, n_o, n_v, sigma=1e-3)
    log('inputs generated in %.1f s' % (time.time() - t0))
    return L, h, eps, t2


def cpu_omega_seconds(n_o, n_v, n_J, seed, threads, ladder_batch=80):
    """One full residual of the oracle (eT's factorisation restated on the host:
    numpy products on the host BLAS, eT's batched ladder with OpenMP gathers) on the
    synthetic workload, `threads` BLAS/OpenMP threads, after a small warm-up."""
    from threadpoolctl import threadpool_limits
    from oracle import ccsd as O
    with threadpool_limits(limits=threads):
        w = syn.system(1, 6, 30, 60, t2_sigma=1e-3, t1_sigma=0.0, J_decay=20.0, L_scale=0.5)
        Lt = O.t1_cholesky(w['L'], w['t1'], 6)
        O.omega(w['L'], Lt, O.fock_t1(w['h'], Lt, w['t1'], 6), w['t1'], w['t2'], 6, ladder_batch=16)
        L = syn.cholesky_vectors(seed, n_o + n_v, n_J)
        eps = syn.orbital_energies(seed, n_o, n_v)
        h = syn.canonical_core_hamiltonian(L, eps, n_o)
        t2 = syn.doubles_amplitudes(seed, n_o, n_v, sigma=1e-3)
        t1 = np.zeros((n_v, n_o))
        t0 = time.perf_counter()
        Lt = O.t1_cholesky(L, t1, n_o)
        F = O.fock_t1(h, Lt, t1, n_o)
        om = O.omega(L, Lt, F, t1, t2, n_o, ladder_batch=ladder_batch)
        dt = time.perf_counter() - t0
    return dt, float(np.linalg.norm(om))


def cpu_baseline_worker(args):
    """Child process of the bench (no GPU): the timed CPU residuals, as one JSON line."""
    import resource
    w = WORKLOAD
    threads = int(args.cpu_threads)
    dt, norm = cpu_omega_seconds(w['n_o'], w['n_v'], w['n_J'], w['seed'], threads)
    out = {'seconds': dt, 'omega_norm': norm, 'threads': threads}
    if not args.no_cpu_single:
        c2 = CPU_SINGLE_THREAD
        dt1, _ = cpu_omega_seconds(c2['n_o'], c2['n_v'], c2['n_J'], c2['seed'], 1)
        out['single_thread'] = dict(c2, seconds=dt1, omega_per_s=1.0 / dt1)
    out['peak_rss_gb'] = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1e6
    print(json.dumps(out), flush=True)


def start_cpu_baseline(args):
    """Starts the CPU residual timing in a child process before the GPU is touched
    (it runs beside the GPU legs on all host cores but one)."""
    import subprocess
    cores = min(os.cpu_count() or 1, int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1)))
    threads = max(1, cores - 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS=str(threads))
    cmd = [sys.executable, os.path.abspath(__file__), '--cpu-baseline-worker', '--cpu-threads', str(threads)]
    if args.no_cpu_single:
        cmd.append('--no-cpu-single')
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, env=env, text=True), threads


def finish_cpu_baseline(proc, threads):
    outs, _ = proc.communicate()
    if proc.returncode != 0:
        log('CPU baseline failed with status %d' % proc.returncode)
        return None
    r = json.loads(outs.strip().splitlines()[-1])
    w = WORKLOAD
    f_omega = flopcount.omega_flops(w['n_o'], w['n_v'], w['n_J'])
    out = {
        'value': 1.0 / r['seconds'],
        'unit': 'omega evals/s',
        'cores': threads,
        'kind': 'port',
        'sample': ('one full oracle residual (eT factorisation, batched ladder) at o=%d v=%d J=%d seed %d '
                   'after a warm-up: %.1f s on %d threads (host BLAS + OpenMP), %.0f GFLOP/s over F_omega; '
                   'ran beside the GPU legs' % (w['n_o'], w['n_v'], w['n_J'], w['seed'], r['seconds'],
                                                threads, f_omega / r['seconds'] / 1e9)),
        'peak_rss_gb': r.get('peak_rss_gb'),
    }
    if 'single_thread' in r:
        s1 = r['single_thread']
        out['single_thread'] = {'workload': 'omega_o%d_v%d_J%d_seed%d' % (s1['n_o'], s1['n_v'], s1['n_J'], s1['seed']),
                                'omega_evals_per_s': s1['omega_per_s'], 'seconds': s1['seconds'], 'cores': 1}
    log('CPU baseline: %s' % json.dumps(out))
    return out


class DeviceVectors:
    """Column-major float64 arrays in HBM through the HIP runtime the library loaded
    (the same process-wide instance), for calls on device-resident trial vectors."""

    def __init__(self, X):
        import ctypes
        ccsd.load_library()
        with open('/proc/self/maps') as f:
            paths = {ln.split()[-1] for ln in f if 'libamdhip64' in ln and ln.split()[-1].startswith('/')}
        self._hip = hip = ctypes.CDLL(sorted(paths)[0])
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        X = np.asfortranarray(X, dtype=np.float64)
        self.x, self.r = ctypes.c_void_p(), ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(self.x), X.nbytes) != 0 or hip.hipMalloc(ctypes.byref(self.r), X.nbytes) != 0:
            raise RuntimeError('hipMalloc of the trial vectors failed')
        if hip.hipMemcpy(self.x, X.ctypes.data, X.nbytes, 1) != 0:
            raise RuntimeError('hipMemcpy of the trial vectors failed')

    def close(self):
        self._hip.hipFree(self.x)
        self._hip.hipFree(self.r)


def jacobian_throughput(wf):
    """Right and left Jacobian transformations of one trial vector on the same workload.

    The residual program is recorded once (etg_prepare_for_jacobian); a right
    transformation replays it forward, a left one backward with the g_acbd
    blocks recomputed.  Each rate is the second call of its (side, precision) on a
    trial vector resident in HBM (the first sizes the buffers and captures the
    pass); the same call on host arrays (two PCIe copies of n_t1 + n_t2 doubles
    of pageable memory) is reported beside it.
    """
    t0 = time.perf_counter()
    wf.prepare_for_jacobian()
    prep = time.perf_counter() - t0
    X = np.random.default_rng(5).standard_normal((wf.n_gs_amplitudes, 1))
    dv = DeviceVectors(X)
    out = {'prepare_s': prep}
    for side in ('right', 'left'):
        for prec, pname in ((ccsd.FP64, 'fp64'), (ccsd.BF16X3, 'split-bf16x3')):
            wf.jacobian_transformations_at(dv.x.value, dv.r.value, 1, side, prec)
            wf.synchronize()
            t0 = time.perf_counter()
            wf.jacobian_transformations_at(dv.x.value, dv.r.value, 1, side, prec)
            wf.synchronize()
            dt = time.perf_counter() - t0
            t0 = time.perf_counter()
            wf.jacobian_transformations(X, side, precision=prec)
            dt_host = time.perf_counter() - t0
            out['%s_%s_transforms_per_s' % (side, pname)] = 1.0 / dt
            out['%s_%s_transforms_per_s_host_vectors' % (side, pname)] = 1.0 / dt_host
            log('%s transformation (%s): %.3f s (%.3f s with host vectors)' % (side, pname, dt, dt_host))
    dv.close()
    return out


def jacobian_sweep(r, n_vecs=(1, 7, 12, 24, 36)):
    """Batched Jacobian transformations on the system r after an FP64 ground state:
    for each side, precision and trial-vector count (7: an uneven group), the wall
    time of one etg_jacobian_transformations call on trial vectors resident in HBM
    (graph replays; as the Davidson solver calls it) and on host vectors (two PCIe
    copies of pageable memory per vector included), and from the per-kernel ledger
    of an eager pass of the same call the products' FLOPs per transformation and
    their rate against the matrix-core roofline of the precision."""
    from et_amd import solvers
    n_o = r['n_o']
    n_v = r['L_mo'].shape[1] - n_o
    out = {'workload': 'h2o8_ccpvdz_seed0', 'n_o': n_o, 'n_v': n_v, 'n_J': r['n_J'], 'rows': []}
    with ccsd.CCSD(n_o, n_v, r['n_J']) as wf:
        load(wf, r)
        solvers.DiisCCGroundState(wf, residual_threshold=1e-8).run()
        rng = np.random.default_rng(11)
        X = rng.standard_normal((wf.n_gs_amplitudes, max(n_vecs)))
        dv = DeviceVectors(X)                                  # the same vectors in HBM
        for side in ('right', 'left'):
            wf.prepare_for_jacobian(side)
            for prec, pname in ((ccsd.FP64, 'fp64'), (ccsd.BF16X3, 'bf16x3')):
                for nv in n_vecs:
                    wf.jacobian_transformations_at(dv.x.value, dv.r.value, nv, side, prec)  # sizes, captures
                    wf.synchronize()
                    t0 = time.perf_counter()
                    wf.jacobian_transformations_at(dv.x.value, dv.r.value, nv, side, prec)
                    wf.synchronize()
                    dt = time.perf_counter() - t0
                    t0 = time.perf_counter()
                    wf.jacobian_transformations(X[:, :nv], side, prec)
                    wf.synchronize()
                    dt_host = time.perf_counter() - t0
                    wf.set_kernel_ledger(True)
                    wf.jacobian_transformations_at(dv.x.value, dv.r.value, nv, side, prec)
                    ledger = wf.kernel_ledger()
                    wf.set_kernel_ledger(False)
                    fl = sum(e[3] for e in ledger)
                    sec = sum(e[2] for e in ledger)
                    row = {'side': side, 'precision': pname, 'n_vec': nv,
                           'transforms_per_s': nv / dt, 'ms_per_transform': dt / nv * 1e3,
                           'ms_per_transform_host_vectors': dt_host / nv * 1e3,
                           'product_flops_per_transform': fl / nv,
                           'product_tflops': fl / sec / 1e12 if sec > 0 else None,
                           'product_frac': fl / sec / PEAKS[pname] if sec > 0 else None,
                           'product_ms_per_transform': sec / nv * 1e3}
                    out['rows'].append(row)
                    log('jacobian %s %s n_vec %d: %.1f ms per transform (%.1f with host vectors), products '
                        '%.1f TF/s (%.0f %% of peak)'
                        % (side, pname, nv, row['ms_per_transform'], row['ms_per_transform_host_vectors'],
                           row['product_tflops'] or 0, 100 * (row['product_frac'] or 0)))
        dv.close()
    return out


def water_octamer(seed=0, cholesky_threshold=1e-8):
    """(H2O)_8 / cc-pVDZ (o = 40, v = 152): eight waters on a 2 x 2 x 2 lattice of
    2.9 Angstrom with seeded uniform rotations (1.5 Angstrom rejection), through
    the device producer (integrals, Cholesky at 1e-8, RHF)."""
    from et_amd import molecule
    t0 = time.perf_counter()
    with molecule.Molecule(molecule.water_lattice((2, 2, 2), 2.9, seed)) as m:
        r = m.run(cholesky_threshold=cholesky_threshold, scf_threshold=1e-10)
    r['front_end_s'] = time.perf_counter() - t0
    log('(H2O)8 producer: %.2f s, n_J %d, E_HF %.10f, %d RHF iterations'
        % (r['front_end_s'], r['n_J'], r['e_hf'], r['scf_iterations']))
    return r


def load(wf, r):
    wf.set_cholesky(r['L_mo'])
    wf.set_oei(r['h_mo'])
    wf.set_orbital_energies(r['eps'])


def time_to_solution(r, threshold=1e-10):
    """Ground state from the MP2 guess to ||Omega_FP64|| <= threshold, FP64 vs mixed
    precision (split-bf16, FP32, then FP64 residuals), on the system r."""
    from et_amd import solvers
    n_o = r['n_o']
    n_v = r['L_mo'].shape[1] - n_o
    out = {'workload': 'h2o8_ccpvdz_seed0', 'n_o': n_o, 'n_v': n_v, 'n_J': r['n_J'], 'e_hf': r['e_hf'],
           'residual_threshold': threshold, 'producer_s': r['front_end_s']}
    with ccsd.CCSD(n_o, n_v, r['n_J']) as wf:
        load(wf, r)
        wf.set_initial_amplitudes_guess()
        for p in (ccsd.FP64, ccsd.FP32, ccsd.BF16X3):        # size workspaces, choose tiles once
            wf.construct_omega(p, copy_out=False)
        results = {}
        for name in ('fp64', 'mixed'):
            wf.synchronize()
            t0 = time.perf_counter()
            if name == 'fp64':
                sv = solvers.DiisCCGroundState(wf, residual_threshold=threshold)
                sv.run()
                iters = {'fp64': len(sv.history)}
            else:
                sv = solvers.MixedPrecisionGroundState(wf, residual_threshold=threshold)
                sv.run()
                iters = {}
                for p, _, _ in sv.history:
                    k = {ccsd.FP64: 'fp64', ccsd.FP32: 'fp32', ccsd.BF16X3: 'bf16x3'}[p]
                    iters[k] = iters.get(k, 0) + 1
            wf.synchronize()
            dt = time.perf_counter() - t0
            results[name] = {'seconds': dt, 'iterations': iters, 'correlation_energy': wf.correlation_energy}
            log('time to solution %s: %.2f s, %s, E_corr %.12f' % (name, dt, iters, wf.correlation_energy))
    results['mixed']['abs_err_vs_fp64_gpu'] = abs(results['mixed']['correlation_energy']
                                                  - results['fp64']['correlation_energy'])
    out.update(results)
    out['speedup_mixed_vs_fp64'] = results['fp64']['seconds'] / results['mixed']['seconds']
    return out


def eom_time_to_solution(r, n_states=12, threshold=1e-5):
    """12 EOM roots, right and left Davidson, FP64 vs split-bf16, FP32 then FP64
    transformations, on the system r after an FP64 ground state; excitation
    energies compared in eV (CODATA 2014 factor).  Residual threshold 1e-5 (eT's
    default is 1e-3, cc_es_amplitudes_solver_factory_class.F90:176): at 1e-4 the
    two runs' energies still differ by ~3e-5 eV, above the 1e-5 eV target."""
    from et_amd import solvers
    n_o = r['n_o']
    n_v = r['L_mo'].shape[1] - n_o
    out = {'workload': 'h2o8_ccpvdz_seed0', 'n_o': n_o, 'n_v': n_v, 'n_J': r['n_J'],
           'n_states': n_states, 'residual_threshold': threshold}
    vectors = {}
    with ccsd.CCSD(n_o, n_v, r['n_J']) as wf:
        load(wf, r)
        solvers.DiisCCGroundState(wf, residual_threshold=1e-10).run()
        for side in ('right', 'left'):
            wf.prepare_for_jacobian(side)
            # size the group buffers and choose tiles for every group size before timing
            ones = np.ones((wf.n_gs_amplitudes, n_states))
            for p in (ccsd.FP64, ccsd.FP32, ccsd.BF16X3):
                for nv in (n_states, 8, 4, 2, 1):
                    wf.jacobian_transformations(ones[:, :nv], side, p)
            res = {}
            for name in ('fp64', 'mixed'):
                wf.synchronize()
                wf.davidson_timing(reset=True)
                t0 = time.perf_counter()
                if name == 'fp64':
                    sv = solvers.EigenDavidson(wf, n_states, side, residual_threshold=threshold)
                    w = sv.run()
                    iters = {'fp64': len(sv.history)}
                else:
                    sv = solvers.MixedPrecisionDavidson(wf, n_states, side, residual_threshold=threshold)
                    w = sv.run()
                    names = {ccsd.FP64: 'fp64', ccsd.BF16X3: 'bf16x3', ccsd.FP32: 'fp32'}
                    iters = {}
                    for p, k, _, _ in sv.phases:
                        iters[names.get(p, p)] = iters.get(names.get(p, p), 0) + k
                    phases = [{'precision': names.get(p, p), 'iterations': k, 'stalled': st, 'max_residual': rr}
                              for p, k, st, rr in sv.phases]
                wf.synchronize()
                dt = time.perf_counter() - t0
                tr_s, dav_s, n_tr = wf.davidson_timing(reset=True)
                res[name] = {'seconds': dt, 'iterations': iters, 'energies_ev': [float(x) * HARTREE_TO_EV for x in w],
                             'davidson_seconds': dav_s, 'transformation_seconds': tr_s, 'transformations': n_tr}
                if name == 'mixed':
                    res[name]['phases'] = phases
                    res[name]['noise'] = {names[p]: v for p, v in sv.noise.items()}
                    res[name]['fp64_checks_max_residual'] = [float(c.max()) for c in sv.checks]
                    vectors[side] = sv.eigenvectors
                log('EOM %s %s: %.2f s, %s; inside the Davidson solves %.2f s, of which %d transformations %.2f s'
                    % (side, name, dt, iters, dav_s, n_tr, tr_s))
            res['mixed']['max_abs_err_vs_fp64_gpu_ev'] = float(
                np.abs(np.array(res['mixed']['energies_ev']) - np.array(res['fp64']['energies_ev'])).max())
            res['speedup_mixed_vs_fp64'] = res['fp64']['seconds'] / res['mixed']['seconds']
            log('EOM %s: max |w_mixed - w_fp64| %.2e eV' % (side, res['mixed']['max_abs_err_vs_fp64_gpu_ev']))
            out[side] = res
        # Certification against the two-sided Rayleigh quotient of the mixed-precision
        # right and left vectors, w2 = l^T A r / l^T r (one FP64 transformation of the
        # right vectors): its error is of second order in the two residuals (~1e-10 Eh
        # here), so it stands for the converged FP64 eigenvalue.
        R, Lv = vectors['right'], vectors['left']
        wf.prepare_for_jacobian('right')
        AR = wf.jacobian_transformations(R, 'right', ccsd.FP64)
        w2 = np.einsum('ij,ij->j', Lv, AR) / np.einsum('ij,ij->j', Lv, R)
        w2_ev = w2 * HARTREE_TO_EV
        cert = {'two_sided_energies_ev': w2_ev.tolist()}
        for side in ('right', 'left'):
            for name in ('fp64', 'mixed'):
                cert['max_abs_%s_%s_vs_two_sided_ev' % (name, side)] = float(
                    np.abs(np.array(out[side][name]['energies_ev']) - w2_ev).max())
        out['two_sided_check'] = cert
        log('EOM two-sided check (eV): %s' % json.dumps({k: v for k, v in cert.items() if k.startswith('max')}))
    return out


def transforms_profile(side, precision, system='h2o_he_x4', n_vec=12):
    """One group of n_vec right or left transformations (eager passes, graphs off)
    after a short FP64 ground state, timed after a first group that sizes the
    buffers: the launch sequence the EOM Davidson replays, for rocprofv3 kernel
    traces.  system: 'h2o_he_x4' (four H2O-He copies from the committed inputs) or
    'h2o8' ((H2O)_8 / cc-pVDZ from the device producer)."""
    from et_amd import solvers, workloads
    prec = ccsd.PRECISIONS[precision]
    if system == 'h2o8':
        r = water_octamer()
    else:
        r = workloads.replicate(workloads.h2o_he_inputs(), 4)
    n_o = r['n_o']
    n_v = r['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, r['n_J']) as wf:
        wf.set_cholesky(r['L_mo'])
        wf.set_oei(r['h_mo'])
        wf.set_orbital_energies(r['eps'])
        del r
        solvers.DiisCCGroundState(wf, residual_threshold=1e-4).run()
        wf.set_transform_graphs(False)
        wf.prepare_for_jacobian(side)
        X = np.random.default_rng(0).standard_normal((wf.n_gs_amplitudes, n_vec))
        wf.jacobian_transformations(X, side, prec)            # sizes the buffers
        wf.synchronize()
        t0 = time.perf_counter()
        wf.jacobian_transformations(X, side, prec)
        wf.synchronize()
        dt = time.perf_counter() - t0
    print(json.dumps({'transforms_profile': side, 'precision': precision, 'system': system, 'n_vec': n_vec,
                      'ms_per_transform': dt / n_vec * 1e3}), flush=True)


def eom_profile(side, precision, n_states=12, threshold=1e-5):
    """One 12-root EOM solve on (H2O)_8 / cc-pVDZ after an FP64 ground state, with the
    transformation passes launched eagerly (graphs off) so that a kernel trace sees
    every kernel: FP64 Davidson ('fp64') or the mixed-precision solver ('mixed')."""
    from et_amd import solvers
    r = water_octamer()
    n_o = r['n_o']
    n_v = r['L_mo'].shape[1] - n_o
    with ccsd.CCSD(n_o, n_v, r['n_J']) as wf:
        load(wf, r)
        del r
        solvers.DiisCCGroundState(wf, residual_threshold=1e-10).run()
        wf.set_transform_graphs(False)
        wf.prepare_for_jacobian(side)
        wf.davidson_timing(reset=True)
        t0 = time.perf_counter()
        if precision == 'mixed':
            solvers.MixedPrecisionDavidson(wf, n_states, side, residual_threshold=threshold).run()
        else:
            solvers.EigenDavidson(wf, n_states, side, residual_threshold=threshold).run()
        dt = time.perf_counter() - t0
        tr_s, dav_s, n_tr = wf.davidson_timing()
    print(json.dumps({'eom_profile': side, 'precision': precision, 'seconds': dt, 'davidson_seconds': dav_s,
                      'transformation_seconds': tr_s, 'transformations': n_tr}), flush=True)


def lambda_dipole(shape=(4, 2, 2), seed=0, gs_threshold=1e-10, lambda_threshold=1e-8, fp64=False):
    """Ground state, Lambda equations and the CC-Lagrangian dipole of a seeded water
    lattice ((H2O)_16 / cc-pVDZ for the 4 x 2 x 2 default: o = 80, v = 304): the
    device producer, then split-bf16 -> FP32 -> FP64 ground-state and multiplier
    iterations (FP64 decides convergence), then <mu> from one tangent replay.  With
    fp64, the same ground state and multipliers with FP64 products throughout, and
    the differences."""
    from et_amd import molecule, properties, solvers
    n_mol = shape[0] * shape[1] * shape[2]
    out = {'workload': 'h2o%d_ccpvdz_seed%d' % (n_mol, seed), 'gs_threshold': gs_threshold,
           'lambda_threshold': lambda_threshold}
    t0 = time.perf_counter()
    mol = molecule.Molecule(molecule.water_lattice(shape, 2.9, seed))
    r = mol.run(cholesky_threshold=1e-8, scf_threshold=1e-10)
    n_o = r['n_o']
    n_v = r['L_mo'].shape[1] - n_o
    out.update(n_o=n_o, n_v=n_v, n_J=r['n_J'], e_hf=r['e_hf'], producer_s=time.perf_counter() - t0)
    log('(H2O)%d producer: %.1f s, o %d v %d n_J %d, E_HF %.10f' % (n_mol, out['producer_s'], n_o, n_v,
                                                                   r['n_J'], r['e_hf']))
    runs = [('mixed', True)] + ([('fp64', False)] if fp64 else [])
    tbar_mixed = None
    for name, mixed in runs:
        res = {}
        with ccsd.CCSD(n_o, n_v, r['n_J']) as wf:
            load(wf, r)
            t0 = time.perf_counter()
            if mixed:
                gs = solvers.MixedPrecisionGroundState(wf, residual_threshold=gs_threshold, log=log)
                gs.run()
                res['gs_iterations'] = {_PN[p]: sum(1 for h in gs.history if h[0] == p) for p in _PN}
            else:
                gs = solvers.DiisCCGroundState(wf, residual_threshold=gs_threshold, log=log)
                gs.run()
                res['gs_iterations'] = {'fp64': len(gs.history)}
            wf.synchronize()
            res['gs_s'] = time.perf_counter() - t0
            res['correlation_energy'] = wf.correlation_energy
            log('%s ground state: %.1f s, E_corr %.12f' % (name, res['gs_s'], wf.correlation_energy))
            t0 = time.perf_counter()
            if mixed:
                lam = solvers.MixedPrecisionMultipliers(wf, residual_threshold=lambda_threshold, log=log)
                tbar = lam.run()
                res['lambda_iterations'] = {_PN[p]: sum(1 for h in lam.history if h[0] == p) for p in _PN}
                res['lambda_noise'] = {_PN[p]: v for p, v in lam.noise.items()}
            else:
                lam = solvers.DiisCCMultipliers(wf, residual_threshold=lambda_threshold, log=log)
                tbar = lam.run()
                res['lambda_iterations'] = {'fp64': len(lam.history)}
            wf.synchronize()
            res['lambda_s'] = time.perf_counter() - t0
            res['jacobian_memory_gb'] = {k: v / 1e9 for k, v in wf.jacobian_memory().items()}
            t0 = time.perf_counter()
            mu = properties.dipole_moment(wf, mol, r['C'])
            res['dipole_s'] = time.perf_counter() - t0
            res['dipole_debye'] = [float(x) for x in mu['total_debye']]
            log('%s Lambda: %.1f s; dipole %s D (%.1f s)' % (name, res['lambda_s'], res['dipole_debye'],
                                                             res['dipole_s']))
        if tbar_mixed is None:
            tbar_mixed = tbar
        else:
            res['max_abs_tbar_diff_vs_mixed'] = float(np.abs(tbar - tbar_mixed).max())
        out[name] = res
    mol.close()
    if fp64:
        out['abs_e_corr_diff'] = abs(out['mixed']['correlation_energy'] - out['fp64']['correlation_energy'])
        out['max_abs_dipole_diff_debye'] = max(abs(a - b) for a, b in zip(out['mixed']['dipole_debye'],
                                                                         out['fp64']['dipole_debye']))
        for k in ('gs_s', 'lambda_s'):
            out['speedup_' + k[:-2]] = out['fp64'][k] / out['mixed'][k]
    return out


def copy_calibration(nbytes=1 << 30, reps=3):
    """Plain copies of known size: each reads and writes `nbytes` of HBM once.

    Run under rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) to scale the HBM
    counters of the residual's kernels.
    """
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({'copy_bytes': nbytes, 'copies': reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--precision', choices=['fp64', 'fp32', 'bf16x3'], default='fp64')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-mixed', action='store_true')
    ap.add_argument('--no-tts', action='store_true', help='skip the time-to-solution comparison')
    ap.add_argument('--no-jacobian', action='store_true', help='skip the Jacobian transformation rates')
    ap.add_argument('--transforms-profile', nargs=2, metavar=('SIDE', 'PRECISION'),
                    help='only run two groups of 12 EOM transformations (right|left, fp64|fp32|bf16x3) '
                         'on four H2O-He copies, graphs off (for kernel traces)')
    ap.add_argument('--transforms-system', choices=['h2o_he_x4', 'h2o8'], default='h2o_he_x4',
                    help='system of --transforms-profile')
    ap.add_argument('--eom-profile', nargs=2, metavar=('SIDE', 'PRECISION'),
                    help='one 12-root EOM solve on (H2O)_8 with eager passes (fp64 or mixed), for kernel traces')
    ap.add_argument('--lambda-dipole', action='store_true',
                    help='only run the (H2O)16 / cc-pVDZ ground state, Lambda equations and dipole '
                         '(mixed precision; a few minutes)')
    ap.add_argument('--fp64-reference', action='store_true',
                    help='with --lambda-dipole: also the FP64 ground state and Lambda, and the differences')
    ap.add_argument('--cpu-baseline-worker', action='store_true', help=argparse.SUPPRESS)
    ap.add_argument('--cpu-threads', type=int, default=1, help=argparse.SUPPRESS)
    ap.add_argument('--no-cpu-single', action='store_true',
                    help='skip the 1-thread CPU residual at the (H2O)8 dimensions')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only run three 1 GiB device-to-device copies (HBM counter calibration)')
    args = ap.parse_args()
    if args.cpu_baseline_worker:
        cpu_baseline_worker(args)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.transforms_profile:
        transforms_profile(*args.transforms_profile, system=args.transforms_system)
        return
    if args.eom_profile:
        eom_profile(*args.eom_profile)
        return
    if args.lambda_dipole:
        print(json.dumps({'lambda_dipole': lambda_dipole(fp64=args.fp64_reference)}), flush=True)
        return
    if args.gpus != 1:
        print('bench.py measures one GPU (a single MI355X); --gpus %d is not supported: the '
              'ladder and ring products would shard over GPUs by output blocks, which this '
              'library does not implement.' % args.gpus, file=sys.stderr)
        sys.exit(2)

    cpu_job = None if args.no_cpu else start_cpu_baseline(args)
    w = WORKLOAD
    n_o, n_v, n_J = w['n_o'], w['n_v'], w['n_J']
    L, h, eps, t2 = build_system(n_o, n_v, n_J, w['seed'])
    wf = ccsd.CCSD(n_o, n_v, n_J)
    t0 = time.time()
    wf.set_cholesky(L)
    wf.set_oei(h)
    wf.set_orbital_energies(eps)
    t = np.zeros(wf.n_gs_amplitudes)
    t[wf.n_t1:] = t2
    wf.set_amplitudes(t)
    wf.construct_t1_cholesky()
    wf.construct_fock()
    del L
    log('uploaded in %.1f s' % (time.time() - t0))

    prec = ccsd.PRECISIONS[args.precision]
    for _ in range(args.warmup):
        wf.construct_omega(prec, copy_out=False)
    wf.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        wf.construct_omega(prec, copy_out=False)
        log('step %d done' % k)
    wf.synchronize()
    dt = (time.perf_counter() - t0) / args.steps

    # per-kernel ledger over one more evaluation: events on the library's stream
    wf.set_kernel_ledger(True)
    wf.construct_omega(prec, copy_out=False)
    ledger = wf.kernel_ledger()
    wf.set_kernel_ledger(False)
    dom = max(ledger, key=lambda e: e[2])
    traffic = MEASURED_TRAFFIC.get((WORKLOAD_NAME, dom[0])) if args.precision == 'fp64' else None
    peak = PEAKS[args.precision]
    f_omega = flopcount.omega_flops(n_o, n_v, n_J)
    roofline = {
        'bound': 'mfma',
        'kernel': dom[0],
        'launches': dom[1],
        'achieved': dom[3] / dom[2] / 1e12,
        'peak': peak / 1e12,
        'unit': 'TFLOP/s',
        'frac': dom[3] / dom[2] / peak,
        'traffic': traffic,
        'algorithmic_bytes': dom[4] / dom[1],
        'avg_launch_ms': dom[2] / dom[1] * 1e3,
        'omega_achieved': f_omega / dt / 1e12,
        'omega_frac': f_omega / dt / peak,
    }
    log('ledger:', json.dumps([(e[0], e[1], round(e[2], 4), '%.3e' % e[3], '%.3e' % e[4]) for e in ledger]))

    extra = {}
    if not args.no_jacobian and args.precision == 'fp64':
        extra['jacobian'] = jacobian_throughput(wf)
    if not args.no_mixed and args.precision == 'fp64':
        # the same residual in the reduced-precision forms, against the FP64 one
        ref = wf.get_omega()
        forms = []
        for name, p in (('fp32-mfma', ccsd.FP32), ('split-bf16x3', ccsd.BF16X3)):
            wf.construct_omega(p, copy_out=False)           # first call sizes the workspaces
            wf.synchronize()
            t0 = time.perf_counter()
            for _ in range(2):
                wf.construct_omega(p, copy_out=False)
            wf.synchronize()
            dtp = (time.perf_counter() - t0) / 2
            om = wf.get_omega()
            err = np.abs(om - ref)
            forms.append({
                'precision': name,
                'omega_evals_per_s': 1.0 / dtp,
                'speedup_vs_fp64': dt / dtp,
                'max_abs_err_vs_fp64': float(err.max()),
                'rel_err_vs_fp64': float(err.max() / np.abs(ref).max()),
            })
            log('%s: %.3f s per residual, rel err %.2e' % (name, dtp, forms[-1]['rel_err_vs_fp64']))
        extra['mixed'] = forms
    wf.close()

    if not args.no_tts and args.precision == 'fp64':
        octamer = water_octamer()
        extra['time_to_solution'] = time_to_solution(octamer)
        extra['eom_time_to_solution'] = eom_time_to_solution(octamer)
        if not args.no_jacobian:
            extra['jacobian_sweep'] = jacobian_sweep(octamer)
        del octamer

    cpu = finish_cpu_baseline(*cpu_job) if cpu_job else None

    out = {
        'metric': METRIC,
        'value': 1.0 / dt,
        'unit': 'omega evals/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': dt * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': {'fp64': 'f64', 'fp32': 'f32', 'bf16x3': 'bf16x3-f32acc'}[args.precision],
        'data': 'synthetic',
        'config': {'workload': WORKLOAD_NAME, 'n_o': n_o, 'n_v': n_v, 'n_J': n_J,
                   'seed': w['seed'], 'precision': args.precision, 'F_omega': f_omega},
        'roofline': roofline,
        'cpu_baseline': cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
