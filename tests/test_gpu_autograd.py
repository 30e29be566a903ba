"""The residual as a torch.autograd.Function inside the test harness: forward
Omega(t) through the C ABI (FP64 residual on the GPU), backward the library's
hand-written left transformation (the adjoint replay of the recorded residual
program, etg_jacobian_transpose_transformation).

eT's Jacobian is A = (dOmega/dt) D2, with D2 doubling the packed doubles
diagonal of the trial vector (jacobian_ccsd.F90:123), so the left
transformation returns A^T g = D2 (dOmega/dt)^T g; the vector-Jacobian product
is D2^-1 A^T g.  Checks:

* torch.autograd.gradcheck: every row of the backward against central finite
  differences of the forward (FP64, tiny systems, t1 != 0);
* the vjp against the transpose of the oracle's complex-step Jacobian of eT's
  residual factorisation (oracle/ccsd.py), 1e-11 relative;
* the vjp of a loss through torch: d(sum w * Omega)/dt equals D2^-1 A^T w.

torch holds only host tensors here (no torch.cuda call): the device work is
the library's.  The library is loaded before torch is imported, so the
process's HIP runtime is the one the library was built against.
"""
import numpy as np
import pytest

from et_amd import ccsd
from et_amd import synthetic as syn
from oracle import ccsd as O

pytestmark = pytest.mark.gpu


def torch_after_library():
    ccsd.load_library()
    import torch
    return torch


def system(n_o, n_v, n_J, seed):
    s = syn.system(seed, n_o, n_v, n_J, t2_sigma=0.03, t1_sigma=0.03, J_decay=6.0, L_scale=0.3)
    wf = ccsd.CCSD(n_o, n_v, n_J)
    wf.set_cholesky(s['L'])
    wf.set_oei(s['h'])
    wf.set_orbital_energies(s['eps'])
    t = np.concatenate([s['t1'].ravel(order='F'), s['t2']])
    return s, wf, t


def doubles_diagonal(n_t1):
    ai = np.arange(n_t1)
    return n_t1 + ai * (ai + 1) // 2 + ai


def residual_function(torch, wf):
    """torch.autograd.Function for Omega(t) on the wavefunction `wf`."""
    diag = doubles_diagonal(wf.n_t1)

    class Residual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, t):
            x = np.ascontiguousarray(t.detach().numpy(), dtype=np.float64)
            wf.set_amplitudes(x)
            wf.construct_t1_cholesky()
            wf.construct_fock()
            ctx.save_for_backward(t.detach().clone())
            return torch.from_numpy(wf.construct_omega(ccsd.FP64).copy())

        @staticmethod
        def backward(ctx, g):
            (t,) = ctx.saved_tensors
            wf.set_amplitudes(np.ascontiguousarray(t.numpy(), dtype=np.float64))
            wf.construct_t1_cholesky()
            wf.prepare_for_jacobian('left')
            sigma = wf.jacobian_transpose_transformation(np.ascontiguousarray(g.numpy(), dtype=np.float64))
            sigma[diag] *= 0.5                       # A^T g = D2 (dOmega/dt)^T g
            return torch.from_numpy(sigma)

    return Residual


@pytest.mark.parametrize('n_o,n_v,n_J,seed', [(1, 2, 4, 1), (2, 3, 6, 2), (3, 2, 5, 3)])
def test_gradcheck_left_transformation(n_o, n_v, n_J, seed):
    print('case', n_o, n_v, n_J, seed, flush=True)
    torch = torch_after_library()
    s, wf, t = system(n_o, n_v, n_J, seed)
    with wf:
        f = residual_function(torch, wf)
        x = torch.tensor(t, dtype=torch.float64, requires_grad=True)
        assert torch.autograd.gradcheck(f.apply, (x,), eps=1e-6, atol=1e-8, rtol=1e-6)


@pytest.mark.parametrize('n_o,n_v,n_J,seed', [(2, 3, 6, 4), (3, 4, 9, 5), (4, 3, 7, 6)])
def test_vjp_against_oracle_transpose(n_o, n_v, n_J, seed):
    print('case', n_o, n_v, n_J, seed, flush=True)
    torch = torch_after_library()
    s, wf, t = system(n_o, n_v, n_J, seed)
    A = O.jacobian_matrix(s['L'], s['h'], s['t1'], s['t2'], n_o)      # A = (dOmega/dt) D2
    d2 = np.ones(A.shape[0])
    d2[doubles_diagonal(n_o * n_v)] = 2.0
    rng = np.random.default_rng(seed)
    with wf:
        f = residual_function(torch, wf)
        for _ in range(2):
            x = torch.tensor(t, dtype=torch.float64, requires_grad=True)
            w = rng.normal(size=A.shape[0])
            loss = (f.apply(x) * torch.from_numpy(w)).sum()
            loss.backward()
            ref = (A.T @ w) / d2
            got = x.grad.numpy()
            assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()
