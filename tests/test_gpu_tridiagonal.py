"""The stand-alone tridiagonal operators on the GPU (reference
``autodiff/tridiagonal_factor.py`` / ``tridiagonal_solve.py``).

* The reference's own known-answer system (``tests/test_derivative_identity.py:34-37``)
  against the dense solve of the matrix the two sweeps invert (numpy.linalg.solve).
* The reference's JVP/VJP identity with its term-wise 8 gamma_n bound and its
  failing zero-VJP control (``test_derivative_identity.py:8-67``), plus a
  sign-flipped control; then the same identity for every input of both operators.
* torch.autograd.gradcheck (central differences, forward and reverse mode) on
  small random systems.
* The factorization's invariants: b = 1/alpha + a gamma(k-1) and c = gamma/alpha.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch
import torch.autograd.forward_ad as fw_ad

pytestmark = pytest.mark.gpu

DEV = "cuda"
EPS = float(np.finfo(np.float64).eps)


def _identity(jvp, weight, direction, vjp):
    """(relative residual, 8 gamma_n) of <J d, w> against <d, J^T w> in the reference's term-wise form."""
    lt = (jvp.reshape(-1) * weight.reshape(-1)).double()
    rt = (direction.reshape(-1) * vjp.reshape(-1)).double()
    den = max(float(lt.abs().sum() + rt.abs().sum()), np.finfo(np.float64).tiny)
    n = max(lt.numel(), rt.numel())
    gamma = n * EPS / (1.0 - n * EPS)
    return abs(float(lt.sum() - rt.sum())) / den, 8.0 * gamma


def _dense(a, alpha, gamma):
    """The (L+1) x (L+1) matrix whose inverse the elimination + back substitution apply."""
    L = a.shape[0]
    M = np.zeros((L + 1, L + 1))
    M[0, 0] = 1.0
    M[0, 1] = gamma[0]
    for k in range(1, L):
        M[k, k - 1] = a[k]
        M[k, k] = 1.0 / alpha[k] + a[k] * gamma[k - 1]
        M[k, k + 1] = gamma[k] / alpha[k]
    M[L, L] = 1.0
    return M


def _kat():
    a = torch.tensor([[0.0, 0.08, -0.05, 0.03]], dtype=torch.float64, device=DEV)
    alpha = torch.tensor([[1.0, 0.95, 1.03, 0.98]], dtype=torch.float64, device=DEV)
    gamma = torch.tensor([[0.04, -0.02, 0.06, 0.01]], dtype=torch.float64, device=DEV)
    rhs = torch.tensor([[1.0, -0.5, 0.75, 0.2, -0.1]], dtype=torch.float64, device=DEV)
    return a, alpha, gamma, rhs


def test_known_answer_system_against_the_dense_solve():
    from mpas_dycore_pytorch_amd.tridiagonal import tridiagonal_solve

    a, alpha, gamma, rhs = _kat()
    x = tridiagonal_solve(a, alpha, gamma, rhs).cpu().numpy()[0]
    M = _dense(a.cpu().numpy()[0], alpha.cpu().numpy()[0], gamma.cpu().numpy()[0])
    ref = np.linalg.solve(M, rhs.cpu().numpy()[0])
    print(f"[tridiag] known-answer solve {x.tolist()} dense {ref.tolist()}", flush=True)
    assert np.allclose(x, ref, rtol=1e-14, atol=1e-15)
    # and the residual of the dense system itself
    assert np.allclose(M @ x, rhs.cpu().numpy()[0], rtol=0, atol=1e-15)


def test_reference_identity_and_failing_controls():
    """The reference's identity test, run on the GPU operators."""
    from mpas_dycore_pytorch_amd.tridiagonal import TridiagonalSolve

    a, alpha, gamma, primal = _kat()
    direction = torch.randn(primal.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(7301)).to(DEV)
    weight = torch.randn(primal.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(9109)).to(DEV)
    with fw_ad.dual_level():
        out = TridiagonalSolve.apply(a, alpha, gamma, fw_ad.make_dual(primal, direction))
        jvp = fw_ad.unpack_dual(out).tangent
    assert jvp is not None
    r = primal.clone().requires_grad_(True)
    vjp = torch.autograd.grad(torch.sum(TridiagonalSolve.apply(a, alpha, gamma, r) * weight), r)[0]
    residual, bound = _identity(jvp, weight, direction, vjp)
    print(f"[tridiag] identity residual {residual:.2e} bound {bound:.2e}", flush=True)
    assert residual <= bound
    zero, _ = _identity(jvp, weight, direction, torch.zeros_like(vjp))
    flip, _ = _identity(jvp, weight, direction, -vjp)
    print(f"[tridiag] controls: zero VJP {zero:.2e}, sign-flipped VJP {flip:.2e}", flush=True)
    assert zero > 1.0e-6 and flip > 1.0e-6


def _system(n, L, seed):
    g = torch.Generator().manual_seed(seed)
    a = 0.2 * torch.randn(n, L, dtype=torch.float64, generator=g)
    c = 0.2 * torch.randn(n, L, dtype=torch.float64, generator=g)
    b = 1.5 + torch.rand(n, L, dtype=torch.float64, generator=g)
    rhs = torch.randn(n, L + 1, dtype=torch.float64, generator=g)
    a[:, 0] = 0.0
    return a.to(DEV), b.to(DEV), c.to(DEV), rhs.to(DEV)


@pytest.mark.parametrize("n,L", [(1, 1), (3, 4), (4096, 55), (257, 130)])
def test_factorization_invariants_and_dense_solve(n, L):
    from mpas_dycore_pytorch_amd.tridiagonal import tridiagonal_factorize, tridiagonal_solve

    a, b, c, rhs = _system(n, L, 11 + n + L)
    alpha, gamma = tridiagonal_factorize(a, b, c)
    al, ga = alpha.cpu().numpy(), gamma.cpu().numpy()
    an, bn, cn = a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy()
    assert np.all(al[:, 0] == 0.0) and np.all(ga[:, 0] == 0.0)
    if L > 1:
        gprev = np.concatenate([np.zeros((n, 1)), ga[:, :-1]], axis=1)
        assert np.allclose(1.0 / al[:, 1:] + an[:, 1:] * gprev[:, 1:], bn[:, 1:], rtol=1e-13, atol=0)
        assert np.allclose(ga[:, 1:] / al[:, 1:], cn[:, 1:], rtol=1e-13, atol=1e-300)
    x = tridiagonal_solve(a, alpha, gamma, rhs).cpu().numpy()
    for col in sorted({0, n // 2, n - 1}):
        M = _dense(an[col], al[col], ga[col])
        ref = np.linalg.solve(M, rhs.cpu().numpy()[col])
        assert np.allclose(x[col], ref, rtol=1e-12, atol=1e-13), col


def test_gradcheck_forward_and_reverse():
    from mpas_dycore_pytorch_amd.tridiagonal import TridiagonalFactorize, TridiagonalSolve

    a, b, c, rhs = _system(3, 5, 5)
    ins = tuple(t.clone().requires_grad_(True) for t in (a, b, c))
    assert torch.autograd.gradcheck(TridiagonalFactorize.apply, ins, check_forward_ad=True)
    alpha, gamma = TridiagonalFactorize.apply(a, b, c)
    ins = tuple(t.detach().clone().requires_grad_(True) for t in (a, alpha, gamma, rhs))
    assert torch.autograd.gradcheck(TridiagonalSolve.apply, ins, check_forward_ad=True)


def test_identity_for_every_input_at_column_scale():
    """<J d, w> = <d, J^T w> for random directions on every input of both
    operators, 2048 columns x 55 levels, within the term-wise bound."""
    from mpas_dycore_pytorch_amd.tridiagonal import TridiagonalFactorize, TridiagonalSolve

    a, b, c, rhs = _system(2048, 55, 3)
    g = torch.Generator().manual_seed(17)
    for name, fn, ins in (("factorize", lambda *x: TridiagonalFactorize.apply(*x), (a, b, c)),
                          ("solve", lambda *x: TridiagonalSolve.apply(*x), None)):
        if ins is None:
            alpha, gamma = TridiagonalFactorize.apply(a, b, c)
            ins = (a, alpha.detach(), gamma.detach(), rhs)
        dirs = tuple(torch.randn(t.shape, dtype=torch.float64, generator=g).to(DEV) for t in ins)
        with fw_ad.dual_level():
            outs = fn(*[fw_ad.make_dual(t, d) for t, d in zip(ins, dirs)])
            outs = outs if isinstance(outs, tuple) else (outs,)
            jvps = [fw_ad.unpack_dual(o).tangent for o in outs]
        weights = [torch.randn(o.shape, dtype=torch.float64, generator=g).to(DEV) for o in outs]
        leaves = [t.detach().clone().requires_grad_(True) for t in ins]
        outs2 = fn(*leaves)
        outs2 = outs2 if isinstance(outs2, tuple) else (outs2,)
        vjps = torch.autograd.grad(sum(torch.sum(o * w) for o, w in zip(outs2, weights)), leaves)
        residual, bound = _identity(torch.cat([j.reshape(-1) for j in jvps]),
                                    torch.cat([w.reshape(-1) for w in weights]),
                                    torch.cat([d.reshape(-1) for d in dirs]),
                                    torch.cat([v.reshape(-1) for v in vjps]))
        print(f"[tridiag] {name}: identity residual {residual:.2e} bound {bound:.2e}", flush=True)
        assert residual <= bound, name


def test_shape_and_device_errors():
    from mpas_dycore_pytorch_amd.errors import MpasDeviceError
    from mpas_dycore_pytorch_amd.tridiagonal import tridiagonal_factorize, tridiagonal_solve

    a, b, c, rhs = _system(2, 4, 1)
    with pytest.raises(ValueError):
        tridiagonal_factorize(a, b[:, :3], c)
    with pytest.raises(ValueError):
        tridiagonal_solve(a, b, c, rhs[:, :4])
    with pytest.raises(MpasDeviceError):
        tridiagonal_factorize(a.cpu(), b.cpu(), c.cpu())
