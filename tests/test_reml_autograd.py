"""Test-only torch.autograd binding of the REML adjoint (eqtl_gpu_reml_grad_hess).

RemlNegLL maps theta = ln(delta) (one entry per prepared gene) to the REML
objective f(theta) of reml_negll (model_lmm.cpp:10-37); its backward uses the
adjoint gradient df/dtheta from the same call, and the gradient's own backward
uses the adjoint second derivative (genes are independent, so the Hessian in
theta is diagonal). The library computes everything on the GPU; torch only
carries the values, in float64 CPU tensors.

Checks:
  * torch.autograd.gradcheck / gradgradcheck of the binding (its analytic
    first and second derivatives against torch's finite differences of its
    own f and gradient), at delta, delta e^-1 and delta e^+1;
  * the gradient against central differences of the ORACLE's f (step 1e-4 in
    ln delta), so the binding is tied to the reference's objective;
  * Newton steps driven through the binding (gradient by backward, Hessian
    by a second backward) reach the library's batched Newton delta and are no
    worse than the reference search's delta (optimize_delta,
    model_lmm.cpp:39-86) on every gene.
"""
import numpy as np
import pytest

import oracle as O
import eqtl_amd

from _cases import lmm_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


class _RemlGrad(torch.autograd.Function):
    """theta -> df/dtheta (with d2f/dtheta2 as its derivative)."""

    @staticmethod
    def forward(ctx, theta, engine):
        _, g, h = engine.reml_grad_hess(theta.detach().numpy())
        ctx.save_for_backward(torch.from_numpy(h))
        return torch.from_numpy(g)

    @staticmethod
    def backward(ctx, grad_out):
        (h,) = ctx.saved_tensors
        return grad_out * h, None


class RemlNegLL(torch.autograd.Function):
    """theta = ln(delta) per gene -> f(theta) per gene."""

    @staticmethod
    def forward(ctx, theta, engine):
        f, _, _ = engine.reml_grad_hess(theta.detach().numpy())
        ctx.engine = engine
        ctx.save_for_backward(theta)
        return torch.from_numpy(f)

    @staticmethod
    def backward(ctx, grad_out):
        (theta,) = ctx.saved_tensors
        return grad_out * _RemlGrad.apply(theta, ctx.engine), None


@pytest.fixture(scope="module")
def prepared(gpu_ctx):
    ds, rows, (Q, lam), win, ref = lmm_case()
    gpu_ctx.set_samples(ds.n)
    gpu_ctx.set_basis(Q, lam)
    gpu_ctx.lmm_prep(ds.Y, ds.X(), delta_mode=eqtl_amd.DELTA_GIVEN, delta_in=ref["state"]["delta"])
    return ds, Q, lam, ref


def test_gradcheck_and_gradgradcheck(gpu_ctx, prepared):
    """On the genes whose delta is interior (1e-4 < delta < 1e4): at the range's
    ends the likelihood is flat and torch's fixed-step differences of f are
    rounding noise (the noise-aware comparison below covers every gene)."""
    ds, Q, lam, ref = prepared
    base = np.log(ref["state"]["delta"])
    sub = np.nonzero((ref["state"]["delta"] > 1e-4) & (ref["state"]["delta"] < 1e4))[0]
    assert sub.size >= 10
    for shift in (-1.0, 0.0, 1.0):
        full = torch.tensor(base + shift, dtype=torch.float64)
        idx = torch.from_numpy(sub)

        def fn(t):
            th = full.clone().index_put((idx,), t)
            return RemlNegLL.apply(th, gpu_ctx)[idx]

        t0 = full[idx].clone().requires_grad_(True)
        assert torch.autograd.gradcheck(fn, (t0,), eps=1e-5, atol=1e-5, rtol=1e-4)
        assert torch.autograd.gradgradcheck(fn, (t0,), eps=1e-5, atol=1e-4, rtol=1e-3)


def test_gradient_against_oracle_central_differences(gpu_ctx, prepared):
    ds, Q, lam, ref = prepared
    Yt, Xt = O.rotate(Q, ds.Y), O.rotate(Q, ds.X())
    th = torch.tensor(np.log(ref["state"]["delta"]) + 0.5, dtype=torch.float64, requires_grad=True)
    f = RemlNegLL.apply(th, gpu_ctx)
    (g,) = torch.autograd.grad(f.sum(), th)
    hstep = 1e-4
    for k in range(ds.Y.shape[1]):
        F = lambda t: O.reml_negll(np.exp(t), Yt[:, k], Xt, lam)
        t0 = float(th[k])
        fo = F(t0)
        gfd = (F(t0 + hstep) - F(t0 - hstep)) / (2 * hstep)
        eps_f = 64 * np.finfo(float).eps * abs(fo)
        assert abs(float(f[k]) - fo) <= 1e-9 * (1 + abs(fo))
        assert abs(float(g[k]) - gfd) <= 1e-5 * (1 + abs(gfd)) + 2 * eps_f / hstep, (k, float(g[k]), gfd)


def test_newton_through_the_binding_reaches_the_library_newton_delta(gpu_ctx, prepared):
    """Newton steps on ln(delta) driven by torch: the gradient from backward()
    and the (diagonal) Hessian from a second backward through the gradient,
    steps clipped to +-1 and halved until f decreases. On every curved
    interior minimum this reaches the library's batched Newton delta, and on
    every gene it is no worse than the reference search's delta."""
    ds, Q, lam, ref = prepared
    Yt, Xt = O.rotate(Q, ds.Y), O.rotate(Q, ds.X())
    th = torch.tensor(np.log(ref["state"]["delta"]) + 0.7, dtype=torch.float64)
    lo, hi = np.log(1e-8), np.log(1e6)
    for _ in range(40):
        t = th.clone().requires_grad_(True)
        f = RemlNegLL.apply(t, gpu_ctx)
        (g,) = torch.autograd.grad(f.sum(), t, create_graph=True)
        (h,) = torch.autograd.grad(g.sum(), t)  # diagonal Hessian: each f_k depends on theta_k only
        g, h, f = g.detach(), h.detach(), f.detach()
        step = torch.where(h > 0, -g / h, -0.5 * torch.sign(g)).clamp(-1.0, 1.0)
        for _half in range(30):
            cand = (th + step).clamp(lo, hi)
            fc = RemlNegLL.apply(cand, gpu_ctx)
            better = fc <= f + 1e-13 * f.abs()
            th = torch.where(better, cand, th)
            if bool(better.all()):
                break
            step = torch.where(better, torch.zeros_like(step), 0.5 * step)
        if float(g.abs().max()) < 1e-9:
            break
    d_torch = np.exp(th.numpy())
    gpu_ctx.set_basis(Q, lam)
    d_newton, _ = gpu_ctx.lmm_prep(ds.Y, ds.X(), delta_mode=eqtl_amd.DELTA_NEWTON)
    _, _, hh = gpu_ctx.reml_grad_hess(np.log(d_newton))
    curved = 0
    for k in range(ds.Y.shape[1]):
        ft = O.reml_negll(d_torch[k], Yt[:, k], Xt, lam)
        fn = O.reml_negll(d_newton[k], Yt[:, k], Xt, lam)
        fr = O.reml_negll(ref["state"]["delta"][k], Yt[:, k], Xt, lam)
        assert ft <= fr + 1e-9 * abs(fr) and fn <= fr + 1e-9 * abs(fr), (k, ft, fn, fr)
        if hh[k] > 1e-3 and 1e-6 < d_newton[k] < 1e5:  # a curved interior minimum
            curved += 1
            assert abs(np.log(d_torch[k]) - np.log(d_newton[k])) < 1e-6, (k, d_torch[k], d_newton[k])
    assert curved >= 10
