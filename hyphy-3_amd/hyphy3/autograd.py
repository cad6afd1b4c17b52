"""PyTorch autograd integration (reference: python/hyphy3/autograd.py:1-55).

``PhylogeneticTreeLogLikelihood`` keeps the reference's contract: forward
returns ln L(t) as a 0-d tensor and caches d lnL/d t computed by the engine's
inside-outside adjoint; backward returns ``grad_output * grads``.

``MixtureLogLikelihood`` extends it to branch lengths, omegas, mixture
weights and GTR exchangeabilities (all five differentiable), computed in one
MI355X evaluation (forward + outside pass + spectral Frechet projections).
"""
import os

os.environ.setdefault("KMP_DUPLICATE_LIB_OK", "TRUE")

try:
    import torch
    TORCH_AVAILABLE = True
except ImportError:  # pragma: no cover
    TORCH_AVAILABLE = False

import _hyphy3

if TORCH_AVAILABLE:

    class PhylogeneticTreeLogLikelihood(torch.autograd.Function):
        @staticmethod
        def forward(ctx, branch_lengths, tree, aln, params):
            bl = branch_lengths.detach().to("cpu", torch.float64).tolist()
            ll, grads = _hyphy3.compute_branch_length_gradients(tree, aln, params, bl)
            ctx.save_for_backward(torch.tensor(grads, dtype=branch_lengths.dtype, device=branch_lengths.device))
            return torch.tensor(ll, dtype=branch_lengths.dtype, device=branch_lengths.device)

        @staticmethod
        def backward(ctx, grad_output):
            (grads,) = ctx.saved_tensors
            return grad_output * grads, None, None, None

    def tree_log_likelihood(branch_lengths, tree, aln, params=None):
        if params is None:
            params = _hyphy3.MG94Parameters()
        return PhylogeneticTreeLogLikelihood.apply(branch_lengths, tree, aln, params)

    class MixtureLogLikelihood(torch.autograd.Function):
        """ln L(bl, omegas, weights, theta) of a branch-site omega mixture."""

        @staticmethod
        def forward(ctx, branch_lengths, omegas, weights, theta, tree, aln, alpha=1.0, time_convention=0):
            p = _hyphy3.MG94Parameters()
            p.alpha = float(alpha)
            th = theta.detach().to("cpu", torch.float64).tolist()
            p.theta_AC, p.theta_AT, p.theta_CG, p.theta_CT, p.theta_GT = th
            r = _hyphy3.compute_mixture_gradients(
                tree, aln, p, omegas.detach().to("cpu", torch.float64).tolist(),
                weights.detach().to("cpu", torch.float64).tolist(),
                branch_lengths.detach().to("cpu", torch.float64).tolist(), int(time_convention), True)
            dev, dt = branch_lengths.device, branch_lengths.dtype
            ctx.save_for_backward(torch.tensor(r["grad_branches"], dtype=dt, device=dev),
                                  torch.tensor(r["grad_omegas"], dtype=dt, device=dev),
                                  torch.tensor(r["grad_weights"], dtype=dt, device=dev),
                                  torch.tensor(r["grad_theta"], dtype=dt, device=dev))
            return torch.tensor(r["log_likelihood"], dtype=dt, device=dev)

        @staticmethod
        def backward(ctx, g):
            gb, go, gw, gt = ctx.saved_tensors
            return g * gb, g * go, g * gw, g * gt, None, None, None, None

    def mixture_log_likelihood(branch_lengths, omegas, weights, theta, tree, aln, alpha=1.0, time_convention=0):
        return MixtureLogLikelihood.apply(branch_lengths, omegas, weights, theta, tree, aln, alpha, time_convention)

    class EngineLogLikelihood(torch.autograd.Function):
        """ln L(bl, omegas, weights, theta) of a branch-site omega mixture on a
        resident ``hyphy3.engine.Engine`` (problem, model and mixture already
        set).  Forward runs one MI355X evaluation (inside pass, outside pass,
        adjoint and spectral Frechet projections); backward scales the
        gradients it returned.  ``branch_lengths`` holds one entry per node
        (the root's is ignored)."""

        @staticmethod
        def forward(ctx, branch_lengths, omegas, weights, theta, engine, precision):
            def host(t):
                return t.detach().to("cpu", torch.float64).numpy()

            r = engine.eval(host(branch_lengths), host(omegas), host(weights), host(theta), precision=precision)
            dev, dt = branch_lengths.device, branch_lengths.dtype
            ctx.save_for_backward(*(torch.as_tensor(r[k], dtype=dt, device=dev)
                                    for k in ("grad_bl", "grad_omega", "grad_w", "grad_theta")))
            return torch.tensor(r["log_likelihood"], dtype=dt, device=dev)

        @staticmethod
        def backward(ctx, g):
            gb, go, gw, gt = ctx.saved_tensors
            return g * gb, g * go, g * gw, g * gt, None, None

    def engine_log_likelihood(branch_lengths, omegas, weights, theta, engine, precision):
        return EngineLogLikelihood.apply(branch_lengths, omegas, weights, theta, engine, precision)
