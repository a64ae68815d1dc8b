"""Differentiable permanent (no reference counterpart: theboss has no gradients).

forward:  perm(A) for a batch of n x n complex128 matrices. When A requires grad, one
          gradient launch (tb_perm_grad_batch) yields the permanents and all n^2 minors
          d perm / d A_ij = perm(A without row i and column j); the minors (B x n x n) are
          kept for the backward pass. Otherwise the Glynn kernel computes perm alone.
backward: grad_out times the kept minors. Per-subset sums are recomputed inside the
          sweep, never stored.
The permanent is holomorphic, so in PyTorch's conjugate-Wirtinger convention the
incoming gradient g maps to g * conj(d perm / d A).
"""
from . import _device as D
from . import ops


class PermanentFunction:
    """Namespace holder so importing this module does not require torch at import time."""


def _make_function():
    T = D.torch()

    class _Permanent(T.autograd.Function):
        @staticmethod
        def forward(ctx, A):
            if ctx.needs_input_grad[0]:
                perm, G = ops.perm_grad_batch(A.detach())
                ctx.save_for_backward(G)
                return perm
            return ops.glynn_batch(A.detach())

        @staticmethod
        def backward(ctx, grad_out):
            (G,) = ctx.saved_tensors
            return grad_out.reshape(-1, 1, 1) * G.conj()

    return _Permanent


_FN = None


def permanent(A):
    """Permanent of a square matrix or a batch (..., n, n) of complex128 GPU matrices,
    differentiable with respect to A."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    batch_shape = A.shape[:-2]
    n = A.shape[-1]
    flat = A.reshape(-1, n, n)
    if not flat.is_contiguous():
        flat = flat.contiguous()
    out = _FN.apply(flat)
    return out.reshape(batch_shape)
