"""Differentiable encoding operator: y[c, m] = A_c(x; k_m) as a torch.autograd.Function.

Forward is the operator's image-to-k-space gather (gridFFT::backward,
gridFFT.cpp:824-906).  Backward uses two hand-written HIP paths behind the C
ABI:

* w.r.t. the image: the exact adjoint A^H (the scatter without density or
  taper weights: the operator is built with k_rad = inf and the density
  compensation enters only as a 0/1 sample mask), mrn_kspace_to_image;
* w.r.t. the trajectory: mrn_traj_grad, the derivative of the gather through
  the kernel weights (table slope, or the analytic Kaiser-Bessel derivative in
  exact_kb mode).  It needs the oversampled grid of x: when nothing has run on
  the operator since the forward call (same image, same stream) the grid that
  call left is used, otherwise it is recomputed from x.

Gradients follow torch's convention for complex tensors (the returned image
gradient is dL/dRe x + i dL/dIm x), so a real loss built from y gives
x.grad = A^H(dL/dy) and k.grad[m, a] = Re sum_c conj(dL/dy[c, m]) dy[c, m]/dk[m, a].

The trajectory is in the reference's units (cycles per FOV, |k| <= N/2,
src/mri_data.h:50-57), float32 of shape (M, 3) ordered (kx, ky, kz).  Samples
with weight 0 are skipped, as in the reference (gridFFT.cpp:1196-1197).
"""
import math

import torch
from torch.autograd.function import once_differentiable

from .gridfft import GridFFT


def _plain(t):
    """Contiguous memory holding t's values (torch may hand over lazily
    conjugated or negated views, e.g. the gradient of conj(y))."""
    return t.resolve_conj().resolve_neg().contiguous()


class _EncodeFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, k, mask, owner):
        if k is not None:
            owner._bind(k, mask)
        y = owner.op.backward(_plain(x))
        ctx.save_for_backward(x, owner._k, owner._mask)
        ctx.owner = owner
        ctx.token = owner._token
        return y

    @staticmethod
    @once_differentiable
    def backward(ctx, gy):
        x, k, mask = ctx.saved_tensors
        owner = ctx.owner
        if owner._token is not ctx.token:  # another trajectory was bound since forward
            owner._bind(k, mask)
            ctx.token = owner._token
        gy = _plain(gy)
        gx = gk = None
        # the trajectory gradient first: straight after the forward call, on
        # the same image and stream, the operator's grid still holds that
        # image's oversampled grid and mrn_traj_grad reuses it (the adjoint
        # below overwrites the grid)
        if ctx.needs_input_grad[1]:
            gk = owner.op.traj_grad(_plain(x), gy)
        if ctx.needs_input_grad[0]:
            gx = owner.op.forward(gy, accumulate=False)
        return gx, gk, None, None


class SenseEncoding:
    """Callable A(x; k): complex64 image (Nz, Ny, Nx) -> k-space (C, M).

    ``smaps`` (C, Nz, Ny, Nx) complex64 on the device.  ``kernel_mode`` is
    "table" (the reference's 600-point lerp table; piecewise-linear in k) or
    "exact_kb" (direct Kaiser-Bessel evaluation, smooth in k inside the
    support, for trajectory optimisation).  Other keyword arguments go to
    GridFFT.

    ``enc(x, k, kw)`` sorts the samples of k into bins on every call (k may
    change between calls, e.g. while it is being optimised); for a fixed
    trajectory call ``enc.set_traj(k, kw)`` once and then ``enc(x)``.  At a
    tap exactly on the edge of the kernel's support (|dk - l| = dwin, where
    the truncated kernel jumps to zero) the trajectory derivative is the one
    from inside the support."""

    def __init__(self, N, smaps, overgrid=1.5, dwin=-1.0, kernel_mode="exact_kb", **kw):
        C = smaps.shape[0]
        self.op = GridFFT(N, ncoils=C, overgrid=overgrid, dwin=dwin, k_rad=math.inf,
                          kernel_mode=kernel_mode, **kw)
        self.op.set_smaps(smaps.contiguous())
        self._k = self._mask = None
        self._token = None

    @staticmethod
    def _mask_of(k, kw):
        if k.dim() != 2 or k.shape[1] != 3 or k.dtype != torch.float32:
            raise ValueError("k must be float32 of shape (M, 3)")
        if kw is None:
            return torch.ones(k.shape[0], dtype=torch.float32, device=k.device)
        return (kw != 0).to(torch.float32)

    def _bind(self, k, mask):
        kc = k.detach().t().contiguous()
        self.op.set_traj(kc[0], kc[1], kc[2], mask.detach().contiguous())
        self._k, self._mask = k.detach(), mask.detach()
        self._token = object()

    def set_traj(self, k, kw=None):
        """Bind a fixed trajectory (M, 3) for later calls without k."""
        self._bind(k, self._mask_of(k, kw))

    def __call__(self, x, k=None, kw=None):
        if k is None:
            if self._k is None:
                raise ValueError("no trajectory: pass k or call set_traj first")
            return _EncodeFn.apply(x, None, None, self)
        return _EncodeFn.apply(x, k, self._mask_of(k, kw), self)

    def close(self):
        self.op.close()
