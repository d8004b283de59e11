"""torch.autograd.Functions over the gfx950 FFT kernels, with adjoint FFTs.

Adjoints (unnormalised forward F, conjugate-Wirtinger gradients as torch uses):
  c2c(x, s)     y = F_s x              grad_x = F_{-s} grad_y          (F^H = F_{-s})
  c2c(x, s, a)  y = a F_s x (a real)   grad_x = a F_{-s} grad_y        (numpy's `norm` scales)
  rfft(x)       y = (F x)[0..n/2]      grad_x = n * irfft(w . grad_y), w = 1 at bins 0 and n/2,
                                       1/2 elsewhere (the Hermitian weights of the real input)
  irfft(h, n)   x = irfft(h)           grad_h = rfft(grad_x) * v / n, v = 1 at bins 0 and n/2,
                                       2 elsewhere, imaginary parts of bins 0 and n/2 zero
Every adjoint is one launch of the library's kernels with the weighting done
inside the kernel (bfly_real_ex with BFLY_REAL_ADJOINT; no torch.fft call, no
separate weighting pass).
"""
from __future__ import annotations

import torch

from . import ops


def _as_interleaved(z: torch.Tensor) -> torch.Tensor:
    return torch.view_as_real(z.contiguous()).reshape(*z.shape[:-1], 2 * z.shape[-1])


def _as_complex(f: torch.Tensor) -> torch.Tensor:
    return torch.view_as_complex(f.reshape(*f.shape[:-1], f.shape[-1] // 2, 2))


class C2C(torch.autograd.Function):
    """y = scale * F_sign x along the last axis (scale None: unnormalised);
    the backward is the same scale times the conjugate transform, both with
    the scale applied in the kernel's last store."""
    @staticmethod
    def forward(ctx, x: torch.Tensor, sign: int = -1, scale: float | None = None):
        ctx.sign, ctx.scale = sign, 1.0 if scale is None else scale
        return ops.c2c(x.contiguous(), sign, scale=ctx.scale)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        return ops.c2c(g.contiguous(), -ctx.sign, scale=ctx.scale), None, None


class RFFT(torch.autograd.Function):
    """float32 [..., n] -> complex64 [..., n/2+1]."""

    @staticmethod
    def forward(ctx, x: torch.Tensor):
        ctx.n = x.shape[-1]
        return _as_complex(ops.rfft(x.contiguous()))

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        return ops.irfft(_as_interleaved(g), ctx.n, adjoint=True)


class IRFFT(torch.autograd.Function):
    """complex64 [..., n/2+1] -> float32 [..., n] (1/n)."""

    @staticmethod
    def forward(ctx, h: torch.Tensor, n: int):
        ctx.n = n
        return ops.irfft(_as_interleaved(h), n)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        return _as_complex(ops.rfft(g.contiguous(), adjoint=True)), None


class RFFTModes(torch.autograd.Function):
    """float32 [..., n] -> complex64 [..., K]: the first K bins of the R2C."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, modes: int):
        ctx.n = x.shape[-1]
        return _as_complex(ops.rfft_modes(x.contiguous(), modes))

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        return ops.irfft_modes(_as_interleaved(g), ctx.n, adjoint=True), None


class IRFFTModes(torch.autograd.Function):
    """complex64 [..., K] (higher bins zero) -> float32 [..., n] (1/n)."""

    @staticmethod
    def forward(ctx, h: torch.Tensor, n: int):
        ctx.modes = h.shape[-1]
        return ops.irfft_modes(_as_interleaved(h), n)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        return _as_complex(ops.rfft_modes(g.contiguous(), ctx.modes, adjoint=True)), None


def rfft_modes(x: torch.Tensor, modes: int) -> torch.Tensor:
    return RFFTModes.apply(x, modes)


def irfft_modes(h: torch.Tensor, n: int) -> torch.Tensor:
    return IRFFTModes.apply(h, n)


def fft(x: torch.Tensor, norm: str | None = None) -> torch.Tensor:
    """Differentiable forward C2C of complex64 [..., n]; `norm` as in
    torch.fft.fft (None and "backward": unscaled, "ortho": 1/sqrt(n),
    "forward": 1/n)."""
    return C2C.apply(x, -1, ops.norm_scale(x.shape[-1], -1, norm))


def ifft(x: torch.Tensor, norm: str | None = None) -> torch.Tensor:
    """Differentiable inverse C2C; `norm` as in torch.fft.ifft (None and
    "backward": 1/n, "ortho": 1/sqrt(n), "forward": unscaled)."""
    return C2C.apply(x, 1, ops.norm_scale(x.shape[-1], 1, norm))


def ifft_unnormalized(x: torch.Tensor) -> torch.Tensor:
    """The reference's unnormalised inverse (n times numpy's ifft)."""
    return C2C.apply(x, 1)


def rfft(x: torch.Tensor) -> torch.Tensor:
    return RFFT.apply(x)


def irfft(h: torch.Tensor, n: int) -> torch.Tensor:
    return IRFFT.apply(h, n)
