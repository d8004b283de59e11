"""1-D spectral convolution on the gfx950 FFT kernels.

y = irfft( pad( einsum('bik,iok->bok', rfft(x)[..., :K], W) ), N )

x [B, C_in, N] float32, W complex64 [C_in, C_out, K]; the first K modes of
the R2C are mixed across channels, the remaining N/2+1-K modes are zero, and
the C2R returns [B, C_out, N].  Forward and backward run through the
adjoint-FFT autograd Functions; the mode mixing (forward, input gradient,
weight gradient) is the library's per-mode mixing kernel working in the
modes' own layout [B, C, K], so no transposed copies are made.
"""
from __future__ import annotations

import torch

from . import _lib, ops
from . import autograd as bag


def mix_modes_torch(xm: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    """xm [B, C_in, K] x w [C_in, C_out, K] -> [B, C_out, K] (complex64) with
    torch: transposed copies and one batched complex GEMM per mode."""
    a = xm.permute(2, 0, 1).contiguous()          # [K, B, C_in]
    b = w.permute(2, 0, 1).contiguous()           # [K, C_in, C_out]
    return torch.bmm(a, b).permute(1, 2, 0)       # [B, C_out, K]


class MixModes(torch.autograd.Function):
    """y[b,o,k] = sum_i x[b,i,k] w[i,o,k] on the mode-mixing kernel; the
    backward is the same kernel for gx = gy w^H and gw = x^H gy per mode."""

    @staticmethod
    def forward(ctx, xm: torch.Tensor, w: torch.Tensor):
        xm, w = xm.contiguous(), w.contiguous()
        B, C_in, K = xm.shape
        if w.shape[0] != C_in or w.shape[2] != K:
            raise ValueError(f"weights {tuple(w.shape)} do not match modes {tuple(xm.shape)}")
        ctx.save_for_backward(xm, w)
        ctx.dims = (B, C_in, w.shape[1], K)
        return ops.mode_mix(_lib.MIX_FORWARD, xm, w, B, C_in, w.shape[1], K)

    @staticmethod
    def backward(ctx, gy: torch.Tensor):
        xm, w = ctx.saved_tensors
        B, C_in, C_out, K = ctx.dims
        gy = gy.contiguous()
        gx = gw = None
        if ctx.needs_input_grad[0]:
            gx = ops.mode_mix(_lib.MIX_GRAD_INPUT, gy, w, B, C_in, C_out, K)
        if ctx.needs_input_grad[1]:
            gw = ops.mode_mix(_lib.MIX_GRAD_WEIGHT, xm, gy, B, C_in, C_out, K)
        return gx, gw


def mix_modes(xm: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    """xm [B, C_in, K] x w [C_in, C_out, K] -> [B, C_out, K] (complex64)."""
    return MixModes.apply(xm, w)


def spectral_conv1d(x: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    """The layer on the mode-truncated kernels: the R2C writes only the K kept
    bins and the C2R reads only them, so neither the discarded modes nor the
    zero fill ever touch HBM."""
    N = x.shape[-1]
    K = w.shape[-1]
    X = bag.rfft_modes(x, K)                      # [B, C_in, K]
    Y = mix_modes(X, w)                           # [B, C_out, K]
    return bag.irfft_modes(Y, N)


def spectral_conv1d_full(x: torch.Tensor, w: torch.Tensor, rfft=bag.rfft, irfft=bag.irfft,
                         mix=mix_modes) -> torch.Tensor:
    """The layer on full-spectrum transforms (plain R2C, slice, zero-pad, C2R)."""
    B, C_in, N = x.shape
    K = w.shape[-1]
    X = rfft(x)                                   # [B, C_in, N/2+1]
    Y = mix(X[..., :K], w)                        # [B, C_out, K]
    pad = torch.zeros(B, w.shape[1], N // 2 + 1 - K, dtype=Y.dtype, device=Y.device)
    return irfft(torch.cat([Y, pad], dim=-1), N)


class SpectralConv1d(torch.nn.Module):
    def __init__(self, c_in: int, c_out: int, modes: int, generator=None, device=None):
        super().__init__()
        scale = (1.0 / (c_in * 2)) ** 0.5
        re = torch.randn(c_in, c_out, modes, generator=generator) * scale
        im = torch.randn(c_in, c_out, modes, generator=generator) * scale
        self.weight = torch.nn.Parameter(torch.complex(re, im).to(device))

    def forward(self, x):
        return spectral_conv1d(x, self.weight)


def torch_fft_reference(x: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    """The same layer on torch.fft (hipFFT on ROCm), for comparison."""
    return spectral_conv1d_full(x, w, rfft=lambda t: torch.fft.rfft(t),
                                irfft=lambda h, n: torch.fft.irfft(h, n), mix=mix_modes_torch)
