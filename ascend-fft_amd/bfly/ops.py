"""Device FFTs on torch tensors through the C ABI (no torch types cross it).

Tensors stay in HBM; the kernels run on torch's current stream of the
tensor's device.  Layouts are the reference's transform contract:
  c2c   complex64 [..., n]  or float32 interleaved [..., 2n]
  rfft  float32 [..., n]       -> float32 half spectrum [..., n+2]
  irfft float32 [..., n+2], n  -> float32 [..., n]   (scaled 1/n)
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib

_tables: dict = {}


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def supported(kind: int, n: int) -> bool:
    return bool(_lib.hip().bfly_supported(kind, n))


def table(kind: int, n: int, device) -> torch.Tensor:
    """Device copy of the FP64-generated twiddle table of (kind, n), cached."""
    device = torch.device(device)
    key = (device.index, kind, n)
    t = _tables.get(key)
    if t is None:
        lib = _lib.hip()
        nf = lib.bfly_table_floats(kind, n)
        if nf == 0:
            raise _lib.BflyError(_lib.ERR_SHAPE, f"no kernel for kind={kind} n={n}")
        host = np.empty(nf, np.float32)
        _lib.check(lib.bfly_table_fill(kind, n, host.ctypes.data), "bfly_table_fill")
        t = torch.from_numpy(host).to(device)
        _tables[key] = t
    return t


def workspace(kind: int, n: int, batch: int, device) -> torch.Tensor | None:
    nbytes = _lib.hip().bfly_workspace_bytes(kind, n, batch)
    if nbytes == 0:
        return None
    return torch.empty(nbytes // 4, dtype=torch.float32, device=device)


def _check_device(t: torch.Tensor, name: str) -> None:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def c2c_raw(x: torch.Tensor, out: torch.Tensor, n: int, batch: int, sign: int = -1,
            work: torch.Tensor | None = None, scale: float | None = None) -> None:
    """bfly_c2c on interleaved float32 buffers (no allocation except tables);
    bfly_c2c_scaled when `scale` is given."""
    lib = _lib.hip()
    tab = table(_lib.KIND_C2C, n, x.device)
    if work is None and lib.bfly_workspace_bytes(_lib.KIND_C2C, n, batch):
        work = workspace(_lib.KIND_C2C, n, batch, x.device)
    wp = work.data_ptr() if work is not None else None
    if scale is None:
        rc = lib.bfly_c2c(x.data_ptr(), out.data_ptr(), n, batch, sign, tab.data_ptr(), wp, _stream(x.device))
        if rc:
            _lib.check(rc, f"bfly_c2c(n={n}, batch={batch}, sign={sign})")
    else:
        rc = lib.bfly_c2c_scaled(x.data_ptr(), out.data_ptr(), n, batch, sign, scale, tab.data_ptr(), wp,
                                 _stream(x.device))
        if rc:
            _lib.check(rc, f"bfly_c2c_scaled(n={n}, batch={batch}, sign={sign}, scale={scale})")


def norm_scale(n: int, sign: int, norm: str | None) -> float | None:
    """Output scale of a length-n C2C under numpy's / torch.fft's `norm`
    conventions: "backward" (also None, their default) scales the inverse by
    1/n, "ortho" both directions by 1/sqrt(n), "forward" the forward by 1/n.
    Returns None where the transform is left unscaled."""
    norm = "backward" if norm is None else norm
    if norm not in ("backward", "ortho", "forward"):
        raise ValueError(f"norm must be None, 'backward', 'ortho' or 'forward', not {norm!r}")
    if norm == "ortho":
        return 1.0 / math.sqrt(n)
    inverse = sign > 0
    return 1.0 / n if (norm == "backward") == inverse else None


def c2c(x: torch.Tensor, sign: int = -1, norm: str | None = None, *, scale: float | None = None) -> torch.Tensor:
    """C2C along the last axis (sign -1 forward, +1 inverse) with numpy's
    `norm` convention (None = "backward": the forward unscaled, the inverse
    scaled 1/n).  `scale` instead gives the factor outright, applied in the
    transform's last store; scale=1.0 is the reference's unnormalised
    transform in both directions."""
    if norm is not None and scale is not None:
        raise ValueError("give norm or scale, not both")
    _check_device(x, "x")
    if x.dtype == torch.complex64:
        n = x.shape[-1]
        xf = torch.view_as_real(x)
    elif x.dtype == torch.float32:
        if x.shape[-1] % 2:
            raise ValueError("interleaved input needs an even last dimension")
        n = x.shape[-1] // 2
        xf = x
    else:
        raise TypeError("c2c takes complex64 or interleaved float32")
    batch = x.numel() // (n * (1 if x.dtype == torch.complex64 else 2))
    out = torch.empty_like(x)
    if batch == 0:
        return out
    of = torch.view_as_real(out) if x.dtype == torch.complex64 else out
    factor = norm_scale(n, sign, norm) if scale is None else scale
    c2c_raw(xf, of, n, batch, sign, scale=None if factor == 1.0 else factor)
    return out


def c2c_rows(x: torch.Tensor, sign: int = -1, out: torch.Tensor | None = None) -> torch.Tensor:
    """Batched C2C over the rows of a 2-D complex64 view whose rows may be
    padded: stride(1) == 1 and stride(0) >= n, e.g. buf[:, :n] of a
    [batch][dist] buffer.  The result goes to `out` (a [batch][n] view of the
    same kind, its gaps untouched) or to a new contiguous tensor
    (bfly_c2c_rows; unnormalised, sign -1 forward, +1 inverse)."""
    if not x.is_cuda:
        raise ValueError("x must be a device tensor")
    if x.dtype != torch.complex64 or x.dim() != 2:
        raise TypeError("c2c_rows takes a 2-D complex64 tensor")
    batch, n = x.shape
    if out is None:
        out = torch.empty((batch, n), dtype=torch.complex64, device=x.device)
    if out.dtype != torch.complex64 or tuple(out.shape) != (batch, n) or out.device != x.device:
        raise ValueError("out must be a complex64 [batch][n] tensor on the input's device")
    if batch == 0:
        return out
    for t, name in ((x, "x"), (out, "out")):
        if t.stride(1) != 1 or (batch > 1 and t.stride(0) < n):
            raise ValueError(f"{name} rows must be unit-stride and at least n apart")
    idist = x.stride(0) if batch > 1 else n
    odist = out.stride(0) if batch > 1 else n
    tab = table(_lib.KIND_C2C, n, x.device)
    work = workspace(_lib.KIND_C2C, n, batch, x.device) if n > 16384 else None
    rc = _lib.hip().bfly_c2c_rows(x.data_ptr(), idist, out.data_ptr(), odist, n, batch, sign, tab.data_ptr(),
                                  work.data_ptr() if work is not None else None, _stream(x.device))
    _lib.check(rc, f"bfly_c2c_rows(n={n}, batch={batch}, idist={idist}, odist={odist}, sign={sign})")
    return out


def table2d_cols(n1: int, device) -> torch.Tensor:
    """Device copy of the 2-D column table of n1 (bfly_fft2d_table_fill), cached."""
    device = torch.device(device)
    key = (device.index, "fft2d_cols", n1)
    t = _tables.get(key)
    if t is None:
        lib = _lib.hip()
        nf = lib.bfly_fft2d_table_floats(n1)
        if nf == 0:
            raise _lib.BflyError(_lib.ERR_SHAPE, f"no 2-D column kernel for n1={n1}")
        host = np.empty(nf, np.float32)
        _lib.check(lib.bfly_fft2d_table_fill(n1, host.ctypes.data), "bfly_fft2d_table_fill")
        t = torch.from_numpy(host).to(device)
        _tables[key] = t
    return t


def fft2_raw(x: torch.Tensor, out: torch.Tensor, n1: int, n2: int, batch: int, sign: int = -1,
             work: torch.Tensor | None = None) -> None:
    """bfly_fft2d on interleaved float32 buffers [batch][n1][2 n2]."""
    lib = _lib.hip()
    if work is None:
        work = torch.empty(max(1, lib.bfly_fft2d_workspace_bytes(n1, n2, batch) // 4), dtype=torch.float32,
                           device=x.device)
    rc = lib.bfly_fft2d(x.data_ptr(), out.data_ptr(), n1, n2, batch, sign,
                        table(_lib.KIND_C2C, n2, x.device).data_ptr(), table2d_cols(n1, x.device).data_ptr(),
                        work.data_ptr(), _stream(x.device))
    _lib.check(rc, f"bfly_fft2d(n1={n1}, n2={n2}, batch={batch}, sign={sign})")


def fft_columns(x: torch.Tensor, sign: int = -1) -> torch.Tensor:
    """Unnormalised C2C along axis -2 of a complex64 tensor [..., n, ncols]
    (element stride ncols) = torch.fft.fft(x, dim=-2) for sign -1."""
    _check_device(x, "x")
    if x.dtype != torch.complex64 or x.dim() < 2:
        raise TypeError("fft_columns takes a complex64 tensor [..., n, ncols]")
    n, ncols = x.shape[-2], x.shape[-1]
    batch = x.numel() // (n * ncols) if n * ncols else 0
    out = torch.empty_like(x)
    if batch == 0:
        return out
    lib = _lib.hip()
    nbytes = lib.bfly_c2c_columns_workspace_bytes(n, ncols, batch)
    work = torch.empty(nbytes // 4, dtype=torch.float32, device=x.device) if nbytes else None
    rc = lib.bfly_c2c_columns(x.data_ptr(), out.data_ptr(), n, ncols, batch, sign,
                              table2d_cols(n, x.device).data_ptr(), work.data_ptr() if work is not None else None,
                              _stream(x.device))
    _lib.check(rc, f"bfly_c2c_columns(n={n}, ncols={ncols}, batch={batch}, sign={sign})")
    return out


def fft2(x: torch.Tensor, sign: int = -1) -> torch.Tensor:
    """Unnormalised 2-D C2C over the last two axes of a complex64 tensor
    [..., n1, n2] (sign -1 forward, +1 inverse) = torch.fft.fft2 (norm
    "backward") for sign -1."""
    _check_device(x, "x")
    if x.dtype != torch.complex64 or x.dim() < 2:
        raise TypeError("fft2 takes a complex64 tensor [..., n1, n2]")
    n1, n2 = x.shape[-2], x.shape[-1]
    batch = x.numel() // (n1 * n2) if n1 * n2 else 0
    out = torch.empty_like(x)
    if batch == 0:
        return out
    fft2_raw(torch.view_as_real(x), torch.view_as_real(out), n1, n2, batch, sign)
    return out


def _real(kind: int, src: torch.Tensor, dst: torch.Tensor, n: int, batch: int, bins: int, stride: int,
          adjoint: bool, what: str) -> None:
    tab = table(kind, n, src.device)
    rc = _lib.hip().bfly_real_ex(kind, src.data_ptr(), dst.data_ptr(), n, batch, bins, stride,
                                 _lib.VARIANT_AUTO, _lib.REAL_ADJOINT if adjoint else 0, tab.data_ptr(),
                                 _stream(src.device))
    _lib.check(rc, f"{what}(n={n}, batch={batch}, bins={bins}, adjoint={adjoint})")


def rfft(x: torch.Tensor, adjoint: bool = False) -> torch.Tensor:
    """R2C along the last axis: float32 [..., n] -> half spectrum [..., n+2].

    adjoint=True returns the adjoint of irfft instead: bin k times v_k / n,
    v = 1 at bins 0 and n/2 (imaginary part 0) and 2 elsewhere."""
    _check_device(x, "x")
    if x.dtype != torch.float32:
        raise TypeError("rfft takes float32")
    n = x.shape[-1]
    batch = x.numel() // n if n else 0
    out = torch.empty(*x.shape[:-1], n + 2, dtype=torch.float32, device=x.device)
    if batch == 0:
        return out
    _real(_lib.KIND_R2C, x, out, n, batch, n // 2 + 1, n + 2, adjoint, "bfly_r2c")
    return out


def irfft(h: torch.Tensor, n: int, adjoint: bool = False) -> torch.Tensor:
    """C2R along the last axis: half spectrum [..., n+2] -> float32 [..., n], 1/n.

    adjoint=True returns the adjoint of rfft instead: n * irfft(w h), w = 1
    at bins 0 and n/2 and 1/2 elsewhere."""
    _check_device(h, "h")
    if h.dtype != torch.float32 or h.shape[-1] != n + 2:
        raise ValueError("irfft takes float32 [..., n+2]")
    batch = h.numel() // (n + 2)
    out = torch.empty(*h.shape[:-1], n, dtype=torch.float32, device=h.device)
    if batch == 0:
        return out
    _real(_lib.KIND_C2R, h, out, n, batch, n // 2 + 1, n + 2, adjoint, "bfly_c2r")
    return out


def rfft_modes(x: torch.Tensor, modes: int, adjoint: bool = False) -> torch.Tensor:
    """First `modes` bins of the R2C along the last axis: float32 [..., n] ->
    float32 [..., 2*modes] (interleaved); the other bins are never written.
    adjoint=True: the adjoint of irfft_modes (weights as for rfft)."""
    _check_device(x, "x")
    n = x.shape[-1]
    batch = x.numel() // n if n else 0
    out = torch.empty(*x.shape[:-1], 2 * modes, dtype=torch.float32, device=x.device)
    if batch == 0:
        return out
    _real(_lib.KIND_R2C, x, out, n, batch, modes, 2 * modes, adjoint, "bfly_r2c_modes")
    return out


def irfft_modes(h: torch.Tensor, n: int, adjoint: bool = False) -> torch.Tensor:
    """C2R of a spectrum given by its first modes bins (float32 [..., 2*modes],
    interleaved; every higher bin zero) -> float32 [..., n], 1/n.
    adjoint=True: the adjoint of rfft_modes (weights as for irfft)."""
    _check_device(h, "h")
    modes = h.shape[-1] // 2
    batch = h.numel() // (2 * modes)
    out = torch.empty(*h.shape[:-1], n, dtype=torch.float32, device=h.device)
    if batch == 0:
        return out
    _real(_lib.KIND_C2R, h, out, n, batch, modes, 2 * modes, adjoint, "bfly_c2r_modes")
    return out


def mode_mix(op: int, a: torch.Tensor, b: torch.Tensor, batch: int, c_in: int, c_out: int,
             modes: int) -> torch.Tensor:
    """bfly_mode_mix on complex64 tensors (mode index last, contiguous):
    op MIX_FORWARD   a = x [B, C_in, K],  b = w [C_in, C_out, K]  -> y  [B, C_out, K]
    op MIX_GRAD_INPUT  a = gy [B, C_out, K], b = w             -> gx [B, C_in, K]
    op MIX_GRAD_WEIGHT a = x [B, C_in, K], b = gy [B, C_out, K] -> gw [C_in, C_out, K]"""
    for t, name in ((a, "a"), (b, "b")):
        _check_device(t, name)
        if t.dtype != torch.complex64:
            raise TypeError(f"mode_mix takes complex64 ({name} is {t.dtype})")
    shape = {_lib.MIX_FORWARD: (batch, c_out, modes), _lib.MIX_GRAD_INPUT: (batch, c_in, modes),
             _lib.MIX_GRAD_WEIGHT: (c_in, c_out, modes)}.get(op)
    if shape is None:
        raise ValueError(f"unknown mode_mix op {op}")
    out = torch.empty(shape, dtype=torch.complex64, device=a.device)
    rc = _lib.hip().bfly_mode_mix(op, a.data_ptr(), b.data_ptr(), out.data_ptr(), batch, c_in, c_out, modes,
                                  _stream(a.device))
    _lib.check(rc, f"bfly_mode_mix(op={op}, batch={batch}, c_in={c_in}, c_out={c_out}, modes={modes})")
    return out


def probe_copy(src: torch.Tensor, dst: torch.Tensor) -> None:
    rc = _lib.hip().bfly_probe_copy(src.data_ptr(), dst.data_ptr(), src.numel() * src.element_size(),
                                    _stream(src.device))
    _lib.check(rc, "bfly_probe_copy")


def device_facts(device: int = 0) -> str:
    import ctypes
    buf = ctypes.create_string_buffer(1024)
    _lib.check(_lib.hip().bfly_probe_device(device, buf, 1024), "bfly_probe_device")
    return buf.value.decode()
