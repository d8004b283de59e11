"""GPU: adjoint-FFT autograd Functions and the spectral-convolution layer.

Gradients are checked against FP64 references on the CPU (torch.fft in
float64: central finite differences of the FP64 layer, and torch's own
autograd of the FP64 transforms), and at full size against torch.fft autograd.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


def _rel(a, b):
    a = np.asarray(a, np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b, np.complex128 if np.iscomplexobj(b) else np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n", [64, 1024, 4096])
def test_c2c_adjoint_matches_fp64_autograd(n):
    import torch
    from bfly import autograd as bag
    g = torch.Generator().manual_seed(0)
    x = torch.randn(3, n, dtype=torch.complex64, generator=g)
    up = torch.randn(3, n, dtype=torch.complex64, generator=g)
    xd = x.cuda().requires_grad_(True)
    y = bag.fft(xd)
    (y * up.cuda().conj()).real.sum().backward()
    x64 = x.to(torch.complex128).requires_grad_(True)
    (torch.fft.fft(x64) * up.to(torch.complex128).conj()).real.sum().backward()
    assert _rel(y.detach().cpu().numpy(), torch.fft.fft(x.to(torch.complex128)).numpy()) < 1e-6
    assert _rel(xd.grad.cpu().numpy(), x64.grad.numpy()) < 1e-6


@pytest.mark.parametrize("n", [128, 1024, 4096])
def test_rfft_irfft_adjoints_match_fp64_autograd(n):
    import torch
    from bfly import autograd as bag
    g = torch.Generator().manual_seed(1)
    x = torch.randn(4, n, generator=g)
    up = torch.randn(4, n // 2 + 1, dtype=torch.complex64, generator=g)
    xd = x.cuda().requires_grad_(True)
    (bag.rfft(xd) * up.cuda().conj()).real.sum().backward()
    x64 = x.double().requires_grad_(True)
    (torch.fft.rfft(x64) * up.to(torch.complex128).conj()).real.sum().backward()
    assert _rel(xd.grad.cpu().numpy(), x64.grad.numpy()) < 1e-6

    h = torch.randn(4, n // 2 + 1, dtype=torch.complex64, generator=g)
    upr = torch.randn(4, n, generator=g)
    hd = h.cuda().requires_grad_(True)
    (bag.irfft(hd, n) * upr.cuda()).sum().backward()
    h64 = h.to(torch.complex128).requires_grad_(True)
    (torch.fft.irfft(h64, n) * upr.double()).sum().backward()
    assert _rel(hd.grad.cpu().numpy(), h64.grad.numpy()) < 1e-6


@pytest.mark.parametrize("n,modes", [(64, 1), (256, 32), (4096, 256), (4096, 2049), (8192, 100)])
def test_mode_truncated_real_transforms(n, modes):
    # bfly_r2c_ex / bfly_c2r_ex: first `modes` bins of the R2C; C2R of a
    # spectrum whose higher bins are zero.  Checked against the FP64 oracle.
    import torch
    import oracle
    from bfly import ops
    x = oracle.uniform_input(6, (5, n))
    got = ops.rfft_modes(torch.from_numpy(x).cuda(), modes).cpu().numpy()
    ref = oracle.r2c_f64(x, n)[:, :2 * modes]
    assert oracle.rel_l2_rows(got, np.ascontiguousarray(ref)) <= 1e-5
    h = oracle.uniform_input(7, (5, 2 * modes))
    if modes == n // 2 + 1:
        h[:, -1] = 0.0
    full = np.zeros((5, n + 2), np.float32)
    full[:, :2 * modes] = h
    if n <= 4096:
        back = ops.irfft_modes(torch.from_numpy(h).cuda(), n).cpu().numpy()
        assert oracle.rel_l2_rows(back, oracle.c2r_f64(full, n)) <= 1e-5


def _fp64_layer(x, w):
    import torch
    K = w.shape[-1]
    X = torch.fft.rfft(x)
    Y = torch.einsum("bik,iok->bok", X[..., :K], w)
    full = torch.zeros(x.shape[0], w.shape[1], x.shape[-1] // 2 + 1, dtype=Y.dtype)
    full[..., :K] = Y
    return torch.fft.irfft(full, x.shape[-1])


def test_spectral_layer_gradients_fp64_finite_differences():
    # B=2, C=4, N=256 (the differentiation pattern's FD scale), K=32 modes.
    import torch
    from bfly.spectral import spectral_conv1d
    B, C, N, K = 2, 4, 256, 32
    g = torch.Generator().manual_seed(4)
    x = torch.randn(B, C, N, generator=g)
    w = torch.complex(torch.randn(C, C, K, generator=g), torch.randn(C, C, K, generator=g)) / 8
    up = torch.randn(B, C, N, generator=g)
    xd = x.cuda().requires_grad_(True)
    wd = w.cuda().requires_grad_(True)
    (spectral_conv1d(xd, wd) * up.cuda()).sum().backward()
    x64, w64, up64 = x.double(), w.to(torch.complex128), up.double()

    def loss(xx, ww):
        return float((_fp64_layer(xx, ww) * up64).sum())
    eps = 1e-6
    for trial in range(4):
        vx = torch.randn(B, C, N, generator=g, dtype=torch.float64)
        fd = (loss(x64 + eps * vx, w64) - loss(x64 - eps * vx, w64)) / (2 * eps)
        an = float((xd.grad.double().cpu() * vx).sum())
        assert abs(fd - an) <= 1e-4 * max(1.0, abs(fd)), (trial, fd, an)
        vw = torch.complex(torch.randn(C, C, K, generator=g, dtype=torch.float64),
                           torch.randn(C, C, K, generator=g, dtype=torch.float64))
        fd = (loss(x64, w64 + eps * vw) - loss(x64, w64 - eps * vw)) / (2 * eps)
        # torch's complex gradient: dL = Re(sum(conj(grad) * dw))
        an = float((wd.grad.to(torch.complex128).cpu().conj() * vw).real.sum())
        assert abs(fd - an) <= 1e-4 * max(1.0, abs(fd)), (trial, fd, an)


@pytest.mark.parametrize("n,modes", [(256, 32), (4096, 256)])
def test_mode_truncated_adjoints_match_fp64_autograd(n, modes):
    import torch
    from bfly import autograd as bag
    g = torch.Generator().manual_seed(2)
    x = torch.randn(4, n, generator=g)
    up = torch.randn(4, modes, dtype=torch.complex64, generator=g)
    xd = x.cuda().requires_grad_(True)
    (bag.rfft_modes(xd, modes) * up.cuda().conj()).real.sum().backward()
    x64 = x.double().requires_grad_(True)
    (torch.fft.rfft(x64)[..., :modes] * up.to(torch.complex128).conj()).real.sum().backward()
    assert _rel(xd.grad.cpu().numpy(), x64.grad.numpy()) < 1e-6
    h = torch.randn(4, modes, dtype=torch.complex64, generator=g)
    upr = torch.randn(4, n, generator=g)
    hd = h.cuda().requires_grad_(True)
    (bag.irfft_modes(hd, n) * upr.cuda()).sum().backward()
    h64 = h.to(torch.complex128).requires_grad_(True)
    pad = torch.zeros(4, n // 2 + 1 - modes, dtype=torch.complex128)
    (torch.fft.irfft(torch.cat([h64, pad], -1), n) * upr.double()).sum().backward()
    assert _rel(hd.grad.cpu().numpy(), h64.grad.numpy()) < 1e-6


def test_spectral_layer_full_size_matches_torch_fft():
    # cfg: x [64, 128, 4096] ~ N(0,1), W [128, 128, 256] ~ N(0, 1/256) per part, seed 4.
    import torch
    from bfly.spectral import spectral_conv1d, spectral_conv1d_full, torch_fft_reference
    torch.manual_seed(4)
    B, C, N, K = 64, 128, 4096, 256
    x = torch.randn(B, C, N, device="cuda")
    s = (1.0 / (C * 2)) ** 0.5
    w = torch.complex(torch.randn(C, C, K, device="cuda") * s, torch.randn(C, C, K, device="cuda") * s)
    up = torch.randn(B, C, N, device="cuda")
    xa, wa = x.clone().requires_grad_(True), w.clone().requires_grad_(True)
    ya = spectral_conv1d(xa, wa)
    (ya * up).sum().backward()
    xb, wb = x.clone().requires_grad_(True), w.clone().requires_grad_(True)
    yb = torch_fft_reference(xb, wb)
    (yb * up).sum().backward()
    torch.cuda.synchronize()
    assert _rel(ya.detach().cpu().numpy(), yb.detach().cpu().numpy()) < 1e-5
    assert _rel(xa.grad.cpu().numpy(), xb.grad.cpu().numpy()) < 1e-5
    assert _rel(wa.grad.cpu().numpy(), wb.grad.cpu().numpy()) < 1e-5
    xc, wc = x.clone().requires_grad_(True), w.clone().requires_grad_(True)
    yc = spectral_conv1d_full(xc, wc)
    (yc * up).sum().backward()
    assert _rel(yc.detach().cpu().numpy(), yb.detach().cpu().numpy()) < 1e-5
    assert _rel(xc.grad.cpu().numpy(), xb.grad.cpu().numpy()) < 1e-5
