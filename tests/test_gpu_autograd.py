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


@pytest.mark.parametrize("norm", [None, "backward", "ortho", "forward"])
@pytest.mark.parametrize("n", [64, 4096, 16384])
def test_normalised_c2c_gradients_match_torch_fft_fp64(n, norm):
    # fft / ifft with torch.fft's `norm` conventions: values and gradients
    # against torch.fft.fft / ifft autograd in FP64 on the CPU.
    import torch
    from bfly import autograd as bag
    g = torch.Generator().manual_seed(1)
    x = torch.randn(3, n, dtype=torch.complex64, generator=g)
    up = torch.randn(3, n, dtype=torch.complex64, generator=g)
    for ours, ref in ((bag.fft, torch.fft.fft), (bag.ifft, torch.fft.ifft)):
        xd = x.cuda().requires_grad_(True)
        y = ours(xd, norm=norm)
        (y * up.cuda().conj()).real.sum().backward()
        x64 = x.to(torch.complex128).requires_grad_(True)
        y64 = ref(x64, norm=norm)
        (y64 * up.to(torch.complex128).conj()).real.sum().backward()
        assert _rel(y.detach().cpu().numpy(), y64.detach().numpy()) < 1e-6, (ours.__name__, norm)
        assert _rel(xd.grad.cpu().numpy(), x64.grad.numpy()) < 1e-6, (ours.__name__, norm)


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


def _real_ex(kind, src, dst, n, b, bins, stride, variant, flags):
    import torch
    from bfly import _lib, ops
    rc = _lib.hip().bfly_real_ex(kind, src.data_ptr(), dst.data_ptr(), n, b, bins, stride, variant, flags,
                                 ops.table(kind, n, src.device).data_ptr(),
                                 torch.cuda.current_stream().cuda_stream)
    assert rc == 0, (kind, n, b, bins, variant, flags, rc)


@pytest.mark.parametrize("n", [32, 256, 1024, 4096, 8192])
def test_adjoint_real_kernels_equal_weighted_forward_bitwise(n):
    # bfly_real_ex with BFLY_REAL_ADJOINT in every kernel form equals the
    # forward launch of the same form with the adjoint weights applied
    # outside (all powers of two, so exact), bit for bit, full and truncated.
    import torch
    import oracle
    from bfly import _lib
    M = n // 2
    for b, bins in ((3, M + 1), (257, M + 1), (130, max(1, M // 8))):
        x = torch.from_numpy(oracle.uniform_input(40 + b, (b, n))).cuda()
        h = torch.from_numpy(oracle.uniform_input(50 + b, (b, 2 * bins))).cuda()
        v = torch.full((bins,), 1.0 / M, device="cuda")
        v[0] = 0.5 / M
        if bins == M + 1:
            v[M] = 0.5 / M
        w = torch.full((bins,), 0.5, device="cuda")
        w[0] = 1.0
        if bins == M + 1:
            w[M] = 1.0
        for variant in (_lib.VARIANT_PLAIN, _lib.VARIANT_STREAM, _lib.VARIANT_STREAM_LT, _lib.VARIANT_AUTO):
            fwd = torch.full((b, 2 * bins), float("nan"), device="cuda")
            adj = torch.full((b, 2 * bins), float("nan"), device="cuda")
            _real_ex(1, x, fwd, n, b, bins, 2 * bins, variant, 0)
            _real_ex(1, x, adj, n, b, bins, 2 * bins, variant, _lib.REAL_ADJOINT)
            want = (fwd.view(b, bins, 2) * v[:, None]).reshape(b, 2 * bins)
            want[:, 1] = 0.0
            if bins == M + 1:
                want[:, 2 * M + 1] = 0.0
            torch.cuda.synchronize()
            assert torch.equal(adj, want), ("r2c", n, b, bins, variant)
            hw = (h.view(b, bins, 2) * w[:, None]).reshape(b, 2 * bins).contiguous()
            back = torch.full((b, n), float("nan"), device="cuda")
            badj = torch.full((b, n), float("nan"), device="cuda")
            _real_ex(2, hw, back, n, b, bins, 2 * bins, variant, 0)
            _real_ex(2, h, badj, n, b, bins, 2 * bins, variant, _lib.REAL_ADJOINT)
            torch.cuda.synchronize()
            assert torch.equal(badj, back * n), ("c2r", n, b, bins, variant)


@pytest.mark.parametrize("n,bins", [(64, 33), (1024, 513), (1024, 40), (8192, 4097)])
def test_adjoint_real_kernels_satisfy_the_adjoint_identity_fp64(n, bins):
    # <C2R h, g> = <h, R2C_adj g> and <R2C x, k> = <x, C2R_adj k> with the
    # left sides from the FP64 oracle (real inner products; complex bins as
    # pairs of reals).  A wrong weight, a missed edge or a wrong scale breaks
    # them.
    import torch
    import oracle
    from bfly import _lib
    b = 6
    M = n // 2
    h = oracle.uniform_input(60, (b, 2 * bins)).astype(np.float64)
    h[:, 1] = 0.0  # the C2R reads only the real part of bin 0 (and of bin M)
    if bins == M + 1:
        h[:, 2 * M + 1] = 0.0
    g = oracle.uniform_input(61, (b, n)).astype(np.float64)
    full = np.zeros((b, n + 2), np.float64)
    full[:, :2 * bins] = h
    lhs = np.sum(oracle.c2r_f64(full.astype(np.float32), n).astype(np.float64) * g, axis=1)
    ga = torch.from_numpy(g.astype(np.float32)).cuda()
    adj = torch.empty(b, 2 * bins, device="cuda")
    _real_ex(1, ga, adj, n, b, bins, 2 * bins, _lib.VARIANT_AUTO, _lib.REAL_ADJOINT)
    rhs = np.sum(h * adj.cpu().numpy().astype(np.float64), axis=1)
    scale = np.sqrt(np.sum(lhs ** 2))
    assert np.max(np.abs(lhs - rhs)) <= 1e-5 * scale, (lhs, rhs)

    x = oracle.uniform_input(62, (b, n)).astype(np.float64)
    k = oracle.uniform_input(63, (b, 2 * bins)).astype(np.float64)
    spec = oracle.r2c_f64(x.astype(np.float32), n).astype(np.float64)[:, :2 * bins]
    lhs = np.sum(spec * k, axis=1)
    ka = torch.from_numpy(k.astype(np.float32)).cuda()
    cadj = torch.empty(b, n, device="cuda")
    _real_ex(2, ka, cadj, n, b, bins, 2 * bins, _lib.VARIANT_AUTO, _lib.REAL_ADJOINT)
    rhs = np.sum(x * cadj.cpu().numpy().astype(np.float64), axis=1)
    scale = np.sqrt(np.sum(lhs ** 2))
    assert np.max(np.abs(lhs - rhs)) <= 1e-5 * scale, (lhs, rhs)


@pytest.mark.parametrize("B,C_in,C_out,K", [(2, 4, 4, 32), (3, 5, 7, 19), (17, 40, 24, 37), (64, 128, 128, 256)])
def test_mode_mix_kernel_matches_fp64(B, C_in, C_out, K):
    # bfly_mode_mix forward / input gradient / weight gradient against
    # complex128 einsums of the same operands (ragged tiles at every edge)
    import torch
    from bfly import _lib, ops
    g = torch.Generator().manual_seed(B * 1000 + K)

    def crand(*shape):
        return torch.complex(torch.randn(*shape, generator=g), torch.randn(*shape, generator=g))
    x, w, gy = crand(B, C_in, K), crand(C_in, C_out, K), crand(B, C_out, K)
    x64, w64, gy64 = x.to(torch.complex128), w.to(torch.complex128), gy.to(torch.complex128)
    want = {_lib.MIX_FORWARD: torch.einsum("bik,iok->bok", x64, w64),
            _lib.MIX_GRAD_INPUT: torch.einsum("bok,iok->bik", gy64, w64.conj()),
            _lib.MIX_GRAD_WEIGHT: torch.einsum("bik,bok->iok", x64.conj(), gy64)}
    args = {_lib.MIX_FORWARD: (x, w), _lib.MIX_GRAD_INPUT: (gy, w), _lib.MIX_GRAD_WEIGHT: (x, gy)}
    for op, (a, b) in args.items():
        out = ops.mode_mix(op, a.cuda(), b.cuda(), B, C_in, C_out, K)
        torch.cuda.synchronize()
        assert _rel(out.cpu().numpy(), want[op].numpy()) < 1e-6, op
    # the output must not overlap an input; unknown ops are refused
    a, b = x.cuda(), w.cuda()
    assert _lib.hip().bfly_mode_mix(0, a.data_ptr(), b.data_ptr(), a.data_ptr(), B, C_in, C_out, K,
                                    None) == _lib.ERR_SHAPE
    assert _lib.hip().bfly_mode_mix(3, a.data_ptr(), b.data_ptr(), a.data_ptr(), B, C_in, C_out, K,
                                    None) == _lib.ERR_SHAPE


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
