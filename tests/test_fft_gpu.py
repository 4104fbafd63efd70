"""The path's 3-D transforms (ettk_fft3) against numpy's FP64 FFTs.

r2c = numpy.fft.rfftn over the three axes; c2r = irfftn * n^3 (FFTW's unnormalised backward transform,
nbfFourierFilter.h:270-358); the two fused CCC transforms against their composition in FP64:
    ETTK_FFT_MASK:   r2c(c2r(X) * m)
    ETTK_FFT_SQUARE: r2c(c2r(X)^2 * (-1)^(i+j+k))
Bars: relative L2 error below 5e-6 over the whole result, and separately over the k = 0 and k = n/2 planes of a
half spectrum (the real-valued DC and Nyquist columns of every row, which the plane kernels handle apart from
the other columns), so a mistake confined to those planes cannot hide behind the rest of the volume.  FP32
transforms of these lengths are accurate to about 1e-7 relative; the inputs are preserved.
"""
import numpy as np
import pytest

BAR = 5e-6


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _check_half(got, ref):
    assert _rel(got, ref) < BAR, _rel(got, ref)
    h = ref.shape[-1] - 1
    for k in (0, h):
        assert _rel(got[..., k], ref[..., k]) < BAR, (k, _rel(got[..., k], ref[..., k]))


def _checker(n):
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return np.where((i + j + k) & 1, -1.0, 1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("n,batch", [(16, 3), (32, 2), (48, 2), (64, 3), (96, 1), (128, 1), (192, 1), (256, 1)])
def test_r2c_and_c2r_match_numpy(torch_gpu, ctx, n, batch):
    torch = torch_gpu
    print(f"[fft] r2c / c2r n={n} batch={batch}", flush=True)
    g = np.random.default_rng(n)
    x = g.standard_normal((batch, n, n, n)).astype(np.float32)
    h = n // 2 + 1
    dx = torch.from_numpy(x).cuda()
    dX = torch.empty((batch, n, n, h), dtype=torch.complex64, device="cuda")
    ctx.fft3(ctx.FFT_R2C, n, batch, dx, dX)
    ref = np.fft.rfftn(x.astype(np.float64), axes=(1, 2, 3))
    _check_half(dX.cpu().numpy(), ref)
    assert np.array_equal(dx.cpu().numpy(), x)
    # c2r of the spectrum of a real volume
    Xc = ref.astype(np.complex64)
    dXc = torch.from_numpy(Xc).cuda()
    dy = torch.empty((batch, n, n, n), dtype=torch.float32, device="cuda")
    ctx.fft3(ctx.FFT_C2R, n, batch, dXc, dy)
    yref = np.fft.irfftn(Xc.astype(np.complex128), s=(n, n, n), axes=(1, 2, 3)) * float(n) ** 3
    assert _rel(dy.cpu().numpy(), yref) < BAR
    assert np.array_equal(dXc.cpu().numpy(), Xc)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [16, 32, 48, 64])
def test_fused_transforms_match_numpy(torch_gpu, ctx, n):
    torch = torch_gpu
    print(f"[fft] fused mask / square n={n}", flush=True)
    batch = 2
    g = np.random.default_rng(100 + n)
    x = g.standard_normal((batch, n, n, n))
    X = np.fft.rfftn(x, axes=(1, 2, 3)).astype(np.complex64)
    m = g.uniform(0.0, 1.0, (n, n, n)).astype(np.float32)
    r = np.fft.irfftn(X.astype(np.complex128), s=(n, n, n), axes=(1, 2, 3)) * float(n) ** 3
    dX = torch.from_numpy(X).cuda()
    dm = torch.from_numpy(m).cuda()
    out = torch.empty_like(dX)
    ctx.fft3(ctx.FFT_MASK, n, batch, dX, out, dm)
    _check_half(out.cpu().numpy(), np.fft.rfftn(r * m.astype(np.float64), axes=(1, 2, 3)))
    ctx.fft3(ctx.FFT_SQUARE, n, batch, dX, out)
    _check_half(out.cpu().numpy(), np.fft.rfftn(r * r * _checker(n), axes=(1, 2, 3)))
    assert np.array_equal(dX.cpu().numpy(), X)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [16, 48, 64, 128])
def test_c2r_of_a_non_hermitian_spectrum_keeps_the_real_part(torch_gpu, ctx, n):
    """A half spectrum that is not exactly Hermitian on its k = 0 and k = n/2 planes (as the CCC's wedge-weighted
    spectra are near the borders) transforms as FFTW's and numpy's c2r do: complex inverse transforms along i and
    j, then the half-complex transform along k with the imaginary parts of those two planes dropped."""
    torch = torch_gpu
    print(f"[fft] c2r of a non-Hermitian spectrum n={n}", flush=True)
    g = np.random.default_rng(200 + n)
    h = n // 2 + 1
    X = (g.standard_normal((1, n, n, h)) + 1j * g.standard_normal((1, n, n, h))).astype(np.complex64)
    r = np.fft.irfftn(X.astype(np.complex128), s=(n, n, n), axes=(1, 2, 3)) * float(n) ** 3
    dX = torch.from_numpy(X).cuda()
    dy = torch.empty((1, n, n, n), dtype=torch.float32, device="cuda")
    ctx.fft3(ctx.FFT_C2R, n, 1, dX, dy)
    assert _rel(dy.cpu().numpy(), r) < BAR
    if n <= 64:
        out = torch.empty_like(dX)
        ctx.fft3(ctx.FFT_SQUARE, n, 1, dX, out)
        _check_half(out.cpu().numpy(), np.fft.rfftn(r * r * _checker(n), axes=(1, 2, 3)))
        m = g.uniform(0.0, 1.0, (n, n, n)).astype(np.float32)
        ctx.fft3(ctx.FFT_MASK, n, 1, dX, out, torch.from_numpy(m).cuda())
        _check_half(out.cpu().numpy(), np.fft.rfftn(r * m.astype(np.float64), axes=(1, 2, 3)))


@pytest.mark.gpu
def test_unsupported_requests_are_rejected(torch_gpu, ctx):
    torch = torch_gpu
    from ettk_amd.lib import EttkError

    d = torch.zeros((1, 80, 80, 41), dtype=torch.complex64, device="cuda")
    with pytest.raises(EttkError):
        ctx.fft3(ctx.FFT_R2C, 80, 1, d, d)
    with pytest.raises(EttkError):
        ctx.fft3(ctx.FFT_SQUARE, 96, 1, d, d)
    with pytest.raises(EttkError):
        ctx.fft3(ctx.FFT_MASK, 64, 1, d, d)  # no mask
