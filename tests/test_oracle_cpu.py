"""CPU checks of the FP64 oracle: golden vectors of soft-1.0, the compiled soft-1.0 checker (when present),
and closed forms for the parts no fixture covers.  No GPU is used."""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

from ettk_amd.abi import Particle, default_params
from oracle_lib import OraclePlan, oracle, s2_forward, so3_correlate, softref, vp

GOLD = Path(__file__).resolve().parent / "golden" / "soft"


def _wigner_D_grid(B: int, l: int, m1: int, m2: int) -> np.ndarray:
    """Normalized Wigner D sqrt((2l+1)/8pi^2) e^{-i m1 a} d^l_{m1 m2}(b) e^{-i m2 g} on the SO(3) grid,
    [beta][alpha][gamma], from the oracle's d^l tables."""
    o = oracle()
    n, W = 2 * B, 2 * B - 1
    a = 2 * np.pi * np.arange(n) / n
    out = np.zeros((n, n, n), complex)
    for kb in range(n):
        d = np.zeros(B * W * W)
        o.orc_wigner_d(B, math.pi * (2 * kb + 1) / (4 * B), vp(d))
        dv = d[(l * W + (m1 + B - 1)) * W + (m2 + B - 1)]
        out[kb] = np.exp(-1j * m1 * a)[:, None] * dv * np.exp(-1j * m2 * a)[None, :]
    return out * math.sqrt((2 * l + 1) / (8 * math.pi ** 2))


def _load_cx(stem: str) -> np.ndarray:
    return (np.loadtxt(GOLD / f"{stem}real_bw4.dat") + 1j * np.loadtxt(GOLD / f"{stem}imag_bw4.dat")).reshape(8, 8, 8)


# ------------------------------------------------------------------------------------- soft-1.0 fixtures
def test_wigner_d_matches_soft_fixtures():
    """soft-1.0's sampled D^1_{0,1} (x (2+i)) and D^3_{-1,1} at bw = 4 pin the sign convention, the
    normalisation sqrt((2l+1)/8 pi^2) and the [beta][alpha][gamma] sample order of the oracle's Wigner-d."""
    assert np.abs(_load_cx("D101") - (2 + 1j) * _wigner_D_grid(4, 1, 0, 1)).max() < 1e-14
    assert np.abs(_load_cx("D3-11") - _wigner_D_grid(4, 3, -1, 1)).max() < 1e-14


def test_wigner_sum_matches_soft_fixture():
    """dSum = (2 + sqrt2 i) D^1_{1,0} + (-sqrt5 + 11 i) D^2_{2,2} + (7 + sqrt3 i) D^3_{0,-2}."""
    want = ((2 + math.sqrt(2) * 1j) * _wigner_D_grid(4, 1, 1, 0) + (-math.sqrt(5) + 11j) * _wigner_D_grid(4, 2, 2, 2)
            + (7 + math.sqrt(3) * 1j) * _wigner_D_grid(4, 3, 0, -2))
    assert np.abs(_load_cx("dSum_") - want).max() < 1e-12


def test_so3_correlation_matches_soft_output():
    """The oracle's SO(3) correlation of soft-1.0's two random bw = 8 signals equals soft-1.0's own output."""
    sig = np.loadtxt(GOLD / "randomS2sig_bw8.dat")[0::2]
    pat = np.loadtxt(GOLD / "randomS2sigA_bw8.dat")[0::2]
    want = np.load(GOLD / "correlate_bw8.npy")  # [beta][alpha][gamma]
    got = so3_correlate(8, sig, pat).transpose(1, 0, 2)  # oracle: [alpha][beta][gamma]
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()
    assert np.argmax(got) == np.argmax(want)


# ------------------------------------------------------------------------------------- compiled soft-1.0
def _bandlimited(B: int, seed: int) -> np.ndarray:
    o = oracle()
    g = np.random.default_rng(seed)
    re, im = g.standard_normal(B * B), g.standard_normal(B * B)
    # real signal: coefficient (l, -m) = (-1)^m conj(l, m); enforce through a forward/inverse pass
    f = np.zeros((2 * B) * (2 * B))
    o.orc_s2_inverse(B, vp(re), vp(im), vp(f))
    return np.ascontiguousarray(f)


@pytest.mark.skipif(softref() is None, reason="soft-1.0 checker not built (reference tree absent)")
@pytest.mark.parametrize("B", [4, 8, 16])
def test_so3_correlation_matches_compiled_soft(B):
    r = softref()
    sig, pat = _bandlimited(B, 1 + B), _bandlimited(B, 2 + B)
    n = 2 * B
    want = np.zeros(n ** 3)
    r.ref_soft_correlate(B, vp(sig), vp(pat), vp(want))
    got = so3_correlate(B, sig, pat).transpose(1, 0, 2).ravel()
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


@pytest.mark.skipif(softref() is None, reason="soft-1.0 checker not built (reference tree absent)")
def test_s2_rotation_matches_compiled_soft():
    r = softref()
    o = oracle()
    B = 8
    f = _bandlimited(B, 7)
    want = np.zeros_like(f)
    got = np.zeros_like(f)
    r.ref_rotate(B, vp(f), vp(want), 0.3, 1.1, -0.7)
    o.orc_s2_rotate(B, vp(f), vp(got), 0.3, 1.1, -0.7)
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


# ------------------------------------------------------------------------------------- closed forms
def test_s2_transform_round_trip_and_orthonormality():
    """Forward(inverse(e_lm)) on the Driscoll-Healy grid returns the coefficients of Re(Y_l^m)
    = (Y_l^m + (-1)^m Y_l^-m) / 2 (the synthesis keeps the real part); coefficients are stored at l^2 + l + m."""
    o = oracle()
    B = 8
    ix = lambda l, m: l * l + l + m  # noqa: E731
    for (l, m) in [(0, 0), (4, 0), (3, 2), (7, -5), (5, 1)]:
        re, im = np.zeros(B * B), np.zeros(B * B)
        re[ix(l, m)] = 1.0
        f = np.zeros(4 * B * B)
        o.orc_s2_inverse(B, vp(re), vp(im), vp(f))
        c = s2_forward(B, f)
        e = np.zeros(B * B, complex)
        if m == 0:
            e[ix(l, 0)] = 1.0
        else:
            e[ix(l, m)] = 0.5
            e[ix(l, -m)] = 0.5 * (-1) ** m
        assert np.abs(c - e).max() < 1e-12, (l, m)


def test_s2_rotation_preserves_norm_and_inverts():
    o = oracle()
    B = 8
    f = _bandlimited(B, 11)
    g = np.zeros_like(f)
    h = np.zeros_like(f)
    a, b, c = 0.4, 0.9, -1.3
    o.orc_s2_rotate(B, vp(f), vp(g), a, b, c)
    o.orc_s2_rotate(B, vp(g), vp(h), -c, -b, -a)  # R(a,b,c)^-1 = R(-c,-b,-a)
    assert np.abs(h - f).max() < 1e-10 * np.abs(f).max()
    cf, cg = s2_forward(B, f), s2_forward(B, g)
    assert abs(np.linalg.norm(cf) - np.linalg.norm(cg)) < 1e-10 * np.linalg.norm(cf)


def test_so3_correlation_peaks_at_the_rotation():
    """Correlating f with f rotated by a grid rotation g0 peaks at the bin of g0."""
    o = oracle()
    B = 8
    n = 2 * B
    f = _bandlimited(B, 21)
    ia, ib, ig = 3, 5, 12
    al, be, ga = 2 * np.pi * ia / n, np.pi * (2 * ib + 1) / (4 * B), 2 * np.pi * ig / n
    g = np.zeros_like(f)
    o.orc_s2_rotate(B, vp(f), vp(g), al, be, ga)
    Cc = so3_correlate(B, g, f)  # [alpha][beta][gamma]
    assert np.unravel_index(np.argmax(Cc), Cc.shape) == (ia, ib, ig)


def test_so3_scaled_correlation_takes_the_weighted_inner_products():
    """softScaled (ProjRot.h:1457-1589) for g = R(g0) f: the correlation peaks at g0's bin, the pattern rotated to
    that bin is g itself (up to the float rounding of the bin's angles), so the rescaled maximum is the
    sin(theta)-weighted inner product of g with itself; the rescaled volume is an increasing affine map of the
    correlation."""
    o = oracle()
    B = 8
    n = 2 * B
    f = _bandlimited(B, 23)
    ia, ib, ig = 5, 2, 9
    al, be, ga = 2 * np.pi * ia / n, np.pi * (2 * ib + 1) / (4 * B), 2 * np.pi * ig / n
    g = np.zeros_like(f)
    o.orc_s2_rotate(B, vp(f), vp(g), al, be, ga)
    Cc = so3_correlate(B, g, f)
    Cs = so3_correlate(B, g, f, scaled=True)
    assert np.unravel_index(np.argmax(Cs), Cs.shape) == (ia, ib, ig)
    th = np.sin(np.pi * (2 * np.arange(n) + 1) / (4 * B))
    gg = g.reshape(n, n)
    wself = float(np.sum(gg * gg * th[:, None]))
    assert abs(Cs.max() - wself) <= 1e-6 * abs(wself), (Cs.max(), wself)
    slope, icpt = np.polyfit(Cc.ravel(), Cs.ravel(), 1)
    assert slope > 0
    assert np.max(np.abs(Cs - (slope * Cc + icpt))) <= 1e-9 * (Cs.max() - Cs.min())


def test_local_maxima_brute_force():
    o = oracle()
    n = 10
    g = np.random.default_rng(3)
    Cv = g.uniform(-1, 0, (n, n, n))
    peaks = [(1, 2, 3), (5, 5, 5), (9, 0, 9), (0, 9, 4)]
    for k, p in enumerate(peaks):
        Cv[p] = 1.0 + k
    out = np.zeros(n ** 3)
    o.orc_local_maxima(vp(np.ascontiguousarray(Cv.ravel())), n, vp(out))
    out = out.reshape(n, n, n)
    # brute force: v >= all 26 neighbours (border -FLT_MAX)
    P = np.pad(Cv, 1, constant_values=-np.finfo(np.float32).max)
    want = np.zeros_like(Cv)
    for a in range(n):
        for b in range(n):
            for c in range(n):
                blk = P[a:a + 3, b:b + 3, c:c + 3]
                if Cv[a, b, c] >= blk.max():
                    want[a, b, c] = Cv[a, b, c]
    assert np.array_equal(out, want)
    for p in peaks:
        assert out[p] == Cv[p]


def test_powell_restatement_on_quadratics():
    """One Powell sweep (maxfev = 1) with the library's line-search settings minimises a separable quadratic
    to the line-search tolerance; a coupled one still decreases strictly."""
    o = oracle()
    o.orc_powell_quadratic.restype = C.c_int
    o.orc_powell_quadratic.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_double, C.c_double,
                                       C.c_double, C.c_void_p, C.c_void_p]
    a = np.array([1.0, 2.0, 0.5])
    c = np.array([3.0, -2.0, 7.5])
    x0 = np.zeros(3)
    xm = np.zeros(3)
    fm = C.c_double()
    nev = o.orc_powell_quadratic(vp(a), vp(c), 0.0, vp(x0), 5.0, 1e-2, 1e-1, vp(xm), C.byref(fm))
    assert np.abs(xm - c).max() < 5e-2
    assert fm.value < 1e-3
    assert 6 <= nev <= 80
    f0 = float(np.sum(a * c * c) + 0.8 * c[0] * c[1])
    o.orc_powell_quadratic(vp(a), vp(c), 0.8, vp(x0), 5.0, 1e-2, 1e-1, vp(xm), C.byref(fm))
    assert fm.value < 0.05 * f0


_FN3 = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double))


def _powell_fn(f, x0, step=5.0, xtol=1e-2, ftol=1e-1):
    """The oracle's Powell sweep (maxfev = 1) on a Python function of three variables."""
    o = oracle()
    o.orc_powell_fn.restype = C.c_int
    o.orc_powell_fn.argtypes = [_FN3, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    cb = _FN3(lambda p: float(f(np.array([p[0], p[1], p[2]]))))
    x = np.ascontiguousarray(x0, np.float64)
    xm = np.zeros(3)
    fm = C.c_double()
    nev = o.orc_powell_fn(cb, vp(x), step, xtol, ftol, vp(xm), C.byref(fm))
    return xm, fm.value, nev


def _exact_line_min(f, p, d):
    """Minimiser of a quadratic f along p + s d, from three exact samples."""
    f0, fp, fm = f(p), f(p + d), f(p - d)
    return p - 0.5 * (fp - fm) / (fp + fm - 2 * f0) * d


def _closed_form_sweep(f, x0, ftol=1e-1, order=(0, 1, 2), last_search=True):
    """One Powell iteration with exact line minima: along the axes in turn, then, unless the decrease already meets
    the tolerance, along the net displacement when the extrapolated point is lower and the criterion of Powell's
    direction-set method (Numerical Recipes 10.5, which vnl_powell.cxx:60-175 implements) accepts it."""
    p = np.array(x0, float)
    pt, fret = p.copy(), f(p)
    fp, big = fret, 0.0
    for i in order:
        f_prev = fret
        p = _exact_line_min(f, p, np.eye(3)[i])
        fret = f(p)
        big = max(big, abs(f_prev - fret))
    if 2 * abs(fp - fret) <= ftol * (abs(fp) + abs(fret)) or not last_search:
        return p, False
    fe = f(2 * p - pt)
    if fe < fp and 2 * (fp - 2 * fret + fe) * (fp - fret - big) ** 2 - big * (fp - fe) ** 2 < 0:
        return _exact_line_min(f, p, p - pt), True
    return p, False


def test_powell_sweep_equals_closed_form_line_minima():
    """On coupled quadratics (random positive-definite Hessians) one sweep of the restated vnl Powell lands where
    exact line minima put it: the three axis searches in order, then the search along the net displacement.  The
    cases include sweeps that take that last search; reversing the axis order or leaving out the last search moves
    the closed-form endpoint far beyond the tolerance, so the check fixes both."""
    rng = np.random.default_rng(17)
    with_last = reordered_far = 0
    for _ in range(40):
        Q = rng.standard_normal((3, 3))
        H = Q @ Q.T + 0.3 * np.eye(3)
        c = rng.uniform(-6, 6, 3)
        x0 = rng.uniform(-3, 3, 3)
        f = lambda x, H=H, c=c: float((x - c) @ H @ (x - c))
        xm, fm, nev = _powell_fn(f, x0)
        want, took_last = _closed_form_sweep(f, x0)
        assert np.abs(xm - want).max() < 2e-2, (xm, want)
        assert abs(fm - f(xm)) <= 1e-9 * max(1.0, abs(fm))
        if took_last:
            with_last += 1
            assert np.abs(_closed_form_sweep(f, x0, last_search=False)[0] - want).max() > 0.1
        reordered_far += np.abs(_closed_form_sweep(f, x0, order=(2, 1, 0))[0] - want).max() > 0.1
    assert with_last >= 5 and reordered_far >= 30, (with_last, reordered_far)


def test_powell_line_searches_on_a_non_quadratic():
    """Separable even functions with their minimum far from the start (log cosh plus a quartic per axis): the bracket
    has to expand several initial steps and Brent's golden-section steps have to close in, and each axis search
    ends within the line tolerance of the minimiser; by symmetry the extrapolated point is never lower, so no
    fourth search runs."""
    g = lambda d: np.logaddexp(d, -d) - np.log(2.0) + 0.01 * d ** 4  # log cosh + quartic
    c = np.array([23.0, -17.5, 2.25])
    for x0 in (np.zeros(3), np.array([-4.0, 6.0, 30.0])):
        xm, fm, nev = _powell_fn(lambda x: float(np.sum(g(x - c))), x0)
        assert np.abs(xm - c).max() < 3e-2, (xm, c)
        assert fm < 1e-3
        assert nev > 20
    # Brent's parabolic steps converge superlinearly on a smooth function: over minimisers up to 35 from the start
    # a sweep takes 32-43 evaluations, where a copy of the restatement whose parabolic step has the wrong sign
    # (every such step rejected, golden-section steps only) takes 53-62
    rng = np.random.default_rng(4)
    for _ in range(30):
        c = rng.uniform(-20, 20, 3)
        xm, fm, nev = _powell_fn(lambda x, c=c: float(np.sum(g(x - c))), np.zeros(3))
        assert np.abs(xm - c).max() < 2e-2 and nev <= 48, (c, xm, nev)


def _quadratic_volume(N: int) -> np.ndarray:
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return (0.01 * x * x - 0.02 * y * y + 0.015 * z * z + 0.005 * x * y - 0.3 * z + 0.2 * x + 1.0).astype(np.float32)


def _reslice(vol, t):
    o = oracle()
    N = vol.shape[0]
    out = np.zeros(N ** 3)
    tt = np.ascontiguousarray(np.asarray(t, np.float64).ravel())
    o.orc_reslice(vp(np.ascontiguousarray(vol)), N, C.byref(Particle()), vp(tt), vp(out))
    return out.reshape(N, N, N)  # [x][y][z]


def test_keys_reslice_reproduces_quadratics():
    N = 16
    v = _quadratic_volume(N)
    t = np.eye(4)
    t[:3, 3] = [0.3, -0.7, 0.45]
    out = _reslice(v, t)
    f = lambda x, y, z: 0.01 * x * x - 0.02 * y * y + 0.015 * z * z + 0.005 * x * y - 0.3 * z + 0.2 * x + 1.0  # noqa
    sl = slice(2, N - 3)
    xi, yi, zi = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")  # output [x][y][z]
    ref = f(xi + 0.3, yi - 0.7, zi + 0.45)
    d = (out - ref)[sl, sl, sl]
    assert np.ptp(d) < 2e-5  # equal up to the mean-of-valid offset and float storage


def test_reslice_identity_and_quarter_turn():
    N = 16
    g = np.random.default_rng(5)
    v = g.standard_normal((N, N, N)).astype(np.float32)
    out = _reslice(v, np.eye(4))
    ref = v.transpose(2, 1, 0).astype(np.float64)
    # the reslice floors with a tolerance of 2^-17 (vtkInterpolationMath::Floor) that stays in the fraction:
    # the identity map samples at +7.6e-6 voxel, an error of that order times the local slope
    assert np.abs(out - (ref - np.float32(ref.astype(np.float32).mean(dtype=np.float64)))).max() < 2e-4
    # quarter turn about z about voxel N/2: input (x, y) = (N/2 - (y - N/2), N/2 + (x - N/2))
    t = np.eye(4)
    t[:2, :2] = [[0, -1], [1, 0]]
    out = _reslice(v, t)
    xi, yi = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    px, py = N - yi, xi
    ok = (px >= 0) & (px < N)
    vals = np.zeros((N, N, N))
    vals[ok] = v.transpose(2, 1, 0)[px[ok], py[ok], :]
    inside = ok[:, :, None] & np.ones((1, 1, N), bool)
    mean = vals[inside].astype(np.float32).mean(dtype=np.float64)
    assert np.abs(out[inside] - (vals[inside] - mean)).max() < 2e-4


def test_reslice_half_voxel_border():
    """vtkImageReslice with its border on (the reference's setting): a sample is valid within half a voxel beyond the
    outermost voxel centres, [-1/2, N - 1/2] per axis.  On a ramp in y that is constant in x, shifts of +-0.4 voxel
    along x keep the edge plane valid with the ramp's exact values (the Keys weights over taps that reach past the
    edge still sum to one), shifts of +-0.6 make it invalid (0), and the mean of the valid samples is what is
    subtracted."""
    N = 12
    zi, yi, xi = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    v = (0.25 * yi + 3.0).astype(np.float32)  # [z][y][x]
    ramp = 0.25 * np.arange(N)
    for s in (0.4, -0.4, 0.6, -0.6):
        t = np.eye(4)
        t[0, 3] = s
        out = _reslice(v, t)  # [x][y][z]
        edge = N - 1 if s > 0 else 0
        valid = np.ones(N, bool)
        if abs(s) > 0.5:
            valid[edge] = False
        want = ramp - ramp.mean()  # the valid samples hold every y value equally often
        for i in range(N):
            if valid[i]:
                assert np.abs(out[i] - want[None, :, None]).max() < 1e-5, (s, i)
            else:
                assert np.all(out[i] == 0.0), (s, i)


def test_wedge_full_and_plane():
    o = oracle()
    N = 16
    H = N // 2 + 1
    w = np.zeros(N * N * H, np.float32)
    o.orc_wedge_half(N, C.byref(Particle(lwedge=-90, uwedge=90)), None, vp(w))
    assert np.all(w == 1.0)
    o.orc_wedge_half(N, C.byref(Particle(lwedge=-60, uwedge=60)), None, vp(w))
    w = w.reshape(N, N, H)
    # smoothed binary planes: the tilt axis ky and the kx axis (0 deg) are measured, the beam axis kz is not
    assert np.all(w[N // 2, :, H - 1] > 0.5)
    assert np.all(w[:, N // 2, H - 1] > 0.5)
    assert w[N // 2, N // 2, 0] < 0.01
    assert np.all(w[N // 2, N // 2, : H - 3] < 0.5)


def test_ccc_self_and_shift_recovery():
    N = 32
    p = default_params(N, 16)
    p.missing_wedge = 0
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    blob = lambda cx, cy, cz: (np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2 * 2.5 ** 2))  # noqa
                               + 0.6 * np.exp(-((x - cx - 4) ** 2 + (y - cy + 3) ** 2 + (z - cz - 2) ** 2) / 8.0))
    ref = blob(16, 16, 16).astype(np.float32)
    orc = OraclePlan(p)
    orc.set_reference(ref)
    d0, t0 = orc.ccc(ref, Particle(), np.eye(4))
    assert np.abs(t0[:3, 3]).max() == 0.0
    for s in [(2.0, -1.5, 0.5), (-3.0, 1.0, 2.5)]:
        part = blob(16 + s[0], 16 + s[1], 16 + s[2]).astype(np.float32)
        d, t = orc.ccc(part, Particle(), np.eye(4))
        assert np.allclose(t[:3, 3], s), (s, t[:3, 3])
        assert d > d0


def test_average_of_identical_copies():
    """The average of aligned identical copies is the RMS-normalised copy, whatever the (positive) weights."""
    o = oracle()
    N = 16
    p = default_params(N, 8)
    orc = OraclePlan(p)
    g = np.random.default_rng(9)
    v = g.standard_normal((N, N, N)).astype(np.float32)
    vols = np.ascontiguousarray(np.stack([v, v, v]))
    from ettk_amd.abi import particle_array
    mats = np.ascontiguousarray(np.tile(np.eye(4).ravel(), (3, 1)))
    out1 = np.zeros((N, N, N))
    out2 = np.zeros((N, N, N))
    o.orc_average(orc.h, 3, vp(vols), particle_array([Particle()] * 3), vp(np.array([1, 1, 1], np.float32)),
                  vp(mats), vp(out1))
    o.orc_average(orc.h, 3, vp(vols), particle_array([Particle()] * 3), vp(np.array([0.2, 3, 1.5], np.float32)),
                  vp(mats), vp(out2))
    assert np.abs(out1 - out2).max() < 1e-9
    assert abs(np.sqrt(np.mean(out1 ** 2)) - 1.0) < 1e-6
    # the copy itself, mean removed (DC := 0) and RMS-normalised
    c = v.astype(np.float64) - v.astype(np.float32).mean(dtype=np.float64)
    c = c - c.mean()
    c /= np.sqrt(np.mean(c ** 2))
    assert np.abs(out1 - c).max() < 1e-4


def test_refine_score_is_one_at_the_identity():
    """The continuous score of the reference against itself at the identity pose is 1 and stationary."""
    o = oracle()
    o.orc_refine_eval.restype = C.c_double
    o.orc_refine_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
    from ettk_amd.synth import blob_reference
    N = 16
    p = default_params(N, 8)
    p.missing_wedge = 0
    ref = blob_reference(N, seed=3)
    orc = OraclePlan(p)
    orc.set_reference(ref)
    g = np.zeros(6)
    S = o.orc_refine_eval(orc.h, vp(ref), C.byref(Particle()), vp(np.eye(4).ravel().copy()), vp(np.zeros(6)), 1e-4,
                          vp(g))
    assert abs(S - 1.0) < 1e-12
    assert np.abs(g).max() < 1e-6
    S2 = o.orc_refine_eval(orc.h, vp(ref), C.byref(Particle()), vp(np.eye(4).ravel().copy()),
                           vp(np.array([1.0, 0, 0, 0.2, 0, 0])), 1e-4, None)
    assert S2 < S


# ------------------------------------------------------------------------------------- harmonics(A, t, R)
def _rz(a):
    c, s = math.cos(a), math.sin(a)
    M = np.eye(4)
    M[:2, :2] = [[c, -s], [s, c]]
    return M


def _ry(b):
    c, s = math.cos(b), math.sin(b)
    M = np.eye(4)
    M[0, 0], M[0, 2], M[2, 0], M[2, 2] = c, s, -s, c
    return M


def _harmonics_t(B, A, t):
    out = np.zeros((2 * B, 2 * B), np.float32)
    a32 = np.ascontiguousarray(A, dtype=np.float32)
    oracle().orc_harmonics_t(B, vp(a32), vp(np.ascontiguousarray(t, dtype=np.float64)), vp(out))
    return out


def _rotate(B, A, a, b, g):
    f = np.ascontiguousarray(np.asarray(A, dtype=np.float32).astype(np.float64))
    out = np.zeros_like(f)
    oracle().orc_s2_rotate(B, vp(f), vp(out), a, b, g)
    return out.astype(np.float32)


@pytest.mark.parametrize("a,b,c", [(0.4, 1.1, -0.7), (2.5, 0.6, 1.9), (-2.2, 2.3, -2.8), (1.0, -0.8, 0.3),
                                   (-1.3, -2.0, 2.6), (0.4, 1.1, 0.0), (-2.0, 0.7, 0.0), (0.0, 0.9, 1.2)])
def test_harmonics_of_a_transform_is_the_inverse_zyz_rotation(a, b, c):
    """For t = Rz(a) Ry(b) Rz(c) the Euler angles read back from t give rotateFct(-c, -b, -a): the rotation by
    the inverse of t, whichever branch the sign tests select (c = 0 makes M21 vanish and takes the theta < 0
    branch).  With a = 0 exactly, M12 = 0 and the reference's tests pick phi = pi instead of 0
    (phi_1 = 0 is not > 0, ProjRot.h:1409-1424): the image is then turned by an extra pi about z."""
    B = 8
    A = _bandlimited(B, 21).reshape(2 * B, 2 * B)
    t = _rz(a) @ _ry(b) @ _rz(c)
    t[:3, 3] = (3.0, -1.0, 2.0)  # the translation is removed first
    got = _harmonics_t(B, A, t)
    want = _rotate(B, A, -c, -b, -math.pi if a == 0.0 else -a)
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    # and it is not the rotation by t itself
    assert np.abs(got - _rotate(B, A, a, b, c)).max() > 1e-2 * np.abs(want).max()


def test_harmonics_degenerate_transforms_keep_the_image():
    """Identity, and a rotation about z alone: every ratio's denominator and numerator vanish, psi = phi = 0 and
    theta = 0, so the reference leaves the image unrotated (ProjRot.h:1371-1397)."""
    B = 8
    A = _bandlimited(B, 22).reshape(2 * B, 2 * B).astype(np.float32)
    for t in (np.eye(4), _rz(0.9)):
        got = _harmonics_t(B, A, t)
        assert np.abs(got - A).max() <= 1e-5 * np.abs(A).max()


# ------------------------------------------------------------------------------------- wedge overlap
def test_wedge_overlap_closed_forms():
    """getWedgeOverlap: 1 without effective wedges; a +-60 degree wedge keeps about 2/3 of a radial band
    (the angular fraction 120/180); swapping the roles of two unrotated wedges leaves the product unchanged;
    turning the particle's wedge by 90 degrees about z moves its missing region away from the reference's."""
    N = 32
    p = default_params(N, 8, bp_lo=0.1, bp_hi=0.45, bp_lo_decay=0.02, bp_hi_decay=0.02)
    full, w60 = Particle(lwedge=-90, uwedge=90), Particle(lwedge=-60, uwedge=60)
    I = np.eye(4)
    o = OraclePlan(p)
    o.set_reference(np.zeros((N, N, N), np.float32), full)
    assert o.wedge_overlap(full, I) == 1.0
    assert o.wedge_overlap(Particle(lwedge=-120, uwedge=130), I) == 1.0  # clamped to -90 / 90
    one = o.wedge_overlap(w60, I)
    assert abs(one - 2.0 / 3.0) < 0.03, one
    w45 = Particle(lwedge=-45, uwedge=50)
    a = o.wedge_overlap(w45, I)
    o2 = OraclePlan(p)
    o2.set_reference(np.zeros((N, N, N), np.float32), w60)
    b = o2.wedge_overlap(w45, I)
    o3 = OraclePlan(p)
    o3.set_reference(np.zeros((N, N, N), np.float32), w45)
    c = o3.wedge_overlap(w60, I)
    assert abs(b - c) < 1e-6 and b < a
    rz = np.eye(4)
    rz[:2, :2] = [[0, -1], [1, 0]]
    assert o2.wedge_overlap(w60, rz) < o2.wedge_overlap(w60, I) - 0.05


def test_variance_map_closed_forms():
    """getImageVarianceMPI at the identity alignment, where the resliced particle is the input minus its mean:
    around that image itself the variance vanishes; around zero it is the |w|-weighted mean of T^2 (negative
    weights count by magnitude, zero weights not at all)."""
    N = 16
    orc = OraclePlan(default_params(N, 8))
    g = np.random.default_rng(10)
    v1, v2 = g.standard_normal((2, N, N, N)).astype(np.float32)
    T1 = v1.astype(np.float64) - v1.mean(dtype=np.float64)
    T2 = v2.astype(np.float64) - v2.mean(dtype=np.float64)
    mats = np.tile(np.eye(4), (3, 1, 1))
    descs = [Particle()] * 3
    var0 = orc.variance(np.stack([v1, v1, v1]), descs, [0.2, 3.0, -1.5], mats, T1)
    assert np.abs(var0).max() < 1e-8
    var = orc.variance(np.stack([v1, v2, v2]), descs, [2.0, -0.5, 0.0], mats, np.zeros((N, N, N)))
    want = (2.0 * T1 ** 2 + 0.5 * T2 ** 2) / 2.5
    # the identity reslice reproduces the input to the 2^-17 floor tolerance of VTK's index arithmetic
    assert np.abs(var - want).max() <= 2e-4 * want.max()
