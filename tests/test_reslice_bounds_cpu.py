"""Box sizing of the LDS-staged reslice kernels, replayed on the CPU over random rigid maps.

k_reslice (ccc.hip) stages, per 8^3 output tile, the box [floor(mn + tol) - 1, floor(mn + ext + tol) + 2] of the
tile's mapped corners in a 17^3 LDS array and clamps every tap origin into [lo + 1, hi - 2]; the trilinear
kernels of the refinement (refine.hip) stage [floor(mn), floor(mn + ext) + 1] in a 15^3 array and clamp the tap
origin into [lo, hi - 1].  For a rigid map both boxes must always fit, and the clamp must only ever move a tap
whose sample lies within rounding of a voxel boundary (otherwise the kernels would interpolate a different
cubic/linear piece and the result would change by more than a few ulps).
"""
import numpy as np

TOL = 7.62939453125e-06


def _rotations(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                     2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                     2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1).reshape(n, 3, 3)


def _tiles(N):
    t = np.arange(0, N, 8)
    return np.stack(np.meshgrid(t, t, t, indexing="ij"), -1).reshape(-1, 3)


def _check(kind, N, seed):
    rng = np.random.default_rng(seed)
    R = _rotations(rng, 24)
    sh = rng.uniform(-6, 6, (24, 3))
    off = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)
    kb = 17 if kind == "tricubic" else 15
    widths, moved = [], 0
    for r in range(24):
        A = R[r].astype(np.float32).astype(np.float64)  # the kernels map with float32 coefficients
        c = N / 2.0
        t0 = c - A @ np.full(3, c) + sh[r]
        for tile in _tiles(N):
            pts = (tile + off) @ A.T + t0  # input position (x, y, z) of every voxel of the tile, FP64
            c0 = tile @ A.T + t0
            ext = 7 * np.abs(A).sum(1)
            mn = c0 + 7 * np.minimum(A, 0).sum(1)
            if kind == "tricubic":
                lo = np.floor(mn + TOL) - 1
                hi = np.floor(mn + ext + TOL) + 2
                f = np.floor(pts + TOL)
                inside = (f >= lo + 1) & (f <= hi - 2)
            else:
                lo = np.floor(mn)
                hi = np.floor(mn + ext) + 1
                f = np.floor(pts)
                inside = (f >= lo) & (f <= hi - 1)
            widths.append(hi - lo + 1)
            frac = pts - np.round(pts)
            # a tap the clamp moves must belong to a sample within rounding of an integer
            bad = ~inside & (np.abs(frac) > 1e-4)
            assert not bad.any(), (kind, N, r, tile, pts[bad.any(1)][:3], lo, hi)
            moved += int((~inside).sum())
    widths = np.array(widths)
    assert widths.max() <= kb, (kind, widths.max())
    return widths, moved


def test_tricubic_box_fits_17_and_taps_stay_inside():
    for N, seed in ((64, 1), (128, 2)):
        w, moved = _check("tricubic", N, seed)
        assert 11 <= w.mean() <= 16


def test_trilinear_box_fits_15_and_taps_stay_inside():
    for N, seed in ((64, 3), (96, 4)):
        w, moved = _check("trilinear", N, seed)
        assert 9 <= w.mean() <= 14


def test_worst_case_rotation_reaches_the_bound():
    # the body diagonal turned onto an axis maximises 7 * sum|row|: 7 sqrt(3) = 12.12 -> 13 voxel steps
    v = np.ones(3) / np.sqrt(3)
    ext = 7 * np.abs(v).sum()
    assert 12 < ext < 13
    assert np.floor(0.99 + ext) - np.floor(0.99) + 4 == 17  # tricubic width at an unlucky phase
    assert np.floor(0.99 + ext) - np.floor(0.99) + 2 == 15  # trilinear width
