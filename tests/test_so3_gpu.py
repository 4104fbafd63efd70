"""Parity of the GPU S^2 projection, SO(3) correlation and full pairwise alignment with the FP64 oracle."""
import numpy as np
import pytest

from conftest import to_dev
from ettk_amd.abi import Particle, default_params
from ettk_amd.synth import blob_reference, make_particles, rotation_error_deg
from oracle_lib import OraclePlan, so3_correlate


@pytest.mark.gpu
@pytest.mark.parametrize("B", [8, 16, 32, 64])
def test_so3_correlation_matches_oracle(torch_gpu, ctx, B):
    """FP64 S^2 transforms + SO(3) synthesis on the GPU against the oracle (itself pinned to soft-1.0); B = 32
    runs the synthesis with wavefront-uniform coefficient loads and sixteen degrees per round trip."""
    from ettk_amd.lib import Plan

    print(f"[so3 correlate] B={B}", flush=True)
    plan = Plan(ctx, default_params(32 if B <= 32 else 64, B))
    try:
        g = np.random.default_rng(B)
        n = 2 * B
        m = 3 if B <= 32 else 2
        sig = g.standard_normal((m, n, n))
        pat = g.standard_normal((m, n, n))
        out = plan.so3_correlate(sig, pat)
        for i in range(m):
            ref = so3_correlate(B, sig[i], pat[i])
            assert np.max(np.abs(out[i] - ref)) <= 1e-10 * np.max(np.abs(ref))
            assert np.argmax(out[i]) == np.argmax(ref)
    finally:
        plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("B", [8, 16, 32, 64])
def test_so3_scaled_correlation_matches_oracle(torch_gpu, ctx, B):
    """softScaled (ProjRot.h:1457-1589): the correlation rescaled so that its minimum and maximum equal the S^2 inner
    products of the signal with the pattern rotated to the first minimum's and first maximum's rotation.  The GPU
    (FP64 extrema, Wigner-D rotation of the pattern's coefficients, inverse S^2 transform, weighted sums) against
    the oracle's restatement; the extrema positions are identical and the volume agrees to 1e-9 of its range."""
    from ettk_amd.lib import Plan

    print(f"[so3 correlate scaled] B={B}", flush=True)
    plan = Plan(ctx, default_params(32 if B <= 32 else 64, B))
    try:
        g = np.random.default_rng(100 + B)
        n = 2 * B
        m = 3 if B <= 32 else 2
        sig = g.standard_normal((m, n, n))
        pat = g.standard_normal((m, n, n))
        out = plan.so3_correlate(sig, pat, scaled=True)
        plain = plan.so3_correlate(sig, pat)
        for i in range(m):
            ref = so3_correlate(B, sig[i], pat[i], scaled=True)
            assert np.argmax(out[i]) == np.argmax(ref) and np.argmin(out[i]) == np.argmin(ref)
            span = np.max(ref) - np.min(ref)
            assert np.max(np.abs(out[i] - ref)) <= 1e-9 * span, (i, np.max(np.abs(out[i] - ref)), span)
            # an affine map of the unscaled volume: the same positions order the same way
            assert np.argmax(plain[i]) == np.argmax(out[i])
    finally:
        plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("B", [16, 32])
def test_spherical_rotation_matches_oracle(torch_gpu, ctx, B):
    """harmonics(A, t, R): Euler angles read from t on the host, FP64 S^2 transform / rotateFct / inverse on the
    GPU, against the oracle (its rotation pinned to soft-1.0, its angle reading to the ZYZ closed form)."""
    import ctypes as C

    from ettk_amd.lib import Plan
    from ettk_amd.synth import random_rotation
    from oracle_lib import oracle, vp

    print(f"[sph rotate] B={B}", flush=True)
    plan = Plan(ctx, default_params(32, B))
    try:
        g = np.random.default_rng(B + 1)
        n = 2 * B
        imgs = (g.random((5, n, n)) > 0.4).astype(np.float32)  # binary wedge-like images
        ts = []
        for i in range(5):
            t = np.eye(4)
            t[:3, :3] = random_rotation(g)
            t[:3, 3] = g.uniform(-2, 2, 3)
            ts.append(t)
        ts[4] = np.eye(4)  # degenerate branch: zero angles
        out = plan.sph_rotate(imgs, ts)
        o = oracle()
        for i in range(5):
            want = np.zeros((n, n), np.float32)
            o.orc_harmonics_t(B, vp(np.ascontiguousarray(imgs[i])), vp(np.ascontiguousarray(ts[i])), vp(want))
            assert np.max(np.abs(out[i] - want)) <= 1e-5 * max(1.0, np.max(np.abs(want))), i
    finally:
        plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [3, 16])
def test_projection_matches_oracle(torch_gpu, ctx, n):
    """cartesianToSpherical: FP32 (2N)^3 FFT + tricubic S^2 sampling against the FP64 oracle.  With 16 particles
    in one launch the sampling kernel deals its workgroups one particle per XCD; with 3 it keeps grid order."""
    from ettk_amd.lib import Plan

    print(f"[projection] N=32 B=16 particles={n}", flush=True)
    p = default_params(32, 16)
    ref = blob_reference(32, seed=3)
    ps = make_particles(ref, n, seed=4, half_wedge=(55, 65), snr=(0.2, 0.2))
    plan = Plan(ctx, p)
    orc = OraclePlan(p)
    try:
        plan.set_reference(to_dev(torch_gpu, ref))
        pp = plan.project(to_dev(torch_gpu, ps.vols), ps.descs)
        for i in range(n):
            pr = orc.project(ps.vols[i], ps.descs[i])
            assert np.max(np.abs(pp[i] - pr)) <= 2e-4 * np.max(np.abs(pr)), i
    finally:
        plan.close()


def _compare_alignment(res_gpu, res_orc, tol_dist=2e-3, exact=False, orc=None, vol=None, desc=None):
    """Same SO(3) candidate bins and candidate count.  exact (no Powell: the pose is a function of the integer
    bins and the CCC peak): identical rotation and CCC evaluation count, identical shift unless the FP64 surface
    proves a near-tie.  Otherwise (Powell): rotation within 0.5 degree, shift within half a voxel."""
    from parity import peak_tie
    assert res_gpu.use_window == res_orc.use_window
    assert res_gpu.n_cand == res_orc.n_cand
    for c in range(res_orc.n_cand):
        assert tuple(res_gpu.so3_angles[c]) == tuple(res_orc.so3_angles[c])
    Mg, Mo = res_gpu.transform(), res_orc.transform()
    if exact:
        assert res_gpu.n_ccc == res_orc.n_ccc
        assert np.allclose(Mg[:3, :3], Mo[:3, :3], rtol=0, atol=1e-12)
        if not np.allclose(Mg[:3, 3], Mo[:3, 3], atol=1e-9):
            t_in = np.eye(4)
            t_in[:3, :3] = Mo[:3, :3]
            tie, margin, so, sg = peak_tie(orc, vol, desc, t_in, Mg)
            assert tie, ("shift differs beyond a FP64-proven tie", so, sg, margin)
            return 1
    else:
        assert rotation_error_deg(Mg, Mo) <= 0.5
        assert np.max(np.abs(Mg[:3, 3] - Mo[:3, 3])) <= 0.5 + 1e-6
    assert abs(res_gpu.distance - res_orc.distance) <= tol_dist * abs(res_orc.distance) + 1e-7
    return 0


@pytest.mark.gpu
@pytest.mark.parametrize("refine", [0, 1])
def test_alignment_plain_matches_oracle(torch_gpu, ctx, refine):
    from ettk_amd.lib import Plan

    print(f"[align plain] N=32 B=16 refine={refine}", flush=True)
    p = default_params(32, 16, refine=refine)
    ref = blob_reference(32, seed=21)
    ps = make_particles(ref, 3, seed=22, half_wedge=(55, 65), snr=(0.5, 0.5))
    plan = Plan(ctx, p)
    orc = OraclePlan(p)
    orc.set_reference(ref)
    try:
        plan.set_reference(to_dev(torch_gpu, ref))
        d = to_dev(torch_gpu, ps.vols)
        res = plan.align(d, ps.descs, mode=0)
        exact_ccc = 0
        for i in range(len(res)):
            ro = orc.align(ps.vols[i], ps.descs[i])
            _compare_alignment(res[i], ro, exact=refine == 0, orc=orc, vol=ps.vols[i], desc=ps.descs[i])
            exact_ccc += res[i].n_ccc == ro.n_ccc
        # Powell's path can branch on FP32-level differences of a line search; most pairs follow it exactly
        assert exact_ccc >= len(res) - 1
    finally:
        plan.close()


@pytest.mark.gpu
def test_batched_equals_plain(torch_gpu, ctx):
    """Mode 1 runs every particle's control flow against batched launches: results identical to mode 0."""
    from ettk_amd.lib import Plan

    print("[align batched == plain] N=32 B=16", flush=True)
    p = default_params(32, 16)
    ref = blob_reference(32, seed=31)
    ps = make_particles(ref, 6, seed=32, half_wedge=(45, 70), snr=(0.3, 0.3))
    plan = Plan(ctx, p)
    try:
        plan.set_reference(to_dev(torch_gpu, ref))
        d = to_dev(torch_gpu, ps.vols)
        r0 = plan.align(d, ps.descs, mode=0)
        r1 = plan.align(d, ps.descs, mode=1)
        for a, b in zip(r0, r1):
            assert a.n_ccc == b.n_ccc
            assert a.distance == b.distance
            assert np.array_equal(a.transform(), b.transform())
    finally:
        plan.close()
