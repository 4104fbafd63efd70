"""The continuous constrained CCC, its hand-written adjoint and the batched refinement.

Checks: (1) adjoint gradient against central finite differences of the GPU score (h = 1e-3 deg / voxel);
(2) score and gradient against the FP64 oracle (gradient by FP64 central differences, h = 1e-4);
(3) the torch.autograd wrapper; (4) the lock-step refinement against the FP64 restatement of the same
ascent: refined poses within 0.5 degree and 0.1 voxel, and a higher score than at the seed; (5) the
batched + adjoint alignment mode against the Powell mode; (6) the weighted Fourier average.
"""
import numpy as np
import pytest
import torch

from conftest import to_dev
from ettk_amd.abi import default_params
from ettk_amd.synth import blob_reference, make_particles, rotation_error_deg
from oracle_lib import OraclePlan, oracle, vp

import ctypes as C


def _perturbed_seeds(truth, deg, vox, seed):
    g = np.random.default_rng(seed)
    seeds = []
    for T in truth:
        ax = g.standard_normal(3)
        ax /= np.linalg.norm(ax)
        a = np.radians(deg)
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        Rp = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K
        S = T.copy()
        S[:3, :3] = T[:3, :3] @ Rp
        S[:3, 3] = T[:3, 3] + g.uniform(-vox, vox, 3)
        seeds.append(S)
    return np.array(seeds)


def _setup(torch_gpu, ctx, N=32, n=3, snr=0.5):
    from ettk_amd.lib import Plan

    p = default_params(N, 16)
    ref = blob_reference(N, seed=41)
    ps = make_particles(ref, n, seed=42, half_wedge=(50, 65), snr=(snr, snr))
    plan = Plan(ctx, p)
    plan.set_reference(to_dev(torch_gpu, ref))
    return p, ref, ps, plan, to_dev(torch_gpu, ps.vols)


@pytest.mark.gpu
def test_adjoint_gradient_vs_finite_differences(torch_gpu, ctx):
    print("[refine] adjoint vs FD", flush=True)
    p, ref, ps, plan, d = _setup(torch_gpu, ctx)
    try:
        seeds = _perturbed_seeds(ps.truth, 2.0, 0.7, 1)
        n = len(seeds)
        x0 = np.random.default_rng(2).uniform(-0.5, 0.5, (n, 6))
        S, g = plan.ccc_grad(d, ps.descs, list(range(n)), seeds, x0)
        h = 1e-3
        for r in range(6):
            xp, xm = x0.copy(), x0.copy()
            xp[:, r] += h
            xm[:, r] -= h
            sp, _ = plan.ccc_grad(d, ps.descs, list(range(n)), seeds, xp)
            sm, _ = plan.ccc_grad(d, ps.descs, list(range(n)), seeds, xm)
            fd = (sp - sm) / (2 * h)
            for e in range(n):
                scale = np.linalg.norm(g[e])
                # per component: 5% of the component plus a floor of 1% of the gradient's norm (the FP32 score's
                # rounding over the step h); components above 5% of the norm must agree in sign
                assert abs(fd[e] - g[e, r]) <= 0.05 * abs(g[e, r]) + 0.01 * scale + 1e-6, (r, e, fd[e], g[e, r])
                if abs(g[e, r]) > 0.05 * scale:
                    assert np.sign(fd[e]) == np.sign(g[e, r]), (r, e, fd[e], g[e, r])
        # directional derivatives along random directions
        u = np.random.default_rng(9).standard_normal((4, 6))
        for k in range(4):
            xp, xm = x0 + h * u[k], x0 - h * u[k]
            sp, _ = plan.ccc_grad(d, ps.descs, list(range(n)), seeds, xp)
            sm, _ = plan.ccc_grad(d, ps.descs, list(range(n)), seeds, xm)
            for e in range(n):
                dd, gu = (sp[e] - sm[e]) / (2 * h), float(g[e] @ u[k])
                assert abs(dd - gu) <= 0.05 * abs(gu) + 0.01 * np.linalg.norm(g[e]) * np.linalg.norm(u[k]), (k, e, dd, gu)
    finally:
        plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("N", [32, 64, 96])
def test_score_and_gradient_vs_fp64(torch_gpu, ctx, N):
    # 32 and 64 run different pencil transforms inside the score and adjoint i passes (LDS radix passes; the
    # 8 x 8 register transform)
    print(f"[refine] score/gradient vs FP64 oracle N={N}", flush=True)
    p, ref, ps, plan, d = _setup(torch_gpu, ctx, N=N)
    orc = OraclePlan(p)
    orc.set_reference(ref)
    o = oracle()
    o.orc_refine_eval.restype = C.c_double
    o.orc_refine_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
    try:
        seeds = _perturbed_seeds(ps.truth, 2.5, 0.8, 3)
        n = len(seeds)
        x0 = np.random.default_rng(4).uniform(-0.5, 0.5, (n, 6))
        S, g = plan.ccc_grad(d, ps.descs, list(range(n)), seeds, x0)
        for e in range(n):
            grad64 = np.zeros(6)
            sd = np.ascontiguousarray(seeds[e].ravel())
            xe = np.ascontiguousarray(x0[e])
            vol = np.ascontiguousarray(ps.vols[e])
            s64 = o.orc_refine_eval(orc.h, vp(vol), C.byref(ps.descs[e]), vp(sd), vp(xe), 1e-4, vp(grad64))
            assert abs(S[e] - s64) <= 2e-4 * max(1.0, abs(s64)), (e, S[e], s64)
            assert np.linalg.norm(g[e] - grad64) <= 0.03 * np.linalg.norm(grad64) + 1e-6, (e, g[e], grad64)
            # per component against the FP64 gradient: 3% of the component plus 0.5% of the norm
            nrm = np.linalg.norm(grad64)
            assert np.all(np.abs(g[e] - grad64) <= 0.03 * np.abs(grad64) + 0.005 * nrm + 1e-6), (e, g[e], grad64)
    finally:
        plan.close()


@pytest.mark.gpu
def test_autograd_wrapper(torch_gpu, ctx):
    from ettk_amd.autograd import ccc_score

    print("[refine] torch.autograd wrapper", flush=True)
    p, ref, ps, plan, d = _setup(torch_gpu, ctx)
    try:
        seeds = _perturbed_seeds(ps.truth, 2.0, 0.5, 5)
        n = len(seeds)
        x = torch.zeros((n, 6), dtype=torch.float64, requires_grad=True)
        S = ccc_score(plan, d, ps.descs, list(range(n)), seeds, x)
        (2.0 * S).sum().backward()
        _, g = plan.ccc_grad(d, ps.descs, list(range(n)), seeds, np.zeros((n, 6)))
        assert np.allclose(x.grad.numpy(), 2.0 * g)
    finally:
        plan.close()


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("N,n", [(32, 3), (96, 2)])
def test_refinement_matches_fp64_ascent(torch_gpu, ctx, N, n):
    """Box 96 is the size of the refinement workload (seeds up to 3 degrees and 1 voxel off, 20 steps); the
    FP64 ascents run in parallel host threads (the oracle plan is only read)."""
    from concurrent.futures import ThreadPoolExecutor

    print(f"[refine] lock-step ascent vs FP64 ascent N={N}", flush=True)
    p, ref, ps, plan, d = _setup(torch_gpu, ctx, N=N, n=n, snr=1.0)
    orc = OraclePlan(p)
    orc.set_reference(ref)
    o = oracle()
    o.orc_refine.restype = C.c_double
    o.orc_refine.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    try:
        seeds = _perturbed_seeds(ps.truth, 3.0, 1.0, 6)
        n = len(seeds)
        poses, scores = plan.refine(d, ps.descs, list(range(n)), seeds, n_iter=20)
        s0, _ = plan.ccc_grad(d, ps.descs, list(range(n)), seeds, np.zeros((n, 6)))

        def fp64(e):
            out = np.zeros(16)
            vol = np.ascontiguousarray(ps.vols[e])
            s64 = o.orc_refine(orc.h, vp(vol), C.byref(ps.descs[e]), vp(np.ascontiguousarray(seeds[e].ravel())), 20,
                               vp(out))
            return s64, out.reshape(4, 4)
        with ThreadPoolExecutor(max_workers=n) as ex:
            ref64 = list(ex.map(fp64, range(n)))
        for e in range(n):
            s64, P64 = ref64[e]
            assert rotation_error_deg(poses[e], P64) <= 0.5, e
            assert np.max(np.abs(poses[e][:3, 3] - P64[:3, 3])) <= 0.1, e
            assert abs(scores[e] - s64) <= 2e-4, (e, scores[e], s64)
            # the ascent raised the score
            assert scores[e] > s0[e], (e, scores[e], s0[e])
    finally:
        plan.close()


@pytest.mark.gpu
def test_adjoint_mode_alignment(torch_gpu, ctx):
    """Mode 2 (batched + adjoint) aligns particles at least as well as the Powell modes on clean data."""
    from ettk_amd.lib import Plan

    print("[align] mode 2 vs mode 1", flush=True)
    p = default_params(32, 16)
    ref = blob_reference(32, seed=51)
    ps = make_particles(ref, 6, seed=52, half_wedge=(50, 65), snr=(1.0, 1.0))
    plan = Plan(ctx, p)
    try:
        plan.set_reference(to_dev(torch_gpu, ref))
        d = to_dev(torch_gpu, ps.vols)
        r1 = plan.align(d, ps.descs, mode=1)
        r2 = plan.align(d, ps.descs, mode=2)
        e1 = [rotation_error_deg(r.transform(), ps.truth[i]) for i, r in enumerate(r1)]
        e2 = [rotation_error_deg(r.transform(), ps.truth[i]) for i, r in enumerate(r2)]
        assert np.median(e2) <= np.median(e1) + 0.5, (e1, e2)
        assert all(r.n_ccc < q.n_ccc for r, q in zip(r2, r1))
    finally:
        plan.close()


@pytest.mark.gpu
def test_average_matches_oracle(torch_gpu, ctx):
    from ettk_amd.lib import Plan

    print("[average] vs FP64 oracle", flush=True)
    p = default_params(32, 16)
    ref = blob_reference(32, seed=61)
    ps = make_particles(ref, 5, seed=62, half_wedge=(50, 65), snr=(0.5, 0.5))
    w = np.array([1.0, 0.5, -0.25, 0.0, 2.0], np.float32)
    plan = Plan(ctx, p)
    orc = OraclePlan(p)
    try:
        d = to_dev(torch_gpu, ps.vols)
        out = torch_gpu.zeros((32, 32, 32), dtype=torch_gpu.float32, device="cuda")
        plan.average(d, ps.descs, w, ps.truth, out)
        avg = out.cpu().numpy()
        ref64 = np.zeros((32, 32, 32))
        o = oracle()
        from ettk_amd.abi import particle_array
        vols = np.ascontiguousarray(ps.vols)
        mats = np.ascontiguousarray(ps.truth.reshape(5, 16))
        o.orc_average(orc.h, 5, vp(vols), particle_array(ps.descs), vp(w), vp(mats), vp(ref64))
        assert np.max(np.abs(avg - ref64)) <= 1e-4 * np.max(np.abs(ref64)) + 1e-5
        # with positive weights, the average of the aligned copies correlates with the reference better than
        # any single aligned copy does
        wp = np.array([1.0, 0.5, 0.25, 1.5, 2.0], np.float32)
        plan.average(d, ps.descs, wp, ps.truth, out)
        c_all = np.corrcoef(out.cpu().numpy().ravel(), ref.ravel())[0, 1]
        c_one = []
        for i in range(5):
            plan.average(d[i:i + 1], ps.descs[i:i + 1], wp[i:i + 1], ps.truth[i:i + 1], out)
            c_one.append(np.corrcoef(out.cpu().numpy().ravel(), ref.ravel())[0, 1])
        assert c_all > max(c_one), (c_all, c_one)
    finally:
        plan.close()


@pytest.mark.gpu
def test_variance_map_matches_oracle(torch_gpu, ctx):
    """getImageVariance around the GPU average, against the FP64 restatement around the same average."""
    from ettk_amd.lib import Plan

    print("[variance] vs FP64 oracle", flush=True)
    N = 32
    p = default_params(N, 16)
    ref = blob_reference(N, seed=63)
    ps = make_particles(ref, 5, seed=64, half_wedge=(50, 65), snr=(0.5, 0.5))
    w = np.array([1.0, 0.5, -0.25, 0.0, 2.0], np.float32)
    plan = Plan(ctx, p)
    orc = OraclePlan(p)
    try:
        d = to_dev(torch_gpu, ps.vols)
        avg = torch_gpu.zeros((N, N, N), dtype=torch_gpu.float32, device="cuda")
        plan.average(d, ps.descs, w, ps.truth, avg)
        var = torch_gpu.zeros_like(avg)
        plan.variance(d, ps.descs, w, ps.truth, avg, var)
        got = var.cpu().numpy()
        want = orc.variance(ps.vols, ps.descs, w, ps.truth, avg.cpu().numpy().astype(np.float64))
        assert np.max(np.abs(got - want)) <= 2e-4 * np.max(want), np.max(np.abs(got - want))
        assert got.min() >= 0.0
    finally:
        plan.close()
