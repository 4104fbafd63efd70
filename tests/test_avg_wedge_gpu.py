"""Accumulated wedge of an average on the GPU (ettk_avg_wedge) and a reference that is an average
(ettk_reference_set_average) against the FP64 oracle.

Bars: the sum of smoothed wedges within 2e-5 (float sums of FP32 against FP64-rounded smoothing); the binarised
image may differ only where the sum sits within 2e-5 of the threshold; the spherical window and the
effective flag exactly; CCC distances within 2e-3 relative and the same half-voxel shift; the search picks the
same rotation (0.5 degree) with the same number of CCCs; the continuous score within 2e-4.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import to_dev
from ettk_amd.abi import Particle, default_params
from ettk_amd.synth import blob_reference, make_particles, random_rotation, rotation_error_deg
from oracle_lib import OraclePlan, avg_wedge, oracle, vp


def _entries(n, seed):
    g = np.random.default_rng(seed)
    descs, mats, w = [], [], []
    for e in range(n):
        h = g.uniform(40, 80)
        descs.append(Particle(lwedge=-h, uwedge=g.uniform(40, 80), normal=tuple(g.uniform(-40, 40, 3))))
        M = np.eye(4)
        M[:3, :3] = random_rotation(g)
        M[:3, 3] = g.uniform(-3, 3, 3)
        mats.append(M)
        w.append(0.0 if e == 2 else g.uniform(0.2, 1.0))
    return descs, np.array(mats), np.array(w, np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("N,B,n", [(32, 16, 7), (64, 32, 3)])
def test_avg_wedge_matches_oracle(torch_gpu, ctx, N, B, n):
    from ettk_amd.lib import Plan

    print(f"[avg wedge] N={N} B={B} entries={n}", flush=True)
    descs, mats, w = _entries(n, seed=N + n)
    ref = avg_wedge(N, B, descs, w, mats)
    plan = Plan(ctx, default_params(N, B))
    try:
        acc = torch.zeros((N, N, N), dtype=torch.float32, device="cuda")
        wed = torch.zeros((N, N, N), dtype=torch.float32, device="cuda")
        sph, eff = plan.avg_wedge(descs, w, mats, acc, wed)
        acc, wed = acc.cpu().numpy(), wed.cpu().numpy()
        np.testing.assert_allclose(acc, ref["accum"], atol=2e-5)
        assert eff == ref["effective"]
        thr = 0.5 * ref["accum"].max()
        near = np.abs(ref["accum"] - thr) <= 2e-5
        gbin = (acc > 0.5 * acc.max())
        assert np.array_equal(gbin[~near], ref["bin"][~near] == 1)
        if not near.any():
            np.testing.assert_allclose(wed, ref["wedge"], atol=1e-5)
        np.testing.assert_array_equal(sph, ref["sph"])
    finally:
        plan.close()


def _average_setup(torch_gpu, ctx, N=32, B=16, seed=131, n=3):
    from ettk_amd.lib import Plan

    p = default_params(N, B)
    ref = blob_reference(N, seed=seed)
    ps = make_particles(ref, n, seed=seed + 1, half_wedge=(50, 65), snr=(0.4, 0.4))
    # the average's own particles: limited wedges about Y at two alignments, so the accumulated wedge is effective
    adescs = [Particle(lwedge=-55, uwedge=50), Particle(lwedge=-60, uwedge=62)]
    amats = np.array([np.eye(4), np.eye(4)])
    amats[1][:3, :3] = random_rotation(np.random.default_rng(seed))
    aw = np.array([1.0, 0.7], np.float32)
    plan = Plan(ctx, p)
    plan.set_reference_average(to_dev(torch_gpu, ref), adescs, aw, amats)
    orc = OraclePlan(p)
    orc.set_reference_average(ref, adescs, aw, amats)
    assert avg_wedge(N, B, adescs, aw, amats)["effective"]
    return p, ref, ps, plan, orc, to_dev(torch_gpu, ps.vols)


@pytest.mark.gpu
def test_average_reference_ccc_matches_oracle(torch_gpu, ctx):
    print("[avg wedge] CCC against an average", flush=True)
    p, ref, ps, plan, orc, d = _average_setup(torch_gpu, ctx)
    try:
        g = np.random.default_rng(5)
        cases = []
        for i in range(len(ps.descs)):
            cases.append((i, ps.truth[i].copy()))
            t = np.eye(4)
            t[:3, :3] = random_rotation(g)
            cases.append((i, t))
        dist, tout = plan.ccc(d, ps.descs, [c[0] for c in cases], [c[1] for c in cases])
        for e, (i, t) in enumerate(cases):
            d_ref, t_ref = orc.ccc(ps.vols[i], ps.descs[i], t)
            assert abs(dist[e] - d_ref) <= 2e-3 * abs(d_ref) + 1e-7, (e, dist[e], d_ref)
            assert np.max(np.abs(tout[e][:3, 3] - t_ref[:3, 3])) <= 1e-6 + 0.5 * (abs(dist[e] - d_ref) > 1e-4)
    finally:
        plan.close()


@pytest.mark.gpu
def test_average_reference_search_matches_oracle(torch_gpu, ctx):
    print("[avg wedge] alignment against an average", flush=True)
    p, ref, ps, plan, orc, d = _average_setup(torch_gpu, ctx, seed=141, n=2)
    try:
        res = plan.align(d, ps.descs, mode=1)
        for i, r in enumerate(res):
            ro = orc.align(ps.vols[i], ps.descs[i])
            assert r.use_window == ro.use_window and (r.use_window & 1)
            assert r.n_ccc == ro.n_ccc, (i, r.n_ccc, ro.n_ccc)
            assert rotation_error_deg(r.transform(), ro.transform()) <= 0.5
            assert abs(r.distance - ro.distance) <= 2e-3 * abs(ro.distance) + 1e-7
    finally:
        plan.close()


@pytest.mark.gpu
def test_average_reference_continuous_score_matches_fp64(torch_gpu, ctx):
    print("[avg wedge] continuous score against an average", flush=True)
    p, ref, ps, plan, orc, d = _average_setup(torch_gpu, ctx, seed=151, n=3)
    o = oracle()
    o.orc_refine_eval.restype = C.c_double
    o.orc_refine_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
    try:
        n = len(ps.descs)
        x0 = np.random.default_rng(6).uniform(-0.5, 0.5, (n, 6))
        S, g = plan.ccc_grad(d, ps.descs, list(range(n)), ps.truth, x0)
        for e in range(n):
            grad64 = np.zeros(6)
            sd = np.ascontiguousarray(ps.truth[e].ravel())
            xe = np.ascontiguousarray(x0[e])
            vol = np.ascontiguousarray(ps.vols[e])
            s64 = o.orc_refine_eval(orc.h, vp(vol), C.byref(ps.descs[e]), vp(sd), vp(xe), 1e-4, vp(grad64))
            assert abs(S[e] - s64) <= 2e-4 * max(1.0, abs(s64)), (e, S[e], s64)
            assert np.linalg.norm(g[e] - grad64) <= 0.03 * np.linalg.norm(grad64) + 1e-6, (e, g[e], grad64)
    finally:
        plan.close()
