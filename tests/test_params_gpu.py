"""Parameters the other parity tests leave at their defaults, against the FP64 oracle.

* CCC with a B-factor filter (buildBFactorHalf, FFilt.h:738-797), and particles whose descriptors carry an
  Euler normal and a cut offset (the frame fin = Euler(normal) * matrix of WSub.h:603-802 moves the particle's
  wedge): identical half-voxel peaks except at FP64-proven near-ties, distances within 2e-3.
* The search without Powell under a rotation restriction of 40 degrees (isTransformValid, nbfImageMetric.h:
  416-459: candidates whose z axis turns further are rejected), three candidates per particle and three peaks per
  SO(3) volume: identical candidate counts, CCC counts and rotations; a candidate ranked differently is accepted
  only where the FP64 CCC distance of the GPU's candidate ties the oracle's at that rank (within 2e-3); shifts
  identical except at proven ties.
"""
import os

import numpy as np
import pytest

from conftest import to_dev
from ettk_amd.abi import default_params
from ettk_amd.synth import blob_reference, make_particles, random_rotation
from oracle_lib import OraclePlan
from parity import DIST_RTOL, peak_tie

N = 32


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


@pytest.mark.gpu
@pytest.mark.parametrize("bfactor", [0.0, 25.0])
def test_ccc_bfactor_normal_cut_offset(torch_gpu, ctx, bfactor):
    from ettk_amd.lib import Plan

    print(f"[params ccc] N={N} bfactor={bfactor}, particles with Euler normals and cut offsets", flush=True)
    p = default_params(N, 16, bfactor=bfactor)
    ref = blob_reference(N, seed=21)
    ps = make_particles(ref, 4, seed=22, half_wedge=(50, 65), snr=(0.3, 0.3))
    normals = [(0.0, 0.0, 0.0), (20.0, 35.0, -10.0), (-45.0, 15.0, 60.0), (90.0, -30.0, 5.0)]
    for d, nrm, cut in zip(ps.descs, normals, (0.0, 1.5, -2.0, 3.25)):
        for a in range(3):
            d.normal[a] = nrm[a]
        d.cut_offset = cut
    g = np.random.default_rng(23)
    cases = []
    for i in range(4):
        cases.append((i, ps.truth[i].copy()))
        for _ in range(3):
            t = np.eye(4)
            t[:3, :3] = random_rotation(g)
            cases.append((i, t))
    orc = OraclePlan(p)
    orc.set_reference(ref)
    plan = Plan(ctx, p)
    try:
        plan.set_reference(to_dev(torch_gpu, ref))
        dist, tout = plan.ccc(to_dev(torch_gpu, ps.vols), ps.descs, [c[0] for c in cases], [c[1] for c in cases])
        n_tie = 0
        for e, (i, t) in enumerate(cases):
            d_ref, t_ref = orc.ccc(ps.vols[i], ps.descs[i], t)
            assert abs(dist[e] - d_ref) <= DIST_RTOL * abs(d_ref) + 1e-7, (e, dist[e], d_ref)
            if not np.allclose(tout[e][:3, 3], t_ref[:3, 3], atol=1e-9):
                tie, margin, so, sg = peak_tie(orc, ps.vols[i], ps.descs[i], t, tout[e])
                assert tie, (e, "peak differs beyond a FP64-proven tie", so, sg, margin)
                n_tie += 1
        print(f"  {len(cases)} evaluations, {n_tie} near-ties", flush=True)
    finally:
        plan.close()


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_search_rotation_restriction_and_candidates(torch_gpu, ctx):
    from ettk_amd.lib import Plan

    print(f"[params search] N={N} B=16 rot_restrict=40 n_candidates=3 n_peaks=3, no Powell", flush=True)
    p = default_params(N, 16, refine=0, rot_restrict=40.0, n_candidates=3, n_peaks=3)
    ref = blob_reference(N, seed=31)
    ps = make_particles(ref, 12, seed=32, half_wedge=(50, 65), snr=(0.3, 0.3))
    orc = OraclePlan(p)
    orc.set_reference(ref)
    plan = Plan(ctx, p)
    try:
        plan.set_reference(to_dev(torch_gpu, ref))
        res = plan.align(to_dev(torch_gpu, ps.vols), ps.descs, mode=1)
        ro = orc.align_many(ps.vols, ps.descs, _threads())
        n_tie = multi = n_rank = 0
        for i, (gr, o) in enumerate(zip(res, ro)):
            assert gr.use_window == o.use_window, i
            assert gr.n_cand == o.n_cand, (i, gr.n_cand, o.n_cand)
            multi += o.n_cand > 1
            for c in range(o.n_cand):
                if tuple(gr.so3_angles[c]) == tuple(o.so3_angles[c]):
                    continue
                # the candidates are ranked by their CCC distance: a different candidate at rank c is accepted
                # only when the FP64 distance of the GPU's candidate rotation ties the oracle's at that rank
                t_in = np.eye(4)
                t_in[:3, :3] = np.array(gr.cand_matrix[c][:], np.float64).reshape(4, 4)[:3, :3]
                d64, _ = orc.ccc(ps.vols[i], ps.descs[i], t_in)
                assert abs(d64 - o.cand_distance[c]) <= DIST_RTOL * abs(o.cand_distance[c]), \
                    (i, c, tuple(gr.so3_angles[c]), tuple(o.so3_angles[c]), d64, o.cand_distance[c])
                n_rank += 1
            assert gr.n_ccc == o.n_ccc, (i, gr.n_ccc, o.n_ccc)
            Mg, Mo = gr.transform(), o.transform()
            if tuple(gr.so3_angles[0]) != tuple(o.so3_angles[0]):  # the best candidate itself is a proven tie
                assert abs(gr.distance - o.distance) <= DIST_RTOL * abs(o.distance) + 1e-7, i
                continue
            assert np.allclose(Mg[:3, :3], Mo[:3, :3], rtol=0, atol=1e-12), i
            if not np.allclose(Mg[:3, 3], Mo[:3, 3], atol=1e-9):
                t_in = np.eye(4)
                t_in[:3, :3] = Mo[:3, :3]
                tie, margin, so, sg = peak_tie(orc, ps.vols[i], ps.descs[i], t_in, Mg)
                assert tie, (i, "shift differs beyond a FP64-proven tie", so, sg, margin)
                n_tie += 1
            assert abs(gr.distance - o.distance) <= DIST_RTOL * abs(o.distance) + 1e-7, (i, gr.distance, o.distance)
        print(f"  12 particles: {multi} with more than one candidate, {n_rank} candidate-rank near-ties, "
              f"{n_tie} shift near-ties", flush=True)
        assert multi > 0
    finally:
        plan.close()
