"""Execution choices and uneven box sizes.

* The CCC batch gives identical results for every execution choice (evaluations per chunk, eager launches or
  chunks replayed from captured HIP graphs): the tuner's candidates differ in scheduling only.
* ettk_plan_tune picks one of its candidates, records it in the tuning cache, and a new plan of the same shape
  starts from the recorded choice.
* ettk_align_buckets over box sizes 48, 64, 96 and 128 (bandwidths 16, 32, 32 and 64, one reference per size, the
  uneven-size workload's mix) gives, per bucket, the results of that bucket's plan run alone, and the search
  without Powell matches the FP64 oracle per size: identical SO(3) bins and CCC counts, identical rotations, shifts
  identical or FP64-proven near-ties, distances within 2e-3.
"""
import os

import numpy as np
import pytest

from conftest import to_dev
from ettk_amd.abi import ETTK_BUCKETS_GRAPHS, Particle, default_params
from ettk_amd.synth import blob_reference, make_particles, random_rotation
from oracle_lib import OraclePlan
from parity import DIST_RTOL, peak_tie


@pytest.mark.gpu
def test_execution_choices_give_identical_results(torch_gpu, ctx):
    from ettk_amd.lib import Plan

    print("[execution] CCC batch at 64^3 for chunks 16/64/256, eager and graph replays", flush=True)
    N = 64
    p = default_params(N, 32)
    ref = blob_reference(N, seed=71)
    ps = make_particles(ref, 6, seed=72, half_wedge=(45, 70), snr=(0.2, 0.2))
    g = np.random.default_rng(73)
    pidx = [i % 6 for i in range(100)]
    ts = []
    for _ in range(100):
        t = np.eye(4)
        t[:3, :3] = random_rotation(g)
        ts.append(t)
    plan = Plan(ctx, p)
    try:
        plan.set_reference(to_dev(torch_gpu, ref))
        d = to_dev(torch_gpu, ps.vols)
        base = None
        for chunk in (16, 64, 256):
            for graphs in (False, True):
                plan.set_execution(chunk, graphs)
                for rep in range(2):  # the second run of a graph choice replays captured chunks
                    dist, tout = plan.ccc(d, ps.descs, pidx, ts)
                    if base is None:
                        base = (dist, tout)
                    assert np.array_equal(dist, base[0]), (chunk, graphs, rep)
                    assert np.array_equal(tout, base[1]), (chunk, graphs, rep)
    finally:
        plan.close()


@pytest.mark.gpu
def test_tuner_records_and_reuses_its_choice(torch_gpu, ctx, tmp_path):
    from ettk_amd.lib import Plan

    print("[tune] CCC execution choice at 32^3 recorded and reused", flush=True)
    os.environ["ETTK_TUNE_CACHE"] = str(tmp_path / "tune.txt")
    try:
        N = 32
        p = default_params(N, 16)
        ref = blob_reference(N, seed=81)
        ps = make_particles(ref, 8, seed=82, half_wedge=(45, 70), snr=(0.2, 0.2))
        plan = Plan(ctx, p)
        try:
            plan.set_reference(to_dev(torch_gpu, ref))
            chunk, graphs, us = plan.tune(to_dev(torch_gpu, ps.vols), ps.descs)
            print(f"  chose chunk {chunk}, graphs {graphs}: {us:.1f} us per evaluation", flush=True)
            assert chunk in (16, 32, 64, 128, 256) and us > 0
            assert plan.execution() == (chunk, graphs)
        finally:
            plan.close()
        text = (tmp_path / "tune.txt").read_text()
        assert f"N=32 B=16" in text and f"chunk={chunk}" in text
        plan2 = Plan(ctx, p)
        try:
            assert plan2.execution() == (chunk, graphs)
        finally:
            plan2.close()
    finally:
        del os.environ["ETTK_TUNE_CACHE"]


SIZES = [(48, 16, 6), (64, 32, 8), (96, 32, 3), (128, 64, 2)]


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_uneven_buckets_match_single_plans_and_oracle(torch_gpu, ctx):
    from concurrent.futures import ThreadPoolExecutor

    from ettk_amd.lib import Plan

    print("[buckets] boxes 48/64/96/128 (B 16/32/32/64) in one call vs single plans and the oracle", flush=True)
    refs, sets, params, buckets = [], [], [], []
    for k, (N, B, n) in enumerate(SIZES):
        p = default_params(N, B, refine=0)
        ref = blob_reference(N, seed=500 + k)
        ps = make_particles(ref, n, seed=510 + k, half_wedge=(45, 70), snr=(0.2, 0.2))
        refs.append(ref)
        sets.append(ps)
        params.append(p)
        buckets.append((p, to_dev(torch_gpu, ref), Particle(), to_dev(torch_gpu, ps.vols), ps.descs))
    res = ctx.align_buckets(buckets, mode=1, flags=ETTK_BUCKETS_GRAPHS)

    def oracle_run(k):
        o = OraclePlan(params[k])
        o.set_reference(refs[k])
        return [o.align(v, d) for v, d in zip(sets[k].vols, sets[k].descs)], o

    with ThreadPoolExecutor(max_workers=4) as ex:
        orc = list(ex.map(oracle_run, range(len(SIZES))))
    for k, (N, B, n) in enumerate(SIZES):
        plan = Plan(ctx, params[k])
        try:
            plan.set_reference(buckets[k][1])
            alone = plan.align(buckets[k][3], sets[k].descs, mode=1)
        finally:
            plan.close()
        ro, o = orc[k]
        ties = 0
        for i in range(n):
            g, a, r = res[k][i], alone[i], ro[i]
            assert g.distance == a.distance and np.array_equal(g.transform(), a.transform()), (N, i)
            assert g.n_ccc == r.n_ccc and g.use_window == r.use_window, (N, i)
            assert tuple(g.so3_angles[0]) == tuple(r.so3_angles[0]), (N, i)
            Mg, Mo = g.transform(), r.transform()
            assert np.allclose(Mg[:3, :3], Mo[:3, :3], rtol=0, atol=1e-12), (N, i)
            if not np.allclose(Mg[:3, 3], Mo[:3, 3], atol=1e-9):
                t_in = np.eye(4)
                t_in[:3, :3] = Mo[:3, :3]
                tie, margin, so, sg = peak_tie(o, sets[k].vols[i], sets[k].descs[i], t_in, Mg)
                assert tie, (N, i, so, sg, margin)
                ties += 1
            assert abs(g.distance - r.distance) <= DIST_RTOL * abs(r.distance) + 1e-7, (N, i, g.distance, r.distance)
        print(f"  box {N} B={B}: {n} particles identical bins, {ties} shift near-ties", flush=True)
