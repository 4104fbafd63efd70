"""ettk_ccc_eval_batch_noalloc (the CCC entry point that allocates nothing) against ettk_ccc_eval_batch.

Both run the same kernels; the no-allocation form uses the library's fixed scratch region, in chunks of 16
evaluations at box 64.  Two calls in a row, the second with fewer particles and evaluations, each compared
with the allocating entry point: identical shifts, distances within FP32 rounding.  The case also runs after
a call of the allocating entry point has sized the plan's own workspace, and before it (a fresh plan).
"""
import numpy as np
import pytest

from conftest import to_dev
from ettk_amd.abi import default_params
from ettk_amd.synth import blob_reference, make_particles, random_rotation


def _cases(ps, n_per, seed):
    g = np.random.default_rng(seed)
    pidx, ts = [], []
    for i in range(len(ps.descs)):
        pidx.append(i)
        ts.append(ps.truth[i].copy())
        for _ in range(n_per - 1):
            t = np.eye(4)
            t[:3, :3] = random_rotation(g)
            pidx.append(i)
            ts.append(t)
    return pidx, ts


@pytest.mark.gpu
@pytest.mark.parametrize("N", [64, 32])
def test_noalloc_matches_allocating_entry_point(torch_gpu, ctx, N):
    from ettk_amd.lib import Plan

    print(f"[ccc noalloc] N={N}", flush=True)
    p = default_params(N, 16)
    ref = blob_reference(N, seed=41)
    ps = make_particles(ref, 5, seed=42, half_wedge=(50, 65), snr=(0.2, 0.2))
    plan = Plan(ctx, p)
    try:
        plan.set_reference(to_dev(torch_gpu, ref))
        for call, (n_part, n_per) in enumerate([(5, 7), (2, 3)]):
            d = to_dev(torch_gpu, ps.vols[:n_part])
            pidx, ts = _cases(ps, n_per, seed=call)
            pidx = [q for q in pidx if q < n_part]
            ts = ts[:len(pidx)]
            d0, t0 = plan.ccc_noalloc(d, ps.descs[:n_part], pidx, ts)
            d1, t1 = plan.ccc(d, ps.descs[:n_part], pidx, ts)
            d2, t2 = plan.ccc_noalloc(d, ps.descs[:n_part], pidx, ts)
            for dd, tt in ((d0, t0), (d2, t2)):
                assert np.array_equal(tt[:, :3, :3], t1[:, :3, :3])
                assert np.max(np.abs(tt[:, :3, 3] - t1[:, :3, 3])) <= 1e-9, (call, tt[:, :3, 3], t1[:, :3, 3])
                assert np.max(np.abs(dd - d1) / np.maximum(np.abs(d1), 1e-12)) <= 1e-5, (call, dd, d1)
    finally:
        plan.close()
