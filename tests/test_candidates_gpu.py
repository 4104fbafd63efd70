"""Candidate extraction on the device (k_lm_candidates: sort by index, isTransformValid, equivalent-angle
elimination of ProjRot.h:1593-1702) against the host's extraction of the same local maxima: the search without
Powell gives identical candidate lists, CCC counts and poses with either, at box 32 / B = 16 (also under a 40-degree
rotation restriction, where the restricted bins form a plateau of maxima) and box 64 / B = 32."""
import os

import numpy as np
import pytest

from conftest import to_dev
from ettk_amd.abi import default_params
from ettk_amd.synth import blob_reference, make_particles


def _align(plan, d, descs, host):
    if host:
        os.environ["ETTK_HOST_CANDIDATES"] = "1"
    try:
        return plan.align(d, descs, mode=1)
    finally:
        os.environ.pop("ETTK_HOST_CANDIDATES", None)


@pytest.mark.gpu
@pytest.mark.parametrize("N,B,rot,n", [(32, 16, 180.0, 16), (32, 16, 40.0, 8), (64, 32, 180.0, 16)])
def test_device_candidates_equal_host_candidates(torch_gpu, ctx, N, B, rot, n):
    from ettk_amd.lib import Plan

    print(f"[candidates] N={N} B={B} rot_restrict={rot}: device vs host extraction", flush=True)
    p = default_params(N, B, refine=0, rot_restrict=rot, n_candidates=3, n_peaks=5)
    ref = blob_reference(N, seed=51)
    ps = make_particles(ref, n, seed=52, half_wedge=(45, 70), snr=(0.1, 0.3))
    plan = Plan(ctx, p)
    try:
        plan.set_reference(to_dev(torch_gpu, ref))
        d = to_dev(torch_gpu, ps.vols)
        dev = _align(plan, d, ps.descs, host=False)
        host = _align(plan, d, ps.descs, host=True)
        for i, (a, b) in enumerate(zip(dev, host)):
            assert a.n_cand == b.n_cand and a.n_ccc == b.n_ccc, (i, a.n_cand, b.n_cand, a.n_ccc, b.n_ccc)
            for c in range(a.n_cand):
                assert tuple(a.so3_angles[c]) == tuple(b.so3_angles[c]), (i, c)
                assert a.cand_distance[c] == b.cand_distance[c], (i, c)
            assert np.array_equal(a.transform(), b.transform()), i
            assert a.distance == b.distance, i
        print(f"  {n} particles: identical candidates, CCC counts and poses", flush=True)
    finally:
        plan.close()
