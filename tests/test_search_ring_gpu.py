"""The SO(3) search runs particles in chunks of 32 whose device list buffers come from a ring of four slots; a chunk
reuses a slot only after the host has finished extracting the candidates of the chunk four places earlier.

One call over 167 particles (six chunks, the last one partial: every slot reused at least once) must give the same
results as separate calls over each chunk's particles: the per-particle arithmetic does not depend on which chunk
or slot a particle ran in.  Box 32, B = 16, batched + adjoint form and the search without Powell."""
import numpy as np
import pytest

from conftest import to_dev
from ettk_amd.abi import default_params
from ettk_amd.synth import blob_reference, make_particles


@pytest.mark.gpu
@pytest.mark.parametrize("mode,refine", [(2, 1), (1, 0)])
def test_more_chunks_than_slots_equal_separate_calls(torch_gpu, ctx, mode, refine):
    from ettk_amd.lib import Plan

    n, chunk = 167, 32
    print(f"[ring] {n} particles in chunks of {chunk}, mode {mode}: one call vs one call per chunk", flush=True)
    p = default_params(32, 16, refine=refine)
    ref = blob_reference(32, seed=61)
    ps = make_particles(ref, n, seed=62, half_wedge=(45, 70), snr=(0.1, 0.3))
    plan = Plan(ctx, p)
    try:
        plan.set_reference(to_dev(torch_gpu, ref))
        d = to_dev(torch_gpu, ps.vols)
        whole = plan.align(d, ps.descs, mode=mode)
        for c0 in range(0, n, chunk):
            part = plan.align(d[c0:c0 + chunk].contiguous(), ps.descs[c0:c0 + chunk], mode=mode)
            for q, r in enumerate(part):
                w = whole[c0 + q]
                assert r.n_ccc == w.n_ccc and r.n_cand == w.n_cand, (c0 + q, r.n_ccc, w.n_ccc)
                assert np.array_equal(r.transform(), w.transform()), (c0 + q, r.transform(), w.transform())
                assert r.distance == w.distance, (c0 + q, r.distance, w.distance)
    finally:
        plan.close()
