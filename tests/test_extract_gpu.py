"""Sub-volume extraction from a tomogram on the GPU (ettk_extract_batch) against the FP64 oracle's extraction
(itself pinned by the closed forms of test_extract_cpu.py): random positions, Euler normals, alignments and cut
offsets, including boxes that reach outside the tomogram.  Bar: the same valid voxels (a sample within FP32
rounding of the border may flip: at most 0.05%), values within 2e-5 of the volume's range."""
import numpy as np
import pytest
from scipy import ndimage

from ettk_amd.abi import Particle
from ettk_amd.synth import random_rotation
from oracle_lib import extract


@pytest.mark.gpu
def test_extract_matches_oracle(torch_gpu, ctx):
    print("[extract] 12 sub-volumes of 32^3 from a 96 x 80 x 64 tomogram", flush=True)
    g = np.random.default_rng(17)
    tomo = ndimage.gaussian_filter(g.standard_normal((64, 80, 96)), 1.5).astype(np.float32)
    N, n = 32, 12
    pos = np.stack([g.uniform(4, 92, n), g.uniform(4, 76, n), g.uniform(4, 60, n)], 1)
    descs = []
    for i in range(n):
        t = np.eye(4)
        t[:3, :3] = random_rotation(g)
        t[:3, 3] = g.uniform(-2, 2, 3)
        descs.append(Particle(matrix=t, normal=tuple(g.uniform(-40, 40, 3)), cut_offset=float(g.uniform(-3, 3))))
    d_tomo = torch_gpu.from_numpy(tomo).cuda()
    out = torch_gpu.empty((n, N, N, N), dtype=torch_gpu.float32, device="cuda")
    ctx.extract(d_tomo, pos, descs, N, out)
    got = out.cpu().numpy()
    want = extract(tomo, pos, descs, N)
    scale = np.abs(tomo).max()
    for i in range(n):
        vg, vw = got[i] != 0, want[i] != 0
        flips = np.count_nonzero(vg != vw)
        assert flips <= 0.0005 * N ** 3, (i, flips)
        both = vg & vw
        assert np.max(np.abs(got[i][both] - want[i][both])) <= 2e-5 * scale, i
    assert np.count_nonzero(want == 0) > 0  # some boxes reach outside the tomogram
