"""The oracle's sub-volume extraction (nbfWedgedSubImage3D::getImage with a tomogram, WSub.h:603-802) pinned by
closed forms (to 1e-4: VTK adds 2^-17 before its floor, so an integer sample point carries a fractional part of 7.6e-6): an integer position with no rotation cuts the box [p - N/2, p + N/2) exactly (minus its mean);
Euler(normal) = Rz(-n2) Rx(-n0) Rz(-n1) with n1 = 90 turns the box by a quarter about z, an exact permutation;
a cut offset moves the box along its own z; samples that fall outside the tomogram are invalid (0 after the mean
of the valid ones is removed); a tomogram equal to one particle at its centre gives the particle reslice."""
import ctypes as C

import numpy as np

from ettk_amd.abi import Particle
from oracle_lib import extract, oracle, vp


def _tomo(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


def test_integer_position_cuts_the_box():
    T = _tomo((40, 36, 44), 1)  # (nz, ny, nx)
    N = 16
    pos = [(20, 18, 22), (9, 20, 30)]
    out = extract(T, pos, [Particle(), Particle()], N)
    for i, (x, y, z) in enumerate(pos):
        box = T[z - N // 2:z + N // 2, y - N // 2:y + N // 2, x - N // 2:x + N // 2].astype(np.float64)
        assert np.allclose(out[i], box - np.float32(box.mean()), atol=1e-4)


def test_quarter_turn_and_cut_offset():
    T = _tomo((48, 48, 48), 2)
    N = 16
    x, y, z = 24, 24, 24
    # normal (0, 90, 0): fin = Rz(-90) maps output (u, v, w) - N/2 to input (v, -u, w) about the position
    out = extract(T, [(x, y, z)], [Particle(normal=(0.0, 90.0, 0.0))], N)[0]
    box = T[z - 8:z + 8, :, :]
    want = np.zeros((N, N, N))
    for w in range(N):
        for v in range(N):
            for u in range(N):
                du, dv = u - N // 2, v - N // 2
                want[w, v, u] = T[z + w - N // 2, y - du, x + dv]
    assert np.allclose(out, want - np.float32(want.mean()), atol=1e-4)
    # cut offset c: output voxel reads z - N/2 - c
    out = extract(T, [(x, y, z)], [Particle(cut_offset=3.0)], N)[0]
    box = T[z - 8 - 3:z + 8 - 3, y - 8:y + 8, x - 8:x + 8].astype(np.float64)
    assert np.allclose(out, box - np.float32(box.mean()), atol=1e-4)


def test_outside_the_tomogram_is_invalid():
    T = _tomo((20, 20, 20), 3)
    N = 16
    out = extract(T, [(4, 10, 10)], [Particle()], N)[0]
    # x inputs 4 - 8 .. 4 + 7 = -4 .. 11: columns 0..3 fall outside [-0.5, 19.5]
    assert np.all(out[:, :, :4] == 0.0)
    inside = T[2:18, 2:18, 0:12].astype(np.float64)
    assert np.allclose(out[:, :, 4:], inside - np.float32(inside.mean()), atol=1e-4)


def test_particle_reslice_is_the_centred_case():
    N = 16
    vol = _tomo((N, N, N), 4)
    t = np.eye(4)
    th = np.radians(23.0)
    t[:3, :3] = [[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]]
    t[:3, 3] = [0.3, -1.2, 0.7]
    d = Particle(matrix=t, normal=(5.0, 10.0, -7.0), cut_offset=0.5)
    out = extract(vol, [(N / 2, N / 2, N / 2)], [d], N)[0]
    ref = np.zeros(N ** 3)
    oracle().orc_reslice(vp(np.ascontiguousarray(vol)), N, C.byref(d), vp(np.eye(4).ravel().copy()), vp(ref))
    assert np.array_equal(out, ref.reshape(N, N, N).transpose(2, 1, 0))
