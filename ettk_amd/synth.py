"""Deterministic synthetic sub-volumes for the alignment workloads.

Reference volumes are sums of isotropic Gaussian blobs; particles are the reference rotated by a uniform
random rotation (trilinear), shifted, band-limited by a missing wedge about Y and given additive white
Gaussian noise at a target SNR.  Everything is seeded through numpy's counter-based Philox generator so that
every run (and every machine) sees the same inputs.

Volumes are float32 arrays indexed [z, y, x] (x fastest in memory), the order the library's C ABI expects.
The ground-truth transform of each particle is T = [R | s]: reslicing the particle with T about the voxel
N/2 (nbfWedgedSubImage3D::getImage, inc/em/nbfWedgedSubImage3D.h:712-726) gives the reference back.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy import ndimage

from .abi import Particle


def _rng(seed: int, stream: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=int(seed) & 0xFFFFFFFFFFFFFFFF, counter=int(stream)))


def blob_reference(N: int, seed: int, n_blobs: int = 40, radius: float | None = None,
                   sigma=(1.5, 4.0), amp=(0.5, 1.0)) -> np.ndarray:
    """Sum of n_blobs Gaussians (sigma U[1.5,4] vox, amplitude U[0.5,1]) centred uniformly in a ball of
    radius 24 (scaled by N/64) around the voxel N/2."""
    g = _rng(seed, 1)
    if radius is None:
        radius = 24.0 * N / 64.0
    c = N / 2.0
    # uniform in the ball
    d = g.standard_normal((n_blobs, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * g.random(n_blobs) ** (1.0 / 3.0)
    centres = c + d * r[:, None]
    sig = g.uniform(sigma[0], sigma[1], n_blobs)
    a = g.uniform(amp[0], amp[1], n_blobs)
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    vol = np.zeros((N, N, N), np.float64)
    for b in range(n_blobs):
        cx, cy, cz = centres[b]
        vol += a[b] * np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2.0 * sig[b] ** 2))
    return vol.astype(np.float32)


def random_rotation(g: np.random.Generator) -> np.ndarray:
    """Uniform rotation from a Shoemake unit quaternion."""
    u1, u2, u3 = g.random(3)
    q = np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def wedge_mask_full(N: int, lwedge: float, uwedge: float) -> np.ndarray:
    """Binary missing-wedge mask on the full numpy FFT grid [kz, ky, kx]: 1 where data was measured,
    with the plane normals of nbfWedge3D::set (inc/em/nbfWedge3D.h:143-184)."""
    l = max(lwedge, -90.0)
    u = min(uwedge, 90.0)
    if u - l >= 180:
        return np.ones((N, N, N), np.float32)
    n1 = np.array([math.cos(-math.radians(90.0 - u)), 0.0, -math.sin(-math.radians(90.0 - u))])
    n2 = np.array([math.cos(-math.radians(90.0 - l)), 0.0, -math.sin(-math.radians(90.0 - l))])
    f = np.fft.fftfreq(N) * N
    kz, ky, kx = np.meshgrid(f, f, f, indexing="ij")
    a = n1[0] * kx + n1[1] * ky + n1[2] * kz
    b = n2[0] * kx + n2[1] * ky + n2[2] * kz
    return (a * b <= 0).astype(np.float32)


@dataclass
class ParticleSet:
    vols: np.ndarray            # (n, N, N, N) float32, [z, y, x]
    descs: list                 # Particle descriptors (wedges, identity alignment)
    truth: np.ndarray           # (n, 4, 4) ground-truth transforms [R | s]
    snr: np.ndarray


def make_particles(ref: np.ndarray, n: int, seed: int, half_wedge=(60.0, 60.0), snr=(0.1, 0.1),
                   max_shift: float = 4.0) -> ParticleSet:
    N = ref.shape[0]
    c = N / 2.0
    vols = np.empty((n, N, N, N), np.float32)
    truth = np.zeros((n, 4, 4))
    descs = []
    snrs = np.zeros(n)
    grid = np.stack(np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij"), 0).reshape(3, -1)
    xyz = grid[::-1].astype(np.float64)  # rows x, y, z
    for p in range(n):
        g = _rng(seed, 1000 + p)
        R = random_rotation(g)
        s = g.uniform(-max_shift, max_shift, 3)
        h = g.uniform(half_wedge[0], half_wedge[1])
        snr_p = g.uniform(snr[0], snr[1])
        src = c + R.T @ (xyz - c - s[:, None])          # source (x, y, z) in the reference
        vol = ndimage.map_coordinates(ref.astype(np.float64), src[::-1], order=1, mode="constant", cval=0.0)
        vol = vol.reshape(N, N, N)
        W = wedge_mask_full(N, -h, h)
        vol = np.real(np.fft.ifftn(np.fft.fftn(vol) * W))
        var = vol.var()
        noise = g.standard_normal((N, N, N)) * math.sqrt(var / snr_p) if snr_p > 0 else 0.0
        vols[p] = (vol + noise).astype(np.float32)
        T = np.eye(4)
        T[:3, :3] = R
        T[:3, 3] = s
        truth[p] = T
        snrs[p] = snr_p
        descs.append(Particle(lwedge=-h, uwedge=h))
    return ParticleSet(vols, descs, truth, snrs)


def make_particles_device(ref: np.ndarray, n: int, seed: int, half_wedge=(45.0, 70.0), snr=(0.05, 0.2),
                          max_shift: float = 4.0, chunk: int = 64):
    """Same generator as make_particles (same rotations, shifts, wedges and SNRs from the same Philox
    streams), evaluated on the GPU with torch: trilinear rotation by grid_sample, wedge by torch.fft, noise
    from a seeded torch generator.  Returns (device tensor (n, N, N, N) float32, descs, truth)."""
    import torch
    import torch.nn.functional as F

    N = ref.shape[0]
    c = N / 2.0
    dev = torch.device("cuda")
    out = torch.empty((n, N, N, N), dtype=torch.float32, device=dev)
    refd = torch.from_numpy(ref.astype(np.float32)).to(dev)[None, None]
    ar = torch.arange(N, dtype=torch.float64, device=dev)
    zz, yy, xx = torch.meshgrid(ar, ar, ar, indexing="ij")
    xyz = torch.stack([xx, yy, zz], -1).reshape(-1, 3)  # (N^3, 3) x, y, z
    f = torch.fft.fftfreq(N, device=dev, dtype=torch.float64) * N
    kz, ky, kx = torch.meshgrid(f, f, f, indexing="ij")
    gen = torch.Generator(device=dev)
    truth = np.zeros((n, 4, 4))
    descs = []
    params = []
    for p in range(n):
        g = _rng(seed, 1000 + p)
        R = random_rotation(g)
        s = g.uniform(-max_shift, max_shift, 3)
        h = g.uniform(half_wedge[0], half_wedge[1])
        snr_p = g.uniform(snr[0], snr[1])
        params.append((R, s, h, snr_p))
        T = np.eye(4)
        T[:3, :3] = R
        T[:3, 3] = s
        truth[p] = T
        descs.append(Particle(lwedge=-h, uwedge=h))
    for c0 in range(0, n, chunk):
        cn = min(chunk, n - c0)
        grids = []
        for p in range(c0, c0 + cn):
            R, s, _, _ = params[p]
            Rt = torch.from_numpy(R.T.copy()).to(dev)
            src = c + (xyz - c - torch.from_numpy(s).to(dev)) @ Rt.T  # (N^3, 3) source x, y, z
            grids.append((2.0 * src / (N - 1) - 1.0).reshape(N, N, N, 3))
        grid = torch.stack(grids).to(torch.float32)
        vol = F.grid_sample(refd.expand(cn, 1, N, N, N), grid, mode="bilinear", padding_mode="zeros",
                            align_corners=True)[:, 0].to(torch.float64)
        for q in range(cn):
            _, _, h, snr_p = params[c0 + q]
            l, u = -h, h
            n1 = (math.cos(-math.radians(90.0 - u)), 0.0, -math.sin(-math.radians(90.0 - u)))
            n2 = (math.cos(-math.radians(90.0 - l)), 0.0, -math.sin(-math.radians(90.0 - l)))
            W = ((n1[0] * kx + n1[2] * kz) * (n2[0] * kx + n2[2] * kz) <= 0).to(torch.float64)
            v = torch.fft.ifftn(torch.fft.fftn(vol[q]) * W).real
            gen.manual_seed(int(seed) * 1000003 + c0 + q)
            noise = torch.randn((N, N, N), generator=gen, device=dev, dtype=torch.float64)
            out[c0 + q] = (v + noise * math.sqrt(float(v.var()) / snr_p)).to(torch.float32)
    torch.cuda.synchronize()
    return out, descs, truth


def rotation_error_deg(A: np.ndarray, B: np.ndarray) -> float:
    """Geodesic angle between the rotation parts of two 4x4 transforms."""
    R = A[:3, :3].T @ B[:3, :3]
    c = (np.trace(R) - 1.0) / 2.0
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))
