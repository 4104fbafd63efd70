"""Shared comparison rules of the GPU parity tests (test infrastructure).

Index parity is exact except at a near-tie proven on the FP64 surface: when the GPU's half-voxel peak differs
from the oracle's, the oracle's upsampled correlation T (before the spherical mask) is read at both peaks, and
the GPU's choice is accepted only if T there is within `rtol` of the distance scale (k1_m - T_max) below the
FP64 maximum, i.e. taking it changes the distance by less than the FP32 tolerance the tests state.
"""
from __future__ import annotations

import numpy as np

DIST_RTOL = 2e-3  # FP32 chain of five N^3 transforms and the upsampled partial DFT against FP64


def shift_of(t_in: np.ndarray, t_out: np.ndarray) -> np.ndarray:
    """The CCC shift s of t_out = t_in Translate(-s) (executeFourierNewHalf, FIM.h:1312-1322)."""
    return -np.linalg.solve(t_in[:3, :3], t_out[:3, 3] - t_in[:3, 3])


def peak_tie(orc, vol, desc, t_in: np.ndarray, t_gpu: np.ndarray, rtol: float = DIST_RTOL):
    """(is_tie, margin, oracle shift, gpu shift) for a GPU CCC whose shift differs from the oracle's."""
    _, t_orc, T, k1m = orc.ccc_trace(vol, desc, t_in)
    ext = T.shape[0]
    c = (ext - 1) // 2
    sg, so = shift_of(t_in, t_gpu), shift_of(t_in, t_orc)
    ig = np.rint(c + 2 * sg).astype(int)
    io = np.rint(c + 2 * so).astype(int)
    if np.any(ig < 0) or np.any(ig >= ext):
        return False, np.inf, so, sg
    t_max, t_g = T[tuple(io)], T[tuple(ig)]
    margin = t_max - t_g
    scale = max(np.float32(k1m) - t_max, 1e-30)
    return bool(margin <= rtol * scale), float(margin / scale), so, sg
