"""ctypes mirror of include/ettk_align.h (structs and constants)."""
from __future__ import annotations

import ctypes as C

import numpy as np

ETTK_MAX_CANDIDATES = 8


class Params(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("B", C.c_int),
        ("bp_lo", C.c_float), ("bp_hi", C.c_float), ("bp_lo_decay", C.c_float), ("bp_hi_decay", C.c_float),
        ("bfactor", C.c_float), ("mask_radius", C.c_float), ("mask_sigma", C.c_float),
        ("square_mask", C.c_int), ("mcf", C.c_int), ("overlap_norm", C.c_int), ("missing_wedge", C.c_int),
        ("n_peaks", C.c_int), ("n_candidates", C.c_int),
        ("rot_restrict", C.c_float), ("trans_restrict", C.c_float), ("refine", C.c_int),
    ]


def default_params(N: int = 64, B: int = 32, **kw) -> Params:
    """Parameters of the alignment workloads: band-pass lo=0.05 hi=0.5 decays 0.01, MCF on, missing-wedge
    compensation on, spherical mask radius N/2-5 sigma 2, square FFT window 5, Kp=5, Kc=1, translation
    restriction 8 voxels, rotation restriction 180 degrees, Powell refinement on."""
    p = Params(N=N, B=B, bp_lo=0.05, bp_hi=0.5, bp_lo_decay=0.01, bp_hi_decay=0.01, bfactor=0.0,
               mask_radius=N / 2 - 5, mask_sigma=2.0, square_mask=5, mcf=1, overlap_norm=1, missing_wedge=1,
               n_peaks=5, n_candidates=1, rot_restrict=180.0, trans_restrict=8.0, refine=1)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Particle(C.Structure):
    _fields_ = [
        ("lwedge", C.c_float), ("uwedge", C.c_float), ("normal", C.c_float * 3), ("cut_offset", C.c_float),
        ("matrix", C.c_double * 16),
    ]

    def __init__(self, lwedge=-90.0, uwedge=90.0, matrix=None, **kw):
        super().__init__()
        self.lwedge = lwedge
        self.uwedge = uwedge
        m = np.eye(4) if matrix is None else np.asarray(matrix, dtype=np.float64).reshape(4, 4)
        for i in range(16):
            self.matrix[i] = float(m.flat[i])
        for k, v in kw.items():
            setattr(self, k, v)


class Result(C.Structure):
    _fields_ = [
        ("distance", C.c_float), ("matrix", C.c_double * 16), ("n_cand", C.c_int),
        ("cand_distance", C.c_float * ETTK_MAX_CANDIDATES),
        ("cand_matrix", (C.c_double * 16) * ETTK_MAX_CANDIDATES),
        ("so3_angles", (C.c_float * 3) * ETTK_MAX_CANDIDATES),
        ("n_ccc", C.c_int), ("use_window", C.c_int), ("overlap", C.c_float),
    ]

    def transform(self) -> np.ndarray:
        return np.array(self.matrix[:], dtype=np.float64).reshape(4, 4)


def particle_array(descs) -> C.Array:
    arr = (Particle * len(descs))()
    for i, d in enumerate(descs):
        arr[i] = d
    return arr


ETTK_BUCKETS_GRAPHS = 1
ETTK_BUCKETS_TUNE = 2


class Bucket(C.Structure):
    """ettk_bucket: particles of one box size and bandwidth with their reference (ettk_align_buckets)."""
    _fields_ = [
        ("params", Params), ("d_reference", C.c_void_p), ("reference", Particle), ("d_particles", C.c_void_p),
        ("n", C.c_int), ("descs", C.POINTER(Particle)), ("out", C.POINTER(Result)),
    ]
