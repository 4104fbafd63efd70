"""Alignment + classification driver over ettk_mra_iteration, with the figures a classification run reports.

classify() repeats the multi-reference iteration of nbfLoopClustering::doMra (inc/em/nbfLoopClustering.h:
1218-1647) as the classification driver of MPI_Classification does (src/MPI_Classification.cpp:880-1166): each
iteration aligns every particle to every class reference, assigns it to the reference whose alignment gives the
lowest summed zero-shift distance, keeps the best top_percentage of each class and rebuilds each class
reference as the wedge-weighted average of its kept particles; the next iteration's references carry the
accumulated wedge of those particles.

agreement() is the fraction of particles whose class matches a known label under the best one-to-one matching
of classes to labels; fsc() is the Fourier shell correlation of two volumes (reported against ground truth).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Iteration:
    classes: np.ndarray
    selected: np.ndarray
    distances: np.ndarray
    matrices: np.ndarray


@dataclass
class Classification:
    classes: np.ndarray
    iterations: list = field(default_factory=list)


def classify(plan, d_refs, d_parts, descs, iterations: int, type: int = 0, form: int = 2,
             top_percentage: float = 1.0) -> Classification:
    """d_refs: (K, N, N, N) float32 device tensor of initial references, replaced by the final class averages."""
    members = mats = None
    out = Classification(classes=np.zeros(d_parts.shape[0], np.int32))
    for _ in range(iterations):
        cls, sel, dist, m = plan.mra_iteration(d_refs, d_parts, descs, type=type, form=form,
                                               top_percentage=top_percentage, ref_members=members,
                                               member_matrices=mats)
        members = np.where(sel == 1, cls, -1).astype(np.int32)
        mats = m
        out.iterations.append(Iteration(cls, sel, dist, m))
        out.classes = cls
    return out


def agreement(classes: np.ndarray, labels: np.ndarray, K: int) -> float:
    """Fraction of particles whose class equals its label under the best matching of classes to labels."""
    from scipy.optimize import linear_sum_assignment

    conf = np.zeros((K, K), np.int64)
    for c, t in zip(classes, labels):
        conf[c, t] += 1
    r, c = linear_sum_assignment(-conf)
    return float(conf[r, c].sum()) / max(1, len(classes))


def fsc(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Fourier shell correlation of two N^3 volumes over integer shells 0 .. N/2 - 1 (FP64)."""
    N = a.shape[0]
    A = np.fft.fftn(a.astype(np.float64))
    B = np.fft.fftn(b.astype(np.float64))
    f = np.fft.fftfreq(N) * N
    kz, ky, kx = np.meshgrid(f, f, f, indexing="ij")
    r = np.rint(np.sqrt(kx * kx + ky * ky + kz * kz)).astype(int)
    out = np.zeros(N // 2)
    for s in range(N // 2):
        m = r == s
        num = np.real(np.sum(A[m] * np.conj(B[m])))
        den = np.sqrt(np.sum(np.abs(A[m]) ** 2) * np.sum(np.abs(B[m]) ** 2))
        out[s] = num / den if den > 0 else 0.0
    return out


def fsc_resolution(curve: np.ndarray, threshold: float = 0.5) -> float:
    """First shell (in Fourier voxels) where the FSC drops below the threshold; len(curve) if it never does."""
    below = np.nonzero(curve < threshold)[0]
    return float(below[0]) if len(below) else float(len(curve))
