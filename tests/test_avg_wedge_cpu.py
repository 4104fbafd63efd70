"""Pins of the oracle's accumulated wedge of an average (WAvg.h:49-70, 496-538, 614-667).

Each check is fixed by something other than the restatement itself: an entry without a missing wedge gives all
ones; one entry reproduces the half-grid particle wedge (already pinned by the analytic plane tests) and the
binary spherical wedge; an alignment about the tilt axis Y by theta equals the wedge with both limits shifted by
-theta; two entries 90 degrees apart binarise to their intersection; zero-weight entries are skipped.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ettk_amd.abi import Particle
from oracle_lib import avg_wedge, oracle, vp

N, B = 16, 8


def _rot(axis: str, deg: float) -> np.ndarray:
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    R = {"x": [[1, 0, 0], [0, c, -s], [0, s, c]], "y": [[c, 0, s], [0, 1, 0], [-s, 0, c]],
         "z": [[c, -s, 0], [s, c, 0], [0, 0, 1]]}[axis]
    M = np.eye(4)
    M[:3, :3] = R
    return M


def _desc(l, u, normal=(0.0, 0.0, 0.0), matrix=None) -> Particle:
    return Particle(lwedge=l, uwedge=u, normal=normal, matrix=matrix)


def test_entries_without_wedge_are_not_effective():
    aw = avg_wedge(N, B, [_desc(-90, 90), _desc(-90, 90)], [1.0, 0.5], [np.eye(4), _rot("z", 30)])
    assert np.all(aw["accum"] == 2.0) and np.all(aw["bin"] == 1.0)
    assert not aw["effective"]
    assert np.all(aw["wedge"] == 1.0) and np.all(aw["sph"] == 1.0)


def test_nearly_full_wedge_is_not_effective():
    aw = avg_wedge(N, B, [_desc(-89, 89)], [1.0], [np.eye(4)])
    assert not aw["effective"] and np.all(aw["wedge"] == 1.0)


def test_single_entry_matches_particle_wedge():
    o = oracle()
    store = _rot("z", 20) @ _rot("x", 35) @ _rot("z", -50)
    store[:3, 3] = (1.5, -2.0, 0.25)  # the translation does not move a wedge
    garbage = _rot("y", 70)           # the particle's own matrix is replaced by the alignment
    d = _desc(-55, 60, normal=(10.0, 20.0, 30.0), matrix=garbage)
    aw = avg_wedge(N, B, [d], [1.0], [store])
    half = np.zeros((N, N, N // 2 + 1), np.float32)
    d_store = _desc(-55, 60, normal=(10.0, 20.0, 30.0), matrix=store)
    o.orc_wedge_half(N, C.byref(d_store), None, vp(half))
    # below the last half plane both grids smooth over the same neighbours
    np.testing.assert_allclose(aw["accum"][:, :, : N // 2], half[:, :, : N // 2], atol=1e-6)
    np.testing.assert_array_equal(aw["bin"], (aw["accum"] > 0.5 * aw["accum"].max()).astype(np.float32))
    assert aw["effective"]
    sph = np.zeros((2 * B, 2 * B), np.float32)
    o.orc_spherical_wedge(B, C.byref(d_store), 0, vp(sph))
    np.testing.assert_array_equal(aw["sph"], sph)


def test_alignment_about_the_tilt_axis_shifts_the_limits():
    theta = 15.0
    a = avg_wedge(N, B, [_desc(-50, 50)], [1.0], [_rot("y", theta)])
    b = avg_wedge(N, B, [_desc(-50 - theta, 50 - theta)], [1.0], [np.eye(4)])
    diff = np.abs(a["accum"] - b["accum"]) > 1e-5
    assert diff.mean() < 1e-3, diff.sum()
    c = avg_wedge(N, B, [_desc(-50 + theta, 50 + theta)], [1.0], [np.eye(4)])
    assert (np.abs(a["accum"] - c["accum"]) > 1e-5).mean() > 0.01  # the other direction differs


def test_two_entries_binarise_to_the_intersection_and_skip_zero_weights():
    da, db = _desc(-60, 60), _desc(-60, 60)
    wa = avg_wedge(N, B, [da], [1.0], [np.eye(4)])["accum"]
    wb = avg_wedge(N, B, [db], [1.0], [_rot("z", 90)])["accum"]
    aw = avg_wedge(N, B, [da, _desc(-10, 10), db], [1.0, 0.0, -2.0], [np.eye(4), _rot("x", 40), _rot("z", 90)])
    np.testing.assert_allclose(aw["accum"], wa + wb, atol=1e-6)
    crisp = np.isin(wa, (0.0, 1.0)) & np.isin(wb, (0.0, 1.0))
    assert crisp.mean() > 0.3
    np.testing.assert_array_equal(aw["bin"][crisp], ((wa == 1) & (wb == 1))[crisp].astype(np.float32))
    assert aw["effective"]
    # smoothed binarised image: the 3-tap Gaussian keeps values in [0, 1] and ones deep inside the region
    assert aw["wedge"].min() >= 0.0 and aw["wedge"].max() <= 1.0 + 1e-6
