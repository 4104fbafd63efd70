"""The on-disk formats of include/ettk_io.hpp against independent writers of the reference's layouts.

* MRC (inc/io/nbfMrcWriter.h header() and write(), inc/io/nbfMrcReader.h): the 1024-byte header built field by
  field here, the data with x fastest and each section's rows in reverse y order, byte for byte; reading a
  mode-1 (int16) file with an extended header gives its values back in memory order.
* Volume lists (nbfWedgedSubImage3D::write / read, inc/em/nbfWedgedSubImage3D.h:343-600): 33 header names and
  tab-separated rows with C++'s default 6-significant-digit formatting, identical text after a read and write;
  a row whose wedge limits are 0 0 loads its "<file>.wedge.mrc" sidecar.
* "matlab" arrays (nbfMatlabWriter::write): int32 extents, then float32 values with the first index fastest.
"""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "tests" / "cpp" / "test_io"


def _mrc_bytes(vol_zyx: np.ndarray, mode: int = 2, ext: bytes = b"") -> bytes:
    """The reference writer's file for a [z][y][x] volume (its header() and the reversed-row data loop)."""
    nz, ny, nx = vol_zyx.shape
    zc = vol_zyx[nz // 2].astype(np.float32)
    h = struct.pack("<4i3i3i3f3f3i3f", nx, ny, nz, mode, 0, 0, 0, nx, ny, nz, 410.0, 410.0, 410.0, 90.0, 90.0, 90.0,
                    1, 2, 3, float(zc.min()), float(zc.max()), float(np.float32(zc.astype(np.float64).mean())))
    h += b"\0" * (1024 - len(h))
    dt = {0: np.uint8, 1: np.int16, 2: np.float32}[mode]
    return h + ext + vol_zyx[:, ::-1, :].astype(dt).tobytes()


def _fmt(x: float) -> str:
    """C++ ostream default formatting (%g with 6 significant digits)."""
    return "%g" % x


@pytest.fixture(scope="module")
def exe():
    if not EXE.exists():
        from ettk_amd.build import build_io_test
        build_io_test()
    return EXE


def test_formats_round_trip(exe, tmp_path):
    g = np.random.default_rng(3)
    nx, ny, nz = 12, 10, 7
    vol = g.standard_normal((nz, ny, nx)).astype(np.float32)
    (tmp_path / "dims.txt").write_text(f"{nx} {ny} {nz}\n")
    vol.tofile(tmp_path / "vol.raw")
    short = g.integers(-3000, 3000, (5, 6, 4)).astype(np.int16)
    (tmp_path / "in_short.mrc").write_bytes(_mrc_bytes(short, mode=1, ext=b"\x07" * 512))

    # a list with one stored-wedge row (limits 0 0) whose sidecar exists
    wedge = (g.random((4, 4, 4)) > 0.5).astype(np.float32)
    side = tmp_path / "part2.mrc"
    (tmp_path / "part2.mrc.wedge.mrc").write_bytes(_mrc_bytes(wedge))
    rows = []
    mats = [np.eye(4).ravel(), np.array([0.36, -0.48, 0.8, 1.5, 0.8, 0.6, 0.0, -2.25, -0.48, 0.64, 0.6, 0.125,
                                         0, 0, 0, 1])]
    header = ("number\tlwedge\tuwedge\tposX\tposY\tposZ\tgeomX\tgeomY\tgeomZ\tnormalX\tnormalY\tnormalZ\t" +
              "".join(f"matrix[{i}]\t" for i in range(16)) +
              "magnification[0]\tmagnification[1]\tmagnification[2]\tcutOffset\tfilename")
    recs = [(-60.0, 60.0, (120.5, 88.0, 40.25), (64, 64, 64), (10.0, -20.0, 33.3333), mats[0], (1, 1, 1), 0.0,
             str(tmp_path / "part1.mrc")),
            (0.0, 0.0, (1.0, 2.0, 3.0), (32, 32, 32), (0.0, 0.0, 0.0), mats[1], (1, 1, 1), 2.5, str(side))]
    for i, (lw, uw, pos, geo, nrm, m, mag, cut, fn) in enumerate(recs):
        vals = [lw, uw, *pos, *geo, *nrm, *m, *mag, cut]
        rows.append("\t".join([str(i + 1)] + [_fmt(v) for v in vals] + [fn]))
    text = header + "\n" + "\n".join(rows) + "\n"
    (tmp_path / "in_list.txt").write_text(text)

    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr

    assert (tmp_path / "out.mrc").read_bytes() == _mrc_bytes(vol)
    assert np.array_equal(np.fromfile(tmp_path / "roundtrip.raw", np.float32).reshape(vol.shape), vol)
    assert np.array_equal(np.fromfile(tmp_path / "in_short.raw", np.float32).reshape(short.shape),
                          short.astype(np.float32))
    assert (tmp_path / "out_list.txt").read_text() == text
    back = (tmp_path / "out_list17.txt").read_text().splitlines()
    assert back[0] == header and len(back) == 3
    assert float(back[2].split("\t")[12]) == 0.36  # matrix[0]: 17 digits keep the double
    assert np.array_equal(np.fromfile(tmp_path / "out_wedge.raw", np.float32).reshape(wedge.shape), wedge)

    m = (tmp_path / "out.matlab").read_bytes()
    assert struct.unpack("<4i", m[:16]) == (2, 3, 17, 1)
    a = np.frombuffer(m[16:], np.float32)
    assert np.array_equal(a, 0.5 * np.arange(2 * 3 * 17, dtype=np.float32) - 3.0)
