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


def _volume_list_text(recs) -> str:
    header = ("number\tlwedge\tuwedge\tposX\tposY\tposZ\tgeomX\tgeomY\tgeomZ\tnormalX\tnormalY\tnormalZ\t" +
              "".join(f"matrix[{i}]\t" for i in range(16)) +
              "magnification[0]\tmagnification[1]\tmagnification[2]\tcutOffset\tfilename")
    rows = ["\t".join([str(i + 1)] + [_fmt(v) for v in vals] + [fn]) for i, (vals, fn) in enumerate(recs)]
    return header + "\n" + "\n".join(rows) + "\n"


def _record(lw, uw, pos, mat, cut, fn):
    return [lw, uw, *pos, 64, 64, 64, 0.0, 0.0, 1.0, *mat, 1, 1, 1, cut], fn


@pytest.mark.parametrize("cols", [17, 19])
def test_classification_outputs(exe, tmp_path, cols):
    """nbfLoopClustering::doMra's per-iteration files (nbfLoopClustering.h:1595-1645) and saveState
    (:5542-5637), written from a known distance array, classes and averages: the .matlab array with the reference
    index fastest; per reference, the volume list with cutOffset = |distance| and the matrix read from column 1
    (17 columns) or 2 (longer rows, the reference's choice); the volume and class lists; the averages negated in
    float MRC files named by level and index, and their list at position N/2."""
    g = np.random.default_rng(cols)
    K, n, N = 2, 3, 4
    D = g.uniform(-1, 1, (K, n, cols)).astype(np.float32)
    D[:, :, 0] = -g.uniform(0.2, 0.9, (K, n)).astype(np.float32)
    avg = g.standard_normal((K, N, N, N)).astype(np.float32)
    cls = [1, -1, 0]
    recs = [_record(-60.0, 60.0, (10.5, 20.0, 30.25), np.eye(4).ravel(), 0.0, "p1.mrc"),
            _record(-45.0, 70.0, (1.0, 2.0, 3.0), np.eye(4).ravel(), 2.5, "p2.mrc"),
            _record(-90.0, 90.0, (7.0, 8.0, 9.0), np.eye(4).ravel(), 0.0, "p3.mrc")]
    text = _volume_list_text(recs)
    (tmp_path / "vols.txt").write_text(text)
    (tmp_path / "state_in.txt").write_text(f"{K} {n} {cols} {N}\n" + " ".join(map(str, cls)) + "\n")
    D.tofile(tmp_path / "D.raw")
    avg.tofile(tmp_path / "avg.raw")
    r = subprocess.run([str(exe), str(tmp_path), "state"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr

    m = (tmp_path / "run_iteration_007_alignmentToReferences.matlab").read_bytes()
    assert struct.unpack("<4i", m[:16]) == (K, n, cols, 1)
    assert np.array_equal(np.frombuffer(m[16:], np.float32), D.transpose(2, 1, 0).ravel())

    off = 2 if cols > 17 else 1
    for k in range(K):
        lines = (tmp_path / f"run_iteration_007_alignments_to_reference_{k}.txt").read_text().splitlines()
        assert lines[0] == text.splitlines()[0] and len(lines) == n + 1
        for j in range(n):
            f = lines[1 + j].split("\t")
            orig = text.splitlines()[1 + j].split("\t")
            assert f[:12] == orig[:12] and f[28:31] == orig[28:31] and f[32] == orig[32]
            assert f[31] == _fmt(abs(float(D[k, j, 0])))
            assert f[12:28] == [_fmt(float(x)) for x in D[k, j, off:off + 16]]

    assert (tmp_path / "run_state_volumes.txt").read_text() == text
    classes = (tmp_path / "run_state_classes.txt").read_text().splitlines()
    assert [ln.split("\t")[31] for ln in classes[1:]] == ["1", "0", "0"]
    assert [ln.split("\t")[:31] for ln in classes[1:]] == [ln.split("\t")[:31] for ln in text.splitlines()[1:]]

    names = [str(tmp_path / f"run_state_level_3_average_00{k}.mrc") for k in range(K)]
    for k in range(K):
        assert Path(names[k]).read_bytes() == _mrc_bytes(-avg[k])
        assert (tmp_path / f"run_idx_state_average_3_{4 + 5 * k}.mrc").read_bytes() == _mrc_bytes(-avg[k])
    lst = (tmp_path / "run_state_averages.txt").read_text()
    # default-constructed records: wedge -90 90, geometry and normal 0, identity, magnification 1, cutOffset 0
    want = [([-90.0, 90.0, N / 2, N / 2, N / 2, 0, 0, 0, 0, 0, 0, *np.eye(4).ravel(), 1, 1, 1, 0], names[k])
            for k in range(K)]
    assert lst == _volume_list_text(want)
