"""Writes the soft-1.0 fixtures under tests/golden/soft/ (run in a checkout that has the reference tree).

* D101{real,imag}_bw4.dat, D3-11{real,imag}_bw4.dat, dSum_{real,imag}_bw4.dat: sample files shipped with the
  reference's soft-1.0 (external/soft-1.0): Wigner-D functions sampled on the bw = 4 SO(3) grid, copied as data.
* randomS2sig_bw8.dat, randomS2sigA_bw8.dat: the two real S^2 signals soft-1.0 ships for its correlation test,
  copied as data.
* correlate_bw8.npy: the SO(3) correlation of those two signals, [beta][alpha][gamma], computed by soft-1.0
  itself (so3CombineCoef + Inverse_SO3_Naive_sym), compiled from the reference sources by oracle/Makefile's
  `ref` target into oracle/_ref/libsoftref.so.

Usage: python tests/golden/make_soft_fixtures.py [/root/reference]
"""
import ctypes as C
import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
FILES = ["D101real_bw4.dat", "D101imag_bw4.dat", "D3-11real_bw4.dat", "D3-11imag_bw4.dat", "dSum_real_bw4.dat",
         "dSum_imag_bw4.dat", "randomS2sig_bw8.dat", "randomS2sigA_bw8.dat"]


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "external" / "soft-1.0"
    out = HERE / "soft"
    out.mkdir(exist_ok=True)
    for f in FILES:
        shutil.copyfile(ref / f, out / f)
    lib = C.CDLL(str(ROOT / "oracle" / "_ref" / "libsoftref.so"))
    bw = 8
    n = 2 * bw
    sig = np.ascontiguousarray(np.loadtxt(out / "randomS2sig_bw8.dat")[0::2])
    pat = np.ascontiguousarray(np.loadtxt(out / "randomS2sigA_bw8.dat")[0::2])
    res = np.zeros(n * n * n)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    lib.ref_soft_correlate(bw, vp(sig), vp(pat), vp(res))
    np.save(out / "correlate_bw8.npy", res.reshape(n, n, n))


if __name__ == "__main__":
    main()
