"""CPU checks of the C ABI boundary: the library loads, exports every entry point include/ettk_align.h declares,
and the ctypes mirror of the header structs has the C layout.  No GPU call is made."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ettk_amd import abi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "ettk_align.h"
LIB = ROOT / "ettk_amd" / "libettk_align.so"


def _declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ettk_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not LIB.exists():
        subprocess.run(["python", "-m", "ettk_amd.build", "--no-oracle"], cwd=ROOT, check=True)
    return C.CDLL(str(LIB))


def test_library_exports_every_declared_entry_point(lib):
    names = _declared_functions()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_is_callable_without_a_device(lib):
    lib.ettk_last_error.restype = C.c_char_p
    assert isinstance(lib.ettk_last_error(), bytes)


def test_null_arguments_are_rejected_without_a_device(lib):
    """Argument checks run before any device call: null handles come back as ETTK_ERR_ARG."""
    err_arg = _constant("ETTK_ERR_ARG")
    assert lib.ettk_plan_destroy(None) == 0  # destroying nothing is allowed
    assert lib.ettk_reference_set(None, None, None) == err_arg
    assert lib.ettk_ccc_eval_batch(None, None, 0, None, 0, None, None, None, None) == err_arg
    assert lib.ettk_align_batch(None, None, 0, None, 0, None) == err_arg
    assert lib.ettk_avg_finalize(None, None) == err_arg
    assert lib.ettk_avg_wedge(None, 0, None, None, None, None, None, None, None) == err_arg
    assert lib.ettk_reference_set_average(None, None, 0, None, None, None) == err_arg
    assert lib.ettk_sph_rotate_batch(None, 0, None, None, None) == err_arg
    assert lib.ettk_avg_variance(None, None, 0, None, None, None, None, None) == err_arg
    assert lib.ettk_fft3(None, 64, 1, 0, None, None, None) == err_arg
    lib.ettk_last_error.restype = C.c_char_p
    assert b"null" in lib.ettk_last_error() or b"bad" in lib.ettk_last_error()


def _constant(name: str) -> int:
    m = re.search(rf"#define\s+{name}\s+\(?(-?\d+)\)?", HEADER.read_text())
    assert m, name
    return int(m.group(1))


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """Compile a probe against the header with gcc and compare sizeof/offsetof with the ctypes mirror."""
    fields = {
        "ettk_params": [f for f, _ in abi.Params._fields_],
        "ettk_particle": [f for f, _ in abi.Particle._fields_],
        "ettk_result": [f for f, _ in abi.Result._fields_],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ettk_align.h"', "int main(void) {"]
    for s, fs in fields.items():
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f in fs:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                               check=True).stdout.splitlines())
    for s, cls in (("ettk_params", abi.Params), ("ettk_particle", abi.Particle), ("ettk_result", abi.Result)):
        assert int(out[s]) == C.sizeof(cls), s
        for f in fields[s]:
            assert int(out[f"{s}.{f}"]) == getattr(cls, f).offset, (s, f)
    assert _constant("ETTK_MAX_CANDIDATES") == abi.ETTK_MAX_CANDIDATES


def test_autograd_wrapper_backward_uses_the_returned_gradient():
    """The torch.autograd.Function returns grad_out * (gradient computed with the forward pass)."""
    torch = pytest.importorskip("torch")
    from ettk_amd.autograd import ccc_score

    class FakePlan:
        def ccc_grad(self, d, descs, pidx, seeds, x):
            val = np.sin(x).sum(1)
            return val, np.cos(x)

    x = torch.tensor(np.random.default_rng(0).uniform(-1, 1, (4, 6)), requires_grad=True)
    S = ccc_score(FakePlan(), None, None, list(range(4)), None, x)
    w = torch.arange(4, dtype=torch.float64)
    (w * S).sum().backward()
    assert torch.allclose(x.grad, w[:, None] * torch.cos(x.detach()))


def test_synthetic_particles_are_deterministic():
    from ettk_amd.synth import blob_reference, make_particles, rotation_error_deg
    ref = blob_reference(16, seed=1)
    a = make_particles(ref, 2, seed=5, snr=(0.5, 0.5))
    b = make_particles(ref, 2, seed=5, snr=(0.5, 0.5))
    assert np.array_equal(a.vols, b.vols) and np.array_equal(a.truth, b.truth)
    assert rotation_error_deg(a.truth[0], a.truth[0]) < 1e-6
    R = np.eye(4)
    R[:3, :3] = [[0, -1, 0], [1, 0, 0], [0, 0, 1]]
    assert abs(rotation_error_deg(np.eye(4), R) - 90.0) < 1e-9
