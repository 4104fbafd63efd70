"""ctypes binding of libettk_align.so (the C ABI of include/ettk_align.h).

Device buffers are torch tensors on the HIP device; the library receives their data pointers.  Importing
this module never falls back to another implementation: if the shared library is missing, loading fails.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .abi import Bucket, Params, Particle, Result, particle_array

LIB_PATH = Path(__file__).resolve().parent / "libettk_align.so"
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m ettk_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        vp, ip, dp, fp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_float)
        sig = {
            "ettk_last_error": ([], C.c_char_p),
            "ettk_ctx_create": ([C.c_int, C.POINTER(vp)], C.c_int),
            "ettk_ctx_destroy": ([vp], C.c_int),
            "ettk_plan_create": ([vp, C.POINTER(Params), C.POINTER(vp)], C.c_int),
            "ettk_plan_destroy": ([vp], C.c_int),
            "ettk_plan_set_restrictions": ([vp, C.c_float, C.c_float], C.c_int),
            "ettk_reference_set": ([vp, vp, C.POINTER(Particle)], C.c_int),
            "ettk_ccc_eval_batch": ([vp, vp, C.c_int, C.POINTER(Particle), C.c_int, ip, dp, fp, dp], C.c_int),
            "ettk_ccc_eval_batch_noalloc": ([vp, vp, C.c_int, C.POINTER(Particle), C.c_int, ip, dp, fp, dp],
                                            C.c_int),
            "ettk_project_batch": ([vp, vp, C.c_int, C.POINTER(Particle), fp], C.c_int),
            "ettk_so3_correlate_batch": ([vp, C.c_int, dp, dp, dp], C.c_int),
            "ettk_so3_correlate_scaled_batch": ([vp, C.c_int, dp, dp, dp], C.c_int),
            "ettk_sph_rotate_batch": ([vp, C.c_int, fp, dp, fp], C.c_int),
            "ettk_align_batch": ([vp, vp, C.c_int, C.POINTER(Particle), C.c_int, C.POINTER(Result)], C.c_int),
            "ettk_distances_batch": ([vp, vp, C.c_int, C.POINTER(Particle), C.c_int, C.c_int, C.POINTER(Result)],
                                     C.c_int),
            "ettk_ccc_grad_batch": ([vp, vp, C.c_int, C.POINTER(Particle), C.c_int, ip, dp, dp, dp, dp], C.c_int),
            "ettk_avg_reset": ([vp], C.c_int),
            "ettk_avg_accumulate": ([vp, vp, C.c_int, C.POINTER(Particle), fp, dp], C.c_int),
            "ettk_avg_finalize": ([vp, vp], C.c_int),
            "ettk_avg_variance": ([vp, vp, C.c_int, C.POINTER(Particle), fp, dp, vp, vp], C.c_int),
            "ettk_avg_wedge": ([vp, C.c_int, C.POINTER(Particle), fp, dp, vp, vp, fp, ip], C.c_int),
            "ettk_reference_set_average": ([vp, vp, C.c_int, C.POINTER(Particle), fp, dp], C.c_int),
            "ettk_last_timing": ([vp, dp, C.POINTER(C.c_longlong), dp], C.c_int),
            "ettk_refine_batch": ([vp, vp, C.c_int, C.POINTER(Particle), C.c_int, ip, dp, C.c_int, dp, dp], C.c_int),
            "ettk_fft3": ([vp, C.c_int, C.c_int, C.c_int, vp, vp, vp], C.c_int),
            "ettk_mra_iteration": ([vp, C.c_int, vp, ip, dp, vp, C.c_int, C.POINTER(Particle), C.c_int, C.c_int,
                                    C.c_float, ip, ip, fp, dp], C.c_int),
            "ettk_plan_tune": ([vp, vp, C.c_int, C.POINTER(Particle), ip, ip, dp], C.c_int),
            "ettk_plan_set_execution": ([vp, C.c_int, C.c_int], C.c_int),
            "ettk_plan_get_execution": ([vp, ip, ip], C.c_int),
            "ettk_align_buckets": ([vp, C.c_int, C.POINTER(Bucket), C.c_int, C.c_int], C.c_int),
            "ettk_extract_batch": ([vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.POINTER(Particle), C.c_int, vp],
                                   C.c_int),
            "ettk_profile": ([vp, C.c_int], C.c_int),
            "ettk_profile_query": ([vp, C.c_char_p, dp, C.POINTER(C.c_longlong), dp], C.c_int),
            "ettk_profile_only": ([vp, C.c_char_p], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


class EttkError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc != 0:
        raise EttkError(f"ettk error {rc}: {load().ettk_last_error().decode()}")


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Context:
    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = C.c_void_p()
        check(self.lib.ettk_ctx_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            check(self.lib.ettk_ctx_destroy(self.h))
            self.h = C.c_void_p()

    FFT_R2C, FFT_C2R, FFT_MASK, FFT_SQUARE = 0, 1, 2, 3

    def fft3(self, kind: int, n: int, batch: int, d_in, d_out, d_aux=None) -> None:
        """ettk_fft3 on device tensors: real volumes [batch][n][n][n] float32, half spectra
        [batch][n][n][n/2+1] complex64 (as the kind expects), d_aux the n^3 float32 mask of FFT_MASK."""
        aux = C.c_void_p(d_aux.data_ptr()) if d_aux is not None else None
        check(self.lib.ettk_fft3(self.h, n, batch, kind, C.c_void_p(d_in.data_ptr()), C.c_void_p(d_out.data_ptr()),
                                 aux))

    def extract(self, d_tomo, positions, descs, N: int, d_out) -> None:
        """ettk_extract_batch: d_tomo (nz, ny, nx) float32 tensor, positions (n, 3) voxel coordinates (x, y, z),
        d_out (n, N, N, N) float32 tensor."""
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
        nz, ny, nx = d_tomo.shape
        check(self.lib.ettk_extract_batch(self.h, C.c_void_p(d_tomo.data_ptr()), nx, ny, nz, pos.shape[0], _dp(pos),
                                          particle_array(descs), N, C.c_void_p(d_out.data_ptr())))

    def align_buckets(self, buckets, mode: int = 1, flags: int = 0):
        """ettk_align_buckets over a list of (params, d_reference, reference Particle, d_particles, descs);
        returns one list of results per bucket."""
        arr = (Bucket * len(buckets))()
        keep = []
        outs = []
        for i, (params, d_ref, rdesc, d_parts, descs) in enumerate(buckets):
            n = d_parts.shape[0] if d_parts is not None else 0
            res = (Result * max(n, 1))()
            da = particle_array(descs) if n else None
            keep += [res, da]
            arr[i].params = params
            arr[i].d_reference = d_ref.data_ptr()
            arr[i].reference = rdesc
            arr[i].d_particles = d_parts.data_ptr() if n else None
            arr[i].n = n
            arr[i].descs = C.cast(da, C.POINTER(Particle)) if n else None
            arr[i].out = C.cast(res, C.POINTER(Result))
            outs.append((res, n))
        check(self.lib.ettk_align_buckets(self.h, len(buckets), arr, mode, flags))
        return [list(res)[:n] for res, n in outs]

    def profile(self, enable: bool) -> None:
        check(self.lib.ettk_profile(self.h, int(enable)))

    def profile_only(self, kernel: str | None) -> None:
        """Time only the launches of `kernel` (None: every instrumented kernel)."""
        check(self.lib.ettk_profile_only(self.h, kernel.encode() if kernel else None))

    def profile_query(self, kernel: str):
        """(total ms, launches, algorithmic bytes) of one instrumented kernel since ettk_profile(ctx, 1)."""
        ms = C.c_double()
        n = C.c_longlong()
        by = C.c_double()
        check(self.lib.ettk_profile_query(self.h, kernel.encode(), C.byref(ms), C.byref(n), C.byref(by)))
        return ms.value, n.value, by.value


class Plan:
    """One box size / bandwidth / parameter set (ettk_plan)."""

    def __init__(self, ctx: Context, params: Params):
        self.lib = ctx.lib
        self.ctx = ctx
        self.params = params
        self.h = C.c_void_p()
        check(self.lib.ettk_plan_create(ctx.h, C.byref(params), C.byref(self.h)))
        self._keep = []

    def close(self):
        if self.h:
            check(self.lib.ettk_plan_destroy(self.h))
            self.h = C.c_void_p()

    def set_reference(self, d_ref, desc: Particle | None = None):
        desc = desc or Particle()
        self._keep = [d_ref]
        check(self.lib.ettk_reference_set(self.h, C.c_void_p(d_ref.data_ptr()), C.byref(desc)))

    def ccc(self, d_parts, descs, pidx, transforms):
        n = len(pidx)
        arr = particle_array(descs)
        pidx = np.ascontiguousarray(pidx, dtype=np.int32)
        t = np.ascontiguousarray(np.asarray(transforms, dtype=np.float64).reshape(n, 16))
        dist = np.zeros(n, np.float32)
        tout = np.zeros((n, 16), np.float64)
        check(self.lib.ettk_ccc_eval_batch(self.h, C.c_void_p(d_parts.data_ptr()), d_parts.shape[0], arr, n,
                                           _ip(pidx), _dp(t), _fp(dist), _dp(tout)))
        return dist, tout.reshape(n, 4, 4)

    def ccc_noalloc(self, d_parts, descs, pidx, transforms):
        """ettk_ccc_eval_batch_noalloc: as ccc(), with the library's fixed scratch region as workspace."""
        n = len(pidx)
        arr = particle_array(descs)
        pidx = np.ascontiguousarray(pidx, dtype=np.int32)
        t = np.ascontiguousarray(np.asarray(transforms, dtype=np.float64).reshape(n, 16))
        dist = np.zeros(n, np.float32)
        tout = np.zeros((n, 16), np.float64)
        check(self.lib.ettk_ccc_eval_batch_noalloc(self.h, C.c_void_p(d_parts.data_ptr()), d_parts.shape[0], arr, n,
                                                   _ip(pidx), _dp(t), _fp(dist), _dp(tout)))
        return dist, tout.reshape(n, 4, 4)

    def project(self, d_parts, descs):
        n = d_parts.shape[0]
        B = self.params.B
        out = np.zeros((n, 2 * B, 2 * B), np.float32)
        check(self.lib.ettk_project_batch(self.h, C.c_void_p(d_parts.data_ptr()), n, particle_array(descs), _fp(out)))
        return out

    def so3_correlate(self, sig: np.ndarray, pat: np.ndarray, scaled: bool = False) -> np.ndarray:
        """soft (ettk_so3_correlate_batch) or softScaled (ettk_so3_correlate_scaled_batch) of signal/pattern pairs."""
        B = self.params.B
        sig = np.ascontiguousarray(sig, dtype=np.float64).reshape(-1, 2 * B, 2 * B)
        pat = np.ascontiguousarray(pat, dtype=np.float64).reshape(-1, 2 * B, 2 * B)
        n = sig.shape[0]
        out = np.zeros((n, 2 * B, 2 * B, 2 * B), np.float64)
        fn = self.lib.ettk_so3_correlate_scaled_batch if scaled else self.lib.ettk_so3_correlate_batch
        check(fn(self.h, n, _dp(sig), _dp(pat), _dp(out)))
        return out

    def sph_rotate(self, images: np.ndarray, transforms) -> np.ndarray:
        """harmonics(A, t, R) of n (2B, 2B) spherical images (ettk_sph_rotate_batch)."""
        img = np.ascontiguousarray(images, dtype=np.float32)
        n = img.shape[0]
        t = np.ascontiguousarray(np.asarray(transforms, dtype=np.float64).reshape(n, 16))
        out = np.zeros_like(img)
        check(self.lib.ettk_sph_rotate_batch(self.h, n, _fp(img), _dp(t), _fp(out)))
        return out

    def align(self, d_parts, descs, mode: int = 0):
        n = d_parts.shape[0]
        res = (Result * n)()
        check(self.lib.ettk_align_batch(self.h, C.c_void_p(d_parts.data_ptr()), n, particle_array(descs), mode, res))
        return list(res)

    def set_restrictions(self, rot_restrict: float, trans_restrict: float):
        check(self.lib.ettk_plan_set_restrictions(self.h, rot_restrict, trans_restrict))
        self.params.rot_restrict = rot_restrict
        self.params.trans_restrict = trans_restrict

    def distances(self, d_parts, descs, type: int = 0, form: int = 1):
        """getDistances semantics by metric type (0 search+refine, 1 refine only, 2 shift only, 3 search)."""
        n = d_parts.shape[0]
        res = (Result * n)()
        check(self.lib.ettk_distances_batch(self.h, C.c_void_p(d_parts.data_ptr()), n, particle_array(descs), type,
                                            form, res))
        return list(res)

    def ccc_grad(self, d_parts, descs, pidx, seeds, params6):
        n = len(pidx)
        pidx = np.ascontiguousarray(pidx, dtype=np.int32)
        seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.float64).reshape(n, 16))
        x = np.ascontiguousarray(np.asarray(params6, dtype=np.float64).reshape(n, 6))
        val = np.zeros(n, np.float64)
        grad = np.zeros((n, 6), np.float64)
        check(self.lib.ettk_ccc_grad_batch(self.h, C.c_void_p(d_parts.data_ptr()), d_parts.shape[0],
                                           particle_array(descs), n, _ip(pidx), _dp(seeds), _dp(x), _dp(val),
                                           _dp(grad)))
        return val, grad

    def refine(self, d_parts, descs, pidx, seeds, n_iter: int = 20):
        n = len(pidx)
        pidx = np.ascontiguousarray(pidx, dtype=np.int32)
        seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.float64).reshape(n, 16))
        poses = np.zeros((n, 16), np.float64)
        scores = np.zeros(n, np.float64)
        check(self.lib.ettk_refine_batch(self.h, C.c_void_p(d_parts.data_ptr()), d_parts.shape[0], particle_array(descs),
                                         n, _ip(pidx), _dp(seeds), int(n_iter), _dp(poses), _dp(scores)))
        return poses.reshape(n, 4, 4), scores

    def average(self, d_parts, descs, weights, matrices, d_out):
        n = d_parts.shape[0]
        w = np.ascontiguousarray(weights, dtype=np.float32)
        m = np.ascontiguousarray(np.asarray(matrices, dtype=np.float64).reshape(n, 16))
        check(self.lib.ettk_avg_reset(self.h))
        check(self.lib.ettk_avg_accumulate(self.h, C.c_void_p(d_parts.data_ptr()), n, particle_array(descs), _fp(w),
                                           _dp(m)))
        check(self.lib.ettk_avg_finalize(self.h, C.c_void_p(d_out.data_ptr())))

    def variance(self, d_parts, descs, weights, matrices, d_average, d_out):
        """Variance map of an average around d_average (ettk_avg_variance)."""
        n = d_parts.shape[0]
        w = np.ascontiguousarray(weights, dtype=np.float32)
        m = np.ascontiguousarray(np.asarray(matrices, dtype=np.float64).reshape(n, 16))
        check(self.lib.ettk_avg_variance(self.h, C.c_void_p(d_parts.data_ptr()), n, particle_array(descs), _fp(w),
                                         _dp(m), C.c_void_p(d_average.data_ptr()), C.c_void_p(d_out.data_ptr())))

    def avg_wedge(self, descs, weights, matrices, d_accum=None, d_wedge=None):
        """Accumulated wedge of an average (ettk_avg_wedge): fills the optional device arrays (N^3 float,
        k fastest) and returns (spherical (2B, 2B) window, effective)."""
        n = len(descs)
        w = np.ascontiguousarray(weights, dtype=np.float32)
        m = np.ascontiguousarray(np.asarray(matrices, dtype=np.float64).reshape(n, 16))
        B = self.params.B
        sph = np.zeros((2 * B, 2 * B), np.float32)
        eff = C.c_int()
        ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        check(self.lib.ettk_avg_wedge(self.h, n, particle_array(descs), _fp(w), _dp(m), ptr(d_accum), ptr(d_wedge),
                                      _fp(sph), C.byref(eff)))
        return sph, bool(eff.value)

    def set_reference_average(self, d_ref, descs, weights, matrices):
        """Reference that is an average of the given aligned particles (ettk_reference_set_average)."""
        n = len(descs)
        w = np.ascontiguousarray(weights, dtype=np.float32)
        m = np.ascontiguousarray(np.asarray(matrices, dtype=np.float64).reshape(n, 16))
        self._keep = [d_ref]
        check(self.lib.ettk_reference_set_average(self.h, C.c_void_p(d_ref.data_ptr()), n, particle_array(descs),
                                                  _fp(w), _dp(m)))

    def mra_iteration(self, d_refs, d_parts, descs, type: int = 0, form: int = 2, top_percentage: float = 1.0,
                      ref_members=None, member_matrices=None):
        """ettk_mra_iteration: d_refs (K, N, N, N) float32 tensor, replaced by the new class averages.  Returns
        (class, selected, distance, matrices (n, 4, 4))."""
        K = d_refs.shape[0]
        n = d_parts.shape[0]
        cls = np.zeros(n, np.int32)
        sel = np.zeros(n, np.int32)
        dist = np.zeros(n, np.float32)
        mats = np.zeros((n, 16), np.float64)
        rm = mm = None
        if ref_members is not None:
            rm_a = np.ascontiguousarray(ref_members, dtype=np.int32)
            mm_a = np.ascontiguousarray(np.asarray(member_matrices, dtype=np.float64).reshape(n, 16))
            rm, mm = _ip(rm_a), _dp(mm_a)
        check(self.lib.ettk_mra_iteration(self.h, K, C.c_void_p(d_refs.data_ptr()), rm, mm,
                                          C.c_void_p(d_parts.data_ptr()), n, particle_array(descs), type, form,
                                          float(top_percentage), _ip(cls), _ip(sel), _fp(dist), _dp(mats)))
        return cls, sel, dist, mats.reshape(n, 4, 4)

    def tune(self, d_parts, descs):
        """ettk_plan_tune on these particles: (chunk, graphs, microseconds per evaluation)."""
        ch, gr, us = C.c_int(), C.c_int(), C.c_double()
        check(self.lib.ettk_plan_tune(self.h, C.c_void_p(d_parts.data_ptr()), d_parts.shape[0], particle_array(descs),
                                      C.byref(ch), C.byref(gr), C.byref(us)))
        return ch.value, bool(gr.value), us.value

    def set_execution(self, chunk: int, graphs: bool) -> None:
        check(self.lib.ettk_plan_set_execution(self.h, int(chunk), int(graphs)))

    def execution(self):
        ch, gr = C.c_int(), C.c_int()
        check(self.lib.ettk_plan_get_execution(self.h, C.byref(ch), C.byref(gr)))
        return ch.value, bool(gr.value)

    def last_timing(self):
        ms = C.c_double()
        n = C.c_longlong()
        tot = C.c_double()
        check(self.lib.ettk_last_timing(self.h, C.byref(ms), C.byref(n), C.byref(tot)))
        return ms.value, n.value, tot.value
