"""Benchmark of the sub-volume alignment hot path on one MI355X.

Workload (the plain GPU scale-up configuration): 1024 particles of 64^3 fp32 from a 40-blob reference,
per-particle missing wedges +-U[45,70] degrees about Y, SNR U[0.05,0.2], SH bandwidth B=32, aligned to the
reference (S^2 projection, SO(3) search, direct/mirror constrained CCC at Kp=5 peaks, Kc=1 refined candidate).
Inputs are generated on the GPU and resident in HBM before timing starts.

A step is one alignment pass of the headline form over the whole batch.  Forms measured:
  plain            - the reference's control flow particle after particle (one CCC per launch), on a subset
  batched          - every stage over all particles; per-particle control flow served by batched launches
  batched+adjoint  - the batched search and candidate CCCs, then 20 steps of the continuous refinement with
                     the hand-written adjoint (the headline form)
The JSON line reports the headline form's particles/s as `value`, the other forms and the ratios beside it,
a roofline record for the dominant kernel (HIP events inside the library, on its stream), a pass-model
roofline per form, the headline form's pose errors against the same form in FP64 on the first --cpu-n
particles, and the FP64 restatement of the reference method timed on this host's cores as a reported
baseline (all threads of the process's share, and one core).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "aligned particles/s, 64^3 B=32: batched+adjoint vs plain GPU at ≤0.5° error"
HBM_PEAK_GBS = 8000.0


def measured_traffic_ratio(kernel: str):
    """HBM traffic / algorithmic bytes of `kernel` from the committed rocprof counter summaries (FETCH_SIZE and
    WRITE_SIZE passes, calibrated on a copy kernel); (None, None) when no summary covers the kernel."""
    prof = os.path.join(ROOT, "profiles")
    if not os.path.isdir(prof):
        return None, None
    for f in sorted(os.listdir(prof)):
        if not (f.startswith("pmc_traffic") and f.endswith(".md")):
            continue
        for line in open(os.path.join(prof, f)):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 6 and cells[0] == kernel:
                try:
                    return float(cells[5]), f"profiles/{f}"
                except ValueError:
                    pass
    return None, None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024, help="particles per step")
    ap.add_argument("--plain-n", type=int, default=16, help="particles timed in the plain form")
    ap.add_argument("--cpu-n", type=int, default=32,
                    help="particles of the FP64 oracle legs: accuracy of the headline form and the CPU baseline "
                         "(0 = skip)")
    ap.add_argument("--cpu-one-core", type=int, default=1, help="particles of the single-core CPU baseline")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--forms", default="plain,batched,adjoint",
                    help="forms to run (the headline form 'adjoint' always runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 256 MiB device copy (a known-size kernel for rocprof counter calibration)")
    ap.add_argument("--workload", default="scale-up", choices=["scale-up", "uneven", "refine96", "classify128"],
                    help="scale-up: the headline measurement; uneven: particles of box sizes 48/64/96/128; "
                         "refine96: continuous refinement of 96^3 particles from seeds 3 deg / 1 voxel off; "
                         "classify128: multi-reference classification of 128^3 particles from 8 references")
    ap.add_argument("--iters", type=int, default=5, help="classify128: classification iterations")
    args = ap.parse_args()
    forms = set(args.forms.split(","))
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU: --gpus must be 1 (sharding across GPUs is designed, not built)"}))
        return 2
    if args.workload == "uneven":
        return uneven(args)
    if args.workload == "refine96":
        return refine96(args)
    if args.workload == "classify128":
        return classify128(args)

    import torch

    from ettk_amd.abi import default_params
    from ettk_amd.lib import Context, Plan
    from ettk_amd.synth import blob_reference, make_particles_device, rotation_error_deg

    if args.copy_calibration:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    N, B = 64, 32
    p = default_params(N, B)
    ref = blob_reference(N, seed=args.seed)
    t0 = time.time()
    d_parts, descs, truth = make_particles_device(ref, args.n, seed=args.seed + 1, half_wedge=(45.0, 70.0),
                                                  snr=(0.05, 0.2))
    gen_s = time.time() - t0
    ctx = Context(0)
    plan = Plan(ctx, p)
    d_ref = torch.from_numpy(ref).cuda()
    torch.cuda.synchronize()
    plan.set_reference(d_ref)

    # ---------------------------------------------------------------- plain form on a subset
    plain_pps, res_plain = None, []
    if "plain" in forms:
        sub = d_parts[: args.plain_n].contiguous()
        torch.cuda.synchronize()
        plan.align(sub, descs[: args.plain_n], mode=0)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res_plain = plan.align(sub, descs[: args.plain_n], mode=0)
        torch.cuda.synchronize()
        plain_pps = args.plain_n / (time.perf_counter() - t0)
        plain_evals = plan.last_timing()[1] / args.plain_n

    # ---------------------------------------------------------------- batched form (Powell, the reference method)
    batched_pps, res_b = None, []
    if "batched" in forms:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res_b = plan.align(d_parts, descs, mode=1)
        torch.cuda.synchronize()
        batched_pps = args.n / (time.perf_counter() - t0)
        batched_evals = plan.last_timing()[1] / args.n

    # ---------------------------------------------------------------- batched + adjoint form (headline)
    for _ in range(args.warmup):
        plan.align(d_parts, descs, mode=2)
    torch.cuda.synchronize()
    # one untimed step with HIP events around every instrumented launch: the per-kernel breakdown, the host stages
    # and the dominant kernel; the timed steps then carry events around that kernel's launches only
    ctx.profile(True)
    ctx.profile_only(None)
    plan.align(d_parts, descs, mode=2)
    torch.cuda.synchronize()
    kernels = {}
    for name in ("k_fft_strided<64>", "k_r2c_plane<64>", "k_c2r_plane<64>", "k_c2r_r2c_plane<64>", "k_ccc_ipass<64>",
                 "k_r2c_rows<64>", "k_c2r_rows<64>", "k_reslice", "k_spec_filter", "k_mask_ref",
                 "k_combine_dftA", "k_dftBC_argmax", "k_combine_fftA", "k_dftBC_fft", "k_fft_strided<128>", "k_r2c_rows<128>",
                 "k_reslice_ref", "k_refine_dot", "k_refine_adj", "k_refine_ipass_dot", "k_refine_ipass_adj", "k_refine_moments", "k_sph_sample",
                 "k_so3_synth_fft", "k_so3_synth_mfma", "k_so3_fft2", "k_rotate_coef", "k_extrema", "k_combine",
                 "k_localmax"):
        ms, nl, by = ctx.profile_query(name)
        if nl:
            kernels[name] = {"ms": ms, "launches": nl, "bytes": by}
    host = {}
    for name in ("search_stage_wall", "search_inputs", "search_enqueue", "search_host", "search_download", "search_tail",
                 "candidate_stage_wall", "refine_wall",
                 "ccc_batch_wall", "lm_count"):
        ms, nl, _ = ctx.profile_query(name)
        host[name] = {"total": round(ms, 2), "count": nl}
    dom_name = max(kernels.items(), key=lambda kv: kv[1]["ms"])[0]
    ctx.profile(True)
    ctx.profile_only(dom_name)
    step_ms = []
    res = None
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = plan.align(d_parts, descs, mode=2)
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - t0) * 1e3)
    adjoint_evals = plan.last_timing()[1] / args.n  # CCC evaluations the device ran in the last timed step
    ms, nl, by = ctx.profile_query(dom_name)
    timed_dom = {"ms": ms, "launches": nl, "bytes": by}
    ctx.profile_only(None)
    ctx.profile(False)
    ms_per_step = float(np.mean(step_ms))
    value = args.n / (ms_per_step / 1e3)

    # consistency of the two forms on the shared subset, and accuracy against the ground truth
    same = sum(1 for a, b in zip(res_plain, res_b[: args.plain_n])
               if a.n_ccc == b.n_ccc and np.array_equal(a.transform(), b.transform()))
    rot_err = np.array([rotation_error_deg(r.transform(), truth[i]) for i, r in enumerate(res)])
    sh_err = np.array([np.abs(r.transform()[:3, 3] - truth[i][:3, 3]).max() for i, r in enumerate(res)])
    n_ccc = np.array([r.n_ccc for r in res])
    batched = None
    if res_b:
        rot_err_b = np.array([rotation_error_deg(r.transform(), truth[i]) for i, r in enumerate(res_b)])
        sh_err_b = np.array([np.abs(r.transform()[:3, 3] - truth[i][:3, 3]).max() for i, r in enumerate(res_b)])
        batched = {"particles_per_s": round(batched_pps, 3), "particles": args.n,
                   "ccc_per_particle": round(float(np.mean([r.n_ccc for r in res_b])), 2),
                   "rot_err_deg_median": round(float(np.median(rot_err_b)), 3),
                   "shift_err_vox_median": round(float(np.median(sh_err_b)), 3)}

    # ---------------------------------------------------------------- roofline of the dominant kernel
    # (its launches in the timed steps, timed by HIP events on the plan's stream)
    dom = (dom_name, timed_dom)
    avg_ms = dom[1]["ms"] / dom[1]["launches"]
    bytes_per_launch = dom[1]["bytes"] / dom[1]["launches"]
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
    ratio, traffic_src = measured_traffic_ratio(dom[0])
    roofline = {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(ratio * bytes_per_launch) if ratio else None,
                "traffic_source": traffic_src,
                "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": bytes_per_launch}
    if dom[0] == "k_reslice":
        # the tricubic reslice is arithmetic-bound: 64 taps x 2 FLOP per output voxel (33.5 MFLOP per
        # evaluation at 64^3); evaluations per launch from its algorithmic bytes (2 R per evaluation)
        n3 = 64 ** 3
        evals = bytes_per_launch / (2 * n3 * 4)
        flops = evals * n3 * 128
        roofline["compute"] = {"achieved": round(flops / (avg_ms / 1e3) / 1e12, 2), "peak": 157.3,
                               "unit": "TFLOP/s", "frac": round(flops / (avg_ms / 1e3) / 157.3e12, 4),
                               "flops_per_launch": flops, "basis": "64-tap tricubic, 128 FLOP per output voxel, FP32 VALU"}

    # ---------------------------------------------------------------- per-form roofline (pass model)
    # bytes per particle of the algorithm each form executes, with the CCC evaluations the device ran (the library's
    # count; the reference method's count, ccc_per_particle, also includes the batched + adjoint form's refine(t, 0)
    # evaluation, which reuses the candidate's own evaluation):
    # a pruned constrained CCC (the upsampled inverse transform evaluated on the search window only) moves
    # 5R + 11H; the S^2 projection R + 2H2; the SO(3) correlations 3 x 11.9 MB at B = 32 (scaled by (B/32)^3);
    # one continuous-refinement iteration 2R (reslice) + R + H (r2c) + 3H (spectral) + H + R (c2r) + 3R (moments)
    R, H = N ** 3 * 4.0, N * N * (N // 2 + 1) * 8.0
    R2, H2 = (2 * N) ** 3 * 4.0, (2 * N) ** 2 * (N + 1) * 8.0
    search_b = R + 2 * H2 + 3 * 11.9e6 * (B / 32.0) ** 3
    ccc_b = 5 * R + 11 * H
    refine_it_b = 7 * R + 5 * H

    def form_roofline(ccc_per_particle, refine_iters, pps):
        bpp = search_b + ccc_per_particle * ccc_b + refine_iters * refine_it_b
        ach = bpp * pps / 1e9
        return {"bytes_per_particle": round(bpp), "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4)}
    forms_roofline = {"basis": "pass model: search R+2H2+3x11.9MB(B/32)^3, pruned CCC 5R+11H per evaluation the "
                               "device ran, refinement iteration 7R+5H"}
    if plain_pps:
        forms_roofline["plain"] = form_roofline(plain_evals, 0, plain_pps)
    if batched_pps:
        forms_roofline["batched"] = form_roofline(batched_evals, 0, batched_pps)
    forms_roofline["batched+adjoint"] = form_roofline(adjoint_evals, 20, value)

    # ---------------------------------------------------------------- FP64 oracle legs (host cores)
    cpu, acc = None, None
    if args.cpu_n > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import OraclePlan

        try:
            aff = len(os.sched_getaffinity(0))
        except AttributeError:
            aff = os.cpu_count() or 1
        try:
            import subprocess
            lines = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True).stdout.splitlines()
            physical = len({ln for ln in lines if ln and not ln.startswith("#")})
        except OSError:
            physical = aff
        threads = max(1, min(physical, aff, 16))  # the process's share of the host (16 per GPU)
        vols = d_parts[: args.cpu_n].cpu().numpy()
        orc = OraclePlan(p)
        orc.set_reference(ref)
        # accuracy of the headline form: the same particles through the batched + adjoint form in FP64
        t0 = time.perf_counter()
        r64, _ = orc.align_adjoint_many(vols, descs[: args.cpu_n], 20, threads)
        acc_s = time.perf_counter() - t0
        ea = np.array([rotation_error_deg(res[i].transform(), r64[i].transform()) for i in range(args.cpu_n)])
        es = np.array([np.abs(res[i].transform()[:3, 3] - r64[i].transform()[:3, 3]).max() for i in range(args.cpu_n)])
        acc = {"particles": args.cpu_n, "rot_err_deg_max": round(float(ea.max()), 4),
               "rot_err_deg_median": round(float(np.median(ea)), 4),
               "shift_err_vox_max": round(float(es.max()), 4),
               "frac_rot_within_0.5deg": round(float(np.mean(ea <= 0.5)), 4),
               "frac_shift_within_0.1vox": round(float(np.mean(es <= 0.1)), 4),
               "oracle": "FP64 batched+adjoint form: search, candidate CCCs, 20 ascent steps with FP64 gradients, "
                         "final CCC", "oracle_s": round(acc_s, 1)}
        # CPU baseline: the reference method (search + Powell refinement) in FP64, one particle per thread
        t0 = time.perf_counter()
        orc.align_many(vols, descs[: args.cpu_n], nthreads=threads)
        cpu_s = time.perf_counter() - t0
        n1 = max(1, args.cpu_one_core)
        t0 = time.perf_counter()
        orc.align_many(vols[:n1], descs[:n1], nthreads=1)
        one_s = time.perf_counter() - t0
        cpu = {"value": round(args.cpu_n / cpu_s, 4), "unit": "particles/s", "cores": threads, "kind": "port",
               "sample": f"first {args.cpu_n} particles of this workload through the reference method (S2/SO(3) "
                         f"search, direct/mirror CCCs, Powell refinement) in the FP64 restatement, one particle per "
                         f"OpenMP thread on {threads} threads, {cpu_s:.1f} s wall",
               "one_core": {"value": round(n1 / one_s, 4), "particles": n1, "wall_s": round(one_s, 1)},
               "host": {"physical_cores": physical, "affinity_cpus": aff}}

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (40-blob reference, rotated/shifted/wedged/noisy particles generated on the GPU)",
        "config": {"workload": "plain GPU scale-up: 1024 particles 64^3, B=32, wedges +-U[45,70], SNR U[0.05,0.2]",
                   "particles": args.n, "box": N, "bandwidth": B, "form": "batched+adjoint",
                   "Kp": p.n_peaks, "Kc": p.n_candidates, "trans_restrict": p.trans_restrict},
        "forms": {"plain": {"particles_per_s": round(plain_pps, 3), "particles": args.plain_n} if plain_pps else None,
                  "batched": batched,
                  "batched+adjoint": {"particles_per_s": round(value, 3), "particles": args.n,
                                      "ccc_per_particle": round(float(n_ccc.mean()), 2),
                                      "ccc_evaluated_per_particle": round(adjoint_evals, 2),
                                      "refine_iterations": 20,
                                      "rot_err_deg_median": round(float(np.median(rot_err)), 3),
                                      "shift_err_vox_median": round(float(np.median(sh_err)), 3),
                                      "frac_rot_err_below_5deg": round(float(np.mean(rot_err < 5.0)), 3)},
                  "batched+adjoint_over_plain": round(value / plain_pps, 2) if plain_pps else None,
                  "batched_over_plain": round(batched_pps / plain_pps, 2) if (plain_pps and batched_pps) else None,
                  "plain_batched_identical": f"{same}/{args.plain_n}" if (res_plain and res_b) else None},
        "kernels": kernels,  # one untimed step with every launch timed
        "host_stages": host,
        "roofline": roofline,
        "forms_roofline": forms_roofline,
        "accuracy_vs_fp64": acc,
        "cpu_baseline": cpu,
        "gen_s": round(gen_s, 1),
    }
    print(json.dumps(line))
    plan.close()
    ctx.close()
    return 0


def uneven(args) -> int:
    """Particles of uneven box sizes: --n particles with boxes 48/64/96/128 drawn with probabilities
    .3/.4/.2/.1 (bandwidths 16/32/32/64), one 40-blob reference per size, wedges +-U[45,70], SNR U[0.05,0.2];
    one step aligns all of them in one ettk_align_buckets call (batched + adjoint form, CCC chunks replayed from
    captured graphs, execution choices tuned per size on the first call)."""
    import torch

    from ettk_amd.abi import ETTK_BUCKETS_GRAPHS, ETTK_BUCKETS_TUNE, Particle, default_params
    from ettk_amd.lib import Context
    from ettk_amd.synth import _rng, blob_reference, make_particles_device

    sizes = [(48, 16, 0.3), (64, 32, 0.4), (96, 32, 0.2), (128, 64, 0.1)]
    g = _rng(args.seed, 7)
    draw = g.choice(len(sizes), size=args.n, p=[s[2] for s in sizes])
    counts = [int(np.sum(draw == k)) for k in range(len(sizes))]
    t0 = time.time()
    buckets = []
    for k, (N, B, _) in enumerate(sizes):
        ref = blob_reference(N, seed=args.seed + 10 * k)
        d, descs, _ = make_particles_device(ref, counts[k], seed=args.seed + 10 * k + 1, half_wedge=(45.0, 70.0),
                                            snr=(0.05, 0.2), max_shift=4.0 * N / 64)
        buckets.append((default_params(N, B), torch.from_numpy(ref).cuda(), Particle(), d, descs))
    gen_s = time.time() - t0
    ctx = Context(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.align_buckets(buckets, mode=2, flags=ETTK_BUCKETS_GRAPHS | ETTK_BUCKETS_TUNE)  # warm-up: plans, tuning
    torch.cuda.synchronize()
    first_s = time.perf_counter() - t0
    step_ms = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.align_buckets(buckets, mode=2, flags=ETTK_BUCKETS_GRAPHS)
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.mean(step_ms))
    line = {"metric": "aligned particles/s, uneven box sizes 48/64/96/128 (B 16/32/32/64), batched+adjoint",
            "value": round(args.n / (ms / 1e3), 2), "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "ms_per_step": round(ms, 2), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (per-size 40-blob references)",
            "config": {"workload": "uneven box sizes", "particles": args.n, "counts": dict(zip(["48", "64", "96", "128"],
                                                                                              counts))},
            "first_call_s": round(first_s, 2), "gen_s": round(gen_s, 1)}
    print(json.dumps(line))
    ctx.close()
    return 0


def _perturb(truth, deg, vox, seed):
    """Each pose turned by `deg` degrees about a random axis and shifted by U[-vox, vox] per axis."""
    g = np.random.default_rng(seed)
    out = []
    for T in truth:
        ax = g.standard_normal(3)
        ax /= np.linalg.norm(ax)
        a = np.radians(deg)
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        S = T.copy()
        S[:3, :3] = T[:3, :3] @ (np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K)
        S[:3, 3] = T[:3, 3] + g.uniform(-vox, vox, 3)
        out.append(S)
    return np.array(out)


def refine96(args) -> int:
    """The differentiable local refinement workload: --n particles of 96^3 (40-blob reference, wedges
    +-U[45,70], SNR U[0.05,0.2]) seeded 3 degrees (random axis) and up to 1 voxel per axis off their true pose, as
    a coarse B = 32 search leaves them; one step is 20 lock-step ascent iterations of the continuous constrained
    CCC with the hand-written adjoint for every particle (ettk_refine_batch)."""
    import torch

    from ettk_amd.abi import default_params
    from ettk_amd.lib import Context, Plan
    from ettk_amd.synth import blob_reference, make_particles_device, rotation_error_deg

    N, B = 96, 32
    n = args.n
    ref = blob_reference(N, seed=args.seed)
    t0 = time.time()
    d_parts, descs, truth = make_particles_device(ref, n, seed=args.seed + 1, half_wedge=(45.0, 70.0),
                                                  snr=(0.05, 0.2), max_shift=6.0)
    gen_s = time.time() - t0
    seeds = _perturb(truth, 3.0, 1.0, args.seed + 2)
    ctx = Context(0)
    plan = Plan(ctx, default_params(N, B))
    plan.set_reference(torch.from_numpy(ref).cuda())
    pidx = list(range(n))
    for _ in range(args.warmup):
        plan.refine(d_parts, descs, pidx, seeds, n_iter=20)
    step_ms = []
    poses = scores = None
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        poses, scores = plan.refine(d_parts, descs, pidx, seeds, n_iter=20)
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.mean(step_ms))
    s0, _ = plan.ccc_grad(d_parts, descs, pidx, seeds, np.zeros((n, 6)))
    e0 = np.array([rotation_error_deg(seeds[i], truth[i]) for i in range(n)])
    e1 = np.array([rotation_error_deg(poses[i], truth[i]) for i in range(n)])
    t1 = np.array([np.abs(poses[i][:3, 3] - truth[i][:3, 3]).max() for i in range(n)])
    R, H = N ** 3 * 4.0, N * N * (N // 2 + 1) * 8.0
    bpp = 20 * (7 * R + 5 * H)
    ach = bpp * n / (ms / 1e3) / 1e9
    line = {"metric": "refined particles/s, 96^3, 20 adjoint ascent iterations", "value": round(n / (ms / 1e3), 2),
            "unit": "particles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (40-blob reference, seeds 3 deg / 1 voxel off the true pose)",
            "config": {"workload": "differentiable local refinement: 96^3, 20 iterations", "particles": n, "box": N},
            "seed_rot_err_deg_median": round(float(np.median(e0)), 3),
            "rot_err_deg_median": round(float(np.median(e1)), 3),
            "shift_err_vox_median": round(float(np.median(t1)), 3),
            "frac_score_raised": round(float(np.mean(scores > s0)), 4),
            "roofline_pass_model": {"bytes_per_particle": round(bpp), "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                                    "basis": "per iteration 7R + 5H"},
            "gen_s": round(gen_s, 1)}
    print(json.dumps(line))
    plan.close()
    ctx.close()
    return 0


def classify128(args) -> int:
    """The alignment + classification workload: --n particles of 128^3 drawn equally from K = 8 blob references
    (wedges +-U[40,70], SNR 0.05), aligned to every class reference at B = 64 (search, candidate CCCs, adjoint
    refinement), assigned to the best reference, averaged into 8 class averages with their accumulated wedges,
    over --iters iterations (ettk_mra_iteration); the first references are the generating ones.  Reports
    particle-alignments per second, class agreement with the generating labels and the FSC of each class
    average against its generating reference."""
    import torch

    from ettk_amd.abi import default_params
    from ettk_amd.classify import agreement, classify, fsc, fsc_resolution
    from ettk_amd.lib import Context, Plan
    from ettk_amd.synth import blob_reference, make_particles_device

    N, B, K = 128, 64, 8
    per = max(1, args.n // K)
    n = per * K
    refs = [blob_reference(N, seed=args.seed + 100 * k) for k in range(K)]
    t0 = time.time()
    parts, descs, labels = [], [], []
    for k in range(K):
        d, ds, _ = make_particles_device(refs[k], per, seed=args.seed + 100 * k + 1, half_wedge=(40.0, 70.0),
                                         snr=(0.05, 0.05), max_shift=8.0)
        parts.append(d)
        descs += ds
        labels += [k] * per
    d_parts = torch.cat(parts).contiguous()
    del parts
    order = np.random.default_rng(args.seed).permutation(n)
    d_parts = d_parts[torch.from_numpy(order).cuda()].contiguous()
    descs = [descs[i] for i in order]
    labels = np.array(labels)[order]
    gen_s = time.time() - t0
    ctx = Context(0)
    plan = Plan(ctx, default_params(N, B))
    d_refs = torch.from_numpy(np.stack(refs)).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = classify(plan, d_refs, d_parts, descs, iterations=args.iters, type=0, form=2)
    torch.cuda.synchronize()
    s = time.perf_counter() - t0
    avgs = d_refs.cpu().numpy()
    agree = [round(agreement(it.classes, labels, K), 4) for it in out.iterations]
    res = [1 + fsc_resolution(fsc(avgs[k], refs[k])[1:]) for k in range(K)]  # shells 1 .. N/2 - 1 (no DC term)
    line = {"metric": "particle-reference alignments/s, 128^3 B=64, classification into 8 classes",
            "value": round(n * K * args.iters / s, 2), "unit": "alignments/s", "n_gpus": 1, "steps": args.iters,
            "warmup": 0, "ms_per_step": round(s / args.iters * 1e3, 2), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (8 blob references)",
            "config": {"workload": "alignment + classification: 128^3, B=64, K=8", "particles": n, "box": N,
                       "bandwidth": B, "classes": K, "iterations": args.iters, "form": "batched+adjoint"},
            "particles_per_s_per_iteration": round(n * args.iters / s, 3),
            "class_agreement_per_iteration": agree,
            "fsc_0.5_shell_per_class": res, "fsc_shells": N // 2,
            "note": "the first iteration includes plan creation", "gen_s": round(gen_s, 1)}
    print(json.dumps(line))
    plan.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
