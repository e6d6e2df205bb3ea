"""Benchmark: voxel-iterations/s of one VPFFT strain step on one MI355X.

Workload (the headline configuration): one strain step of a 256^3
periodic-Voronoi polycrystal, 4096 grains with Haar-random orientations, FCC,
n=20, tau0=1, gdot0=1, a seeded random symmetric traceless strain rate of
Frobenius norm 1, tolerance 1e-6, at most 100 fixed-point iterations,
dt=0.005.  Every timed step is that same step: the state it starts from is
saved on the device once and restored (outside the timed region) before each
repetition, so the whole step -- prologue, fixed-point loop, epilogue --
runs every time on inputs resident in HBM.  value = N * iterations / wall
time summed over the timed steps (plain FP64 form).  The mixed form (FP32 FFT
+ Green, FP64 Newton) runs the same step and reports its speedup and its
macroscopic-stress error against the FP64 form.  --consecutive runs the
reference's RunVPFFT loop instead (each step from the previous one's state).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256] ...
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

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic work per unit (the per-unit figures of the measurement plan):
# the reference's constrained Newton solve (Solvers.cpp:111-222) is taken as
# 1.5 kflop per Newton step (12 resolved shears, the power law, the 5x5
# Jacobian sum and J' = -I - L0 J, the residual, the 5x5 LU solve) plus
# 0.6 kflop per voxel-iteration outside the Newton steps (the rescale, the
# constitutive law, phi and lambda, the norms); a 1-D transform of a real
# field costs 2.5 log2(n) flop per point, so the forward and inverse
# transforms of the 5 components along one axis cost 25 log2(n) per voxel.
# The Newton steps are counted by the kernel's own histogram.  The flops the
# kernel actually issues (its crystal-frame form needs fewer per step and
# more outside them) are read from the FP64 counter profile below.
NR_ALG_FLOPS_FIXED = 600.0      # per voxel-iteration, outside the Newton steps
NR_ALG_FLOPS_PER_STEP = 1500.0  # per Newton step


def fft_axis_flops(n):
    """Algorithmic flops per voxel of the forward + inverse transforms of the
    five components along an axis of n points."""
    return 25.0 * float(np.log2(n))


# d~, sigma, lambda (40 B each), R (72 B) and one CRSS (8 B) read; sigma,
# lambda, phi written (40 B each): the uniform-CRSS variant the workload runs
# (every voxel's CRSS equal and kept equal by the hardening law)
NR_BYTES = 320.0
# the fused inverse-z / Newton / forward-z kernel (Nz = 256 or 512): the Newton
# kernel's bytes without d~ and phi, plus the half spectrum read and written
# (5 components x (Nz/2 + 1) complex bins per z row, each way)
# measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a
# plain copy of known size) for the default workload
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "strain_step_256cubed_hbm_traffic.md")
# FP64 flops the Newton kernel issues per launch (rocprofv3 SQ_INSTS_VALU_*_F64)
# for the default workload
FP64_PROFILE = os.path.join(ROOT, "profiles", "nr_update_256cubed_fp64_counters.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fft_bytes(kernel, n, nc, elt):
    """Algorithmic HBM bytes of one FFT pass launch (elt = bytes per real)."""
    if kernel in ("fft_z_r2c", "fft_z_c2r"):
        return 5 * n * elt + 5 * nc * 2 * elt
    return 2 * 5 * nc * 2 * elt  # y / x passes: read + write the half spectrum


def profile_entry(args, path, kernel):
    """`kernel`'s entry of the JSON block of a committed counter profile, when
    the run is the workload that profile measured (else None)."""
    if (args.size, args.grains, args.nexp, args.seed) != (256, 4096, 20, 0) or not os.path.exists(path):
        return None
    import re
    m = re.search(r"```json\n(.*?)\n```", open(path).read(), re.S)
    return json.loads(m.group(1)).get(kernel) if m else None


def measured_traffic(args, kernel):
    """HBM bytes per launch of `kernel` from the committed counter profile."""
    e = profile_entry(args, TRAFFIC_PROFILE, kernel)
    return None if e is None else e["read_bytes"] + e["write_bytes"]


def run_form(V, R, mat, edot, args, precision, steps, warmup, profile, field=False):
    """Warm up, then time `steps` strain steps.  The timed steps replay the
    fixed-point loop as captured HIP graphs.  Unless --consecutive, every step
    starts from the same saved initial state (restored on the device between
    steps, untimed).  With profile=True, min(steps, --profile-steps) steps run
    once more with every kernel launched from the host between HIP events on
    the grid's stream, which gives the per-kernel durations."""
    n = R.shape[0]
    g = V.MaterialGrid(args.size, args.size, args.size, mat, precision=precision)
    g.SetVoxelLength(1.0, 1.0, 1.0)
    g.SendDataAndInitialize(R)
    g.SaveState()

    def one(prof_steps=None, timed=True):
        k = steps if prof_steps is None else prof_steps
        iters, avgs, infos, wall = 0, [], [], 0.0
        g.RestoreState()
        for _ in range(k):
            if not args.consecutive:
                g.RestoreState()
            t0 = time.perf_counter()
            a = g.RunSingleStrainStep(edot, args.max_iter, args.dt, args.tol)
            wall += time.perf_counter() - t0   # RunSingleStrainStep returns after the device is done
            avgs.append(a)
            infos.append(g.last_info.as_dict())
            iters += g.last_info.iterations
        return dict(wall=wall, iters=iters, vox_iters=n * iters, avgs=avgs, infos=infos)

    g.SetProfiling(False)
    for _ in range(warmup):
        g.RestoreState()
        g.RunSingleStrainStep(edot, args.max_iter, args.dt, args.tol)
    out = one()
    if profile:
        g.SetProfiling(True)
        p = one(prof_steps=min(steps, args.profile_steps))
        out["ktimes"] = g.KernelTimes()
        out["profiled_wall"] = p["wall"] / len(p["avgs"])
        out["profiled_same_bits"] = all(np.array_equal(a, b) for a, b in zip(out["avgs"], p["avgs"]))
        g.SetProfiling(False)
    if field:
        # the stress field after the first step (untimed), for the per-voxel error
        g.RestoreState()
        g.RunSingleStrainStep(edot, args.max_iter, args.dt, args.tol)
        out["sigma1"] = g.Stress()
    g.close()
    return out


def field_stats(a, b):
    """Per voxel, the largest stress-component difference relative to the
    max magnitude of b: maximum, median, 99.99th percentile and the number of
    voxels above 1e-3 (at n=20 a few voxels of the FP64 form move by ~1e-3
    under a 1e-12 relative change of the strain rate)."""
    dv = np.abs(a - b).max(axis=1) / np.abs(b).max()
    return {"stress_field_rel_err_step1": float(dv.max()),
            "stress_field_rel_err_step1_median": float(np.median(dv)),
            "stress_field_rel_err_step1_p9999": float(np.quantile(dv, 0.9999)),
            "stress_field_voxels_above_1e-3": int((dv > 1e-3).sum())}


def cpu_baseline(V, args, edot):
    """The oracle's fixed-point loop on a bounded sample of the workload."""
    import ctypes as C

    from oracle import oracle as O
    s = args.cpu_size
    grains = max(8, int(round(args.grains * (s / args.size) ** 3)))
    R, _, _ = V.generate_voronoi(s, s, s, grains, args.seed)
    n = R.shape[0]
    mat = V.Material.uniform(1.0, args.nexp)
    lib = O.load()
    m = O.make_material(s, s, s, mat)
    mall = O.schmid_all(R)
    crss = np.ones((n, 12))
    sig, lam = np.zeros((n, 5)), np.zeros((n, 5))
    phi, dtl = np.zeros((5, n)), np.zeros((5, n))
    p = O._p
    lib.vo_initialize(C.byref(m), p(mall), p(crss), p(edot), p(sig), p(lam), p(phi))
    L0 = np.zeros(25)
    lib.vo_build_L0(C.byref(m), p(mall), p(crss), p(sig), p(L0))
    its = 0
    t0 = time.perf_counter()
    while its < args.cpu_max_iters and (time.perf_counter() - t0) < args.cpu_seconds:
        lib.vo_strain_rate_evolution(C.byref(m), p(L0), p(phi), p(dtl))
        S, mx = np.zeros(12), np.zeros(1)
        lib.vo_voxel_update(C.byref(m), p(mall), p(crss), p(edot), p(L0), p(dtl), p(sig), p(lam), p(phi), p(S), p(mx))
        its += 1
    wall = time.perf_counter() - t0
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": n * its / wall, "unit": "voxel-iters/s", "cores": O.num_threads(), "kind": "port",
        "cpu_model": model, "nproc": os.cpu_count(),
        "sample": f"oracle fixed-point loop (FFT + literal Green + per-voxel NR) on a {s}^3 grid, "
                  f"{grains} grains, n={args.nexp}, same strain rate; {its} iterations in {wall:.1f} s; "
                  "prologue excluded",
    }


def batched_workload(V, args):
    """Uneven RVEs advancing together: edges drawn from {16,24,32,48,64,96},
    grains log-uniform in 8-512, n from {5,10,20}, each its own random unit
    traceless strain rate; FP32 FFT + FP64 Newton; FFT block shapes and
    radix orders picked by the batch's autotuner;
    several strain steps of dt 0.005.  Timed once as one batch and once
    grid by grid through the single-grid entry point."""
    rng = np.random.default_rng(args.seed + 4)
    cases = []
    for i in range(args.batch_rves):
        e = int(rng.choice([16, 24, 32, 48, 64, 96]))
        grains = int(round(float(np.exp(rng.uniform(np.log(8.0), np.log(512.0))))))
        nexp = int(rng.choice([5, 10, 20]))
        ed = rng.standard_normal(5)
        cases.append((e, grains, nexp, ed / np.linalg.norm(ed), 1000 + i))
    micro = [V.generate_voronoi(e, e, e, gr, sd)[0] for (e, gr, _, _, sd) in cases]

    def make():
        gs = []
        for (e, gr, nexp, _, _), R in zip(cases, micro):
            g = V.MaterialGrid(e, e, e, V.Material.uniform(1.0, nexp), precision=V.MIXED)
            g.SetVoxelLength(1.0, 1.0, 1.0)
            g.SendDataAndInitialize(R)
            gs.append(g)
        return gs

    edots = np.array([c[3] for c in cases])
    voxels = sum(c[0] ** 3 for c in cases)
    with V.Batch(tune="radix") as b:
        gs = make()  # warm-up pass: tunes the schedules, loads the kernels
        b.RunSingleStrainStep(gs, edots, args.max_iter, args.dt, args.batch_tol)
        for g in gs:
            g.close()
        gs = make()
        iters = 0
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            b.RunSingleStrainStep(gs, edots, args.max_iter, args.dt, args.batch_tol)
            iters += sum(inf.iterations * c[0] ** 3 for inf, c in zip(b.last_infos, cases))
        wall = time.perf_counter() - t0
        shapes = sorted({(c[0], tuple(b.block_shape(g, p) for p in range(3))) for g, c in zip(gs, cases)})
        radix = sorted({(c[0], tuple(".".join(map(str, b.radix_plan(g, p))) for p in range(3)))
                        for g, c in zip(gs, cases)})
        for g in gs:
            g.close()
    rve_steps = args.batch_rves * args.batch_steps
    wall_seq = None
    if not getattr(args, "no_one_by_one", False):
        gs = make()
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            for g, ed in zip(gs, edots):
                g.RunSingleStrainStep(ed, args.max_iter, args.dt, args.batch_tol)
        wall_seq = time.perf_counter() - t0
        for g in gs:
            g.close()
    return {
        "rves": args.batch_rves, "voxels": voxels, "steps": args.batch_steps, "tol": args.batch_tol,
        "precision": "fp32 FFT + fp64 Newton",
        "rve_steps_per_s": rve_steps / wall, "voxel_iters_per_s": iters / wall, "wall_s": wall,
        "one_by_one_rve_steps_per_s": rve_steps / wall_seq if wall_seq else None,
        "speedup_vs_one_by_one": wall_seq / wall if wall_seq else None,
        "tuned_block_shapes_zyx": {str(e): list(sh) for e, sh in shapes},
        "tuned_radix_plans_zyx": {str(e): list(r) for e, r in radix},
    }


def serving_workload(V, args):
    """Poisson request streams at a few rates: each request one RVE (edge
    32-128 in steps of 16, 8-256 grains, n in {10, 20}) and 1-10 increments;
    admitted by an HBM budget, resident between increments, mixed precision
    with FP64 reruns of steps that do not converge."""
    from vpfft_amd import serve as S
    out = []
    for rate in args.serve_rates:
        reqs = S.poisson_requests(args.serve_requests, rate, seed=args.seed + 11)
        res, st, _ = S.serve(reqs, hbm_budget_bytes=int(args.serve_budget_gib * (1 << 30)), max_active=64,
                             precision=V.MIXED, fp64_fallback=True, max_iter=args.max_iter, tol=args.batch_tol)
        bad = sum(1 for r in res if r["status"] != 0)
        failures = [{"request": i, "status": V.ERRORS.get(r["status"], r["status"]), "edge": q["edge"],
                     "grains": q["grains"], "n": q["n"], "increments": len(q["edots"]),
                     "increments_done": r["increments_done"]}
                    for i, (q, r) in enumerate(zip(reqs, res)) if r["status"] != 0]
        out.append({"rate_per_s": rate, "requests": len(reqs), "failed": bad, "failures": failures,
                    "rve_steps_per_s": st["rve_steps_per_s"], "voxel_iters_per_s": st["voxel_iters_per_s"],
                    "p50_latency_s": st["p50_latency_s"], "p99_latency_s": st["p99_latency_s"],
                    "mean_latency_s": st["mean_latency_s"], "wall_s": st["wall_s"],
                    "max_active": st["max_active_seen"], "max_gib": st["max_bytes_seen"] / (1 << 30),
                    "batched_steps": st["batched_steps"], "fp64_reruns": st["fp64_reruns"]})
        log(f"[bench] serving at {rate}/s: p50 {st['p50_latency_s']:.3f} s, p99 {st['p99_latency_s']:.3f} s, "
            f"{st['rve_steps_per_s']:.1f} RVE-steps/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--grains", type=int, default=4096)
    ap.add_argument("--nexp", type=int, default=20)
    ap.add_argument("--max-iter", type=int, default=100)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--dt", type=float, default=0.005)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--consecutive", action="store_true",
                    help="time consecutive strain steps (RunVPFFT) instead of repetitions of the first step")
    ap.add_argument("--profile-steps", type=int, default=3,
                    help="steps of the host-launched pass that gives the per-kernel HIP-event times")
    ap.add_argument("--sensitivity", type=float, default=0.0,
                    help="also run the FP64 form with the strain rate moved by this relative amount and report "
                         "its per-voxel stress change beside the mixed form's error")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-only", action="store_true",
                    help="time only the oracle's fixed-point loop (the cpu_baseline leg) and print its JSON; "
                         "no GPU is used")
    ap.add_argument("--cpu-size", type=int, default=96)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-max-iters", type=int, default=20)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched workload and print its JSON (for a kernel trace of it)")
    ap.add_argument("--no-one-by-one", action="store_true",
                    help="batched workload without the grid-by-grid comparison run")
    ap.add_argument("--batch-rves", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-tol", type=float, default=1e-5)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[4.0, 16.0])
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only copy this many MiB device to device (torch) three times and exit: a kernel of "
                         "known traffic that calibrates FETCH_SIZE / WRITE_SIZE counters read under rocprofv3")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU: the whole voxel grid lives in one MI355X's HBM and there is no "
              "multi-GPU decomposition in this library (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mib > 0:
        import torch
        a = torch.ones(args.copy_probe_mib << 17, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 3 x {args.copy_probe_mib} MiB read + written")
        return

    import vpfft_amd as V

    if args.cpu_only:
        rng = np.random.default_rng(args.seed)
        edot = rng.standard_normal(5)
        print(json.dumps(cpu_baseline(V, args, edot / np.linalg.norm(edot))), flush=True)
        return

    V.load_library()

    if args.batched_only:
        log(f"[bench] batched: {args.batch_rves} uneven RVEs")
        print(json.dumps(batched_workload(V, args)), flush=True)
        return

    log(f"[bench] generating {args.size}^3 microstructure, {args.grains} grains")
    R, _, _ = V.generate_voronoi(args.size, args.size, args.size, args.grains, args.seed)
    rng = np.random.default_rng(args.seed)
    edot = rng.standard_normal(5)
    edot /= np.linalg.norm(edot)
    mat = V.Material.uniform(1.0, args.nexp)
    n = R.shape[0]

    log("[bench] FP64 form")
    f64 = run_form(V, R, mat, edot, args, V.FP64, args.steps, args.warmup, profile=True, field=not args.no_mixed)
    log(f"[bench] FP64: {f64['iters']} iterations in {f64['wall']:.3f} s")
    value = f64["vox_iters"] / f64["wall"]

    # roofline of the dominant kernel, from the HIP-event times of the same steps
    kt = f64["ktimes"]
    loop_kernels = ("nr_update", "z_nr_fused", "fft_z_r2c", "fft_y_fwd", "fft_x_green", "fft_y_inv", "fft_z_c2r")
    dom = max(loop_kernels, key=lambda k: kt.get(k, (0.0, 0))[0])
    nc = args.size * args.size * (args.size // 2 + 1)
    nr_steps_total = sum(i["nr_steps_total"] for i in f64["infos"])
    per_kernel = {}
    for k, (ms, cnt) in kt.items():
        if cnt == 0:
            continue
        entry = {"ms_total": ms, "launches": cnt, "ms_avg": ms / cnt}
        if k.startswith("fft"):
            b = fft_bytes(k, n, nc, 8)
            entry["GBps"] = b / (ms / cnt) / 1e6
        if k in ("nr_update", "z_nr_fused"):
            # per launch (one fixed-point iteration over the grid): the fixed
            # work of every voxel, the Newton steps of an average iteration
            # and, in the fused kernel, the z transforms
            flops = n * NR_ALG_FLOPS_FIXED + nr_steps_total / f64["iters"] * NR_ALG_FLOPS_PER_STEP
            if k == "z_nr_fused":
                flops += n * fft_axis_flops(args.size)
            entry["flops_per_launch"] = flops
            entry["TFLOPs"] = flops / (ms / cnt / 1e3) / 1e12
            ex = profile_entry(args, FP64_PROFILE, k)
            if ex is not None:
                entry["issued_fp64_flops_per_launch"] = ex["fp64_flops_issued"]
                entry["issued_TFLOPs"] = ex["fp64_flops_issued"] / (ms / cnt / 1e3) / 1e12
            b = NR_BYTES * n if k == "nr_update" else (NR_BYTES - 80.0) * n + 2 * 5 * nc * 16
            entry["GBps"] = b / (ms / cnt) / 1e6
        per_kernel[k] = entry
    if dom in ("nr_update", "z_nr_fused"):
        ach = per_kernel[dom]["TFLOPs"]
        roof = {"bound": "fp64_valu", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / FP64_PEAK_TFLOPS, "traffic": measured_traffic(args, dom), "kernel": dom,
                "note": "algorithmic flops: the reference's Newton solve at 1.5 kflop per Newton step (the "
                        "kernel's own step histogram) + 0.6 kflop per voxel-iteration, + the z transforms; "
                        "FP64 vector ALU work (no MFMA): MI355X's FP64 vector and matrix peaks are both "
                        "78.6 TFLOP/s"}
        if "issued_TFLOPs" in per_kernel[dom]:
            roof["issued_fp64_tflops"] = per_kernel[dom]["issued_TFLOPs"]
            roof["issued_frac"] = per_kernel[dom]["issued_TFLOPs"] / FP64_PEAK_TFLOPS
    else:
        ach = per_kernel[dom]["GBps"]
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic(args, dom), "kernel": dom}

    mixed = None
    if not args.no_mixed:
        log("[bench] mixed form")
        mx = run_form(V, R, mat, edot, args, V.MIXED, args.steps, args.warmup, profile=False, field=True)
        v_mx = mx["vox_iters"] / mx["wall"]
        err = float(np.linalg.norm(mx["avgs"][0] - f64["avgs"][0]) / np.linalg.norm(f64["avgs"][0]))
        st = field_stats(mx["sigma1"], f64["sigma1"])
        del mx["sigma1"]
        mixed = {"value": v_mx, "speedup_vs_fp64": v_mx / value, "iterations": mx["iters"],
                 "macro_stress_rel_err_step1": err, **st, "ms_per_step": 1e3 * mx["wall"] / args.steps}
        log(f"[bench] mixed: {mx['iters']} iterations in {mx['wall']:.3f} s, <sigma> err {err:.2e}, "
            f"per-voxel stress err {st['stress_field_rel_err_step1']:.2e}")
        if args.sensitivity > 0:
            # the FP64 form against itself with the strain rate moved by a
            # relative `sensitivity` (a seeded unit direction): the scale of
            # per-voxel change the algorithm itself shows for such a change
            u = np.random.default_rng(args.seed + 7).standard_normal(5)
            ed2 = edot + args.sensitivity * u / np.linalg.norm(u)
            log(f"[bench] FP64 form with the strain rate moved by {args.sensitivity:g}")
            f2 = run_form(V, R, mat, ed2, args, V.FP64, 1, 0, profile=False, field=True)
            mixed["fp64_sensitivity"] = {
                "strain_rate_rel_change": args.sensitivity,
                "macro_stress_rel_change_step1": float(np.linalg.norm(f2["avgs"][0] - f64["avgs"][0]) /
                                                       np.linalg.norm(f64["avgs"][0])),
                **field_stats(f2["sigma1"], f64["sigma1"])}
        del f64["sigma1"]

    batched = None
    if not args.no_batched:
        log(f"[bench] batched: {args.batch_rves} uneven RVEs")
        batched = batched_workload(V, args)
        log(f"[bench] batched: {batched['rve_steps_per_s']:.1f} RVE-steps/s, "
            f"{batched['speedup_vs_one_by_one'] or 0:.2f}x the one-by-one rate")

    serving = None
    if not args.no_serving:
        log("[bench] serving: Poisson request streams")
        serving = serving_workload(V, args)

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(V, args, edot)

    loop_ms = sum(i["ms_loop"] for i in f64["infos"])
    out = {
        "metric": "voxel-iters/s at 256^3 (plain FP64 vs mixed @1e-5 stress err); RVE p99 latency",
        "value": value,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * f64["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"one VPFFT strain step per step: {args.size}^3 periodic Voronoi, {args.grains} grains, "
                        f"FCC n={args.nexp}, random unit traceless strain rate, tol {args.tol}, "
                        f"max {args.max_iter} iterations, dt {args.dt}; "
                        + ("consecutive steps (RunVPFFT)" if args.consecutive else
                           "every step from the same initial state (restored on the device, untimed)"),
            "consecutive": bool(args.consecutive),
            "grid": args.size, "grains": args.grains, "rate_sensitivity": args.nexp, "tol": args.tol,
            "max_iter": args.max_iter, "iterations": f64["iters"], "precision": "fp64",
        },
        "loop_only_voxel_iters_per_s": f64["vox_iters"] / (loop_ms / 1e3),
        "nr_steps_per_voxel_iter": nr_steps_total / f64["vox_iters"],
        "nr_steps_max": max(i["max_nr_steps"] for i in f64["infos"]),
        "nr_lane_efficiency": nr_steps_total / sum(i["nr_wave_steps_total"] for i in f64["infos"]),
        "taylor_steps_per_voxel": sum(i["taylor_steps_total"] for i in f64["infos"]) / (n * args.steps),
        "ms_prologue": sum(i["ms_prologue"] for i in f64["infos"]) / args.steps,
        "ms_per_step_host_launched": 1e3 * f64["profiled_wall"],
        "kernel_pass_same_bits": f64["profiled_same_bits"],
        "roofline": roof,
        "kernels": per_kernel,
        "mixed": mixed,
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
