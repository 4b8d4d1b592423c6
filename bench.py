"""GSE reciprocal-space benchmark on one MI355X.

One step = bin -> spread sigma1 (fixed point, bit-reproducible) -> brick
reduce -> r2c -> G(k) k-pass (FP64 energy) -> c2r -> gather (reference force
+ per-particle potential), on device-resident synthetic input, FP32 grids and
kernels with FP64 Green's table and reductions.  Workload: 16,000 rigid
water-like triplets (48,000 charges) in L = 78.3 A, M = 80, sigma = 2.12.

Prints ONE JSON line (rank 0, N = 1).  Extra figures (FP64 mode, adjoint
cost, the separable convolution at M = 80 and M = 216, the mode with the
on-grid convolution, batched systems/s, the 1,000,002-charge M = 216 system)
go into "extras".
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GSE energy+force steps/s (48k atoms, M=80) + HBM roofline frac; batched systems/s"
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch measured by two rocprofv3 --pmc passes (FETCH_SIZE,
# WRITE_SIZE), scaled by the same counters read for a 256 MiB device copy
# (FETCH x2.000, WRITE x1.000); the commands and per-kernel table are in
# profiles/k_gather_k_spread_water48k_M80_fp32_hbm_traffic.md.
MEASURED_TRAFFIC = {
    ("water_48k_M80", "fp32", "gather"): 10650317,
    ("water_48k_M80", "fp32", "spread"): 43789312,
}
# Wave-level VALU instructions per launch (rocprofv3 --pmc SQ_INSTS_VALU, same
# file).  The gather and the spread issue far more vector instructions than
# their bytes need; their bound is VALU issue: 256 CUs x 2.4 GHz x one wave64
# VALU instruction per CU per clock (4 SIMDs, 16 lanes each) = 614.4 G/s.
MEASURED_VALU = {
    ("water_48k_M80", "fp32", "gather"): 15666063,
    ("water_48k_M80", "fp32", "spread"): 8510291,
}
VALU_PEAK_GINST = 614.4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--workload", default="water_48k_M80")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget for the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--conv-only", action="store_true",
                    help="only the separable convolution timing (for counter passes)")
    ap.add_argument("--rows-only", action="store_true",
                    help="only the comparator timings (SPME, Ewald, GSE-r, GSE-u)")
    return ap.parse_args()


def stage_bytes(N, M, ct_bytes):
    """Algorithmic HBM bytes per launch of each stage (one system)."""
    nk = (M // 2 + 1) * M * M
    return {
        "spread": 4 * ct_bytes * N + 8 * M ** 3,          # xyzq in, int64 grid written once
        "convert": (8 + ct_bytes) * M ** 3,              # int64 in, grid out (zeroing is extra)
        "kpass": 2 * (2 * ct_bytes) * nk + ct_bytes * nk,  # spectrum in/out, G table in
        "gather": ct_bytes * M ** 3 + 4 * ct_bytes * N + 4 * ct_bytes * N,  # grid, xyzq, F+phi_i
    }


def note(msg):
    """Progress line on stderr (stdout carries only the result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_conv(api, torch, dev):
    """The separable three-pass sigma2 convolution of mode R (spreadChargeOnGrid,
    11 taps per axis at h ~ 1 A) on its own, FP32 with compensated sums, at
    M = 80 and M = 216: HBM roofline with 3 passes x (4 B read + 4 B write)
    per grid point as the algorithmic bytes."""
    out = {}
    for M, L in ((80, 78.3), (216, 215.3)):
        h = L / M
        s2 = 2.12 * math.sqrt(0.2)
        R = int(math.ceil(3.0 * math.sqrt(2.0) * s2 / h))
        k = np.exp(-(h * np.arange(-R, R + 1)) ** 2 / (2 * s2 * s2))
        conv = api.CpuConvolution(torch.float32, M, M, M)
        conv.setupKernel(R, R, R, k, k, k)
        a = torch.rand((M, M, M), device=dev, dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            conv.runXYZ(a, b)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 50
        e0.record()
        for _ in range(reps):
            conv.runXYZ(a, b)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nbytes = 3 * 8 * M ** 3
        out[f"M{M}"] = {"taps": 2 * R + 1, "ms": round(ms, 5),
                        "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                        "frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    return out


def bench_large(api, systems, torch, dev):
    """The 1,000,002-charge water box at M = 216 (FP32, same step as the
    headline): forward steps/s and the adjoint sweep's cost relative to it
    (the adjoint recomputes bin, spread and the Poisson solve)."""
    xyzq_np, L, M, sigma = systems.make("water_1M_M216")
    x = torch.from_numpy(xyzq_np).to(device=dev, dtype=torch.float32).contiguous()
    gs = api.CpuGSEk(torch.float32, torch.float32, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
    out = {}
    sf = time_steps(lambda: gs.step(x, mode=0, force=True, potential=True, out=out), 20, 3, torch)
    sa = time_steps(lambda: gs.adjoint(x), 20, 3, torch)
    return {"workload": "water_1M_M216", "n_charges": len(xyzq_np), "grid": M,
            "steps_per_s": round(1.0 / sf, 2), "ms_per_step": round(sf * 1e3, 4),
            "adjoint_ms": round(sa * 1e3, 4), "adjoint_over_forward": round(sa / sf, 4)}


def bench_batched(api, systems, torch, dev, nsys=256):
    """256 independent water boxes (600..8000 charges, M in {24, 32, 40}),
    FP32: the batched step (autotuned, replayed from a HIP graph) against
    looping the plain single-system step over the same systems."""
    bs = systems.batched_systems(nsys, seed=4)
    xs = [torch.from_numpy(x).to(device=dev, dtype=torch.float32).contiguous() for x, _, _ in bs]
    xc = torch.cat(xs).contiguous()
    batch = api.GSEBatch(torch.float32, torch.float32,
                         [((M, M, M), (L, L, L), len(x)) for x, L, M in bs], 2.12)
    cfg, tuned_ms = batch.autotune(xc, reps=3)
    ob = {}
    t_graph = time_steps(lambda: batch.step(xc, force=True, graph=True, out=ob), 50, 5, torch)
    t_eager = time_steps(lambda: batch.step(xc, force=True, out=ob), 50, 5, torch)
    handles = [api.CpuGSEk(torch.float32, torch.float32, M, M, M, 2.12, 0.4, 3.0, 3.0, L, L, L)
               for _, L, M in bs]
    outs = [{} for _ in bs]

    def loop():
        for h, x, o in zip(handles, xs, outs):
            h.step(x, force=True, potential=False, out=o)

    t_loop = time_steps(loop, 10, 2, torch)
    # the batched forces against the looped single-system forces, every system
    f_loop = torch.cat([o["force"] for o in outs], dim=1).double().cpu().numpy()
    f_batch = ob["force"].double().cpu().numpy()
    e_loop = torch.cat([o["energy"] for o in outs]).cpu().numpy()
    e_batch = ob["energy"].cpu().numpy()
    erms = float(np.sqrt(np.sum((f_batch - f_loop) ** 2) / np.sum(f_loop ** 2)))
    de = float(np.max(np.abs(e_batch - e_loop) / np.abs(e_loop)))
    return {"systems": nsys, "charges": int(xc.shape[0]),
            "force_erms_batched_vs_looped": erms, "max_energy_rel_batched_vs_looped": de,
            "systems_per_s_graph": round(nsys / t_graph, 1),
            "systems_per_s_eager": round(nsys / t_eager, 1),
            "systems_per_s_looped": round(nsys / t_loop, 1),
            "speedup_graph_vs_looped": round(t_loop / t_graph, 3),
            "ms_per_batched_step": round(t_graph * 1e3, 4),
            "autotuned": cfg}


def bench_rows(api, systems, torch, dev):
    """The comparators and pipelines beside the GSE-k step, on the same 48k
    water box (M = 80, sigma = 2.12): SPME order 4 (CudaPMERecip, kappa =
    1/(sqrt(2) sigma)), the Ewald direct-space sum (rcut 9 A, FP64 cell
    list), GSE-r (exact HERMITIAN solve) and GSE-u (N = 4, b = 1.168), plus
    the exact Ewald reciprocal sum on the 3k box (kcut 10).  Steps per second
    of each whole call, inputs resident in HBM."""
    xyzq_np, L, M, sigma = systems.make("water_48k_M80")
    n = len(xyzq_np)
    out = {"workload": "water_48k_M80"}
    kappa = 1.0 / (math.sqrt(2.0) * sigma)
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        x = torch.from_numpy(xyzq_np).to(device=dev, dtype=dt).contiguous()
        pm = api.CudaPMERecip(dt, dt, M, M, M, 4)
        o = {}
        s = time_steps(lambda: pm.step(x, (L, L, L), kappa, out=o), 50, 5, torch)
        out[f"spme_order4_{name}_steps_per_s"] = round(1.0 / s, 2)
        for cls, key, extra in ((api.CpuGSEr, "gse_r", ()), (api.CpuGSEu, "gse_u_N4", (4, 1.168))):
            g = cls(dt, dt, *extra, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
            og = {}
            s = time_steps(lambda: g.step(x, force=True, out=og), 30, 3, torch)
            out[f"{key}_{name}_steps_per_s"] = round(1.0 / s, 2)
    x64 = torch.from_numpy(xyzq_np).to(device=dev, dtype=torch.float64).contiguous()
    d = api.CpuEwaldDirect(sigma, 9.0, L, L, L)
    f = torch.zeros((3, n), dtype=torch.float64, device=dev)
    s = time_steps(lambda: d.calcForceEnergy(n, x64, f[0], f[1], f[2]), 20, 3, torch)
    out["ewald_direct_rcut9_fp64_ms"] = round(s * 1e3, 4)
    x3, L3, _, s3 = systems.make("water_3k_M32")
    x3d = torch.from_numpy(x3).to(device=dev, dtype=torch.float64).contiguous()
    er = api.CpuEwaldRecip(s3, 10, L3, L3, L3)
    s = time_steps(lambda: er.calcForceEnergy(len(x3), x3d), 10, 2, torch)
    out["ewald_recip_3k_kcut10_fp64_ms"] = round(s * 1e3, 4)
    return out


def host_info():
    """Core count and CPU model of this host (for the baseline record)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"nproc": affinity, "logical_cpus": os.cpu_count(), "cpu_model": model}


def cpu_baseline(args, M, L, sigma, xyzq_np, f_gpu, e_gpu, N):
    """The oracle's GSE-k step (spread, FFT solve, energy, reference force
    and phi_i: the GPU step's work) timed on this host: one core, the
    reference's own threading, and OpenMP across the cores the process may
    use (OMP_NUM_THREADS when set).  Each leg runs up to 5 steps within
    --cpu-seconds and reports the median.  Accuracy of the GPU step is taken
    against the serial FP64 oracle."""
    from oracle import oracle
    p = oracle.Params(M, L, sigma)
    info = host_info()
    omp_n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["nproc"]

    def timed(threads):
        oracle.set_omp_threads(threads)
        ts = []
        t_end = time.perf_counter() + args.cpu_seconds
        while not ts or (time.perf_counter() < t_end and len(ts) < 5):
            t0 = time.perf_counter()
            oracle.step_omp(p, xyzq_np)
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), len(ts)

    t1, n1 = timed(1)
    tn, nn = timed(omp_n)
    ref = oracle.gse_step(p, xyzq_np, mode=0)
    cpu = {"value": round(1.0 / t1, 5), "unit": "steps/s", "cores": 1, "kind": "port",
           "sample": f"{n1} FP64 oracle GSE-k step(s) of the {N}-charge M={M} box on one "
                     f"host core, median {t1:.3f} s",
           "all_cores": {"value": round(1.0 / tn, 5), "unit": "steps/s", "cores": omp_n,
                         "sample": f"{nn} step(s), OpenMP over {omp_n} threads, "
                                   f"median {tn:.3f} s"},
           **info}
    return cpu, oracle.erms(ref.force, f_gpu), abs(e_gpu - ref.energy) / abs(ref.energy)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: it measures the single-GPU step of one "
              "system; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from gse_amd import api, systems

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.conv_only:
        print(json.dumps({"conv_roofline": bench_conv(api, torch, dev)}), flush=True)
        return
    if args.rows_only:
        print(json.dumps({"comparators": bench_rows(api, systems, torch, dev)}), flush=True)
        return
    ct = torch.float32 if args.precision == "fp32" else torch.float64
    xyzq_np, L, M, sigma = systems.make(args.workload)
    N = len(xyzq_np)
    x = torch.from_numpy(xyzq_np).to(device=dev, dtype=ct).contiguous()
    gs = api.CpuGSEk(ct, ct, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
    out = {}

    def step():
        gs.step(x, mode=0, force=True, potential=True, out=out)

    sec = time_steps(step, args.steps, args.warmup, torch)
    value = 1.0 / sec
    note(f"{args.steps} timed steps: {sec * 1e3:.4f} ms per step")

    # per-stage HIP events over a second pass of the same number of steps
    gs.setProfiling(True)
    for _ in range(args.steps):
        step()
    stages, nrec = gs.stageTimes()
    gs.setProfiling(False)
    ctb = 4 if ct == torch.float32 else 8
    sb = stage_bytes(N, M, ctb)
    ours = {k: v for k, v in stages.items() if k in sb}
    dom = max(ours, key=ours.get)
    achieved = sb[dom] / (ours[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": MEASURED_TRAFFIC.get((args.workload, args.precision, dom)),
            "bytes_per_launch": sb[dom], "avg_ms": round(ours[dom], 5)}
    nvalu = MEASURED_VALU.get((args.workload, args.precision, dom))
    if nvalu:
        g = nvalu / (ours[dom] * 1e-3) / 1e9
        roof["valu"] = {"insts_per_launch": nvalu, "achieved_Ginst_s": round(g, 1),
                        "peak_Ginst_s": VALU_PEAK_GINST, "frac": round(g / VALU_PEAK_GINST, 4)}

    # accuracy against the FP64 oracle and the CPU baseline (oracle timing)
    torch.cuda.synchronize()
    f_gpu = out["force"].double().cpu().numpy().T
    e_gpu = float(out["energy"].cpu()[0])
    cpu = None
    erms = None
    de = None
    if not args.no_cpu:
        cpu, erms, de = cpu_baseline(args, M, L, sigma, xyzq_np, f_gpu, e_gpu, N)
        note("CPU baseline done")

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if ct == torch.float32 else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "n_charges": N, "box": L, "grid": M,
                   "sigma": sigma, "mode": "K", "precision": args.precision +
                   (" spread/FFT/gather, fp64 G table and energy" if ct == torch.float32 else "")},
        "roofline": roof, "cpu_baseline": cpu,
        "stages_ms": {k: round(v, 5) for k, v in stages.items()},
        "force_erms_vs_fp64_oracle": erms, "energy_rel_err_vs_fp64_oracle": de,
    }

    if not args.no_extras:
        extras = {}
        # FP64 mode at the same workload
        x64 = torch.from_numpy(xyzq_np).to(device=dev, dtype=torch.float64).contiguous()
        g64 = api.CpuGSEk(torch.float64, torch.float64, M, M, M, sigma, 0.4, 3.0, 3.0, L, L, L)
        o64 = {}
        s64 = time_steps(lambda: g64.step(x64, force=True, potential=True, out=o64),
                         max(20, args.steps // 5), 5, torch)
        extras["fp64_steps_per_s"] = round(1.0 / s64, 3)
        # adjoint sweep (recomputes spread and k-pass; dE/dq, dE/dr, virial)
        sa = time_steps(lambda: gs.adjoint(x), max(20, args.steps // 5), 5, torch)
        sf = time_steps(step, max(20, args.steps // 5), 5, torch)
        extras["adjoint_ms"] = round(sa * 1e3, 5)
        extras["forward_ms"] = round(sf * 1e3, 5)
        extras["adjoint_over_forward"] = round(sa / sf, 4)
        # conv + gather roofline as the survey defines it (mode R includes the conv)
        extras["stage_roofline_frac"] = {
            k: round(sb[k] / (ours[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 5) for k in sb}
        extras["conv_roofline"] = bench_conv(api, torch, dev)
        note("FP64, adjoint and convolution extras done")
        om = {}
        sm = time_steps(lambda: gs.step(x, mode=1, force=True, potential=True, out=om),
                        max(20, args.steps // 5), 5, torch)
        extras["mode_R_steps_per_s"] = round(1.0 / sm, 3)
        extras["batched"] = bench_batched(api, systems, torch, dev)
        note("batched extra done")
        extras["large"] = bench_large(api, systems, torch, dev)
        note("large-system extra done")
        extras["comparators"] = bench_rows(api, systems, torch, dev)
        note("comparator extras done")
        line["extras"] = extras

    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
