#!/usr/bin/env python3
"""Benchmark of the BFM hot path on one MI355X.

Workload: bridge deck NN=1000, NM=30, u+w turbulence, Classic grid 1025 x 1025
over [-4, 4] Hz (the 1024^2 configuration, realised as the reference's odd
1025-point grid).  One step = one full Classic pass: nodal PSDs, projection
A(f), node contraction over the whole (f1,f2) plane with FRF weighting and
quadrature -> m3mf, m3mr (NM^3 each), plus the modal PSD path -> m2mf, m2mr,
m2o2mr.  An "eval" is one (f1,f2) point of the plane with its full NM^3 modal
bispectrum accumulated into m3 (1025^2 evals per step).

Prints ONE JSON line (see keys below).  --gpus N with N > 1 exits: this build
runs on a single GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "modal-bispectrum evals/s (NN=1000,NM=30) at skew err≤1e-4: mixed vs FP64 GPU"
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
# HBM bytes per k_gsym_bf16x3 launch on the deck workload (1025^2, NM=30):
# rocprofv3 --pmc FETCH_SIZE x 2 + WRITE_SIZE x 1 (KiB), the scales read off a
# 1 GiB device copy (bench.py --copy-calibration); commands and counts in
# profiles/gsym_bf16x3_traffic_deck_nn1000_nm30.md
GSYM_BF16X3_TRAFFIC_B = int((2 * 4184124.25 + 410062.5) * 1024)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def gsym_flops(prob, nm):
    """Algorithmic FP64 flops of one k_gsym_f64 launch: every distinct value
    G[x](y,u; z,v) (u, v on the non-negative half grid, (y,u) <= (z,v)) is a
    dot product of length NTC*NNL (2 flops per term)."""
    nh = prob.nfreqs
    n = nm * nh
    return 2.0 * prob.ntc * prob.nnl * nm * n * (n + 1) / 2.0


def host_description() -> dict:
    """CPU of the host the baselines run on: logical CPUs, the ones this process
    may use, model and socket count (from /proc/cpuinfo)."""
    model, sockets = None, set()
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k == "model name" and model is None:
                model = v.strip()
            elif k == "physical id":
                sockets.add(v.strip())
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable": usable, "model": model,
            "sockets": len(sockets) or None}


REF_THREADS = 8   # getFM_full_tnm_vect_cls_ opens num_threads(8) (functionsImpl.F90:1490)


def cpu_reference_baseline(budget_s: float = 25.0):
    """Reference's own getFM_full_tnm_vect_cls_ (oracle/_ref, amdflang -O2,
    OpenMP with its hard-coded 8 threads) on bounded samples of the deck:
    NNL = 6, 8, 10, ... nodes spread over the span, NM=30, NTC=2, 17-point grid,
    until the samples have taken budget_s.  Its loop count is NTC*NNL^3*NM^3 NF-long
    updates per grid row: the largest sample's update rate is extrapolated by that
    exact count to NNL=1000 and divided by its NF^2 points, the reference's evals/s
    at the benchmark size.  The NNL exponent fitted on the samples (log t vs log NNL)
    is reported as a check of the cube law, with the extrapolation it would give."""
    from bsalib_amd import synth
    from oracle import oracle as O
    if not O.reference_available():
        return None
    samples = []
    spent = 0.0
    nnl = 6
    while True:
        p = synth.deck(nn=nnl, nm=30, nfreqs=9, df=0.5, span=1000.0)
        t = O.reference_time(p, REF_THREADS)
        spent += t
        samples.append({"nnl": nnl, "nf": p.nf, "seconds": t,
                        "updates_per_s": p.ntc * nnl ** 3 * 30 ** 3 * p.nf ** 2 / t})
        if len(samples) >= 3 and spent + t * ((nnl + 2) / nnl) ** 3 > budget_s:
            break
        nnl += 2
    x = np.log([s["nnl"] for s in samples])
    y = np.log([s["seconds"] for s in samples])
    expo = float(np.polyfit(x, y, 1)[0])
    last = samples[-1]
    per_eval_fit = last["seconds"] * (1000.0 / last["nnl"]) ** expo / last["nf"] ** 2
    per_eval_cube = 2 * 1000 ** 3 * 30 ** 3 / last["updates_per_s"]
    return {"value": 1.0 / per_eval_cube, "unit": "evals/s", "cores": REF_THREADS,
            "kind": "reference", "threads": REF_THREADS, "host": host_description(),
            "sample": f"getFM_full_tnm_vect_cls_ on the deck at NNL="
                      f"{','.join(str(s['nnl']) for s in samples)}, NM=30, NTC=2, NF=17 "
                      f"({spent:.1f} s); the largest sample's update rate extrapolated to "
                      f"NNL=1000 by the loop count NTC*NNL^3*NM^3 (extrapolated)",
            "samples": samples, "nnl_exponent": expo,
            "value_fitted_exponent": 1.0 / per_eval_fit,
            "seconds": spent}


def cpu_factorised_baseline(prob, threads: int, npts: int = 64):
    """Factorised FP64 oracle (A(f) projection + three node contractions per
    point, OpenMP over points) on a sample of npts points of the benchmark grid."""
    from oracle import oracle as O
    nf = prob.nf
    rng = np.random.default_rng(7)
    ii = rng.integers(0, nf, npts)
    jj = rng.integers(0, nf, npts)
    t0 = time.perf_counter()
    O.classic_factorised(prob, rule="vect", truncate=True, nthreads=threads, point_sample=(ii, jj))
    t = time.perf_counter() - t0
    return {"value": npts / t, "unit": "evals/s", "cores": threads, "kind": "port",
            "threads": threads, "host": host_description(),
            "sample": f"{npts} random points of the 1025^2 grid incl. the projection table "
                      f"({t:.2f} s)"}


def precision_errors(prob, r64: dict, r: dict, prec: str, b64=None, bound=None) -> dict:
    """Errors of a mixed-precision run against the FP64 run of the same plan,
    per result: third moments and skewness normwise (max|d| / max|ref|) and per
    entry (max |d|/|ref| over entries with |ref| >= 1e-3 max|ref|), and the
    bispectrum L2 error ||B - B64|| / ||B64|| over 27 modal entries (modes 1,
    NM/2, NM in every position) at every point of the grid, force and response
    (bsahip_classic_bispectra_sel: the mixed G values with the G kernel's
    operand rounding and FP32 accumulation)."""
    from bsalib_amd import abi
    from bsalib_amd.host import skewness
    nm = prob.nm

    def norm_entry(a, ref):
        a, ref = np.asarray(a), np.asarray(ref)
        big = np.abs(ref) >= 1e-3 * np.abs(ref).max()
        return (float(np.abs(a - ref).max() / np.abs(ref).max()),
                float((np.abs(a - ref)[big] / np.abs(ref[big])).max()))
    out = {}
    for k in ("m3mf", "m3mr", "m2mr"):
        out[k + "_normwise"], out[k + "_per_entry"] = norm_entry(r[k], r64[k])
    sk64 = skewness(r64["m2mr"], r64["m3mr"], nm)
    sk = skewness(r["m2mr"], r["m3mr"], nm)
    out["skewness_normwise"], out["skewness_per_entry"] = norm_entry(sk, sk64)
    modes = np.array([0, nm // 2, nm - 1])
    sel = (modes[:, None, None] + nm * modes[None, :, None] + nm * nm * modes[None, None, :])
    sel = np.unique(sel.ravel()) + 1
    if b64 is None:
        b64 = abi.classic_bispectra_sel(prob, sel, bound=bound)
    bm = abi.classic_bispectra_sel(prob, sel, bound=bound, prec=prec)
    for k in ("bisp_f", "bisp_r"):
        out[k + "_l2"] = float(np.linalg.norm(bm[k] - b64[k]) / np.linalg.norm(b64[k]))
    out["bisp_entries"] = int(sel.size)
    out["bisp_points"] = int(prob.nf ** 2)
    return out, b64


def batch_leg(log):
    """Batches of uneven structures (bsahip_batch_*): the 256-structure mix of
    configuration 4 (NN 50-800, NM 4-24, 1-3 components, grids 257-1025) and 256
    small towers (NN 20-60, NM 3-6, grids 65-129), split-bf16x3.  evals/s from
    the device time of a step (tables + contraction + PSD moments) for one G
    launch per structure (mode 1), the persistent launch replayed from a HIP
    graph (mode 0) and the persistent launch with tables on worker streams
    (mode 2)."""
    from bsalib_amd import abi, synth
    out = {}
    rng = np.random.default_rng(5)
    small = []
    for s in range(256):
        nm = int(rng.integers(3, 7))
        small.append(synth.tower(nn=int(rng.integers(20, 61)), nm=nm,
                                 natf=np.sort(rng.uniform(0.15, 2.0, nm)),
                                 xi=rng.uniform(0.005, 0.02, nm), nfreqs=int(rng.choice([33, 65])),
                                 df=1 / 16, seed=s, zone_per_node=False))
    for name, probs in (("c4_mix_256", synth.batch(256, seed=0)), ("small_towers_256", small)):
        B = abi.Batch(probs, prec="bf16x3")
        grid_mult, chunk = B.autotune()
        res = {"structures": len(probs), "evals_per_step": B.evals, "items": B.info()["items"],
               "autotuned": {"grid_mult": grid_mult, "chunk": chunk}}
        for mode, key in ((1, "per_structure"), (0, "graph_persistent"), (2, "streams_persistent")):
            B.run(mode)
            ms = []
            for _ in range(2):
                B.run(mode)
                ms.append(B.info()["ms_step"])
            res[key] = {"ms_per_step": min(ms), "evals_per_s": B.evals / (min(ms) * 1e-3)}
        res["speedup_best_vs_per_structure"] = res["per_structure"]["ms_per_step"] / min(
            res["graph_persistent"]["ms_per_step"], res["streams_persistent"]["ms_per_step"])
        B.close()
        log(f"batch {name}: {res}")
        out[name] = res
    return out


def mesher_batch_leg(log, n: int = 64):
    """Post-mesh of n small towers (NN 20-60, NM 3-6, rule-based levels) as n
    bsahip_mesher_run calls and as one bsahip_mesher_run_batch (the FRF tables
    and quadrature of every Mesher's tiles in one launch each); wall time of
    each form after a warmup, and the largest relative difference of m3mr."""
    import time
    from bsalib_amd import abi, proposer, synth
    rng = np.random.default_rng(7)
    ms = []
    for s in range(n):
        nm = int(rng.integers(3, 7))
        p = synth.tower(nn=int(rng.integers(20, 61)), nm=nm, natf=np.sort(rng.uniform(0.15, 2.0, nm)),
                        xi=rng.uniform(0.005, 0.02, nm), nfreqs=65, df=1 / 16, seed=100 + s,
                        zone_per_node=False)
        ms.append(abi.Mesher(p, **proposer.mesher_settings(p)))
    seq = [M.run() for M in ms]
    abi.mesher_run_batch(ms)
    t0 = time.perf_counter()
    seq = [M.run() for M in ms]
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    bat = abi.mesher_run_batch(ms)
    t_bat = time.perf_counter() - t0
    diff = max(float(np.abs(b["m3mr"] - a["m3mr"]).max() / np.abs(a["m3mr"]).max())
               for a, b in zip(seq, bat))
    out = {"structures": n, "tiles": int(sum(M.ntiles for M in ms)),
           "evals": int(sum(r["evals_pre"] + r["evals_post"] for r in seq)),
           "brm_points": int(sum(r["brm_points"] for r in seq)),
           "separate_ms": 1e3 * t_seq, "batched_ms": 1e3 * t_bat, "speedup": t_seq / t_bat,
           "max_rel_diff_m3mr": diff}
    for M in ms:
        M.close()
    log(f"mesher batch: {out}")
    return out


def mesher_leg(log):
    """Learned refinement proposer on the deck of configuration 5 (NN=500,
    NM=20): BFM evaluations and third-moment error against the converged Mesher
    (every tile at level 8), next to every plain setting (uniform levels 1, 2,
    4, 8, the rule-based levels, the a-posteriori adaptive levels for 1e-3) and
    the reduction against the cheapest plain setting that meets 1e-3."""
    from bsalib_amd import proposer, synth
    P = proposer.Proposer()
    r = proposer.evaluate(synth.config5(), P, log=log)
    P.net.close()
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the structure-batch leg")
    ap.add_argument("--no-mesher", action="store_true", help="skip the learned-Mesher leg")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the per-result precision report (bispectrum L2 etc.)")
    ap.add_argument("--config", default="config3", choices=["config1", "config2", "config3"])
    ap.add_argument("--prec", default="both", choices=["both", "fp64", "fp32", "bf16x3"],
                    help="profiling aid: run only one precision (the JSON line needs both)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="profiling aid: one 1 GiB device-to-device copy (HBM counter scale)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(2 ** 28, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copied 1 GiB")
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build targets one GPU; multi-GPU runs are not part of it "
              "(the (f1,f2) plane would be sharded by tile pairs with one all-reduce of m3).",
              file=sys.stderr)
        sys.exit(2)

    from bsalib_amd import abi, synth
    prob = getattr(synth, args.config)()
    nm = prob.nm
    log(f"workload {prob.name} NN={prob.nn} NM={nm} NTC={prob.ntc} NF={prob.nf}")
    abi.init(0)
    bound = abi.Bound(prob)
    from bsalib_amd.host import skewness

    def timed(prec):
        # the structure is uploaded once (plan); a step is one full run on it
        plan = abi.Plan(prob, quad="vect", prec=prec, bound=bound)
        for _ in range(args.warmup):
            plan.run()
        kms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            r = plan.run()
            kms.append(abi.last_kernel_ms()[0])
        t = time.perf_counter() - t0
        plan.close()
        return r, t, float(np.mean(kms))

    evals = prob.nf ** 2
    flops = gsym_flops(prob, nm)
    if args.prec != "both":
        r1, t1, k1 = timed(args.prec)
        log(f"{args.prec}: step {1e3 * t1 / args.steps:.2f} ms, kernel {k1:.2f} ms, "
            f"{flops / k1 / 1e9:.2f} TFLOP/s (algorithmic)")
        return
    r64, t64, k64 = timed("fp64")
    log(f"fp64: step {1e3 * t64 / args.steps:.2f} ms, k_gsym_f64 {k64:.2f} ms, "
        f"{flops / k64 / 1e9:.2f} TFLOP/s")
    r32, t32, k32 = timed("fp32")
    log(f"fp32: step {1e3 * t32 / args.steps:.2f} ms, k_gsym_f32 {k32:.2f} ms, "
        f"{flops / k32 / 1e9:.2f} TFLOP/s")
    r, t, k_ms = timed("bf16x3")
    ms_step = 1e3 * t / args.steps
    value = evals * args.steps / t
    value64 = evals * args.steps / t64
    value32 = evals * args.steps / t32
    achieved = flops / (k_ms * 1e-3) / 1e12
    sk64 = skewness(r64["m2mr"], r64["m3mr"], nm)
    sk16 = skewness(r["m2mr"], r["m3mr"], nm)
    sk_err = float(np.abs(sk16 - sk64).max() / np.abs(sk64).max())
    m3_err = float(np.abs(r["m3mr"] - r64["m3mr"]).max() / np.abs(r64["m3mr"]).max())
    log(f"bf16x3: step {ms_step:.2f} ms, k_gsym_bf16x3 {k_ms:.2f} ms, {achieved:.2f} TFLOP/s, "
        f"speedup vs fp64 {value / value64:.2f}x, skewness err {sk_err:.2e}")
    errors = None
    if not args.no_errors:
        e16, b64 = precision_errors(prob, r64, r, "bf16x3", bound=bound)
        e32, _ = precision_errors(prob, r64, r32, "fp32", b64=b64, bound=bound)
        errors = {"bf16x3": e16, "fp32": e32}
        log(f"errors vs fp64: {errors}")

    batch = None if args.no_batch else batch_leg(log)
    if batch is not None:
        batch["mesher_small_towers_64"] = mesher_batch_leg(log)
    mesher = None if args.no_mesher else mesher_leg(log)
    cpu = None
    cpu2 = None
    if not args.no_cpu:
        threads = int(os.environ.get("BENCH_CPU_THREADS",
                                     str(min(16, host_description()["usable"] or 1))))
        try:
            cpu = cpu_reference_baseline()
            log(f"cpu reference baseline {cpu}")
        except Exception as e:  # noqa: BLE001
            log(f"cpu reference baseline failed: {e}")
        try:
            cpu2 = cpu_factorised_baseline(prob, threads)
            log(f"cpu factorised baseline {cpu2}")
        except Exception as e:  # noqa: BLE001
            log(f"cpu factorised baseline failed: {e}")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3",
        "data": "synthetic (seeded bridge-deck generator)",
        "config": {"workload": "bridge deck NN=1000 NM=30 NTC=2 (u,w), Classic grid 1025x1025 "
                               "over [-4,4] Hz; node contraction in split-bf16x3 on MFMA with "
                               "FP32 accumulation, FRF/quadrature/projection in FP64",
                   "nn": prob.nn, "nm": nm, "ntc": prob.ntc, "grid": prob.nf,
                   "precision": "bf16x3"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / BF16_PEAK_TFLOPS,
                     "traffic": GSYM_BF16X3_TRAFFIC_B if args.config == "config3" else None,
                     "kernel": "k_gsym_bf16x3", "kernel_ms": k_ms, "flops_per_launch": flops,
                     "note": "algorithmic flops counted once; split-bf16x3 issues 3 MFMA "
                             "products per term, so its attainable fraction is <= 1/3"},
        "fp64": {"value": value64, "unit": "evals/s", "ms_per_step": 1e3 * t64 / args.steps,
                 "kernel": "k_gsym_f64", "kernel_ms": k64,
                 "achieved_tflops": flops / (k64 * 1e-3) / 1e12, "peak_tflops": FP64_PEAK_TFLOPS,
                 "frac": flops / (k64 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS},
        "fp32": {"value": value32, "unit": "evals/s", "ms_per_step": 1e3 * t32 / args.steps,
                 "kernel": "k_gsym_f32", "kernel_ms": k32,
                 "achieved_tflops": flops / (k32 * 1e-3) / 1e12, "peak_tflops": FP32_PEAK_TFLOPS,
                 "frac": flops / (k32 * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                 "vs_fp64": value32 / value64},
        "mixed_vs_fp64": value / value64,
        "skewness_rel_err_vs_fp64": sk_err,
        "meets_target": bool(sk_err <= 1e-4 and value / value64 >= 4.0),
        "m3mr_rel_err_vs_fp64": m3_err,
        "errors": errors,
        "cpu_baseline": cpu,
        "cpu_baseline_factorised": cpu2,
        "batch": batch,
        "mesher_proposer": mesher,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
