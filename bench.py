"""Headline benchmark: paths*entries/s of the 4096^2 speckle covariance.

Workload (the mixed-precision covariance configuration): slab of thickness 1
(x, y in +-5e4), sigma = 20 (albedo 1), HG g = 0.9, 64 illumination x 64
view far-field directions uniform in solid angle on a 15 degree cone (seeded),
CBS on, exponential first-scatter sampling, lambda = 1: a 4096 x 4096
complex covariance.  Inputs are synthetic and generated on the device by the
tracer (Philox key 0).

A "step" is one MCcov pass over --paths-per-step paths (trace -> NEE
factors -> MFMA covariance update into the device accumulators of one
session); K timed steps follow W warm-up steps on disjoint path ranges.  Each
step returns only after its work has completed on the GPU (the library
synchronises its stream), so the timed region is bracketed by device
synchronisation on both sides.  value = K * paths_per_step * 4096^2 / time,
mixed precision (FP16 hi/lo MFMA, FP32 accumulation, FP64 tracer/phase).

Also reported: the plain FP64 GPU path on the same workload (fewer paths),
the error of the mixed result against the FP64 result on identical paths,
the roofline of the dominant kernel (the mixed covariance update, timed with
HIP events on its stream), and the FP64 CPU restatement timed on the host on a
bounded sample (the oracle; test infrastructure used here only as the
baseline being timed).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP16 = 2.5e15      # MI355X dense FP16 MFMA, flop/s
PEAK_FP64 = 78.6e12     # MI355X FP64, flop/s


def workload(seed_dirs=2024):
    from mcspeckle_amd import engines as E
    rng = np.random.default_rng(seed_dirs)

    def cone(n, deg):
        ct = rng.uniform(np.cos(np.deg2rad(deg)), 1.0, n)
        st = np.sqrt(1 - ct ** 2)
        ph = rng.uniform(0, 2 * np.pi, n)
        return np.vstack([st * np.cos(ph), st * np.sin(ph), ct])

    l, v = cone(64, 15), cone(64, 15)
    return E.make_problem(20.0, 1.0, [-5e4, -5e4, 0], [5e4, 5e4, 1], l, v, 1, 1, 1.0, True, 2, 3, 0.9, seed=0)


def cs_err(A, B):
    nv, _, nl, _ = B.shape
    C = np.transpose(A, (0, 2, 1, 3)).reshape(nv * nl, nv * nl, order="F")
    R = np.transpose(B, (0, 2, 1, 3)).reshape(nv * nl, nv * nl, order="F")
    d = np.sqrt(np.abs(np.real(np.diag(R))))
    return float(np.max(np.abs(C - R) / np.maximum(np.outer(d, d), 1e-300)))


def cpu_baseline(spec, n_paths, threads):
    from oracle import oracle as O
    t0 = time.perf_counter()
    O.mccov_sums(spec, 0, n_paths, threads)
    dt = time.perf_counter() - t0
    N = spec.nv * spec.nl
    return n_paths * N * N / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--paths-per-step", type=int, default=50000)
    ap.add_argument("--fp64-paths", type=int, default=10000)
    ap.add_argument("--cpu-paths", type=int, default=32)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(paths would shard by Philox path id with one final reduce).", file=sys.stderr)
        sys.exit(2)

    from mcspeckle_amd import _native
    from mcspeckle_amd import engines as E

    if _native.lib().scmc_device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    spec = workload()
    N = spec.nv * spec.nl
    P = args.paths_per_step

    # ---- mixed precision, the headline
    sess = E.Session(spec, "cov", "mixed")
    for w in range(args.warmup):
        sess.run(w * P, P)
    st0 = sess.stats()
    first = args.warmup * P
    t0 = time.perf_counter()
    for k in range(args.steps):
        sess.run(first + k * P, P)
    dt = time.perf_counter() - t0
    st1 = sess.stats()
    value = args.steps * P * N * N / dt
    herk_ms = st1["ms_herk"] - st0["ms_herk"]
    herk_flops = st1["herk_flops"] - st0["herk_flops"]
    herk_launches = st1["herk_launches"] - st0["herk_launches"]
    verts = st1["vertices"] - st0["vertices"]
    achieved = herk_flops / (herk_ms * 1e-3) if herk_ms > 0 else 0.0
    sess.close()

    out = {
        "metric": "paths*entries/s at 4096^2 covariance (MFMA mixed)",
        "value": value, "unit": "paths*entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp16x3+fp32acc (fp64 tracer/phase)", "data": "synthetic (device-sampled Philox paths, seed 0)",
        "config": {"workload": "mixed-precision covariance 4096^2: slab T=1 sigma=20 albedo=1 HG g=0.9, "
                               "64 lights x 64 views on a 15 deg cone, CBS on",
                   "paths_per_step": P, "vertices_per_path": verts / (args.steps * P)},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP16 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP16, "traffic": None,
                     "kernel": "herk_mixed_kernel", "expand_ms_per_launch": (st1["ms_expand"] - st0["ms_expand"]) / max(herk_launches, 1), "launches": herk_launches,
                     "avg_launch_ms": herk_ms / max(herk_launches, 1),
                     "flops_per_launch": herk_flops / max(herk_launches, 1),
                     "issued_mfma_flops_factor": 1.5},
    }

    # ---- plain FP64 GPU path on the same workload, and the error of mixed vs FP64 on the same paths
    if not args.no_fp64:
        Pf = args.fp64_paths
        s64 = E.Session(spec, "cov", "fp64")
        s64.run(0, min(Pf, 1000))      # warm-up range re-used below
        s64.close()
        s64 = E.Session(spec, "cov", "fp64")
        t1 = time.perf_counter()
        s64.run(0, Pf)
        d64 = time.perf_counter() - t1
        stf = s64.stats()
        Ms64, Mm64 = s64.result(Pf)
        s64.close()
        smx = E.Session(spec, "cov", "mixed")
        smx.run(0, Pf)
        Msm, Mmm = smx.result(Pf)
        smx.close()
        out["fp64_plain"] = {
            "value": Pf * N * N / d64, "unit": "paths*entries/s", "paths": Pf,
            "herk_tflops": stf["herk_flops"] / (stf["ms_herk"] * 1e-3) / 1e12 if stf["ms_herk"] > 0 else None,
            "herk_frac_of_fp64_peak": (stf["herk_flops"] / (stf["ms_herk"] * 1e-3) / PEAK_FP64)
            if stf["ms_herk"] > 0 else None,
        }
        out["mixed_vs_fp64_speedup"] = value / out["fp64_plain"]["value"]
        out["mixed_vs_fp64_max_cs_rel_err"] = cs_err(Msm + Mmm, Ms64 + Mm64)

    # ---- CPU baseline: FP64 restatement with the reference's loop structure
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        v, t = cpu_baseline(spec, args.cpu_paths, threads)
        out["cpu_baseline"] = {"value": v, "unit": "paths*entries/s", "cores": threads, "kind": "port",
                               "sample": f"first {args.cpu_paths} paths of the same workload (seed 0), "
                                         f"{t:.1f} s, OpenMP over path shards"}
        out["vs_cpu_baseline"] = value / v
    print(json.dumps(out))


if __name__ == "__main__":
    main()
