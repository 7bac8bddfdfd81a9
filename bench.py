#!/usr/bin/env python3
"""Free energies per second of batched BTRG (chi = 32) on one MI355X.

Workload: a beta scan of independent BTRG runs on the Trivial 2D Ising tensor
(src/models/ising.jl:44-55), beta drawn from a seeded uniform distribution on
[0.35, 0.55], truncrank(32), maxiter(30) -> 31 norms -> one free energy per
run.  One bench step is one batched solve of the whole scan on the device
(every coarse-graining step of every run issued as grouped launches).

Reported beside the throughput:
* accuracy gate: |f_batch - f_plain| / |f_plain| over a sample of runs,
  where f_plain is the plain one-run-at-a-time FP64 GPU path (must be
  <= 1e-9), and that path's own throughput;
* roofline of the dominant kernel class (device time from HIP events on the
  library's stream, algorithmic flops / bytes per launch);
* cpu_baseline: the FP64 numpy oracle (OpenBLAS dgemm / LAPACK gesdd, the
  library family TensorKit reaches) on the host cores, over a bounded
  sample of the same workload, scaled to free energies per second.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("free energies/sec (BTRG χ=32) at |Δf/f|≤1e-9 vs plain FP64 GPU; "
          "p99 latency")
SEED = 20260512
PEAK = {"fp64_mfma": 78.6e12, "fp32_mfma": 157.3e12, "hbm": 8.0e12}
KIND_NAMES = ["gemm_f64", "gemm_split32", "gemm_reduce", "jacobi", "permute", "colnorm",
              "gather", "trace_normalize", "fill", "misc"]


def device_sync(tk):
    tk.lib().tnr_device_sync()


def profile_batch(tk, reqs, precision):
    L = tk.lib()
    L.tnr_profile_enable(1)
    tk.run_batch(reqs, precision)
    out = {}
    for k, name in enumerate(KIND_NAMES):
        n, ms, fl, by = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        L.tnr_profile_read(k, C.byref(n), C.byref(ms), C.byref(fl), C.byref(by))
        out[name] = {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}
    L.tnr_profile_enable(0)
    return out


def roofline_from(prof, precision):
    name, d = max(prof.items(), key=lambda kv: kv[1]["ms"])
    total_ms = sum(v["ms"] for v in prof.values())
    if name.startswith("gemm"):
        peak = PEAK["fp32_mfma"] if name == "gemm_split32" else PEAK["fp64_mfma"]
        flops = d["flops"] * (3 if name == "gemm_split32" else 1)
        ach = flops / (d["ms"] * 1e-3) if d["ms"] else 0.0
        return {"kernel": name, "bound": "mfma", "achieved": ach / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s", "frac": ach / peak, "traffic": None,
                "share_of_device_time": d["ms"] / total_ms if total_ms else None}
    if name in ("permute", "gather"):
        ach = d["bytes"] / (d["ms"] * 1e-3) if d["ms"] else 0.0
        return {"kernel": name, "bound": "hbm", "achieved": ach / 1e9, "peak": PEAK["hbm"] / 1e9,
                "unit": "GB/s", "frac": ach / PEAK["hbm"], "traffic": None,
                "share_of_device_time": d["ms"] / total_ms if total_ms else None}
    # small-matrix solver kernels: report the GEMM class as the MFMA roofline
    g = prof["gemm_f64"]
    ach = g["flops"] / (g["ms"] * 1e-3) if g["ms"] else 0.0
    return {"kernel": "gemm_f64", "dominant_kernel": name, "bound": "mfma",
            "achieved": ach / 1e12, "peak": PEAK["fp64_mfma"] / 1e12, "unit": "TFLOP/s",
            "frac": ach / PEAK["fp64_mfma"], "traffic": None,
            "share_of_device_time": g["ms"] / total_ms if total_ms else None}


def cpu_baseline(chi, maxiter, beta, budget_s):
    """FP64 oracle on the host: time steps until the budget is spent, then
    scale the steady-state step time to one free energy (maxiter steps)."""
    from oracle import models as om
    from oracle import schemes as osch
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    sc = osch.BTRG(om.classical_ising_trivial(beta))
    sc.finalize()
    times = []
    t_all = time.perf_counter()
    for _ in range(maxiter):
        t0 = time.perf_counter()
        sc.step(chi)
        sc.finalize()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_all > budget_s:
            break
    n = len(times)
    if n >= maxiter:
        per_fe = sum(times)
        sample = f"one full BTRG chi={chi} run ({maxiter} steps)"
    else:
        steady = float(np.median(times[-3:]))
        per_fe = sum(times) + steady * (maxiter - n)
        sample = (f"first {n} of {maxiter} BTRG chi={chi} steps, remaining steps at the "
                  f"steady-state step time {steady:.2f} s")
    return {"value": 1.0 / per_fe, "unit": "free energies/s", "cores": threads, "kind": "port",
            "sample": sample + f"; {time.perf_counter() - t_all:.1f} s measured"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--chi", type=int, default=32)
    ap.add_argument("--maxiter", type=int, default=30)
    ap.add_argument("--precision", choices=["fp64", "split32"], default="fp64")
    ap.add_argument("--check", type=int, default=4, help="runs checked against the plain path")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent runs would shard by request across GPUs)", file=sys.stderr)
        sys.exit(2)
    import __graft_entry__
    tk = __graft_entry__.package()
    prec = tk.FP64 if a.precision == "fp64" else tk.SPLIT32
    rng = np.random.default_rng(SEED)
    betas = rng.uniform(0.35, 0.55, a.batch)
    reqs = [tk.Request("BTRG", tk.classical_ising(float(b), symmetry="Trivial"), a.chi, a.maxiter)
            for b in betas]

    for _ in range(a.warmup):
        tk.run_batch(reqs, prec)
    device_sync(tk)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = tk.run_batch(reqs, prec)
    device_sync(tk)
    dt = time.perf_counter() - t0
    ms_per_step = dt / a.steps * 1e3
    value = a.batch * a.steps / dt
    f_batch = [tk.free_energy(r, float(b)) for r, b in zip(res, betas)]

    # plain one-run-at-a-time FP64 path: accuracy reference and its throughput
    errs, t_plain = [], 0.0
    for i in range(min(a.check, a.batch)):
        sc = tk.BTRG(reqs[i].T)
        t1 = time.perf_counter()
        d = tk.run(sc, tk.truncrank(a.chi), tk.maxiter(a.maxiter), verbosity=0)
        t_plain += time.perf_counter() - t1
        sc.close()
        fp = tk.free_energy(d, float(betas[i]))
        errs.append(abs(f_batch[i] - fp) / abs(fp))
    plain_value = min(a.check, a.batch) / t_plain if t_plain else None

    out = {
        "metric": METRIC, "value": value, "unit": "free energies/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32x3-split+f64",
        "data": "synthetic: Trivial Ising tensors at seeded uniform beta in [0.35, 0.55]",
        "config": {"workload": f"BTRG beta-scan, {a.batch} runs, chi={a.chi}, maxiter={a.maxiter}",
                   "batch": a.batch, "chi": a.chi, "maxiter": a.maxiter, "seed": SEED,
                   "precision": a.precision},
        "accuracy": {"max_rel_df_vs_plain_fp64": max(errs) if errs else None,
                     "runs_checked": len(errs), "gate": 1e-9,
                     "pass": bool(errs) and max(errs) <= 1e-9},
        "plain_fp64_one_at_a_time": {"value": plain_value, "unit": "free energies/s"},
        "speedup_vs_plain": value / plain_value if plain_value else None,
        "latency_s": {"p50": ms_per_step / 1e3, "p99": ms_per_step / 1e3,
                      "note": "every run of a batch completes with the batch"},
    }
    if not a.no_profile:
        prof = profile_batch(tk, reqs, prec)
        out["roofline"] = roofline_from(prof, a.precision)
        out["kernel_time_ms"] = {k: round(v["ms"], 3) for k, v in prof.items() if v["launches"]}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.chi, a.maxiter, float(betas[0]), a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
