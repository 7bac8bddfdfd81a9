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
* roofline of the dominant kernel class over the timed steps: device time
  from each launch's dispatch start / end timestamps (HIP events of
  hipExtLaunchKernelGGL on the launching stream), algorithmic flops / bytes
  per launch.  Beside it the GEMM's rate on one untimed batch run on a
  single lane;
* latency p50/p99 from a seeded Poisson stream of BTRG solves (Z2 Ising,
  Z3 Potts, Z6 clock; chi 8-64) through the continuous-batching server;
* cpu_baseline: the FP64 numpy oracle (OpenBLAS dgemm / LAPACK gesdd, the
  library family TensorKit reaches) as a beta scan, one single-threaded
  process per host core, over a bounded sample of the same workload, scaled
  to free energies per second.
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


def read_profile(tk):
    """Per kernel class since tnr_profile_enable(1): launches, summed launch
    durations (each launch's dispatch start / end timestamps), the union of
    the launch intervals (device-busy time of the class; launches on the two
    lanes' streams overlap), algorithmic flops (GEMM: counted on the device
    for the problems that ran) and bytes; switches the recording off."""
    L = tk.lib()
    out = {}
    for k, name in enumerate(KIND_NAMES):
        n, ms, fl, by = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        busy, planned = C.c_double(), C.c_double()
        L.tnr_profile_read(k, C.byref(n), C.byref(ms), C.byref(fl), C.byref(by), C.byref(busy),
                           C.byref(planned))
        out[name] = {"launches": n.value, "ms": ms.value, "busy_ms": busy.value,
                     "flops": fl.value, "flops_planned": planned.value, "bytes": by.value}
    L.tnr_profile_enable(0)
    return out


def profile_batch(tk, reqs, precision):
    tk.lib().tnr_profile_enable(1)
    tk.run_batch(reqs, precision)
    return read_profile(tk)


def _gemm_roofline(prof, name, steps=1):
    """achieved = device-counted GEMM flops of the problems that ran / the
    union of the GEMM launch intervals (the time some GEMM was on the device;
    launches of the lanes overlap, so summed durations would double-count
    shared time); beside it the same flops over the summed launch durations,
    the per-launch figure a kernel trace gives."""
    d = prof[name]
    peak = PEAK["fp32_mfma"] if name == "gemm_split32" else PEAK["fp64_mfma"]
    flops = d["flops"] * (3 if name == "gemm_split32" else 1)   # 3 MFMA products per FMA
    ach = flops / (d["busy_ms"] * 1e-3) if d["busy_ms"] else 0.0
    per_launch = flops / (d["ms"] * 1e-3) if d["ms"] else 0.0
    n = max(1, d["launches"])
    return {"kernel": "k_gemm_split32" if name == "gemm_split32" else "fp64_gemm",
            "kernels": "k_gemm_split32" if name == "gemm_split32" else
                       "k_gemm_dl (LDS-DMA tiles) + k_gemm_f64 (32-deep register tile)",
            "bound": "mfma", "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": ach / peak, "launches": d["launches"],
            "flops_per_launch": d["flops"] / n,
            "flops_planned_per_launch": d["flops_planned"] / n,
            "busy_ms_per_step": d["busy_ms"] / steps,
            "avg_launch_ms": d["ms"] / n,
            "achieved_over_summed_launch_time": per_launch / 1e12}


def roofline_from(prof, precision, traffic, steps=1):
    """Roofline of the kernel class with the most device-busy time (union of
    its launch intervals, HIP dispatch timestamps on the lanes' streams)."""
    total = sum(v["busy_ms"] for v in prof.values())
    name, d = max(prof.items(), key=lambda kv: kv[1]["busy_ms"])
    n = max(1, d["launches"])
    if name.startswith("gemm"):
        r = _gemm_roofline(prof, name, steps)
    elif name == "jacobi":
        ach = d["flops"] / (d["busy_ms"] * 1e-3) if d["busy_ms"] else 0.0
        r = {"kernel": "k_jacobi", "bound": "mfma", "achieved": ach / 1e12,
             "peak": PEAK["fp64_mfma"] / 1e12, "unit": "TFLOP/s", "frac": ach / PEAK["fp64_mfma"],
             "launches": d["launches"], "avg_launch_ms": d["ms"] / n,
             "busy_ms_per_step": d["busy_ms"] / steps,
             "note": "FP64 VALU rotations (no MFMA); MI355X FP64 vector and matrix peaks are "
                     "both 78.6 TF/s; work = 18n flops per rotation, sweeps read back"}
    else:
        ach = d["bytes"] / (d["busy_ms"] * 1e-3) if d["busy_ms"] else 0.0
        r = {"kernel": name, "bound": "hbm", "achieved": ach / 1e9, "peak": PEAK["hbm"] / 1e9,
             "unit": "GB/s", "frac": ach / PEAK["hbm"], "launches": d["launches"],
             "avg_launch_ms": d["ms"] / n, "busy_ms_per_step": d["busy_ms"] / steps}
    r["share_of_device_busy_time"] = d["busy_ms"] / total if total else None
    r["traffic"] = traffic.get(r["kernel"]) if traffic else None
    return r


def load_traffic():
    """HBM bytes per launch per kernel from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE + WRITE_SIZE, calibrated on a copy
    kernel); None when absent."""
    p = os.path.join(ROOT, "profiles", "traffic_btrg_chi32_b128.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    return {k: v["bytes_per_launch"] for k, v in t.get("kernels", {}).items()}


def host_cpu_info():
    """What the host offers: nproc, the CPUs this process may run on, the
    cgroup's CPU quota (cores) where one is set, and the CPU model."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity"] = os.cpu_count()
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    info["cgroup_quota_cores"] = quota
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    info["model"] = model
    return info


def host_cpu_share(info, cap=16):
    """Worker processes for the CPU leg: the CPUs this process may use, the
    cgroup quota, and at most `cap` (on a host shared by several GPUs, one
    GPU's share is 16 cores, whatever nproc reports)."""
    n = info["affinity"] or 1
    if info["cgroup_quota_cores"]:
        n = min(n, max(1, int(info["cgroup_quota_cores"])))
    return max(1, min(n, cap))


def _cpu_run(args):
    """One oracle BTRG run: every step up to the steady bond dimension is
    timed once; once D = chi every step costs the same, so after one warm-up
    steady step the best of five steady steps prices the remaining ones
    (each timed step still advances the run).  Returns seconds per free
    energy and the steps actually run."""
    chi, maxiter, beta, budget_s = args
    from oracle import models as om
    from oracle import schemes as osch
    sc = osch.BTRG(om.classical_ising_trivial(beta))
    t0 = time.perf_counter()
    sc.finalize()
    total = time.perf_counter() - t0
    D, done, steady = 2, 0, []
    t_all = time.perf_counter()
    while done < maxiter:
        t0 = time.perf_counter()
        sc.step(chi)
        sc.finalize()
        dt = time.perf_counter() - t0
        done += 1
        D2 = min(chi, D * D)
        if D2 < chi or D < chi:
            total += dt            # bond dimension still growing: timed once
        else:
            steady.append(dt)      # steady state: warm-up + best of five
            if len(steady) == 6:
                break
        D = D2
        if time.perf_counter() - t_all > budget_s and len(steady) >= 2:
            break
    if steady:
        best = min(steady[1:]) if len(steady) > 1 else steady[0]
        total += best * (maxiter - (done - len(steady)))
    return total, done


def cpu_baseline(chi, maxiter, betas, budget_s):
    """FP64 oracle (numpy -> OpenBLAS dgemm / LAPACK gesdd) on the host:
    * beta-scan throughput: one single-threaded process per core of this
      process's CPU share, each on its own beta; value = sum over processes
      of 1 / (seconds per free energy);
    * a single run with OpenBLAS threads = that share (the latency figure)."""
    import multiprocessing as mp
    info = host_cpu_info()
    procs = min(host_cpu_share(info), len(betas))
    saved = {k: os.environ.get(k) for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS",
                                             "MKL_NUM_THREADS")}
    try:
        for k in saved:
            os.environ[k] = "1"
        with mp.get_context("spawn").Pool(procs) as pool:
            res = pool.map(_cpu_run, [(chi, maxiter, float(betas[i]), budget_s)
                                      for i in range(procs)])
        for k in saved:
            os.environ[k] = str(procs)
        with mp.get_context("spawn").Pool(1) as pool:
            single = pool.map(_cpu_run, [(chi, maxiter, float(betas[0]), budget_s)])[0]
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    value = sum(1.0 / r[0] for r in res)
    steps = min(r[1] for r in res)
    sample = (f"{procs} single-threaded oracle BTRG chi={chi} runs on distinct betas, each "
              f"timed over its first {steps}+ of {maxiter} steps: steps before the bond "
              f"dimension reaches chi timed once, then a warm-up steady step and the best of "
              f"five steady steps pricing the rest")
    return {"value": value, "unit": "free energies/s", "cores": procs, "kind": "port",
            "sample": sample, "host": info,
            "single_run": {"free_energies_per_s": 1.0 / single[0], "openblas_threads": procs,
                           "seconds_per_free_energy": single[0]}}


def oracle_scan_free_energies(chi, maxiter, betas):
    """(index, f) of the headline scan's runs whose FP64 oracle free energy
    is stored in tests/golden/oracle_configs.json (matched by beta, chi and
    maxiter)."""
    p = os.path.join(ROOT, "tests", "golden", "oracle_configs.json")
    if not os.path.exists(p):
        return []
    with open(p) as f:
        g = json.load(f)
    out = []
    for c in g.values():
        if c["scheme"] != "BTRG" or c["model"] != "ising_trivial" or c["chi"] != chi or \
                c["maxiter"] != maxiter:
            continue
        for i, b in enumerate(betas):
            if float(b) == c["beta"]:
                out.append((i, c["f"]))
                break
    return out


def btrg_model_flops(chi, maxiter):
    """Algorithmic FP64 flops of one BTRG free energy on the Trivial Ising
    tensor: bond schedule D0 = 2, D' = min(chi, D^2); per step the optimal
    pairwise contraction 2(chi'^4 D^2 + 2 D^3 chi'^2 + 4 chi' D^3) plus two
    dense truncated SVDs of n = D^2 at 8n^3/3 + 8n^2 chi' each."""
    D, total = 2, 0.0
    for _ in range(max(maxiter, 1)):
        cp = min(chi, D * D)
        n = D * D
        total += 2.0 * (cp ** 4 * D ** 2 + 2 * D ** 3 * cp ** 2 + 4 * cp * D ** 3)
        total += 2.0 * (8.0 * n ** 3 / 3.0 + 8.0 * n * n * cp)
        D = cp
    return total


def _timed(fn):
    t0 = time.perf_counter()
    r = fn()
    return r, time.perf_counter() - t0


def workload_c2(tk, maxiter):
    """TRG and HOTRG on Trivial Ising at beta_c, chi 32 and 48: plain FP64
    against split-FP32 contractions, per-step norm and free-energy error."""
    cases = []
    T = tk.classical_ising(None, symmetry="Trivial")
    fo = tk.f_onsager
    for name, sf in (("TRG", 2.0), ("HOTRG", 4.0)):
        for chi in (32, 48):
            res = {}
            for prec in (tk.FP64, tk.SPLIT32):
                sc = getattr(tk, name)(T, precision=prec)
                d, dt = _timed(lambda: tk.run(sc, tk.truncrank(chi), tk.maxiter(maxiter),
                                              verbosity=0))
                sc.close()
                res[prec] = (np.asarray(d), dt)
            d64, t64 = res[tk.FP64]
            d32, t32 = res[tk.SPLIT32]
            f64 = tk.free_energy(d64, tk.ising_betac, scalefactor=sf)
            f32 = tk.free_energy(d32, tk.ising_betac, scalefactor=sf)
            cases.append({"scheme": name, "chi": chi, "maxiter": maxiter,
                          "fp64_s_per_run": t64, "split32_s_per_run": t32,
                          "steps_per_s_fp64": maxiter / t64,
                          "max_rel_norm_diff_split_vs_fp64": float(np.max(np.abs(d32 - d64) /
                                                                         np.abs(d64))),
                          "rel_df_split_vs_fp64": abs(f32 - f64) / abs(f64),
                          "rel_df_fp64_vs_onsager": abs(f64 - fo) / abs(fo)})
    return {"workload": "C2: TRG/HOTRG Trivial Ising beta_c, chi 32/48, plain FP64 vs split",
            "cases": cases}


def workload_plain(tk, chi, maxiter, n):
    """The plain one-run-at-a-time FP64 path on the headline's first n betas:
    each BTRG run is its own scheme handle, created, run and destroyed."""
    betas = np.random.default_rng(SEED).uniform(0.35, 0.55, n)
    times = []
    for b in betas:
        sc = tk.BTRG(tk.classical_ising(float(b), symmetry="Trivial"))
        _, dt = _timed(lambda: tk.run(sc, tk.truncrank(chi), tk.maxiter(maxiter), verbosity=0))
        sc.close()
        times.append(dt)
    return {"workload": f"plain BTRG runs one at a time, chi={chi}, maxiter={maxiter}",
            "runs": n, "s_per_run": times, "free_energies_per_s": n / sum(times)}


def workload_c3(tk, maxiter):
    """BTRG with Z2 Ising at beta_c and Z3 Potts at ln(1+sqrt 3), chi = 64."""
    cases = []
    for label, T, beta in (("Z2 Ising", tk.classical_ising(None, symmetry="Z2"), tk.ising_betac),
                           ("Z3 Potts", tk.classical_potts(3), tk.potts_betac(3))):
        sc = tk.BTRG(T)
        d, dt = _timed(lambda: tk.run(sc, tk.truncrank(64), tk.maxiter(maxiter), verbosity=0))
        sc.close()
        f = tk.free_energy(d, beta)
        c = {"model": label, "chi": 64, "maxiter": maxiter, "s_per_run": dt,
             "steps_per_s": maxiter / dt, "f": f}
        if label == "Z2 Ising":
            c["rel_df_vs_onsager"] = abs(f - tk.f_onsager) / abs(tk.f_onsager)
        # the same two runs batched together
        cases.append(c)
    reqs = [tk.Request("BTRG", tk.classical_ising(None, symmetry="Z2"), 64, maxiter),
            tk.Request("BTRG", tk.classical_potts(3), 64, maxiter)]
    _, dt = _timed(lambda: tk.run_batch(reqs))
    return {"workload": "C3: BTRG Z2 Ising beta_c and Z3 Potts ln(1+sqrt3), chi 64",
            "cases": cases, "both_batched_s": dt}


def _oracle_impurity_seconds(kind, chi, maxiter, budget_s):
    """Single-threaded oracle time of one impurity run (run in a spawned
    process with one BLAS thread), steps past budget_s extrapolated."""
    import multiprocessing as mp
    saved = {k: os.environ.get(k) for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS",
                                             "MKL_NUM_THREADS")}
    for k in saved:
        os.environ[k] = "1"
    try:
        with mp.get_context("spawn").Pool(1) as pool:
            return pool.apply(_cpu_impurity, (kind, chi, maxiter, budget_s))
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _cpu_impurity(kind, chi, maxiter, budget_s):
    from oracle import models as om
    from oracle import schemes as osch
    T, Ti = om.classical_ising_trivial(om.ising_betac_mp), om.classical_ising_impurity(om.ising_betac_mp)
    sc = (osch.ImpurityTRG(T, Ti, T, T, T, gauge="pivoted") if kind == "ImpurityTRG"
          else osch.ImpurityHOTRG(T, Ti, Ti, T))
    t0 = time.perf_counter()
    sc.finalize()
    last, steps = 0.0, 0
    while steps < maxiter:
        t1 = time.perf_counter()
        sc.step(chi)
        sc.finalize()
        last = time.perf_counter() - t1
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    return (time.perf_counter() - t0) + (maxiter - steps) * last, steps


def workload_imp(tk, maxiter, cpu_seconds):
    """ImpurityTRG (sites 1-4, magnetisation impurity at site 1) and
    ImpurityHOTRG (first- and second-order impurities) on Trivial Ising at
    beta_c: time per run and per step on the device, free energy against
    Onsager, <m> (ImpurityTRG) and <m^2> (ImpurityHOTRG) of the last step,
    and the single-threaded oracle on the same run beside it (at the
    reference's test bond dimensions)."""
    cases = []
    T, Ti = tk.classical_ising(None, "Trivial"), tk.classical_ising_impurity(None)
    for kind, sf, chis in (("ImpurityTRG", 2.0, (24, 32, 48)), ("ImpurityHOTRG", 4.0, (16, 24, 32))):
        for chi in chis:
            mk = ((lambda: tk.ImpurityTRG(T, Ti, T, T, T)) if kind == "ImpurityTRG"
                  else (lambda: tk.ImpurityHOTRG(T, Ti, Ti, T)))
            w = mk()
            tk.run(w, tk.truncrank(chi), tk.maxiter(2), verbosity=0)   # warm-up
            w.close()
            sc = mk()
            d, dt = _timed(lambda: tk.run(sc, tk.truncrank(chi), tk.maxiter(maxiter),
                                          verbosity=0))
            sc.close()
            f = tk.free_energy([x[0] for x in d], tk.ising_betac, scalefactor=sf)
            c = {"scheme": kind, "chi": chi, "maxiter": maxiter, "s_per_run": dt,
                 "steps_per_s": maxiter / dt, "rel_df_vs_onsager": abs(f - tk.f_onsager) /
                 abs(tk.f_onsager)}
            if kind == "ImpurityTRG":
                c["m_last"] = d[-1][1] / d[-1][0]
            else:
                c["m2_last"] = d[-1][3] / d[-1][0]
            if chi == chis[0] and cpu_seconds > 0:
                cs, steps = _oracle_impurity_seconds(kind, chi, maxiter, cpu_seconds)
                c["oracle_cpu_s_per_run"] = cs
                c["oracle_cpu_steps_timed"] = steps
                c["gpu_speedup_vs_oracle_1core"] = cs / dt
            cases.append(c)
    return {"workload": "Impurity schemes: Trivial Ising beta_c, magnetisation impurities",
            "cases": cases}


def _median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def workload_obs(tk, reps, with_cpu):
    """Observables on the fixed-point tensor after a run (cft_data,
    ground_state_degeneracy, gu_wen_ratio, central_charge, the 2x2 finalize):
    median seconds per call on the device at the reference's CFT test
    configurations and at larger Trivial chi, the single-threaded oracle on
    the same state beside the small ones."""
    import os as _os
    cases = []
    configs = (("TRG", "Z2", 24, 10), ("BTRG", "Z2", 24, 10), ("HOTRG", "Z2", 16, 4),
               ("TRG", "Trivial", 64, 10), ("BTRG", "Trivial", 80, 8))
    for name, sym, chi, n in configs:
        sc = getattr(tk, name)(tk.classical_ising(symmetry=sym))
        d = tk.run(sc, tk.truncrank(chi), tk.maxiter(n), verbosity=0)
        c = {"scheme": name, "symmetry": sym, "chi": chi, "maxiter": n}
        tk.cft_data(sc)                                          # warm-up
        c["cft_data_s"] = _median_time(lambda: tk.cft_data(sc), reps)
        c["ground_state_degeneracy_s"] = _median_time(lambda: tk.ground_state_degeneracy(sc), reps)
        c["gu_wen_ratio_s"] = _median_time(lambda: tk.gu_wen_ratio(sc), reps)
        c["central_charge_s"] = _median_time(lambda: tk.central_charge(sc, d[-1]), reps)
        c["finalize_two_by_two_s"] = _median_time(lambda: tk.finalize_two_by_two(sc), reps)
        cft = tk.cft_data(sc)
        c["cft_1_2"] = [float(cft[1]), float(cft[2])] if len(cft) > 2 else None
        if with_cpu and chi <= 24:
            _os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
            from oracle import models as om
            from oracle import observables as ob
            from oracle import schemes as osch
            T = om.classical_ising_z2(om.ising_betac_mp) if sym == "Z2" else \
                om.classical_ising_trivial(om.ising_betac_mp)
            o = getattr(osch, name)(T)
            osch.run(o, chi, n)
            c["oracle_cft_data_s"] = _median_time(lambda: ob.cft_data(o), reps)
            c["oracle_gu_wen_ratio_s"] = _median_time(lambda: ob.gu_wen_ratio(o), reps)
            c["oracle_finalize_two_by_two_s"] = _median_time(lambda: ob.finalize_two_by_two(o), reps)
        sc.close()
        cases.append(c)
    return {"workload": "Observables on the Ising fixed point, per call", "reps": reps,
            "cases": cases}


def _cpu_atrg(chi, maxiter, budget_s):
    from oracle import models as om
    from oracle import schemes as osch
    sc = osch.ATRG(om.classical_ising_z2(om.ising_betac_mp))
    t0 = time.perf_counter()
    sc.finalize()
    last, steps = 0.0, 0
    while steps < maxiter:
        t1 = time.perf_counter()
        sc.step(chi)
        sc.finalize()
        last = time.perf_counter() - t1
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    return (time.perf_counter() - t0) + (maxiter - steps) * last, steps


def workload_atrg(tk, maxiter, batch, cpu_seconds):
    """ATRG on Z2 Ising at beta_c (one run at a time, chi 24 / 32 / 48) and a
    batched beta scan of Trivial ATRG runs at chi 24; the single-threaded
    oracle on the chi-24 run beside it."""
    cases = []
    for chi in (24, 32, 48):
        w = tk.ATRG(tk.classical_ising(symmetry="Z2"))
        tk.run(w, tk.truncrank(chi), tk.maxiter(2), verbosity=0)
        w.close()
        sc = tk.ATRG(tk.classical_ising(symmetry="Z2"))
        d, dt = _timed(lambda: tk.run(sc, tk.truncrank(chi), tk.maxiter(maxiter), verbosity=0))
        sc.close()
        f = tk.free_energy(d, tk.ising_betac, scalefactor=4.0)
        c = {"chi": chi, "maxiter": maxiter, "s_per_run": dt, "steps_per_s": maxiter / dt,
             "rel_df_vs_onsager": abs(f - tk.f_onsager) / abs(tk.f_onsager)}
        if chi == 24 and cpu_seconds > 0:
            import multiprocessing as mp
            with mp.get_context("spawn").Pool(1) as pool:
                cs, steps = pool.apply(_cpu_atrg, (chi, maxiter, cpu_seconds))
            c["oracle_cpu_s_per_run"] = cs
            c["oracle_cpu_steps_timed"] = steps
            c["gpu_speedup_vs_oracle_1core"] = cs / dt
        cases.append(c)
    rng = np.random.default_rng(SEED + 7)
    betas = rng.uniform(0.35, 0.55, batch)
    reqs = [tk.Request("ATRG", tk.classical_ising(float(b), symmetry="Trivial"), 24, maxiter)
            for b in betas]
    tk.run_batch(reqs[:4])
    _, dt = _timed(lambda: tk.run_batch(reqs))
    return {"workload": "ATRG: Z2 Ising beta_c single runs; Trivial beta scan batch at chi 24",
            "cases": cases, "batch": {"runs": batch, "chi": 24, "maxiter": maxiter,
                                      "free_energies_per_s": batch / dt, "s": dt}}


def _cpu_hotrg3d(chi, maxiter, budget_s):
    from oracle import models as om
    from oracle import schemes as osch
    sc = osch.HOTRG_3D(om.classical_ising_3d_z2(om.ising_betac_3d))
    t0 = time.perf_counter()
    sc.finalize()
    last, steps = 0.0, 0
    while steps < maxiter:
        t1 = time.perf_counter()
        sc.step(chi)
        sc.finalize()
        last = time.perf_counter() - t1
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    return (time.perf_counter() - t0) + (maxiter - steps) * last, steps


def workload_3d(tk, maxiter, cpu_seconds):
    """HOTRG_3D on the Z2 3D Ising tensor at beta_c, chi 8 / 10 / 12, one
    run at a time; free energy against -3.507; the single-threaded oracle
    on the chi-8 run beside it."""
    cases = []
    for chi in (8, 10, 12):
        w = tk.HOTRG_3D(tk.classical_ising_3D())
        tk.run(w, tk.truncrank(chi), tk.maxiter(1), verbosity=0)
        w.close()
        sc = tk.HOTRG_3D(tk.classical_ising_3D())
        d, dt = _timed(lambda: tk.run(sc, tk.truncrank(chi), tk.maxiter(maxiter), verbosity=0))
        sc.close()
        f = tk.free_energy(d, tk.ising_betac_3D, scalefactor=8.0)
        c = {"chi": chi, "maxiter": maxiter, "s_per_run": dt, "steps_per_s": maxiter / dt,
             "f": f, "rel_df_vs_benchmark": abs(f / tk.f_benchmark3D - 1)}
        if chi == 8 and cpu_seconds > 0:
            import multiprocessing as mp
            with mp.get_context("spawn").Pool(1) as pool:
                cs, steps = pool.apply(_cpu_hotrg3d, (chi, maxiter, cpu_seconds))
            c["oracle_cpu_s_per_run"] = cs
            c["oracle_cpu_steps_timed"] = steps
            c["gpu_speedup_vs_oracle_1core"] = cs / dt
        cases.append(c)
    return {"workload": "HOTRG_3D: Z2 3D Ising beta_c", "cases": cases}


def workload_copy():
    """Calibration of the HBM-traffic counters: three device-to-device
    copies of 256 MiB (the HIP runtime's copy kernel), so rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE of a known byte count give the counters' scale."""
    h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    nbytes = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    if h.hipMalloc(C.byref(a), nbytes) or h.hipMalloc(C.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    try:
        for _ in range(3):
            if h.hipMemcpy(b, a, nbytes, 3):          # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        h.hipDeviceSynchronize()
    finally:
        h.hipFree(a)
        h.hipFree(b)
    return {"workload": "copy calibration: 3 device copies of 256 MiB", "bytes": nbytes}


def workload_gemm(tk, shapes, modes, repeat=5):
    """FP64 GEMM microbenchmark through tnr_kernel_gemm: C = A B on
    device-resident copies, the device time of the last of `repeat` launches
    (HIP events).  modes: 2 / 3 / 4 = 64x64 / 128x64 / 128x128 tiles, | 8 no
    split-K, | 32 register-staged loads instead of LDS-DMA, 0 = the engine's
    choice."""
    D = C.POINTER(C.c_double)
    rows = []
    for (M, N, K) in shapes:
        rng = np.random.default_rng(M + N + K)
        A = np.asfortranarray(rng.standard_normal((M, K)))
        B = np.asfortranarray(rng.standard_normal((K, N)))
        Cm = np.zeros((M, N), order="F")
        for mode in modes:
            ms = C.c_double()
            st = tk.lib().tnr_kernel_gemm(M, N, K, A.ctypes.data_as(D), M, 0, B.ctypes.data_as(D),
                                          K, 0, Cm.ctypes.data_as(D), M, 1.0, 0.0, mode, repeat,
                                          C.byref(ms))
            if st != 0:
                raise RuntimeError(tk.lib().tnr_last_error())
            rows.append({"M": M, "N": N, "K": K, "mode": mode, "ms": ms.value,
                         "tflops": 2.0 * M * N * K / (ms.value * 1e-3) / 1e12})
    return {"workload": "FP64 GEMM through tnr_kernel_gemm, device-resident operands",
            "cases": rows}


def workload_c4(tk, n, maxiter, check):
    """512 BTRG runs, beta ~ U[0.35, 0.55], chi ~ U{16, 24, 32, 48}."""
    rng = np.random.default_rng(SEED + 4)
    betas = rng.uniform(0.35, 0.55, n)
    chis = rng.choice([16, 24, 32, 48], n)
    reqs = [tk.Request("BTRG", tk.classical_ising(float(b), symmetry="Trivial"), int(c), maxiter)
            for b, c in zip(betas, chis)]
    out = {"workload": f"C4: {n} BTRG runs, beta U[0.35,0.55], chi U{{16,24,32,48}}, "
                       f"maxiter {maxiter}", "seed": SEED + 4}
    fs = {}
    for name, prec in (("fp64", tk.FP64), ("split32", tk.SPLIT32)):
        tk.run_batch(reqs[:8], prec)
        res, dt = _timed(lambda: tk.run_batch(reqs, prec))
        fs[name] = [tk.free_energy(r, float(b)) for r, b in zip(res, betas)]
        out[name] = {"free_energies_per_s": n / dt, "s": dt}
    idx = list(range(min(check, n)))
    errs = {"fp64": [], "split32": []}
    t_plain = 0.0
    for i in idx:
        sc = tk.BTRG(reqs[i].T)
        d, dt = _timed(lambda: tk.run(sc, tk.truncrank(int(chis[i])), tk.maxiter(maxiter),
                                      verbosity=0))
        t_plain += dt
        sc.close()
        fp = tk.free_energy(d, float(betas[i]))
        for k in errs:
            errs[k].append(abs(fs[k][i] - fp) / abs(fp))
    if idx:   # --check 0 runs no plain reference
        out["plain_fp64_one_at_a_time_per_s"] = len(idx) / t_plain
        for k in errs:
            out[k]["max_rel_df_vs_plain_fp64"] = max(errs[k])
    return out


def serving_stream(tk, seed, n_req, rate, precision):
    """Poisson stream of BTRG solves through the continuous-batching server.

    Request mix: Z2 Ising beta ~ U[0.35, 0.55], Z3 Potts beta ~ U[0.8, 1.2],
    Z6 clock beta ~ U[0.5, 1.5]; chi ~ U{8, 16, 32, 64}; maxiter ~ U[20, 40].
    Latency of a request = completion seen by poll - its arrival time (both
    on the host clock, so queueing is included)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n_req):
        m = int(rng.integers(3))
        if m == 0:
            T = tk.classical_ising(float(rng.uniform(0.35, 0.55)), symmetry="Z2")
        elif m == 1:
            T = tk.classical_potts(3, float(rng.uniform(0.8, 1.2)))
        else:
            T = tk.classical_clock(6, float(rng.uniform(0.5, 1.5)))
        chi = int(rng.choice([8, 16, 32, 64]))
        reqs.append(tk.Request("BTRG", T, chi, int(rng.integers(20, 41))))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    srv = tk.Server(0, 1024, precision)
    lat = np.zeros(n_req)
    inflight = {}
    nxt, ndone = 0, 0
    t0 = time.perf_counter()
    while ndone < n_req:
        now = time.perf_counter() - t0
        while nxt < n_req and arrivals[nxt] <= now:
            inflight[srv.submit(reqs[nxt])] = nxt
            nxt += 1
        if not inflight:
            time.sleep(max(0.0, arrivals[nxt] - now))
            continue
        srv.pump()
        t_poll = time.perf_counter() - t0
        for rid in list(inflight):
            if srv.poll(rid) is not None:
                i = inflight.pop(rid)
                lat[i] = t_poll - arrivals[i]
                ndone += 1
    span = time.perf_counter() - t0 - arrivals[0]
    srv.close()
    return {"requests": n_req, "rate_per_s": rate, "seed": seed, "solves_per_s": n_req / span,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mix": "BTRG; Z2 Ising / Z3 Potts / Z6 clock; chi in {8,16,32,64}; maxiter 20-40"}


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
    ap.add_argument("--gemm-shapes", default="4096x4096x1024,8192x8192x2048",
                    help="--workload gemm: MxNxK,...")
    ap.add_argument("--gemm-modes", default="2,3,4,34,35,36",
                    help="--workload gemm: tnr_kernel_gemm modes")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--serve-rate", type=float, default=6.0, help="Poisson arrivals per second")
    ap.add_argument("--workload", choices=["scan", "c2", "c3", "c4", "imp", "obs", "atrg", "3d",
                                           "plain", "copy", "gemm"],
                    default="scan",
                    help="scan = the headline beta scan; c2-c4, imp, obs, atrg, 3d and plain print a "
                         "report line of their own")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent runs would shard by request across GPUs)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "copy":
        print(json.dumps(workload_copy()))
        return
    if a.workload != "scan":
        import __graft_entry__
        tk = __graft_entry__.package()
        if a.workload == "c2":
            r = workload_c2(tk, a.maxiter)
        elif a.workload == "c3":
            r = workload_c3(tk, a.maxiter)
        elif a.workload == "imp":
            r = workload_imp(tk, a.maxiter, 0.0 if a.no_cpu else a.cpu_seconds)
        elif a.workload == "obs":
            r = workload_obs(tk, 5, not a.no_cpu)
        elif a.workload == "3d":
            r = workload_3d(tk, a.maxiter, 0.0 if a.no_cpu else a.cpu_seconds)
        elif a.workload == "plain":
            r = workload_plain(tk, a.chi, a.maxiter, 6)
        elif a.workload == "gemm":
            shapes = [tuple(int(x) for x in s.split("x")) for s in a.gemm_shapes.split(",")]
            r = workload_gemm(tk, shapes, [int(m) for m in a.gemm_modes.split(",")])
        elif a.workload == "atrg":
            r = workload_atrg(tk, a.maxiter, 64 if a.batch == 128 else a.batch,
                              0.0 if a.no_cpu else a.cpu_seconds)
        else:
            r = workload_c4(tk, 512 if a.batch == 128 else a.batch, a.maxiter, a.check)
        print(json.dumps(r))
        return
    rng = np.random.default_rng(SEED)
    betas = rng.uniform(0.35, 0.55, a.batch)
    # the CPU leg runs first, in fresh single-threaded processes, before this
    # process touches the GPU
    cpu = None if a.no_cpu else cpu_baseline(a.chi, a.maxiter, betas, a.cpu_seconds)
    import __graft_entry__
    tk = __graft_entry__.package()
    prec = tk.FP64 if a.precision == "fp64" else tk.SPLIT32
    reqs = [tk.Request("BTRG", tk.classical_ising(float(b), symmetry="Trivial"), a.chi, a.maxiter)
            for b in betas]

    for _ in range(a.warmup):
        tk.run_batch(reqs, prec)
    device_sync(tk)
    if not a.no_profile:
        # every launch of the timed steps carries its dispatch's start / end
        # timestamps; nothing synchronises the host
        tk.lib().tnr_profile_enable(1)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = tk.run_batch(reqs, prec)
    device_sync(tk)
    dt = time.perf_counter() - t0
    prof = None if a.no_profile else read_profile(tk)
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
    # the FP64 oracle's free energies of the scan's first runs (committed
    # fixtures, computed on the CPU by tests/golden/make_golden.py configs)
    oracle_errs = []
    for i, f_or in oracle_scan_free_energies(a.chi, a.maxiter, betas):
        oracle_errs.append(abs(f_batch[i] - f_or) / abs(f_or))

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
                     "runs_checked": len(errs),
                     "max_rel_df_vs_fp64_oracle": max(oracle_errs) if oracle_errs else None,
                     "oracle_runs_checked": len(oracle_errs), "gate": 1e-9,
                     "pass": bool(errs) and max(errs) <= 1e-9 and
                             (not oracle_errs or max(oracle_errs) <= 1e-9)},
        "plain_fp64_one_at_a_time": {"value": plain_value, "unit": "free energies/s"},
        "speedup_vs_plain": value / plain_value if plain_value else None,
        "steps_per_s": value * max(a.maxiter, 1),   # coarse-graining steps of all runs
        "batch_latency_s": ms_per_step / 1e3,
        "model_flops_per_free_energy": btrg_model_flops(a.chi, a.maxiter),
        "end_to_end_model_tflops": btrg_model_flops(a.chi, a.maxiter) * value / 1e12,
        "end_to_end_frac_fp64_peak": btrg_model_flops(a.chi, a.maxiter) * value / PEAK["fp64_mfma"],
    }
    if a.serve_requests > 0:
        sv = serving_stream(tk, SEED + 1, a.serve_requests, a.serve_rate, prec)
        out["latency_s"] = {"p50": sv["p50_s"], "p99": sv["p99_s"], "source": "serving stream"}
        out["serving"] = sv
    if prof is not None:
        out["roofline"] = roofline_from(prof, a.precision, load_traffic(), a.steps)
        gname = "gemm_split32" if a.precision == "split32" else "gemm_f64"
        out["roofline_gemm"] = _gemm_roofline(prof, gname, a.steps)
        out["kernel_busy_ms_per_step"] = {k: round(v["busy_ms"] / a.steps, 3)
                                          for k, v in prof.items() if v["launches"]}
        # the same batch once more on a single lane: no launch then shares the
        # device with the other lane's kernels, so this is the GEMM's own rate
        prev = os.environ.get("TNR_LANES")
        os.environ["TNR_LANES"] = "1"
        try:
            p1 = profile_batch(tk, reqs, prec)
        finally:
            if prev is None:
                del os.environ["TNR_LANES"]
            else:
                os.environ["TNR_LANES"] = prev
        g1 = _gemm_roofline(p1, gname)
        g1["note"] = ("one untimed batch on one lane (TNR_LANES=1); the timed steps run "
                      "several lanes whose kernels overlap, which lengthens each launch")
        out["roofline_gemm_one_lane"] = g1
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
