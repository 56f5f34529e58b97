#!/usr/bin/env python3
"""Benchmark of the H2* fast matvec driving GMRES on one MI355X.

Workload (the configuration the metric is quoted on): 3D H2* with the repo's
RBF kernel (choice=1) on the reference's uniform grid 64^3 (N = 262144) in
[-1,1]^3, n_max = 64 (4 levels), ACA/NCA tolerance 1e-10, GMRES tolerance
1e-10, x_true ~ U(-1,1) (seed 2), b = A x_true from the GPU's dense FP64
product, x0 = 1 (as the reference's main).

A step is one GMRES solve to 1e-10 inside FP64 iterative refinement: residuals
on the stored FP64 operators, corrections from GMRES on the FP32 inner
operator (every box truncated to the leading pivots at which its cross met
1e-6; P2P/M2L in FP32 with FP32 accumulation), b and x resident in HBM.  The
plain FP64 solve on the stored FP64 operators, the refinement around the full
FP32 operators and around on-the-fly FP32 P2P/M2L are timed beside it.  Matvec bandwidth is the reference's stored
element count M (findMemory_n) times the element size plus the two N-vectors,
per HIP-event-timed matvec.  The CPU baseline is the oracle (a C++/OpenMP
restatement of the reference) on a bounded sample.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GMRES time-to-1e-10 (s) & matvec GB/s: FP32+refine batched vs FP64 GPU, 1 GPU"
HBM_PEAK_GBPS = 8000.0
FP32_PEAK_TFLOPS = 157.3
MFMA_PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}  # dense matrix-core peaks of MI355X
# flops of one 3D RBF entry evaluated on the fly (add/mul = 1, fma = 2, sqrt/log = 1):
# 3 differences, R^2 (mul + 2 fma), sqrt, log, scale, accumulate (fma)
FLOP_PER_ENTRY = 3 + 5 + 1 + 1 + 1 + 2
HEADLINE = "fp32_inner_ir"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


class Hip:
    """Device buffers through the HIP runtime (plumbing for device-resident inputs)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so.7")

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        return p

    def h2d(self, dst, arr):
        assert self.lib.hipMemcpy(dst, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes), 1) == 0

    def d2d(self, dst, src, nbytes):
        assert self.lib.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0

    def d2h(self, arr, src):
        assert self.lib.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), src, ctypes.c_size_t(arr.nbytes), 2) == 0

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def free(self, p):
        self.lib.hipFree(p)

    def copy_gbps(self, nbytes=4 << 30, reps=5):
        """HBM bandwidth of a plain device-to-device copy (read + write bytes / time), the
        achievable ceiling a streaming kernel is compared with."""
        src, dst = self.malloc(nbytes), self.malloc(nbytes)
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        assert self.lib.hipEventCreate(ctypes.byref(e0)) == 0 and self.lib.hipEventCreate(ctypes.byref(e1)) == 0
        n = ctypes.c_size_t(nbytes)
        assert self.lib.hipMemcpyAsync(dst, src, n, 3, None) == 0
        assert self.lib.hipEventRecord(e0, None) == 0
        for _ in range(reps):
            assert self.lib.hipMemcpyAsync(dst, src, n, 3, None) == 0
        assert self.lib.hipEventRecord(e1, None) == 0
        assert self.lib.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert self.lib.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        self.lib.hipEventDestroy(e0)
        self.lib.hipEventDestroy(e1)
        self.free(src)
        self.free(dst)
        return 2.0 * nbytes * reps / (ms.value * 1e-3) * 1e-9


def _omp_env():
    """OpenMP placement for the CPU baseline, as the reference's runs would use it: one thread per
    core, threads bound close to each other (set before the oracle's OpenMP runtime starts)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_solve(side, tol_pow, gtol, b=None, reps=3):
    """The oracle (C++/OpenMP restatement of the reference's CPU algorithm) on the 3D H2* RBF grid
    side^3: assembly, then GMRES to gtol from x0 = 1 timed with a wall clock around the solve only
    (as the reference brackets it, 1D/codes/main.cpp:37-40), median of reps solves."""
    _omp_env()
    N = side ** 3
    if N > 100000:
        os.environ.setdefault("H2W_ORACLE_PROGRESS", "1")  # (the oracle reports each finished level)
    import oracle
    t0 = time.time()
    o = oracle.Oracle(3, N, 64, 1.0, tol_pow, 1, 0)
    o.tree_initialisation()
    t_asm = time.time() - t0
    log(f"oracle {side}^3: assembly {t_asm:.1f} s")
    xt = np.random.default_rng(2).uniform(-1, 1, N)
    if b is None:
        b = o.dense_matvec(xt)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        x, it, resid, nmv = o.gmres(b, None, 500, gtol)
        ts.append(time.perf_counter() - t0)
        log(f"oracle {side}^3: GMRES {it} iterations in {ts[-1]:.3f} s")
    # one oracle matvec, for its bandwidth on the reference's storage count
    xr = np.random.default_rng(3).normal(size=N)
    t0 = time.perf_counter()
    o.mat_vec_prod(xr)
    t_mv = time.perf_counter() - t0
    err = float(np.linalg.norm(x - xt) / np.linalg.norm(xt))
    return {"N": N, "side": side, "assembly_s": round(t_asm, 2), "gmres_s": float(np.median(ts)),
            "gmres_runs_s": [round(v, 4) for v in ts], "iterations": it, "matvecs": nmv, "resid": resid,
            "rel_error": err, "memory": o.memory, "matvec_s": t_mv,
            "matvec_effective_GBps": round((8 * o.memory + 16 * N) / t_mv * 1e-9, 2),
            "threads": oracle.threads(), "cpu_model": _cpu_model(),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}


def gpu_solve_at(side, tol_pow, gtol, reps=3):
    """The GPU path on the same grid as a CPU sample: plain FP64 GMRES and the FP32-inner refinement
    (default inner tolerance, no tuning), b from the GPU's dense FP64 product, x0 = 1, median of reps."""
    import h2weak_amd as h2w
    N = side ** 3
    A = h2w.Initialisation(3, N, 64, 1.0, tol_pow, h2w.CHOICE_RBF, h2w.NESTED)
    A.tree_initialisation()
    xt = np.random.default_rng(2).uniform(-1, 1, N)
    b = A.dense_rows(None, xt)
    out = {}
    for name, mode in (("fp64", h2w.GMRES_FP64), ("fp32_inner_ir", h2w.GMRES_FP32_INNER_IR)):
        A.gmres(b, None, 500, gtol, mode)  # warm-up (graph capture)
        runs = [A.gmres(b, None, 500, gtol, mode) for _ in range(reps)]
        out[name] = {"s": float(np.median([r[1]["seconds"] for r in runs])), "iterations": runs[-1][1]["iterations"],
                     "rel_error": float(np.linalg.norm(runs[-1][0] - xt) / np.linalg.norm(xt))}
    return out, b


def cpu_baseline(side, tol_pow, gtol):
    """CPU baseline on a bounded sample of the workload: the oracle on the 3D H2* RBF grid side^3, its
    GMRES time measured (not scaled), beside the GPU's time on the same grid and the same b."""
    gpu, b = gpu_solve_at(side, tol_pow, gtol)
    c = cpu_solve(side, tol_pow, gtol, b=b)
    return {"value": round(c["gmres_s"], 6), "unit": "s", "cores": c["threads"], "kind": "port",
            "sample": (f"oracle (C++/OpenMP restatement of the reference) on 3D H2* RBF grid {side}^3 (N={c['N']}), "
                       f"n_max=64, tol 1e-{tol_pow}, same b as the GPU run on that grid: GMRES to {gtol:g} from x0 = 1, "
                       f"{c['iterations']} iterations, median of 3 solves; {c['threads']} OpenMP threads on "
                       f"{c['cpu_model']}, OMP_PROC_BIND={c['omp']['OMP_PROC_BIND']} OMP_PLACES={c['omp']['OMP_PLACES']}"),
            "sample_N": c["N"], "sample_assembly_s": c["assembly_s"], "sample_iterations": c["iterations"],
            "sample_rel_error": c["rel_error"], "sample_runs_s": c["gmres_runs_s"],
            "sample_matvec_effective_GBps": c["matvec_effective_GBps"], "cpu_model": c["cpu_model"],
            "gpu_same_sample": gpu,
            "speedup_gpu_fp32_inner_ir_same_sample": round(c["gmres_s"] / gpu["fp32_inner_ir"]["s"], 2),
            "speedup_gpu_fp64_same_sample": round(c["gmres_s"] / gpu["fp64"]["s"], 2)}


def pmc_reduce(d, side, tp):
    """HBM bytes per dispatch of the M2L + near-field launches from FETCH_SIZE / WRITE_SIZE (KiB per
    dispatch), each scaled so that the 1 GiB device copy of the probe reads and writes exactly 1 GiB."""
    import csv
    import re
    per = {}
    for tag in ("fetch", "write"):
        vals = {}
        for r in csv.DictReader(open(os.path.join(d, f"{tag}_counter_collection.csv"))):
            k = int(r["Dispatch_Id"])
            name, v = vals.get(k, (r["Kernel_Name"], 0.0))
            vals[k] = (name, v + float(r["Counter_Value"]) * 1024.0)
        per[tag] = vals
    copy = [k for k, (n, v) in per["write"].items() if "copyBuffer" in n]
    kc = max(copy, key=lambda k: per["write"][k][1])
    nb = float(1 << 30)
    f_fetch, f_write = nb / per["fetch"][kc][1], nb / per["write"][kc][1]
    out = {"calibration": {"copy_bytes": int(nb), "fetch_factor": round(f_fetch, 5), "write_factor": round(f_write, 5),
                           "note": "bytes = counter KiB * 1024 * factor; factors from the 1 GiB device copy"}}
    probe = None  # the probe's own line (launch bytes), from the logs of the two runs
    for fn in sorted(os.listdir(d)):
        if fn.endswith(".log"):
            for line in open(os.path.join(d, fn)):
                if line.startswith('{"pmc_probe"'):
                    probe = json.loads(line)
    for prec, key in (("double", "fp64_m2l_near"), ("float", "fp32inner_m2l_near"), ("sym", "fp64_16wide_m2l_near")):
        pat = re.compile(r"k_bgemm_sym<double>" if prec == "sym" else r"k_bgemv<%s, \d+, 1[,>]" % prec)
        ks = sorted(k for k, (n, v) in per["fetch"].items() if pat.search(n) and k in per["write"])
        if not ks:
            continue
        rd = [per["fetch"][k][1] * f_fetch for k in ks]
        wr = [per["write"][k][1] * f_write for k in ks]
        tot = sorted(a + b for a, b in zip(rd, wr))
        rec = {"dispatches": len(ks), "fetch_bytes": int(sorted(rd)[len(rd) // 2]),
               "write_bytes": int(sorted(wr)[len(wr) // 2]), "hbm_bytes_per_launch": int(tot[len(tot) // 2])}
        if probe:
            tag, es = ("fp32_inner", 4) if prec == "float" else ("fp64", 8)
            q = probe[f"launch_role_{tag}"].index(1)
            # the bench's algorithmic bytes: the stored blocks once, x and y once (16 columns each when 16-wide)
            alg = probe[f"launch_bytes_{tag}"][q] + 2 * es * probe["N"] * (16 if prec == "sym" else 1)
            rec["algorithmic_bytes"] = int(alg)
            rec["traffic_over_algorithmic"] = round(rec["hbm_bytes_per_launch"] / alg, 4)
        out[f"3d_rbf_{side}_{tp}_{key}"] = rec
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    json.dump(out, open(path, "w"), indent=1)
    print(json.dumps(out))


def kernel_table(path, top=25):
    """Markdown table of a rocprofv3 --stats kernel summary (<prefix>_kernel_stats.csv), top kernels by
    total time, names shortened to the template-id (no GPU)."""
    import csv
    rows = []
    for r in csv.DictReader(open(path)):
        name = r["Name"].replace("(anonymous namespace)::", "")
        depth, cut = 0, len(name)
        for i, ch in enumerate(name):  # the argument list: the first '(' outside template brackets
            depth += ch == "<"
            depth -= ch == ">"
            if ch == "(" and depth == 0:
                cut = i
                break
        name = name[:cut].replace("void ", "").replace("h2w::", "").strip()
        rows.append((float(r["TotalDurationNs"]), int(r["Calls"]), float(r["AverageNs"]), float(r["Percentage"]), name))
    rows.sort(reverse=True)
    print("| total ms | calls | average us | % | kernel |")
    print("|---:|---:|---:|---:|---|")
    for tot, calls, avg, pct, name in rows[:top]:
        print(f"| {tot * 1e-6:.1f} | {calls} | {avg * 1e-3:.1f} | {pct:.1f} | `{name}` |")


def bench_batch(S=32, tol_pow=8, seed=5, defl=((128, 48), (128, 64)), tune_inner=(6.0, 5.5, 5.0, 4.5)):
    """32 independent 3D H2* RBF systems, N log-uniform in [4096, 65536] (seed 5), points U([-1,1]^3),
    n_max 64, ACA and GMRES tolerance 1e-8, 16 right-hand sides each (X_true ~ N(0,1), B = A X_true):
    batched products (one launch per phase for all systems, 16 right-hand sides per pass) and lockstep
    solves of all 512 columns, against plain FP64 solves one right-hand side at a time.  The refinement's
    inner tolerance and deflation spaces are chosen on held-out right-hand sides before the timed solves."""
    import h2weak_amd as h2w
    rng = np.random.default_rng(seed)
    Ns = [int(v) for v in np.rint(np.exp(rng.uniform(np.log(4096), np.log(65536), S)))]
    gtol = 10.0 ** (-tol_pow)
    t0 = time.time()
    systems = []
    for q, N in enumerate(Ns):
        P = np.random.default_rng(100 + q).uniform(-1.0, 1.0, (3, N))
        systems.append(h2w.Initialisation(3, N, 64, 1.0, tol_pow, h2w.CHOICE_RBF, h2w.NESTED, points=P))
    h2w.assemble_many(systems, 4)  # four host threads, one stream per system
    t_asm = time.time() - t0
    log(f"batch: {S} systems, N {min(Ns)}..{max(Ns)} (sum {sum(Ns)}), assembly {t_asm:.1f} s (4 threads)")
    B = h2w.Batch(systems)
    out = {"systems": S, "N": Ns, "nrhs": 16, "tol": gtol, "assembly_s": round(t_asm, 2), "assembly_threads": 4,
           "nlevels": sorted({g.stats()["nlevels"] for g in systems}), "matvec": {}, "gmres": {}}
    sts = [g.stats() for g in systems]
    for prec, name, key, s_op in ((h2w.FP64, "fp64", "stored_bytes_fp64", 8), (h2w.FP32, "fp32", "stored_bytes_fp32", 4),
                                  (h2w.FP32_INNER, "fp32_inner", "stored_bytes_inner_fp32", 4)):
        r = B.bench_matvec(5, prec)
        # every launch's operator bytes (a separate transposed launch reads the symmetric blocks again)
        nbytes = sum(r["launch_bytes"]) + 2 * s_op * 16 * sum(Ns)
        out["matvec"][name] = {"ms_per_16": round(r["ms_per_pass"], 4), "launches": len(r["launch_role"]),
                               "streamed_GBps": round(nbytes / (r["ms_per_pass"] * 1e-3) * 1e-9, 1)}
        log(f"batch 16-wide product {name}: {r['ms_per_pass']:.3f} ms, {len(r['launch_role'])} launches")
    Xt = [np.random.default_rng(200 + q).normal(size=(16, g.N)) for q, g in enumerate(systems)]
    Bs = B.matvec(Xt)
    def solves(name, mode):
        B.gmres([b[:1] for b in Bs], None, 300, gtol, mode)  # warm-up (graph capture)
        Xs, st = B.gmres(Bs, None, 300, gtol, mode)
        cols = [c for row in st for c in row]
        err = max(float(np.linalg.norm(Xs[q][c] - Xt[q][c]) / np.linalg.norm(Xt[q][c]))
                  for q in range(S) for c in range(16))
        out["gmres"][name] = {"s_512": round(cols[0]["seconds"], 4), "ms_per_solve": round(cols[0]["seconds"] / 512 * 1e3, 4),
                              "all_converged": all(c["converged"] for c in cols),
                              "max_iterations": max(c["iterations"] for c in cols), "max_rel_error": err}
        log(f"batch solves {name}: {cols[0]['seconds'] * 1e3:.1f} ms for 512 columns, max error {err:.2e}")

    solves("fp64", h2w.GMRES_FP64)
    solves("fp32_ir", h2w.GMRES_FP32_IR)
    if defl:
        solves("fp32_inner_ir_no_deflation", h2w.GMRES_FP32_INNER_IR)
        # the inner tolerance and every system's deflation space for its inner operator, chosen by timing
        # lockstep solves of held-out right-hand sides (X ~ N(0,1), seeds 300 + q), never the timed ones
        t0 = time.time()
        Bh = B.matvec([np.random.default_rng(300 + q).normal(size=(16, g.N)) for q, g in enumerate(systems)])
        tune = B.autotune(Bh, gtol, 300, tune_inner, defl)
        cands = [{"inner_tol_exp": c["inner_tol_exp"], "m": c["m"], "k": c["k"], "s_512": round(c["seconds"], 4),
                  "max_iterations": c["max_iterations"], "converged": c["converged"]} for c in tune["candidates"]]
        out["deflation"] = {"m": tune["m"], "k": tune["k"], "inner_tol_exp": tune["inner_tol_exp"],
                            "tune_s": round(time.time() - t0, 2), "tuned_on": "held-out right-hand sides (seeds 300 + q)",
                            "candidates": cands}
        log(f"batch: inner 10^-{tune['inner_tol_exp']}, spaces ({tune['m']}, {tune['k']}) chosen in "
            f"{out['deflation']['tune_s']} s from " + ", ".join(f"10^-{c['inner_tol_exp']} ({c['m']},{c['k']}): "
                                                                f"{c['s_512'] * 1e3:.1f} ms" for c in cands))
    solves("fp32_inner_ir", h2w.GMRES_FP32_INNER_IR)
    # plain FP64, one right-hand side at a time: every 4th system's first column, scaled to 512 solves
    ts = []
    for q in range(0, S, 4):
        _, s1 = systems[q].gmres(Bs[q][0], None, 300, gtol, h2w.GMRES_FP64)
        ts.append(s1["seconds"])
    seq = float(np.mean(ts)) * 512
    out["sequential_fp64_single_rhs_s"] = round(seq, 3)
    out["sequential_sample"] = "mean of one single-RHS FP64 solve on every 4th system, x512"
    out["speedup_fp32_inner_ir_vs_sequential_fp64"] = round(seq / out["gmres"]["fp32_inner_ir"]["s_512"], 2)
    out["speedup_fp32_inner_ir_vs_batched_fp64"] = round(out["gmres"]["fp64"]["s_512"]
                                                         / out["gmres"]["fp32_inner_ir"]["s_512"], 3)
    log(f"batch: sequential FP64 estimate {seq:.2f} s; batched FP32-inner IR x{out['speedup_fp32_inner_ir_vs_sequential_fp64']}")
    del B, systems
    return out


# The other benchmark configurations: (dim, variant, choice, N, n_max, tol_pow, gmres tol, points, x_true, modes)
CONFIGS = {
    "c1_1d_rbf_65536": dict(dim=1, variant=0, choice=1, N=65536, n_max=100, tol_pow=12, gtol=1e-12,
                            points="U(-1,1) seed 1", xt="U(-1,1) seed 2"),
    "c2_1d_ie_250000": dict(dim=1, variant=0, choice=0, N=250000, n_max=100, tol_pow=12, gtol=1e-12,
                            points="U(-1,1) seed 1", xt="U(-1,1) seed 2"),
    "c2_1d_rbf_250000": dict(dim=1, variant=0, choice=1, N=250000, n_max=100, tol_pow=12, gtol=1e-12,
                             points="U(-1,1) seed 1", xt="U(-1,1) seed 2"),
    "c3_2d_semi_ie_400x400": dict(dim=2, variant=1, choice=0, N=160000, n_max=64, tol_pow=10, gtol=1e-10,
                                  points="grid", xt="N(0,1) seed 3"),
}


def bench_configs(names=None, cpu=True):
    """The 1D and 2D benchmark configurations at their full sizes: on the GPU the FP64 product's
    bandwidth, plain FP64 GMRES and the FP32-inner refinement; beside them the oracle (the reference's
    CPU algorithm) on the host cores, its FP64 GMRES timed on the same b.  b = A x_true from the GPU's
    dense FP64 product; x0 = 1; errors against x_true, and 1024 sampled product rows against the dense
    FP64 rows."""
    import h2weak_amd as h2w
    out = {}
    for name, c in CONFIGS.items():
        if names and name not in names:
            continue
        N, dim = c["N"], c["dim"]
        P = None if c["points"] == "grid" else np.random.default_rng(1).uniform(-1, 1, (dim, N))
        t0 = time.time()
        A = h2w.Initialisation(dim, N, c["n_max"], 1.0, c["tol_pow"], c["choice"], c["variant"], points=P)
        A.tree_initialisation()
        t_asm = time.time() - t0
        st = A.stats()
        xt = (np.random.default_rng(3).normal(size=N) if c["xt"].startswith("N(")
              else np.random.default_rng(2).uniform(-1.0, 1.0, N))
        b = A.dense_rows(None, xt)
        rows = np.sort(np.random.default_rng(4).choice(N, 1024, replace=False))
        y = A.mat_vec_prod(xt)
        row_err = float(np.linalg.norm(y[rows] - b[rows]) / np.linalg.norm(b[rows]))
        r = A.bench_matvec(10, h2w.FP64)
        streamed = sum(r["launch_bytes"]) + 2 * 8 * N
        row = {"config": c, "nlevels": st["nlevels"], "gpu_assembly_s": round(t_asm, 3), "memory": st["memory"],
               "sampled_rows_rel_error": row_err,
               "gpu_matvec_fp64": {"ms": round(r["ms_per_matvec"], 4),
                                   "streamed_GBps": round(streamed / (r["ms_per_matvec"] * 1e-3) * 1e-9, 1),
                                   "hbm_frac": round(streamed / (r["ms_per_matvec"] * 1e-3) * 1e-9 / HBM_PEAK_GBPS, 4),
                                   "effective_GBps": round((8 * st["memory"] + 16 * N) / (r["ms_per_matvec"] * 1e-3) * 1e-9, 1)},
               "gpu_gmres": {}}
        for mode, nm in ((h2w.GMRES_FP64, "fp64"), (h2w.GMRES_FP32_INNER_IR, "fp32_inner_ir")):
            A.gmres(b, None, 500, c["gtol"], mode)  # warm-up (graph capture)
            runs = [A.gmres(b, None, 500, c["gtol"], mode) for _ in range(3)]
            x, gs = runs[-1]
            row["gpu_gmres"][nm] = {"s": float(np.median([q[1]["seconds"] for q in runs])),
                                    "iterations": gs["iterations"], "resid": gs["resid"],
                                    "rel_error": float(np.linalg.norm(x - xt) / np.linalg.norm(xt))}
        log(f"{name}: GPU assembly {t_asm:.2f} s, matvec {row['gpu_matvec_fp64']}, GMRES {row['gpu_gmres']}")
        del A
        if cpu:
            _omp_env()
            import oracle
            t0 = time.time()
            o = oracle.Oracle(dim, N, c["n_max"], 1.0, c["tol_pow"], c["choice"], c["variant"], points=P)
            o.tree_initialisation()
            t_oasm = time.time() - t0
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                x, it, resid, nmv = o.gmres(b, None, 500, c["gtol"])
                ts.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            o.mat_vec_prod(xt)
            t_mv = time.perf_counter() - t0
            row["cpu_oracle"] = {"assembly_s": round(t_oasm, 2), "gmres_s": float(np.median(ts)),
                                 "gmres_runs_s": [round(v, 4) for v in ts], "iterations": it, "resid": resid,
                                 "rel_error": float(np.linalg.norm(x - xt) / np.linalg.norm(xt)),
                                 "matvec_s": round(t_mv, 4),
                                 "matvec_effective_GBps": round((8 * o.memory + 16 * N) / t_mv * 1e-9, 2),
                                 "memory": o.memory, "threads": oracle.threads(), "cpu_model": _cpu_model(),
                                 "omp": {k: os.environ.get(k) for k in ("OMP_PROC_BIND", "OMP_PLACES")}}
            row["speedup_gpu_fp64_vs_cpu"] = round(row["cpu_oracle"]["gmres_s"] / row["gpu_gmres"]["fp64"]["s"], 1)
            row["speedup_gpu_fp32_inner_ir_vs_cpu"] = round(row["cpu_oracle"]["gmres_s"]
                                                            / row["gpu_gmres"]["fp32_inner_ir"]["s"], 1)
            log(f"{name}: CPU oracle {row['cpu_oracle']}")
            del o
        out[name] = row
    print(json.dumps({"configs": out}), flush=True)


def bench_complex():
    """The complex kernels (the reference's -DUSE_Helm / -DUSE_Hankel builds): 3D Helmholtz e^{iR}/(4 pi) on
    the 32^3 grid and 2D Hankel (i/4) H0^(1) on the 256^2 grid, H2*, n_max 64, ACA tol 1e-8, FP64 complex GMRES
    to 1e-8 from x0 = 1 with b = A x_true from the GPU's dense complex product (x_true complex N(0,1), seed 3):
    assembly, the product's streamed bandwidth (16 bytes per stored complex element), GMRES time and error,
    and 1024 sampled rows of the product against the dense rows."""
    import h2weak_amd as h2w
    out = {}
    for name, dim, side, choice in (("helmholtz_3d_grid32", 3, 32, h2w.CHOICE_HELMHOLTZ),
                                    ("hankel_2d_grid256", 2, 256, h2w.CHOICE_HANKEL)):
        N = side ** dim
        t0 = time.time()
        A = h2w.Initialisation(dim, N, 64, 1.0, 8, choice, h2w.NESTED)
        A.tree_initialisation()
        t_asm = time.time() - t0
        st = A.stats()
        rng = np.random.default_rng(3)
        xt = rng.normal(size=N) + 1j * rng.normal(size=N)
        b = A.dense_rows(None, xt)
        rows = np.sort(np.random.default_rng(4).choice(N, 1024, replace=False))
        y = A.mat_vec_prod(xt)
        row_err = float(np.linalg.norm(y[rows] - b[rows]) / np.linalg.norm(b[rows]))
        r = A.bench_matvec(10, h2w.FP64)
        streamed = sum(r["launch_bytes"]) + 2 * 16 * N
        q = r["launch_role"].index(1)
        A.gmres(b, None, 500, 1e-8, h2w.GMRES_FP64)  # warm-up
        runs = [A.gmres(b, None, 500, 1e-8, h2w.GMRES_FP64) for _ in range(3)]
        x, gs = runs[-1]
        out[name] = {"N": N, "dim": dim, "choice": choice, "n_max": 64, "tol": 1e-8, "nlevels": st["nlevels"],
                     "assembly_s": round(t_asm, 2), "memory": st["memory"],
                     "stored_bytes_complex": st["stored_bytes_fp64"],
                     "sampled_rows_rel_error": row_err,
                     "matvec": {"ms": round(r["ms_per_matvec"], 4),
                                "streamed_GBps": round(streamed / (r["ms_per_matvec"] * 1e-3) * 1e-9, 1),
                                "m2l_near_launch_ms": round(r["launch_ms"][q], 4),
                                "m2l_near_launch_GBps": round(r["launch_bytes"][q] / (r["launch_ms"][q] * 1e-3) * 1e-9, 1)},
                     "gmres_fp64": {"s": float(np.median([g[1]["seconds"] for g in runs])),
                                    "iterations": gs["iterations"], "resid": gs["resid"],
                                    "rel_error": float(np.linalg.norm(x - xt) / np.linalg.norm(xt))}}
        log(f"{name}: {out[name]}")
        del A
    print(json.dumps({"complex": out}), flush=True)


# The reference's eight published runs (uniform grids, x0 = 1): parameters, and the printed iteration count,
# storage (GB) and wall times of assembly and GMRES on its CPU (hardware unstated).
TRANSCRIPTS = {
    "1d_nested": dict(dim=1, variant=0, N=250000, n_max=100, tol_pow=12, choice=1, gtol=1e-12, max_iter=500,
                      iters=11, storage_gb=0.658597, asm_s=47.0822, gmres_s=1.49966, src="1D/results/nhodlr.txt"),
    "1d_semi": dict(dim=1, variant=1, N=250000, n_max=100, tol_pow=12, choice=1, gtol=1e-12, max_iter=500,
                    iters=11, storage_gb=1.61, asm_s=26.4311, gmres_s=2.37944, src="1D/results/sn_hodlr.txt"),
    "2d_nested": dict(dim=2, variant=0, N=90000, n_max=400, tol_pow=12, choice=1, gtol=1e-12, max_iter=500,
                      iters=28, storage_gb=2.47347, asm_s=68.1836, gmres_s=4.57271, src="2D/results/nhodlr.txt"),
    "2d_semi": dict(dim=2, variant=1, N=90000, n_max=400, tol_pow=12, choice=1, gtol=1e-12, max_iter=500,
                    iters=28, storage_gb=3.03009, asm_s=37.1335, gmres_s=4.75639, src="2D/results/sn_hodlr.txt"),
    "2d_hstar": dict(dim=2, variant=2, N=90000, n_max=400, tol_pow=12, choice=1, gtol=1e-12, max_iter=500,
                     iters=28, storage_gb=4.17922, asm_s=39.9024, gmres_s=7.14276, src="2D/results/hodlr.txt"),
    "3d_nested": dict(dim=3, variant=0, N=42875, n_max=512, tol_pow=7, choice=0, gtol=1e-10, max_iter=1000,
                      iters=5, storage_gb=2.87549, asm_s=50.4586, gmres_s=1.59765, src="3D/results/nhodlr.txt"),
    "3d_semi": dict(dim=3, variant=1, N=42875, n_max=512, tol_pow=7, choice=0, gtol=1e-10, max_iter=1000,
                    iters=5, storage_gb=3.02412, asm_s=55.2427, gmres_s=2.11972, src="3D/results/sn_hodlr.txt"),
    "3d_hstar": dict(dim=3, variant=2, N=42875, n_max=512, tol_pow=7, choice=0, gtol=1e-10, max_iter=1000,
                     iters=5, storage_gb=4.39326, asm_s=129.971, gmres_s=3.54523, src="3D/results/hodlr.txt"),
}


def bench_transcripts():
    """The reference's published runs on the GPU: assembly and GMRES wall time beside the published CPU
    times, with the storage count, iterations and solution error of each."""
    import h2weak_amd as h2w
    out = {}
    for name, c in TRANSCRIPTS.items():
        t0 = time.time()
        A = h2w.Initialisation(c["dim"], c["N"], c["n_max"], 1.0, c["tol_pow"], c["choice"], c["variant"])
        A.tree_initialisation()
        t_asm = time.time() - t0
        st = A.stats()
        xt = np.random.default_rng(2).uniform(-1.0, 1.0, c["N"])
        b = A.dense_rows(None, xt)
        row = {"nlevels": st["nlevels"], "storage_gb": st["memory"] * 8e-9, "published_storage_gb": c["storage_gb"],
               "assembly_s": round(t_asm, 3), "published_assembly_s": c["asm_s"],
               "assembly_speedup": round(c["asm_s"] / t_asm, 1), "source": c["src"]}
        for mode, nm in ((h2w.GMRES_FP64, "fp64"), (h2w.GMRES_FP32_INNER_IR, "fp32_inner_ir")):
            A.gmres(b, None, c["max_iter"], c["gtol"], mode)  # warm-up (graph capture)
            x, gs = A.gmres(b, None, c["max_iter"], c["gtol"], mode)
            err = float(np.linalg.norm(x - xt) / np.linalg.norm(xt))
            row[nm] = {"s": round(gs["seconds"], 5), "iterations": gs["iterations"], "resid": gs["resid"],
                       "rel_error": err, "converged": bool(gs["converged"]),
                       "speedup_vs_published": round(c["gmres_s"] / gs["seconds"], 1)}
        row["published_iterations"] = c["iters"]
        row["published_gmres_s"] = c["gmres_s"]
        log(f"{name}: storage {row['storage_gb']:.6g} GB (published {c['storage_gb']}), assembly {t_asm:.2f} s "
            f"(published {c['asm_s']} s), GMRES FP64 {row['fp64']['s'] * 1e3:.1f} ms / {row['fp64']['iterations']} "
            f"its (published {c['gmres_s']} s / {c['iters']}), FP32-inner IR {row['fp32_inner_ir']['s'] * 1e3:.1f} ms")
        out[name] = row
        del A
    print(json.dumps({"reference_transcripts_on_gpu": out}), flush=True)


def bench_c5(N=262144, tol_pow=10, seed=7):
    """Learned rank predictor at N = 262144, 3D RBF, tol 1e-10, uniform and Gaussian (sigma 0.2) points:
    assembly with the heuristic workspace estimates and with the shipped model, and the model's accuracy on
    every ACA block (these two point sets are not in its training harvest)."""
    import h2weak_amd as h2w
    from h2weak_amd import rank_mlp
    params, meta = rank_mlp.load(rank_mlp.default_model_path(3, 1, tol_pow))
    out = {"workload": f"3D H2* RBF N={N} n_max=64 tol 1e-{tol_pow}, uniform and gauss(0.2) points (seed {seed})",
           "model": {k: meta[k] for k in ("features", "hidden", "activation", "params", "train_samples", "train_sets")},
           "sets": {}}
    for kind in ("uniform", "gauss"):
        P = rank_mlp.points(kind, N, seed)
        row = {}
        mem = None
        for use in (False, True):
            g = h2w.Initialisation(3, N, 64, 1.0, tol_pow, h2w.CHOICE_RBF, h2w.NESTED, points=P)
            if use:
                g.set_rank_model(params)
            g.tree_initialisation()
            st = g.stats()
            mem = st["memory"] if mem is None else mem
            assert st["memory"] == mem  # the model never changes the operator
            r = {"assembly_s": round(st["assembly_seconds"], 2), "aca_batches": st["aca_batches"],
                 "aca_reruns": st["aca_reruns"], "aca_pool_GB": round(st["aca_pool_bytes"] * 1e-9, 2)}
            if use:
                F, R = g.rank_samples()
                pl = g.rank_predict(F)
                est = np.ceil(1.2 * np.maximum(0, np.exp2(pl) - 1)) + 4
                r.update({"blocks": int(len(R)), "mae_log2": round(float(np.mean(np.abs(pl - np.log2(1 + R)))), 4),
                          "under_estimates": st["rank_pred_under"],
                          "reserved_over_actual": round(float(np.sum(est) / max(1, np.sum(R))), 3),
                          "fixed_rank_blocks": st["rank_fixed_blocks"],
                          "fixed_rank_fallback_resumed": st["rank_fixed_resumed"],
                          "fixed_rank_extra_steps": st["rank_fixed_extra_steps"],
                          "fixed_rank_sampled_rel_err_max": st["rank_fixed_sampled_err"]})
            row["model" if use else "heuristic"] = r
            log(f"C5 {kind} {'model' if use else 'heuristic'}: {r}")
            del g
        out["sets"][kind] = row
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=64, help="grid points per axis (N = side^3)")
    ap.add_argument("--tol-pow", type=int, default=10)
    ap.add_argument("--cpu-side", type=int, default=32, help="grid side of the CPU-baseline sample")
    ap.add_argument("--cpu-full", action="store_true",
                    help="only the CPU baseline at the full workload (the oracle at --side^3, several minutes); "
                         "prints its JSON line (kept under profiles/ and cited by the default run)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-multi", action="store_true", help="skip the 16-right-hand-side section")
    ap.add_argument("--no-batch", action="store_true", help="skip the 32-system batch section")
    ap.add_argument("--no-rank-model", action="store_true", help="assemble with heuristic workspace estimates")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default launch shapes, inner 1e-6)")
    ap.add_argument("--no-deflation", action="store_true", help="tune without the refinement's deflation space")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="assemble, then run 3 FP64 and 3 FP32-inner matvecs and a 1 GiB device copy "
                         "(the calibration kernel) for rocprofv3 --pmc; prints what it ran")
    ap.add_argument("--inner-tol", type=float, default=6.5,
                    help="inner-operator tolerance exponent of --pmc-probe (the tuner's usual choice at 64^3)")
    ap.add_argument("--pmc-reduce", metavar="DIR",
                    help="read DIR/fetch_counter_collection.csv and DIR/write_counter_collection.csv written by "
                         "two rocprofv3 --pmc runs of --pmc-probe and write profiles/pmc_traffic.json (no GPU)")
    ap.add_argument("--assembly-only", action="store_true",
                    help="assemble the workload once (heuristic rank estimates) and print the time")
    ap.add_argument("--c5", action="store_true",
                    help="rank-predictor measurement (N = 262144 uniform and Gaussian, heuristic vs model)")
    ap.add_argument("--complex", action="store_true", help="the complex kernels' measurement lines only")
    ap.add_argument("--configs", action="store_true",
                    help="the 1D / 2D benchmark configurations at full size, GPU and CPU oracle side by side")
    ap.add_argument("--transcripts", action="store_true",
                    help="the reference's eight published runs on the GPU, beside the published CPU times")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 32-system batch section (16 right-hand sides each, tol 1e-8); prints its JSON")
    ap.add_argument("--kernel-table", metavar="CSV",
                    help="print a rocprofv3 --stats kernel summary CSV as a markdown table (no GPU)")
    args = ap.parse_args()
    if args.kernel_table:
        kernel_table(args.kernel_table)
        return
    if args.assembly_only:
        import h2weak_amd as h2w
        N = args.side ** 3
        A = h2w.Initialisation(3, N, 64, 1.0, args.tol_pow, h2w.CHOICE_RBF, h2w.NESTED)
        t0 = time.time()
        A.tree_initialisation()
        st = A.stats()
        print(json.dumps({"workload": f"3D H2* RBF grid {args.side}^3 tol 1e-{args.tol_pow}",
                          "assembly_s": round(time.time() - t0, 3), "memory": st["memory"],
                          "max_rank_far": st["max_rank_far"], "max_rank_vertex": st["max_rank_vertex"]}), flush=True)
        return
    if args.transcripts:
        bench_transcripts()
        return
    if args.configs:
        bench_configs(cpu=not args.no_cpu)
        return
    if args.complex:
        bench_complex()
        return
    if args.c5:
        bench_c5()
        return
    if args.batch_only:
        print(json.dumps({"batch_c4": bench_batch()}), flush=True)
        return
    if args.pmc_reduce:
        pmc_reduce(args.pmc_reduce, args.side, args.tol_pow)
        return
    if args.cpu_full:
        import h2weak_amd as h2w
        N = args.side ** 3
        G = h2w.Initialisation(3, N, 64, 1.0, args.tol_pow, h2w.CHOICE_RBF, h2w.NESTED)  # (dense product only)
        b = G.dense_rows(None, np.random.default_rng(2).uniform(-1, 1, N))
        del G
        c = cpu_solve(args.side, args.tol_pow, 10.0 ** -args.tol_pow, b=b)
        print(json.dumps({"cpu_full": c, "workload": f"3D H2* RBF grid {args.side}^3 N={N} n_max=64 tol 1e-{args.tol_pow}",
                          "command": "python bench.py --cpu-full", "b": "GPU dense FP64 product of x_true ~ U(-1,1) seed 2"}),
              flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); sharding the operator by "
              "subtree over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    import h2weak_amd as h2w

    side, tp = args.side, args.tol_pow
    N = side ** 3
    gtol = 10.0 ** (-tp)
    log(f"workload: 3D H2* RBF grid {side}^3 N={N} n_max=64 tol 1e-{tp}")
    t0 = time.time()
    A = h2w.Initialisation(3, N, 64, 1.0, tp, h2w.CHOICE_RBF, h2w.NESTED)
    from h2weak_amd import rank_mlp
    mpath = None if args.no_rank_model else rank_mlp.default_model_path(3, 1, tp)
    if mpath:
        A.set_rank_model(rank_mlp.load(mpath)[0])  # workspace estimates only; the operator is unchanged
    A.tree_initialisation()
    st = A.stats()
    log(f"assembly {time.time() - t0:.1f} s, M = {st['memory']:.4e} elements ({st['memory'] * 8e-9:.2f} GB fp64)")
    if args.pmc_probe:
        hip = Hip()
        A.set_inner_tolerance(args.inner_tol)
        A.bench_matvec(3, h2w.FP64)
        A.bench_matvec(3, h2w.FP32_INNER)
        A.bench_matvec_multi(2, h2w.FP64)  # the 16-wide fused launch (k_bgemm_sym)
        nb = 1 << 30
        src, dst = hip.malloc(nb), hip.malloc(nb)
        hip.d2d(dst, src, nb)
        hip.sync()
        r64, r32 = A.bench_matvec(1, h2w.FP64), A.bench_matvec(1, h2w.FP32_INNER)
        print(json.dumps({"pmc_probe": True, "copy_bytes": nb, "workload": f"3D RBF grid {side}^3 tol 1e-{tp}",
                          "inner_tol_exp": args.inner_tol,
                          "N": N, "launch_bytes_fp64": r64["launch_bytes"], "launch_role_fp64": r64["launch_role"],
                          "launch_bytes_fp32_inner": r32["launch_bytes"],
                          "launch_role_fp32_inner": r32["launch_role"]}), flush=True)
        return
    xt = np.random.default_rng(2).uniform(-1.0, 1.0, N)
    b = A.dense_rows(None, xt)
    log("b = A x_true from the GPU dense FP64 product")
    tune = None
    if not args.no_tune:
        # the tuner solves a right-hand side of its own (x_tune ~ U(-1,1), seed 12), never the timed b
        b_tune = A.dense_rows(None, np.random.default_rng(12).uniform(-1.0, 1.0, N))
        t0 = time.time()
        flags = h2w.TUNE_UNROLL | h2w.TUNE_INNER | (0 if args.no_deflation else h2w.TUNE_DEFLATION)
        tune = A.autotune(b_tune, tol=gtol, max_iter=500, flags=flags)
        tune["seconds"] = round(time.time() - t0, 2)
        tune["rhs"] = "b_tune = A x_tune, x_tune ~ U(-1,1) seed 12 (the timed b uses seed 2)"
        log(f"autotune {time.time() - t0:.1f} s: inner tolerance 10^-{tune['ir_tol_exp']:g}; "
            + ", ".join(f"10^-{c['ir_tol_exp']:g}: {c['seconds'] * 1e3:.1f} ms ({c['iterations']}+{c['outer']})"
                        for c in tune["candidates"])
            + f"; matvec ms FP64/FP32/inner {['%.3f' % v for v in tune['matvec_ms_before']]} -> "
            f"{['%.3f' % v for v in tune['matvec_ms_after']]}")
        d = tune["deflation"]
        log(f"deflation space m={d['m']} k={d['k']} (build {d['build_seconds'] * 1e3:.0f} ms); "
            + ", ".join(f"({c['m']},{c['k']}) 10^-{c['ir_tol_exp']:g}: {c['seconds'] * 1e3:.1f} ms "
                        f"({c['iterations']}+{c['outer']})" for c in d["candidates"]))

    hip = Hip()
    db, dx = hip.malloc(8 * N), hip.malloc(8 * N)
    hip.h2d(db, b)
    ones = np.ones(N)
    dones = hip.malloc(8 * N)
    hip.h2d(dones, ones)
    L = h2w.lib()

    def solve(mode):
        hip.d2d(dx, dones, 8 * N)
        s = h2w.GmresStats()
        rc = L.h2w_gmres(A._h, db, dx, 500, gtol, mode, h2w.DEVICE, ctypes.byref(s))
        assert rc == h2w.H2W_OK, (rc, L.h2w_last_error())
        return s.as_dict()

    modes = {"fp32_inner_ir": h2w.GMRES_FP32_INNER_IR, "fp32_ir": h2w.GMRES_FP32_IR,
             "fp32_otf_ir": h2w.GMRES_FP32_OTF_IR, "fp64": h2w.GMRES_FP64}
    for _ in range(args.warmup):
        for m in modes.values():
            solve(m)
    res = {}

    def timed(name, m, k):
        hip.sync()
        t0 = time.perf_counter()
        for _ in range(k):
            st_ = solve(m)
        hip.sync()
        t = (time.perf_counter() - t0) / k
        xs = np.empty(N)
        hip.d2h(xs, dx)
        res[name] = {"s": t, "stats": st_, "err": float(np.linalg.norm(xs - xt) / np.linalg.norm(xt)), "runs": k}
        log(f"{name}: {t * 1e3:.2f} ms/solve, error {res[name]['err']:.2e}, {st_}")

    for name, m in modes.items():
        timed(name, m, args.steps if name == HEADLINE else max(1, min(args.steps, 3)))
    defl = tune["deflation"] if tune else {"m": 0, "k": 0}
    if defl["m"] > 0:
        # the same refinement without the deflation space, then the space rebuilt for the later sections
        A.set_deflation(0)
        solve(modes[HEADLINE])
        timed(HEADLINE + "_no_deflation", modes[HEADLINE], max(1, min(args.steps, 3)))
        A.set_deflation(defl["m"], defl["k"])
    t_step = res[HEADLINE]["s"]
    t64 = res["fp64"]["s"]

    M = st["memory"]
    mv = {}
    kinfo = {}
    for prec, name, s_op in ((h2w.FP64, "fp64", 8), (h2w.FP32, "fp32", 4), (h2w.FP32_INNER, "fp32_inner", 4),
                             (h2w.FP32_OTF, "fp32_otf", 4), (h2w.FP64_OTF, "fp64_otf", 8)):
        r = A.bench_matvec(10, prec)
        t = r["ms_per_matvec"] * 1e-3
        # streamed: every launch's operator bytes; effective: the reference's
        # storage count M (both directions of every pair) over the same time
        streamed = sum(r["launch_bytes"]) + 2 * 8 * N
        mv[name] = {"ms": r["ms_per_matvec"],
                    "streamed_GBps": None if "otf" in name else round(streamed / t * 1e-9, 1),
                    "effective_GBps": round((s_op * M + 2 * 8 * N) / t * 1e-9, 1),
                    "launch_ms": [round(v, 4) for v in r["launch_ms"]],
                    "launch_GB": [round(v * 1e-9, 4) for v in r["launch_bytes"]],
                    "launch_role": r["launch_role"]}
        q = r["launch_role"].index(1)
        kinfo[name] = (r["launch_ms"][q], r["launch_bytes"][q] // s_op)  # (ms, entries) of the M2L + near launch
        log(f"matvec {name}: {r['ms_per_matvec']:.3f} ms ({mv[name]['streamed_GBps']} GB/s streamed, "
            f"{mv[name]['effective_GBps']:.0f} GB/s effective); M2L+near launch {r['launch_ms'][q]:.3f} ms")
    # dominant kernel of the measured solve: the inner FP32 M2L + near launches (truncated operator)
    # against the outer FP64 stored M2L + near launches
    so = res[HEADLINE]["stats"]
    t_in = so["matvecs_fp32"] * kinfo["fp32_inner"][0]
    t_st = so["matvecs_fp64"] * kinfo["fp64"][0]
    ms_i, ent_i = kinfo["fp32_inner"]
    nb_i = 4 * ent_i + 2 * 4 * N
    roof_in = {"bound": "hbm", "achieved": round(nb_i / (ms_i * 1e-3) * 1e-9, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
               "frac": round(nb_i / (ms_i * 1e-3) * 1e-9 / HBM_PEAK_GBPS, 4), "traffic": None,
               "kernel": "k_bgemv<float, 1> (M2L + near field of the truncated FP32 inner operator)",
               "launch_ms": round(ms_i, 4), "algorithmic_bytes": int(nb_i)}
    ms_s, ent_s = kinfo["fp64"]
    nbytes = 8 * ent_s + 2 * 8 * N
    ach = nbytes / (ms_s * 1e-3) * 1e-9
    roof_st = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
               "frac": round(ach / HBM_PEAK_GBPS, 4), "traffic": None,
               "kernel": "k_bgemv<double, 1> (M2L + near field, stored FP64)", "launch_ms": round(ms_s, 4),
               "algorithmic_bytes": int(nbytes)}
    ms_o, ent = kinfo["fp32_otf"]
    flop = FLOP_PER_ENTRY * ent
    roof_otf = {"bound": "valu", "achieved": round(flop / (ms_o * 1e-3) * 1e-12, 2), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(flop / (ms_o * 1e-3) * 1e-12 / FP32_PEAK_TFLOPS, 4),
                "kernel": "k_otf<float> (M2L + near field, entries evaluated in FP32)",
                "launch_ms": round(ms_o, 4), "entries": int(ent), "flop_per_entry": FLOP_PER_ENTRY,
                # the library is built without packed FP32 instructions: its scalar FP32 VALU ceiling is half
                "frac_of_unpacked_fp32": round(flop / (ms_o * 1e-3) * 1e-12 / (FP32_PEAK_TFLOPS / 2), 4)}
    tr_path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tr_path):
        tr = json.load(open(tr_path))
        # measured HBM bytes of the same launch; when the tuner chose another inner operator than
        # the counter runs measured, the measured traffic/algorithmic ratio times this launch's bytes
        for roof, key in ((roof_st, f"3d_rbf_{side}_{tp}_fp64_m2l_near"), (roof_in, f"3d_rbf_{side}_{tp}_fp32inner_m2l_near")):
            rec = tr.get(key)
            if not rec or "algorithmic_bytes" not in rec:
                continue
            if abs(rec["algorithmic_bytes"] - roof["algorithmic_bytes"]) <= 1e-3 * roof["algorithmic_bytes"]:
                roof["traffic"] = rec["hbm_bytes_per_launch"]
                roof["traffic_source"] = "rocprofv3 FETCH_SIZE + WRITE_SIZE of this launch"
            else:
                roof["traffic"] = int(rec["traffic_over_algorithmic"] * roof["algorithmic_bytes"])
                roof["traffic_source"] = (f"measured traffic/algorithmic ratio {rec['traffic_over_algorithmic']} of the "
                                          f"same kernel on a launch of {rec['algorithmic_bytes'] / 1e9:.2f} GB, scaled")
    roof, roof2 = (roof_in, roof_st) if t_in >= t_st else (roof_st, roof_in)
    copy_bw = hip.copy_gbps()
    read_bw = h2w.bench_stream_read(4 << 30, 5)
    log(f"device-to-device copy: {copy_bw:.0f} GB/s (read + write); read-only stream: {read_bw:.0f} GB/s")

    # 16 right-hand sides per pass on the matrix cores: the product and 16 lockstep solves
    multi = None
    if not args.no_multi:
        multi = {"nrhs": 16, "matvec": {}, "gmres": {}}
        peak = {"fp64": MFMA_PEAK_TFLOPS["f64"], "fp32": MFMA_PEAK_TFLOPS["f32"], "fp32_inner": MFMA_PEAK_TFLOPS["f32"]}
        for prec, name, s_op in ((h2w.FP64, "fp64", 8), (h2w.FP32, "fp32", 4), (h2w.FP32_INNER, "fp32_inner", 4)):
            r = A.bench_matvec_multi(5, prec)
            ms = r["ms_per_pass"]
            elems = r["launch_bytes"][0] // s_op  # operator elements multiplied (symmetric blocks twice)
            # streamed: every launch's operator bytes (each stored block once when the fused launch applies
            # it both ways) and the 16-wide input and output vectors
            nbytes = sum(r["launch_bytes"][1:]) + 2 * s_op * N * 16
            tflops = 2.0 * elems * 16 / (ms * 1e-3) * 1e-12
            multi["matvec"][name] = {"ms_per_16": round(ms, 4), "ms_per_rhs": round(ms / 16, 4),
                                     "streamed_GBps": round(nbytes / (ms * 1e-3) * 1e-9, 1),
                                     "TFLOPs": round(tflops, 2), "mfma_frac": round(tflops / peak[name], 4),
                                     "speedup_vs_16_single": round(16 * mv[name]["ms"] / ms, 2),
                                     "launch_ms": [round(v, 4) for v in r["launch_ms"][1:]],
                                     "launch_GB": [round(v * 1e-9, 4) for v in r["launch_bytes"][1:]],
                                     "launch_role": r["launch_role"][1:]}
            log(f"16-wide matvec {name}: {ms:.3f} ms per pass ({ms / 16:.3f} ms per right-hand side, "
                f"{tflops:.1f} TFLOP/s, x{16 * mv[name]['ms'] / ms:.2f} vs 16 single products)")
        XT = np.random.default_rng(6).uniform(-1.0, 1.0, (16, N))
        BM = np.empty_like(XT)
        for c in range(16):
            BM[c] = A.dense_rows(None, XT[c])
        for name, mode in (("fp64", h2w.GMRES_FP64), ("fp32_inner_ir", h2w.GMRES_FP32_INNER_IR)):
            A.gmres_multi(BM[:2], None, 500, gtol, mode)  # warm-up (graph capture)
            t0 = time.time()
            XM, sts = A.gmres_multi(BM, None, 500, gtol, mode)
            dt = time.time() - t0
            errs = [float(np.linalg.norm(XM[c] - XT[c]) / np.linalg.norm(XT[c])) for c in range(16)]
            multi["gmres"][name] = {"s_16": round(sts[0]["seconds"], 4), "s_16_host_incl": round(dt, 4),
                                    "max_rel_error": max(errs), "all_converged": all(x["converged"] for x in sts),
                                    "max_iterations": max(x["iterations"] for x in sts),
                                    "matvecs_fp64": sts[0]["matvecs_fp64"], "matvecs_fp32": sts[0]["matvecs_fp32"]}
            log(f"16 solves {name}: {sts[0]['seconds'] * 1e3:.1f} ms ({sts[0]['seconds'] * 1e3 / 16:.2f} ms per "
                f"right-hand side), max error {max(errs):.2e}")
        multi["gmres"]["fp32_inner_ir_vs_fp64"] = round(multi["gmres"]["fp64"]["s_16"]
                                                       / multi["gmres"]["fp32_inner_ir"]["s_16"], 3)
        multi["gmres"]["vs_16_single_fp64_solves"] = round(16 * t64 / multi["gmres"]["fp32_inner_ir"]["s_16"], 2)
    for r in (roof_in, roof_st):
        r["copy_GBps"] = round(copy_bw, 1)
        r["frac_of_copy"] = round(r["achieved"] / copy_bw, 4)
        r["stream_read_GBps"] = round(read_bw, 1)
        r["frac_of_stream_read"] = round(r["achieved"] / read_bw, 4)

    inner_bytes = A.stats()["stored_bytes_inner_fp32"]
    del A  # frees the large system's operators before the batch section
    import gc
    gc.collect()
    batch = None if args.no_batch else bench_batch()

    out = {"metric": METRIC, "value": round(t_step, 6), "unit": "s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(t_step * 1e3, 3), "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 operators, f64 refinement",
           "data": "synthetic (reference uniform grid, x_true ~ U(-1,1) seed 2)",
           "config": {"workload": f"3D H2* RBF grid {side}^3 N={N} n_max=64 ACA tol 1e-{tp} GMRES tol 1e-{tp}",
                      "global_batch": 1, "seq_len": N, "parallelism": "single GPU"},
           "fp64_gmres_s": round(t64, 6), "fp32_stored_ir_gmres_s": round(res["fp32_ir"]["s"], 6),
           "speedup_vs_fp64": round(t64 / t_step, 3),
           "gmres": {k: {"s": round(v["s"], 6), "rel_solution_error": v["err"],
                         **{kk: v["stats"][kk] for kk in ("iterations", "outer_iterations", "matvecs_fp64",
                                                          "matvecs_fp32", "resid")}} for k, v in res.items()},
           "matvec": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                      for k, v in mv.items()},
           "assembly_s": round(st["assembly_seconds"], 2), "memory_elements": M,
           "assembly_rank_model": {k: st[k] for k in ("rank_pred_blocks", "rank_pred_under", "rank_pred_mae_log2",
                                                      "aca_batches", "aca_reruns")},
           "autotune": tune, "multi_rhs": multi, "batch_c4": batch,
           "roofline": roof, "roofline_other": roof2, "roofline_otf": roof_otf,
           "inner_operator_bytes_fp32": inner_bytes, "cpu_baseline": None}
    # GPU time-to-tolerance over FP64 GPU at equal batching: one right-hand side, 16 in lockstep, 512 in a batch
    if HEADLINE + "_no_deflation" in res:
        out["deflation"] = {"m": defl["m"], "k": defl["k"], "build_s": round(defl["build_seconds"], 3),
                            "note": "m Arnoldi steps of the inner operator from a fixed pseudo-random vector, "
                                    "its k Ritz vectors farthest from the diagonal value kept (k = 0: all); "
                                    "built once per operator, independent of b, outside the timed solves",
                            "fp32_inner_ir_no_deflation_s": round(res[HEADLINE + "_no_deflation"]["s"], 6),
                            "speedup_vs_fp64_no_deflation": round(t64 / res[HEADLINE + "_no_deflation"]["s"], 3)}
    out["speedup_vs_fp64_like_for_like"] = {
        "single_rhs": round(t64 / t_step, 3),
        "rhs16_lockstep": multi["gmres"]["fp32_inner_ir_vs_fp64"] if multi else None,
        "batch512": batch["speedup_fp32_inner_ir_vs_batched_fp64"] if batch else None}
    if not args.no_cpu:
        log(f"CPU baseline (oracle) on grid {args.cpu_side}^3")
        out["cpu_baseline"] = cpu_baseline(args.cpu_side, tp, gtol)
        out["speedup_vs_cpu_same_sample"] = out["cpu_baseline"]["speedup_gpu_fp32_inner_ir_same_sample"]
    # the oracle at the full workload, timed once on a GPU box's host cores (bench.py --cpu-full)
    full = os.path.join(ROOT, "profiles", f"cpu_oracle_3d_rbf_grid{side}.json")
    if os.path.exists(full):
        c = json.load(open(full))["cpu_full"]
        if c["N"] == N:
            out["cpu_full_size"] = {"gmres_s": c["gmres_s"], "threads": c["threads"], "cpu_model": c["cpu_model"],
                                    "iterations": c["iterations"], "rel_error": c["rel_error"],
                                    "source": os.path.relpath(full, ROOT)}
            out["speedup_vs_cpu"] = round(c["gmres_s"] / t_step, 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
