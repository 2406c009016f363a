"""Benchmark: kriged nodes/s for 3-D ordinary kriging of 4,194,304 nodes on one
MI355X, FP32 factorisation + FP64 refinement against the plain FP64 GPU path.

Workload (krigekit_amd.synth.ok3d): 50,000 observations from 20 Gaussian
clusters in a 12.8 km x 12.8 km x 320 m box; 256 x 256 x 64 nodes at
(50, 50, 5) m; anisotropic exponential variogram (sill 1.0, nugget 0.02,
ranges 2000/800/50 m, azimuth 30 deg); octant search, at most 4 per octant
(<= 32 neighbours out of the 128 nearest).  One step = one full pass of the
hot path (neighbour search, bucketing, per-node solve, estimate, variance)
over all nodes, with observations, tree, tables and outputs resident in HBM.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "kriged nodes/s, 3D OK 4.2M nodes: FP32+refine vs FP64 GPU at 1e-6 rel err"
HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6


def algorithmic_search_bytes(nnear):
    """per launch of the search kernel: query coordinates (3 x FP64) read,
    neighbour count, flag and ascending id list (int32) written"""
    nb = len(nnear)
    return 24.0 * nb + 8.0 * nb + 4.0 * float(np.sum(nnear))


def solve_flops(nnear, c=30, p=1, q=1):
    """algorithmic flops of assembly + factorisation + solves + outputs per node
    (covariance c flops per evaluation, p constraint rows, q right-hand sides)"""
    n = nnear.astype(np.float64)
    f = c * (n * (n - 1) / 2 + n * q) + n ** 3 / 3 + n ** 2 / 2 + n / 6 + p ** 3 / 3
    f += 2 * n ** 2 * (p + q) + 2 * n * p ** 2 + 4 * n * p * q + 2 * n * q + 2 * (n + p) * q ** 2
    return float(np.sum(f))


def time_path(spec, precision, steps, warmup, reorder=1, graph=False):
    """reorder 1: the default pipeline (two search chunks, the first chunk's
    solve overlapping the second search, unless factor reuse runs the search
    in one launch); 2: one search launch, so the solve phase is timed on its
    own.  graph: after the host-orchestrated passes, the pass captured into a
    HIP graph is replayed `steps` times, and the wall time returned is the
    replays'."""
    from krigekit_amd.kriging import device_synchronize, from_spec
    k = from_spec(spec, precision=precision)
    k.set_search_mode(True, reorder)
    t0 = time.perf_counter()
    k.solve()  # uploads inputs, first full pass, results copied back
    e2e = time.perf_counter() - t0
    for _ in range(warmup):
        k.solve_resident()
    device_synchronize()
    phases = []
    t0 = time.perf_counter()
    for _ in range(steps):
        k.solve_resident()
        phases.append(k.timing)
    device_synchronize()
    wall = (time.perf_counter() - t0) / steps
    k.fetch()
    est, var = k.get_results(copy=True)
    nnear, _ = k.neighbours()
    agg = {key: float(np.mean([p[key] for p in phases])) for key in phases[0]}
    agg["end_to_end_ms"] = e2e * 1e3
    agg["pipeline_ms_per_step"] = wall * 1e3
    if graph:
        # the same pass captured into a HIP graph (every count consumed on the
        # device) and replayed: the production path, timed the same way
        k.graph_capture()
        for _ in range(warmup):
            k.graph_launch()
        device_synchronize()
        gms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            k.graph_launch()
            gms.append(k.timing["total_ms"])
        device_synchronize()
        wall = (time.perf_counter() - t0) / steps
        agg["graph_ms_per_step"] = wall * 1e3
        agg["graph_device_ms"] = float(np.mean(gms))
        k.fetch()
        eg, vg = k.get_results(copy=True)
        agg["graph_bitwise_equal"] = bool(np.array_equal(eg, est, equal_nan=True) and
                                          np.array_equal(vg, var, equal_nan=True))
    k.close()
    return wall, agg, est, var, nnear


def rel_stats(x, ref, nnear, floor=1e-4):
    ok = nnear != 1
    e = np.abs(x[ok] - ref[ok]) / np.maximum(np.abs(ref[ok]), floor)
    return dict(max=float(np.max(e)), p99=float(np.percentile(e, 99)))


def host_cpu():
    """host CPU model and the number of online CPUs"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(spec, stride=4):
    """the compiled reference engine on every `stride`-th node, with as many
    threads as the host has cores up to 16 (the CPU share of one GPU on the
    GPU hosts); returns the baseline and the reference's results on the sample"""
    nthread = min(os.cpu_count() or 1, 16)
    sub = dict(spec)
    sub["grid"] = spec["grid"][::stride]
    n = len(sub["grid"])
    try:
        from oracle import refdrv
        if not refdrv.ref_available("table"):
            raise FileNotFoundError
        from oracle.refdrv import RefKriging
        k = RefKriging(ndim=3)
        o = sub["obs"][0]
        k.set_obs(1, o["coord"], o["value"], nmax=o["nmax"])
        k.set_grid(sub["grid"])
        for v in sub["vgm"]:
            k.set_vgm(**v)
        for s in sub["search"]:
            k.set_search(**s)
        t = time.perf_counter()
        k.solve(nthread=nthread)
        t = time.perf_counter() - t
        e, v = k.results()
        ref = dict(estimate=e[:, 0], variance=v[:, 0, 0])
        k.close()
        kind = "reference"
    except (FileNotFoundError, OSError):
        from oracle import okrige
        t = time.perf_counter()
        ref = okrige.solve(sub, nthread=nthread)
        t = time.perf_counter() - t
        kind = "port"
    model, ncpu = host_cpu()
    return dict(value=n / t, unit="nodes/s", cores=nthread, kind=kind, host_cpu=model, host_nproc=ncpu,
                sample=f"every {stride}th node of the 256x256x64 grid ({n} nodes), solve() only, "
                       f"kd-tree built beforehand; {t:.2f} s wall"), ref


def err_stats(x, ref, floor=1e-4):
    """max and p99 of |x - ref| / max(|ref|, floor) over the values the
    reference defines (blocks with a single neighbour, whose reference estimate
    depends on thread scheduling, and failed blocks are NaN on both sides)"""
    x = np.asarray(x, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    ok = np.isfinite(ref) & np.isfinite(x)
    if not ok.any():
        return dict(max=None, p99=None, n=0)
    e = np.abs(x[ok] - ref[ok]) / np.maximum(np.abs(ref[ok]), floor)
    return dict(max=float(np.max(e)), p99=float(np.percentile(e, 99)), n=int(ok.sum()))


def errors_vs_reference(runs, c2_ref, c2_nnear, stride=4):
    """the checker leg: estimate and variance errors of every workload's GPU
    results (FP64 and FP32+refine) against the compiled reference where it
    runs, else the oracle, on a bounded sample of each workload"""
    from krigekit_amd import synth
    out = {}
    keep = c2_nnear[::stride] != 1
    c2 = {}
    for prec, (e, v) in runs["ok3d"].items():
        c2[prec] = dict(estimate=err_stats(np.where(keep, e[::stride], np.nan), c2_ref["estimate"]),
                        variance=err_stats(np.where(keep, v[::stride], np.nan), c2_ref["variance"]))
    out["3d_octant_4.2M"] = dict(against="compiled reference, table mode", sample=f"every {stride}th node", **c2)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from runners import run_ref
    from oracle import ckrige, okrige
    for name, spec_fn, label in (("2d_ok_40k", synth.ok2d, "all 40,000 nodes"),
                                 ("loo_100k", synth.loo2d, "all 100,000 leave-one-out estimates")):
        spec = spec_fn()
        r = run_ref(spec, mode="table" if spec_fn is synth.ok2d else "analytic", nthread=16)
        keep = r["nnear"] != 1
        res = {}
        for prec, (e, v) in runs[name].items():
            res[prec] = dict(estimate=err_stats(np.where(keep, e, np.nan), r["estimate"]),
                             variance=err_stats(np.where(keep, v, np.nan), r["variance"]))
        out[name] = dict(against="compiled reference, " + ("table" if spec_fn is synth.ok2d else "analytic")
                         + " mode", sample=label, **res)
    # cokriging: the numpy oracle (Cholesky and Schur complement) on every
    # 4999th block
    spec = synth.ck2d()
    idx = np.arange(0, len(spec["grid"]), 4999)
    sub = dict(spec)
    sub["grid"] = spec["grid"][idx]
    sub["sub_coord"] = spec["sub_coord"].reshape(-1, 16, 2)[idx].reshape(-1, 2)
    sub["sub_weight"] = spec["sub_weight"].reshape(-1, 16)[idx].ravel()
    o = ckrige.solve(sub)
    res = {}
    for prec, (e, v) in runs["cokriging"].items():
        res[prec] = dict(estimate=err_stats(e[idx], o["estimate"]), variance=err_stats(v[idx], o["variance"]))
    out["cokriging_1M_blocks"] = dict(against="numpy oracle", sample=f"{len(idx)} blocks (every 4999th)", **res)
    # space-time: the C oracle on every 7499th node (the reference's
    # space-time entry points cannot be run here)
    spec = synth.st()
    idx = np.arange(0, len(spec["grid"]), 7499)
    sub = dict(spec)
    sub["grid"] = spec["grid"][idx]
    o = okrige.solve(sub, nthread=16)
    keep = o["nnear"] != 1
    res = {}
    for prec, (e, v) in runs["space_time"].items():
        res[prec] = dict(estimate=err_stats(np.where(keep, e[idx], np.nan), o["estimate"]),
                         variance=err_stats(np.where(keep, v[idx], np.nan), o["variance"]))
    out["space_time_7.5M"] = dict(against="C oracle", sample=f"{len(idx)} nodes (every 7499th)", **res)
    return out


def loo_gradient(steps=10):
    """leave-one-out MSE gradient on 100,000 observations (90,000 uniform +
    10,000 duplicates jittered by 0.01 m), nmax 16, anisotropic spherical
    (nugget 0.1, sill 1, ranges 1500/600, azimuth 45): adjoint kernel timing
    and its agreement with central differences (step 1e-4 relative) of the
    forward MSE, analytic covariance, FP64"""
    from krigekit_amd import synth
    from krigekit_amd.kriging import device_synchronize, from_spec
    spec = synth.loo2d()
    names = ["nugget", "sill", "a_major", "a_minor1", "azimuth"]
    k = from_spec(spec, table_mode=False)
    v0 = dict(spec["vgm"][0])
    vset = {key: v0[key] for key in ("vtype", "nugget", "sill", "a_major", "a_minor1", "azimuth")}
    device_synchronize()
    t0 = time.perf_counter()
    k.loo_grad()  # first evaluation: the search and a full pass
    first = time.perf_counter() - t0
    # an optimiser's loop: a new variogram per evaluation, the neighbour
    # lists kept (model rebuild, adjoint kernel, device reduction)
    t0 = time.perf_counter()
    for i in range(steps):
        k.set_vgm(1, 1, append=False, **dict(vset, sill=v0["sill"] * (1.0 + 1e-3 * (i % 2 + 1))))
        k.loo_grad()
    wall = (time.perf_counter() - t0) / steps
    tm = k.timing
    k.set_vgm(1, 1, append=False, **vset)
    mse, g = k.loo_grad()
    k.close()

    def mse_at(name, val):
        v = dict(spec["vgm"][0])
        v[name] = val
        s2 = dict(spec)
        s2["vgm"] = [v]
        kk = from_spec(s2, table_mode=False)
        kk.solve()
        e, _ = kk.get_results(copy=True)
        kk.close()
        z = spec["obs"][0]["value"]
        ok = np.isfinite(e)
        return float(np.mean((e[ok] - z[ok]) ** 2))

    rel = {}
    for n in names:
        h = 1e-4 * abs(v0[n])
        fd = (mse_at(n, v0[n] + h) - mse_at(n, v0[n] - h)) / (2 * h)
        rel[n] = abs(g[n] - fd) / abs(fd)
    return dict(obs=len(spec["obs"][0]["value"]), mse=mse, grad={n: g[n] for n in names},
                ms_per_eval=wall * 1e3, grad_kernel_ms=tm["loo_grad_kernel_ms"], first_eval_ms=first * 1e3,
                fd_rel_err=rel, fd_max_rel_err=max(rel.values()))


def resident_rate(k, nodes, steps=3):
    """nodes/s of repeated device passes over resident inputs (one warm pass
    first), with the last pass's per-phase device timings and the results"""
    from krigekit_amd.kriging import device_synchronize
    k.solve()
    res = k.get_results(copy=True)
    k.solve_resident()
    device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        k.solve_resident()
    device_synchronize()
    wall = (time.perf_counter() - t0) / steps
    tm = k.timing
    return dict(nodes=nodes, nodes_per_s=nodes / wall, ms_per_pass=wall * 1e3,
                phase_ms={key: round(float(v), 4) for key, v in tm.items()}), res


def space_time(precision="fp64"):
    """space-time product-sum kriging: 500 stations x 365 days (182,500
    observations), 500 x 500 km grid at 1 km x 30 times (7.5 million nodes),
    Cs exponential (nugget 0.05, sill 1, 50 km), Ct exponential (sill 0.5,
    5 days), k 0.5, time_at 10 km/day, nmax 48, maxdist 40"""
    from krigekit_amd import synth
    from krigekit_amd.kriging_st import from_st_spec
    spec = synth.st()
    k = from_st_spec(spec, precision=precision)
    out, res = resident_rate(k, len(spec["grid"]))
    k.close()
    return dict(precision=precision, **out), res


def cokriging(precision="fp64"):
    """2-variable standard cokriging (linear model of coregionalisation):
    20,000 + 200,000 observations, 10^6 blocks of 4 x 4 Gauss-Legendre nodes,
    24 + 48 neighbours (74 unknowns); fp32r = FP32 factorisation with the
    MFMA trailing update and two FP64 refinement steps"""
    from krigekit_amd import synth
    from krigekit_amd.kriging import from_spec
    spec = synth.ck2d()
    k = from_spec(spec, precision=precision)
    out, res = resident_rate(k, len(spec["grid"]))
    k.close()
    return dict(precision=precision, **out), res


def small_run(spec_fn, precision):
    """one full solve of a small workload (results only)"""
    from krigekit_amd.kriging import from_spec
    k = from_spec(spec_fn(), precision=precision, table_mode=spec_fn.__name__ != "loo2d")
    k.solve()
    res = k.get_results(copy=True)
    k.close()
    return res


def load_schedules():
    """the per-bucket solve schedules the library loaded (tuning.txt beside
    it): width and looping grid per (coordinates, precision, bucket)"""
    path = os.path.join(ROOT, "krigekit_amd", "tuning.txt")
    if not os.path.exists(path):
        return None
    rows = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith("#")]
    return [dict(dim=int(r[0]), precision=("fp64", "fp32r")[int(r[1])], table=bool(int(r[2])),
                 bucket_rows=(8, 12, 16, 24, 32, 48, 64)[int(r[3]) - 1], width=int(r[4]), grid=int(r[5]),
                 tuned_ms=float(r[6])) for r in rows if int(r[0]) == 3]


def load_issue():
    """VALU and SALU wave-instructions per search launch from the committed
    SQ counter summary (profiles/search_issue.json), or None"""
    path = os.path.join(ROOT, "profiles", "search_issue.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def load_traffic():
    path = os.path.join(ROOT, "profiles", "search_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-loo-gradient", action="store_true")
    ap.add_argument("--no-solve-alone", action="store_true",
                    help="skip the two passes with one search launch that time the solve roofline")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the space-time and cokriging workload lines")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a plain device copy of this many bytes (counter calibration)")
    ap.add_argument("--only", choices=["space-time", "cokriging"], default=None,
                    help="only time this secondary workload and print its line")
    args = ap.parse_args()
    if args.copy_probe > 0:
        from krigekit_amd.kriging import copy_probe
        ms = copy_probe(args.copy_probe, reps=5)
        print(json.dumps({"copy_bytes": args.copy_probe, "copy_ms": ms,
                          "copy_GBps": 2 * args.copy_probe / (ms * 1e-3) / 1e9}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        sys.exit("bench.py runs on one GPU only: this build has no multi-GPU path "
                 "(nodes are independent: each GPU would take a contiguous range of the Morton-ordered nodes).")
    if args.only:
        out, _ = space_time() if args.only == "space-time" else cokriging()
        print(json.dumps({"workload": args.only, **out}), flush=True)
        return
    from krigekit_amd import synth
    from krigekit_amd.kriging import from_spec
    # one small solve first: the process's one-time device set-up stays out
    # of the end-to-end figures
    kw = from_spec(synth.ok3d(scale=1.0 / 1024))
    kw.solve()
    kw.close()
    spec = synth.ok3d()
    nb = len(spec["grid"])
    w64, t64, e64, v64, nn = time_path(spec, "fp64", args.steps, args.warmup, graph=True)
    w32, t32, e32, v32, _ = time_path(spec, "fp32r", args.steps, args.warmup, graph=True)
    # the solves timed apart from the search (no overlapped chunk)
    if args.no_solve_alone:
        t64s, t32s = t64, t32
    else:
        _, t64s, _, _, _ = time_path(spec, "fp64", 2, 1, reorder=2)
        _, t32s, _, _, _ = time_path(spec, "fp32r", 2, 1, reorder=2)
    value = nb / w32
    search_ms = t32["fast_search_kernel_ms"]
    # the search runs as fast_search_launches launches over equal query
    # chunks: per launch, both the algorithmic bytes and the time are that
    # share of the pass, so their ratio is the pass's
    nlaunch = max(1, int(t32["fast_search_launches"]))
    alg_bytes = algorithmic_search_bytes(nn)
    achieved = alg_bytes / (search_ms * 1e-3) / 1e9
    traffic = load_traffic()
    flops = solve_flops(nn)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": w32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factorisation + f64 refinement (f64 assembly, distances, outputs)",
        "data": "synthetic (seeded clustered observations, regular target grid)",
        "config": {"workload": "3-D ordinary kriging, octant search (<=4 per octant, 128 candidates)",
                   "nodes": nb, "grid": [256, 256, 64], "obs": 50000,
                   "variogram": "exponential sill 1.0 nugget 0.02 ranges 2000/800/50 m azimuth 30",
                   "covariance": "lookup table (reference default)", "parallelism": "1 GPU"},
        "fp64_nodes_per_s": nb / w64,
        "fp64_ms_per_step": w64 * 1e3,
        "speedup_fp32r_vs_fp64": w64 / w32,
        "solve_kernel_speedup_fp32r_vs_fp64": t64s["solve_ms"] / t32s["solve_ms"],
        "err_fp32r_vs_fp64": {"estimate": rel_stats(e32, e64, nn), "variance": rel_stats(v32, v64, nn),
                              "floor": 1e-4},
        "step": "one graph replay of the captured pass (search, set grouping, solve per bucket, exact search of "
                "flagged queries, re-solves, fallback) over all nodes; the host-orchestrated pass is "
                "pipeline_ms_per_step in the phase objects",
        "tuned_schedules": load_schedules(),
        "phase_ms_fp32r": {k: round(v, 4) for k, v in t32.items()},
        "phase_ms_fp64": {k: round(v, 4) for k, v in t64.items()},
        "roofline": {
            "kernel": "k_search_sector (octant neighbour search: one kd-tree traversal, per-octant register lists, key histogram bound, exact FP64 ranks)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic.get("bytes_per_launch") if traffic else None,
            "launches_per_pass": nlaunch,
            "algorithmic_bytes_per_launch": alg_bytes / nlaunch,
            "launch_ms": search_ms / nlaunch,
        },
        "search_issue": None,
        "solve_roofline": {
            "kernel": "k_solve (assembly + LDL^T + forward solves, FP64)",
            "bound": "valu_fp64",
            "achieved": flops / (t64s["solve_ms"] * 1e-3) / 1e12,
            "peak": FP64_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": flops / (t64s["solve_ms"] * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS,
            "solve_ms": t64s["solve_ms"],
            "note": "solve timed in a pass with one search launch (no overlapped chunk)" if not args.no_solve_alone
                    else "solve of the default pipeline: part of it overlaps the search and is not counted",
        },
    }
    issue = load_issue()
    if issue:
        # wave-instructions issue at one per 4 cycles per SIMD: 1,024 SIMDs at
        # 2.4 GHz; the search is bound by this issue rate, not by HBM
        cap = 1024 * 2.4e9 / 4 * (search_ms / nlaunch) * 1e-3
        line["search_issue"] = {"bound": "valu+salu instruction issue",
                                "valu_per_launch": issue["valu_per_launch"],
                                "salu_per_launch": issue["salu_per_launch"],
                                "valu_issue_frac": issue["valu_per_launch"] / cap,
                                "salu_issue_frac": issue["salu_per_launch"] / cap,
                                "source": issue.get("source")}
    line["end_to_end"] = {
        "note": "Kriging.solve() on host arrays: inputs copied to HBM, one pass, results copied back (PCIe "
                "included; never the value)",
        "fp32r_ms": t32["end_to_end_ms"], "fp64_ms": t64["end_to_end_ms"],
        "fp32r_nodes_per_s": nb / (t32["end_to_end_ms"] * 1e-3),
        "fp64_nodes_per_s": nb / (t64["end_to_end_ms"] * 1e-3)}
    if not args.no_loo_gradient:
        line["loo_gradient"] = loo_gradient()
    runs = {"ok3d": {"fp64": (e64, v64), "fp32r": (e32, v32)}}
    if not args.no_secondary:
        st64, r = space_time("fp64")
        runs["space_time"] = {"fp64": r}
        st32, r = space_time("fp32r")
        runs["space_time"]["fp32r"] = r
        line["space_time_product_sum"] = st64
        line["space_time_product_sum_fp32r"] = st32
        ck64, r = cokriging("fp64")
        runs["cokriging"] = {"fp64": r}
        ck32, r = cokriging("fp32r")
        runs["cokriging"]["fp32r"] = r
        line["cokriging_2var_blocks"] = ck64
        line["cokriging_2var_blocks_fp32r_mfma"] = ck32
    if not args.no_cpu_baseline:
        line["cpu_baseline"], c2_ref = cpu_baseline(spec)
        if not args.no_secondary:
            from krigekit_amd import synth as _s
            runs["2d_ok_40k"] = {p: small_run(_s.ok2d, p) for p in ("fp64", "fp32r")}
            runs["loo_100k"] = {p: small_run(_s.loo2d, p) for p in ("fp64", "fp32r")}
            line["cpu_baseline"]["errors_vs_reference"] = errors_vs_reference(runs, c2_ref, nn)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
