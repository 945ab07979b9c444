#!/usr/bin/env python3
"""Benchmark of the batched eigensolver of one NRG step on one MI355X.

Workload: 256 real symmetric blocks, orders log-uniform in [16, 1024] from a fixed seed, with
clusters of eigenvalues spaced 1e-10 and exactly degenerate pairs (nrgljubljana_amd/workloads.py).
A step is one nrg_syev_batched call on the whole batch (eigenvalues + eigenvectors), preceded by a
device-to-device restore of the input arena (the call overwrites it, as dsyev does). Inputs are
resident in HBM before the timed region starts.

value = algorithmic GFLOP/s = sum_b (10/3) n_b^3 / time per step (the FLOP count of Householder
reduction + back-transformation, fixed for every algorithm and precision).

Prints one JSON line. One process on one GPU.
"""
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TAIL = 128   # trailing order the library reduces inside one workgroup (k_tail)


def main_panels(n, nb=32, tail=TAIL):
    """Panels of a block of order n reduced by the column kernels: those starting with more than
    `tail` trailing rows (the rest goes to k_tail)."""
    return -(-(n - tail) // nb) if n > tail else 0


def reduction_bytes(sizes, nb=32, groups=1):
    """Algorithmic HBM bytes of the two per-column kernels of the blocked Householder reduction
    (LAPACK dsytrd/dlatrd, lower), summed over the launches of one call:

    k_col_symv, column jl of panel p of a block (trailing order t = m - jl - 1, m = n - 32 p):
        the lower triangle of the trailing matrix read once, the panel's V and W columns [0, jl)
        over the t trailing rows read once (for W^T v and V^T v), v read, y written:
        8 (t (t + 1) / 2 + 2 jl t + 2 t)
    k_col_prep, same column: the panel's V and W columns [0, jl) over the r = m - jl rows read once,
        the ceil(r / 64) symmetric-product partials and 64 ceil(r / 64) dot-product partials of the
        previous column read, the column read and written and W's new column written:
        8 (2 jl r + ceil(r / 64) (r + 64) + 3 r)

    Returns {kernel: (total bytes, launches)}; launch counts follow the library's schedule: the blocks
    are dealt round-robin, largest first, into `groups` groups (the library's concurrent column
    chains), and each group has one launch per (panel, column) while any of its blocks is still in
    its column-kernel panels."""
    sizes = [int(n) for n in sizes]
    order = sorted(range(len(sizes)), key=lambda b: -sizes[b])
    grp = {b: k % groups for k, b in enumerate(order)}
    pmax = max(main_panels(n, nb) for n in sizes)
    prep_b = symv_b = 0.0
    prep_l = symv_l = 0
    for p in range(pmax):
        act = [b for b in range(len(sizes)) if main_panels(sizes[b], nb) > p]
        ng = len({grp[b] for b in act})
        for jl in range(nb + 1):
            for b in act:
                m = sizes[b] - p * nb
                r = m - jl
                prep_b += 8.0 * (2 * jl * r + ((r + 63) // 64) * (r + 64) + 3 * r)
                if jl < nb:
                    t = m - jl - 1
                    symv_b += 8.0 * (t * (t + 1) / 2.0 + 2 * jl * t + 2 * t)
            prep_l += ng
            symv_l += ng if jl < nb else 0
    return {"k_col_prep": (prep_b, prep_l), "k_col_symv": (symv_b, symv_l)}


def red_groups():
    """Concurrent column chains of the library's reduction: 2, or NRG_RED_STREAMS (1 .. 4)."""
    return max(1, min(int(os.environ.get("NRG_RED_STREAMS", "2")), 4))


def measured_traffic(kernel, args):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter runs (profiles/), when
    they were measured on this workload; None otherwise."""
    path = os.path.join(ROOT, "profiles", "traffic_per_launch.json")
    if args.mode != "fp64" or args.nblk != 256 or (args.nmin, args.nmax) != (16, 1024) or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    return k["bytes"] if k else None


def host_cores():
    """CPU threads this job may use: the affinity mask, bounded by OMP_NUM_THREADS when the host sets it
    (a shared host grants each GPU job a share of its cores that way)."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        cores = min(cores, int(omp))
    return max(1, cores)


def host_info():
    """nproc, the affinity mask's size, OMP_NUM_THREADS and the CPU model (lscpu's 'Model name')."""
    model = None
    try:
        with open("/proc/cpuinfo") as F:
            for line in F:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def cpu_baseline(sizes, off_H, arena, off_w, flops, budget_s=30.0):
    """The oracle (C restatement: OpenMP over blocks, largest first, LAPACK dsyev per block) timed on
    this host's cores in the two thread layouts of the reference; the better one is reported."""
    import oracle
    cores = host_cores()
    res = {}
    for label, threads, blas in (("diagth=ncores,blas=1", cores, 1), ("diagth=1,blas=ncores", 1, cores)):
        a = arena.copy()
        t0 = time.perf_counter()
        w, info = oracle.syev_batched(sizes, off_H, a, off_w, threads=threads, blas_threads=blas)
        dt = time.perf_counter() - t0
        res[label] = dt
        log(f"cpu baseline {label}: {dt:.3f} s ({flops / dt / 1e9:.2f} GFLOP/s)")
        if dt > budget_s:
            break
    best = min(res, key=res.get)
    return {"value": flops / res[best] / 1e9, "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "sample": f"the full 256-block batch once, oracle dsyev per block ({best}); seconds: "
                      + ", ".join(f"{k}={v:.3f}" for k, v in res.items()),
            "host": host_info()}


def bench_eig(args):
    import torch
    from nrgljubljana_amd import Context, NRG_SYEV_FP64, NRG_SYEV_MIXED, workloads

    if args.mode not in ("fp64", "mixed"):
        sys.exit("bench.py: the eigensolver forms are fp64 and mixed")
    mode = NRG_SYEV_FP64 if args.mode == "fp64" else NRG_SYEV_MIXED
    t0 = time.perf_counter()
    sizes, off_H, arena, off_w, lams = workloads.eig_batch(nblk=args.nblk, nmin=args.nmin, nmax=args.nmax)
    log(f"workload generated in {time.perf_counter() - t0:.1f} s: {len(sizes)} blocks, sum n = {int(sizes.sum())}, "
        f"max n = {int(sizes.max())}")
    flops = workloads.eig_flops(sizes)
    abytes = workloads.eig_bytes(sizes)

    dev = torch.device("cuda:0")
    src = torch.from_numpy(arena).to(dev)
    H = torch.empty_like(src)
    w = torch.zeros(int(sizes.sum()), dtype=torch.float64, device=dev)
    ctx = Context(0)
    stream = torch.cuda.current_stream()

    def step():
        H.copy_(src, non_blocking=True)
        ctx.syev_batched(H, sizes, off_H, w, off_w, mode)

    # the library's autotuner measures each candidate chain count on the first calls of this histogram
    # signature: let it settle before the warm-up and the timed region
    tuning_calls = 0
    while tuning_calls < 12 and any(e["chains"] < 0 for e in ctx.syev_tuning().values()) or tuning_calls == 0:
        step()
        torch.cuda.synchronize()
        tuning_calls += 1
    tuning = ctx.syev_tuning()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # accuracy of the last warm-up result against the generator's closed-form spectra
    wn = w.cpu().numpy()
    eig_err = max(float(np.max(np.abs(wn[off_w[b]:off_w[b] + n] - lams[b]))) for b, n in enumerate(sizes)) / 20.0
    orth_err = None
    if not args.no_check:
        Vn = H.cpu().numpy()
        orth_err = 0.0
        for b, n in enumerate(sizes):
            if n > 400 or b % 4 == 0:
                V = Vn[off_H[b]:off_H[b] + n * n].reshape(n, n)
                orth_err = max(orth_err, float(np.max(np.abs(V @ V.T - np.eye(n)))))

    # timed region: K steps, synchronised on both sides
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = 1e3 * elapsed / args.steps
    syev_ms = ctx.last_time_ms(0)
    from nrgljubljana_amd.core import phase_name
    phases = {phase_name(k): round(ctx.last_time_ms(k), 3) for k in range(1, 5) if ctx.last_time_ms(k) > 0}

    # the same call with the input coming from, and the eigenpairs going back to, pinned host memory
    # (what a caller holding host matrices pays; not part of `value`)
    host_in = torch.from_numpy(arena).pin_memory()
    host_vec = torch.empty_like(host_in).pin_memory()
    host_w = torch.empty(w.shape, dtype=torch.float64).pin_memory()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        H.copy_(host_in, non_blocking=True)
        ctx.syev_batched(H, sizes, off_H, w, off_w, mode)
        host_vec.copy_(H, non_blocking=True)
        host_w.copy_(w, non_blocking=True)
    torch.cuda.synchronize()
    pcie_ms = 1e3 * (time.perf_counter() - t0) / 2

    # live per-kernel timing (HIP events on the library's stream) over two more steps
    ctx.kernel_timing(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.kernel_timing(False)
    for k, (tms, cnt) in sorted(kt.items(), key=lambda x: -x[1][0]):
        log(f"  {k:14s} {tms / 2:9.3f} ms/step  {cnt // 2:5d} launches/step")
    dom = max(kt.items(), key=lambda x: x[1][0])[0] if kt else None
    # the call against the FP64 peak: algorithmic flops (10/3 n^3 per block) over the call's wall time
    # (HIP events at its first and last launch, on the library's stream)
    ach_tf = flops / (syev_ms * 1e-3) / 1e12 if syev_ms > 0 else None
    roofline = {"bound": "mfma", "kernel": "nrg_syev_batched (whole call)",
                "achieved": round(ach_tf, 4) if ach_tf else None, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach_tf / FP64_PEAK_TFLOPS, 5) if ach_tf else None, "traffic": None,
                "call_ms": round(syev_ms, 3), "flops_per_call": flops,
                "note": "algorithmic flops / call wall time / FP64 peak; the per-launch HBM model of the dominant "
                        "kernel is roofline_dominant_kernel"}
    roofdom = None
    # launch counts follow the chain count the calls ran with: the autotuner's choice for this batch's
    # signature, else NRG_RED_STREAMS or the default
    chosen = next(iter(tuning.values()))["chains"] if tuning else -1
    rb = reduction_bytes(sizes, groups=chosen if chosen > 0 else red_groups())
    if dom in rb:
        tot_bytes, nl = rb[dom]
        tms, cnt = kt[dom]
        if cnt != 2 * nl:
            log(f"warning: {dom} launched {cnt // 2} times per step, schedule model says {nl}")
        avg_s = tms / cnt * 1e-3
        ach = (tot_bytes / nl) / avg_s / 1e9
        roofdom = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, args),
                   "bytes_per_launch": round(tot_bytes / nl), "launches_per_step": nl,
                   "avg_launch_us_events": round(avg_s * 1e6, 3),
                   "note": "average launch time from HIP events, which add about 2 us per launch over the kernel "
                           "trace's durations (an upper bound on time, so a lower bound on the fraction)"}
    # the other reduction kernel, same accounting
    others = {}
    for k, (tot_bytes, nl) in rb.items():
        if k != dom and k in kt and nl:
            tms, cnt = kt[k]
            gbs = (tot_bytes / nl) / (tms / cnt * 1e-3) / 1e9
            others[k] = {"achieved_GBs": round(gbs, 2), "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "traffic": measured_traffic(k, args), "bytes_per_launch": round(tot_bytes / nl),
                         "avg_launch_us": round(tms / cnt * 1e3, 3)}

    gflops = flops / (ms_step * 1e-3) / 1e9
    out = {
        "metric": "NRG step wall-time & batched-eig GFLOP/s, mixed vs FP64 GPU at 1e-12 rel err",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if mode == NRG_SYEV_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"batched symmetric eigensolve, {len(sizes)} blocks, n log-uniform in "
                               f"[{args.nmin},{args.nmax}], seed 20261203, "
                               "clusters 1e-10 + degenerate pairs",
                   "form": args.mode, "blocks": int(len(sizes)), "sum_n": int(sizes.sum()), "max_n": int(sizes.max()),
                   "flops_per_step": flops, "alg_bytes_per_step": abytes},
        "syev_call_ms": round(syev_ms, 3),
        "phases_ms": phases,   # of the last timed call, from HIP events on the library's stream
        "pcie_inclusive_ms_per_step": round(pcie_ms, 3),
        "frac_fp64_peak": round(gflops / 1e3 / FP64_PEAK_TFLOPS, 4),
        "max_eig_err_rel": eig_err,
        "max_orth_err": orth_err,
        "kernels_ms_per_step": {k: round(v[0] / 2, 3) for k, v in kt.items()},
        "autotune": autotune_summary(tuning, tuning_calls),
        "roofline": roofline,
        "roofline_dominant_kernel": roofdom,
        "roofline_other": others,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sizes, off_H, arena, off_w, flops)
    ctx.close()
    print(json.dumps(out), flush=True)


def autotune_summary(table, calls):
    """The eigensolver autotuner's choice for the benchmark's signature and its gain over the fixed
    default of two column chains (call times measured by the library on the tuning calls)."""
    if not table:
        return {"tuning_calls": calls, "table": table, "note": "fixed by NRG_RED_STREAMS or NRG_AUTOTUNE=0"}
    sig, e = next(iter(table.items()))
    ms = {int(k): v for k, v in e["ms"].items()}
    out = {"tuning_calls": calls, "signature": sig, "chosen_chains": e["chains"],
           "best_call_ms_per_candidate": e["ms"]}
    if e["chains"] in ms and 2 in ms:
        out["gain_over_default_2_chains"] = round(ms[2] / ms[e["chains"]], 4)
    return out


INT8_PEAK_TOPS = 5000.0


def timed_calls(step, steps, warmup, sync):
    for _ in range(warmup):
        step()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    sync()
    return 1e3 * (time.perf_counter() - t0) / steps


def bench_recalc(args):
    """Grouped U^T O U recalculation (recalc_general over 16-row two-channel QSZ doublet tables):
    FP64 and split-precision forms on the same table; value = FP64-equivalent GFLOP/s of --mode."""
    import torch
    import oracle
    from nrgljubljana_amd import Context, NRG_GEMM_FP64, NRG_GEMM_OZAKI, Gemm3Term, workloads
    t0 = time.perf_counter()
    table = workloads.recalc_table(npairs=args.npairs)
    f1, f2 = workloads.recalc_flops(table)
    flops = f1 + f2
    log(f"recalc table: {len(table)} outputs x 16 rows, {flops / 1e9:.2f} GFLOP per call, "
        f"generated in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda:0")
    keep, terms, outs = [], [], []
    for t in table:
        U1 = torch.from_numpy(t["U1"]).to(dev)
        Up = torch.from_numpy(t["Up"]).to(dev)
        C = torch.zeros(t["dim1"], t["dimp"], dtype=torch.float64, device=dev)
        keep += [U1, Up]
        outs.append(C)
        o1 = op = 0
        for i, (a, b) in enumerate(zip(t["r1"], t["rp"])):
            if a and b:
                O = torch.from_numpy(t["O"][i]).to(dev)
                keep.append(O)
                s = Gemm3Term()
                s.m, s.n, s.k, s.l, s.f = t["dim1"], t["dimp"], int(a), int(b), float(t["f"][i])
                s.A, s.lda, s.transA = U1.data_ptr() + 8 * o1, t["dim1"], 0
                s.O, s.ldo = O.data_ptr(), int(b)
                s.B, s.ldb, s.transB = Up.data_ptr() + 8 * op, t["dimp"], 0
                s.C, s.ldc = C.data_ptr(), t["dimp"]
                terms.append(s)
            o1 += int(a)
            op += int(b)
    ctx = Context(0)
    res, kts = {}, {}
    for name, prec in (("fp64", NRG_GEMM_FP64), ("ozaki", NRG_GEMM_OZAKI)):
        def step():
            for C in outs:
                C.zero_()
            ctx.gemm3_grouped(terms, prec)
        ms = timed_calls(step, args.steps, args.warmup, torch.cuda.synchronize)
        ctx.kernel_timing(True)
        for _ in range(2):
            step()
        torch.cuda.synchronize()
        kts[name] = ctx.kernel_times()
        ctx.kernel_timing(False)
        res[name] = (ms, [C.cpu().numpy() for C in outs])
        log(f"  {name}: {ms:.3f} ms per call, {flops / ms / 1e6:.1f} GFLOP/s (FP64-equivalent); kernels "
            + ", ".join(f"{k} {v[0] / 2:.3f} ms" for k, v in kts[name].items()))
    # normwise error of the split-precision form against FP64, per output, scaled as the C4 criterion
    worst = 0.0
    for t, cf, co in zip(table, res["fp64"][1], res["ozaki"][1]):
        o1 = op = 0
        scale = 0.0
        for i, (a, b) in enumerate(zip(t["r1"], t["rp"])):
            if a and b:
                scale += np.linalg.norm(t["U1"][:, o1:o1 + a]) * np.linalg.norm(t["O"][i]) * \
                    np.linalg.norm(t["Up"][:, op:op + b])
            o1 += int(a)
            op += int(b)
        worst = max(worst, float(np.linalg.norm(co - cf)) / scale)
    mode = args.mode if args.mode in ("fp64", "ozaki") else "ozaki"
    ms = res[mode][0]
    # the product kernels' time per call (two timed calls): both FP64 tile kernels, or the slice kernel
    if mode == "fp64":
        kname = " + ".join(k for k in sorted(kts[mode]) if k.startswith("k_gemm"))
        kms = sum(v[0] for k, v in kts[mode].items() if k.startswith("k_gemm")) / 2
    else:
        kname = "k_oz_gemm"
        kms = kts[mode][kname][0] / 2
    ach = flops / (kms * 1e-3) / 1e12 if kms > 0 else None
    roof = {"bound": "mfma", "kernel": kname, "achieved": round(ach, 3), "unit": "TFLOP/s", "traffic": None}
    if mode == "fp64":
        roof.update(peak=FP64_PEAK_TFLOPS, frac=round(ach / FP64_PEAK_TFLOPS, 4))
    else:
        slice_tops = ach * 36   # 36 INT8 slice products per FP64 product (8 slices, p + q < 8)
        roof.update(peak=INT8_PEAK_TOPS, unit_peak="TOP/s", frac=round(slice_tops / INT8_PEAK_TOPS, 4),
                    note="achieved is FP64-equivalent; frac is the INT8 slice-product rate (36 slice "
                         "products per FP64 product) over the 5.0 POP/s dense INT8 peak")
    cpu = None
    if not args.no_cpu:
        # oracle dgemm on the host cores over a bounded sample of the table
        L = oracle.lib()
        cores = host_cores()
        L.oracle_blas_threads(cores)
        sample = table
        sf = flops
        hterms, hold = [], []
        for t in sample:
            C = np.zeros((t["dim1"], t["dimp"]))
            U1, Up = np.ascontiguousarray(t["U1"]), np.ascontiguousarray(t["Up"])
            hold += [C, U1, Up]
            o1 = op = 0
            for i, (a, b) in enumerate(zip(t["r1"], t["rp"])):
                if a and b:
                    s = Gemm3Term()
                    s.m, s.n, s.k, s.l, s.f = t["dim1"], t["dimp"], int(a), int(b), float(t["f"][i])
                    s.A, s.lda, s.transA = U1.ctypes.data + 8 * o1, t["dim1"], 0
                    O = np.ascontiguousarray(t["O"][i])
                    hold.append(O)
                    s.O, s.ldo = O.ctypes.data, int(b)
                    s.B, s.ldb, s.transB = Up.ctypes.data + 8 * op, t["dimp"], 0
                    s.C, s.ldc = C.ctypes.data, t["dimp"]
                    hterms.append(s)
                o1 += int(a)
                op += int(b)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and reps < 200:
            oracle.gemm3(hterms)
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": round(reps * sf / dt / 1e9, 2), "unit": "GFLOP/s", "cores": cores, "kind": "port",
               "sample": f"the whole table ({len(table)} outputs x 16 rows) {reps} times, oracle dgemm pairs in the "
                         f"reference's order, {dt:.1f} s"}
    out = {"metric": "batched U^T O U recalc GFLOP/s (FP64-equivalent), split-precision vs FP64 at 1e-12 normwise",
           "value": round(flops / ms / 1e6, 2), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int8-slices+i32+f64" if mode == "ozaki" else "f64", "data": "synthetic",
           "config": {"workload": f"recalc_general, {len(table)} outputs, dims log-uniform [16,2000], 16-row QSZ "
                                  "2ch doublet tables, Haar U, N(0,1) O, seed 20261205", "form": mode,
                      "flops_per_step": flops},
           "fp64_ms": round(res["fp64"][0], 3), "ozaki_ms": round(res["ozaki"][0], 3),
           "ozaki_vs_fp64_normwise_err": worst, "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_fdm(args):
    """FDM spectral weights + log binning of one synthetic step (nrg_fdm_bin); value = elements/s."""
    import torch
    import oracle
    from nrgljubljana_amd import Context, FdmPair, BinGrid, workloads
    pairs = workloads.fdm_step(npairs=args.npairs)
    nel = workloads.fdm_elements(pairs)
    dev = torch.device("cuda:0")
    grid = BinGrid(-12.0, 1000.0, 14001, 1e-16)
    keep, structs = [], []
    for p in pairs:
        t = {k: torch.from_numpy(np.ascontiguousarray(p[k])).to(dev) for k in ("E_i", "E_j", "op", "rho_i", "rho_j")}
        keep.append(t)
        structs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], t["E_i"].data_ptr(), t["E_j"].data_ptr(),
                               t["op"].data_ptr(), t["op"].data_ptr(), p["dim_i"], t["rho_i"].data_ptr(), p["K_i"],
                               t["rho_j"].data_ptr(), p["K_j"], 1.0))
    bins = [torch.zeros(grid.nbins, dtype=torch.float64, device=dev) for _ in range(4)]
    ctx = Context(0)
    T, wnf = 1e-3, 0.37

    def step():
        for b in bins:
            b.zero_()
        ctx.fdm_bin(structs, grid, T, wnf, -1, *bins)
    ms = timed_calls(step, args.steps, args.warmup, torch.cuda.synchronize)
    ctx.kernel_timing(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.kernel_timing(False)
    log("  " + ", ".join(f"{k} {v[0] / 2:.3f} ms ({v[1] // 2} launches)" for k, v in kt.items()))
    abytes = 16.0 * nel + 8.0 * sum(p["dim_i"] + p["dim_j"] for p in pairs)
    # kernel time per step (over the two steps timed)
    tms, cnt = kt["k_fdm_bin"]
    ach = abytes / (tms / 2 * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "k_fdm_bin", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "launches_per_step": cnt // 2,
            "note": "algorithmic bytes = 16 B per element (op1 + op2 or the rho product) + 8 B per energy, over "
                    "the step's k_fdm_bin time"}
    cpu = None
    if not args.no_cpu:
        sample = pairs
        hs, hold = [], []
        for p in sample:
            arrs = {k: np.ascontiguousarray(p[k]) for k in ("E_i", "E_j", "op", "rho_i", "rho_j")}
            hold.append(arrs)
            hs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], arrs["E_i"].ctypes.data,
                              arrs["E_j"].ctypes.data, arrs["op"].ctypes.data, arrs["op"].ctypes.data, p["dim_i"],
                              arrs["rho_i"].ctypes.data, p["K_i"], arrs["rho_j"].ctypes.data, p["K_j"], 1.0))
        oracle.lib().oracle_blas_threads(1)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and reps < 500:
            oracle.fdm_bin(hs, grid, T, wnf, -1, grid.nbins)
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": round(reps * workloads.fdm_elements(sample) / dt / 1e6, 2), "unit": "Melements/s",
               "cores": 1, "kind": "port",
               "sample": f"the whole step ({len(pairs)} pairs) {reps} times, scalar term loops + serial binning, "
                         f"{dt:.1f} s"}
    out = {"metric": "FDM spectral weights + log binning, elements per second", "value": round(nel / ms / 1e3, 2),
           "unit": "Melements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"one FDM step, {len(pairs)} subspace pairs, dims log-uniform [16,2000], 40 % kept, "
                                  "1000 bins/decade over 14 decades, seed 20261206", "elements_per_half": nel},
           "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_cfs(args):
    """CFS spectral weights + log binning of one synthetic step below the last (nrg_cfs_bin, Eqs. 15-16);
    value = elements/s."""
    import torch
    import oracle
    from nrgljubljana_amd import Context, FdmPair, BinGrid, workloads
    pairs = workloads.fdm_step(npairs=args.npairs)
    nel = workloads.cfs_elements(pairs)
    flops = workloads.cfs_flops(pairs)
    dev = torch.device("cuda:0")
    grid = BinGrid(-12.0, 1000.0, 14001, 1e-16)
    keep, structs = [], []
    for p in pairs:
        t = {k: torch.from_numpy(np.ascontiguousarray(p[k])).to(dev) for k in ("E_i", "E_j", "op", "rho_i", "rho_j")}
        keep.append(t)
        structs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], t["E_i"].data_ptr(), t["E_j"].data_ptr(),
                               t["op"].data_ptr(), t["op"].data_ptr(), p["dim_i"], t["rho_i"].data_ptr(), p["K_i"],
                               t["rho_j"].data_ptr(), p["K_j"], 1.0))
    bins = [torch.zeros(grid.nbins, dtype=torch.float64, device=dev) for _ in range(4)]
    ctx = Context(0)
    T, Z = 1e-3, 1.0

    def step():
        for b in bins:
            b.zero_()
        ctx.cfs_bin(structs, grid, T, Z, False, -1, *bins)
    ms = timed_calls(step, args.steps, args.warmup, torch.cuda.synchronize)
    ctx.kernel_timing(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.kernel_timing(False)
    log(f"cfs step: {len(pairs)} pairs, {nel / 1e6:.2f} M elements, {flops / 1e9:.2f} GFLOP of products, "
        f"{ms:.3f} ms per call; " + ", ".join(f"{k} {v[0] / 2:.3f} ms ({v[1] // 2} launches)" for k, v in kt.items()))
    gemm_ms = sum(v[0] for k, v in kt.items() if k.startswith("k_gemm")) / 2
    bin_ms, bin_cnt = kt.get("k_cfs_bin", (0.0, 2))
    if gemm_ms >= bin_ms / 2:
        ach = flops / (gemm_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": " + ".join(sorted(k for k in kt if k.startswith("k_gemm"))),
                "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None}
    else:
        abytes = 16.0 * nel + 8.0 * sum(p["dim_i"] + p["dim_j"] for p in pairs)
        ach = abytes / (bin_ms / bin_cnt * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": "k_cfs_bin", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "note": "algorithmic bytes = 16 B per element (an operator and a product) + 8 B per energy"}
    cpu = None
    if not args.no_cpu:
        hs, hold = [], []
        for p in pairs:
            arrs = {k: np.ascontiguousarray(p[k]) for k in ("E_i", "E_j", "op", "rho_i", "rho_j")}
            hold.append(arrs)
            hs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], arrs["E_i"].ctypes.data,
                              arrs["E_j"].ctypes.data, arrs["op"].ctypes.data, arrs["op"].ctypes.data, p["dim_i"],
                              arrs["rho_i"].ctypes.data, p["K_i"], arrs["rho_j"].ctypes.data, p["K_j"], 1.0))
        oracle.lib().oracle_blas_threads(1)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and reps < 500:
            oracle.cfs_bin(hs, grid, T, Z, False, -1, grid.nbins)
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": round(reps * nel / dt / 1e6, 2), "unit": "Melements/s", "cores": 1, "kind": "port",
               "sample": f"the whole step ({len(pairs)} pairs) {reps} times, dgemm products (1 thread) + scalar "
                         f"term loops + serial binning, {dt:.1f} s"}
    out = {"metric": "CFS spectral weights + log binning, elements per second", "value": round(nel / ms / 1e3, 2),
           "unit": "Melements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"one CFS step below the last, {len(pairs)} subspace pairs, dims log-uniform "
                                  "[16,2000], 40 % kept, 1000 bins/decade over 14 decades, seed 20261206",
                      "elements": nel, "product_flops": flops},
           "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_ft(args):
    """FT spectral weights + log binning of one synthetic step (nrg_ft_bin, kept x kept); value = elements/s."""
    import torch
    import oracle
    from nrgljubljana_amd import Context, FdmPair, BinGrid, workloads
    pairs = workloads.dmnrg_step(npairs=args.npairs)
    nel = sum(p["K_i"] * p["K_j"] for p in pairs)
    dev = torch.device("cuda:0")
    grid = BinGrid(-12.0, 1000.0, 14001, 1e-16)
    T, Z = 0.5, 40.0
    keep, structs = [], []
    for p in pairs:
        t = {k: torch.from_numpy(np.ascontiguousarray(p[k])).to(dev) for k in ("E_i", "E_j", "op")}
        keep.append(t)
        structs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], t["E_i"].data_ptr(), t["E_j"].data_ptr(),
                               t["op"].data_ptr(), t["op"].data_ptr(), p["dim_i"], None, 0, None, 0, 1.0))
    bins = [torch.zeros(grid.nbins, dtype=torch.float64, device=dev) for _ in range(2)]
    ctx = Context(0)

    def step():
        for b in bins:
            b.zero_()
        ctx.ft_bin(structs, grid, T, Z, -1, *bins)
    ms = timed_calls(step, args.steps, args.warmup, torch.cuda.synchronize)
    ctx.kernel_timing(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.kernel_timing(False)
    log(f"ft step: {len(pairs)} pairs, {nel / 1e6:.2f} M elements, {ms:.3f} ms per call; "
        + ", ".join(f"{k} {v[0] / 2:.3f} ms" for k, v in kt.items()))
    tms, cnt = kt["k_cfs_bin"]
    abytes = 16.0 * nel + 8.0 * sum(p["K_i"] + p["K_j"] for p in pairs)
    ach = abytes / (tms / cnt * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "k_cfs_bin", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "note": "algorithmic bytes = 16 B per element (op1, op2) + 8 B per energy"}
    cpu = None
    if not args.no_cpu:
        hs, hold, el = [], [], 0
        for p in sorted(pairs, key=lambda q: q["K_i"] * q["K_j"])[:len(pairs) // 2]:
            arrs = {k: np.ascontiguousarray(p[k]) for k in ("E_i", "E_j", "op")}
            hold.append(arrs)
            hs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], arrs["E_i"].ctypes.data,
                              arrs["E_j"].ctypes.data, arrs["op"].ctypes.data, arrs["op"].ctypes.data, p["dim_i"],
                              None, 0, None, 0, 1.0))
            el += p["K_i"] * p["K_j"]
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and reps < 5000:
            oracle.ft_bin(hs, grid, T, Z, -1, grid.nbins)
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": round(reps * el / dt / 1e6, 3), "unit": "Melements/s", "cores": 1, "kind": "port",
               "sample": f"the {len(hs)} smallest pairs ({el / 1e6:.2f} M elements) {reps} times, the reference's "
                         f"loop + serial binning, {dt:.1f} s"}
    out = {"metric": "FT spectral weights + log binning, elements per second", "value": round(nel / ms / 1e3, 2),
           "unit": "Melements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"one FT step, {len(pairs)} subspace pairs, kept dims log-uniform [16,1000], "
                                  "1000 bins/decade over 14 decades, seed 20261208", "elements": nel},
           "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_temp(args):
    """chi(T) temperature sum of one synthetic step (nrg_temp_sum, Algo_CHIT over kept x kept);
    value = elements/s."""
    import torch
    import oracle
    from nrgljubljana_amd import Context, FdmPair, workloads
    from nrgljubljana_amd.core import NRG_TEMP_CHIT
    pairs = workloads.dmnrg_step(npairs=args.npairs)
    nel = sum(p["K_i"] * p["K_j"] for p in pairs)
    dev = torch.device("cuda:0")
    T, Z = 0.5, 40.0
    keep, structs = [], []
    for p in pairs:
        t = {k: torch.from_numpy(np.ascontiguousarray(p[k])).to(dev) for k in ("E_i", "E_j", "op")}
        keep.append(t)
        structs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], t["E_i"].data_ptr(), t["E_j"].data_ptr(),
                               t["op"].data_ptr(), t["op"].data_ptr(), p["dim_i"], None, 0, None, 0, 1.0))
    out = torch.zeros(1, dtype=torch.float64, device=dev)
    ctx = Context(0)

    def step():
        out.zero_()
        ctx.temp_sum(NRG_TEMP_CHIT, structs, T, Z, out)
    ms = timed_calls(step, args.steps, args.warmup, torch.cuda.synchronize)
    ctx.kernel_timing(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.kernel_timing(False)
    log(f"temp step: {len(pairs)} pairs, {nel / 1e6:.2f} M elements, {ms:.3f} ms per call; "
        + ", ".join(f"{k} {v[0] / 2:.3f} ms" for k, v in kt.items()))
    tms, cnt = kt["k_temp_sum"]
    abytes = 16.0 * nel + 8.0 * sum(p["K_i"] + p["K_j"] for p in pairs)
    ach = abytes / (tms / cnt * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "k_temp_sum", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "note": "algorithmic bytes = 16 B per element (op1, op2) + 8 B per energy"}
    cpu = None
    if not args.no_cpu:
        hs, hold, el = [], [], 0
        for p in sorted(pairs, key=lambda q: q["K_i"] * q["K_j"])[:len(pairs) // 2]:
            arrs = {k: np.ascontiguousarray(p[k]) for k in ("E_i", "E_j", "op")}
            hold.append(arrs)
            hs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], arrs["E_i"].ctypes.data,
                              arrs["E_j"].ctypes.data, arrs["op"].ctypes.data, arrs["op"].ctypes.data, p["dim_i"],
                              None, 0, None, 0, 1.0))
            el += p["K_i"] * p["K_j"]
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and reps < 5000:
            oracle.temp_sum("chit", hs, T, Z)
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": round(reps * el / dt / 1e6, 3), "unit": "Melements/s", "cores": 1, "kind": "port",
               "sample": f"the {len(hs)} smallest pairs ({el / 1e6:.2f} M elements) {reps} times, the reference's "
                         f"loop, {dt:.1f} s"}
    res = {"metric": "chi(T) temperature sum, elements per second", "value": round(nel / ms / 1e3, 2),
           "unit": "Melements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"one chi(T) step, {len(pairs)} subspace pairs, kept dims log-uniform [16,1000], "
                                  "seed 20261208", "elements": nel},
           "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(res), flush=True)


def bench_mats(args):
    """Matsubara Green's function of one synthetic FDM step (nrg_fdm_mats, 100 frequencies, the
    reference's default P.mats); value = (element, frequency) evaluations per second."""
    import torch
    import oracle
    from nrgljubljana_amd import Context, FdmPair, workloads
    pairs = workloads.fdm_step(npairs=args.npairs)
    nel = workloads.fdm_elements(pairs)
    mats = args.mats
    evals = float(nel) * mats
    dev = torch.device("cuda:0")
    T, wnf = 0.05, 0.37
    keep, structs = [], []
    for p in pairs:
        t = {k: torch.from_numpy(np.ascontiguousarray(p[k])).to(dev) for k in ("E_i", "E_j", "op", "rho_i", "rho_j")}
        keep.append(t)
        structs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], t["E_i"].data_ptr(), t["E_j"].data_ptr(),
                               t["op"].data_ptr(), t["op"].data_ptr(), p["dim_i"], t["rho_i"].data_ptr(), p["K_i"],
                               t["rho_j"].data_ptr(), p["K_j"], 1.0))
    G = torch.zeros(2 * mats, dtype=torch.float64, device=dev)
    ctx = Context(0)

    def step():
        G.zero_()
        ctx.mats("fdm", structs, mats, T, -1, G, wnf=wnf)
    ms = timed_calls(step, args.steps, args.warmup, torch.cuda.synchronize)
    ctx.kernel_timing(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.kernel_timing(False)
    log(f"mats step: {len(pairs)} pairs, {nel / 1e6:.2f} M elements x {mats} frequencies, {ms:.3f} ms per call; "
        + ", ".join(f"{k} {v[0] / 2:.3f} ms" for k, v in kt.items()))
    tms, cnt = kt["k_mats_fdm"]
    # issued FP64 work per (element, frequency): E^2 + w^2, two Newton steps of the reciprocal (2 FMA
    # each) and the two accumulating FMAs: 13 flops, plus one v_rcp_f64
    ach = 13.0 * evals / (tms / cnt * 1e-3) / 1e12
    roof = {"bound": "valu", "kernel": "k_mats_fdm", "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None,
            "note": "FP64 vector work, 13 flops + 1 reciprocal per (element, frequency); peak = FP64 vector"}
    cpu = None
    if not args.no_cpu:
        hs, hold, el = [], [], 0
        for p in sorted(pairs, key=lambda q: q["dim_i"] * q["dim_j"])[:len(pairs) // 4]:
            arrs = {k: np.ascontiguousarray(p[k]) for k in ("E_i", "E_j", "op", "rho_i", "rho_j")}
            hold.append(arrs)
            hs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], arrs["E_i"].ctypes.data,
                              arrs["E_j"].ctypes.data, arrs["op"].ctypes.data, arrs["op"].ctypes.data, p["dim_i"],
                              arrs["rho_i"].ctypes.data, p["K_i"], arrs["rho_j"].ctypes.data, p["K_j"], 1.0))
            el += p["dim_i"] * p["dim_j"] - p["K_i"] * p["K_j"]
        oracle.lib().oracle_blas_threads(1)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and reps < 5000:
            oracle.mats("fdm", hs, mats, T, -1, wnf=wnf)
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": round(reps * el * mats / dt / 1e6, 2), "unit": "Mevaluations/s", "cores": 1, "kind": "port",
               "sample": f"the {len(hs)} smallest pairs ({el / 1e6:.3f} M elements x {mats} frequencies) {reps} "
                         f"times, the reference's term loops with C complex division, {dt:.1f} s"}
    out = {"metric": "FDM Matsubara Green's function, (element, frequency) evaluations per second",
           "value": round(evals / ms / 1e3, 2), "unit": "Mevaluations/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"one FDM step, {len(pairs)} subspace pairs, dims log-uniform [16,2000], 40 % "
                                  f"kept, {mats} fermionic Matsubara frequencies, seed 20261206",
                      "elements": nel, "frequencies": mats},
           "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_dmnrg(args):
    """DM-NRG spectral weights + log binning of one synthetic step (nrg_dmnrg_bin); value = elements/s."""
    import torch
    import oracle
    from nrgljubljana_amd import Context, FdmPair, BinGrid, workloads
    pairs = workloads.dmnrg_step(npairs=args.npairs)
    nel = sum(p["K_i"] * p["K_j"] for p in pairs)
    flops = workloads.dmnrg_flops(pairs)
    dev = torch.device("cuda:0")
    grid = BinGrid(-12.0, 1000.0, 14001, 1e-16)
    emin, emax = 2.0, 8.0   # the window of a step of scale 1 with goodE = 2, Lambda = 2
    keep, structs = [], []
    for p in pairs:
        t = {k: torch.from_numpy(np.ascontiguousarray(p[k])).to(dev) for k in ("E_i", "E_j", "op", "rho_i", "rho_j")}
        keep.append(t)
        structs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], t["E_i"].data_ptr(), t["E_j"].data_ptr(),
                               t["op"].data_ptr(), t["op"].data_ptr(), p["dim_i"], t["rho_i"].data_ptr(), p["K_i"],
                               t["rho_j"].data_ptr(), p["K_j"], 1.0))
    bins = [torch.zeros(grid.nbins, dtype=torch.float64, device=dev) for _ in range(2)]
    ctx = Context(0)

    def step():
        for b in bins:
            b.zero_()
        ctx.dmnrg_bin(structs, grid, emin, emax, -1, *bins)
    ms = timed_calls(step, args.steps, args.warmup, torch.cuda.synchronize)
    ctx.kernel_timing(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.kernel_timing(False)
    log(f"dmnrg step: {len(pairs)} pairs, {nel / 1e6:.2f} M kept x kept elements, {flops / 1e9:.2f} GFLOP of products, "
        f"{ms:.3f} ms per call; " + ", ".join(f"{k} {v[0] / 2:.3f} ms" for k, v in kt.items()))
    gemm_ms = sum(v[0] for k, v in kt.items() if k.startswith("k_gemm")) / 2
    bin_ms = kt["k_dmnrg_bin"][0] / 2 if "k_dmnrg_bin" in kt else 0.0
    if gemm_ms >= bin_ms:
        ach = flops / (gemm_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": " + ".join(sorted(k for k in kt if k.startswith("k_gemm"))),
                "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None}
    else:
        ach = 32.0 * nel / (bin_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": "k_dmnrg_bin", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "note": "algorithmic bytes = 32 B per element (op1, op2 and the two products)"}
    cpu = None
    if not args.no_cpu:
        # the reference's loops as they stand (inner sums per element), one core, on the smallest pairs
        order = sorted(range(len(pairs)), key=lambda i: pairs[i]["K_i"] * pairs[i]["K_j"])
        hs, hold, el = [], [], 0
        for i in order:
            p = pairs[i]
            arrs = {k: np.ascontiguousarray(p[k]) for k in ("E_i", "E_j", "op", "rho_i", "rho_j")}
            hold.append(arrs)
            hs.append(FdmPair(p["dim_i"], p["dim_j"], p["K_i"], p["K_j"], arrs["E_i"].ctypes.data,
                              arrs["E_j"].ctypes.data, arrs["op"].ctypes.data, arrs["op"].ctypes.data, p["dim_i"],
                              arrs["rho_i"].ctypes.data, p["K_i"], arrs["rho_j"].ctypes.data, p["K_j"], 1.0))
            el += p["K_i"] * p["K_j"]
            if len(hs) >= len(pairs) // 2:
                break
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and reps < 5000:
            oracle.dmnrg_bin(hs, grid, emin, emax, -1, grid.nbins)
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": round(reps * el / dt / 1e6, 3), "unit": "Melements/s", "cores": 1, "kind": "port",
               "sample": f"the {len(hs)} smallest pairs ({el / 1e6:.2f} M elements) {reps} times, the reference's "
                         f"loops with per-element inner sums + serial binning, {dt:.1f} s"}
    out = {"metric": "DM-NRG spectral weights + log binning, elements per second", "value": round(nel / ms / 1e3, 2),
           "unit": "Melements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"one DM-NRG step, {len(pairs)} subspace pairs, kept dims log-uniform [16,1000], "
                                  "window [2, 8], 1000 bins/decade over 14 decades, seed 20261208",
                      "elements": nel, "product_flops": flops},
           "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def device_uploader(dev):
    """ptr(array) -> device address of a copy of the array. Every array is copied once and both the
    array and its copy are held until the caller drops `keep`: a cache keyed by id() alone would hand a
    later temporary (which can reuse a freed array's id) the copy of an array of another shape."""
    import torch
    keep = {}

    def ptr(a):
        hit = keep.get(id(a))
        if hit is None or hit[0] is not a:
            hit = (a, torch.from_numpy(np.ascontiguousarray(a)).to(dev))
            keep[id(a)] = hit
        assert tuple(hit[1].shape) == tuple(np.shape(a))
        return hit[1].data_ptr()
    return ptr, keep


def ham_terms(sizes, blocks, ptr):
    """nrg_ham_term list of a ham_step, in hamiltonian()'s order: energies, then per block the table's
    offdiag calls and the zeta shifts (c++/core.hpp:72-88, c++/matrix.hpp:32-111)."""
    from nrgljubljana_amd import HamTerm, NRG_HAM_BLOCK, NRG_HAM_DIAG
    scale, sf, zeta = 0.37, 2.0 ** 0.5, 0.05
    terms = []
    for b, blk in enumerate(blocks):
        dims = blk["dims"]
        offs = np.concatenate([[0], np.cumsum(dims)[:-1]]).astype(int)
        for i, d in enumerate(dims):
            if d:
                terms.append(HamTerm(NRG_HAM_DIAG, 0, b, int(offs[i]), int(offs[i]), d, d, ptr(blk["corr"][i]), 0, sf, 0.0))
        for (i, j), m in blk["mats"].items():
            if i > j:
                terms.append(HamTerm(NRG_HAM_BLOCK, 1, b, int(offs[j - 1]), int(offs[i - 1]), dims[j - 1], dims[i - 1],
                                     ptr(m), dims[j - 1], 0.6 / scale, 0.0))
            else:
                terms.append(HamTerm(NRG_HAM_BLOCK, 0, b, int(offs[i - 1]), int(offs[j - 1]), dims[i - 1], dims[j - 1],
                                     ptr(m), dims[j - 1], 0.6 / scale, 0.0))
        for i, d in enumerate(dims):
            if d:
                terms.append(HamTerm(NRG_HAM_DIAG, 0, b, int(offs[i]), int(offs[i]), d, d, None, 0, 0.0,
                                     zeta * (i - 0.5) / scale))
    return terms


def bench_ham(args):
    """Assembly of one step's Hamiltonian blocks in the eigensolver's arena (nrg_hamiltonian); value =
    GB/s of algorithmic traffic (every arena element written once, every source element read once)."""
    import torch
    import oracle
    from nrgljubljana_amd import Context, workloads
    sizes, blocks = workloads.ham_step(nblk=args.nblk)
    off = np.concatenate([[0], np.cumsum(sizes.astype(np.int64) ** 2)[:-1]]).astype(np.int64)
    tot = int(np.sum(sizes.astype(np.int64) ** 2))
    dev = torch.device("cuda:0")
    dptr, keep = device_uploader(dev)
    from nrgljubljana_amd import HamTerm
    terms = ham_terms(sizes, blocks, dptr)
    terms_arr = (HamTerm * len(terms))(*terms)   # marshalled once, as a caller keeping its plan would
    src_elems = sum(m.size for b in blocks for m in b["mats"].values()) + sum(int(np.sum(b["dims"])) for b in blocks)
    abytes = 8.0 * (tot + src_elems)
    H = torch.empty(tot, dtype=torch.float64, device=dev)
    ctx = Context(0)

    def step():
        ctx.hamiltonian(sizes, off, H, terms_arr)
    ms = timed_calls(step, args.steps, args.warmup, torch.cuda.synchronize)
    ctx.kernel_timing(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.kernel_timing(False)
    kms = kt["k_ham_tile"][0] / 2
    ach = abytes / (kms * 1e-3) / 1e9
    # what the kernel replaces: building on the host and copying the arena over PCIe (pinned)
    hbuf = torch.empty(tot, dtype=torch.float64).pin_memory()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        H.copy_(hbuf, non_blocking=True)
    torch.cuda.synchronize()
    h2d_ms = (time.perf_counter() - t0) / 3 * 1e3
    log(f"hamiltonian: {len(sizes)} blocks, {tot * 8 / 1e6:.1f} MB arena, {len(terms)} terms, {ms:.3f} ms per call "
        f"(k_ham_tile {kms:.3f} ms); the same arena over PCIe from pinned memory: {h2d_ms:.3f} ms")
    roof = {"bound": "hbm", "kernel": "k_ham_tile", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "note": "algorithmic bytes = 8 B per arena element written + 8 B per source element read"}
    cpu = None
    if not args.no_cpu:
        hk = {}

        def hptr(a):
            hk[id(a)] = np.ascontiguousarray(a)
            return hk[id(a)].ctypes.data
        hterms = ham_terms(sizes, blocks, hptr)
        Hh = np.empty(tot)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and reps < 1000:
            oracle.hamiltonian(sizes, off, Hh, hterms)
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        cpu = {"value": round(abytes / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"the whole step {reps} times (zero fill + term loops, hamiltonian()'s order), "
                         f"{dt * 1e3:.1f} ms each"}
    out = {"metric": "Hamiltonian assembly of one NRG step, algorithmic GB/s", "value": round(abytes / (ms * 1e-3) / 1e9, 2),
           "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{len(sizes)} blocks, orders log-uniform [16,1024] (seed 20261203), 4 sub-blocks "
                                  "each, 4 couplings (one transposed), energies and zeta shifts, seed 20261209",
                      "arena_bytes": tot * 8, "terms": len(terms)},
           "pcie_h2d_ms_same_arena": round(h2d_ms, 3), "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_step(args):
    """One synthetic NRG step, every stage on the GPU and device-resident: Hamiltonian assembly of the
    256 blocks of the eigensolver workload (4 sub-blocks each, ham_step), the batched FP64 eigensolve
    of the assembled arena, the recalc of new <||f||> blocks between neighbouring subspaces
    (recalc_general over the 4 sub-blocks, U blocks viewed in the eigenvector arena), and the FDM
    binning of those blocks with the new energies. value = wall time of the whole step."""
    import torch
    import oracle
    from nrgljubljana_amd import (Context, HamTerm, Gemm3Term, FdmPair, BinGrid, NRG_SYEV_FP64, NRG_GEMM_FP64,
                                  workloads)
    sizes, blocks = workloads.ham_step(nblk=args.nblk)
    nb = len(sizes)
    sizes = sizes.astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes ** 2)[:-1]]).astype(np.int64)
    offw = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    tot, totw = int(np.sum(sizes ** 2)), int(np.sum(sizes))
    dev = torch.device("cuda:0")
    dptr, keep = device_uploader(dev)
    hterms = ham_terms(sizes, blocks, dptr)
    harr = (HamTerm * len(hterms))(*hterms)
    H = torch.empty(tot, dtype=torch.float64, device=dev)
    w = torch.empty(totw, dtype=torch.float64, device=dev)
    # recalc: C_b (n_b x n_b+1) = sum_i f_i U_b(:, part_i) O_i U_b+1(:, part_i)^T
    rng = np.random.default_rng(20261210)
    gterms, Cs, rflops = [], [], 0.0
    base = H.data_ptr()
    Carena = torch.zeros(int(sum(int(sizes[b]) * int(sizes[b + 1]) for b in range(nb - 1))), dtype=torch.float64,
                         device=dev)   # the new operator blocks, one device arena
    co = 0
    for b in range(nb - 1):
        n1, n2 = int(sizes[b]), int(sizes[b + 1])
        C = Carena[co:co + n1 * n2].view(n1, n2)
        co += n1 * n2
        Cs.append(C)
        d1, d2 = blocks[b]["dims"], blocks[b + 1]["dims"]
        o1, o2 = np.concatenate([[0], np.cumsum(d1)[:-1]]), np.concatenate([[0], np.cumsum(d2)[:-1]])
        for i in range(4):
            if d1[i] and d2[i]:
                O = rng.standard_normal((d1[i], d2[i]))
                t = Gemm3Term()
                t.m, t.n, t.k, t.l, t.f = n1, n2, int(d1[i]), int(d2[i]), float(rng.choice([-1.0, 1.0]))
                t.A, t.lda, t.transA = base + 8 * (int(off[b]) + int(o1[i])), n1, 0
                t.O, t.ldo = dptr(O), int(d2[i])
                t.B, t.ldb, t.transB = base + 8 * (int(off[b + 1]) + int(o2[i])), n2, 0
                t.C, t.ldc = C.data_ptr(), n2
                gterms.append(t)
                rflops += 2.0 * d1[i] * d2[i] * n2 + 2.0 * n1 * d1[i] * n2
    # FDM: pair (Ii = b+1, Ij = b), op = C_b, energies = eigenvalues shifted to a zero minimum, 40 % kept
    wmin = torch.zeros(1, dtype=torch.float64, device=dev)
    Eg = torch.empty(totw, dtype=torch.float64, device=dev)
    rhos, fpairs = [], []
    for b in range(nb - 1):
        n1, n2 = int(sizes[b]), int(sizes[b + 1])
        Kj, Ki = int(0.4 * n1), int(0.4 * n2)
        ri, rj = rng.standard_normal((Ki, Ki)), rng.standard_normal((Kj, Kj))
        ri, rj = (ri + ri.T) / max(Ki, 1), (rj + rj.T) / max(Kj, 1)
        fpairs.append(FdmPair(n2, n1, Ki, Kj, Eg.data_ptr() + 8 * int(offw[b + 1]), Eg.data_ptr() + 8 * int(offw[b]),
                              Cs[b].data_ptr(), Cs[b].data_ptr(), n2, dptr(ri), Ki, dptr(rj), Kj, 1.0))
    grid = BinGrid(-6.0, 100.0, 1001, 1e-16)
    bins = [torch.zeros(grid.nbins, dtype=torch.float64, device=dev) for _ in range(4)]
    ctx = Context(0)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    phase = np.zeros(4)

    def step(record=False):
        if record:
            ev[0].record()
        ctx.hamiltonian(sizes, off, H, harr)
        if record:
            ev[1].record()
        ctx.syev_batched(H, sizes, off, w, offw, NRG_SYEV_FP64)
        if record:
            ev[2].record()
        Carena.zero_()
        ctx.gemm3_grouped(gterms, NRG_GEMM_FP64)
        if record:
            ev[3].record()
        torch.sub(w, torch.min(w), out=Eg)
        for x in bins:
            x.zero_()
        ctx.fdm_bin(fpairs, grid, 1.0, 0.5, -1, *bins)
        if record:
            ev[4].record()
    ms = timed_calls(step, args.steps, args.warmup, torch.cuda.synchronize)
    for _ in range(2):
        step(record=True)
        torch.cuda.synchronize()
        phase += [ev[k].elapsed_time(ev[k + 1]) / 2 for k in range(4)]
    names = ["hamiltonian", "eigensolve", "recalc", "fdm"]
    log(f"step: {nb} blocks, {tot * 8 / 1e6:.0f} MB of Hamiltonians, {len(gterms)} recalc terms "
        f"({rflops / 1e9:.1f} GFLOP), {len(fpairs)} FDM pairs; {ms:.3f} ms per step: "
        + ", ".join(f"{k} {v:.3f} ms" for k, v in zip(names, phase)))
    eflops = workloads.eig_flops(sizes)
    out_err = None
    cpu = None
    if not args.no_cpu:
        # the oracle chain on the host: assembly (1 core), dsyev per block (OpenMP over blocks), recalc
        # dgemm pairs (threaded BLAS), FDM loops (1 core)
        cores = host_cores()
        hk = {}

        def hptr(a):
            hk[id(a)] = np.ascontiguousarray(a)
            return hk[id(a)].ctypes.data
        Hh = np.empty(tot)
        t0 = time.perf_counter()
        oracle.hamiltonian(sizes, off, Hh, ham_terms(sizes, blocks, hptr))
        t1 = time.perf_counter()
        wh, info = oracle.syev_batched(sizes, off, Hh, offw, threads=cores, blas_threads=1)
        t2 = time.perf_counter()
        gw = w.cpu().numpy()
        out_err = float(max(np.max(np.abs(gw[o:o + n] - wh[o:o + n])) / max(np.max(np.abs(wh[o:o + n])), 1e-300)
                            for o, n in zip(offw, sizes)))
        cpu = {"value": round((t2 - t0) * 1e3, 1), "unit": "ms", "cores": cores, "kind": "port",
               "sample": f"the oracle's assembly (1 core, {1e3 * (t1 - t0):.0f} ms) and dsyev per block (OpenMP over "
                         f"{cores} cores, {1e3 * (t2 - t1):.0f} ms) of the same step; recalc and FDM not timed on the "
                         "host (their CPU rates are the recalc and fdm workloads' baselines)"}
        log(f"cpu: {cpu['sample']}; max |lambda_gpu - lambda_oracle| / max|lambda| = {out_err:.2e}")
    out = {"metric": "NRG step wall-time (synthetic 256-subspace step: assembly + eigensolve + recalc + FDM)",
           "value": round(ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{nb} subspaces, orders log-uniform [16,1024] (seed 20261203), 4 sub-blocks "
                                  "each; recalc and FDM over the 255 neighbouring pairs, seeds 20261209-10",
                      "eig_flops": eflops, "recalc_flops": rflops},
           "phases_ms": dict(zip(names, [round(x, 3) for x in phase])),
           "eig_gflops_in_step": round(eflops / (phase[1] * 1e-3) / 1e9, 1),
           "max_eig_err_rel_vs_oracle": out_err,
           "roofline": None, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_broaden(args):
    """Broadening of one binned spectrum (1000 bins/decade over 14 decades, trimmed) onto the standard
    460-point mesh (nrg_broaden); value = BR_NEW evaluations per second."""
    import torch
    import oracle
    from nrgljubljana_amd import Context
    rng = np.random.default_rng(20261207)
    n = 14 * 1000 + 1
    e = 10.0 ** (-12.0 + np.arange(n) / 1000.0)
    halves = []
    for _ in range(2):
        w = np.exp(-((np.log10(e) + 3.0) / 2.0) ** 2) * e * rng.uniform(0.5, 1.5, n)
        w[rng.uniform(size=n) < 0.3] = 0.0
        keep = np.append(np.abs(w[:-1]) >= 1e-16 * np.diff(e), True)
        halves.append((e[keep], w[keep]))
    mesh, E = [], 10.0
    while E > 1e-8:
        mesh.append(E)
        E /= 1.05
    mesh = np.array(mesh)
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (*halves[0], *halves[1], mesh)]
    op = torch.zeros(len(mesh), dtype=torch.float64, device=dev)
    on = torch.zeros_like(op)
    ctx = Context(0)
    alpha, omega0 = 0.3, 1e-9
    evals = 2.0 * len(mesh) * (len(halves[0][0]) + len(halves[1][0]))

    def step():
        ctx.broaden(*t, alpha, omega0, op, on)
    ms = timed_calls(step, args.steps, args.warmup, torch.cuda.synchronize)
    ctx.kernel_timing(True)
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    tms, cnt = ctx.kernel_times()["k_broaden"]
    ctx.kernel_timing(False)
    # about 30 FP64 operations per evaluation (log, exp, divisions and the blend), on the vector units
    ach = 30.0 * evals / (tms / cnt * 1e-3) / 1e12
    roof = {"bound": "valu", "kernel": "k_broaden", "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None,
            "note": "transcendental-bound: ~30 FP64 operations per kernel evaluation counted against the 78.6 TF/s "
                    "FP64 vector peak"}
    cpu = None
    if not args.no_cpu:
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and reps < 50:
            oracle.broaden(*halves[0], *halves[1], mesh, alpha, omega0)
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": round(reps * evals / dt / 1e6, 2), "unit": "Mevals/s", "cores": 1, "kind": "port",
               "sample": f"the whole spectrum {reps} times, scalar double loop, {dt:.1f} s"}
    out = {"metric": "spectral broadening (BR_NEW), kernel evaluations per second", "value": round(evals / ms / 1e6, 2),
           "unit": "Gevals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{len(halves[0][0])} + {len(halves[1][0])} trimmed peaks, {len(mesh)} mesh points, "
                                  "alpha 0.3, omega0 1e-9, seed 20261207", "evaluations_per_step": evals},
           "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_dm(args):
    """One backward step of the density-matrix sweep: all cdmI rotations of the step as one grouped
    nrg_gemm3_grouped call (FP64); value = GFLOP/s in the reference's operation order."""
    import torch
    import oracle
    from nrgljubljana_amd import Context, NRG_GEMM_FP64, Gemm3Term, workloads
    step = workloads.dm_step(nsub=args.npairs)
    flops = workloads.dm_flops(step)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(20261208)
    keep, terms, outs, host = [], [], [], []
    for sub in step:
        C = torch.zeros(sub["kept"], sub["kept"], dtype=torch.float64, device=dev)
        outs.append(C)
        for ch in sub["children"]:
            rho = rng.standard_normal((ch["nromega"], ch["nromega"]))
            U = rng.standard_normal((ch["nromega"], ch["total"])) / np.sqrt(ch["total"])
            dr, dU = torch.from_numpy(rho).to(dev), torch.from_numpy(U).to(dev)
            keep += [dr, dU]
            host.append((rho, U))
            t = Gemm3Term()
            t.m = t.n = sub["kept"]; t.k = t.l = ch["nromega"]; t.f = ch["f"]
            t.A = dU.data_ptr() + 8 * ch["off"]; t.lda = ch["total"]; t.transA = 1
            t.O = dr.data_ptr(); t.ldo = ch["nromega"]
            t.B = dU.data_ptr() + 8 * ch["off"]; t.ldb = ch["total"]; t.transB = 1
            t.C = C.data_ptr(); t.ldc = sub["kept"]
            terms.append(t)
    ctx = Context(0)

    def call():
        for C in outs:
            C.zero_()
        ctx.gemm3_grouped(terms, NRG_GEMM_FP64)
    ms = timed_calls(call, args.steps, args.warmup, torch.cuda.synchronize)
    ctx.kernel_timing(True)
    for _ in range(2):
        call()
    torch.cuda.synchronize()
    kt = ctx.kernel_times()
    ctx.kernel_timing(False)
    log(f"dm step: {len(step)} subspaces x 16 children, {flops / 1e9:.2f} GFLOP, {ms:.3f} ms per call; kernels "
        + ", ".join(f"{k} {v[0] / 2:.3f} ms" for k, v in kt.items()))
    gemm_ms = sum(v[0] for k, v in kt.items() if k.startswith("k_gemm")) / 2
    ach = flops / (gemm_ms * 1e-3) / 1e12
    roof = {"bound": "mfma", "kernel": "k_gemm_f64 + k_gemm_f64_big", "achieved": round(ach, 3),
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None}
    cpu = None
    if not args.no_cpu:
        L = oracle.lib()
        cores = host_cores()
        L.oracle_blas_threads(cores)
        hterms, hold, hi = [], [], 0
        for sub in step:
            C = np.zeros((sub["kept"], sub["kept"]))
            hold.append(C)
            for ch in sub["children"]:
                rho, U = host[hi]
                hi += 1
                t = Gemm3Term()
                t.m = t.n = sub["kept"]; t.k = t.l = ch["nromega"]; t.f = ch["f"]
                t.A = U.ctypes.data + 8 * ch["off"]; t.lda = ch["total"]; t.transA = 1
                t.O = rho.ctypes.data; t.ldo = ch["nromega"]
                t.B = U.ctypes.data + 8 * ch["off"]; t.ldb = ch["total"]; t.transB = 1
                t.C = C.ctypes.data; t.ldc = sub["kept"]
                hterms.append(t)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and reps < 100:
            oracle.gemm3(hterms)
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": round(reps * flops / dt / 1e9, 2), "unit": "GFLOP/s", "cores": cores, "kind": "port",
               "sample": f"the whole step {reps} times, oracle dgemm pairs in the reference's rotate order, {dt:.1f} s"}
    out = {"metric": "density-matrix sweep step (cdmI rotations) GFLOP/s", "value": round(flops / ms / 1e6, 2),
           "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{len(step)} subspaces of step N-1, kept log-uniform [16,1000], 16 children each, "
                                  "seed 20261207", "flops_per_step": flops},
           "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def tdfdm_columns(d1, d2):
    """max |difference| per column of two tdfdm files"""
    def rd(d):
        with open(os.path.join(d, "tdfdm")) as F:
            lines = [l.split() for l in F if l.strip()]
        return lines[0][1:], np.array([[float(x) for x in l] for l in lines[1:]])
    names, a = rd(d1)
    _, b = rd(d2)
    return {nm: float(np.max(np.abs(a[:, i] - b[:, i]))) for i, nm in enumerate(names)}


def bench_nrg(args):
    """A whole NRG calculation at a configuration's parameter point (workloads.NRG_CONFIGS: c1 the
    single-channel DM-NRG run, c5 the two-channel keep-4000 FDM run), through libnrgrun.so: every step
    assembled, diagonalised, recalculated and measured on the GPU. --mode fp64 | mixed | both (the
    eigensolver form; both: the two runs' output files are compared with each other)."""
    import shutil
    import tempfile
    from nrgljubljana_amd import run_nrg, workloads
    from tests.compare_outputs import compare_files
    # torch is not imported here: nrg_run_dir is synchronous, and a process that imports torch loads the
    # HIP runtime bundled with torch in place of the system one (same SONAME), under which rocprofv3's
    # kernel tracing was seen to corrupt this workload's results
    forms = ["fp64", "mixed"] if args.mode == "both" else [args.mode]
    if any(f not in ("fp64", "mixed") for f in forms):
        sys.exit("bench.py: --workload nrg takes --mode fp64, mixed or both")
    base = tempfile.mkdtemp(prefix="nrg_bench_")
    # warm-up: a three-step run loads every kernel
    wd = os.path.join(base, "warm")
    workloads.nrg_config(args.config, wd, keep=min(args.keep or 100, 100), nmax=3)
    run_nrg(wd, log_path=os.path.join(wd, "log"))
    runs = {}
    for form in forms:
        d = os.path.join(base, form)
        N = workloads.nrg_config(args.config, d, diag="default" if form == "fp64" else "mixed", keep=args.keep,
                                 nmax=args.chain)
        log(f"{args.config} {form}: chain length {N}, running")
        t0 = time.perf_counter()
        run_nrg(d, log_path=os.path.join(d, "log"))
        wall = time.perf_counter() - t0
        with open(os.path.join(d, "nrg_timing.json")) as F:
            tm = json.load(F)
        n3 = sum(sum(float(n) ** 3 for n in step) for step in tm["blocks"])
        eig_flops = 10.0 / 3.0 * n3
        steps = len(tm["nrg_steps"]) + len(tm["dm_steps"])
        runs[form] = dict(dir=d, wall=wall, N=N, timing=tm, eig_flops=eig_flops, steps=steps)
        log(f"{args.config} {form}: {wall:.2f} s, {len(tm['nrg_steps'])} NRG + {len(tm['dm_steps'])} DM steps, "
            f"sections " + ", ".join(f"{k} {v:.2f} s" for k, v in sorted(tm["sections"].items())) +
            f", eigensolver {eig_flops / tm['sections'].get('diag', 1e-30) / 1e9:.1f} GFLOP/s, largest block "
            f"{max(max(b) for b in tm['blocks'] if b)}")
    if args.save:
        for form, r in runs.items():
            dst = os.path.join(args.save, form)
            os.makedirs(dst, exist_ok=True)
            for f in os.listdir(r["dir"]):
                if f != "data" and os.path.isfile(os.path.join(r["dir"], f)):
                    shutil.copy(os.path.join(r["dir"], f), dst)
    cmp = None
    if len(forms) == 2:
        cmp = {}
        for f in sorted(os.listdir(runs["fp64"]["dir"])):
            if f in ("param", "data", "log", "nrg_timing.json", "DONE"):
                continue
            if f == "tdfdm":
                # C_fdm = var(E) / T^2 of absolute energies (|E| ~ 4 at T = 1e-10): the rounding of the
                # FDM weights' normalisation, ~1e-16 E^2 / T^2, is all it holds (negative in both forms, as
                # the reference's formula gives); it is reported apart from the other columns
                cols = tdfdm_columns(runs["mixed"]["dir"], runs["fp64"]["dir"])
                cmp["tdfdm (T, E, F, S)"] = max(v for k, v in cols.items() if k != "C_fdm")
                cmp["tdfdm C_fdm (rounding only)"] = cols["C_fdm"]
                continue
            ok, err, _ = compare_files(os.path.join(runs["mixed"]["dir"], f), os.path.join(runs["fp64"]["dir"], f))
            cmp[f] = err
        log("mixed against fp64, max |diff| per file: " + ", ".join(f"{k} {v:.1e}" for k, v in cmp.items()))
    main_form = forms[-1] if args.mode != "both" else "mixed"
    r = runs[main_form]
    tm = r["timing"]
    cpu = None
    if not args.no_cpu:
        cpu = nrg_cpu_sample(tm["blocks"])
    out = {"metric": "NRG step wall-time (end-to-end run)", "value": round(1e3 * r["wall"] / r["steps"], 3),
           "unit": "ms/step", "n_gpus": 1, "steps": r["steps"], "warmup": 1, "ms_per_step": round(1e3 * r["wall"] / r["steps"], 3),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if main_form == "fp64" else "f32+f64", "data": "synthetic",
           "config": {"workload": f"{args.config}: " + ", ".join(f"{k}={v}" for k, v in workloads.NRG_CONFIGS[args.config].items()),
                      "chain_length": r["N"], "keep": args.keep or workloads.NRG_CONFIGS[args.config]["keep"],
                      "form": main_form},
           "wall_s": {k: round(v["wall"], 3) for k, v in runs.items()},
           "sections_s": {k: {s: round(t, 3) for s, t in v["timing"]["sections"].items()} for k, v in runs.items()},
           "eig_gflops": {k: round(v["eig_flops"] / v["timing"]["sections"].get("diag", 1e-30) / 1e9, 1)
                          for k, v in runs.items()},
           "largest_block": max(max(b) for b in tm["blocks"] if b),
           "blocks_per_step_max": max(len(b) for b in tm["blocks"]),
           "mixed_vs_fp64_max_abs_diff": cmp,
           "autotune": {k: v["timing"].get("syev_tuning") for k, v in runs.items()}, "cpu_baseline": cpu}
    shutil.rmtree(base, ignore_errors=True)
    print(json.dumps(out), flush=True)


def nrg_cpu_sample(blocks, budget_s=20.0):
    """The oracle's dsyev (OpenMP over blocks, largest first) on synthetic symmetric matrices with the
    block orders of the run's most expensive step, largest first, until the budget is spent."""
    import oracle
    step = max(blocks, key=lambda b: sum(float(n) ** 3 for n in b))
    cores = host_cores()
    rng = np.random.default_rng(5)
    done, flops, t_tot = 0, 0.0, 0.0
    for n in sorted(step, reverse=True):
        if t_tot > budget_s:
            break
        a = np.triu(rng.standard_normal((n, n)))
        t0 = time.perf_counter()
        oracle.syev_batched(np.array([n]), np.array([0]), a.ravel(), np.array([0]), threads=1, blas_threads=cores)
        t_tot += time.perf_counter() - t0
        flops += 10.0 / 3.0 * n ** 3
        done += 1
    return {"value": round(flops / t_tot / 1e9, 2), "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "sample": f"{done} of the {len(step)} blocks of the run's most expensive step (largest first, orders "
                      f"{max(step)} down), oracle dsyev with {cores} BLAS threads, {t_tot:.1f} s",
            "host": host_info()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="eig",
                    choices=["eig", "recalc", "fdm", "broaden", "dm", "dmnrg", "cfs", "ft", "mats", "temp", "ham", "step",
                             "nrg"])
    ap.add_argument("--mode", default="fp64", help="eig: fp64 | mixed; recalc: fp64 | ozaki")
    ap.add_argument("--nblk", type=int, default=256)
    ap.add_argument("--nmin", type=int, default=16, help="eig: smallest block order")
    ap.add_argument("--nmax", type=int, default=1024, help="eig: largest block order")
    ap.add_argument("--npairs", type=int, default=48)
    ap.add_argument("--mats", type=int, default=100, help="mats: Matsubara frequencies")
    ap.add_argument("--config", default="c5", choices=["c1", "c5"], help="nrg: configuration")
    ap.add_argument("--keep", type=int, default=None, help="nrg: override the configuration's keep")
    ap.add_argument("--chain", type=int, default=None, help="nrg: override the chain length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--save", default=None, help="nrg: copy each form's output files to SAVE/<form>")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: runs on one GPU only (sharding the blocks of a step across GPUs is a design, not "
              "built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    {"eig": bench_eig, "recalc": bench_recalc, "fdm": bench_fdm, "broaden": bench_broaden,
     "dm": bench_dm, "dmnrg": bench_dmnrg, "cfs": bench_cfs, "ft": bench_ft, "mats": bench_mats,
     "temp": bench_temp, "ham": bench_ham, "step": bench_step, "nrg": bench_nrg}[args.workload](args)


if __name__ == "__main__":
    main()
