"""Benchmark of the GPU-HeiPa multilevel hot path on one MI355X.

Default workload ("headline"): the metric the reference's benchmark
configuration names, input edges/s of the batched and served forms against
the plain drop-in, with p50/p99 latency at equal cut:

  * batched form: the 4096 independent graphs of the batched configuration
    (RGG of average degree 8 or R-MAT of edge factor 8, each with probability
    0.5, n log-uniform in [1024, 65536], k from {2,4,8,16,32}, eps 0.03,
    `default`) through the worker pool.  The K timed steps together process
    every graph exactly once (ceil(4096 / K) graphs per step); `value` is
    their undirected input edges / the time of the K steps.  Before the
    warmup, the pool's autotuner picks the lane-group caps and the
    single-workgroup threshold on a sample of 64 of the graphs (launch shapes
    only: results do not change).  The plain path
    (one context, graphs one after another) is timed on the graphs of the
    first timed step beside it, and every one of those partitions must be
    bit-identical in both forms (equal cut);
  * served form: a Poisson-arrival sample of the served configuration's
    request mix (partition and HM requests, n capped, both stated in the
    line) at 0.8x the plain path's measured request capacity: p50/p99
    latency, edges/s, saturated capacity;
  * the plain drop-in on the RGG n = 2^20 `ultra` k = 64 workload: its
    edges/s, the roofline of its dominant kernel and the per-kernel table
    (from event-timed steps outside the timed region);
  * cpu_baseline: the serial C++ oracle on the first 1024 batched graphs,
    one graph per host thread.

Other workloads (--workload) time one configuration alone: the plain
drop-in (rgg20-ultra-k64, grid512-default-k32, rgg16-ultra-k64), IM and HM
process mapping (im-rgg20-486, hm-rmat22-486), and the pool forms alone
(batched, served).

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-heipa_amd", "python"))

WORKLOADS = {
    # batched form vs plain, served sample, plain rgg20-ultra-k64 roofline (the default)
    "headline": dict(kind="headline"),
    "rgg20-ultra-k64": dict(kind="rgg", n=1 << 20, deg=12.0, seed=1, wmod=10, k=64, ultra=True),
    "grid512-default-k32": dict(kind="grid", rows=512, cols=512, k=32, ultra=False),
    "rgg16-ultra-k64": dict(kind="rgg", n=1 << 16, deg=12.0, seed=1, wmod=10, k=64, ultra=True),
    # IM process mapping (GPU-HeiProMap --config IM) onto 4:8:6 with distances 1:10:100
    "im-rgg20-486": dict(kind="rgg", n=1 << 20, deg=12.0, seed=1, wmod=10, k=192, ultra=False,
                         hier=[4, 8, 6], dist=[1, 10, 100]),
    # GPU-HeiProMap --config HM onto 4:8:6 / 1:10:100 on an R-MAT graph of scale 22, edge factor 16
    "hm-rmat22-486": dict(kind="rmat", scale=22, ef=16.0, seed=1, wmod=100, k=192, ultra=False,
                          hier=[4, 8, 6], dist=[1, 10, 100], hm=True, cpu_scale=18),
    # batched form: independent small graphs through a worker pool
    "batched": dict(kind="batched"),
    # served form: Poisson arrivals of partition and mapping requests
    "served": dict(kind="served"),
}


def make_graph(w, scale=None):
    import heipa
    if w["kind"] == "rgg":
        return heipa.gen_rgg(w["n"], w["deg"], w["seed"], w["wmod"])
    if w["kind"] == "rmat":
        return heipa.gen_rmat(scale or w["scale"], 0, w["ef"], 0.57, 0.19, 0.19, w["seed"], w["wmod"], False)
    return heipa.gen_grid(w["rows"], w["cols"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="headline", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-all", action="store_true", help="print the per-kernel event profile to stderr")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 256 MiB device to device (the known-size kernel that scales HBM counters)")
    ap.add_argument("--slots", type=int, default=8, help="batched: pool workers (and served, alone)")
    ap.add_argument("--served-slots", type=int, default=12,
                    help="headline: the served leg's pool workers (measured best for the served mix: 21.2M "
                         "edges/s at 12 against 20.1M at 8, 20.7M at 16, 16.5M at 6)")
    ap.add_argument("--hm-workers", type=int, default=None,
                    help="HM: worker contexts for the subtrees below the first split (default: the library's)")
    ap.add_argument("--block-jet-n", type=int, default=None,
                    help="largest level run by the single-workgroup Jet (default: the library's)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="headline: skip the pool's launch-shape autotuner (defaults stay)")
    ap.add_argument("--tune-graphs", type=int, default=64, help="headline: batched graphs the autotuner times")
    ap.add_argument("--batch-count", type=int, default=4096, help="batched: number of graphs")
    ap.add_argument("--plain-sample", type=int, default=256, help="batched: graphs timed through the plain path")
    ap.add_argument("--batch-order", choices=["seq", "size"], default="seq",
                    help="headline: graphs in generation order or largest first")
    ap.add_argument("--requests", type=int, default=200, help="served: number of requests")
    ap.add_argument("--served-nmax", type=float, default=1e6, help="served: largest request n")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads for generating graphs and for the CPU baseline")
    ap.add_argument("--cpu-graphs", type=int, default=1024, help="headline: batched graphs the CPU baseline solves")
    ap.add_argument("--served-repeat", type=int, default=0,
                    help="served: draw the requests' graphs from this many distinct graphs and keep their "
                         "hierarchies resident (0: every request on its own graph)")
    ap.add_argument("--dump-peaks", default=None, help="write (n, m, spec, estimate, arena peak) of every pool job "
                                                      "to PATH.batched.json / PATH.served.json")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path runs on one "
              f"GPU; independent graphs and requests would shard one context per GPU)", file=sys.stderr)
        sys.exit(2)
    if WORKLOADS[args.workload]["kind"] == "headline":
        return run_headline(args)
    if WORKLOADS[args.workload]["kind"] == "served":
        r = served_sample(args)
        r.update({"metric": "input edges/sec on 1 MI355X, served form", "n_gpus": 1, "steps": 1, "warmup": 1,
                  "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                  "dtype": "int32/int64", "data": "synthetic (seeded generators)"})
        print(json.dumps(r), flush=True)
        return
    if WORKLOADS[args.workload]["kind"] == "batched":
        return run_pool_workload(args)

    if args.calibrate_copy:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    res, g, stats, ctx = measure_single(args, args.workload, args.steps, args.warmup)
    w = WORKLOADS[args.workload]
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_single(args, w, g, stats, ctx)
    ctx.close()
    print(json.dumps(res), flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measure_single(args, wname, steps, warmup, profile=True):
    """Times `steps` solves of one workload (after `warmup` untimed ones) with
    the shipping settings (Jet batches replayed from HIP graphs, no event
    timing), then profiles one more solve for the per-kernel table and the
    dominant kernel's roofline.  Returns (result dict, host graph, stats of the
    last timed solve, context)."""
    import numpy as np
    import torch
    import heipa

    w = WORKLOADS[wname]
    g = make_graph(w)
    m_und = g.m // 2
    ctx = heipa.Context(0, 16 << 30)
    if args.block_jet_n is not None:
        ctx.set_option(heipa.OPT_BLOCK_JET, args.block_jet_n)
    if args.hm_workers is not None:
        ctx.set_option(heipa.OPT_HM_WORKERS, args.hm_workers)
    dg = heipa.DeviceGraph(g, "cuda:0")
    part = torch.empty(g.n, dtype=torch.int32, device="cuda:0")
    k, ultra = w["k"], w["ultra"]
    hier = w.get("hier")
    hm = w.get("hm", False)

    def step(dgraph=dg, out=part):
        if hm:
            return ctx.hm_partition(dgraph, hier, 0.03, args.seed, False, part=out), None, None
        if hier:
            return ctx.im_partition(dgraph, hier, w["dist"], 0.03, args.seed, part=out)
        return ctx.partition(dgraph, k, 0.03, args.seed, ultra, part=out)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = None
    for _ in range(steps):
        _, stats, _ = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / steps
    value = m_und / (ms_per_step / 1000.0)

    roof, per_kernel = None, {}
    if profile:
        # one solve with every launch event-timed: the per-kernel table and the
        # dominant kernel (graph replay is off while profiling, so short
        # launches read slower here than in the timed steps)
        ctx.prof_enable(True)
        ctx.prof_filter("")
        ctx.prof_reset()
        step()
        prof_all = ctx.prof_stats()
        ctx.prof_enable(False)
        if args.profile_all:
            for name, s in sorted(prof_all.items(), key=lambda kv: -kv[1]["ms"]):
                print(f"  {name:32s} launches={s['launches']:7d} ms={s['ms']:9.3f} "
                      f"GB/s={s['bytes'] / max(s['ms'], 1e-9) / 1e6:9.1f}", file=sys.stderr)
        # launches of the finest level also have entries "<name>@L0": every
        # level together, and level 0 alone
        combined = {}
        for name, s in prof_all.items():
            c = combined.setdefault(name[:-3] if name.endswith("@L0") else name,
                                    {"launches": 0, "ms": 0.0, "bytes": 0.0})
            for f in c:
                c[f] += s[f]
        accounted = {n: s for n, s in combined.items() if s["bytes"] > 0}
        dom = max(accounted, key=lambda n: accounted[n]["ms"]) if accounted else None
        top = sorted(combined.items(), key=lambda kv: -kv[1]["ms"])[:10]
        level0 = sorted((kv for kv in prof_all.items() if kv[0].endswith("@L0")), key=lambda kv: -kv[1]["ms"])
        for name, s in top + level0:
            gbs = s["bytes"] / max(s["ms"], 1e-9) / 1e6 if s["bytes"] > 0 else None
            per_kernel[name] = {"launches": s["launches"], "ms": round(s["ms"], 3),
                                "avg_us": round(s["ms"] * 1000.0 / max(s["launches"], 1), 3),
                                "alg_GBps": None if gbs is None else round(gbs, 1),
                                "frac": None if gbs is None else round(gbs / 8000.0, 5)}
        if dom:
            # the dominant kernel alone: HIP events on the context stream around
            # about one launch in 8 (its algorithmic bytes are counted for every
            # launch), over two solves
            ctx.prof_filter(dom)
            ctx.set_option(heipa.OPT_PROF_SAMPLE, 8)
            ctx.prof_reset()
            ctx.prof_enable(True)
            for _ in range(2):
                step()
            prof = ctx.prof_stats()
            ctx.prof_enable(False)
            ctx.set_option(heipa.OPT_PROF_SAMPLE, 1)
            s = None
            for name in (dom, dom + "@L0"):  # every level of the dominant kernel
                if name in prof:
                    s = dict(prof[name]) if s is None else {f: s[f] + prof[name][f] for f in s}
            if s and s["ms"] > 0:
                achieved = s["bytes"] / (s["ms"] / 1000.0) / 1e9
                # HBM bytes per launch of this kernel from the committed rocprofv3
                # FETCH_SIZE / WRITE_SIZE passes over this workload (bench.py cannot
                # read hardware counters itself)
                traffic = None
                tpath = os.path.join(ROOT, "profiles", f"traffic-{wname}.json")
                if os.path.exists(tpath):
                    with open(tpath) as fh:
                        t = json.load(fh)
                    if t.get("kernel") == dom:
                        traffic = t["hbm_bytes_per_launch"]
                roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(achieved / 8000.0, 5), "traffic": traffic, "kernel": dom,
                        "launches": s["launches"], "timed_fraction": 1 / 8,
                        "avg_us": round(s["ms"] * 1000.0 / max(s["launches"], 1), 3),
                        "alg_bytes_per_launch": round(s["bytes"] / max(s["launches"], 1), 1),
                        "workload": wname}

    # the same solve from host CSR to host partition array (upload, solve,
    # download): the PCIe-inclusive rate, reported beside `value`, never as it
    torch.cuda.synchronize()
    th0 = time.perf_counter()
    dg_h = heipa.DeviceGraph(g, "cuda:0")
    part_h, _, _ = step(dg_h, None)
    host_part = heipa.to_numpy_u32(part_h)
    th1 = time.perf_counter()
    del dg_h, part_h
    pcie = {"ms_per_step": round((th1 - th0) * 1000.0, 3), "value": round(m_und / (th1 - th0), 1),
            "includes": "H2D upload of the CSR, solve, D2H download of the partition (one step)"}

    # quality beside the number
    p = heipa.to_numpy_u32(part)
    assert np.array_equal(p, host_part)
    vw = np.ones(g.n, np.int64) if g.vwgt is None else g.vwgt.astype(np.int64)
    bw = np.bincount(p, weights=vw, minlength=k)
    if hm:  # the HM recursion reports no solver statistics: J and balance from the result
        import math
        D = heipa.distance_matrix(hier, w["dist"])
        Dd = heipa.from_numpy_u32(D.reshape(-1).view(np.uint32))
        lmax = math.ceil(1.03 * int(vw.sum()) / k)
        quality = {"comm_cost": int(ctx.comm_cost(dg, part, Dd, k)), "maxw": int(bw.max()), "lmax": lmax,
                   "balanced": bool(bw.max() <= lmax)}
    else:
        quality = {"comm_cost" if hier else "cut": int(stats.cut), "maxw": int(stats.maxw),
                   "lmax": int(stats.lmax), "balanced": bool(bw.max() <= stats.lmax), "levels": int(stats.levels),
                   "jet_iterations": int(stats.jet_iterations)}
    result = {
        "metric": "input edges/sec on 1 MI355X (" + ("HM process mapping" if hm else "IM process mapping" if hier
                                                     else "plain drop-in") + ")",
        "value": round(value, 1),
        "unit": "input edges/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32/int64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": wname, "n": g.n, "undirected_edges": m_und, "k": k,
                   "config": "HM" if hm else "IM" if hier else ("ultra" if ultra else "default"),
                   "imbalance": 0.03, "seed": args.seed,
                   **({"hierarchy": hier, "distance": w["dist"]} if hier else {}),
                   "input": "device-resident CSR", "phase_ms": None if stats is None else {
                       "coarsening": round(stats.ms_coarsening, 2), "contraction": round(stats.ms_contraction, 2),
                       "initial": round(stats.ms_initial, 2), "uncontraction": round(stats.ms_uncontraction, 2),
                       "refinement": round(stats.ms_refinement, 2)}},
        "quality": quality,
        "roofline": roof,
        "per_kernel": per_kernel,
        "pcie_inclusive": pcie,
    }
    return result, g, stats, ctx


def cpu_single(args, w, g, stats, ctx):
    """The serial oracle on the same workload (HM: on a smaller R-MAT graph of
    the same generator), on one host core."""
    import numpy as np
    import heipa
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import heipa_oracle as orc
    hier, hm, k, ultra = w.get("hier"), w.get("hm", False), w["k"], w["ultra"]
    if hm:
        gs = make_graph(w, w["cpu_scale"])
        tc = time.perf_counter()
        cp = orc.hm_partition(gs, hier, 0.03, args.seed, False)
        dt = time.perf_counter() - tc
        dgs = heipa.DeviceGraph(gs, "cuda:0")
        gp = heipa.to_numpy_u32(ctx.hm_partition(dgs, hier, 0.03, args.seed, False))
        del dgs
        return {"value": round((gs.m // 2) / dt, 1), "unit": "input edges/s", "cores": 1, "kind": "port",
                "cpu": cpu_model(),
                "sample": f"R-MAT scale {w['cpu_scale']} (same generator parameters, {gs.m // 2} edges), one HM "
                          f"solve by the serial C++ oracle (deterministic mode), {dt:.1f} s",
                "same_result_as_gpu": bool(np.array_equal(cp, gp))}
    tc = time.perf_counter()
    if hier:
        _, ccut, cmaxw, _ = orc.im_partition(g, hier, w["dist"], 0.03, args.seed)
    else:
        _, ccut, cmaxw, _ = orc.partition(g, k, 0.03, args.seed, ultra)
    dt = time.perf_counter() - tc
    return {"value": round((g.m // 2) / dt, 1), "unit": "input edges/s", "cores": 1, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"the full workload, one solve by the serial C++ oracle (deterministic mode), {dt:.1f} s",
            "cut": int(ccut), "maxw": int(cmaxw),
            "same_result_as_gpu": bool(ccut == stats.cut and cmaxw == stats.maxw)}


def run_headline(args):
    """The default line: batched vs plain on the batched configuration's 4096
    graphs, a served sample, the plain rgg20-ultra-k64 roofline and the CPU
    baseline (see the module docstring)."""
    import math
    import numpy as np
    import torch
    import heipa
    from heipa import workloads

    K, W = args.steps, args.warmup
    tg = time.perf_counter()
    reqs = workloads.batched(args.batch_count, threads=args.threads)
    dgs = [heipa.DeviceGraph(g, "cuda:0") for g, _, _ in reqs]
    gen_s = time.perf_counter() - tg
    edges = [g.m // 2 for g, _, _ in reqs]
    # steps take the graphs in generation order (--batch-order size: largest
    # first) and are submitted back to back: a step's graphs start as workers
    # free up, the timed region ends when all are done
    order = (sorted(range(len(reqs)), key=lambda i: -reqs[i][0].m) if args.batch_order == "size"
             else list(range(len(reqs))))
    per = math.ceil(len(reqs) / K)
    batches = [order[s * per:(s + 1) * per] for s in range(K)]
    probe = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **spec) for dg, (_, spec, _) in zip(dgs, reqs)]
    pool0 = heipa.Pool(1, 1 << 20)
    arena = int(max(pool0.estimate(j) for j in probe) * 1.05) + (32 << 20)
    pool0.close()
    del probe
    parts = [torch.empty(max(dg.n, 1), dtype=torch.int32, device="cuda:0") for dg in dgs]

    def jobs_of(idx):
        return [heipa.Pool.job(dgs[i], imbalance=0.03, seed=1, part=parts[i], **reqs[i][1]) for i in idx]

    pool = heipa.Pool(args.slots, arena)
    if args.block_jet_n is not None:
        pool.set_option(heipa.OPT_BLOCK_JET, args.block_jet_n)
    tune = None
    if not args.no_autotune:
        # launch shapes that never change results, timed on an evenly spaced
        # sample of the graphs before the warmup; the plain path below gets the
        # same setting, so the speedup is the batching alone
        tt = time.perf_counter()
        stride = max(1, len(reqs) // max(1, args.tune_graphs))
        tune = pool.autotune(jobs_of(list(range(0, len(reqs), stride))[:args.tune_graphs]))
        tune["seconds"] = round(time.perf_counter() - tt, 2)
        tune["sample_graphs"] = min(args.tune_graphs, len(range(0, len(reqs), stride)))
    for s in range(W):
        pool.run(jobs_of(batches[s % K]))
    timed = [jobs_of(b) for b in batches]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for jobs in timed:
        for j in jobs:
            pool.submit(j)
    for jobs in timed:
        for j in jobs:
            pool.wait(j)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = sum(edges) / wall
    done = [j for jobs in timed for j in jobs]
    done_idx = [i for b in batches for i in b]
    failed = sum(1 for j in done if j.status != heipa.OK)
    unbalanced = sum(1 for j, i in zip(done, done_idx)
                     if j.maxw > math.ceil(1.03 * reqs[i][0].g_weight / reqs[i][1]["k"]))
    est_ratio = estimate_ratio(done)
    if args.dump_peaks:
        dump_peaks(args.dump_peaks + ".batched.json", [(reqs[i][0], reqs[i][1]) for i in done_idx], done)

    # the same graphs in both forms: every 20th graph of the sequence (sizes
    # across the whole range), batched through the pool, then plain (one
    # context, one graph after another)
    sample = list(range(0, len(reqs), 20))
    es = sum(edges[i] for i in sample)
    batched_parts = {i: heipa.to_numpy_u32(parts[i])[:dgs[i].n].copy() for i in sample}
    sjobs = jobs_of(sorted(sample, key=lambda i: -reqs[i][0].m))
    torch.cuda.synchronize()
    ts = time.perf_counter()
    pool.run(sjobs)
    sample_batched_s = time.perf_counter() - ts
    batched_cut = sum(int(j.cut) for j in sjobs)

    # PCIe-inclusive: the sample from host CSR to host partition arrays
    # (upload, pool run, download)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    hdg = [heipa.DeviceGraph(reqs[i][0], "cuda:0") for i in sample]
    hjobs = [heipa.Pool.job(d, imbalance=0.03, seed=1, **reqs[i][1]) for d, i in zip(hdg, sample)]
    pool.run(hjobs)
    hparts = [heipa.to_numpy_u32(j._part)[:d.n] for j, d in zip(hjobs, hdg)]
    pcie_s = time.perf_counter() - tp
    pcie_same = all(np.array_equal(a, batched_parts[i]) for a, i in zip(hparts, sample))
    del hdg, hjobs, hparts
    pool.close()

    ctx = heipa.Context(0, arena)
    _apply_tune(ctx, tune)
    for i in sample[:2]:
        ctx.partition(dgs[i], reqs[i][1]["k"], 0.03, 1, False)
    torch.cuda.synchronize()
    tq = time.perf_counter()
    plain_cut, same = 0, 0
    for i in sample:
        p, st, _ = ctx.partition(dgs[i], reqs[i][1]["k"], 0.03, 1, False)
        plain_cut += int(st.cut)
        same += int(np.array_equal(heipa.to_numpy_u32(p), batched_parts[i]))
    plain_s = time.perf_counter() - tq
    ctx.close()
    plain_rate = es / plain_s
    same_graph_batched = es / sample_batched_s

    served = served_sample(args, tune, args.served_slots)
    res1, _, _, ctx1 = measure_single(args, "rgg20-ultra-k64", 5, 2)
    ctx1.close()
    cpu = None if args.no_cpu_baseline else cpu_batched(args, reqs, edges)

    result = {
        "metric": "input edges/sec on 1 MI355X, batched+served vs plain; p50/p99 latency at equal cut",
        "value": round(value, 1),
        "unit": "input edges/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(wall * 1000.0 / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32/int64",
        "data": "synthetic (seeded generators)",
        "config": {"workload": f"batched-{len(reqs)}: RGG (avg degree 8) or R-MAT (edge factor 8) with "
                               f"probability 0.5, n log-uniform in [1024, 65536], k from {{2,4,8,16,32}}, "
                               f"eps 0.03, default; {per} graphs per step ({args.batch_order} order), the {K} timed "
                               f"steps cover every graph once and are submitted back to back",
                   "graphs": len(reqs), "graphs_per_step": per, "undirected_edges": sum(edges),
                   "slots": args.slots, "arena_bytes_per_slot": arena, "input": "device-resident CSR",
                   "autotune": tune,
                   "mix": {"rgg": sum(1 for r in reqs if r[2] == "rgg"),
                           "rmat": sum(1 for r in reqs if r[2] == "rmat")}},
        "plain": {"value": round(plain_rate, 1), "unit": "input edges/s", "graphs": len(sample),
                  "seconds": round(plain_s, 3), "undirected_edges": es,
                  "batched_value_same_graphs": round(same_graph_batched, 1),
                  "what": "every 20th batched graph, one after another on one context (with the "
                          "autotuned launch shapes); the same graphs through the pool give "
                          "batched_value_same_graphs"},
        "speedup_vs_plain": round(same_graph_batched / plain_rate, 3),
        "admission_estimate_over_peak": est_ratio,
        "quality": {"failed": failed, "unbalanced": unbalanced,
                    "identical_to_plain": f"{same}/{len(sample)}",
                    "sum_cut_batched": batched_cut, "sum_cut_plain": plain_cut},
        "served": served,
        "plain_rgg20_ultra_k64": {"value": res1["value"], "unit": "input edges/s",
                                  "ms_per_step": res1["ms_per_step"], "quality": res1["quality"],
                                  "phase_ms": res1["config"]["phase_ms"],
                                  "pcie_inclusive": res1["pcie_inclusive"]},
        "roofline": res1["roofline"],
        "per_kernel": res1["per_kernel"],
        "cpu_baseline": cpu,
        "pcie_inclusive": {"value": round(es / pcie_s, 1), "seconds": round(pcie_s, 3),
                           "includes": "the same-graph sample: H2D upload of each CSR, pool run, D2H "
                                       "download of each partition", "identical": pcie_same},
        "generation_s": round(gen_s, 2),
    }
    print(json.dumps(result), flush=True)


def estimate_ratio(jobs):
    """admission estimate / measured arena peak over finished jobs"""
    import numpy as np
    r = np.array([j.estimate / j.peak for j in jobs if j.peak > 0])
    if not len(r):
        return None
    return {"min": round(float(r.min()), 3), "median": round(float(np.median(r)), 3), "max": round(float(r.max()), 3)}


def dump_peaks(path, reqs, jobs):
    rows = [{"n": g.n, "m": g.m, "spec": spec, "estimate": j.estimate, "peak": j.peak}
            for (g, spec), j in zip(reqs, jobs)]
    with open(path, "w") as fh:
        json.dump(rows, fh)


def cpu_batched(args, reqs, edges):
    """The serial oracle on the first --cpu-graphs batched graphs, one graph
    per host thread (--threads at a time)."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import heipa_oracle as orc
    orc.lib()
    idx = list(range(min(args.cpu_graphs, len(reqs))))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(args.threads) as ex:
        list(ex.map(lambda i: orc.partition(reqs[i][0], reqs[i][1]["k"], 0.03, 1, False), idx))
    dt = time.perf_counter() - t0
    return {"value": round(sum(edges[i] for i in idx) / dt, 1), "unit": "input edges/s", "cores": args.threads,
            "kind": "port", "cpu": cpu_model(),
            "sample": f"the first {len(idx)} of the {len(reqs)} batched graphs, one serial C++ oracle solve "
                      f"(deterministic mode) per graph, {args.threads} graphs at a time on {args.threads} host "
                      f"threads, {dt:.1f} s"}


def _apply_tune(obj, tune):
    """the autotuned launch shapes on a context or a pool (results do not change)"""
    import heipa
    if tune:
        obj.set_option(heipa.OPT_GROUP_CAP, tune["group_cap"])
        obj.set_option(heipa.OPT_BLOCK_GROUP_CAP, tune["block_group_cap"])
        obj.set_option(heipa.OPT_BLOCK_JET, tune["block_jet"])


def served_sample(args, tune=None, slots=None):
    """Poisson arrivals of the served request mix at 0.8x the plain path's
    request capacity (measured on the first 24 requests): latency
    percentiles, edges/s over the makespan, saturated capacity.  `tune`: the
    launch shapes the batched leg's autotuner chose, set on the pool and on
    the plain context alike."""
    import numpy as np
    import torch
    import heipa
    from heipa import workloads

    reqs = workloads.served(args.requests, args.served_nmax, threads=args.threads)
    keys = [0] * len(reqs)
    if args.served_repeat:
        # requests on D distinct graphs (request i on graph i mod D, with its own
        # k or hierarchy), named by graph keys: the pool keeps their hierarchies
        D = args.served_repeat
        reqs = [(reqs[i % D][0], spec, reqs[i % D][2]) for i, (_, spec, _) in enumerate(reqs)]
        keys = [1 + i % D for i in range(len(reqs))]
    dgs = [heipa.DeviceGraph(g, "cuda:0") for g, _, _ in reqs]
    edges = [g.m // 2 for g, _, _ in reqs]
    probe = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **spec) for dg, (_, spec, _) in zip(dgs, reqs)]
    pool0 = heipa.Pool(1, 1 << 20)
    arena = int(max(pool0.estimate(j) for j in probe) * 1.05) + (32 << 20)
    pool0.close()
    del probe
    ctx = heipa.Context(0, arena)
    _apply_tune(ctx, tune)
    sample = list(range(min(24, len(reqs))))
    for i in sample[:2]:
        _run_alone(ctx, dgs[i], reqs[i][1])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    alone = {i: _run_alone(ctx, dgs[i], reqs[i][1]) for i in sample}
    plain_s = time.perf_counter() - t0
    ctx.close()
    plain_rate = sum(edges[i] for i in sample) / plain_s
    capacity = len(sample) / plain_s
    lam = 0.8 * capacity
    slots = slots or args.slots
    pool = heipa.Pool(slots, arena)
    _apply_tune(pool, tune)
    pool.run([heipa.Pool.job(dgs[i], imbalance=0.03, seed=1, **reqs[i][1]) for i in range(min(2 * slots, len(reqs)))])
    if args.served_repeat:
        pool.set_option(heipa.OPT_HIER_CACHE, 64 << 30)
    jobs = [heipa.Pool.job(dg, imbalance=0.03, seed=1, graph_key=key, **spec)
            for dg, (_, spec, _), key in zip(dgs, reqs, keys)]
    arrivals = workloads.poisson_arrivals(len(reqs), lam)
    t0 = time.perf_counter()
    for j, a in zip(jobs, arrivals):
        dt = a - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        pool.submit(j)
    for j in jobs:
        pool.wait(j)
    sat = [heipa.Pool.job(dg, imbalance=0.03, seed=1, graph_key=key, **spec)
           for dg, (_, spec, _), key in zip(dgs, reqs, keys)]
    torch.cuda.synchronize()
    ts = time.perf_counter()
    pool.run(sat)
    sat_s = time.perf_counter() - ts
    cache = pool.hier_cache_stats() if args.served_repeat else None
    pool.close()
    lat = np.array([j.t_done - j.t_submit for j in jobs])
    makespan = (max(j.t_done for j in jobs) - min(j.t_submit for j in jobs)) / 1000.0
    same_alone = sum(1 for i in sample if np.array_equal(heipa.to_numpy_u32(jobs[i]._part)[:dgs[i].n], alone[i]))
    same_sat = sum(1 for a, b, dg in zip(sat, jobs, dgs)
                   if np.array_equal(heipa.to_numpy_u32(a._part)[:dg.n], heipa.to_numpy_u32(b._part)[:dg.n]))
    value = sum(edges) / makespan
    if args.dump_peaks:
        dump_peaks(args.dump_peaks + ".served.json", [(g, spec) for g, spec, _ in reqs], jobs)
    return {"value": round(value, 1), "unit": "input edges/s",
            "admission_estimate_over_peak": estimate_ratio(jobs),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                           "p99": round(float(np.percentile(lat, 99)), 3), "max": round(float(lat.max()), 3)},
            "capacity": {"value": round(sum(edges) / sat_s, 1), "unit": "input edges/s",
                         "seconds": round(sat_s, 3)},
            "plain": {"value": round(plain_rate, 1), "unit": "input edges/s", "requests": len(sample)},
            "capacity_speedup_vs_plain": round(sum(edges) / sat_s / plain_rate, 3),
            "resident_hierarchies": None if cache is None else {
                "distinct_graphs": args.served_repeat, "graphs_held": cache[0], "bytes": cache[1],
                "levels_reused": cache[2], "levels_built": cache[3]},
            "config": {"requests": len(reqs), "n_min": 1e4, "n_max": args.served_nmax,
                       "launch_shapes": ({key: tune[key] for key in ("group_cap", "block_group_cap", "block_jet")}
                                         if tune else "defaults"),
                       "offered_rate_per_s": round(lam, 3), "plain_capacity_per_s": round(capacity, 3),
                       "undirected_edges": sum(edges), "slots": slots,
                       "partition_requests": sum(1 for r in reqs if "k" in r[1]),
                       "mapping_requests": sum(1 for r in reqs if "hier" in r[1]),
                       "mix": "partition k uniform in 8..256 (70%), HM 4:8:6 or 2:4:8 (30%); grid / RGG "
                              "(avg degree 12) / R-MAT (edge factor 16) each 1/3; n log-uniform"},
            "quality": {"failed": sum(1 for j in jobs if j.status != heipa.OK),
                        "identical_to_alone": f"{same_alone}/{len(sample)}",
                        "identical_saturated_vs_poisson": f"{same_sat}/{len(jobs)}"}}


def _run_alone(ctx, dg, spec):
    import heipa
    if "hier" in spec:
        return heipa.to_numpy_u32(ctx.hm_partition(dg, spec["hier"], 0.03, 1, False))
    part, _, _ = ctx.partition(dg, spec["k"], 0.03, 1, False)
    return heipa.to_numpy_u32(part)


def run_pool_workload(args):
    import numpy as np
    import torch
    import heipa

    served = args.workload == "served"
    tg = time.perf_counter()
    from heipa import workloads
    reqs = workloads.served(args.requests, args.served_nmax) if served else workloads.batched(args.batch_count)
    dgs = [heipa.DeviceGraph(g, "cuda:0") for g, _, _ in reqs]
    gen_s = time.perf_counter() - tg
    edges = [g.m // 2 for g, _, _ in reqs]
    probe = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **spec) for dg, (_, spec, _) in zip(dgs, reqs)]
    pool0 = heipa.Pool(1, 1 << 20)
    est = [pool0.estimate(j) for j in probe]
    pool0.close()
    del probe
    arena = int(max(est) * 1.05) + (32 << 20)
    ctx = heipa.Context(0, arena)

    # plain path: requests one after another on one context (the drop-in)
    sample = list(range(min(args.plain_sample if not served else 24, len(reqs))))
    for i in sample[:2]:
        _run_alone(ctx, dgs[i], reqs[i][1])
    alone = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    service = []
    for i in sample:
        ts = time.perf_counter()
        alone[i] = _run_alone(ctx, dgs[i], reqs[i][1])
        service.append(time.perf_counter() - ts)
    plain_s = time.perf_counter() - t0
    plain_rate = sum(edges[i] for i in sample) / plain_s

    pool = heipa.Pool(args.slots, arena)
    if args.block_jet_n is not None:
        pool.set_option(heipa.OPT_BLOCK_JET, args.block_jet_n)
    warm = [heipa.Pool.job(dgs[i], imbalance=0.03, seed=1, **reqs[i][1]) for i in range(min(2 * args.slots, len(reqs)))]
    pool.run(warm)
    del warm
    jobs = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **spec) for dg, (_, spec, _) in zip(dgs, reqs)]
    result = {"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
              "dtype": "int32/int64", "data": "synthetic (seeded generators)", "steps": 1, "warmup": 1}
    if not served:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pool.run(jobs)
        wall = time.perf_counter() - t0
        value = sum(edges) / wall
        result.update({"metric": "input edges/sec on 1 MI355X, batched form", "value": round(value, 1),
                       "unit": "input edges/s", "ms_per_step": round(wall * 1000.0, 3),
                       "config": {"workload": "batched", "graphs": len(reqs), "undirected_edges": sum(edges),
                                  "slots": args.slots, "arena_bytes_per_slot": arena,
                                  "mix": {"rgg": sum(1 for r in reqs if r[2] == "rgg"),
                                          "rmat": sum(1 for r in reqs if r[2] == "rmat")}}})
    else:
        capacity = len(sample) / plain_s  # requests/s of the plain path on the sample
        lam = 0.8 * capacity
        arrivals = workloads.poisson_arrivals(len(reqs), lam)
        t0 = time.perf_counter()
        for j, a in zip(jobs, arrivals):
            dt = a - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            pool.submit(j)
        for j in jobs:
            pool.wait(j)
        # capacity: the same requests all at once (arrivals faster than service)
        sat = [heipa.Pool.job(dg, imbalance=0.03, seed=1, **spec) for dg, (_, spec, _) in zip(dgs, reqs)]
        torch.cuda.synchronize()
        ts = time.perf_counter()
        pool.run(sat)
        sat_s = time.perf_counter() - ts
        sat_same = sum(1 for a, b, dg in zip(sat, jobs, dgs)
                       if np.array_equal(heipa.to_numpy_u32(a._part)[:dg.n], heipa.to_numpy_u32(b._part)[:dg.n]))
        del sat
        lat = np.array([j.t_done - j.t_submit for j in jobs])
        makespan = (max(j.t_done for j in jobs) - min(j.t_submit for j in jobs)) / 1000.0
        value = sum(edges) / makespan
        result.update({"metric": "input edges/sec on 1 MI355X, served form", "value": round(value, 1),
                       "unit": "input edges/s", "ms_per_step": round(makespan * 1000.0, 3),
                       "capacity": {"value": round(sum(edges) / sat_s, 1), "unit": "input edges/s",
                                    "seconds": round(sat_s, 3), "identical_to_poisson_run": f"{sat_same}/{len(jobs)}"},
                       "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                      "p99": round(float(np.percentile(lat, 99)), 3),
                                      "max": round(float(lat.max()), 3)},
                       "config": {"workload": "served", "requests": len(reqs), "undirected_edges": sum(edges),
                                  "offered_rate_per_s": round(lam, 3),
                                  "plain_capacity_per_s": round(capacity, 3), "n_max": args.served_nmax,
                                  "slots": args.slots, "arena_bytes_per_slot": arena,
                                  "partition_requests": sum(1 for r in reqs if "k" in r[1]),
                                  "mapping_requests": sum(1 for r in reqs if "hier" in r[1])}})
    pool.close()
    bad = [j for j in jobs if j.status != heipa.OK]
    same = sum(1 for i in sample if np.array_equal(heipa.to_numpy_u32(jobs[i]._part)[:dgs[i].n], alone[i]))
    balanced = 0
    for (g, spec, _), j, dg in zip(reqs, jobs, dgs):
        kk = spec["k"] if "k" in spec else int(np.prod(spec["hier"]))
        p = heipa.to_numpy_u32(j._part)[:dg.n]
        lmax = int(np.ceil(1.03 * g.g_weight / kk)) if "k" in spec else None
        w = np.bincount(p, minlength=kk)
        balanced += int(lmax is None or w.max() <= lmax)
    result["plain"] = {"value": round(plain_rate, 1), "unit": "input edges/s", "requests": len(sample),
                       "seconds": round(plain_s, 3)}
    result["speedup_vs_plain"] = round(result["value"] / plain_rate, 3)
    if served:
        result["capacity_speedup_vs_plain"] = round(result["capacity"]["value"] / plain_rate, 3)
    result["quality"] = {"failed": len(bad), "identical_to_alone": f"{same}/{len(sample)}",
                         "partition_jobs_balanced": balanced,
                         "sum_cut": int(sum(j.cut for j in jobs))}
    result["generation_s"] = round(gen_s, 2)
    unbalanced = []
    for i, ((g, spec, kind), j, dg) in enumerate(zip(reqs, jobs, dgs)):
        if "k" in spec and j.maxw > int(np.ceil(1.03 * g.g_weight / spec["k"])):
            unbalanced.append({"i": i, "kind": kind, "n": g.n, "k": spec["k"], "maxw": int(j.maxw),
                               "lmax": int(np.ceil(1.03 * g.g_weight / spec["k"]))})
    result["quality"]["unbalanced"] = unbalanced[:8]
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
