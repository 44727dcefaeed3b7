#!/usr/bin/env python3
"""Benchmark of the relational hot path on one MI355X.

One step = one pass of the config-1 plan (filter build 1.5M rows and probe 6M rows -> sort-based
equi-join -> 264k-group sum -> top-10) through the plain per-operator HIP path, with the base
columns already resident in HBM.  Prints one JSON line: input rows/s (value), per-kernel
achieved HBM bandwidth from HIP events over the timed region (the dominant kernel as
`roofline`), and the CPU baseline (the torch-CPU restatement of the reference operators on this
host's cores, on a bounded sample of the same workload).
"""
import argparse
import collections
import copy
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "tqp-inmem_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("input rows/s + % HBM peak per op; batched rows/s & p50/p99 vs plain, 1 GPU")
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-runs", type=int, default=6, help="pipelines timed for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the config-3 batched comparison")
    ap.add_argument("--batch-pipes", type=int, default=256)
    ap.add_argument("--serve", action="store_true", help="add the served-form sweep")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the config-0 and config-2 plain-form lines")
    ap.add_argument("--op-sync", action="store_true",
                    help="synchronise after every operator, as the reference executor does")
    ap.add_argument("--inflight", type=int, default=2,
                    help="queries queued on the stream at once (1: each completes before the "
                         "next is prepared)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB copied once by a plain copy kernel (PMC traffic calibration)")
    ap.add_argument("--only", type=int, choices=(0, 2, 3, 5), default=None,
                    help="run only the config-0, config-2, config-3 (batched) or the extra "
                         "(5: repeated-key join, large sorts) section and print its JSON (for "
                         "per-workload profiles)")
    return ap.parse_args()


def kernel_table(calls, launches):
    """Per kernel: launches and mean device duration (dispatch-stamped start/stop events); for
    the principal kernel of every entry point that states its algorithmic bytes, bytes per
    launch and achieved GB/s against the HBM peak."""
    agg = {}
    for name, ms in launches:
        a = agg.setdefault(name, {"n": 0, "ms": 0.0, "bn": 0, "bms": 0.0, "bytes": 0})
        a["n"] += 1
        a["ms"] += ms
    for _entry, kern, nbytes, c0, c1 in calls:
        if nbytes is None or kern is None:
            continue
        hits = [i for i in range(c0, c1) if launches[i][0] == kern]
        if not hits:
            continue
        a = agg[kern]
        a["bn"] += len(hits)
        a["bms"] += sum(launches[i][1] for i in hits)
        a["bytes"] += nbytes
    out = {}
    for name, a in sorted(agg.items(), key=lambda kv: -kv[1]["ms"]):
        row = {"launches": a["n"], "avg_us": round(a["ms"] / a["n"] * 1e3, 2),
               "total_ms": round(a["ms"], 3)}
        if a["bn"] and a["bms"] > 0:
            gbs = a["bytes"] / (a["bms"] * 1e-3) / 1e9
            row.update({"alg_bytes_per_launch": int(a["bytes"] / a["bn"]), "GBps": round(gbs, 1),
                        "frac": round(gbs / HBM_PEAK_GBS, 4)})
        out[name] = row
    return out


JOIN_AGG_ENTRIES = {"tq_join_build", "tq_join_build_ranged", "tq_join_count", "tq_join_write",
                    "tq_join_probe_deferred", "tq_groupby_bounds", "tq_groupby_run",
                    "tq_groupby_run_deferred", "tq_groupby_accumulate", "tq_groupby_count",
                    "tq_groupby_write", "tq_groupjoin_run"}


def join_aggregate_span(calls, launches, steps, tables, pairs, groups):
    """The config-1 join + aggregate against the HBM roofline: the two operators' boundary bytes
    (every input column they read once, every output written once) over the summed device
    durations of the kernels their entry points launched."""
    seen = set()
    ms = 0.0
    for entry, _kern, _b, c0, c1 in calls:
        if entry in JOIN_AGG_ENTRIES and (c0, c1) not in seen:
            seen.add((c0, c1))
            ms += sum(launches[i][1] for i in range(c0, c1))
    nb, np_ = tables[101][1].shape[0], tables[201][1].shape[0]
    join_bytes = (nb * (4 + 1)            # build key + build mask
                  + np_ * (4 + 1)         # probe key + probe mask
                  + np_ * (8 + 8)         # price, discount gathered from the probe table
                  + pairs * (4 + 4 + 8 + 8))   # key, fk, price, discount written per pair
    agg_bytes = pairs * (4 + 8 + 8) + groups * (4 + 8)   # key, price, discount in; key, sum out
    total = join_bytes + agg_bytes
    # what the two operators cannot avoid moving: the build key and mask, the probe key and
    # mask, the gathered probe columns (a 10% gather touches nearly every line), the groups
    compulsory = nb * (4 + 1) + np_ * (4 + 1) + np_ * (8 + 8) + groups * (4 + 8)
    us = ms / steps * 1e3
    gbs = total / (us * 1e-6) / 1e9
    return {"boundary_bytes": total, "join_bytes": join_bytes, "aggregate_bytes": agg_bytes,
            "pairs": pairs, "groups": groups, "kernel_us_per_step": round(us, 2),
            "launches_per_step": round(sum(c1 - c0 for c0, c1 in seen) / steps, 2),
            "achieved_GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
            "compulsory_bytes": compulsory,
            "compulsory_frac": round(compulsory / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
            "target_frac": 0.5}


def cpu_model():
    """The host CPU's model name (what `lscpu` prints as 'Model name')."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(tables, plan, runs):
    """The torch-CPU restatement of the reference operators on this host: one thread per core
    of the process's CPU share (OMP_NUM_THREADS when the host sets one, else every core)."""
    import torch
    from oracle import ref_ops as R
    from tqp_amd import workloads
    ncpu = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", ncpu) or ncpu)
    threads = max(1, min(ncpu, share))
    torch.set_num_threads(threads)
    rows = workloads.input_rows(tables, 1)
    cols = R.to_cols(tables)
    R.execute(copy.deepcopy(plan), dict(cols))          # warm-up
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        R.execute(copy.deepcopy(plan), dict(cols))
        times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
    return {"value": rows / med, "unit": "input rows/s", "cores": torch.get_num_threads(),
            "kind": "port", "os_cpu_count": ncpu, "torch_threads": torch.get_num_threads(),
            "cpu_model": cpu_model(),
            "sample": f"median of {runs} full config-1 pipelines (7.5M input rows each) through "
                      f"the torch-CPU restatement of the reference operators, after one warm-up; "
                      f"{sum(times):.1f} s in all, {torch.get_num_threads()} threads of "
                      f"{ncpu} cores"}


CONFIG_WORKLOADS = {
    0: "filter -> 6-group aggregate (7 aggregates) -> sort; 1,000,000 rows",
    2: "3-join star (fact 60,000,000 rows, Zipf fk_b) -> ~2,500-group sum/count -> full sort",
}


def configs_section(dev, steps=20, warmup=5, cfgs=(0, 2)):
    """Configs 0 and 2 in the plain per-operator form: input rows/s, ms per pipeline and the
    dominant kernel's achieved HBM bandwidth (dispatch-stamped events, as for config 1)."""
    import torch
    from tqp_amd import _lib, executor, workloads
    out = {}
    for cfg in cfgs:
        tables = getattr(workloads, f"config{cfg}_tables")()
        plan = getattr(workloads, f"config{cfg}_plan")()
        base = workloads.to_tensor_group(tables, dev)
        rows = workloads.input_rows(tables, cfg)
        del tables
        plans = [copy.deepcopy(plan) for _ in range(steps + warmup)]
        for _ in range(warmup):
            executor.execute(plans.pop(), dict(base), sync=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            executor.execute(plans.pop(), dict(base), sync=False)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        plans = [copy.deepcopy(plan) for _ in range(steps)]
        _lib.start_trace()
        for _ in range(steps):
            executor.execute(plans.pop(), dict(base), sync=False)
        calls, launches = _lib.stop_trace()
        kt = kernel_table(calls, launches)
        dom = max(kt, key=lambda k: kt[k]["total_ms"]) if kt else None
        out[f"config{cfg}"] = {
            "workload": CONFIG_WORKLOADS[cfg], "input_rows": rows, "steps": steps,
            "rows_per_s": rows / dt, "ms_per_step": dt * 1e3,
            "kernel_ms_per_step": round(sum(v["total_ms"] for v in kt.values()) / steps, 4),
            "dominant_kernel": dict(kt[dom], kernel=dom) if dom else None,
            "kernels": dict(list(kt.items())[:8])}
        del base
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    return out


EXTRA_WORKLOADS = {
    "join_repeated_keys": "inner join, build 2,000,000 i32 keys (~4 rows per key, one hot key on "
                          "4,000 rows), probe 20,000,000 i32 keys (1,000 on the hot key); "
                          "outputs build key + i64 payload, probe key + f64 payload",
    "sort_16M_f64_i32": "full stable sort of 16,000,000 rows by (f64 descending, i32 ascending)",
    "sort_64M_i64": "full stable sort of 64,000,000 i64 keys",
}


def extra_section(dev, reps=5):
    """Workloads no config line covers: an inner join whose build keys repeat (with a hot key)
    and full LSD radix sorts above a million rows.  Per workload: wall time of one call
    (synchronised), device time of its kernels (dispatch-stamped) and the achieved bandwidth of
    the kernels with stated algorithmic bytes.  A radix scatter pass reads and writes a (key,
    row) pair per row: n * 2 * (key bytes + 4)."""
    import numpy as np
    import torch
    from tqp_amd import _lib, kernels
    g = np.random.default_rng(11)
    out = {}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps
        _lib.start_trace()
        for _ in range(reps):
            fn()
        calls, launches = _lib.stop_trace()
        return wall, kernel_table(calls, launches), launches

    nb, np_ = 2_000_000, 20_000_000
    bk = g.integers(1, 500_000, nb, dtype=np.int32)
    bk[g.choice(nb, 4000, replace=False)] = 0
    pk = g.integers(1, 600_000, np_, dtype=np.int32)
    pk[g.choice(np_, 1000, replace=False)] = 0
    bk_t, pk_t = torch.from_numpy(bk).to(dev), torch.from_numpy(pk).to(dev)
    bv = torch.from_numpy(g.integers(0, 1 << 40, nb, dtype=np.int64)).to(dev)
    pv = torch.from_numpy(g.random(np_)).to(dev)
    pairs = [0]

    def join():
        idx = kernels.JoinIndex(bk_t)
        lo, ro, _, _ = idx.probe(pk_t, None, [bk_t, bv], [pk_t, pv])
        pairs[0] = int(lo[0].shape[0])

    wall, kt, _ = timed(join)
    out["join_repeated_keys"] = {"workload": EXTRA_WORKLOADS["join_repeated_keys"],
                                 "pairs": pairs[0], "ms_per_join": wall * 1e3,
                                 "kernel_ms_per_join": round(sum(v["total_ms"] for v in
                                                                 kt.values()) / reps, 4),
                                 "kernels": dict(list(kt.items())[:6])}
    del bk_t, pk_t, bv, pv
    for name, n, make in (
            ("sort_16M_f64_i32", 16_000_000,
             lambda n: [(torch.from_numpy(g.random(n)).to(dev), True),
                        (torch.from_numpy(g.integers(-1000, 1000, n, dtype=np.int32)).to(dev),
                         False)]),
            ("sort_64M_i64", 64_000_000,
             lambda n: [(torch.from_numpy(g.integers(-(1 << 62), 1 << 62, n,
                                                     dtype=np.int64)).to(dev), False)])):
        keys = make(n)
        wall, kt, launches = timed(lambda: kernels.sort_perm(keys, n))
        # a scatter pass of 64-bit keys moves 8 + 4 bytes a row in and out (the f64 + i32 sort
        # mixes 64- and 32-bit passes, so only its rows/s is stated)
        kb = 8
        sc = [ms for k, ms in launches if k.startswith("k_radix_scatter")] \
            if name == "sort_64M_i64" else []
        row = {"workload": EXTRA_WORKLOADS[name], "rows": n, "ms_per_sort": wall * 1e3,
               "rows_per_s": n / wall,
               "kernel_ms_per_sort": round(sum(v["total_ms"] for v in kt.values()) / reps, 4),
               "kernels": dict(list(kt.items())[:6])}
        if sc:
            avg = sum(sc) / len(sc)
            gbs = n * 2 * (kb + 4) / (avg * 1e-3) / 1e9
            row["scatter_pass"] = {"passes_per_sort": len(sc) / reps, "avg_us": round(avg * 1e3, 2),
                                   "alg_bytes": n * 2 * (kb + 4), "GBps": round(gbs, 1),
                                   "frac": round(gbs / HBM_PEAK_GBS, 4)}
        out[name] = row
        del keys
        torch.cuda.empty_cache()
    return out


def batched_section(pipes_n, dev):
    """Config 3: the same pipelines as one persistent batched launch and as plain per-operator
    launches; rows/s of both and whether every result matched."""
    import numpy as np
    import torch
    from tqp_amd import batched, executor, workloads
    probe, build, consts = workloads.config3_sizes(n_pipes=pipes_n)
    pipes = [workloads.config3_pipeline(i, probe[i], build[i], consts[i][0], consts[i][1], dev)
             for i in range(pipes_n)]
    rows = int(probe.sum() + build.sum())
    # plain per-operator launches of the same pipelines, timed first: after the autotuner has
    # built and dropped its candidate engines the same loop measured slower in this process
    def plain_all():
        res = []
        for g, p in pipes:
            tg = dict(g)
            executor.execute(copy.deepcopy(p), tg)
            res.append(tg)
        return res
    plain_all()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref = plain_all()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    specs = [batched.match_plan(p, g) for g, p in pipes]
    tune = batched.autotune(specs, key=f"{torch.cuda.get_device_name(0)}|config3|{pipes_n}")
    eng = batched.BatchedEngine(specs, tile_rows=tune["tile_rows"], grid=tune["grid"],
                                records=tune.get("records", False),
                                lds_table_max=tune.get("lds_table_max"),
                                reduce_lds=tune.get("reduce_lds"),
                                fused_min_rows=tune.get("fused_min_rows"))
    eng.run()
    out, rerun = eng.results()
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.run()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    out, rerun = eng.results()
    same = True
    for i, tg in enumerate(ref):
        if i in rerun:
            continue
        same &= torch.equal(out[i][101].cpu(), tg[101].tensor.cpu())
        same &= torch.equal(out[i][302].cpu(), tg[302].tensor.cpu())
        a, b = out[i][301].cpu().numpy(), tg[301].tensor.cpu().numpy()
        same &= a.shape == b.shape and bool(np.all(np.abs(a - b) <= 1e-12 * np.abs(b)))
    return {"pipelines": pipes_n, "input_rows": rows,
            "batched_rows_per_s": rows / t_batch, "plain_rows_per_s": rows / t_plain,
            "speedup": t_plain / t_batch, "batched_ms": t_batch * 1e3, "plain_ms": t_plain * 1e3,
            "identical_results": bool(same), "rerun_plain": rerun,
            "tile_rows": tune["tile_rows"], "grid": tune["grid"],
            "large_tables": "records" if tune.get("records") else "atomics",
            "lds_table_max": tune.get("lds_table_max"), "reduce_lds": tune.get("reduce_lds"),
            "fused_min_rows": tune.get("fused_min_rows"), "fused_pipelines": len(eng.fused),
            "tuning_ms": {k: round(v * 1e3, 3) for k, v in tune.get("timings", {}).items()}}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              f"(the path runs on one MI355X; sharding across GPUs is a design, not built)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from tqp_amd import _lib, executor, workloads

    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        x = torch.empty(args.calibrate_copy << 20, dtype=torch.uint8, device=dev)
        y = x.clone()
        torch.cuda.synchronize()
        del x, y
    if args.only is not None:
        if args.only == 3:
            sec = batched_section(args.batch_pipes, dev)
        elif args.only == 5:
            sec = extra_section(dev)
        else:
            sec = configs_section(dev, steps=args.steps, warmup=args.warmup, cfgs=(args.only,))
        print(json.dumps({"config3": sec} if args.only == 3 else sec), flush=True)
        return
    tables = workloads.config1_tables()
    plan = workloads.config1_plan()
    base = workloads.to_tensor_group(tables, dev)
    rows = workloads.input_rows(tables, 1)

    # the executor edits plan arguments in place (as main.py's operators do), so every step gets
    # its own copy of the plan, made before the timed region
    copies = []

    def fresh_plans(k):
        copies[:] = [copy.deepcopy(plan) for _ in range(k)]

    def run(k, op_sync=args.op_sync, inflight=args.inflight):
        """k queries; with inflight > 1 the next query's kernels are queued before the previous
        one's counts are collected (the host prepares query i+1 while the device runs query i).
        Every query runs whole: its results are settled when the loop ends."""
        pend = collections.deque()
        for _ in range(k):
            tg = dict(base)
            q = executor.execute(copies.pop() if copies else copy.deepcopy(plan), tg,
                                 sync=op_sync, wait=op_sync or inflight <= 1)
            if isinstance(q, executor.PendingQuery):
                pend.append(q)
                while len(pend) >= inflight:
                    pend.popleft().result()
        while pend:
            pend.popleft().result()

    run(args.warmup)
    run(args.warmup, inflight=1)
    torch.cuda.synchronize()
    # timed region: the whole pipeline, kernels untimed
    fresh_plans(args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(args.steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # the same queries one at a time: each completes before the next is prepared
    fresh_plans(args.steps)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    run(args.steps, inflight=1)
    torch.cuda.synchronize()
    dt_seq = time.perf_counter() - t3
    # the same steps again with every kernel launch carrying device-stamped start/stop events
    # (the kernel durations and achieved bandwidths), then with the reference executor's
    # per-operator synchronisation (main.py:273-274), for comparison
    fresh_plans(args.steps)
    _lib.start_trace()
    t1 = time.perf_counter()
    run(args.steps)
    torch.cuda.synchronize()
    dt_traced = time.perf_counter() - t1
    calls, launches = _lib.stop_trace()
    kernels = kernel_table(calls, launches)
    # pair and group counts of one step, for the join + aggregate boundary bytes
    ja_plan = copy.deepcopy(plan[:4])
    ja_group = dict(base)
    executor.execute(ja_plan, ja_group, sync=True)
    pairs = int(ja_group[201].tensor.shape[0])
    groups = int(ja_group[301].tensor.shape[0])
    ja = join_aggregate_span(calls, launches, args.steps, tables, pairs, groups)
    # the dominant kernel: the largest device time of the step
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    d = kernels[dom]
    fresh_plans(args.steps)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    run(args.steps, op_sync=True)
    torch.cuda.synchronize()
    side = {"kernel_timed": round(dt_traced / args.steps * 1e3, 4),
            "op_sync": round((time.perf_counter() - t2) / args.steps * 1e3, 4)}
    # HBM bytes per launch of that kernel from the committed rocprofv3 counter runs of this
    # workload (FETCH_SIZE and WRITE_SIZE scaled by a 256 MiB copy); counters cannot be read in
    # this process, so the figure is labelled with its source
    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        tj = json.load(open(tf))
        if dom in tj.get("kernels", {}):
            traffic = tj["kernels"][dom]
            traffic_src = f"profiles/traffic.json ({tj.get('source', 'rocprofv3 --pmc')})"
    line = {
        "metric": METRIC,
        "value": rows * args.steps / dt,
        "unit": "input rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 tables of the config-1 shape)",
        "config": {"workload": "filter -> sort-based equi-join -> group-by sum -> top-10; build "
                               "1,500,000 rows, probe 6,000,000 rows, ~264k groups",
                   "input_rows": rows,
                   "form": "plain form (one query after another, no batching); the two filters "
                           "evaluated inside the fused join + group-by (groupjoin) kernels",
                   "inflight": 1 if args.op_sync else args.inflight,
                   "op_sync": bool(args.op_sync)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": d.get("GBps"),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d.get("frac"),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "alg_bytes_per_launch": d.get("alg_bytes_per_launch"),
                     "avg_us": d["avg_us"]},
        "join_aggregate": ja,
        "ms_per_step_sequential": round(dt_seq / args.steps * 1e3, 4),
        "ms_per_step_kernel_timed": side["kernel_timed"],
        "ms_per_step_op_sync": side["op_sync"],
        "kernels": kernels,
    }
    if not args.no_configs:
        line["configs"] = configs_section(dev)
    if not args.no_batched:
        line["batched"] = batched_section(args.batch_pipes, dev)
    if not args.no_configs:
        # after the batched comparison: its multi-GB inputs and workspaces stay cached by the
        # allocator and the role workspaces, which the plain path's timing must not inherit
        line["extra"] = extra_section(dev)
    if args.serve:
        from tqp_amd import serving_bench
        line["served"] = serving_bench.sweep(dev)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(tables, plan, args.cpu_runs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
