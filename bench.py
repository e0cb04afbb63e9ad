"""Benchmark: MIPS of the HIP reduction engine on one MI355X.

A step is one reduction of one net: Net::new (reset) + boot(main) + normal()
over the workload's book, which is already resident on the device.  Prints
one JSON line (see the fields below).  Workloads:

  bitonic   bitonic sort of 2**k seeded u24 keys, single net (default, k=16)
  tree_sum  sum of a complete binary tree of depth d (default d=18)
  church    Church exponentiation c2 ^ c_e applied to +1 on 0 (default e=16;
            the configuration's e=22 is a 4M-step sequential chain)

The CPU baseline is the oracle (a C++ restatement of the reference evaluator)
timed on this machine's host cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hvm-64_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBPS = 8000.0


def workload(name, size, seed):
    from hvm64 import configs
    if name == "bitonic":
        size = 16 if size is None else size
        return (size, configs.bitonic(size, seed), configs.bitonic_expected(size, seed),
                f"bitonic sort of 2^{size} u24 keys (seed {seed}), single net")
    if name == "tree_sum":
        size = 18 if size is None else size
        return size, configs.tree_sum(size), configs.tree_sum_expected(size), f"tree-sum depth {size}, single net"
    if name == "church":
        # c2^c_e applies inc 2^e times in sequence: the net's average
        # parallelism is about 3 redexes (CPU level simulation), so a GPU
        # pays one dependent-latency step per level; e=16 keeps a run short
        size = 16 if size is None else size
        return (size, configs.church(size, 16), configs.church_expected(size),
                f"Church c2^c{size} applied to +1 on 0, 16 DUP labels, single net")
    raise SystemExit(f"unknown workload {name}")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# Host cores one GPU's share of the machine may use (the box's worker budget);
# the reference's parallel_normal takes 2^floor(log2(available cores)) threads
# (runtime/src/parallel.rs:55-57) of that share.
CPU_SHARE = 16


def _median(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2]


def cpu_baseline(name, size, seed, runs=5):
    """Times the oracle (the reference evaluator restated in C++) on this
    machine's host cores, median of `runs` runs each:

    * parallel_normal (runtime/src/parallel.rs restated) with
      T = 2^floor(log2(cores)) threads, cores = min(affinity, CPU_SHARE), on
      the same workload at the same size (bitonic 2^16: ~1.7 s a run);
    * single-threaded normal() (runtime/src/net.rs:55-60) on the same program
      two sizes smaller (a quarter of the work), so five runs stay bounded.

    Also returns the full-size run's normal form and counts (the GPU's are
    compared with them)."""
    try:
        import oracle_lib
        o = oracle_lib.Oracle()
    except Exception as e:  # the oracle is optional for the GPU figure
        return {"value": None, "unit": "MIPS", "cores": 0, "kind": "port", "sample": f"unavailable: {e}"}, None
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    cores = max(1, min(affinity, CPU_SHARE))
    threads = 1
    while threads * 2 <= cores:
        threads *= 2
    text = workload(name, size, seed)[1]
    multi = [o.run(text, heap_words=1 << 29, threads=threads) for _ in range(runs)]
    small = max(1, size - 2)
    small_text = workload(name, small, seed)[1]
    single = [o.run(small_text, heap_words=1 << 28, threads=0) for _ in range(runs)]
    mips_t = _median([r["rwts"]["rwts"] / r["seconds"] / 1e6 for r in multi])
    mips_1 = _median([r["rwts"]["rwts"] / r["seconds"] / 1e6 for r in single])
    sec_t = _median([r["seconds"] for r in multi])
    sec_1 = _median([r["seconds"] for r in single])
    out = {"value": round(mips_t, 2), "unit": "MIPS", "cores": threads, "kind": "port",
           "sample": f"{name} size {size}: {multi[0]['rwts']['rwts']} rewrites, parallel_normal with {threads} threads, "
                     f"median of {runs} runs ({sec_t:.3f} s); single-thread normal() on size {small} "
                     f"({single[0]['rwts']['rwts']} rewrites): {mips_1:.2f} MIPS, median of {runs} ({sec_1:.3f} s)",
           "single_thread_mips": round(mips_1, 2), "runs": runs, "nproc": nproc, "affinity": affinity,
           "cpu_share": CPU_SHARE, "cpu_model": cpu_model()}
    # (iii) the compiled path: the book's definitions as C++ functions
    # (src/compile.rs:110-161, do_switch for do_mat), same two runs
    try:
        import tempfile
        work = tempfile.mkdtemp(prefix="hvm64_compiled_")
        exe = o.compile_book(text, work)
        nf_c, runs_t = o.run_compiled(exe, text, threads=threads, runs=runs, heap_words=1 << 29)
        _, runs_1 = o.run_compiled(exe, small_text, threads=0, runs=runs, heap_words=1 << 28)
        ok_c = nf_c == multi[0]["nf"] and runs_t[0][1] == multi[0]["rwts"]
        out["compiled"] = {
            "mips": round(_median([c["rwts"] / s / 1e6 for s, c in runs_t]), 2), "threads": threads,
            "single_thread_mips": round(_median([c["rwts"] / s / 1e6 for s, c in runs_1]), 2),
            "matches_interpreter": ok_c,
            "sample": f"definitions compiled to C++ and built with g++ -O2; parallel_normal on size {size}, "
                      f"normal() on size {small}, median of {runs} runs each"}
    except Exception as e:  # g++ or the oracle sources missing: report it, keep the rest
        out["compiled"] = {"mips": None, "sample": f"unavailable: {e}"}
    return out, multi[0]


def copy_calibration(nbytes):
    """One device-to-device copy of `nbytes` (the HIP runtime's copy kernel,
    reading and writing exactly nbytes) and one fill of `nbytes`: known-size
    kernels that calibrate the FETCH_SIZE / WRITE_SIZE counters of a profile."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)


def recorded_traffic(name, size, seed):
    """HBM bytes per k_reduce launch recorded by a rocprofv3 FETCH_SIZE /
    WRITE_SIZE profile of this workload (profiles/k_reduce_*.md, scaled by
    a known-size copy), or (None, None) when no profile covers it."""
    import glob
    key = f"{name}:{size}:{seed}"
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "k_reduce_*.md"))):
        fields = {}
        for line in open(path):
            if " = " in line:
                k, v = line.strip().split(" = ", 1)
                fields[k] = v
        if fields.get("workload_key") == key and "traffic_bytes_per_launch" in fields:
            return int(fields["traffic_bytes_per_launch"]), os.path.relpath(path, ROOT)
    return None, None


def algorithmic_bytes_per_rewrite(name, size, seed):
    """Heap-word traffic of the reference algorithm per rewrite, counted by
    the metered oracle (8 B per load or store, 16 B per atomic RMW) on a
    smaller instance of the workload (the ratio is size-stable)."""
    try:
        import oracle_lib
        o = oracle_lib.Oracle()
        small = {"bitonic": min(size, 10), "tree_sum": min(size, 12), "church": min(size, 12)}[name]
        r = o.run(workload(name, small, seed)[1], heap_words=1 << 26, meter=True)
        m = r["meter"]
        return (8 * m["loads"] + 8 * m["stores"] + 16 * m["rmws"]) / r["rwts"]["rwts"]
    except Exception:
        return None


def bench_batch(args):
    """Config 4: args.nets independent nets (tree-sum depth U[8,16] or bitonic
    k U[6,12], seeded), each allocating from an arena of its own and all
    reduced by one replay of a HIP graph (reset, boot, expand, reduce, root
    ports) per step.  Beside it: the same nets through the unplanned batch
    calls (boot_batch + normal_batch, shared allocation rings), and one
    launch sequence per net (Net::new + boot + normal) on a seeded random
    sample of the nets."""
    import random

    import hvm64
    from hvm64 import configs
    book, nets, table = configs.batch_mix(args.nets, args.seed)
    heap = hvm64.Heap(int(args.heap_gb * (1 << 30)))
    heap.set_timeout(60)
    host = hvm64.Host(book, heap)
    n_compiled = heap.set_compiled(not args.interpreted)
    defs = [host.defs[table[v][0]] for v in nets]
    expected = [hvm64.u24_port(int(table[v][1])) for v in nets]

    # arenas: first from each variant's allocation count alone, then from
    # what the plan's own first run asked of each arena
    def note(msg):
        print(f"[bench batch {time.perf_counter() - t_plan:7.2f} s] {msg}", file=sys.stderr, flush=True)

    t_plan = time.perf_counter()
    sizes = hvm64.arena_sizes(heap, defs, margin=1.3)
    note(f"variants measured: {len(set(d.idx for d in defs))} definitions, {sum(sizes) * 16 / 2**30:.1f} GiB")
    plan = hvm64.PlannedBatch(heap, defs, sizes)
    note(f"planned: {plan.info()}")
    plan.run()
    note(f"first run: {plan.last_kernel_ms():.1f} ms, {plan.info()['pool_nodes']} nodes taken from the pool")
    used = plan.used()
    sizes = [max(1024, int(u * 1.1)) for u in used]
    plan = hvm64.PlannedBatch(heap, defs, sizes)
    plan_s = time.perf_counter() - t_plan
    note(f"replanned: {sum(sizes) * 16 / 2**30:.1f} GiB of arenas")
    tune_s = 0.0
    if not args.no_tune:
        t_tune = time.perf_counter()
        plan.tune()
        tune_s = time.perf_counter() - t_tune
        note(f"tuned in {tune_s:.1f} s")
    info = plan.info()
    note(f"plan: {info}")

    for _ in range(args.warmup):
        plan.run()
    ok = True
    t0 = time.perf_counter()
    total = 0
    kernel_ms = []
    for _ in range(args.steps):
        ports = plan.run()
        ok = ok and ports == expected
        total += plan.rwts.total()
        kernel_ms.append(plan.last_kernel_ms())
    elapsed = time.perf_counter() - t0
    note(f"timed {args.steps} steps: {elapsed:.2f} s")
    rw = plan.rwts
    pool_nodes = plan.info()["pool_nodes"]
    arena_mib = sum(sizes) * 16 / (1 << 20)
    batched_mips = total / elapsed / 1e6

    # the same nets through the unplanned calls (shared allocation rings)
    t2 = time.perf_counter()
    b = hvm64.Batch(heap)
    b.boot(defs)
    b.normal()
    un_elapsed = time.perf_counter() - t2
    un_ok = b.ports() == expected
    unplanned_mips = b.rwts.total() / un_elapsed / 1e6
    note(f"unplanned batch: {un_elapsed:.2f} s")

    sample = min(args.per_net_sample, len(nets))
    picks = sorted(random.Random(args.seed + 1).sample(range(len(nets)), sample))
    t1 = time.perf_counter()
    per_total = 0
    for i in picks:
        net = hvm64.Net(heap)
        net.boot(defs[i])
        net.normal()
        per_total += net.rwts.total()
    per_elapsed = time.perf_counter() - t1
    per_net_mips = per_total / per_elapsed / 1e6
    blocks, waves = heap.grid()
    k_ms = sum(kernel_ms) / len(kernel_ms)
    roofline = batch_roofline(book, table, nets, rw.total(), k_ms, args)
    out = {
        "metric": "MIPS on 1 MI355X: batched vs per-net launches",
        "value": round(batched_mips, 2),
        "unit": "MIPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded mix of in-net generated trees and keys)",
        "config": {"workload": f"{args.nets} independent nets, one graph replay per step: tree-sum depth U[8,16] "
                               f"or bitonic 2^k keys k U[6,12] (seed {args.seed})",
                   "nets": args.nets, "seed": args.seed, "heap_gb": args.heap_gb, "grid_blocks": blocks, "waves": waves,
                   "definitions": "interpreted" if args.interpreted else f"compiled ({n_compiled})"},
        "rewrites_per_step": rw.total(),
        "rewrites": rw.as_dict(),
        "kernel_ms": round(k_ms, 3),
        "result_ok": ok,
        "roofline": roofline,
        "plan": {"tuning": info["tuning"], "graph": info["graph"], "tuned": not args.no_tune,
                 "tune_seconds": round(tune_s, 2), "plan_seconds": round(plan_s, 2),
                 "arena_mib": round(arena_mib, 1), "pool_nodes": pool_nodes},
        "unplanned": {"mips": round(unplanned_mips, 2), "result_ok": un_ok, "seconds": round(un_elapsed, 4)},
        "per_net": {"mips": round(per_net_mips, 2), "nets_timed": sample, "sample": f"seeded random (seed {args.seed + 1})",
                    "rewrites": per_total, "seconds": round(per_elapsed, 3)},
        "batched_over_per_net": round(batched_mips / per_net_mips, 2),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = batch_cpu_baseline(book, table, nets, args.seed)
    print(json.dumps(out), flush=True)
    heap.close()
    return 0 if ok and un_ok else 1


def batch_roofline(book, table, nets, rewrites, k_ms, args):
    """Roofline of one replay's k_reduce: the algorithmic heap traffic of
    every net (the metered oracle on each variant: 8 B per load or store,
    16 B per atomic RMW) over the replay's kernel time."""
    try:
        import oracle_lib
        o = oracle_lib.Oracle()
        per = {}
        for v in table:
            r = o.run(book, entry=table[v][0], heap_words=1 << 27, meter=True)
            m = r["meter"]
            per[v] = 8 * m["loads"] + 8 * m["stores"] + 16 * m["rmws"]
    except Exception:
        return None
    alg = sum(per[v] for v in nets)
    achieved = alg / (k_ms / 1e3) / 1e9
    traffic, source = recorded_traffic("batch", args.nets, args.seed)
    out = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBPS, 5), "traffic": traffic,
           "bytes_per_rewrite": round(alg / rewrites, 2), "kernel": "k_reduce (per-net arenas)"}
    if source:
        out["traffic_source"] = source
    return out


def batch_cpu_baseline(book, table, nets, seed, sample=256):
    """The oracle on a seeded random sample of the batch's nets, each net
    reduced single-threaded (normal(), runtime/src/net.rs:55-60) and the nets
    spread over the host cores (one net per worker at a time): the CPU's way
    to batch independent nets."""
    import random
    from concurrent.futures import ThreadPoolExecutor
    try:
        import oracle_lib
        o = oracle_lib.Oracle()
    except Exception as e:
        return {"value": None, "unit": "MIPS", "cores": 0, "kind": "port", "sample": f"unavailable: {e}"}
    try:
        cores = max(1, min(len(os.sched_getaffinity(0)), CPU_SHARE))
    except AttributeError:
        cores = max(1, min(os.cpu_count() or 1, CPU_SHARE))
    picks = random.Random(seed + 2).sample(range(len(nets)), min(sample, len(nets)))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:  # ctypes calls release the GIL
        runs = list(ex.map(lambda i: o.run(book, entry=table[nets[i]][0], heap_words=1 << 27, threads=0), picks))
    sec = time.perf_counter() - t0
    ok = all(r["nf"] == table[nets[i]][1] for r, i in zip(runs, picks))
    total = sum(r["rwts"]["rwts"] for r in runs)
    return {"value": round(total / sec / 1e6, 2), "unit": "MIPS", "cores": cores, "kind": "port",
            "sample": f"{len(picks)} nets of the batch (seeded random, seed {seed + 2}), {total} rewrites, "
                      f"single-threaded normal() per net over {cores} workers: {sec:.2f} s",
            "results_ok": ok, "cpu_model": cpu_model()}


def bench_serve(args):
    """Config 5: seeded Poisson arrivals of the tree-sum / bitonic mix (sizes
    log-uniform), served by one persistent launch that admits requests into
    free slot arenas itself (hvm_gpu_serve_*); reports p50/p99 latency
    (arrival to the host poll that saw the result) and aggregate MIPS."""
    from hvm64 import serve
    heap_bytes = int(args.heap_gb * (1 << 30))
    srv = serve.Server(slots=args.slots, heap_bytes=heap_bytes, book_seed=args.seed)
    srv.heap.set_sched_stats(args.sched_stats)
    try:
        if args.warmup:
            srv.run(serve.poisson_requests(args.rate, 0.2, seed=args.seed + 1))
        reqs = serve.poisson_requests(args.rate, args.duration, seed=args.seed)
        st = srv.run(reqs)
    finally:
        srv.close()
    out = {
        "metric": "p99 latency on 1 MI355X: serving (Poisson arrivals, admission inside the running kernel)",
        "value": round(st.p99_ms, 2),
        "unit": "ms",
        "n_gpus": 1,
        "steps": 1,
        "warmup": args.warmup,
        "ms_per_step": round(st.wall_s * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded Poisson arrivals of in-net generated trees and keys)",
        "config": {"workload": f"Poisson arrivals at {args.rate:g}/s for {args.duration:g} s of tree-sum depth / "
                               f"bitonic k log-uniform over [8,16] / [6,12] (seed {args.seed})",
                   "rate": args.rate, "duration_s": args.duration, "slots": st.slots,
                   "slot_nodes": srv.slot_nodes, "heap_gb": args.heap_gb, "grid_blocks": st.grid_blocks},
        "p50_ms": round(st.p50_ms, 2),
        "p99_ms": round(st.p99_ms, 2),
        "max_ms": round(st.max_ms, 2),
        "mips": round(st.mips, 2),
        "requests": st.requests,
        "completed": st.completed,
        "result_ok": st.all_ok,
        "submits": st.submits,
        "mean_submit": round(st.mean_submit, 1),
        "rewrites": st.rewrites,
        "wall_s": round(st.wall_s, 3),
        "kernel_ms": round(st.kernel_ms, 3),
        "sched": srv.last_sched,
    }
    print(json.dumps(out), flush=True)
    return 0 if st.all_ok else 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="bitonic", choices=["bitonic", "tree_sum", "church", "batch", "serve"])
    ap.add_argument("--nets", type=int, default=4096, help="batch: number of nets")
    ap.add_argument("--rate", type=float, default=2000.0, help="serve: Poisson arrivals per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--slots", type=int, default=1024, help="serve: request arenas (requests in flight at most)")
    ap.add_argument("--per-net-sample", type=int, default=256,
                    help="batch: nets timed with one launch sequence each (the per-net baseline)")
    ap.add_argument("--no-tune", action="store_true", help="batch: keep the default launch tuning")
    ap.add_argument("--size", type=int, default=None)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--heap-gb", type=float, default=None,
                    help="heap size (default 192 GiB of the MI355X's 288 GB, batch 224 GiB)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sched-stats", action="store_true",
                    help="collect the scheduler statistics of a single net's launch (diagnostic; costs 2-4%%)")
    ap.add_argument("--interpreted", action="store_true",
                    help="run every definition interpreted (compiled definitions off)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="before the steps, copy a buffer of this many MiB on the device (a known-size "
                         "kernel that calibrates HBM byte counters in a profile)")
    args = ap.parse_args()
    if args.heap_gb is None:
        # a larger heap keeps each lane's allocation slice on its first lap
        # (no reads to find free slots): bitonic 2^16 runs at ~1640 MIPS with
        # 32 GiB and ~1890 with 192 GiB
        # batch: 4,096 arenas hold ~181 GiB (every node a net allocates,
        # plus the ends of lanes' chunks); the pool grows them
        args.heap_gb = 224.0 if args.workload == "batch" else 192.0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the engine runs one net per "
              f"device and this build has no multi-GPU path).", file=sys.stderr)
        return 2

    import hvm64
    if args.copy_mib:
        copy_calibration(args.copy_mib << 20)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)
    size, text, expected, desc = workload(args.workload, args.size, args.seed)
    heap = hvm64.Heap(int(args.heap_gb * (1 << 30)))
    heap.set_timeout(300)
    heap.set_sched_stats(args.sched_stats)
    t_up = time.perf_counter()
    host = hvm64.Host(text, heap)  # parse, encode and upload the book
    n_compiled = heap.set_compiled(not args.interpreted)
    upload_ms = (time.perf_counter() - t_up) * 1e3
    main_def = host.defs["main"]

    def step():
        net = hvm64.Net(heap)
        net.boot(main_def)
        net.normal()
        return net

    for _ in range(args.warmup):
        net = step()
    t0 = time.perf_counter()
    kernel_ms = []
    total = 0
    for _ in range(args.steps):
        net = step()
        kernel_ms.append(net.last_kernel_ms())
        total += net.rwts.total()
    elapsed = time.perf_counter() - t0
    t_rb = time.perf_counter()
    nf = host.readback(net)
    readback_ms = (time.perf_counter() - t_rb) * 1e3
    rw = net.rwts
    stats = net.sched_stats()
    ok = nf == expected

    value = total / elapsed / 1e6
    k_ms = sum(kernel_ms) / len(kernel_ms)
    bpr = algorithmic_bytes_per_rewrite(args.workload, size, args.seed)
    roofline = None
    if bpr is not None:
        achieved = bpr * rw.total() / (k_ms / 1e3) / 1e9
        traffic, source = recorded_traffic(args.workload, size, args.seed)
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBPS, 5), "traffic": traffic,
                    "bytes_per_rewrite": round(bpr, 2), "kernel": "k_reduce"}
        if source:
            roofline["traffic_source"] = source
    blocks, waves = heap.grid()
    out = {
        "metric": f"MIPS on 1 MI355X: single net ({ {'bitonic': 'bitonic sort', 'tree_sum': 'tree-sum', 'church': 'Church exponentiation'}[args.workload] })",
        "value": round(value, 2),
        "unit": "MIPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded in-net key generator)",
        "config": {"workload": desc, "size": size, "seed": args.seed, "heap_gb": args.heap_gb,
                   "grid_blocks": blocks, "waves": waves,
                   "definitions": "interpreted" if args.interpreted else f"compiled ({n_compiled})"},
        "rewrites_per_step": rw.total(),
        "rewrites": rw.as_dict(),
        "kernel_ms": round(k_ms, 3),
        "kernel_mips": round(rw.total() / k_ms / 1e3, 2),
        # one net end to end: book upload + reset/boot/normal + readback
        "e2e": {"book_upload_ms": round(upload_ms, 3), "step_ms": round(elapsed / args.steps * 1e3, 3),
                "readback_ms": round(readback_ms, 3),
                "mips": round(rw.total() / ((upload_ms + readback_ms) / 1e3 + elapsed / args.steps) / 1e6, 2)},
        "result_ok": ok,
        "sched": stats,
        "roofline": roofline,
    }
    if not args.no_cpu:
        base, ref = cpu_baseline(args.workload, size, args.seed)
        out["cpu_baseline"] = base
        if ref is not None:
            # the oracle's full-size run: same normal form, same five counts
            out["oracle_match"] = ref["nf"] == nf and ref["rwts"] == rw.as_dict()
            out["result_ok"] = ok = ok and out["oracle_match"]
    print(json.dumps(out), flush=True)
    heap.close()
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
