"""Planned batches (hvm_gpu_batch_plan / _run / _tune): every net allocates
from an arena of its own and a run is one replay of a HIP graph.  Each net
must reach the same normal form as Net::new + Net::boot + Net::normal alone
(runtime/src/net.rs:19-60) and the batch's five counts must equal the sum of
the oracle's per-variant counts, on every replay."""
import pytest

pytestmark = pytest.mark.gpu


def _reference(oracle, book, table):
    ref = {}
    for v in table:
        r = oracle.run(book, entry=table[v][0], heap_words=1 << 27)
        assert r["nf"] == table[v][1]
        ref[v] = r
    return ref


def _check(plan, host, nets, table, ref, ports):
    import hvm64
    assert ports == [hvm64.u24_port(int(table[v][1])) for v in nets]
    total = {k: sum(ref[v]["rwts"][k] for v in nets) for k in ["rwts", "anni", "comm", "eras", "dref", "oper"]}
    assert plan.rwts.as_dict() == total
    for i in range(0, len(nets), 23):
        assert host.readback_root(plan.roots[i]) == ref[nets[i]]["nf"]


def test_planned_batch_against_oracle(gpu_heap, oracle):
    import hvm64
    from hvm64 import configs
    # without reuse a net of the mix allocates up to ~3M nodes: 160 nets fit
    # the 16 GiB heap
    book, nets, table = configs.batch_mix(160, seed=11)
    print("case planned batch 160 seed 11", flush=True)
    ref = _reference(oracle, book, table)
    host = hvm64.Host(book, gpu_heap)
    defs = [host.defs[table[v][0]] for v in nets]
    sizes = hvm64.arena_sizes(gpu_heap, defs)
    plan = hvm64.PlannedBatch(gpu_heap, defs, sizes)
    info = plan.info()
    print("plan", info, flush=True)
    for rep in range(3):  # replays of the same graph
        ports = plan.run()
        print("replay", rep, f"{plan.last_kernel_ms():.1f} ms", flush=True)
        _check(plan, host, nets, table, ref, ports)
    # arenas sized from the first run's use
    plan = hvm64.PlannedBatch(gpu_heap, defs, [int(u * 1.25) + 1024 for u in plan.used()])
    ports = plan.run()
    _check(plan, host, nets, table, ref, ports)
    print("replanned", plan.info(), flush=True)


def test_planned_batch_grows_small_arenas(gpu_heap, oracle):
    """Arenas far too small: every net grows its arena from the pool, granule
    by granule, and the results are unchanged; the use reported per net
    covers at least each net's first granule."""
    import hvm64
    from hvm64 import configs
    book, nets, table = configs.batch_mix(150, seed=12)
    print("case planned batch 150 seed 12, 64-node arenas", flush=True)
    ref = _reference(oracle, book, table)
    host = hvm64.Host(book, gpu_heap)
    defs = [host.defs[table[v][0]] for v in nets]
    plan = hvm64.PlannedBatch(gpu_heap, defs, [64] * len(defs),
                              tuning={"blocks": 0, "mbox_keep": 128, "fill_below": 16, "arena_chunk": 16})
    gpu_heap.set_timeout(30)  # (plans again: a run's state is read before the next change)
    try:
        for rep in range(2):
            ports = plan.run()
            info = plan.info()
            print("replay", rep, f"{plan.last_kernel_ms():.1f} ms", info, flush=True)
            _check(plan, host, nets, table, ref, ports)
            assert info["pool_nodes"] > 0
            assert min(plan.used()) > 1024  # every net outgrew its one-granule arena
    finally:
        gpu_heap.set_timeout(120)


def test_planned_batch_tune_keeps_results(gpu_heap, oracle):
    import hvm64
    from hvm64 import configs
    book, nets, table = configs.batch_mix(120, seed=13)
    print("case planned batch 120 seed 13, tuned", flush=True)
    ref = _reference(oracle, book, table)
    host = hvm64.Host(book, gpu_heap)
    defs = [host.defs[table[v][0]] for v in nets]
    plan = hvm64.PlannedBatch(gpu_heap, defs, hvm64.arena_sizes(gpu_heap, defs))
    plan.run()
    plan = hvm64.PlannedBatch(gpu_heap, defs, [int(u * 1.25) + 1024 for u in plan.used()])
    best = plan.tune()
    print("tuned", best, flush=True)
    assert plan.info()["tuning"] == best
    ports = plan.run()
    _check(plan, host, nets, table, ref, ports)


def test_planned_batch_then_single_net(gpu_heap, oracle):
    """A plan overwrites the heap with arenas; a fresh net afterwards still
    reduces to the oracle's normal form and counts."""
    import hvm64
    from hvm64 import configs
    book, nets, table = configs.batch_mix(64, seed=14)
    host = hvm64.Host(book, gpu_heap)
    defs = [host.defs[table[v][0]] for v in nets]
    plan = hvm64.PlannedBatch(gpu_heap, defs, hvm64.arena_sizes(gpu_heap, defs))
    plan.run()
    text = configs.bitonic(9, 4)
    print("case single net after a plan: bitonic 2^9", flush=True)
    host2 = hvm64.Host(text, gpu_heap)
    net = hvm64.Net(gpu_heap)
    net.boot(host2.defs["main"])
    net.normal()
    r = oracle.run(text, heap_words=1 << 26)
    assert host2.readback(net) == r["nf"]
    assert net.rwts.as_dict() == r["rwts"]


def test_planned_batch_refuses_arenas_beyond_heap(gpu_heap):
    import hvm64
    from hvm64 import configs
    book, nets, table = configs.batch_mix(4, seed=15)
    host = hvm64.Host(book, gpu_heap)
    defs = [host.defs[table[v][0]] for v in nets]
    with pytest.raises(hvm64.HvmError) as e:
        hvm64.PlannedBatch(gpu_heap, defs, [1 << 40] * len(defs))
    assert e.value.kind == "OOM"


def test_planned_batch_4096_against_oracle(oracle):
    """The batched configuration at full size as a planned batch: 4,096 nets
    in arenas of their own, sized from a first run's use, replayed twice;
    every result is its variant's, and the five counts equal the sum of the
    oracle's per-variant counts on both replays."""
    import hvm64
    from hvm64 import configs
    print("case planned batch 4096 seed 7", flush=True)
    book, nets, table = configs.batch_mix(4096, seed=7)
    ref = {}
    for v in table:
        r = oracle.run(book, entry=table[v][0], heap_words=1 << 27)
        assert r["nf"] == table[v][1]
        ref[v] = r["rwts"]
    total = {k: sum(ref[v][k] for v in nets) for k in ["rwts", "anni", "comm", "eras", "dref", "oper"]}
    expected = [hvm64.u24_port(int(table[v][1])) for v in nets]
    heap = hvm64.Heap(200 << 30)
    try:
        heap.set_timeout(60)
        host = hvm64.Host(book, heap)
        defs = [host.defs[table[v][0]] for v in nets]
        plan = hvm64.PlannedBatch(heap, defs, hvm64.arena_sizes(heap, defs, margin=1.3))
        plan.run()
        plan = hvm64.PlannedBatch(heap, defs, [int(u * 1.1) + 1024 for u in plan.used()])
        for rep in range(2):
            ports = plan.run()
            print("replay", rep, f"{plan.last_kernel_ms():.1f} ms", plan.info(), flush=True)
            assert ports == expected
            assert plan.rwts.as_dict() == total
    finally:
        heap.close()
