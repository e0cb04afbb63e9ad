"""Serving (continuous batching): host logic on the CPU, results on the GPU."""
import pytest

from hvm64 import configs, serve


def test_poisson_requests_are_seeded():
    a = serve.poisson_requests(2000, 0.5, seed=11)
    b = serve.poisson_requests(2000, 0.5, seed=11)
    assert [(r.arrival, r.kind, r.size) for r in a] == [(r.arrival, r.kind, r.size) for r in b]
    assert 800 < len(a) < 1200  # 2000/s over 0.5 s
    assert all(0 <= r.arrival < 0.5 for r in a)
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))
    kinds = {r.kind for r in a}
    assert kinds == {"tree_sum", "bitonic"}
    for r in a:
        lo, hi = (8, 16) if r.kind == "tree_sum" else (6, 12)
        assert lo <= r.size <= hi


def test_percentile():
    xs = list(range(1, 101))
    assert serve.percentile(xs, 50) == 50
    assert serve.percentile(xs, 99) == 99
    assert serve.percentile(xs, 100) == 100
    assert serve.percentile([7], 99) == 7


@pytest.mark.parametrize("heap_gb,slot_nodes,slots", [(16, 1 << 20, 1024), (64, 28864928, 128),
                                                     (192, 28864928, 1024), (192, 5000, 64)])
def test_slot_layout_fits_the_heap(heap_gb, slot_nodes, slots):
    """Every slot holds at least slot_nodes nodes in eight sub-arenas of whole
    granules, the slots and the fixed-cell granule fit the heap, and one more
    slot would not (when fewer than asked for fit)."""
    heap_nodes = (heap_gb << 30) // 16
    n, per_slot, gran = serve.slot_layout(heap_gb << 30, slot_nodes, slots)
    assert heap_nodes // gran <= 4 << 20 and (gran == 1024 or heap_nodes // (gran // 2) > 4 << 20)
    assert per_slot >= slot_nodes and per_slot % (8 * gran) == 0 and per_slot - slot_nodes < 8 * gran
    assert 0 < n <= slots
    assert gran + n * per_slot <= heap_nodes
    if n < slots:
        assert gran + (n + 1) * per_slot > (heap_nodes // gran) * gran


def test_slot_layout_rejects_a_slot_beyond_the_heap():
    assert serve.slot_layout(1 << 30, 1 << 30, 4)[0] == 0


@pytest.mark.gpu
def test_serving_results_alone_and_in_the_mix():
    """Requests admitted by the running kernel (hvm_gpu_serve_*) at 1,200
    requests/s for one second, many in flight at once in one launch, give the
    same result as each request reduced alone (integer results: exact), and
    the launch's rewrite counts equal the sum of each request's counts."""
    import hvm64
    print("case serving 1200/s for 1 s, 128 slots", flush=True)
    srv = serve.Server(slots=128, heap_bytes=64 << 30, timeout_s=60)
    print("slots", srv.n_slots, "of", srv.slot_nodes, "nodes", flush=True)
    assert srv.n_slots >= 64
    try:
        reqs = serve.poisson_requests(1200, 1.0, seed=5)
        st = srv.run(reqs)
        print("served", st.completed, "p50", st.p50_ms, "p99", st.p99_ms, "mips", st.mips, flush=True)
        assert st.completed == len(reqs) and st.all_ok
        assert st.p99_ms >= st.p50_ms > 0
        served_total = st.rewrites
        # alone: every variant that was served, each on a fresh net
        alone = {}
        for v in sorted({(r.kind, r.size) for r in reqs}):
            net = hvm64.Net(srv.heap)
            net.boot(srv.host.defs[srv.table[v][0]])
            net.normal()
            alone[v] = (net.rwts.total(), hvm64.u24_port(int(srv.host.readback(net))))
        for r in reqs:
            assert r.port == alone[(r.kind, r.size)][1] == r.expected_port
        assert served_total == sum(alone[(r.kind, r.size)][0] for r in reqs)
    finally:
        srv.close()


@pytest.mark.gpu
def test_serving_refuses_other_calls_and_restarts():
    """While serving, the handle's other entry points are refused; after
    stop the handle reduces single nets again, and a second serving launch
    on the same handle serves correctly."""
    import hvm64
    srv = serve.Server(slots=16, heap_bytes=8 << 30, timeout_s=60)
    try:
        s = hvm64.Serving(srv.heap, 16, srv.slot_nodes)
        with pytest.raises(hvm64.HvmError) as e:
            hvm64.Net(srv.heap)
        assert e.value.kind == "ARG"
        v = ("tree_sum", 9)
        s.submit([srv.host.defs[srv.table[v][0]]] * 5, [1, 2, 3, 4, 5])
        got = []
        import time
        t0 = time.time()
        while len(got) < 5 and time.time() - t0 < 30:
            got += s.poll()
        s.stop()
        assert sorted(t for t, _ in got) == [1, 2, 3, 4, 5]
        assert all(p == hvm64.u24_port(int(srv.table[v][1])) for _, p in got)
        reqs = serve.poisson_requests(300, 0.3, seed=6)
        st = srv.run(reqs)
        assert st.completed == len(reqs) and st.all_ok
    finally:
        srv.close()
