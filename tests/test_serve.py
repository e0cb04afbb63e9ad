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


@pytest.mark.parametrize("kind,size", [("tree_sum", 8), ("tree_sum", 12), ("bitonic", 6), ("bitonic", 9)])
def test_node_bound_covers_sequential_peak(oracle, kind, size):
    """The admission bound is at least the oracle's peak live nodes."""
    text, expected = configs.variant(kind, size)
    r = oracle.run(text, heap_words=1 << 25)
    assert r["nf"] == expected
    assert serve.node_bound(kind, size) >= serve.PARALLEL_MARGIN * r["peak_nodes"]


@pytest.mark.gpu
def test_serving_results_alone_and_in_the_mix():
    """Requests served in mixed batches on two concurrent slots give the
    same result as each request reduced alone (integer results: exact)."""
    import hvm64
    print("case serving rate 400/s 0.5 s, 2 slots", flush=True)
    srv = serve.Server(slots=2, heap_bytes=4 << 30, max_batch=256)
    try:
        reqs = serve.poisson_requests(400, 0.5, seed=5)
        st = srv.run(reqs)
        assert st.completed == len(reqs) and st.all_ok
        assert st.batches >= 1 and st.p99_ms >= st.p50_ms > 0
        heap, host = srv.slots[0]
        for r in reqs[:6]:
            net = hvm64.Batch(heap)
            net.boot([host.defs[srv.table[(r.kind, r.size)][0]]])
            net.normal()
            assert net.ports() == [r.port] == [r.expected_port]
    finally:
        srv.close()
