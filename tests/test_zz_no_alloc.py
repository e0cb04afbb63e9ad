"""hvm_gpu_boot_batch_noalloc, the allocation-free batch boot, called through
the C ABI: the nets it boots reduce to the same normal forms with the same
five counts as the nets hvm_gpu_boot_batch boots, on two calls in a row (the
second smaller); a batch above its maximum is refused before any device
work."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def _run(heap, host, table, nets, allocation_free):
    import hvm64
    b = hvm64.Batch(heap)
    b.boot([host.defs[table[v][0]] for v in nets], allocation_free=allocation_free)
    b.normal()
    texts = [host.readback_root(b.roots[i]) for i in range(0, len(nets), 97)]
    return b.ports(), b.rwts.as_dict(), texts


def test_boot_batch_noalloc_matches_boot_batch():
    import hvm64
    from hvm64 import configs
    book, nets, table = configs.batch_mix(700, seed=5)
    heap = hvm64.Heap(8 << 30)
    try:
        heap.set_timeout(120)
        host = hvm64.Host(book, heap)
        for n in (700, 233):  # two calls in a row, the second smaller
            print("case noalloc boot, nets", n, flush=True)
            sub = nets[:n]
            ref_ports, ref_rw, ref_texts = _run(heap, host, table, sub, False)
            ports, rw, texts = _run(heap, host, table, sub, True)
            assert ports == ref_ports
            assert rw == ref_rw
            assert texts == ref_texts
            assert ports == [hvm64.u24_port(int(table[v][1])) for v in sub]
    finally:
        heap.close()


def test_boot_batch_noalloc_refuses_oversized_batches():
    import hvm64
    from hvm64 import configs
    heap = hvm64.Heap(1 << 30)
    try:
        host = hvm64.Host(configs.tree_sum(2), heap)
        n = (1 << 16) + 1
        ids = (ctypes.c_uint32 * n)(*([host.defs["main"].idx] * n))
        roots = (ctypes.c_uint64 * n)()
        hvm64.Net(heap)
        rc = heap._lib.hvm_gpu_boot_batch_noalloc(heap._h, n, ids, roots)
        assert rc == 10  # HVM_E_ARG
        assert "65536" in heap._lib.hvm_gpu_last_error(heap._h).decode()
    finally:
        heap.close()
