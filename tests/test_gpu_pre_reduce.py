"""Book::pre_reduce on the device (hvm_gpu_pre_reduce) against the
reference's own vectors (transform/src/pre_reduce.rs:38-164):

  * each `pre_reduce_run@*` snapshot (tests/tests.rs:44-55) whose
    pre-reduction is not one long sequential chain: every definition but main
    pre-reduced on the GPU gives the snapshot's pre-reduce counts, and the new
    book (now loaded on the heap) runs main on the GPU to the snapshot's
    normal form and run counts; the same book text also runs on the oracle to
    the snapshot's result, so what the device wrote back is the right book;
  * test_fast_pre_reduce (tests/transform.rs:14-46): black_box removed, so
    calls on it are inert and stay pending; only two definitions reduced;
    pre-reduce plus the run of the entry point give the Fast and Slow count
    sets exactly;
  * a budget of one interaction per definition (reduce(limit), net.rs:41-52):
    the pending redexes are written back into each definition and the book
    still runs main to the same normal form."""
import re

import pytest

from test_pre_reduce import FIX, SNAPS, _text, split_out

pytestmark = pytest.mark.gpu

# the two whose pre-reduction is a chain of 18M / 33M sequential interactions
# (burn, apelacion) are covered by the oracle tests only
GPU_PRE = sorted(k for k, v in SNAPS.items() if v["pre_rwts"]["rwts"] <= 1_000_000)
RUN_ON_GPU = {k for k in GPU_PRE if SNAPS[k]["run_rwts"]["rwts"] <= 3_000_000}


@pytest.mark.parametrize("key", GPU_PRE)
def test_gpu_pre_reduce_snapshot(gpu_heap, oracle, key):
    import hvm64
    v = SNAPS[key]
    print("case pre-reduce", key, flush=True)
    book, rw = hvm64.pre_reduce(gpu_heap, _text(v["program"]), skip=["main"])
    assert rw.as_dict() == v["pre_rwts"]
    if key in RUN_ON_GPU:
        host = hvm64.Host(book, gpu_heap)
        net = hvm64.Net(gpu_heap)
        net.boot(host.defs["main"])
        net.normal()
        assert host.readback(net) == v["normal_form"]
        assert net.rwts.as_dict() == v["run_rwts"]
        r = oracle.run(book, heap_words=1 << 26)
        assert r["nf"] == v["normal_form"] and r["rwts"] == v["run_rwts"]


@pytest.mark.parametrize("mode", ["fast", "slow"])
def test_gpu_fast_pre_reduce(gpu_heap, mode):
    import hvm64
    fp = FIX["fast_pre_reduce"]
    m = fp[mode]
    print("case fast_pre_reduce", mode, flush=True)
    rest, removed = split_out(_text(fp["program"]), fp["removed"])
    book, rw = hvm64.pre_reduce(gpu_heap, rest, only=m["reduce"])
    assert not re.search(r"^@" + fp["removed"] + " = ", book, re.M)
    host = hvm64.Host(book + "\n" + removed, gpu_heap)
    net = hvm64.Net(gpu_heap)
    net.boot(host.defs[m["entry"]])
    net.normal()
    total = {k: rw.as_dict()[k] + net.rwts.as_dict()[k] for k in m["rwts"]}
    assert total == m["rwts"]


def test_gpu_pre_reduce_budget(gpu_heap, oracle):
    import hvm64
    text = _text(SNAPS["examples::sort::bitonic::bitonic_sort_lam"]["program"])
    print("case pre-reduce budget 1", flush=True)
    book, rw = hvm64.pre_reduce(gpu_heap, text, skip=["main"], max_rwts=1)
    # every definition but main is booted once (a dereference each); which
    # redex the device pops first is its own choice, so only the bounds hold
    n_defs = len(re.findall(r"^@", text, re.M)) - 1
    assert rw.dref >= n_defs
    assert "&" in book
    full = oracle.run(text, heap_words=1 << 26)
    assert oracle.run(book, heap_words=1 << 26)["nf"] == full["nf"]
