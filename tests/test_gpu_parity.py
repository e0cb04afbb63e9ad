"""GPU parity: the HIP engine against the reference's golden snapshots and the
oracle.  Normal forms must be identical text and all five rewrite counts
identical (integer work: bit-exact)."""
import os
import random
import re

import pytest

from conftest import GOLDEN, load_snapshots

pytestmark = pytest.mark.gpu

SNAPS = load_snapshots()
# Programs whose reduction is parallel enough to finish in seconds on the GPU;
# the tail-recursive stress tests (sum_tail, tuple_rots, boom, ...) are long
# sequential chains and are exercised by the oracle tests instead.
GPU_CASES = sorted((k for k, v in SNAPS.items() if v["rwts"]["rwts"] <= 3_000_000), key=lambda k: SNAPS[k]["rwts"]["rwts"])
GPU_CASES += ["tests::programs::stress_tests::fib_rec", "tests::programs::stress_tests::sum_rec"]


def run_gpu(heap, text, entry="main"):
    import hvm64
    host = hvm64.Host(text, heap)
    net = hvm64.Net(heap)
    net.boot(host.defs[entry])
    net.normal()
    return host.readback(net), net.rwts.as_dict(), net


@pytest.mark.parametrize("name", GPU_CASES)
def test_snapshot(gpu_heap, name):
    snap = SNAPS[name]
    print("case", name, flush=True)
    nf, rw, _ = run_gpu(gpu_heap, snap["text"])
    assert nf == snap["normal_form"]
    assert rw == snap["rwts"]


@pytest.mark.parametrize("cfg,size", [("tree_sum", 10), ("tree_sum", 14), ("tree_sum", 18),
                                      ("bitonic", 6), ("bitonic", 9), ("bitonic", 12),
                                      ("church", 4), ("church", 10), ("church", 14)])
def test_config_against_oracle(gpu_heap, oracle, cfg, size):
    from hvm64 import configs
    print("case", cfg, size, flush=True)
    text = {"tree_sum": lambda: configs.tree_sum(size), "bitonic": lambda: configs.bitonic(size, 7),
            "church": lambda: configs.church(size, 16)}[cfg]()
    expected = {"tree_sum": lambda: configs.tree_sum_expected(size), "bitonic": lambda: configs.bitonic_expected(size, 7),
                "church": lambda: configs.church_expected(size)}[cfg]()
    nf, rw, net = run_gpu(gpu_heap, text)
    ref = oracle.run(text, heap_words=1 << 28)
    assert nf == expected == ref["nf"]
    assert rw == ref["rwts"]
    assert net.shortcuts == (0, 0, 0)


@pytest.mark.parametrize("k,seed", [(5, 3), (10, 2024)])
def test_bitonic_tree_sorted(gpu_heap, oracle, k, seed):
    from hvm64 import configs
    text = configs.bitonic(k, seed, result="tree")
    nf, rw, _ = run_gpu(gpu_heap, text)
    keys = [int(x) for x in re.findall(r"\(\((\d+) [a-z]+\)", nf)]
    assert keys == sorted(configs.bitonic_keys(k, seed))
    ref = oracle.run(text, heap_words=1 << 26)
    assert nf == ref["nf"] and rw == ref["rwts"]


APPLY_CASES = [
    ("(a a)", ["(a a)"], "(a a)"),
    ("b & (a b) ~ a", ["(a a)"], "a"),
    ("(z0 z0)", ["(z1 z1)"], "(a a)"),
    ("(* 1)", ["(a a)"], "1"),
    ("($([+] $(a b)) (a b))", ["1", "2"], "3"),
    ("($([*] $(a b)) (a b))", ["2", "3"], "6"),
    ("($([*] $(a b)) (a b))", ["2"], "($([*2] a) a)"),
]


@pytest.mark.parametrize("fun,args,expected", APPLY_CASES)
def test_apply_tree_encode_net(gpu_heap, fun, args, expected):
    """tests/cli.rs:206-251 through Host::encode_net + normal() on the GPU."""
    import hvm64
    net_text = fun
    for a in args:
        net_text = hvm64.apply_tree(net_text, a)
    host = hvm64.Host("", gpu_heap)
    net = hvm64.Net(gpu_heap)
    host.encode_net(net, net_text)
    net.normal()
    assert host.readback(net) == expected


ARITH = open(os.path.join(GOLDEN, "programs/examples/arithmetic.hvm")).read()


@pytest.mark.parametrize("book,expr,expected", [
    ("", "1", "1"), ("", "a & 3 ~ $([*] $(4 a))", "12"), (ARITH, "a & @mul ~ (3 (4 a))", "12"),
    (ARITH, "a & @div ~ (64 (2 a))", "32"),
    (ARITH, "a\n  & @main ~ (64 a)", "(#2{$([/64] a) $([%64] b)} {a b})"),
])
def test_reduce_expressions(gpu_heap, oracle, book, expr, expected):
    """tests/cli.rs:41-131 (`hvm64 reduce` / `run`): encode_net, normal, readback."""
    import hvm64
    host = hvm64.Host(book, gpu_heap)
    net = hvm64.Net(gpu_heap)
    host.encode_net(net, expr)
    net.normal()
    nf = host.readback(net)
    ref_nf, ref_counts = oracle.reduce_expr(book, expr)
    assert nf == expected == ref_nf
    r = net.rwts
    assert [r.anni, r.comm, r.eras, r.dref, r.oper] == ref_counts


@pytest.mark.parametrize("kind,cases,expected", [
    ("got", [(0,), (3,), (31,)], [594, 681, 1493]),
    ("put", [(0, 2), (7, 16), (31, 0)], [585, 795, 1515]),
])
def test_lists_counts(gpu_heap, kind, cases, expected):
    """tests/lists.rs:29-77 count vectors."""
    import hvm64
    text = open(os.path.join(GOLDEN, "programs/tests/programs/list_put_got.hvm")).read()
    totals = []
    for args in cases:
        if kind == "got":
            book = hvm64.book_apply(text, "GenGotIndex", f"@S{args[0]}")
            book = hvm64.book_apply(book, "main", "@GenGotIndex")
        else:
            book = hvm64.book_apply(text, "GenPutIndexValue", f"@S{args[0]}")
            book = hvm64.book_apply(book, "GenPutIndexValue", f"@S{args[1]}")
            book = hvm64.book_apply(book, "main", "@GenPutIndexValue")
        totals.append(run_gpu(gpu_heap, book)[1]["rwts"])
    assert totals == expected


def numeric_book(seed, n=120):
    """A chain of random binary operations on random u24 / i24 / f24 operands,
    in the style of tests/programs/u24.hvm."""
    rng = random.Random(seed)
    ops = ["+", "-", ":-", "*", "/", ":/", "%", ":%", "&", "|", "^", "<<", ":<<", ">>", ":>>", "=", "!", "<", ">"]
    lines = ["@main = x & @t0 ~ (* x)"]
    for i in range(n):
        ty = rng.choice(["u", "i", "f"])
        op = rng.choice(ops)
        if ty == "u":
            a, b = rng.randrange(0, 1 << 24), rng.randrange(1, 1 << 24)
            if op in (":/", ":%") and a == 0:
                a = 1
            a, b = str(a), str(b)
        elif ty == "i":
            a, b = rng.randrange(-(1 << 23), 1 << 23), rng.randrange(-(1 << 23), 1 << 23)
            b = b or 1
            a = a or 1
            a, b = f"{a:+d}", f"{b:+d}"
        else:
            if op in ("&", "|", "^", "<<", ":<<", ">>", ":>>"):
                op = "+"
            a, b = f"{rng.uniform(-1e4, 1e4):.3f}", f"{rng.uniform(-1e4, 1e4):.3f}"
        lines.append(f"@t{i} = (* {{n x}}) & @t{i + 1} ~ (* x) & {a} ~ $([{op}] $({b} n))")
    lines.append(f"@t{n} = *")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_numeric_ops_against_oracle(gpu_heap, oracle, seed):
    """u24 / i24 / f24 operators on the device give bit-identical results."""
    text = numeric_book(seed)
    nf, rw, _ = run_gpu(gpu_heap, text)
    ref = oracle.run(text, heap_words=1 << 22)
    assert nf == ref["nf"]
    assert rw == ref["rwts"]


def test_reduce_limit_then_normal(gpu_heap, oracle):
    """Net::reduce(limit) (net.rs:41-52) stops with redexes left (None) and a
    later normal() reaches the same normal form with the same counts."""
    import hvm64
    from hvm64 import configs
    text = configs.bitonic(8, 5)
    host = hvm64.Host(text, gpu_heap)
    net = hvm64.Net(gpu_heap)
    net.boot(host.defs["main"])
    assert net.reduce(1000) is None
    popped = net.rwts.total()
    assert popped > 0
    net.normal()
    ref = oracle.run(text, heap_words=1 << 24)
    assert host.readback(net) == ref["nf"]
    assert net.rwts.as_dict() == ref["rwts"]


@pytest.mark.parametrize("text,kind", [
    ("@main = a & 1 ~ $([/] $(0 a))", "DIV0"),
    ("@main = a & +5 ~ $([%] $(+0 a))", "DIV0"),
    ("@main = * & $(a b) ~ ?(a b)", "UNIMPL_PAIR"),
])
def test_reference_panics_become_errors(gpu_heap, text, kind):
    """Rust panics of the reference (num.rs:183-212, interact.rs:39-42) are
    reported as error codes; the handle stays usable."""
    import hvm64
    with pytest.raises(hvm64.HvmError) as e:
        run_gpu(gpu_heap, text)
    assert e.value.kind == kind
    nf, rw, _ = run_gpu(gpu_heap, SNAPS["examples::machine_u32::num_add"]["text"])
    assert nf == "223"


def test_batch_mix_against_oracle(gpu_heap, oracle):
    """Config 4 (batched): 256 independent nets of the tree-sum / bitonic mix
    in one launch.  Every net's result equals its expected value and the
    oracle's normal form, and the batch's per-rule counts equal the sum of
    the oracle's per-net counts (each net reduced alone)."""
    import hvm64
    from hvm64 import configs
    book, nets, table = configs.batch_mix(256, seed=3)
    print("case batch_mix 256 seed 3", flush=True)
    host = hvm64.Host(book, gpu_heap)
    batch = hvm64.Batch(gpu_heap)
    batch.boot([host.defs[table[v][0]] for v in nets])
    batch.normal()
    ports = batch.ports()
    ref = {}
    for v in table:
        text, expected = configs.variant(*v, seed=3)
        r = oracle.run(text, heap_words=1 << 26)
        assert r["nf"] == expected
        ref[v] = r
    for i, v in enumerate(nets):
        assert ports[i] == hvm64.u24_port(int(table[v][1])), (i, v)
    for i in range(0, 256, 37):
        assert host.readback_root(batch.roots[i]) == ref[nets[i]]["nf"]
    total = {k: 0 for k in ["rwts", "anni", "comm", "eras", "dref", "oper"]}
    for v in nets:
        for k in total:
            total[k] += ref[v]["rwts"][k]
    assert batch.rwts.as_dict() == total


@pytest.mark.parametrize("k", [13, 14])
def test_small_heap_reuses_memory(oracle, k):
    """A heap far smaller than the allocations of the reduction: every lane's
    allocation slice wraps several times and the shared overflow ring is
    used, so freed nodes are reused while the net is live."""
    import hvm64
    from hvm64 import configs
    print("case small heap bitonic", k, flush=True)
    heap = hvm64.Heap(1 << 30)
    try:
        heap.set_timeout(120)
        text = configs.bitonic(k, 11)
        rw, nf, _ = hvm64.normal_with(text, heap=heap)
        assert nf == configs.bitonic_expected(k, 11)
        ref = oracle.run(text, heap_words=1 << 28)
        assert nf == ref["nf"] and rw.as_dict() == ref["rwts"]
    finally:
        heap.close()


def _image_text(img, read, port):
    """Prints the tree a port leads to in the raw heap image, following var
    cells as host/src/readback.rs:38-87 does (lab-0 Ctr, ERA and u24 only)."""
    NONE, MASK = 0xFFFF_FFFF_FFFF_FFF0, 0x0000_FFFF_FFFF_FFF8
    while port & 7 in (0, 1):  # a wire endpoint: the cell holds its target once linked
        v = read(port & MASK)
        assert v != NONE and v != 1 << 63, "unlinked or freed wire in a closed normal form"
        port = v
    tag = port & 7
    if port == 2:
        return "*"
    if tag == 3:
        raw = port & 0xFFFF_FFFF
        assert (raw >> 3) & 31 == 0, "only u24 numbers are printed here"
        return str(raw >> 8)
    assert tag == 7 and port >> 48 == 0, f"unexpected port {port:#x}"
    a = port & MASK
    return f"({_image_text(img, read, read(a))} {_image_text(img, read, read(a + 8))})"


@pytest.mark.parametrize("text,expected", [
    ("@main = (1 (2 (* 3)))", "(1 (2 (* 3)))"),
    ("@main = a\n& @f ~ (1 a)\n@f = (x (x 2))", "(1 2)"),
    ("@main = a\n& @f ~ (4 (5 a))\n@f = (x (y (y (x *))))", "(5 (4 *))"),
])
def test_snapshot_image_matches_readback(gpu_heap, text, expected):
    """hvm_gpu_snapshot: walking the raw heap image on the host from the root
    cell gives the same tree as the device readback."""
    print("case snapshot", repr(text), flush=True)
    nf, _, net = run_gpu(gpu_heap, text)
    assert nf == expected
    img, _ = net.snapshot()
    assert img.heap_base == 1 << 32 and img.heap_words > 0
    cache = {}

    def read(addr):
        w = (addr - img.heap_base) >> 3
        assert 0 <= w < img.heap_words
        if w not in cache:
            _, words = net.snapshot(w, 1)
            cache[w] = words[0]
        return cache[w]

    assert _image_text(img, read, img.root_cell | 1) == expected


# ---------------------------------------------------------------------------
# F24 `&` `|` `^` (num/src/num.rs:243-248) on the device
# ---------------------------------------------------------------------------
def _f24_gpu_items(heap, cases):
    import f24_transcendental as ft
    nf, rw, _ = run_gpu(heap, ft.book([c[:3] for c in cases]))
    items = ft.tuple_items(nf)
    assert len(items) == len(cases)
    assert rw["oper"] == len(cases)
    return items


def test_f24_transcendental_identities_gpu(gpu_heap, oracle):
    """Results fixed by mathematics (x^1 = x, x^0 = 1, 2^k, log_b(b^k) = k,
    atan2 special values), rounded to f24, on the device's atan2f / logf /
    powf."""
    import f24_transcendental as ft
    cases = ft.identity_cases()
    items = _f24_gpu_items(gpu_heap, cases)
    for (op, a, b, want), got in zip(cases, items):
        assert got == oracle.num_display(ft.f24_raw(want)), (op, a, b, want, got)


@pytest.mark.parametrize("seed", [5, 6])
def test_f24_transcendental_random_gpu(gpu_heap, oracle, seed):
    """Random operands whose exact result is away from an f24 rounding
    boundary: the device's f24 result is the exact result's, and so is the
    oracle's (the same book reduced on the CPU gives the same text)."""
    import f24_transcendental as ft
    cases = ft.random_cases(seed=seed, n=90)
    items = _f24_gpu_items(gpu_heap, cases)
    for (op, a, b, v), got in zip(cases, items):
        assert got == oracle.num_display(ft.f24_raw(v)), (op, a, b, v, got)
    ref = oracle.run(ft.book([c[:3] for c in cases]), heap_words=1 << 20)
    assert ft.tuple_items(ref["nf"]) == items


def test_f24_transcendental_nan_gpu(gpu_heap):
    import f24_transcendental as ft
    items = _f24_gpu_items(gpu_heap, ft.nan_cases())
    assert all(x.lstrip("+-") == "NaN" for x in items), items


# ---------------------------------------------------------------------------
# ERA shortcuts of do_op / do_op_num / do_switch (instruction.rs:144-146,
# 163-165, 194-197): books where the called definition's pattern reaches an
# eraser that is already in the argument when the call runs
# ---------------------------------------------------------------------------
SHORTCUT_BOOKS = [
    ("@main = r\n& @f ~ (* r)\n@f = ($([+1] y) y)\n", (0, 1, 0)),
    ("@main = r\n& @f ~ (* r)\n@f = ($(a y) (a y))\n", (1, 0, 0)),
    ("@main = r\n& @f ~ (* (5 r))\n@f = ($(a y) (a y))\n", (1, 0, 0)),
    ("@main = r\n& @f ~ (* r)\n@f = (?(a b) (a b))\n", (0, 0, 1)),
    ("@main = r\n& @f ~ (* (1 r))\n@f = (?((a b) c) (a (b c)))\n", (0, 0, 1)),
    # the eraser reaches the argument through a wire linked before the call
    ("@main = r\n& @f ~ (x r)\n& * ~ x\n@f = ($([*3] y) y)\n", (0, 1, 0)),
    ("@main = r\n& @f ~ (x (y r))\n& * ~ x\n& * ~ y\n@f = (?(a b) ($(c d) (a (b (c d)))))\n", (1, 0, 1)),
]


@pytest.mark.parametrize("text,shortcuts", SHORTCUT_BOOKS)
def test_era_shortcuts_fire(gpu_heap, oracle, text, shortcuts):
    print("case shortcut", repr(text), flush=True)
    nf, rw, net = run_gpu(gpu_heap, text)
    ref = oracle.run(text, heap_words=1 << 16)
    assert ref["shortcuts"] == shortcuts
    assert nf == ref["nf"]
    assert rw == ref["rwts"]
    assert net.shortcuts == ref["shortcuts"]


# ---------------------------------------------------------------------------
# book loading through the encoded forms (hvm_gpu_load_book /
# hvm_gpu_load_book_image), as a host that encodes books itself would
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("via", ["image", "blob"])
@pytest.mark.parametrize("name", ["examples::sort::bitonic::bitonic_sort_lam", "tests::programs::f24",
                                  "examples::arithmetic", "tests::programs::stress_tests::sum_tree"])
def test_encoded_book_paths(gpu_heap, via, name):
    import hvm64
    snap = SNAPS.get(name) or next(v for k, v in SNAPS.items() if k.endswith(name.split("::")[-1]))
    if snap["rwts"]["rwts"] > 200_000_000:
        from hvm64 import configs
        text = configs.tree_sum(12)
        ref_nf = configs.tree_sum_expected(12)
        ref_rw = None
    else:
        text, ref_nf, ref_rw = snap["text"], snap["normal_form"], snap["rwts"]
    host = hvm64.Host(text, gpu_heap, via=via)
    net = hvm64.Net(gpu_heap)
    net.boot(host.defs["main"])
    net.normal()
    assert host.readback(net) == ref_nf
    if ref_rw is not None:
        assert net.rwts.as_dict() == ref_rw


# ---------------------------------------------------------------------------
# full-size configurations against the oracle
# ---------------------------------------------------------------------------
def test_bitonic_full_size_counts(oracle):
    """Bitonic sort of 2^16 keys (the single-net benchmark): normal form and
    all five counts equal the oracle's."""
    import hvm64
    from hvm64 import configs
    print("case bitonic 2^16 seed 7", flush=True)
    heap = hvm64.Heap(64 << 30)
    try:
        heap.set_timeout(120)
        text = configs.bitonic(16, 7)
        rw, nf, _ = hvm64.normal_with(text, heap=heap)
    finally:
        heap.close()
    ref = oracle.run(text, heap_words=1 << 29)
    assert nf == ref["nf"] == configs.bitonic_expected(16, 7)
    assert rw.as_dict() == ref["rwts"]


def test_church_full_size(oracle):
    """Church c2^c22 applied to +1 on 0 with 16 DUP labels (a sequential
    chain of about 4M increments): result 2^22 and the oracle's counts."""
    import hvm64
    from hvm64 import configs
    print("case church c2^c22", flush=True)
    heap = hvm64.Heap(32 << 30)
    try:
        heap.set_timeout(110)
        text = configs.church(22, 16)
        rw, nf, _ = hvm64.normal_with(text, heap=heap)
    finally:
        heap.close()
    ref = oracle.run(text, heap_words=1 << 28)
    assert nf == ref["nf"] == configs.church_expected(22)
    assert rw.as_dict() == ref["rwts"]


def test_batch_4096_against_oracle(oracle):
    """The batched configuration at full size: 4,096 nets of the mix in one
    launch; every result is its variant's, and the five counts equal the sum
    of the oracle's per-variant counts."""
    import hvm64
    from hvm64 import configs
    print("case batch 4096 seed 7", flush=True)
    book, nets, table = configs.batch_mix(4096, seed=7)
    heap = hvm64.Heap(64 << 30)
    try:
        heap.set_timeout(120)
        host = hvm64.Host(book, heap)
        batch = hvm64.Batch(heap)
        batch.boot([host.defs[table[v][0]] for v in nets])
        batch.normal()
        ports = batch.ports()
        got = batch.rwts.as_dict()
    finally:
        heap.close()
    ref = {}
    for v in table:
        r = oracle.run(book, entry=table[v][0], heap_words=1 << 27)
        assert r["nf"] == table[v][1]
        ref[v] = r["rwts"]
    for i, v in enumerate(nets):
        assert ports[i] == hvm64.u24_port(int(table[v][1])), (i, v)
    total = {k: sum(ref[v][k] for v in nets) for k in got}
    assert got == total


@pytest.mark.parametrize("limit", [1, 7, 1000, 50_000])
def test_reduce_limit_bounds_pops(gpu_heap, oracle, limit):
    """Net::reduce(limit) (net.rs:41-52) pops at most `limit` redexes per
    call; repeated calls reach the normal form with the oracle's counts."""
    import hvm64
    from hvm64 import configs
    text = configs.bitonic(7, 3)
    host = hvm64.Host(text, gpu_heap)
    net = hvm64.Net(gpu_heap)
    net.boot(host.defs["main"])
    calls = 0
    while True:
        r = net.reduce(limit)
        assert net.last_popped <= limit
        calls += 1
        if r is not None:
            break
        assert net.last_popped == limit
        assert calls < 2_000_000 // limit + 10
    net.normal()
    ref = oracle.run(text, heap_words=1 << 24)
    assert host.readback(net) == ref["nf"]
    assert net.rwts.as_dict() == ref["rwts"]


def _with_main(book, body):
    """The book with @main's definition (its tree and its redex lines)
    replaced by `body`."""
    out, skip = [], False
    for line in book.splitlines():
        if line.startswith("@"):
            skip = re.match(r"@main\s*=", line) is not None
            if skip:
                out.append("@main = " + body)
                continue
        if not skip:
            out.append(line)
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("limit", [1, 1000, 20_000])  # (47,181 pops in all)
def test_readback_after_reduce_limit(gpu_heap, oracle, limit):
    """Host::readback after Net::reduce(limit) (readback.rs:28-30): the root's
    tree and every pending redex as "& a ~ b".  Which redexes are pending
    depends on the order of reduction, so the text is checked by what it
    must mean: booted in place of @main, the partial net reduces on the
    oracle to the full program's normal form, and the counts of the GPU's
    first part plus the oracle's rest equal the oracle's counts for the
    whole (less the one DREF of booting the partial net)."""
    import hvm64
    from hvm64 import configs
    text = configs.bitonic(7, 3)
    print("case readback after reduce", limit, flush=True)
    host = hvm64.Host(text, gpu_heap)
    net = hvm64.Net(gpu_heap)
    net.boot(host.defs["main"])
    assert net.reduce(limit) is None
    part = host.readback(net)
    assert "\n  & " in part
    done = net.rwts.as_dict()
    full = oracle.run(text, heap_words=1 << 24)
    rest = oracle.run(_with_main(text, part), heap_words=1 << 24)
    assert rest["nf"] == full["nf"]
    for k in ["anni", "comm", "eras", "oper"]:
        assert done[k] + rest["rwts"][k] == full["rwts"][k], k
    assert done["dref"] + rest["rwts"]["dref"] - 1 == full["rwts"]["dref"]
    # and the net itself still reaches the normal form with the full counts
    net.normal()
    assert host.readback(net) == full["nf"]
    assert net.rwts.as_dict() == full["rwts"]


# ---------------------------------------------------------------------------
# interpreted and compiled definitions give the same results
# ---------------------------------------------------------------------------
INTERP_CASES = ["examples::sort::bitonic::bitonic_sort_lam", "tests::programs::f24", "tests::programs::u24",
                "examples::church_encoding::church", "tests::programs::list_put_got"]


@pytest.mark.parametrize("name", [n for n in INTERP_CASES if n in SNAPS])
def test_interpreted_definitions_match_snapshot(gpu_heap, name):
    """The same programs with compiled definitions off: every definition is
    interpreted instruction by instruction (def.rs:217-289)."""
    import hvm64
    snap = SNAPS[name]
    host = hvm64.Host(snap["text"], gpu_heap)
    assert gpu_heap.set_compiled(False) > 0
    try:
        net = hvm64.Net(gpu_heap)
        net.boot(host.defs["main"])
        net.normal()
        assert host.readback(net) == snap["normal_form"]
        assert net.rwts.as_dict() == snap["rwts"]
    finally:
        gpu_heap.set_compiled(True)


@pytest.mark.parametrize("text,shortcuts", SHORTCUT_BOOKS[:4])
def test_era_shortcuts_interpreted(gpu_heap, oracle, text, shortcuts):
    import hvm64
    host = hvm64.Host(text, gpu_heap)
    gpu_heap.set_compiled(False)
    try:
        net = hvm64.Net(gpu_heap)
        net.boot(host.defs["main"])
        net.normal()
        ref = oracle.run(text, heap_words=1 << 16)
        assert host.readback(net) == ref["nf"]
        assert net.rwts.as_dict() == ref["rwts"]
        assert net.shortcuts == ref["shortcuts"] == shortcuts
    finally:
        gpu_heap.set_compiled(True)


def test_large_normal_form_readback(gpu_heap, oracle):
    """A normal form that is a wide tree (the 2^12 keys the bitonic sort
    generates, about 28k nodes) read back level by level on the device is the
    oracle's text (readback.rs:23-87 order, variable names by first
    occurrence), with the same counts."""
    from hvm64 import configs
    text = configs.bitonic(12, 7)
    main = "@main = a\n& @sum ~ (b a)\n& @sort ~ (c (0 b))\n& @rev ~ (d c)\n& @gen ~ (12 (0 d))"
    assert main in text
    text = text.replace(main, "@main = a\n& @gen ~ (12 (0 a))")
    print("case large readback: the generated tree of 2^12 keys", flush=True)
    r = oracle.run(text, heap_words=1 << 26)
    assert r["nf"].count("(") > 4000
    nf, rw, _ = run_gpu(gpu_heap, text)
    assert nf == r["nf"]
    assert rw == r["rwts"]


def test_sched_stats_switch(gpu_heap, oracle):
    """A single net's scheduler statistics only when asked for
    (hvm_gpu_set_sched_stats); the pop count either way, and the same
    result."""
    import hvm64
    from hvm64 import configs
    text = configs.tree_sum(12)
    print("case sched timing on/off: tree-sum 12", flush=True)
    r = oracle.run(text, heap_words=1 << 26)
    host = hvm64.Host(text, gpu_heap)
    out = {}
    try:
        for on in (False, True):
            gpu_heap.set_sched_stats(on)
            net = hvm64.Net(gpu_heap)
            net.boot(host.defs["main"])
            net.normal()
            assert host.readback(net) == r["nf"]
            assert net.rwts.as_dict() == r["rwts"]
            out[on] = net.sched_stats()
    finally:
        gpu_heap.set_sched_stats(False)
    assert out[False]["pops"] > 0
    assert out[False]["busy_iters"] == 0 and out[False]["busy_ticks"] == 0 and out[False]["call_ticks"] == 0
    assert out[True]["busy_iters"] > 0 and out[True]["busy_ticks"] > 0 and out[True]["interact_ticks"] > 0
