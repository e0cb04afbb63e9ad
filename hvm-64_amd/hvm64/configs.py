"""Seeded generators of the benchmark workloads, as .hvm book text.

1. tree_sum(depth): the reference's tree-sum stress test
   (tests/programs/stress_tests/sum_tree.hvm) at a chosen depth: `@gen`
   builds a complete binary tree of CTR label-0 nodes with Num leaves 1 and
   `@sum` folds it with u24 `+`.  Result 2**depth mod 2**24.
2. bitonic(k, seed): the reference's bitonic sort
   (examples/sort/bitonic/bitonic_sort_lam.hvm) over 2**k leaves whose keys
   are a seeded u24 hash of the leaf position, computed in the net by `@hash`.
   `@main` returns the u24 sum of the sorted keys (or the sorted tree).
3. church(exp, labels): Church exponentiation c2 ^ c_exp applied to a Num
   increment on 0, every numeral a pure CTR/DUP net; the duplicators of c_exp
   cycle through `labels` distinct DUP labels.  Result 2**exp mod 2**24.
4. batch_mix(n, seed): n independent nets, each a seeded choice of tree_sum
   with depth uniform in [8, 16] or bitonic with k uniform in [6, 12], as one
   book holding every (program, size) variant under its own name prefix.
5. serving_sizes: the same mix with sizes log-uniform over those ranges.
"""
import math
import os
import random
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
# The two reference programs the workloads are built from (sum_tree.hvm and
# bitonic_sort_lam.hvm of HigherOrderCO/hvm-64, Apache-2.0), shipped with
# the package.
_PROGRAMS = os.path.join(_HERE, "programs")

U24 = 1 << 24
HASH_M1 = 0x5D2B93
HASH_M2 = 0x3C6EF3


def _program(name):
    with open(os.path.join(_PROGRAMS, name)) as f:
        return f.read()


def tree_sum(depth):
    """Book whose @main sums a complete binary tree of 2**depth ones."""
    text = _program("sum_tree.hvm")
    text, n = re.subn(r"@gen ~ \(24 b\)", f"@gen ~ ({int(depth)} b)", text)
    assert n == 1
    return text


def tree_sum_expected(depth):
    return str((1 << depth) % U24)


def key_hash(i, seed):
    """u24 hash of a leaf position, as computed by the net's @hash."""
    t = (i * HASH_M1 + seed) % U24
    u = t ^ (t >> 11)
    v = (u * HASH_M2) % U24
    return v ^ (v >> 13)


def _hash_def(seed):
    return (f"@hash = (x k)\n"
            f"& x ~ $([*{HASH_M1}] $([+{seed % U24}] {{t1 t2}}))\n"
            f"& t1 ~ $([>>] $(11 s))\n"
            f"& t2 ~ $([^] $(s u))\n"
            f"& u ~ $([*{HASH_M2}] {{v1 v2}})\n"
            f"& v1 ~ $([>>] $(13 w))\n"
            f"& v2 ~ $([^] $(w k))\n")


def bitonic(k, seed=1, result="sum"):
    """Book whose @main bitonic-sorts 2**k hashed keys.

    result="sum": the u24 sum of the sorted keys; result="tree": the sorted
    Scott-encoded tree itself (for sortedness checks at small k)."""
    text = _program("bitonic_sort_lam.hvm")
    text, n1 = re.subn(r"@gen\$S0 = \(a b\)\n& @Leaf ~ \(a b\)\n",
                       "@gen$S0 = (a b)\n& @Leaf ~ (h b)\n& @hash ~ (a h)\n", text)
    text, n2 = re.subn(r"& @gen ~ \(10 \(0 d\)\)", f"& @gen ~ ({int(k)} (0 d))", text)
    assert n1 == 1 and n2 == 1
    if result == "tree":
        text, n3 = re.subn(r"@main = a\n& @sum ~ \(b a\)\n& @sort ~ \(c \(0 b\)\)",
                           "@main = b\n& @sort ~ (c (0 b))", text)
        assert n3 == 1
    return text + _hash_def(seed)


def bitonic_keys(k, seed=1):
    return [key_hash(i, seed) for i in range(1 << k)]


def bitonic_expected(k, seed=1):
    return str(sum(bitonic_keys(k, seed)) % U24)


def church(exp=22, labels=16, base=2):
    """Book computing (c_base ^ c_exp) inc 0 with Church numerals.

    c_n = ({d1 ... (x1 x2) ...} (x0 R)) (x_n R): f is copied n-1 times by a
    chain of DUP nodes.  c_exp's duplicators use labels 2..labels+1 (cycling),
    c_base's use label 1 (`{}`)."""
    def numeral(n, lab_of):
        # f applied n times: (f1 ... ) composition x0 -> x1 -> ... -> xn
        if n == 1:
            return "((a R) (a R))"
        apps = [f"(x{i} x{i + 1})" for i in range(n)]
        tree = apps[0]
        for i in range(1, n):
            lab = lab_of(i)
            tree = f"#{lab}{{{tree} {apps[i]}}}"
        return f"({tree} (x0 x{n}))"
    cb = numeral(base, lambda i: 0)
    cb = cb.replace("#0{", "{")
    ce = numeral(exp, lambda i: 1 + (i - 1) % labels)
    return (f"@base = {cb}\n"
            f"@exp = {ce}\n"
            f"@inc = ($([+1] r) r)\n"
            f"@main = r\n"
            f"& @exp ~ (@base (@inc (0 r)))\n")


def church_expected(exp=22, base=2):
    return str((base ** exp) % U24)


_NAME = re.compile(r"@([A-Za-z0-9_.$]+)")


def prefixed(book_text, prefix):
    """The book with every definition name prefixed (refs included)."""
    return _NAME.sub(lambda m: "@" + prefix + m.group(1), book_text)


def variant(kind, size, seed=7):
    """(book_text, expected_result) of one (program, size) variant."""
    if kind == "tree_sum":
        return tree_sum(size), tree_sum_expected(size)
    if kind == "bitonic":
        key_seed = (seed * 31 + size) % U24
        return bitonic(size, key_seed), bitonic_expected(size, key_seed)
    raise ValueError(kind)


def variant_prefix(kind, size):
    return f"{'t' if kind == 'tree_sum' else 'b'}{size}_"


def mix_choice(rng, log_uniform=False):
    """One net of the mix: a coin picks tree_sum (depth in [8, 16]) or bitonic
    (k in [6, 12]); sizes are uniform, or log-uniform over the same ranges."""
    kind = "tree_sum" if rng.random() < 0.5 else "bitonic"
    lo, hi = (8, 16) if kind == "tree_sum" else (6, 12)
    if log_uniform:
        size = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
    else:
        size = rng.randint(lo, hi)
    return kind, size


def split_defs(book_text):
    """{name: body} of a book's definitions: a definition is a line
    `@name = net` and the `& a ~ b` lines that follow it."""
    defs, name = {}, None
    for line in book_text.splitlines():
        m = re.match(r"@([A-Za-z0-9_.$]+)\s*=(.*)$", line)
        if m:
            name = m.group(1)
            defs[name] = [m.group(2).strip()]
        elif line.strip() and name is not None:
            defs[name].append(line.strip())
    return {k: "\n".join(v) for k, v in defs.items()}


def mix_book(variants, seed=7):
    """One book holding each (kind, size) variant; returns (book_text,
    {(kind, size): (main_name, expected)}).

    A definition whose text is the same in every size of a kind, and whose
    references all lead to such definitions, is written once under the
    kind's prefix (`t_` / `b_`); the others (main, and for bitonic sort the
    key hash and the generator that calls it) are written per variant under
    its prefix.  Every net reduces exactly as it would with its own book."""
    table, by_kind = {}, {}
    for kind, size in sorted(set(variants)):
        text, expected = variant(kind, size, seed)
        by_kind.setdefault(kind, {})[size] = split_defs(text)
        table[(kind, size)] = (variant_prefix(kind, size) + "main", expected)
    parts = []
    for kind, sizes in sorted(by_kind.items()):
        books = list(sizes.values())
        names = set(books[0])
        # main is always per variant: the table names it by variant prefix
        shared = {n for n in names if n != "main" and all(b.get(n) == books[0][n] for b in books)}
        changed = True
        while changed:  # a shared definition may only reach shared ones
            changed = False
            for n in sorted(shared):
                if any(r not in shared for r in _NAME.findall(books[0][n])):
                    shared.discard(n)
                    changed = True
        kpre = variant_prefix(kind, 0)[0] + "_"

        def emit(name, body, pre):
            body = _NAME.sub(lambda m: "@" + (kpre if m.group(1) in shared else pre) + m.group(1), body)
            lines = body.split("\n")
            return "\n".join([f"@{pre}{name} = {lines[0]}"] + lines[1:])

        for n in sorted(shared):
            parts.append(emit(n, books[0][n], kpre))
        for size, defs in sorted(sizes.items()):
            pre = variant_prefix(kind, size)
            for n in sorted(defs):
                if n not in shared:
                    parts.append(emit(n, defs[n], pre))
    return "\n".join(parts) + "\n", table


def batch_mix(n, seed=7, log_uniform=False):
    """n nets of the batched workload: (book_text, nets, table) where nets[i]
    = (kind, size) and table maps each variant to (main_name, expected)."""
    rng = random.Random(seed)
    nets = [mix_choice(rng, log_uniform) for _ in range(n)]
    book, table = mix_book(nets, seed)
    return book, nets, table
