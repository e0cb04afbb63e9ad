"""The reference's own unit vectors, checked against the oracle."""
import os

import pytest

from conftest import GOLDEN


def test_create_var(oracle):
    """util/src/create_var.rs:35-49."""
    for i, s in [(0, "a"), (1, "b"), (25, "z"), (26, "aa"), (27, "ab"), (51, "az"), (52, "ba"), (676, "za"),
                 (701, "zz"), (702, "aaa"), (703, "aab"), (728, "aba"), (1351, "ayz"), (1352, "aza"), (1378, "baa")]:
        assert oracle.create_var(i) == s


CALC_LABELS_BOOK = """
@a = {0 @b @c}
@b = {1 @a *}
@c = {2 @a *}
@p = {3 @q {4 x x}}
@q = {5 @r @t}
@r = {6 @s *}
@s = {7 @r *}
@t = {8 @q {9 @t @u}}
@u = {10 @u @s}
""".replace("{0 ", "#0{").replace("{1 ", "#1{").replace("{2 ", "#2{").replace("{3 ", "#3{").replace(
    "{4 ", "#4{").replace("{5 ", "#5{").replace("{6 ", "#6{").replace("{7 ", "#7{").replace("{8 ", "#8{").replace(
    "{9 ", "#9{").replace("{10 ", "#10{")


def calc_labels_expected():
    """host/src/calc_labels.rs:196-234, with `{k ...}` read as label k.

    The reference test parses `{k x y}` through the test's own parser; here the
    same label numbers are written with the `#k{...}` syntax, whose label is
    2k+1 (ast/src/parser.rs:52-67)."""
    raw = {"a": [0, 1, 2], "b": [0, 1, 2], "c": [0, 1, 2], "p": [3, 4, 5, 6, 7, 8, 9, 10],
           "q": [5, 6, 7, 8, 9, 10], "r": [6, 7], "s": [6, 7], "t": [5, 6, 7, 8, 9, 10], "u": [6, 7, 10]}
    return {k: [2 * x + 1 for x in v] for k, v in raw.items()}


def test_calculate_labels(oracle):
    assert oracle.label_sets(CALC_LABELS_BOOK) == calc_labels_expected()


@pytest.mark.parametrize("kind,cases,expected", [
    ("got", [(0,), (1,), (3,), (7,), (15,), (31,)], [594, 623, 681, 797, 1029, 1493]),
    ("put", [(0, 2), (1, 4), (3, 8), (7, 16), (15, 32), (31, 0)], [585, 615, 675, 795, 1035, 1515]),
])
def test_lists_counts(oracle, kind, cases, expected):
    """tests/lists.rs:29-77 (single-threaded normal())."""
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
        totals.append(oracle.run(book, heap_words=1 << 22)["rwts"]["rwts"])
    assert totals == expected
    delta = totals[1] - totals[0]
    for i, mult in zip(range(1, 5), [2, 4, 8, 16]):
        assert totals[i] + delta * mult == totals[i + 1]


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
def test_apply_tree(oracle, fun, args, expected):
    """tests/cli.rs:206-251: apply_tree + normal() on a 16-word heap."""
    assert oracle.apply_tree_normal(fun, args, heap_words=16) == expected


ARITH = open(os.path.join(GOLDEN, "programs/examples/arithmetic.hvm")).read()
REDUCE_CASES = [("", "1", "1"), ("", "a & 3 ~ $([*] $(4 a))", "12"), ("", "a & 64 ~ $([/] $(2 a))", "32"),
                (ARITH, "a & @mul ~ (3 (4 a))", "12"), (ARITH, "a & @div ~ (64 (2 a))", "32")]


@pytest.mark.parametrize("book,expr,expected", REDUCE_CASES)
def test_reduce_expressions(oracle, book, expr, expected):
    """tests/cli.rs:41-90 (`hvm64 reduce`)."""
    assert oracle.reduce_expr(book, expr)[0] == expected


RUN_CASES = [([], "(#1{$([/] $(a b)) $([%] $(c d))} (#2{a c} {b d}))"), (["64"], "(#2{$([/64] a) $([%64] b)} {a b})"),
             (["64", "3"], "{21 1}")]


@pytest.mark.parametrize("args,expected", RUN_CASES)
def test_run_with_args(oracle, args, expected):
    """tests/cli.rs:93-131 (`hvm64 run` with arguments, src/main.rs:93-101)."""
    import hvm64
    net = "@main"
    for a in args:
        net = hvm64.apply_tree(net, a)
    assert oracle.reduce_expr(ARITH, net)[0] == expected
