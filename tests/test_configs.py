"""Benchmark workload generators: the oracle's results match the closed forms
(2**d for tree-sum, the host-computed key sum for bitonic, 2**e for Church
exponentiation), and the bitonic output tree is sorted."""
import re

import pytest

from hvm64 import configs


@pytest.mark.parametrize("depth", [1, 4, 8, 12])
def test_tree_sum(oracle, depth):
    r = oracle.run(configs.tree_sum(depth), heap_words=1 << 24)
    assert r["nf"] == configs.tree_sum_expected(depth)


def test_tree_sum_counts_are_linear(oracle):
    """Rewrites per leaf are constant (the reference's sum_tree at depth 24
    does 771,751,906 rewrites: 46 per leaf, plus a constant)."""
    t = [oracle.run(configs.tree_sum(d), heap_words=1 << 24)["rwts"]["rwts"] for d in (8, 9, 10)]
    assert t[2] - t[1] == 2 * (t[1] - t[0])
    big = 771_751_906
    per_leaf = (t[1] - t[0]) // (1 << 8)
    assert per_leaf * (1 << 24) + (t[0] - per_leaf * (1 << 8)) == big


@pytest.mark.parametrize("k,seed", [(3, 1), (6, 7), (9, 12345)])
def test_bitonic_sum(oracle, k, seed):
    r = oracle.run(configs.bitonic(k, seed), heap_words=1 << 24)
    assert r["nf"] == configs.bitonic_expected(k, seed)


def test_bitonic_reference_size_matches_snapshot(oracle, snapshots):
    """With identity keys the generator's program is the reference's example."""
    snap = snapshots["examples::sort::bitonic::bitonic_sort_lam"]
    r = oracle.run(snap["text"], heap_words=1 << 24)
    assert r["rwts"] == snap["rwts"]


@pytest.mark.parametrize("k,seed", [(4, 3), (7, 99)])
def test_bitonic_tree_is_sorted(oracle, k, seed):
    r = oracle.run(configs.bitonic(k, seed, result="tree"), heap_words=1 << 24)
    keys = [int(x) for x in re.findall(r"\(\((\d+) [a-z]+\)", r["nf"])]
    assert keys == sorted(configs.bitonic_keys(k, seed))


@pytest.mark.parametrize("exp", [2, 3, 5, 8])
def test_church(oracle, exp):
    r = oracle.run(configs.church(exp, 16), heap_words=1 << 24)
    assert r["nf"] == configs.church_expected(exp)


def test_mix_book_shares_definitions_without_changing_results(oracle):
    """The batched/serving book writes definitions common to every size of a
    program once; each variant's main still reduces to the same normal form
    with the same five counts as that variant's own book."""
    variants = [("tree_sum", 8), ("tree_sum", 10), ("bitonic", 6), ("bitonic", 8)]
    book, table = configs.mix_book(variants, seed=7)
    own = sum(len(configs.split_defs(configs.variant(k, s, 7)[0])) for k, s in variants)
    assert len(configs.split_defs(book)) < own
    for kind, size in variants:
        main, expected = table[(kind, size)]
        shared = oracle.run(book, heap_words=1 << 24, entry=main)
        alone = oracle.run(configs.variant(kind, size, 7)[0], heap_words=1 << 24)
        assert shared["nf"] == alone["nf"] == expected
        assert shared["rwts"] == alone["rwts"]


@pytest.mark.parametrize("variants", [[("tree_sum", 10)], [("bitonic", 7)], [("tree_sum", 9), ("bitonic", 6)]])
def test_mix_book_with_one_size_per_kind(oracle, variants):
    """A kind present at a single size still names its main by the variant
    prefix that the table gives."""
    book, table = configs.mix_book(variants, seed=7)
    defs = configs.split_defs(book)
    for kind, size in variants:
        main, expected = table[(kind, size)]
        assert main in defs
        assert oracle.run(book, heap_words=1 << 24, entry=main)["nf"] == expected


@pytest.mark.parametrize("n,seed", [(1, 7), (8, 7), (3, 11)])
def test_batch_mix_small(oracle, n, seed):
    book, nets, table = configs.batch_mix(n, seed)
    defs = configs.split_defs(book)
    for v in set(nets):
        main, expected = table[v]
        assert main in defs
        assert oracle.run(book, heap_words=1 << 24, entry=main)["nf"] == expected
