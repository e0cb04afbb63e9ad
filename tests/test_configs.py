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
