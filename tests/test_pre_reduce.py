"""Book::pre_reduce (transform/src/pre_reduce.rs:38-164): the oracle's
restatement against the reference's own vectors, on the CPU.

  * every `pre_reduce_run@*` snapshot (tests/tests.rs:44-55): pre-reducing
    every definition but main gives the snapshot's pre-reduce counts, and the
    pre-reduced book then runs main to the snapshot's normal form and run
    counts (the runs up to 2M rewrites here; the GPU tests run the rest);
  * test_fast_pre_reduce (tests/transform.rs:14-46): with black_box removed
    (an inert definition that captures the redexes it meets) and only two
    definitions pre-reduced, pre-reduce plus the run of the entry point give
    the Fast and Slow count sets exactly.  Reducing `expensive` once for both
    of its uses (Fast) against two copies (Slow) is what separates them, so a
    driver that re-reduced a definition already seen, or skipped the wrong
    ones, fails here."""
import json
import os
import re

import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = json.load(open(os.path.join(GOLDEN, "pre_reduce_snapshots.json")))
SNAPS = FIX["snapshots"]
RUN_LIMIT = 2_000_000


def _text(rel):
    with open(os.path.join(GOLDEN, rel)) as f:
        return f.read()


def split_out(text, name):
    """(book without @name, @name's definition text)."""
    defs = re.split(r"\n(?=@)", text.strip())
    keep = [d for d in defs if not d.startswith("@" + name + " ")]
    gone = [d for d in defs if d.startswith("@" + name + " ")]
    assert len(gone) == 1
    return "\n".join(keep) + "\n", gone[0] + "\n"


@pytest.mark.parametrize("key", sorted(SNAPS))
def test_oracle_pre_reduce_counts(oracle, key):
    v = SNAPS[key]
    book, rw = oracle.pre_reduce(_text(v["program"]), skip=["main"])
    assert rw == v["pre_rwts"]
    assert "@main = " in book


@pytest.mark.parametrize("key", sorted(k for k in SNAPS if SNAPS[k]["run_rwts"]["rwts"] <= RUN_LIMIT))
def test_oracle_pre_reduced_book_runs_to_snapshot(oracle, key):
    v = SNAPS[key]
    book, _ = oracle.pre_reduce(_text(v["program"]), skip=["main"])
    r = oracle.run(book, heap_words=1 << 26)
    assert r["nf"] == v["normal_form"]
    assert r["rwts"] == v["run_rwts"]


@pytest.mark.parametrize("mode", ["fast", "slow"])
def test_oracle_fast_pre_reduce(oracle, mode):
    fp = FIX["fast_pre_reduce"]
    m = fp[mode]
    rest, removed = split_out(_text(fp["program"]), fp["removed"])
    book, rw = oracle.pre_reduce(rest, only=m["reduce"])
    assert not re.search(r"^@" + fp["removed"] + " = ", book, re.M)  # an inert definition is not written out
    assert "@" + fp["removed"] in book  # the redexes it captured are
    r = oracle.run(book + "\n" + removed, entry=m["entry"], heap_words=1 << 26)
    assert {k: rw[k] + r["rwts"][k] for k in rw} == m["rwts"]


def test_oracle_pre_reduce_budget_leaves_redexes(oracle):
    """A budget of one interaction per definition (net.rs:41-52): each
    definition is booted (one dereference) and interacts once, keeps its
    other redexes (read back into its net, readback.rs:28-30), and the book
    still runs main to the same normal form."""
    text = _text(SNAPS["examples::sort::bitonic::bitonic_sort_lam"]["program"])
    book, rw = oracle.pre_reduce(text, skip=["main"], max_rwts=1)
    n_defs = len(re.findall(r"^@", text, re.M)) - 1
    assert rw["dref"] >= n_defs and rw["rwts"] <= 2 * n_defs
    assert "&" in book
    full = oracle.run(text, heap_words=1 << 26)
    part = oracle.run(book, heap_words=1 << 26)
    assert part["nf"] == full["nf"]
