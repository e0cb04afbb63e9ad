"""The oracle reproduces the reference's run@ snapshots (tests/tests.rs:28-42):
normal form and all five rewrite counts, for every golden program.

Programs above 4e7 rewrites take minutes single-threaded; set HVM64_FULL=1 to
include them."""
import os

import pytest

from conftest import load_snapshots

SNAPS = load_snapshots()
LIMIT = 2_000_000_000 if os.environ.get("HVM64_FULL") else 40_000_000
CASES = sorted((k for k, v in SNAPS.items() if v["rwts"]["rwts"] <= LIMIT), key=lambda k: SNAPS[k]["rwts"]["rwts"])


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_snapshot(oracle, name):
    snap = SNAPS[name]
    r = oracle.run(snap["text"], heap_words=1 << 26)
    assert r["nf"] == snap["normal_form"]
    assert r["rwts"] == snap["rwts"]


@pytest.mark.parametrize("name", ["examples::sort::bitonic::bitonic_sort_lam", "tests::programs::list_put_got",
                                  "examples::lambda_calculus::hoas", "tests::programs::f24"])
def test_parallel_baseline_matches_snapshot(oracle, name):
    """parallel_normal (runtime/src/parallel.rs) with 4 threads gives the same
    normal form and counts: the snapshots themselves came from it."""
    snap = SNAPS[name]
    r = oracle.run(snap["text"], heap_words=1 << 26, threads=4)
    assert r["nf"] == snap["normal_form"]
    assert r["rwts"] == snap["rwts"]


def test_no_shortcut_sensitivity(oracle):
    """The interpreter's ERA shortcuts (instruction.rs:144-146, 163-165,
    194-197) never fire on the golden programs, so their counts do not depend
    on reduction order."""
    for name in CASES:
        r = oracle.run(SNAPS[name]["text"], heap_words=1 << 26)
        assert r["shortcuts"] == (0, 0, 0), name
