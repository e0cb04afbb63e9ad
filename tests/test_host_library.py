"""The engine library (libhvm64_amd.so) on the CPU: it loads, exports every
entry point of include/hvm64_amd.h, and its host side (parser, printer, label
sets, apply_tree) agrees with the oracle.  No device call is made here."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, load_snapshots

HEADER = os.path.join(ROOT, "include", "hvm64_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(hvm_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    import hvm64
    lib = ctypes.CDLL(hvm64.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(hvm64.EXPORTS)


SNAPS = load_snapshots()


@pytest.mark.parametrize("name", sorted(SNAPS))
def test_parse_print_agrees_with_oracle(oracle, name):
    import hvm64
    text = SNAPS[name]["text"]
    assert hvm64.book_print(text) == oracle.print_book(text)


@pytest.mark.parametrize("name", sorted(SNAPS))
def test_label_sets_agree_with_oracle(oracle, name):
    import hvm64
    text = SNAPS[name]["text"]
    assert hvm64.label_sets(text) == oracle.label_sets(text)


def test_label_sets_reference_vector():
    import hvm64
    from test_oracle_units import CALC_LABELS_BOOK, calc_labels_expected
    assert hvm64.label_sets(CALC_LABELS_BOOK) == calc_labels_expected()


def test_apply_tree_text():
    """ast.rs:346-357: `x & ...` becomes `y & ... & x ~ (arg y)`, y fresh."""
    import hvm64
    assert hvm64.apply_tree("(a a)", "(a a)") == "b\n  & (a a) ~ ((a a) b)"
    assert hvm64.apply_tree("@main", "64") == "b\n  & @main ~ (64 b)"


def test_parse_errors_are_reported():
    import hvm64
    for bad in ["@main = (a", "@main = a & b", "@ = *", "@main = [?]", "@main = 1e5"]:
        with pytest.raises(hvm64.HvmError) as e:
            hvm64.book_print(bad)
        assert e.value.kind == "PARSE"


def test_missing_reference_is_reported():
    import hvm64
    with pytest.raises(hvm64.HvmError) as e:
        hvm64.label_sets("@main = @nope")
    assert e.value.kind == "NOREF"
