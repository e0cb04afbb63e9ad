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
    """ast.rs:112-123: `x & ...` becomes `y & ... & x ~ (arg y)`, y fresh."""
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


@pytest.mark.parametrize("name", ["examples::sort::bitonic::bitonic_sort_lam", "tests::programs::f24",
                                  "examples::arithmetic"])
def test_book_image_round_trip(oracle, name):
    """hvm_book_encode writes every definition of the book (Host::insert_book,
    host.rs:41-73) into one image; the blob view of it has the definitions,
    label sets and instruction counts the text describes."""
    import hvm64
    text = SNAPS[name]["text"]
    img = hvm64.encode_book(text)
    assert img[:8] == b"HVM64BK1" and len(img) % 8 == 0
    blob, keep = hvm64.image_blobs(img)
    names = [blob.defs[i].name.decode() for i in range(blob.n_defs)]
    assert names == sorted(oracle.label_sets(text))
    labs = hvm64.label_sets(text)
    for i, n in enumerate(names):
        d = blob.defs[i]
        bits = [d.lab_bits[k] for k in range(d.n_lab_words)]
        got = [l for l in range(d.min_safe) if l // 64 < len(bits) and (bits[l // 64] >> (l % 64)) & 1]
        assert got == labs[n], n
        assert d.n_instr > 0 and d.n_slots > 0
    assert hvm64.encode_book(text) is img  # cached by the text's hash


def test_book_image_rejects_garbage():
    import hvm64
    lib = hvm64.load_library()
    assert lib.hvm_book_encode(b"@main = (", None, 0, None) == 1
