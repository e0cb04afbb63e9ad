"""The CPU baseline of the reference's compiled path (src/compile.rs:110-161):
a book's definitions written out as C++ functions that call the
interpreter's do_* methods in instruction order, built with g++.  Its
normal forms and five counts equal the interpreter's (the oracle) on the
benchmark programs, the numeric chains and the books whose ERA shortcuts
fire (instruction.rs:134-203)."""
import os

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("compiled"))


def books():
    import test_gpu_parity as tg
    from hvm64 import configs
    out = [("tree_sum", configs.tree_sum(8)), ("bitonic", configs.bitonic(7, 3)),
           ("church", configs.church(6, 16)), ("numeric", tg.numeric_book(4, 40))]
    out += [(f"shortcut{i}", t) for i, (t, _) in enumerate(tg.SHORTCUT_BOOKS)]
    return out


@pytest.mark.parametrize("name,book", books())
def test_compiled_book_matches_interpreter(oracle, workdir, name, book):
    exe = oracle.compile_book(book, os.path.join(workdir, name), opt="-O0")
    ref = oracle.run(book, heap_words=1 << 22)
    for threads in (0, 2):  # normal(), parallel_normal()
        nf, runs = oracle.run_compiled(exe, book, threads=threads, heap_words=1 << 22)
        assert nf == ref["nf"]
        assert runs[0][1] == ref["rwts"]
