"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: the oracle is the checker, never the thing
measured or shipped."""
import ctypes
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")


class OracleError(RuntimeError):
    pass


class Oracle:
    def __init__(self, path=LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        self.lib = lib = ctypes.CDLL(path)
        lib.oracle_last_error.restype = ctypes.c_char_p
        u64 = ctypes.c_uint64
        lib.oracle_run.argtypes = [ctypes.c_char_p, ctypes.c_char_p, u64, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                   u64, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64),
                                   ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_double)]
        lib.oracle_reduce_expr.argtypes = [ctypes.c_char_p, ctypes.c_char_p, u64, ctypes.c_char_p, u64,
                                           ctypes.POINTER(u64), ctypes.POINTER(u64)]
        lib.oracle_apply_tree.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, u64,
                                          ctypes.c_char_p, u64, ctypes.POINTER(u64)]
        lib.oracle_print_book.argtypes = [ctypes.c_char_p, ctypes.c_char_p, u64, ctypes.POINTER(u64)]
        lib.oracle_label_sets.argtypes = [ctypes.c_char_p, ctypes.c_char_p, u64, ctypes.POINTER(u64)]
        lib.oracle_operate_binary.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        lib.oracle_operate_binary.restype = ctypes.c_uint32
        lib.oracle_operate_unary.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        lib.oracle_operate_unary.restype = ctypes.c_uint32
        lib.oracle_new_f24.argtypes = [ctypes.c_float]
        lib.oracle_new_f24.restype = ctypes.c_uint32
        lib.oracle_num_display.argtypes = [ctypes.c_uint32, ctypes.c_char_p, u64]
        lib.oracle_parse_num.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32)]
        lib.oracle_create_var.argtypes = [u64, ctypes.c_char_p, u64]
        lib.oracle_compile_book.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, u64, ctypes.POINTER(u64)]
        lib.oracle_pre_reduce.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_uint32,
                                          ctypes.c_int, u64, u64, ctypes.c_char_p, u64, ctypes.POINTER(u64),
                                          ctypes.POINTER(u64)]

    def _err(self, rc):
        raise OracleError(f"oracle error {rc}: {self.lib.oracle_last_error().decode()}")

    def run(self, book, entry="main", heap_words=1 << 27, threads=0, meter=False):
        """Host::new + boot(entry) + normal() (threads=0) or parallel_normal().

        Returns dict(nf, rwts, shortcuts, peak_nodes, meter, seconds)."""
        cap = 1 << 20
        while True:
            buf = ctypes.create_string_buffer(cap)
            n = ctypes.c_uint64()
            c = (ctypes.c_uint64 * 5)()
            e = (ctypes.c_uint64 * 4)()
            m = (ctypes.c_uint64 * 3)()
            s = ctypes.c_double()
            rc = self.lib.oracle_run(book.encode(), entry.encode(), heap_words, threads, int(meter), buf, cap,
                                     ctypes.byref(n), c, e, m, ctypes.byref(s))
            if rc == 7 and n.value + 1 > cap:
                cap = n.value + 1
                continue
            if rc:
                self._err(rc)
            rw = dict(anni=c[0], comm=c[1], eras=c[2], dref=c[3], oper=c[4])
            rw["rwts"] = sum(c)
            return dict(nf=buf.value.decode(), rwts=rw, shortcuts=tuple(e[:3]), peak_nodes=e[3],
                        meter=dict(loads=m[0], stores=m[1], rmws=m[2]), seconds=s.value)

    def reduce_expr(self, book, expr, heap_words=1 << 24):
        buf = ctypes.create_string_buffer(1 << 20)
        n = ctypes.c_uint64()
        c = (ctypes.c_uint64 * 5)()
        rc = self.lib.oracle_reduce_expr(book.encode(), expr.encode(), heap_words, buf, 1 << 20, ctypes.byref(n), c)
        if rc:
            self._err(rc)
        return buf.value.decode(), list(c)

    def apply_tree_normal(self, fun, args, heap_words=16):
        arr = (ctypes.c_char_p * max(len(args), 1))(*[a.encode() for a in args])
        buf = ctypes.create_string_buffer(1 << 16)
        n = ctypes.c_uint64()
        rc = self.lib.oracle_apply_tree(fun.encode(), arr, len(args), heap_words, buf, 1 << 16, ctypes.byref(n))
        if rc:
            self._err(rc)
        return buf.value.decode()

    def _text(self, fn, *args):
        buf = ctypes.create_string_buffer(1 << 22)
        n = ctypes.c_uint64()
        rc = fn(*[a.encode() for a in args], buf, 1 << 22, ctypes.byref(n))
        if rc:
            self._err(rc)
        return buf.value.decode()

    def pre_reduce(self, book, skip=(), only=None, max_rwts=(1 << 64) - 1, heap_words=1 << 26):
        """Book::pre_reduce (transform/src/pre_reduce.rs:38-75): reduces every
        definition except those in `skip` (or, with `only`, just those).
        Returns (book_text, rwts dict)."""
        names = list(only if only is not None else skip)
        arr = (ctypes.c_char_p * max(len(names), 1))(*[x.encode() for x in names])
        cap = 1 << 22
        while True:
            buf = ctypes.create_string_buffer(cap)
            n = ctypes.c_uint64()
            c = (ctypes.c_uint64 * 5)()
            rc = self.lib.oracle_pre_reduce(book.encode(), arr, len(names), int(only is not None), max_rwts, heap_words,
                                            buf, cap, ctypes.byref(n), c)
            if rc == 7 and n.value + 1 > cap:
                cap = n.value + 1
                continue
            if rc:
                self._err(rc)
            rw = dict(anni=c[0], comm=c[1], eras=c[2], dref=c[3], oper=c[4])
            rw["rwts"] = sum(c)
            return buf.value.decode(), rw

    def print_book(self, book):
        return self._text(self.lib.oracle_print_book, book)

    def label_sets(self, book):
        out = {}
        for line in self._text(self.lib.oracle_label_sets, book).splitlines():
            name, _, labs = line.partition(":")
            out[name] = [int(x) for x in labs.split()]
        return out

    def operate_binary(self, a, op, b):
        return self.lib.oracle_operate_binary(a, op, b)

    def operate_unary(self, op, a):
        return self.lib.oracle_operate_unary(op, a)

    def new_f24(self, f):
        return self.lib.oracle_new_f24(f)

    def num_display(self, raw):
        buf = ctypes.create_string_buffer(64)
        self.lib.oracle_num_display(raw, buf, 64)
        return buf.value.decode()

    def parse_num(self, text):
        raw = ctypes.c_uint32()
        rc = self.lib.oracle_parse_num(text.encode(), ctypes.byref(raw))
        if rc:
            self._err(rc)
        return raw.value

    def create_var(self, i):
        buf = ctypes.create_string_buffer(64)
        self.lib.oracle_create_var(i, buf, 64)
        return buf.value.decode()

    def compile_book(self, book, workdir, opt="-O2"):
        """The book's definitions compiled to C++ (src/compile.rs:110-161,
        with do_switch), built with g++ into `workdir`; returns the path of
        the executable (FILE [-t THREADS] [-e ENTRY] [-m WORDS] [-r RUNS])."""
        import subprocess
        src_path = os.path.join(ROOT, "oracle", "hvm64_oracle.cpp")
        text = None
        cap = 1 << 22
        while text is None:
            buf = ctypes.create_string_buffer(cap)
            n = ctypes.c_uint64()
            rc = self.lib.oracle_compile_book(book.encode(), src_path.encode(), buf, cap, ctypes.byref(n))
            if rc == 7 and n.value + 1 > cap:
                cap = n.value + 1
                continue
            if rc:
                self._err(rc)
            text = buf.value.decode()
        os.makedirs(workdir, exist_ok=True)
        cpp = os.path.join(workdir, "compiled_book.cpp")
        exe = os.path.join(workdir, "compiled_book")
        with open(cpp, "w") as f:
            f.write(text)
        subprocess.run(["g++", opt, "-std=c++20", "-pthread", "-w", cpp, "-o", exe], check=True)
        return exe

    @staticmethod
    def run_compiled(exe, book, threads=0, runs=1, entry="main", heap_words=1 << 28):
        """Runs a compiled book; returns (normal form, [(seconds, counts)])."""
        import subprocess
        import tempfile
        with tempfile.NamedTemporaryFile("w", suffix=".hvm", delete=False) as f:
            f.write(book)
            path = f.name
        try:
            r = subprocess.run([exe, path, "-t", str(threads), "-r", str(runs), "-e", entry, "-m", str(heap_words)],
                               capture_output=True, text=True, check=True)
        finally:
            os.unlink(path)
        lines = r.stdout.split("\n")
        out, nf, k = [], None, 0
        for _ in range(runs):
            parts = lines[k].split()
            c = [int(x) for x in parts[1:]]
            out.append((float(parts[0]), dict(anni=c[0], comm=c[1], eras=c[2], dref=c[3], oper=c[4], rwts=sum(c))))
            k += 1
            if nf is None:
                nf = lines[k]
                k += 1
        return nf, out
