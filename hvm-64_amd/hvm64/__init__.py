"""Python mirror of the hvm-64 host/runtime API over the engine's C ABI.

The reference's Rust API (HigherOrderCO/hvm-64 v0.3.0):

    let heap = Heap::new(bytes)                 runtime/src/allocator.rs:22
    let host = Host::new(&book)                 host/src/host.rs:25
    let mut net = Net::new(&heap)               runtime/src/net.rs:19
    net.boot(&host.defs["main"])                runtime/src/net.rs:30
    net.normal()                                runtime/src/net.rs:55
    net.reduce(limit) -> Option<usize>          runtime/src/net.rs:41
    net.rwts                                    runtime/src/runtime.rs:72
    host.readback(&net)                         host/src/readback.rs:23

maps here to the same names.  Every call runs on the GPU through
libhvm64_amd.so; there is no CPU fallback: if the library or the device is
missing, the constructors raise.
"""
import ctypes
import os
from dataclasses import dataclass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libhvm64_amd.so")

HVM_E = {
    1: "PARSE", 2: "OOM", 3: "UNIMPL_PAIR", 4: "DIV0", 5: "UNBOUND", 6: "NOREF",
    7: "QUEUE", 8: "READBACK", 9: "TRUNC", 10: "ARG", 11: "DEVICE", 12: "TIMEOUT",
}

EXPORTS = [
    "hvm_gpu_open", "hvm_gpu_close", "hvm_gpu_last_error", "hvm_gpu_load_book_text", "hvm_gpu_load_book",
    "hvm_gpu_def_id", "hvm_gpu_reset", "hvm_gpu_boot", "hvm_gpu_encode_net_text", "hvm_gpu_normal",
    "hvm_gpu_reduce", "hvm_gpu_rewrites", "hvm_gpu_shortcuts", "hvm_gpu_readback", "hvm_gpu_last_kernel_ms",
    "hvm_gpu_grid", "hvm_gpu_set_timeout", "hvm_gpu_sched_stats", "hvm_gpu_open_grid",
    "hvm_gpu_boot_batch", "hvm_gpu_normal_batch", "hvm_gpu_root_ports", "hvm_gpu_readback_root",
    "hvm_gpu_boot_batch_noalloc", "hvm_gpu_batch_plan", "hvm_gpu_batch_run", "hvm_gpu_batch_tune",
    "hvm_gpu_batch_info", "hvm_gpu_batch_used", "hvm_gpu_alloc_count",
    "hvm_gpu_serve_start", "hvm_gpu_serve_submit", "hvm_gpu_serve_poll", "hvm_gpu_serve_stop",
    "hvm_gpu_snapshot", "hvm_gpu_load_book_image", "hvm_gpu_set_compiled", "hvm_gpu_pre_reduce",
    "hvm_gpu_set_sched_stats",
    "hvm_book_print", "hvm_book_label_sets", "hvm_net_apply_tree", "hvm_book_encode",
]

_lib = None


class NetImage(ctypes.Structure):
    """hvm_net_image: root var cell, port address of heap word 0, heap words."""
    _fields_ = [("root_cell", ctypes.c_uint64), ("heap_base", ctypes.c_uint64), ("heap_words", ctypes.c_uint64)]


class Instr(ctypes.Structure):
    """hvm_instr: one device instruction (two words)."""
    _fields_ = [("w0", ctypes.c_uint64), ("w1", ctypes.c_uint64)]


class DefBlob(ctypes.Structure):
    """hvm_def_blob: one encoded definition."""
    _fields_ = [("name", ctypes.c_char_p), ("min_safe", ctypes.c_uint32), ("n_lab_words", ctypes.c_uint32),
                ("lab_bits", ctypes.POINTER(ctypes.c_uint64)), ("n_instr", ctypes.c_uint32),
                ("instr", ctypes.POINTER(Instr)), ("n_slots", ctypes.c_uint32)]


class BookBlob(ctypes.Structure):
    """hvm_book_blob: an encoded book."""
    _fields_ = [("n_defs", ctypes.c_uint32), ("defs", ctypes.POINTER(DefBlob))]


class BatchTuning(ctypes.Structure):
    """hvm_batch_tuning: launch parameters of a planned batch."""
    _fields_ = [("blocks", ctypes.c_uint32), ("mbox_keep", ctypes.c_uint32), ("fill_below", ctypes.c_uint32),
                ("arena_chunk", ctypes.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class HvmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{HVM_E.get(code, code)}: {msg}")
        self.code = code
        self.kind = HVM_E.get(code, str(code))


def load_library(path=None):
    """Loads libhvm64_amd.so (raises if it has not been built).  The
    HVM64_LIB environment variable names another build of the library."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("HVM64_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with hvm-64_amd/build.py (or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    u64 = ctypes.c_uint64
    lib.hvm_gpu_open.argtypes = [ctypes.c_int, u64, ctypes.POINTER(P)]
    lib.hvm_gpu_close.argtypes = [P]
    lib.hvm_gpu_close.restype = None
    lib.hvm_gpu_last_error.argtypes = [P]
    lib.hvm_gpu_last_error.restype = ctypes.c_char_p
    lib.hvm_gpu_load_book_text.argtypes = [P, ctypes.c_char_p]
    lib.hvm_gpu_def_id.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32)]
    lib.hvm_gpu_reset.argtypes = [P]
    lib.hvm_gpu_boot.argtypes = [P, ctypes.c_uint32]
    lib.hvm_gpu_encode_net_text.argtypes = [P, ctypes.c_char_p]
    lib.hvm_gpu_normal.argtypes = [P]
    lib.hvm_gpu_reduce.argtypes = [P, u64, ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_int)]
    lib.hvm_gpu_rewrites.argtypes = [P, ctypes.POINTER(u64)]
    lib.hvm_gpu_shortcuts.argtypes = [P, ctypes.POINTER(u64)]
    lib.hvm_gpu_readback.argtypes = [P, ctypes.c_char_p, u64, ctypes.POINTER(u64)]
    lib.hvm_gpu_last_kernel_ms.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    lib.hvm_gpu_grid.argtypes = [P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    lib.hvm_gpu_set_timeout.argtypes = [P, ctypes.c_double]
    lib.hvm_gpu_sched_stats.argtypes = [P, ctypes.POINTER(u64)]
    lib.hvm_gpu_open_grid.argtypes = [ctypes.c_int, u64, ctypes.c_uint32, ctypes.POINTER(P)]
    lib.hvm_gpu_boot_batch.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(u64)]
    lib.hvm_gpu_boot_batch_noalloc.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(u64)]
    lib.hvm_gpu_normal_batch.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(u64)]
    lib.hvm_gpu_batch_plan.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(u64),
                                       ctypes.POINTER(BatchTuning)]
    lib.hvm_gpu_batch_run.argtypes = [P, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    lib.hvm_gpu_batch_tune.argtypes = [P, ctypes.POINTER(BatchTuning)]
    lib.hvm_gpu_batch_info.argtypes = [P, ctypes.POINTER(BatchTuning), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(u64)]
    lib.hvm_gpu_alloc_count.argtypes = [P, ctypes.POINTER(u64)]
    lib.hvm_gpu_set_sched_stats.argtypes = [P, ctypes.c_int]
    lib.hvm_gpu_pre_reduce.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_uint32,
                                       ctypes.c_int, u64, ctypes.c_char_p, u64, ctypes.POINTER(u64),
                                       ctypes.POINTER(u64)]
    lib.hvm_gpu_batch_used.argtypes = [P, ctypes.POINTER(u64)]
    lib.hvm_gpu_serve_start.argtypes = [P, ctypes.c_uint32, u64, u64]
    lib.hvm_gpu_serve_submit.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(u64)]
    lib.hvm_gpu_serve_poll.argtypes = [P, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_uint32)]
    lib.hvm_gpu_serve_stop.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    lib.hvm_gpu_root_ports.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    lib.hvm_gpu_readback_root.argtypes = [P, u64, ctypes.c_char_p, u64, ctypes.POINTER(u64)]
    lib.hvm_gpu_snapshot.argtypes = [P, ctypes.POINTER(NetImage), u64, u64, ctypes.POINTER(u64)]
    lib.hvm_gpu_load_book.argtypes = [P, ctypes.POINTER(BookBlob)]
    lib.hvm_gpu_load_book_image.argtypes = [P, ctypes.c_char_p, u64]
    lib.hvm_gpu_set_compiled.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
    lib.hvm_book_encode.argtypes = [ctypes.c_char_p, ctypes.c_char_p, u64, ctypes.POINTER(u64)]
    for name in ("hvm_book_print", "hvm_book_label_sets"):
        getattr(lib, name).argtypes = [ctypes.c_char_p, ctypes.c_char_p, u64, ctypes.POINTER(u64)]
    lib.hvm_net_apply_tree.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, u64, ctypes.POINTER(u64)]
    _lib = lib
    return lib


def pretty_num(n):
    """util/src/pretty_num.rs: 1234567 -> '1_234_567'."""
    s = str(n)
    parts = []
    while len(s) > 3:
        parts.append(s[-3:])
        s = s[:-3]
    parts.append(s)
    return "_".join(reversed(parts))


@dataclass
class Rewrites:
    """runtime/src/runtime.rs:72-141."""
    anni: int = 0
    comm: int = 0
    eras: int = 0
    dref: int = 0
    oper: int = 0

    def total(self):
        return self.anni + self.comm + self.eras + self.dref + self.oper

    def as_dict(self):
        return {"rwts": self.total(), "anni": self.anni, "comm": self.comm, "eras": self.eras,
                "dref": self.dref, "oper": self.oper}

    def __str__(self):
        rows = [("RWTS   ", self.total()), ("- ANNI ", self.anni), ("- COMM ", self.comm),
                ("- ERAS ", self.eras), ("- DREF ", self.dref), ("- OPER ", self.oper)]
        return "".join(f"{k}: {pretty_num(v):>15}\n" for k, v in rows)


class Heap:
    """Device heap + engine handle (Heap::new, runtime/src/allocator.rs:22)."""

    def __init__(self, nbytes=1 << 30, device=0, max_blocks=0):
        """max_blocks caps the reduction grid (0: the whole device), so that
        several heaps can reduce concurrently from separate host threads."""
        lib = load_library()
        h = ctypes.c_void_p()
        rc = lib.hvm_gpu_open_grid(device, nbytes, max_blocks, ctypes.byref(h))
        if rc != 0:
            raise HvmError(rc, lib.hvm_gpu_last_error(None).decode())
        self._lib = lib
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise HvmError(rc, self._lib.hvm_gpu_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.hvm_gpu_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timeout(self, seconds):
        """Wall-clock limit of one reduction launch (HVM_E_TIMEOUT past it)."""
        self._check(self._lib.hvm_gpu_set_timeout(self._h, float(seconds)))

    def set_compiled(self, on=True):
        """Compiled definitions on or off (default on); returns how many
        definitions of the loaded book have a compiled form."""
        n = ctypes.c_uint32()
        self._check(self._lib.hvm_gpu_set_compiled(self._h, 1 if on else 0, ctypes.byref(n)))
        return n.value

    def set_sched_stats(self, on=True):
        """Scheduler statistics for single-net launches (off by default)."""
        self._check(self._lib.hvm_gpu_set_sched_stats(self._h, 1 if on else 0))

    def grid(self):
        b, w = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self._lib.hvm_gpu_grid(self._h, ctypes.byref(b), ctypes.byref(w)))
        return b.value, w.value


class DefRef:
    def __init__(self, name, idx):
        self.name, self.idx = name, idx

    def __repr__(self):
        return f"@{self.name}"


class _Defs:
    def __init__(self, host):
        self._host = host

    def __getitem__(self, name):
        h = self._host.heap
        i = ctypes.c_uint32()
        h._check(h._lib.hvm_gpu_def_id(h._h, name.encode(), ctypes.byref(i)))
        return DefRef(name, i.value)

    def get(self, name):
        try:
            return self[name]
        except HvmError:
            return None


_IMAGES = {}


def encode_book(book_text):
    """Book::from_str + Host::insert_book into a book image (bytes), the
    flat encoded form hvm_gpu_load_book_image takes.  Images are cached by
    the book text's hash for the life of the process."""
    import hashlib
    key = hashlib.sha256(book_text.encode()).hexdigest()
    img = _IMAGES.get(key)
    if img is not None:
        return img
    lib = load_library()
    n = ctypes.c_uint64()
    rc = lib.hvm_book_encode(book_text.encode(), None, 0, ctypes.byref(n))
    if rc not in (0, 9):
        raise HvmError(rc, lib.hvm_gpu_last_error(None).decode())
    buf = ctypes.create_string_buffer(n.value)
    rc = lib.hvm_book_encode(book_text.encode(), buf, n.value, ctypes.byref(n))
    if rc != 0:
        raise HvmError(rc, lib.hvm_gpu_last_error(None).decode())
    img = buf.raw[:n.value]
    _IMAGES[key] = img
    return img


def image_blobs(image):
    """The hvm_book_blob structures describing a book image: what a host
    that encodes books itself (the Rust Host's Def<InterpretedDef> list)
    passes to hvm_gpu_load_book.  Returns (BookBlob, keepalive)."""
    import struct
    words = struct.unpack(f"<{len(image) // 8}Q", image)
    assert words[0] == 0x314B4234364D5648, "not a book image"
    n, pos = words[1], 2
    defs = (DefBlob * max(n, 1))()
    keep = [defs]
    for i in range(n):
        name_len, n_instr = words[pos] & 0xFFFFFFFF, words[pos] >> 32
        n_slots, min_safe = words[pos + 1] & 0xFFFFFFFF, words[pos + 1] >> 32
        n_lab = words[pos + 2]
        pos += 3
        name = image[pos * 8:pos * 8 + name_len]
        pos += (name_len + 7) // 8
        labs = (ctypes.c_uint64 * max(n_lab, 1))(*words[pos:pos + n_lab])
        pos += n_lab
        ins = (Instr * max(n_instr, 1))(*[Instr(words[pos + 2 * k], words[pos + 2 * k + 1]) for k in range(n_instr)])
        pos += 2 * n_instr
        keep += [name, labs, ins]
        defs[i] = DefBlob(name, min_safe, n_lab, labs, n_instr, ins, n_slots)
    return BookBlob(n, defs), keep


class Host:
    """Host::new(&book) (host/src/host.rs:25): loads a book into the device.

    `via`: "text" parses and encodes the book on every load, "image" loads
    the cached book image (encode_book), "blob" passes the image's
    definitions as hvm_def_blob structures (hvm_gpu_load_book)."""

    def __init__(self, book_text, heap, via="text"):
        self.heap = heap
        if via == "text":
            heap._check(heap._lib.hvm_gpu_load_book_text(heap._h, book_text.encode()))
        elif via == "image":
            img = encode_book(book_text)
            heap._check(heap._lib.hvm_gpu_load_book_image(heap._h, img, len(img)))
        elif via == "blob":
            blob, keep = image_blobs(encode_book(book_text))
            heap._check(heap._lib.hvm_gpu_load_book(heap._h, ctypes.byref(blob)))
            del keep
        else:
            raise ValueError(via)
        self.defs = _Defs(self)

    def readback(self, net):
        """Host::readback + Display (host/src/readback.rs:23-33): normal form text."""
        h = self.heap
        n = ctypes.c_uint64()
        cap = 1 << 16
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = h._lib.hvm_gpu_readback(h._h, buf, cap, ctypes.byref(n))
            if rc == 9 and n.value + 1 > cap:
                cap = n.value + 1
                continue
            h._check(rc)
            return buf.value.decode()

    def readback_root(self, root):
        """Host::readback of one net of a Batch (its root cell)."""
        h = self.heap
        n = ctypes.c_uint64()
        cap = 1 << 12
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = h._lib.hvm_gpu_readback_root(h._h, root, buf, cap, ctypes.byref(n))
            if rc == 9 and n.value + 1 > cap:
                cap = n.value + 1
                continue
            h._check(rc)
            return buf.value.decode()

    def encode_net(self, net, net_text):
        """Host::encode_net(net, Var(root), &ast) (host/src/encode.rs:31-35)."""
        h = self.heap
        h._check(h._lib.hvm_gpu_encode_net_text(h._h, net_text.encode()))


class Net:
    """Net::new(&heap) (runtime/src/net.rs:19): a fresh net on the heap."""

    def __init__(self, heap):
        self.heap = heap
        heap._check(heap._lib.hvm_gpu_reset(heap._h))

    def boot(self, d):
        self.heap._check(self.heap._lib.hvm_gpu_boot(self.heap._h, d.idx))

    def normal(self):
        self.heap._check(self.heap._lib.hvm_gpu_normal(self.heap._h))

    def reduce(self, limit):
        """Returns the number of redexes popped if the net is now normal, else
        None; `last_popped` holds the count in both cases."""
        popped = ctypes.c_uint64()
        normal = ctypes.c_int()
        h = self.heap
        h._check(h._lib.hvm_gpu_reduce(h._h, limit, ctypes.byref(popped), ctypes.byref(normal)))
        self.last_popped = popped.value
        return popped.value if normal.value else None

    @property
    def rwts(self):
        out = (ctypes.c_uint64 * 5)()
        self.heap._check(self.heap._lib.hvm_gpu_rewrites(self.heap._h, out))
        return Rewrites(*[int(x) for x in out])

    @property
    def shortcuts(self):
        out = (ctypes.c_uint64 * 3)()
        self.heap._check(self.heap._lib.hvm_gpu_shortcuts(self.heap._h, out))
        return tuple(int(x) for x in out)

    def sched_stats(self):
        """Scheduler statistics of the last reduce launch (summed over waves)."""
        out = (ctypes.c_uint64 * 20)()
        self.heap._check(self.heap._lib.hvm_gpu_sched_stats(self.heap._h, out))
        keys = ["busy_iters", "idle_iters", "busy_ticks", "spills", "spilled", "claims", "claimed",
                "interact_ticks", "spill_ticks", "alloc_miss", "ov_alloc",
                "anni_iters", "comm_iters", "call_iters", "other_iters",
                "anni_ticks", "comm_ticks", "call_ticks", "other_ticks", "pops"]
        return dict(zip(keys, [int(x) for x in out]))

    def snapshot(self, first=0, n=0):
        """(NetImage, words): the raw device image, with heap words
        [first, first + n) copied to the host as a list of ints."""
        img = NetImage()
        buf = (ctypes.c_uint64 * n)() if n else None
        self.heap._check(self.heap._lib.hvm_gpu_snapshot(self.heap._h, ctypes.byref(img), first, n, buf))
        return img, (list(buf) if n else [])

    def last_kernel_ms(self):
        ms = ctypes.c_double()
        self.heap._check(self.heap._lib.hvm_gpu_last_kernel_ms(self.heap._h, ctypes.byref(ms)))
        return ms.value


class Batch(Net):
    """Independent nets reduced together in one launch: each net i behaves as
    Net::new + Net::boot(defs[i]) + Net::normal() would on its own
    (runtime/src/net.rs:19-60); rwts are the sums over the batch."""

    def __init__(self, heap):
        super().__init__(heap)
        self.roots = []

    def boot(self, defs, allocation_free=False):
        """Boots one net per entry of `defs` (Host.defs items or indices).
        allocation_free: through hvm_gpu_boot_batch_noalloc (at most 65,536
        nets, staged through a static device array)."""
        n = len(defs)
        ids = (ctypes.c_uint32 * n)(*[d.idx if hasattr(d, "idx") else int(d) for d in defs])
        roots = (ctypes.c_uint64 * n)()
        fn = self.heap._lib.hvm_gpu_boot_batch_noalloc if allocation_free else self.heap._lib.hvm_gpu_boot_batch
        self.heap._check(fn(self.heap._h, n, ids, roots))
        self.roots = list(roots)

    def normal(self):
        n = len(self.roots)
        roots = (ctypes.c_uint64 * n)(*self.roots)
        self.heap._check(self.heap._lib.hvm_gpu_normal_batch(self.heap._h, n, roots))
        self.roots = list(roots)

    def ports(self):
        """The port each net's root leads to (a Num port for numeric results)."""
        n = len(self.roots)
        roots = (ctypes.c_uint64 * n)(*self.roots)
        out = (ctypes.c_uint64 * n)()
        self.heap._check(self.heap._lib.hvm_gpu_root_ports(self.heap._h, n, roots, out))
        return list(out)


class PlannedBatch:
    """A fixed list of nets (Net::new + Net::boot(defs[i]) + Net::normal()
    each, runtime/src/net.rs:19-60) planned once and run again and again:
    every net allocates from an arena of its own, and one run is one
    replay of a HIP graph (reset, boot, expand, reduce, root ports)."""

    def __init__(self, heap, defs, arena_nodes, tuning=None):
        self.heap = heap
        n = len(defs)
        self.n = n
        ids = (ctypes.c_uint32 * n)(*[d.idx if hasattr(d, "idx") else int(d) for d in defs])
        nodes = (ctypes.c_uint64 * n)(*[int(x) for x in arena_nodes])
        t = None
        if tuning is not None:
            t = BatchTuning(**tuning) if isinstance(tuning, dict) else tuning
        heap._check(heap._lib.hvm_gpu_batch_plan(heap._h, n, ids, nodes, ctypes.byref(t) if t else None))
        self.roots = []

    def run(self):
        """Runs the batch once; returns the port each net's root leads to."""
        n = self.n
        ports = (ctypes.c_uint64 * n)()
        roots = (ctypes.c_uint64 * n)()
        self.heap._check(self.heap._lib.hvm_gpu_batch_run(self.heap._h, ports, roots))
        self.roots = list(roots)
        return list(ports)

    def used(self):
        """Per net, the nodes the last run took from its arena (with the
        extents it took from the pool)."""
        out = (ctypes.c_uint64 * self.n)()
        self.heap._check(self.heap._lib.hvm_gpu_batch_used(self.heap._h, out))
        return list(out)

    def tune(self):
        best = BatchTuning()
        self.heap._check(self.heap._lib.hvm_gpu_batch_tune(self.heap._h, ctypes.byref(best)))
        return best.as_dict()

    def info(self):
        t = BatchTuning()
        graph = ctypes.c_int()
        pool = ctypes.c_uint64()
        self.heap._check(self.heap._lib.hvm_gpu_batch_info(self.heap._h, ctypes.byref(t), ctypes.byref(graph),
                                                           ctypes.byref(pool)))
        return {"tuning": t.as_dict(), "graph": bool(graph.value), "pool_nodes": int(pool.value)}

    @property
    def rwts(self):
        out = (ctypes.c_uint64 * 5)()
        self.heap._check(self.heap._lib.hvm_gpu_rewrites(self.heap._h, out))
        return Rewrites(*[int(x) for x in out])

    @property
    def shortcuts(self):
        out = (ctypes.c_uint64 * 3)()
        self.heap._check(self.heap._lib.hvm_gpu_shortcuts(self.heap._h, out))
        return tuple(int(x) for x in out)

    def last_kernel_ms(self):
        ms = ctypes.c_double()
        self.heap._check(self.heap._lib.hvm_gpu_last_kernel_ms(self.heap._h, ctypes.byref(ms)))
        return ms.value


class Serving:
    """A serving launch on `heap` (hvm_gpu_serve_*): the running reduction
    kernel admits submitted requests into `slots` arenas of `slot_nodes`
    nodes and reports each request's root port when its net is normal."""

    def __init__(self, heap, slots, slot_nodes, queue=1 << 16):
        self.heap = heap
        heap._check(heap._lib.hvm_gpu_serve_start(heap._h, slots, slot_nodes, queue))
        self.running = True
        self._cap = 1 << 16
        self._tags = (ctypes.c_uint64 * self._cap)()
        self._ports = (ctypes.c_uint64 * self._cap)()

    def submit(self, defs, tags):
        """Queues one request per (definition, tag); HvmError QUEUE when the
        request ring lacks room (nothing queued)."""
        n = len(defs)
        ids = (ctypes.c_uint32 * n)(*[d.idx if hasattr(d, "idx") else int(d) for d in defs])
        tg = (ctypes.c_uint64 * n)(*[int(t) for t in tags])
        self.heap._check(self.heap._lib.hvm_gpu_serve_submit(self.heap._h, n, ids, tg))

    def poll(self):
        """[(tag, port)] of the requests finished since the last poll."""
        out = []
        while True:
            n = ctypes.c_uint32()
            self.heap._check(self.heap._lib.hvm_gpu_serve_poll(self.heap._h, self._tags, self._ports, self._cap,
                                                               ctypes.byref(n)))
            out.extend(zip(self._tags[:n.value], self._ports[:n.value]))
            if n.value < self._cap:
                return out

    def stop(self):
        """Ends the launch once every admitted request is finished; returns
        the launch's duration in ms."""
        ms = ctypes.c_double()
        self.running = False
        self.heap._check(self.heap._lib.hvm_gpu_serve_stop(self.heap._h, ctypes.byref(ms)))
        return ms.value


def pre_reduce(heap, book, skip=(), only=None, max_rwts=(1 << 64) - 1):
    """Book::pre_reduce (transform/src/pre_reduce.rs:38-75) with every
    definition's reduction on the device: reduces the definitions of `book`
    (text) except those in `skip` (or, with `only`, just those), each for at
    most `max_rwts` interactions.  Returns (new book text, Rewrites of the
    pre-reduction); the heap then holds the new book."""
    names = list(only if only is not None else skip)
    arr = (ctypes.c_char_p * max(len(names), 1))(*[x.encode() for x in names])
    cap = 4 * len(book) + (1 << 20)
    while True:
        buf = ctypes.create_string_buffer(cap)
        n = ctypes.c_uint64()
        c = (ctypes.c_uint64 * 5)()
        rc = heap._lib.hvm_gpu_pre_reduce(heap._h, book.encode(), arr, len(names), int(only is not None), max_rwts,
                                          buf, cap, ctypes.byref(n), c)
        if rc == 9 and n.value + 1 > cap:
            cap = n.value + 1
            continue
        heap._check(rc)
        return buf.value.decode(), Rewrites(*[int(x) for x in c])


def alloc_count(heap):
    """Nodes the allocation rings handed out since the last reset."""
    out = ctypes.c_uint64()
    heap._check(heap._lib.hvm_gpu_alloc_count(heap._h, ctypes.byref(out)))
    return int(out.value)


def arena_sizes(heap, defs, margin=1.25, extra=1024):
    """Arena sizes (nodes) for a planned batch of `defs`: each distinct
    definition is reduced once alone with the allocation rings and its
    allocation count, times `margin`, plus `extra`, sizes the arenas of its
    nets.  The handle's own net is reset afterwards."""
    need = {}
    for d in defs:
        idx = d.idx if hasattr(d, "idx") else int(d)
        if idx in need:
            continue
        net = Net(heap)
        net.boot(DefRef("", idx))
        net.normal()
        need[idx] = int(alloc_count(heap) * margin) + extra
    Net(heap)
    return [need[d.idx if hasattr(d, "idx") else int(d)] for d in defs]


def u24_port(value):
    """The Num port of a u24 value (runtime/src/port.rs:185-187, num/src/num.rs:86)."""
    return ((value % (1 << 24)) << 8) | 3


def _host_text(fn, *args):
    lib = load_library()
    cap = 1 << 16
    while True:
        buf = ctypes.create_string_buffer(cap)
        n = ctypes.c_uint64()
        rc = fn(*[a.encode() for a in args], buf, cap, ctypes.byref(n))
        if rc == 9 and n.value + 1 > cap:
            cap = n.value + 1
            continue
        if rc != 0:
            raise HvmError(rc, lib.hvm_gpu_last_error(None).decode())
        return buf.value.decode()


def book_print(book_text):
    """Book::from_str + Display (ast/src/parser.rs:178-183, ast/src/ast.rs:221-232)."""
    return _host_text(load_library().hvm_book_print, book_text)


def label_sets(book_text):
    """calculate_label_sets (host/src/calc_labels.rs:78-93) as {name: [labels]}."""
    out = {}
    for line in _host_text(load_library().hvm_book_label_sets, book_text).splitlines():
        name, _, labs = line.partition(":")
        out[name] = [int(x) for x in labs.split()]
    return out


def apply_tree(net_text, tree_text):
    """Net::apply_tree (ast/src/ast.rs:112-123), as net text."""
    return _host_text(load_library().hvm_net_apply_tree, net_text, tree_text)


def book_apply(book_text, name, tree_text):
    """Returns the book with definition `name` applied to `tree_text`
    (the tests/lists.rs pattern: book.get_mut(name).apply_tree(tree))."""
    printed = book_print(book_text)
    defs = {}
    for chunk in printed.split("\n\n"):
        head, _, body = chunk.partition(" = ")
        defs[head[1:]] = body
    defs[name] = apply_tree(defs[name], tree_text)
    return "\n\n".join(f"@{k} = {v}" for k, v in defs.items())


def normal_with(book_text, entry="main", heap_bytes=1 << 30, heap=None):
    """tests/loaders.rs:28-38: Host::new + boot(entry) + normal() + readback.

    Returns (Rewrites, normal_form_text, kernel_ms)."""
    own = heap is None
    heap = heap or Heap(heap_bytes)
    try:
        host = Host(book_text, heap)
        net = Net(heap)
        net.boot(host.defs[entry])
        net.normal()
        return net.rwts, host.readback(net), net.last_kernel_ms()
    finally:
        if own:
            heap.close()
