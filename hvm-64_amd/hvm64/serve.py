"""Serving: requests admitted by the running reduction kernel.

Requests arrive over time (seeded Poisson arrivals).  One persistent
reduction launch serves them all (hvm_gpu_serve_*, include/hvm64_amd.h): the
heap is cut into `slots` arenas of `slot_nodes` nodes, a wave of the launch
admits waiting requests into free slots between its iterations, and a
request is finished when no redex of its net is left, which frees its slot
for the next one.  A request's latency is the time from its arrival to the
host's poll that saw it finished.

The host loop submits each request when its arrival time has come and polls
for finished ones; the book, the heap and the launch stay resident on the
device for the whole run.
"""
import math
import random
import time
from dataclasses import dataclass

import ctypes

from . import Heap, HvmError, Host, Serving, arena_sizes, u24_port
from . import configs

NODE_BYTES = 16
SUB_ARENAS = 8  # a slot's arena is cut in eight, taken from by lane index


def slot_layout(heap_bytes, slot_nodes, slots):
    """(slots that fit, nodes per slot, granule nodes) for a serving heap, as
    hvm_gpu_serve_start lays it out: granules of 2^10 nodes, doubled until the
    heap has at most 4M of them; the first granule holds the fixed cells;
    every slot is eight sub-arenas of whole granules."""
    heap_nodes = heap_bytes // NODE_BYTES
    gran = 1 << 10
    while heap_nodes // gran > (4 << 20):
        gran <<= 1
    sub = -(-slot_nodes // (SUB_ARENAS * gran))
    per_slot = sub * SUB_ARENAS * gran
    room = heap_nodes // gran - 1
    return max(0, min(slots, room // (sub * SUB_ARENAS))), per_slot, gran


@dataclass
class Request:
    id: int
    arrival: float          # seconds after the start of the run
    kind: str
    size: int
    expected_port: int = 0
    start: float = -1.0     # when it was submitted
    done: float = -1.0      # when a poll saw it finished
    port: int = 0

    @property
    def latency(self):
        return self.done - self.arrival

    @property
    def ok(self):
        return self.done >= 0 and self.port == self.expected_port


def poisson_requests(rate, duration, seed=7, log_uniform=True):
    """Seeded Poisson arrivals at `rate` per second over `duration` seconds,
    each request a net of the tree-sum / bitonic mix."""
    rng = random.Random(seed)
    out, t = [], 0.0
    while True:
        t += rng.expovariate(rate)
        if t >= duration:
            return out
        kind, size = configs.mix_choice(rng, log_uniform)
        out.append(Request(len(out), t, kind, size))


@dataclass
class ServeStats:
    requests: int
    completed: int
    all_ok: bool
    wall_s: float
    rewrites: int
    mips: float
    p50_ms: float
    p99_ms: float
    max_ms: float
    submits: int            # hvm_gpu_serve_submit calls
    mean_submit: float      # requests per call
    slots: int
    grid_blocks: int
    kernel_ms: float = 0.0


def percentile(xs, q):
    if not xs:
        return float("nan")
    ys = sorted(xs)
    k = min(len(ys) - 1, max(0, int(math.ceil(q / 100.0 * len(ys))) - 1))
    return ys[k]


class Server:
    """One heap of `heap_bytes` holding the book of mix variants, served by
    one persistent launch with `slots` request arenas."""

    def __init__(self, slots=1024, heap_bytes=192 << 30, book_seed=7, device=0, timeout_s=300, slot_nodes=None,
                 queue=1 << 16):
        self.heap = Heap(heap_bytes, device=device)
        self.heap.set_timeout(timeout_s)
        variants = [("tree_sum", d) for d in range(8, 17)] + [("bitonic", k) for k in range(6, 13)]
        book, self.table = configs.mix_book(variants, book_seed)
        self.host = Host(book, self.heap)
        if slot_nodes is None:
            # every variant alone, allocation counted; lanes take nodes in
            # chunks (a serving lane drops its chunk at every task), so a
            # slot holds twice the largest count (1.3 times ran dry)
            defs = [self.host.defs[self.table[v][0]] for v in variants]
            slot_nodes = 2 * max(arena_sizes(self.heap, defs, margin=1.0, extra=0)) + 4096
        self.slot_nodes = slot_nodes
        # as many slots as asked for, as far as the heap holds them
        self.n_slots, _, _ = slot_layout(heap_bytes, slot_nodes, slots)
        if self.n_slots == 0:
            raise HvmError(2, f"one slot of {slot_nodes} nodes exceeds the heap")  # HVM_E_OOM
        self.queue = queue
        self.blocks, _ = self.heap.grid()

    def close(self):
        if self.heap is not None:
            self.heap.close()
            self.heap = None

    def prepare(self, requests):
        for r in requests:
            r.expected_port = u24_port(int(self.table[(r.kind, r.size)][1]))

    def run(self, requests):
        """Serves `requests` (arrival times relative to now); returns ServeStats."""
        self.prepare(requests)
        pending = sorted(requests, key=lambda r: r.arrival)
        by_id = {r.id: r for r in requests}
        defs = {v: self.host.defs[self.table[v][0]] for v in self.table}
        srv = Serving(self.heap, self.n_slots, self.slot_nodes, self.queue)
        t0 = time.perf_counter()
        nxt, finished, submits, sizes = 0, 0, 0, []
        try:
            while finished < len(requests):
                t = time.perf_counter() - t0
                k = nxt
                while k < len(pending) and pending[k].arrival <= t:
                    k += 1
                if k > nxt:
                    batch = pending[nxt:k]
                    try:
                        srv.submit([defs[(r.kind, r.size)] for r in batch], [r.id for r in batch])
                        for r in batch:
                            r.start = t
                        sizes.append(len(batch))
                        submits += 1
                        nxt = k
                    except HvmError as e:
                        if e.kind != "QUEUE":
                            raise
                done = srv.poll()
                t = time.perf_counter() - t0
                for tag, port in done:
                    r = by_id[tag]
                    r.done, r.port = t, port
                finished += len(done)
                if not done and (nxt >= len(pending) or pending[nxt].arrival > t):
                    time.sleep(0.0001)
        finally:
            kernel_ms = srv.stop()
        wall = time.perf_counter() - t0
        done = [r for r in requests if r.done >= 0]
        lat = [r.latency * 1e3 for r in done]
        rewrites = self.heap_rewrites()
        self.last_sched = self.sched_stats()
        return ServeStats(
            requests=len(requests), completed=len(done), all_ok=all(r.ok for r in requests),
            wall_s=wall, rewrites=rewrites, mips=rewrites / wall / 1e6 if wall > 0 else 0.0,
            p50_ms=percentile(lat, 50), p99_ms=percentile(lat, 99), max_ms=max(lat) if lat else float("nan"),
            submits=submits, mean_submit=sum(sizes) / len(sizes) if sizes else 0.0,
            slots=self.n_slots, grid_blocks=self.blocks, kernel_ms=kernel_ms)

    def sched_stats(self):
        """Scheduler statistics of the last serving launch (summed over waves)."""
        out = (ctypes.c_uint64 * 20)()
        self.heap._check(self.heap._lib.hvm_gpu_sched_stats(self.heap._h, out))
        keys = ["busy_iters", "idle_iters", "busy_ticks", "spills", "spilled", "claims", "claimed",
                "interact_ticks", "spill_ticks", "alloc_miss", "ov_alloc",
                "anni_iters", "comm_iters", "call_iters", "other_iters",
                "anni_ticks", "comm_ticks", "call_ticks", "other_ticks", "pops"]
        return dict(zip(keys, [int(x) for x in out]))

    def heap_rewrites(self):
        out = (ctypes.c_uint64 * 5)()
        self.heap._check(self.heap._lib.hvm_gpu_rewrites(self.heap._h, out))
        return sum(int(x) for x in out)
