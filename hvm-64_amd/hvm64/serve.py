"""Serving: continuous batching of independent nets on one device.

Requests arrive over time (seeded Poisson arrivals).  The device is split
into `slots`: each slot is its own heap (Heap with a capped reduction grid)
driven by its own host thread, so the kernels of different slots overlap on
the device.  Whenever a slot is free it admits every request that has
arrived and is still waiting, oldest first, while the admitted requests'
node bounds fit the slot's arena bytes (and at most `max_batch` of them),
and reduces them as one Batch.  A request's latency is the time from its
arrival to the end of the batch that reduced it.

The heaps, books and grids stay resident on the device for the whole run;
only the batch's root cells move between host and device.
"""
import math
import random
import threading
import time
from dataclasses import dataclass, field

from . import Batch, Heap, Host, u24_port
from . import configs

NODE_BYTES = 16
# Parallel reduction keeps more of a net expanded at once than the
# sequential normal() whose peak the bound starts from.
PARALLEL_MARGIN = 16


def node_bound(kind, size):
    """Static node bound of one request: the oracle's sequential peak for
    these programs is about 7 * 2**size nodes (tree-sum depth `size`,
    bitonic 2**size keys), times PARALLEL_MARGIN."""
    return (7 * (1 << size) + 256) * PARALLEL_MARGIN


@dataclass
class Request:
    id: int
    arrival: float          # seconds after the start of the run
    kind: str
    size: int
    expected_port: int = 0
    bound_bytes: int = 0
    start: float = -1.0
    done: float = -1.0
    port: int = 0
    slot: int = -1
    batch: int = -1

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
    batches: int
    mean_batch: float
    slots: int
    per_slot_blocks: int
    slot_rewrites: list = field(default_factory=list)


def percentile(xs, q):
    if not xs:
        return float("nan")
    ys = sorted(xs)
    k = min(len(ys) - 1, max(0, int(math.ceil(q / 100.0 * len(ys))) - 1))
    return ys[k]


class Server:
    """`slots` heaps of `heap_bytes` each, every one with 1/slots of the
    device's reduction grid, all holding the same book of mix variants."""

    def __init__(self, slots=4, heap_bytes=16 << 30, max_batch=1024, book_seed=7, device=0, timeout_s=300):
        probe = Heap(1 << 30, device=device)
        full_blocks, _ = probe.grid()
        probe.close()
        self.blocks = max(1, full_blocks // slots)
        self.slots = []
        variants = [("tree_sum", d) for d in range(8, 17)] + [("bitonic", k) for k in range(6, 13)]
        book, self.table = configs.mix_book(variants, book_seed)
        for _ in range(slots):
            heap = Heap(heap_bytes, device=device, max_blocks=self.blocks)
            heap.set_timeout(timeout_s)
            host = Host(book, heap)
            self.slots.append((heap, host))
        self.heap_bytes = heap_bytes
        self.max_batch = max_batch

    def close(self):
        for heap, _ in self.slots:
            heap.close()
        self.slots = []

    def prepare(self, requests):
        for r in requests:
            r.expected_port = u24_port(int(self.table[(r.kind, r.size)][1]))
            r.bound_bytes = node_bound(r.kind, r.size) * NODE_BYTES

    def run(self, requests):
        """Serves `requests` (arrival times relative to now); returns ServeStats."""
        self.prepare(requests)
        pending = sorted(requests, key=lambda r: r.arrival)
        cond = threading.Condition()
        state = {"next": 0, "batches": 0, "sizes": []}
        slot_rewrites = [0] * len(self.slots)
        errors = []
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        def admit():
            # called with cond held: the longest prefix of arrived requests that fits
            i = state["next"]
            t = now()
            taken, used = [], 0
            while i < len(pending) and pending[i].arrival <= t and len(taken) < self.max_batch:
                r = pending[i]
                if taken and used + r.bound_bytes > self.heap_bytes:
                    break
                taken.append(r)
                used += r.bound_bytes
                i += 1
            state["next"] = i
            return taken

        def worker(slot_id):
            heap, host = self.slots[slot_id]
            try:
                while True:
                    with cond:
                        while True:
                            if state["next"] >= len(pending):
                                return
                            batch = admit()
                            if batch:
                                state["batches"] += 1
                                state["sizes"].append(len(batch))
                                batch_id = state["batches"]
                                break
                            wait = pending[state["next"]].arrival - now()
                            cond.wait(timeout=max(wait, 0.0005))
                    start = now()
                    b = Batch(heap)
                    b.boot([host.defs[self.table[(r.kind, r.size)][0]] for r in batch])
                    b.normal()
                    ports = b.ports()
                    end = now()
                    slot_rewrites[slot_id] += b.rwts.total()
                    for r, p in zip(batch, ports):
                        r.start, r.done, r.port, r.slot, r.batch = start, end, p, slot_id, batch_id
            except Exception as e:  # surfaced after the run
                errors.append(e)
                with cond:
                    cond.notify_all()

        threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(self.slots))]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        wall = now()
        done = [r for r in requests if r.done >= 0]
        lat = [r.latency * 1e3 for r in done]
        rewrites = sum(slot_rewrites)
        sizes = state["sizes"]
        return ServeStats(
            requests=len(requests), completed=len(done), all_ok=all(r.ok for r in requests),
            wall_s=wall, rewrites=rewrites, mips=rewrites / wall / 1e6 if wall > 0 else 0.0,
            p50_ms=percentile(lat, 50), p99_ms=percentile(lat, 99), max_ms=max(lat) if lat else float("nan"),
            batches=len(sizes), mean_batch=sum(sizes) / len(sizes) if sizes else 0.0,
            slots=len(self.slots), per_slot_blocks=self.blocks, slot_rewrites=slot_rewrites)
