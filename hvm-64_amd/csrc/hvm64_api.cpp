// Engine object and the C ABI (include/hvm64_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/hvm64_amd.h"
#include "hvm64_engine.h"
#include "hvm64_host.h"

namespace hvm {
hipError_t launch_fill(u64 *words, u64 n, u64 value, hipStream_t st);
hipError_t launch_exec(const Params &P, u32 def, u64 t0, u32 count_dref, hipStream_t st);
hipError_t launch_expand(const Params &P, u64 old_root_cell, hipStream_t st);
hipError_t launch_reduce(const Params &P, u32 blocks, hipStream_t st);
hipError_t launch_boot_batch(const Params &P, u32 n, const u32 *defs, u64 *roots, u32 max_blocks, hipStream_t st);
hipError_t launch_expand_batch(const Params &P, u32 n, u64 *roots, u32 max_blocks, hipStream_t st);
hipError_t launch_root_ports(const Params &P, u32 n, const u64 *roots, u64 *out, hipStream_t st);
hipError_t launch_batch_reset(const Params &P, u64 pool_start, u64 ov_start, hipStream_t st);
hipError_t launch_readback(const Params &P, u64 root_cell, u64 q_head, u64 q_tail, u64 *stack, u64 stack_cap,
                           u64 *out, u64 out_cap, hipStream_t st);
u32 reduce_waves_per_block();
u32 reduce_book_lds_words();
u32 reduce_lds_slots();
hipError_t reduce_occupancy(int *blocks_per_cu);
hipError_t noalloc_scratch(u64 **out);
u32 noalloc_nets();
}  // namespace hvm

using namespace hvm;
using hvm64::Error;

struct hvm_gpu {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  u64 *heap = nullptr;
  u64 heap_words = 0;
  Globals *g = nullptr;
  u64 *queue = nullptr;
  u64 qslots = 0;
  DefRec *d_defs = nullptr;
  u64 *d_instrs = nullptr;
  u64 *d_labbits = nullptr;
  u64 *d_cdefs = nullptr;  // compiled-definition records
  u64 *d_tmpl = nullptr;   // templates of the compiled definitions
  u32 dev_instr_words = 0, tmpl_words = 0, n_compiled = 0;
  bool compiled = true;    // calls run compiled definitions where there are some
  u32 ndefs = 0;
  u64 *d_trg = nullptr;
  u32 trg_slots = 0;
  u32 max_nslots = 0;   // widest definition of the loaded book
  u32 blocks = 0, waves = 0;
  u64 *d_rb_stack = nullptr, *d_rb_out = nullptr;
  u64 *d_batch = nullptr;  // batch staging: roots, def ids, ports
  u64 *d_captured = nullptr;  // redexes recorded by calls on inert definitions (pairs)
  u64 captured_cap = 0;       // pairs
  u64 *scratch = nullptr;  // this device's static staging array (hvm_gpu_boot_batch_noalloc)
  // the planned batch (hvm_gpu_batch_plan): device arrays in one buffer
  struct Plan {
    u32 n = 0;
    u64 *buf = nullptr;  // roots | ports | arena_word | arena_base0 | pool_next | ids | arena_len | gmap
    u64 *roots = nullptr, *ports = nullptr, *arena_word = nullptr, *arena_base0 = nullptr;
    u64 *pool_next = nullptr;  // [0]: next pool granule, [1]: next overflow node
    u32 *ids = nullptr, *arena_len = nullptr, *gmap = nullptr;
    u64 gmap_len = 0, pool_start = 0, pool_end = 0;  // granules
    u64 ov_start = 0;                                 // node index: the overflow region
    u32 gran_shift = 0, refill = 1;
    std::vector<u64> base0;        // per net: first granule of its initial extent
    std::vector<u32> len;          // per net: granules of its initial extent
    std::vector<u64> arena_nodes;  // as planned, per net
    hvm_batch_tuning tuning{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;  // null: the steps are launched one by one
  } plan;
  u64 batch_cap = 0;       // words
  u64 rb_cap = 0;
  hvm64::Host host;
  std::vector<DefRec> defs_host;
  std::vector<u64> instrs_host, labbits_host;
  u64 root_cell = 0;
  u64 ring_node0 = 0, wave_nodes = 0, ov_node0 = 0, ov_nodes = 0;
  u64 *alloc_pos = nullptr;
  u64 *idle_bits = nullptr, *mbox = nullptr;
  u32 idle_words = 0;
  Globals last{};
  double last_ms = 0;
  // the serving launch (hvm_gpu_serve_start .. _stop)
  struct Serve {
    bool running = false;
    u32 slots = 0;
    u64 *dev = nullptr;        // pending | slot_root | slot_tag | free_ring | ctl | words | base0 | len | gmap | inl
    u64 *host = nullptr;       // pinned: host_ctl | requests | completions
    u64 *host_dev = nullptr;   // its device address
    u64 req_cap = 0, done_cap = 0;
    u64 req_tail = 0, done_read = 0;
    u64 *req = nullptr, *done = nullptr, *ctl = nullptr;  // host views
  } serve;
  bool heap_dirty = false;  // arenas left words the allocation rings would read as live
  u64 timeout_ticks = 0;  // wall-clock ticks a reduction may run (0 = unbounded)
  bool sched_stats_on = false;  // a single net's launches collect scheduler statistics
  u64 wall_khz = 100000;
  std::string err;
};

namespace {

thread_local std::string g_open_error;

// A planned batch captures the book's and the handle's device addresses:
// whatever replaces them drops the plan.
void drop_plan(hvm_gpu *h);

// Granule-map entry of a granule no net owns (the device's NO_NET).
const u32 NO_NET_HOST = 0xFFFFFFFFu;

// First word of a book image written by hvm_book_encode: "HVM64BK1".
const u64 BOOK_IMAGE_MAGIC = 0x314B4234364D5648ull;

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw Error(HVM_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

Params make_params(hvm_gpu *h, u64 limit) {
  Params P{};
  P.heap = h->heap;
  P.heap_words = h->heap_words;
  P.ring_node0 = h->ring_node0;
  P.wave_nodes = h->wave_nodes;
  P.ov_node0 = h->ov_node0;
  P.ov_nodes = h->ov_nodes;
  P.alloc_pos = h->alloc_pos;
  P.defs = h->d_defs;
  P.instrs = h->d_instrs;
  P.labbits = h->d_labbits;
  P.timing = h->sched_stats_on ? 1 : 0;
  P.captured = h->d_captured;
  P.captured_cap = h->captured_cap;
  P.trg_buf = h->d_trg;
  P.queue = h->queue;
  P.qmask = h->qslots - 1;
  P.g = h->g;
  P.ndefs = h->ndefs;
  P.trg_slots = h->trg_slots;
  P.total_waves = h->waves;
  P.instr_words = h->dev_instr_words;
  P.lab_words = (u32)h->labbits_host.size();
  P.cdefs = h->d_cdefs;
  P.tmpl = h->d_tmpl;
  P.tmpl_words = h->tmpl_words;
  P.compiled = h->compiled ? 1 : 0;
  P.idle_bits = h->idle_bits;
  P.mbox = h->mbox;
  P.idle_words = h->idle_words;
  P.t_lds = h->max_nslots <= reduce_lds_slots() ? 1 : 0;
  P.book_lds =
      (u64)h->ndefs * 5 + P.instr_words + P.lab_words + P.tmpl_words <= reduce_book_lds_words() ? 1 : 0;
  P.limit = limit;
  P.timeout_ticks = h->timeout_ticks;
  return P;
}

void sync_globals(hvm_gpu *h) {
  hip_check(hipStreamSynchronize(h->stream), "kernel");
  hip_check(hipMemcpy(&h->last, h->g, sizeof(Globals), hipMemcpyDeviceToHost), "read globals");
  if (h->last.err) {
    u32 e = h->last.err;
    int code = e == ERR_OOM ? HVM_E_OOM
               : e == ERR_UNIMPL ? HVM_E_UNIMPL_PAIR
               : e == ERR_DIV0 ? HVM_E_DIV0
               : e == ERR_QUEUE ? HVM_E_QUEUE
               : e == ERR_READBACK ? HVM_E_READBACK
               : e == ERR_TIMEOUT ? HVM_E_TIMEOUT
                                   : HVM_E_DEVICE;
    const char *msg = e == ERR_OOM ? "OOM"
                      : e == ERR_UNIMPL ? "unimplemented interaction (Switch-Ctr(0) or Op-Switch)"
                      : e == ERR_DIV0 ? "integer division by zero"
                      : e == ERR_QUEUE ? "global redex ring overflow"
                      : e == ERR_READBACK ? "readback failed"
                      : e == ERR_TIMEOUT ? "reduction exceeded its time limit"
                      : e == ERR_BADPTR ? "invalid heap or book address (engine bug)"
                                          : "device error";
    if (e == ERR_BADPTR) {
      char what[32];
      snprintf(what, sizeof what, "0x%llx", (unsigned long long)h->last.result);
      throw Error(code, std::string(msg) + ", check " + std::to_string(h->last.err_info) + " (" + what + ")");
    }
    if (e == ERR_OOM && h->last.err_info == OOM_POOL) throw Error(code, "OOM: the arena pool ran dry");
    if (e == ERR_OOM && h->last.err_info == OOM_OVERFLOW) throw Error(code, "OOM: the arena overflow region ran dry");
    throw Error(code, msg);
  }
}

// (Re)size the per-wave trg register file to the book's widest definition.
void ensure_trg(hvm_gpu *h, u32 slots) {
  if (slots <= h->trg_slots && h->d_trg) return;
  u32 s = std::max<u32>(slots, 16);
  dfree(h->d_trg);
  hip_check(hipMalloc(&h->d_trg, (size_t)h->waves * s * 64 * sizeof(u64)), "hipMalloc trg");
  h->trg_slots = s;
}

// The device trusts the book's instructions (no per-step checks), so every
// operand register index is checked here against its definition's size.
void validate_book(const hvm_gpu *h) {
  for (size_t i = 0; i < h->defs_host.size(); i++) {
    const DefRec &d = h->defs_host[i];
    if (d.nslots < 1 || (u64)d.ioff + d.ilen > h->instrs_host.size() / 2)
      throw Error(HVM_E_ARG, "book: definition " + std::to_string(i) + " is malformed");
    for (u32 k = 0; k < d.ilen; k++) {
      u64 w0 = h->instrs_host[2 * ((u64)d.ioff + k)], w1 = h->instrs_host[2 * ((u64)d.ioff + k) + 1];
      u32 kind = (u32)(w0 & 0xFF);
      u32 a = (u32)(w0 >> 16) & 0xFFFF, b = (u32)(w0 >> 32) & 0xFFFF, c = (u32)(w0 >> 48);
      u32 d4 = kind == I_WIRES ? (u32)(w1 & 0xFFFF) : 0;
      if (kind > I_WIRES || a >= d.nslots || b >= d.nslots || c >= d.nslots || d4 >= d.nslots)
        throw Error(HVM_E_ARG, "book: definition " + std::to_string(i) + " has an invalid instruction");
    }
  }
}

void upload_book(hvm_gpu *h) {
  validate_book(h);
  // The device book: the definitions' own instructions, then their compiled
  // programs and templates (hvm64_defcomp.cpp).
  hvm64::DeviceBook b = hvm64::build_device_book(h->defs_host, h->instrs_host);
  drop_plan(h);
  dfree(h->d_defs);
  dfree(h->d_instrs);
  dfree(h->d_labbits);
  dfree(h->d_cdefs);
  dfree(h->d_tmpl);
  h->ndefs = (u32)h->defs_host.size();
  auto up = [&](auto *&dst, const auto &vec) {
    size_t bytes = std::max<size_t>(vec.size() * sizeof(vec[0]), 16);
    hip_check(hipMalloc(&dst, bytes), "hipMalloc book");
    if (!vec.empty()) hip_check(hipMemcpy(dst, vec.data(), vec.size() * sizeof(vec[0]), hipMemcpyHostToDevice), "upload book");
  };
  up(h->d_defs, h->defs_host);
  up(h->d_instrs, b.instrs);
  up(h->d_labbits, h->labbits_host);
  up(h->d_cdefs, b.cdefs);
  up(h->d_tmpl, b.tmpl);
  h->dev_instr_words = (u32)b.instrs.size();
  h->tmpl_words = (u32)b.tmpl.size();
  h->n_compiled = b.n_compiled;
  h->max_nslots = b.max_slots;
  ensure_trg(h, b.max_slots);
}

// Net::new.  The heap is filled with FREE once, when the handle opens; a
// reset rewinds every allocation ring instead.  That is enough for a fresh
// net: a ring's first lap overwrites its slots without reading them, and
// later laps only reuse slots whose words read FREE, so words left by a
// previous net are either overwritten or skipped, never read as live.
void reset_net(hvm_gpu *h, bool fill) {
  fill = fill || h->heap_dirty;
  h->heap_dirty = false;
  if (fill) hip_check(launch_fill(h->heap, h->heap_words, P_FREE, h->stream), "fill heap");
  hip_check(launch_fill(h->queue, h->qslots * 2, 0, h->stream), "clear ring");
  Globals g0{};
  hip_check(hipMemsetAsync(h->alloc_pos, 0, (size_t)h->waves * 64 * sizeof(u64), h->stream), "clear rings");
  hip_check(hipMemcpyAsync(h->g, &g0, sizeof(Globals), hipMemcpyHostToDevice, h->stream), "init globals");
  u64 none = P_NONE;
  hip_check(hipMemcpyAsync(h->heap + ROOT0_WORD, &none, sizeof(u64), hipMemcpyHostToDevice, h->stream), "root");
  h->root_cell = HVM_HEAP_OFF + ROOT0_WORD * 8;
  hip_check(hipStreamSynchronize(h->stream), "reset");
  h->last = g0;
}

// Run the persistent reduction kernel with a redex budget.
// `fill`: for batched nets, where nearly every wave is busy and the global
// ring holds the excess, a wave whose rings run low claims from the global
// ring without going idle first (fuller iterations; a single net, with many
// idle waves, gains nothing from it).
void run_reduce(hvm_gpu *h, u64 limit, bool fill = false) {
  Params P = make_params(h, limit);
  P.fill = fill ? 1 : 0;
  u64 zero = 0;
  u64 zst[ST_SLOTS] = {};
  hip_check(hipMemcpyAsync(&h->g->done, &zero, sizeof(u64), hipMemcpyHostToDevice, h->stream), "budget");
  hip_check(hipMemcpyAsync(h->g->st, zst, sizeof(zst), hipMemcpyHostToDevice, h->stream), "stats");
  hip_check(hipMemsetAsync(h->idle_bits, 0, (size_t)h->idle_words * 8, h->stream), "clear idle set");
  hip_check(hipMemsetAsync(h->mbox, 0, (size_t)h->waves * MBOX_WORDS * 8, h->stream), "clear mailboxes");
  hip_check(hipEventRecord(h->ev0, h->stream), "event");
  hip_check(launch_reduce(P, h->blocks, h->stream), "launch reduce");
  hip_check(hipEventRecord(h->ev1, h->stream), "event");
  sync_globals(h);
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, h->ev0, h->ev1), "elapsed");
  h->last_ms = ms;
}

// Host::readback (readback.rs:23-33): the root's tree, then each pending
// redex as "\n  & a ~ b" (the Display of ast::Net, ast.rs:234-244), with one
// variable naming across all of them.  with_redexes: the net's pending
// redexes are in the global ring (a single net after reduce(limit)).
std::string readback(hvm_gpu *h, u64 root_cell, bool with_redexes = false) {
  u64 q_head = 0, q_tail = 0;
  if (with_redexes) {
    Globals g;
    hip_check(hipMemcpy(&g, h->g, sizeof(Globals), hipMemcpyDeviceToHost), "read globals");
    q_head = g.q_head;
    q_tail = std::max(g.q_head, g.q_tail);
    if (q_tail - q_head > h->qslots) throw Error(HVM_E_READBACK, "readback: redex ring inconsistent");
  }
  for (;;) {
    if (!h->d_rb_out) {
      h->rb_cap = std::max<u64>(h->rb_cap, 1 << 20);
      hip_check(hipMalloc(&h->d_rb_out, h->rb_cap * sizeof(u64)), "hipMalloc readback");
      hip_check(hipMalloc(&h->d_rb_stack, h->rb_cap * sizeof(u64)), "hipMalloc readback");
    }
    Params P = make_params(h, ~0ull);
    hip_check(launch_readback(P, root_cell, q_head, q_tail, h->d_rb_stack, h->rb_cap, h->d_rb_out, h->rb_cap,
                              h->stream),
              "launch readback");
    hip_check(hipStreamSynchronize(h->stream), "readback");
    Globals g;
    hip_check(hipMemcpy(&g, h->g, sizeof(Globals), hipMemcpyDeviceToHost), "read globals");
    if (g.err == ERR_READBACK && g.err_info == 1 && h->rb_cap < (1ull << 31)) {
      u32 zero = 0;
      hip_check(hipMemcpy(&h->g->err, &zero, sizeof(u32), hipMemcpyHostToDevice), "clear err");
      dfree(h->d_rb_out);
      dfree(h->d_rb_stack);
      h->rb_cap *= 4;
      continue;
    }
    if (g.err) throw Error(HVM_E_READBACK, "readback: malformed net");
    std::vector<u64> stream(g.result);
    if (g.result)
      hip_check(hipMemcpy(stream.data(), h->d_rb_out, g.result * sizeof(u64), hipMemcpyDeviceToHost), "copy readback");
    std::string text;
    hvm64::with_big_stack([&] {
      size_t pos = 0;
      std::map<u64, u64> vars;
      u64 next = 0;
      text = hvm64::readback_text(h->host, stream, pos, vars, next);
      for (u64 r = q_head; r < q_tail; r++) {
        const std::string a = hvm64::readback_text(h->host, stream, pos, vars, next);
        const std::string b = hvm64::readback_text(h->host, stream, pos, vars, next);
        text += "\n  & " + a + " ~ " + b;
      }
    });
    return text;
  }
}

// Device staging for batch calls: at least `words` 8-byte words.
u64 *batch_buffer(hvm_gpu *h, u64 words) {
  if (words > h->batch_cap) {
    dfree(h->d_batch);
    h->batch_cap = std::max<u64>(words, 1 << 16);
    hip_check(hipMalloc(&h->d_batch, h->batch_cap * sizeof(u64)), "hipMalloc batch");
  }
  return h->d_batch;
}

// Every entry point runs with the handle's device current: allocations,
// copies and launches made from a thread other than the one that opened the
// handle must land on the handle's GPU.
template <class F>
int guarded(hvm_gpu *h, F fn, bool while_serving = false) {
  if (!h) return HVM_E_ARG;
  if (h->serve.running && !while_serving) {
    h->err = "the handle is serving (hvm_gpu_serve_stop first)";
    return HVM_E_ARG;
  }
  try {
    hip_check(hipSetDevice(h->device), "hipSetDevice");
    fn();
    return 0;
  } catch (Error &e) {
    h->err = e.what();
    return e.code;
  } catch (std::exception &e) {
    h->err = e.what();
    return HVM_E_DEVICE;
  }
}

// ---------------------------------------------------------------------------
// planned batches: per-net arenas, one graph launch per batch
// ---------------------------------------------------------------------------
void drop_plan(hvm_gpu *h) {
  hvm_gpu::Plan &p = h->plan;
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  if (p.graph) (void)hipGraphDestroy(p.graph);
  dfree(p.buf);
  p = hvm_gpu::Plan{};
}

hvm_batch_tuning default_tuning(const hvm_gpu *h) {
  hvm_batch_tuning t{};
  t.blocks = h->blocks;
  t.mbox_keep = 256;
  t.fill_below = 32;
  t.arena_chunk = 32;
  return t;
}

Params plan_params(hvm_gpu *h) {
  hvm_gpu::Plan &p = h->plan;
  Params P = make_params(h, ~0ull);
  P.fill = 1;
  P.total_waves = p.tuning.blocks * reduce_waves_per_block();
  P.mbox_keep = p.tuning.mbox_keep;
  P.fill_below = p.tuning.fill_below;
  P.gmap = p.gmap;
  P.gmap_len = p.gmap_len;
  P.gran_shift = p.gran_shift;
  P.arena_word = p.arena_word;
  P.arena_base0 = p.arena_base0;
  P.arena_len = p.arena_len;
  P.pool_next = p.pool_next;
  P.pool_end = p.pool_end;
  P.ov_next = p.pool_next + 1;
  P.ov_end = p.gmap_len << p.gran_shift;
  P.n_nets = p.n;
  P.refill_granules = p.refill;
  P.arena_chunk = p.tuning.arena_chunk;
  return P;
}

// The steps of one planned batch, on the handle's stream: reset, boot every
// net into its arena, expand the roots, reduce, read the root ports.
void plan_steps(hvm_gpu *h) {
  hvm_gpu::Plan &p = h->plan;
  Params P = plan_params(h);
  hip_check(launch_batch_reset(P, p.pool_start, p.ov_start, h->stream), "launch batch reset");
  hip_check(launch_boot_batch(P, p.n, p.ids, p.roots, p.tuning.blocks * reduce_waves_per_block(), h->stream),
            "launch boot batch");
  hip_check(launch_expand_batch(P, p.n, p.roots, p.tuning.blocks * reduce_waves_per_block(), h->stream),
            "launch expand batch");
  hip_check(launch_reduce(P, p.tuning.blocks, h->stream), "launch reduce");
  hip_check(launch_root_ports(P, p.n, p.roots, p.ports, h->stream), "launch root ports");
}

// Lays out the arenas of n nets (each net's node bound rounded up to whole
// granules, one after another; the rest of the heap is the pool their
// arenas grow from), uploads the plan and captures its steps into a graph.
void make_plan(hvm_gpu *h, u32 n, const u32 *def_ids, const u64 *arena_nodes, const hvm_batch_tuning &tuning) {
  drop_plan(h);
  hvm_gpu::Plan &p = h->plan;
  p.n = n;
  p.tuning = tuning;
  p.arena_nodes.assign(arena_nodes, arena_nodes + n);
  const u64 heap_nodes = h->heap_words / 2;
  // granules: the smallest power of two of at least 1,024 nodes that keeps
  // the granule map within 4M entries
  u32 gs = 10;
  while ((heap_nodes >> gs) > (4u << 20)) gs++;
  p.gran_shift = gs;
  p.gmap_len = heap_nodes >> gs;  // whole granules only
  if (p.tuning.arena_chunk > (1u << gs)) p.tuning.arena_chunk = 1u << gs;
  // refill extents: half a net's initial arena, within [16, 64] granules
  p.refill = 64;
  std::vector<u64> base0(n);
  std::vector<u32> len(n);
  std::vector<u32> gmap(p.gmap_len, NO_NET_HOST);
  // granule 0 holds the heap's fixed cells
  u64 cur = (RESERVED_WORDS / 2 + (1ull << gs) - 1) >> gs;
  for (u32 i = 0; i < n; i++) {
    const u64 gr = (std::max<u64>(arena_nodes[i], 1) + (1ull << gs) - 1) >> gs;
    if (gr > 0xFFFFFFFFull || cur + gr > p.gmap_len)
      throw Error(HVM_E_OOM, "the arenas of the batch exceed the heap (" + std::to_string(i) + " of " +
                                 std::to_string(n) + " nets placed)");
    base0[i] = cur;
    len[i] = (u32)gr;
    for (u64 g = cur; g < cur + gr; g++) gmap[g] = i;
    cur += gr;
  }
  p.pool_start = cur;
  // the last sixteenth of the free granules (at least 64) is the overflow
  // region, served node by node while an extent is being replaced
  const u64 free_g = p.gmap_len - cur;
  const u64 ov_g = std::max<u64>(free_g / 16, 64);
  if (ov_g >= free_g) throw Error(HVM_E_OOM, "the arenas of the batch leave no pool in the heap");
  p.pool_end = p.gmap_len - ov_g;
  p.ov_start = p.pool_end << gs;
  const u64 words = 2 * (u64)n + 2 * (u64)n + 2 + (n + 1) / 2 * 2 + (p.gmap_len + 1) / 2 + 8;
  hip_check(hipMalloc(&p.buf, words * 8), "hipMalloc batch plan");
  u64 *w = p.buf;
  p.roots = w;
  w += n;
  p.ports = w;
  w += n;
  p.arena_word = w;
  w += n;
  p.arena_base0 = w;
  w += n;
  p.pool_next = w;
  w += 2;
  p.ids = (u32 *)w;
  w += (n + 1) / 2;
  p.arena_len = (u32 *)w;
  w += (n + 1) / 2;
  p.gmap = (u32 *)w;
  std::vector<u64> word0(n);
  for (u32 i = 0; i < n; i++) word0[i] = base0[i] << 32;
  const u64 pool0[2] = {p.pool_start, p.ov_start};
  hip_check(hipMemcpy(p.arena_word, word0.data(), (size_t)n * 8, hipMemcpyHostToDevice), "upload arenas");
  hip_check(hipMemcpy(p.arena_base0, base0.data(), (size_t)n * 8, hipMemcpyHostToDevice), "upload arenas");
  hip_check(hipMemcpy(p.pool_next, pool0, 16, hipMemcpyHostToDevice), "upload arenas");
  hip_check(hipMemcpy(p.arena_len, len.data(), (size_t)n * 4, hipMemcpyHostToDevice), "upload arenas");
  hip_check(hipMemcpy(p.ids, def_ids, (size_t)n * 4, hipMemcpyHostToDevice), "upload batch");
  hip_check(hipMemcpy(p.gmap, gmap.data(), p.gmap_len * 4, hipMemcpyHostToDevice), "upload granule map");
  p.base0 = std::move(base0);
  p.len = std::move(len);
  // the global redex ring starts empty (a completed reduction leaves it so)
  hip_check(launch_fill(h->queue, h->qslots * 2, 0, h->stream), "clear ring");
  hip_check(hipStreamSynchronize(h->stream), "clear ring");
  // the arenas overwrite the heap: the handle's own net is gone
  h->heap_dirty = true;
  // capture the steps; without a graph they are launched one by one
  hipGraph_t graph = nullptr;
  if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    bool ok = true;
    try {
      plan_steps(h);
    } catch (Error &) {
      ok = false;
    }
    if (hipStreamEndCapture(h->stream, &graph) != hipSuccess || !ok) {
      if (graph) (void)hipGraphDestroy(graph);
      graph = nullptr;
    }
  }
  (void)hipGetLastError();
  if (graph && hipGraphInstantiate(&p.exec, graph, nullptr, nullptr, 0) == hipSuccess) {
    p.graph = graph;
  } else {
    if (graph) (void)hipGraphDestroy(graph);
    p.exec = nullptr;
    (void)hipGetLastError();
  }
}

// Per net, the nodes the last run took from its arena (every extent it
// holds, less what is left of its current one), and the nodes the arenas
// took from the pool and the overflow region.
void plan_usage(hvm_gpu *h, std::vector<u64> &used, u64 &pool_nodes) {
  hvm_gpu::Plan &p = h->plan;
  std::vector<u64> word(p.n);
  u64 next[2] = {0, 0};
  hip_check(hipMemcpy(word.data(), p.arena_word, (size_t)p.n * 8, hipMemcpyDeviceToHost), "read arenas");
  hip_check(hipMemcpy(next, p.pool_next, 16, hipMemcpyDeviceToHost), "read arenas");
  const u64 pool_g = std::min(next[0], p.pool_end) - p.pool_start;
  const u64 ov_end = p.gmap_len << p.gran_shift;
  pool_nodes = (pool_g << p.gran_shift) + (std::min(next[1], ov_end) - p.ov_start);
  std::vector<u32> gmap(pool_g);
  if (pool_g) hip_check(hipMemcpy(gmap.data(), p.gmap + p.pool_start, pool_g * 4, hipMemcpyDeviceToHost), "read map");
  std::vector<u64> extra(p.n, 0);  // pool granules per net
  for (u32 net : gmap)
    if (net < p.n) extra[net]++;
  used.assign(p.n, 0);
  const u64 G = 1ull << p.gran_shift;
  for (u32 i = 0; i < p.n; i++) {
    const u64 refill = std::min<u64>(std::max<u64>(p.len[i] >> 1, 16), p.refill);  // the device's refill_of
    const u64 base = word[i] >> 32, off = word[i] & 0xFFFFFFFFull;
    const u64 cur_len = (base == p.base0[i] ? p.len[i] : refill) * G;
    const u64 total = (p.len[i] + extra[i]) * G;
    used[i] = total - cur_len + std::min(off, cur_len);
  }
}

// Plans the planned nets again (new tuning, or launch parameters changed).
void replan(hvm_gpu *h, const hvm_batch_tuning t) {  // a copy: make_plan clears the plan first
  hvm_gpu::Plan &p = h->plan;
  if (!p.n) return;
  const u32 n = p.n;
  std::vector<u32> ids(n);
  hip_check(hipMemcpy(ids.data(), p.ids, (size_t)n * 4, hipMemcpyDeviceToHost), "read batch");
  const std::vector<u64> nodes = p.arena_nodes;
  make_plan(h, n, ids.data(), nodes.data(), t);
}

// Runs the planned batch once; the duration of its steps goes to last_ms.
void run_plan(hvm_gpu *h) {
  hvm_gpu::Plan &p = h->plan;
  if (!p.n) throw Error(HVM_E_ARG, "no batch planned");
  hip_check(hipEventRecord(h->ev0, h->stream), "event");
  if (p.exec) hip_check(hipGraphLaunch(p.exec, h->stream), "graph launch");
  else plan_steps(h);
  hip_check(hipEventRecord(h->ev1, h->stream), "event");
  try {
    sync_globals(h);
  } catch (Error &) {
    // a failed reduction may leave redexes in the global ring
    (void)launch_fill(h->queue, h->qslots * 2, 0, h->stream);
    (void)hipStreamSynchronize(h->stream);
    throw;
  }
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, h->ev0, h->ev1), "elapsed");
  h->last_ms = ms;
}

}  // namespace

extern "C" {

int hvm_gpu_batch_plan(hvm_gpu *h, uint32_t n, const uint32_t *def_ids, const uint64_t *arena_nodes,
                       const hvm_batch_tuning *tuning) {
  return guarded(h, [&] {
    if (n == 0 || !def_ids || !arena_nodes) throw Error(HVM_E_ARG, "empty batch");
    for (u32 i = 0; i < n; i++)
      if (def_ids[i] >= h->ndefs) throw Error(HVM_E_ARG, "no such definition in batch");
    hvm_batch_tuning t = tuning ? *tuning : default_tuning(h);
    if (t.blocks == 0 || t.blocks > h->blocks) t.blocks = h->blocks;
    if (t.arena_chunk < 4) t.arena_chunk = 4;
    if (t.arena_chunk > 4096) t.arena_chunk = 4096;
    make_plan(h, n, def_ids, arena_nodes, t);
  });
}

int hvm_gpu_batch_run(hvm_gpu *h, uint64_t *ports, uint64_t *roots) {
  return guarded(h, [&] {
    run_plan(h);
    hvm_gpu::Plan &p = h->plan;
    if (ports) hip_check(hipMemcpy(ports, p.ports, (size_t)p.n * 8, hipMemcpyDeviceToHost), "download ports");
    if (roots) hip_check(hipMemcpy(roots, p.roots, (size_t)p.n * 8, hipMemcpyDeviceToHost), "download roots");
  });
}

int hvm_gpu_batch_info(hvm_gpu *h, hvm_batch_tuning *tuning, int *graph, uint64_t *pool_nodes) {
  return guarded(h, [&] {
    hvm_gpu::Plan &p = h->plan;
    if (tuning) *tuning = p.tuning;
    if (graph) *graph = p.exec != nullptr;
    if (pool_nodes) {
      *pool_nodes = 0;
      if (p.n) {
        std::vector<u64> used;
        plan_usage(h, used, *pool_nodes);
      }
    }
  });
}

int hvm_gpu_batch_used(hvm_gpu *h, uint64_t *used) {
  return guarded(h, [&] {
    hvm_gpu::Plan &p = h->plan;
    if (!p.n) throw Error(HVM_E_ARG, "no batch planned");
    if (!used) throw Error(HVM_E_ARG, "null argument");
    std::vector<u64> u;
    u64 pool = 0;
    plan_usage(h, u, pool);
    std::copy(u.begin(), u.end(), used);
  });
}

int hvm_gpu_batch_tune(hvm_gpu *h, hvm_batch_tuning *best) {
  return guarded(h, [&] {
    hvm_gpu::Plan &p = h->plan;
    if (!p.n) throw Error(HVM_E_ARG, "no batch planned");
    hvm_batch_tuning cur = p.tuning;
    const auto t0 = std::chrono::steady_clock::now();
    double best_ms = 1e300;
    // a candidate's first replay uploads its graph; the second is timed,
    // unless the first was already far slower than the best so far
    auto time_of = [&](const hvm_batch_tuning &t) {
      replan(h, t);
      run_plan(h);
      if (h->last_ms > 4 * best_ms) return h->last_ms;
      run_plan(h);
      return h->last_ms;
    };
    best_ms = time_of(cur);
    // one parameter at a time, the others fixed at their best so far; at
    // most TUNE_SECONDS of trials
    const double TUNE_SECONDS = 30;
    const u32 full = h->blocks;
    const std::vector<std::vector<u32>> cand = {
        {full, full * 3 / 4, full / 2}, {64, 128, 256, 512}, {16, 32, 48}, {8, 16, 32}};
    for (size_t k = 0; k < cand.size(); k++) {
      for (u32 v : cand[k]) {
        hvm_batch_tuning t = cur;
        u32 *field = k == 0 ? &t.blocks : k == 1 ? &t.mbox_keep : k == 2 ? &t.fill_below : &t.arena_chunk;
        if (*field == v || v == 0) continue;
        if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > TUNE_SECONDS) break;
        *field = v;
        const double ms = time_of(t);
        if (ms < best_ms) {
          best_ms = ms;
          cur = t;
        }
      }
    }
    replan(h, cur);
    if (best) *best = cur;
  });
}

int hvm_gpu_serve_start(hvm_gpu *h, uint32_t slots, uint64_t slot_nodes, uint64_t queue) {
  return guarded(h, [&] {
    if (slots == 0 || slot_nodes == 0 || queue == 0) throw Error(HVM_E_ARG, "empty serving configuration");
    if (slots > (1u << 20)) throw Error(HVM_E_ARG, "too many slots");
    if (!h->ndefs) throw Error(HVM_E_ARG, "no book loaded");
    drop_plan(h);
    hvm_gpu::Serve &sv = h->serve;
    sv = hvm_gpu::Serve{};
    const u64 heap_nodes = h->heap_words / 2;
    u32 gs = 10;
    while ((heap_nodes >> gs) > (4u << 20)) gs++;
    const u64 gmap_len = heap_nodes >> gs;
    // whole granules per sub-arena
    const u64 sub_g = (slot_nodes + SERVE_SUB * (1ull << gs) - 1) / (SERVE_SUB * (1ull << gs));
    const u64 slot_g = sub_g * SERVE_SUB;
    const u64 first = (RESERVED_WORDS / 2 + (1ull << gs) - 1) >> gs;  // granule 0: the fixed cells
    if (slot_g > 0xFFFFFFFFull || first + (u64)slots * slot_g > gmap_len)
      throw Error(HVM_E_OOM, "the slots' arenas exceed the heap");
    u64 rq = 1, dq = 1;
    while (rq < queue) rq <<= 1;
    while (dq < 4 * rq) dq <<= 1;
    sv.slots = slots;
    sv.req_cap = rq;
    sv.done_cap = dq;
    // device: pending | root | tag | free ring | counters | words | base0 | len | gmap | inl
    const u64 inl_words = (u64)h->waves * 64 * SERVE_INL * 2;
    const u64 nsub = (u64)slots * SERVE_SUB;
    const u64 words = (slots + 1) / 2 + 3 * (u64)slots + SV_WORDS + 2 * nsub + (nsub + 1) / 2 + (gmap_len + 1) / 2 +
                      inl_words + 64;
    hip_check(hipMalloc(&sv.dev, words * 8), "hipMalloc serving");
    u64 *w = sv.dev;
    u32 *pending = (u32 *)w;
    w += (slots + 1) / 2;
    u64 *slot_root = w;
    w += slots;
    u64 *slot_tag = w;
    w += slots;
    u64 *free_ring = w;
    w += slots;
    u64 *ctl = w;
    w += SV_WORDS;
    u64 *word = w;
    w += nsub;
    u64 *base0 = w;
    w += nsub;
    u32 *len = (u32 *)w;
    w += (nsub + 1) / 2;
    u32 *gmap = (u32 *)w;
    w += (gmap_len + 1) / 2;
    u64 *inl = w;
    w += inl_words;
    u64 *nopool = w;  // the serving arenas do not grow: a pool and an overflow region of 0
    std::vector<u64> h_word(nsub), h_base0(nsub), h_free(slots), h_ctl(SV_WORDS, 0);
    std::vector<u32> h_len(nsub, (u32)sub_g), h_gmap(gmap_len, NO_NET_HOST);
    for (u32 i = 0; i < slots; i++) {
      const u64 b = first + (u64)i * slot_g;
      for (u32 k = 0; k < SERVE_SUB; k++) {
        h_base0[(u64)i * SERVE_SUB + k] = b + k * sub_g;
        h_word[(u64)i * SERVE_SUB + k] = (b + k * sub_g) << 32;
      }
      h_free[i] = (u64)i + 1;
      for (u64 g = b; g < b + slot_g; g++) h_gmap[g] = i;
    }
    h_ctl[SV_FREE_TAIL] = slots;
    h_ctl[SV_FREE_COUNT] = slots;
    hip_check(hipMemset(pending, 0, (size_t)slots * 4), "init serving");
    hip_check(hipMemcpy(free_ring, h_free.data(), (size_t)slots * 8, hipMemcpyHostToDevice), "init serving");
    hip_check(hipMemcpy(ctl, h_ctl.data(), SV_WORDS * 8, hipMemcpyHostToDevice), "init serving");
    hip_check(hipMemcpy(word, h_word.data(), nsub * 8, hipMemcpyHostToDevice), "init serving");
    hip_check(hipMemcpy(base0, h_base0.data(), nsub * 8, hipMemcpyHostToDevice), "init serving");
    hip_check(hipMemcpy(len, h_len.data(), nsub * 4, hipMemcpyHostToDevice), "init serving");
    hip_check(hipMemcpy(gmap, h_gmap.data(), gmap_len * 4, hipMemcpyHostToDevice), "init serving");
    hip_check(hipMemset(nopool, 0, 64 * 8), "init serving");
    // pinned host memory: control words | requests | completions
    const u64 hwords = SH_WORDS + 2 * rq + 4 * dq;
    hip_check(hipHostMalloc((void **)&sv.host, hwords * 8, hipHostMallocMapped | hipHostMallocCoherent),
              "hipHostMalloc serving rings");
    std::memset(sv.host, 0, hwords * 8);
    hip_check(hipHostGetDevicePointer((void **)&sv.host_dev, sv.host, 0), "serving rings' device address");
    sv.ctl = sv.host;
    sv.req = sv.host + SH_WORDS;
    sv.done = sv.req + 2 * rq;
    // a clean control block, idle set, mailboxes and global ring
    hip_check(launch_fill(h->queue, h->qslots * 2, 0, h->stream), "clear ring");
    hip_check(hipMemsetAsync(h->idle_bits, 0, (size_t)h->idle_words * 8, h->stream), "clear idle set");
    hip_check(hipMemsetAsync(h->mbox, 0, (size_t)h->waves * MBOX_WORDS * 8, h->stream), "clear mailboxes");
    Globals g0{};
    hip_check(hipMemcpyAsync(h->g, &g0, sizeof(Globals), hipMemcpyHostToDevice, h->stream), "init globals");
    hip_check(hipStreamSynchronize(h->stream), "init serving");
    h->last = g0;
    h->heap_dirty = true;
    Params P = make_params(h, ~0ull);
    P.fill = 1;
    P.gmap = gmap;
    P.gmap_len = gmap_len;
    P.gran_shift = gs;
    P.arena_word = word;
    P.arena_base0 = base0;
    P.arena_len = len;
    P.pool_next = nopool;
    P.pool_end = 0;
    P.ov_next = nopool + 16;
    P.ov_end = 0;
    P.n_nets = slots;
    P.refill_granules = 1;
    P.arena_chunk = 32;
    P.serve = 1;
    P.serve_slots = slots;
    P.pending = pending;
    P.slot_root = slot_root;
    P.slot_tag = slot_tag;
    P.free_ring = free_ring;
    P.serve_ctl = ctl;
    P.host_ctl = sv.host_dev;
    P.req = sv.host_dev + SH_WORDS;
    P.req_mask = rq - 1;
    P.done = sv.host_dev + SH_WORDS + 2 * rq;
    P.done_mask = dq - 1;
    P.inl = inl;
    hip_check(hipEventRecord(h->ev0, h->stream), "event");
    hip_check(launch_reduce(P, h->blocks, h->stream), "launch serving kernel");
    hip_check(hipEventRecord(h->ev1, h->stream), "event");
    sv.running = true;
  });
}

int hvm_gpu_serve_submit(hvm_gpu *h, uint32_t n, const uint32_t *def_ids, const uint64_t *tags) {
  return guarded(h, [&] {
    hvm_gpu::Serve &sv = h->serve;
    if (!sv.running) throw Error(HVM_E_ARG, "not serving");
    if (n && (!def_ids || !tags)) throw Error(HVM_E_ARG, "null argument");
    const u64 dev_err = __atomic_load_n(&sv.ctl[SH_ERR], __ATOMIC_ACQUIRE);
    if (dev_err) throw Error(HVM_E_DEVICE, "the serving kernel stopped with device error " + std::to_string(dev_err));
    for (u32 i = 0; i < n; i++)
      if (def_ids[i] >= h->ndefs) throw Error(HVM_E_ARG, "no such definition");
    const u64 head = __atomic_load_n(&sv.ctl[SH_REQ_HEAD], __ATOMIC_ACQUIRE);
    if (sv.req_tail - head + n > sv.req_cap) throw Error(HVM_E_QUEUE, "request ring full");
    for (u32 i = 0; i < n; i++) {
      u64 *e = sv.req + ((sv.req_tail + i) & (sv.req_cap - 1)) * 2;
      e[0] = (u64)def_ids[i] + 1;
      e[1] = tags[i];
    }
    sv.req_tail += n;
    __atomic_store_n(&sv.ctl[SH_REQ_TAIL], sv.req_tail, __ATOMIC_RELEASE);
  }, true);
}

int hvm_gpu_serve_poll(hvm_gpu *h, uint64_t *tags, uint64_t *ports, uint32_t cap, uint32_t *n) {
  return guarded(h, [&] {
    hvm_gpu::Serve &sv = h->serve;
    if (!sv.running) throw Error(HVM_E_ARG, "not serving");
    if (!n || (cap && (!tags || !ports))) throw Error(HVM_E_ARG, "null argument");
    u32 k = 0;
    while (k < cap) {
      const u64 *e = sv.done + (sv.done_read & (sv.done_cap - 1)) * 4;
      if (__atomic_load_n(&e[0], __ATOMIC_ACQUIRE) != sv.done_read + 1) break;
      tags[k] = e[1];
      ports[k] = e[2];
      k++;
      sv.done_read++;
    }
    __atomic_store_n(&sv.ctl[SH_DONE_READ], sv.done_read, __ATOMIC_RELEASE);
    *n = k;
    const u64 dev_err = __atomic_load_n(&sv.ctl[SH_ERR], __ATOMIC_ACQUIRE);
    if (!k && dev_err)
      throw Error(HVM_E_DEVICE, "the serving kernel stopped with device error " + std::to_string(dev_err));
  }, true);
}

int hvm_gpu_serve_stop(hvm_gpu *h, double *ms) {
  return guarded(h, [&] {
    hvm_gpu::Serve &sv = h->serve;
    if (!sv.running) throw Error(HVM_E_ARG, "not serving");
    __atomic_store_n(&sv.ctl[SH_STOP], 1ull, __ATOMIC_RELEASE);
    const hipError_t e = hipStreamSynchronize(h->stream);
    float k_ms = 0;
    (void)hipEventElapsedTime(&k_ms, h->ev0, h->ev1);
    h->last_ms = k_ms;
    if (ms) *ms = k_ms;
    sv.running = false;
    if (sv.host) (void)hipHostFree(sv.host);
    dfree(sv.dev);
    sv = hvm_gpu::Serve{};
    hip_check(e, "serving kernel");
    sync_globals(h);
  }, true);
}

int hvm_gpu_alloc_count(hvm_gpu *h, uint64_t *nodes) {
  return guarded(h, [&] {
    if (!nodes) throw Error(HVM_E_ARG, "null argument");
    std::vector<u64> pos((size_t)h->waves * 64);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    hip_check(hipMemcpy(pos.data(), h->alloc_pos, pos.size() * 8, hipMemcpyDeviceToHost), "read rings");
    // ring positions advanced, less the slots found in use, plus the
    // overflow ring's allocations and the nodes lanes reused from their free
    // caches (an arena reuses nothing, so it needs all of them)
    hip_check(hipMemcpy(&h->last, h->g, sizeof(Globals), hipMemcpyDeviceToHost), "read globals");
    u64 sum = h->last.rw[RW_OV_ALLOC] + h->last.rw[RW_REUSE];
    for (u64 v : pos) sum += v;
    *nodes = sum - std::min(sum, (u64)h->last.rw[RW_ALLOC_MISS]);
  });
}

int hvm_gpu_open(int device, uint64_t heap_bytes, hvm_gpu **out) {
  return hvm_gpu_open_grid(device, heap_bytes, 0, out);
}

int hvm_gpu_open_grid(int device, uint64_t heap_bytes, uint32_t max_blocks, hvm_gpu **out) {
  if (!out) return HVM_E_ARG;
  *out = nullptr;
  auto h = std::make_unique<hvm_gpu>();
  try {
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw Error(HVM_E_ARG, "no such device");
    h->device = device;
    hip_check(hipSetDevice(device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "stream");
    hip_check(hipEventCreate(&h->ev0), "event");
    hip_check(hipEventCreate(&h->ev1), "event");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, device), "props");
    int per_cu = 0;
    hip_check(reduce_occupancy(&per_cu), "occupancy");
    if (per_cu < 1) per_cu = 1;
    h->blocks = (u32)(prop.multiProcessorCount * per_cu);
    if (max_blocks && max_blocks < h->blocks) h->blocks = max_blocks;
    h->waves = h->blocks * reduce_waves_per_block();
    u64 words = heap_bytes / 8;
    u64 min_words = RESERVED_WORDS + (u64)h->waves * 64 * 2 * 8 + 4096;
    if (words < min_words) words = min_words;
    h->heap_words = words;
    hip_check(hipMalloc(&h->heap, words * 8), "hipMalloc heap");
    // Node layout: reserved words, one ring per wave (each lane allocates
    // every 64th node of its wave's ring), then the shared overflow ring of
    // 2^j nodes, the largest power of two within a sixteenth of the heap.
    // The rings share the rest evenly (a multiple of 64 nodes each): a lane
    // laps its slice only once it has allocated that many nodes.
    u64 nodes = (words - RESERVED_WORDS) / 2;
    u64 ov = 64;
    while (ov * 2 <= nodes / 16) ov *= 2;
    if (nodes < ov) throw std::runtime_error("heap too small for the grid");
    const u64 per_wave = (nodes - ov) / h->waves / 64 * 64;
    if (per_wave < 64) throw std::runtime_error("heap too small for the grid");
    h->ring_node0 = RESERVED_WORDS / 2;
    h->wave_nodes = per_wave;
    h->ov_node0 = h->ring_node0 + per_wave * h->waves;
    h->ov_nodes = ov;
    hip_check(hipMalloc(&h->alloc_pos, (size_t)h->waves * 64 * sizeof(u64)), "hipMalloc rings");
    h->idle_words = (h->waves + 63) / 64;
    if (h->idle_words > 64) throw Error(HVM_E_ARG, "grid too large");
    hip_check(hipMalloc(&h->idle_bits, (size_t)h->idle_words * 8), "hipMalloc idle set");
    hip_check(hipMalloc(&h->mbox, (size_t)h->waves * MBOX_WORDS * 8), "hipMalloc mailboxes");
    hip_check(hipMalloc(&h->g, sizeof(Globals)), "hipMalloc globals");
    // Global redex ring: 2^26 slots (1 GiB).  A massively parallel net keeps
    // about (lanes x recursion depth) redexes pending; the LDS rings hold
    // 256 per wave and the rest waits here.
    h->qslots = 1ull << 26;
    while (h->qslots < (u64)h->waves * 1024) h->qslots <<= 1;
    hip_check(hipMalloc(&h->queue, h->qslots * 2 * sizeof(u64)), "hipMalloc ring");
    int wall_khz = 0;
    hip_check(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device), "wall clock rate");
    h->wall_khz = wall_khz > 0 ? (u64)wall_khz : 100000;
    h->timeout_ticks = h->wall_khz * 1000 * 600;  // 10 minutes unless the caller sets one
    hip_check(noalloc_scratch(&h->scratch), "staging array address");
    upload_book(h.get());
    reset_net(h.get(), true);
  } catch (Error &e) {
    g_open_error = e.what();
    hvm_gpu_close(h.release());
    return e.code;
  }
  *out = h.release();
  return 0;
}

void hvm_gpu_close(hvm_gpu *h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->serve.running) (void)hvm_gpu_serve_stop(h, nullptr);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_plan(h);
  dfree(h->heap);
  dfree(h->alloc_pos);
  dfree(h->idle_bits);
  dfree(h->mbox);
  dfree(h->d_batch);
  dfree(h->d_captured);
  dfree(h->g);
  dfree(h->queue);
  dfree(h->d_defs);
  dfree(h->d_instrs);
  dfree(h->d_labbits);
  dfree(h->d_cdefs);
  dfree(h->d_tmpl);
  dfree(h->d_trg);
  dfree(h->d_rb_out);
  dfree(h->d_rb_stack);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

const char *hvm_gpu_last_error(hvm_gpu *h) { return h ? h->err.c_str() : g_open_error.c_str(); }

int hvm_gpu_load_book_text(hvm_gpu *h, const char *text) {
  return guarded(h, [&] {
    if (!text) throw Error(HVM_E_ARG, "null book");
    hvm64::with_big_stack([&] {
      hvm64::Book book = hvm64::parse_book(text);
      hvm64::Host host;
      host.insert_book(book);
      h->host = std::move(host);
    });
    h->defs_host = h->host.def_records();
    h->instrs_host = h->host.all_instrs();
    h->labbits_host = h->host.all_labbits();
    upload_book(h);
  });
}

namespace {
// The handle's book from `host` (device records, instructions, label bits),
// uploaded.
void set_book(hvm_gpu *h, hvm64::Host host) {
  h->host = std::move(host);
  h->defs_host = h->host.def_records();
  h->instrs_host = h->host.all_instrs();
  h->labbits_host = h->host.all_labbits();
  upload_book(h);
}
}  // namespace

// Book::pre_reduce (transform/src/pre_reduce.rs:38-164).  The order in which
// definitions are reduced depends on the book alone, so it is found first
// (pre_reduce.rs:120-163: a definition's references, in tree order, before
// the definition; a definition met again while its own references are
// visited keeps its net); then each definition in turn is booted on a fresh
// net and reduced on the device, the redexes its inert calls captured go
// back among the pending ones, and the read-back net becomes both the
// definition's instructions (label sets unchanged) and, in the new book, its
// net.
int hvm_gpu_pre_reduce(hvm_gpu *h, const char *book_text, const char *const *names, uint32_t n_names, int only,
                       uint64_t max_rwts, char *out, uint64_t cap, uint64_t *len, uint64_t rwts[5]) {
  int rc = 0;
  int g = guarded(h, [&] {
    if (!book_text || (n_names && !names)) throw Error(HVM_E_ARG, "null argument");
    std::set<std::string> listed;
    for (u32 i = 0; i < n_names; i++) {
      if (!names[i]) throw Error(HVM_E_ARG, "null name");
      listed.insert(names[i]);
    }
    auto skip = [&](const std::string &x) { return only ? listed.count(x) == 0 : listed.count(x) != 0; };
    hvm64::Book book;
    hvm64::Host host;
    std::vector<std::string> order;
    hvm64::with_big_stack([&] {
      book = hvm64::parse_book(book_text);
      host.insert_book_with_inert_default(book);
      std::set<std::string> seen;
      std::function<void(const std::string &)> visit_def;
      std::function<void(u32)> visit_tree = [&](u32 t) {
        const hvm64::TNode &n = book.arena.nodes[t];
        if (n.kind == hvm64::TK::Ref) visit_def(book.arena.names[n.name]);
        if (n.kind == hvm64::TK::Ctr || n.kind == hvm64::TK::Op || n.kind == hvm64::TK::Switch) {
          visit_tree(n.a);
          visit_tree(n.b);
        }
      };
      visit_def = [&](const std::string &name) {
        if (seen.count(name) || skip(name) || !book.nets.count(name)) return;
        seen.insert(name);
        const hvm64::Net &net = book.nets.at(name);
        visit_tree(net.root);
        for (auto &r : net.redexes) {
          visit_tree(r.first);
          visit_tree(r.second);
        }
        order.push_back(name);
      };
      for (auto &kv : book.nets) visit_def(kv.first);
    });
    if (!h->d_captured) {
      h->captured_cap = 1 << 20;
      hip_check(hipMalloc(&h->d_captured, h->captured_cap * 16), "hipMalloc captured redexes");
    }
    set_book(h, host);
    u64 total[5] = {};
    std::map<std::string, std::string> reduced;
    bool stale = false;  // the device book lags the host's instructions
    for (const std::string &name : order) {
      if (stale) {
        h->defs_host = h->host.def_records();
        h->instrs_host = h->host.all_instrs();
        upload_book(h);
        stale = false;
      }
      reset_net(h, false);
      Params P = make_params(h, ~0ull);
      hip_check(launch_exec(P, h->host.def_id(name), mkport(T_VAR, 0, h->root_cell), 1, h->stream), "launch boot");
      sync_globals(h);
      // reduce(limit) interacts once before it compares (net.rs:41-52)
      run_reduce(h, std::max<u64>(max_rwts, 1));
      for (int k = 0; k < 5; k++) total[k] += h->last.rw[k];
      // the captured redexes after the pending ones (pre_reduce.rs:150-153)
      Globals gl;
      hip_check(hipMemcpy(&gl, h->g, sizeof(Globals), hipMemcpyDeviceToHost), "read globals");
      if (gl.captured_n > h->captured_cap) throw Error(HVM_E_OOM, "pre-reduce: too many captured redexes");
      const u64 tail = std::max(gl.q_head, gl.q_tail);
      if (tail - gl.q_head + gl.captured_n > h->qslots) throw Error(HVM_E_OOM, "pre-reduce: redex ring full");
      for (u64 i = 0; i < gl.captured_n; i++)
        hip_check(hipMemcpy(h->queue + ((tail + i) % h->qslots) * 2, h->d_captured + 2 * i, 16,
                            hipMemcpyDeviceToDevice),
                  "captured redexes");
      if (gl.captured_n) {
        const u64 nt = tail + gl.captured_n;
        hip_check(hipMemcpy(&h->g->q_tail, &nt, 8, hipMemcpyHostToDevice), "captured redexes");
      }
      std::string text = readback(h, h->root_cell, true);
      hvm64::with_big_stack([&] {
        hvm64::Arena arena;
        hvm64::Net net = hvm64::parse_net(arena, text);
        h->host.set_code(h->host.def_id(name), h->host.encode(arena, net));
      });
      reduced[name] = text;
      stale = true;
    }
    std::string result;
    hvm64::with_big_stack([&] {
      for (auto &kv : book.nets) {  // ast.rs:221-232
        if (!result.empty()) result += "\n\n";
        auto it = reduced.find(kv.first);
        result += "@" + kv.first + " = " + (it != reduced.end() ? it->second : hvm64::show_net(book.arena, kv.second));
      }
      // the handle now holds the new book, its label sets recomputed (the
      // reference's Host::new on the pre-reduced book); names it still
      // refers to but does not define stay inert
      hvm64::Book nb = hvm64::parse_book(result);
      hvm64::Host nh;
      nh.insert_book_with_inert_default(nb);
      host = std::move(nh);
    });
    set_book(h, std::move(host));
    reset_net(h, false);
    if (rwts)
      for (int k = 0; k < 5; k++) rwts[k] = total[k];
    if (len) *len = result.size();
    if (!out || cap < result.size() + 1) {
      rc = HVM_E_TRUNC;
      if (out && cap) {
        std::memcpy(out, result.data(), cap - 1);
        out[cap - 1] = 0;
      }
      return;
    }
    std::memcpy(out, result.data(), result.size() + 1);
  });
  return g ? g : rc;
}

int hvm_gpu_load_book(hvm_gpu *h, const hvm_book_blob *book) {
  return guarded(h, [&] {
    if (!book) throw Error(HVM_E_ARG, "null book");
    if (book->n_defs && !book->defs) throw Error(HVM_E_ARG, "null definition array");
    for (u32 i = 0; i < book->n_defs; i++) {
      const hvm_def_blob &d = book->defs[i];
      if ((d.n_lab_words && !d.lab_bits) || (d.n_instr && !d.instr))
        throw Error(HVM_E_ARG, "definition " + std::to_string(i) + " has a null array");
    }
    std::vector<std::string> names;
    std::vector<hvm64::LabSet> labs;
    std::vector<hvm64::DefCode> code;
    for (u32 i = 0; i < book->n_defs; i++) {
      const hvm_def_blob &d = book->defs[i];
      if (!d.name) throw Error(HVM_E_ARG, "definition without a name");
      names.push_back(d.name);
      hvm64::LabSet l;
      l.min_safe = d.min_safe;
      l.bits.assign(d.lab_bits, d.lab_bits + d.n_lab_words);
      labs.push_back(l);
      hvm64::DefCode c;
      c.nslots = d.n_slots ? d.n_slots : 1;
      for (u32 k = 0; k < d.n_instr; k++) {
        c.instr.push_back(d.instr[k].w0);
        c.instr.push_back(d.instr[k].w1);
      }
      code.push_back(std::move(c));
    }
    hvm64::Host host;
    host.adopt(std::move(names), std::move(labs), std::move(code));
    h->host = std::move(host);
    h->defs_host.clear();
    h->instrs_host.clear();
    h->labbits_host.clear();
    for (u32 i = 0; i < book->n_defs; i++) {
      const hvm_def_blob &d = book->defs[i];
      DefRec r{};
      r.ioff = (u32)(h->instrs_host.size() / 2);
      r.ilen = d.n_instr;
      r.nslots = d.n_slots ? d.n_slots : 1;
      r.labs_off = (u32)h->labbits_host.size();
      r.labs_words = d.n_lab_words;
      r.min_safe = d.min_safe;
      for (u32 k = 0; k < d.n_instr; k++) {
        h->instrs_host.push_back(d.instr[k].w0);
        h->instrs_host.push_back(d.instr[k].w1);
      }
      for (u32 k = 0; k < d.n_lab_words; k++) h->labbits_host.push_back(d.lab_bits[k]);
      h->defs_host.push_back(r);
    }
    upload_book(h);
  });
}

int hvm_gpu_load_book_image(hvm_gpu *h, const void *image, uint64_t len) {
  std::vector<hvm_def_blob> defs;
  std::vector<std::string> names;
  int rc = guarded(h, [&] {
    if (!image) throw Error(HVM_E_ARG, "null image");
    const u64 *w = (const u64 *)image;
    const u64 nw = len / 8;
    if (len % 8 || nw < 2 || w[0] != BOOK_IMAGE_MAGIC) throw Error(HVM_E_ARG, "not a book image");
    const u64 n = w[1];
    u64 pos = 2;
    auto need = [&](u64 words) {
      if (words > nw || pos > nw - words) throw Error(HVM_E_ARG, "truncated book image");
    };
    defs.resize(n);
    names.resize(n);
    for (u64 i = 0; i < n; i++) {
      need(3);
      const u32 name_len = (u32)w[pos], n_instr = (u32)(w[pos] >> 32);
      const u32 n_slots = (u32)w[pos + 1], min_safe = (u32)(w[pos + 1] >> 32);
      const u32 n_lab = (u32)w[pos + 2];
      pos += 3;
      const u64 name_words = ((u64)name_len + 7) / 8;
      need(name_words + n_lab + 2 * (u64)n_instr);
      names[i].assign((const char *)(w + pos), name_len);
      pos += name_words;
      hvm_def_blob &d = defs[i];
      d.min_safe = min_safe;
      d.n_lab_words = n_lab;
      d.lab_bits = w + pos;
      pos += n_lab;
      d.n_instr = n_instr;
      d.instr = (const hvm_instr *)(w + pos);
      pos += 2 * (u64)n_instr;
      d.n_slots = n_slots;
    }
    for (u64 i = 0; i < n; i++) defs[i].name = names[i].c_str();
  });
  if (rc) return rc;
  hvm_book_blob book{(uint32_t)defs.size(), defs.data()};
  return hvm_gpu_load_book(h, &book);
}

int hvm_gpu_def_id(hvm_gpu *h, const char *name, uint32_t *id) {
  return guarded(h, [&] {
    if (!name || !id) throw Error(HVM_E_ARG, "null argument");
    *id = h->host.def_id(name);
  }, true);  // host-only: allowed while serving
}

int hvm_gpu_reset(hvm_gpu *h) {
  return guarded(h, [&] { reset_net(h, false); });
}

int hvm_gpu_boot(hvm_gpu *h, uint32_t def_id) {
  return guarded(h, [&] {
    if (def_id >= h->ndefs) throw Error(HVM_E_ARG, "no such definition");
    Params P = make_params(h, ~0ull);
    hip_check(launch_exec(P, def_id, mkport(T_VAR, 0, h->root_cell), 1, h->stream), "launch boot");
    sync_globals(h);
  });
}

int hvm_gpu_encode_net_text(hvm_gpu *h, const char *net_text) {
  return guarded(h, [&] {
    if (!net_text) throw Error(HVM_E_ARG, "null net");
    hvm64::DefCode code;
    hvm64::with_big_stack([&] {
      hvm64::Arena arena;
      hvm64::Net net = hvm64::parse_net(arena, net_text);
      code = h->host.encode(arena, net);
    });
    // Append the net's code as a temporary definition after the book.
    DefRec r{};
    r.ioff = (u32)(h->instrs_host.size() / 2);
    r.ilen = code.ninstr();
    r.nslots = code.nslots;
    r.labs_off = 0;
    r.labs_words = 0;
    r.min_safe = 0;
    std::vector<DefRec> defs = h->defs_host;
    std::vector<u64> instrs = h->instrs_host;
    defs.push_back(r);
    instrs.insert(instrs.end(), code.instr.begin(), code.instr.end());
    std::swap(defs, h->defs_host);
    std::swap(instrs, h->instrs_host);
    upload_book(h);
    Params P = make_params(h, ~0ull);
    hip_check(launch_exec(P, (u32)h->defs_host.size() - 1, mkport(T_VAR, 0, h->root_cell), 0, h->stream),
              "launch encode");
    sync_globals(h);
    std::swap(defs, h->defs_host);
    std::swap(instrs, h->instrs_host);
    upload_book(h);
  });
}

int hvm_gpu_normal(hvm_gpu *h) {
  return guarded(h, [&] {
    Params P = make_params(h, ~0ull);
    hip_check(launch_expand(P, h->root_cell, h->stream), "launch expand");
    sync_globals(h);
    h->root_cell = h->last.result;
    run_reduce(h, ~0ull);
  });
}

int hvm_gpu_boot_batch(hvm_gpu *h, uint32_t n, const uint32_t *def_ids, uint64_t *roots) {
  return guarded(h, [&] {
    if (n == 0) return;
    if (!def_ids || !roots) throw Error(HVM_E_ARG, "null batch array");
    for (u32 i = 0; i < n; i++)
      if (def_ids[i] >= h->ndefs) throw Error(HVM_E_ARG, "no such definition in batch");
    u64 *buf = batch_buffer(h, (u64)n + (n + 1) / 2);
    u64 *d_roots = buf;
    u32 *d_defs = (u32 *)(buf + n);
    hip_check(hipMemcpyAsync(d_defs, def_ids, (size_t)n * 4, hipMemcpyHostToDevice, h->stream), "upload batch");
    Params P = make_params(h, ~0ull);
    hip_check(launch_boot_batch(P, n, d_defs, d_roots, h->waves, h->stream), "launch boot batch");
    hip_check(hipMemcpyAsync(roots, d_roots, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream), "download roots");
    sync_globals(h);
  });
}

// hvm_gpu_boot_batch staged through the device's static array instead of
// the handle's staging buffer: no device or host allocation.  Calls that
// overlap on one GPU share that array and race.
int hvm_gpu_boot_batch_noalloc(hvm_gpu *h, uint32_t n, const uint32_t *def_ids, uint64_t *roots) {
  return guarded(h, [&] {
    if (n == 0) return;
    if (!def_ids || !roots) throw Error(HVM_E_ARG, "null batch array");
    if (n > noalloc_nets())
      throw Error(HVM_E_ARG, "batch of " + std::to_string(n) + " nets exceeds the allocation-free maximum of " +
                                 std::to_string(noalloc_nets()));
    for (u32 i = 0; i < n; i++)
      if (def_ids[i] >= h->ndefs) throw Error(HVM_E_ARG, "no such definition in batch");
    u64 *d_roots = h->scratch;
    u32 *d_defs = (u32 *)(h->scratch + noalloc_nets());
    hip_check(hipMemcpyAsync(d_defs, def_ids, (size_t)n * 4, hipMemcpyHostToDevice, h->stream), "upload batch");
    Params P = make_params(h, ~0ull);
    hip_check(launch_boot_batch(P, n, d_defs, d_roots, h->waves, h->stream), "launch boot batch");
    hip_check(hipMemcpyAsync(roots, d_roots, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream), "download roots");
    sync_globals(h);
  });
}

int hvm_gpu_normal_batch(hvm_gpu *h, uint32_t n, uint64_t *roots) {
  return guarded(h, [&] {
    if (n && !roots) throw Error(HVM_E_ARG, "null batch array");
    if (n) {
      u64 *d_roots = batch_buffer(h, n);
      hip_check(hipMemcpyAsync(d_roots, roots, (size_t)n * 8, hipMemcpyHostToDevice, h->stream), "upload roots");
      Params P = make_params(h, ~0ull);
      hip_check(launch_expand_batch(P, n, d_roots, h->waves, h->stream), "launch expand batch");
      hip_check(hipMemcpyAsync(roots, d_roots, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream), "download roots");
      sync_globals(h);
    }
    run_reduce(h, ~0ull, true);
  });
}

int hvm_gpu_root_ports(hvm_gpu *h, uint32_t n, const uint64_t *roots, uint64_t *ports) {
  return guarded(h, [&] {
    if (n == 0) return;
    if (!roots || !ports) throw Error(HVM_E_ARG, "null batch array");
    u64 *buf = batch_buffer(h, 2 * (u64)n);
    hip_check(hipMemcpyAsync(buf, roots, (size_t)n * 8, hipMemcpyHostToDevice, h->stream), "upload roots");
    Params P = make_params(h, ~0ull);
    hip_check(launch_root_ports(P, n, buf, buf + n, h->stream), "launch root ports");
    hip_check(hipMemcpyAsync(ports, buf + n, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream), "download ports");
    sync_globals(h);
  });
}

int hvm_gpu_readback_root(hvm_gpu *h, uint64_t root, char *buf, uint64_t cap, uint64_t *len) {
  int rc = 0;
  int g = guarded(h, [&] {
    if (root < HVM_HEAP_OFF || root >= HVM_HEAP_OFF + h->heap_words * 8) throw Error(HVM_E_ARG, "not a heap cell");
    std::string text = readback(h, root);
    if (len) *len = text.size();
    if (!buf || cap < text.size() + 1) {
      rc = HVM_E_TRUNC;
      if (buf && cap) {
        std::memcpy(buf, text.data(), cap - 1);
        buf[cap - 1] = 0;
      }
      return;
    }
    std::memcpy(buf, text.data(), text.size());
    buf[text.size()] = 0;
  });
  return g ? g : rc;
}

int hvm_gpu_reduce(hvm_gpu *h, uint64_t limit, uint64_t *popped, int *is_normal) {
  return guarded(h, [&] {
    u64 before = h->last.rw[RW_POPS];
    if (limit == 0) {
      if (popped) *popped = 0;
      if (is_normal) *is_normal = (h->last.S >> 32) == 0;
      return;
    }
    run_reduce(h, limit);
    const u64 n = h->last.rw[RW_POPS] - before;
    if (popped) *popped = n;
    // net.rs:41-52 returns None as soon as `limit` redexes were popped, even
    // when none is left, and Some(count) only when the queue ran dry first
    if (is_normal) *is_normal = n < limit && (h->last.S >> 32) == 0;
  });
}

int hvm_gpu_rewrites(hvm_gpu *h, uint64_t out[5]) {
  return guarded(h, [&] {
    for (int k = 0; k < 5; k++) out[k] = h->last.rw[k];
  });
}

int hvm_gpu_shortcuts(hvm_gpu *h, uint64_t out[3]) {
  return guarded(h, [&] {
    for (int k = 0; k < 3; k++) out[k] = h->last.rw[SC_OP_ERA + k];
  });
}

int hvm_gpu_snapshot(hvm_gpu *h, hvm_net_image *img, uint64_t first, uint64_t n, uint64_t *dst) {
  return guarded(h, [&] {
    if (!img) throw Error(HVM_E_ARG, "null image");
    if (first > h->heap_words || n > h->heap_words - first) throw Error(HVM_E_ARG, "word range outside the heap");
    img->root_cell = h->root_cell;
    img->heap_base = HVM_HEAP_OFF;
    img->heap_words = h->heap_words;
    if (dst && n) {
      hip_check(hipStreamSynchronize(h->stream), "snapshot");
      hip_check(hipMemcpy(dst, h->heap + first, n * sizeof(u64), hipMemcpyDeviceToHost), "copy heap image");
    }
  });
}

int hvm_gpu_readback(hvm_gpu *h, char *buf, uint64_t cap, uint64_t *len) {
  int rc = 0;
  int g = guarded(h, [&] {
    std::string text = readback(h, h->root_cell, true);
    if (len) *len = text.size();
    if (!buf || cap < text.size() + 1) {
      rc = HVM_E_TRUNC;
      if (buf && cap) {
        std::memcpy(buf, text.data(), cap - 1);
        buf[cap - 1] = 0;
      }
      return;
    }
    std::memcpy(buf, text.data(), text.size());
    buf[text.size()] = 0;
  });
  return g ? g : rc;
}

int hvm_gpu_last_kernel_ms(hvm_gpu *h, double *ms) {
  return guarded(h, [&] { *ms = h->last_ms; });
}

int hvm_gpu_set_compiled(hvm_gpu *h, int on, uint32_t *n_compiled) {
  return guarded(h, [&] {
    h->compiled = on != 0;
    if (n_compiled) *n_compiled = h->n_compiled;
    replan(h, h->plan.tuning);
  });
}

int hvm_gpu_set_timeout(hvm_gpu *h, double seconds) {
  return guarded(h, [&] {
    if (seconds < 0) throw Error(HVM_E_ARG, "negative timeout");
    h->timeout_ticks = (u64)(seconds * (double)h->wall_khz * 1000.0);
    replan(h, h->plan.tuning);
  });
}

// ---------------------------------------------------------------------------
// host-only book utilities (no device needed)
// ---------------------------------------------------------------------------
static int host_text(const std::function<std::string()> &fn, char *buf, uint64_t cap, uint64_t *len) {
  try {
    std::string text;
    hvm64::with_big_stack([&] { text = fn(); });
    if (len) *len = text.size();
    if (!buf || cap < text.size() + 1) {
      if (buf && cap) {
        std::memcpy(buf, text.data(), cap - 1);
        buf[cap - 1] = 0;
      }
      return HVM_E_TRUNC;
    }
    std::memcpy(buf, text.data(), text.size() + 1);
    return 0;
  } catch (Error &e) {
    g_open_error = e.what();
    return e.code;
  } catch (std::exception &e) {
    g_open_error = e.what();
    return HVM_E_ARG;
  }
}

int hvm_book_print(const char *book_text, char *buf, uint64_t cap, uint64_t *len) {
  if (!book_text) return HVM_E_ARG;
  return host_text([&] { return hvm64::show_book(hvm64::parse_book(book_text)); }, buf, cap, len);
}

int hvm_book_encode(const char *book_text, void *buf, uint64_t cap, uint64_t *len) {
  if (!book_text) return HVM_E_ARG;
  try {
    std::vector<u64> img;
    hvm64::with_big_stack([&] {
      hvm64::Book book = hvm64::parse_book(book_text);
      hvm64::Host host;
      host.insert_book(book);
      img.push_back(BOOK_IMAGE_MAGIC);
      img.push_back(host.ndefs());
      for (u32 i = 0; i < host.ndefs(); i++) {
        const std::string &name = host.def_name(i);
        const hvm64::DefCode &c = host.code(i);
        const hvm64::LabSet &l = host.labs(i);
        img.push_back((u64)name.size() | ((u64)c.ninstr() << 32));
        img.push_back((u64)c.nslots | ((u64)l.min_safe << 32));
        img.push_back((u64)l.bits.size());
        std::vector<u64> nw((name.size() + 7) / 8, 0);
        std::memcpy(nw.data(), name.data(), name.size());
        img.insert(img.end(), nw.begin(), nw.end());
        img.insert(img.end(), l.bits.begin(), l.bits.end());
        img.insert(img.end(), c.instr.begin(), c.instr.end());
      }
    });
    const u64 bytes = img.size() * 8;
    if (len) *len = bytes;
    if (!buf || cap < bytes) return HVM_E_TRUNC;
    std::memcpy(buf, img.data(), bytes);
    return 0;
  } catch (Error &e) {
    g_open_error = e.what();
    return e.code;
  } catch (std::exception &e) {
    g_open_error = e.what();
    return HVM_E_ARG;
  }
}

int hvm_book_label_sets(const char *book_text, char *buf, uint64_t cap, uint64_t *len) {
  if (!book_text) return HVM_E_ARG;
  return host_text(
      [&] {
        auto sets = hvm64::calculate_label_sets(hvm64::parse_book(book_text));
        std::string out;
        for (auto &kv : sets) {
          out += kv.first + ":";
          for (u32 l = 0; l < kv.second.min_safe; l++)
            if (kv.second.has((u16)l)) out += " " + std::to_string(l);
          out += "\n";
        }
        return out;
      },
      buf, cap, len);
}

int hvm_net_apply_tree(const char *net_text, const char *tree_text, char *buf, uint64_t cap, uint64_t *len) {
  if (!net_text || !tree_text) return HVM_E_ARG;
  return host_text(
      [&] {
        hvm64::Arena arena;
        hvm64::Net net = hvm64::parse_net(arena, net_text);
        u32 arg = hvm64::parse_tree(arena, tree_text);
        hvm64::apply_tree(arena, net, arg);
        return hvm64::show_net(arena, net);
      },
      buf, cap, len);
}

int hvm_gpu_sched_stats(hvm_gpu *h, uint64_t out[20]) {
  return guarded(h, [&] {
    for (u32 k = 0; k < ST_CLS_IT; k++) out[k] = h->last.st[k];
    out[9] = h->last.rw[RW_ALLOC_MISS];
    out[10] = h->last.rw[RW_OV_ALLOC];
    for (u32 k = 0; k < 8; k++) out[11 + k] = h->last.st[ST_CLS_IT + k];
    out[19] = h->last.rw[RW_POPS];
  });
}

int hvm_gpu_set_sched_stats(hvm_gpu *h, int on) {
  return guarded(h, [&] {
    h->sched_stats_on = on != 0;
    replan(h, h->plan.tuning);  // a plan's graph captured the launch's parameters
  });
}

int hvm_gpu_grid(hvm_gpu *h, uint32_t *blocks, uint32_t *waves) {
  return guarded(h, [&] {
    if (blocks) *blocks = h->blocks;
    if (waves) *waves = h->waves;
  }, true);
}

}  // extern "C"
