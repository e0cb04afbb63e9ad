// gfx950 reduction kernels for HVM-64 interaction nets.
//
// Hot path: runtime/src/net.rs:41-60 (Net::reduce / Net::normal) with the
// rules of runtime/src/interact.rs, the instruction interpreter of
// runtime/src/def.rs:217-289 and instruction.rs:110-210, and ExpandDef
// (net.rs:72-109).
//
// Memory model.  Heap words are shared by every wave on all 8 XCDs, so every
// heap access is a device-scope operation: relaxed agent-scope loads/stores
// (global_load/store ... sc1) and 64-bit atomic exchange.  A port is
// published (exchanged into a var cell or handed to another wave) only after
// `s_waitcnt vmcnt(0)` has retired the stores that initialised its node.
//
// Linking.  Node words are written once when the node is created and read
// once when it is consumed; all concurrent wiring goes through var cells.  A
// var cell joins exactly two endpoints.  The first endpoint to link swaps its
// port into the cell and leaves; the second swaps, finds that port, frees
// the cell and keeps linking.  There is no lock and no spin on another lane's
// progress, so divergent lanes of one wave can never wait on each other.
#include <hip/hip_runtime.h>

#include <atomic>

#include "hvm64_rules.h"

namespace hvm {

// Sum the per-lane counters of a wave into the global counters.
DEV void wave_flush_counters(Lane &L, u32 lane) {
#pragma unroll
  for (u32 k = 0; k < RW_SLOTS; k++) {
    u64 v = L.rw[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0 && v) __hip_atomic_fetch_add(&L.P->g->rw[k], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    L.rw[k] = 0;
  }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Fill the heap with FREE words.
__global__ void k_fill(u64 *words, u64 n, u64 value) {
  u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  u64 stride = (u64)gridDim.x * blockDim.x;
  for (; i < n; i += stride) words[i] = value;
}

// Run one definition's instructions with t0 = `t0` on lane 0 of wave 0:
// Net::boot (net.rs:30-32, counts one dref) or Host::encode_net
// (encode.rs:31-35, counts nothing).  Redexes go to the global ring.
__global__ void k_exec(Params P, u32 def, u64 t0, u32 count_dref) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  __shared__ u64 s_pl[2 * LINK_CAP * 64];
  __shared__ u64 s_t[LDS_SLOTS * 64];
  Lane L;
  lane_init(L, P, 0, 0);
  L.PL = s_pl;
  L.T_lds = s_t;
  if (count_dref) L.rw[RW_DREF]++;
  run_task(L, false, 0, 0, def, t0);
  publish();
  lane_finish(L);
  for (u32 k = 0; k < RW_SLOTS; k++)
    if (L.rw[k]) __hip_atomic_fetch_add(&P.g->rw[k], (u64)L.rw[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A batch of independent nets: net i gets a fresh root cell (roots[i]) and
// is booted with definition defs[i] (Net::boot per net, net.rs:30-32, one
// DREF each).  One wave per block; each lane boots nets lane, lane + 64*B,
// ...  Redexes go to the global ring.
__global__ __launch_bounds__(64) void k_boot_batch(Params P, u32 n, const u32 *defs, u64 *roots) {
  __shared__ u64 s_pl[2 * LINK_CAP * 64];
  __shared__ u64 s_t[LDS_SLOTS * 64];
  const u32 lane = threadIdx.x;
  Lane L;
  lane_init(L, P, blockIdx.x, lane);
  L.PL = &s_pl[lane];
  L.T_lds = &s_t[lane];
  for (u32 i = blockIdx.x * 64 + lane; i < n; i += gridDim.x * 64) {
    u32 di = defs[i];
    if (di >= P.ndefs) {
      fail(L, ERR_BADPTR, BAD_DEF_CALL);
      roots[i] = 0;
      continue;
    }
    L.cur_net = i;  // with arenas, net i allocates from arena i
    u64 r = alloc_var1(L);
    roots[i] = r;
    L.rw[RW_DREF]++;
    run_task(L, false, 0, 0, di, var_port(r));
  }
  publish();
  lane_finish(L);
  wave_flush_counters(L, lane);
}

// Net::expand for every net of a batch: roots[i] is replaced by the new root.
__global__ __launch_bounds__(64) void k_expand_batch(Params P, u32 n, u64 *roots) {
  __shared__ u64 s_pl[2 * LINK_CAP * 64];
  const u32 lane = threadIdx.x;
  Lane L;
  lane_init(L, P, blockIdx.x, lane);
  L.PL = &s_pl[lane];
  for (u32 i = blockIdx.x * 64 + lane; i < n; i += gridDim.x * 64) {
    u64 old = roots[i];
    L.cur_net = i;
    u64 r2 = alloc_node(L), e = r2 + 8;
    st2(P, r2, P_NONE, var_port(r2));
    L.dirty = true;
    u32 np = 0;
    push_link(L, np, var_port(old), expand_ref(e));
    flush_links(L, np);
    roots[i] = r2;
  }
  publish();
  lane_finish(L);
  wave_flush_counters(L, lane);
}

// Resets everything a batched launch reads before it starts, with stores
// only (a HIP graph replays it ahead of every batch): the control block, the
// idle set, the mailboxes, every net's arena (back to its initial extent),
// the pool (from granule pool_start, owned by no net) and the overflow
// region (from node ov_start).
__global__ void k_batch_reset(Params P, u64 pool_start, u64 ov_start) {
  const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x, stride = (u64)gridDim.x * blockDim.x;
  u64 *gw = (u64 *)P.g;
  for (u64 i = tid; i < sizeof(Globals) / 8; i += stride) gw[i] = 0;
  for (u64 i = tid; i < P.idle_words; i += stride) P.idle_bits[i] = 0;
  for (u64 i = tid; i < (u64)P.total_waves * MBOX_WORDS; i += stride) P.mbox[i] = 0;
  for (u64 i = tid; i < P.n_nets; i += stride) P.arena_word[i] = P.arena_base0[i] << 32;
  for (u64 i = pool_start + tid; i < P.pool_end; i += stride) P.gmap[i] = NO_NET;
  if (tid == 0) {
    *P.pool_next = pool_start;
    *P.ov_next = ov_start;
  }
}

// The port each root cell leads to (following parked ports without
// consuming them); a root whose wire is still open yields Var(root).
__global__ void k_root_ports(Params P, u32 n, const u64 *roots, u64 *out) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    u64 cell = roots[i];
    u64 p = var_port(cell);
    for (u32 hop = 0; hop < MAX_CHAIN && is_endpoint(p); hop++) {
      u64 v = ld(P, paddr(p));
      if (v == P_NONE) break;
      p = v;
    }
    out[i] = p;
  }
}

// Net::expand (net.rs:65-69): a fresh root cell R2 and an ExpandDef whose
// `out` is R2's other endpoint, linked to the old root.  Writes R2's address
// to g->result.
__global__ void k_expand(Params P, u64 old_root_cell) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  __shared__ u64 s_pl[2 * LINK_CAP * 64];
  Lane L;
  lane_init(L, P, 0, 0);
  L.PL = s_pl;
  u64 r2 = alloc_node(L), e = r2 + 8;
  st2(P, r2, P_NONE, var_port(r2));
  L.dirty = true;
  u32 np = 0;
  push_link(L, np, var_port(old_root_cell), expand_ref(e));
  flush_links(L, np);
  publish();
  lane_finish(L);
  P.g->result = r2;
  for (u32 k = 0; k < RW_SLOTS; k++)
    if (L.rw[k]) __hip_atomic_fetch_add(&P.g->rw[k], (u64)L.rw[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The persistent reduction kernel (net.rs:41-60).  Every wave keeps its
// redexes in LDS rings, one per rule class (redex_class); each iteration it
// picks the fullest class and its 64 lanes pop up to 64 redexes of it (LIFO)
// and interact them, so the lanes of an iteration run the same rule.  New
// redexes are pushed into the rings by class.  A wave that ends an iteration
// with more than one iteration's worth of redexes spills the excess (the
// oldest redexes of its fullest class) to the global ring while other waves
// are idle; an empty wave claims up to 64 redexes from the global ring.
// S = (queued << 32) | active waves; the reduction is complete when S == 0,
// a state no wave can leave.  S is one word shared by every wave and its
// device-scope accesses all go to one memory channel, so busy waves touch it
// only to spill (and read it every 64 iterations).  Idle waves poll their own
// mailbox and read S only every IDLE_CHECK polls (for redexes on the global
// ring); the wave that brings S to 0 writes MBOX_DONE into every mailbox, so
// the end is seen within one poll.  Reading S more often slows every other
// memory access of the reduction (tree-sum depth 18: 1,260 MIPS reading it
// every 64 polls, 1,850 every 1,024; without mailboxes, a 50 us backoff cap
// made bitonic sort of 2^16 keys 2.6x slower).  Under a reduce(limit) budget
// mailboxes are off and idle waves poll S with a growing, jittered backoff
// capped near 1 ms.
static const u32 REFRESH = 64;        // busy iterations between control-word reads
static const u32 MAX_BACKOFF = 16384;  // idle backoff cap, in s_sleep(2) units (~0.9 ms)
static const u32 MBOX_POLL = 16;       // s_sleep(2) units between mailbox polls (~1 us)
static const u32 IDLE_CHECK = 1024;    // mailbox polls between reads of S (a power of two)
// Defaults of the tunable thresholds (Params::mbox_keep, Params::fill_below).
static const u32 MBOX_KEEP = 256;      // redexes a busy wave keeps before handing some off
static const u32 FILL_BELOW = 32;      // batched nets: pending redexes under which a busy wave claims

DEV void wave_sleep(u32 units) {
  for (u32 z = 0; z < units; z++) __builtin_amdgcn_s_sleep(2);
}

DEV u32 uniform(u32 v) { return __builtin_amdgcn_readfirstlane(v); }

// Moves `cnt` redexes, the oldest of class `c`, from the wave's rings to the
// global ring (all lanes; lane 0 updates the counters).  Returns S after.
// Takes the next k (<= 64) redexes of the global ring, whose count in S this
// wave has already claimed, into its rings (the caller checks they fit).
DEV void claim_global(const Params &P, Lane &L, u64 *bag, u32 *top_c, u32 k, u32 lane, u64 t_end) {
  u64 base = 0;
  if (lane == 0) base = __hip_atomic_fetch_add(&P.g->q_head, (u64)k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  base = __shfl(base, 0, 64);
  if (lane < k) {
    u64 *slot = P.queue + ((base + lane) & P.qmask) * 2;
    u64 a, b;
    for (;;) {  // the producer reserved this slot and is writing it
      a = gld(slot);
      b = gld(slot + 1);
      if (a != 0 && b != 0) break;
      if (wall_clock64() > t_end) {
        fail(L, ERR_TIMEOUT);
        a = b = P_ERA;  // skippable pair: dropped without effect
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    gst(slot, 0);
    gst(slot + 1, 0);
    const u32 c = redex_class(a, b, L.book, has_book(L), P.ndefs);
    const u32 idx = HVM_LDS_FAA(&top_c[c], 1u);
    const u32 i = (ring_base(c) + (idx & (ring_cap(c) - 1))) * 2;
    bag[i] = a;
    bag[i + 1] = b;
  }
}

DEV u64 spill_to_global(const Params &P, u64 *bag, u32 *bot_c, u32 c, u32 cnt, u32 lane, u32 bot, u32 &err) {
  // the nodes these redexes reach must be in memory before a claimer can
  // read the ring slots (which it may do before S counts them)
  publish();
  u64 base = 0;
  if (lane == 0) base = __hip_atomic_fetch_add(&P.g->q_tail, (u64)cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  base = __shfl(base, 0, 64);
  for (u32 i = lane; i < cnt; i += 64) {
    u32 s2 = (ring_base(c) + ((bot + i) & (ring_cap(c) - 1))) * 2;
    u64 *q = P.queue + ((base + i) & P.qmask) * 2;
    gst(q, bag[s2]);
    gst(q + 1, bag[s2 + 1]);
  }
  publish();
  u64 old = 0;
  if (lane == 0) {
    old = __hip_atomic_fetch_add(&P.g->S, (u64)cnt << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((old >> 32) + cnt > (P.qmask + 1) / 2) err = ERR_QUEUE;
    *bot_c = bot + cnt;
  }
  return __shfl(old, 0, 64) + ((u64)cnt << 32);
}

// Serving: wave 0 admits waiting requests, one per lane (at most 64, and at
// most as many as there are free slots), booting each into a free slot.
// Returns how many it booted.
DEV u32 serve_admit(const Params &P, Lane &L, u32 lane, u64 t_end) {
  u64 *const sv = P.serve_ctl;
  u64 *const hc = P.host_ctl;
  u64 k = 0, head = 0, fhead = 0;
  if (lane == 0) {
    const u64 tail = HVM_LD_SYS(&hc[SH_REQ_TAIL]);
    head = HVM_LD(&sv[SV_REQ_HEAD]);  // only wave 0 moves the admission counters
    const u64 avail = tail > head ? tail - head : 0;
    const u64 free = HVM_LD(&sv[SV_FREE_COUNT]);
    k = avail < free ? avail : free;
    k = k < 64 ? k : 64;
    if (k) {
      HVM_FAA(&sv[SV_FREE_COUNT], (u64)0 - k);
      fhead = HVM_LD(&sv[SV_FREE_HEAD]);
      HVM_ST(&sv[SV_FREE_HEAD], fhead + k);
      HVM_ST(&sv[SV_REQ_HEAD], head + k);
    }
  }
  k = __shfl(k, 0, 64);
  if (k == 0) return 0;
  head = __shfl(head, 0, 64);
  fhead = __shfl(fhead, 0, 64);
  if (lane < k) {
    // the entry a completing lane wrote before it counted the slot free
    u64 *slot = &P.free_ring[(fhead + lane) % P.serve_slots];
    u64 v = 0;
    for (;;) {
      v = HVM_LD(slot);
      if (v != 0) break;
      if (wall_clock64() > t_end) {
        fail(L, ERR_TIMEOUT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (v != 0) {
      HVM_ST(slot, 0);
      const u64 *rq = P.req + ((head + lane) & P.req_mask) * 2;
      const u64 d = HVM_LD_SYS(rq), tag = HVM_LD_SYS(rq + 1);
      if (d == 0 || d - 1 >= P.ndefs) fail(L, ERR_BADPTR, BAD_DEF_CALL);
      else serve_boot(L, (u32)(v - 1), (u32)(d - 1), tag);
    }
  }
  if (lane == 0) HVM_ST_SYS(&hc[SH_REQ_HEAD], head + k);  // the host sees the ring's room
  return (u32)k;
}

// Serving: requests waiting and slots free (wave 0, lane 0's view).
DEV bool serve_waiting(const Params &P, u32 lane) {
  u64 w = 0;
  if (lane == 0) {
    const u64 tail = HVM_LD_SYS(&P.host_ctl[SH_REQ_TAIL]);
    w = tail > HVM_LD(&P.serve_ctl[SV_REQ_HEAD]) && HVM_LD(&P.serve_ctl[SV_FREE_COUNT]) > 0;
  }
  return __shfl(w, 0, 64) != 0;
}

// After pushes outside an iteration (admission), trim each ring whose push
// counter ran past its capacity: those pushes went to the global ring.
DEV void clamp_rings(u32 *top_c, const u32 *bot_c, u32 lane) {
  if (lane < NCLS && top_c[lane] - bot_c[lane] > ring_cap(lane)) top_c[lane] = bot_c[lane] + ring_cap(lane);
  __builtin_amdgcn_wave_barrier();
}

// ARENAS: compiled for per-net arenas (P.gmap set) or for the per-lane
// allocation rings; each instantiation carries only its mode's allocator.
// SERVE: the serving launch (arenas are request slots; the kernel runs until
// the host sets the stop flag and no work is left).
// TIMING: the launch collects scheduler statistics (iterations, spills,
// claims, wall-clock ticks): the batched and serving instantiations always;
// the single-net one when asked for, as keeping them costs it 2-4%.
// LDS: the book and every definition's trg registers fit in LDS (P.book_lds
// and P.t_lds are set), so the instantiation reads them only from there.
template <bool ARENAS, bool SERVE, bool TIMING, bool LDS>
__global__ __launch_bounds__(WAVES_PER_BLOCK * 64) void k_reduce(Params P) {
  __shared__ u64 s_bag[WAVES_PER_BLOCK][RING_PAIRS * 2];
  __shared__ u64 s_T[WAVES_PER_BLOCK][(LDS_SLOTS + 2 * LINK_CAP) * 64];
  __shared__ u32 s_top[WAVES_PER_BLOCK][NCLS];
  __shared__ u32 s_bot[WAVES_PER_BLOCK][NCLS];
  __shared__ u64 s_book[BOOK_LDS_WORDS];
  const u32 wib = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // the book, copied into LDS when it fits (the host decides: P.book_lds)
  const u32 nd = P.ndefs * 4;
  const u32 b_cdef = nd + P.instr_words + P.lab_words, b_tmpl = b_cdef + P.ndefs;
  if (LDS || P.book_lds) {
    const u64 *defs = (const u64 *)P.defs;
    for (u32 i = threadIdx.x; i < nd; i += blockDim.x) s_book[i] = defs[i];
    for (u32 i = threadIdx.x; i < P.instr_words; i += blockDim.x) s_book[nd + i] = P.instrs[i];
    for (u32 i = threadIdx.x; i < P.lab_words; i += blockDim.x) s_book[nd + P.instr_words + i] = P.labbits[i];
    for (u32 i = threadIdx.x; i < P.ndefs; i += blockDim.x) s_book[b_cdef + i] = P.cdefs[i];
    for (u32 i = threadIdx.x; i < P.tmpl_words; i += blockDim.x) s_book[b_tmpl + i] = P.tmpl[i];
  }
  __syncthreads();
  const u32 gwave = blockIdx.x * WAVES_PER_BLOCK + wib;
  Globals *g = P.g;
  u64 *bag = s_bag[wib];
  u32 *top_c = s_top[wib];
  u32 *bot_c = s_bot[wib];
  if (lane < NCLS) {
    top_c[lane] = 0;
    bot_c[lane] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  Lane L;
  lane_init(L, P, gwave, lane);
  L.arenas = ARENAS;
  L.serve = SERVE;
  L.lds_all = LDS;
  L.T_lds = &s_T[wib][lane];
  L.PL = &s_T[wib][LDS_SLOTS * 64 + lane];
  if (LDS || P.book_lds) {
    L.book = HVM_TO_LDS(s_book);
    L.book_ins = nd;
    L.book_lab = nd + P.instr_words;
    L.book_cdef = b_cdef;
    L.book_tmpl = b_tmpl;
  }
  L.bag = bag;
  L.bag_top = top_c;
  L.bag_bot = bot_c;
  __builtin_amdgcn_wave_barrier();
  bool active = false;
  const bool limited = P.limit != ~0ull;
  const u32 mbox_keep = P.mbox_keep ? P.mbox_keep : MBOX_KEEP;
  const u32 fill_below = P.fill_below ? P.fill_below : FILL_BELOW;
  const u64 t_end = P.timeout_ticks ? wall_clock64() + P.timeout_ticks : ~0ull;
  u32 backoff = 1 + (gwave & 7);
  u32 busy = 0;                // busy iterations since the last control read
  // The lanes that run an iteration's redexes start where the last
  // iteration's left off (wave-uniform), so every lane's allocation slice
  // fills at the same pace (with lane i taking the i-th redex, low lanes ran
  // nearly every task and lapped their slices long before the others).
  u32 rot = 0;
  u64 S_cache = 0;             // last value of S seen by this wave
  u64 S_next = 0;              // lane 0: control words read at iteration end
  u32 err_next = 0;
  bool fresh = false;          // S_next / err_next hold a read to consume
  u64 st[ST_SLOTS] = {};
  // mailboxes: off under a reduce(limit) budget (a wave that stops there
  // must not have a delivery in flight)
  const bool use_mbox = !limited && P.mbox != nullptr && P.idle_bits != nullptr;
  u64 *my_box = use_mbox ? P.mbox + (u64)gwave * MBOX_WORDS : nullptr;
  bool in_set = false;         // this wave's bit is set in P.idle_bits
  u32 idle_polls = 0;
  u32 admit_tick = 0;          // serving, wave 0: busy iterations since the last admission
  if (use_mbox) {  // every wave starts idle: in the idle set, not counted in S
    if (lane == 0)
      __hip_atomic_fetch_or(&P.idle_bits[gwave >> 6], 1ull << (gwave & 63), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    in_set = true;
  }

  for (;;) {
    if (fresh) {
      S_cache = __shfl(S_next, 0, 64);
      u32 err = __shfl(err_next, 0, 64);
      fresh = false;
      if (err) {
        if (SERVE && lane == 0) HVM_ST_SYS(&P.host_ctl[SH_ERR], (u64)err);
        break;
      }
      if (wall_clock64() > t_end) {
        fail(L, ERR_TIMEOUT);
        break;
      }
    }
    // ---- serving: the last task's count change (a net it completed), then
    // wave 0 admits requests every fourth busy iteration
    if (SERVE) serve_settle(L);
    if (SERVE && gwave == 0 && active && (++admit_tick & 3) == 0) {
      __builtin_amdgcn_wave_barrier();
      serve_admit(P, L, lane, t_end);
      __builtin_amdgcn_wave_barrier();
      clamp_rings(top_c, bot_c, lane);
    }
    // ring occupancy: lane c < NCLS reads class c
    u32 my_top = lane < NCLS ? top_c[lane] : 0, my_bot = lane < NCLS ? bot_c[lane] : 0;
    u32 my_n = my_top - my_bot;
    u32 n = my_n;
#pragma unroll
    for (u32 o = 1; o < 16; o <<= 1) n += __shfl_xor(n, o, 64);
    n = uniform(n);

    // ---- empty rings: idle.  An idle wave is in the idle set (its bit in
    // P.idle_bits) and polls its own mailbox, where a busy wave can hand it
    // redexes directly and where the end of the reduction is announced
    // (MBOX_DONE); every IDLE_CHECK polls it also reads S, for redexes on
    // the global ring, errors and the deadline.
    if (n == 0) {
      if (active) {
        if (use_mbox && lane == 0)
          __hip_atomic_fetch_or(&P.idle_bits[gwave >> 6], 1ull << (gwave & 63), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
        in_set = use_mbox;
        u64 old = 0;
        if (lane == 0) old = __hip_atomic_fetch_sub(&g->S, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        S_cache = __shfl(old, 0, 64) - 1;
        active = false;
        idle_polls = 0;
        // nothing active, nothing queued: done (serving: once the host has
        // asked to stop)
        bool end = S_cache == 0;
        if (SERVE && end) {
          u64 stop = 0;
          if (lane == 0) stop = HVM_LD_SYS(&P.host_ctl[SH_STOP]);
          end = __shfl(stop, 0, 64) != 0;
        }
        if (end) {
          // no wave is active, so no delivery is in flight: tell the idle
          // waves through their mailboxes
          if (use_mbox)
            for (u32 w = lane; w < P.total_waves; w += 64) gst(P.mbox + (u64)w * MBOX_WORDS, MBOX_DONE);
          break;
        }
      }
      if (TIMING) st[ST_IDLE_IT]++;
      if (use_mbox) {
        u64 k = 0;
        if (lane == 0) k = gld(my_box);
        k = __shfl(k, 0, 64);
        if (k == MBOX_DONE) break;
        if (k > 0) {  // a delivery: the sender already counted this wave active
          if (lane < k) {
            const u64 a = gld(my_box + 16 + 2 * lane), b = gld(my_box + 16 + 2 * lane + 1);
            const u32 c = redex_class(a, b, L.book, has_book(L), P.ndefs);  // the rings are empty: k <= 64 entries fit
            const u32 idx = HVM_LDS_FAA(&top_c[c], 1u);
            const u32 i = (ring_base(c) + (idx & (ring_cap(c) - 1))) * 2;
            bag[i] = a;
            bag[i + 1] = b;
          }
          if (lane == 0) gst(my_box, 0);
          if (TIMING) st[ST_CLAIMS]++;
          if (TIMING) st[ST_CLAIMED] += k;
          active = true;
          in_set = false;
          backoff = 1 + (gwave & 7);
          __builtin_amdgcn_wave_barrier();
          continue;
        }
      }
      // serving: wave 0 takes waiting requests while idle
      // (only a wave still in the idle set may leave it to admit: one that is
      // out of it has a delivery on the way, and must take that first, or a
      // second sender could write its mailbox while it is read)
      if (SERVE && gwave == 0 && (in_set || !use_mbox) && serve_waiting(P, lane)) {
        bool left = true;
        if (use_mbox) {
          u64 old = 0;
          if (lane == 0)
            old = __hip_atomic_fetch_and(&P.idle_bits[gwave >> 6], ~(1ull << (gwave & 63)), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
          left = ((__shfl(old, 0, 64) >> (gwave & 63)) & 1) != 0;
          in_set = false;  // either we left, or a sender took us out and a delivery is coming
        }
        if (left) {
          if (lane == 0) __hip_atomic_fetch_add(&g->S, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          active = true;
          serve_admit(P, L, lane, t_end);
          __builtin_amdgcn_wave_barrier();
          clamp_rings(top_c, bot_c, lane);
          continue;
        }
      }
      if (!use_mbox || (idle_polls++ & (IDLE_CHECK - 1)) == 0) {
        u64 s = 0;
        u32 err = 0, stop = 0;
        if (lane == 0) {
          s = gld(&g->S);
          err = gld32(&g->err);
          if (SERVE) stop = (u32)HVM_LD_SYS(&P.host_ctl[SH_STOP]);
        }
        s = __shfl(s, 0, 64);
        err = __shfl(err, 0, 64);
        stop = __shfl(stop, 0, 64);
        S_cache = s;
        if (err) {
          if (SERVE && lane == 0) HVM_ST_SYS(&P.host_ctl[SH_ERR], (u64)err);
          break;
        }
        if (s == 0 && (!SERVE || stop)) break;
        if (limited && gld(&g->done) >= P.limit) break;
        if (wall_clock64() > t_end) {
          fail(L, ERR_TIMEOUT);
          break;
        }
        // redexes on the global ring: leave the idle set, then claim
        if ((s >> 32) > 0 && (!use_mbox || in_set)) {
          bool left = true;
          if (use_mbox) {
            u64 old = 0;
            if (lane == 0)
              old = __hip_atomic_fetch_and(&P.idle_bits[gwave >> 6], ~(1ull << (gwave & 63)), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            left = ((__shfl(old, 0, 64) >> (gwave & 63)) & 1) != 0;
            in_set = false;  // either we left, or a sender took us out and a delivery is coming
          }
          if (left) {
            u32 k = 0;
            if (lane == 0) {
              for (int tries = 0; tries < 4; tries++) {
                u64 q = s >> 32;
                if (q == 0) break;
                u64 want = q < 64 ? q : 64;
                u64 ns = s - (want << 32) + 1;
                if (__hip_atomic_compare_exchange_strong(&g->S, &s, ns, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)) {
                  k = (u32)want;
                  break;
                }
              }
            }
            k = __shfl(k, 0, 64);
            if (k > 0) {
              backoff = 1 + (gwave & 7);
              active = true;
              if (TIMING) st[ST_CLAIMS]++;
              if (TIMING) st[ST_CLAIMED] += k;
              claim_global(P, L, bag, top_c, k, lane, t_end);  // the rings are empty: k <= 64 fit
              __builtin_amdgcn_wave_barrier();
              continue;
            }
            if (use_mbox) {  // nothing claimed: back into the idle set
              if (lane == 0)
                __hip_atomic_fetch_or(&P.idle_bits[gwave >> 6], 1ull << (gwave & 63), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
              in_set = true;
            }
          }
        }
      }
      if (use_mbox) {
        wave_sleep(MBOX_POLL);
      } else {
        wave_sleep(backoff);
        backoff = backoff < MAX_BACKOFF ? backoff * 2 + (gwave & 3) : MAX_BACKOFF - (gwave & 63);
      }
      continue;
    }

    // ---- choose a ring: the fullest of the anni / comm / other rings if any
    // holds redexes (they are cheap and release the work behind them, and
    // definition calls meanwhile gather into fuller iterations), else the
    // fullest definition-call ring; lowest index on ties
    u32 key = lane < NCLS ? ((lane < CLS_CALL && my_n > 0) ? (1u << 24) : 0u) | (my_n << 4) | (15 - lane) : 0;
#pragma unroll
    for (u32 o = 1; o < 16; o <<= 1) {
      u32 other = __shfl_xor(key, o, 64);
      key = other > key ? other : key;
    }
    key = uniform(key);
    const u32 c = 15 - (key & 15);
    const u32 nc = (key >> 4) & 0xFFFFF;
    const u32 top = uniform(__shfl(my_top, c, 64));
    const u32 bot = uniform(__shfl(my_bot, c, 64));

    // ---- budget for reduce(limit) (net.rs:41-52)
    u32 take = nc < 64 ? nc : 64;
    if (limited) {
      u32 granted = 0;
      if (lane == 0) {
        u64 d = gld(&g->done);
        for (;;) {
          if (d >= P.limit) {
            granted = 0;
            break;
          }
          u64 want = P.limit - d < take ? P.limit - d : take;
          if (__hip_atomic_compare_exchange_strong(&g->done, &d, d + want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)) {
            granted = (u32)want;
            break;
          }
        }
      }
      take = uniform(__shfl(granted, 0, 64));
      if (take == 0) {
        // budget spent: hand every ring back to the global queue and stop
        u32 qerr = 0;
        for (u32 cc = 0; cc < NCLS; cc++) {
          u32 t2 = uniform(__shfl(my_top, cc, 64)), b2 = uniform(__shfl(my_bot, cc, 64));
          if (t2 != b2) S_cache = spill_to_global(P, bag, &bot_c[cc], cc, t2 - b2, lane, b2, qerr);
        }
        if (qerr) fail(L, qerr);
        if (lane == 0 && active) __hip_atomic_fetch_sub(&g->S, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        active = false;
        break;
      }
    }

    // ---- pop (LIFO) and interact
    const bool timing = TIMING;
    const u64 t_it = timing ? wall_clock64() : 0;
    if (TIMING) st[ST_BUSY_IT]++;
    const u32 nth = (lane - rot) & 63;  // this lane runs the nth-newest redex
    bool has = nth < take;
    u64 ra = 0, rb = 0;
    if (has) {
      u32 i = (ring_base(c) + ((top - 1 - nth) & (ring_cap(c) - 1))) * 2;
      ra = bag[i];
      rb = bag[i + 1];
    }
    rot = (rot + take) & 63;
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) top_c[c] = top - take;
    __builtin_amdgcn_wave_barrier();
    const u64 t_in = timing ? wall_clock64() : 0;
    if (has) L.rw[RW_POPS]++;
    run_task(L, has, ra, rb, NO_DEF, 0);
    {
      const u64 dt = timing ? wall_clock64() - t_in : 0;
      if (TIMING) st[ST_INTERACT_TICKS] += dt;
      // c is wave-uniform; unrolled so st stays in registers
      const u32 sk = ring_stat(c);
#pragma unroll
      for (u32 k = 0; k < 4; k++)
        if (sk == k) {
          if (TIMING) st[ST_CLS_IT + k]++;
          if (TIMING) st[ST_CLS_TICKS + k] += dt;
        }
    }
    if (++busy >= REFRESH) {  // in flight while this iteration's stores retire
      busy = 0;
      fresh = true;
      if (lane == 0) {
        S_next = gld(&g->S);
        err_next = gld32(&g->err);
      }
    }
    // No fence here: a node this wave created is visible to other waves only
    // through a link (flush_links publishes before its swaps) or a spill
    // (spill_to_global publishes before writing the global ring).
    __builtin_amdgcn_wave_barrier();

    // ---- clamp the rings after overflow pushes, then spill the excess
    my_top = lane < NCLS ? top_c[lane] : 0;
    my_bot = lane < NCLS ? bot_c[lane] : 0;
    if (lane < NCLS && my_top - my_bot > ring_cap(lane)) {
      my_top = my_bot + ring_cap(lane);
      top_c[lane] = my_top;
    }
    my_n = my_top - my_bot;
    n = my_n;
#pragma unroll
    for (u32 o = 1; o < 16; o <<= 1) n += __shfl_xor(n, o, 64);
    n = uniform(n);
    key = lane < NCLS ? (my_n << 4) | (15 - lane) : 0;
#pragma unroll
    for (u32 o = 1; o < 16; o <<= 1) {
      u32 other = __shfl_xor(key, o, 64);
      key = other > key ? other : key;
    }
    key = uniform(key);
    const u32 sc = 15 - (key & 15);
    const u32 snc = key >> 4;
    u32 active_waves = (u32)(S_cache & 0xFFFFFFFFu);
    u32 idle_waves = P.total_waves > active_waves ? P.total_waves - active_waves : 0;
    u32 queued = (u32)(S_cache >> 32);
    u32 spill = 0;
    if (snc > ring_cap(sc) * 3 / 4) spill = snc - ring_cap(sc) / 2;  // keep room for one iteration's pushes
    else if (n > 64 + 16 && idle_waves > 0 && queued < idle_waves * 64) spill = snc < n - 64 ? snc : n - 64;
    // with idle waves waiting on their mailboxes while the global ring is
    // well stocked (the rule above did not fire), a wave holding more than
    // MBOX_KEEP redexes hands the excess to one of them; keeping more makes
    // fuller iterations (batched mix: 29% faster keeping 256 than 80)
    if (use_mbox && idle_waves > 0 && n > mbox_keep && spill == 0) spill = snc < n - mbox_keep ? snc : n - mbox_keep;
    const u32 spill0 = spill;
    if (spill && use_mbox && idle_waves > 0) {
      // hand up to MBOX_PAIRS of the excess to one idle wave
      const u32 give = spill < MBOX_PAIRS ? spill : MBOX_PAIRS;
      const u32 start = (gwave * 7u) % P.idle_words;
      u32 wi = start + lane;
      wi = wi >= P.idle_words ? wi - P.idle_words : wi;
      const u64 w = lane < P.idle_words ? gld(&P.idle_bits[wi]) : 0;
      const u64 ball = __ballot(w != 0);
      if (ball) {
        const u32 first = (u32)__builtin_ctzll(ball);
        const u32 word = uniform(__shfl(wi, first, 64));
        const u64 bits = __shfl(w, first, 64);
        const u32 r = gwave & 63;
        const u64 rot = (bits >> r) | (r ? bits << (64 - r) : 0);
        const u32 bit = ((u32)__builtin_ctzll(rot) + r) & 63;
        u64 old = 0;
        if (lane == 0)
          old = __hip_atomic_fetch_and(&P.idle_bits[word], ~(1ull << bit), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool got = ((__shfl(old, 0, 64) >> bit) & 1) != 0;
        const u32 target = word * 64 + bit;
        if (got && target < P.total_waves) {
          if (lane == 0) __hip_atomic_fetch_add(&g->S, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          publish();  // the nodes these redexes reach, before the mailbox entries
          u64 *box = P.mbox + (u64)target * MBOX_WORDS;
          const u32 bot0 = uniform(__shfl(my_bot, sc, 64));
          if (lane < give) {
            const u32 s2 = (ring_base(sc) + ((bot0 + lane) & (ring_cap(sc) - 1))) * 2;
            gst(box + 16 + 2 * lane, bag[s2]);
            gst(box + 16 + 2 * lane + 1, bag[s2 + 1]);
          }
          publish();
          if (lane == 0) {
            gst(box, (u64)give);
            bot_c[sc] = bot0 + give;
          }
          if (TIMING) st[ST_SPILLS]++;
          if (TIMING) st[ST_SPILLED] += give;
          spill = 0;
        }
      }
    }
    if (spill) {
      const u64 t_sp = timing ? wall_clock64() : 0;
      if (TIMING) st[ST_SPILLS]++;
      if (TIMING) st[ST_SPILLED] += spill;
      u32 qerr = 0;
      S_cache = spill_to_global(P, bag, &bot_c[sc], sc, spill, lane, uniform(__shfl(my_bot, sc, 64)), qerr);
      if (qerr) fail(L, qerr);
      if (timing) st[ST_SPILL_TICKS] += wall_clock64() - t_sp;
    }
    // ---- batched nets: a busy wave short of work also claims from the
    // global ring, while few waves are idle.  With n < 64 pending, every
    // ring has room for the 64 - n it may take.
    if (P.fill && !limited && spill0 == 0 && n < fill_below && (S_cache >> 32) > 0 &&
        idle_waves * 8 < P.total_waves) {
      u32 k = 0;
      u64 s_new = S_cache;
      if (lane == 0) {
        u64 s = S_cache;
        for (int tries = 0; tries < 2; tries++) {
          const u64 q = s >> 32;
          if (q == 0) break;
          const u64 room = 64 - n;
          const u64 want = q < room ? q : room;
          const u64 ns = s - (want << 32);
          if (__hip_atomic_compare_exchange_strong(&g->S, &s, ns, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)) {
            k = (u32)want;
            s_new = ns;
            break;
          }
          s_new = s;
        }
      }
      k = uniform(__shfl(k, 0, 64));
      S_cache = __shfl(s_new, 0, 64);
      if (k > 0) {
        if (TIMING) st[ST_CLAIMS]++;
        if (TIMING) st[ST_CLAIMED] += k;
        claim_global(P, L, bag, top_c, k, lane, t_end);
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (timing) st[ST_BUSY_TICKS] += wall_clock64() - t_it;
  }

  lane_finish(L);
  publish();
  wave_flush_counters(L, lane);
  __builtin_amdgcn_wave_barrier();
  if (lane == 0) {
    for (u32 k = 0; k < ST_SLOTS; k++)
      if (st[k]) __hip_atomic_fetch_add(&g->st[k], st[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Serialise the tree hanging from `root_cell` in preorder (readback.rs:23-87
// walk order: aux port 1 before aux port 2), then the two sides of each
// pending redex in global-ring slots [q_head, q_tail) (after a reduce(limit),
// every pending redex is there).  Principal ports are emitted as they are; a
// wire between two aux ports is emitted as Var(cell) of the var cell that
// joins them.  g->result = number of entries, or ERR_READBACK (err_info 1:
// `work` or `out` too small; 2: a dangling or cyclic wire).
//
// One block, level by level, so a wide tree is read with its lanes in
// parallel:
//  1. breadth first: the lanes of a level resolve its entries (following
//     parked ports) and append each binary node's two children, aux port 1
//     first, as a pair;
//  2. bottom up: the size of each entry's subtree;
//  3. top down: each entry's preorder position (the trees one after another;
//     a node's first child right after it, its second child after the first
//     child's subtree), where its port is written.
// `work` holds four arrays of M = (work_cap - 1) / 4 words: ports, first
// child (NO_KID: none), sizes (then positions), and the levels' first
// entries.
static const u32 RB_THREADS = 256;
static const u64 NO_KID = ~0ull;

__global__ __launch_bounds__(RB_THREADS) void k_readback(Params P, u64 root_cell, u64 q_head, u64 q_tail, u64 *work,
                                                         u64 work_cap, u64 *out, u64 out_cap) {
  const u32 t = threadIdx.x;
  const u64 M = (work_cap - 1) / 4;
  u64 *const port = work, *const kid = work + M, *const sz = work + 2 * M, *const lev = work + 3 * M;
  __shared__ unsigned long long s_next;  // entries appended so far
  __shared__ u32 s_err;
  __shared__ u64 s_n;
  const u64 nroots = 1 + 2 * (q_tail - q_head);
  if (t == 0) {
    s_next = nroots;
    s_err = nroots > M ? 1 : 0;
  }
  __syncthreads();
  if (s_err) {
    if (t == 0) {
      P.g->err_info = 1;
      P.g->err = ERR_READBACK;
      P.g->result = 0;
    }
    return;
  }
  for (u64 k = t; k < nroots; k += RB_THREADS) {
    u64 p;
    if (k == 0) {
      p = var_port(root_cell);
    } else {
      const u64 *slot = P.queue + ((q_head + (k - 1) / 2) & P.qmask) * 2;
      p = gld((u64 *)slot + ((k - 1) & 1));
    }
    HVM_ST(&port[k], p);
  }
  __syncthreads();
  // 1. breadth first
  u64 a = 0, b = nroots, L = 0;
  while (a < b) {
    if (t == 0) HVM_ST(&lev[L], a);
    for (u64 i = a + t; i < b; i += RB_THREADS) {
      u64 p = HVM_LD(&port[i]);
      u64 hops = 0;
      bool bad = false;
      while (is_endpoint(p)) {
        const u64 v = ld(P, paddr(p));
        if (v == P_NONE) break;
        if (v == P_FREE || ++hops > (1ull << 24)) {
          bad = true;
          break;
        }
        p = v;
      }
      u64 k = NO_KID;
      if (bad) {
        atomicMax(&s_err, 2u);
      } else if (is_endpoint(p)) {
        p = var_port(paddr(p));
      } else if (is_binary(p)) {
        k = atomicAdd(&s_next, 2ull);
        if (k + 2 > M) {
          atomicMax(&s_err, 1u);
          k = NO_KID;
        } else {
          const u64 ad = paddr(p);
          HVM_ST(&port[k], ld(P, ad));
          HVM_ST(&port[k + 1], ld(P, ad + 8));
        }
      }
      HVM_ST(&port[i], p);
      HVM_ST(&kid[i], k);
    }
    __syncthreads();
    if (s_err) break;
    a = b;
    b = s_next;
    L++;
    __syncthreads();
  }
  if (s_err) {
    if (t == 0) {
      P.g->err_info = s_err;
      P.g->err = ERR_READBACK;
      P.g->result = 0;
    }
    return;
  }
  if (t == 0) HVM_ST(&lev[L], a);  // a: every entry
  __syncthreads();
  // 2. bottom up: subtree sizes
  for (u64 l = L; l-- > 0;) {
    const u64 s0 = HVM_LD(&lev[l]), s1 = HVM_LD(&lev[l + 1]);
    for (u64 i = s0 + t; i < s1; i += RB_THREADS) {
      const u64 k = HVM_LD(&kid[i]);
      u64 z = 1;
      if (k != NO_KID) z += HVM_LD(&sz[k]) + HVM_LD(&sz[k + 1]);
      HVM_ST(&sz[i], z);
    }
    __syncthreads();
  }
  // 3. top down: positions (the trees in order), each port to its place
  if (t == 0) {
    u64 off = 0;
    for (u64 r = 0; r < nroots; r++) {
      const u64 z = HVM_LD(&sz[r]);
      HVM_ST(&sz[r], off);
      off += z;
    }
    s_n = off;
    s_err = off > out_cap ? 1 : 0;
  }
  __syncthreads();
  if (s_err) {
    if (t == 0) {
      P.g->err_info = 1;
      P.g->err = ERR_READBACK;
      P.g->result = s_n;
    }
    return;
  }
  for (u64 l = 0; l < L; l++) {
    const u64 s0 = HVM_LD(&lev[l]), s1 = HVM_LD(&lev[l + 1]);
    for (u64 i = s0 + t; i < s1; i += RB_THREADS) {
      const u64 pos = HVM_LD(&sz[i]);
      HVM_ST(&out[pos], HVM_LD(&port[i]));
      const u64 k = HVM_LD(&kid[i]);
      if (k != NO_KID) {
        const u64 zk = HVM_LD(&sz[k]);  // still the first child's size
        HVM_ST(&sz[k], pos + 1);
        HVM_ST(&sz[k + 1], pos + 1 + zk);
      }
    }
    __syncthreads();
  }
  if (t == 0) P.g->result = s_n;
}

// ---------------------------------------------------------------------------
// staging of the allocation-free batch boot (hvm_gpu_boot_batch_noalloc)
// ---------------------------------------------------------------------------
// NOALLOC_NETS root words, then NOALLOC_NETS 32-bit definition ids (in
// NOALLOC_NETS / 2 words): 768 KiB, held by every GPU the library runs on
// until the process ends.  Host code never passes `scratch` itself to a
// kernel (on the host it is a stand-in, not a device address): it uses the
// address hipGetSymbolAddress reports for the current device.
static const u32 NOALLOC_NETS = 1u << 16;
static __device__ __align__(256) u64 scratch[NOALLOC_NETS + NOALLOC_NETS / 2];
static const int NOALLOC_MAX_DEVICES = 64;
static std::atomic<u64 *> scratch_ptr[NOALLOC_MAX_DEVICES];

// The device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address).
hipError_t noalloc_scratch(u64 **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  u64 *p = scratch_ptr[dev].load();
  if (!p) {
    void *addr = nullptr;
    e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    p = (u64 *)addr;
    scratch_ptr[dev].store(p);
  }
  *out = p;
  return hipSuccess;
}
u32 noalloc_nets() { return NOALLOC_NETS; }

// ---------------------------------------------------------------------------
// launch wrappers used by the host engine
// ---------------------------------------------------------------------------
hipError_t launch_fill(u64 *words, u64 n, u64 value, hipStream_t st) {
  u64 blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_fill, dim3((u32)blocks), dim3(256), 0, st, words, n, value);
  return hipGetLastError();
}
hipError_t launch_exec(const Params &P, u32 def, u64 t0, u32 count_dref, hipStream_t st) {
  hipLaunchKernelGGL(k_exec, dim3(1), dim3(64), 0, st, P, def, t0, count_dref);
  return hipGetLastError();
}
hipError_t launch_expand(const Params &P, u64 old_root_cell, hipStream_t st) {
  hipLaunchKernelGGL(k_expand, dim3(1), dim3(64), 0, st, P, old_root_cell);
  return hipGetLastError();
}
hipError_t launch_boot_batch(const Params &P, u32 n, const u32 *defs, u64 *roots, u32 max_blocks, hipStream_t st) {
  u32 blocks = (n + 63) / 64;
  if (blocks > max_blocks) blocks = max_blocks;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_boot_batch, dim3(blocks), dim3(64), 0, st, P, n, defs, roots);
  return hipGetLastError();
}
hipError_t launch_expand_batch(const Params &P, u32 n, u64 *roots, u32 max_blocks, hipStream_t st) {
  u32 blocks = (n + 63) / 64;
  if (blocks > max_blocks) blocks = max_blocks;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_expand_batch, dim3(blocks), dim3(64), 0, st, P, n, roots);
  return hipGetLastError();
}
hipError_t launch_batch_reset(const Params &P, u64 pool_start, u64 ov_start, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1024), dim3(256), 0, st, P, pool_start, ov_start);
  return hipGetLastError();
}
hipError_t launch_root_ports(const Params &P, u32 n, const u64 *roots, u64 *out, hipStream_t st) {
  u32 blocks = (n + 255) / 256;
  if (blocks == 0) blocks = 1;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_root_ports, dim3(blocks), dim3(256), 0, st, P, n, roots, out);
  return hipGetLastError();
}
hipError_t launch_reduce(const Params &P, u32 blocks, hipStream_t st) {
  const dim3 grid(blocks), block(WAVES_PER_BLOCK * 64);
  // the LDS instantiation when the book and every trg register fit there
  const bool lds = P.book_lds && P.t_lds;
#define HVM_REDUCE(A, S, T)                                                     \
  do {                                                                          \
    if (lds) hipLaunchKernelGGL((k_reduce<A, S, T, true>), grid, block, 0, st, P); \
    else hipLaunchKernelGGL((k_reduce<A, S, T, false>), grid, block, 0, st, P); \
  } while (0)
  if (P.serve) HVM_REDUCE(true, true, true);
  else if (P.gmap) HVM_REDUCE(true, false, true);
  else if (P.timing) HVM_REDUCE(false, false, true);
  else HVM_REDUCE(false, false, false);
#undef HVM_REDUCE
  return hipGetLastError();
}
hipError_t launch_readback(const Params &P, u64 root_cell, u64 q_head, u64 q_tail, u64 *stack, u64 stack_cap,
                           u64 *out, u64 out_cap, hipStream_t st) {
  hipLaunchKernelGGL(k_readback, dim3(1), dim3(RB_THREADS), 0, st, P, root_cell, q_head, q_tail, stack, stack_cap,
                     out, out_cap);
  return hipGetLastError();
}
u32 reduce_waves_per_block() { return WAVES_PER_BLOCK; }
u32 reduce_book_lds_words() { return BOOK_LDS_WORDS; }
u32 reduce_lds_slots() { return LDS_SLOTS; }
hipError_t reduce_occupancy(int *blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_reduce<false, false, false, true>, WAVES_PER_BLOCK * 64, 0);
}

}  // namespace hvm
