// Lane-level reduction rules of the engine (see hvm64_kernels.hip for the
// memory model and the linking protocol).  Compiled for gfx950 by the
// kernels; the memory primitives are the macros below.
#pragma once
#include "hvm64_engine.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define DEV __device__ __forceinline__
#define NOINL __device__ __noinline__
#define HVM_LD(p) __hip_atomic_load((p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define HVM_ST(p, v) __hip_atomic_store((p), (v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define HVM_XCHG(p, v) __hip_atomic_exchange((p), (v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define HVM_FAA(p, v) __hip_atomic_fetch_add((p), (v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define HVM_CAS(p, e, v) \
  __hip_atomic_compare_exchange_strong((p), (e), (v), __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define HVM_LDS_FAA(p, v) atomicAdd((p), (v))
// Pinned host memory (serving rings): system-scope accesses.
#define HVM_LD_SYS(p) __hip_atomic_load((p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)
#define HVM_ST_SYS(p, v) __hip_atomic_store((p), (v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)
#define HVM_FENCE_SYS() __builtin_amdgcn_fence(__ATOMIC_RELEASE, "")
// A pointer known to address LDS (ds_* instructions, never flat ones).
typedef __attribute__((address_space(3))) unsigned long long lds_u64;
#define HVM_TO_LDS(p) ((lds_u64 *)(p))
// Retire every outstanding vector memory operation of this wave (vmcnt(0)).
#define HVM_PUBLISH() __builtin_amdgcn_s_waitcnt(0x0F70)
// The same wait, placed where nothing is outstanding but loads whose results
// are dead on some paths: it clears the compiler's record of them, which
// would otherwise put a vmcnt wait (that also waits for the stores issued
// since) at the top of every interpreter step.
#define HVM_SETTLE_LOADS() __builtin_amdgcn_s_waitcnt(0x0F70)
#else
#define DEV static inline
#define NOINL static
#define HVM_LD(p) __atomic_load_n((p), __ATOMIC_RELAXED)
#define HVM_ST(p, v) __atomic_store_n((p), (v), __ATOMIC_RELAXED)
#define HVM_XCHG(p, v) __atomic_exchange_n((p), (v), __ATOMIC_RELAXED)
#define HVM_FAA(p, v) __atomic_fetch_add((p), (v), __ATOMIC_RELAXED)
#define HVM_CAS(p, e, v) __atomic_compare_exchange_n((p), (e), (v), false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)
#define HVM_LDS_FAA(p, v) __atomic_fetch_add((p), (v), __ATOMIC_RELAXED)
#define HVM_LD_SYS(p) __atomic_load_n((p), __ATOMIC_RELAXED)
#define HVM_ST_SYS(p, v) __atomic_store_n((p), (v), __ATOMIC_RELAXED)
#define HVM_FENCE_SYS() __atomic_thread_fence(__ATOMIC_RELEASE)
typedef unsigned long long lds_u64;
#define HVM_TO_LDS(p) ((lds_u64 *)(p))
#define HVM_PUBLISH() ((void)0)
#define HVM_SETTLE_LOADS() ((void)0)
#endif

namespace hvm {



static const u32 WAVES_PER_BLOCK = 4;
// Every wave keeps its pending redexes in LDS, in one ring per rule class, so
// that an iteration can run one rule on all 64 lanes.
// Definition calls go to NCALL rings by definition, so that the lanes of a
// call iteration mostly run the same definition in lockstep.
enum RedexClass : u32 { CLS_ANNI = 0, CLS_COMM = 1, CLS_OTHER = 2, CLS_CALL = 3, NCALL = 4, NCLS = CLS_CALL + NCALL };
static const u32 CLS_CAP = 128;    // redexes per anni / comm / other ring (power of two)
static const u32 CALL_CAP = 128;   // redexes per definition-call ring (power of two)
static const u32 RING_PAIRS = CLS_CALL * CLS_CAP + NCALL * CALL_CAP;  // per wave
DEV u32 ring_cap(u32 c) { return c < CLS_CALL ? CLS_CAP : CALL_CAP; }
DEV u32 ring_base(u32 c) { return c < CLS_CALL ? c * CLS_CAP : CLS_CALL * CLS_CAP + (c - CLS_CALL) * CALL_CAP; }
// Slot of a ring in the per-class statistics: anni, comm, call, other.
DEV u32 ring_stat(u32 c) { return c == CLS_OTHER ? 3u : c >= CLS_CALL ? 2u : c; }
static const u32 LDS_SLOTS = 10;   // trg registers per lane kept in LDS
static const u32 LINK_CAP = 6;     // deferred links per lane (LDS)
// A book of at most BOOK_LDS_WORDS words (definition records, instructions,
// label bits) is copied into every block's LDS by the reduce kernel.
static const u32 BOOK_LDS_WORDS = 6144;

// ---------------------------------------------------------------------------
// heap access
// ---------------------------------------------------------------------------
// Every heap address is checked: a bad port (an engine bug) is redirected to
// the sink words and recorded, so it can never touch memory outside the heap.
// Records the first error of a launch; `site` says which check found it.
NOINL void report_error(Globals *g, u32 code, u32 site = 0, u64 what = 0) {
  u32 zero = 0;
  if (HVM_CAS(&g->err, &zero, code)) {
    HVM_ST(&g->err_info, site);
    HVM_ST(&g->result, what);  // the offending address or port, for the message
  }
}

DEV u64 *hw(const Params &P, u64 addr) {
  u64 off = addr - HVM_HEAP_OFF;
  if (off >= P.heap_words * 8) {
    report_error(P.g, ERR_BADPTR, BAD_HEAP_ADDR, addr);
    off = DUMP_WORD * 8;
  }
  return P.heap + (off >> 3);
}
DEV u64 ld(const Params &P, u64 addr) { return HVM_LD(hw(P, addr)); }
// Both words of a node (16-byte aligned) in one agent-scope store.  A store
// ahead of a load delays that load until the store is acknowledged (the
// vector memory counter retires in order), so a node costs one store, not
// two.  The compiler does not count this store in its waits; since the
// counter retires in order, an uncounted store only makes later waits
// longer, never shorter, and a full wait (publish) covers it.
// gfx940 and later need two wait states before a VALU instruction may
// overwrite the data registers of a store wider than 64 bits; the compiler
// does not insert them after inline asm, so the asm ends with `s_nop 1`.
DEV void st16(u64 *p, u64 w0, u64 w1) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef u64 u64x2 __attribute__((ext_vector_type(2)));
  const u64x2 d = {w0, w1};
  asm volatile("global_store_dwordx4 %0, %1, off sc1\n\ts_nop 1" ::"v"(p), "v"(d) : "memory");
#else
  HVM_ST(p, w0);
  HVM_ST(p + 1, w1);
#endif
}
DEV void st(const Params &P, u64 addr, u64 v) { HVM_ST(hw(P, addr), v); }
DEV u64 xchg(const Params &P, u64 addr, u64 v) { return HVM_XCHG(hw(P, addr), v); }
DEV u64 gld(u64 *p) { return HVM_LD(p); }
DEV void gst(u64 *p, u64 v) { HVM_ST(p, v); }
DEV u32 gld32(u32 *p) { return HVM_LD(p); }
DEV void publish() { HVM_PUBLISH(); }

DEV u64 var_port(u64 cell) { return mkport(T_VAR, 0, cell); }
DEV u64 wire_port(u64 cell) { return mkport(T_RED, 0, cell); }
// Endpoints are stored as Var ports; the Red tag only marks trg registers
// that the interpreter may resolve (the reference's Trg::wire, linker.rs:233-281).
DEV u64 norm_endpoint(u64 p) { return is_endpoint(p) ? (paddr(p) | T_VAR) : p; }

// ---------------------------------------------------------------------------
// per-lane state
// ---------------------------------------------------------------------------
struct Lane {
  const Params *P;
  // allocator: this lane's slice of its wave's ring: nodes
  // sub_node0 + 64*i for i in [0, P->wave_nodes / 64)
  u64 sub_node0;
  u64 apos;            // allocations from the slice since the heap was reset
  u64 *apos_home;      // where apos is kept between launches
  u32 slice_rest;      // allocations left to take from the overflow ring
  u32 ov_left;         // nodes left in the current overflow run
  u64 ov_cur;          // next overflow-ring position of that run
  // free cache: nodes this lane consumed, newest first (fc0 .. fc{fcn-1})
  u64 fc0, fc1, fc2, fc3;
  u32 fcn;
  // arenas: the net of the running task, and the lane's chunk of an arena
  // (nodes [a_pos, a_end) of net a_net; a_net = NO_NET: none)
  // (arenas: allocation from per-net arenas; a kernel that knows its mode
  // sets it to a constant after lane_init, and the compiler drops the other
  // mode's code from the inlined rules)
  bool arenas;
  // (serve: a serving launch, the arenas being request slots; also a
  // compile-time constant in the kernel that serves)
  bool serve;
  // the book and every definition's trg registers are in LDS: a constant in
  // the kernel instantiation for such books, which then carries no global
  // fallback reads nor the per-access branches that choose between the two
  bool lds_all;
  u32 lane_id;         // the lane's index in its wave
  u32 pushed;          // serving: redexes this task put in the wave's rings
  u32 inl_n;           // serving: redexes without a heap port waiting to run in this task
  u64 *inl;            // serving: their storage, SERVE_INL pairs
  u32 cur_net, a_net;
  u32 a_take;          // nodes of the lane's last take from a_net's arena
  // serving: the last task's change to its net's count, whose old value is
  // looked at only at the next loop pass (acct_delta 0: none)
  u32 acct_old, acct_delta, acct_net;
  u64 a_pos, a_end;
  bool dirty;          // stores of new nodes/cells not yet retired
  // links deferred by the running definition: pair i at PL[2i * 64],
  // PL[(2i + 1) * 64] (LDS; nullptr: link at once)
  u64 *PL;
  // trg registers: T[s * 64] (LDS for small definitions, else global)
  u64 *T_lds;
  u64 *T_glb;
  // LDS redex rings, RING_PAIRS pairs (nullptr: push straight to the
  // global ring); per-ring push counters and oldest entries
  u64 *bag;
  u32 *bag_top;
  const u32 *bag_bot;
  // the book in LDS (nullptr: read it from global memory): definition
  // records at 0 (4 words each), instructions at book_ins, label bits at
  // book_lab, compiled-definition records at book_cdef, templates at
  // book_tmpl
  lds_u64 *book;
  u32 book_ins, book_lab, book_cdef, book_tmpl;
  // counters
  u32 rw[RW_SLOTS];
};

DEV void fail(Lane &L, u32 code, u32 site = 0) { report_error(L.P->g, code, site); }

// The book is read from LDS (L.book) rather than from global memory.
DEV bool has_book(const Lane &L) { return L.lds_all || L.book != nullptr; }

DEV u64 dump_addr() { return HVM_HEAP_OFF + DUMP_WORD * 8; }

// Counters of rare events (interpreter shortcuts, allocation misses) go to
// the global counters at once: a per-lane register each would be kept for
// the whole launch.
DEV void count_rare(Lane &L, u32 k, u32 n) {
  if (n) HVM_FAA(&L.P->g->rw[k], (u64)n);
}

// Ring allocation.  Every wave owns a ring (a multiple of 64 nodes) and
// every lane a 1/64 slice of it (every 64th node, so lanes allocating
// together write adjacent nodes).  A slice's position counts its allocations since the heap
// was reset, so on the first lap every slot is known to be free and is taken
// without reading it.  Later laps take the next slot whose two words are
// FREE (frees write FREE in place, from any wave).  When SCAN_MISSES slots in
// a row are taken, the slice is treated as full for the next SLICE_REST
// allocations, which come from the shared overflow ring in runs of OV_RUN
// nodes (one atomic per run; a run in the overflow ring's first lap is free
// without reads, later laps check every slot).
static const u32 SCAN_MISSES = 8;
static const u32 FREE_CACHE = 4;  // nodes a lane keeps for reuse (registers)
static const u32 SLICE_REST = 256;
static const u32 OV_RUN = 64;

// The slow path takes and returns the cursor by value (the lane state stays
// in registers).
struct AllocCursor {
  u64 apos, ov_cur;
  u32 slice_rest, ov_left;
  u32 misses, ov_allocs;  // counter increments
  u64 addr;               // the node allocated (dump_addr() and err on OOM)
  u32 err;
};

// Ring slots are in the heap by construction of the layout (no bounds check).
NOINL AllocCursor alloc_slow(u64 *heap, Globals *g, u64 wave_nodes, u64 ov_node0, u64 ov_nodes, u64 sub_node0,
                             AllocCursor c) {
  const u64 n = wave_nodes / 64;
  c.err = 0;
  if (c.slice_rest == 0) {
    for (u32 miss = 0; miss < SCAN_MISSES; miss++) {
      u64 node = sub_node0 + (c.apos % n) * 64;
      c.apos++;
      u64 w0 = HVM_LD(heap + 2 * node), w1 = HVM_LD(heap + 2 * node + 1);
      if (w0 == P_FREE && w1 == P_FREE) {
        c.addr = HVM_HEAP_OFF + node * 16;
        return c;
      }
      c.misses++;
    }
    c.slice_rest = SLICE_REST;
  } else {
    c.slice_rest--;
  }
  c.ov_allocs++;
  // Overflow slots are shared: a slot is taken by swapping its first word
  // from FREE (so two lanes whose runs alias the same slot a lap apart can
  // never both take it), and kept only if its second word is FREE too.
  for (u64 miss = 0; miss <= ov_nodes; miss++) {
    if (c.ov_left == 0) {
      c.ov_cur = HVM_FAA(&g->ov_pos, (u64)OV_RUN);
      c.ov_left = OV_RUN;
    }
    u64 pos = c.ov_cur++;
    c.ov_left--;
    u64 node = ov_node0 + (pos & (ov_nodes - 1));
    u64 expect = P_FREE;
    if (HVM_CAS(heap + 2 * node, &expect, P_NONE)) {
      if (HVM_LD(heap + 2 * node + 1) == P_FREE) {
        c.addr = HVM_HEAP_OFF + node * 16;
        return c;
      }
      HVM_ST(heap + 2 * node, P_FREE);  // second word in use: not a free node
    }
  }
  c.err = ERR_OOM;
  c.addr = dump_addr();
  return c;
}

// Per-net arenas.  Net i's arena is a list of extents: its initial extent
// (arena_len[i] granules from granule arena_base0[i]) and refill extents of
// refill_of(i) granules taken from the shared pool.  One word per net,
// arena_word[i] = base granule of the current extent << 32 | nodes handed
// out of it, serves a lane's request with one atomic add.  The one lane
// whose add crosses the extent's end takes the next extent from the pool,
// names the net in the granule map for its granules and installs it; the
// lanes whose adds land past the end before it is installed take their
// nodes from the overflow region (node-granular, owned by no net).  No lane
// waits for another.  Requests never exceed one granule (the host bounds the
// chunk), so a fresh extent always holds the request that installed it.
static const u32 NO_NET = 0xFFFFFFFFu;
static const u32 ARENA_FIRST = 2;     // nodes taken right after the lane's net changed
static const u32 ARENA_MIN_REFILL = 16;  // granules of the smallest refill extent

struct ArenaChunk {
  u64 pos, end;
  u32 err, site;  // site: OOM_POOL or OOM_OVERFLOW
};

// Granules of a net's refill extents: half its initial extent (len
// granules), within [ARENA_MIN_REFILL, max_refill] (an extent must outlast
// the window in which it is replaced, or most requests would go to the
// overflow region).
DEV u64 refill_of(u64 len, u64 max_refill) {
  const u64 h = len >> 1;
  return h < ARENA_MIN_REFILL ? ARENA_MIN_REFILL : h > max_refill ? max_refill : h;
}

// Out of line, with plain arguments: a reference to the kernel's Params
// would make the compiler copy them to scratch memory.
NOINL ArenaChunk arena_take(u64 *word, u64 base0, u64 len0, u64 max_refill, u32 gs, u64 *pool_next, u64 pool_end,
                            u32 *gmap, u32 net, u64 *ov_next, u64 ov_end, u32 n) {
  ArenaChunk c;
  c.err = 0;
  c.site = 0;
  const u64 refill = refill_of(len0, max_refill);
  const u64 w = HVM_FAA(word, (u64)n);
  const u64 base = w >> 32, off = w & 0xFFFFFFFFull;
  const u64 len = (base == base0 ? len0 : refill) << gs;
  if (off + n <= len) {
    c.pos = (base << gs) + off;
    c.end = c.pos + n;
    return c;
  }
  if (off <= len) {  // this add crossed the end: the next extent is this lane's to install
    const u64 g = HVM_FAA(pool_next, refill);
    if (g + refill > pool_end) {
      c.err = ERR_OOM;
      c.site = OOM_POOL;
      c.pos = c.end = 0;
      return c;
    }
    for (u64 k = 0; k < refill; k++) HVM_ST(&gmap[g + k], net);
    HVM_ST(word, (g << 32) | n);
    c.pos = g << gs;
    c.end = c.pos + n;
    return c;
  }
  const u64 x = HVM_FAA(ov_next, (u64)n);
  if (x + n > ov_end) {
    c.err = ERR_OOM;
    c.site = OOM_OVERFLOW;
    c.pos = c.end = 0;
    return c;
  }
  c.pos = x;
  c.end = x + n;
  return c;
}

// A port that reaches a heap node or cell (and so names its net).
DEV bool reaches_heap(u64 p) { return ptag(p) >= T_OP || (ptag(p) == T_REF && addr_in_heap(paddr(p))); }

// The net a redex belongs to: the arena of the node (or ExpandDef cell) it
// reaches.  Two nilary ports allocate nothing, and a granule of the pool
// whose owner this lane does not see yet names no net: the lane keeps its
// net.
DEV u32 net_of_port(const Params &P, u64 p, u32 dflt) {
  if (!reaches_heap(p)) return dflt;
  const u64 g = ((paddr(p) - HVM_HEAP_OFF) >> 4) >> P.gran_shift;
  if (g >= P.gmap_len) {
    report_error(P.g, ERR_BADPTR, BAD_NET, p);
    return dflt;
  }
  const u32 net = HVM_LD(&P.gmap[g]);
  return net < P.n_nets ? net : dflt;
}

// Makes room for `need` consecutive nodes of the running task's net.  Right
// after the lane's net changed it takes little (a lane's next task is often
// another net's, and the rest of a chunk would stay unused); while it stays
// on one net each take doubles, up to the chunk.
DEV void arena_room(Lane &L, u32 need) {
  const bool same = L.a_net == L.cur_net;
  if (same && L.a_pos + need <= L.a_end) return;
  const Params &P = *L.P;
  u32 t = same ? L.a_take * 2 : ARENA_FIRST;
  t = t > P.arena_chunk ? P.arena_chunk : t;
  L.a_take = t;
  const u32 net = L.cur_net;
  // serving: a slot's arena is SERVE_SUB sub-arenas, each with its own word,
  // a lane taking from the one of its lane index (a serving lane takes at
  // nearly every task, and one word per slot was hammered by every lane on
  // that slot)
  const u32 wi = L.serve ? net * SERVE_SUB + (L.lane_id & (SERVE_SUB - 1)) : net;
  const ArenaChunk c = arena_take(&P.arena_word[wi], P.arena_base0[wi], P.arena_len[wi], P.refill_granules,
                                  P.gran_shift, P.pool_next, P.pool_end, P.gmap, net, P.ov_next, P.ov_end,
                                  need > t ? need : t);
  L.a_net = L.cur_net;
  L.a_pos = c.pos;
  L.a_end = c.end;
  if (c.err) {
    fail(L, c.err, c.site);
    L.a_end = L.a_pos;  // every later allocation reports again
  }
}

DEV u64 alloc_node(Lane &L) {
  if (L.arenas) {
    arena_room(L, 1);
    return L.a_pos < L.a_end ? HVM_HEAP_OFF + (L.a_pos++) * 16 : dump_addr();
  }
  if (L.fcn) {
    const u64 a = L.fc0;
    L.fc0 = L.fc1;
    L.fc1 = L.fc2;
    L.fc2 = L.fc3;
    L.fcn--;
    L.rw[RW_REUSE]++;
    return a;
  }
  const u64 n = L.P->wave_nodes / 64;
  if (L.apos < n) return HVM_HEAP_OFF + (L.sub_node0 + (L.apos++) * 64) * 16;
  AllocCursor c;
  c.apos = L.apos;
  c.ov_cur = L.ov_cur;
  c.slice_rest = L.slice_rest;
  c.ov_left = L.ov_left;
  c.misses = 0;
  c.ov_allocs = 0;
  const Params &P = *L.P;
  c = alloc_slow(P.heap, P.g, P.wave_nodes, P.ov_node0, P.ov_nodes, L.sub_node0, c);
  L.apos = c.apos;
  L.ov_cur = c.ov_cur;
  L.slice_rest = c.slice_rest;
  L.ov_left = c.ov_left;
  count_rare(L, RW_ALLOC_MISS, c.misses);
  count_rare(L, RW_OV_ALLOC, c.ov_allocs);
  if (c.err) fail(L, c.err);
  return c.addr;
}

// Two node words.  Agent-scope stores: a plain store may stay in the
// writing XCD's L2, where lanes on the other XCDs would not see it.
DEV void st2(const Params &P, u64 addr, u64 w0, u64 w1) { st16(hw(P, addr), w0, w1); }

// Both words of a published node (immutable until taken) in one 16-byte
// agent-scope load; two nodes: two loads and one wait.  The compiler does not
// track these loads, so the asm waits for them itself.
DEV void ld16x2(const u64 *p, const u64 *q, u64 &p0, u64 &p1, u64 &q0, u64 &q1) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef u64 u64x2 __attribute__((ext_vector_type(2)));
  u64x2 dp, dq;
  asm volatile("global_load_dwordx4 %0, %2, off sc1\n\tglobal_load_dwordx4 %1, %3, off sc1\n\ts_waitcnt vmcnt(0)"
               : "=&v"(dp), "=&v"(dq)
               : "v"(p), "v"(q)
               : "memory");
  p0 = dp.x;
  p1 = dp.y;
  q0 = dq.x;
  q1 = dq.y;
#else
  p0 = HVM_LD(p);
  p1 = HVM_LD(p + 1);
  q0 = HVM_LD(q);
  q1 = HVM_LD(q + 1);
#endif
}
DEV void ld16(const u64 *p, u64 &p0, u64 &p1) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef u64 u64x2 __attribute__((ext_vector_type(2)));
  u64x2 dp;
  asm volatile("global_load_dwordx4 %0, %1, off sc1\n\ts_waitcnt vmcnt(0)" : "=v"(dp) : "v"(p) : "memory");
  p0 = dp.x;
  p1 = dp.y;
#else
  p0 = HVM_LD(p);
  p1 = HVM_LD(p + 1);
#endif
}

// A fresh pair of var cells (both words of one node): cells n and n + 8.
DEV u64 alloc_vars2(Lane &L) {
  u64 n = alloc_node(L);
  st2(*L.P, n, P_NONE, P_NONE);
  L.dirty = true;
  return n;
}

// One fresh var cell; the node's other word stays FREE.
DEV u64 alloc_var1(Lane &L) {
  u64 n = alloc_node(L);
  st2(*L.P, n, P_NONE, P_FREE);
  L.dirty = true;
  return n;
}

DEV u64 new_node(Lane &L, u32 tag, u32 lab, u64 p1, u64 p2) {
  u64 n = alloc_node(L);
  st2(*L.P, n, norm_endpoint(p1), norm_endpoint(p2));
  L.dirty = true;
  return mkport(tag, lab, n);
}

// A node this lane has consumed (its words already read).  Per-net arenas
// never reuse a node, so it is simply dropped; with the allocation rings it
// goes to the lane's free cache, from which the lane's next allocations are
// taken, or, when the cache is full, it is marked FREE for the ring scan.  A
// cached node is never marked: the scan skips it as live until it is reused.
DEV void drop_node(Lane &L, u64 *w) {
  if (L.arenas) return;
  const u64 off = (u64)(w - L.P->heap);
  if (L.fcn < FREE_CACHE && off >= RESERVED_WORDS) {
    L.fc3 = L.fc2;
    L.fc2 = L.fc1;
    L.fc1 = L.fc0;
    L.fc0 = HVM_HEAP_OFF + off * 8;
    L.fcn++;
    return;
  }
  st16(w, P_FREE, P_FREE);
}

// Consume a binary node: read both aux words, then drop the node.
DEV void take(Lane &L, u64 p, u64 &x, u64 &y) {
  u64 a = paddr(p);
  u64 *w = hw(*L.P, a & ~15ull);
  ld16(w, x, y);
  drop_node(L, w);
}

// take() of two nodes, their loads in flight together.
DEV void take2(Lane &L, u64 p, u64 q, u64 &p1, u64 &p2, u64 &q1, u64 &q2) {
  u64 *wp = hw(*L.P, paddr(p) & ~15ull), *wq = hw(*L.P, paddr(q) & ~15ull);
  ld16x2(wp, wq, p1, p2, q1, q2);
  drop_node(L, wp);
  drop_node(L, wq);
}

// A var cell both of whose sides have arrived: marked FREE for the ring scan
// (arenas never reuse it).
DEV void free_cell(Lane &L, u64 c) {
  if (!L.arenas) st(*L.P, c, P_FREE);
}

// Retire this lane's stores before a port to new memory becomes visible.
DEV void publish_if_dirty(Lane &L) {
  if (L.dirty) {
    publish();
    L.dirty = false;
  }
}

// ---------------------------------------------------------------------------
// redexes
// ---------------------------------------------------------------------------
// Push one redex to the global ring (retires the lane's stores first).
NOINL void push_global(u64 *queue, u64 qmask, Globals *g, u64 a, u64 b) {
  publish();
  u64 pos = HVM_FAA(&g->q_tail, 1ull);
  u64 *s = queue + (pos & qmask) * 2;
  gst(s, a);
  gst(s + 1, b);
  publish();
  u64 old = HVM_FAA(&g->S, 1ull << 32);
  if ((old >> 32) + 1 > (qmask + 1) / 2) report_error(g, ERR_QUEUE);
}

DEV void global_push1(Lane &L, u64 a, u64 b) {
  // serving: a redex in the global ring can be taken and finished by another
  // wave before this task ends, so it is counted for its net at once
  if (L.serve) HVM_FAA(&L.P->pending[L.cur_net], 1u);
  push_global(L.P->queue, L.P->qmask, L.P->g, a, b);
  L.dirty = false;
}

// The rule a redex will take, grouped: the arms of interact() that run
// anni2, comm22, a definition call (a non-ERA Ref), and everything else.
// Calls go to the ring of the definition's group (DefRec::group, by
// instruction count), read from the book in LDS; a wave runs one call ring
// per iteration and waits for its longest definition, so definitions of
// similar length share a ring.  Without the LDS book: definition modulo NCALL.
DEV u32 redex_class(u64 a, u64 b, const lds_u64 *book, bool lds, u32 ndefs) {
  u32 at = ptag(a), bt = ptag(b);
  const bool aref = at == T_REF && a != P_ERA, bref = bt == T_REF && b != P_ERA;
  if (aref || bref) {
    const u64 addr = paddr(aref ? a : b);
    if (addr_in_heap(addr)) return CLS_CALL;
    const u32 di = def_index(addr);
    const u32 g = lds ? (di < ndefs ? (u32)book[4 * di + 3] : 0u) : di;
    return CLS_CALL + g % NCALL;
  }
  if (at == bt) {
    if (at == T_OP || at == T_SWI) return CLS_ANNI;
    if (at == T_CTR) return plab(a) == plab(b) ? CLS_ANNI : CLS_COMM;
    return CLS_OTHER;
  }
  if ((at == T_CTR && bt == T_OP) || (at == T_OP && bt == T_CTR)) return CLS_COMM;
  if ((at == T_CTR && bt == T_SWI && plab(a) != 0) || (at == T_SWI && bt == T_CTR && plab(b) != 0)) return CLS_COMM;
  return CLS_OTHER;
}

// linker.rs:70-85: skippable pairs count as an erasure and are dropped.
DEV void redux(Lane &L, u64 a, u64 b) {
  if (is_skippable(a) && is_skippable(b)) {
    L.rw[RW_ERAS]++;
    return;
  }
  if (L.serve) {
    if (!reaches_heap(a) && !reaches_heap(b)) {  // no net to count it for: run it in this task
      if (L.inl_n >= SERVE_INL) {
        fail(L, ERR_QUEUE);
        return;
      }
      L.inl[2 * L.inl_n] = a;
      L.inl[2 * L.inl_n + 1] = b;
      L.inl_n++;
      return;
    }
  }
  if (L.bag) {
    u32 c = redex_class(a, b, L.book, has_book(L), L.P->ndefs);
    u32 idx = HVM_LDS_FAA(&L.bag_top[c], 1u);
    if (idx - L.bag_bot[c] < ring_cap(c)) {
      u32 k = (ring_base(c) + (idx & (ring_cap(c) - 1))) * 2;
      L.bag[k] = a;
      L.bag[k + 1] = b;
      if (L.serve) L.pushed++;  // counted when the task ends (a full ring's redex is counted at once, below)
      return;
    }
  }
  global_push1(L, a, b);
}

// Linking.  A link step either meets two principal ports (a redex) or swaps a
// port into the var cell of an endpoint: the first endpoint to arrive leaves
// its port there; the second finds it, frees the cell and links on.
//
// Rules never link directly: they append (a, b) pairs to the lane's link
// list (LDS, LINK_CAP pairs), and the lane's task loop (run_task) links the
// whole list at one place in the code, four swaps in flight at a time.  A
// swap that finds the other endpoint's port frees the cell and appends the
// continuation link to the list.  Every continuation consumes a var cell, so
// a longer run than MAX_CHAIN rounds means a corrupted net.
static const u32 MAX_CHAIN = 1u << 24;

// Every producer checks the room it needs (interp_run: one link per
// instruction; a rule: at most four, into an empty list).
DEV void push_link(Lane &L, u32 &np, u64 a, u64 b) {
  L.PL[(u64)(2 * np) * 64] = a;
  L.PL[(u64)(2 * np + 1) * 64] = b;
  np++;
}

// First step of one link: true with `cell` set when a swap of
// norm_endpoint(b) into `cell` is due, false when the link is complete.
DEV bool link_begin(Lane &L, u64 &a, u64 &b, u64 &cell) {
  if (!is_endpoint(a)) {
    if (!is_endpoint(b)) {
      redux(L, a, b);
      return false;
    }
    u64 t = a;
    a = b;
    b = t;
  }
  cell = paddr(a);
  return true;
}

// Links every pair of the list (the list is empty afterwards).
DEV void flush_links(Lane &L, u32 &np) {
  const Params &P = *L.P;
  u64 *PL = L.PL;
  for (u32 round = 0; np > 0; round++) {
    if (round > MAX_CHAIN) {
      fail(L, ERR_BADPTR, BAD_CHAIN);
      np = 0;
      return;
    }
    // take the last (up to) four pairs
    const u32 cnt = np < 4 ? np : 4;
    np -= cnt;
    u64 a[4], b[4], c[4], o[4];
    bool p[4];
#pragma unroll
    for (u32 j = 0; j < 4; j++) {
      a[j] = j < cnt ? PL[(u64)(2 * (np + j)) * 64] : 0;
      b[j] = j < cnt ? PL[(u64)(2 * (np + j) + 1) * 64] : 0;
      c[j] = 0;
    }
#pragma unroll
    for (u32 j = 0; j < 4; j++) p[j] = j < cnt && link_begin(L, a[j], b[j], c[j]);
    if (p[0] || p[1] || p[2] || p[3]) publish_if_dirty(L);
#pragma unroll
    for (u32 j = 0; j < 4; j++) o[j] = p[j] ? xchg(P, c[j], norm_endpoint(b[j])) : P_NONE;
#pragma unroll
    for (u32 j = 0; j < 4; j++) {
      if (o[j] != P_NONE) {  // second arrival: free the cell, link on
        free_cell(L, c[j]);
        PL[(u64)(2 * np) * 64] = o[j];
        PL[(u64)(2 * np + 1) * 64] = b[j];
        np++;
      }
    }
  }
}

// Follow a wire endpoint whose other side has already linked (the
// reference's Trg::target on a wire, linker.rs:278-280).  The caller owns the
// endpoint, so once a cell holds a port the cell is ours to free.
DEV u64 resolve(Lane &L, u64 p) {
  const Params &P = *L.P;
  for (u32 hop = 0; is_endpoint(p); hop++) {
    if (hop > MAX_CHAIN) {
      fail(L, ERR_BADPTR, BAD_RESOLVE);
      return P_ERA;
    }
    u64 c = paddr(p);
    u64 v = ld(P, c);
    if (v == P_NONE) break;
    free_cell(L, c);
    p = v;
  }
  return p;
}

DEV u64 as_wire(u64 p) { return is_endpoint(p) ? wire_port(paddr(p)) : p; }

// ---------------------------------------------------------------------------
// interaction rules (interact.rs); each appends at most 4 links
// ---------------------------------------------------------------------------
DEV void anni2(Lane &L, u32 &np, u64 a, u64 b) {
  L.rw[RW_ANNI]++;
  u64 a1, a2, b1, b2;
  take2(L, a, b, a1, a2, b1, b2);
  push_link(L, np, a1, b1);
  push_link(L, np, a2, b2);
}

DEV void comm22(Lane &L, u32 &np, u64 a, u64 b) {
  L.rw[RW_COMM]++;
  u64 a1, a2, b1, b2;
  take2(L, a, b, a1, a2, b1, b2);
  if (L.arenas) arena_room(L, 6);  // the six nodes below, in one take
  u64 v1 = alloc_vars2(L), v3 = alloc_vars2(L);
  u64 v2 = v1 + 8, v4 = v3 + 8;
  u64 A1 = new_node(L, ptag(a), plab(a), var_port(v1), var_port(v2));
  u64 A2 = new_node(L, ptag(a), plab(a), var_port(v3), var_port(v4));
  u64 B1 = new_node(L, ptag(b), plab(b), var_port(v1), var_port(v3));
  u64 B2 = new_node(L, ptag(b), plab(b), var_port(v2), var_port(v4));
  push_link(L, np, a1, B1);
  push_link(L, np, a2, B2);
  push_link(L, np, b1, A1);
  push_link(L, np, b2, A2);
}

DEV void comm02(Lane &L, u32 &np, u64 a, u64 b) {
  L.rw[RW_COMM]++;
  u64 b1, b2;
  take(L, b, b1, b2);
  push_link(L, np, b1, a);
  push_link(L, np, b2, a);
}

DEV void switch_num(Lane &L, u32 &np, u64 a, u64 b) {  // interact.rs:201-223
  L.rw[RW_OPER]++;
  u64 a1, a2;
  take(L, a, a1, a2);
  u32 n = pnum(b);
  u32 v = num_tag(n) == N_U24 ? num_u24(n) : 0;
  u64 x;
  if (v == 0) {
    x = new_node(L, T_CTR, 0, a2, P_ERA);
  } else {
    u64 y = new_node(L, T_CTR, 0, numport(num_new_u24(v - 1)), a2);
    x = new_node(L, T_CTR, 0, P_ERA, y);
  }
  push_link(L, np, a1, x);
}

DEV void op_num(Lane &L, u32 &np, u64 a, u64 b) {  // interact.rs:251-270
  u32 op = plab(a);
  u64 a1, a2;
  take(L, a, a1, a2);
  a1 = resolve(L, a1);
  if (ptag(a1) == T_NUM) {
    L.rw[RW_OPER]++;
    bool div0 = false;
    u32 out = operate_binary(pnum(b), op, pnum(a1), &div0);
    if (div0) fail(L, ERR_DIV0);
    push_link(L, np, a2, numport(out));
  } else {
    u64 x = new_node(L, T_OP, N_SYM, numport(operate_unary(op, pnum(b))), a2);
    push_link(L, np, a1, x);
  }
}

// ---------------------------------------------------------------------------
// definitions: call, instruction interpreter, ExpandDef
// ---------------------------------------------------------------------------
// A definition record, from the LDS copy of the book when there is one.
DEV DefRec load_def(const Lane &L, u32 di) {
  if (has_book(L)) {
    DefRec d;
    u64 w0 = L.book[4 * di], w1 = L.book[4 * di + 1], w2 = L.book[4 * di + 2];
    d.ioff = (u32)w0;
    d.ilen = (u32)(w0 >> 32);
    d.nslots = (u32)w1;
    d.labs_off = (u32)(w1 >> 32);
    d.labs_words = (u32)w2;
    d.min_safe = (u32)(w2 >> 32);
    d.group = (u32)L.book[4 * di + 3];
    d.flags = (u32)(L.book[4 * di + 3] >> 32);
    return d;
  }
  return L.P->defs[di];
}

DEV bool labs_has(const Lane &L, const DefRec &d, u32 lab) {  // def.rs:30-37
  if (lab >= d.min_safe) return false;
  u32 idx = lab >> 6;
  if (idx >= d.labs_words) return false;
  u64 w = has_book(L) ? L.book[L.book_lab + d.labs_off + idx] : L.P->labbits[d.labs_off + idx];
  return (w >> (lab & 63)) & 1;
}

// Definition instructions (def.rs:217-289, instruction.rs:110-210).
//
// A definition runs as a sequence of steps with the same shape for every
// instruction kind, so that lanes running different definitions stay
// converged on the memory operations: (1) the trg operand is read and, if it
// is a wire, followed (Trg::target); (2) a CTR meeting a CTR of its own label
// takes that node (the annihilation fast path of do_ctr); (3) the kind's own
// work, which only allocates and stores; (4) the link it asks for, if any,
// is appended to the lane's link list.  The list is linked when it is full
// and when the definition ends.  Deferring a link inside a definition changes
// nothing a later instruction relies on: a trg read in (1) sees either the
// port the link would have parked or the unresolved wire, and both lead to
// the same net with the same rewrite counts (each fast path counts what the
// interaction it replaces counts).
struct Interp {
  u32 ipw;        // next instruction's first word in the instruction array
                  // (of the LDS copy of the book, from L.book_ins)
  lds_u64 *Tl;    // trg registers 0 .. LDS_SLOTS-1 (LDS, stride 64)
  u64 *Tg;        // trg registers LDS_SLOTS .. (global, stride 64)
  u32 left;       // instructions left; 0: no definition running
  bool t_lds;     // every trg register is in LDS (else all in global memory)
  u64 tbase;      // compiled definition: address of its first template slot
  u64 tstride;    // bytes between its slots
};

// Trg register s of this lane: in LDS when every definition of the book
// fits there (a kernel-wide choice, so the branch is uniform), else in the
// lane's global trg area.
DEV u64 t_get(const Interp &I, u32 s) { return I.t_lds ? I.Tl[s * 64] : I.Tg[(u64)s * 64]; }
DEV void t_set(Interp &I, u32 s, u64 v) {
  if (I.t_lds) I.Tl[s * 64] = v;
  else I.Tg[(u64)s * 64] = v;
}

// Two words of a node this lane has just allocated (always in the heap).
DEV void st2_new(const Params &P, u64 addr, u64 w0, u64 w1) {
  st16(P.heap + ((addr - HVM_HEAP_OFF) >> 3), w0, w1);
}

// Two fresh nodes: consecutive slice positions on the first lap.
DEV void alloc2(Lane &L, u64 &n0, u64 &n1) {
  if (!L.arenas && L.fcn == 0 && L.apos + 2 <= L.P->wave_nodes / 64) {
    n0 = HVM_HEAP_OFF + (L.sub_node0 + L.apos * 64) * 16;
    n1 = n0 + 64 * 16;
    L.apos += 2;
    return;
  }
  n0 = alloc_node(L);
  n1 = alloc_node(L);
}

// The rare instruction kinds (Op, OpNum, Switch; instruction.rs:134-203).
DEV void exec_num(Lane &L, Interp &I, u32 kind, u32 b, u32 c, u64 w1, u64 port, u32 &np) {
  const Params &P = *L.P;
  const u32 lab = (u32)(w1 & 0xFFFF);
  if (kind == I_OPNUM) {  // instruction.rs:154-172
    const u32 rhs = (u32)(w1 >> 32);
    if (ptag(port) == T_NUM) {
      L.rw[RW_OPER]++;
      bool div0 = false;
      const u32 out = operate_binary(pnum(port), lab, rhs, &div0);
      if (div0) fail(L, ERR_DIV0);
      t_set(I, c, numport(out));
    } else if (port == P_ERA) {
      count_rare(L, SC_OPNUM_ERA, 1);
      t_set(I, c, P_ERA);
    } else {
      u64 v, n;
      alloc2(L, v, n);
      st2_new(P, v, P_NONE, P_FREE);
      st2_new(P, n, numport(rhs), var_port(v));
      L.dirty = true;
      push_link(L, np, port, mkport(T_OP, lab, n));
      t_set(I, c, var_port(v));
    }
    return;
  }
  if (kind == I_OP) {  // instruction.rs:134-152
    if (ptag(port) == T_NUM) {
      u64 v, n;
      alloc2(L, v, n);
      st2_new(P, v, P_NONE, P_FREE);
      st2_new(P, n, numport(operate_unary(lab, pnum(port))), var_port(v));
      L.dirty = true;
      t_set(I, b, mkport(T_OP, N_SYM, n));
      t_set(I, c, var_port(v));
    } else if (port == P_ERA) {
      count_rare(L, SC_OP_ERA, 1);
      t_set(I, b, P_ERA);
      t_set(I, c, P_ERA);
    } else {
      u64 v, n;
      alloc2(L, v, n);
      st2_new(P, v, P_NONE, P_NONE);
      st2_new(P, n, var_port(v), var_port(v + 8));
      L.dirty = true;
      push_link(L, np, port, mkport(T_OP, lab, n));
      t_set(I, b, var_port(v));
      t_set(I, c, var_port(v + 8));
    }
    return;
  }
  // I_SWITCH, instruction.rs:174-203
  if (ptag(port) == T_NUM) {
    L.rw[RW_OPER]++;
    const u32 n = pnum(port);
    const u32 v = num_tag(n) == N_U24 ? num_u24(n) : 0;
    u64 w, x;
    alloc2(L, w, x);
    st2_new(P, w, P_NONE, P_FREE);
    if (v == 0) {
      st2_new(P, x, var_port(w), P_ERA);
      t_set(I, b, mkport(T_CTR, 0, x));
    } else {
      const u64 y = alloc_node(L);
      st2_new(P, x, numport(num_new_u24(v - 1)), var_port(w));
      st2_new(P, y, P_ERA, mkport(T_CTR, 0, x));
      t_set(I, b, mkport(T_CTR, 0, y));
    }
    L.dirty = true;
    t_set(I, c, wire_port(w));
  } else if (port == P_ERA) {
    count_rare(L, SC_SWITCH_ERA, 1);
    L.rw[RW_ERAS]++;
    t_set(I, b, P_ERA);
    t_set(I, c, P_ERA);
  } else {
    u64 v, m;
    alloc2(L, v, m);
    st2_new(P, v, P_NONE, P_NONE);
    st2_new(P, m, var_port(v), var_port(v + 8));
    L.dirty = true;
    push_link(L, np, port, mkport(T_SWI, 0, m));
    t_set(I, b, var_port(v));
    t_set(I, c, var_port(v + 8));
  }
}

// A template word spec (hvm64_engine.h) as the word it stands for.  Every
// candidate is computed and one is selected, with no branch: the lanes of a
// wave write different templates, and a branch per tag costs each word the
// exec-mask bookkeeping of every arm.  The register read uses register 0
// for words that name none (every definition has it: t0).
DEV u64 tspec(const Interp &I, u64 s) {
  const u32 t = ptag(s);
  const u64 k = paddr(s) >> 3;
  const bool var = t == T_VAR;
  const u64 addr = I.tbase + (var ? k >> 1 : k) * I.tstride + (var ? (k & 1) * 8 : 0);
  const u64 rel = (s & ~ADDR_MASK) | addr;  // a slot's port; a Var spec's label is 0, so var_port(addr)
  const bool reg = t == T_RED && s != P_NONE && s != P_FREE;
  const u64 rv = norm_endpoint(t_get(I, reg ? (u32)k : 0u));
  const bool lit = t == T_REF || t == T_NUM || (t == T_RED && !reg);
  return lit ? s : reg ? rv : rel;
}

// I_TEMPLATE: writes every slot of the running definition's template (one
// 16-byte store each), pushes its redexes and adds the counts of the
// instructions the compiler evaluated.  The slots were reserved when the
// call began (interp_begin).
DEV void run_template(Lane &L, Interp &I, u32 off) {
  const Params &P = *L.P;
  const bool lds = has_book(L);
  const u32 base = lds ? L.book_tmpl + off : off;
#define HVM_TW(i) (lds ? (u64)L.book[base + (i)] : P.tmpl[base + (i)])
  const u64 h0 = HVM_TW(0), h1 = HVM_TW(1);
  const u32 S = (u32)(h0 & 0xFF), R = (u32)((h0 >> 16) & 0xFF);
  for (u32 j = 0; j < S; j++)
    st2_new(P, I.tbase + j * I.tstride, tspec(I, HVM_TW(2 + 2 * j)), tspec(I, HVM_TW(3 + 2 * j)));
  if (S) L.dirty = true;
  for (u32 r = 0; r < R; r++) redux(L, tspec(I, HVM_TW(2 + 2 * S + 2 * r)), tspec(I, HVM_TW(3 + 2 * S + 2 * r)));
#undef HVM_TW
  if (h1) {
#pragma unroll
    for (u32 k = 0; k < 5; k++) L.rw[k] += (u32)((h1 >> (8 * k)) & 0xFF);
    if (h1 >> 40) {  // the interpreter shortcuts the compiler took (rare)
#pragma unroll
      for (u32 k = 5; k < 8; k++) count_rare(L, k, (u32)((h1 >> (8 * k)) & 0xFF));
    }
  }
}

// One instruction (instruction.rs:110-210).  The kinds that read a trg
// (Ctr first, the most common) are tested first and the others after: a
// wave executes the union of its lanes' paths, and each test of a per-lane
// kind costs every lane the exec-mask bookkeeping of its branch.
DEV void exec_step(Lane &L, Interp &I, u64 w0, u64 w1, u32 &np) {
  const Params &P = *L.P;
  const u32 kind = (u32)(w0 & 0xFF);
  const u32 a = (u32)(w0 >> 16) & 0xFFFF, b = (u32)(w0 >> 32) & 0xFFFF, c = (u32)(w0 >> 48);
  const u64 t_a = t_get(I, a);  // every kind's `a` is a register of the definition (checked at book load)
  if (kind - I_CTR <= I_SWITCH - I_CTR) {
    // Ctr / Op / OpNum / Switch: the trg, followed if it is a wire
    const u64 ta = (w0 & I_ASWIRE) ? as_wire(t_a) : t_a;
    const u64 port = ptag(ta) == T_RED ? resolve(L, ta) : ta;
    if (kind != I_CTR) {
      exec_num(L, I, kind, b, c, w1, port, np);
      return;
    }
    const u32 lab = (u32)(w1 & 0xFFFF);
    if (ptag(port) == T_CTR && plab(port) == lab) {  // do_ctr's fast path (instruction.rs:116-121)
      u64 x, y;
      take(L, port, x, y);
      L.rw[RW_ANNI]++;
      t_set(I, b, as_wire(x));
      t_set(I, c, as_wire(y));
      return;
    }
    u64 v, n;  // instruction.rs:126-131
    alloc2(L, v, n);
    st2_new(P, v, P_NONE, P_NONE);
    st2_new(P, n, var_port(v), var_port(v + 8));
    L.dirty = true;
    push_link(L, np, port, mkport(T_CTR, lab, n));
    t_set(I, b, var_port(v));
    t_set(I, c, var_port(v + 8));
    return;
  }
  if (kind == I_LINK || kind == I_LINKCONST || kind == I_LINKT) {
    push_link(L, np, t_a, kind == I_LINK ? t_get(I, b) : kind == I_LINKT ? tspec(I, w1) : w1);
    return;
  }
  if (kind == I_TEMPLATE) {
    run_template(L, I, (u32)w1);
    return;
  }
  if (kind == I_CONST) {
    t_set(I, a, w1);
    return;
  }
  // I_WIRES
  const u64 v = alloc_node(L);
  st2_new(P, v, P_NONE, P_NONE);
  L.dirty = true;
  t_set(I, a, var_port(v));
  t_set(I, b, wire_port(v));
  t_set(I, c, var_port(v + 8));
  t_set(I, (u32)(w1 & 0xFFFF), wire_port(v + 8));
}

// def.rs:217-289: start the definition whose record is `d`, with t0 = `t0`.
// The compiled-definition record of definition di (0: none, or compiled
// definitions are off).
DEV u64 load_cdef(const Lane &L, u32 di) {
  if (!L.P->compiled) return 0;
  return has_book(L) ? (u64)L.book[L.book_cdef + di] : L.P->cdefs[di];
}

// A compiled definition runs when its template's slots fit, as one block, in
// what is left of the first lap of the lane's allocation slice (consecutive
// slice positions are TSLOT_STRIDE bytes apart), or in the lane's chunk of
// the net's arena (consecutive nodes); otherwise the definition's
// own instructions run.
DEV void interp_begin(Lane &L, Interp &I, const DefRec &d, u64 t0, u64 cdef) {
  const Params &P = *L.P;
  I.Tl = HVM_TO_LDS(L.T_lds);
  I.Tg = L.T_glb;
  I.t_lds = L.lds_all || (P.t_lds && L.T_lds);
  u32 ioff = d.ioff, ilen = d.ilen, nslots = d.nslots;
  I.tbase = 0;
  I.tstride = TSLOT_STRIDE;
  if (cdef) {
    const u32 ts = cdef_tslots(cdef);
    bool fits = false;
    if (L.arenas) {  // arenas: consecutive nodes of the lane's chunk
      if (ts) arena_room(L, ts);
      fits = L.a_pos + ts <= L.a_end;
      if (fits) {
        I.tbase = HVM_HEAP_OFF + L.a_pos * 16;
        I.tstride = 16;
        L.a_pos += ts;
      }
    } else if (L.apos + ts <= P.wave_nodes / 64) {
      I.tbase = HVM_HEAP_OFF + (L.sub_node0 + L.apos * 64) * 16;
      L.apos += ts;
      fits = true;
    }
    if (fits) {
      ioff = cdef_ioff(cdef);
      ilen = cdef_ilen(cdef);
      nslots = cdef_nslots(cdef);
    }
  }
  I.ipw = ioff * 2;
  I.left = ilen;
  if (nslots > P.trg_slots) {  // the trg area is sized for the book
    fail(L, ERR_BADPTR, BAD_SLOTS);
    I.left = 0;
    return;
  }
  t_set(I, 0, t0);
}

// Runs instructions until the definition ends or the link list is full
// (every instruction appends at most one link), one per loop pass:
// fetching two per pass and stepping through them in an inner loop cost
// register copies at the inner loop's edges (bitonic 2^16 -2%).
DEV void interp_run(Lane &L, Interp &I, u32 &np) {
  HVM_SETTLE_LOADS();
  while (I.left > 0 && np < LINK_CAP) {
    u64 w0, w1;
    if (has_book(L)) {
      w0 = L.book[L.book_ins + I.ipw];
      w1 = L.book[L.book_ins + I.ipw + 1];
    } else {
      w0 = L.P->instrs[I.ipw];
      w1 = L.P->instrs[I.ipw + 1];
      HVM_SETTLE_LOADS();
    }
    exec_step(L, I, w0, w1, np);
    I.ipw += 2;
    I.left -= 1;
  }
}

DEV u64 expand_ref(u64 cell) { return mkport(T_REF, 0xFFFF, cell); }

// net.rs:82-109.  `cell` holds the ExpandDef's `out` endpoint.
DEV void expand_call(Lane &L, Interp &I, u32 &np, u64 cell, u64 port) {
  const Params &P = *L.P;
  if (ptag(port) == T_REF && port != P_ERA) {
    u64 oaddr = paddr(port);
    if (addr_in_heap(oaddr)) {
      u64 x = ld(P, cell), y = ld(P, oaddr);
      st(P, cell, P_FREE);
      st(P, oaddr, P_FREE);
      push_link(L, np, x, y);
      return;
    }
    // net.call(port, Ref(self)): the book def meets this ExpandDef, which is
    // never a Ctr, so the call always dereferences (def.rs:182-188).
    if (oaddr == 0 || def_index(oaddr) >= P.ndefs) {
      fail(L, ERR_BADPTR, BAD_DEF_EXPAND);
      return;
    }
    L.rw[RW_DREF]++;
    interp_begin(L, I, load_def(L, def_index(oaddr)), expand_ref(cell), load_cdef(L, def_index(oaddr)));
    return;
  }
  u64 out = ld(P, cell);
  st(P, cell, P_FREE);
  u32 t = ptag(port);
  if (t == T_REF || t == T_NUM || t == T_VAR || t == T_RED) {
    push_link(L, np, out, port);
    return;
  }
  u64 p1, p2;
  take(L, port, p1, p2);
  u64 v1 = alloc_vars2(L), e1 = alloc_node(L);
  u64 v2 = v1 + 8, e2 = e1 + 8;
  st2(P, e1, var_port(v1), var_port(v2));
  L.dirty = true;
  u64 n = new_node(L, t, plab(port), var_port(v1), var_port(v2));
  push_link(L, np, out, n);
  push_link(L, np, p1, expand_ref(e1));
  push_link(L, np, p2, expand_ref(e2));
}

// def.rs:177-189
// Records (port, trg) in the captured list; the host puts the list back
// among the net's pending redexes after the launch.
DEV void capture_redex(Lane &L, u64 port, u64 trg) {
  const Params &P = *L.P;
  const u64 i = HVM_FAA(&P.g->captured_n, 1ull);
  if (i >= P.captured_cap) {
    fail(L, ERR_OOM);
    return;
  }
  HVM_ST(&P.captured[2 * i], port);
  HVM_ST(&P.captured[2 * i + 1], trg);
}

DEV void call(Lane &L, Interp &I, u32 &np, u64 port, u64 trg) {
  const Params &P = *L.P;
  u64 addr = paddr(port);
  if (addr_in_heap(addr)) {  // ExpandDef: LabSet::ALL
    if (ptag(trg) == T_CTR && plab(trg) >= 0xFFFF) {
      comm02(L, np, port, trg);
      return;
    }
    L.rw[RW_DREF]++;
    expand_call(L, I, np, addr, trg);
    return;
  }
  u32 di = def_index(addr);
  if (addr == 0 || di >= P.ndefs) {
    fail(L, ERR_BADPTR, BAD_DEF_CALL);
    return;
  }
  const DefRec d = load_def(L, di);
  if (ptag(trg) == T_CTR && !labs_has(L, d, plab(trg))) {
    comm02(L, np, port, trg);
    return;
  }
  L.rw[RW_DREF]++;
  if (d.flags & DEF_INERT) {  // InertDef::call (pre_reduce.rs:108-115)
    capture_redex(L, port, trg);
    return;
  }
  interp_begin(L, I, d, trg, load_cdef(L, di));
}

// interact.rs:6-44.  The arms are tested in the reference's order; each arm
// picks a rule and which side goes first, and every rule has a single call
// site (code size: the rules are inlined once).
enum Rule : u32 { R_UNIMPL, R_CALL, R_ERAS, R_COMM22, R_ANNI, R_COMM02, R_OPNUM, R_SWITCHNUM };

DEV u32 choose_rule(u64 a, u64 b, bool &swap) {
  u32 at = ptag(a), bt = ptag(b);
  bool anil = at == T_REF || at == T_NUM, bnil = bt == T_REF || bt == T_NUM;
  swap = false;
  if (at <= T_VAR || bt <= T_VAR) return R_UNIMPL;
  if (at == T_REF && bnil && !is_skippable(a)) return R_CALL;
  if (anil && bt == T_REF && !is_skippable(b)) { swap = true; return R_CALL; }
  if (anil && bnil) return R_ERAS;
  if (at == T_CTR && bt == T_SWI && plab(a) != 0) return R_COMM22;
  if (at == T_SWI && bt == T_CTR && plab(b) != 0) return R_COMM22;
  if ((at == T_CTR && bt == T_OP) || (at == T_OP && bt == T_CTR)) return R_COMM22;
  if (at == T_CTR && bt == T_CTR && plab(a) != plab(b)) return R_COMM22;
  if (at == bt && (at == T_SWI || at == T_OP || at == T_CTR)) return R_ANNI;
  if (at == T_REF && bt == T_CTR && plab(b) >= plab(a)) return R_COMM02;
  if (at == T_CTR && bt == T_REF && plab(a) >= plab(b)) { swap = true; return R_COMM02; }
  if (at == T_NUM && bt == T_CTR) return R_COMM02;
  if (at == T_CTR && bt == T_NUM) { swap = true; return R_COMM02; }
  if (a == P_ERA) return R_COMM02;
  if (b == P_ERA) { swap = true; return R_COMM02; }
  if (at == T_REF) return R_CALL;
  if (bt == T_REF) { swap = true; return R_CALL; }
  if (at == T_OP && bt == T_NUM) return R_OPNUM;
  if (at == T_NUM && bt == T_OP) { swap = true; return R_OPNUM; }
  if (at == T_SWI && bt == T_NUM) return R_SWITCHNUM;
  if (at == T_NUM && bt == T_SWI) { swap = true; return R_SWITCHNUM; }
  return R_UNIMPL;  // Switch-Ctr(0), Op-Switch: unimplemented!() in the reference
}

DEV void interact(Lane &L, Interp &I, u32 &np, u64 a, u64 b) {
  bool swap;
  const u32 rule = choose_rule(a, b, swap);
  if (swap) {
    u64 t = a;
    a = b;
    b = t;
  }
  switch (rule) {
    case R_CALL: call(L, I, np, a, b); break;
    case R_ERAS: L.rw[RW_ERAS]++; break;
    case R_COMM22: comm22(L, np, a, b); break;
    case R_ANNI: anni2(L, np, a, b); break;
    case R_COMM02: comm02(L, np, a, b); break;
    case R_OPNUM: op_num(L, np, a, b); break;
    case R_SWITCHNUM: switch_num(L, np, a, b); break;
    default: fail(L, ERR_UNIMPL); break;
  }
}

// One lane's whole task: interact the redex (a, b) when `has`, else run
// definition `di` with t0 (when di != NO_DEF), then link everything the task
// asks for.  The loop alternates instructions and link rounds until the
// definition (if any) has ended and the link list is empty; the link code is
// shared by every rule and instruction.
static const u32 NO_DEF = 0xFFFFFFFFu;

// Serving: slot `s`'s net is normal.  Its root port (following parked
// ports) and tag go to the completion ring in host memory (the entry's
// sequence word last), then the slot is reset and handed back for admission.
DEV void serve_complete(Lane &L, u32 s) {
  const Params &P = *L.P;
  u64 p = var_port(HVM_LD(&P.slot_root[s]));
  for (u32 hop = 0; hop < MAX_CHAIN && is_endpoint(p); hop++) {
    const u64 v = ld(P, paddr(p));
    if (v == P_NONE) break;
    p = v;
  }
  u64 *const sv = P.serve_ctl;
  const u64 idx = HVM_FAA(&sv[SV_DONE_TAIL], 1ull);
  if (idx - HVM_LD_SYS(&P.host_ctl[SH_DONE_READ]) > P.done_mask) {  // the host fell behind by a whole ring
    fail(L, ERR_QUEUE);
    return;
  }
  u64 *e = P.done + (idx & P.done_mask) * 4;
  HVM_ST_SYS(e + 1, HVM_LD(&P.slot_tag[s]));
  HVM_ST_SYS(e + 2, p);
  HVM_FENCE_SYS();
  HVM_ST_SYS(e, idx + 1);
  // the slot again: its arena restarts (lanes drop their chunks at every
  // task, so no chunk of the old net is used again), then it is freed
  for (u32 k = 0; k < SERVE_SUB; k++)
    HVM_ST(&P.arena_word[s * SERVE_SUB + k], P.arena_base0[s * SERVE_SUB + k] << 32);
  publish();
  const u64 pos = HVM_FAA(&sv[SV_FREE_TAIL], 1ull);
  HVM_ST(&P.free_ring[pos % P.serve_slots], (u64)s + 1);
  publish();
  HVM_FAA(&sv[SV_FREE_COUNT], 1ull);
}

// Serving: the running task changed its net's count of redexes by `delta`
// (the redexes it put in rings, less the one it consumed); the task that
// brings the count to 0 completes the net.
// Serving, at the end of a task: the change is sent now, its old value
// looked at in serve_settle at the wave's next loop pass, so the lane does not
// wait for the atomic here.  Every path that shows this task's redexes to
// another wave first retires the wave's memory operations (this one too),
// and the wave pops its own rings only after serve_settle, so no consumer's
// decrement can overtake it.
DEV void serve_account_deferred(Lane &L, u32 delta) {
  L.acct_delta = delta;
  if (delta == 0) return;
  L.acct_net = L.cur_net;
  L.acct_old = HVM_FAA(&L.P->pending[L.cur_net], delta);
}

DEV void serve_settle(Lane &L) {
  if (L.acct_delta != 0 && L.acct_old + L.acct_delta == 0) serve_complete(L, L.acct_net);
  L.acct_delta = 0;
}

DEV void serve_account(Lane &L, u32 delta) {
  if (delta == 0) return;
  // always waited for, increments too: an increment still in flight when a
  // lane of another iteration consumes the redex could let the count reach
  // 0 early (measured: early completions, then a reused slot's stale redexes)
  const u32 old = HVM_FAA(&L.P->pending[L.cur_net], delta);
  if (old + delta == 0) serve_complete(L, L.cur_net);
}

// The instructions and link rounds of a task, then (serving) the redexes
// without a heap port it made, each interacted here in turn.
DEV void task_loop(Lane &L, Interp &I, u32 &np) {
  for (;;) {
    for (;;) {
      if (I.left > 0) interp_run(L, I, np);
      if (np > 0) flush_links(L, np);
      if (I.left == 0) break;
    }
    if (!L.serve || L.inl_n == 0) break;
    L.inl_n--;
    interact(L, I, np, L.inl[2 * L.inl_n], L.inl[2 * L.inl_n + 1]);
  }
}

DEV void run_task(Lane &L, bool has, u64 a, u64 b, u32 di, u64 t0) {
  Interp I;
  I.left = 0;
  I.ipw = 0;
  I.Tl = nullptr;
  I.Tg = nullptr;
  I.t_lds = L.lds_all;  // a constant where the kernel knows it
  u32 np = 0;
  if (has && L.arenas) L.cur_net = net_of_port(*L.P, a, net_of_port(*L.P, b, L.cur_net));
  if (L.serve) {
    L.pushed = 0;
    L.a_net = NO_NET;  // a slot's arena restarts when its net completes
  }
  if (has) interact(L, I, np, a, b);
  else if (di != NO_DEF) interp_begin(L, I, load_def(L, di), t0, load_cdef(L, di));
  task_loop(L, I, np);
  // nodes a template wrote may be reached only through the redexes it
  // pushed: retire their stores before any lane reads them
  publish_if_dirty(L);
  if (L.serve && has) serve_account_deferred(L, L.pushed - 1u);
}

// Serving: boots request `tag` (definition di) into free slot s, as
// Net::boot + the expansion of Net::normal (net.rs:30-32, 55-69) would.  The
// slot's count is held at 1 while the boot runs, so that a redex the boot
// sends to the global ring cannot complete the net early.
DEV void serve_boot(Lane &L, u32 s, u32 di, u64 tag) {
  const Params &P = *L.P;
  L.cur_net = s;
  L.a_net = NO_NET;
  HVM_FAA(&P.pending[s], 1u);
  HVM_ST(&P.slot_tag[s], tag);
  const u64 r2 = alloc_node(L);
  st2(P, r2, P_NONE, var_port(r2));
  HVM_ST(&P.slot_root[s], r2);
  const u64 r = alloc_var1(L);
  L.rw[RW_DREF]++;
  Interp I;
  I.left = 0;
  I.ipw = 0;
  I.Tl = nullptr;
  I.Tg = nullptr;
  I.t_lds = L.lds_all;  // a constant where the kernel knows it
  u32 np = 0;
  L.pushed = 0;
  interp_begin(L, I, load_def(L, di), var_port(r), load_cdef(L, di));
  task_loop(L, I, np);
  push_link(L, np, var_port(r), expand_ref(r2 + 8));
  task_loop(L, I, np);
  publish_if_dirty(L);
  publish();
  serve_account(L, L.pushed - 1u);
}

// ---------------------------------------------------------------------------
// lane setup / teardown
// ---------------------------------------------------------------------------
DEV void lane_init(Lane &L, const Params &P, u32 gwave, u32 lane) {
  L.P = &P;
  L.sub_node0 = P.ring_node0 + (u64)gwave * P.wave_nodes + lane;
  L.apos_home = &P.alloc_pos[(u64)gwave * 64 + lane];
  L.apos = *L.apos_home;
  L.slice_rest = 0;
  L.ov_left = 0;
  L.ov_cur = 0;
  L.fcn = 0;
  L.fc0 = L.fc1 = L.fc2 = L.fc3 = 0;
  L.dirty = false;
  L.arenas = P.gmap != nullptr;
  L.cur_net = 0;
  L.serve = false;  // set by the serving kernel
  L.lane_id = lane;
  L.pushed = 0;
  L.inl_n = 0;
  L.inl = P.inl ? P.inl + ((u64)gwave * 64 + lane) * SERVE_INL * 2 : nullptr;
  L.a_net = NO_NET;
  L.a_take = 0;
  L.acct_old = L.acct_delta = L.acct_net = 0;
  L.a_pos = 0;
  L.a_end = 0;
  L.T_lds = nullptr;
  L.PL = nullptr;
  L.T_glb = P.trg_buf + (u64)gwave * P.trg_slots * 64 + lane;
  L.bag = nullptr;
  L.bag_top = nullptr;
  L.bag_bot = nullptr;
  L.book = nullptr;
  L.lds_all = false;
  L.book_ins = 0;
  L.book_lab = 0;
  L.book_cdef = 0;
  L.book_tmpl = 0;
#pragma unroll
  for (u32 k = 0; k < RW_SLOTS; k++) L.rw[k] = 0;
}

DEV void lane_finish(Lane &L) { *L.apos_home = L.apos; }

}  // namespace hvm
