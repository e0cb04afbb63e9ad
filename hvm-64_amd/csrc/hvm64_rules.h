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
// Retire every outstanding vector memory operation of this wave (vmcnt(0)).
#define HVM_PUBLISH() __builtin_amdgcn_s_waitcnt(0x0F70)
#else
#define DEV static inline
#define NOINL static
#define HVM_LD(p) __atomic_load_n((p), __ATOMIC_RELAXED)
#define HVM_ST(p, v) __atomic_store_n((p), (v), __ATOMIC_RELAXED)
#define HVM_XCHG(p, v) __atomic_exchange_n((p), (v), __ATOMIC_RELAXED)
#define HVM_FAA(p, v) __atomic_fetch_add((p), (v), __ATOMIC_RELAXED)
#define HVM_CAS(p, e, v) __atomic_compare_exchange_n((p), (e), (v), false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)
#define HVM_LDS_FAA(p, v) __atomic_fetch_add((p), (v), __ATOMIC_RELAXED)
#define HVM_PUBLISH() ((void)0)
#endif

namespace hvm {



static const u32 WAVES_PER_BLOCK = 4;
static const u32 BAG_CAP = 256;    // redexes per wave in LDS (power of two)
static const u32 LDS_SLOTS = 10;   // trg registers per lane kept in LDS

// ---------------------------------------------------------------------------
// heap access
// ---------------------------------------------------------------------------
// Every heap address is checked: a bad port (an engine bug) is redirected to
// the sink words and recorded, so it can never touch memory outside the heap.
DEV u64 *hw(const Params &P, u64 addr) {
  u64 off = addr - HVM_HEAP_OFF;
  if (off >= P.heap_words * 8) {
    u32 zero = 0;
    HVM_CAS(&P.g->err, &zero, (u32)ERR_BADPTR);
    off = DUMP_WORD * 8;
  }
  return P.heap + (off >> 3);
}
DEV u64 ld(const Params &P, u64 addr) { return HVM_LD(hw(P, addr)); }
DEV void st(const Params &P, u64 addr, u64 v) { HVM_ST(hw(P, addr), v); }
DEV u64 xchg(const Params &P, u64 addr, u64 v) { return HVM_XCHG(hw(P, addr), v); }
DEV u64 gld(u64 *p) { return HVM_LD(p); }
DEV void gst(u64 *p, u64 v) { HVM_ST(p, v); }
DEV u32 gld32(u32 *p) { return HVM_LD(p); }
DEV void publish() { HVM_PUBLISH(); }

DEV u64 var_port(u64 cell) { return mkport(T_VAR, 0, cell); }
DEV u64 wire_port(u64 cell) { return mkport(T_RED, 0, cell); }
// Endpoints are stored as Var ports; the Red tag only marks trg registers
// that the interpreter may resolve (the reference's Trg::wire, linker.rs:352-405).
DEV u64 norm_endpoint(u64 p) { return is_endpoint(p) ? (paddr(p) | T_VAR) : p; }

// ---------------------------------------------------------------------------
// per-lane state
// ---------------------------------------------------------------------------
struct Lane {
  const Params *P;
  // allocator: the wave's ring of nodes [wave_node0, wave_node0 + P->wave_nodes)
  u64 wave_node0;
  u64 *alloc_ctr;      // the ring's position counter (LDS in the reduce kernel)
  u64 spare;           // reserved second word of a var pair, or 0
  bool dirty;          // stores of new nodes/cells not yet retired
  // trg registers: T[s * 64] (LDS for small definitions, else global)
  u64 *T_lds;
  u64 *T_glb;
  // LDS redex bag (nullptr: push straight to the global ring)
  u64 *bag;
  u32 *bag_top;
  u32 bag_bot;
  // counters
  u32 rw[RW_SLOTS];
};

DEV void fail(Lane &L, u32 code) {
  u32 zero = 0;
  HVM_CAS(&L.P->g->err, &zero, code);
}

DEV u64 dump_addr() { return HVM_HEAP_OFF + DUMP_WORD * 8; }

// Ring allocation.  Every wave owns a ring of 2^k nodes; the ring position
// counts allocations since the heap was reset, so on the first lap every slot
// is known to be free and is taken without reading it.  Later laps take the
// next slot whose two words are FREE (frees write FREE in place, from any
// wave).  A ring with no free slot in a full lap falls back to the shared
// overflow ring, which works the same way.
DEV u64 ring_take(Lane &L, u64 *ctr, u64 node0, u64 nodes, bool &ok) {
  const Params &P = *L.P;
  for (u64 miss = 0; miss <= nodes; miss++) {
    u64 pos = HVM_FAA(ctr, 1ull);
    u64 addr = HVM_HEAP_OFF + (node0 + (pos & (nodes - 1))) * 16;
    if (pos < nodes) {
      ok = true;
      return addr;
    }
    u64 w0 = ld(P, addr), w1 = ld(P, addr + 8);
    if (w0 == P_FREE && w1 == P_FREE) {
      ok = true;
      return addr;
    }
  }
  ok = false;
  return dump_addr();
}

NOINL u64 alloc_node(Lane &L) {
  const Params &P = *L.P;
  bool ok;
  u64 a = ring_take(L, L.alloc_ctr, L.wave_node0, P.wave_nodes, ok);
  if (ok) return a;
  a = ring_take(L, &P.g->ov_pos, P.ov_node0, P.ov_nodes, ok);
  if (ok) return a;
  fail(L, ERR_OOM);
  return dump_addr();
}

// Var cells come in pairs from one node; both words are reserved at once.
DEV u64 alloc_var(Lane &L) {
  if (L.spare) {
    u64 a = L.spare;
    L.spare = 0;
    return a;
  }
  u64 n = alloc_node(L);
  st(*L.P, n, P_NONE);
  st(*L.P, n + 8, P_NONE);
  L.dirty = true;
  L.spare = n + 8;
  return n;
}

DEV u64 new_node(Lane &L, u32 tag, u32 lab, u64 p1, u64 p2) {
  u64 n = alloc_node(L);
  st(*L.P, n, norm_endpoint(p1));
  st(*L.P, n + 8, norm_endpoint(p2));
  L.dirty = true;
  return mkport(tag, lab, n);
}

// Consume a binary node: read both aux words, then free the node.
DEV void take(Lane &L, u64 p, u64 &x, u64 &y) {
  u64 a = paddr(p);
  x = ld(*L.P, a);
  y = ld(*L.P, a + 8);
  st(*L.P, a, P_FREE);
  st(*L.P, a + 8, P_FREE);
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
NOINL void global_push1(Lane &L, u64 a, u64 b) {
  const Params &P = *L.P;
  publish_if_dirty(L);
  u64 pos = HVM_FAA(&P.g->q_tail, 1ull);
  u64 *s = P.queue + (pos & P.qmask) * 2;
  gst(s, a);
  gst(s + 1, b);
  publish();
  u64 old = HVM_FAA(&P.g->S, 1ull << 32);
  if ((old >> 32) + 1 > (P.qmask + 1) / 2) fail(L, ERR_QUEUE);
}

// linker.rs:194-209: skippable pairs count as an erasure and are dropped.
DEV void redux(Lane &L, u64 a, u64 b) {
  if (is_skippable(a) && is_skippable(b)) {
    L.rw[RW_ERAS]++;
    return;
  }
  if (L.bag) {
    u32 idx = HVM_LDS_FAA(L.bag_top, 1u);
    if (idx - L.bag_bot < BAG_CAP) {
      u32 k = (idx & (BAG_CAP - 1)) * 2;
      L.bag[k] = a;
      L.bag[k + 1] = b;
      return;
    }
  }
  global_push1(L, a, b);
}

// Linking.  A link step either meets two principal ports (a redex) or swaps a
// port into the var cell of an endpoint: the first endpoint to arrive leaves
// its port there; the second finds it, frees the cell and links on.  Each
// step consumes a var cell, so a longer chain means a corrupted net.
static const u32 MAX_CHAIN = 1u << 24;

// First step: returns true with `cell` set when a swap into `cell` of
// norm_endpoint(b) is pending, false when the link is already complete.
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

NOINL void link(Lane &L, u64 a, u64 b) {
  const Params &P = *L.P;
  for (u32 hop = 0;; hop++) {
    if (hop > MAX_CHAIN) {
      fail(L, ERR_BADPTR);
      return;
    }
    u64 cell;
    if (!link_begin(L, a, b, cell)) return;
    publish_if_dirty(L);
    u64 old = xchg(P, cell, norm_endpoint(b));
    if (old == P_NONE) return;
    st(P, cell, P_FREE);
    a = old;
  }
}

// Two / four independent links whose first swaps are in flight together.
DEV void link2(Lane &L, u64 a1, u64 b1, u64 a2, u64 b2) {
  const Params &P = *L.P;
  u64 c1 = 0, c2 = 0;
  bool p1 = link_begin(L, a1, b1, c1);
  bool p2 = link_begin(L, a2, b2, c2);
  if (p1 || p2) publish_if_dirty(L);
  u64 o1 = p1 ? xchg(P, c1, norm_endpoint(b1)) : P_NONE;
  u64 o2 = p2 ? xchg(P, c2, norm_endpoint(b2)) : P_NONE;
  if (o1 != P_NONE) {
    st(P, c1, P_FREE);
    link(L, o1, b1);
  }
  if (o2 != P_NONE) {
    st(P, c2, P_FREE);
    link(L, o2, b2);
  }
}

DEV void link4(Lane &L, u64 a1, u64 b1, u64 a2, u64 b2, u64 a3, u64 b3, u64 a4, u64 b4) {
  const Params &P = *L.P;
  u64 c1 = 0, c2 = 0, c3 = 0, c4 = 0;
  bool p1 = link_begin(L, a1, b1, c1);
  bool p2 = link_begin(L, a2, b2, c2);
  bool p3 = link_begin(L, a3, b3, c3);
  bool p4 = link_begin(L, a4, b4, c4);
  if (p1 || p2 || p3 || p4) publish_if_dirty(L);
  u64 o1 = p1 ? xchg(P, c1, norm_endpoint(b1)) : P_NONE;
  u64 o2 = p2 ? xchg(P, c2, norm_endpoint(b2)) : P_NONE;
  u64 o3 = p3 ? xchg(P, c3, norm_endpoint(b3)) : P_NONE;
  u64 o4 = p4 ? xchg(P, c4, norm_endpoint(b4)) : P_NONE;
  if (o1 != P_NONE) { st(P, c1, P_FREE); link(L, o1, b1); }
  if (o2 != P_NONE) { st(P, c2, P_FREE); link(L, o2, b2); }
  if (o3 != P_NONE) { st(P, c3, P_FREE); link(L, o3, b3); }
  if (o4 != P_NONE) { st(P, c4, P_FREE); link(L, o4, b4); }
}

// Follow a wire endpoint whose other side has already linked (the
// reference's Trg::target on a wire, linker.rs:401-404).  The caller owns the
// endpoint, so once a cell holds a port the cell is ours to free.
NOINL u64 resolve(Lane &L, u64 p) {
  const Params &P = *L.P;
  for (u32 hop = 0; is_endpoint(p); hop++) {
    if (hop > MAX_CHAIN) {
      fail(L, ERR_BADPTR);
      return P_ERA;
    }
    u64 c = paddr(p);
    u64 v = ld(P, c);
    if (v == P_NONE) break;
    st(P, c, P_FREE);
    p = v;
  }
  return p;
}

DEV u64 as_wire(u64 p) { return is_endpoint(p) ? wire_port(paddr(p)) : p; }

// ---------------------------------------------------------------------------
// interaction rules (interact.rs)
// ---------------------------------------------------------------------------
NOINL void anni2(Lane &L, u64 a, u64 b) {
  L.rw[RW_ANNI]++;
  u64 a1, a2, b1, b2;
  take(L, a, a1, a2);
  take(L, b, b1, b2);
  link2(L, a1, b1, a2, b2);
}

NOINL void comm22(Lane &L, u64 a, u64 b) {
  L.rw[RW_COMM]++;
  u64 a1, a2, b1, b2;
  take(L, a, a1, a2);
  take(L, b, b1, b2);
  u64 v1 = alloc_var(L), v2 = alloc_var(L), v3 = alloc_var(L), v4 = alloc_var(L);
  u64 A1 = new_node(L, ptag(a), plab(a), var_port(v1), var_port(v2));
  u64 A2 = new_node(L, ptag(a), plab(a), var_port(v3), var_port(v4));
  u64 B1 = new_node(L, ptag(b), plab(b), var_port(v1), var_port(v3));
  u64 B2 = new_node(L, ptag(b), plab(b), var_port(v2), var_port(v4));
  link4(L, a1, B1, a2, B2, b1, A1, b2, A2);
}

NOINL void comm02(Lane &L, u64 a, u64 b) {
  L.rw[RW_COMM]++;
  u64 b1, b2;
  take(L, b, b1, b2);
  link2(L, b1, a, b2, a);
}

NOINL void switch_num(Lane &L, u64 a, u64 b) {  // interact.rs:201-223
  L.rw[RW_OPER]++;
  u64 a1, a2;
  take(L, a, a1, a2);
  u32 n = pnum(b);
  u32 v = num_tag(n) == N_U24 ? num_u24(n) : 0;
  if (v == 0) {
    u64 x = new_node(L, T_CTR, 0, a2, P_ERA);
    link(L, a1, x);
  } else {
    u64 y = new_node(L, T_CTR, 0, numport(num_new_u24(v - 1)), a2);
    u64 x = new_node(L, T_CTR, 0, P_ERA, y);
    link(L, a1, x);
  }
}

NOINL void op_num(Lane &L, u64 a, u64 b) {  // interact.rs:251-270
  u32 op = plab(a);
  u64 a1, a2;
  take(L, a, a1, a2);
  a1 = resolve(L, a1);
  if (ptag(a1) == T_NUM) {
    L.rw[RW_OPER]++;
    bool div0 = false;
    u32 out = operate_binary(pnum(b), op, pnum(a1), &div0);
    if (div0) fail(L, ERR_DIV0);
    link(L, a2, numport(out));
  } else {
    u64 x = new_node(L, T_OP, N_SYM, numport(operate_unary(op, pnum(b))), a2);
    link(L, a1, x);
  }
}

// ---------------------------------------------------------------------------
// definitions: call, instruction interpreter, ExpandDef
// ---------------------------------------------------------------------------
DEV bool labs_has(const Params &P, const DefRec &d, u32 lab) {  // def.rs:30-37
  if (lab >= d.min_safe) return false;
  u32 idx = lab >> 6;
  if (idx >= d.labs_words) return false;
  return (P.labbits[d.labs_off + idx] >> (lab & 63)) & 1;
}


// instruction.rs:403-422
NOINL void do_ctr(Lane &L, u32 lab, u64 trg, u64 &o1, u64 &o2) {
  u64 port = ptag(trg) == T_RED ? resolve(L, trg) : trg;
  if (ptag(port) == T_CTR && plab(port) == lab) {
    u64 x, y;
    take(L, port, x, y);
    L.rw[RW_ANNI]++;
    o1 = as_wire(x);
    o2 = as_wire(y);
    return;
  }
  u64 v1 = alloc_var(L), v2 = alloc_var(L);
  u64 n = new_node(L, T_CTR, lab, var_port(v1), var_port(v2));
  link(L, port, n);
  o1 = var_port(v1);
  o2 = var_port(v2);
}

// instruction.rs:425-442
NOINL void do_op(Lane &L, u32 op, u64 trg, u64 &o1, u64 &o2) {
  u64 port = ptag(trg) == T_RED ? resolve(L, trg) : trg;
  if (ptag(port) == T_NUM) {
    u64 v = alloc_var(L);
    u64 n = new_node(L, T_OP, N_SYM, numport(operate_unary(op, pnum(port))), var_port(v));
    o1 = n;
    o2 = var_port(v);
    return;
  }
  if (port == P_ERA) {
    L.rw[SC_OP_ERA]++;
    o1 = P_ERA;
    o2 = P_ERA;
    return;
  }
  u64 v1 = alloc_var(L), v2 = alloc_var(L);
  u64 n = new_node(L, T_OP, op, var_port(v1), var_port(v2));
  link(L, port, n);
  o1 = var_port(v1);
  o2 = var_port(v2);
}

// instruction.rs:445-462
NOINL u64 do_op_num(Lane &L, u32 op, u64 trg, u32 rhs) {
  u64 port = ptag(trg) == T_RED ? resolve(L, trg) : trg;
  if (ptag(port) == T_NUM) {
    L.rw[RW_OPER]++;
    bool div0 = false;
    u32 out = operate_binary(pnum(port), op, rhs, &div0);
    if (div0) fail(L, ERR_DIV0);
    return numport(out);
  }
  if (port == P_ERA) {
    L.rw[SC_OPNUM_ERA]++;
    return P_ERA;
  }
  u64 v = alloc_var(L);
  u64 n = new_node(L, T_OP, op, numport(rhs), var_port(v));
  link(L, port, n);
  return var_port(v);
}

// instruction.rs:465-493
NOINL void do_switch(Lane &L, u64 trg, u64 &o1, u64 &o2) {
  u64 port = ptag(trg) == T_RED ? resolve(L, trg) : trg;
  if (ptag(port) == T_NUM) {
    L.rw[RW_OPER]++;
    u32 n = pnum(port);
    u32 v = num_tag(n) == N_U24 ? num_u24(n) : 0;
    u64 w = alloc_var(L);
    if (v == 0) {
      o1 = new_node(L, T_CTR, 0, var_port(w), P_ERA);
    } else {
      u64 c2 = new_node(L, T_CTR, 0, numport(num_new_u24(v - 1)), var_port(w));
      o1 = new_node(L, T_CTR, 0, P_ERA, c2);
    }
    o2 = wire_port(w);
    return;
  }
  if (port == P_ERA) {
    L.rw[SC_SWITCH_ERA]++;
    L.rw[RW_ERAS]++;
    o1 = P_ERA;
    o2 = P_ERA;
    return;
  }
  u64 v1 = alloc_var(L), v2 = alloc_var(L);
  u64 m = new_node(L, T_SWI, 0, var_port(v1), var_port(v2));
  link(L, port, m);
  o1 = var_port(v1);
  o2 = var_port(v2);
}

// def.rs:217-289.  Instructions are fetched four at a time (one 64-byte
// line); small definitions keep their trg registers in LDS.
NOINL void exec_instr(Lane &L, u64 *T, u64 w0, u64 w1);
NOINL void interpret(Lane &L, u32 di, u64 t0) {
  const Params &P = *L.P;
  const DefRec d = P.defs[di];
  u64 *T = (d.nslots <= LDS_SLOTS && L.T_lds) ? L.T_lds : L.T_glb;
  if (d.nslots > P.trg_slots && T == L.T_glb) {
    fail(L, ERR_BADPTR);
    return;
  }
  T[0] = t0;
  const u64 *ip = P.instrs + (u64)d.ioff * 2;
  for (u32 k = 0; k < d.ilen; k += 4) {
    u64 w[8];
#pragma unroll
    for (u32 j = 0; j < 4; j++) {
      bool in = k + j < d.ilen;
      w[2 * j] = in ? ip[2 * (k + j)] : 0;
      w[2 * j + 1] = in ? ip[2 * (k + j) + 1] : 0;
    }
    u32 cnt = d.ilen - k < 4 ? d.ilen - k : 4;
    for (u32 j = 0; j < cnt; j++) exec_instr(L, T, w[2 * j], w[2 * j + 1]);
  }
}

#define T_SET(s, v) (T[(u64)(s) * 64] = (v))
NOINL void exec_instr(Lane &L, u64 *T, u64 w0, u64 w1) {
  const Params &P = *L.P;
  {
    u32 kind = (u32)(w0 & 0xFF);
    u32 a = (u32)(w0 >> 16) & 0xFFFF, b = (u32)(w0 >> 32) & 0xFFFF, c = (u32)(w0 >> 48);
    u32 d4 = kind == I_WIRES ? (u32)(w1 & 0xFFFF) : 0;
    u32 lim = T == L.T_lds ? LDS_SLOTS : P.trg_slots;
    if (a >= lim || b >= lim || c >= lim || d4 >= lim) {
      fail(L, ERR_BADPTR);
      return;
    }
    switch (kind) {
      case I_CONST: T_SET(a, w1); break;
      case I_LINK: link(L, T[(u64)a * 64], T[(u64)b * 64]); break;
      case I_LINKCONST: link(L, T[(u64)a * 64], w1); break;
      case I_CTR: {
        u64 x, y;
        do_ctr(L, (u32)(w1 & 0xFFFF), T[(u64)a * 64], x, y);
        T_SET(b, x);
        T_SET(c, y);
        break;
      }
      case I_OP: {
        u64 x, y;
        do_op(L, (u32)(w1 & 0xFFFF), T[(u64)a * 64], x, y);
        T_SET(b, x);
        T_SET(c, y);
        break;
      }
      case I_OPNUM: T_SET(c, do_op_num(L, (u32)(w1 & 0xFFFF), T[(u64)a * 64], (u32)(w1 >> 32))); break;
      case I_SWITCH: {
        u64 x, y;
        do_switch(L, T[(u64)a * 64], x, y);
        T_SET(b, x);
        T_SET(c, y);
        break;
      }
      case I_WIRES: {
        u64 v1 = alloc_var(L), v2 = alloc_var(L);
        T_SET(a, var_port(v1));
        T_SET(b, wire_port(v1));
        T_SET(c, var_port(v2));
        T_SET((u32)(w1 & 0xFFFF), wire_port(v2));
        break;
      }
      default: fail(L, ERR_UNIMPL); return;
    }
  }
}
#undef T_SET

DEV u64 expand_ref(u64 cell) { return mkport(T_REF, 0xFFFF, cell); }

// net.rs:353-380.  `cell` holds the ExpandDef's `out` endpoint.
NOINL void expand_call(Lane &L, u64 cell, u64 port) {
  const Params &P = *L.P;
  if (ptag(port) == T_REF && port != P_ERA) {
    u64 oaddr = paddr(port);
    if (addr_in_heap(oaddr)) {
      u64 x = ld(P, cell), y = ld(P, oaddr);
      st(P, cell, P_FREE);
      st(P, oaddr, P_FREE);
      link(L, x, y);
      return;
    }
    // net.call(port, Ref(self)): the book def meets this ExpandDef, which is
    // never a Ctr, so the call always dereferences (def.rs:182-188).
    if (oaddr == 0 || def_index(oaddr) >= P.ndefs) {
      fail(L, ERR_BADPTR);
      return;
    }
    L.rw[RW_DREF]++;
    interpret(L, def_index(oaddr), expand_ref(cell));
    return;
  }
  u64 out = ld(P, cell);
  st(P, cell, P_FREE);
  u32 t = ptag(port);
  if (t == T_REF || t == T_NUM || t == T_VAR || t == T_RED) {
    link(L, out, port);
    return;
  }
  u64 p1, p2;
  take(L, port, p1, p2);
  u64 v1 = alloc_var(L), v2 = alloc_var(L), e1 = alloc_var(L), e2 = alloc_var(L);
  st(P, e1, var_port(v1));
  st(P, e2, var_port(v2));
  u64 n = new_node(L, t, plab(port), var_port(v1), var_port(v2));
  link(L, out, n);
  link2(L, p1, expand_ref(e1), p2, expand_ref(e2));
}

// def.rs:177-189
NOINL void call(Lane &L, u64 port, u64 trg) {
  const Params &P = *L.P;
  u64 addr = paddr(port);
  if (addr_in_heap(addr)) {  // ExpandDef: LabSet::ALL
    if (ptag(trg) == T_CTR && plab(trg) >= 0xFFFF) {
      comm02(L, port, trg);
      return;
    }
    L.rw[RW_DREF]++;
    expand_call(L, addr, trg);
    return;
  }
  u32 di = def_index(addr);
  if (addr == 0 || di >= P.ndefs) {
    fail(L, ERR_BADPTR);
    return;
  }
  const DefRec &d = P.defs[di];
  if (ptag(trg) == T_CTR && !labs_has(P, d, plab(trg))) {
    comm02(L, port, trg);
    return;
  }
  L.rw[RW_DREF]++;
  interpret(L, di, trg);
}

// interact.rs:6-44 (arm order preserved)
NOINL void interact(Lane &L, u64 a, u64 b) {
  u32 at = ptag(a), bt = ptag(b);
  bool anil = at == T_REF || at == T_NUM, bnil = bt == T_REF || bt == T_NUM;
  if (at <= T_VAR || bt <= T_VAR) { fail(L, ERR_UNIMPL); return; }
  if (at == T_REF && bnil && !is_skippable(a)) { call(L, a, b); return; }
  if (anil && bt == T_REF && !is_skippable(b)) { call(L, b, a); return; }
  if (anil && bnil) { L.rw[RW_ERAS]++; return; }
  if (at == T_CTR && bt == T_SWI && plab(a) != 0) { comm22(L, a, b); return; }
  if (at == T_SWI && bt == T_CTR && plab(b) != 0) { comm22(L, a, b); return; }
  if ((at == T_CTR && bt == T_OP) || (at == T_OP && bt == T_CTR)) { comm22(L, a, b); return; }
  if (at == T_CTR && bt == T_CTR && plab(a) != plab(b)) { comm22(L, a, b); return; }
  if (at == bt && (at == T_SWI || at == T_OP || at == T_CTR)) { anni2(L, a, b); return; }
  if (at == T_REF && bt == T_CTR && plab(b) >= plab(a)) { comm02(L, a, b); return; }
  if (at == T_CTR && bt == T_REF && plab(a) >= plab(b)) { comm02(L, b, a); return; }
  if (at == T_NUM && bt == T_CTR) { comm02(L, a, b); return; }
  if (at == T_CTR && bt == T_NUM) { comm02(L, b, a); return; }
  if (a == P_ERA) { comm02(L, a, b); return; }
  if (b == P_ERA) { comm02(L, b, a); return; }
  if (at == T_REF) { call(L, a, b); return; }
  if (bt == T_REF) { call(L, b, a); return; }
  if (at == T_OP && bt == T_NUM) { op_num(L, a, b); return; }
  if (at == T_NUM && bt == T_OP) { op_num(L, b, a); return; }
  if (at == T_SWI && bt == T_NUM) { switch_num(L, a, b); return; }
  if (at == T_NUM && bt == T_SWI) { switch_num(L, b, a); return; }
  fail(L, ERR_UNIMPL);  // Switch-Ctr(0), Op-Switch: unimplemented!() in the reference
}

// ---------------------------------------------------------------------------
// lane setup / teardown
// ---------------------------------------------------------------------------
DEV void lane_init(Lane &L, const Params &P, u32 gwave, u32 lane, u64 *alloc_ctr) {
  L.P = &P;
  L.wave_node0 = P.ring_node0 + (u64)gwave * P.wave_nodes;
  L.alloc_ctr = alloc_ctr;
  L.spare = 0;
  L.dirty = false;
  L.T_lds = nullptr;
  L.T_glb = P.trg_buf + (u64)gwave * P.trg_slots * 64 + lane;
  L.bag = nullptr;
  L.bag_top = nullptr;
  L.bag_bot = 0;
#pragma unroll
  for (u32 k = 0; k < RW_SLOTS; k++) L.rw[k] = 0;
}

DEV void lane_finish(Lane &L) {
  if (L.spare) st(*L.P, L.spare, P_FREE);
  L.spare = 0;
}

}  // namespace hvm
