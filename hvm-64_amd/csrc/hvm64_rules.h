// Lane-level reduction rules of the engine (see hvm64_kernels.hip for the
// memory model and the linking protocol).  Compiled for gfx950 by the
// kernels; the memory primitives are the macros below.
#pragma once
#include "hvm64_engine.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define DEV __device__ __forceinline__
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
static const u32 BAG_CAP = 256;  // redexes per wave in LDS (power of two)

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
  // allocator: this lane owns nodes lane_node0 + (region*nodes_per_lane + k)*64 + lane
  u64 my_node0;
  u64 cursor;
  u64 nslots;
  u64 ov_cur, ov_end;  // overflow bump chunk (word indices)
  u64 spare;           // reserved second word of a var pair, or 0
  bool wrapped, ov_mode;
  // trg registers
  u64 *T;
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

// Node allocation: scan this lane's interleaved slots for a free node; after
// one full wrap, bump-allocate 64-node chunks from the shared overflow region.
DEV u64 alloc_node(Lane &L) {
  const Params &P = *L.P;
  while (!L.ov_mode) {
    if (L.cursor >= L.nslots) {
      if (L.wrapped || L.nslots == 0) {
        L.ov_mode = true;
        break;
      }
      L.wrapped = true;
      L.cursor = 0;
    }
    u64 addr = HVM_HEAP_OFF + (L.my_node0 + L.cursor * 64) * 16;
    L.cursor++;
    u64 w0 = ld(P, addr), w1 = ld(P, addr + 8);
    if (w0 == P_FREE && w1 == P_FREE) return addr;
  }
  if (L.ov_cur >= L.ov_end) {
    const u64 chunk = 128;  // words
    u64 start = HVM_FAA(&P.g->ov_top, chunk);
    if (start + chunk > P.ov_end) {
      fail(L, ERR_OOM);
      return dump_addr();
    }
    L.ov_cur = start;
    L.ov_end = start + chunk;
  }
  u64 addr = HVM_HEAP_OFF + L.ov_cur * 8;
  L.ov_cur += 2;
  return addr;
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
  L.spare = n + 8;
  return n;
}

DEV u64 new_node(Lane &L, u32 tag, u32 lab, u64 p1, u64 p2) {
  u64 n = alloc_node(L);
  st(*L.P, n, norm_endpoint(p1));
  st(*L.P, n + 8, norm_endpoint(p2));
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

// ---------------------------------------------------------------------------
// redexes
// ---------------------------------------------------------------------------
DEV void global_push1(Lane &L, u64 a, u64 b) {
  const Params &P = *L.P;
  publish();
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

// Connect two endpoints/ports.  Each step consumes one var cell, so a chain
// longer than the heap can only come from a corrupted net: stop with an error.
static const u32 MAX_CHAIN = 1u << 24;
DEV void link(Lane &L, u64 a, u64 b) {
  const Params &P = *L.P;
  for (u32 hop = 0;; hop++) {
    if (hop > MAX_CHAIN) {
      fail(L, ERR_BADPTR);
      return;
    }
    if (!is_endpoint(a)) {
      if (!is_endpoint(b)) {
        redux(L, a, b);
        return;
      }
      u64 t = a;
      a = b;
      b = t;
    }
    u64 cell = paddr(a);
    publish();
    u64 old = xchg(P, cell, norm_endpoint(b));
    if (old == P_NONE) return;
    st(P, cell, P_FREE);
    a = old;
  }
}

// Follow a wire endpoint whose other side has already linked (the
// reference's Trg::target on a wire, linker.rs:401-404).  The caller owns the
// endpoint, so once a cell holds a port the cell is ours to free.
DEV u64 resolve(Lane &L, u64 p) {
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
DEV void anni2(Lane &L, u64 a, u64 b) {
  L.rw[RW_ANNI]++;
  u64 a1, a2, b1, b2;
  take(L, a, a1, a2);
  take(L, b, b1, b2);
  link(L, a1, b1);
  link(L, a2, b2);
}

DEV void comm22(Lane &L, u64 a, u64 b) {
  L.rw[RW_COMM]++;
  u64 a1, a2, b1, b2;
  take(L, a, a1, a2);
  take(L, b, b1, b2);
  u64 v1 = alloc_var(L), v2 = alloc_var(L), v3 = alloc_var(L), v4 = alloc_var(L);
  u64 A1 = new_node(L, ptag(a), plab(a), var_port(v1), var_port(v2));
  u64 A2 = new_node(L, ptag(a), plab(a), var_port(v3), var_port(v4));
  u64 B1 = new_node(L, ptag(b), plab(b), var_port(v1), var_port(v3));
  u64 B2 = new_node(L, ptag(b), plab(b), var_port(v2), var_port(v4));
  link(L, a1, B1);
  link(L, a2, B2);
  link(L, b1, A1);
  link(L, b2, A2);
}

DEV void comm02(Lane &L, u64 a, u64 b) {
  L.rw[RW_COMM]++;
  u64 b1, b2;
  take(L, b, b1, b2);
  link(L, b1, a);
  link(L, b2, a);
}

DEV void switch_num(Lane &L, u64 a, u64 b) {  // interact.rs:201-223
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

DEV void op_num(Lane &L, u64 a, u64 b) {  // interact.rs:251-270
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

DEV u64 T_get(Lane &L, u32 s) { return L.T[(u64)s * 64]; }
DEV void T_set(Lane &L, u32 s, u64 v) { L.T[(u64)s * 64] = v; }

// instruction.rs:403-422
DEV void do_ctr(Lane &L, u32 lab, u64 trg, u64 &o1, u64 &o2) {
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
DEV void do_op(Lane &L, u32 op, u64 trg, u64 &o1, u64 &o2) {
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
DEV u64 do_op_num(Lane &L, u32 op, u64 trg, u32 rhs) {
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
DEV void do_switch(Lane &L, u64 trg, u64 &o1, u64 &o2) {
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

// def.rs:217-289
DEV void interpret(Lane &L, u32 di, u64 t0) {
  const Params &P = *L.P;
  const DefRec d = P.defs[di];
  T_set(L, 0, t0);
  const u64 *ip = P.instrs + (u64)d.ioff * 2;
  for (u32 k = 0; k < d.ilen; k++) {
    u64 w0 = ip[2 * k], w1 = ip[2 * k + 1];
    u32 kind = (u32)(w0 & 0xFF);
    u32 a = (u32)(w0 >> 16) & 0xFFFF, b = (u32)(w0 >> 32) & 0xFFFF, c = (u32)(w0 >> 48);
    u32 d4 = kind == I_WIRES ? (u32)(w1 & 0xFFFF) : 0;
    if (a >= P.trg_slots || b >= P.trg_slots || c >= P.trg_slots || d4 >= P.trg_slots) {
      fail(L, ERR_BADPTR);
      return;
    }
    switch (kind) {
      case I_CONST: T_set(L, a, w1); break;
      case I_LINK: link(L, T_get(L, a), T_get(L, b)); break;
      case I_LINKCONST: link(L, T_get(L, a), w1); break;
      case I_CTR: {
        u64 x, y;
        do_ctr(L, (u32)(w1 & 0xFFFF), T_get(L, a), x, y);
        T_set(L, b, x);
        T_set(L, c, y);
        break;
      }
      case I_OP: {
        u64 x, y;
        do_op(L, (u32)(w1 & 0xFFFF), T_get(L, a), x, y);
        T_set(L, b, x);
        T_set(L, c, y);
        break;
      }
      case I_OPNUM: T_set(L, c, do_op_num(L, (u32)(w1 & 0xFFFF), T_get(L, a), (u32)(w1 >> 32))); break;
      case I_SWITCH: {
        u64 x, y;
        do_switch(L, T_get(L, a), x, y);
        T_set(L, b, x);
        T_set(L, c, y);
        break;
      }
      case I_WIRES: {
        u64 v1 = alloc_var(L), v2 = alloc_var(L);
        T_set(L, a, var_port(v1));
        T_set(L, b, wire_port(v1));
        T_set(L, c, var_port(v2));
        T_set(L, (u32)(w1 & 0xFFFF), wire_port(v2));
        break;
      }
      default: fail(L, ERR_UNIMPL); return;
    }
  }
}

DEV u64 expand_ref(u64 cell) { return mkport(T_REF, 0xFFFF, cell); }

// net.rs:353-380.  `cell` holds the ExpandDef's `out` endpoint.
DEV void expand_call(Lane &L, u64 cell, u64 port) {
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
  link(L, p1, expand_ref(e1));
  link(L, p2, expand_ref(e2));
}

// def.rs:177-189
DEV void call(Lane &L, u64 port, u64 trg) {
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
DEV void interact(Lane &L, u64 a, u64 b) {
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
DEV void lane_init(Lane &L, const Params &P, u32 gwave, u32 lane) {
  L.P = &P;
  L.my_node0 = P.lane_node0 + (u64)gwave * P.nodes_per_lane * 64 + lane;
  L.cursor = 0;
  L.nslots = P.nodes_per_lane;
  L.wrapped = false;
  L.ov_mode = false;
  L.ov_cur = L.ov_end = 0;
  L.spare = 0;
  L.T = P.trg_buf + (u64)gwave * P.trg_slots * 64 + lane;
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
