// gfx950 reduction kernels for HVM-64 interaction nets.
//
// Hot path: runtime/src/net.rs:41-60 (Net::reduce / Net::normal) with the
// rules of runtime/src/interact.rs, the instruction interpreter of
// runtime/src/def.rs:217-289 and instruction.rs:400-500, and ExpandDef
// (net.rs:335-380).
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
// (encode.rs:116-120, counts nothing).  Redexes go to the global ring.
__global__ void k_exec(Params P, u32 def, u64 t0, u32 count_dref) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Lane L;
  lane_init(L, P, 0, 0);
  if (count_dref) L.rw[RW_DREF]++;
  interpret(L, def, t0);
  publish();
  lane_finish(L);
  for (u32 k = 0; k < RW_SLOTS; k++)
    if (L.rw[k]) __hip_atomic_fetch_add(&P.g->rw[k], (u64)L.rw[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Net::expand (net.rs:65-69): a fresh root cell R2 and an ExpandDef whose
// `out` is R2's other endpoint, linked to the old root.  Writes R2's address
// to g->result.
__global__ void k_expand(Params P, u64 old_root_cell) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Lane L;
  lane_init(L, P, 0, 0);
  u64 r2 = alloc_var(L);
  u64 e = alloc_var(L);
  st(P, e, var_port(r2));
  link(L, var_port(old_root_cell), expand_ref(e));
  publish();
  lane_finish(L);
  P.g->result = r2;
  for (u32 k = 0; k < RW_SLOTS; k++)
    if (L.rw[k]) __hip_atomic_fetch_add(&P.g->rw[k], (u64)L.rw[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The persistent reduction kernel (net.rs:41-60).  Every wave owns an LDS
// ring of redexes; each iteration its 64 lanes pop up to 64 redexes (LIFO)
// and interact them; new redexes are pushed back into the ring.  Rings spill
// their oldest half to the global ring while other waves are idle, and an
// empty wave claims work from the global ring.  S = (queued << 32) | active;
// the reduction is complete when S == 0, a state no wave can leave.
__global__ __launch_bounds__(WAVES_PER_BLOCK * 64) void k_reduce(Params P) {
  __shared__ u64 s_bag[WAVES_PER_BLOCK][BAG_CAP * 2];
  __shared__ u32 s_top[WAVES_PER_BLOCK];
  __shared__ u32 s_bot[WAVES_PER_BLOCK];
  const u32 wib = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const u32 gwave = blockIdx.x * WAVES_PER_BLOCK + wib;
  Globals *g = P.g;
  u64 *bag = s_bag[wib];
  Lane L;
  lane_init(L, P, gwave, lane);
  L.bag = bag;
  L.bag_top = &s_top[wib];
  if (lane == 0) {
    s_top[wib] = 0;
    s_bot[wib] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  bool active = false;
  const bool limited = P.limit != ~0ull;
  u32 idle_sleep = 1;
  const u64 t_end = P.timeout_ticks ? wall_clock64() + P.timeout_ticks : ~0ull;

  for (;;) {
    // ---- control words (one lane reads, the wave shares)
    u64 S = 0;
    u32 err = 0;
    if (lane == 0) {
      err = gld32(&g->err);
      S = gld(&g->S);
    }
    err = __shfl(err, 0, 64);
    S = __shfl(S, 0, 64);
    if (err) break;
    if (wall_clock64() > t_end) {
      fail(L, ERR_TIMEOUT);
      break;
    }

    u32 top = s_top[wib], bot = s_bot[wib];
    u32 n = top - bot;

    // ---- claim work from the global ring when the local ring is empty
    if (n == 0) {
      u32 k = 0;
      if (lane == 0) {
        u64 s = S;
        for (int tries = 0; tries < 4; tries++) {
          u64 q = s >> 32;
          if (q == 0) break;
          u64 want = q > 128 ? 64 : (q + 1) / 2;
          u64 ns = s - (want << 32) + (active ? 0 : 1);
          if (__hip_atomic_compare_exchange_strong(&g->S, &s, ns, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)) {
            k = (u32)want;
            break;
          }
        }
      }
      k = __shfl(k, 0, 64);
      if (k == 0) {
        if (active) {
          if (lane == 0)
            __hip_atomic_fetch_sub(&g->S, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          active = false;
        }
        if (S == 0) break;
        if (limited && gld(&g->done) >= P.limit) break;
        for (u32 z = 0; z < idle_sleep; z++) __builtin_amdgcn_s_sleep(2);
        idle_sleep = idle_sleep < 64 ? idle_sleep * 2 : 64;
        continue;
      }
      idle_sleep = 1;
      active = true;
      u64 base = 0;
      if (lane == 0) base = __hip_atomic_fetch_add(&g->q_head, (u64)k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        u32 i = ((bot + lane) & (BAG_CAP - 1)) * 2;
        bag[i] = a;
        bag[i + 1] = b;
      }
      if (lane == 0) s_top[wib] = bot + k;
      __builtin_amdgcn_wave_barrier();
      top = bot + k;
      n = k;
    }

    // ---- budget for reduce(limit) (net.rs:41-52)
    u32 take = n < 64 ? n : 64;
    if (limited) {
      u32 granted = 0;
      if (lane == 0) {
        u64 d = gld(&g->done);
        for (;;) {
          if (d >= P.limit) { granted = 0; break; }
          u64 want = P.limit - d < take ? P.limit - d : take;
          if (__hip_atomic_compare_exchange_strong(&g->done, &d, d + want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)) {
            granted = (u32)want;
            break;
          }
        }
      }
      take = __shfl(granted, 0, 64);
      if (take == 0) {
        // budget spent: hand the whole ring back to the global queue and stop
        for (u32 i = lane; i < n; i += 64) {
          u32 s = ((bot + i) & (BAG_CAP - 1)) * 2;
          u64 a = bag[s], b = bag[s + 1];
          publish();
          u64 pos = __hip_atomic_fetch_add(&g->q_tail, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          u64 *q = P.queue + (pos & P.qmask) * 2;
          gst(q, a);
          gst(q + 1, b);
        }
        publish();
        if (lane == 0) {
          s_bot[wib] = top;
          __hip_atomic_fetch_add(&g->S, ((u64)n << 32) - (active ? 1 : 0), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
        }
        active = false;
        break;
      }
    }

    // ---- pop (LIFO) and interact
    bool has = lane < take;
    u64 ra = 0, rb = 0;
    if (has) {
      u32 i = ((top - 1 - lane) & (BAG_CAP - 1)) * 2;
      ra = bag[i];
      rb = bag[i + 1];
    }
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) s_top[wib] = top - take;
    __builtin_amdgcn_wave_barrier();
    L.bag_bot = bot;
    if (has) {
      L.rw[RW_POPS]++;
      interact(L, ra, rb);
    }
    publish();
    __builtin_amdgcn_wave_barrier();

    // ---- clamp the ring after overflow pushes, then balance
    top = s_top[wib];
    if (top - bot > BAG_CAP) top = bot + BAG_CAP;
    n = top - bot;
    u32 idle_waves = P.total_waves - (u32)(S & 0xFFFFFFFFu);
    u32 queued = (u32)(S >> 32);
    u32 spill = 0;
    if (n > BAG_CAP * 3 / 4) spill = n - BAG_CAP / 2;  // keep room for one iteration's pushes
    else if (n >= 2 && idle_waves > 0 && queued < idle_waves * 2) spill = n / 2;
    if (spill) {
      u64 base = 0;
      if (lane == 0) base = __hip_atomic_fetch_add(&g->q_tail, (u64)spill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      base = __shfl(base, 0, 64);
      for (u32 i = lane; i < spill; i += 64) {
        u32 s = ((bot + i) & (BAG_CAP - 1)) * 2;
        u64 *q = P.queue + ((base + i) & P.qmask) * 2;
        gst(q, bag[s]);
        gst(q + 1, bag[s + 1]);
      }
      publish();
      if (lane == 0) {
        u64 old = __hip_atomic_fetch_add(&g->S, (u64)spill << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((old >> 32) + spill > (P.qmask + 1) / 2) fail(L, ERR_QUEUE);
      }
      bot += spill;
    }
    if (lane == 0) {
      s_top[wib] = top;
      s_bot[wib] = bot;
    }
    __builtin_amdgcn_wave_barrier();
  }

  lane_finish(L);
  publish();
  wave_flush_counters(L, lane);
}

// Serialise the tree hanging from `root_cell` in preorder (readback.rs:288-365
// walk order: aux port 1 before aux port 2).  Principal ports are emitted as
// they are; a wire between two aux ports is emitted as Var(cell) of the var
// cell that joins them.  g->result = number of entries, or ERR_READBACK.
__global__ void k_readback(Params P, u64 root_cell, u64 *stack, u64 stack_cap, u64 *out, u64 out_cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  u64 sp = 0, n = 0;
  stack[sp++] = var_port(root_cell);
  while (sp > 0) {
    u64 p = stack[--sp];
    u64 hops = 0;
    while (is_endpoint(p)) {
      u64 v = ld(P, paddr(p));
      if (v == P_NONE) break;
      if (v == P_FREE || ++hops > (1ull << 24)) {
        P.g->err_info = 2;  // dangling or cyclic wire
        P.g->err = ERR_READBACK;
        P.g->result = n;
        return;
      }
      p = v;
    }
    if (n >= out_cap) {
      P.g->err_info = 1;  // buffer too small
      P.g->err = ERR_READBACK;
      P.g->result = n;
      return;
    }
    if (is_endpoint(p)) {
      out[n++] = var_port(paddr(p));
      continue;
    }
    out[n++] = p;
    if (is_binary(p)) {
      if (sp + 2 > stack_cap) {
        P.g->err_info = 1;
        P.g->err = ERR_READBACK;
        P.g->result = n;
        return;
      }
      u64 a = paddr(p);
      stack[sp++] = ld(P, a + 8);
      stack[sp++] = ld(P, a);
    }
  }
  P.g->result = n;
}

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
hipError_t launch_reduce(const Params &P, u32 blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce, dim3(blocks), dim3(WAVES_PER_BLOCK * 64), 0, st, P);
  return hipGetLastError();
}
hipError_t launch_readback(const Params &P, u64 root_cell, u64 *stack, u64 stack_cap, u64 *out, u64 out_cap,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_readback, dim3(1), dim3(64), 0, st, P, root_cell, stack, stack_cap, out, out_cap);
  return hipGetLastError();
}
u32 reduce_waves_per_block() { return WAVES_PER_BLOCK; }
hipError_t reduce_occupancy(int *blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_reduce, WAVES_PER_BLOCK * 64, 0);
}

}  // namespace hvm
