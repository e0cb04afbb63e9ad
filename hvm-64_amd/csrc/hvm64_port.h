// Port / Num encodings and the 24-bit numeric semantics, shared by the host
// library and the gfx950 kernels.
//
// Port layout (runtime/src/port.rs:8-14, 120-122): bits 0-2 tag, bits 3-47
// address, bits 48-63 label.  Num ports keep a 32-bit Num in the low half
// (port.rs:53-57, num/src/num.rs:61-63): payload<<8 | numtag<<3 | 3.
//
// Addresses are byte offsets in the engine's address space rather than host
// pointers: book definition records live below HVM_HEAP_OFF ((i+1)<<5 for
// definition i, so the null address stays the eraser), heap words live at
// HVM_HEAP_OFF + 8*word_index.  The bit layout of every port is unchanged.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define HVM_HD __host__ __device__ __forceinline__
// Rarely executed, large: kept out of line on the device (code size).
#define HVM_HD_COLD __host__ __device__ inline __attribute__((noinline))
#else
#define HVM_HD inline
#define HVM_HD_COLD inline
#endif

namespace hvm {

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint16_t u16;
typedef int32_t i32;

// tags (port.rs:22-86)
enum : u32 { T_RED = 0, T_VAR = 1, T_REF = 2, T_NUM = 3, T_OP = 5, T_SWI = 6, T_CTR = 7 };

// sentinels (port.rs:105-116)
static const u64 P_ERA = 2;
static const u64 P_FREE = 0x8000000000000000ull;
static const u64 P_LOCK = 0xFFFFFFFFFFFFFFF0ull;
// An allocated var cell whose first endpoint has not arrived yet holds the
// same word the reference's allocator writes into fresh words (allocator.rs:107).
static const u64 P_NONE = P_LOCK;

static const u64 HVM_HEAP_OFF = 1ull << 32;
static const u64 ADDR_MASK = 0x0000FFFFFFFFFFF8ull;

HVM_HD u64 mkport(u32 tag, u32 lab, u64 addr) { return ((u64)lab << 48) | addr | (u64)tag; }
HVM_HD u32 ptag(u64 p) { return (u32)(p & 7); }
HVM_HD u32 plab(u64 p) { return (u32)(p >> 48); }
HVM_HD u64 paddr(u64 p) { return p & ADDR_MASK; }
HVM_HD bool is_principal(u64 p) { return ptag(p) >= T_REF; }
HVM_HD bool is_endpoint(u64 p) { return ptag(p) <= T_VAR; }  // Var or Red (wire) endpoint
HVM_HD bool is_skippable(u64 p) { return ptag(p) == T_NUM || (ptag(p) == T_REF && plab(p) != 0xFFFF); }
HVM_HD bool is_binary(u64 p) { return ptag(p) > T_NUM; }
HVM_HD u32 pnum(u64 p) { return (u32)p; }
HVM_HD u64 numport(u32 raw) { return (u64)raw; }
HVM_HD u64 def_addr(u32 idx) { return ((u64)idx + 1) << 5; }
HVM_HD u32 def_index(u64 addr) { return (u32)(addr >> 5) - 1; }
HVM_HD bool addr_in_heap(u64 addr) { return addr >= HVM_HEAP_OFF; }

// ---------------------------------------------------------------------------
// Numbers (num/src/num.rs)
// ---------------------------------------------------------------------------
enum : u32 {
  N_U24 = 0, N_I24, N_F24, N_SYM, N_ADD, N_SUB, N_SUBS, N_MUL, N_DIV, N_DIVS, N_REM, N_REMS,
  N_AND, N_OR, N_XOR, N_SHL, N_SHLS, N_SHR, N_SHRS, N_EQ, N_NE, N_LT, N_GT, N_TAGS
};

HVM_HD u32 num_make(u32 tag, u32 payload) { return (payload << 8) | (tag << 3) | 3u; }
HVM_HD u32 num_tag(u32 raw) { return (raw >> 3) & 0x1f; }
HVM_HD u32 num_u24(u32 raw) { return raw >> 8; }
HVM_HD i32 num_i24(u32 raw) { return ((i32)raw) >> 8; }
HVM_HD float num_f24(u32 raw) { return __builtin_bit_cast(float, raw & ~0xFFu); }
HVM_HD u32 num_new_u24(u32 v) { return num_make(N_U24, v); }
HVM_HD u32 num_new_i24(i32 v) { return num_make(N_I24, (u32)v); }
HVM_HD u32 num_new_sym(u32 t) { return num_make(N_SYM, t); }
static const u32 NUM_INVALID = 3u;  // u24 0

// f32 -> f24, round to nearest even on the dropped low 8 bits; NaN stays NaN.
HVM_HD u32 num_new_f24(float val) {
  u32 bits = __builtin_bit_cast(u32, val);
  u32 shifted = bits >> 8;
  u32 lost = bits & 0xFF;
  u32 notnan = (val != val) ? 0u : 1u;
  shifted += notnan & ((lost - ((lost >> 7) & ~shifted)) >> 7);
  shifted |= (1u - notnan);
  return num_make(N_F24, shifted);
}

// Saturating float->int conversions with NaN -> 0.
HVM_HD u32 sat_f2u(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return (u32)f;
}
HVM_HD i32 sat_f2i(float f) {
  if (f != f) return 0;
  if (f <= -2147483648.0f) return (i32)0x80000000u;
  if (f >= 2147483648.0f) return 0x7FFFFFFF;
  return (i32)f;
}

HVM_HD_COLD u32 operate_unary(u32 op, u32 num) {
  const u32 U24_MAX = (1u << 24) - 1;
  const i32 I24_MAX = (1 << 23) - 1, I24_MIN = -(1 << 23);
  u32 t = num_tag(num);
  if (op == N_SYM) return num;
  if (op == N_U24) {
    if (t == N_U24) return num;
    if (t == N_I24) return num_new_u24(num_u24(num));
    if (t == N_F24) {
      u32 v = sat_f2u(num_f24(num));
      return num_new_u24(v > U24_MAX ? U24_MAX : v);
    }
    return NUM_INVALID;
  }
  if (op == N_I24) {
    if (t == N_U24) return num_new_i24(num_i24(num));
    if (t == N_I24) return num;
    if (t == N_F24) {
      i32 v = sat_f2i(num_f24(num));
      v = v < I24_MIN ? I24_MIN : (v > I24_MAX ? I24_MAX : v);
      return num_new_i24(v);
    }
    return NUM_INVALID;
  }
  if (op == N_F24) {
    if (t == N_U24) return num_new_f24((float)num_u24(num));
    if (t == N_I24) return num_new_f24((float)num_i24(num));
    if (t == N_F24) return num;
    return NUM_INVALID;
  }
  return num_make(op, num >> 8);
}

// `x` holds the partial operator (its payload is the left operand), `y` the
// typed value; `*div0` is set on an integer division or remainder by zero,
// which aborts the reference (Rust panic).
HVM_HD_COLD u32 operate_typed(u32 op, u32 ty, u32 x, u32 y, bool *div0) {
  if (ty == N_U24) {
    u32 p = num_u24(x), q = num_u24(y);
    switch (op) {
      case N_ADD: return num_new_u24(p + q);
      case N_SUB: return num_new_u24(p - q);
      case N_SUBS: return num_new_u24(q - p);
      case N_MUL: return num_new_u24(p * q);
      case N_DIV: if (q == 0) { *div0 = true; return NUM_INVALID; } return num_new_u24(p / q);
      case N_DIVS: if (p == 0) { *div0 = true; return NUM_INVALID; } return num_new_u24(q / p);
      case N_REM: if (q == 0) { *div0 = true; return NUM_INVALID; } return num_new_u24(p % q);
      case N_REMS: if (p == 0) { *div0 = true; return NUM_INVALID; } return num_new_u24(q % p);
      case N_EQ: return num_new_u24(p == q);
      case N_NE: return num_new_u24(p != q);
      case N_LT: return num_new_u24(p < q);
      case N_GT: return num_new_u24(p > q);
      case N_AND: return num_new_u24(p & q);
      case N_OR: return num_new_u24(p | q);
      case N_XOR: return num_new_u24(p ^ q);
      case N_SHL: return num_new_u24(p << (q & 31));
      case N_SHLS: return num_new_u24(q << (p & 31));
      case N_SHR: return num_new_u24(p >> (q & 31));
      case N_SHRS: return num_new_u24(q >> (p & 31));
      default: return NUM_INVALID;
    }
  }
  if (ty == N_I24) {
    i32 p = num_i24(x), q = num_i24(y);
    u32 up = (u32)p, uq = (u32)q;
    switch (op) {
      case N_ADD: return num_new_i24((i32)(up + uq));
      case N_SUB: return num_new_i24((i32)(up - uq));
      case N_SUBS: return num_new_i24((i32)(uq - up));
      case N_MUL: return num_new_i24((i32)(up * uq));
      case N_DIV: if (q == 0) { *div0 = true; return NUM_INVALID; } return num_new_i24(p / q);
      case N_DIVS: if (p == 0) { *div0 = true; return NUM_INVALID; } return num_new_i24(q / p);
      case N_REM: if (q == 0) { *div0 = true; return NUM_INVALID; } return num_new_i24(p % q);
      case N_REMS: if (p == 0) { *div0 = true; return NUM_INVALID; } return num_new_i24(q % p);
      case N_EQ: return num_new_u24(p == q);
      case N_NE: return num_new_u24(p != q);
      case N_LT: return num_new_u24(p < q);
      case N_GT: return num_new_u24(p > q);
      case N_AND: return num_new_i24(p & q);
      case N_OR: return num_new_i24(p | q);
      case N_XOR: return num_new_i24(p ^ q);
      case N_SHL: return num_new_i24((i32)(up << (uq & 31)));
      case N_SHLS: return num_new_i24((i32)(uq << (up & 31)));
      case N_SHR: return num_new_i24(p >> (uq & 31));
      case N_SHRS: return num_new_i24(q >> (up & 31));
      default: return NUM_INVALID;
    }
  }
  if (ty == N_F24) {
    float p = num_f24(x), q = num_f24(y);
    switch (op) {
      case N_ADD: return num_new_f24(p + q);
      case N_SUB: return num_new_f24(p - q);
      case N_SUBS: return num_new_f24(q - p);
      case N_MUL: return num_new_f24(p * q);
      case N_DIV: return num_new_f24(p / q);
      case N_DIVS: return num_new_f24(q / p);
      case N_REM: return num_new_f24(fmodf(p, q));
      case N_REMS: return num_new_f24(fmodf(q, p));
      case N_EQ: return num_new_u24(p == q);
      case N_NE: return num_new_u24(p != q);
      case N_LT: return num_new_u24(p < q);
      case N_GT: return num_new_u24(p > q);
      case N_AND: return num_new_f24(atan2f(p, q));
      case N_OR: return num_new_f24(logf(q) / logf(p));
      case N_XOR: return num_new_f24(powf(p, q));
      default: return NUM_INVALID;
    }
  }
  return NUM_INVALID;
}

HVM_HD bool tag_is_op(u32 t) { return t > N_SYM; }

HVM_HD u32 operate_sym(u32 a, u32 b, bool *div0) {
  u32 at = num_tag(a), bt = num_tag(b);
  if (at == N_SYM && bt == N_SYM) return NUM_INVALID;
  if (at == N_SYM) return operate_unary(num_u24(a), b);
  if (bt == N_SYM) return operate_unary(num_u24(b), a);
  bool aop = tag_is_op(at), bop = tag_is_op(bt);
  if (aop == bop) return NUM_INVALID;
  return aop ? operate_typed(at, bt, a, b, div0) : operate_typed(bt, at, b, a, div0);
}

HVM_HD_COLD u32 operate_binary(u32 a, u32 op, u32 b, bool *div0) {
  if (op == N_SYM) return operate_sym(a, b, div0);
  return operate_sym(operate_sym(a, num_new_sym(op), div0), b, div0);
}

}  // namespace hvm
