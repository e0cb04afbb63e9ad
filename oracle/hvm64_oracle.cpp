// HVM-64 CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A single-file C++20 restatement of the reference evaluator
// (HigherOrderCO/hvm-64 v0.3.0) used as the correctness checker for the GPU
// engine and as the timed CPU baseline.  Nothing in the product links it:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it, and only as the checker.
//
// Each section cites the reference file it restates.  What pins each part:
//   num        num/src/num.rs                -- exhaustive u24/i24 closed forms and
//                                               f24 "f32 op then RNE" closed forms
//                                               (tests/test_oracle_num.py) plus the
//                                               u24/i24/f24/numeric_casts snapshots.
//                                               F24 `&` `|` `^` (atan2/log/powf): no
//                                               reference vector; pinned by exact
//                                               identities and by seeded operands whose
//                                               exact result lies far from an f24
//                                               rounding boundary
//                                               (tests/f24_transcendental.py).
//   ast/parser ast/src/{ast,parser}.rs       -- every snapshot program parses and
//                                               prints back to the snapshot text.
//   host       host/src/*.rs                 -- calc_labels unit vector
//                                               (calc_labels.rs:196-234), snapshots.
//   runtime    runtime/src/*.rs              -- 37 run@ snapshots (normal form + 5
//                                               counts), lists.rs count vectors,
//                                               cli.rs apply_tree normal forms.
//   parallel   runtime/src/parallel.rs       -- same snapshots under T threads.
//   readback   host/src/readback.rs          -- snapshot text.
#include <algorithm>
#include <atomic>
#include <barrier>
#include <cassert>
#include <charconv>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <memory>
#include <pthread.h>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace oracle {

struct HvmError : std::runtime_error {
  int code;
  HvmError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
enum { E_PARSE = 1, E_OOM = 2, E_UNIMPL = 3, E_DIV0 = 4, E_UNBOUND = 5, E_NOREF = 6 };

// ---------------------------------------------------------------------------
// num  (num/src/num.rs)
// ---------------------------------------------------------------------------
enum NumTag : uint8_t {
  U24 = 0, I24, F24, Sym, Add, Sub, SubS, Mul, Div, DivS, Rem, RemS,
  And, Or, Xor, Shl, ShlS, Shr, ShrS, Eq, Ne, Lt, Gt, NUM_TAGS
};
// Declaration order matters for from_str_prefix (num.rs:14-42, bi_enum.rs:44-48).
static const char *TAG_STR[NUM_TAGS] = {"u24", "i24", "f24", " ", "+",  "-",   ":-", "*",
                                        "/",   ":/",  "%",   ":%", "&", "|",   "^",  "<<",
                                        ":<<", ">>",  ":>>", "=",  "!", "<",   ">"};

struct Num {
  uint32_t raw;
  static Num make(NumTag t, uint32_t payload) { return Num{(payload << 8) | ((uint32_t)t << 3) | 3u}; }
  static Num u24(uint32_t v) { return make(U24, v); }
  static Num i24(int32_t v) { return make(I24, (uint32_t)v); }
  static Num sym(NumTag t) { return make(Sym, (uint32_t)t); }
  static Num f24(float val) {  // num.rs:81-90
    uint32_t bits;
    std::memcpy(&bits, &val, 4);
    uint32_t shifted = bits >> 8;
    uint32_t lost = bits & 0xFF;
    uint32_t notnan = std::isnan(val) ? 0u : 1u;
    shifted += notnan & ((lost - ((lost >> 7) & ~shifted)) >> 7);
    shifted |= (1u - notnan);
    return make(F24, shifted);
  }
  NumTag tag() const { return (NumTag)((raw >> 3) & 0x1f); }
  uint32_t payload() const { return raw >> 8; }
  uint32_t get_u24() const { return payload(); }
  int32_t get_i24() const { return ((int32_t)raw) >> 8; }
  float get_f24() const {
    uint32_t b = raw & ~0xFFu;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
  }
  NumTag get_sym() const { return (NumTag)get_u24(); }
};
static const Num NUM_INVALID = Num::u24(0);

static inline bool is_op(NumTag t) { return t > Sym; }
static inline bool is_ty(NumTag t) { return t < Sym; }

// Rust `f as u32` / `f as i32`: saturating, NaN -> 0.
static uint32_t f2u32(float f) {
  if (std::isnan(f) || f <= 0.0f) return 0;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)f;
}
static int32_t f2i32(float f) {
  if (std::isnan(f)) return 0;
  if (f <= -2147483648.0f) return INT32_MIN;
  if (f >= 2147483648.0f) return INT32_MAX;
  return (int32_t)f;
}

static Num operate_unary(NumTag op, Num num) {  // num.rs:133-161
  const uint32_t U24_MAX = (1u << 24) - 1;
  const int32_t I24_MAX = (1 << 23) - 1, I24_MIN = -(1 << 23);
  switch (op) {
    case Sym: return num;
    case U24:
      switch (num.tag()) {
        case U24: return num;
        case I24: return Num::u24(num.get_u24());
        case F24: return Num::u24(std::min(f2u32(num.get_f24()), U24_MAX));
        default: return NUM_INVALID;
      }
    case I24:
      switch (num.tag()) {
        case U24: return Num::i24(num.get_i24());
        case I24: return num;
        case F24: return Num::i24(std::clamp(f2i32(num.get_f24()), I24_MIN, I24_MAX));
        default: return NUM_INVALID;
      }
    case F24:
      switch (num.tag()) {
        case U24: return Num::f24((float)num.get_u24());
        case I24: return Num::f24((float)num.get_i24());
        case F24: return num;
        default: return NUM_INVALID;
      }
    default: return Num::make(op, num.payload());
  }
}

static Num operate_sym(Num a, Num b) {  // num.rs:164-254
  NumTag at = a.tag(), bt = b.tag();
  if (at == Sym && bt == Sym) return NUM_INVALID;
  if (at == Sym) return operate_unary(a.get_sym(), b);
  if (bt == Sym) return operate_unary(b.get_sym(), a);
  bool aop = is_op(at), bop = is_op(bt);
  if (aop == bop) return NUM_INVALID;
  NumTag op, ty;
  Num x, y;  // x: the partial op (left operand), y: the typed value
  if (aop) { op = at; ty = bt; x = a; y = b; } else { op = bt; ty = at; x = b; y = a; }
  if (ty == U24) {
    uint32_t p = x.get_u24(), q = y.get_u24();
    switch (op) {
      case Add: return Num::u24(p + q);
      case Sub: return Num::u24(p - q);
      case SubS: return Num::u24(q - p);
      case Mul: return Num::u24(p * q);
      case Div: if (q == 0) throw HvmError(E_DIV0, "u24 division by zero"); return Num::u24(p / q);
      case DivS: if (p == 0) throw HvmError(E_DIV0, "u24 division by zero"); return Num::u24(q / p);
      case Rem: if (q == 0) throw HvmError(E_DIV0, "u24 remainder by zero"); return Num::u24(p % q);
      case RemS: if (p == 0) throw HvmError(E_DIV0, "u24 remainder by zero"); return Num::u24(q % p);
      case Eq: return Num::u24(p == q);
      case Ne: return Num::u24(p != q);
      case Lt: return Num::u24(p < q);
      case Gt: return Num::u24(p > q);
      case And: return Num::u24(p & q);
      case Or: return Num::u24(p | q);
      case Xor: return Num::u24(p ^ q);
      case Shl: return Num::u24(p << (q & 31));
      case ShlS: return Num::u24(q << (p & 31));
      case Shr: return Num::u24(p >> (q & 31));
      case ShrS: return Num::u24(q >> (p & 31));
      default: break;
    }
  } else if (ty == I24) {
    int32_t p = x.get_i24(), q = y.get_i24();
    auto w = [](int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; };
    switch (op) {
      case Add: return Num::i24(w((int64_t)p + q));
      case Sub: return Num::i24(w((int64_t)p - q));
      case SubS: return Num::i24(w((int64_t)q - p));
      case Mul: return Num::i24(w((int64_t)p * q));
      case Div: if (q == 0) throw HvmError(E_DIV0, "i24 division by zero"); return Num::i24(p / q);
      case DivS: if (p == 0) throw HvmError(E_DIV0, "i24 division by zero"); return Num::i24(q / p);
      case Rem: if (q == 0) throw HvmError(E_DIV0, "i24 remainder by zero"); return Num::i24(p % q);
      case RemS: if (p == 0) throw HvmError(E_DIV0, "i24 remainder by zero"); return Num::i24(q % p);
      case Eq: return Num::u24(p == q);
      case Ne: return Num::u24(p != q);
      case Lt: return Num::u24(p < q);
      case Gt: return Num::u24(p > q);
      case And: return Num::i24(p & q);
      case Or: return Num::i24(p | q);
      case Xor: return Num::i24(p ^ q);
      case Shl: return Num::i24((int32_t)((uint32_t)p << ((uint32_t)q & 31)));
      case ShlS: return Num::i24((int32_t)((uint32_t)q << ((uint32_t)p & 31)));
      case Shr: return Num::i24(p >> ((uint32_t)q & 31));
      case ShrS: return Num::i24(q >> ((uint32_t)p & 31));
      default: break;
    }
  } else if (ty == F24) {
    float p = x.get_f24(), q = y.get_f24();
    switch (op) {
      case Add: return Num::f24(p + q);
      case Sub: return Num::f24(p - q);
      case SubS: return Num::f24(q - p);
      case Mul: return Num::f24(p * q);
      case Div: return Num::f24(p / q);
      case DivS: return Num::f24(q / p);
      case Rem: return Num::f24(std::fmod(p, q));
      case RemS: return Num::f24(std::fmod(q, p));
      case Eq: return Num::u24(p == q);
      case Ne: return Num::u24(p != q);
      case Lt: return Num::u24(p < q);
      case Gt: return Num::u24(p > q);
      // Rust std atan2 / log / powf in f32 (num.rs:243-248); pinned by
      // identities and margin-checked operands, not by reference vectors
      case And: return Num::f24(atan2f(p, q));
      case Or: return Num::f24(logf(q) / logf(p));
      case Xor: return Num::f24(powf(p, q));
      default: return NUM_INVALID;
    }
  }
  return NUM_INVALID;
}

static Num operate_binary(Num a, NumTag op, Num b) {  // num.rs:257-263
  if (op == Sym) return operate_sym(a, b);
  return operate_sym(operate_sym(a, Num::sym(op)), b);
}

// Rust `{:?}` of a finite f32: shortest round-trip digits; exponent form when
// |x| < 1e-4 (x != 0) or |x| >= 1e16; otherwise at least one fractional digit.
static std::string rust_f32_debug(float v) {
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
  std::string s(buf, r.ptr);
  bool neg = false;
  if (!s.empty() && s[0] == '-') { neg = true; s = s.substr(1); }
  size_t epos = s.find('e');
  std::string mant = s.substr(0, epos);
  int exp10 = std::atoi(s.c_str() + epos + 1);
  std::string digits;
  for (char c : mant) if (c != '.') digits += c;
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  float a = std::fabs(v);
  std::string out = neg ? "-" : "";
  if ((a != 0.0f && a < 1e-4f) || a >= 1e16f) {
    out += digits[0];
    if (digits.size() > 1) { out += '.'; out += digits.substr(1); }
    out += 'e';
    out += std::to_string(exp10);
    return out;
  }
  if (a == 0.0f) return out + "0.0";
  int point = exp10 + 1;  // digits before the decimal point
  if (point <= 0) {
    out += "0.";
    out += std::string(-point, '0');
    out += digits;
  } else if ((size_t)point >= digits.size()) {
    out += digits;
    out += std::string(point - digits.size(), '0');
    out += ".0";
  } else {
    out += digits.substr(0, point);
    out += '.';
    out += digits.substr(point);
  }
  return out;
}

static std::string num_display(Num n) {  // num.rs:287-306
  switch (n.tag()) {
    case Sym: return std::string("[") + TAG_STR[n.get_sym() < NUM_TAGS ? n.get_sym() : 0] + "]";
    case U24: return std::to_string(n.get_u24());
    case I24: {
      int32_t v = n.get_i24();
      return (v >= 0 ? "+" : "") + std::to_string(v);
    }
    case F24: {
      float v = n.get_f24();
      if (std::isinf(v)) return v > 0 ? "+inf" : "-inf";
      if (std::isnan(v)) return "+NaN";
      return rust_f32_debug(v);
    }
    default: return std::string("[") + TAG_STR[n.tag()] + std::to_string(n.payload()) + "]";
  }
}

// ---------------------------------------------------------------------------
// ast + parser + Display  (ast/src/ast.rs, ast/src/parser.rs)
// ---------------------------------------------------------------------------
struct Tree {
  enum Kind : uint8_t { Era, NumK, RefK, CtrK, OpK, SwitchK, VarK } kind = Era;
  uint16_t lab = 0;
  Num num{0};
  std::string name;
  std::unique_ptr<Tree> a, b;  // Ctr: p1 p2 ; Op: rhs out ; Switch: arms out
  Tree() = default;
  Tree(Tree &&) = default;
  Tree &operator=(Tree &&) = default;
  ~Tree() {  // iterative drop (ast.rs:195-219 avoids stack overflow the same way)
    std::vector<std::unique_ptr<Tree>> stack;
    if (a) stack.push_back(std::move(a));
    if (b) stack.push_back(std::move(b));
    while (!stack.empty()) {
      std::unique_ptr<Tree> t = std::move(stack.back());
      stack.pop_back();
      if (t->a) stack.push_back(std::move(t->a));
      if (t->b) stack.push_back(std::move(t->b));
    }
  }
  static Tree leaf(Kind k) { Tree t; t.kind = k; return t; }
  Tree clone() const {
    Tree t;
    t.kind = kind; t.lab = lab; t.num = num; t.name = name;
    if (a) t.a = std::make_unique<Tree>(a->clone());
    if (b) t.b = std::make_unique<Tree>(b->clone());
    return t;
  }
};

struct AstNet {
  Tree root;
  std::vector<std::pair<Tree, Tree>> redexes;
};
using Book = std::map<std::string, AstNet>;

static std::string create_var(size_t id) {  // util/src/create_var.rs:5-15
  std::string txt;
  id += 1;
  while (id > 0) {
    id -= 1;
    txt.push_back((char)('a' + id % 26));
    id /= 26;
  }
  std::reverse(txt.begin(), txt.end());
  return txt;
}

static bool var_to_num(const std::string &s, size_t &out) {  // create_var.rs:21-33
  size_t n = 0;
  for (char c : s) {
    if (c < 'a') return false;
    size_t i = (size_t)(c - 'a');
    if (i > 'z') return false;
    n = n * 26 + i + 1;
  }
  if (n == 0) return false;
  out = n - 1;
  return true;
}

static void print_tree(std::string &o, const Tree &t) {  // ast.rs:246-263
  switch (t.kind) {
    case Tree::Era: o += '*'; break;
    case Tree::CtrK: {
      uint16_t lab = t.lab;
      if (lab == 0) o += '(';
      else if (lab == 1) o += '{';
      else { o += '#'; o += std::to_string(lab / 2); o += (lab % 2 == 0) ? '(' : '{'; }
      print_tree(o, *t.a);
      o += ' ';
      print_tree(o, *t.b);
      o += (lab % 2 == 0) ? ')' : '}';
      break;
    }
    case Tree::VarK: o += t.name; break;
    case Tree::RefK: o += '@'; o += t.name; break;
    case Tree::NumK: o += num_display(t.num); break;
    case Tree::OpK: o += "$("; print_tree(o, *t.a); o += ' '; print_tree(o, *t.b); o += ')'; break;
    case Tree::SwitchK: o += "?("; print_tree(o, *t.a); o += ' '; print_tree(o, *t.b); o += ')'; break;
  }
}

static std::string print_net(const AstNet &n) {
  std::string o;
  print_tree(o, n.root);
  for (auto &r : n.redexes) {
    o += "\n  & ";
    print_tree(o, r.first);
    o += " ~ ";
    print_tree(o, r.second);
  }
  return o;
}

struct Parser {
  const std::string &in;
  size_t i = 0;
  explicit Parser(const std::string &s) : in(s) {}
  [[noreturn]] void fail(const std::string &what) {
    throw HvmError(E_PARSE, "parse error at byte " + std::to_string(i) + ": expected " + what);
  }
  void skip_trivia() {
    for (;;) {
      while (i < in.size() && isspace((unsigned char)in[i])) i++;
      if (i + 1 < in.size() && in[i] == '/' && in[i + 1] == '/') {
        while (i < in.size() && in[i] != '\n') i++;
        continue;
      }
      break;
    }
  }
  int peek() { return i < in.size() ? (unsigned char)in[i] : -1; }
  bool try_consume(const char *s) {
    skip_trivia();
    size_t n = strlen(s);
    if (in.compare(i, n, s) == 0) { i += n; return true; }
    return false;
  }
  void consume(const char *s) { if (!try_consume(s)) fail(std::string("`") + s + "`"); }
  static bool name_char(int c) { return isalnum(c) || c == '_' || c == '.' || c == '$'; }
  std::string parse_name() {
    size_t s = i;
    while (i < in.size() && name_char((unsigned char)in[i])) i++;
    if (i == s) fail("name");
    return in.substr(s, i - s);
  }
  Book parse_book() {
    Book book;
    while (try_consume("@")) {
      std::string name = parse_name();
      consume("=");
      AstNet net = parse_net();
      book[name] = std::move(net);
    }
    return book;
  }
  AstNet parse_net() {
    AstNet net;
    net.root = parse_tree();
    while (try_consume("&")) {
      Tree a = parse_tree();
      consume("~");
      Tree b = parse_tree();
      net.redexes.emplace_back(std::move(a), std::move(b));
    }
    return net;
  }
  Tree parse_tree() {
    skip_trivia();
    int c = peek();
    Tree t;
    if (c == '*') { i++; t.kind = Tree::Era; return t; }
    if (c == '(' || c == '{' || c == '#') {
      i++;
      uint16_t tag = 0;
      if (c == '#') {
        skip_trivia();
        size_t s = i;
        while (i < in.size() && isdigit((unsigned char)in[i])) i++;
        if (s == i) fail("label");
        tag = (uint16_t)std::stoull(in.substr(s, i - s));
        skip_trivia();
        c = peek();
        i++;
      }
      const char *close;
      uint16_t kind;
      if (c == '(') { close = ")"; kind = 0; }
      else if (c == '{') { close = "}"; kind = 1; }
      else fail("`(` or `{`");
      t.kind = Tree::CtrK;
      t.lab = (uint16_t)(tag * 2 + kind);
      t.a = std::make_unique<Tree>(parse_tree());
      t.b = std::make_unique<Tree>(parse_tree());
      consume(close);
      return t;
    }
    if (c == '@') {
      i++;
      skip_trivia();
      t.kind = Tree::RefK;
      t.name = parse_name();
      return t;
    }
    if (c == '$') {
      consume("$(");
      t.kind = Tree::OpK;
      t.a = std::make_unique<Tree>(parse_tree());
      t.b = std::make_unique<Tree>(parse_tree());
      consume(")");
      return t;
    }
    if (c == '?') {
      consume("?(");
      t.kind = Tree::SwitchK;
      t.a = std::make_unique<Tree>(parse_tree());
      t.b = std::make_unique<Tree>(parse_tree());
      consume(")");
      return t;
    }
    if ((c >= '0' && c <= '9') || c == '+' || c == '-' || c == '[') {
      t.kind = Tree::NumK;
      t.num = parse_num();
      return t;
    }
    t.kind = Tree::VarK;
    t.name = parse_name();
    return t;
  }
  Num parse_num() {  // parser.rs:116-140
    skip_trivia();
    if (peek() != '[') return parse_num_lit();
    consume("[");
    int op = -1;
    for (int k = 0; k < NUM_TAGS; k++) {
      size_t n = strlen(TAG_STR[k]);
      if (in.compare(i, n, TAG_STR[k]) == 0) { op = k; break; }
    }
    if (op < 0) fail("operator");
    i += strlen(TAG_STR[op]);
    skip_trivia();
    if (is_ty((NumTag)op) || peek() == ']') {
      consume("]");
      return Num::sym((NumTag)op);
    }
    Num lhs = parse_num_lit();
    consume("]");
    return operate_unary((NumTag)op, lhs);
  }
  static uint64_t parse_int(const std::string &s, Parser *p) {
    try {
      size_t used = 0;
      uint64_t v;
      if (s.rfind("0x", 0) == 0) v = std::stoull(s.substr(2), &used, 16), used += 2;
      else if (s.rfind("0b", 0) == 0) v = std::stoull(s.substr(2), &used, 2), used += 2;
      else {
        if (s.empty() || !isdigit((unsigned char)s[0])) p->fail("integer");
        v = std::stoull(s, &used, 10);
      }
      if (used != s.size()) p->fail("integer");
      return v;
    } catch (std::logic_error &) {
      p->fail("integer");
    }
  }
  Num parse_num_lit() {  // parser.rs:142-154
    size_t s = i;
    while (i < in.size() && (isalnum((unsigned char)in[i]) || in[i] == '+' || in[i] == '-' || in[i] == '.')) i++;
    std::string num = in.substr(s, i - s);
    if (num.find('.') != std::string::npos || num.find("inf") != std::string::npos ||
        num.find("NaN") != std::string::npos) {
      char *end = nullptr;
      float v = strtof(num.c_str(), &end);
      if (end != num.c_str() + num.size()) fail("float");
      return Num::f24(v);
    }
    if (!num.empty() && (num[0] == '+' || num[0] == '-')) {
      int32_t v = (int32_t)(uint32_t)parse_int(num.substr(1), this);
      return Num::i24(num[0] == '-' ? (int32_t)(0u - (uint32_t)v) : v);
    }
    return Num::u24((uint32_t)parse_int(num, this));
  }
};

template <class F>
static auto parse_eof(const std::string &s, F f) {
  Parser p(s);
  auto out = f(p);
  p.skip_trivia();
  if (p.i != s.size()) throw HvmError(E_PARSE, "Unable to parse the whole input. Is this not an hvm64 file?");
  return out;
}
static Book parse_book(const std::string &s) { return parse_eof(s, [](Parser &p) { return p.parse_book(); }); }
static AstNet parse_net(const std::string &s) { return parse_eof(s, [](Parser &p) { return p.parse_net(); }); }
static Tree parse_tree(const std::string &s) { return parse_eof(s, [](Parser &p) { return p.parse_tree(); }); }

static void tree_fresh(const Tree &t, size_t &fresh) {  // ast.rs:169-176
  if (t.kind == Tree::VarK) {
    size_t n;
    if (var_to_num(t.name, n)) fresh = std::max(fresh, n);
  }
  if (t.a) tree_fresh(*t.a, fresh);
  if (t.b) tree_fresh(*t.b, fresh);
}

static void apply_tree(AstNet &net, Tree arg) {  // ast.rs:112-123
  size_t fresh = 0;
  tree_fresh(net.root, fresh);
  for (auto &r : net.redexes) { tree_fresh(r.first, fresh); tree_fresh(r.second, fresh); }
  tree_fresh(arg, fresh);
  std::string fs = create_var(fresh + 1);
  Tree fun = std::move(net.root);
  Tree app;
  app.kind = Tree::CtrK;
  app.lab = 0;
  app.a = std::make_unique<Tree>(std::move(arg));
  app.b = std::make_unique<Tree>();
  app.b->kind = Tree::VarK;
  app.b->name = fs;
  net.root = Tree();
  net.root.kind = Tree::VarK;
  net.root.name = fs;
  net.redexes.emplace_back(std::move(fun), std::move(app));
}

// ---------------------------------------------------------------------------
// runtime: Port / Addr / Wire  (runtime/src/port.rs, addr.rs, wire.rs)
// ---------------------------------------------------------------------------
enum Tag : uint8_t { Red = 0, Var = 1, Ref = 2, NumT = 3, Op = 5, Switch = 6, Ctr = 7 };
using Lab = uint16_t;

static const uint64_t P_ERA = 2, P_FREE = 0x8000000000000000ull, P_LOCK = 0xFFFFFFFFFFFFFFF0ull,
                      P_GONE = 0xFFFFFFFFFFFFFFFFull;

struct Port {
  uint64_t v;
  static Port make(Tag t, Lab lab, uint64_t addr) { return Port{((uint64_t)lab << 48) | addr | (uint64_t)t}; }
  static Port var(uint64_t addr) { return make(Var, 0, addr); }
  static Port num(Num n) { return Port{(uint64_t)n.raw}; }
  Tag tag() const { return (Tag)(v & 7); }
  Lab lab() const { return (Lab)(v >> 48); }
  uint64_t addr() const { return v & 0x0000FFFFFFFFFFF8ull; }
  NumTag op() const { return (NumTag)lab(); }
  Num getnum() const { return Num{(uint32_t)v}; }
  bool is_principal() const { return tag() >= Ref; }
  bool is_skippable() const { return tag() == NumT || (tag() == Ref && lab() != 0xFFFF); }
  bool is_era() const { return v == P_ERA; }
  bool operator==(const Port &o) const { return v == o.v; }
  bool operator!=(const Port &o) const { return v != o.v; }
  Port redirect() const { return make(Red, 0, addr()); }
};
static const Port ERA{P_ERA};

// Optional metering of heap word traffic (used to derive algorithmic bytes per
// interaction; loads 8 B, stores 8 B, read-modify-writes 16 B).
struct Meter {
  uint64_t loads = 0, stores = 0, rmws = 0;
};
static thread_local Meter *g_meter = nullptr;
#define METER(field) \
  do {                \
    if (g_meter) g_meter->field++; \
  } while (0)

static inline std::atomic<uint64_t> &word(uint64_t addr) { return *reinterpret_cast<std::atomic<uint64_t> *>(addr); }
static inline uint64_t other_half(uint64_t a) { return a ^ 8; }
static inline uint64_t left_half(uint64_t a) { return a & ~(uint64_t)8; }

struct Wire {
  uint64_t a;  // address of the AtomicU64 holding the target
  Port load_target() const { METER(loads); return Port{word(a).load(std::memory_order_relaxed)}; }
  void set_target(Port p) const { METER(stores); word(a).store(p.v, std::memory_order_relaxed); }
  bool cas_target(Port expected, Port value, Port &got) const {
    METER(rmws);
    uint64_t e = expected.v;
    bool ok = word(a).compare_exchange_strong(e, value.v, std::memory_order_relaxed);
    got = Port{e};
    return ok;
  }
  Port swap_target(Port value) const { METER(rmws); return Port{word(a).exchange(value.v, std::memory_order_relaxed)}; }
  Port lock_target() const {  // wire.rs:67-75
    for (;;) {
      Port got = swap_target(Port{P_LOCK});
      if (got.v != P_LOCK) return got;
    }
  }
  Port as_var() const { return Port::var(a); }
};
static inline Wire port_wire(Port p) { return Wire{p.addr()}; }

// Trg: a port or a wire; wires reuse the Red tag (linker.rs:233-281).
struct Trg {
  uint64_t v;
  static Trg port(Port p) { return Trg{p.v}; }
  static Trg wire(Wire w) { return Trg{w.a}; }
  bool is_wire() const { return (v & 7) == Red; }
  Wire as_wire() const { return Wire{v}; }
  Port as_port() const { return Port{v}; }
  Port target() const { return is_wire() ? as_wire().load_target() : as_port(); }
};

struct Rewrites {
  uint64_t anni = 0, comm = 0, eras = 0, dref = 0, oper = 0;
  uint64_t total() const { return anni + comm + eras + dref + oper; }
};

// ---------------------------------------------------------------------------
// defs + instructions  (runtime/src/def.rs, instruction.rs)
// ---------------------------------------------------------------------------
struct LabSet {  // def.rs:7-62
  Lab min_safe = 0;
  std::vector<uint64_t> bits;
  bool all = false;
  void add(Lab lab) {
    min_safe = std::max<Lab>(min_safe, (Lab)(lab + 1));
    size_t idx = lab >> 6;
    if (idx >= bits.size()) bits.resize(idx + 1, 0);
    bits[idx] |= 1ull << (lab & 63);
  }
  bool has(Lab lab) const {
    if (lab >= min_safe) return false;
    if (all) return true;
    return (bits[lab >> 6] >> (lab & 63)) & 1;
  }
  void unite(const LabSet &o) {
    min_safe = std::max(min_safe, o.min_safe);
    if (o.bits.size() > bits.size()) bits.resize(o.bits.size(), 0);
    for (size_t k = 0; k < o.bits.size(); k++) bits[k] |= o.bits[k];
  }
  static LabSet ALL() {
    LabSet s;
    s.min_safe = 0xFFFF;
    s.all = true;
    return s;
  }
  bool operator==(const LabSet &o) const {
    if (min_safe != o.min_safe) return false;
    size_t n = std::max(bits.size(), o.bits.size());
    for (size_t k = 0; k < n; k++) {
      uint64_t x = k < bits.size() ? bits[k] : 0, y = k < o.bits.size() ? o.bits[k] : 0;
      if (x != y) return false;
    }
    return true;
  }
};

enum InstrKind : uint8_t { I_CONST, I_LINK, I_LINKCONST, I_CTR, I_OP, I_OPNUM, I_SWITCH, I_WIRES };
struct Instruction {  // instruction.rs:34-72
  InstrKind kind;
  uint32_t a = 0, b = 0, c = 0, d = 0;  // TrgId indices (meaning depends on kind)
  Lab lab = 0;                           // Ctr label / Op op
  Port port{0};                          // Const / LinkConst
  Num rhs{0};                            // OpNum
};

struct InterpretedDef {
  std::vector<Instruction> instr;
  uint32_t trgs = 1;
  uint32_t new_trg() { return trgs++; }
};

struct Net;
struct alignas(32) Def {  // def.rs:86-95
  LabSet labs;
  int kind = 0;  // 0 interpreted, 1 expand, 2 inert (pre_reduce.rs:99-116)
  InterpretedDef data;
  Port out{0};  // ExpandDef payload (net.rs:72-74)
  // A compiled definition (AsDef::call of the structs src/compile.rs:110-161
  // writes): when set, calls run it instead of interpreting `data`.
  void (*compiled)(Net &net, Port port) = nullptr;
};
static inline Port new_ref(const Def *d) { return Port::make(Ref, d->labs.min_safe, (uint64_t)d); }
static inline Def *port_def(Port p) { return reinterpret_cast<Def *>(p.addr()); }

// ---------------------------------------------------------------------------
// Heap / Allocator  (runtime/src/allocator.rs)
// ---------------------------------------------------------------------------
struct alignas(16) Node {
  std::atomic<uint64_t> w0, w1;
};

struct Heap {
  Node *nodes = nullptr;
  size_t len = 0;
  explicit Heap(size_t words) {
    len = words / 2;
    if (len == 0) throw HvmError(E_OOM, "empty heap");
    nodes = static_cast<Node *>(std::aligned_alloc(16, len * sizeof(Node)));
    if (!nodes) throw HvmError(E_OOM, "memory allocation failed");
  }
  Heap(Node *n, size_t l) : nodes(n), len(l) {}
  ~Heap() {}
};

struct RedexQueue {  // linker.rs:307-344
  std::vector<std::pair<Port, Port>> fast, slow;
  bool pop(Port &a, Port &b) {
    if (!fast.empty()) { a = fast.back().first; b = fast.back().second; fast.pop_back(); return true; }
    if (!slow.empty()) { a = slow.back().first; b = slow.back().second; slow.pop_back(); return true; }
    return false;
  }
  size_t len() const { return fast.size() + slow.size(); }
  bool empty() const { return fast.empty() && slow.empty(); }
};

static bool redex_would_shrink(Port a, Port b) {  // linker.rs:347-351
  return a == ERA || b == ERA || (a.tag() == Ctr && b.tag() == Ctr && a.lab() == b.lab()) ||
         (!(a.tag() == Ref || b.tag() == Ref) && (a.tag() == NumT || b.tag() == NumT));
}

struct TraverseNode {
  Tag tag;
  Lab lab;
  Wire p1, p2;
};
static inline TraverseNode traverse(Port p) { return TraverseNode{p.tag(), p.lab(), Wire{p.addr()}, Wire{other_half(p.addr())}}; }
struct CreatedNode {
  Port p0, p1, p2;
};

// Per-net shortcut counters: interpreter fast paths whose firing depends on
// whether a wire target is already resolved (instruction.rs:144-146,163-165,194-197).
struct Shortcuts {
  uint64_t op_era = 0, opnum_era = 0, switch_era = 0;
};

// ---------------------------------------------------------------------------
// Net: allocator + linker + interactions + call  (allocator.rs, linker.rs,
// interact.rs, def.rs, instruction.rs, net.rs)
// ---------------------------------------------------------------------------
struct Net {
  // allocator
  Heap *heap;
  size_t next = 0;
  uint64_t head = 0;
  // linker
  Rewrites rwts;
  RedexQueue redexes;
  Shortcuts shortcuts;
  // Redexes an inert definition was called on (InertDef::call,
  // transform/src/pre_reduce.rs:108-115), in call order.
  std::vector<std::pair<Port, Port>> captured;
  std::vector<Trg> trgs;
  Wire root{0};
  size_t tid = 0, tids = 1;

  explicit Net(Heap *h, bool with_root = true) : heap(h) {
    trgs.resize(1 << 16);
    if (with_root) root = Wire{alloc()};  // net.rs:19-23
  }

  // --- allocator.rs:71-110
  void half_free(uint64_t addr) {
    METER(stores);
    word(addr).store(P_FREE, std::memory_order_relaxed);
    METER(loads);
    if (word(other_half(addr)).load(std::memory_order_relaxed) == P_FREE) {
      uint64_t l = left_half(addr);
      uint64_t e = P_FREE;
      METER(rmws);
      if (word(l).compare_exchange_strong(e, head, std::memory_order_relaxed)) head = l;
    }
  }
  uint64_t alloc() {
    uint64_t addr;
    if (head != 0) {
      addr = head;
      METER(loads);
      head = word(head).load(std::memory_order_relaxed);
    } else {
      if (next >= heap->len) throw HvmError(E_OOM, "OOM");
      addr = (uint64_t)&heap->nodes[next++].w0;
    }
    METER(stores);
    METER(stores);
    word(addr).store(P_LOCK, std::memory_order_relaxed);
    word(other_half(addr)).store(P_LOCK, std::memory_order_relaxed);
    return addr;
  }
  void free_wire(Wire w) { half_free(w.a); }
  void free_trg(Trg t) { if (t.is_wire()) free_wire(t.as_wire()); }

  // --- node.rs:36-47
  CreatedNode create_node(Tag tag, Lab lab) {
    uint64_t a = alloc();
    return CreatedNode{Port::make(tag, lab, a), Port::var(a), Port::var(other_half(a))};
  }
  Wire create_wire_to(Port p) {
    uint64_t a = alloc();
    half_free(other_half(a));
    Wire w{a};
    link_port_port(p, Port::var(w.a));
    return w;
  }

  // --- linker.rs:147-209
  void link_port_port(Port a, Port b) {
    if (a.is_principal() && b.is_principal()) {
      redux(a, b);
    } else {
      half_link_port_port(a, b);
      half_link_port_port(b, a);
    }
  }
  void link_wire_wire(Wire aw, Wire bw) {
    Port ap = aw.lock_target();
    Port bp = bw.lock_target();
    if (ap.is_principal() && bp.is_principal()) {
      free_wire(aw);
      free_wire(bw);
      redux(ap, bp);
    } else {
      half_link_wire_port(ap, aw, bp);
      half_link_wire_port(bp, bw, ap);
    }
  }
  void link_wire_port(Wire aw, Port bp) {
    Port ap = aw.lock_target();
    if (ap.is_principal() && bp.is_principal()) {
      free_wire(aw);
      redux(ap, bp);
    } else {
      half_link_wire_port(ap, aw, bp);
      half_link_port_port(bp, ap);
    }
  }
  void redux(Port a, Port b) {
    if (a.is_skippable() && b.is_skippable()) {
      rwts.eras++;
    } else if (redex_would_shrink(a, b)) {
      redexes.fast.emplace_back(a, b);
    } else {
      redexes.slow.emplace_back(a, b);
    }
  }
  void half_link_port_port(Port a, Port b) {
    if (a.tag() == Var) port_wire(a).set_target(b);
  }
  void half_link_wire_port(Port ap, Wire aw, Port bp) {  // linker.rs:224-251
    if (ap.tag() == Var) {
      Port got;
      if (port_wire(ap).cas_target(aw.as_var(), bp, got)) {
        free_wire(aw);
      } else {
        if (bp.tag() == Var) {
          aw.set_target(bp.redirect());
        } else if (bp.is_principal()) {
          aw.set_target(bp);
          resolve_redirect_pri(ap, aw, bp);
        } else {
          throw HvmError(E_UNIMPL, "unreachable link state");
        }
      }
    } else {
      free_wire(aw);
    }
  }
  void resolve_redirect_pri(Port ap, Wire aw, Port bp) {  // linker.rs:254-327
    for (;;) {
      Wire tw = port_wire(ap);
      Port tp = tw.load_target();
      if (tp.v == P_LOCK) continue;
      if (tp.tag() == Red) {
        free_wire(tw);
        ap = tp;
        continue;
      }
      if (tp.tag() == Var) {
        Port got;
        if (tw.cas_target(tp, bp, got)) {
          tw = port_wire(tp);
          tp = tw.load_target();
          while (tp.v != P_LOCK && tp.tag() == Red) {
            free_wire(tw);
            tw = port_wire(tp);
            tp = tw.load_target();
          }
          return;
        }
        continue;
      }
      if (tp.is_principal() || tp.v == P_GONE) {
        Wire xw = aw.a < tw.a ? aw : tw;
        Wire yw = aw.a < tw.a ? tw : aw;
        Port xp = xw.swap_target(Port{P_GONE});
        if (xp.v != P_GONE) {
          Port yp = yw.swap_target(Port{P_GONE});
          redux(xp, yp);
          return;
        } else {
          free_wire(xw);
          Port got;
          while (!yw.cas_target(Port{P_GONE}, Port{P_LOCK}, got)) {
          }
          free_wire(yw);
          return;
        }
      }
      throw HvmError(E_UNIMPL, "unreachable redirect state");
    }
  }
  void link_trg_port(Trg a, Port b) {
    if (a.is_wire()) link_wire_port(a.as_wire(), b);
    else link_port_port(a.as_port(), b);
  }
  void link_trg(Trg a, Trg b) {
    bool aw = a.is_wire(), bw = b.is_wire();
    if (aw && bw) link_wire_wire(a.as_wire(), b.as_wire());
    else if (aw) link_wire_port(a.as_wire(), b.as_port());
    else if (bw) link_wire_port(b.as_wire(), a.as_port());
    else link_port_port(a.as_port(), b.as_port());
  }

  // --- interact.rs:6-44
  void interact(Port a, Port b) {
    Tag at = a.tag(), bt = b.tag();
    if (at == Var || at == Red || bt == Var || bt == Red) throw HvmError(E_UNIMPL, "not an active pair");
    bool anil = (at == Ref || at == NumT), bnil = (bt == Ref || bt == NumT);
    if (at == Ref && bnil && !a.is_skippable()) return call(a, b);
    if (anil && bt == Ref && !b.is_skippable()) return call(b, a);
    if (anil && bnil) { rwts.eras++; return; }
    if (at == Ctr && bt == Switch && a.lab() != 0) return comm22(a, b);
    if (at == Switch && bt == Ctr && b.lab() != 0) return comm22(a, b);
    if ((at == Ctr && bt == Op) || (at == Op && bt == Ctr)) return comm22(a, b);
    if (at == Ctr && bt == Ctr && a.lab() != b.lab()) return comm22(a, b);
    if ((at == Switch && bt == Switch) || (at == Op && bt == Op) || (at == Ctr && bt == Ctr)) return anni2(a, b);
    if (at == Ref && bt == Ctr && b.lab() >= a.lab()) return comm02(a, b);
    if (at == Ctr && bt == Ref && a.lab() >= b.lab()) return comm02(b, a);
    if (at == NumT && bt == Ctr) return comm02(a, b);
    if (at == Ctr && bt == NumT) return comm02(b, a);
    if (at == Ref && a == ERA) return comm02(a, b);
    if (bt == Ref && b == ERA) return comm02(b, a);
    if (at == Ref) return call(a, b);
    if (bt == Ref) return call(b, a);
    if (at == Op && bt == NumT) return op_num(a, b);
    if (at == NumT && bt == Op) return op_num(b, a);
    if (at == Switch && bt == NumT) return switch_num(a, b);
    if (at == NumT && bt == Switch) return switch_num(b, a);
    throw HvmError(E_UNIMPL, "unimplemented interaction");
  }
  void anni2(Port a, Port b) {
    rwts.anni++;
    TraverseNode x = traverse(a), y = traverse(b);
    link_wire_wire(x.p1, y.p1);
    link_wire_wire(x.p2, y.p2);
  }
  void comm22(Port a, Port b) {
    rwts.comm++;
    TraverseNode x = traverse(a), y = traverse(b);
    CreatedNode A1 = create_node(x.tag, x.lab), A2 = create_node(x.tag, x.lab);
    CreatedNode B1 = create_node(y.tag, y.lab), B2 = create_node(y.tag, y.lab);
    link_port_port(A1.p1, B1.p1);
    link_port_port(A1.p2, B2.p1);
    link_port_port(A2.p1, B1.p2);
    link_port_port(A2.p2, B2.p2);
    link_wire_port(x.p1, B1.p0);
    link_wire_port(x.p2, B2.p0);
    link_wire_port(y.p1, A1.p0);
    link_wire_port(y.p2, A2.p0);
  }
  void comm02(Port a, Port b) {
    rwts.comm++;
    TraverseNode y = traverse(b);
    link_wire_port(y.p1, a);
    link_wire_port(y.p2, a);
  }
  void switch_num(Port a, Port b) {  // interact.rs:201-223
    rwts.oper++;
    TraverseNode x = traverse(a);
    Num n = b.getnum();
    uint32_t v = n.tag() == U24 ? n.get_u24() : 0;
    if (v == 0) {
      CreatedNode c = create_node(Ctr, 0);
      link_port_port(c.p2, ERA);
      link_wire_port(x.p2, c.p1);
      link_wire_port(x.p1, c.p0);
    } else {
      CreatedNode c = create_node(Ctr, 0);
      CreatedNode d = create_node(Ctr, 0);
      link_port_port(c.p1, ERA);
      link_port_port(c.p2, d.p0);
      link_port_port(d.p1, Port::num(Num::u24(v - 1)));
      link_wire_port(x.p2, d.p2);
      link_wire_port(x.p1, c.p0);
    }
  }
  void op_num(Port a, Port b) {  // interact.rs:251-270
    NumTag op = a.op();
    TraverseNode x = traverse(a);
    Port a1 = x.p1.load_target();
    if (a1.tag() == NumT) {
      rwts.oper++;
      half_free(x.p1.a);
      Num out = operate_binary(b.getnum(), op, a1.getnum());
      link_wire_port(x.p2, Port::num(out));
    } else {
      CreatedNode c = create_node(Op, (Lab)Sym);
      link_port_port(c.p1, Port::num(operate_unary(op, b.getnum())));
      link_wire_port(x.p2, c.p2);
      link_wire_port(x.p1, c.p0);
    }
  }

  // --- def.rs:177-189
  void call(Port port, Port trg) {
    Def *def = port_def(port);
    if (trg.tag() == Ctr && !def->labs.has(trg.lab())) return comm02(port, trg);
    rwts.dref++;
    if (def->kind == 1) return expand_call(def, trg);
    if (def->kind == 2) return captured.emplace_back(port, trg), void();
    if (def->compiled) return def->compiled(*this, trg);
    interpret(def->data, Trg::port(trg));
  }

  // --- def.rs:217-289
  void interpret(const InterpretedDef &d, Trg t0) {
    if (d.trgs >= trgs.size()) trgs.resize(d.trgs);
    Trg *t = trgs.data();
    t[0] = t0;
    for (const Instruction &i : d.instr) {
      switch (i.kind) {
        case I_CONST: t[i.a] = Trg::port(i.port); break;
        case I_LINK: link_trg(t[i.a], t[i.b]); break;
        case I_LINKCONST: link_trg_port(t[i.a], i.port); break;
        case I_CTR: {
          auto r = do_ctr(i.lab, t[i.a]);
          t[i.b] = r.first;
          t[i.c] = r.second;
          break;
        }
        case I_OP: {
          auto r = do_op((NumTag)i.lab, t[i.a]);
          t[i.b] = r.first;
          t[i.c] = r.second;
          break;
        }
        case I_OPNUM: t[i.c] = do_op_num((NumTag)i.lab, t[i.a], i.rhs); break;
        case I_SWITCH: {
          auto r = do_switch(t[i.a]);
          t[i.b] = r.first;
          t[i.c] = r.second;
          break;
        }
        case I_WIRES: {
          uint64_t a = alloc();
          uint64_t b = other_half(a);
          t[i.a] = Trg::port(Port::var(a));
          t[i.b] = Trg::wire(Wire{a});
          t[i.c] = Trg::port(Port::var(b));
          t[i.d] = Trg::wire(Wire{b});
          break;
        }
      }
    }
  }

  // --- instruction.rs:110-210
  std::pair<Trg, Trg> do_ctr(Lab lab, Trg trg) {
    Port port = trg.target();
    if (port.tag() == Ctr && port.lab() == lab) {
      free_trg(trg);
      TraverseNode n = traverse(port);
      rwts.anni++;
      return {Trg::wire(n.p1), Trg::wire(n.p2)};
    }
    CreatedNode n = create_node(Ctr, lab);
    link_trg_port(trg, n.p0);
    return {Trg::port(n.p1), Trg::port(n.p2)};
  }
  std::pair<Trg, Trg> do_op(NumTag op, Trg trg) {
    Port port = trg.target();
    if (port.tag() == NumT) {
      free_trg(trg);
      CreatedNode n = create_node(Op, (Lab)Sym);
      port_wire(n.p1).set_target(Port::num(operate_unary(op, port.getnum())));
      return {Trg::port(n.p0), Trg::port(n.p2)};
    } else if (port == ERA) {
      shortcuts.op_era++;
      free_trg(trg);
      return {Trg::port(ERA), Trg::port(ERA)};
    }
    CreatedNode n = create_node(Op, (Lab)op);
    link_trg_port(trg, n.p0);
    return {Trg::port(n.p1), Trg::port(n.p2)};
  }
  Trg do_op_num(NumTag op, Trg trg, Num rhs) {
    Port port = trg.target();
    if (port.tag() == NumT) {
      rwts.oper++;
      free_trg(trg);
      return Trg::port(Port::num(operate_binary(port.getnum(), op, rhs)));
    } else if (port == ERA) {
      shortcuts.opnum_era++;
      free_trg(trg);
      return Trg::port(ERA);
    }
    CreatedNode n = create_node(Op, (Lab)op);
    link_trg_port(trg, n.p0);
    port_wire(n.p1).set_target(Port::num(rhs));
    return Trg::port(n.p2);
  }
  std::pair<Trg, Trg> do_switch(Trg trg) {
    Port port = trg.target();
    if (port.tag() == NumT) {
      rwts.oper++;
      free_trg(trg);
      Num n = port.getnum();
      uint32_t v = n.tag() == U24 ? n.get_u24() : 0;
      CreatedNode c1 = create_node(Ctr, 0);
      if (v == 0) {
        link_port_port(c1.p2, ERA);
        return {Trg::port(c1.p0), Trg::wire(create_wire_to(c1.p1))};
      }
      CreatedNode c2 = create_node(Ctr, 0);
      link_port_port(c1.p1, ERA);
      link_port_port(c1.p2, c2.p0);
      link_port_port(c2.p1, Port::num(Num::u24(v - 1)));
      return {Trg::port(c1.p0), Trg::wire(create_wire_to(c2.p2))};
    } else if (port == ERA) {
      shortcuts.switch_era++;
      rwts.eras++;
      free_trg(trg);
      return {Trg::port(ERA), Trg::port(ERA)};
    }
    CreatedNode m = create_node(Switch, 0);
    link_trg_port(trg, m.p0);
    return {Trg::port(m.p1), Trg::port(m.p2)};
  }

  // --- net.rs:72-109  (ExpandDef)
  Port new_expand(Port out) {
    Def *d = new Def();
    d->labs = LabSet::ALL();
    d->kind = 1;
    d->out = out;
    return new_ref(d);
  }
  void expand_call(Def *def, Port port) {
    if (port.tag() == Ref && port != ERA) {
      Def *other = port_def(port);
      if (other->kind == 1) {
        Port a = def->out, b = other->out;
        delete def;
        delete other;
        return link_port_port(a, b);
      }
      return call(port, new_ref(def));
    }
    Port out = def->out;
    delete def;
    switch (port.tag()) {
      case Red: throw HvmError(E_UNIMPL, "unreachable");
      case Ref:
      case NumT:
      case Var: return link_port_port(out, port);
      default: {
        TraverseNode old = traverse(port);
        CreatedNode n = create_node(old.tag, old.lab);
        link_port_port(out, n.p0);
        link_wire_port(old.p1, new_expand(n.p1));
        link_wire_port(old.p2, new_expand(n.p2));
      }
    }
  }
  void expand() {  // net.rs:65-69
    uint64_t a = alloc();
    half_free(other_half(a));
    Wire new_root{a};
    Wire old_root = root;
    root = new_root;
    link_wire_port(old_root, new_expand(Port::var(a)));
  }

  // --- net.rs:30-60
  void boot(const Def *def) { call(new_ref(def), root.as_var()); }
  bool reduce(size_t limit, size_t &count) {
    count = 0;
    Port a, b;
    while (redexes.pop(a, b)) {
      interact(a, b);
      count++;
      if (count >= limit) return false;
    }
    return true;
  }
  void normal() {
    expand();
    size_t c;
    while (!redexes.empty()) reduce(SIZE_MAX, c);
  }
  void parallel_normal(size_t threads);
};

// ---------------------------------------------------------------------------
// parallel_normal  (runtime/src/parallel.rs)
// ---------------------------------------------------------------------------
void Net::parallel_normal(size_t cores) {
  expand();
  const size_t SHARE_LIMIT = 1 << 12, LOCAL_LIMIT = 1 << 18;
  size_t tlog2 = 0;
  while ((size_t(2) << tlog2) <= cores) tlog2++;
  size_t tids = size_t(1) << tlog2;
  if (tids == 1) {  // a single thread reduces the whole net (parallel.rs with tids = 1)
    size_t c;
    while (!redexes.empty()) reduce(LOCAL_LIMIT, c);
    return;
  }
  std::vector<std::atomic<uint64_t>> share(SHARE_LIMIT * tids * 2);
  std::vector<std::atomic<size_t>> rlens(tids);
  std::atomic<size_t> total{0};
  std::barrier<> barry((std::ptrdiff_t)tids);
  std::atomic<uint64_t> d_anni{0}, d_comm{0}, d_eras{0}, d_dref{0}, d_oper{0};
  // fork (parallel.rs:12-32)
  std::vector<std::unique_ptr<Net>> nets;
  std::vector<std::unique_ptr<Heap>> areas;
  std::vector<std::pair<Port, Port>> all;
  for (auto &r : redexes.fast) all.push_back(r);
  for (auto &r : redexes.slow) all.push_back(r);
  redexes.fast.clear();
  redexes.slow.clear();
  size_t taken = 0;
  for (size_t tid = 0; tid < tids; tid++) {
    size_t heap_size = (heap->len / tids) & ~(size_t)63;
    size_t heap_start = heap_size * tid;
    areas.push_back(std::make_unique<Heap>(heap->nodes + heap_start, heap_size));
    auto n = std::make_unique<Net>(areas.back().get(), false);
    n->root = root;
    n->next = next > heap_start ? next - heap_start : 0;
    n->head = 0;  // the fresh net's own (empty) free list, as parallel.rs:25 reads it
    n->tid = tid;
    n->tids = tids;
    size_t count = all.size() / (tids - tid);  // parallel.rs:29 divides the original length
    for (size_t k = 0; k < count && taken < all.size(); k++, taken++) n->redux(all[taken].first, all[taken].second);
    nets.push_back(std::move(n));
  }
  std::vector<std::exception_ptr> errs(tids);
  auto count = [&](Net &net) -> size_t {
    barry.arrive_and_wait();
    total.store(0, std::memory_order_relaxed);
    barry.arrive_and_wait();
    rlens[net.tid].store(net.redexes.slow.size(), std::memory_order_relaxed);
    total.fetch_add(net.redexes.len(), std::memory_order_relaxed);
    barry.arrive_and_wait();
    return total.load(std::memory_order_relaxed);
  };
  auto split = [&](Net &net, size_t tick) {
    size_t plog2 = tlog2;
    size_t side = (net.tid >> (plog2 - 1 - (tick % plog2))) & 1;
    size_t shift = (size_t(1) << (plog2 - 1)) >> (tick % plog2);
    size_t a_tid = net.tid;
    size_t b_tid = side == 1 ? a_tid - shift : a_tid + shift;
    size_t a_len = net.redexes.slow.size();
    size_t b_len = rlens[b_tid].load(std::memory_order_relaxed);
    size_t send = a_len > b_len ? (a_len - b_len) / 2 : 0;
    size_t recv = b_len > a_len ? (b_len - a_len) / 2 : 0;
    send = std::min(send, SHARE_LIMIT);
    recv = std::min(recv, SHARE_LIMIT);
    for (size_t i = 0; i < send; i++) {
      size_t init = a_len - send * 2;
      auto rdx0 = net.redexes.slow[init + i * 2];
      auto rdx1 = net.redexes.slow[init + i * 2 + 1];
      size_t k = (b_tid * SHARE_LIMIT + i) * 2;
      net.redexes.slow[init + i] = rdx0;
      share[k].store(rdx1.first.v, std::memory_order_relaxed);
      share[k + 1].store(rdx1.second.v, std::memory_order_relaxed);
    }
    net.redexes.slow.resize(a_len - send);
    barry.arrive_and_wait();
    for (size_t i = 0; i < recv; i++) {
      size_t k = (a_tid * SHARE_LIMIT + i) * 2;
      net.redexes.slow.emplace_back(Port{share[k].load(std::memory_order_relaxed)},
                                    Port{share[k + 1].load(std::memory_order_relaxed)});
    }
  };
  std::vector<std::thread> ths;
  for (size_t tid = 0; tid < tids; tid++) {
    ths.emplace_back([&, tid] {
      Net &net = *nets[tid];
      size_t tick = 0;
      for (;;) {
        for (;;) {
          size_t c;
          net.reduce(LOCAL_LIMIT, c);
          if (count(net) == 0) break;
          split(net, tick);
          tick++;
        }
        if (count(net) == 0) break;
      }
      d_anni += net.rwts.anni;
      d_comm += net.rwts.comm;
      d_eras += net.rwts.eras;
      d_dref += net.rwts.dref;
      d_oper += net.rwts.oper;
    });
  }
  for (auto &t : ths) t.join();
  rwts.anni += d_anni;
  rwts.comm += d_comm;
  rwts.eras += d_eras;
  rwts.dref += d_dref;
  rwts.oper += d_oper;
  for (auto &n : nets) {
    shortcuts.op_era += n->shortcuts.op_era;
    shortcuts.opnum_era += n->shortcuts.opnum_era;
    shortcuts.switch_era += n->shortcuts.switch_era;
  }
}

// ---------------------------------------------------------------------------
// host: label sets, encoder, readback  (host/src/{host,calc_labels,encode,readback}.rs)
// ---------------------------------------------------------------------------
struct LabelCalc {  // calc_labels.rs:78-194
  const Book &book;
  struct State {
    bool done;
    size_t depth;  // Cycle(depth) when !done
    LabSet labs;
  };
  std::map<std::string, State> labels;
  // calc_labels.rs:183-190: a reference to a name that is not in the book
  // takes its label set from `lookup` (host.rs:53-59), or is an error.
  std::function<LabSet(const std::string &)> lookup;
  explicit LabelCalc(const Book &b) : book(b) {}
  static const size_t NONE = SIZE_MAX;  // "depth = None" (second pass)

  size_t visit_def(const std::string &key, size_t depth, LabSet *out) {
    auto it = labels.find(key);
    if (it == labels.end()) {
      labels[key] = State{false, depth, {}};
      return calc_def(key, depth, out);
    }
    State &st = it->second;
    if (st.done) {
      if (out) out->unite(st.labs);
      return SIZE_MAX;
    }
    if (depth != NONE) return st.depth;
    labels[key] = State{true, 0, {}};
    return calc_def(key, depth, out);
  }
  size_t calc_def(const std::string &key, size_t depth, LabSet *out) {
    LabSet labs;
    size_t head = visit_within_def(key, depth, &labs);
    if (out) out->unite(labs);
    if (depth != NONE && depth > head) {
      labels[key] = State{false, head, {}};
    } else {
      labels[key] = State{true, 0, labs};
      if (depth != NONE && depth == head) visit_within_def(key, NONE, nullptr);
    }
    return head;
  }
  size_t visit_within_def(const std::string &key, size_t depth, LabSet *out) {
    const AstNet &def = book.at(key);
    size_t head = visit_tree(def.root, depth, out);
    for (auto &r : def.redexes) {
      head = std::min(head, visit_tree(r.first, depth, out));
      head = std::min(head, visit_tree(r.second, depth, out));
    }
    return head;
  }
  size_t visit_tree(const Tree &t, size_t depth, LabSet *out) {
    if (t.kind == Tree::CtrK && out) out->add(t.lab);
    if (t.kind == Tree::RefK) {
      if (book.count(t.name)) return visit_def(t.name, depth == NONE ? NONE : depth + 1, out);
      if (!lookup) throw HvmError(E_NOREF, "Found reference @" + t.name + ", which is not in the book!");
      if (out) out->unite(lookup(t.name));
    }
    size_t m = SIZE_MAX;
    if (t.a) m = std::min(m, visit_tree(*t.a, depth, out));
    if (t.b) m = std::min(m, visit_tree(*t.b, depth, out));
    return m;
  }
  std::map<std::string, LabSet> run() {
    for (auto &kv : book) visit_def(kv.first, 0, nullptr);
    std::map<std::string, LabSet> r;
    for (auto &kv : labels) r[kv.first] = kv.second.labs;
    return r;
  }
};

struct Host {
  std::map<std::string, std::unique_ptr<Def>> defs;
  std::unordered_map<uint64_t, std::string> back;

  // host.rs:41-73 with the default of pre_reduce.rs:44-45: every name the
  // book refers to but does not define becomes an inert definition
  // (LabSet::ALL) that captures the redexes it is called on.
  void insert_book_with_inert_default(const Book &book) {
    LabelCalc calc(book);
    calc.lookup = [&](const std::string &name) {
      auto it = defs.find(name);
      if (it == defs.end()) {
        auto d = std::make_unique<Def>();
        d->labs = LabSet::ALL();
        d->kind = 2;
        back[new_ref(d.get()).addr()] = name;
        it = defs.emplace(name, std::move(d)).first;
      }
      return it->second->labs;
    };
    insert_sets(book, calc.run());
  }
  void insert_book(const Book &book) { insert_sets(book, LabelCalc(book).run()); }
  void insert_sets(const Book &book, const std::map<std::string, LabSet> &sets) {  // host.rs:41-73
    for (auto &kv : sets) {
      auto d = std::make_unique<Def>();
      d->labs = kv.second;
      back[new_ref(d.get()).addr()] = kv.first;
      defs[kv.first] = std::move(d);
    }
    for (auto &kv : book) defs[kv.first]->data = encode_def(kv.second);
  }
  Port ref_port(const std::string &name) const {
    auto it = defs.find(name);
    if (it == defs.end()) throw HvmError(E_NOREF, "Found reference @" + name + ", which is not in the book!");
    return new_ref(it->second.get());
  }

  // Encoder over InterpretedDef (encode.rs:119-163)
  struct DefEnc {
    InterpretedDef &d;
    using T = uint32_t;
    void link_const(T t, Port p) { d.instr.push_back(Instruction{I_LINKCONST, t, 0, 0, 0, 0, p, {}}); }
    void link(T a, T b) { d.instr.push_back(Instruction{I_LINK, a, b, 0, 0, 0, {}, {}}); }
    T make_const(Port p) {
      T t = d.new_trg();
      d.instr.push_back(Instruction{I_CONST, t, 0, 0, 0, 0, p, {}});
      return t;
    }
    std::pair<T, T> ctr(Lab lab, T t) {
      T p1 = d.new_trg(), p2 = d.new_trg();
      d.instr.push_back(Instruction{I_CTR, t, p1, p2, 0, lab, {}, {}});
      return {p1, p2};
    }
    std::pair<T, T> op(NumTag op, T t) {
      T r = d.new_trg(), o = d.new_trg();
      d.instr.push_back(Instruction{I_OP, t, r, o, 0, (Lab)op, {}, {}});
      return {r, o};
    }
    T op_num(NumTag op, T t, Num rhs) {
      T o = d.new_trg();
      d.instr.push_back(Instruction{I_OPNUM, t, 0, o, 0, (Lab)op, {}, rhs});
      return o;
    }
    std::pair<T, T> sw(T t) {
      T a = d.new_trg(), o = d.new_trg();
      d.instr.push_back(Instruction{I_SWITCH, t, a, o, 0, 0, {}, {}});
      return {a, o};
    }
    std::tuple<T, T, T, T> wires() {
      T av = d.new_trg(), aw = d.new_trg(), bv = d.new_trg(), bw = d.new_trg();
      d.instr.push_back(Instruction{I_WIRES, av, aw, bv, bw, 0, {}, {}});
      return {av, aw, bv, bw};
    }
  };
  // Encoder over a live Net (encode.rs:165-192)
  struct NetEnc {
    Net &n;
    using T = Trg;
    void link_const(T t, Port p) { n.link_trg_port(t, p); }
    void link(T a, T b) { n.link_trg(a, b); }
    T make_const(Port p) { return Trg::port(p); }
    std::pair<T, T> ctr(Lab lab, T t) { return n.do_ctr(lab, t); }
    std::pair<T, T> op(NumTag op, T t) { return n.do_op(op, t); }
    T op_num(NumTag op, T t, Num rhs) { return n.do_op_num(op, t, rhs); }
    std::pair<T, T> sw(T t) { return n.do_switch(t); }
    std::tuple<T, T, T, T> wires() {
      uint64_t a = n.alloc();
      uint64_t b = other_half(a);
      return {Trg::port(Port::var(a)), Trg::wire(Wire{a}), Trg::port(Port::var(b)), Trg::wire(Wire{b})};
    }
  };
  template <class E>
  struct EncState {  // encode.rs:38-104
    const Host &host;
    E &enc;
    std::map<std::string, typename E::T> scope;
    void visit_net(const AstNet &net, typename E::T t) {
      visit_tree(net.root, t);
      for (auto &r : net.redexes) visit_redex(r.first, r.second);
    }
    void visit_redex(const Tree &a, const Tree &b) {
      Port port;
      const Tree *t;
      auto nil = [&](const Tree &x, Port &p) -> bool {
        if (x.kind == Tree::Era) { p = ERA; return true; }
        if (x.kind == Tree::RefK) { p = host.ref_port(x.name); return true; }
        if (x.kind == Tree::NumK) { p = Port::num(x.num); return true; }
        return false;
      };
      // (Era, t) | (t, Era) first, then Ref, then Num (encode.rs:52-66)
      if (a.kind == Tree::Era) { port = ERA; t = &b; }
      else if (b.kind == Tree::Era) { port = ERA; t = &a; }
      else if (a.kind == Tree::RefK) { nil(a, port); t = &b; }
      else if (b.kind == Tree::RefK) { nil(b, port); t = &a; }
      else if (a.kind == Tree::NumK) { nil(a, port); t = &b; }
      else if (b.kind == Tree::NumK) { nil(b, port); t = &a; }
      else {
        auto [av, aw, bv, bw] = enc.wires();
        visit_tree(a, av);
        visit_tree(b, bv);
        enc.link(aw, bw);
        return;
      }
      auto trg = enc.make_const(port);
      visit_tree(*t, trg);
    }
    void visit_tree(const Tree &tree, typename E::T trg) {
      switch (tree.kind) {
        case Tree::Era: enc.link_const(trg, ERA); break;
        case Tree::NumK: enc.link_const(trg, Port::num(tree.num)); break;
        case Tree::RefK: enc.link_const(trg, host.ref_port(tree.name)); break;
        case Tree::CtrK: {
          auto [l, r] = enc.ctr(tree.lab, trg);
          visit_tree(*tree.a, l);
          visit_tree(*tree.b, r);
          break;
        }
        case Tree::OpK: {
          NumTag op = Sym;
          const Tree *rhs = tree.a.get(), *out = tree.b.get();
          if (rhs->kind == Tree::NumK && rhs->num.tag() == Sym && out->kind == Tree::OpK) {
            op = rhs->num.get_sym();
            rhs = out->a.get();
            out = out->b.get();
          }
          if (rhs->kind == Tree::NumK) {
            auto o = enc.op_num(op, trg, rhs->num);
            visit_tree(*out, o);
          } else {
            auto [r, o] = enc.op(op, trg);
            visit_tree(*rhs, r);
            visit_tree(*out, o);
          }
          break;
        }
        case Tree::SwitchK: {
          auto [a, o] = enc.sw(trg);
          visit_tree(*tree.a, a);
          visit_tree(*tree.b, o);
          break;
        }
        case Tree::VarK: {
          auto it = scope.find(tree.name);
          if (it != scope.end()) {
            auto other = it->second;
            scope.erase(it);
            enc.link(other, trg);
          } else {
            scope.emplace(tree.name, trg);
          }
          break;
        }
      }
    }
    void finish() {
      if (!scope.empty()) throw HvmError(E_UNBOUND, "unbound variables: " + scope.begin()->first);
    }
  };

  InterpretedDef encode_def(const AstNet &net) const {
    InterpretedDef d;
    d.trgs = 1;
    DefEnc e{d};
    EncState<DefEnc> st{*this, e, {}};
    st.visit_net(net, 0);
    st.finish();
    return d;
  }
  void encode_net(Net &n, Trg trg, const AstNet &net) const {
    NetEnc e{n};
    EncState<NetEnc> st{*this, e, {}};
    st.visit_net(net, trg);
    st.finish();
  }

  // readback.rs:23-87
  struct RB {
    const Host &host;
    std::unordered_map<uint64_t, size_t> vars;
    size_t next = 0;
    Tree read_wire(Wire w) { return read_port(w.load_target(), &w); }
    Tree read_port(Port port, const Wire *w) {
      Tree t;
      switch (port.tag()) {
        case Var:
        case Red: {
          if (!w) throw HvmError(E_UNIMPL, "var in redex");
          uint64_t key = std::min(w->a, port.addr());
          auto it = vars.find(key);
          size_t id;
          if (it != vars.end()) { id = it->second; vars.erase(it); }
          else { id = next++; vars[key] = id; }
          t.kind = Tree::VarK;
          t.name = create_var(id);
          return t;
        }
        case Ref:
          if (port == ERA) { t.kind = Tree::Era; return t; }
          t.kind = Tree::RefK;
          t.name = host.back.count(port.addr()) ? host.back.at(port.addr()) : "<expand>";
          return t;
        case NumT: t.kind = Tree::NumK; t.num = port.getnum(); return t;
        case Op: {
          NumTag op = port.op();
          TraverseNode n = traverse(port);
          Tree node;
          node.kind = Tree::OpK;
          node.a = std::make_unique<Tree>(read_wire(n.p1));
          node.b = std::make_unique<Tree>(read_wire(n.p2));
          if (op == Sym) return node;
          t.kind = Tree::OpK;
          t.a = std::make_unique<Tree>();
          t.a->kind = Tree::NumK;
          t.a->num = Num::sym(op);
          t.b = std::make_unique<Tree>(std::move(node));
          return t;
        }
        case Ctr: {
          TraverseNode n = traverse(port);
          t.kind = Tree::CtrK;
          t.lab = n.lab;
          t.a = std::make_unique<Tree>(read_wire(n.p1));
          t.b = std::make_unique<Tree>(read_wire(n.p2));
          return t;
        }
        case Switch: {
          TraverseNode n = traverse(port);
          t.kind = Tree::SwitchK;
          t.a = std::make_unique<Tree>(read_wire(n.p1));
          t.b = std::make_unique<Tree>(read_wire(n.p2));
          return t;
        }
        default: throw HvmError(E_UNIMPL, "bad port in readback");
      }
    }
  };
  AstNet readback(const Net &n) const {
    RB rb{*this, {}, 0};
    AstNet out;
    out.root = rb.read_wire(n.root);
    for (auto &r : n.redexes.fast) out.redexes.emplace_back(rb.read_port(r.first, nullptr), rb.read_port(r.second, nullptr));
    for (auto &r : n.redexes.slow) out.redexes.emplace_back(rb.read_port(r.first, nullptr), rb.read_port(r.second, nullptr));
    return out;
  }
};

static std::string pretty_num(uint64_t n) {  // util/src/pretty_num.rs
  std::string s = std::to_string(n), o;
  int cnt = 0;
  for (size_t i = s.size(); i-- > 0;) {
    o.push_back(s[i]);
    if (++cnt % 3 == 0 && i != 0) o.push_back('_');
  }
  std::reverse(o.begin(), o.end());
  return o;
}

static std::string rewrites_display(const Rewrites &r) {  // runtime.rs:131-141
  char buf[512];
  snprintf(buf, sizeof buf,
           "RWTS   : %15s\n- ANNI : %15s\n- COMM : %15s\n- ERAS : %15s\n- DREF : %15s\n- OPER : %15s\n",
           pretty_num(r.total()).c_str(), pretty_num(r.anni).c_str(), pretty_num(r.comm).c_str(),
           pretty_num(r.eras).c_str(), pretty_num(r.dref).c_str(), pretty_num(r.oper).c_str());
  return buf;
}

// Run `fn` on a thread with a large stack (deep trees recurse in encode/readback).
template <class F>
static void with_big_stack(F fn) {
  struct Ctx {
    F *f;
    std::exception_ptr err;
  } ctx{&fn, nullptr};
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, (size_t)1 << 30);
  pthread_t th;
  auto entry = [](void *p) -> void * {
    Ctx *c = static_cast<Ctx *>(p);
    try {
      (*c->f)();
    } catch (...) {
      c->err = std::current_exception();
    }
    return nullptr;
  };
  if (pthread_create(&th, &attr, entry, &ctx) != 0) throw std::runtime_error("pthread_create failed");
  pthread_join(th, nullptr);
  pthread_attr_destroy(&attr);
  if (ctx.err) std::rethrow_exception(ctx.err);
}

struct RunResult {
  std::string normal_form;
  Rewrites rwts;
  Shortcuts shortcuts;
  double seconds = 0;
  Meter meter;
  size_t peak_nodes = 0;
};

// mode: 0 = single-threaded normal(), >0 = parallel_normal() with that many cores.
static RunResult run_book(const std::string &text, const std::string &entry, size_t heap_words, int mode,
                          bool meter) {
  RunResult res;
  with_big_stack([&] {
    Book book = parse_book(text);
    Host host;
    host.insert_book(book);
    auto it = host.defs.find(entry);
    if (it == host.defs.end()) throw HvmError(E_NOREF, "no entry point @" + entry);
    Heap heap(heap_words);
    Net net(&heap);
    Meter m;
    if (meter) g_meter = &m;
    net.boot(it->second.get());
    auto t0 = std::chrono::steady_clock::now();
    if (mode <= 0) net.normal();
    else net.parallel_normal((size_t)mode);
    auto t1 = std::chrono::steady_clock::now();
    g_meter = nullptr;
    res.seconds = std::chrono::duration<double>(t1 - t0).count();
    res.rwts = net.rwts;
    res.shortcuts = net.shortcuts;
    res.meter = m;
    res.peak_nodes = net.next;
    res.normal_form = print_net(host.readback(net));
    std::free(heap.nodes);
  });
  return res;
}

// ---------------------------------------------------------------------------
// Book::pre_reduce (transform/src/pre_reduce.rs:38-164): every definition not
// skipped is booted on a fresh net (all nets share one heap, as there) and
// reduced for at most `max_rwts` interactions, its dependencies first
// (references in tree order: root, then each redex's two sides; a definition
// met again while its own dependencies are visited is left as it is); the
// redexes an inert definition captured go back into the net; the read-back
// net replaces the definition's instructions (label sets stay those of the
// original book) and, in the returned book, its net.  Missing definitions
// are inert (pre_reduce.rs:44-45).
// ---------------------------------------------------------------------------
struct PreReduceResult {
  std::string book;
  Rewrites rwts;
  std::vector<std::string> not_normal;
};

static PreReduceResult pre_reduce_book(const std::string &text, const std::function<bool(const std::string &)> &skip,
                                       size_t heap_words, uint64_t max_rwts) {
  PreReduceResult res;
  with_big_stack([&] {
    Book book = parse_book(text);
    Host host;
    host.insert_book_with_inert_default(book);
    Heap heap(heap_words);
    std::map<std::string, int> seen;  // 1: being visited (Cycled), 2: reduced
    std::map<std::string, AstNet> reduced;
    std::function<void(const std::string &)> pre;
    std::function<void(const Tree &)> visit_tree = [&](const Tree &t) {  // pre_reduce.rs:120-128
      if (t.kind == Tree::RefK) pre(t.name);
      if (t.a) visit_tree(*t.a);
      if (t.b) visit_tree(*t.b);
    };
    pre = [&](const std::string &name) {  // pre_reduce.rs:136-163
      if (seen.count(name) || skip(name) || !book.count(name)) return;
      seen[name] = 1;
      const AstNet &src = book.at(name);
      visit_tree(src.root);
      for (auto &r : src.redexes) {
        visit_tree(r.first);
        visit_tree(r.second);
      }
      Net rt(&heap);
      rt.boot(host.defs.at(name).get());
      size_t count = 0;
      const bool normal = rt.reduce((size_t)std::min<uint64_t>(max_rwts, SIZE_MAX), count);
      res.rwts.anni += rt.rwts.anni;
      res.rwts.comm += rt.rwts.comm;
      res.rwts.eras += rt.rwts.eras;
      res.rwts.dref += rt.rwts.dref;
      res.rwts.oper += rt.rwts.oper;
      for (auto &c : rt.captured) rt.redux(c.first, c.second);
      rt.captured.clear();
      AstNet net = host.readback(rt);
      host.defs.at(name)->data = host.encode_def(net);
      if (!normal) res.not_normal.push_back(name);
      reduced[name] = std::move(net);
      seen[name] = 2;
    };
    for (auto &kv : book) pre(kv.first);
    bool first = true;
    for (auto &kv : book) {  // ast.rs:221-232
      if (!first) res.book += "\n\n";
      first = false;
      auto it = reduced.find(kv.first);
      res.book += "@" + kv.first + " = " + print_net(it != reduced.end() ? it->second : kv.second);
    }
    std::free(heap.nodes);
  });
  return res;
}

// ---------------------------------------------------------------------------
// Compiled definitions (src/compile.rs:31-161): one C++ function per
// definition, the instruction list written out as straight-line calls of
// the interpreter's own do_* methods, as the reference's compiler writes
// Rust (compile.rs:110-161), with do_switch where compile.rs:150 names a
// do_mat that the runtime does not have.  The text is a translation unit
// that includes this file and reduces a book with the compiled definitions
// registered (main at the end): the CPU baseline of the compiled path.
// ---------------------------------------------------------------------------
static std::string compile_book_cpp(const std::string &text, const std::string &oracle_path) {
  std::string out;
  with_big_stack([&] {
    Book book = parse_book(text);
    Host host;
    host.insert_book(book);
    std::vector<std::string> names;
    std::map<std::string, size_t> index;
    for (auto &kv : host.defs) {
      index[kv.first] = names.size();
      names.push_back(kv.first);
    }
    auto port_text = [&](Port p) -> std::string {
      char buf[64];
      if (p.tag() == Ref && !p.is_era()) return "REF[" + std::to_string(index.at(host.back.at(p.addr()))) + "]";
      snprintf(buf, sizeof buf, "Port{0x%llxull}", (unsigned long long)p.v);
      return buf;
    };
    out += "// Generated: the definitions of a book compiled to C++ (src/compile.rs:110-161).\n";
    out += "#include \"" + oracle_path + "\"\n";
    out += "namespace oracle {\n";
    out += "static Port REF[" + std::to_string(std::max<size_t>(names.size(), 1)) + "];\n";
    for (size_t k = 0; k < names.size(); k++) {
      const InterpretedDef &d = host.defs.at(names[k])->data;
      out += "// @" + names[k] + "\n";
      out += "static void def_" + std::to_string(k) + "(Net &net, Port port) {\n";
      out += "  Trg t[" + std::to_string(std::max<uint32_t>(d.trgs, 1)) + "];\n";
      out += "  t[0] = Trg::port(port);\n";
      for (const Instruction &i : d.instr) {
        std::string a = "t[" + std::to_string(i.a) + "]", b = "t[" + std::to_string(i.b) + "]";
        std::string c = "t[" + std::to_string(i.c) + "]", dd = "t[" + std::to_string(i.d) + "]";
        char buf[256];
        switch (i.kind) {
          case I_CONST: out += "  " + a + " = Trg::port(" + port_text(i.port) + ");\n"; break;
          case I_LINK: out += "  net.link_trg(" + a + ", " + b + ");\n"; break;
          case I_LINKCONST: out += "  net.link_trg_port(" + a + ", " + port_text(i.port) + ");\n"; break;
          case I_CTR:
            snprintf(buf, sizeof buf, "  { auto r = net.do_ctr(%u, %s); %s = r.first; %s = r.second; }\n", (unsigned)i.lab,
                     a.c_str(), b.c_str(), c.c_str());
            out += buf;
            break;
          case I_OP:
            snprintf(buf, sizeof buf, "  { auto r = net.do_op((NumTag)%u, %s); %s = r.first; %s = r.second; }\n",
                     (unsigned)i.lab, a.c_str(), b.c_str(), c.c_str());
            out += buf;
            break;
          case I_OPNUM:
            snprintf(buf, sizeof buf, "  %s = net.do_op_num((NumTag)%u, %s, Num{0x%xu});\n", c.c_str(), (unsigned)i.lab,
                     a.c_str(), (unsigned)i.rhs.raw);
            out += buf;
            break;
          case I_SWITCH:
            out += "  { auto r = net.do_switch(" + a + "); " + b + " = r.first; " + c + " = r.second; }\n";
            break;
          case I_WIRES:
            out += "  { uint64_t x = net.alloc(), y = other_half(x); " + a + " = Trg::port(Port::var(x)); " + b +
                   " = Trg::wire(Wire{x}); " + c + " = Trg::port(Port::var(y)); " + dd + " = Trg::wire(Wire{y}); }\n";
            break;
        }
      }
      out += "}\n";
    }
    out += "static const char *NAMES[] = {";
    for (auto &n : names) out += "\"" + n + "\", ";
    out += "nullptr};\n";
    out += "static void (*FNS[])(Net &, Port) = {";
    for (size_t k = 0; k < names.size(); k++) out += "def_" + std::to_string(k) + ", ";
    out += "nullptr};\n";
    out += "static void register_compiled(Host &host) {\n";
    out += "  for (size_t k = 0; NAMES[k]; k++) {\n";
    out += "    Def *d = host.defs.at(NAMES[k]).get();\n";
    out += "    d->compiled = FNS[k];\n";
    out += "    REF[k] = new_ref(d);\n";
    out += "  }\n";
    out += "}\n";
    out += "}  // namespace oracle\n";
    out += "ORACLE_COMPILED_MAIN\n";
  });
  return out;
}

// The main of a compiled book: FILE [-t THREADS] [-e ENTRY] [-m WORDS]
// [-r RUNS]; prints one line per run: seconds, the five counts, and (first
// run) the normal form.  THREADS 0: normal(), else parallel_normal().
#define ORACLE_COMPILED_MAIN                                                                        \
  int main(int argc, char **argv) {                                                                \
    using namespace oracle;                                                                        \
    std::string file = argc > 1 ? argv[1] : "", entry = "main";                                    \
    int threads = 0, runs = 1;                                                                     \
    size_t words = (size_t)1 << 28;                                                                \
    for (int k = 2; k + 1 < argc; k += 2) {                                                        \
      std::string a = argv[k];                                                                     \
      if (a == "-t") threads = atoi(argv[k + 1]);                                                  \
      if (a == "-e") entry = argv[k + 1];                                                          \
      if (a == "-m") words = strtoull(argv[k + 1], nullptr, 10);                                   \
      if (a == "-r") runs = atoi(argv[k + 1]);                                                     \
    }                                                                                              \
    std::ifstream f(file);                                                                         \
    std::stringstream ss;                                                                          \
    ss << f.rdbuf();                                                                               \
    int rc = 0;                                                                                    \
    with_big_stack([&] {                                                                           \
      try {                                                                                        \
        Book book = parse_book(ss.str());                                                          \
        Host host;                                                                                 \
        host.insert_book(book);                                                                    \
        register_compiled(host);                                                                   \
        for (int run = 0; run < runs; run++) {                                                     \
          Heap heap(words);                                                                        \
          Net net(&heap);                                                                          \
          net.boot(host.defs.at(entry).get());                                                     \
          auto t0 = std::chrono::steady_clock::now();                                              \
          if (threads <= 0) net.normal();                                                          \
          else net.parallel_normal((size_t)threads);                                               \
          auto t1 = std::chrono::steady_clock::now();                                              \
          const Rewrites &r = net.rwts;                                                            \
          printf("%.6f %llu %llu %llu %llu %llu\n", std::chrono::duration<double>(t1 - t0).count(), \
                 (unsigned long long)r.anni, (unsigned long long)r.comm, (unsigned long long)r.eras,  \
                 (unsigned long long)r.dref, (unsigned long long)r.oper);                           \
          if (run == 0) printf("%s\n", print_net(host.readback(net)).c_str());                     \
          std::free(heap.nodes);                                                                   \
        }                                                                                          \
      } catch (HvmError &e) {                                                                      \
        fprintf(stderr, "error: %s\n", e.what());                                                  \
        rc = 1;                                                                                    \
      }                                                                                            \
    });                                                                                            \
    return rc;                                                                                     \
  }

// Reduce an expression net against a book, as `hvm64 reduce` does (main.rs:176-193).
static RunResult run_expr(const std::string &book_text, const std::string &expr, size_t heap_words) {
  RunResult res;
  with_big_stack([&] {
    Book book = book_text.empty() ? Book{} : parse_book(book_text);
    Host host;
    host.insert_book(book);
    AstNet net_ast = parse_net(expr);
    Heap heap(heap_words);
    Net net(&heap);
    host.encode_net(net, Trg::port(Port::var(net.root.a)), net_ast);
    net.normal();
    res.rwts = net.rwts;
    res.normal_form = print_net(host.readback(net));
    std::free(heap.nodes);
  });
  return res;
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// C ABI for the tests / bench (ctypes)
// ---------------------------------------------------------------------------
using namespace oracle;

static thread_local std::string g_err;

static int copy_out(const std::string &s, char *out, uint64_t cap, uint64_t *len) {
  if (len) *len = s.size();
  if (!out || cap == 0) return 0;
  size_t n = std::min<size_t>(s.size(), cap - 1);
  std::memcpy(out, s.data(), n);
  out[n] = 0;
  return s.size() + 1 > cap ? 7 : 0;
}

extern "C" {

const char *oracle_last_error() { return g_err.c_str(); }

// counts: anni comm eras dref oper ; extra: op_era opnum_era switch_era peak_nodes
// meter: loads stores rmws
int oracle_run(const char *book, const char *entry, uint64_t heap_words, int mode, int meter, char *out,
               uint64_t cap, uint64_t *len, uint64_t counts[5], uint64_t extra[4], uint64_t meters[3],
               double *seconds) {
  try {
    RunResult r = run_book(book, entry, heap_words, mode, meter != 0);
    counts[0] = r.rwts.anni; counts[1] = r.rwts.comm; counts[2] = r.rwts.eras;
    counts[3] = r.rwts.dref; counts[4] = r.rwts.oper;
    if (extra) {
      extra[0] = r.shortcuts.op_era; extra[1] = r.shortcuts.opnum_era;
      extra[2] = r.shortcuts.switch_era; extra[3] = r.peak_nodes;
    }
    if (meters) { meters[0] = r.meter.loads; meters[1] = r.meter.stores; meters[2] = r.meter.rmws; }
    if (seconds) *seconds = r.seconds;
    return copy_out(r.normal_form, out, cap, len);
  } catch (HvmError &e) {
    g_err = e.what();
    return 100 + e.code;
  } catch (std::exception &e) {
    g_err = e.what();
    return 99;
  }
}

int oracle_reduce_expr(const char *book, const char *expr, uint64_t heap_words, char *out, uint64_t cap,
                       uint64_t *len, uint64_t counts[5]) {
  try {
    RunResult r = run_expr(book ? book : "", expr, heap_words);
    counts[0] = r.rwts.anni; counts[1] = r.rwts.comm; counts[2] = r.rwts.eras;
    counts[3] = r.rwts.dref; counts[4] = r.rwts.oper;
    return copy_out(r.normal_form, out, cap, len);
  } catch (HvmError &e) {
    g_err = e.what();
    return 100 + e.code;
  } catch (std::exception &e) {
    g_err = e.what();
    return 99;
  }
}

// Applies each argument tree to `fun` (ast.rs apply_tree) and reduces with
// normal() on an exact-size heap, as cli.rs test_apply_tree does.
int oracle_apply_tree(const char *fun, const char *const *args, int nargs, uint64_t heap_words, char *out,
                      uint64_t cap, uint64_t *len) {
  try {
    std::string text;
    with_big_stack([&] {
      AstNet net = parse_net(fun);
      for (int k = 0; k < nargs; k++) apply_tree(net, parse_tree(args[k]));
      Host host;
      Heap heap(heap_words);
      Net rn(&heap);
      host.encode_net(rn, Trg::port(Port::var(rn.root.a)), net);
      rn.normal();
      text = print_net(host.readback(rn));
      std::free(heap.nodes);
    });
    return copy_out(text, out, cap, len);
  } catch (HvmError &e) {
    g_err = e.what();
    return 100 + e.code;
  } catch (std::exception &e) {
    g_err = e.what();
    return 99;
  }
}

// Book -> canonical text (parse + Display), for parser round-trip checks.
// Book::pre_reduce on `book`: `names` (n of them) are the definitions to
// skip, or, with only != 0, the only ones to reduce.  Writes the new book's
// text and the pre-reduction's five counts (anni comm eras dref oper).
int oracle_pre_reduce(const char *book, const char *const *names, uint32_t n, int only, uint64_t max_rwts,
                      uint64_t heap_words, char *out, uint64_t cap, uint64_t *len, uint64_t rwts[5]) {
  try {
    std::set<std::string> listed;
    for (uint32_t i = 0; i < n; i++) listed.insert(names[i]);
    auto skip = [&](const std::string &x) { return only ? listed.count(x) == 0 : listed.count(x) != 0; };
    PreReduceResult r = pre_reduce_book(book, skip, heap_words, max_rwts);
    if (rwts) {
      rwts[0] = r.rwts.anni;
      rwts[1] = r.rwts.comm;
      rwts[2] = r.rwts.eras;
      rwts[3] = r.rwts.dref;
      rwts[4] = r.rwts.oper;
    }
    return copy_out(r.book, out, cap, len);
  } catch (HvmError &e) {
    g_err = e.what();
    return 100 + e.code;
  }
}

int oracle_print_book(const char *book, char *out, uint64_t cap, uint64_t *len) {
  try {
    std::string text;
    Book b = parse_book(book);
    bool first = true;
    for (auto &kv : b) {
      if (!first) text += "\n\n";
      first = false;
      text += "@" + kv.first + " = " + print_net(kv.second);
    }
    return copy_out(text, out, cap, len);
  } catch (HvmError &e) {
    g_err = e.what();
    return 100 + e.code;
  }
}

// A book's definitions compiled to C++ (compile_book_cpp); `oracle_path` is
// the path of this file for the generated #include.
int oracle_compile_book(const char *book, const char *oracle_path, char *out, uint64_t cap, uint64_t *len) {
  try {
    return copy_out(compile_book_cpp(book, oracle_path), out, cap, len);
  } catch (HvmError &e) {
    g_err = e.what();
    return 100 + e.code;
  }
}

// Label sets (calc_labels.rs), one line per def: "name: l1 l2 ..."
int oracle_label_sets(const char *book, char *out, uint64_t cap, uint64_t *len) {
  try {
    Book b = parse_book(book);
    auto sets = LabelCalc(b).run();
    std::string text;
    for (auto &kv : sets) {
      text += kv.first + ":";
      for (uint32_t l = 0; l < kv.second.min_safe; l++)
        if (kv.second.has((Lab)l)) text += " " + std::to_string(l);
      text += "\n";
    }
    return copy_out(text, out, cap, len);
  } catch (HvmError &e) {
    g_err = e.what();
    return 100 + e.code;
  }
}

// Numeric primitives, for closed-form checks.
uint32_t oracle_operate_binary(uint32_t a, uint32_t op, uint32_t b) {
  try {
    return operate_binary(Num{a}, (NumTag)op, Num{b}).raw;
  } catch (HvmError &) {
    return 0xFFFFFFFFu;
  }
}
uint32_t oracle_operate_unary(uint32_t op, uint32_t a) { return operate_unary((NumTag)op, Num{a}).raw; }
uint32_t oracle_new_f24(float f) { return Num::f24(f).raw; }
int oracle_num_display(uint32_t raw, char *out, uint64_t cap) { return copy_out(num_display(Num{raw}), out, cap, nullptr); }
int oracle_parse_num(const char *text, uint32_t *raw) {
  try {
    Tree t = parse_tree(text);
    if (t.kind != Tree::NumK) return 1;
    *raw = t.num.raw;
    return 0;
  } catch (HvmError &e) {
    g_err = e.what();
    return 100 + e.code;
  }
}
int oracle_create_var(uint64_t id, char *out, uint64_t cap) { return copy_out(create_var(id), out, cap, nullptr); }

}  // extern "C"

#ifdef ORACLE_MAIN
// CLI: hvm64_oracle FILE [-1|-t N] [-s] [-e ENTRY] [-m WORDS]
int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s FILE [-1|-t THREADS] [-s] [-e ENTRY] [-m HEAP_WORDS]\n", argv[0]);
    return 2;
  }
  std::string file = argv[1], entry = "main";
  int mode = (int)std::thread::hardware_concurrency();
  bool stats = false;
  size_t words = (size_t)1 << 30;
  for (int k = 2; k < argc; k++) {
    std::string a = argv[k];
    if (a == "-1") mode = 0;
    else if (a == "-t" && k + 1 < argc) mode = atoi(argv[++k]);
    else if (a == "-s") stats = true;
    else if (a == "-e" && k + 1 < argc) entry = argv[++k];
    else if (a == "-m" && k + 1 < argc) words = strtoull(argv[++k], nullptr, 10);
  }
  std::ifstream f(file);
  if (!f) {
    fprintf(stderr, "Input file \"%s\" not found\n", file.c_str());
    return 1;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  try {
    RunResult r = run_book(ss.str(), entry, words, mode, false);
    printf("%s\n", r.normal_form.c_str());
    if (stats) {
      fprintf(stderr, "%s", rewrites_display(r.rwts).c_str());
      fprintf(stderr, "TIME   : %.3f s\n", r.seconds);
      fprintf(stderr, "RPS    : %.3f M\n", (double)r.rwts.total() / r.seconds / 1e6);
    }
  } catch (HvmError &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
#endif
