// Host side: text format, book loading and readback (see hvm64_host.h).
#include "hvm64_host.h"

#include <pthread.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <exception>

namespace hvm64 {

using namespace hvm;

// Spelling of every NumTag, in declaration order (num/src/num.rs:14-42); the
// parser takes the first spelling that prefixes the input, as bi_enum's
// from_str_prefix does (util/src/bi_enum.rs:44-48).
static const char *const NUM_TAG_STR[N_TAGS] = {"u24", "i24", "f24", " ",  "+",   "-",  ":-",  "*",
                                                "/",   ":/",  "%",   ":%", "&",   "|",  "^",   "<<",
                                                ":<<", ">>",  ":>>", "=",  "!",   "<",  ">"};

u32 Arena::intern(const std::string &s) {
  auto it = name_ids.find(s);
  if (it != name_ids.end()) return it->second;
  u32 id = (u32)names.size();
  names.push_back(s);
  name_ids.emplace(s, id);
  return id;
}

// ---------------------------------------------------------------------------
// numbers: Display (num.rs:287-306)
// ---------------------------------------------------------------------------

// Rust's `{:?}` for a finite f32: the shortest digits that round-trip, in
// exponent form when 0 < |x| < 1e-4 or |x| >= 1e16, otherwise positional with
// at least one fractional digit.
static std::string f32_debug(float v) {
  char buf[48];
  auto res = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
  std::string sci(buf, res.ptr);
  bool neg = sci[0] == '-';
  if (neg) sci.erase(0, 1);
  size_t e = sci.find('e');
  int exp10 = std::atoi(sci.c_str() + e + 1);
  std::string digits;
  for (size_t i = 0; i < e; i++)
    if (sci[i] != '.') digits.push_back(sci[i]);
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string out = neg ? "-" : "";
  float mag = std::fabs(v);
  if (mag == 0.0f) return out + "0.0";
  if (mag < 1e-4f || mag >= 1e16f) {
    out.push_back(digits[0]);
    if (digits.size() > 1) out += "." + digits.substr(1);
    return out + "e" + std::to_string(exp10);
  }
  int int_digits = exp10 + 1;
  if (int_digits <= 0) return out + "0." + std::string((size_t)-int_digits, '0') + digits;
  if ((size_t)int_digits >= digits.size()) return out + digits + std::string(int_digits - digits.size(), '0') + ".0";
  return out + digits.substr(0, int_digits) + "." + digits.substr(int_digits);
}

std::string show_num(u32 raw) {
  u32 t = num_tag(raw);
  switch (t) {
    case N_SYM: {
      u32 s = num_u24(raw);
      return std::string("[") + (s < N_TAGS ? NUM_TAG_STR[s] : "?") + "]";
    }
    case N_U24: return std::to_string(num_u24(raw));
    case N_I24: {
      i32 v = num_i24(raw);
      return (v >= 0 ? "+" : "") + std::to_string(v);
    }
    case N_F24: {
      float v = num_f24(raw);
      if (std::isinf(v)) return v > 0 ? "+inf" : "-inf";
      if (std::isnan(v)) return "+NaN";
      return f32_debug(v);
    }
    default: return std::string("[") + (t < N_TAGS ? NUM_TAG_STR[t] : "?") + std::to_string(num_u24(raw)) + "]";
  }
}

std::string var_name(u64 id) {
  std::string s;
  id += 1;
  while (id > 0) {
    id -= 1;
    s.push_back((char)('a' + id % 26));
    id /= 26;
  }
  std::reverse(s.begin(), s.end());
  return s;
}

// ---------------------------------------------------------------------------
// parser (ast/src/parser.rs)
// ---------------------------------------------------------------------------
namespace {
struct Reader {
  const std::string &src;
  size_t at = 0;
  Arena &arena;
  Reader(const std::string &s, Arena &a) : src(s), arena(a) {}

  [[noreturn]] void error(const char *what) const {
    throw Error(HVM_E_PARSE, std::string("parse error at byte ") + std::to_string(at) + ": expected " + what);
  }
  void trivia() {
    while (at < src.size()) {
      if (isspace((unsigned char)src[at])) {
        at++;
      } else if (src.compare(at, 2, "//") == 0) {
        while (at < src.size() && src[at] != '\n') at++;
      } else {
        break;
      }
    }
  }
  bool eat(const char *tok) {
    trivia();
    size_t n = std::strlen(tok);
    if (src.compare(at, n, tok) != 0) return false;
    at += n;
    return true;
  }
  void expect(const char *tok) {
    if (!eat(tok)) error(tok);
  }
  int peek() const { return at < src.size() ? (unsigned char)src[at] : -1; }
  static bool name_char(int c) { return isalnum(c) || c == '_' || c == '.' || c == '$'; }
  std::string name() {
    size_t s = at;
    while (at < src.size() && name_char((unsigned char)src[at])) at++;
    if (s == at) error("name");
    return src.substr(s, at - s);
  }

  u32 tree() {
    trivia();
    int c = peek();
    TNode n{};
    if (c == '*') {
      at++;
      n.kind = TK::Era;
      return arena.add(n);
    }
    if (c == '(' || c == '{' || c == '#') {
      at++;
      u32 label = 0;
      if (c == '#') {
        trivia();
        size_t s = at;
        while (at < src.size() && isdigit((unsigned char)src[at])) at++;
        if (s == at) error("label");
        label = (u32)std::stoull(src.substr(s, at - s));
        trivia();
        c = peek();
        at++;
      }
      const char *close = c == '(' ? ")" : c == '{' ? "}" : nullptr;
      if (!close) error("`(` or `{`");
      n.kind = TK::Ctr;
      n.lab = (u16)(label * 2 + (c == '{' ? 1 : 0));
      u32 a = tree();
      u32 b = tree();
      expect(close);
      n.a = a;
      n.b = b;
      return arena.add(n);
    }
    if (c == '@') {
      at++;
      trivia();
      n.kind = TK::Ref;
      n.name = arena.intern(name());
      return arena.add(n);
    }
    if (c == '$' || c == '?') {
      expect(c == '$' ? "$(" : "?(");
      n.kind = c == '$' ? TK::Op : TK::Switch;
      u32 a = tree();
      u32 b = tree();
      expect(")");
      n.a = a;
      n.b = b;
      return arena.add(n);
    }
    if ((c >= '0' && c <= '9') || c == '+' || c == '-' || c == '[') {
      n.kind = TK::Num;
      n.num = number();
      return arena.add(n);
    }
    n.kind = TK::Var;
    n.name = arena.intern(name());
    return arena.add(n);
  }

  u32 number() {  // parser.rs:116-140
    trivia();
    if (peek() != '[') return literal();
    expect("[");
    int op = -1;
    for (int k = 0; k < (int)N_TAGS; k++) {
      size_t len = std::strlen(NUM_TAG_STR[k]);
      if (src.compare(at, len, NUM_TAG_STR[k]) == 0) {
        op = k;
        break;
      }
    }
    if (op < 0) error("operator");
    at += std::strlen(NUM_TAG_STR[op]);
    trivia();
    if ((u32)op < N_SYM || peek() == ']') {
      expect("]");
      return num_new_sym((u32)op);
    }
    u32 lhs = literal();
    expect("]");
    return operate_unary((u32)op, lhs);
  }

  u64 integer(const std::string &s) {
    int base = 10;
    size_t off = 0;
    if (s.rfind("0x", 0) == 0) base = 16, off = 2;
    else if (s.rfind("0b", 0) == 0) base = 2, off = 2;
    if (off >= s.size()) error("integer");
    u64 v = 0;
    auto r = std::from_chars(s.data() + off, s.data() + s.size(), v, base);
    if (r.ec != std::errc() || r.ptr != s.data() + s.size()) error("integer");
    return v;
  }

  u32 literal() {  // parser.rs:142-154
    size_t s = at;
    while (at < src.size() && (isalnum((unsigned char)src[at]) || src[at] == '+' || src[at] == '-' || src[at] == '.'))
      at++;
    std::string lit = src.substr(s, at - s);
    if (lit.find('.') != std::string::npos || lit.find("inf") != std::string::npos ||
        lit.find("NaN") != std::string::npos) {
      char *end = nullptr;
      float f = std::strtof(lit.c_str(), &end);
      if (end != lit.c_str() + lit.size()) error("float");
      return num_new_f24(f);
    }
    if (!lit.empty() && (lit[0] == '+' || lit[0] == '-')) {
      u32 mag = (u32)integer(lit.substr(1));
      return num_new_i24((i32)(lit[0] == '-' ? 0u - mag : mag));
    }
    return num_new_u24((u32)integer(lit));
  }

  Net net() {
    Net n;
    n.root = tree();
    while (eat("&")) {
      u32 a = tree();
      expect("~");
      u32 b = tree();
      n.redexes.emplace_back(a, b);
    }
    return n;
  }

  void finish() {
    trivia();
    if (at != src.size()) throw Error(HVM_E_PARSE, "Unable to parse the whole input. Is this not an hvm64 file?");
  }
};
}  // namespace

Book parse_book(const std::string &text) {
  Book book;
  Reader r(text, book.arena);
  while (r.eat("@")) {
    std::string nm = r.name();
    r.expect("=");
    book.nets[nm] = r.net();
  }
  r.finish();
  return book;
}

Net parse_net(Arena &arena, const std::string &text) {
  Reader r(text, arena);
  Net n = r.net();
  r.finish();
  return n;
}

u32 parse_tree(Arena &arena, const std::string &text) {
  Reader r(text, arena);
  u32 t = r.tree();
  r.finish();
  return t;
}

// ---------------------------------------------------------------------------
// printer (ast.rs:221-263), iterative
// ---------------------------------------------------------------------------
static void ctr_brackets(u32 lab, std::string &open, const char *&close) {
  if (lab == 0) open = "(";
  else if (lab == 1) open = "{";
  else open = "#" + std::to_string(lab / 2) + (lab % 2 == 0 ? "(" : "{");
  close = lab % 2 == 0 ? ")" : "}";
}

std::string show_tree(const Arena &A, u32 root) {
  std::string out;
  struct Item {
    u32 node;
    const char *text;  // when non-null, emit this literal instead of a node
  };
  std::vector<Item> stack{{root, nullptr}};
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    if (it.text) {
      out += it.text;
      continue;
    }
    const TNode &n = A.nodes[it.node];
    switch (n.kind) {
      case TK::Era: out += "*"; break;
      case TK::Num: out += show_num(n.num); break;
      case TK::Ref: out += "@" + A.names[n.name]; break;
      case TK::Var: out += A.names[n.name]; break;
      case TK::Ctr: {
        std::string open;
        const char *close;
        ctr_brackets(n.lab, open, close);
        out += open;
        stack.push_back({0, close});
        stack.push_back({n.b, nullptr});
        stack.push_back({0, " "});
        stack.push_back({n.a, nullptr});
        break;
      }
      case TK::Op:
      case TK::Switch:
        out += n.kind == TK::Op ? "$(" : "?(";
        stack.push_back({0, ")"});
        stack.push_back({n.b, nullptr});
        stack.push_back({0, " "});
        stack.push_back({n.a, nullptr});
        break;
    }
  }
  return out;
}

std::string show_net(const Arena &A, const Net &net) {
  std::string s = show_tree(A, net.root);
  for (auto &r : net.redexes) s += "\n  & " + show_tree(A, r.first) + " ~ " + show_tree(A, r.second);
  return s;
}

std::string show_book(const Book &book) {
  std::string s;
  bool first = true;
  for (auto &kv : book.nets) {
    if (!first) s += "\n\n";
    first = false;
    s += "@" + kv.first + " = " + show_net(book.arena, kv.second);
  }
  return s;
}

// apply_tree (ast.rs:112-123) with ensure_no_conflicts (ast.rs:125-132, 169-176).
static bool var_index(const std::string &s, u64 &out) {  // util/src/create_var.rs:21-33
  u64 n = 0;
  for (char ch : s) {
    if (ch < 'a') return false;
    u64 i = (u64)(ch - 'a');
    if (i > 'z') return false;
    n = n * 26 + i + 1;
  }
  if (n == 0) return false;
  out = n - 1;
  return true;
}

void apply_tree(Arena &A, Net &net, u32 arg) {
  u64 fresh = 0;
  std::vector<u32> todo{net.root, arg};
  for (auto &r : net.redexes) {
    todo.push_back(r.first);
    todo.push_back(r.second);
  }
  while (!todo.empty()) {
    u32 t = todo.back();
    todo.pop_back();
    const TNode &n = A.nodes[t];
    if (n.kind == TK::Var) {
      u64 v;
      if (var_index(A.names[n.name], v)) fresh = std::max(fresh, v);
    }
    if (n.kind == TK::Ctr || n.kind == TK::Op || n.kind == TK::Switch) {
      todo.push_back(n.a);
      todo.push_back(n.b);
    }
  }
  u32 fv = A.intern(var_name(fresh + 1));
  TNode var{};
  var.kind = TK::Var;
  var.name = fv;
  u32 out_var = A.add(var);
  TNode app{};
  app.kind = TK::Ctr;
  app.lab = 0;
  app.a = arg;
  app.b = out_var;
  u32 app_id = A.add(app);
  u32 root_var = A.add(var);
  net.redexes.emplace_back(net.root, app_id);
  net.root = root_var;
}

// ---------------------------------------------------------------------------
// label sets (calc_labels.rs:78-194): two-phase memoised traversal
// ---------------------------------------------------------------------------
void LabSet::add(u16 lab) {
  min_safe = std::max<u32>(min_safe, (u32)lab + 1);
  size_t w = lab >> 6;
  if (w >= bits.size()) bits.resize(w + 1, 0);
  bits[w] |= 1ull << (lab & 63);
}
void LabSet::unite(const LabSet &o) {
  min_safe = std::max(min_safe, o.min_safe);
  if (o.bits.size() > bits.size()) bits.resize(o.bits.size(), 0);
  for (size_t i = 0; i < o.bits.size(); i++) bits[i] |= o.bits[i];
}
LabSet LabSet::all() {
  LabSet s;
  s.min_safe = 0xFFFF;
  s.bits.assign(1024, ~0ull);
  return s;
}
bool LabSet::has(u16 lab) const {
  if (lab >= min_safe) return false;
  size_t w = lab >> 6;
  return w < bits.size() && ((bits[w] >> (lab & 63)) & 1);
}

namespace {
struct LabelPass {
  const Book &book;
  const std::function<LabSet(const std::string &)> &missing;
  static constexpr size_t SECOND = SIZE_MAX;  // depth "None": the second phase
  struct Memo {
    bool done;
    size_t head;  // head depth of the cycle this def sits in (while !done)
    LabSet set;
  };
  std::map<std::string, Memo> memo;

  size_t def(const std::string &key, size_t depth, LabSet *out) {
    auto it = memo.find(key);
    if (it == memo.end()) {
      memo[key] = Memo{false, depth, {}};
      return calc(key, depth, out);
    }
    if (it->second.done) {
      if (out) out->unite(it->second.set);
      return SIZE_MAX;
    }
    if (depth != SECOND) return it->second.head;
    it->second = Memo{true, 0, {}};
    return calc(key, depth, out);
  }
  size_t calc(const std::string &key, size_t depth, LabSet *out) {
    LabSet labs;
    size_t head = within(key, depth, &labs);
    if (out) out->unite(labs);
    if (depth != SECOND && depth > head) {
      memo[key] = Memo{false, head, {}};
    } else {
      memo[key] = Memo{true, 0, labs};
      if (depth != SECOND && depth == head) within(key, SECOND, nullptr);
    }
    return head;
  }
  size_t within(const std::string &key, size_t depth, LabSet *out) {
    const Net &n = book.nets.at(key);
    size_t head = tree(n.root, depth, out);
    for (auto &r : n.redexes) {
      head = std::min(head, tree(r.first, depth, out));
      head = std::min(head, tree(r.second, depth, out));
    }
    return head;
  }
  size_t tree(u32 t, size_t depth, LabSet *out) {
    const TNode &n = book.arena.nodes[t];
    if (n.kind == TK::Ctr && out) out->add(n.lab);
    if (n.kind == TK::Ref) {
      const std::string &nm = book.arena.names[n.name];
      if (book.nets.count(nm)) return def(nm, depth == SECOND ? SECOND : depth + 1, out);
      if (!missing) throw Error(HVM_E_NOREF, "Found reference @" + nm + ", which is not in the book!");
      if (out) out->unite(missing(nm));
      return SIZE_MAX;
    }
    if (n.kind == TK::Ctr || n.kind == TK::Op || n.kind == TK::Switch)
      return std::min(tree(n.a, depth, out), tree(n.b, depth, out));
    return SIZE_MAX;
  }
};
}  // namespace

std::map<std::string, LabSet> calculate_label_sets(const Book &book,
                                                   const std::function<LabSet(const std::string &)> &missing) {
  LabelPass pass{book, missing, {}};
  for (auto &kv : book.nets) pass.def(kv.first, 0, nullptr);
  std::map<std::string, LabSet> out;
  for (auto &kv : pass.memo) out[kv.first] = kv.second.set;
  return out;
}

// ---------------------------------------------------------------------------
// encoder (encode.rs:38-163) + register compaction
// ---------------------------------------------------------------------------
namespace {
struct RawInstr {
  u32 kind, a, b, c, d;  // TrgIds (full width)
  u64 w1;                // label / op / port / rhs
};

struct Enc {
  const Host &host;
  const Arena &A;
  std::vector<RawInstr> code;
  u32 trgs = 1;
  std::map<u32, u32> scope;  // var name id -> trg

  u32 fresh() { return trgs++; }
  void emit(u32 kind, u32 a, u32 b, u32 c, u32 d, u64 w1) { code.push_back(RawInstr{kind, a, b, c, d, w1}); }
  u64 nil_port(const TNode &n) const {
    if (n.kind == TK::Era) return P_ERA;
    if (n.kind == TK::Ref) return host.ref_port(A.names[n.name]);
    return numport(n.num);
  }
  void net(const Net &n, u32 t) {
    tree(n.root, t);
    for (auto &r : n.redexes) redex(r.first, r.second);
  }
  void redex(u32 a, u32 b) {  // encode.rs:52-66
    const TNode &x = A.nodes[a], &y = A.nodes[b];
    u32 nil, other;
    if (x.kind == TK::Era) nil = a, other = b;
    else if (y.kind == TK::Era) nil = b, other = a;
    else if (x.kind == TK::Ref) nil = a, other = b;
    else if (y.kind == TK::Ref) nil = b, other = a;
    else if (x.kind == TK::Num) nil = a, other = b;
    else if (y.kind == TK::Num) nil = b, other = a;
    else {
      u32 av = fresh(), aw = fresh(), bv = fresh(), bw = fresh();
      emit(I_WIRES, av, aw, bv, bw, 0);
      tree(a, av);
      tree(b, bv);
      emit(I_LINK, aw, bw, 0, 0, 0);
      return;
    }
    u32 t = fresh();
    emit(I_CONST, t, 0, 0, 0, nil_port(A.nodes[nil]));
    tree(other, t);
  }
  void tree(u32 id, u32 t) {  // encode.rs:68-104
    const TNode &n = A.nodes[id];
    switch (n.kind) {
      case TK::Era:
      case TK::Num:
      case TK::Ref: emit(I_LINKCONST, t, 0, 0, 0, nil_port(n)); break;
      case TK::Ctr: {
        u32 p1 = fresh(), p2 = fresh();
        emit(I_CTR, t, p1, p2, 0, n.lab);
        tree(n.a, p1);
        tree(n.b, p2);
        break;
      }
      case TK::Op: {
        u32 op = N_SYM, rhs = n.a, out = n.b;
        const TNode &r = A.nodes[n.a], &o = A.nodes[n.b];
        if (r.kind == TK::Num && num_tag(r.num) == N_SYM && o.kind == TK::Op) {
          op = num_u24(r.num);
          rhs = o.a;
          out = o.b;
        }
        const TNode &rn = A.nodes[rhs];
        if (rn.kind == TK::Num) {
          u32 o2 = fresh();
          emit(I_OPNUM, t, 0, o2, 0, (u64)op | ((u64)rn.num << 32));
          tree(out, o2);
        } else {
          u32 r2 = fresh(), o2 = fresh();
          emit(I_OP, t, r2, o2, 0, op);
          tree(rhs, r2);
          tree(out, o2);
        }
        break;
      }
      case TK::Switch: {
        u32 arms = fresh(), out = fresh();
        emit(I_SWITCH, t, arms, out, 0, 0);
        tree(n.a, arms);
        tree(n.b, out);
        break;
      }
      case TK::Var: {
        auto it = scope.find(n.name);
        if (it != scope.end()) {
          u32 other = it->second;
          scope.erase(it);
          emit(I_LINK, other, t, 0, 0, 0);
        } else {
          scope.emplace(n.name, t);
        }
        break;
      }
    }
  }
};

// Every TrgId is written by one instruction and read by exactly one later
// instruction, so registers can be reused: an instruction's inputs are
// released before its outputs are assigned (the interpreter reads every input
// before it writes an output).
struct SlotMap {
  std::map<u32, u32> where;        // live TrgId -> slot
  std::vector<u32> free_slots;     // min-heap of released slots
  u32 nslots = 1;
  SlotMap() { where[0] = 0; }
  u32 read(u32 id) {
    auto it = where.find(id);
    if (it == where.end()) throw Error(HVM_E_UNBOUND, "trg register read before it was written");
    u32 s = it->second;
    where.erase(it);
    free_slots.push_back(s);
    std::push_heap(free_slots.begin(), free_slots.end(), std::greater<u32>());
    return s;
  }
  u32 write(u32 id) {
    u32 s;
    if (!free_slots.empty()) {
      std::pop_heap(free_slots.begin(), free_slots.end(), std::greater<u32>());
      s = free_slots.back();
      free_slots.pop_back();
    } else {
      s = nslots++;
    }
    where[id] = s;
    return s;
  }
};

DefCode compact(const std::vector<RawInstr> &raw) {
  DefCode out;
  SlotMap m;
  out.instr.reserve(raw.size() * 2);
  for (const RawInstr &r : raw) {
    u32 a = 0, b = 0, c = 0, d = 0;
    switch (r.kind) {
      case I_CONST: a = m.write(r.a); break;
      case I_LINK: a = m.read(r.a); b = m.read(r.b); break;
      case I_LINKCONST: a = m.read(r.a); break;
      case I_CTR:
      case I_OP:
      case I_SWITCH: a = m.read(r.a); b = m.write(r.b); c = m.write(r.c); break;
      case I_OPNUM: a = m.read(r.a); c = m.write(r.c); break;
      case I_WIRES: a = m.write(r.a); b = m.write(r.b); c = m.write(r.c); d = m.write(r.d); break;
    }
    if (m.nslots > 0xFFFF) throw Error(HVM_E_ARG, "definition needs more than 65535 live trg registers");
    out.instr.push_back(instr_w0(r.kind, a, b, c));
    out.instr.push_back(r.kind == I_WIRES ? (u64)d : r.w1);
  }
  out.nslots = m.nslots;
  return out;
}
}  // namespace

DefCode Host::encode(const Arena &A, const Net &n) const {
  Enc e{*this, A, {}, 1, {}};
  e.net(n, 0);
  if (!e.scope.empty()) throw Error(HVM_E_UNBOUND, "unbound variable " + A.names[e.scope.begin()->first]);
  return compact(e.code);
}

void Host::insert_book(const Book &book) {  // host.rs:41-73
  insert_sets(book, calculate_label_sets(book));
}

void Host::insert_book_with_inert_default(const Book &book) {
  // the default definition is inserted when the label pass first meets the
  // name, as host.rs:53-59 does
  auto sets = calculate_label_sets(book, [&](const std::string &name) {
    auto it = index_.find(name);
    if (it != index_.end()) return labs_[it->second];
    const u32 id = (u32)names_.size();
    names_.push_back(name);
    index_[name] = id;
    labs_.push_back(LabSet::all());
    code_.emplace_back();
    inert_.push_back(true);
    return labs_[id];
  });
  insert_sets(book, sets);
}

void Host::insert_sets(const Book &book, const std::map<std::string, LabSet> &sets) {
  for (auto &kv : sets) {
    if (index_.count(kv.first)) throw Error(HVM_E_ARG, "definition @" + kv.first + " already loaded");
    u32 id = (u32)names_.size();
    names_.push_back(kv.first);
    index_[kv.first] = id;
    labs_.push_back(kv.second);
    code_.emplace_back();
    inert_.push_back(false);
  }
  for (auto &kv : book.nets) code_[index_.at(kv.first)] = encode(book.arena, kv.second);
}

u32 Host::def_id(const std::string &name) const {
  auto it = index_.find(name);
  if (it == index_.end()) throw Error(HVM_E_NOREF, "no definition @" + name);
  return it->second;
}

u64 Host::ref_port(const std::string &name) const {
  auto it = index_.find(name);
  if (it == index_.end()) throw Error(HVM_E_NOREF, "Found reference @" + name + ", which is not in the book!");
  return mkport(T_REF, labs_[it->second].min_safe, def_addr(it->second));
}

std::vector<DefRec> Host::def_records() const {
  std::vector<DefRec> recs;
  u32 ioff = 0, loff = 0;
  for (size_t i = 0; i < names_.size(); i++) {
    DefRec r{};
    r.ioff = ioff;
    r.ilen = code_[i].ninstr();
    r.nslots = code_[i].nslots;
    r.labs_off = loff;
    r.labs_words = (u32)labs_[i].bits.size();
    r.min_safe = labs_[i].min_safe;
    r.flags = inert_[i] ? hvm::DEF_INERT : 0;
    ioff += r.ilen;
    loff += r.labs_words;
    recs.push_back(r);
  }
  return recs;
}

std::vector<u64> Host::all_instrs() const {
  std::vector<u64> v;
  for (auto &c : code_) v.insert(v.end(), c.instr.begin(), c.instr.end());
  return v;
}

std::vector<u64> Host::all_labbits() const {
  std::vector<u64> v;
  for (auto &l : labs_) v.insert(v.end(), l.bits.begin(), l.bits.end());
  return v;
}

void Host::adopt(std::vector<std::string> names, std::vector<LabSet> labs, std::vector<DefCode> code) {
  names_ = std::move(names);
  labs_ = std::move(labs);
  code_ = std::move(code);
  inert_.assign(names_.size(), false);
  index_.clear();
  for (u32 i = 0; i < names_.size(); i++) {
    if (index_.count(names_[i])) throw Error(HVM_E_ARG, "duplicate definition @" + names_[i]);
    index_[names_[i]] = i;
  }
}

u32 Host::max_slots() const {
  u32 m = 1;
  for (auto &c : code_) m = std::max(m, c.nslots);
  return m;
}

// ---------------------------------------------------------------------------
// readback (readback.rs:23-87) from the device's preorder stream
// ---------------------------------------------------------------------------
std::string readback_text(const Host &host, const std::vector<u64> &stream, size_t &pos,
                          std::map<u64, u64> &vars, u64 &next_var) {
  std::string out;
  std::vector<const char *> closers;  // pending text once a node's children are done
  std::vector<int> left;              // children still to print per open node
  auto child_done = [&]() {
    while (!left.empty()) {
      if (--left.back() == 1) {
        out += " ";
        return;
      }
      out += closers.back();
      closers.pop_back();
      left.pop_back();
    }
  };
  do {
    if (pos >= stream.size()) throw Error(HVM_E_READBACK, "truncated readback stream");
    u64 p = stream[pos++];
    switch (ptag(p)) {
      case T_VAR:
      case T_RED: {
        u64 key = paddr(p);
        auto it = vars.find(key);
        u64 id;
        if (it != vars.end()) {
          id = it->second;
          vars.erase(it);
        } else {
          id = next_var++;
          vars[key] = id;
        }
        out += var_name(id);
        child_done();
        break;
      }
      case T_REF:
        if (p == P_ERA) out += "*";
        else if (addr_in_heap(paddr(p))) out += "@<expand>";
        else out += "@" + host.def_name(def_index(paddr(p)));
        child_done();
        break;
      case T_NUM:
        out += show_num(pnum(p));
        child_done();
        break;
      case T_CTR: {
        std::string open;
        const char *close;
        ctr_brackets(plab(p), open, close);
        out += open;
        closers.push_back(close);
        left.push_back(2);
        break;
      }
      case T_OP:
        if (plab(p) == N_SYM) {
          out += "$(";
          closers.push_back(")");
        } else {
          out += std::string("$([") + (plab(p) < N_TAGS ? NUM_TAG_STR[plab(p)] : "?") + "] $(";
          closers.push_back("))");
        }
        left.push_back(2);
        break;
      case T_SWI:
        out += "?(";
        closers.push_back(")");
        left.push_back(2);
        break;
      default: throw Error(HVM_E_READBACK, "bad port in readback stream");
    }
  } while (!left.empty());
  return out;
}

void with_big_stack(const std::function<void()> &fn) {
  struct Ctx {
    const std::function<void()> *fn;
    std::exception_ptr err;
  } ctx{&fn, nullptr};
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, (size_t)512 << 20);
  pthread_t th;
  auto entry = [](void *p) -> void * {
    Ctx *c = static_cast<Ctx *>(p);
    try {
      (*c->fn)();
    } catch (...) {
      c->err = std::current_exception();
    }
    return nullptr;
  };
  if (pthread_create(&th, &attr, entry, &ctx) != 0) {
    pthread_attr_destroy(&attr);
    fn();
    return;
  }
  pthread_join(th, nullptr);
  pthread_attr_destroy(&attr);
  if (ctx.err) std::rethrow_exception(ctx.err);
}

}  // namespace hvm64
