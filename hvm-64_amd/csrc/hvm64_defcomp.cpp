// Definition compiler: splits an interpreted definition (def.rs:217-289,
// instruction.rs:110-210) into a residual program and a template.
//
// A definition's instructions run against the port that met its Ref (t0).
// Every instruction whose operand does not derive from t0 behaves the same
// on every call: its fast path (or slow path) is decided by constants and by
// nodes and wires the definition itself creates.  Those instructions are
// evaluated here, once, on an abstract heap:
//
//   * nodes they create become template slots;
//   * var cells joining two template words stay cells (P_NONE); a cell whose
//     far side was linked to a value is dissolved and the value is written
//     into the node word instead (the engine's var-cell protocol would park
//     it there and hand it over on the second arrival);
//   * links of two principal ports become template redexes (or an erasure,
//     counted, when both are skippable, linker.rs:70-85);
//   * the counts of the fast paths they take (anni / oper / eras and the
//     three ERA shortcuts) are added once per call.
//
// The instructions whose operand derives from t0 (the pattern the definition
// matches against its argument) are kept, in order, as the residual program,
// with the reference's fast paths intact.  Their results that the template
// needs are written into template words as register values ("holes"), and
// links between them and template objects follow the template.
//
// Each call then runs: the residual instructions; one I_TEMPLATE that writes
// every slot with single stores and pushes the template redexes; the
// remaining links.  Rewrite counts are those of the reference: every static
// instruction counts exactly what its interpretation would.
#include <algorithm>
#include <tuple>
#include <map>
#include <stdexcept>
#include <vector>

#include "hvm64_host.h"

namespace hvm64 {

using namespace hvm;

namespace {

enum class AK : uint8_t { Lit, Node, End, Wire, Dyn, DynW };
struct AV {
  AK k;
  u64 v;  // Lit: port; Node: node index; End / Wire: cell index; Dyn / DynW: token
};
AV lit(u64 p) { return {AK::Lit, p}; }

enum class CS : uint8_t { Empty, Parked, Done };
struct Cell {
  CS st = CS::Empty;
  AV parked{AK::Lit, 0};
};

struct ANode {
  u32 tag, lab;
  AV w[2];
  bool alive = true;
};

// Residual operation on runtime tokens.
struct ROp {
  u32 kind;
  bool aswire = false;
  u64 in = 0;            // input token
  u64 out1 = ~0ull, out2 = ~0ull;
  u64 w1 = 0;            // label / op / rhs / port
  AV other{AK::Lit, 0};  // links: the second operand
};

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Compiler {
  std::vector<AV> T;
  std::vector<Cell> cells;
  std::vector<ANode> nodes;
  std::vector<std::pair<AV, AV>> redexes;
  // Order keys of the redexes: the interpreter defers a definition's links
  // to a list of LINK_CAP entries and links a full list the last four
  // first, so its redexes appear in that order; the template pushes its
  // redexes in the same order (seq: the link request that made the redex,
  // counted in instruction order; depth: continuations after it).
  std::vector<std::pair<u64, u64>> redex_seq;
  u64 link_seq = 0, cur_seq = 0;
  std::vector<ROp> pattern;  // residual instructions (before the template)
  std::vector<ROp> links;    // residual links (after the template)
  u64 ntok = 1;              // token 0 = t0
  u64 cnt[8] = {};           // anni comm eras dref oper sc_op sc_opnum sc_switch

  static bool principal(const AV &a) { return a.k == AK::Lit || a.k == AK::Node; }
  static AV norm(AV a) {
    if (a.k == AK::Wire) a.k = AK::End;
    if (a.k == AK::DynW) a.k = AK::Dyn;
    return a;
  }
  static AV as_wire(AV a) {
    if (a.k == AK::End) a.k = AK::Wire;
    if (a.k == AK::Dyn) a.k = AK::DynW;
    return a;
  }
  bool skippable(const AV &a) const { return a.k == AK::Lit && is_skippable(a.v); }

  u64 new_cell() {
    cells.emplace_back();
    return cells.size() - 1;
  }
  u64 new_node(u32 tag, u32 lab, AV w0, AV w1) {
    ANode n;
    n.tag = tag;
    n.lab = lab;
    n.w[0] = w0;
    n.w[1] = w1;
    nodes.push_back(n);
    return nodes.size() - 1;
  }
  AV node_port(u64 i) { return {AK::Node, i}; }

  // The engine's link (hvm64_rules.h link_begin / flush_links): two
  // principal ports make a redex; otherwise the value is parked in an
  // endpoint's var cell, or, when the cell already holds the far side's
  // value, the two are linked on.  Runtime tokens that meet a principal
  // port or another token are linked at run time.
  void link(AV a, AV b, int depth = 0) {
    if (depth > 1 << 20) throw Fail("link chain");
    a = norm(a);
    b = norm(b);
    if (principal(a) && principal(b)) {
      if (skippable(a) && skippable(b)) cnt[2]++;
      else {
        redexes.push_back({a, b});
        redex_seq.push_back({cur_seq, (u64)depth});
      }
      return;
    }
    if (a.k != AK::End && b.k == AK::End) std::swap(a, b);
    if (a.k == AK::End) {
      Cell &c = cells[a.v];
      if (c.st == CS::Empty) {
        c.st = CS::Parked;
        c.parked = b;
        return;
      }
      if (c.st == CS::Parked) {
        c.st = CS::Done;
        AV z = c.parked;
        link(z, b, depth + 1);
        return;
      }
      throw Fail("cell linked three times");
    }
    // a or b is a runtime token, the other a token or a principal port
    if (a.k != AK::Dyn) std::swap(a, b);
    ROp r;
    r.kind = I_LINK;
    r.in = a.v;
    r.other = b;
    links.push_back(r);
  }

  // A link an instruction asks for (one entry of the interpreter's list).
  void ilink(AV a, AV b) {
    cur_seq = link_seq++;
    link(a, b);
  }

  // The interpreter's order of the redexes: full lists of LINK_CAP links in
  // turn; within a list the last four links first, then the rest (and the
  // continuations of earlier rounds).
  std::vector<size_t> redex_order() const {
    const u64 cap = 6;
    std::vector<std::tuple<u64, u64, u64, size_t>> keys;
    for (size_t i = 0; i < redexes.size(); i++) {
      const u64 s = redex_seq[i].first, d = redex_seq[i].second;
      const u64 batch = s / cap, pos = s % cap;
      const u64 m = std::min<u64>(cap, link_seq - batch * cap);
      const u64 round = (pos + 4 >= m ? 0 : 1) + d;
      keys.push_back({batch, round, pos, i});
    }
    std::sort(keys.begin(), keys.end());
    std::vector<size_t> order;
    for (auto &k : keys) order.push_back(std::get<3>(k));
    return order;
  }

  // Trg::target of an operand: a principal port, an unresolved endpoint
  // (returned as End(c)), or a runtime token (Dyn / DynW).
  AV resolve(AV x) {
    for (int hop = 0; hop < 1 << 20; hop++) {
      if (x.k != AK::Wire) return x;
      Cell &c = cells[x.v];
      if (c.st == CS::Empty) return {AK::End, x.v};
      if (c.st == CS::Done) throw Fail("wire followed twice");
      c.st = CS::Done;
      AV z = c.parked;
      if (z.k == AK::End) {
        x = {AK::Wire, z.v};
        continue;
      }
      if (z.k == AK::Dyn) return {AK::DynW, z.v};
      return z;
    }
    throw Fail("resolve chain");
  }

  u64 tok() { return ntok++; }

  void residual(u32 kind, const AV &in, u64 w1, u32 b, u32 c, bool two_outputs) {
    ROp r;
    r.kind = kind;
    r.in = in.v;
    r.aswire = in.k == AK::DynW;
    r.w1 = w1;
    r.out1 = tok();
    if (two_outputs) r.out2 = tok();
    pattern.push_back(r);
    if (kind == I_OPNUM) {
      T[c] = {AK::Dyn, r.out1};
    } else {
      T[b] = {AK::Dyn, r.out1};
      T[c] = {AK::Dyn, r.out2};
    }
  }

  void run(const u64 *instr, u32 n, u32 nslots) {
    T.assign(nslots, lit(P_ERA));
    T[0] = {AK::Dyn, 0};
    for (u32 k = 0; k < n; k++) {
      const u64 w0 = instr[2 * k], w1 = instr[2 * k + 1];
      const u32 kind = (u32)(w0 & 0xFF);
      const u32 a = (u32)(w0 >> 16) & 0xFFFF, b = (u32)(w0 >> 32) & 0xFFFF, c = (u32)(w0 >> 48);
      switch (kind) {
        case I_CONST: T[a] = lit(w1); break;
        case I_LINK: ilink(T[a], T[b]); break;
        case I_LINKCONST: ilink(T[a], lit(w1)); break;
        case I_WIRES: {
          const u64 c1 = new_cell(), c2 = new_cell();
          const u32 d = (u32)(w1 & 0xFFFF);
          T[a] = {AK::End, c1};
          T[b] = {AK::Wire, c1};
          T[c] = {AK::End, c2};
          T[d] = {AK::Wire, c2};
          break;
        }
        case I_CTR: ctr((u32)(w1 & 0xFFFF), a, b, c); break;
        case I_OP: op((u32)(w1 & 0xFFFF), a, b, c); break;
        case I_OPNUM: opnum(w1, a, c); break;
        case I_SWITCH: sw(a, b, c); break;
        default: throw Fail("instruction kind");
      }
    }
  }

  // do_ctr (instruction.rs:113-132)
  void ctr(u32 lab, u32 a, u32 b, u32 c) {
    const AV t = resolve(T[a]);
    if (t.k == AK::Dyn || t.k == AK::DynW) return residual(I_CTR, t, lab, b, c, true);
    if (t.k == AK::Node && nodes[t.v].tag == T_CTR && nodes[t.v].lab == lab) {
      ANode &n = nodes[t.v];
      if (!n.alive) throw Fail("node taken twice");
      n.alive = false;
      cnt[0]++;
      T[b] = as_wire(n.w[0]);
      T[c] = as_wire(n.w[1]);
      return;
    }
    const u64 v1 = new_cell(), v2 = new_cell();
    const u64 m = new_node(T_CTR, lab, {AK::End, v1}, {AK::End, v2});
    ilink(t, node_port(m));
    T[b] = {AK::End, v1};
    T[c] = {AK::End, v2};
  }

  // do_op (instruction.rs:134-152)
  void op(u32 opc, u32 a, u32 b, u32 c) {
    const AV t = resolve(T[a]);
    if (t.k == AK::Dyn || t.k == AK::DynW) return residual(I_OP, t, opc, b, c, true);
    if (t.k == AK::Lit && ptag(t.v) == T_NUM) {
      const u64 v = new_cell();
      const u64 m = new_node(T_OP, N_SYM, lit(numport(operate_unary(opc, pnum(t.v)))), {AK::End, v});
      T[b] = node_port(m);
      T[c] = {AK::End, v};
      return;
    }
    if (t.k == AK::Lit && t.v == P_ERA) {
      cnt[5]++;
      T[b] = lit(P_ERA);
      T[c] = lit(P_ERA);
      return;
    }
    const u64 v1 = new_cell(), v2 = new_cell();
    const u64 m = new_node(T_OP, opc, {AK::End, v1}, {AK::End, v2});
    ilink(t, node_port(m));
    T[b] = {AK::End, v1};
    T[c] = {AK::End, v2};
  }

  // do_op_num (instruction.rs:154-172)
  void opnum(u64 w1, u32 a, u32 c) {
    const u32 opc = (u32)(w1 & 0xFFFF), rhs = (u32)(w1 >> 32);
    const AV t = resolve(T[a]);
    if (t.k == AK::Dyn || t.k == AK::DynW) return residual(I_OPNUM, t, w1, 0, c, false);
    if (t.k == AK::Lit && ptag(t.v) == T_NUM) {
      bool div0 = false;
      const u32 out = operate_binary(pnum(t.v), opc, rhs, &div0);
      if (div0) throw Fail("division by zero in a definition");  // left to the device's error path
      cnt[4]++;
      T[c] = lit(numport(out));
      return;
    }
    if (t.k == AK::Lit && t.v == P_ERA) {
      cnt[6]++;
      T[c] = lit(P_ERA);
      return;
    }
    const u64 v = new_cell();
    const u64 m = new_node(T_OP, opc, lit(numport(rhs)), {AK::End, v});
    ilink(t, node_port(m));
    T[c] = {AK::End, v};
  }

  // do_switch (instruction.rs:174-203)
  void sw(u32 a, u32 b, u32 c) {
    const AV t = resolve(T[a]);
    if (t.k == AK::Dyn || t.k == AK::DynW) return residual(I_SWITCH, t, 0, b, c, true);
    if (t.k == AK::Lit && ptag(t.v) == T_NUM) {
      cnt[4]++;
      const u32 n = pnum(t.v);
      const u32 v = num_tag(n) == N_U24 ? num_u24(n) : 0;
      const u64 w = new_cell();
      if (v == 0) {
        const u64 x = new_node(T_CTR, 0, {AK::End, w}, lit(P_ERA));
        T[b] = node_port(x);
      } else {
        const u64 x = new_node(T_CTR, 0, lit(numport(num_new_u24(v - 1))), {AK::End, w});
        const u64 y = new_node(T_CTR, 0, lit(P_ERA), node_port(x));
        T[b] = node_port(y);
      }
      T[c] = {AK::Wire, w};
      return;
    }
    if (t.k == AK::Lit && t.v == P_ERA) {
      cnt[7]++;
      cnt[2]++;
      T[b] = lit(P_ERA);
      T[c] = lit(P_ERA);
      return;
    }
    const u64 v1 = new_cell(), v2 = new_cell();
    const u64 m = new_node(T_SWI, 0, {AK::End, v1}, {AK::End, v2});
    ilink(t, node_port(m));
    T[b] = {AK::End, v1};
    T[c] = {AK::End, v2};
  }

  // Redexes of the template that interact (interact.rs:6-44) without a
  // definition call are applied here, with the links and the count the
  // reference's rule makes: two nodes of the same kind (Ctr of one label,
  // Op, Switch) annihilate (anni2, interact.rs:72-79); an eraser meeting a
  // node, or a number meeting a Ctr, is copied onto both of the node's aux
  // ports (comm02, interact.rs:162-168).  The template then holds what is
  // left, and a call makes those interactions no more.
  void prereduce() {
    // redexes made here come after every list
    cur_seq = (link_seq / 6 + 1) * 6;
    link_seq = cur_seq + 1;
    for (u64 guard = 0;; guard++) {
      if (guard > 1 << 16) throw Fail("prereduce");
      bool done = false;
      for (size_t i = 0; i < redexes.size() && !done; i++) {
        AV a = redexes[i].first, b = redexes[i].second;
        if (a.k == AK::Node && b.k != AK::Node) std::swap(a, b);  // a nilary, b a node
        if (b.k != AK::Node) continue;
        ANode &nb = nodes[b.v];
        if (a.k == AK::Node) {
          ANode &na = nodes[a.v];
          const bool same = na.tag == nb.tag && (na.tag != T_CTR || na.lab == nb.lab);
          if (!same) continue;
          if (!na.alive || !nb.alive) throw Fail("redex of a taken node");
          redexes.erase(redexes.begin() + (long)i);
          redex_seq.erase(redex_seq.begin() + (long)i);
          na.alive = nb.alive = false;
          cnt[0]++;
          link(na.w[0], nb.w[0]);
          link(na.w[1], nb.w[1]);
          done = true;
        } else if (a.k == AK::Lit && (a.v == P_ERA || (ptag(a.v) == T_NUM && nb.tag == T_CTR))) {
          if (!nb.alive) throw Fail("redex of a taken node");
          redexes.erase(redexes.begin() + (long)i);
          redex_seq.erase(redex_seq.begin() + (long)i);
          nb.alive = false;
          cnt[1]++;
          link(nb.w[0], a);
          link(nb.w[1], a);
          done = true;
        }
      }
      if (!done) return;
    }
  }

  // Dissolve cells that a value was parked in: the one node word that is
  // the cell's other endpoint takes the value.  Repeats until every cell
  // left joins two node words.
  void dissolve() {
    for (int round = 0;; round++) {
      if (round > 1 << 16) throw Fail("dissolve");
      std::vector<std::vector<std::pair<u64, int>>> refs(cells.size());
      for (u64 i = 0; i < nodes.size(); i++) {
        if (!nodes[i].alive) continue;
        for (int h = 0; h < 2; h++)
          if (nodes[i].w[h].k == AK::End) refs[nodes[i].w[h].v].push_back({i, h});
      }
      bool changed = false;
      for (u64 c = 0; c < cells.size() && !changed; c++) {
        const Cell &cl = cells[c];
        if (cl.st == CS::Parked && refs[c].size() == 1) {
          nodes[refs[c][0].first].w[refs[c][0].second] = cl.parked;
          cells[c].st = CS::Done;
          changed = true;
        }
      }
      if (changed) continue;
      for (u64 c = 0; c < cells.size(); c++) {
        const Cell &cl = cells[c];
        if (cl.st == CS::Parked) throw Fail("parked cell without a single node word");
        if (cl.st == CS::Empty && refs[c].size() != 2) throw Fail("open cell");
        if (cl.st == CS::Done && !refs[c].empty()) throw Fail("node word names a consumed cell");
      }
      return;
    }
  }
};

}  // namespace

CompiledDef compile_def(const u64 *instr, u32 ninstr, u32 nslots) {
  CompiledDef out;
  try {
    Compiler C;
    C.run(instr, ninstr, std::max<u32>(nslots, 1));
    C.prereduce();
    C.dissolve();

    // template slots: live nodes, then the cells left (two per slot)
    std::vector<u64> node_slot(C.nodes.size(), ~0ull), cell_slot(C.cells.size(), ~0ull);
    u32 slots = 0;
    for (u64 i = 0; i < C.nodes.size(); i++)
      if (C.nodes[i].alive) node_slot[i] = slots++;
    u32 half = 0;
    for (u64 c = 0; c < C.cells.size(); c++) {
      if (C.cells[c].st != CS::Empty) continue;
      if (half == 0) slots++;
      cell_slot[c] = (u64)(slots - 1) * 2 + half;
      half ^= 1;
    }
    if (slots > 255 || C.redexes.size() > 255) throw Fail("template too large");
    for (u64 k : C.cnt)
      if (k > 255) throw Fail("count too large");

    // runtime registers: linear scan over token live ranges
    // positions: pattern ops 0..P-1, template P, links P+1..
    const u64 P = C.pattern.size();
    std::vector<u64> def_pos(C.ntok, 0), last_use(C.ntok, 0);
    std::vector<bool> used(C.ntok, false);
    auto use = [&](u64 t, u64 pos) {
      last_use[t] = std::max(last_use[t], pos);
      used[t] = true;
    };
    for (u64 i = 0; i < P; i++) {
      const ROp &r = C.pattern[i];
      use(r.in, i);
      if (r.out1 != ~0ull) def_pos[r.out1] = i;
      if (r.out2 != ~0ull) def_pos[r.out2] = i;
    }
    auto use_av = [&](const AV &a, u64 pos) {
      if (a.k == AK::Dyn || a.k == AK::DynW) use(a.v, pos);
    };
    for (auto &n : C.nodes)
      if (n.alive)
        for (auto &w : n.w) use_av(w, P);
    for (auto &r : C.redexes) {
      use_av(r.first, P);
      use_av(r.second, P);
    }
    for (u64 i = 0; i < C.links.size(); i++) {
      use(C.links[i].in, P + 1 + i);
      use_av(C.links[i].other, P + 1 + i);
    }
    // a token defined but never used keeps its register until its own
    // instruction (outputs are always written)
    for (u64 t = 0; t < C.ntok; t++)
      if (!used[t]) last_use[t] = def_pos[t];
    // Instruction i reads its operand before it writes its outputs, so an
    // output may take a register whose token is last used at i or before.
    std::vector<u64> reg(C.ntok, ~0ull);
    std::vector<u64> holder{0};  // token in each register; register 0 holds t0
    reg[0] = 0;
    for (u64 i = 0; i < P; i++) {
      const ROp &r = C.pattern[i];
      std::vector<u64> taken;
      for (u64 t : {r.out1, r.out2}) {
        if (t == ~0ull) continue;
        u64 q = 0;
        for (; q < holder.size(); q++) {
          if (last_use[holder[q]] > i) continue;
          if (std::find(taken.begin(), taken.end(), q) != taken.end()) continue;
          break;
        }
        if (q == holder.size()) holder.push_back(t);
        holder[q] = t;
        reg[t] = q;
        taken.push_back(q);
      }
    }
    const u32 nregs = (u32)std::max<size_t>(holder.size(), 1);
    if (nregs > 255) throw Fail("too many registers");

    auto spec = [&](const AV &a) -> u64 {
      switch (a.k) {
        case AK::Lit: return a.v;
        case AK::Node: {
          const ANode &n = C.nodes[a.v];
          if (!n.alive) throw Fail("template names a taken node");
          return mkport(n.tag, n.lab, node_slot[a.v] << 3);
        }
        case AK::End: {
          if (cell_slot[a.v] == ~0ull) throw Fail("template names a dissolved cell");
          return mkport(T_VAR, 0, cell_slot[a.v] << 3);
        }
        case AK::Dyn:
        case AK::DynW: return mkport(T_RED, 0, reg[a.v] << 3);
        default: throw Fail("spec of a wire");
      }
    };

    // template words
    std::vector<u64> tmpl;
    tmpl.push_back((u64)slots | ((u64)C.redexes.size() << 16));
    u64 packed = 0;
    for (int k = 0; k < 8; k++) packed |= C.cnt[k] << (8 * k);
    tmpl.push_back(packed);
    std::vector<u64> slot_words(2 * (u64)slots, P_FREE);
    for (u64 i = 0; i < C.nodes.size(); i++) {
      if (!C.nodes[i].alive) continue;
      slot_words[2 * node_slot[i]] = spec(C.nodes[i].w[0]);
      slot_words[2 * node_slot[i] + 1] = spec(C.nodes[i].w[1]);
    }
    for (u64 c = 0; c < C.cells.size(); c++)
      if (cell_slot[c] != ~0ull) slot_words[cell_slot[c]] = P_NONE;
    tmpl.insert(tmpl.end(), slot_words.begin(), slot_words.end());
    // Redexes are pushed in the order the interpreter's link list would
    // deliver them: which redexes are oldest in a wave's rings decides which
    // ones are handed to idle waves.
    for (size_t i : C.redex_order()) {
      tmpl.push_back(spec(C.redexes[i].first));
      tmpl.push_back(spec(C.redexes[i].second));
    }

    // residual program
    std::vector<u64> prog;
    for (const ROp &r : C.pattern) {
      const u32 a = (u32)reg[r.in];
      const u32 b = r.out2 == ~0ull ? 0 : (u32)reg[r.out1];
      const u32 c = r.out2 == ~0ull ? (u32)reg[r.out1] : (u32)reg[r.out2];
      prog.push_back(instr_w0(r.kind, a, b, c) | (r.aswire ? I_ASWIRE : 0));
      prog.push_back(r.w1);
    }
    const bool has_template = slots > 0 || !C.redexes.empty() || packed != 0;
    if (has_template) {
      prog.push_back(instr_w0(I_TEMPLATE, 0, 0, 0));
      prog.push_back(0);  // template offset, set by the book builder
    }
    for (const ROp &r : C.links) {
      const u32 a = (u32)reg[r.in];
      const AV &o = r.other;
      if (o.k == AK::Dyn) {
        prog.push_back(instr_w0(I_LINK, a, (u32)reg[o.v], 0));
        prog.push_back(0);
      } else if (o.k == AK::Lit) {
        prog.push_back(instr_w0(I_LINKCONST, a, 0, 0));
        prog.push_back(o.v);
      } else {
        prog.push_back(instr_w0(I_LINKT, a, 0, 0));
        prog.push_back(spec(o));
      }
    }
    if (prog.size() / 2 > 0xFFFF) throw Fail("program too long");
    out.ok = true;
    out.instr = std::move(prog);
    out.nslots = nregs;
    out.tslots = slots;
    out.tmpl = has_template ? std::move(tmpl) : std::vector<u64>{};
  } catch (Fail &f) {
    out = CompiledDef{};
    out.why = f.what();
  }
  return out;
}

namespace {

// Call ring group of a definition by its length: the lanes of a call
// iteration wait for the longest definition among them.
u32 call_group(u32 len) { return len <= 4 ? 0 : len <= 8 ? 1 : len <= 16 ? 2 : 3; }

// A template word spec names slots below `tslots` and registers below
// `nslots` (the device reads them without checks).
bool valid_spec(u64 s, u32 tslots, u32 nslots) {
  const u32 t = ptag(s);
  if (t == T_REF || t == T_NUM) return true;
  if (t >= T_OP) return (paddr(s) >> 3) < tslots;
  if (t == T_VAR) return (paddr(s) >> 4) < tslots && plab(s) == 0;
  if (t != T_RED) return false;
  if (s == P_NONE || s == P_FREE) return true;
  return (paddr(s) >> 3) < nslots && plab(s) == 0;
}

// Checks a compiled definition's program and template before the device
// runs them; false: the definition stays interpreted.
bool valid_compiled(const CompiledDef &c) {
  bool template_seen = false;
  for (size_t k = 0; k < c.instr.size() / 2; k++) {
    const u64 w0 = c.instr[2 * k], w1 = c.instr[2 * k + 1];
    const u32 kind = (u32)(w0 & 0xFF);
    const u32 a = (u32)(w0 >> 16) & 0xFFFF, b = (u32)(w0 >> 32) & 0xFFFF, cc = (u32)(w0 >> 48);
    if (a >= c.nslots || b >= c.nslots || cc >= c.nslots) return false;
    switch (kind) {
      case I_LINK: case I_LINKCONST: case I_CTR: case I_OP: case I_OPNUM: case I_SWITCH: break;
      case I_LINKT:
        if (!valid_spec(w1, c.tslots, c.nslots) || ptag(w1) < T_VAR) return false;
        break;
      case I_TEMPLATE:
        if (template_seen || w1 != 0) return false;
        template_seen = true;
        break;
      default: return false;
    }
  }
  if (template_seen != !c.tmpl.empty()) return false;
  if (c.tmpl.empty()) return c.tslots == 0;
  if (c.tmpl.size() < 2) return false;
  const u64 S = c.tmpl[0] & 0xFF, R = (c.tmpl[0] >> 16) & 0xFF;
  if (S != c.tslots || c.tmpl.size() != 2 + 2 * S + 2 * R) return false;
  for (size_t i = 2; i < c.tmpl.size(); i++)
    if (!valid_spec(c.tmpl[i], c.tslots, c.nslots)) return false;
  return true;
}

}  // namespace

DeviceBook build_device_book(std::vector<DefRec> &defs, const std::vector<u64> &instrs) {
  DeviceBook b;
  b.instrs = instrs;
  b.cdefs.assign(defs.size(), 0);
  b.max_slots = 1;
  for (size_t i = 0; i < defs.size(); i++) {
    DefRec &d = defs[i];
    d.group = call_group(d.ilen);
    b.max_slots = std::max(b.max_slots, d.nslots);
    CompiledDef c = compile_def(&instrs[2 * (u64)d.ioff], d.ilen, d.nslots);
    if (!c.ok || !valid_compiled(c)) continue;
    const u64 ioff = b.instrs.size() / 2, toff = b.tmpl.size();
    if (ioff + c.instr.size() / 2 >= (1u << 24) || c.instr.size() / 2 > 0xFFFF || c.nslots > 255 || c.tslots > 255)
      continue;
    for (size_t k = 0; k < c.instr.size() / 2; k++)
      if ((c.instr[2 * k] & 0xFF) == I_TEMPLATE) c.instr[2 * k + 1] = toff;
    b.instrs.insert(b.instrs.end(), c.instr.begin(), c.instr.end());
    b.tmpl.insert(b.tmpl.end(), c.tmpl.begin(), c.tmpl.end());
    b.cdefs[i] = ioff | ((u64)(c.instr.size() / 2) << 24) | ((u64)c.nslots << 40) | ((u64)c.tslots << 48);
    // the call ring stays the one of the definition's own length: grouping
    // by the compiled program's length measured slower (bitonic sort of 2^16
    // keys, 2,056 against 2,238 MIPS)
    b.max_slots = std::max(b.max_slots, c.nslots);
    b.n_compiled++;
  }
  return b;
}

}  // namespace hvm64
