// Host side of the engine: the text format (ast/src), book loading
// (host/src/host.rs, calc_labels.rs, encode.rs) and readback
// (host/src/readback.rs), in C++.
#pragma once
#include <cstdint>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hvm64_amd.h"
#include "hvm64_engine.h"

namespace hvm64 {

using hvm::u16;
using hvm::u32;
using hvm::u64;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Error codes: the HVM_E_* values of the C ABI.


// ---------------------------------------------------------------------------
// AST (ast/src/ast.rs:29-96).  Trees live in an arena and refer to children
// by index, so deep trees need no recursion to build or free.
// ---------------------------------------------------------------------------
enum class TK : uint8_t { Era, Num, Ref, Ctr, Op, Switch, Var };

struct TNode {
  TK kind;
  u16 lab = 0;    // Ctr label
  u32 num = 0;    // Num raw
  u32 name = 0;   // Ref / Var: index into Arena::names
  u32 a = 0, b = 0;  // children (Ctr p1 p2, Op rhs out, Switch arms out)
};

struct Arena {
  std::vector<TNode> nodes;
  std::vector<std::string> names;
  std::map<std::string, u32> name_ids;
  u32 intern(const std::string &s);
  u32 add(const TNode &n) {
    nodes.push_back(n);
    return (u32)nodes.size() - 1;
  }
};

struct Net {
  u32 root = 0;
  std::vector<std::pair<u32, u32>> redexes;
};

struct Book {
  Arena arena;
  std::map<std::string, Net> nets;  // BTreeMap order (ast.rs:29-31)
};

Book parse_book(const std::string &text);
// Parses a net into an existing arena.
Net parse_net(Arena &arena, const std::string &text);
u32 parse_tree(Arena &arena, const std::string &text);

std::string show_tree(const Arena &arena, u32 t);
std::string show_net(const Arena &arena, const Net &net);
std::string show_book(const Book &book);
std::string show_num(u32 raw);
std::string var_name(u64 id);  // util/src/create_var.rs:5-15

// Net::apply_tree (ast.rs:112-123).
void apply_tree(Arena &arena, Net &net, u32 arg);

// ---------------------------------------------------------------------------
// Host: definitions with label sets and device instructions.
// ---------------------------------------------------------------------------
struct LabSet {
  u32 min_safe = 0;
  std::vector<u64> bits;
  void add(u16 lab);
  void unite(const LabSet &o);
  bool has(u16 lab) const;
  static LabSet all();  // LabSet::ALL (def.rs:17): every label below 0xFFFF
};

struct DefCode {
  std::vector<u64> instr;  // 2 words per instruction (device format)
  u32 nslots = 1;          // trg registers after slot compaction
  u32 ninstr() const { return (u32)(instr.size() / 2); }
};

class Host {
 public:
  // insert_book_with_default (host.rs:41-73): label sets first, then code.
  // insert_book: a reference to a name the book does not define is an
  // error; insert_book_with_inert_default: it becomes an inert definition
  // (LabSet::ALL, no instructions; transform/src/pre_reduce.rs:44-45).
  void insert_book(const Book &book);
  void insert_book_with_inert_default(const Book &book);
  bool inert(u32 id) const { return inert_[id]; }
  // Replaces a definition's instructions; its label set stays
  // (pre_reduce.rs:156-157).
  void set_code(u32 id, DefCode code) { code_[id] = std::move(code); }
  bool has_def(const std::string &name) const { return index_.count(name) != 0; }
  u32 def_id(const std::string &name) const;
  const std::string &def_name(u32 id) const { return names_[id]; }
  u32 ndefs() const { return (u32)names_.size(); }
  u64 ref_port(const std::string &name) const;
  const LabSet &labs(u32 id) const { return labs_[id]; }
  const DefCode &code(u32 id) const { return code_[id]; }
  // Compile an AST net into instructions whose register 0 is the net's root
  // (encode.rs:13-19 for definitions, encode.rs:31-35 for encode_net).
  DefCode encode(const Arena &arena, const Net &net) const;

  // Flattened device book.
  std::vector<hvm::DefRec> def_records() const;
  std::vector<u64> all_instrs() const;
  std::vector<u64> all_labbits() const;
  u32 max_slots() const;
  // Adopt definitions encoded elsewhere (the C ABI's pre-encoded books).
  void adopt(std::vector<std::string> names, std::vector<LabSet> labs, std::vector<DefCode> code);

 private:
  std::vector<std::string> names_;
  std::map<std::string, u32> index_;
  std::vector<LabSet> labs_;
  std::vector<DefCode> code_;
  std::vector<bool> inert_;
  void insert_sets(const Book &book, const std::map<std::string, LabSet> &sets);
};

// A definition compiled into a residual program and a template
// (hvm64_defcomp.cpp).  ok = false: the definition runs interpreted.
struct CompiledDef {
  bool ok = false;
  std::vector<u64> instr;  // residual program, device format (2 words each)
  u32 nslots = 1;          // registers of the residual program
  u32 tslots = 0;          // heap slots of the template
  std::vector<u64> tmpl;   // template words (empty: no I_TEMPLATE)
  std::string why;         // ok = false: the reason
};
CompiledDef compile_def(const u64 *instr, u32 ninstr, u32 nslots);

// The book as the device reads it: the definitions' own instructions, then
// the compiled programs; one compiled-definition record per definition (0:
// interpreted); the template table.  Sets each definition's call ring group.
struct DeviceBook {
  std::vector<u64> instrs, cdefs, tmpl;
  u32 max_slots = 1;   // registers of the widest program
  u32 n_compiled = 0;  // definitions with a compiled form
};
DeviceBook build_device_book(std::vector<hvm::DefRec> &defs, const std::vector<u64> &instrs);

// Label sets of every definition (calc_labels.rs:78-194).  A reference to a
// name the book does not define takes its set from `missing`
// (calc_labels.rs:183-190), or is an error when `missing` is empty.
std::map<std::string, LabSet> calculate_label_sets(
    const Book &book, const std::function<LabSet(const std::string &)> &missing = {});

// Rebuild normal-form text from the device's preorder serialisation
// (k_readback): principal ports in preorder, Var(cell) for aux-aux wires.
// Variable names follow first occurrence (readback.rs:53-62).
std::string readback_text(const Host &host, const std::vector<u64> &stream, size_t &pos,
                          std::map<u64, u64> &vars, u64 &next_var);

// Runs `fn` on a thread with a large stack.
void with_big_stack(const std::function<void()> &fn);

}  // namespace hvm64
