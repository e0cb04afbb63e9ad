// Device-side data layout of the reduction engine, shared between the host
// library and the kernels.
#pragma once
#include <stdint.h>
#include "hvm64_port.h"

namespace hvm {

// Instruction kinds (runtime/src/instruction.rs:34-72).  On the device an
// instruction is two words:
//   w0 = kind | a<<16 | b<<32 | c<<48      (a, b, c: trg register slots)
//   w1 = Const/LinkConst: port; Ctr: label; Op: op; OpNum: op | rhs<<32;
//        Wires: d (fourth slot)
enum : u32 { I_CONST = 0, I_LINK, I_LINKCONST, I_CTR, I_OP, I_OPNUM, I_SWITCH, I_WIRES,
              // kinds that only compiled definitions use (hvm64_defcomp.cpp):
              I_TEMPLATE,  // w1 = word offset of the definition's template in the template table
              I_LINKT,     // link register a with a template object; w1 = template spec
              I_KINDS };
// w0 bit 8: the operand register holds a port to be read as a wire (the
// compiled form of a trg that reached a runtime value through a var cell)
static const u32 I_ASWIRE = 1u << 8;

HVM_HD u64 instr_w0(u32 kind, u32 a, u32 b, u32 c) {
  return (u64)kind | ((u64)a << 16) | ((u64)b << 32) | ((u64)c << 48);
}

// Compiled definitions.  A definition compiled at book load is a residual
// program (the instructions whose operands depend on the port that met the
// definition, run as interpreted instructions) and a template: heap slots
// whose words are known up to the addresses of the slots and the values of
// registers (written by one I_TEMPLATE instruction), redexes between them,
// and the rewrite counts of the instructions evaluated at load time.
//
// Template word specs are port-shaped: Ref / Num ports are literals; a
// binary tag (Op, Switch, Ctr) with slot s in the address field is the
// principal port of the node in slot s; a Var port with 2s+h in the address
// field is the var cell in half h of slot s; P_NONE and P_FREE are literals;
// any other port with the Red tag and r in the address field is the value of
// register r (as a var endpoint).
//
// Template layout: word 0 = slots | redexes << 16; word 1 = the eight
// counts (anni, comm, eras, dref, oper, and the three ERA shortcuts) as
// bytes; then two specs per slot, then two specs per redex.
// Compiled-definition record (one word per definition): program offset
// (instruction index, 24 bits) | program length << 24 (16 bits) |
// registers << 40 (8 bits) | template slots << 48 (8 bits); 0 = not compiled.
HVM_HD u32 cdef_ioff(u64 c) { return (u32)(c & 0xFFFFFF); }
HVM_HD u32 cdef_ilen(u64 c) { return (u32)((c >> 24) & 0xFFFF); }
HVM_HD u32 cdef_nslots(u64 c) { return (u32)((c >> 40) & 0xFF); }
HVM_HD u32 cdef_tslots(u64 c) { return (u32)((c >> 48) & 0xFF); }
// Template slot j sits TSLOT_STRIDE bytes after slot j - 1 (consecutive
// positions of a lane's allocation slice).
static const u64 TSLOT_STRIDE = 64 * 16;

// One book definition (def.rs:86-95 + InterpretedDef, def.rs:194-199).
struct DefRec {
  u32 ioff;       // first instruction
  u32 ilen;       // instruction count
  u32 nslots;     // trg registers needed (after slot compaction)
  u32 labs_off;   // first word of the label bitset in the label array
  u32 labs_words; // bitset length in words
  u32 min_safe;   // LabSet::min_safe (also the label of Ref ports to this def)
  u32 group;      // call ring of this definition's calls (by instruction count)
  u32 flags;      // DEF_INERT: a name the book refers to but does not define
};
// An inert definition (transform/src/pre_reduce.rs:99-116) has no
// instructions; a call on it counts a dereference and records the redex in
// the captured list instead of expanding anything.
static const u32 DEF_INERT = 1;

// Rewrite counter slots (runtime.rs:72-79) and interpreter shortcut counters
// (ERA targets of do_op / do_op_num / do_switch, instruction.rs:144-197),
// and the number of redexes popped.
enum : u32 {
  RW_ANNI = 0, RW_COMM, RW_ERAS, RW_DREF, RW_OPER, SC_OP_ERA, SC_OPNUM_ERA, SC_SWITCH_ERA, RW_POPS,
  RW_ALLOC_MISS,  // ring slots found in use on a later lap
  RW_OV_ALLOC,    // allocations served by the overflow ring
  RW_REUSE,       // allocations served by a lane's free cache
  RW_SLOTS
};

// Which arena store ran dry (Globals::err_info of ERR_OOM in a planned batch).
enum : u32 { OOM_POOL = 1, OOM_OVERFLOW = 2 };

// Which consistency check reported ERR_BADPTR (Globals::err_info).
enum : u32 { BAD_HEAP_ADDR = 1, BAD_LINK_LIST, BAD_CHAIN, BAD_RESOLVE, BAD_SLOTS, BAD_DEF_CALL, BAD_DEF_EXPAND,
             BAD_NET };  // BAD_NET: a port beyond the granule map

// Scheduler statistics (summed over waves): iterations with work, idle
// iterations, wall-clock ticks (100 MHz) spent in iterations with work,
// spills to the global ring / redexes spilled, claims / redexes claimed.
// A single net's launch collects them only when Params::timing is set
// (keeping them costs it 2-4%); batched and serving launches always do.
enum : u32 {
  ST_BUSY_IT = 0, ST_IDLE_IT, ST_BUSY_TICKS, ST_SPILLS, ST_SPILLED, ST_CLAIMS, ST_CLAIMED,
  ST_INTERACT_TICKS,  // ticks inside interact() (the divergent section)
  ST_SPILL_TICKS,     // ticks spilling to the global ring
  ST_CLS_IT,          // 4 slots: iterations that ran each redex class
  ST_CLS_TICKS = ST_CLS_IT + 4,  // 4 slots: ticks inside the rules, per class
  ST_SLOTS = ST_CLS_TICKS + 4
};

// Error codes reported by the device (mirrored in the C ABI header).
enum : u32 {
  ERR_NONE = 0, ERR_OOM = 2, ERR_UNIMPL = 3, ERR_DIV0 = 4, ERR_QUEUE = 7, ERR_READBACK = 8,
  ERR_BADPTR = 12,   // a port addressed memory outside the heap or book (engine bug)
  ERR_TIMEOUT = 13,  // the kernel ran past its deadline
};

// Global control block.  Hot words sit on separate 128-byte lines.
struct alignas(128) Globals {
  u64 S;           // (queued redexes << 32) | active waves
  u64 _p0[15];
  u64 q_head;      // next slot to consume in the global redex ring
  u64 _p1[15];
  u64 q_tail;      // next slot to fill
  u64 _p2[15];
  u64 ov_pos;      // allocations taken from the shared overflow ring
  u64 _p3[15];
  u64 done;        // redexes popped against a reduce(limit) budget
  u64 _p4[15];
  u32 err;         // first error code
  u32 err_info;
  u64 result;      // kernel-specific result word (new root cell, readback length)
  u64 captured_n;  // redexes recorded by calls on inert definitions
  u64 _p5[13];
  u64 rw[RW_SLOTS];  // cumulative counters (RW_POPS: redexes popped)
  u64 _p6[16 - RW_SLOTS];
  u64 st[ST_SLOTS];  // scheduler statistics of the last reduce launch
};

// Launch-time parameters of every kernel.
struct Params {
  u64 *heap;           // heap[(addr - HVM_HEAP_OFF) >> 3] is the word at engine address `addr`
  u64 heap_words;
  u64 ring_node0;      // node index of wave 0's allocation ring
  u64 wave_nodes;      // nodes per wave ring (a multiple of 64)
  u64 ov_node0;        // node index of the shared overflow ring
  u64 ov_nodes;        // nodes in the overflow ring (a power of two)
  u64 *alloc_pos;      // per-wave ring positions, persisted across launches
  const DefRec *defs;
  const u64 *instrs;   // 2 words per instruction
  const u64 *labbits;
  u64 *captured;       // pairs recorded by calls on inert definitions
  u64 captured_cap;    // pairs the list holds
  u64 *trg_buf;        // per-wave trg registers: slot-major, 64 lanes interleaved
  u64 *queue;          // global redex ring, 2 words per slot
  u64 qmask;           // ring slots - 1
  Globals *g;
  u32 ndefs;
  u32 trg_slots;       // registers per lane
  u32 total_waves;     // waves in the reduce grid
  u32 book_lds;        // 1: the reduce kernel copies the book into LDS
  u32 instr_words;     // words in `instrs`
  u32 lab_words;       // words in `labbits`
  u32 t_lds;           // 1: every definition's trg registers fit in LDS
  const u64 *cdefs;    // compiled-definition records (one word per definition)
  const u64 *tmpl;     // template table of the compiled definitions
  u32 tmpl_words;      // words in `tmpl`
  u32 compiled;        // 1: calls run the compiled form where there is one
  u64 *idle_bits;      // one bit per wave of the reduce grid: the wave is idle
  u64 *mbox;           // per-wave mailboxes (MBOX_WORDS each)
  u32 idle_words;      // words in idle_bits
  u32 fill;            // busy waves short of work also claim from the global ring
  u32 timing;          // single net: collect the scheduler statistics (ST_*)
  u64 limit;           // redex budget for reduce(limit); ~0 = unlimited
  u64 timeout_ticks;   // wall-clock (100 MHz) budget of a launch; 0 = none
  // scheduler thresholds (tunable): redexes a busy wave keeps before handing
  // some to an idle wave; pending redexes under which a busy wave of a
  // batched launch claims from the global ring
  u32 mbox_keep, fill_below;
  // Per-net arenas (batched nets; gmap == nullptr: the per-lane rings).  The
  // heap is cut into granules of 2^gran_shift nodes; gmap names each
  // granule's net (or none).  Net i's arena starts as arena_len[i] granules
  // from granule arena_base0[i] and grows by extents taken from the pool of
  // granules [*pool_next, pool_end); arena_word[i] = current extent's base
  // granule << 32 | nodes handed out of it.  Requests that arrive while an
  // extent is being replaced take nodes [*ov_next, ov_end).
  u32 *gmap;
  u64 gmap_len;        // granules
  u64 *arena_word;
  const u64 *arena_base0;
  const u32 *arena_len;
  u64 *pool_next;      // next pool granule
  u64 pool_end;
  u64 *ov_next;        // next overflow node
  u64 ov_end;
  u32 n_nets;
  u32 refill_granules; // largest refill extent (granules)
  u32 arena_chunk;     // nodes a lane takes from its net's arena while it stays on the net
  u32 gran_shift;
  // Serving (serve != 0): the arenas are request slots (net = slot), reused
  // once a slot's net is normal.  pending[s] counts the slot's redexes not
  // yet consumed; the task that brings it to 0 reports the net's root port.
  // Request and completion rings live in pinned host memory.
  u32 serve;
  u32 serve_slots;
  u32 *pending;
  u64 *slot_root;      // root cell of each slot's net
  u64 *slot_tag;       // the request's tag
  u64 *free_ring;      // slot + 1 entries (0: not yet written), serve_slots entries
  u64 *serve_ctl;      // device counters, each on its own 128-byte line (SV_*)
  const u64 *req;      // host: requests, (definition + 1, tag) pairs
  u64 req_mask;        // request ring entries - 1
  u64 *host_ctl;       // host: control words, each on its own 128-byte line (SH_*)
  u64 *done;           // host: completions, (seq, tag, port, 0) per entry
  u64 done_mask;       // completion ring entries - 1
  u64 *inl;            // per lane: SERVE_INL pairs of redexes run inline
};

// Device counters of a serving launch (Params::serve_ctl, in words).
enum : u32 { SV_FREE_HEAD = 0, SV_FREE_TAIL = 16, SV_FREE_COUNT = 32, SV_REQ_HEAD = 48, SV_DONE_TAIL = 64,
             SV_WORDS = 80 };
// Host-side control words of a serving launch (Params::host_ctl, in words):
// requests written (host), requests admitted (device), completions read
// (host), stop flag (host), device error code (device).
enum : u32 { SH_REQ_TAIL = 0, SH_REQ_HEAD = 16, SH_DONE_READ = 32, SH_STOP = 48, SH_ERR = 64, SH_WORDS = 80 };
// Redexes between two ports that reach no heap node carry no net; a serving
// lane runs them at once, inside the task that made them (at most this many
// pending per lane).
static const u32 SERVE_INL = 32;
// A serving slot's arena is this many sub-arenas (consecutive, whole
// granules each), taken from by the lanes of matching index.
static const u32 SERVE_SUB = 8;

// Mailbox of a wave: a count word on its own 128-byte line, then up to
// MBOX_PAIRS redex pairs handed to the wave by another wave.
static const u32 MBOX_PAIRS = 64;
static const u32 MBOX_WORDS = 16 + 2 * MBOX_PAIRS;
// Count word value written to every mailbox by the wave that ends the
// reduction (S reached 0), so idle waves see the end without reading S.
static const u64 MBOX_DONE = ~0ull;

// Heap layout: the first RESERVED_WORDS words are fixed cells.
static const u64 RESERVED_WORDS = 64;
static const u64 DUMP_WORD = 0;       // 2-word sink written after an error
static const u64 ROOT0_WORD = 2;      // the root var cell of a fresh net

}  // namespace hvm
