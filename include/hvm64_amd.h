/*
 * hvm64_amd.h -- C ABI of the MI355X (gfx950) reduction engine for HVM-64
 * interaction nets.
 *
 * The reference (HigherOrderCO/hvm-64 v0.3.0) exposes this path as a Rust
 * API, not a C ABI; each entry point below names the Rust item it stands in
 * for.  Conventions: every call returns 0 on success or an HVM_E_* code;
 * hvm_gpu_last_error() describes the last failure.  Calls are blocking and a
 * handle must be used by one host thread at a time.  The library owns all
 * device memory; the caller owns the handle.
 *
 * Ports keep the reference's 64-bit layout (runtime/src/port.rs:8-14): tag in
 * bits 0-2 (Red 0, Var 1, Ref 2, Num 3, Op 5, Switch 6, Ctr 7), address in
 * bits 3-47, label in bits 48-63, a 32-bit Num in the low half of Num ports.
 * Addresses are byte offsets in the engine's address space: definition i of
 * the loaded book is at (i+1)<<5 (so the eraser stays the null Ref, 0x2), and
 * heap word w is at (1<<32) + 8*w.
 */
#ifndef HVM64_AMD_H
#define HVM64_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define HVM_OK 0
#define HVM_E_PARSE 1        /* text parse error (ast/src/parser.rs:169-176) */
#define HVM_E_OOM 2          /* heap exhausted (panic "OOM", runtime/src/allocator.rs:104) */
#define HVM_E_UNIMPL_PAIR 3  /* Switch-Ctr(0) / Op-Switch (unimplemented!, runtime/src/interact.rs:39-42) */
#define HVM_E_DIV0 4         /* integer / or % by zero (Rust panic, num/src/num.rs:183-212) */
#define HVM_E_UNBOUND 5      /* unbound variable (host/src/encode.rs:46) */
#define HVM_E_NOREF 6        /* reference to a missing definition (host/src/host.rs:35) */
#define HVM_E_QUEUE 7        /* global redex ring overflow */
#define HVM_E_READBACK 8     /* malformed net during readback */
#define HVM_E_TRUNC 9        /* output buffer too small (length reported) */
#define HVM_E_ARG 10         /* invalid argument */
#define HVM_E_DEVICE 11      /* HIP runtime error, or an invalid address inside the engine */
#define HVM_E_TIMEOUT 12     /* a reduction ran past the handle's time limit */

typedef struct hvm_gpu hvm_gpu;

/* One interpreted-definition instruction in device form
 * (runtime/src/instruction.rs:34-72):
 *   w0 = kind | a<<16 | b<<32 | c<<48   kind: 0 Const, 1 Link, 2 LinkConst,
 *        3 Ctr, 4 Op, 5 OpNum, 6 Switch, 7 Wires; a, b, c: trg registers
 *   w1 = Const/LinkConst: port; Ctr: label; Op: op tag; OpNum: op | rhs<<32;
 *        Wires: fourth register                                           */
typedef struct {
  uint64_t w0, w1;
} hvm_instr;

/* One definition: Def<InterpretedDef> (runtime/src/def.rs:86-95, 194-199). */
typedef struct {
  const char *name;
  uint32_t min_safe;          /* LabSet::min_safe; label of Ref ports to it */
  uint32_t n_lab_words;
  const uint64_t *lab_bits;   /* LabSet::bits */
  uint32_t n_instr;
  const hvm_instr *instr;
  uint32_t n_slots;           /* trg registers (InterpretedDef::trgs, may be compacted) */
} hvm_def_blob;

typedef struct {
  uint32_t n_defs;
  const hvm_def_blob *defs;   /* Ref ports to definition i use address (i+1)<<5 */
} hvm_book_blob;

/* ~ Heap::new(Some(bytes)) (runtime/src/allocator.rs:22): opens device
 * `device`, allocates a heap of `heap_bytes` and an empty net. */
int hvm_gpu_open(int device, uint64_t heap_bytes, hvm_gpu **out);

/* hvm_gpu_open with the reduction grid capped at `max_blocks` blocks of 4
 * waves (0: as many as fit on the device).  Handles opened with smaller
 * grids can reduce concurrently on one device, each from its own host
 * thread.  No counterpart in the reference (one Heap per Net there). */
int hvm_gpu_open_grid(int device, uint64_t heap_bytes, uint32_t max_blocks, hvm_gpu **out);

/* Frees all device memory of the handle. */
void hvm_gpu_close(hvm_gpu *h);

/* Text of the last error on this handle (or of the last failed open). */
const char *hvm_gpu_last_error(hvm_gpu *h);

/* ~ Host::new(&book) / Host::insert_book (host/src/host.rs:25-73): parses a
 * book, computes label sets, encodes every definition and uploads it. */
int hvm_gpu_load_book_text(hvm_gpu *h, const char *book_text);

/* ~ Host::insert_book for a book already encoded by the caller (for example
 * the Rust Host's Def<InterpretedDef> list serialised into hvm_def_blob). */
int hvm_gpu_load_book(hvm_gpu *h, const hvm_book_blob *book);

/* hvm_gpu_load_book for a book image written by hvm_book_encode (a flat,
 * self-contained byte buffer that a host can cache instead of parsing and
 * encoding the book's text again).  The image format: 8-byte words, "HVM64BK1",
 * the definition count, then per definition: name length | instruction
 * count << 32, trg registers | min_safe << 32, label-bit words, the name
 * (zero-padded to 8 bytes), the label bits, the instructions (2 words each,
 * hvm_instr).  No counterpart in the reference, whose Host is built from the
 * text every time (host/src/host.rs:25-29). */
int hvm_gpu_load_book_image(hvm_gpu *h, const void *image, uint64_t len);

/* ~ Host::defs[name] (host/src/host.rs:19): index of a loaded definition. */
int hvm_gpu_def_id(hvm_gpu *h, const char *name, uint32_t *id);

/* ~ Net::new(&heap) (runtime/src/net.rs:19-23): a fresh, empty net on the
 * handle's heap: the allocation rings, the redex queue and the rewrite
 * counters restart and the root is a fresh wire. */
int hvm_gpu_reset(hvm_gpu *h);

/* ~ Net::boot(&def) (runtime/src/net.rs:30-32): calls definition `def_id`
 * with the root wire (counts one DREF). */
int hvm_gpu_boot(hvm_gpu *h, uint32_t def_id);

/* ~ Host::encode_net(net, Trg::port(Var(root)), &ast_net)
 * (host/src/encode.rs:31-35): writes a net given as text into the root. */
int hvm_gpu_encode_net_text(hvm_gpu *h, const char *net_text);

/* ~ Net::normal() (runtime/src/net.rs:55-60): expand, then reduce to normal
 * form on the GPU. */
int hvm_gpu_normal(hvm_gpu *h);

/* ~ Net::reduce(limit) (runtime/src/net.rs:41-52): pops at most `limit`
 * redexes; *popped = redexes popped.  As in the reference, *is_normal = 1
 * (Some(count)) only when the redexes ran out before `limit` were popped,
 * else 0 (None). */
int hvm_gpu_reduce(hvm_gpu *h, uint64_t limit, uint64_t *popped, int *is_normal);

/* Batches of independent nets in one reduction.  No counterpart in the
 * reference, which reduces one Net per Heap; each net behaves as its own
 * Net::new + Net::boot + Net::normal (runtime/src/net.rs:19-60) would.
 *
 * hvm_gpu_boot_batch: nets 0..n-1 each get a fresh root cell, written to
 *   roots[i], and are booted with definition def_ids[i] (one DREF each).
 *   Nets booted after a reset share the handle's heap and counters.
 * hvm_gpu_normal_batch: expands every root (roots[i] is replaced by the
 *   expanded root) and reduces all booted nets to normal form in one launch.
 *   hvm_gpu_rewrites then reports the sums over the batch.  In this launch
 *   a busy wave that runs low on redexes also claims from the global ring
 *   (nearly every wave is busy with many nets). 
 * hvm_gpu_root_ports: the port each root leads to (a Num port for a numeric
 *   result; Var(root) while the root's wire is open).
 * hvm_gpu_readback_root: hvm_gpu_readback of one batch root. */
int hvm_gpu_boot_batch(hvm_gpu *h, uint32_t n, const uint32_t *def_ids, uint64_t *roots);

/* hvm_gpu_boot_batch for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops): the same boot, staged through a static
 * 768 KiB device array instead of a buffer the handle allocates, and with no
 * device or host allocation.  At most 65,536 nets per call (HVM_E_ARG
 * beyond).  Every handle on one GPU shares that array: calls that overlap on
 * one GPU race.  No counterpart in the reference. */
int hvm_gpu_boot_batch_noalloc(hvm_gpu *h, uint32_t n, const uint32_t *def_ids, uint64_t *roots);
int hvm_gpu_normal_batch(hvm_gpu *h, uint32_t n, uint64_t *roots);
int hvm_gpu_root_ports(hvm_gpu *h, uint32_t n, const uint64_t *roots, uint64_t *ports);
int hvm_gpu_readback_root(hvm_gpu *h, uint64_t root, char *buf, uint64_t cap, uint64_t *len);

/* Planned batches: Net::boot + Net::normal per net (net.rs:30-32, 53-58)
 * for a fixed list of nets, run again and again.
 *
 * hvm_gpu_batch_plan: plans n nets, net i booted with definition def_ids[i]
 *   and allocating from an arena of its own: arena_nodes[i] 16-byte nodes
 *   rounded up to whole granules (16 KiB on heaps up to 64 GiB, a larger
 *   power of two above, so that at most 4M granules cover the heap), laid
 *   out one after another.  The rest of the heap is a pool: a net that
 *   outgrows its arena extends it from the pool by extents of half its
 *   arena (16 to 64 granules); requests that arrive while an extent is
 *   being installed are served from an overflow region (the pool's last
 *   sixteenth).  HVM_E_OOM when either runs dry, or at plan time when the
 *   arenas do not fit in the heap.  The run's steps (reset, boot, expand, reduce, root
 *   ports) are captured in a HIP graph and replayed as one launch; where
 *   capture is refused they are launched one by one.  The handle's own net
 *   is gone after a plan (hvm_gpu_reset makes a fresh one).  tuning may be
 *   null (defaults).
 * hvm_gpu_batch_run: runs the plan once; ports[i] (may be null) receives the
 *   port net i's root leads to, roots[i] (may be null) its root cell, for
 *   hvm_gpu_readback_root.  hvm_gpu_rewrites then reports the batch's sums.
 * hvm_gpu_batch_tune: times the plan under candidate tunings, one parameter
 *   at a time (coordinate descent over grid size, mailbox and refill
 *   thresholds and arena chunk), and keeps the fastest; *best receives it.
 *   Trials stop after about 30 s.
 * hvm_gpu_batch_info: the plan's tuning, whether it runs as a graph, and the
 *   nodes the last run's arenas took from the pool and the overflow region.
 * hvm_gpu_batch_used: per net, the nodes the last run took from its arena
 *   and the extents it took from the pool (lanes take nodes in chunks, so
 *   this includes the unused ends of chunks): the measure that sizes the
 *   next plan's arenas.
 * hvm_gpu_alloc_count: nodes allocated since the last reset by the per-lane
 *   allocation rings (a single net or an unplanned batch), counting a node a
 *   lane reused from its free cache as a new allocation (an arena reuses
 *   nothing): the measure that sizes arenas.
 * No counterpart in the reference (whose nets each own a heap). */
typedef struct {
  uint32_t blocks;       /* reduce workgroups (0: the handle's grid) */
  uint32_t mbox_keep;    /* redexes a busy wave keeps before handing some to an idle wave */
  uint32_t fill_below;   /* pending redexes under which a busy wave claims from the global ring */
  uint32_t arena_chunk;  /* nodes a lane takes from its net's arena at a time while it stays on
                            that net (at most one granule) */
} hvm_batch_tuning;
int hvm_gpu_batch_plan(hvm_gpu *h, uint32_t n, const uint32_t *def_ids, const uint64_t *arena_nodes,
                       const hvm_batch_tuning *tuning);
int hvm_gpu_batch_run(hvm_gpu *h, uint64_t *ports, uint64_t *roots);
int hvm_gpu_batch_tune(hvm_gpu *h, hvm_batch_tuning *best);
int hvm_gpu_batch_info(hvm_gpu *h, hvm_batch_tuning *tuning, int *graph, uint64_t *pool_nodes);
int hvm_gpu_batch_used(hvm_gpu *h, uint64_t *used);
int hvm_gpu_alloc_count(hvm_gpu *h, uint64_t *nodes);

/* Serving: requests (a definition to boot, a caller's tag) admitted by the
 * running reduction kernel itself.  No counterpart in the reference (one
 * Net per Heap, reduced to normal form by Net::normal, net.rs:55-60, per
 * request).
 *
 * hvm_gpu_serve_start: starts one persistent reduction launch on the
 *   handle's heap, cut into `slots` arenas of `slot_nodes` 16-byte nodes
 *   each (a request's net must fit its slot: HVM_E_OOM stops the launch
 *   otherwise), with a request ring of at least `queue` entries in pinned
 *   host memory.  The launch keeps running while the handle serves; every
 *   other entry point is refused until hvm_gpu_serve_stop.
 * hvm_gpu_serve_submit: queues n requests: boot definition def_ids[i] as
 *   Net::boot + Net::normal would, answer with tags[i].  HVM_E_QUEUE when
 *   the ring lacks room for all n (nothing is queued then).  One wave of the
 *   launch polls the ring, and admits waiting requests into free slots (one
 *   per lane, up to 64 at a time) between its own iterations.
 * hvm_gpu_serve_poll: up to `cap` finished requests, in the order they
 *   finished: tags[k] and ports[k], the port the request's root leads to
 *   (a Num port for a numeric result; a tree's port is only valid until its
 *   slot is reused).  A request is finished when no redex of its net is
 *   left; its slot is then free for the next request.
 * hvm_gpu_serve_stop: waits until every admitted request is finished and
 *   the launch has ended; *ms (may be null) receives the launch's duration.
 *   hvm_gpu_rewrites then reports the sums over every request served. */
int hvm_gpu_serve_start(hvm_gpu *h, uint32_t slots, uint64_t slot_nodes, uint64_t queue);
int hvm_gpu_serve_submit(hvm_gpu *h, uint32_t n, const uint32_t *def_ids, const uint64_t *tags);
int hvm_gpu_serve_poll(hvm_gpu *h, uint64_t *tags, uint64_t *ports, uint32_t cap, uint32_t *n);
int hvm_gpu_serve_stop(hvm_gpu *h, double *ms);

/* ~ Book::pre_reduce (transform/src/pre_reduce.rs:38-164): reduces each
 * definition of `book` on its own net on the device, for at most max_rwts
 * interactions (reduce(limit), net.rs:41-52), its references first, and
 * writes the book with every reduced definition's read-back net (pending
 * redexes included) in its place.  `names` (n_names of them) are the
 * definitions to leave as they are, or, with only != 0, the only ones to
 * reduce (the reference's `skip` predicate).  A name the book refers to but
 * does not define is inert: a call on it counts a dereference and the redex
 * stays pending (pre_reduce.rs:44-45, 99-116); it is not written out.
 * rwts (may be null) receives the pre-reduction's counts (anni comm eras
 * dref oper).  The handle then holds the new book, label sets recomputed
 * (Host::new on it; names it refers to but does not define stay inert);
 * buffer protocol as hvm_gpu_readback. */
int hvm_gpu_pre_reduce(hvm_gpu *h, const char *book, const char *const *names, uint32_t n_names, int only,
                       uint64_t max_rwts, char *out, uint64_t cap, uint64_t *len, uint64_t rwts[5]);

/* ~ Net.rwts (runtime/src/runtime.rs:72-79): anni, comm, eras, dref, oper. */
int hvm_gpu_rewrites(hvm_gpu *h, uint64_t out[5]);

/* Interpreter shortcuts taken on ERA targets by do_op, do_op_num and
 * do_switch (runtime/src/instruction.rs:144-146, 163-165, 194-197). */
int hvm_gpu_shortcuts(hvm_gpu *h, uint64_t out[3]);

/* ~ Host::readback(&net) + Display (host/src/readback.rs:23-33,
 * ast/src/ast.rs:234-244): the net as text, the root's tree followed by
 * "\n  & a ~ b" for each pending redex (none once the net is normal; after
 * reduce(limit), the redexes still to reduce, in the device's ring order,
 * which differs from the reference's queue order).  If `cap` is too small,
 * returns HVM_E_TRUNC with the needed length (without NUL) in *len. */
int hvm_gpu_readback(hvm_gpu *h, char *buf, uint64_t cap, uint64_t *len);

/* Raw image of the device net, for host-side tools that read the net in
 * place as Host::readback does (host/src/readback.rs:23-87): the root's var
 * cell (a port naming it is where readback starts), the port address of
 * heap word 0 and the heap size in words.  When `dst` is non-null, heap
 * words [first, first + n) are copied into it.  A heap word holds one port;
 * a binary node at port address A is the two words at A and A + 8; a var
 * cell holds the port its wire leads to once the far side has linked. */
typedef struct {
  uint64_t root_cell;
  uint64_t heap_base;
  uint64_t heap_words;
} hvm_net_image;
int hvm_gpu_snapshot(hvm_gpu *h, hvm_net_image *img, uint64_t first, uint64_t n, uint64_t *dst);

/* Duration in milliseconds of the last reduction kernel (HIP events on the
 * kernel's stream), and the reduce grid: blocks and waves. */
int hvm_gpu_last_kernel_ms(hvm_gpu *h, double *ms);

/* Compiled definitions on (on != 0, the default) or off.  At book load each
 * definition is also compiled into a residual program (its instructions that
 * depend on the port meeting it, interpreted as InterpretedDef::call does,
 * runtime/src/def.rs:217-289) and a template of everything else, evaluated
 * on the host (the role of the reference's compiled definitions,
 * src/compile.rs:110-161).  Normal forms and counts are the same either way.
 * *n_compiled (if non-null) receives how many definitions of the loaded book
 * have a compiled form. */
int hvm_gpu_set_compiled(hvm_gpu *h, int on, uint32_t *n_compiled);

/* Wall-clock limit of one reduction launch (default 600 s; 0 = none).  A
 * launch that runs past it stops every wave and returns HVM_E_TIMEOUT. */
int hvm_gpu_set_timeout(hvm_gpu *h, double seconds);
int hvm_gpu_grid(hvm_gpu *h, uint32_t *blocks, uint32_t *waves);

/* Scheduler statistics of the last reduction launch, summed over waves:
 * [0] iterations with work, [1] idle iterations, [2] 100 MHz ticks spent in
 * iterations with work, [3] spills to the global ring, [4] redexes spilled,
 * [5] claims from the global ring, [6] redexes claimed, [7] ticks inside the
 * interaction rules, [8] ticks spilling; since the net was reset: [9] ring
 * slots found in use on a later lap, [10] overflow-ring allocations;
 * [11..14] iterations that ran anni / comm / call / other redexes,
 * [15..18] ticks inside the rules for each of those classes; [19] redexes
 * popped since the net was reset.  For a single net [0..8] and [11..18]
 * are 0 unless hvm_gpu_set_sched_stats turned them on; batched and serving
 * launches always collect them. */
int hvm_gpu_sched_stats(hvm_gpu *h, uint64_t out[20]);
/* Whether a single net's reduction launches collect the scheduler
 * statistics (off by default: keeping them costs 2-4% of its throughput).
 * No counterpart in the reference. */
int hvm_gpu_set_sched_stats(hvm_gpu *h, int on);

/* Host-only book utilities (no device).  Text results follow
 * hvm_gpu_readback's buffer convention; errors are described by
 * hvm_gpu_last_error(NULL). */

/* ~ Book::from_str + Display (ast/src/parser.rs:178-183, ast/src/ast.rs:221-232). */
int hvm_book_print(const char *book_text, char *buf, uint64_t cap, uint64_t *len);

/* ~ Book::from_str + Host::insert_book (ast/src/parser.rs:178-183,
 * host/src/host.rs:41-73) into a book image for hvm_gpu_load_book_image;
 * *len receives the image size in bytes (HVM_E_TRUNC if `cap` is smaller). */
int hvm_book_encode(const char *book_text, void *buf, uint64_t cap, uint64_t *len);

/* ~ calculate_label_sets (host/src/calc_labels.rs:78-93): one line per
 * definition, "name: l1 l2 ...". */
int hvm_book_label_sets(const char *book_text, char *buf, uint64_t cap, uint64_t *len);

/* ~ Net::apply_tree (ast/src/ast.rs:112-123): `net_text` applied to
 * `tree_text`, printed as a net. */
int hvm_net_apply_tree(const char *net_text, const char *tree_text, char *buf, uint64_t cap, uint64_t *len);

#ifdef __cplusplus
}
#endif

#endif /* HVM64_AMD_H */
