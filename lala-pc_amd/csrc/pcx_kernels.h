// pcx_kernels.h -- device data layout and kernel entry points (internal).
//
// HBM layout
//   table   : three int32 SoA streams, one entry per propagator in table
//             order (sorted by (op, y, x, z)):
//               xo[i] = x | op << 28,  y[i],  z[i]          (12 B / propagator)
//   store   : int2 {lb, ub} per variable (interval domain) or one 64-bit
//             word per variable (bitset domain), one slice of nvars per
//             instance, so each operand gather is one 8-byte load from one
//             cache line.
//   flags   : per instance, InstFlags (parity-indexed "changed" words so a
//             sweep reads the previous sweep's word while writing its own).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcx {

constexpr int kOpShift = 28;
constexpr uint32_t kVarMask = (1u << kOpShift) - 1u;
constexpr int kWave = 64;
constexpr int kBlock = 256;            // 4 waves per workgroup
constexpr int kNoFail = 0x7fffffff;
constexpr int kTightSlots = 64;        // striped tightening counters
constexpr int kXcds = 8;               // accelerator complex dies on MI355X

// Kernel instantiations per store domain.
constexpr int kItv = 0;       // int32 intervals, tables without division
constexpr int kItvDiv = 1;    // int32 intervals, any opcode
constexpr int kBits = 2;      // 64-bit bitset words

struct Table {
  const uint32_t* __restrict__ xo;
  const int32_t* __restrict__ y;
  const int32_t* __restrict__ z;
  int64_t n;
};

// Per-instance sweep bookkeeping, all written with plain stores of the sweep
// number (every writer of one word in one sweep writes the same value).
struct InstFlags {
  int32_t changed[2];   // changed[s & 1] = s when sweep s tightened a bound
  int32_t failed_at;    // first sweep that found an empty interval, or kNoFail
  int32_t done_at;      // sweep at which the scheduler retired the instance
  int64_t marked_evals; // evaluations done by marked-only sweeps
  int32_t extra_sweeps; // workgroup-local sweeps run after a sweep (serving bursts)
  int32_t settled_at;   // a burst after sweep s reached the fixpoint: s, else -1
  int32_t all_marked[2];// all_marked[s & 1] = s: every propagator is marked for sweep s
};

// Event-driven sweeps.  After an instance's first (full) sweep, a sweep
// evaluates only the propagators marked in the instance's bitmap for it: a
// tightening of variable v marks every propagator adjacent to v (adj_off /
// adj, table positions) that belongs to the instance (member bitmap, or all
// when member == nullptr) in the bitmap of the next sweep.  Bitmaps are
// indexed by table position, `words` 32-bit words per instance, two per
// instance (sweep parity): dirty[(s & 1) * n_inst * words + k * words ...].
//
// A summary bitmap per instance and parity has one bit per chunk of
// kChunkWords bitmap words (kChunkProps propagators) that may hold marks:
// summary[(s & 1) * n_inst * swords + k * swords ...].  The batched scheduler
// copies an instance's summary to snap (and clears it) when it schedules a
// marked sweep, so the sweep's work items are the chunks set in snap.
constexpr int kChunkWords = 256;
constexpr int kChunkShift = 13;            // propagators per chunk = 1 << 13
struct EventSet {
  const int32_t* adj_off;   // [nvars + 1]
  const int32_t* adj;       // [adj_off[nvars]]
  const uint32_t* member;   // [n_inst * words] or nullptr
  uint32_t* dirty;          // [2 * n_inst * words], or nullptr = no marking
  uint32_t* summary;        // [2 * n_inst * swords]
  uint32_t* snap;           // [n_inst * swords]
  int64_t words;
  int64_t swords;
  int64_t n_inst;
  // Queued marks (q == nullptr: none).  Every mark that sets a clear bit of
  // sweep s's bitmap is also appended to queue plane s & 1 as
  // (instance << 32 | propagator), so a marked sweep is one flat pass over
  // its plane; the bitmaps stay complete, so a plane that overflowed (or an
  // instance whose marks went to its all_marked word) is swept from the
  // bitmaps instead.  A plane is kQLists sub-lists of qcap entries; the
  // workgroup index picks the sub-list a workgroup appends to, so each
  // sub-list's tail counter is updated from one XCD only (one tail for the
  // whole plane serialized the workgroups' reservations across XCDs).
  // qctl: kQ* words.
  unsigned long long* q;    // [2 * kQLists * qcap]
  int32_t* qctl;
  int64_t qcap;             // entries per sub-list
};
constexpr int kQLists = 256;
// Queue control words.
constexpr int kQOver = 2;       // [2] the plane lost marks: its sweep needs the bitmaps
constexpr int kQSweep = 4;      // the sweep the next queue pass runs
constexpr int kQStop = 5;       // 1: nothing marked (fixpoint), 2: bitmaps needed from kQAt
constexpr int kQAt = 6;         // the sweep at which the queue passes stopped
constexpr int kQHeadWords = 8;  // the words above (what the host reads)
constexpr int kQTails = 8;      // [2][kQLists] entries appended to each sub-list (may pass qcap)
constexpr int kQCtlWords = kQTails + 2 * kQLists;

// Per-instance propagator subset: prop_idx[off[k] .. off[k+1]) of the shared
// table, the whole table when off == nullptr, or -- off without idx -- the
// instance's own copy of its subset at [off[k], off[k+1]) of the table.
struct Subsets {
  const int64_t* __restrict__ off;
  const int32_t* __restrict__ idx;
};

// Control words of a batch: [0] n_act, [1] sweep, [2] live, [3] persistent
// queue head, [4] marked chunks per work item of the current sweep,
// [8 .. 8+kXcds) per-XCD item counters of the current sweep.
constexpr int kCtlWords = 16;
constexpr int kCtlXq = 8;

// Device work list built by the scheduler for one sweep of the batched mode.
struct WorkList {
  int32_t* act;          // active instance ids, length n_act
  int64_t* item_base;    // prefix sum of work items per active instance (n_act + 1)
  int32_t* n_act;        // [1]
  int32_t* sweep;        // [1] current sweep number
  int32_t* live;         // [1] instances still running after the last schedule;
                         //     live[1] is the persistent mode's queue head
  int32_t* xq;           // [kXcds] items claimed so far on each XCD (zeroed per sweep)
  int32_t* group;        // [1] marked sweeps: chunks with marks per work item
};
inline WorkList worklist_of(int32_t* act, int64_t* item_base, int32_t* ctl) {
  return WorkList{act, item_base, ctl, ctl + 1, ctl + 2, ctl + kCtlXq, ctl + 4};
}

// One serving slot: a request's own table and store (heterogeneous batch).
struct SlotDesc {
  const uint32_t* xo;
  const int32_t* y;
  const int32_t* z;
  int64_t m;             // propagators
  int2* store;           // the request's store (nvars entries)
  // Marked sweeps (adj == nullptr: every sweep of the slot is a full sweep).
  // marks = [2 * words] mark bitmaps (sweep parity), [2 * swords] summaries,
  // [swords] scheduler snapshot; zeroed at admission.
  const int32_t* adj_off;
  const int32_t* adj;
  uint32_t* marks;
  int64_t words;
  int64_t swords;
  int32_t start;         // the request's first sweep (a full sweep)
  int32_t handover;      // side lane: the slot-batch sweep after which it left the batch
};

// done_at markers of a slot in a side lane (see launch_burst_slots): the
// scheduler never runs a slot whose done_at is set, so the slot batch leaves
// it alone while a burst owns it.
constexpr int32_t kSideLane = 0x7fffff00;   // + lane: in that lane's bursts (kept when the budget ran out)
constexpr int32_t kSideBail = 0x7ffffff1;   // the burst stopped: its marks need the slot batch

}  // namespace pcx

// Host-callable launchers (defined in pcx_kernels.hip).
namespace pcx {
hipError_t launch_sweep_plain(const Table& t, int2* store, const int32_t* idx, int64_t m,
                              int32_t sweep, InstFlags* flags, unsigned long long* tight,
                              int ppw, int dom, bool stream_table, hipStream_t st,
                              const EventSet* ev = nullptr, int64_t inst = 0);
// Marked-only sweep of one instance (plain mode), sweep >= 1.
hipError_t launch_sweep_marked_plain(const Table& t, int2* store, int32_t sweep, InstFlags* flags,
                                     unsigned long long* tight, const EventSet& ev, int64_t inst,
                                     int dom, hipStream_t st);
// Marked-only sweep over every running instance (batched mode).
hipError_t launch_sweep_marked(const Table& t, int2* store, int64_t nvars, InstFlags* flags,
                               WorkList wl, unsigned long long* tight, const EventSet& ev,
                               int dom, int grid, hipStream_t st);
// member bitmaps from subsets
hipError_t launch_build_members(const int64_t* off, const int32_t* idx, int64_t n_inst,
                                uint32_t* member, int64_t words, hipStream_t st);
// Queued marked sweeps: one turn (advances kQSweep, or stops) and one flat
// pass over the plane of that sweep; both return at once after a stop.
hipError_t launch_queue_turn(int32_t* qctl, int32_t max_sweeps, hipStream_t st);
// Instance k's marks for sweep s (bitmap row, or every member propagator
// when its all_marked word is set) appended to queue plane s & 1.
hipError_t launch_queue_from_bitmap(const EventSet& ev, InstFlags* flags, int64_t k, int32_t s,
                                    int64_t n_props, hipStream_t st);
hipError_t launch_sweep_queue(const Table& t, int2* store, int64_t nvars, InstFlags* flags,
                              unsigned long long* tight, const EventSet& ev, int dom, int grid,
                              hipStream_t st);
// marked != nullptr: schedule a marked sweep (items = chunks with marks).
hipError_t launch_schedule(const Subsets& sub, int64_t n_props, InstFlags* flags, int32_t n_inst,
                           WorkList wl, int ppw, int32_t max_sweeps, hipStream_t st,
                           const EventSet* marked = nullptr);
hipError_t launch_sweep_batched(const Table& t, int2* store, int64_t nvars, const Subsets& sub,
                                int64_t n_props, InstFlags* flags, WorkList wl,
                                unsigned long long* tight, int ppw, int dom, bool stream_table,
                                int grid, hipStream_t st, const EventSet* ev = nullptr);
// Lockstep full sweep (gi in {16, 32, 64} instances per group): shadow is the
// instance-interleaved copy of the stores made by launch_interleave (groups
// of gi instances; ceil(n_inst / gi) * gi * nvars words), member the member
// bitmaps of a subset batch (nullptr: every propagator).
hipError_t launch_interleave(const int2* store, int2* shadow, int64_t nvars, int64_t n_inst, int gi,
                             hipStream_t st);
// Where a load lists how its stores differ from a base store: the base
// holds, per variable, the value instances 0, 1 and 2 agree on (two of the
// three; instance 0's when they all differ or the batch is smaller), i.e.
// the value most instances of a search-node batch share.  Entry
// (instance << 32 | variable) for every value that differs from the base,
// appended in no particular order to one of kDiffLists sub-lists (the
// workgroup index picks it, so each sub-list's counter is updated from one
// XCD only: a single counter for the whole device serialized its atomics
// across XCDs).  Sub-list c holds list[c * cap, c * cap + count[c]); a count
// past cap means the list is incomplete (and unused); a workgroup stops
// listing once it sees its sub-list full.  list == nullptr: nothing is
// listed.
constexpr int kDiffLists = 256;
struct DiffSink {
  unsigned long long* list;   // [kDiffLists * cap]
  int32_t* count;             // [kDiffLists]
  int64_t cap;                // entries per sub-list
  int2* base;                 // [nvars]: the base store, written by the load
  int2* vals;                 // [kDiffLists * cap]: each entry's value (nullptr: not kept)
};
// Device-resident lb / ub (n_inst * nvars each) into the stores and their
// interleaved copy in one pass; empty[k] = 1 where instance k has lb > ub
// (empty zeroed by the caller); the base store and the differences from it
// into ds.
hipError_t launch_ingest_interleave(const int32_t* lb, const int32_t* ub, int2* store, int2* shadow,
                                    int64_t nvars, int64_t n_inst, int gi, uint8_t* empty,
                                    const DiffSink& ds, hipStream_t st);
// The base store of device-resident lb / ub rows (rows 0-2) into ds.base,
// before launch_ingest_rows lists the differences from it.
hipError_t launch_make_base_rows(const int32_t* lb, const int32_t* ub, int64_t nvars, int64_t n_inst,
                                 const DiffSink& ds, hipStream_t st);
// Device-resident lb / ub into the stores only (no image copy), with the
// differences from the base store (made first) and their values into ds,
// and the emptiness; a complete list is the image (launch_image_from_diffs).
// held (nullable, nvars values): stores that hold these values in every row
// (launch_stores_to_base), of which only the differing values are written.
hipError_t launch_ingest_rows(const int32_t* lb, const int32_t* ub, int2* store, int64_t nvars,
                              int64_t n_inst, uint8_t* empty, const DiffSink& ds, const int2* held,
                              hipStream_t st);
// store[k][v] = the base value, then every listed difference's value: the
// stores a complete difference list (with values) was taken from.
hipError_t launch_image_from_diffs(const DiffSink& ds, int2* store, int64_t nvars, int64_t n_inst,
                                   hipStream_t st);
// The same base and differences from stores already in device memory
// (n_inst * nvars words, row by row).
hipError_t launch_list_diffs(const int2* store, int64_t nvars, int64_t n_inst, const DiffSink& ds,
                             hipStream_t st);
// Propagators the store `base` (nvars words) does not satisfy:
// those that would tighten it, or that read an empty interval.  Their table
// positions are appended to list (room for t.n), their number to *count.
hipError_t launch_base_probe(const Table& t, const int2* base, int32_t* list, int32_t* count,
                             int dom, hipStream_t st);
// Lists kept per instance: instance k's n[k] entries at list + k * cap
// (entries past cap are only counted: the list is then incomplete).
struct InstLists {
  int32_t* list;     // [n_inst * cap], or nullptr
  int32_t* n;        // [n_inst]
  int32_t cap;
};
// Marks of a batch's first sweep (plane 0) seeded from a base store: for
// every listed difference (k, v) (complete sub-lists) the member
// propagators of instance k adjacent to v, and every propagator of `unsat`
// in every instance whose subset holds it.  With seed lists (their counts
// zeroed by the caller; only without summaries and queues) the seeds are
// listed per instance instead, and only those past a list's cap are marked.
hipError_t launch_seed_from_base(const EventSet& ev, const DiffSink& diffs, int64_t n_diffs,
                                 const int32_t* unsat, int64_t n_unsat, const InstLists& seeds,
                                 hipStream_t st);
// The marked phase of a seeded run (marks in plane 0), one instance per
// one-wave workgroup taken from *next_inst (zeroed by the caller): every
// instance to its fixpoint in rounds separated by workgroup barriers, in
// one launch; flags record the rounds as sweeps.  Round 0 takes the seed
// lists when given (and the row's marks when a list is incomplete).  With a
// trail, each instance also lists the variables it tightens, so that a
// restore rewrites only those.
hipError_t launch_drain_marks(const Table& t, int2* store, int64_t nvars, InstFlags* flags, const EventSet& ev,
                              int32_t* next_inst, unsigned long long* tight, const InstLists& seeds,
                              const InstLists& trail, int dom, int grid, hipStream_t st);
// The stores back to their image after a drained run with a trail: each
// instance's listed variables (its whole row when the list is incomplete)
// from the image, which is either d_image or the base store plus the listed
// differences (ds.vals non-null: written after the base values).
hipError_t launch_restore_trail(const InstLists& trail, const DiffSink& ds, const int2* image, int2* store,
                                int64_t nvars, int64_t n_inst, hipStream_t st);
// Stores that hold their image (base store plus a complete difference list)
// except at a drained run's trail (trail non-null) back to the base store in
// every row: the trail's variables (a row past its cap whole) and the listed
// differences' positions get their base values.
hipError_t launch_stores_to_base(const InstLists* trail, const DiffSink& ds, int2* store, int64_t nvars,
                                 int64_t n_inst, hipStream_t st);
// dst (n_inst * nvars words, row by row) from an interleaved copy
hipError_t launch_deinterleave(const int2* shadow, int2* dst, int64_t nvars, int64_t n_inst, int gi,
                               hipStream_t st);
hipError_t launch_sweep_lockstep(const Table& t, int2* store, int2* shadow, int64_t nvars,
                                 int32_t n_inst, const uint32_t* member, int64_t words,
                                 InstFlags* flags, WorkList wl, unsigned long long* tight, int gi,
                                 int dom, int grid, hipStream_t st, const EventSet* ev = nullptr);
constexpr int kMarkedWordsPerItem = kChunkWords;   // bitmap words per marked-sweep item
// lds: stage each instance's store in LDS (nvars * 8 <= lds_cap, at most
// lds_store_cap(device): the device's LDS per workgroup, 160 KB on gfx950,
// less a reserve for the kernel's static variables; kLdsStoreBytes when the
// runtime does not report it).
constexpr int64_t kLdsStoreBytes = 64 * 1024;
constexpr int64_t kLdsStaticReserve = 256;
int64_t lds_store_cap(int device);
hipError_t launch_persistent(const Table& t, int2* store, int64_t nvars, const Subsets& sub,
                             int64_t n_props, InstFlags* flags, const int32_t* order,
                             int32_t n_inst, int32_t* queue, unsigned long long* tight,
                             int32_t max_sweeps, int dom, int grid, bool lds, int64_t lds_cap,
                             hipStream_t st);
// resident workgroups per CU (lds_bytes of dynamic LDS per workgroup)
int persistent_blocks_per_cu(int dom, int64_t lds_bytes);
// empty (may be nullptr): per-instance bytes, 1 = the store starts empty
hipError_t launch_init_flags(InstFlags* flags, int32_t n_inst, WorkList wl, hipStream_t st,
                             const uint8_t* empty = nullptr);
hipError_t launch_store_empty_map(const int2* store, int64_t nvars, int32_t n_inst, uint8_t* empty,
                                  int dom, hipStream_t st);
hipError_t launch_schedule_slots(const SlotDesc* slots, InstFlags* flags, int32_t n_slots,
                                 WorkList wl, int ppw, hipStream_t st);
hipError_t launch_sweep_slots(const SlotDesc* slots, InstFlags* flags, WorkList wl,
                              unsigned long long* tight, int ppw, int dom, int grid,
                              hipStream_t st);
// Side-lane bursts: one workgroup per slot whose done_at is kSideLane + lane
// propagates it alone for up to `budget` rounds (marked slots while their
// marks fit one workgroup, slots with full sweeps of at most kBurstFullMax
// propagators).  Runs on its own stream beside the slot batch.
constexpr int64_t kBurstFullMax = 8192;
constexpr int64_t kBurstSummaryMax = 512;   // summary words of a marked slot in a burst
hipError_t launch_burst_slots(const SlotDesc* slots, InstFlags* flags, int32_t n_slots,
                              unsigned long long* tight, int32_t budget, int32_t lane, int dom,
                              hipStream_t st);
hipError_t launch_compact_table(const Table& t, const int32_t* idx, int64_t n, uint32_t* xo,
                                int32_t* y, int32_t* z, hipStream_t st);
// Entailment over a batch: out_idx == nullptr counts (live[k] = propagators
// not entailed, empty[k] = store holds an empty interval); otherwise writes
// those propagators' table positions at out_off[k].
hipError_t launch_batch_ask(const Table& t, const int2* store, int64_t nvars, const Subsets& sub,
                            int64_t n_props, int64_t n_inst, int64_t* live, int32_t* empty,
                            const int64_t* out_off, int32_t* out_idx, hipStream_t st);
// Branching: var[k] = split variable of parent instance k (or -1); children
// 2k / 2k+1 written to child and child_image, and their emptiness (that of
// their parent) to child_empty when given; one pass over the parents.
hipError_t launch_branch(const int2* parent, int64_t nvars, int64_t n_inst, int32_t* var, int2* child,
                         int2* child_image, uint8_t* child_empty, int dom, hipStream_t st);
hipError_t launch_branch_seed(const EventSet& ev, const int32_t* var, int64_t n_child, hipStream_t st);
hipError_t launch_ask(const Table& t, const int2* store, uint8_t* ent, int32_t* all_flag,
                      hipStream_t st);
// PC's formula trees (pcx_model.h node lists, validated): ok[k] = 0 where a
// formula is not asked on instance k's store (ok set to 1 by the caller).
hipError_t launch_formula_ask(const int4* nodes, const int64_t* off, int64_t n_formulas, const int2* store,
                              int64_t nvars, int64_t n_inst, int32_t* ok, hipStream_t st);
hipError_t launch_store_pack(const int32_t* lb, const int32_t* ub, int2* store, int64_t n, hipStream_t st);
// image[v] = store[v] = {lb[v], ub[v]} for n device-resident bounds, and
// empty[v / nvars] = 1 where lb > ub (empty zeroed by the caller)
hipError_t launch_ingest(const int32_t* lb, const int32_t* ub, int2* image, int2* store, int64_t n,
                         int64_t nvars, uint8_t* empty, hipStream_t st);
// device copy of n 8-byte words (16-byte aligned buffers)
hipError_t launch_copy_words(const void* src, void* dst, int64_t n, hipStream_t st);
hipError_t launch_store_unpack(const int2* store, int32_t* lb, int32_t* ub, int64_t n, hipStream_t st);
hipError_t launch_store_check_empty(const int2* store, int64_t nvars, int32_t n_inst,
                                    InstFlags* flags, int dom, hipStream_t st);
}  // namespace pcx
