/*
 * pcx.h -- C ABI of the MI355X propagator-completion engine.
 *
 * The engine holds a table of ternary propagators `x = y <op> z` (the PIR
 * bytecode form of lattice-land/lala-pc, include/lala/pir.hpp:36-47) and runs
 * their fixpoint over stores of int32 intervals on a gfx950 GPU.  It replaces,
 * for that path, the reference's
 *     GaussSeidelIteration{}.fixpoint(pir.num_deductions(),
 *                                     [&](size_t i) { return pir.deduce(i); })
 * (tests/pir_test.cpp:60-62; engine in lala-core fixpoint.hpp, not vendored).
 *
 * Conventions
 *  - A store is two int32 arrays lb[nvars], ub[nvars].  PCX_MINF / PCX_INF
 *    (INT32_MIN / INT32_MAX) stand for -oo / +oo (pir.hpp:403-404).  An
 *    interval is empty (bot) when lb > ub.
 *  - Failure is a result field, never an error status (pir.hpp:831-833).
 *  - No function throws; every failure is a pcx_status with a message from
 *    pcx_last_error() (thread-local).
 *  - One host thread per engine handle (pc.hpp:618-623: telling is not
 *    thread-safe in the reference either).  Device work is asynchronous
 *    inside a call but every call returns with its results complete.
 */
#ifndef PCX_H
#define PCX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pcx_status;
#define PCX_OK       0
#define PCX_EINVAL   1   /* bad argument, opcode or variable index           */
#define PCX_ENOMEM   2   /* device or host allocation failed                 */
#define PCX_EDEVICE  3   /* HIP runtime error (no device, launch failure ...) */
#define PCX_EBUSY    4   /* serving: no capacity to admit the request now     */
#define PCX_ENOTSUP  5   /* valid request this build does not implement       */

/* Opcodes of `x = y <op> z`.  Subtraction is an ADD with permuted operands
 * and x != y an EQ whose x is the constant 0, as in the reference
 * (pir.hpp:270-273; pir_test.cpp:277-283, 544-548).  The numbering is this
 * library's; it only fixes the table sort order (see pcx_tell). */
enum {
  PCX_ADD = 0, PCX_MUL = 1, PCX_TDIV = 2, PCX_FDIV = 3, PCX_CDIV = 4,
  PCX_EDIV = 5, PCX_MIN = 6, PCX_MAX = 7, PCX_EQ = 8, PCX_LEQ = 9,
  PCX_NUM_OPS = 10
};

/* Store domains. */
#define PCX_DOMAIN_INTERVAL 0   /* int32 lb/ub per variable                     */
#define PCX_DOMAIN_BITSET   1   /* one 64-bit word per variable, bit v = value v */

/* Opcodes of the bitset domain, numbered in the order PC sorts its formula
 * kinds (pc.hpp:634-641 over the variant order of formula.hpp:905-920). */
enum {
  PCX_BLEQ = 0,   /* y <= z              Inequality   formula.hpp:796-805      */
  PCX_BADD = 1,   /* x = y + z           Equality + Binary<GroupAdd> residuals */
                  /*                     formula.hpp:672-681, terms.hpp:376-397 */
  PCX_BNEQ = 2,   /* y != z              Disequality  formula.hpp:638-662      */
  PCX_BEQR = 3,   /* x <=> (y = z)       Biconditional formula.hpp:421-427     */
  PCX_BITS_NUM_OPS = 4
};

#define PCX_INF  INT32_MAX
#define PCX_MINF INT32_MIN
#define PCX_MAX_VARS (1 << 28)   /* variable indices must be < 2^28 */

typedef struct pcx_engine pcx_engine;
typedef struct pcx_batch pcx_batch;

typedef struct {
  int32_t device;           /* HIP device ordinal                                   */
  int32_t props_per_wave;   /* 64 .. 512 (powers of two; batched mode also 1024 and
                               2048: work items of 2 or 4 rounds of 512); 0 = chosen
                               from the sweep size                                   */
  int32_t lds_mode;         /* persistent mode: stage each store in LDS; 0 never,
                               1 or -1 (auto) whenever it fits in a workgroup's
                               LDS (160 KB on MI355X)                              */
  int32_t sweeps_per_check; /* plain engine: sweeps queued between host checks (0=8) */
  int64_t max_sweeps;       /* 0 = run to the fixpoint                               */
  int32_t timing;           /* 1 = record HIP events around every sweep kernel       */
  int32_t domain;           /* PCX_DOMAIN_INTERVAL (default) or PCX_DOMAIN_BITSET    */
  int32_t compact_tables;   /* batches with subsets: -1 auto, 0 read the shared table
                               through the subset indices, 1 give each instance its
                               own copy of its subset (12 B per active propagator)  */
  int32_t stream_tables;    /* -1 auto (stream per-instance tables only), 0 cached
                               table loads, 1 non-temporal table loads               */
  int32_t event_driven;     /* -1 auto (batched mode; plain mode from 2^20
                               propagators per instance), 0 every sweep evaluates
                               every propagator, 1 sweeps after the first evaluate
                               only propagators with an operand tightened since     */
  int32_t lockstep;         /* batched mode, full sweeps: 0 per-instance work items;
                               16, 32 or 64: one propagator for that many instances
                               per step, over an instance-interleaved copy of the
                               stores (8 B per variable per instance); -1 auto     */
  int32_t mark_queue;       /* marked sweeps: 0 scan the mark bitmaps; k in 10..26:
                               also list every new mark in a queue of 2^k entries
                               per sweep parity and sweep the queue (a full queue
                               falls back to the bitmaps; same fixpoint); -1 auto:
                               a queue sized to the batch for batched runs whose
                               full sweep is lockstep, the bitmaps otherwise     */
  int32_t base_seed;        /* batched mode, event-driven: a run that starts from
                               the loaded stores first evaluates only the
                               propagators adjacent to a value that differs from
                               a base store (per variable the value two of
                               instances 0-2 hold) and those the base store does
                               not satisfy (the others compute in instance k what
                               they compute on the base), when that is much less
                               than a full sweep; 0 off, 1 or -1 (auto) on.  The
                               fixpoints are the same.                           */
} pcx_config;

typedef struct {
  int32_t has_changed;      /* some bound was tightened                      */
  int32_t failed;           /* the store became bot (some lb > ub)           */
  int64_t sweeps;           /* sweeps run, including the final quiet one     */
  int64_t evals;            /* propagator evaluations = sum of sweep sizes   */
} pcx_result;

typedef struct {
  double  wall_ms;          /* host wall time of the call                    */
  double  kernel_ms;        /* summed sweep-kernel durations (config.timing) */
  int64_t kernel_launches;  /* sweep-kernel launches                         */
  int64_t evals;            /* total evaluations over all instances          */
  int64_t tightened;        /* successful bound tightenings (atomic updates) */
  int64_t sweeps_max;       /* largest per-instance sweep count              */
  /* the part of the above spent in full sweeps (every active propagator of
     every running instance; with event-driven sweeps only the first sweep) */
  double  full_ms;
  int64_t full_launches;
  int64_t full_evals;
  int64_t full_tightened;
  int64_t lockstep;         /* instances per lockstep group of the full sweeps
                               (batched mode), 0 for per-instance work items  */
  double  load_ms;          /* device time of the pcx_batch_load_device this
                               run started from (config.timing; reported by
                               the first run after the load, else 0)         */
} pcx_stats;

/* ---- library ---------------------------------------------------------- */
const char* pcx_version(void);
const char* pcx_last_error(void);
void        pcx_config_default(pcx_config* cfg);
/* Number of HIP devices visible (0 when none); never fails. */
int32_t     pcx_device_count(void);

/* ---- engine (propagator table) ----------------------------------------
 * pcx_create: engine on cfg->device with an empty table.
 *   replaces PIR(AType, sub_ptr, alloc) (pir.hpp:158-163).               */
pcx_status pcx_create(const pcx_config* cfg, pcx_engine** out);
void       pcx_destroy(pcx_engine* e);

/* pcx_tell: append n propagators, embed x in [0,1] for EQ/LEQ into the
 * caller's store (lb/ub may be NULL to skip), then stable-sort the whole
 * table by (op, y, x, z).
 *   replaces PIR::interpret_tell + PIR::deduce(tell_type) (pir.hpp:316-352);
 *   formulas arrive already in ternary form, so interpretation reduces to
 *   opcode/index validation (PCX_EINVAL for an unknown opcode or an index
 *   outside [0, nvars) when nvars > 0, or >= PCX_MAX_VARS).              */
pcx_status pcx_tell(pcx_engine* e, const int32_t* op, const int32_t* x,
                    const int32_t* y, const int32_t* z, int64_t n,
                    int32_t* lb, int32_t* ub, int64_t nvars);

/* Read / change an engine's configuration (device and domain cannot change:
 * PCX_EINVAL).  Takes effect at the next run. */
pcx_status pcx_get_config(const pcx_engine* e, pcx_config* cfg);
pcx_status pcx_set_config(pcx_engine* e, const pcx_config* cfg);

/* replaces PIR::num_deductions (pir.hpp:382-384). */
int64_t    pcx_num_deductions(const pcx_engine* e);

/* Copy the table in engine (sorted) order; any pointer may be NULL. */
pcx_status pcx_get_table(const pcx_engine* e, int32_t* op, int32_t* x,
                         int32_t* y, int32_t* z);

/* pcx_snapshot / pcx_restore: the table size is the snapshot; restore pops
 * the propagators told since (pir.hpp:857-870).  Stores are the caller's. */
int64_t    pcx_snapshot(const pcx_engine* e);
pcx_status pcx_restore(pcx_engine* e, int64_t snapshot);

/* pcx_fixpoint: run every propagator to the fixpoint on one store held in
 * host memory (copied in, propagated on the GPU, copied back).
 *   replaces GaussSeidelIteration::fixpoint over PIR::deduce(int)
 *   (tests/pir_test.cpp:60-62, 82-86; pir.hpp:387-390, 721-817).
 * On return lb/ub hold the fixpoint; after a failure they hold some empty
 * store (bounds after failure are order-dependent, only `failed` is fixed). */
pcx_status pcx_fixpoint(pcx_engine* e, int32_t* lb, int32_t* ub, int64_t nvars,
                        pcx_result* out);

/* pcx_fixpoint_noalloc: pcx_fixpoint for callers that cannot allocate
 * (tight serving loops, harnesses that forbid hipMalloc).  No counterpart in
 * the reference, whose fixpoint engine allocates nothing on a device.
 * d_lb / d_ub: nvars int32 bounds each in DEVICE memory, the store in and
 * its fixpoint out (after a failure some empty store).  The call allocates
 * nothing: its working state (the packed store, flags and counters) lives
 * in one fixed device array of the library per GPU, which holds stores of
 * at most PCX_NOALLOC_MAX_VARS variables (PCX_EINVAL beyond that), and the
 * engine's table must fit its device copy (after a tell that grew the
 * table, one pcx_fixpoint call sizes it; PCX_EINVAL otherwise).  Sweeps are
 * full sweeps (no marks, which would need memory per propagator); the
 * fixpoint is the one pcx_fixpoint gives.  Calls that overlap on one GPU
 * share that array and race: one call at a time per GPU. */
#define PCX_NOALLOC_MAX_VARS (1 << 21)
pcx_status pcx_fixpoint_noalloc(pcx_engine* e, int32_t* d_lb, int32_t* d_ub, int64_t nvars,
                                pcx_result* out);
/* PCX_NOALLOC_MAX_VARS, for bindings that do not read this header. */
int64_t    pcx_noalloc_max_vars(void);

/* pcx_ask_all: entailment of every propagator on a store,
 *   replaces PIR::ask(int) (pir.hpp:368-370, 417-438) for all i at once.
 * entailed[i] (may be NULL) receives 0/1 per propagator in table order;
 * *all (may be NULL) receives 1 iff every propagator is entailed.        */
pcx_status pcx_ask_all(pcx_engine* e, const int32_t* lb, const int32_t* ub,
                       int64_t nvars, uint8_t* entailed, int32_t* all);

/* ---- batches: many stores over one shared table ------------------------
 * Instance k owns store slice [k*nvars, (k+1)*nvars) and runs the subset of
 * propagators prop_idx[inst_off[k] .. inst_off[k+1]) (indices into the table
 * in engine order, ascending), or the whole table when inst_off == NULL.
 * This is the reference's shared-table clone model (pir.hpp:182-205:
 * bytecodes shared, store cloned per block), with search-node stores.
 * Stores stay resident in device memory between calls.                   */
pcx_status pcx_batch_create(pcx_engine* e, int64_t n_inst, int64_t nvars,
                            const int64_t* inst_off, const int32_t* prop_idx,
                            pcx_batch** out);
void       pcx_batch_destroy(pcx_batch* b);
/* Upload n_inst*nvars bounds; they also become the reset image. */
pcx_status pcx_batch_load(pcx_batch* b, const int32_t* lb, const int32_t* ub);
/* The same from DEVICE memory: d_lb / d_ub hold n_inst*nvars bounds in the
 * same order (e.g. stores a caller's kernel produced, or inputs staged in HBM
 * ahead of time); no host transfer.  One pass writes the stores and the
 * per-instance emptiness, and keeps the reset image (as a copy, or -- stores
 * that share most values -- as a base store and the values that differ from
 * it).  When the batch's stores hold the last such image (or that image but
 * for the variables a run tightened), the pass writes only the loaded values
 * that differ from the base store.  The source may be reused or freed once
 * the call returns.  Interval batches. */
pcx_status pcx_batch_load_device(pcx_batch* b, const int32_t* d_lb, const int32_t* d_ub);
/* PC's is_extractable on every store of the batch (pc.hpp:727-738: the
 * store is not empty and every formula is asked), for formula trees in the
 * postorder node form of pcx_model.h (nodes: 4 * n_nodes int32, off:
 * n_formulas + 1 offsets; pcx_model_formulas gives a model's trees).  One
 * GPU thread per (store, formula) interprets its tree.  extractable[k] is
 * 1 or 0.  Errors: PCX_EINVAL for a malformed tree (unknown kind, a variable
 * outside the store, a stack deeper than 32 or left with more than one
 * value).  Interval batches. */
pcx_status pcx_batch_ask_formulas(pcx_batch* b, const int32_t* nodes, int64_t n_nodes, const int64_t* off,
                                  int64_t n_formulas, int32_t* extractable);
/* Restore every store from the reset image (device to device).  After a
 * run that drained its marks from the image, only the variables that run
 * tightened are rewritten; stores that are the image are left as they are. */
pcx_status pcx_batch_reset(pcx_batch* b);
pcx_status pcx_batch_read(pcx_batch* b, int32_t* lb, int32_t* ub);

/* Entailment over a batch, on the current stores (after pcx_batch_run):
 *   replaces PIR::ask(int) / PIR::is_extractable (pir.hpp:368-370, 417-438,
 *   873-884) for every instance at once.
 * extractable (n_inst entries, may be NULL): 1 when the store holds no empty
 *   interval and every propagator of the instance is entailed (the interval
 *   store itself is taken as extractable).
 * live_off (n_inst + 1, may be NULL): prefix offsets of the propagators that
 *   are not entailed; live_idx (capacity cap, may be NULL): their table
 *   positions, ascending per instance -- entailed propagators bring no more
 *   information (pir.hpp:886-889), so these are the subsets a next batch of
 *   the same search nodes needs.  PCX_EINVAL when cap is too small
 *   (live_off[n_inst] then holds the size needed).  Interval domain only. */
pcx_status pcx_batch_ask(pcx_batch* b, int32_t* extractable, int64_t* live_off,
                         int32_t* live_idx, int64_t cap);

/* Branching on the device: the search step after a fixpoint (the split of a
 * search node into two, as the reference's tests build splits with embed,
 * pir.hpp:354-356, and snapshot/restore, pir.hpp:857-870).  child must hold
 * 2 * parent's instances over the same engine and nvars (its subsets are the
 * caller's; typically each parent's twice).  For parent instance k, the
 * branching variable is the first variable, in input order, that is not
 * fixed; child 2k gets its lower part and child 2k+1 the rest: intervals are
 * bisected ([lb, mid] / [mid + 1, ub], mid = lb + floor((ub - lb) / 2)),
 * bitsets split into {min} / the other values.  A parent that is empty or
 * fully fixed has two copies as children and var -1.  The children become
 * the child batch's stores and its reset image.  var (n_inst entries of the
 * parent, may be NULL) receives the branching variables.
 * from_fixpoint = 1 declares that every parent store is a fixpoint of its
 * instance's propagators (the parent was run and not changed since) and that
 * each child's subset is contained in its parent's: the child batch's next
 * run then starts from marks on the propagators adjacent to the split
 * variable instead of a full first sweep (same fixpoint, far less work).  The
 * seeds are used by that one run (later runs after a reset sweep fully), and
 * only when the run uses marked sweeps (not the persistent mode, not
 * event_driven = 0). */
pcx_status pcx_batch_branch(pcx_batch* parent, pcx_batch* child, int32_t* var,
                            int32_t from_fixpoint);

#define PCX_MODE_PLAIN      0  /* instance by instance, one sweep launch each   */
#define PCX_MODE_BATCHED    1  /* all instances per sweep launch, device queue  */
#define PCX_MODE_PERSISTENT 2  /* one persistent launch for the whole batch     */
/* Propagate every instance to its fixpoint.  results: n_inst entries or NULL. */
pcx_status pcx_batch_run(pcx_batch* b, int32_t mode, pcx_result* results,
                         pcx_stats* stats);

/* ---- autotuning ---------------------------------------------------------
 * One point of the search: how a batch is swept. */
typedef struct {
  int32_t mode;             /* PCX_MODE_BATCHED or PCX_MODE_PERSISTENT          */
  int32_t props_per_wave;   /* batched: 64 .. 2048                              */
  int32_t lds_mode;         /* persistent: 0 or 1 (stores staged in LDS)        */
  int32_t compact_tables;   /* 0 or 1 (only varied for batches with subsets)    */
  int32_t stream_tables;    /* 0 or 1                                           */
  int32_t event_driven;     /* batched: 0 or 1                                  */
  int32_t lockstep;         /* batched: 0, 16, 32 or 64                         */
  double  ms;               /* best host time of one reset + run; < 0: failed   */
} pcx_tuning;

/* Time every point of {batched: props_per_wave x compact x stream x
 * event_driven, and lockstep 16/32/64 with event_driven} U {persistent: lds (where the store fits) x compact} on the
 * batch's loaded image (one warm-up and `reps` timed runs per point, best
 * kept; the points shape the full sweeps and the sweeps after them, so the
 * runs take a full first sweep whatever base_seed says), set the engine's
 * configuration to the fastest point, and restore the stores to the image.  results (may be NULL): up to max_results points
 * in the order tried; *n_results (may be NULL) receives how many were tried.
 * No counterpart in the reference (its fixpoint has no tuning parameters). */
pcx_status pcx_batch_autotune(pcx_batch* b, int32_t reps, pcx_tuning* best,
                              pcx_tuning* results, int32_t max_results,
                              int32_t* n_results);

/* ---- bitset domain (engine created with config.domain = PCX_DOMAIN_BITSET)
 * The store is one uint64 word per variable (bit v <=> value v in [0,63]);
 * a zero word is empty.  Batches are created with pcx_batch_create and run
 * with pcx_batch_run exactly as interval batches.
 * pcx_bits_tell: append propagators, embed EQR outputs into {0,1} in
 *   `words` (may be NULL), keep the table stably sorted by opcode.
 *   replaces PC::interpret_tell + PC::deduce(tell_type) (pc.hpp:605-645). */
pcx_status pcx_bits_tell(pcx_engine* e, const int32_t* op, const int32_t* x,
                         const int32_t* y, const int32_t* z, int64_t n,
                         uint64_t* words, int64_t nvars);
/* replaces GaussSeidelIteration::fixpoint over PC::deduce(int)
 * (tests/pc_bitset_test.cpp:45-61). */
pcx_status pcx_bits_fixpoint(pcx_engine* e, uint64_t* words, int64_t nvars,
                             pcx_result* out);
pcx_status pcx_bits_batch_load(pcx_batch* b, const uint64_t* words);
pcx_status pcx_bits_batch_read(pcx_batch* b, uint64_t* words);

/* ---- serving: continuous batching of fixpoint requests ------------------
 * The reference has no server; each request here is what one reference
 * caller does per store: GaussSeidelIteration::fixpoint over PIR::deduce
 * (tests/pir_test.cpp:60-62) on a copy of a shared network (pir.hpp:182-205).
 * A server owns one device arena for request stores and a fixed number of
 * slots.  Each step admits queued requests while a slot is free and the
 * arena holds their store (admission by HBM footprint, first fit), runs
 * sweeps_per_tick sweeps over every running request at once (requests of
 * different networks and sizes in one launch), and completes the requests
 * that reached their fixpoint or failed.  A request's result is the one
 * pcx_fixpoint gives for the same network and store, whatever else runs.
 * Interval-domain engines only. */
typedef struct pcx_server pcx_server;

typedef struct {
  int32_t device;
  int32_t max_slots;         /* requests on the GPU at once                  */
  int64_t arena_bytes;       /* device bytes for request stores (8 B / var)   */
  int32_t sweeps_per_tick;   /* sweeps launched between host checks (one in a
                                tick that admitted requests)                   */
  int32_t props_per_wave;    /* 64..512; 0 = 512                              */
  int64_t stage_bytes;       /* pinned host staging for store copies          */
  int32_t event_driven;      /* 0 full sweeps only, 1 after the first sweep
                                evaluate only marked propagators (their bitmaps
                                share the request's arena block), -1 auto: 1 for
                                requests of at least 2^17 propagators          */
  int32_t burst_sweeps;      /* side lanes: a request still changing after
                                burst_after sweeps leaves the slot batch and is
                                propagated by one workgroup of its own, beside
                                the batch, in launches of up to this many rounds
                                (marked requests while their marks fit one
                                workgroup, others up to 8,192 propagators);
                                0 off, -1 auto (1,024)                          */
  int32_t burst_after;       /* sweeps before a request may move to a side lane;
                                -1 auto (256)                                   */
  int32_t reserved[1];
} pcx_server_config;

typedef struct {
  int64_t    tag;            /* the caller's tag (replay: request index)      */
  pcx_result result;
  double     latency_ms;     /* submit (replay: arrival) to completion        */
} pcx_served;

typedef struct {
  int64_t ticks;             /* server steps that launched sweeps or idled    */
  int64_t sweeps;            /* slot-batch sweeps launched                    */
  int64_t admitted;          /* requests moved onto the GPU                   */
  int64_t completed;         /* requests completed (including empty stores)   */
  int64_t waits_slot;        /* admissions deferred: every slot busy          */
  int64_t waits_arena;       /* admissions deferred: store did not fit (HBM)  */
  int64_t arena_peak_bytes;  /* highest arena byte in use                     */
  int32_t max_running;       /* most requests on the GPU in one step          */
  int32_t reserved;
  /* store copies deferred to a later tick because the pinned staging buffer
     (stage_bytes) was full: admissions, hand-backs from the slot batch and
     hand-backs from the side lanes                                           */
  int64_t waits_stage_admit;
  int64_t waits_stage_harvest;
  int64_t waits_stage_side;
} pcx_server_stats;

void       pcx_server_config_default(pcx_server_config* cfg);
pcx_status pcx_server_create(const pcx_server_config* cfg, pcx_server** out);
void       pcx_server_destroy(pcx_server* s);
/* Register an engine's table as network *net.  The engine must outlive the
 * server and tell nothing more while registered. */
pcx_status pcx_server_add_network(pcx_server* s, pcx_engine* e, int32_t* net);
/* Give network `net` its told store (lb/ub, nvars entries, host memory, read
 * during the call): the server propagates it once (the root fixpoint) and
 * starts each later request of the network whose store lies inside the told
 * store from the request's meet with the root fixpoint.  The result is the
 * same -- every fixpoint below such a store is below the root fixpoint -- and
 * a request then needs a few sweeps instead of its whole propagation.  A
 * request outside the told store starts from its own store.  Calling it
 * again replaces the root.  No counterpart in the reference (which has no
 * server); search-tree serving, where requests are subproblems of a root. */
pcx_status pcx_server_set_root(pcx_server* s, int32_t net, const int32_t* lb, const int32_t* ub,
                               int64_t nvars);
/* Queue a request.  lb/ub (nvars entries, host memory) stay the caller's
 * until the request completes; they then hold its fixpoint. */
pcx_status pcx_server_submit(pcx_server* s, int32_t net, int32_t* lb, int32_t* ub,
                             int64_t nvars, int64_t tag);
/* One tick: admit, sweep, complete.  *in_flight (may be NULL) receives the
 * requests still queued or running. */
pcx_status pcx_server_step(pcx_server* s, int32_t* in_flight);
/* Counters since creation. */
pcx_status pcx_server_get_stats(const pcx_server* s, pcx_server_stats* out);
/* Move up to max completed requests to out; returns how many. */
int64_t    pcx_server_poll(pcx_server* s, pcx_served* out, int64_t max);
/* Replay a request trace in real time on an idle server: request i arrives
 * arrival_s[i] seconds after the start (ascending), runs network net[i] on
 * store slice [store_off[i], store_off[i+1]) of lb_all/ub_all (overwritten
 * with the fixpoints), and out[i] receives its result and latency.
 * plain = 1 serves the same trace one request at a time (FIFO), each with
 * pcx_fixpoint on its network's engine, as the comparison point.
 * *wall_s = whole replay. */
pcx_status pcx_server_replay(pcx_server* s, int64_t n, const double* arrival_s,
                             const int32_t* net, int32_t* lb_all, int32_t* ub_all,
                             const int64_t* store_off, int32_t plain, pcx_served* out,
                             double* wall_s);

#ifdef __cplusplus
}
#endif
#endif /* PCX_H */
