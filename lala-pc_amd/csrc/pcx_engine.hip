// pcx_engine.hip -- host side of the C ABI (include/pcx.h).
//
// The engine owns the propagator table (host copy in engine order + device
// SoA copy) and launches the sweep kernels; stores live in device memory
// inside batches.  pcx_fixpoint on host arrays is a batch of one instance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "pcx_internal.h"
#include "../../include/pcx_model.h"

using namespace pcx;

namespace pcx_internal {

thread_local std::string g_err;

pcx_status fail_with(pcx_status st, const std::string& msg) {
  g_err = msg;
  return st;
}

int device_cus(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    return prop.multiProcessorCount;
  return 256;
}

}  // namespace pcx_internal

using pcx_internal::fail_with;
using pcx_internal::g_err;

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool valid_ppw(int p) {
  return p == 64 || p == 128 || p == 256 || p == 512 || p == 1024 || p == 2048;
}

bool valid_lockstep(int g) { return g == -1 || g == 0 || g == 16 || g == 32 || g == 64; }
bool valid_mark_queue(int q) { return q == -1 || q == 0 || (q >= 10 && q <= 26); }
bool valid_base_seed(int v) { return v == -1 || v == 0 || v == 1; }
// d_seedctl: the difference sub-lists' counts, then the unsatisfied count
constexpr int kSeedCtlWords = kDiffLists + 1;
constexpr int32_t kTrailCap = 1024;     // variables a drained instance lists for a restore
constexpr int32_t kSeedCap = 1024;      // seeds listed per instance for a drain

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace

namespace pcx_internal {

pcx_status sync_table(pcx_engine* e) {
  if (!e->dirty) return PCX_OK;
  const int64_t n = (int64_t)e->op.size();
  PCX_HIP(hipSetDevice(e->cfg.device));
  if (n > e->d_cap) {
    dfree(e->d_xo); dfree(e->d_y); dfree(e->d_z);
    e->d_cap = 0;                      // a failed allocation below leaves no stale capacity
    const int64_t cap = std::max<int64_t>(n, 1);
    PCX_HIP(hipMalloc(&e->d_xo, cap * sizeof(uint32_t)));
    PCX_HIP(hipMalloc(&e->d_y, cap * sizeof(int32_t)));
    PCX_HIP(hipMalloc(&e->d_z, cap * sizeof(int32_t)));
    e->d_cap = cap;
  }
  if (n > 0) {
    std::vector<uint32_t> xo(n);
    for (int64_t i = 0; i < n; ++i) xo[i] = (uint32_t)e->x[i] | ((uint32_t)e->op[i] << kOpShift);
    PCX_HIP(hipMemcpy(e->d_xo, xo.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    PCX_HIP(hipMemcpy(e->d_y, e->y.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    PCX_HIP(hipMemcpy(e->d_z, e->z.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  e->dirty = false;
  ++e->version;
  return PCX_OK;
}

int dom_of(const pcx_engine* e) {
  if (e->cfg.domain == PCX_DOMAIN_BITSET) return kBits;
  return e->has_div ? kItvDiv : kItv;
}

// Variable -> table positions adjacency of the engine's table (each
// propagator listed once per distinct variable), built on the host when the
// table changed and uploaded.
pcx_status ensure_adjacency(pcx_engine* e) {
  if (e->d_adj_off && e->adj_version == e->version) return PCX_OK;
  const int64_t n = (int64_t)e->op.size();
  const int64_t V = (int64_t)e->max_var + 1;
  std::vector<int32_t> off(V + 1, 0);
  auto each = [&](int64_t i, auto&& f) {
    const int32_t x = e->x[i], y = e->y[i], z = e->z[i];
    f(x);
    if (y != x) f(y);
    if (z != x && z != y) f(z);
  };
  for (int64_t i = 0; i < n; ++i) each(i, [&](int32_t v) { ++off[v + 1]; });
  for (int64_t v = 0; v < V; ++v) off[v + 1] += off[v];
  std::vector<int32_t> adj(std::max<int32_t>(off[V], 1)), pos(off.begin(), off.end() - 1);
  for (int64_t i = 0; i < n; ++i) each(i, [&](int32_t v) { adj[pos[v]++] = (int32_t)i; });
  dfree(e->d_adj_off); dfree(e->d_adj);
  PCX_HIP(hipMalloc(&e->d_adj_off, (V + 1) * sizeof(int32_t)));
  PCX_HIP(hipMalloc(&e->d_adj, adj.size() * sizeof(int32_t)));
  PCX_HIP(hipMemcpy(e->d_adj_off, off.data(), (V + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  PCX_HIP(hipMemcpy(e->d_adj, adj.data(), adj.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  e->adj_version = e->version;
  return PCX_OK;
}

}  // namespace pcx_internal

using pcx_internal::dom_of;
using pcx_internal::ensure_adjacency;
using pcx_internal::sync_table;

namespace {

Table table_of(const pcx_engine* e) {
  return Table{e->d_xo, e->d_y, e->d_z, (int64_t)e->op.size()};
}

// Propagators per wave.  0 = automatic: the largest of 512/256/128/64 that
// still gives at least two workgroups per CU for an m-propagator sweep (small
// tables need many small blocks to fill the chip, large ones amortize the
// per-block overhead); the batched mode walks a grid-stride item list and
// always uses 512 unless told otherwise.
int ppw_of(const pcx_engine* e, int64_t m = -1) {
  // plain sweeps unroll at most 512 per wave (batched items go to 2048)
  if (e->cfg.props_per_wave) return m < 0 ? e->cfg.props_per_wave : std::min(e->cfg.props_per_wave, 512);
  if (m < 0) return 512;
  int cus = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, e->cfg.device) == hipSuccess) cus = prop.multiProcessorCount;
  for (int ppw = 512; ppw > 64; ppw /= 2)
    if (m / ((int64_t)kBlock * (ppw / kWave)) >= 2LL * cus) return ppw;
  return 64;
}

int64_t inst_size(const pcx_batch* b, int64_t k) {
  return b->h_off.empty() ? b->table_n : b->h_off[k + 1] - b->h_off[k];
}

// Events for kernel timing, reused across runs.
hipEvent_t event_at(pcx_batch* b, size_t i) {
  while (b->ev.size() <= i) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    b->ev.push_back(ev);
  }
  return b->ev[i];
}

// HIP-event timing of sweep launches, split into full and marked sweeps.
struct SweepTimer {
  pcx_batch* b;
  bool on;
  size_t nev = 0;
  std::vector<char> full;             // class of each recorded launch
  double ms[2] = {0.0, 0.0};          // [0] marked, [1] full
  int64_t launches[2] = {0, 0};
  int64_t lockstep = 0;               // group size of the lockstep full sweeps
  pcx_status begin(hipStream_t st) {
    if (on) PCX_HIP(hipEventRecord(event_at(b, nev++), st));
    return PCX_OK;
  }
  pcx_status end(hipStream_t st, bool is_full, int n = 1) {
    launches[is_full] += n;
    if (!on) return PCX_OK;
    PCX_HIP(hipEventRecord(event_at(b, nev++), st));
    full.push_back(is_full);
    return PCX_OK;
  }
  // after the stream has been synchronized
  pcx_status drain() {
    for (size_t i = 0; i + 1 < nev; i += 2) {
      float t = 0.f;
      PCX_HIP(hipEventElapsedTime(&t, b->ev[i], b->ev[i + 1]));
      ms[full[i / 2] ? 1 : 0] += t;
    }
    nev = 0;
    full.clear();
    return PCX_OK;
  }
};

// Readbacks go through pinned host memory: a copy into pageable memory is
// staged by the runtime and cost ~70 us per call here.
constexpr size_t kPinScratch = 2048;     // small readbacks: at most the seeding counters
static_assert(kPinScratch >= (kDiffLists + 1) * sizeof(int32_t), "seeding counters fit the scratch");
size_t pinned_bytes(const pcx_batch* b) {
  return kPinScratch + (size_t)b->n_inst * sizeof(InstFlags) + 2 * kTightSlots * sizeof(unsigned long long);
}
pcx_status ensure_pinned(pcx_batch* b) {
  if (!b->h_pin) PCX_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_pin), pinned_bytes(b)));
  return PCX_OK;
}
// Copy `bytes` of device memory into the pinned scratch and wait for it.
pcx_status read_small(pcx_batch* b, const void* d, size_t bytes, void* out) {
  if (bytes > kPinScratch) return fail_with(PCX_EINVAL, "readback larger than the pinned scratch");
  PCX_HIP(hipMemcpyAsync(b->h_pin, d, bytes, hipMemcpyDeviceToHost, b->e->stream));
  PCX_HIP(hipStreamSynchronize(b->e->stream));
  std::memcpy(out, b->h_pin, bytes);
  return PCX_OK;
}

pcx_status collect(pcx_batch* b, std::vector<InstFlags>& flags, pcx_result* results,
                   pcx_stats* stats, SweepTimer& tm) {
  pcx_engine* e = b->e;
  flags.resize(b->n_inst);
  const size_t fbytes = b->n_inst * sizeof(InstFlags);
  unsigned long long tight[2 * kTightSlots];    // [0, kTightSlots): full sweeps
  PCX_HIP(hipMemcpyAsync(b->h_pin + kPinScratch, b->d_flags, fbytes, hipMemcpyDeviceToHost, e->stream));
  PCX_HIP(hipMemcpyAsync(b->h_pin + kPinScratch + fbytes, b->d_tight, sizeof(tight),
                         hipMemcpyDeviceToHost, e->stream));
  PCX_HIP(hipStreamSynchronize(e->stream));
  pcx_status ts = tm.drain();                    // launches still timed (the stream is idle now)
  if (ts != PCX_OK) return ts;
  std::memcpy(flags.data(), b->h_pin + kPinScratch, fbytes);
  std::memcpy(tight, b->h_pin + kPinScratch + fbytes, sizeof(tight));
  int64_t evals = 0, smax = 0, full_evals = 0;
  for (int64_t k = 0; k < b->n_inst; ++k) {
    const InstFlags& f = flags[k];
    const int32_t last = std::max(f.changed[0], f.changed[1]);
    pcx_result r{};
    r.failed = f.failed_at != kNoFail;
    r.has_changed = last >= 0;
    r.sweeps = r.failed ? (int64_t)f.failed_at + 1 : (int64_t)last + 2;
    if (e->cfg.max_sweeps > 0 && r.sweeps > e->cfg.max_sweeps) r.sweeps = e->cfg.max_sweeps;
    if (b->events) {     // one full sweep (unless empty at the start or seeded) + the marked ones
      const int64_t first = (f.failed_at == -1 || b->seeded_run) ? 0 : inst_size(b, k);
      r.evals = first + f.marked_evals;
      full_evals += first;
    } else {
      r.evals = r.sweeps * inst_size(b, k);
      full_evals += r.evals;
    }
    evals += r.evals;
    smax = std::max(smax, r.sweeps);
    if (results) results[k] = r;
  }
  if (stats) {
    stats->kernel_ms = tm.ms[0] + tm.ms[1];
    stats->kernel_launches = tm.launches[0] + tm.launches[1];
    stats->evals = evals;
    unsigned long long t = 0, tf = 0;
    for (int i = 0; i < 2 * kTightSlots; ++i) t += tight[i];
    for (int i = 0; i < kTightSlots; ++i) tf += tight[i];
    stats->tightened = (int64_t)t;
    stats->sweeps_max = smax;
    stats->full_ms = tm.ms[1];
    stats->full_launches = tm.launches[1];
    stats->full_evals = full_evals;
    stats->full_tightened = (int64_t)tf;
    stats->lockstep = tm.lockstep;
    stats->load_ms = b->run_from_image ? b->load_ms : 0.0;
  }
  b->load_ms = 0.0;
  return PCX_OK;
}

// Per-instance subset copies.  Reading a shared table through subset
// indices gathers 12-byte entries out of 128-byte lines and keeps the shared
// table competing with the instance stores for L2; a contiguous copy per
// instance is read once per sweep as a coalesced stream (non-temporal, so the
// stores stay cached) at the price of 12 B of HBM per active propagator.
// Automatic mode copies when that costs at most a quarter of free memory.
pcx_status prepare_tables(pcx_batch* b) {
  pcx_engine* e = b->e;
  bool want = false;
  const int64_t total = b->h_off.empty() ? 0 : b->h_off.back();
  if (!b->h_off.empty() && total > 0) {
    if (e->cfg.compact_tables > 0) {
      want = true;
    } else if (e->cfg.compact_tables < 0) {
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
        want = b->d_cxo != nullptr || (uint64_t)total * 12 <= free_b / 4;
    }
  }
  if (want && (!b->d_cxo || b->ctab_version != e->version)) {
    if (!b->d_cxo) {
      hipError_t he = hipMalloc(&b->d_cxo, total * sizeof(uint32_t));
      if (he == hipSuccess) he = hipMalloc(&b->d_cy, total * sizeof(int32_t));
      if (he == hipSuccess) he = hipMalloc(&b->d_cz, total * sizeof(int32_t));
      if (he != hipSuccess) {
        dfree(b->d_cxo); dfree(b->d_cy); dfree(b->d_cz);
        (void)hipGetLastError();
        if (e->cfg.compact_tables > 0)
          return fail_with(PCX_ENOMEM, "no device memory for per-instance tables");
        want = false;                  // automatic mode: read through the indices
      }
    }
    if (want) {
      PCX_HIP(launch_compact_table(table_of(e), b->d_idx, total, b->d_cxo, b->d_cy, b->d_cz, e->stream));
      b->ctab_version = e->version;
    }
  }
  b->compact = want;
  b->stream = e->cfg.stream_tables > 0 || (e->cfg.stream_tables < 0 && b->compact);
  return PCX_OK;
}

int64_t bitmap_words(const pcx_batch* b) { return (b->table_n + 31) / 32; }
constexpr int64_t kPlainMarkedMin = (int64_t)1 << 20;
int64_t summary_words(const pcx_batch* b) {
  const int64_t chunks = (bitmap_words(b) + kChunkWords - 1) / kChunkWords;
  return (chunks + 31) / 32;
}

// Member bitmaps of a subset batch (bit p of row k: instance k has
// propagator p), rebuilt when the table changed.
pcx_status ensure_members(pcx_batch* b) {
  pcx_engine* e = b->e;
  if (b->h_off.empty() || (b->d_member && b->member_version == e->version)) return PCX_OK;
  const int64_t words = bitmap_words(b);
  const int64_t rows = b->n_inst * words;
  if (!b->d_member) PCX_HIP(hipMalloc(&b->d_member, rows * sizeof(uint32_t)));
  PCX_HIP(hipMemsetAsync(b->d_member, 0, rows * sizeof(uint32_t), e->stream));
  PCX_HIP(launch_build_members(b->d_off, b->d_idx, b->n_inst, b->d_member, words, e->stream));
  b->member_version = e->version;
  return PCX_OK;
}

// Event-driven sweeps need the adjacency, the member bitmaps of subset
// batches and two dirty bitmaps per instance; automatic mode uses them when
// they take at most a quarter of free device memory.  The persistent mode
// always evaluates every propagator.
// Adjacency, member bitmaps (subset batches) and mark bitmaps of a batch.
pcx_status ensure_event_buffers(pcx_batch* b) {
  pcx_engine* e = b->e;
  const int64_t words = bitmap_words(b);
  const int64_t rows = b->n_inst * words;
  pcx_status st = ensure_adjacency(e);
  if (st != PCX_OK) return st;
  if (!b->d_dirty) PCX_HIP(hipMalloc(&b->d_dirty, 2 * rows * sizeof(uint32_t)));
  const int64_t srows = b->n_inst * summary_words(b);
  if (!b->d_summary) PCX_HIP(hipMalloc(&b->d_summary, 3 * srows * sizeof(uint32_t)));
  return ensure_members(b);
}

pcx_status clear_marks(pcx_batch* b) {
  pcx_engine* e = b->e;
  const int64_t rows = b->n_inst * bitmap_words(b);
  const int64_t srows = b->n_inst * summary_words(b);
  PCX_HIP(hipMemsetAsync(b->d_dirty, 0, 2 * rows * sizeof(uint32_t), e->stream));
  PCX_HIP(hipMemsetAsync(b->d_summary, 0, 2 * srows * sizeof(uint32_t), e->stream));
  b->marks_clean = true;
  return PCX_OK;
}

bool events_possible(const pcx_batch* b, int32_t mode) {
  const pcx_engine* e = b->e;
  return mode != PCX_MODE_PERSISTENT && e->cfg.event_driven != 0 && b->table_n > 0 &&
         3 * b->table_n < ((int64_t)1 << 31);             // adjacency indices are int32
}

pcx_status prepare_events(pcx_batch* b, int32_t mode) {
  pcx_engine* e = b->e;
  b->events = false;
  const bool seeded = b->seeded;
  b->seeded = false;                        // seeds serve one run
  b->seeded_run = false;
  if (!events_possible(b, mode)) return PCX_OK;
  if (seeded && b->d_dirty) {               // marks were seeded by pcx_batch_branch
    b->events = true;
    b->seeded_run = true;
    return PCX_OK;
  }
  // Automatic mode, one instance at a time: a marked sweep gives up the
  // propagation within a sweep that a full sweep gets (marks land at the end
  // of a work item), so small instances converge in fewer, cheap full
  // sweeps; marked sweeps pay from about 2^20 propagators per instance.
  if (e->cfg.event_driven < 0 && mode == PCX_MODE_PLAIN) {
    int64_t mmax = 0;
    for (int64_t k = 0; k < b->n_inst; ++k) mmax = std::max(mmax, inst_size(b, k));
    if (mmax < kPlainMarkedMin) return PCX_OK;
  }
  const int64_t rows = b->n_inst * bitmap_words(b);
  const uint64_t need = (uint64_t)rows * 4 * (b->h_off.empty() ? 2 : 3);
  if (e->cfg.event_driven < 0 && !b->d_dirty) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || need > free_b / 4) return PCX_OK;
  }
  pcx_status st = ensure_event_buffers(b);
  if (st != PCX_OK) return st;
  // A run that drained every instance's marks left the planes clear.
  if (!b->marks_clean) st = clear_marks(b);
  if (st != PCX_OK) return st;
  b->marks_clean = false;                   // until this run ends drained
  b->events = true;
  return PCX_OK;
}

EventSet event_set(const pcx_batch* b) {
  const int64_t srows = b->n_inst * summary_words(b);
  return EventSet{b->e->d_adj_off, b->e->d_adj, b->h_off.empty() ? nullptr : b->d_member,
                  b->d_dirty, b->d_summary, b->d_summary ? b->d_summary + 2 * srows : nullptr,
                  bitmap_words(b), summary_words(b), b->n_inst};
}

// Table and subsets a sweep kernel reads for this batch.
Table run_table(const pcx_batch* b) {
  if (b->compact) return Table{b->d_cxo, b->d_cy, b->d_cz, b->h_off.back()};
  return table_of(b->e);
}
Subsets run_subsets(const pcx_batch* b) {
  if (b->h_off.empty()) return Subsets{nullptr, nullptr};
  return Subsets{b->d_off, b->compact ? nullptr : b->d_idx};
}

pcx_status prepare_run(pcx_batch* b, int32_t mode) {
  pcx_engine* e = b->e;
  if ((int64_t)e->op.size() != b->table_n)
    return fail_with(PCX_EINVAL, "table changed since the batch was created");
  if (e->max_var >= b->nvars)
    return fail_with(PCX_EINVAL, "a propagator names a variable outside the store");
  pcx_status s = sync_table(e);
  if (s != PCX_OK) return s;
  PCX_HIP(hipMemsetAsync(b->d_tight, 0, 2 * kTightSlots * sizeof(unsigned long long), e->stream));
  const WorkList wl = worklist_of(b->d_act, b->d_item_base, b->d_ctl);
  // A store already empty is failed before any sweep; when the stores are the
  // image, its emptiness was found at load time.
  if (b->store_is_image) {
    PCX_HIP(launch_init_flags(b->d_flags, (int32_t)b->n_inst, wl, e->stream, b->d_image_empty));
  } else {
    PCX_HIP(launch_init_flags(b->d_flags, (int32_t)b->n_inst, wl, e->stream));
    PCX_HIP(launch_store_check_empty(b->d_store, b->nvars, (int32_t)b->n_inst, b->d_flags, dom_of(e),
                                     e->stream));
  }
  b->run_from_image = b->store_is_image;
  b->store_is_image = false;
  b->trail_valid = false;                   // until a drained run from the image lists its changes
  pcx_status st = ensure_pinned(b);
  if (st != PCX_OK) return st;
  st = prepare_tables(b);
  if (st != PCX_OK) return st;
  return prepare_events(b, mode);
}

// Queue planes of a batch (queued marked sweeps), allocated on first use:
// room for a sixteenth of the batch's (instance, propagator) pairs per plane,
// between 2^16 and 2^23 entries, spread over kQLists sub-lists; a plane that
// overflows (one sub-list of it) falls back to bitmaps.
pcx_status ensure_queue(pcx_batch* b) {
  const int mq = b->e->cfg.mark_queue;
  const int64_t pairs = b->n_inst * std::max<int64_t>(b->table_n, 1);
  // a sixteenth of the pairs, or one whole instance (a dense first sweep of
  // a single large instance marks every propagator), within [2^16, 2^23]
  const int64_t want = std::max<int64_t>(pairs / 16, std::min<int64_t>(pairs, b->table_n));
  const int64_t cap = mq > 0 ? (int64_t)1 << mq
                             : std::min<int64_t>(std::max<int64_t>(want, (int64_t)1 << 16), (int64_t)1 << 23);
  const int64_t per_list = (cap + kQLists - 1) / kQLists;
  if (b->d_q && b->qcap == per_list) return PCX_OK;
  dfree(b->d_q);
  dfree(b->d_qctl);
  b->qcap = per_list;
  PCX_HIP(hipMalloc(&b->d_q, 2 * kQLists * b->qcap * sizeof(unsigned long long)));
  PCX_HIP(hipMalloc(&b->d_qctl, kQCtlWords * sizeof(int32_t)));
  return PCX_OK;
}

// Plain mode: every instance on its own, one launch per sweep, host checks
// the instance's flags every `group` sweeps.
pcx_status run_plain(pcx_batch* b, pcx_result* results, pcx_stats* stats) {
  pcx_engine* e = b->e;
  const int group = e->cfg.sweeps_per_check > 0 ? e->cfg.sweeps_per_check : 8;
  const Table t = table_of(e);
  EventSet ev = event_set(b);
  ev.summary = nullptr;               // plain marked sweeps scan every chunk
  ev.snap = nullptr;
  SweepTimer tm{b, e->cfg.timing != 0};
  InstFlags hf;
  const int32_t maxs = (int32_t)std::min<int64_t>(e->cfg.max_sweeps, 1 << 30);
  // Queued marked sweeps (as in batched mode), when asked for with a queue
  // capacity (mark_queue > 0; automatic mode keeps the bitmap passes, which
  // advance a long chain several links per sweep where a queue pass
  // advances one): after an instance's full sweep its marks are listed in a
  // queue plane, and groups of (turn, queue pass) launches, captured once,
  // run until no mark is left; a plane that overflows hands the instance
  // back to the bitmap sweeps.
  const bool qmode = b->events && !b->seeded_run && e->cfg.mark_queue > 0;
  EventSet evq = ev;
  const int qgrid = pcx_internal::device_cus(e->cfg.device) * 8;   // queried outside any capture
  if (qmode) {
    pcx_status qs = ensure_queue(b);
    if (qs != PCX_OK) return qs;
    evq.q = b->d_q;
    evq.qctl = b->d_qctl;
    evq.qcap = b->qcap;
    const pcx_batch::MarkedGraphArgs args{e->version, group, dom_of(e), maxs, 0,
                                          e->d_xo, e->d_y, e->d_z, e->d_adj_off, e->d_adj, evq.q};
    if (!b->plain_queue_graph || !(b->plain_queue_graph_args == args)) {
      if (b->plain_queue_graph) (void)hipGraphExecDestroy(b->plain_queue_graph);
      b->plain_queue_graph = nullptr;
      hipGraph_t graph = nullptr;
      PCX_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      hipError_t ce = hipSuccess;
      for (int g = 0; g < group && ce == hipSuccess; ++g) {
        ce = launch_queue_turn(b->d_qctl, maxs, e->stream);
        if (ce == hipSuccess)
          ce = launch_sweep_queue(t, b->d_store, b->nvars, b->d_flags, b->d_tight + kTightSlots, evq,
                                  dom_of(e), qgrid, e->stream);
      }
      const hipError_t ee = hipStreamEndCapture(e->stream, &graph);
      if (ce != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        PCX_HIP(ce != hipSuccess ? ce : ee);
      }
      const hipError_t ie = hipGraphInstantiate(&b->plain_queue_graph, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      PCX_HIP(ie);
      b->plain_queue_graph_args = args;
    }
  }
  for (int64_t k = 0; k < b->n_inst; ++k) {
    const int64_t m = inst_size(b, k);
    const int ppw = ppw_of(e, m);
    Table tk = t;
    const int32_t* idx = nullptr;
    if (!b->h_off.empty()) {
      const int64_t beg = b->h_off[k];
      if (b->compact) tk = Table{b->d_cxo + beg, b->d_cy + beg, b->d_cz + beg, m};
      else idx = b->d_idx + beg;
    }
    int2* st = b->d_store + k * b->nvars;
    InstFlags* f = b->d_flags + k;
    int32_t s = 0;
    bool done = false;
    if (qmode && (maxs <= 0 || maxs > 1)) {
      pcx_status ts = tm.begin(e->stream);
      if (ts != PCX_OK) return ts;
      PCX_HIP(launch_sweep_plain(tk, st, idx, m, 0, f, b->d_tight, ppw, dom_of(e), b->stream, e->stream,
                                 &ev, k));
      ts = tm.end(e->stream, true);
      if (ts != PCX_OK) return ts;
      s = 1;
      PCX_HIP(hipMemsetAsync(b->d_qctl, 0, kQCtlWords * sizeof(int32_t), e->stream));
      PCX_HIP(launch_queue_from_bitmap(evq, b->d_flags, k, 1, b->table_n, e->stream));
      for (;;) {
        ts = tm.begin(e->stream);
        if (ts != PCX_OK) return ts;
        PCX_HIP(hipGraphLaunch(b->plain_queue_graph, e->stream));
        ts = tm.end(e->stream, false, group);
        if (ts != PCX_OK) return ts;
        int32_t q[kQHeadWords];
        ts = read_small(b, b->d_qctl, sizeof(q), q);
        if (ts != PCX_OK) return ts;
        ts = tm.drain();
        if (ts != PCX_OK) return ts;
        if (q[kQStop] == 1) { done = true; break; }
        if (q[kQStop] == 2) { s = q[kQAt]; break; }      // bitmap sweeps from sweep s on
      }
    }
    for (; !done;) {
      // Full sweeps are timed one by one; consecutive marked sweeps as one span.
      int n_marked = 0;
      for (int g = 0; g < group; ++g, ++s) {
        if (e->cfg.max_sweeps > 0 && s >= e->cfg.max_sweeps) break;
        const bool marked = b->events && (s > 0 || b->seeded_run);
        pcx_status ts = PCX_OK;
        if (!marked || n_marked == 0) ts = tm.begin(e->stream);
        if (ts != PCX_OK) return ts;
        if (marked) {
          PCX_HIP(launch_sweep_marked_plain(t, st, s, f, b->d_tight + kTightSlots, ev, k, dom_of(e),
                                            e->stream));
          ++n_marked;
        } else {
          PCX_HIP(launch_sweep_plain(tk, st, idx, m, s, f, b->d_tight, ppw, dom_of(e), b->stream,
                                     e->stream, b->events ? &ev : nullptr, k));
          ts = tm.end(e->stream, true);
          if (ts != PCX_OK) return ts;
        }
      }
      if (n_marked > 0) {
        pcx_status ts = tm.end(e->stream, false, n_marked);
        if (ts != PCX_OK) return ts;
      }
      pcx_status ts = read_small(b, f, sizeof(hf), &hf);
      if (ts != PCX_OK) return ts;
      ts = tm.drain();
      if (ts != PCX_OK) return ts;
      const int32_t last = std::max(hf.changed[0], hf.changed[1]);
      if (hf.failed_at != kNoFail || last + 1 < s) break;       // quiet sweep already ran
      if (e->cfg.max_sweeps > 0 && s >= e->cfg.max_sweeps) break;
    }
  }
  std::vector<InstFlags> flags;
  return collect(b, flags, results, stats, tm);
}

// Instances per lockstep group of this run's full sweeps, 0 for per-instance
// work items.  Automatic mode takes the largest group whose interleaved
// stores fit in kLockGroupBytes (the groups one XCD sweeps at a time stay
// in the memory-side cache; larger groups were slower) and that the batch
// fills four times over, when the interleaved copy takes at most a quarter
// of free device memory.
constexpr int64_t kLockGroupBytes = (int64_t)32 << 20;
int lockstep_of(const pcx_batch* b) {
  const pcx_engine* e = b->e;
  int g = e->cfg.lockstep;
  if (b->n_inst < 2 || b->table_n == 0) return 0;
  // The sweep addresses a group's copy with 32-bit offsets.
  if (g > 0 && (uint64_t)g * (uint64_t)b->nvars * 8u > 0xffffffffull) return 0;
  if (g < 0) {
    g = 0;
    for (int c : {64, 32, 16})
      if (c * b->nvars * 8 <= kLockGroupBytes && b->n_inst >= 4 * c) { g = c; break; }
    if (g == 0) return 0;
    if (!b->d_shadow) {
      size_t free_b = 0, total_b = 0;
      const uint64_t need = (uint64_t)((b->n_inst + g - 1) / g) * g * b->nvars * 8;
      if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || need > free_b / 4) return 0;
    }
  }
  return g;
}

// A reset image that lives only in the interleaved copy is written out to
// d_image before the copy is overwritten.
pcx_status materialize_image(pcx_batch* b) {
  if (!b->image_in_shadow) return PCX_OK;
  PCX_HIP(launch_deinterleave(b->d_shadow, b->d_image, b->nvars, b->n_inst, b->shadow_gi, b->e->stream));
  b->image_in_shadow = false;
  return PCX_OK;
}

pcx_status ensure_shadow(pcx_batch* b, int gi) {
  const int64_t words = (b->n_inst + gi - 1) / gi * gi * b->nvars;
  if (b->shadow_words >= words) return PCX_OK;
  dfree(b->d_shadow);
  b->shadow_words = 0;
  b->shadow_is_image = false;
  PCX_HIP(hipMalloc(&b->d_shadow, words * sizeof(int2)));
  b->shadow_words = words;
  return PCX_OK;
}

// Batched mode: per sweep one scheduler launch and one grid-stride sweep
// launch over the work items of every running instance (or, for full sweeps
// in lockstep, over (group, table block) items).  After a lockstep full
// sweep with marks, the marked sweeps run from queue planes: groups of
// (turn, queue pass) launches replayed as one graph until a turn finds no
// marks, or finds a plane that lost marks, from which sweep on the bitmap
// sweeps take over.
// A run that starts from the loaded stores, whose load listed the values that
// differ from the base store (the values instances 0-2 agree on), needs in
// its first sweep only (a) the member propagators adjacent to a listed value
// of their instance and (b) the propagators the base store does not satisfy,
// in every instance that holds them: any other propagator reads in instance
// k the values it reads in the base store, where it tightens nothing.  These are seeded as the first
// sweep's marks, which makes that sweep a marked sweep, when their count is
// under 1/16 of the pairs a full sweep evaluates (and the list is complete).
// This decides; the run launches the seeding (*n_unsat propagators listed
// in d_unsat) with its own event set.
pcx_status plan_seed(pcx_batch* b, bool* seeded, int64_t* n_diff_out, int64_t* n_unsat_out) {
  *seeded = false;
  *n_diff_out = 0;
  *n_unsat_out = 0;
  pcx_engine* e = b->e;
  if (!b->events || b->seeded_run || !b->run_from_image || !b->diffs_listed || e->cfg.base_seed == 0 ||
      b->table_n == 0)
    return PCX_OK;
  // The probe and the counts depend on the image and the table only: runs
  // from the same image (after resets) reuse them.
  if (!b->probe_valid || b->probe_version != e->version) {
    if (!b->d_unsat) PCX_HIP(hipMalloc(&b->d_unsat, b->table_n * sizeof(int32_t)));
    int32_t* unsat_n = b->d_seedctl + kDiffLists;
    PCX_HIP(hipMemsetAsync(unsat_n, 0, sizeof(int32_t), e->stream));
    PCX_HIP(launch_base_probe(table_of(e), b->d_base, b->d_unsat, unsat_n, dom_of(e), e->stream));
    int32_t ctl[kDiffLists + 1];
    pcx_status st = read_small(b, b->d_seedctl, sizeof(ctl), ctl);
    if (st != PCX_OK) return st;
    int64_t n = 0;
    for (int c = 0; c < kDiffLists; ++c)
      n = (ctl[c] > b->diff_cap || n < 0) ? -1 : n + ctl[c];   // -1: a sub-list lost entries
    b->probe_diff = n;
    b->probe_unsat = ctl[kDiffLists];
    b->probe_version = e->version;
    b->probe_valid = true;
  }
  if (b->probe_diff < 0) return PCX_OK;
  const int64_t n_diff = b->probe_diff, n_unsat = b->probe_unsat;
  const int64_t per_var = 3 * b->table_n / std::max<int64_t>(b->nvars, 1) + 1;   // mean adjacency
  const int64_t pairs = b->h_off.empty() ? b->n_inst * b->table_n : b->h_off.back();
  if ((n_diff * per_var + n_unsat * b->n_inst) * 16 > pairs) return PCX_OK;
  *seeded = true;
  *n_unsat_out = n_unsat;
  *n_diff_out = n_diff;
  return PCX_OK;
}

pcx_status run_batched(pcx_batch* b, pcx_result* results, pcx_stats* stats) {
  pcx_engine* e = b->e;
  bool base_seeded = false;
  int64_t n_diff = 0, n_unsat = 0;
  {
    pcx_status ss = plan_seed(b, &base_seeded, &n_diff, &n_unsat);
    if (ss != PCX_OK) return ss;
    if (base_seeded) b->seeded_run = true;
  }
  const int ppw = ppw_of(e);
  const int gi = lockstep_of(b);
  SweepTimer tm{b, e->cfg.timing != 0};
  tm.lockstep = gi;
  const int group = e->cfg.sweeps_per_check > 0 ? e->cfg.sweeps_per_check : 8;
  const Table t = run_table(b);
  const int grid = pcx_internal::device_cus(e->cfg.device) * 8;   // a multiple of the XCD count
  const Subsets sub = run_subsets(b);
  const WorkList wl = worklist_of(b->d_act, b->d_item_base, b->d_ctl);
  int32_t live = 1, sweeps = 0;
  const int32_t maxs = (int32_t)std::min<int64_t>(e->cfg.max_sweeps, 1 << 30);
  EventSet ev = event_set(b);
  // A seeded run (from a base store, or the split variables of a branching)
  // drains each instance's marks in one launch (one workgroup per instance)
  // when it runs to the fixpoint; with a sweep limit a base-seeded run sweeps
  // from queue planes, a branch-seeded one from the bitmaps.  Queued marked
  // sweeps also follow a lockstep full sweep.
  const bool drain = b->seeded_run && e->cfg.max_sweeps <= 0;
  const bool qmode = b->events && e->cfg.mark_queue != 0 &&
                     ((base_seeded && !drain) || (gi > 0 && !b->seeded_run));
  if (qmode) {
    pcx_status qs = ensure_queue(b);
    if (qs != PCX_OK) return qs;
    ev.q = b->d_q;
    ev.qctl = b->d_qctl;
    ev.qcap = b->qcap;
    PCX_HIP(hipMemsetAsync(b->d_qctl, 0, kQCtlWords * sizeof(int32_t), e->stream));
    // the first turn advances to sweep 0, whose plane the seeds fill
    if (base_seeded) PCX_HIP(hipMemsetAsync(b->d_qctl + kQSweep, 0xff, sizeof(int32_t), e->stream));
  }
  // A drain takes its seeds from per-instance lists (those past a list's
  // cap as plain marks) and keeps a trail; it reads neither the summaries
  // nor queues.
  InstLists seeds{nullptr, nullptr, 0};
  if (drain) {
    if (!b->d_trail) {
      PCX_HIP(hipMalloc(&b->d_trail, b->n_inst * (kTrailCap + kSeedCap) * sizeof(int32_t)));
      PCX_HIP(hipMalloc(&b->d_trail_n, 2 * b->n_inst * sizeof(int32_t)));
    }
    if (base_seeded) {                   // a branching's seeds are bits in the rows
      seeds = InstLists{b->d_trail + b->n_inst * kTrailCap, b->d_trail_n + b->n_inst, kSeedCap};
      PCX_HIP(hipMemsetAsync(seeds.n, 0, b->n_inst * sizeof(int32_t), e->stream));
    }
  }
  if (base_seeded) {
    EventSet sev = ev;
    if (drain) sev.summary = nullptr;
    PCX_HIP(launch_seed_from_base(sev, DiffSink{b->d_diff, b->d_seedctl, b->diff_cap, b->d_base, b->d_diff_val}, n_diff,
                                  b->d_unsat, n_unsat, seeds, e->stream));
  }
  if (drain) {
    PCX_HIP(hipMemsetAsync(b->d_ctl + 3, 0, sizeof(int32_t), e->stream));
    const int dgrid = (int)std::min<int64_t>(b->n_inst, (int64_t)pcx_internal::device_cus(e->cfg.device) * 24);
    pcx_status ts = tm.begin(e->stream);
    if (ts != PCX_OK) return ts;
    PCX_HIP(launch_drain_marks(table_of(e), b->d_store, b->nvars, b->d_flags, ev, b->d_ctl + 3,
                               b->d_tight + kTightSlots, seeds, InstLists{b->d_trail, b->d_trail_n, kTrailCap},
                               dom_of(e), dgrid, e->stream));
    ts = tm.end(e->stream, false);
    if (ts != PCX_OK) return ts;
    std::vector<InstFlags> flags;
    ts = collect(b, flags, results, stats, tm);        // one synchronization for the drain and the flags
    // the drain took and cleared every mark (rows of failed instances too;
    // a branching's seeds also set summary bits, which stay), and it started
    // from the image: the stores differ from it only in the trail's variables
    if (ts == PCX_OK) {
      if (base_seeded) b->marks_clean = true;
      b->trail_valid = b->run_from_image;
    }
    return ts;
  }
  auto launch_marked = [&]() -> pcx_status {
    PCX_HIP(launch_schedule(sub, b->table_n, b->d_flags, (int32_t)b->n_inst, wl, ppw, maxs, e->stream,
                            &ev));
    PCX_HIP(launch_sweep_marked(table_of(e), b->d_store, b->nvars, b->d_flags, wl,
                                b->d_tight + kTightSlots, ev, dom_of(e), grid, e->stream));
    return PCX_OK;
  };
  auto launch_queued = [&]() -> pcx_status {
    PCX_HIP(launch_queue_turn(b->d_qctl, maxs, e->stream));
    PCX_HIP(launch_sweep_queue(table_of(e), b->d_store, b->nvars, b->d_flags, b->d_tight + kTightSlots, ev,
                               dom_of(e), grid, e->stream));
    return PCX_OK;
  };
  // A full sweep (scheduled first): lockstep over the interleaved copy, or
  // per-instance work items.
  auto launch_full = [&]() -> pcx_status {
    PCX_HIP(launch_schedule(sub, b->table_n, b->d_flags, (int32_t)b->n_inst, wl, ppw, maxs, e->stream,
                            nullptr));
    pcx_status ts = tm.begin(e->stream);
    if (ts != PCX_OK) return ts;
    if (gi) {                            // snapshot of the stores, then the sweep reading it
      // A run that starts from the image reuses the image's interleaved
      // copy (the sweep only reads it) until the next load.
      const bool from_image = b->run_from_image && sweeps == 0;
      const bool reuse = from_image && b->shadow_is_image && b->shadow_gi == gi;
      if (!reuse) ts = materialize_image(b);
      if (ts == PCX_OK) ts = ensure_shadow(b, gi);
      if (ts == PCX_OK) ts = ensure_members(b);
      if (ts != PCX_OK) return ts;
      if (!reuse) {
        PCX_HIP(launch_interleave(b->d_store, b->d_shadow, b->nvars, b->n_inst, gi, e->stream));
        b->shadow_is_image = from_image;
        b->shadow_gi = gi;
      }
      PCX_HIP(launch_sweep_lockstep(table_of(e), b->d_store, b->d_shadow, b->nvars,
                                    (int32_t)b->n_inst, b->h_off.empty() ? nullptr : b->d_member,
                                    bitmap_words(b), b->d_flags, wl, b->d_tight, gi, dom_of(e),
                                    grid, e->stream, b->events ? &ev : nullptr));
    } else {
      PCX_HIP(launch_sweep_batched(t, b->d_store, b->nvars, sub, b->table_n, b->d_flags, wl,
                                   b->d_tight, ppw, dom_of(e), b->stream, grid, e->stream,
                                   b->events ? &ev : nullptr));
    }
    return tm.end(e->stream, true);
  };
  // Groups of marked sweeps have the same launch arguments every time (the
  // sweep number lives on the device): each kind is captured once into a
  // graph and replayed while every argument the capture used is unchanged
  // (the batch's own buffers never move).
  auto capture = [&](hipGraphExec_t& gx, pcx_batch::MarkedGraphArgs& have, auto&& body) -> pcx_status {
    const pcx_batch::MarkedGraphArgs args{e->version, group, dom_of(e), maxs, ppw,
                                          e->d_xo, e->d_y, e->d_z, e->d_adj_off, e->d_adj, ev.q};
    if (gx && have == args) return PCX_OK;
    if (gx) (void)hipGraphExecDestroy(gx);
    gx = nullptr;
    hipGraph_t graph = nullptr;
    PCX_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    pcx_status cs = PCX_OK;
    for (int g = 0; g < group && cs == PCX_OK; ++g) cs = body();
    const hipError_t ce = hipStreamEndCapture(e->stream, &graph);
    if (cs != PCX_OK) { if (graph) (void)hipGraphDestroy(graph); return cs; }
    PCX_HIP(ce);
    const hipError_t ie = hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    PCX_HIP(ie);
    have = args;
    return PCX_OK;
  };
  if (b->events) {
    pcx_status cs = capture(b->marked_graph, b->marked_graph_args, launch_marked);
    if (cs == PCX_OK && qmode) cs = capture(b->queue_graph, b->queue_graph_args, launch_queued);
    if (cs != PCX_OK) return cs;
  }
  if (qmode) {
    pcx_status ts = PCX_OK;
    if (!base_seeded) {
      ts = launch_full();
      if (ts != PCX_OK) return ts;
      sweeps = 1;
    }
    for (;;) {
      ts = tm.begin(e->stream);
      if (ts != PCX_OK) return ts;
      PCX_HIP(hipGraphLaunch(b->queue_graph, e->stream));
      ts = tm.end(e->stream, false, group);
      if (ts != PCX_OK) return ts;
      int32_t q[kQHeadWords];
      ts = read_small(b, b->d_qctl, sizeof(q), q);
      if (ts != PCX_OK) return ts;
      ts = tm.drain();
      if (ts != PCX_OK) return ts;
      if (q[kQStop] == 1) { live = 0; break; }
      if (q[kQStop] == 2) {              // bitmap sweeps from sweep q[kQAt] on
        sweeps = q[kQAt];
        const int32_t prev = sweeps - 1;
        PCX_HIP(hipMemcpyAsync(wl.sweep, &prev, sizeof(prev), hipMemcpyHostToDevice, e->stream));
        break;
      }
    }
  }
  while (live > 0) {
    if (b->events && sweeps >= group) {     // groups after the first: all marked sweeps
      pcx_status ts = tm.begin(e->stream);
      if (ts != PCX_OK) return ts;
      PCX_HIP(hipGraphLaunch(b->marked_graph, e->stream));
      ts = tm.end(e->stream, false, group);
      if (ts != PCX_OK) return ts;
      sweeps += group;
      ts = read_small(b, b->d_ctl + 2, sizeof(int32_t), &live);
      if (ts != PCX_OK) return ts;
      ts = tm.drain();
      if (ts != PCX_OK) return ts;
      continue;
    }
    // Full sweeps are timed one by one; consecutive marked sweeps (with their
    // scheduler launches) as one span.
    int n_marked = 0;
    for (int g = 0; g < group; ++g) {
      const bool marked = b->events && (sweeps > 0 || b->seeded_run);
      pcx_status ts = PCX_OK;
      if (marked) {
        if (n_marked == 0) ts = tm.begin(e->stream);
        if (ts == PCX_OK) ts = launch_marked();
        ++n_marked;
      } else {
        ts = launch_full();
      }
      if (ts != PCX_OK) return ts;
      ++sweeps;
    }
    if (n_marked > 0) {
      pcx_status ts = tm.end(e->stream, false, n_marked);
      if (ts != PCX_OK) return ts;
    }
    pcx_status ts = read_small(b, b->d_ctl + 2, sizeof(int32_t), &live);
    if (ts != PCX_OK) return ts;
    ts = tm.drain();
    if (ts != PCX_OK) return ts;
  }
  std::vector<InstFlags> flags;
  return collect(b, flags, results, stats, tm);
}

// Persistent mode: one launch; workgroups pull whole instances from a queue
// ordered largest first and run each to its fixpoint.
pcx_status run_persistent(pcx_batch* b, pcx_result* results, pcx_stats* stats) {
  pcx_engine* e = b->e;
  const Table t = run_table(b);
  const int dev_cus = pcx_internal::device_cus(e->cfg.device);
  // LDS-resident stores whenever the store fits in a workgroup's LDS (lds_mode
  // 1 or -1): even at 128 KB, one workgroup per CU, the LDS sweep was 2.3x
  // faster than the same sweep on global memory.
  const int64_t sbytes = b->nvars * 8;
  const int64_t lds_cap = lds_store_cap(e->cfg.device);
  const bool lds = e->cfg.lds_mode != 0 && sbytes <= lds_cap;
  const int per_cu = persistent_blocks_per_cu(dom_of(e), lds ? sbytes : 0);
  const int grid = (int)std::min<int64_t>((int64_t)dev_cus * per_cu, b->n_inst);
  const Subsets sub = run_subsets(b);
  const int32_t maxs = (int32_t)std::min<int64_t>(e->cfg.max_sweeps, 1 << 30);
  SweepTimer tm{b, e->cfg.timing != 0};
  pcx_status ts = tm.begin(e->stream);
  if (ts != PCX_OK) return ts;
  PCX_HIP(launch_persistent(t, b->d_store, b->nvars, sub, b->table_n, b->d_flags, b->d_order,
                            (int32_t)b->n_inst, b->d_ctl + 3, b->d_tight, maxs, dom_of(e), grid,
                            lds, lds_cap, e->stream));
  ts = tm.end(e->stream, true);
  if (ts != PCX_OK) return ts;
  PCX_HIP(hipStreamSynchronize(e->stream));
  ts = tm.drain();
  if (ts != PCX_OK) return ts;
  std::vector<InstFlags> flags;
  return collect(b, flags, results, stats, tm);
}

}  // namespace

// =========================== C ABI ========================================
extern "C" {

const char* pcx_version(void) { return "pcx 0.1 (gfx950)"; }
const char* pcx_last_error(void) { return g_err.c_str(); }

void pcx_config_default(pcx_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->device = 0;
  c->props_per_wave = 0;
  c->lds_mode = -1;
  c->sweeps_per_check = 8;
  c->max_sweeps = 0;
  c->timing = 0;
  c->compact_tables = -1;
  c->stream_tables = -1;
  c->event_driven = -1;
  c->lockstep = -1;
  c->mark_queue = -1;
  c->base_seed = -1;
}

int32_t pcx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

pcx_status pcx_create(const pcx_config* cfg, pcx_engine** out) {
  if (!out) return fail_with(PCX_EINVAL, "out is NULL");
  *out = nullptr;
  pcx_config c;
  if (cfg) c = *cfg; else pcx_config_default(&c);
  if (c.domain != PCX_DOMAIN_INTERVAL && c.domain != PCX_DOMAIN_BITSET)
    return fail_with(PCX_EINVAL, "unknown domain");
  if (c.props_per_wave != 0 && !valid_ppw(c.props_per_wave))
    return fail_with(PCX_EINVAL, "props_per_wave must be 0 (auto) or a power of two in [64, 2048]");
  if (!valid_lockstep(c.lockstep) || !valid_mark_queue(c.mark_queue) || !valid_base_seed(c.base_seed))
    return fail_with(PCX_EINVAL, "lockstep, mark_queue or base_seed out of range");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail_with(PCX_EDEVICE, "no HIP device");
  if (c.device < 0 || c.device >= n) return fail_with(PCX_EINVAL, "device ordinal out of range");
  PCX_HIP(hipSetDevice(c.device));
  auto* e = new (std::nothrow) pcx_engine();
  if (!e) return fail_with(PCX_ENOMEM, "host allocation failed");
  e->cfg = c;
  hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (he != hipSuccess) {
    delete e;
    return fail_with(PCX_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(he));
  }
  *out = e;
  return PCX_OK;
}

void pcx_destroy(pcx_engine* e) {
  if (!e) return;
  (void)hipSetDevice(e->cfg.device);
  if (e->scratch) pcx_batch_destroy(e->scratch);
  dfree(e->d_xo); dfree(e->d_y); dfree(e->d_z); dfree(e->d_adj_off); dfree(e->d_adj);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

pcx_status pcx_tell(pcx_engine* e, const int32_t* op, const int32_t* x, const int32_t* y,
                    const int32_t* z, int64_t n, int32_t* lb, int32_t* ub, int64_t nvars) {
  if (!e) return fail_with(PCX_EINVAL, "engine is NULL");
  if (e->cfg.domain != PCX_DOMAIN_INTERVAL) return fail_with(PCX_EINVAL, "pcx_tell needs an interval engine");
  if (n < 0 || (n > 0 && (!op || !x || !y || !z))) return fail_with(PCX_EINVAL, "bad table arrays");
  if ((lb == nullptr) != (ub == nullptr)) return fail_with(PCX_EINVAL, "lb and ub must both be given or both be NULL");
  const int64_t lim = nvars > 0 ? std::min<int64_t>(nvars, PCX_MAX_VARS) : PCX_MAX_VARS;
  for (int64_t i = 0; i < n; ++i) {
    if (op[i] < 0 || op[i] >= PCX_NUM_OPS)
      return fail_with(PCX_EINVAL, "unknown opcode at " + std::to_string(i));
    if (x[i] < 0 || x[i] >= lim || y[i] < 0 || y[i] >= lim || z[i] < 0 || z[i] >= lim)
      return fail_with(PCX_EINVAL, "variable index out of range at " + std::to_string(i));
  }
  if (lb && nvars <= 0) return fail_with(PCX_EINVAL, "nvars must be given with a store");
  // Reified outputs are Boolean: embed x in [0,1] (pir.hpp:333-335).
  if (lb) {
    for (int64_t i = 0; i < n; ++i)
      if (op[i] == PCX_EQ || op[i] == PCX_LEQ) {
        lb[x[i]] = std::max(lb[x[i]], 0);
        ub[x[i]] = std::min(ub[x[i]], 1);
      }
  }
  // Append, then keep the table stably sorted by (op, y, x, z)
  // (pir.hpp:343-347): sort the new block and merge it behind the old one.
  const int64_t old = (int64_t)e->op.size();
  std::vector<int64_t> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  auto key_less = [&](int64_t a, int64_t b) {
    if (op[a] != op[b]) return op[a] < op[b];
    if (y[a] != y[b]) return y[a] < y[b];
    if (x[a] != x[b]) return x[a] < x[b];
    return z[a] < z[b];
  };
  std::stable_sort(ord.begin(), ord.end(), key_less);
  struct Bc { int32_t op, x, y, z; };
  std::vector<Bc> all(old + n);
  for (int64_t i = 0; i < old; ++i) all[i] = Bc{e->op[i], e->x[i], e->y[i], e->z[i]};
  for (int64_t i = 0; i < n; ++i) {
    const int64_t j = ord[i];
    all[old + i] = Bc{op[j], x[j], y[j], z[j]};
  }
  std::inplace_merge(all.begin(), all.begin() + old, all.end(), [](const Bc& a, const Bc& b) {
    if (a.op != b.op) return a.op < b.op;
    if (a.y != b.y) return a.y < b.y;
    if (a.x != b.x) return a.x < b.x;
    return a.z < b.z;
  });
  e->op.resize(old + n); e->x.resize(old + n); e->y.resize(old + n); e->z.resize(old + n);
  for (int64_t i = 0; i < old + n; ++i) {
    e->op[i] = all[i].op; e->x[i] = all[i].x; e->y[i] = all[i].y; e->z[i] = all[i].z;
    e->max_var = std::max(e->max_var, std::max(all[i].x, std::max(all[i].y, all[i].z)));
    e->has_div |= all[i].op >= PCX_TDIV && all[i].op <= PCX_EDIV;
  }
  e->dirty = true;
  return PCX_OK;
}

pcx_status pcx_get_config(const pcx_engine* e, pcx_config* cfg) {
  if (!e || !cfg) return fail_with(PCX_EINVAL, "NULL argument");
  *cfg = e->cfg;
  return PCX_OK;
}

pcx_status pcx_set_config(pcx_engine* e, const pcx_config* cfg) {
  if (!e || !cfg) return fail_with(PCX_EINVAL, "NULL argument");
  if (cfg->device != e->cfg.device || cfg->domain != e->cfg.domain)
    return fail_with(PCX_EINVAL, "device and domain are fixed at creation");
  if (cfg->props_per_wave != 0 && !valid_ppw(cfg->props_per_wave))
    return fail_with(PCX_EINVAL, "props_per_wave must be 0 (auto) or a power of two in [64, 2048]");
  if (!valid_lockstep(cfg->lockstep))
    return fail_with(PCX_EINVAL, "lockstep must be -1 (auto), 0 (off), 16, 32 or 64");
  if (!valid_mark_queue(cfg->mark_queue))
    return fail_with(PCX_EINVAL, "mark_queue must be -1 (auto), 0 (off) or a log2 capacity in [10, 26]");
  if (!valid_base_seed(cfg->base_seed))
    return fail_with(PCX_EINVAL, "base_seed must be -1 (auto), 0 (off) or 1 (on)");
  e->cfg = *cfg;
  return PCX_OK;
}

int64_t pcx_num_deductions(const pcx_engine* e) { return e ? (int64_t)e->op.size() : 0; }

pcx_status pcx_get_table(const pcx_engine* e, int32_t* op, int32_t* x, int32_t* y, int32_t* z) {
  if (!e) return fail_with(PCX_EINVAL, "engine is NULL");
  const size_t n = e->op.size();
  if (op) std::copy(e->op.begin(), e->op.end(), op);
  if (x) std::copy(e->x.begin(), e->x.end(), x);
  if (y) std::copy(e->y.begin(), e->y.end(), y);
  if (z) std::copy(e->z.begin(), e->z.end(), z);
  (void)n;
  return PCX_OK;
}

int64_t pcx_snapshot(const pcx_engine* e) { return e ? (int64_t)e->op.size() : 0; }

pcx_status pcx_restore(pcx_engine* e, int64_t snap) {
  if (!e) return fail_with(PCX_EINVAL, "engine is NULL");
  if (snap < 0 || snap > (int64_t)e->op.size()) return fail_with(PCX_EINVAL, "bad snapshot");
  // Like the reference (pir.hpp:863-870), pop entries off the end of the
  // table; the table is kept sorted, so these are the last in table order.
  e->op.resize(snap); e->x.resize(snap); e->y.resize(snap); e->z.resize(snap);
  e->max_var = -1;
  e->has_div = false;
  for (int64_t i = 0; i < snap; ++i) {
    e->max_var = std::max(e->max_var, std::max(e->x[i], std::max(e->y[i], e->z[i])));
    e->has_div |= e->op[i] >= PCX_TDIV && e->op[i] <= PCX_EDIV;
  }
  e->dirty = true;
  return PCX_OK;
}

pcx_status pcx_batch_create(pcx_engine* e, int64_t n_inst, int64_t nvars, const int64_t* inst_off,
                            const int32_t* prop_idx, pcx_batch** out) {
  if (!e || !out) return fail_with(PCX_EINVAL, "NULL argument");
  *out = nullptr;
  if (n_inst <= 0 || n_inst > (1 << 30) || nvars <= 0 || nvars > PCX_MAX_VARS)
    return fail_with(PCX_EINVAL, "bad n_inst or nvars");
  const int64_t tn = (int64_t)e->op.size();
  if (inst_off) {
    if (inst_off[0] != 0) return fail_with(PCX_EINVAL, "inst_off[0] must be 0");
    for (int64_t k = 0; k < n_inst; ++k)
      if (inst_off[k + 1] < inst_off[k]) return fail_with(PCX_EINVAL, "inst_off must be non-decreasing");
    if (inst_off[n_inst] > 0 && !prop_idx) return fail_with(PCX_EINVAL, "prop_idx is NULL");
    for (int64_t i = 0; i < inst_off[n_inst]; ++i)
      if (prop_idx[i] < 0 || prop_idx[i] >= tn) return fail_with(PCX_EINVAL, "prop_idx out of range");
  }
  PCX_HIP(hipSetDevice(e->cfg.device));
  auto* b = new (std::nothrow) pcx_batch();
  if (!b) return fail_with(PCX_ENOMEM, "host allocation failed");
  b->e = e;
  b->device = e->cfg.device;
  b->n_inst = n_inst;
  b->nvars = nvars;
  b->table_n = tn;
  const size_t sbytes = (size_t)n_inst * (size_t)nvars * sizeof(int2);
  hipError_t he = hipSuccess;
  auto chk = [&](hipError_t r) { if (he == hipSuccess) he = r; };
  chk(hipMalloc(&b->d_store, sbytes));
  chk(hipMalloc(&b->d_image, sbytes));
  chk(hipMalloc(&b->d_flags, n_inst * sizeof(InstFlags)));
  chk(hipMalloc(&b->d_image_empty, n_inst));
  chk(hipMalloc(&b->d_tight, 2 * kTightSlots * sizeof(unsigned long long)));
  chk(hipMalloc(&b->d_act, n_inst * sizeof(int32_t)));
  chk(hipMalloc(&b->d_item_base, (n_inst + 1) * sizeof(int64_t)));
  chk(hipMalloc(&b->d_ctl, kCtlWords * sizeof(int32_t)));
  if (inst_off && he == hipSuccess) {
    b->h_off.assign(inst_off, inst_off + n_inst + 1);
    chk(hipMalloc(&b->d_off, (n_inst + 1) * sizeof(int64_t)));
    chk(hipMalloc(&b->d_idx, std::max<int64_t>(inst_off[n_inst], 1) * sizeof(int32_t)));
    if (he == hipSuccess) {
      chk(hipMemcpy(b->d_off, inst_off, (n_inst + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
      if (inst_off[n_inst] > 0)
        chk(hipMemcpy(b->d_idx, prop_idx, inst_off[n_inst] * sizeof(int32_t), hipMemcpyHostToDevice));
    }
  }
  if (he == hipSuccess) {
    // Persistent queue order: largest instances first (longest-processing-
    // time-first keeps the tail short when sizes are uneven).
    std::vector<int32_t> order(n_inst);
    std::iota(order.begin(), order.end(), 0);
    if (inst_off)
      std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) {
        return inst_off[a + 1] - inst_off[a] > inst_off[c + 1] - inst_off[c];
      });
    chk(hipMalloc(&b->d_order, n_inst * sizeof(int32_t)));
    if (he == hipSuccess)
      chk(hipMemcpy(b->d_order, order.data(), n_inst * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  if (he != hipSuccess) {
    pcx_batch_destroy(b);
    return fail_with(he == hipErrorOutOfMemory ? PCX_ENOMEM : PCX_EDEVICE,
                     std::string("batch allocation: ") + hipGetErrorString(he));
  }
  *out = b;
  return PCX_OK;
}

void pcx_batch_destroy(pcx_batch* b) {
  if (!b) return;
  // the engine may be gone already (a garbage collector frees both in any
  // order): only the batch's own device number is read
  (void)hipSetDevice(b->device);
  for (hipEvent_t ev : b->ev) (void)hipEventDestroy(ev);
  dfree(b->d_store); dfree(b->d_image); dfree(b->d_off); dfree(b->d_idx); dfree(b->d_flags);
  dfree(b->d_tight); dfree(b->d_act); dfree(b->d_item_base); dfree(b->d_ctl); dfree(b->d_order);
  dfree(b->d_cxo); dfree(b->d_cy); dfree(b->d_cz); dfree(b->d_member); dfree(b->d_dirty);
  dfree(b->d_summary); dfree(b->d_image_empty); dfree(b->d_unpack); dfree(b->d_shadow);
  if (b->h_pin) (void)hipHostFree(b->h_pin);
  if (b->marked_graph) (void)hipGraphExecDestroy(b->marked_graph);
  if (b->queue_graph) (void)hipGraphExecDestroy(b->queue_graph);
  if (b->plain_queue_graph) (void)hipGraphExecDestroy(b->plain_queue_graph);
  dfree(b->d_q); dfree(b->d_qctl);
  dfree(b->d_diff); dfree(b->d_diff_val); dfree(b->d_unsat); dfree(b->d_trail); dfree(b->d_trail_n); dfree(b->d_bvar); dfree(b->d_seedctl); dfree(b->d_base); dfree(b->d_base_alt);
  for (hipEvent_t ev : b->load_ev)
    if (ev) (void)hipEventDestroy(ev);
  delete b;
}

namespace {
// Room for the differences a load lists: a seeded first sweep pays only
// while they are a small share of the values (see seed_from_base).
pcx_status ensure_diff_list(pcx_batch* b) {
  const int64_t want = std::min<int64_t>(std::max<int64_t>(b->n_inst * b->nvars / 128, 4096), 1 << 24);
  // Twice the even share, and at least the 2,048 values one listing
  // workgroup covers: workgroups fill the sub-lists unevenly when a load has
  // few of them, and an instance's differences may sit in one block.
  const int64_t per_list = std::max<int64_t>(2 * ((want + kDiffLists - 1) / kDiffLists), 2048);
  if (!b->d_seedctl) PCX_HIP(hipMalloc(&b->d_seedctl, kSeedCtlWords * sizeof(int32_t)));
  if (!b->d_base) PCX_HIP(hipMalloc(&b->d_base, b->nvars * sizeof(int2)));
  if (b->diff_cap < per_list) {
    dfree(b->d_diff);
    dfree(b->d_diff_val);
    b->diff_cap = 0;
    PCX_HIP(hipMalloc(&b->d_diff, per_list * kDiffLists * sizeof(unsigned long long)));
    PCX_HIP(hipMalloc(&b->d_diff_val, per_list * kDiffLists * sizeof(int2)));
    b->diff_cap = per_list;
  }
  return PCX_OK;
}

// Where this load lists the differences from the base store (batches of
// several instances with seeding on), after clearing the counts; an empty
// sink otherwise.
pcx_status diff_sink(pcx_batch* b, DiffSink* ds) {
  *ds = DiffSink{nullptr, nullptr, 0, nullptr, nullptr};
  b->diffs_listed = false;
  b->probe_valid = false;
  const pcx_engine* e = b->e;
  if (e->cfg.base_seed == 0 || b->n_inst < 2 || b->nvars <= 0)
    return PCX_OK;
  pcx_status st = ensure_diff_list(b);
  if (st != PCX_OK) return st;
  PCX_HIP(hipMemsetAsync(b->d_seedctl, 0, kSeedCtlWords * sizeof(int32_t), e->stream));
  *ds = DiffSink{b->d_diff, b->d_seedctl, b->diff_cap, b->d_base, b->d_diff_val};
  return PCX_OK;
}

// After a load: which instances' images are empty, and the stores are the image.
pcx_status image_loaded(pcx_batch* b, bool empty_known = false) {
  pcx_engine* e = b->e;
  b->diffs_listed = false;
  if (!empty_known) {
    PCX_HIP(hipMemsetAsync(b->d_image_empty, 0, b->n_inst, e->stream));
    PCX_HIP(launch_store_empty_map(b->d_image, b->nvars, (int32_t)b->n_inst, b->d_image_empty, dom_of(e),
                                   e->stream));
  }
  PCX_HIP(hipStreamSynchronize(e->stream));
  b->store_is_image = true;
  b->shadow_is_image = false;
  b->image_in_shadow = false;
  b->image_in_diffs = false;
  b->trail_valid = false;
  b->probe_valid = false;
  return PCX_OK;
}

// After a load listed its differences: the probe of its base store, then
// one read of the counts (differences per sub-list, unsatisfied
// propagators; the stream is synchronized).  *complete: no sub-list lost
// entries.  The probe and the counts are kept for the runs from this image.
pcx_status probe_loaded(pcx_batch* b, bool* complete) {
  pcx_engine* e = b->e;
  pcx_status st = sync_table(e);
  if (st == PCX_OK) st = ensure_pinned(b);
  if (st != PCX_OK) return st;
  const bool probe = b->table_n > 0 && (int64_t)e->op.size() == b->table_n;
  if (probe) {
    if (!b->d_unsat) PCX_HIP(hipMalloc(&b->d_unsat, b->table_n * sizeof(int32_t)));
    int32_t* unsat_n = b->d_seedctl + kDiffLists;      // the probe appends from zero
    PCX_HIP(hipMemsetAsync(unsat_n, 0, sizeof(int32_t), e->stream));
    PCX_HIP(launch_base_probe(table_of(e), b->d_base, b->d_unsat, unsat_n, dom_of(e), e->stream));
  }
  int32_t ctl[kDiffLists + 1];
  st = read_small(b, b->d_seedctl, sizeof(ctl), ctl);
  if (st != PCX_OK) return st;
  int64_t n = 0;
  for (int c = 0; c < kDiffLists; ++c) n = (ctl[c] > b->diff_cap || n < 0) ? -1 : n + ctl[c];
  *complete = n >= 0;
  if (probe) {
    b->probe_diff = n;                                 // -1: a sub-list lost entries
    b->probe_unsat = ctl[kDiffLists];
    b->probe_version = e->version;
    b->probe_valid = true;
  }
  return PCX_OK;
}

// After a load from host memory (the stores and d_image hold the image):
// list the differences from the base store.  A complete list is then also
// the reset image, which a reset writes without reading d_image.
pcx_status list_load_diffs(pcx_batch* b) {
  DiffSink ds;
  pcx_status st = diff_sink(b, &ds);
  if (st != PCX_OK || !ds.list) return st;
  PCX_HIP(launch_list_diffs(b->d_store, b->nvars, b->n_inst, ds, b->e->stream));
  bool complete = false;
  st = probe_loaded(b, &complete);
  if (st != PCX_OK) return st;
  b->diffs_listed = true;
  b->image_in_diffs = complete;
  return PCX_OK;
}
}  // namespace

pcx_status pcx_batch_load(pcx_batch* b, const int32_t* lb, const int32_t* ub) {
  if (!b || !lb || !ub) return fail_with(PCX_EINVAL, "NULL argument");
  if (b->e->cfg.domain != PCX_DOMAIN_INTERVAL) return fail_with(PCX_EINVAL, "interval batch expected");
  pcx_engine* e = b->e;
  PCX_HIP(hipSetDevice(e->cfg.device));
  const int64_t n = b->n_inst * b->nvars;
  // Stage lb / ub through the store buffer's halves, then interleave.
  int32_t* d_lb = reinterpret_cast<int32_t*>(b->d_store);
  int32_t* d_ub = d_lb + n;
  PCX_HIP(hipMemcpyAsync(d_lb, lb, n * sizeof(int32_t), hipMemcpyHostToDevice, e->stream));
  PCX_HIP(hipMemcpyAsync(d_ub, ub, n * sizeof(int32_t), hipMemcpyHostToDevice, e->stream));
  PCX_HIP(launch_store_pack(d_lb, d_ub, b->d_image, n, e->stream));
  PCX_HIP(hipMemcpyAsync(b->d_store, b->d_image, n * sizeof(int2), hipMemcpyDeviceToDevice, e->stream));
  pcx_status st = image_loaded(b);
  if (st != PCX_OK) return st;
  return list_load_diffs(b);
}

pcx_status pcx_batch_load_device(pcx_batch* b, const int32_t* d_lb, const int32_t* d_ub) {
  if (!b || !d_lb || !d_ub) return fail_with(PCX_EINVAL, "NULL argument");
  if (b->e->cfg.domain != PCX_DOMAIN_INTERVAL) return fail_with(PCX_EINVAL, "interval batch expected");
  pcx_engine* e = b->e;
  PCX_HIP(hipSetDevice(e->cfg.device));
  const int64_t n = b->n_inst * b->nvars;
  // Stores that hold their image as a base store plus a complete difference
  // list (or that image but for the trail of a drained run from it) go back
  // to the base store first: the ingest then writes only the values that
  // differ from it (loads of nodes of one search tree share most values).
  // The new base store goes to the other base buffer.
  const bool held = b->shared_loads && b->image_in_diffs && (b->store_is_image || b->trail_valid) &&
                    e->cfg.base_seed != 0 && b->n_inst >= 2 && b->nvars > 0;
  const bool from_image = b->store_is_image;
  b->store_is_image = false;               // until the load below completes
  b->trail_valid = false;
  b->image_in_diffs = false;
  if (held) {
    if (!b->d_base_alt) PCX_HIP(hipMalloc(&b->d_base_alt, b->nvars * sizeof(int2)));
    const InstLists trail{b->d_trail, b->d_trail_n, kTrailCap};
    PCX_HIP(launch_stores_to_base(from_image ? nullptr : &trail,
                                  DiffSink{b->d_diff, b->d_seedctl, b->diff_cap, b->d_base, b->d_diff_val},
                                  b->d_store, b->nvars, b->n_inst, e->stream));
    std::swap(b->d_base, b->d_base_alt);   // d_base_alt: what every row now holds
  }
  PCX_HIP(hipMemsetAsync(b->d_image_empty, 0, b->n_inst, e->stream));
  DiffSink ds;
  pcx_status dst = diff_sink(b, &ds);
  if (dst != PCX_OK) return dst;
  const int gi = lockstep_of(b);
  const bool rows = ds.list && b->shared_loads;
  if (rows) PCX_HIP(launch_make_base_rows(d_lb, d_ub, b->nvars, b->n_inst, ds, e->stream));
  // config.timing: HIP events around the load's pass over the stores
  hipEvent_t t0 = nullptr, t1 = nullptr;
  bool t1_recorded = false;
  if (e->cfg.timing) {
    for (hipEvent_t& ev : b->load_ev)
      if (!ev) PCX_HIP(hipEventCreate(&ev));
    t0 = b->load_ev[0];
    t1 = b->load_ev[1];
    PCX_HIP(hipEventRecord(t0, e->stream));
  }
  if (rows) {
    // Stores that share most values (the last load's did): write the stores
    // only.  A complete difference list (base store + listed values) is then
    // the image; otherwise the image is copied from the stores now, as the
    // interleaved copy a lockstep full sweep reads when the batch has one,
    // and later loads write it in the ingest pass again.
    PCX_HIP(launch_ingest_rows(d_lb, d_ub, b->d_store, b->nvars, b->n_inst, b->d_image_empty, ds,
                               held ? b->d_base_alt : nullptr, e->stream));
    if (t1) PCX_HIP(hipEventRecord(t1, e->stream));
    t1_recorded = true;                 // (a copy of the image below records it again)
    bool complete = false;
    pcx_status st = probe_loaded(b, &complete);
    if (st != PCX_OK) return st;
    b->shadow_is_image = false;
    b->image_in_shadow = false;
    if (complete) {
      b->image_in_diffs = true;
    } else {
      b->shared_loads = false;
      if (gi > 0) {
        st = ensure_shadow(b, gi);
        if (st != PCX_OK) return st;
        PCX_HIP(launch_interleave(b->d_store, b->d_shadow, b->nvars, b->n_inst, gi, e->stream));
        b->shadow_is_image = true;
        b->shadow_gi = gi;
        b->image_in_shadow = true;
      } else {
        PCX_HIP(launch_copy_words(b->d_store, b->d_image, n, e->stream));
      }
      if (t1) PCX_HIP(hipEventRecord(t1, e->stream));
    }
  } else if (gi > 0) {
    // Batched runs sweep a lockstep copy first: write it in the same pass
    // as the stores.  That copy is then also the reset image (d_image is
    // written from it only if a run is about to overwrite it).
    b->image_in_shadow = false;
    pcx_status st = ensure_shadow(b, gi);
    if (st != PCX_OK) return st;
    PCX_HIP(launch_ingest_interleave(d_lb, d_ub, b->d_store, b->d_shadow, b->nvars, b->n_inst, gi,
                                     b->d_image_empty, ds, e->stream));
    b->shadow_is_image = true;
    b->shadow_gi = gi;
    b->image_in_shadow = true;
  } else {
    PCX_HIP(launch_ingest(d_lb, d_ub, b->d_image, b->d_store, n, b->nvars, b->d_image_empty, e->stream));
    PCX_HIP(launch_list_diffs(b->d_store, b->nvars, b->n_inst, ds, e->stream));
    b->shadow_is_image = false;
    b->image_in_shadow = false;
  }
  if (t1 && !t1_recorded) PCX_HIP(hipEventRecord(t1, e->stream));
  PCX_HIP(hipStreamSynchronize(e->stream));
  if (t1) {
    float ms = 0.f;
    PCX_HIP(hipEventElapsedTime(&ms, t0, t1));
    b->load_ms = ms;
  }
  if (ds.list && !rows) {
    // a load of the other kind whose stores share most values: the next load
    // writes the stores only (a rows load has probed its list already)
    bool complete = false;
    pcx_status st = probe_loaded(b, &complete);
    if (st != PCX_OK) return st;
    b->shared_loads = complete;
  }
  b->store_is_image = true;
  b->trail_valid = false;                // the trail described the last image
  b->diffs_listed = ds.list != nullptr;
  return PCX_OK;
}

pcx_status pcx_batch_reset(pcx_batch* b) {
  if (!b) return fail_with(PCX_EINVAL, "NULL argument");
  pcx_engine* e = b->e;
  if (b->store_is_image) return PCX_OK;
  const DiffSink ds{b->d_diff, b->d_seedctl, b->diff_cap, b->d_base, b->d_diff_val};
  if (b->trail_valid && (b->image_in_diffs || !b->image_in_shadow))
    PCX_HIP(launch_restore_trail(InstLists{b->d_trail, b->d_trail_n, kTrailCap}, ds,
                                 b->image_in_diffs ? nullptr : b->d_image, b->d_store, b->nvars, b->n_inst,
                                 e->stream));
  else if (b->image_in_diffs)
    PCX_HIP(launch_image_from_diffs(DiffSink{b->d_diff, b->d_seedctl, b->diff_cap, b->d_base, b->d_diff_val},
                                    b->d_store, b->nvars, b->n_inst, e->stream));
  else if (b->image_in_shadow)
    PCX_HIP(launch_deinterleave(b->d_shadow, b->d_store, b->nvars, b->n_inst, b->shadow_gi, e->stream));
  else
    PCX_HIP(launch_copy_words(b->d_image, b->d_store, b->n_inst * b->nvars, e->stream));
  b->store_is_image = true;
  return PCX_OK;
}

pcx_status pcx_batch_read(pcx_batch* b, int32_t* lb, int32_t* ub) {
  if (!b || !lb || !ub) return fail_with(PCX_EINVAL, "NULL argument");
  if (b->e->cfg.domain != PCX_DOMAIN_INTERVAL) return fail_with(PCX_EINVAL, "interval batch expected");
  pcx_engine* e = b->e;
  PCX_HIP(hipSetDevice(e->cfg.device));
  const int64_t n = b->n_inst * b->nvars;
  // Unpack into a staging buffer kept with the batch (lb then ub).
  if (!b->d_unpack) PCX_HIP(hipMalloc(&b->d_unpack, 2 * n * sizeof(int32_t)));
  int32_t* d_lb = b->d_unpack;
  PCX_HIP(launch_store_unpack(b->d_store, d_lb, d_lb + n, n, e->stream));
  PCX_HIP(hipMemcpyAsync(lb, d_lb, n * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream));
  PCX_HIP(hipMemcpyAsync(ub, d_lb + n, n * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream));
  PCX_HIP(hipStreamSynchronize(e->stream));
  return PCX_OK;
}

pcx_status pcx_batch_run(pcx_batch* b, int32_t mode, pcx_result* results, pcx_stats* stats) {
  if (!b) return fail_with(PCX_EINVAL, "NULL batch");
  pcx_engine* e = b->e;
  PCX_HIP(hipSetDevice(e->cfg.device));
  const double t0 = now_ms();
  pcx_status s = prepare_run(b, mode);
  if (s != PCX_OK) return s;
  switch (mode) {
    case PCX_MODE_PLAIN: s = run_plain(b, results, stats); break;
    case PCX_MODE_BATCHED: s = run_batched(b, results, stats); break;
    case PCX_MODE_PERSISTENT: s = run_persistent(b, results, stats); break;
    default: return fail_with(PCX_EINVAL, "unknown mode");
  }
  if (stats) stats->wall_ms = now_ms() - t0;
  return s;
}

pcx_status pcx_batch_branch(pcx_batch* parent, pcx_batch* child, int32_t* var, int32_t from_fixpoint) {
  if (!parent || !child) return fail_with(PCX_EINVAL, "NULL batch");
  if (parent->e != child->e) return fail_with(PCX_EINVAL, "parent and child batches must share the engine");
  if (child->n_inst != 2 * parent->n_inst || child->nvars != parent->nvars)
    return fail_with(PCX_EINVAL, "child batch must hold 2 * parent instances of the same nvars");
  if (parent == child) return fail_with(PCX_EINVAL, "parent and child must differ");
  pcx_engine* e = parent->e;
  PCX_HIP(hipSetDevice(e->cfg.device));
  if (!parent->d_bvar)
    PCX_HIP(hipMalloc(&parent->d_bvar, std::max<int64_t>(parent->n_inst, 1) * sizeof(int32_t)));
  int32_t* d_var = parent->d_bvar;
  hipError_t he = launch_branch(parent->d_store, parent->nvars, parent->n_inst, d_var, child->d_store,
                                child->d_image, child->d_image_empty, dom_of(e), e->stream);
  child->seeded = false;
  if (he == hipSuccess && from_fixpoint && events_possible(child, PCX_MODE_BATCHED)) {
    pcx_status st = sync_table(e);
    if (st == PCX_OK) st = ensure_event_buffers(child);
    if (st == PCX_OK) st = clear_marks(child);
    if (st != PCX_OK) return st;
    he = launch_branch_seed(event_set(child), d_var, child->n_inst, e->stream);
    child->marks_clean = false;
    child->seeded = he == hipSuccess;
  }
  if (he == hipSuccess && var)
    he = hipMemcpyAsync(var, d_var, parent->n_inst * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
  if (he != hipSuccess) return fail_with(PCX_EDEVICE, std::string("branch: ") + hipGetErrorString(he));
  const bool seeded = child->seeded;
  pcx_status st = image_loaded(child, true);        // the branch wrote the children's emptiness
  child->seeded = seeded;
  return st;
}

namespace {
// A formula tree list is well formed: known kinds, variables inside the
// store, ternary opcodes, a stack within kTreeMaxDepth that ends each formula
// with one value.
constexpr int kTreeMaxDepth = 32;
bool trees_valid(const int32_t* nodes, int64_t n_nodes, const int64_t* off, int64_t n_formulas,
                 int64_t nvars, std::string* why) {
  if (off[0] != 0 || off[n_formulas] != n_nodes) { *why = "offsets do not cover the nodes"; return false; }
  for (int64_t f = 0; f < n_formulas; ++f) {
    if (off[f + 1] < off[f]) { *why = "offsets decrease"; return false; }
    int depth = 0;
    for (int64_t j = off[f]; j < off[f + 1]; ++j) {
      const int32_t* nd = nodes + 4 * j;
      int pop = 0, push = 1;
      switch (nd[0]) {
        case PCX_FT_VAR:
          if (nd[1] < 0 || nd[1] >= nvars) { *why = "variable outside the store"; return false; }
          break;
        case PCX_FT_CONST: case PCX_FT_TRUE: case PCX_FT_FALSE: break;
        case PCX_FT_NEG: case PCX_FT_ABS: case PCX_FT_NOT: case PCX_FT_LIT: pop = 1; break;
        case PCX_FT_ADD: case PCX_FT_SUB: case PCX_FT_MUL: case PCX_FT_TDIV: case PCX_FT_FDIV:
        case PCX_FT_CDIV: case PCX_FT_EDIV: case PCX_FT_MIN: case PCX_FT_MAX: case PCX_FT_LE: case PCX_FT_LT:
        case PCX_FT_EQ: case PCX_FT_NE: case PCX_FT_EQUIV: case PCX_FT_XOR: pop = 2; break;
        case PCX_FT_AND: case PCX_FT_OR:
          if (nd[1] < 0) { *why = "negative operand count"; return false; }
          pop = nd[1];
          break;
        case PCX_FT_PROP: {
          const int32_t* zn = nd + 4;
          if (j + 1 >= off[f + 1] || zn[0] != PCX_FT_PROP_Z) { *why = "PROP without its PROP_Z node"; return false; }
          if (nd[1] < 0 || nd[1] >= PCX_NUM_OPS) { *why = "unknown opcode in PROP"; return false; }
          for (int32_t v : {nd[2], nd[3], zn[1]})
            if (v < 0 || v >= nvars) { *why = "variable outside the store"; return false; }
          ++j;
          break;
        }
        default: *why = "unknown node kind " + std::to_string(nd[0]); return false;
      }
      if (depth < pop) { *why = "node without its operands"; return false; }
      depth += push - pop;
      if (depth > kTreeMaxDepth) { *why = "tree deeper than 32"; return false; }
    }
    if (depth != 1) { *why = "formula leaves " + std::to_string(depth) + " values"; return false; }
  }
  return true;
}
}  // namespace

pcx_status pcx_batch_ask_formulas(pcx_batch* b, const int32_t* nodes, int64_t n_nodes, const int64_t* off,
                                  int64_t n_formulas, int32_t* extractable) {
  if (!b || !off || !extractable || (n_nodes > 0 && !nodes)) return fail_with(PCX_EINVAL, "NULL argument");
  if (n_nodes < 0 || n_formulas < 0) return fail_with(PCX_EINVAL, "negative size");
  pcx_engine* e = b->e;
  if (e->cfg.domain != PCX_DOMAIN_INTERVAL) return fail_with(PCX_EINVAL, "formula trees are asked on interval batches");
  std::string why;
  if (!trees_valid(nodes, n_nodes, off, n_formulas, b->nvars, &why)) return fail_with(PCX_EINVAL, "formula trees: " + why);
  PCX_HIP(hipSetDevice(e->cfg.device));
  int4* d_nodes = nullptr;
  int64_t* d_off = nullptr;
  int32_t* d_ok = nullptr;
  uint8_t* d_empty = nullptr;
  auto release = [&]() {
    for (void* p : {(void*)d_nodes, (void*)d_off, (void*)d_ok, (void*)d_empty}) if (p) (void)hipFree(p);
  };
  hipError_t he = hipMalloc(&d_nodes, std::max<int64_t>(n_nodes, 1) * sizeof(int4));
  if (he == hipSuccess) he = hipMalloc(&d_off, (n_formulas + 1) * sizeof(int64_t));
  if (he == hipSuccess) he = hipMalloc(&d_ok, b->n_inst * sizeof(int32_t));
  if (he == hipSuccess) he = hipMalloc(&d_empty, b->n_inst);
  if (he == hipSuccess && n_nodes > 0)
    he = hipMemcpyAsync(d_nodes, nodes, n_nodes * sizeof(int4), hipMemcpyHostToDevice, e->stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(d_off, off, (n_formulas + 1) * sizeof(int64_t), hipMemcpyHostToDevice, e->stream);
  std::vector<int32_t> ones(b->n_inst, 1);
  if (he == hipSuccess)
    he = hipMemcpyAsync(d_ok, ones.data(), b->n_inst * sizeof(int32_t), hipMemcpyHostToDevice, e->stream);
  if (he == hipSuccess) he = hipMemsetAsync(d_empty, 0, b->n_inst, e->stream);
  if (he == hipSuccess)
    he = launch_store_empty_map(b->d_store, b->nvars, (int32_t)b->n_inst, d_empty, dom_of(e), e->stream);
  if (he == hipSuccess)
    he = launch_formula_ask(d_nodes, d_off, n_formulas, b->d_store, b->nvars, b->n_inst, d_ok, e->stream);
  std::vector<uint8_t> empty(b->n_inst);
  if (he == hipSuccess)
    he = hipMemcpyAsync(extractable, d_ok, b->n_inst * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream);
  if (he == hipSuccess)
    he = hipMemcpyAsync(empty.data(), d_empty, b->n_inst, hipMemcpyDeviceToHost, e->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
  else (void)hipStreamSynchronize(e->stream);     // nothing queued still uses the buffers
  release();
  if (he != hipSuccess) return fail_with(PCX_EDEVICE, std::string("formula ask: ") + hipGetErrorString(he));
  for (int64_t k = 0; k < b->n_inst; ++k) extractable[k] = extractable[k] && !empty[k];
  return PCX_OK;
}

pcx_status pcx_batch_ask(pcx_batch* b, int32_t* extractable, int64_t* live_off, int32_t* live_idx,
                         int64_t cap) {
  if (!b) return fail_with(PCX_EINVAL, "NULL batch");
  pcx_engine* e = b->e;
  if (e->cfg.domain != PCX_DOMAIN_INTERVAL) return fail_with(PCX_EINVAL, "entailment is defined for interval batches");
  if (live_idx && !live_off) return fail_with(PCX_EINVAL, "live_idx needs live_off");
  if ((int64_t)e->op.size() != b->table_n) return fail_with(PCX_EINVAL, "table changed since the batch was created");
  if (e->max_var >= b->nvars) return fail_with(PCX_EINVAL, "a propagator names a variable outside the store");
  PCX_HIP(hipSetDevice(e->cfg.device));
  pcx_status st = sync_table(e);
  if (st != PCX_OK) return st;
  const int64_t n = b->n_inst;
  const Table t = table_of(e);
  const Subsets sub{b->h_off.empty() ? nullptr : b->d_off, b->d_idx};
  int64_t* d_live = nullptr;
  int32_t* d_empty = nullptr;
  int64_t* d_off = nullptr;
  int32_t* d_out = nullptr;
  auto release = [&]() { dfree(d_live); dfree(d_empty); dfree(d_off); dfree(d_out); };
  std::vector<int64_t> live(n);
  std::vector<int32_t> empty(n);
  hipError_t he = hipMalloc(&d_live, n * sizeof(int64_t));
  if (he == hipSuccess) he = hipMalloc(&d_empty, n * sizeof(int32_t));
  if (he == hipSuccess)
    he = launch_batch_ask(t, b->d_store, b->nvars, sub, b->table_n, n, d_live, d_empty, nullptr, nullptr,
                          e->stream);
  if (he == hipSuccess) he = hipMemcpyAsync(live.data(), d_live, n * sizeof(int64_t), hipMemcpyDeviceToHost, e->stream);
  if (he == hipSuccess) he = hipMemcpyAsync(empty.data(), d_empty, n * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
  if (he != hipSuccess) {
    release();
    return fail_with(PCX_EDEVICE, std::string("batch ask: ") + hipGetErrorString(he));
  }
  std::vector<int64_t> off(n + 1, 0);
  for (int64_t k = 0; k < n; ++k) off[k + 1] = off[k] + live[k];
  if (extractable)
    for (int64_t k = 0; k < n; ++k) extractable[k] = !empty[k] && live[k] == 0;
  if (live_off) std::copy(off.begin(), off.end(), live_off);
  if (live_idx) {
    if (off[n] > cap) {
      release();
      return fail_with(PCX_EINVAL, "live_idx capacity too small (live_off[n_inst] holds the size needed)");
    }
    if (off[n] > 0) {
      he = hipMalloc(&d_off, (n + 1) * sizeof(int64_t));
      if (he == hipSuccess) he = hipMalloc(&d_out, off[n] * sizeof(int32_t));
      if (he == hipSuccess) he = hipMemcpyAsync(d_off, off.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, e->stream);
      if (he == hipSuccess)
        he = launch_batch_ask(t, b->d_store, b->nvars, sub, b->table_n, n, nullptr, nullptr, d_off, d_out,
                              e->stream);
      if (he == hipSuccess) he = hipMemcpyAsync(live_idx, d_out, off[n] * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream);
      if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    }
  }
  release();
  if (he != hipSuccess) return fail_with(PCX_EDEVICE, std::string("batch ask: ") + hipGetErrorString(he));
  return PCX_OK;
}

pcx_status pcx_batch_autotune(pcx_batch* b, int32_t reps, pcx_tuning* best, pcx_tuning* results,
                              int32_t max_results, int32_t* n_results) {
  if (!b || !best) return fail_with(PCX_EINVAL, "NULL argument");
  if (reps < 1 || reps > 100) return fail_with(PCX_EINVAL, "reps must be in [1, 100]");
  pcx_engine* e = b->e;
  PCX_HIP(hipSetDevice(e->cfg.device));
  const pcx_config saved = e->cfg;
  std::vector<pcx_tuning> pts;
  const bool subsets = !b->h_off.empty();
  for (int ppw : {64, 128, 256, 512, 1024, 2048})
    for (int c = 0; c <= (subsets ? 1 : 0); ++c)
      for (int st = 0; st <= 1; ++st)
        for (int ev = 0; ev <= 1; ++ev)
          pts.push_back(pcx_tuning{PCX_MODE_BATCHED, ppw, 0, c, st, ev, 0, 0.0});
  // lockstep full sweeps read the shared table: props_per_wave, compact and
  // stream only shape the marked sweeps that follow
  if (b->n_inst >= 2)
    for (int gi : {16, 32, 64})
      pts.push_back(pcx_tuning{PCX_MODE_BATCHED, 512, 0, 0, 0, 1, gi, 0.0});
  const bool lds_fits = b->nvars * 8 <= lds_store_cap(e->cfg.device);
  for (int l = 0; l <= (lds_fits ? 1 : 0); ++l)
    for (int c = 0; c <= (subsets ? 1 : 0); ++c)
      pts.push_back(pcx_tuning{PCX_MODE_PERSISTENT, 0, l, c, 0, 0, 0, 0.0});
  int best_i = -1;
  for (size_t i = 0; i < pts.size(); ++i) {
    pcx_tuning& p = pts[i];
    e->cfg.props_per_wave = p.props_per_wave;
    e->cfg.lds_mode = p.lds_mode;
    e->cfg.compact_tables = p.compact_tables;
    e->cfg.stream_tables = p.stream_tables;
    e->cfg.event_driven = p.event_driven;
    e->cfg.lockstep = p.lockstep;
    e->cfg.timing = 0;
    e->cfg.base_seed = 0;               // the points shape full sweeps: time those
    double bt = -1.0;
    for (int r = 0; r <= reps; ++r) {
      pcx_status s = pcx_batch_reset(b);
      if (s == PCX_OK) s = hipStreamSynchronize(e->stream) == hipSuccess ? PCX_OK : PCX_EDEVICE;
      const double t0 = now_ms();
      if (s == PCX_OK) s = pcx_batch_run(b, p.mode, nullptr, nullptr);
      const double t = now_ms() - t0;
      if (s != PCX_OK) {
        if (s != PCX_ENOMEM) { e->cfg = saved; return s; }   // a point that does not fit is skipped
        bt = -1.0;
        break;
      }
      if (r > 0 && (bt < 0 || t < bt)) bt = t;              // run 0 warms up
    }
    p.ms = bt;
    if (bt >= 0 && (best_i < 0 || bt < pts[best_i].ms)) best_i = (int)i;
  }
  e->cfg = saved;
  if (results)
    for (size_t i = 0; i < pts.size() && (int32_t)i < max_results; ++i) results[i] = pts[i];
  if (n_results) *n_results = (int32_t)pts.size();
  if (best_i < 0) return fail_with(PCX_ENOMEM, "no configuration could run");
  const pcx_tuning& w = pts[best_i];
  e->cfg.props_per_wave = w.props_per_wave;
  e->cfg.lds_mode = w.lds_mode;
  e->cfg.compact_tables = w.compact_tables;
  e->cfg.stream_tables = w.stream_tables;
  e->cfg.event_driven = w.event_driven;
  e->cfg.lockstep = w.lockstep;
  *best = w;
  pcx_status s = pcx_batch_reset(b);
  if (s == PCX_OK) PCX_HIP(hipStreamSynchronize(e->stream));
  return s;
}

pcx_status pcx_fixpoint(pcx_engine* e, int32_t* lb, int32_t* ub, int64_t nvars, pcx_result* out) {
  if (!e || !lb || !ub) return fail_with(PCX_EINVAL, "NULL argument");
  if (e->cfg.domain != PCX_DOMAIN_INTERVAL) return fail_with(PCX_EINVAL, "pcx_fixpoint needs an interval engine");
  if (nvars <= 0) return fail_with(PCX_EINVAL, "nvars must be positive");
  if (e->max_var >= nvars) return fail_with(PCX_EINVAL, "a propagator names a variable outside the store");
  if (!e->scratch || e->scratch_nvars != nvars || e->scratch_table != (int64_t)e->op.size()) {
    if (e->scratch) pcx_batch_destroy(e->scratch);
    e->scratch = nullptr;
    pcx_status s = pcx_batch_create(e, 1, nvars, nullptr, nullptr, &e->scratch);
    if (s != PCX_OK) return s;
    e->scratch_nvars = nvars;
    e->scratch_table = (int64_t)e->op.size();
  }
  pcx_status s = pcx_batch_load(e->scratch, lb, ub);
  if (s != PCX_OK) return s;
  pcx_result r{};
  s = pcx_batch_run(e->scratch, PCX_MODE_PLAIN, &r, nullptr);
  if (s != PCX_OK) return s;
  s = pcx_batch_read(e->scratch, lb, ub);
  if (s != PCX_OK) return s;
  if (out) *out = r;
  return PCX_OK;
}

pcx_status pcx_ask_all(pcx_engine* e, const int32_t* lb, const int32_t* ub, int64_t nvars,
                       uint8_t* entailed, int32_t* all) {
  if (!e || !lb || !ub) return fail_with(PCX_EINVAL, "NULL argument");
  if (e->cfg.domain != PCX_DOMAIN_INTERVAL) return fail_with(PCX_EINVAL, "pcx_ask_all needs an interval engine");
  if (e->max_var >= nvars) return fail_with(PCX_EINVAL, "a propagator names a variable outside the store");
  PCX_HIP(hipSetDevice(e->cfg.device));
  pcx_status s = sync_table(e);
  if (s != PCX_OK) return s;
  const int64_t n = (int64_t)e->op.size();
  int32_t *d_lb = nullptr, *d_ub = nullptr, *d_all = nullptr;
  int2* d_st = nullptr;
  uint8_t* d_ent = nullptr;
  hipError_t he = hipMalloc(&d_lb, nvars * sizeof(int32_t));
  if (he == hipSuccess) he = hipMalloc(&d_ub, nvars * sizeof(int32_t));
  if (he == hipSuccess) he = hipMalloc(&d_st, nvars * sizeof(int2));
  if (he == hipSuccess) he = hipMalloc(&d_all, sizeof(int32_t));
  if (he == hipSuccess) he = hipMalloc(&d_ent, std::max<int64_t>(n, 1));
  const int32_t one = 1;
  if (he == hipSuccess) he = hipMemcpyAsync(d_lb, lb, nvars * sizeof(int32_t), hipMemcpyHostToDevice, e->stream);
  if (he == hipSuccess) he = hipMemcpyAsync(d_ub, ub, nvars * sizeof(int32_t), hipMemcpyHostToDevice, e->stream);
  if (he == hipSuccess) he = hipMemcpyAsync(d_all, &one, sizeof(int32_t), hipMemcpyHostToDevice, e->stream);
  if (he == hipSuccess) he = launch_store_pack(d_lb, d_ub, d_st, nvars, e->stream);
  if (he == hipSuccess) he = launch_ask(table_of(e), d_st, d_ent, d_all, e->stream);
  int32_t h_all = 1;
  if (he == hipSuccess) he = hipMemcpyAsync(&h_all, d_all, sizeof(int32_t), hipMemcpyDeviceToHost, e->stream);
  if (he == hipSuccess && entailed && n > 0)
    he = hipMemcpyAsync(entailed, d_ent, n, hipMemcpyDeviceToHost, e->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
  (void)hipFree(d_lb); (void)hipFree(d_ub); (void)hipFree(d_st); (void)hipFree(d_all); (void)hipFree(d_ent);
  if (he != hipSuccess) return fail_with(PCX_EDEVICE, std::string("ask: ") + hipGetErrorString(he));
  if (all) *all = h_all;
  return PCX_OK;
}

// ---- bitset domain --------------------------------------------------------
pcx_status pcx_bits_tell(pcx_engine* e, const int32_t* op, const int32_t* x, const int32_t* y,
                         const int32_t* z, int64_t n, uint64_t* words, int64_t nvars) {
  if (!e) return fail_with(PCX_EINVAL, "engine is NULL");
  if (e->cfg.domain != PCX_DOMAIN_BITSET) return fail_with(PCX_EINVAL, "pcx_bits_tell needs a bitset engine");
  if (n < 0 || (n > 0 && (!op || !x || !y || !z))) return fail_with(PCX_EINVAL, "bad table arrays");
  if (words && nvars <= 0) return fail_with(PCX_EINVAL, "nvars must be given with a store");
  const int64_t lim = nvars > 0 ? std::min<int64_t>(nvars, PCX_MAX_VARS) : PCX_MAX_VARS;
  for (int64_t i = 0; i < n; ++i) {
    if (op[i] < 0 || op[i] >= PCX_BITS_NUM_OPS)
      return fail_with(PCX_EINVAL, "unknown bitset opcode at " + std::to_string(i));
    if (x[i] < 0 || x[i] >= lim || y[i] < 0 || y[i] >= lim || z[i] < 0 || z[i] >= lim)
      return fail_with(PCX_EINVAL, "variable index out of range at " + std::to_string(i));
  }
  // Reified outputs are Boolean literals: x <- x & {0,1}.
  if (words)
    for (int64_t i = 0; i < n; ++i)
      if (op[i] == PCX_BEQR) words[x[i]] &= 3ull;
  // PC keeps its propagators ordered by formula kind (pc.hpp:634-641); the
  // bitset opcodes are numbered in that order, so a stable sort by opcode.
  const int64_t old = (int64_t)e->op.size();
  std::vector<int64_t> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return op[a] < op[b]; });
  struct Bc { int32_t op, x, y, z; };
  std::vector<Bc> all(old + n);
  for (int64_t i = 0; i < old; ++i) all[i] = Bc{e->op[i], e->x[i], e->y[i], e->z[i]};
  for (int64_t i = 0; i < n; ++i) all[old + i] = Bc{op[ord[i]], x[ord[i]], y[ord[i]], z[ord[i]]};
  std::inplace_merge(all.begin(), all.begin() + old, all.end(),
                     [](const Bc& a, const Bc& b) { return a.op < b.op; });
  e->op.resize(old + n); e->x.resize(old + n); e->y.resize(old + n); e->z.resize(old + n);
  for (int64_t i = 0; i < old + n; ++i) {
    e->op[i] = all[i].op; e->x[i] = all[i].x; e->y[i] = all[i].y; e->z[i] = all[i].z;
    e->max_var = std::max(e->max_var, std::max(all[i].x, std::max(all[i].y, all[i].z)));
  }
  e->dirty = true;
  return PCX_OK;
}

pcx_status pcx_bits_batch_load(pcx_batch* b, const uint64_t* words) {
  if (!b || !words) return fail_with(PCX_EINVAL, "NULL argument");
  if (b->e->cfg.domain != PCX_DOMAIN_BITSET) return fail_with(PCX_EINVAL, "bitset batch expected");
  pcx_engine* e = b->e;
  PCX_HIP(hipSetDevice(e->cfg.device));
  const int64_t n = b->n_inst * b->nvars;
  PCX_HIP(hipMemcpyAsync(b->d_image, words, n * sizeof(uint64_t), hipMemcpyHostToDevice, e->stream));
  PCX_HIP(hipMemcpyAsync(b->d_store, b->d_image, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, e->stream));
  pcx_status st = image_loaded(b);
  if (st != PCX_OK) return st;
  return list_load_diffs(b);
}

pcx_status pcx_bits_batch_read(pcx_batch* b, uint64_t* words) {
  if (!b || !words) return fail_with(PCX_EINVAL, "NULL argument");
  if (b->e->cfg.domain != PCX_DOMAIN_BITSET) return fail_with(PCX_EINVAL, "bitset batch expected");
  pcx_engine* e = b->e;
  PCX_HIP(hipSetDevice(e->cfg.device));
  PCX_HIP(hipMemcpyAsync(words, b->d_store, b->n_inst * b->nvars * sizeof(uint64_t),
                         hipMemcpyDeviceToHost, e->stream));
  PCX_HIP(hipStreamSynchronize(e->stream));
  return PCX_OK;
}

pcx_status pcx_bits_fixpoint(pcx_engine* e, uint64_t* words, int64_t nvars, pcx_result* out) {
  if (!e || !words) return fail_with(PCX_EINVAL, "NULL argument");
  if (e->cfg.domain != PCX_DOMAIN_BITSET) return fail_with(PCX_EINVAL, "pcx_bits_fixpoint needs a bitset engine");
  if (nvars <= 0) return fail_with(PCX_EINVAL, "nvars must be positive");
  if (e->max_var >= nvars) return fail_with(PCX_EINVAL, "a propagator names a variable outside the store");
  if (!e->scratch || e->scratch_nvars != nvars || e->scratch_table != (int64_t)e->op.size()) {
    if (e->scratch) pcx_batch_destroy(e->scratch);
    e->scratch = nullptr;
    pcx_status s = pcx_batch_create(e, 1, nvars, nullptr, nullptr, &e->scratch);
    if (s != PCX_OK) return s;
    e->scratch_nvars = nvars;
    e->scratch_table = (int64_t)e->op.size();
  }
  pcx_status s = pcx_bits_batch_load(e->scratch, words);
  if (s != PCX_OK) return s;
  pcx_result r{};
  s = pcx_batch_run(e->scratch, PCX_MODE_PLAIN, &r, nullptr);
  if (s != PCX_OK) return s;
  s = pcx_bits_batch_read(e->scratch, words);
  if (s != PCX_OK) return s;
  if (out) *out = r;
  return PCX_OK;
}

}  // extern "C"
