// pcx_kernels.hip -- gfx950 kernels of the fixpoint engine.
//
// One thread evaluates one propagator per step: it reads the 12-byte table
// entry (three coalesced streams), gathers its three operands (one 8-byte
// load each: an {lb, ub} interval or a 64-bit bitset word), runs the
// propagator, and tightens the store with atomics (atomicMax on lb /
// atomicMin on ub, atomicAnd on words) only where its result is tighter than
// what it read.  "changed" and "failed" are reduced per wavefront with
// ballots and written once per wave.  Kernel boundaries order the sweeps: a
// sweep reads the store as the previous kernel left it, so a sweep in which
// no atomic tightened anything saw one constant store and proves the
// fixpoint.  Every kernel is instantiated per domain D:
//   kItv     interval store, no division propagators (fewer registers)
//   kItvDiv  interval store, all ten opcodes
//   kBits    bitset store
#include "pcx_kernels.h"
#include "../../include/pcx_model.h"
#include "pcx_prop.h"

#include <type_traits>

namespace pcx {

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Values the whole wave agrees on (work-item indices, descriptors read
// through them) moved to scalar registers: the compiler cannot prove a load
// from a uniform address returns a uniform value, and would otherwise keep
// them, and everything derived from them, in VGPRs.
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
template <class T>
__device__ __forceinline__ T* uni(T* p) { return reinterpret_cast<T*>(uni((int64_t)(uintptr_t)p)); }

// Subset indices are read once per sweep: stream them past the caches so
// they do not evict the store and table lines that are reused.
__device__ __forceinline__ int64_t prop_at(const int32_t* __restrict__ idx, int64_t q) {
  return idx ? (int64_t)__builtin_nontemporal_load(idx + q) : q;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// ---- store access policies ------------------------------------------------
// Agent scope (sweep kernels: ordered by kernel boundaries) or workgroup
// scope (persistent kernel: one workgroup owns an instance).
template <int SCOPE>
__device__ __forceinline__ u64 load_word(const void* p) {
  if constexpr (SCOPE == __HIP_MEMORY_SCOPE_AGENT) return *reinterpret_cast<const u64*>(p);
  else return __hip_atomic_load(reinterpret_cast<const u64*>(p), __ATOMIC_RELAXED, SCOPE);
}
__device__ __forceinline__ int2 as_itv(u64 v) {
  return make_int2((int32_t)(uint32_t)v, (int32_t)(uint32_t)(v >> 32));
}
__device__ __forceinline__ u64 itv_bits(int2 v) { return (u64)(uint32_t)v.x | ((u64)(uint32_t)v.y << 32); }

// Meet interval `n` into variable v, whose value was read as `cur`.
// Returns true when this call tightened a bound of v.
template <int SCOPE>
__device__ __forceinline__ bool meet_itv(int2* __restrict__ st, int32_t v, int2 cur, int2 n,
                                         bool& chg, bool& fail, int& nt) {
  bool t = false;
  if (n.x > cur.x) {
    const int old = __hip_atomic_fetch_max(&st[v].x, n.x, __ATOMIC_RELAXED, SCOPE);
    if (old < n.x) { t = true; ++nt; }
    if (n.x > cur.y) fail = true;          // cur.y >= current ub
  }
  if (n.y < cur.y) {
    const int old = __hip_atomic_fetch_min(&st[v].y, n.y, __ATOMIC_RELAXED, SCOPE);
    if (old > n.y) { t = true; ++nt; }
    if (n.y < cur.x) fail = true;
  }
  chg |= t;
  return t;
}

// AND mask `n` into word v, whose value was read as `cur`.
template <int SCOPE>
__device__ __forceinline__ bool meet_word(u64* __restrict__ st, int32_t v, u64 cur, u64 n,
                                          bool& chg, bool& fail, int& nt) {
  bool t = false;
  if ((cur & n) != cur) {
    const u64 old = __hip_atomic_fetch_and(&st[v], n, __ATOMIC_RELAXED, SCOPE);
    if ((old & n) != old) { t = true; ++nt; }
    if ((old & n) == 0) fail = true;       // the word is now empty
  }
  chg |= t;
  return t;
}

// Where a tightening of variable v is recorded for the next sweep of one
// instance (dirty == nullptr: nowhere, every sweep is a full sweep).  In a
// marked sweep the current sweep's bitmap (also_now) is marked as well: a
// propagator whose word this sweep has not reached yet then runs in this
// sweep, as it would in a full sweep; a bit that lands behind the sweep is
// an extra evaluation two sweeps later, which never changes the fixpoint.
struct Marker {
  const int32_t* adj_off;
  const int32_t* adj;
  const uint32_t* member;   // instance's row, nullptr = every propagator
  uint32_t* dirty;          // instance's row of the next sweep
  uint32_t* also_now;       // instance's row of this sweep (marked sweeps), or nullptr
  int32_t* list;            // workgroup's deferred variables (LDS), or nullptr
  int32_t* list_n;
  uint32_t* sum;            // summary row of the next sweep
  uint32_t* sum_now;        // summary row of this sweep (with also_now)
  uint32_t* lsum;           // LDS accumulators of sum / sum_now during a flush, or nullptr
  uint32_t* lsum_now;
  int32_t* overflow;        // LDS flag: the list overflowed (see flush_marks)
  int64_t words;            // bitmap words of a row
  int32_t swords;
  int32_t* all_marked;      // instance's all_marked word of the next sweep, or nullptr
  int32_t all_marked_s;     // the next sweep's number
};

// Summary words accumulated in LDS during a flush (larger summaries are
// marked in global memory directly).
constexpr int kLdsSummaryWords = 64;

// Deferred marking: a tightened variable is pushed onto the workgroup's LDS
// list and the whole workgroup marks the list's adjacency together at the
// end of its work item (flush_marks), instead of each thread walking the
// adjacency of every variable it tightened one after the other.
constexpr int kMarkCap = 1024;

__device__ __forceinline__ Marker marker_for(const EventSet* ev, int64_t k, int32_t s,
                                             bool marked_sweep = false) {
  if (!ev || !ev->dirty) return Marker{};
  const int64_t row = k * ev->words, srow = k * ev->swords;
  const int64_t plane = ev->n_inst * ev->words, splane = ev->n_inst * ev->swords;
  const int64_t nx = (s + 1) & 1, cu = s & 1;
  const bool sums = ev->summary != nullptr;
  Marker mk{};
  mk.adj_off = ev->adj_off;
  mk.adj = ev->adj;
  mk.member = ev->member ? ev->member + row : nullptr;
  mk.dirty = ev->dirty + nx * plane + row;
  mk.also_now = marked_sweep ? ev->dirty + cu * plane + row : nullptr;
  mk.sum = sums ? ev->summary + nx * splane + srow : nullptr;
  mk.sum_now = sums && marked_sweep ? ev->summary + cu * splane + srow : nullptr;
  mk.words = ev->words;
  mk.swords = (int32_t)ev->swords;
  return mk;
}

__device__ __forceinline__ void set_summary(uint32_t* lds, uint32_t* glob, int32_t c) {
  const uint32_t cbit = 1u << (c & 31);
  if (lds) {
    atomicOr(&lds[c >> 5], cbit);
  } else if (glob &&
             !(__hip_atomic_load(&glob[c >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & cbit)) {
    // a summary word covers 32 chunks: write it only when its bit is unset
    atomicOr(&glob[c >> 5], cbit);
  }
}

// Mark the propagators adjacent to v: adjacency entries are loaded eight at
// a time (independent loads in flight together), then marked with
// non-returning atomics.
__device__ __forceinline__ void mark_var(const Marker& mk, int32_t v) {
  constexpr int kB = 8;
  const int32_t e = mk.adj_off[v + 1];
  for (int32_t i = mk.adj_off[v]; i < e; i += kB) {
    int32_t p[kB];
#pragma unroll
    for (int j = 0; j < kB; ++j) p[j] = i + j < e ? mk.adj[i + j] : -1;
    uint32_t in[kB];
#pragma unroll
    for (int j = 0; j < kB; ++j)
      in[j] = (p[j] >= 0 && mk.member) ? mk.member[p[j] >> 5] : ~0u;
#pragma unroll
    for (int j = 0; j < kB; ++j) {
      if (p[j] < 0) continue;
      const uint32_t bit = 1u << (p[j] & 31);
      if (!(in[j] & bit)) continue;
      const int32_t c = p[j] >> kChunkShift;
      atomicOr(&mk.dirty[p[j] >> 5], bit);
      if (mk.sum) set_summary(mk.lsum, mk.sum, c);
      if (mk.also_now) {
        atomicOr(&mk.also_now[p[j] >> 5], bit);
        if (mk.sum_now) set_summary(mk.lsum_now, mk.sum_now, c);
      }
    }
  }
}

// Queued marks are staged per workgroup in LDS and reserved in their plane
// with one atomic per flush (a single tail counter per plane would otherwise
// take an atomic per wave from every workgroup of the chip).  Entries that
// find the staging area full take a slot of their own.
constexpr int kQStage = 1024;
struct QueueStage {
  unsigned long long* buf;   // LDS, kQStage entries
  int32_t* n;                // LDS counter
};

// This workgroup's sub-list of plane pl: its tail counter and first entry.
__device__ __forceinline__ int32_t* queue_tail(const EventSet& ev, int pl) {
  return ev.qctl + kQTails + pl * kQLists + blockIdx.x % kQLists;
}
__device__ __forceinline__ unsigned long long* queue_list(const EventSet& ev, int pl) {
  return ev.q + ((int64_t)pl * kQLists + blockIdx.x % kQLists) * ev.qcap;
}

__device__ __forceinline__ void stage_entry(const QueueStage& qs, const EventSet& ev, int pl,
                                            unsigned long long entry) {
  const int at = atomicAdd(qs.n, 1);
  if (at < kQStage) { qs.buf[at] = entry; return; }
  const int64_t g = atomicAdd(queue_tail(ev, pl), 1);
  if (g < ev.qcap) queue_list(ev, pl)[g] = entry;
  else ev.qctl[kQOver + pl] = 1;
}

// Every thread of the workgroup calls this after a flush of marks; the
// caller's barrier orders the staging before it.
__device__ __forceinline__ void flush_stage(const QueueStage& qs, const EventSet& ev, int pl) {
  __shared__ int32_t s_base;
  const int n = min(*qs.n, kQStage);
  if (threadIdx.x == 0) s_base = n ? atomicAdd(queue_tail(ev, pl), n) : 0;
  __syncthreads();
  const int64_t base = s_base;
  unsigned long long* list = queue_list(ev, pl);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (base + i < ev.qcap) list[base + i] = qs.buf[i];
    else ev.qctl[kQOver + pl] = 1;
  }
  __syncthreads();                     // every thread has read the count and the entries
  if (threadIdx.x == 0) *qs.n = 0;
}

// mark_var for queued marks: a propagator whose bit this call sets is
// staged for the next sweep's plane (a returning atomic tells).  An atomic
// load that already shows the bit saves the atomic: the next sweep's bits
// are only set, never cleared, during a launch.  Summary bits are kept for a
// sweep that falls back to the bitmaps.
__device__ __forceinline__ void mark_var_q(const Marker& mk, const EventSet& ev, const QueueStage& qs,
                                           int64_t k, int32_t v, int32_t next) {
  constexpr int kB = 8;
  const int pl = next & 1;
  const int32_t e = mk.adj_off[v + 1];
  for (int32_t i = mk.adj_off[v]; i < e; i += kB) {
    int32_t p[kB];
#pragma unroll
    for (int j = 0; j < kB; ++j) p[j] = i + j < e ? mk.adj[i + j] : -1;
    uint32_t in[kB];
#pragma unroll
    for (int j = 0; j < kB; ++j)
      in[j] = (p[j] >= 0 && mk.member) ? mk.member[p[j] >> 5] : ~0u;
    uint32_t seen[kB];
#pragma unroll
    for (int j = 0; j < kB; ++j)
      seen[j] = p[j] >= 0 ? __hip_atomic_load(&mk.dirty[p[j] >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                          : ~0u;
#pragma unroll
    for (int j = 0; j < kB; ++j) {
      if (p[j] < 0) continue;
      const uint32_t bit = 1u << (p[j] & 31);
      if (!(in[j] & bit) || (seen[j] & bit)) continue;
      if (atomicOr(&mk.dirty[p[j] >> 5], bit) & bit) continue;
      if (mk.sum) set_summary(nullptr, mk.sum, p[j] >> kChunkShift);
      stage_entry(qs, ev, pl, ((unsigned long long)k << 32) | (uint32_t)p[j]);
    }
  }
}

// Every marking kernel gives its workgroup an LDS list (mk.list).  When the
// list is full the variable is not recorded; the flush then marks every
// propagator of the instance for the next sweep instead (a superset of the
// marks the lost variables needed): by one word (mk.all_marked) where the
// kernel sets it, or by OR-ing the member row into the bitmap.
__device__ __forceinline__ void defer_mark(const Marker& mk, int32_t v) {
  const int i = atomicAdd(mk.list_n, 1);
  if (i < kMarkCap) mk.list[i] = v;
  else *mk.overflow = 1;
}

// Every thread of the workgroup calls this at the end of a work item.  The
// summary bits of the flushed marks are gathered in LDS and written with
// one atomic per summary word.
__device__ __forceinline__ void flush_marks(const Marker& mk) {
  if (!mk.list) return;
  __shared__ uint32_t s_sum[2][kLdsSummaryWords];
  const bool lds_sum = mk.sum && mk.swords <= kLdsSummaryWords;
  if (lds_sum)
    for (int j = threadIdx.x; j < 2 * kLdsSummaryWords; j += blockDim.x) s_sum[j >> 6][j & 63] = 0;
  // Another workgroup already marked everything for the next sweep: these
  // marks add nothing (a stale read only means marking them anyway).
  if (mk.all_marked && threadIdx.x == 0 &&
      __hip_atomic_load(mk.all_marked, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == mk.all_marked_s)
    *mk.overflow = 1;
  __syncthreads();
  if (*mk.overflow && mk.all_marked) {       // lost variables: every propagator for the next sweep
    if (threadIdx.x == 0) *mk.all_marked = mk.all_marked_s;
    __syncthreads();                        // every thread has read the flag
    if (threadIdx.x == 0) { *mk.list_n = 0; *mk.overflow = 0; }   // ordered by the caller's barrier
    return;
  }
  Marker m2 = mk;
  if (lds_sum) {
    m2.lsum = s_sum[0];
    m2.lsum_now = mk.sum_now ? s_sum[1] : nullptr;
  }
  const int n = min(*mk.list_n, kMarkCap);
  for (int i = threadIdx.x; i < n; i += blockDim.x) mark_var(m2, mk.list[i]);
  __syncthreads();
  if (lds_sum) {
    for (int j = threadIdx.x; j < mk.swords; j += blockDim.x) {
      if (s_sum[0][j]) atomicOr(&mk.sum[j], s_sum[0][j]);
      if (mk.sum_now && s_sum[1][j]) atomicOr(&mk.sum_now[j], s_sum[1][j]);
    }
  }
  if (*mk.overflow) {                       // lost variables: mark the whole instance
    for (int64_t w = threadIdx.x; w < mk.words; w += blockDim.x)
      atomicOr(&mk.dirty[w], mk.member ? mk.member[w] : ~0u);
    if (mk.sum) {
      const int64_t chunks = (mk.words + kChunkWords - 1) / kChunkWords;
      for (int64_t j = threadIdx.x; j < mk.swords; j += blockDim.x) {
        const int64_t left = chunks - j * 32;
        atomicOr(&mk.sum[j], left >= 32 ? ~0u : ((1u << left) - 1u));
      }
    }
  }
  __syncthreads();                          // every thread has read the flag
  if (threadIdx.x == 0) { *mk.list_n = 0; *mk.overflow = 0; }   // ordered by the caller's barrier
}

// Apply one propagator to operands already loaded.
template <int D, int SCOPE>
__device__ __forceinline__ void apply(int2* __restrict__ st, int op, int32_t x, int32_t y,
                                      int32_t z, u64 wx, u64 wy, u64 wz, bool& chg, bool& fail,
                                      int& nt, const Marker& mk) {
  bool tx, ty, tz;
  if constexpr (D == kBits) {
    u64* ws = reinterpret_cast<u64*>(st);
    if (!wx || !wy || !wz) { fail = true; return; }
    u64 nx, ny, nz;
    deduce_bits(op, wx, wy, wz, nx, ny, nz);
    tx = meet_word<SCOPE>(ws, x, wx, nx, chg, fail, nt);
    ty = meet_word<SCOPE>(ws, y, wy, ny, chg, fail, nt);
    tz = meet_word<SCOPE>(ws, z, wz, nz, chg, fail, nt);
  } else {
    const int2 X = as_itv(wx), Y = as_itv(wy), Z = as_itv(wz);
    if (X.x > X.y || Y.x > Y.y || Z.x > Z.y) { fail = true; return; }
    int2 nx, ny, nz;
    deduce<D == kItvDiv>(op, X, Y, Z, nx, ny, nz);
    if (nx.x > nx.y || ny.x > ny.y || nz.x > nz.y) fail = true;
    tx = meet_itv<SCOPE>(st, x, X, nx, chg, fail, nt);
    ty = meet_itv<SCOPE>(st, y, Y, ny, chg, fail, nt);
    tz = meet_itv<SCOPE>(st, z, Z, nz, chg, fail, nt);
  }
  if (mk.dirty) {
    if (tx) defer_mark(mk, x);
    if (ty) defer_mark(mk, y);
    if (tz) defer_mark(mk, z);
  }
}

// Table loads: cached (a shared table that other instances re-read) or
// non-temporal (NT: a table read once per sweep, streamed past the caches so
// that the store lines, which every propagator gathers from, stay in L2).
template <bool NT, class T>
__device__ __forceinline__ T table_load(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <int D, bool NT, int SCOPE = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ void eval_prop(const Table& t, int2* __restrict__ st, int64_t i,
                                          bool& chg, bool& fail, int& nt, const Marker& mk) {
  const uint32_t xo = table_load<NT>(t.xo + i);
  const int32_t y = table_load<NT>(t.y + i), z = table_load<NT>(t.z + i);
  const int32_t x = (int32_t)(xo & kVarMask);
  const u64 wx = load_word<SCOPE>(st + x);
  const u64 wy = load_word<SCOPE>(st + y);
  const u64 wz = load_word<SCOPE>(st + z);
  apply<D, SCOPE>(st, (int)(xo >> kOpShift), x, y, z, wx, wy, wz, chg, fail, nt, mk);
}

// Publish a wave's verdict: one store per flag, one atomic for the count.
__device__ __forceinline__ void publish(InstFlags* f, int32_t s, bool chg, bool fail, int nt,
                                        unsigned long long* tight) {
  const unsigned long long bc = __ballot(chg);
  const unsigned long long bf = __ballot(fail);
  const int sum = wave_sum(nt);
  if (lane_id() == 0) {
    if (bc) f->changed[s & 1] = s;
    if (bf) f->failed_at = s;
    if (sum) atomicAdd(&tight[(blockIdx.x * 4 + (threadIdx.x >> 6)) & (kTightSlots - 1)],
                       (unsigned long long)sum);
  }
}

// Instance k's propagators: the whole table (no subsets), the shared table
// through subset indices, or its own copy of its subset (off without idx:
// the table argument holds every instance's copy back to back).
struct InstTable {
  Table t;
  const int32_t* idx;
  int64_t m;
};
__device__ __forceinline__ InstTable inst_table(const Table& t, const Subsets& sub, int64_t k,
                                                int64_t n_props) {
  if (!sub.off) return InstTable{t, nullptr, n_props};
  const int64_t beg = sub.off[k], m = sub.off[k + 1] - beg;
  if (sub.idx) return InstTable{t, sub.idx + beg, m};
  return InstTable{Table{t.xo + beg, t.y + beg, t.z + beg, m}, nullptr, m};
}

// Active slot holding work item `item`: the last a with item_base[a] <= item.
// A workgroup claims items in increasing order, so the search starts at the
// slot of its previous item (`hint`): usually one or two loads instead of a
// full binary search of dependent loads.
__device__ __forceinline__ int find_active(const int64_t* item_base, int n_act, int64_t item,
                                           int hint) {
  int lo = hint, hi = n_act - 1;
  if (lo >= hi || item_base[lo + 1] > item) return lo;
  ++lo;
  if (lo >= hi || item_base[lo + 1] > item) return lo;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (item_base[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ---- plain mode: one instance, one sweep per launch -----------------------
template <int PPL, int D, bool NT>
__global__ __launch_bounds__(kBlock) void k_sweep_plain(Table t, int2* __restrict__ st,
                                                        const int32_t* __restrict__ idx, int64_t m,
                                                        int32_t s, InstFlags* f,
                                                        unsigned long long* tight, EventSet ev,
                                                        int64_t inst) {
  // Sweep s runs only if sweep s-1 tightened something and nothing failed.
  if (f->changed[(s - 1) & 1] != s - 1 || f->failed_at != kNoFail) return;
  const int64_t base = (int64_t)blockIdx.x * (kBlock * PPL);
  __shared__ int32_t s_list[kMarkCap];
  __shared__ int32_t s_n, s_over;
  Marker mk = marker_for(&ev, inst, s);
  if (mk.dirty) {
    mk.list = s_list;
    mk.list_n = &s_n;
    mk.overflow = &s_over;
    mk.all_marked = &f->all_marked[(s + 1) & 1];
    mk.all_marked_s = s + 1;
    if (threadIdx.x == 0) { s_n = 0; s_over = 0; }
    __syncthreads();
  }
  bool chg = false, fail = false;
  int nt = 0;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const int64_t k = base + j * kBlock + threadIdx.x;
    if (k < m) eval_prop<D, NT>(t, st, prop_at(idx, k), chg, fail, nt, mk);
  }
  publish(f, s, chg, fail, nt, tight);
  flush_marks(mk);
}

template <int D, bool NT>
hipError_t launch_plain_d(const Table& t, int2* store, const int32_t* idx, int64_t m,
                          int32_t sweep, InstFlags* flags, unsigned long long* tight, int ppl,
                          const EventSet& ev, int64_t inst, hipStream_t stream) {
  const int64_t per_block = (int64_t)kBlock * ppl;
  const int64_t grid = (m + per_block - 1) / per_block;
  if (grid <= 0) return hipSuccess;
  const dim3 g((unsigned)grid), b(kBlock);
#define PCX_PLAIN(P) \
  k_sweep_plain<P, D, NT><<<g, b, 0, stream>>>(t, store, idx, m, sweep, flags, tight, ev, inst)
  switch (ppl) {
    case 1: PCX_PLAIN(1); break;
    case 2: PCX_PLAIN(2); break;
    case 4: PCX_PLAIN(4); break;
    case 8: PCX_PLAIN(8); break;
    default: return hipErrorInvalidValue;
  }
#undef PCX_PLAIN
  return hipGetLastError();
}

template <bool NT>
hipError_t launch_plain_nt(const Table& t, int2* store, const int32_t* idx, int64_t m,
                           int32_t sweep, InstFlags* flags, unsigned long long* tight, int ppl,
                           int dom, const EventSet& ev, int64_t inst, hipStream_t stream) {
  switch (dom) {
    case kItv: return launch_plain_d<kItv, NT>(t, store, idx, m, sweep, flags, tight, ppl, ev, inst, stream);
    case kItvDiv: return launch_plain_d<kItvDiv, NT>(t, store, idx, m, sweep, flags, tight, ppl, ev, inst, stream);
    case kBits: return launch_plain_d<kBits, NT>(t, store, idx, m, sweep, flags, tight, ppl, ev, inst, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sweep_plain(const Table& t, int2* store, const int32_t* idx, int64_t m,
                              int32_t sweep, InstFlags* flags, unsigned long long* tight,
                              int ppw, int dom, bool stream_table, hipStream_t stream,
                              const EventSet* ev, int64_t inst) {
  const int ppl = ppw / kWave;
  const EventSet none{nullptr, nullptr, nullptr, nullptr, 0, 0};
  const EventSet& e = ev ? *ev : none;
  return stream_table
             ? launch_plain_nt<true>(t, store, idx, m, sweep, flags, tight, ppl, dom, e, inst, stream)
             : launch_plain_nt<false>(t, store, idx, m, sweep, flags, tight, ppl, dom, e, inst, stream);
}

// ---- marked-only sweeps ----------------------------------------------------
// A marked chunk swept by the whole workgroup: thread t reads bitmap word
// w0 + t (when `active`) and clears it, a block scan of the popcounts places
// every set bit's propagator in an LDS list, and the workgroup evaluates the
// list spread evenly over its threads (a word with many marks no longer
// serializes one thread).  Every thread of the workgroup calls this.
constexpr int kPropCap = 1024;
constexpr int kMarkedPasses = 16;  // at most, passes over a chunk per marked sweep
constexpr int kBatchedMarkedPasses = 4;    // the same in batched marked sweeps (short chains)
constexpr int kMarkedGroupMax = 8;        // marked chunks per batched work item, at most
constexpr int64_t kMarkedTargetItems = 2048;   // work items a marked sweep keeps, at least
template <int D>
// `fresh`: read the words with device-scope atomics (a pass over a chunk
// this workgroup already swept in this launch, whose lines the CU's L1 may
// hold from before other workgroups marked it).  Returns the number of
// marked propagators the pass evaluated (the same in every thread).
__device__ __forceinline__ int sweep_marked_chunk(const Table& t, int2* __restrict__ st,
                                                  uint32_t* cur_row, int64_t w0, int64_t words,
                                                  bool active, const Marker& mk, bool& chg,
                                                  bool& fail, int& nt, int& ne, int32_t* s_props,
                                                  int32_t* s_scan, bool fresh = false,
                                                  bool all = false) {
  const int64_t w = w0 + threadIdx.x;
  uint32_t bits = 0;
  if (active && w < words) {
    if (all && !fresh) {                   // every member propagator is marked
      bits = mk.member ? mk.member[w] : ~0u;
      if (cur_row[w]) cur_row[w] = 0;
    } else if (fresh) {
      // one round trip: a pass's latency is the chain's per-link latency
      // (a load, then an exchange when the word is set, was 4% slower at
      // config 1)
      bits = __hip_atomic_exchange(&cur_row[w], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      bits = cur_row[w];
      if (bits) cur_row[w] = 0;
    }
  }
  const int c = __builtin_popcount(bits);
  const int lane = lane_id();
  int incl = c;                                   // inclusive scan within the wave
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int v = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += v;
  }
  const int wv = threadIdx.x / kWave;
  if (lane == kWave - 1) s_scan[wv] = incl;
  __syncthreads();
  int off = incl - c, total = 0;
#pragma unroll
  for (int i = 0; i < kBlock / kWave; ++i) {
    const int v = s_scan[i];
    if (i < wv) off += v;
    total += v;
  }
  for (int base = 0; base < total; base += kPropCap) {   // same trip count in every thread
    uint32_t b2 = bits;
    int idx = off;
    while (b2) {
      const int j = __builtin_ctz(b2);
      b2 &= b2 - 1;
      if (idx >= base && idx < base + kPropCap) s_props[idx - base] = (int32_t)(w * 32 + j);
      ++idx;
    }
    __syncthreads();
    const int n = min(kPropCap, total - base);
    for (int i = threadIdx.x; i < n; i += kBlock) {
      const int64_t p = s_props[i];
      if (p < t.n) {
        eval_prop<D, false>(t, st, p, chg, fail, nt, mk);
        ++ne;
      }
    }
    __syncthreads();
  }
  return total;
}

// The marked propagators of ng chunks (ids in LDS), scanned together: thread
// t reads word t of each chunk, one block scan places every set bit in the
// LDS list, and the workgroup evaluates the list spread over its threads.
// Same contract as sweep_marked_chunk.
template <int D>
__device__ __forceinline__ int sweep_marked_group(const Table& t, int2* __restrict__ st,
                                                  uint32_t* cur_row, const int32_t* chunks, int ng,
                                                  int64_t words, bool active, const Marker& mk,
                                                  bool& chg, bool& fail, int& nt, int& ne,
                                                  int32_t* s_props, int32_t* s_scan, bool fresh,
                                                  bool all) {
  uint32_t bits[kMarkedGroupMax];
  int c = 0;
#pragma unroll
  for (int g = 0; g < kMarkedGroupMax; ++g) {
    bits[g] = 0;
    const int64_t w = (int64_t)chunks[g] * kChunkWords + threadIdx.x;
    if (active && g < ng && w < words) {
      if (all && !fresh) {                 // every member propagator is marked
        bits[g] = mk.member ? mk.member[w] : ~0u;
        if (cur_row[w]) cur_row[w] = 0;
      } else if (fresh) {
        bits[g] = __hip_atomic_load(&cur_row[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (bits[g]) bits[g] = __hip_atomic_exchange(&cur_row[w], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        bits[g] = cur_row[w];
        if (bits[g]) cur_row[w] = 0;
      }
    }
    c += __builtin_popcount(bits[g]);
  }
  const int lane = lane_id();
  int incl = c;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int v = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += v;
  }
  const int wv = threadIdx.x / kWave;
  if (lane == kWave - 1) s_scan[wv] = incl;
  __syncthreads();
  int off = incl - c, total = 0;
#pragma unroll
  for (int i = 0; i < kBlock / kWave; ++i) {
    const int v = s_scan[i];
    if (i < wv) off += v;
    total += v;
  }
  for (int base = 0; base < total; base += kPropCap) {   // same trip count in every thread
    int idx = off;
#pragma unroll
    for (int g = 0; g < kMarkedGroupMax; ++g) {
      const int64_t w = (int64_t)chunks[g] * kChunkWords + threadIdx.x;
      for (uint32_t b2 = bits[g]; b2; b2 &= b2 - 1) {
        if (idx >= base && idx < base + kPropCap) s_props[idx - base] = (int32_t)(w * 32 + __builtin_ctz(b2));
        ++idx;
      }
    }
    __syncthreads();
    const int n = min(kPropCap, total - base);
    for (int i = threadIdx.x; i < n; i += kBlock) {
      const int64_t p = s_props[i];
      if (p < t.n) {
        eval_prop<D, false>(t, st, p, chg, fail, nt, mk);
        ++ne;
      }
    }
    __syncthreads();
  }
  return total;
}

__device__ __forceinline__ void count_evals(InstFlags* f, int ne) {
  const int sum = wave_sum(ne);
  if (lane_id() == 0 && sum) atomicAdd((unsigned long long*)&f->marked_evals, (unsigned long long)sum);
}

template <int D>
__global__ __launch_bounds__(kBlock) void k_sweep_marked_plain(Table t, int2* __restrict__ st,
                                                               int32_t s, InstFlags* f,
                                                               unsigned long long* tight,
                                                               EventSet ev, int64_t inst) {
  if (f->changed[(s - 1) & 1] != s - 1 || f->failed_at != kNoFail) return;
  uint32_t* cur = ev.dirty + (int64_t)(s & 1) * ev.n_inst * ev.words + inst * ev.words;
  __shared__ int32_t s_list[kMarkCap];
  __shared__ int32_t s_props[kPropCap];
  __shared__ int32_t s_scan[kBlock / kWave];
  __shared__ int32_t s_n, s_over;
  Marker mk = marker_for(&ev, inst, s, true);
  mk.list = s_list;
  mk.list_n = &s_n;
  mk.overflow = &s_over;
  mk.all_marked = &f->all_marked[(s + 1) & 1];
  mk.all_marked_s = s + 1;
  const bool all = uni((int32_t)(f->all_marked[s & 1] == s)) != 0;
  if (threadIdx.x == 0) { s_n = 0; s_over = 0; }
  __syncthreads();
  bool chg = false, fail = false;
  int nt = 0, ne = 0;
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < ev.words;
       base += (int64_t)gridDim.x * kBlock) {      // same trip count in every thread
    // Passes over the chunk while they find marks: marks that other
    // workgroups (or this one) put in it during this sweep are evaluated in
    // this sweep, so a chain can advance several links per launch.
    for (int pass = 0;; ++pass) {
      const int total = sweep_marked_chunk<D>(t, st, cur, base, ev.words, true, mk, chg, fail, nt, ne,
                                              s_props, s_scan, pass > 0, all);
      flush_marks(mk);
      __syncthreads();
      if (total == 0 || pass + 1 >= kMarkedPasses) break;
    }
  }
  publish(f, s, chg, fail, nt, tight);
  count_evals(f, ne);
}

hipError_t launch_sweep_marked_plain(const Table& t, int2* store, int32_t sweep, InstFlags* flags,
                                     unsigned long long* tight, const EventSet& ev, int64_t inst,
                                     int dom, hipStream_t stream) {
  static_assert(kChunkWords == kBlock, "one bitmap word per thread");
  int64_t grid = (ev.words + kChunkWords - 1) / kChunkWords;
  if (grid > 2048) grid = 2048;
  if (grid <= 0) return hipSuccess;
  switch (dom) {
    case kItv: k_sweep_marked_plain<kItv><<<(unsigned)grid, kBlock, 0, stream>>>(t, store, sweep, flags, tight, ev, inst); break;
    case kItvDiv: k_sweep_marked_plain<kItvDiv><<<(unsigned)grid, kBlock, 0, stream>>>(t, store, sweep, flags, tight, ev, inst); break;
    case kBits: k_sweep_marked_plain<kBits><<<(unsigned)grid, kBlock, 0, stream>>>(t, store, sweep, flags, tight, ev, inst); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---- batched mode: device scheduler + grid-stride sweep over work items ---
// Items are (active instance, chunk of kBlock*PPL propagators).  The
// scheduler (one workgroup) retires converged or failed instances and builds
// the prefix sum of the items of those still running.
constexpr int kSchedThreads = 1024;

__global__ __launch_bounds__(kSchedThreads) void k_schedule(Subsets sub, int64_t n_props,
                                                            InstFlags* flags, int32_t n_inst,
                                                            WorkList wl, int64_t chunk,
                                                            int32_t max_sweeps,
                                                            EventSet ev, bool marked) {
  __shared__ int64_t part[kSchedThreads];
  __shared__ int32_t cnt[kSchedThreads];
  const int32_t s = *wl.sweep + 1;       // the sweep being scheduled
  const int tid = threadIdx.x;
  if (tid < kXcds) wl.xq[tid] = 0;
  if (*wl.live == 0) {                   // every instance retired earlier: n_act is already 0
    __syncthreads();
    if (tid == 0) *wl.sweep = s;
    return;
  }
  const int per = (n_inst + kSchedThreads - 1) / kSchedThreads;
  const int lo = min(n_inst, tid * per), hi = min(n_inst, lo + per);
  int64_t items = 0;
  int32_t na = 0;
  for (int k = lo; k < hi; ++k) {
    InstFlags& f = flags[k];
    bool run = f.failed_at == kNoFail && f.changed[(s - 1) & 1] == s - 1 &&
               (max_sweeps <= 0 || s < max_sweeps);
    if (!run) {
      if (f.done_at < 0) f.done_at = s;
      continue;
    }
    if (marked) {            // the chunks with marks become the items; snapshot and clear
      uint32_t* srow = ev.summary + (int64_t)(s & 1) * ev.n_inst * ev.swords + (int64_t)k * ev.swords;
      uint32_t* snap = ev.snap + (int64_t)k * ev.swords;
      const bool all = f.all_marked[s & 1] == s;    // every chunk (see k_sweep_marked)
      const int64_t chunks = (ev.words + kChunkWords - 1) / kChunkWords;
      for (int64_t j = 0; j < ev.swords; ++j) {
        const int64_t left = chunks - j * 32;
        const uint32_t v = all ? (left >= 32 ? ~0u : ((1u << left) - 1u)) : srow[j];
        snap[j] = v;
        srow[j] = 0;
        items += __builtin_popcount(v);
      }
    } else {
      const int64_t m = sub.off ? sub.off[k + 1] - sub.off[k] : n_props;
      items += (m + chunk - 1) / chunk;
    }
    ++na;
  }
  part[tid] = items;
  cnt[tid] = na;
  __syncthreads();
  // inclusive scans (Hillis-Steele; 1024 entries, runs once per sweep)
  for (int o = 1; o < kSchedThreads; o <<= 1) {
    const int64_t pv = tid >= o ? part[tid - o] : 0;
    const int32_t cv = tid >= o ? cnt[tid - o] : 0;
    __syncthreads();
    part[tid] += pv;
    cnt[tid] += cv;
    __syncthreads();
  }
  // Marked sweeps: when many chunks hold marks (typically few marks each,
  // right after a full sweep), a work item takes up to kMarkedGroupMax of an
  // instance's marked chunks, so the per-item cost (scan, barriers, mark
  // flush) is paid once for several chunks; the items still fill the chip.
  int group = 1;
  if (marked) {
    const int64_t total_chunks = part[kSchedThreads - 1];
    group = (int)max((int64_t)1, min((int64_t)kMarkedGroupMax, total_chunks / kMarkedTargetItems));
    if (group > 1) {
      int64_t gi = 0;
      for (int k = lo; k < hi; ++k) {
        const InstFlags& f = flags[k];
        if (f.done_at >= 0 && f.done_at <= s) continue;
        const uint32_t* snap = ev.snap + (int64_t)k * ev.swords;
        int64_t c = 0;
        for (int64_t j = 0; j < ev.swords; ++j) c += __builtin_popcount(snap[j]);
        gi += (c + group - 1) / group;
      }
      __syncthreads();                   // every thread has read the first scan's total
      part[tid] = gi;
      __syncthreads();
      for (int o = 1; o < kSchedThreads; o <<= 1) {
        const int64_t pv = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += pv;
        __syncthreads();
      }
      items = gi;
    }
  }
  int64_t ib = part[tid] - items;
  int32_t ia = cnt[tid] - na;
  for (int k = lo; k < hi; ++k) {
    const InstFlags& f = flags[k];
    if (f.done_at >= 0 && f.done_at <= s) continue;
    wl.act[ia] = k;
    wl.item_base[ia] = ib;
    if (marked) {
      const uint32_t* snap = ev.snap + (int64_t)k * ev.swords;
      int64_t c = 0;
      for (int64_t j = 0; j < ev.swords; ++j) c += __builtin_popcount(snap[j]);
      ib += (c + group - 1) / group;
    } else {
      const int64_t m = sub.off ? sub.off[k + 1] - sub.off[k] : n_props;
      ib += (m + chunk - 1) / chunk;
    }
    ++ia;
  }
  __syncthreads();
  if (tid == kSchedThreads - 1) {
    wl.item_base[cnt[tid]] = part[tid];
    *wl.n_act = cnt[tid];
    *wl.live = cnt[tid];
    *wl.sweep = s;
    *wl.group = group;
  }
}

hipError_t launch_schedule(const Subsets& sub, int64_t n_props, InstFlags* flags, int32_t n_inst,
                           WorkList wl, int ppw, int32_t max_sweeps, hipStream_t stream,
                           const EventSet* marked) {
  const int64_t chunk = (int64_t)kBlock * (ppw / kWave);
  k_schedule<<<1, kSchedThreads, 0, stream>>>(sub, n_props, flags, n_inst, wl, chunk, max_sweeps,
                                              marked ? *marked : EventSet{}, marked != nullptr);
  return hipGetLastError();
}

// R rounds of the PPL-unrolled body per work item (props per wave above 512):
// larger items spread the per-item costs (claim, slot search, mark flush,
// verdict) over more propagators without more registers.
template <int PPL, int D, bool NT, int R>
__global__ __launch_bounds__(kBlock) void k_sweep_batched(Table t, int2* __restrict__ store,
                                                          int64_t nvars, Subsets sub,
                                                          int64_t n_props, InstFlags* flags,
                                                          WorkList wl, unsigned long long* tight,
                                                          EventSet ev) {
  const int32_t s = *wl.sweep;
  const int32_t n_act = *wl.n_act;
  if (n_act <= 0) return;
  const int64_t total = wl.item_base[n_act];
  constexpr int64_t chunk = (int64_t)kBlock * PPL * R;
  // XCD-aware split: workgroups are dispatched round-robin over the 8 XCDs,
  // so workgroup b runs on XCD b % 8.  Each XCD takes one contiguous eighth
  // of the (instance-major) item list, and its workgroups claim that range's
  // items in order from a per-XCD counter.  Claiming (rather than a fixed
  // stride) keeps the items in flight on an XCD a narrow window even when
  // workgroups run at different speeds, so its L2 holds the stores of only
  // the few instances being swept.  The next item is claimed while the
  // current one runs.
  const int xcd = blockIdx.x % kXcds;
  const int64_t first_item = total * xcd / kXcds;
  const int64_t end_item = total * (xcd + 1) / kXcds;
  int32_t* qc = wl.xq + xcd;
  __shared__ int32_t s_next;
  __shared__ int32_t s_list[kMarkCap];
  __shared__ int32_t s_n, s_over;
  if (threadIdx.x == 0) { s_next = atomicAdd(qc, 1); s_n = 0; s_over = 0; }
  __syncthreads();
  int64_t item = first_item + __builtin_amdgcn_readfirstlane(s_next);
  __syncthreads();
  int hint = 0;
  while (item < end_item) {
    int32_t nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(qc, 1);
    // active slot holding this item: last a with item_base[a] <= item
    const int lo = uni(find_active(wl.item_base, n_act, item, hint));
    hint = lo;
    const int32_t k = uni(wl.act[lo]);
    InstFlags* f = flags + k;
    const int64_t first = (item - uni(wl.item_base[lo])) * chunk;
    InstTable it = inst_table(t, sub, k, n_props);
    it.t.xo = uni(it.t.xo); it.t.y = uni(it.t.y); it.t.z = uni(it.t.z);
    it.idx = uni(it.idx); it.m = uni(it.m);
    int2* st = store + (int64_t)k * nvars;
    Marker mk = marker_for(&ev, k, s);
    if (mk.dirty) {
      mk.list = s_list; mk.list_n = &s_n; mk.overflow = &s_over;
      mk.all_marked = &f->all_marked[(s + 1) & 1];
      mk.all_marked_s = s + 1;
    }
    bool chg = false, fail = false;
    int nt = 0;
    if (f->failed_at == kNoFail) {
      for (int rd = 0; rd < R; ++rd) {
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
          const int64_t q = first + (rd * PPL + j) * kBlock + threadIdx.x;
          if (q < it.m) eval_prop<D, NT>(it.t, st, prop_at(it.idx, q), chg, fail, nt, mk);
        }
      }
    }
    publish(f, s, chg, fail, nt, tight);
    flush_marks(mk);
    if (threadIdx.x == 0) s_next = nxt;
    __syncthreads();
    item = first_item + __builtin_amdgcn_readfirstlane(s_next);
    __syncthreads();
  }
}

template <int D, bool NT>
hipError_t launch_batched_d(const Table& t, int2* store, int64_t nvars, const Subsets& sub,
                            int64_t n_props, InstFlags* flags, WorkList wl,
                            unsigned long long* tight, int ppl, int grid, const EventSet& ev,
                            hipStream_t stream) {
#define PCX_BATCHED(P, R) \
  k_sweep_batched<P, D, NT, R><<<grid, kBlock, 0, stream>>>(t, store, nvars, sub, n_props, flags, wl, tight, ev)
  switch (ppl) {
    case 1: PCX_BATCHED(1, 1); break;
    case 2: PCX_BATCHED(2, 1); break;
    case 4: PCX_BATCHED(4, 1); break;
    case 8: PCX_BATCHED(8, 1); break;
    case 16: PCX_BATCHED(8, 2); break;
    case 32: PCX_BATCHED(8, 4); break;
    default: return hipErrorInvalidValue;
  }
#undef PCX_BATCHED
  return hipGetLastError();
}

template <bool NT>
hipError_t launch_batched_nt(const Table& t, int2* store, int64_t nvars, const Subsets& sub,
                             int64_t n_props, InstFlags* flags, WorkList wl,
                             unsigned long long* tight, int ppl, int dom, int grid,
                             const EventSet& ev, hipStream_t stream) {
  switch (dom) {
    case kItv: return launch_batched_d<kItv, NT>(t, store, nvars, sub, n_props, flags, wl, tight, ppl, grid, ev, stream);
    case kItvDiv: return launch_batched_d<kItvDiv, NT>(t, store, nvars, sub, n_props, flags, wl, tight, ppl, grid, ev, stream);
    case kBits: return launch_batched_d<kBits, NT>(t, store, nvars, sub, n_props, flags, wl, tight, ppl, grid, ev, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sweep_batched(const Table& t, int2* store, int64_t nvars, const Subsets& sub,
                                int64_t n_props, InstFlags* flags, WorkList wl,
                                unsigned long long* tight, int ppw, int dom, bool stream_table,
                                int grid, hipStream_t stream, const EventSet* ev) {
  const int ppl = ppw / kWave;
  const EventSet none{nullptr, nullptr, nullptr, nullptr, 0, 0};
  const EventSet& e = ev ? *ev : none;
  return stream_table
             ? launch_batched_nt<true>(t, store, nvars, sub, n_props, flags, wl, tight, ppl, dom, grid, e, stream)
             : launch_batched_nt<false>(t, store, nvars, sub, n_props, flags, wl, tight, ppl, dom, grid, e, stream);
}

// ---- lockstep full sweeps: GI instances side by side ---------------------
// A full sweep of many instances of one table evaluates the same
// propagators in every instance.  The lockstep sweep runs a propagator for
// GI instances at once (lane = instance slot i + GI * propagator), reading an
// instance-interleaved copy of the stores in which variable v of the group's
// GI instances is one contiguous run of GI words: its three operand gathers
// are each one 8*GI-byte run per propagator instead of GI scattered words.
// The copy is a snapshot taken before the sweep and only read during it;
// tightenings go to the stores (atomics, as in every sweep).  A sweep that
// reads the stores as they were at its start is a valid sweep (what it reads
// is at least as tight as the store at its start), and with a read-only copy
// the operand loads of several propagators are issued together.

// Interleaved copy: group g = instances [g*GI, g*GI + GI), variable v of
// slot i at word (g * nvars + v) * GI + i.  One workgroup transposes a tile
// of GI instances x (4096 / GI) variables through LDS; both sides are
// contiguous runs.
// Appends (k << 32 | v) to sub-list c for the lanes whose d is set, with one
// counter atomic per wave.  Called by every lane of the wave.
__device__ __forceinline__ void diff_append_to(const DiffSink& ds, int c, bool d, int64_t k, int64_t v, u64 val) {
  const u64 mask = __ballot(d);
  if (mask == 0) return;
  const int leader = __builtin_ctzll(mask);
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(&ds.count[c], (int)__popcll(mask));
  base = __builtin_amdgcn_readlane(base, leader);
  const int64_t at = (int64_t)base + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
  if (d && at < ds.cap) {
    ds.list[c * ds.cap + at] = ((u64)k << 32) | (uint32_t)v;
    if (ds.vals) reinterpret_cast<u64*>(ds.vals)[c * ds.cap + at] = val;
  }
}
// The same, to the workgroup's sub-list.
__device__ __forceinline__ void diff_append(const DiffSink& ds, bool d, int64_t k, int64_t v, u64 val) {
  diff_append_to(ds, blockIdx.x % kDiffLists, d, k, v, val);
}
// The base value of one variable: what two of instances 0, 1 and 2 hold
// (instance 0's when no two agree or the batch is smaller).
__device__ __forceinline__ u64 base_value(u64 a, u64 b, u64 c, int64_t n_inst) {
  if (n_inst < 3) return a;
  return (a == b || a == c) ? a : (b == c ? b : a);
}
__device__ __forceinline__ u64 pack_bounds(int32_t l, int32_t u) {
  return (u64)(uint32_t)l | ((u64)(uint32_t)u << 32);
}
// Base values of a tile's variables [v0, v0 + n) into LDS (s_base), read from
// rows 0-2 of lb / ub; the workgroups of group 0 also write them out.
__device__ __forceinline__ void tile_base(const DiffSink& ds, const int32_t* lb, const int32_t* ub,
                                          int64_t nvars, int64_t n_inst, int64_t v0, int n, bool writer,
                                          u64* s_base) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const int64_t v = v0 + t;
    if (v >= nvars) continue;
    const u64 a = pack_bounds(lb[v], ub[v]);
    const u64 b = n_inst > 1 ? pack_bounds(lb[nvars + v], ub[nvars + v]) : a;
    const u64 c = n_inst > 2 ? pack_bounds(lb[2 * nvars + v], ub[2 * nvars + v]) : a;
    const u64 r = base_value(a, b, c, n_inst);
    s_base[t] = r;
    if (writer) reinterpret_cast<u64*>(ds.base)[v] = r;
  }
  __syncthreads();
}

// Whether this workgroup lists differences: a sub-list that has already
// lost entries (count past cap) is not contended for, as the list is unused
// anyway.  (A full sub-list is still appended to, so that its count passes
// cap and records the loss.)
__device__ __forceinline__ bool diff_listing(const DiffSink& ds) {
  __shared__ int s_listing;
  if (threadIdx.x == 0)
    s_listing = ds.list != nullptr && ds.count[blockIdx.x % kDiffLists] <= ds.cap;
  __syncthreads();
  const bool r = s_listing != 0;
  __syncthreads();
  return r;
}

template <int GI>
__global__ __launch_bounds__(kBlock) void k_interleave(const u64* __restrict__ store,
                                                      u64* __restrict__ sh, int64_t nvars,
                                                      int64_t n_inst, int64_t tiles_per_group) {
  constexpr int kTV = 4096 / GI;              // variables per tile
  __shared__ u64 tile[GI][kTV + 1];
  const int64_t g = blockIdx.x / tiles_per_group;
  const int64_t v0 = (blockIdx.x - g * tiles_per_group) * kTV;
#pragma unroll
  for (int r = 0; r < 4096 / kBlock; ++r) {
    const int e = r * kBlock + threadIdx.x;
    const int i = e / kTV, j = e - i * kTV;
    const int64_t k = g * GI + i;
    tile[i][j] = (k < n_inst && v0 + j < nvars) ? store[k * nvars + v0 + j] : 0;
  }
  __syncthreads();
  u64* out = sh + (g * nvars + v0) * GI;
  const int64_t lim = min((int64_t)kTV, nvars - v0) * GI;
#pragma unroll
  for (int r = 0; r < 4096 / kBlock; ++r) {
    const int e = r * kBlock + threadIdx.x;
    if (e < lim) out[e] = tile[e % GI][e / GI];
  }
}

hipError_t launch_interleave(const int2* store, int2* shadow, int64_t nvars, int64_t n_inst, int gi,
                             hipStream_t stream) {
  if (nvars <= 0 || n_inst <= 0) return hipSuccess;
  const int64_t groups = (n_inst + gi - 1) / gi;
  const int64_t tv = 4096 / gi;
  const int64_t tpg = (nvars + tv - 1) / tv;
  const int64_t grid = groups * tpg;
  if (grid > 0x7fffffff) return hipErrorInvalidValue;
  const u64* in = reinterpret_cast<const u64*>(store);
  u64* out = reinterpret_cast<u64*>(shadow);
  switch (gi) {
    case 16: k_interleave<16><<<(unsigned)grid, kBlock, 0, stream>>>(in, out, nvars, n_inst, tpg); break;
    case 32: k_interleave<32><<<(unsigned)grid, kBlock, 0, stream>>>(in, out, nvars, n_inst, tpg); break;
    case 64: k_interleave<64><<<(unsigned)grid, kBlock, 0, stream>>>(in, out, nvars, n_inst, tpg); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Stores loaded from device memory straight into the layout a lockstep sweep
// reads: one pass reads lb / ub (once, non-temporal), writes the packed
// stores row by row, the interleaved copy through the same LDS tile as
// k_interleave, and the per-instance emptiness (empty zeroed first).
template <int GI>
__global__ __launch_bounds__(kBlock) void k_ingest_interleave(const int32_t* __restrict__ lb,
                                                             const int32_t* __restrict__ ub,
                                                             u64* __restrict__ store, u64* __restrict__ sh,
                                                             int64_t nvars, int64_t n_inst,
                                                             int64_t tiles_per_group,
                                                             uint8_t* __restrict__ empty, DiffSink ds) {
  constexpr int kTV = 4096 / GI;
  __shared__ u64 tile[GI][kTV + 1];
  const int64_t g = blockIdx.x / tiles_per_group;
  const int64_t v0 = (blockIdx.x - g * tiles_per_group) * kTV;
  const bool listing = diff_listing(ds);
  __shared__ u64 s_base[kTV];
  if (listing) tile_base(ds, lb, ub, nvars, n_inst, v0, kTV, g == 0, s_base);
#pragma unroll
  for (int r = 0; r < 4096 / kBlock; ++r) {
    const int e = r * kBlock + threadIdx.x;
    const int i = e / kTV, j = e - i * kTV;
    const int64_t k = g * GI + i;
    u64 val = 0;
    bool dif = false;
    if (k < n_inst && v0 + j < nvars) {
      const int64_t at = k * nvars + v0 + j;
      const int32_t l = __builtin_nontemporal_load(lb + at);
      const int32_t u = __builtin_nontemporal_load(ub + at);
      val = (u64)(uint32_t)l | ((u64)(uint32_t)u << 32);     // int2 {lb, ub}
      store[at] = val;
      if (l > u) empty[k] = 1;
      if (listing) dif = val != s_base[j];
    }
    if (listing) diff_append(ds, dif, k, v0 + j, val);
    tile[i][j] = val;
  }
  __syncthreads();
  u64* out = sh + (g * nvars + v0) * GI;
  const int64_t lim = min((int64_t)kTV, nvars - v0) * GI;
#pragma unroll
  for (int r = 0; r < 4096 / kBlock; ++r) {
    const int e = r * kBlock + threadIdx.x;
    if (e < lim) out[e] = tile[e % GI][e / GI];
  }
}

// The inverse of k_interleave: rows of the stores from the interleaved copy
// (a batch whose reset image lives only in the interleaved copy).
template <int GI>
__global__ __launch_bounds__(kBlock) void k_deinterleave(const u64* __restrict__ sh, u64* __restrict__ dst,
                                                        int64_t nvars, int64_t n_inst,
                                                        int64_t tiles_per_group) {
  constexpr int kTV = 4096 / GI;
  __shared__ u64 tile[GI][kTV + 1];
  const int64_t g = blockIdx.x / tiles_per_group;
  const int64_t v0 = (blockIdx.x - g * tiles_per_group) * kTV;
  const u64* in = sh + (g * nvars + v0) * GI;
  const int64_t lim = min((int64_t)kTV, nvars - v0) * GI;
#pragma unroll
  for (int r = 0; r < 4096 / kBlock; ++r) {
    const int e = r * kBlock + threadIdx.x;
    if (e < lim) tile[e % GI][e / GI] = in[e];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4096 / kBlock; ++r) {
    const int e = r * kBlock + threadIdx.x;
    const int i = e / kTV, j = e - i * kTV;
    const int64_t k = g * GI + i;
    if (k < n_inst && v0 + j < nvars) dst[k * nvars + v0 + j] = tile[i][j];
  }
}

typedef int32_t v4i __attribute__((ext_vector_type(4)));

// The same with 16-byte accesses: four variables of one instance per thread
// on the way in (nvars % 4 == 0, 16-byte aligned lb / ub), two slots of the
// interleaved copy per thread on the way out.
template <int GI>
__global__ __launch_bounds__(kBlock) void k_ingest_interleave4(const v4i* __restrict__ lb,
                                                              const v4i* __restrict__ ub,
                                                              v4i* __restrict__ store, v4i* __restrict__ sh,
                                                              int64_t nvars, int64_t n_inst,
                                                              int64_t tiles_per_group,
                                                              uint8_t* __restrict__ empty, DiffSink ds) {
  constexpr int kTV = 4096 / GI;
  constexpr int kQ = kTV / 4;                  // 4-variable groups per tile row
  __shared__ u64 tile[GI][kTV + 1];
  const int64_t g = blockIdx.x / tiles_per_group;
  const int64_t v0 = (blockIdx.x - g * tiles_per_group) * kTV;
  const bool listing = diff_listing(ds);
  __shared__ u64 s_base[kTV];
  if (listing)
    tile_base(ds, reinterpret_cast<const int32_t*>(lb), reinterpret_cast<const int32_t*>(ub), nvars, n_inst,
              v0, kTV, g == 0, s_base);
#pragma unroll
  for (int r = 0; r < 1024 / kBlock; ++r) {
    const int e = r * kBlock + threadIdx.x;
    const int i = e / kQ, j = (e - i * kQ) * 4;
    const int64_t k = g * GI + i;
    u64 w[4] = {0, 0, 0, 0};
    int dif = 0;                               // bit q: variable j + q differs from the base
    if (k < n_inst && v0 + j < nvars) {
      const int64_t at4 = (k * nvars + v0 + j) / 4;
      const v4i a = __builtin_nontemporal_load(lb + at4);
      const v4i b = __builtin_nontemporal_load(ub + at4);
      if (listing)
        dif = (pack_bounds(a.x, b.x) != s_base[j]) | ((pack_bounds(a.y, b.y) != s_base[j + 1]) << 1) |
              ((pack_bounds(a.z, b.z) != s_base[j + 2]) << 2) | ((pack_bounds(a.w, b.w) != s_base[j + 3]) << 3);
      const v4i p0 = {a.x, b.x, a.y, b.y};
      const v4i p1 = {a.z, b.z, a.w, b.w};
      store[2 * at4] = p0;
      store[2 * at4 + 1] = p1;
      if ((a.x > b.x) | (a.y > b.y) | (a.z > b.z) | (a.w > b.w)) empty[k] = 1;
      w[0] = (u64)(uint32_t)a.x | ((u64)(uint32_t)b.x << 32);
      w[1] = (u64)(uint32_t)a.y | ((u64)(uint32_t)b.y << 32);
      w[2] = (u64)(uint32_t)a.z | ((u64)(uint32_t)b.z << 32);
      w[3] = (u64)(uint32_t)a.w | ((u64)(uint32_t)b.w << 32);
    }
    if (listing && __ballot(dif != 0)) {
#pragma unroll
      for (int q = 0; q < 4; ++q) diff_append(ds, (dif >> q) & 1, k, v0 + j + q, w[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) tile[i][j + q] = w[q];
  }
  __syncthreads();
  u64* out = reinterpret_cast<u64*>(sh) + (g * nvars + v0) * GI;
  const int64_t lim = min((int64_t)kTV, nvars - v0) * GI;      // even: GI is
#pragma unroll
  for (int r = 0; r < 2048 / kBlock; ++r) {
    const int e = 2 * (r * kBlock + threadIdx.x);
    if (e < lim) {
      const u64 lo = tile[e % GI][e / GI], hi = tile[(e + 1) % GI][(e + 1) / GI];
      const v4i v = {(int32_t)lo, (int32_t)(lo >> 32), (int32_t)hi, (int32_t)(hi >> 32)};
      *reinterpret_cast<v4i*>(out + e) = v;
    }
  }
}

hipError_t launch_ingest_interleave(const int32_t* lb, const int32_t* ub, int2* store, int2* shadow,
                                    int64_t nvars, int64_t n_inst, int gi, uint8_t* empty,
                                    const DiffSink& ds, hipStream_t stream) {
  if (nvars <= 0 || n_inst <= 0) return hipSuccess;
  const int64_t groups = (n_inst + gi - 1) / gi;
  const int64_t tv = 4096 / gi;
  const int64_t tpg = (nvars + tv - 1) / tv;
  const int64_t grid = groups * tpg;
  if (grid > 0x7fffffff) return hipErrorInvalidValue;
  const bool vec = nvars % 4 == 0 &&
                   ((reinterpret_cast<uintptr_t>(lb) | reinterpret_cast<uintptr_t>(ub) |
                     reinterpret_cast<uintptr_t>(store) | reinterpret_cast<uintptr_t>(shadow)) & 15) == 0;
  if (vec) {
    const v4i* l4 = reinterpret_cast<const v4i*>(lb);
    const v4i* u4 = reinterpret_cast<const v4i*>(ub);
    v4i* s4 = reinterpret_cast<v4i*>(store);
    v4i* o4 = reinterpret_cast<v4i*>(shadow);
    switch (gi) {
      case 16: k_ingest_interleave4<16><<<(unsigned)grid, kBlock, 0, stream>>>(l4, u4, s4, o4, nvars, n_inst, tpg, empty, ds); break;
      case 32: k_ingest_interleave4<32><<<(unsigned)grid, kBlock, 0, stream>>>(l4, u4, s4, o4, nvars, n_inst, tpg, empty, ds); break;
      case 64: k_ingest_interleave4<64><<<(unsigned)grid, kBlock, 0, stream>>>(l4, u4, s4, o4, nvars, n_inst, tpg, empty, ds); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  u64* st = reinterpret_cast<u64*>(store);
  u64* out = reinterpret_cast<u64*>(shadow);
  switch (gi) {
    case 16: k_ingest_interleave<16><<<(unsigned)grid, kBlock, 0, stream>>>(lb, ub, st, out, nvars, n_inst, tpg, empty, ds); break;
    case 32: k_ingest_interleave<32><<<(unsigned)grid, kBlock, 0, stream>>>(lb, ub, st, out, nvars, n_inst, tpg, empty, ds); break;
    case 64: k_ingest_interleave<64><<<(unsigned)grid, kBlock, 0, stream>>>(lb, ub, st, out, nvars, n_inst, tpg, empty, ds); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_deinterleave(const int2* shadow, int2* dst, int64_t nvars, int64_t n_inst, int gi,
                               hipStream_t stream) {
  if (nvars <= 0 || n_inst <= 0) return hipSuccess;
  const int64_t groups = (n_inst + gi - 1) / gi;
  const int64_t tv = 4096 / gi;
  const int64_t tpg = (nvars + tv - 1) / tv;
  const int64_t grid = groups * tpg;
  if (grid > 0x7fffffff) return hipErrorInvalidValue;
  const u64* in = reinterpret_cast<const u64*>(shadow);
  u64* out = reinterpret_cast<u64*>(dst);
  switch (gi) {
    case 16: k_deinterleave<16><<<(unsigned)grid, kBlock, 0, stream>>>(in, out, nvars, n_inst, tpg); break;
    case 32: k_deinterleave<32><<<(unsigned)grid, kBlock, 0, stream>>>(in, out, nvars, n_inst, tpg); break;
    case 64: k_deinterleave<64><<<(unsigned)grid, kBlock, 0, stream>>>(in, out, nvars, n_inst, tpg); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Tightened (instance, variable) pairs are deferred to an LDS list and
// marked by the whole workgroup at the end of the work item; a pair that
// does not fit marks its whole instance for the next sweep instead (its
// all_marked word).
constexpr int kLockMarkCap = 1024;
constexpr int kLockWordsPerWave = 4;       // 32-propagator table words per wave per item
constexpr int kLockUnroll = 2;             // propagators whose operands are loaded together

// Waves per SIMD the lockstep sweep is compiled for.  It waits on memory
// for about half its cycles, so occupancy pays even at the price of a few
// spilled registers in its cold paths: measured best at 6 for intervals
// (config 2: 2.75 -> 2.51 ms) and 7 for bitsets (config 3: 7.5 -> 6.4 ms).
// The division variant's registers do not fit such a bound.
template <int D>
constexpr int lock_min_waves() { return D == kItv ? 6 : (D == kBits ? 7 : 1); }

template <int D, int GI>
__global__ __launch_bounds__(kBlock, lock_min_waves<D>()) void k_sweep_lockstep(Table t, int2* __restrict__ store,
                                                          const int2* __restrict__ shadow, int64_t nvars,
                                                          int32_t n_inst,
                                                          const uint32_t* __restrict__ member,
                                                          int64_t words, InstFlags* flags,
                                                          WorkList wl, unsigned long long* tight,
                                                          EventSet ev) {
  constexpr int kPpi = kWave / GI;                           // propagators per wave step
  constexpr int kWpi = kLockWordsPerWave * (kBlock / kWave);  // table words per item
  const int32_t s = *wl.sweep;
  const int64_t tw = (t.n + 31) / 32;
  const int64_t ipg = (tw + kWpi - 1) / kWpi;
  const int64_t groups = ((int64_t)n_inst + GI - 1) / GI;
  const int64_t total = groups * ipg;
  // XCD-aware: each XCD takes a contiguous range of (group-major) items and
  // claims them in order, so its workgroups share the few groups in flight.
  const int xcd = blockIdx.x % kXcds;
  const int64_t first_item = total * xcd / kXcds;
  const int64_t end_item = total * (xcd + 1) / kXcds;
  int32_t* qc = wl.xq + xcd;
  __shared__ int32_t s_next;
  __shared__ int32_t s_n;
  __shared__ int32_t s_mk[kLockMarkCap], s_mv[kLockMarkCap];
  __shared__ unsigned long long s_qbuf[kQStage];
  __shared__ int32_t s_qn;
  const QueueStage qs{s_qbuf, &s_qn};
  if (threadIdx.x == 0) { s_next = atomicAdd(qc, 1); s_n = 0; s_qn = 0; }
  __syncthreads();
  int64_t item = first_item + __builtin_amdgcn_readfirstlane(s_next);
  __syncthreads();
  const int lane = lane_id();
  const int slot = lane % GI, sub = lane / GI;
  const int wave = threadIdx.x / kWave;
  while (item < end_item) {
    int32_t nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(qc, 1);
    const int64_t g = uni(item / ipg);
    const int64_t w0 = (item - g * ipg) * kWpi + wave * kLockWordsPerWave;
    const int64_t k = g * GI + slot;
    // Per-lane flags carried across the item's rounds are ints: as lane
    // masks every round would merge them under the exec mask.
    bool run = false;
    int am = 1;
    InstFlags* f = flags + (k < n_inst ? k : 0);
    if (k < n_inst) {
      run = !(f->done_at >= 0 && f->done_at <= s) && f->failed_at == kNoFail;
      am = !ev.dirty || f->all_marked[(s + 1) & 1] == s + 1;
    }
    int2* sk = store + k * nvars;
    // The group's copy starts at a wave-uniform address and spans less than
    // 4 GiB (the host checks nvars * GI * 8), so an operand is one 32-bit
    // offset from it.
    const char* sgb = reinterpret_cast<const char*>(shadow + g * nvars * GI);
    const uint32_t slot_off = (uint32_t)slot * 8u;
    const uint32_t* mrow = member ? member + k * words : nullptr;
    int chg = 0, fail = 0;
    int nt = 0;
    auto mark = [&](int32_t v) {
      if (am) return;
      const int i = atomicAdd(&s_n, 1);
      if (i < kLockMarkCap) { s_mk[i] = (int32_t)k; s_mv[i] = v; }
      else {
        f->all_marked[(s + 1) & 1] = s + 1;
        am = 1;
        if (ev.q) ev.qctl[kQOver + ((s + 1) & 1)] = 1;   // these marks are in no queue
      }
    };
    // The member word and the table entries of the next word are loaded
    // while this word is swept.
    const int64_t wend = min(w0 + kLockWordsPerWave, tw);
    uint32_t m_n = 0, txo_n = 0;
    int32_t ty_n = 0, tz_n = 0;
    auto fetch = [&](int64_t w) {
      m_n = 0;
      if (run) {
        m_n = mrow ? mrow[w] : ~0u;
        const int64_t left = t.n - w * 32;
        if (left < 32) m_n &= (1u << left) - 1u;
      }
      const int64_t p = w * 32 + (lane & 31);
      const bool pin = p < t.n;
      txo_n = pin ? t.xo[p] : 0u;
      ty_n = pin ? t.y[p] : 0;
      tz_n = pin ? t.z[p] : 0;
    };
    if (w0 < wend) fetch(w0);
    for (int64_t w = w0; w < wend; ++w) {
      const uint32_t m = m_n, txo = txo_n;
      const int32_t ty = ty_n, tz = tz_n;
      if (w + 1 < wend) fetch(w + 1);
      if (!__ballot(m != 0)) continue;
      // The rounds of this word, for one opcode OPC (a compile-time constant:
      // the propagator is branch-free up to a tightening, which is rare) or
      // for mixed opcodes (OPC < 0, read per propagator).
      auto rounds = [&](auto opc) {
        constexpr int OPC = decltype(opc)::value;
        constexpr bool kFlat = OPC >= 0 && (D == kBits || OPC < kTdiv || OPC > kEdiv);
        constexpr int kStep = kPpi * kLockUnroll;
#pragma unroll 1
        for (int j0 = 0; j0 < 32; j0 += kStep) {
          uint32_t xo[kLockUnroll];
          int32_t y[kLockUnroll], z[kLockUnroll];
          u64 wx[kLockUnroll], wy[kLockUnroll], wz[kLockUnroll];
#pragma unroll
          for (int u = 0; u < kLockUnroll; ++u) {
            const int src = j0 + u * kPpi + sub;
            if constexpr (GI == kWave) {
              xo[u] = (uint32_t)__builtin_amdgcn_readlane((int)txo, j0 + u);
              y[u] = __builtin_amdgcn_readlane(ty, j0 + u);
              z[u] = __builtin_amdgcn_readlane(tz, j0 + u);
            } else {
              xo[u] = (uint32_t)__shfl((int)txo, src, kWave);
              y[u] = __shfl(ty, src, kWave);
              z[u] = __shfl(tz, src, kWave);
            }
          }
          // Every lane loads (its table entry names valid variables and its
          // slot lies inside the padded copy), so the loads issue together.
#pragma unroll
          for (int u = 0; u < kLockUnroll; ++u) {
            const int32_t x = (int32_t)(xo[u] & kVarMask);
            wx[u] = *reinterpret_cast<const u64*>(sgb + ((uint32_t)x * (uint32_t)(GI * 8) + slot_off));
            wy[u] = *reinterpret_cast<const u64*>(sgb + ((uint32_t)y[u] * (uint32_t)(GI * 8) + slot_off));
            wz[u] = *reinterpret_cast<const u64*>(sgb + ((uint32_t)z[u] * (uint32_t)(GI * 8) + slot_off));
          }
          bool on[kLockUnroll];
#pragma unroll
          for (int u = 0; u < kLockUnroll; ++u) on[u] = (m >> (j0 + u * kPpi + sub)) & 1u;
          if constexpr (kFlat) {
            // One opcode, no division: the propagators are evaluated
            // branch-free, and what is rare (a bound tightens, a store is
            // empty) runs under one wave-uniform branch per round, so the
            // common round carries no per-lane control flow.
            using Val = std::conditional_t<D == kBits, u64, int2>;
            Val nx[kLockUnroll], ny[kLockUnroll], nz[kLockUnroll];
            bool need[kLockUnroll];
            bool any = false;
#pragma unroll
            for (int u = 0; u < kLockUnroll; ++u) {
              bool empty_in, tight;
              if constexpr (D == kBits) {
                deduce_bits(OPC, wx[u], wy[u], wz[u], nx[u], ny[u], nz[u]);
                empty_in = !wx[u] | !wy[u] | !wz[u];
                // A deduced empty word is a tightening too.
                tight = ((wx[u] & nx[u]) != wx[u]) | ((wy[u] & ny[u]) != wy[u]) |
                        ((wz[u] & nz[u]) != wz[u]);
              } else {
                const int2 X = as_itv(wx[u]), Y = as_itv(wy[u]), Z = as_itv(wz[u]);
                deduce<D == kItvDiv>(OPC, X, Y, Z, nx[u], ny[u], nz[u]);
                // ub - lb saturated is negative iff the interval is empty.
                empty_in = (ssub(X.y, X.x) | ssub(Y.y, Y.x) | ssub(Z.y, Z.x)) < 0;
                // Bounds only move inward, so an interval differs iff it
                // tightens; a deduced empty interval of a non-empty one
                // tightens too.
                tight = ((itv_bits(nx[u]) ^ wx[u]) | (itv_bits(ny[u]) ^ wy[u]) |
                         (itv_bits(nz[u]) ^ wz[u])) != 0;
              }
              need[u] = on[u] & (empty_in | tight);
              any |= need[u];
            }
            if (__builtin_expect(__ballot(any) != 0, 0)) {
#pragma unroll
              for (int u = 0; u < kLockUnroll; ++u) {
                if (!need[u]) continue;
                const int32_t x = (int32_t)(xo[u] & kVarMask);
                bool tx, ty2, tz2, c = false, fl = false;
                if constexpr (D == kBits) {
                  if (!wx[u] || !wy[u] || !wz[u]) { fail = 1; continue; }
                  u64* sk64 = reinterpret_cast<u64*>(sk);
                  tx = meet_word<__HIP_MEMORY_SCOPE_AGENT>(sk64, x, wx[u], nx[u], c, fl, nt);
                  ty2 = meet_word<__HIP_MEMORY_SCOPE_AGENT>(sk64, y[u], wy[u], ny[u], c, fl, nt);
                  tz2 = meet_word<__HIP_MEMORY_SCOPE_AGENT>(sk64, z[u], wz[u], nz[u], c, fl, nt);
                } else {
                  const int2 X = as_itv(wx[u]), Y = as_itv(wy[u]), Z = as_itv(wz[u]);
                  if (X.x > X.y || Y.x > Y.y || Z.x > Z.y) { fail = 1; continue; }
                  if (nx[u].x > nx[u].y || ny[u].x > ny[u].y || nz[u].x > nz[u].y) fl = true;
                  tx = meet_itv<__HIP_MEMORY_SCOPE_AGENT>(sk, x, X, nx[u], c, fl, nt);
                  ty2 = meet_itv<__HIP_MEMORY_SCOPE_AGENT>(sk, y[u], Y, ny[u], c, fl, nt);
                  tz2 = meet_itv<__HIP_MEMORY_SCOPE_AGENT>(sk, z[u], Z, nz[u], c, fl, nt);
                }
                chg |= c;
                fail |= fl;
                if (tx) mark(x);
                if (ty2) mark(y[u]);
                if (tz2) mark(z[u]);
              }
            }
          } else {
#pragma unroll
            for (int u = 0; u < kLockUnroll; ++u) {
              if (!on[u]) continue;
              const int op = OPC >= 0 ? OPC : (int)(xo[u] >> kOpShift);
              const int32_t x = (int32_t)(xo[u] & kVarMask);
              bool tx, ty2, tz2, c = false, fl = false;
              if constexpr (D == kBits) {
                if (!wx[u] || !wy[u] || !wz[u]) { fail = 1; continue; }
                u64 nx, ny, nz;
                deduce_bits(op, wx[u], wy[u], wz[u], nx, ny, nz);
                u64* sk64 = reinterpret_cast<u64*>(sk);
                tx = meet_word<__HIP_MEMORY_SCOPE_AGENT>(sk64, x, wx[u], nx, c, fl, nt);
                ty2 = meet_word<__HIP_MEMORY_SCOPE_AGENT>(sk64, y[u], wy[u], ny, c, fl, nt);
                tz2 = meet_word<__HIP_MEMORY_SCOPE_AGENT>(sk64, z[u], wz[u], nz, c, fl, nt);
              } else {
                const int2 X = as_itv(wx[u]), Y = as_itv(wy[u]), Z = as_itv(wz[u]);
                if (X.x > X.y || Y.x > Y.y || Z.x > Z.y) { fail = 1; continue; }
                int2 nx, ny, nz;
                deduce<D == kItvDiv>(op, X, Y, Z, nx, ny, nz);
                if (nx.x > nx.y || ny.x > ny.y || nz.x > nz.y) fl = true;
                tx = meet_itv<__HIP_MEMORY_SCOPE_AGENT>(sk, x, X, nx, c, fl, nt);
                ty2 = meet_itv<__HIP_MEMORY_SCOPE_AGENT>(sk, y[u], Y, ny, c, fl, nt);
                tz2 = meet_itv<__HIP_MEMORY_SCOPE_AGENT>(sk, z[u], Z, nz, c, fl, nt);
              }
              chg |= c;
              fail |= fl;
              if (tx) mark(x);
              if (ty2) mark(y[u]);
              if (tz2) mark(z[u]);
            }
          }
        }
      };
      // One opcode in the whole word (the table is sorted by opcode)?
      const int op0 = __builtin_amdgcn_readfirstlane((int)(txo >> kOpShift));
      const bool pin = w * 32 + (lane & 31) < t.n;
      const bool one_op = !__ballot(lane < 32 && pin && (int)(txo >> kOpShift) != op0);
      using std::integral_constant;
      if (!one_op) {
        rounds(integral_constant<int, -1>{});
      } else if constexpr (D == kBits) {
        switch (op0) {
          case kBLeq: rounds(integral_constant<int, kBLeq>{}); break;
          case kBAdd: rounds(integral_constant<int, kBAdd>{}); break;
          case kBNeq: rounds(integral_constant<int, kBNeq>{}); break;
          case kBEqr: rounds(integral_constant<int, kBEqr>{}); break;
          default: rounds(integral_constant<int, -1>{}); break;
        }
      } else {
        switch (op0) {
          case kAdd: rounds(integral_constant<int, kAdd>{}); break;
          case kMul: rounds(integral_constant<int, kMul>{}); break;
          case kMin: rounds(integral_constant<int, kMin>{}); break;
          case kMax: rounds(integral_constant<int, kMax>{}); break;
          case kEq: rounds(integral_constant<int, kEq>{}); break;
          case kLeq: rounds(integral_constant<int, kLeq>{}); break;
          default: rounds(integral_constant<int, -1>{}); break;   // division: per propagator
        }
      }
    }
    if (chg) f->changed[s & 1] = s;
    if (fail) f->failed_at = s;
    const int sum = wave_sum(nt);
    if (lane == 0 && sum)
      atomicAdd(&tight[(blockIdx.x * 4 + wave) & (kTightSlots - 1)], (unsigned long long)sum);
    __syncthreads();                            // the item's marks are listed
    const int n = min(s_n, kLockMarkCap);
    if (ev.q) {
      for (int i = threadIdx.x; i < n; i += kBlock)
        mark_var_q(marker_for(&ev, s_mk[i], s), ev, qs, s_mk[i], s_mv[i], s + 1);
      __syncthreads();                          // the item's new marks are staged
      flush_stage(qs, ev, (s + 1) & 1);
    } else {
      for (int i = threadIdx.x; i < n; i += kBlock) mark_var(marker_for(&ev, s_mk[i], s), s_mv[i]);
    }
    __syncthreads();                            // every thread has read s_n and the list
    if (threadIdx.x == 0) { s_next = nxt; s_n = 0; }
    __syncthreads();
    item = first_item + __builtin_amdgcn_readfirstlane(s_next);
    __syncthreads();
  }
}

// Exclusive scan of v over a workgroup of BT threads; returns the offset,
// *total = the sum.  s_scan holds BT / kWave entries.
template <int BT>
__device__ __forceinline__ int block_scan(int v, int32_t* s_scan, int* total) {
  const int lane = lane_id();
  int incl = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int u = __shfl_up(incl, o, kWave);
    if (lane >= o) incl += u;
  }
  __syncthreads();
  const int wv = threadIdx.x / kWave;
  if (lane == kWave - 1) s_scan[wv] = incl;
  __syncthreads();
  int off = incl - v, t = 0;
#pragma unroll
  for (int i = 0; i < BT / kWave; ++i) {
    const int u = s_scan[i];
    if (i < wv) off += u;
    t += u;
  }
  *total = t;
  return off;
}

// ---- queued marked sweeps --------------------------------------------------
// After a lockstep full sweep of a batch, the marks of each sweep are also
// listed in a queue plane (mark_var_q).  A turn checks the plane of the next
// sweep: empty means no propagator of any instance is marked, which by the
// marked-sweep invariant is every instance's fixpoint; a plane that lost
// marks stops the queue passes (the host resumes with bitmap sweeps); else
// the turn clears the other plane for the sweep's own marks and advances.
// One workgroup, a thread per sub-list.
__global__ __launch_bounds__(kQLists) void k_queue_turn(int32_t* qctl, int32_t max_sweeps) {
  if (qctl[kQStop]) return;                       // uniform: every thread reads the same word
  const int32_t s = qctl[kQSweep] + 1;
  const int cu = s & 1;
  const bool marked = __syncthreads_or(qctl[kQTails + cu * kQLists + threadIdx.x] != 0);
  if (qctl[kQOver + cu]) {
    if (threadIdx.x == 0) { qctl[kQStop] = 2; qctl[kQAt] = s; }
    return;
  }
  if (!marked || (max_sweeps > 0 && s >= max_sweeps)) {
    if (threadIdx.x == 0) { qctl[kQStop] = 1; qctl[kQAt] = s; }
    return;
  }
  qctl[kQTails + (cu ^ 1) * kQLists + threadIdx.x] = 0;
  if (threadIdx.x == 0) { qctl[kQOver + (cu ^ 1)] = 0; qctl[kQSweep] = s; }
}

hipError_t launch_queue_turn(int32_t* qctl, int32_t max_sweeps, hipStream_t stream) {
  k_queue_turn<<<1, kQLists, 0, stream>>>(qctl, max_sweeps);
  return hipGetLastError();
}

// One marked sweep over its queue plane: thread i takes entry i, clears its
// bit, evaluates the propagator on the instance's store and lists what it
// tightened; the workgroup then marks those variables' adjacency (queued)
// for the next sweep.  Entries of failed instances are skipped.
constexpr int kQueueList = 3 * kBlock;     // at most three tightened variables per entry
template <int D>
__global__ __launch_bounds__(kBlock) void k_sweep_queue(Table t, int2* __restrict__ store, int64_t nvars,
                                                      InstFlags* flags, unsigned long long* tight,
                                                      EventSet ev) {
  static_assert(kQLists == kBlock, "one thread per queue sub-list");
  const int32_t* qctl = ev.qctl;
  if (qctl[kQStop]) return;
  const int32_t s = qctl[kQSweep];
  const int cu = s & 1;
  const unsigned long long* plane = ev.q + (int64_t)cu * kQLists * ev.qcap;
  uint32_t* cur = ev.dirty + (int64_t)cu * ev.n_inst * ev.words;
  __shared__ int32_t s_k[kQueueList], s_v[kQueueList];
  __shared__ int32_t s_n;
  __shared__ unsigned long long s_qbuf[kQStage];
  __shared__ int32_t s_qn;
  __shared__ int32_t s_scan[kBlock / kWave];
  __shared__ int32_t s_first[kQLists + 1];          // the plane's entries in sub-list order
  const QueueStage qs{s_qbuf, &s_qn};
  if (threadIdx.x == 0) { s_n = 0; s_qn = 0; }
  int n32;
  s_first[threadIdx.x] =
      block_scan<kBlock>((int)min<int64_t>(qctl[kQTails + cu * kQLists + threadIdx.x], ev.qcap), s_scan, &n32);
  if (threadIdx.x == 0) s_first[kQLists] = n32;
  __syncthreads();
  const int64_t n = n32;
  int nt_all = 0;
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += (int64_t)gridDim.x * kBlock) {
    const int64_t i = base + threadIdx.x;
    if (i < n) {
      int lo = 0, hi = kQLists;                   // sub-list c with s_first[c] <= i < s_first[c + 1]
      while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (s_first[mid] <= i) lo = mid; else hi = mid;
      }
      const unsigned long long en = plane[(int64_t)lo * ev.qcap + (i - s_first[lo])];
      const int32_t k = (int32_t)(en >> 32), p = (int32_t)(uint32_t)en;
      InstFlags* f = flags + k;
      if (f->failed_at == kNoFail) {
        atomicAnd(&cur[(int64_t)k * ev.words + (p >> 5)], ~(1u << (p & 31)));
        int2* st = store + (int64_t)k * nvars;
        const uint32_t xo = t.xo[p];
        const int32_t y = t.y[p], z = t.z[p];
        const int32_t x = (int32_t)(xo & kVarMask);
        const u64 wx = load_word<__HIP_MEMORY_SCOPE_AGENT>(st + x);
        const u64 wy = load_word<__HIP_MEMORY_SCOPE_AGENT>(st + y);
        const u64 wz = load_word<__HIP_MEMORY_SCOPE_AGENT>(st + z);
        bool chg = false, fail = false, tx = false, ty = false, tz = false;
        int nt = 0;
        if constexpr (D == kBits) {
          u64* ws = reinterpret_cast<u64*>(st);
          if (!wx || !wy || !wz) {
            fail = true;
          } else {
            u64 nx, ny, nz;
            deduce_bits((int)(xo >> kOpShift), wx, wy, wz, nx, ny, nz);
            tx = meet_word<__HIP_MEMORY_SCOPE_AGENT>(ws, x, wx, nx, chg, fail, nt);
            ty = meet_word<__HIP_MEMORY_SCOPE_AGENT>(ws, y, wy, ny, chg, fail, nt);
            tz = meet_word<__HIP_MEMORY_SCOPE_AGENT>(ws, z, wz, nz, chg, fail, nt);
          }
        } else {
          const int2 X = as_itv(wx), Y = as_itv(wy), Z = as_itv(wz);
          if (X.x > X.y || Y.x > Y.y || Z.x > Z.y) {
            fail = true;
          } else {
            int2 nx, ny, nz;
            deduce<D == kItvDiv>((int)(xo >> kOpShift), X, Y, Z, nx, ny, nz);
            if (nx.x > nx.y || ny.x > ny.y || nz.x > nz.y) fail = true;
            tx = meet_itv<__HIP_MEMORY_SCOPE_AGENT>(st, x, X, nx, chg, fail, nt);
            ty = meet_itv<__HIP_MEMORY_SCOPE_AGENT>(st, y, Y, ny, chg, fail, nt);
            tz = meet_itv<__HIP_MEMORY_SCOPE_AGENT>(st, z, Z, nz, chg, fail, nt);
          }
        }
        const int32_t vs[3] = {x, y, z};
        const bool ts[3] = {tx, ty, tz};
        for (int j = 0; j < 3; ++j)
          if (ts[j]) {
            const int at = atomicAdd(&s_n, 1);     // < kQueueList: three per entry at most
            s_k[at] = k;
            s_v[at] = vs[j];
          }
        if (chg) f->changed[s & 1] = s;
        if (fail) f->failed_at = s;
        atomicAdd((unsigned long long*)&f->marked_evals, 1ull);
        nt_all += nt;
      }
    }
    __syncthreads();                               // the pass's tightened variables are listed
    const int nl = s_n;
    for (int j = threadIdx.x; j < nl; j += kBlock) {
      const int64_t k = s_k[j];
      if (flags[k].failed_at == kNoFail) mark_var_q(marker_for(&ev, k, s), ev, qs, k, s_v[j], s + 1);
    }
    __syncthreads();                               // every thread has read the list; marks staged
    if (threadIdx.x == 0) s_n = 0;
    flush_stage(qs, ev, (s + 1) & 1);              // ends with a barrier
  }
  const int sum = wave_sum(nt_all);
  if (lane_id() == 0 && sum)
    atomicAdd(&tight[(blockIdx.x * 4 + (threadIdx.x >> 6)) & (kTightSlots - 1)], (unsigned long long)sum);
}

// The marks of instance k for sweep s, from its bitmap row (or, when its
// all_marked word says every member propagator is marked, from its member
// row, whose bits are then set in the bitmap too) into queue plane s & 1:
// a block scan of the popcounts and one reservation per workgroup.
__global__ __launch_bounds__(kBlock) void k_queue_from_bitmap(EventSet ev, InstFlags* flags, int64_t k,
                                                            int32_t s, int64_t n_props) {
  const int pl = s & 1;
  uint32_t* row = ev.dirty + (int64_t)pl * ev.n_inst * ev.words + k * ev.words;
  const uint32_t* mrow = ev.member ? ev.member + k * ev.words : nullptr;
  const bool all = flags[k].all_marked[pl] == s;
  unsigned long long* list = queue_list(ev, pl);
  __shared__ int32_t s_scan[kBlock / kWave];
  __shared__ int32_t s_base;
  for (int64_t w0 = (int64_t)blockIdx.x * kBlock; w0 < ev.words; w0 += (int64_t)gridDim.x * kBlock) {
    const int64_t w = w0 + threadIdx.x;
    uint32_t bits = 0;
    if (w < ev.words) {
      if (all) {
        bits = mrow ? mrow[w] : ~0u;
        const int64_t left = n_props - w * 32;
        if (left < 32) bits &= (1u << left) - 1u;
        if (bits) atomicOr(&row[w], bits);
      } else {
        bits = row[w];
      }
    }
    int total;
    int off = block_scan<kBlock>(__builtin_popcount(bits), s_scan, &total);
    if (threadIdx.x == 0) s_base = total ? atomicAdd(queue_tail(ev, pl), total) : 0;
    __syncthreads();
    for (uint32_t b = bits; b; b &= b - 1, ++off) {
      const int64_t at = (int64_t)s_base + off;
      const unsigned long long en = ((unsigned long long)k << 32) | (uint32_t)(w * 32 + __builtin_ctz(b));
      if (at < ev.qcap) list[at] = en;
      else ev.qctl[kQOver + pl] = 1;
    }
    __syncthreads();                               // s_base is read
  }
}

hipError_t launch_queue_from_bitmap(const EventSet& ev, InstFlags* flags, int64_t k, int32_t s,
                                    int64_t n_props, hipStream_t stream) {
  int64_t grid = (ev.words + kBlock - 1) / kBlock;
  if (grid > 2048) grid = 2048;
  if (grid < 1) grid = 1;
  k_queue_from_bitmap<<<(unsigned)grid, kBlock, 0, stream>>>(ev, flags, k, s, n_props);
  return hipGetLastError();
}

hipError_t launch_sweep_queue(const Table& t, int2* store, int64_t nvars, InstFlags* flags,
                              unsigned long long* tight, const EventSet& ev, int dom, int grid,
                              hipStream_t stream) {
  switch (dom) {
    case kItv: k_sweep_queue<kItv><<<grid, kBlock, 0, stream>>>(t, store, nvars, flags, tight, ev); break;
    case kItvDiv: k_sweep_queue<kItvDiv><<<grid, kBlock, 0, stream>>>(t, store, nvars, flags, tight, ev); break;
    case kBits: k_sweep_queue<kBits><<<grid, kBlock, 0, stream>>>(t, store, nvars, flags, tight, ev); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sweep_lockstep(const Table& t, int2* store, int2* shadow, int64_t nvars,
                                 int32_t n_inst, const uint32_t* member, int64_t words,
                                 InstFlags* flags, WorkList wl, unsigned long long* tight, int gi,
                                 int dom, int grid, hipStream_t stream, const EventSet* ev) {
  const EventSet none{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, 0};
  const EventSet& e = ev ? *ev : none;
  if (t.n <= 0 || n_inst <= 0) return hipSuccess;
#define PCX_LOCK(D, G) \
  k_sweep_lockstep<D, G><<<grid, kBlock, 0, stream>>>(t, store, shadow, nvars, n_inst, member, words, flags, wl, tight, e)
#define PCX_LOCK_GI(D)                   \
  switch (gi) {                          \
    case 16: PCX_LOCK(D, 16); break;     \
    case 32: PCX_LOCK(D, 32); break;     \
    case 64: PCX_LOCK(D, 64); break;     \
    default: return hipErrorInvalidValue; \
  }
  switch (dom) {
    case kItv: PCX_LOCK_GI(kItv); break;
    case kItvDiv: PCX_LOCK_GI(kItvDiv); break;
    case kBits: PCX_LOCK_GI(kBits); break;
    default: return hipErrorInvalidValue;
  }
#undef PCX_LOCK_GI
#undef PCX_LOCK
  return hipGetLastError();
}

// Marked-only sweep over every running instance: items are (instance,
// kMarkedWordsPerItem bitmap words), claimed per XCD as in the full sweep.
// Waves per SIMD of the batched marked sweep: 6 for intervals (config 2's
// marked sweeps 0.98 -> 0.93 ms); the other variants keep their default.
template <int D>
constexpr int marked_min_waves() { return D == kItv ? 6 : 1; }

template <int D>
__global__ __launch_bounds__(kBlock, marked_min_waves<D>()) void k_sweep_marked(Table t, int2* __restrict__ store,
                                                         int64_t nvars, InstFlags* flags,
                                                         WorkList wl, unsigned long long* tight,
                                                         EventSet ev) {
  const int32_t s = *wl.sweep;
  const int32_t n_act = *wl.n_act;
  if (n_act <= 0) return;
  const int64_t total = wl.item_base[n_act];
  const int xcd = blockIdx.x % kXcds;
  const int64_t first_item = total * xcd / kXcds;
  const int64_t end_item = total * (xcd + 1) / kXcds;
  int32_t* qc = wl.xq + xcd;
  uint32_t* cur = ev.dirty + (int64_t)(s & 1) * ev.n_inst * ev.words;
  const int group = uni(*wl.group);
  __shared__ int32_t s_next;
  __shared__ int32_t s_list[kMarkCap];
  __shared__ int32_t s_props[kPropCap];
  __shared__ int32_t s_scan[kBlock / kWave];
  __shared__ int32_t s_chunks[kMarkedGroupMax];
  __shared__ int32_t s_n, s_over, s_ng;
  if (threadIdx.x == 0) { s_next = atomicAdd(qc, 1); s_n = 0; s_over = 0; }
  __syncthreads();
  int64_t item = first_item + __builtin_amdgcn_readfirstlane(s_next);
  __syncthreads();
  int hint = 0;
  while (item < end_item) {
    int32_t nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(qc, 1);
    const int lo = uni(find_active(wl.item_base, n_act, item, hint));
    hint = lo;
    const int32_t k = uni(wl.act[lo]);
    InstFlags* f = flags + k;
    // the item's chunks: set bits [r * group, r * group + group) of the
    // instance's snapshot, listed by thread 0
    if (threadIdx.x == 0) {
      const int64_t first = (item - wl.item_base[lo]) * group;
      const uint32_t* snap = ev.snap + (int64_t)k * ev.swords;
      int64_t seen = 0;
      int ng = 0;
      for (int64_t j = 0; j < ev.swords && ng < group; ++j) {
        for (uint32_t v = snap[j]; v && ng < group; v &= v - 1, ++seen)
          if (seen >= first) s_chunks[ng++] = (int32_t)(j * 32 + __builtin_ctz(v));
      }
      for (int g = ng; g < kMarkedGroupMax; ++g) s_chunks[g] = 0;
      s_ng = ng;
    }
    __syncthreads();
    const int ng = s_ng;
    Marker mk = marker_for(&ev, k, s, true);
    mk.list = s_list;
    mk.list_n = &s_n;
    mk.overflow = &s_over;
    mk.all_marked = &f->all_marked[(s + 1) & 1];
    mk.all_marked_s = s + 1;
    const bool all = uni((int32_t)(f->all_marked[s & 1] == s)) != 0;
    bool chg = false, fail = false;
    int nt = 0, ne = 0;
    const bool active = uni((int32_t)(f->failed_at == kNoFail && ng > 0)) != 0;
    for (int pass = 0;; ++pass) {        // passes while they find marks (see k_sweep_marked_plain)
      const int total = sweep_marked_group<D>(t, store + (int64_t)k * nvars, cur + (int64_t)k * ev.words,
                                              s_chunks, ng, ev.words, active, mk, chg, fail, nt, ne,
                                              s_props, s_scan, pass > 0, all);
      flush_marks(mk);
      __syncthreads();
      if (total == 0 || pass + 1 >= kBatchedMarkedPasses) break;
    }
    publish(f, s, chg, fail, nt, tight);
    count_evals(f, ne);
    if (threadIdx.x == 0) s_next = nxt;
    __syncthreads();
    item = first_item + __builtin_amdgcn_readfirstlane(s_next);
    __syncthreads();
  }
}

hipError_t launch_sweep_marked(const Table& t, int2* store, int64_t nvars, InstFlags* flags,
                               WorkList wl, unsigned long long* tight, const EventSet& ev,
                               int dom, int grid, hipStream_t stream) {
  static_assert(kMarkedWordsPerItem == kBlock, "one bitmap word per thread");
  switch (dom) {
    case kItv: k_sweep_marked<kItv><<<grid, kBlock, 0, stream>>>(t, store, nvars, flags, wl, tight, ev); break;
    case kItvDiv: k_sweep_marked<kItvDiv><<<grid, kBlock, 0, stream>>>(t, store, nvars, flags, wl, tight, ev); break;
    case kBits: k_sweep_marked<kBits><<<grid, kBlock, 0, stream>>>(t, store, nvars, flags, wl, tight, ev); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// member[k][p >> 5] bit p & 31 <=> table position p is in instance k's subset.
// Lanes of a wave take consecutive subset entries; subsets are ascending, so
// neighbouring lanes mostly set bits of the same word.  A segmented OR over
// lanes with equal words (shuffles) leaves each word's bits in the last lane
// of its run, and only that lane writes: a few atomics per wave instead of
// 64.  (Every lane's bits reach a lane that writes them whatever the order.)
__global__ void k_build_members(const int64_t* __restrict__ off, const int32_t* __restrict__ idx,
                                int64_t n_inst, uint32_t* member, int64_t words) {
  const int lane = lane_id();
  for (int64_t k = blockIdx.x; k < n_inst; k += gridDim.x) {
    uint32_t* row = member + k * words;
    const int64_t end = off[k + 1];
    for (int64_t base = off[k]; base < end; base += blockDim.x) {   // same trip count in every thread
      const int64_t i = base + threadIdx.x;
      const bool ok = i < end;
      const int32_t p = ok ? idx[i] : 0;
      const int32_t w = ok ? (p >> 5) : -1;
      uint32_t b = ok ? (1u << (p & 31)) : 0u;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t b2 = __shfl_up(b, o, kWave);
        const int32_t w2 = __shfl_up(w, o, kWave);
        if (lane >= o && w2 == w) b |= b2;
      }
      const int32_t wn = __shfl_down(w, 1, kWave);
      if (ok && (lane == kWave - 1 || wn != w)) atomicOr(&row[w], b);
    }
  }
}

hipError_t launch_build_members(const int64_t* off, const int32_t* idx, int64_t n_inst,
                                uint32_t* member, int64_t words, hipStream_t stream) {
  const int64_t grid = n_inst < 65536 ? n_inst : 65536;
  if (grid > 0) k_build_members<<<(unsigned)grid, 256, 0, stream>>>(off, idx, n_inst, member, words);
  return hipGetLastError();
}

// ---- serving slots: heterogeneous batch, one table and store per slot ----
__device__ __forceinline__ Marker slot_marker(const SlotDesc& sd, int32_t s, bool marked_sweep) {
  if (!sd.adj) return Marker{};
  uint32_t* dirty = sd.marks;
  uint32_t* summary = sd.marks + 2 * sd.words;
  const int64_t nx = (s + 1) & 1, cu = s & 1;
  Marker mk{};
  mk.adj_off = sd.adj_off;
  mk.adj = sd.adj;
  mk.dirty = dirty + nx * sd.words;
  mk.also_now = marked_sweep ? dirty + cu * sd.words : nullptr;
  mk.sum = summary + nx * sd.swords;
  mk.sum_now = marked_sweep ? summary + cu * sd.swords : nullptr;
  mk.words = sd.words;
  mk.swords = (int32_t)sd.swords;
  return mk;
}

// A slot's sweep s is a full sweep: no marks, its first sweep, or a sweep
// for which a workgroup marked everything (all_marked).
__device__ __forceinline__ bool slot_full_sweep(const SlotDesc& sd, const InstFlags& f, int32_t s) {
  return !sd.adj || s == sd.start || f.all_marked[s & 1] == s;
}
// Same scheme as the batched mode: a one-workgroup scheduler builds the item
// prefix sum over the slots still running, then a grid-stride XCD-aware
// sweep.  Free slots carry failed_at != kNoFail and done_at = 0, so they are
// never scheduled.  The sweep counter keeps running across ticks; a slot
// admitted before sweep S is primed with changed[(S-1)&1] = S-1.
__global__ __launch_bounds__(kSchedThreads) void k_schedule_slots(const SlotDesc* slots,
                                                                  InstFlags* flags,
                                                                  int32_t n_slots, WorkList wl,
                                                                  int64_t chunk) {
  __shared__ int64_t part[kSchedThreads];
  __shared__ int32_t cnt[kSchedThreads];
  const int32_t s = *wl.sweep + 1;
  const int tid = threadIdx.x;
  if (tid < kXcds) wl.xq[tid] = 0;
  const int per = (n_slots + kSchedThreads - 1) / kSchedThreads;
  const int lo = min(n_slots, tid * per), hi = min(n_slots, lo + per);
  int64_t items = 0;
  int32_t na = 0;
  for (int k = lo; k < hi; ++k) {
    InstFlags& f = flags[k];
    const bool run = f.done_at < 0 && f.failed_at == kNoFail && f.settled_at < 0 &&
                     f.changed[(s - 1) & 1] == s - 1;
    if (!run) {
      if (f.done_at < 0) f.done_at = s;
      continue;
    }
    const SlotDesc& sd = slots[k];
    if (slot_full_sweep(sd, f, s)) {
      items += (sd.m + chunk - 1) / chunk;
    } else {                 // marked sweep: the chunks holding marks; snapshot and clear
      uint32_t* srow = sd.marks + 2 * sd.words + (int64_t)(s & 1) * sd.swords;
      uint32_t* snap = sd.marks + 2 * sd.words + 2 * sd.swords;
      for (int64_t j = 0; j < sd.swords; ++j) {
        const uint32_t v = srow[j];
        snap[j] = v;
        srow[j] = 0;
        items += __builtin_popcount(v);
      }
    }
    ++na;
  }
  part[tid] = items;
  cnt[tid] = na;
  __syncthreads();
  for (int o = 1; o < kSchedThreads; o <<= 1) {
    const int64_t pv = tid >= o ? part[tid - o] : 0;
    const int32_t cv = tid >= o ? cnt[tid - o] : 0;
    __syncthreads();
    part[tid] += pv;
    cnt[tid] += cv;
    __syncthreads();
  }
  int64_t ib = part[tid] - items;
  int32_t ia = cnt[tid] - na;
  for (int k = lo; k < hi; ++k) {
    const InstFlags& f = flags[k];
    if (f.done_at >= 0) continue;
    wl.act[ia] = k;
    wl.item_base[ia] = ib;
    const SlotDesc& sd = slots[k];
    if (slot_full_sweep(sd, f, s)) {
      ib += (sd.m + chunk - 1) / chunk;
    } else {
      const uint32_t* snap = sd.marks + 2 * sd.words + 2 * sd.swords;
      for (int64_t j = 0; j < sd.swords; ++j) ib += __builtin_popcount(snap[j]);
    }
    ++ia;
  }
  __syncthreads();
  if (tid == kSchedThreads - 1) {
    wl.item_base[cnt[tid]] = part[tid];
    *wl.n_act = cnt[tid];
    *wl.live = cnt[tid];
    *wl.sweep = s;
  }
}

hipError_t launch_schedule_slots(const SlotDesc* slots, InstFlags* flags, int32_t n_slots,
                                 WorkList wl, int ppw, hipStream_t stream) {
  const int64_t chunk = (int64_t)kBlock * (ppw / kWave);
  k_schedule_slots<<<1, kSchedThreads, 0, stream>>>(slots, flags, n_slots, wl, chunk);
  return hipGetLastError();
}

template <int PPL, int D>
__global__ __launch_bounds__(kBlock) void k_sweep_slots(const SlotDesc* slots, InstFlags* flags,
                                                        WorkList wl, unsigned long long* tight) {
  const int32_t s = *wl.sweep;
  const int32_t n_act = *wl.n_act;
  if (n_act <= 0) return;
  const int64_t total = wl.item_base[n_act];
  constexpr int64_t chunk = (int64_t)kBlock * PPL;
  const int xcd = blockIdx.x % kXcds;
  const int64_t first_item = total * xcd / kXcds;
  const int64_t end_item = total * (xcd + 1) / kXcds;
  int32_t* qc = wl.xq + xcd;
  __shared__ int32_t s_next;
  __shared__ int32_t s_list[kMarkCap];
  __shared__ int32_t s_props[kPropCap];
  __shared__ int32_t s_scan[kBlock / kWave];
  __shared__ int32_t s_n, s_over;
  if (threadIdx.x == 0) { s_next = atomicAdd(qc, 1); s_n = 0; s_over = 0; }
  __syncthreads();
  int64_t item = first_item + __builtin_amdgcn_readfirstlane(s_next);
  __syncthreads();
  int hint = 0;
  while (item < end_item) {
    int32_t nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(qc, 1);
    const int lo = uni(find_active(wl.item_base, n_act, item, hint));
    hint = lo;
    const int32_t k = uni(wl.act[lo]);
    SlotDesc sd = slots[k];
    sd.xo = uni(sd.xo); sd.y = uni(sd.y); sd.z = uni(sd.z); sd.m = uni(sd.m);
    sd.store = uni(sd.store); sd.adj_off = uni(sd.adj_off); sd.adj = uni(sd.adj);
    sd.marks = uni(sd.marks); sd.words = uni(sd.words); sd.swords = uni(sd.swords);
    sd.start = uni(sd.start);
    InstFlags* f = flags + k;
    const Table t{sd.xo, sd.y, sd.z, sd.m};
    const bool full = uni((int32_t)slot_full_sweep(sd, *f, s)) != 0;
    Marker mk = slot_marker(sd, s, !full);
    if (mk.dirty) {
      mk.list = s_list; mk.list_n = &s_n; mk.overflow = &s_over;
      mk.all_marked = &f->all_marked[(s + 1) & 1];
      mk.all_marked_s = s + 1;
    }
    const int64_t r = item - uni(wl.item_base[lo]);
    bool chg = false, fail = false;
    int nt = 0, ne = 0;
    if (full) {
      if (f->failed_at == kNoFail) {
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
          const int64_t q = r * chunk + j * kBlock + threadIdx.x;
          if (q < sd.m) eval_prop<D, false>(t, sd.store, q, chg, fail, nt, mk);
        }
      }
    } else {
      // the item's chunk: the r-th set bit of the slot's snapshot
      const uint32_t* snap = sd.marks + 2 * sd.words + 2 * sd.swords;
      int64_t rr = r, c = -1;
      for (int64_t j = 0; j < sd.swords; ++j) {
        uint32_t v = snap[j];
        const int pc = __builtin_popcount(v);
        if (rr < pc) {
          for (int q = 0; q < rr; ++q) v &= v - 1;
          c = j * 32 + __builtin_ctz(v);
          break;
        }
        rr -= pc;
      }
      c = uni(c);
      const bool active = uni((int32_t)(f->failed_at == kNoFail && c >= 0)) != 0;
      // One pass per chunk: repeated passes shorten long requests but lengthen
      // the tick that every request in the batch waits for.
      sweep_marked_chunk<D>(t, sd.store, sd.marks + (int64_t)(s & 1) * sd.words, c * kChunkWords,
                            sd.words, active, mk, chg, fail, nt, ne, s_props, s_scan);
    }
    publish(f, s, chg, fail, nt, tight);
    count_evals(f, ne);
    flush_marks(mk);
    if (threadIdx.x == 0) s_next = nxt;
    __syncthreads();
    item = first_item + __builtin_amdgcn_readfirstlane(s_next);
    __syncthreads();
  }
}

hipError_t launch_sweep_slots(const SlotDesc* slots, InstFlags* flags, WorkList wl,
                              unsigned long long* tight, int ppw, int dom, int grid,
                              hipStream_t stream) {
  const int ppl = ppw / kWave;
#define PCX_SLOTS(P, D) k_sweep_slots<P, D><<<grid, kBlock, 0, stream>>>(slots, flags, wl, tight)
  if (dom == kItvDiv) {
    switch (ppl) {
      case 1: PCX_SLOTS(1, kItvDiv); break;
      case 2: PCX_SLOTS(2, kItvDiv); break;
      case 4: PCX_SLOTS(4, kItvDiv); break;
      case 8: PCX_SLOTS(8, kItvDiv); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (ppl) {
      case 1: PCX_SLOTS(1, kItv); break;
      case 2: PCX_SLOTS(2, kItv); break;
      case 4: PCX_SLOTS(4, kItv); break;
      case 8: PCX_SLOTS(8, kItv); break;
      default: return hipErrorInvalidValue;
    }
  }
#undef PCX_SLOTS
  return hipGetLastError();
}

// ---- serving bursts (side lane) --------------------------------------------
// Some requests need thousands of sweeps, most of which change a few bounds
// each (bounds of a cycle climbing step by step).  In the slot batch each of
// those sweeps costs two launches and a share of a tick's host round trip.
// The server hands such a slot to a side lane: this kernel, on a stream of
// its own, runs one workgroup per handed-over slot while the slot batch goes
// on without it.  The slot left the batch after sweep S = sd.handover.
//  - Marked slots: the propagators marked for sweep S+1 are gathered into
//    an LDS ring (when they fit).  Each round pops up to kBurstPop of them,
//    clears their bits, evaluates them, and marks the adjacency of every
//    variable they tightened: a mark whose bit was clear (returning
//    atomicOr) is appended to the ring, so the ring always holds exactly
//    the set bits.  Every unmarked propagator is at its fixpoint (the
//    marked-sweep invariant), so an empty ring is the fixpoint.  At exit the
//    summary row is rewritten from the ring, so bitmap and summary again
//    describe the marks of sweep S+1.
//  - Slots with full sweeps (at most kBurstFullMax propagators): a round is
//    a sweep over every propagator; one that tightens nothing is the fixpoint.
// Exit: settled (settled_at = S), failed (failed_at = S), budget spent
// (done_at stays kSideLane + lane), or the marks do not fit (done_at = kSideBail:
// the slot goes back to the batch).  The slot's store and marks are touched
// by this workgroup alone, so the store is read and tightened with
// workgroup-scope operations and barriers separate rounds, as in the
// persistent kernel; mark bits are device-scope atomics.
constexpr int kBurstBlock = 512;
constexpr int kBurstQ = 4096;        // ring of marked propagators (power of two)
constexpr int kBurstPop = kMarkCap / 3;   // propagators evaluated per round
constexpr int kBurstChunks = 256;    // marked chunks gathered at the start
static_assert(kBurstSummaryMax <= kBurstBlock, "one summary word per thread");


template <int D>
__global__ __launch_bounds__(kBurstBlock) void k_burst_slots(const SlotDesc* slots,
                                                             InstFlags* flags, int32_t n_slots,
                                                             unsigned long long* tight,
                                                             int32_t budget, int32_t lane) {
  constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
  constexpr int BT = kBurstBlock;
  const int32_t k = blockIdx.x;
  if (k >= n_slots) return;
  InstFlags* f = flags + k;
  if (uni((int32_t)(f->done_at == kSideLane + lane && f->failed_at == kNoFail &&
                    f->settled_at < 0)) == 0)
    return;
  SlotDesc sd = slots[k];
  sd.xo = uni(sd.xo); sd.y = uni(sd.y); sd.z = uni(sd.z); sd.m = uni(sd.m);
  sd.store = uni(sd.store); sd.adj_off = uni(sd.adj_off); sd.adj = uni(sd.adj);
  sd.marks = uni(sd.marks); sd.words = uni(sd.words); sd.swords = uni(sd.swords);
  const int32_t S = uni(sd.handover);
  const bool marked = sd.adj != nullptr;
  const int tid = threadIdx.x;
  if (!(marked ? sd.swords <= kBurstSummaryMax : sd.m <= kBurstFullMax)) {
    if (tid == 0) f->done_at = kSideBail;
    return;
  }

  __shared__ int32_t s_list[2][kMarkCap];          // tightened variables, by round parity
  __shared__ int32_t s_n[2];
  __shared__ int32_t s_q[kBurstQ];
  __shared__ uint32_t s_sum[kBurstSummaryMax];
  __shared__ int32_t s_chunks[kBurstChunks];
  __shared__ int32_t s_scan[BT / kWave];
  __shared__ int32_t s_over, s_head, s_tail;
  const Table t{sd.xo, sd.y, sd.z, sd.m};
  const int64_t plane = (S + 1) & 1;               // the marks of sweep S+1
  uint32_t* cur = sd.marks + plane * sd.words;
  uint32_t* srow = sd.marks + 2 * sd.words + plane * sd.swords;
  if (marked && f->all_marked[(S + 1) & 1] == S + 1) {   // every propagator marked: the batch's work
    if (tid == 0) f->done_at = kSideBail;
    return;
  }
  if (marked) {
    // gather every marked propagator: chunks from the summary row, then
    // their words (counted first; the slot goes back if they do not fit)
    const uint32_t sw = tid < sd.swords ? srow[tid] : 0u;
    int nc;
    int o = block_scan<BT>(__builtin_popcount(sw), s_scan, &nc);
    bool fits = nc <= kBurstChunks;
    if (fits) for (uint32_t b = sw; b; b &= b - 1) s_chunks[o++] = tid * 32 + __builtin_ctz(b);
    __syncthreads();
    const int64_t nw = fits ? (int64_t)nc * kChunkWords : 0;
    int cnt = 0;
    for (int64_t j = tid; j < nw; j += BT) {
      const int64_t w = (int64_t)s_chunks[j / kChunkWords] * kChunkWords + j % kChunkWords;
      if (w < sd.words) cnt += __builtin_popcount(cur[w]);
    }
    int np;
    int base = block_scan<BT>(cnt, s_scan, &np);
    fits = fits && np <= kBurstQ;
    if (!fits) {
      if (tid == 0) f->done_at = kSideBail;
      return;
    }
    for (int64_t j = tid; j < nw; j += BT) {
      const int64_t w = (int64_t)s_chunks[j / kChunkWords] * kChunkWords + j % kChunkWords;
      if (w >= sd.words) continue;
      for (uint32_t bits = cur[w]; bits; bits &= bits - 1)
        s_q[base++] = (int32_t)(w * 32 + __builtin_ctz(bits));
    }
    if (tid == 0) { s_head = 0; s_tail = np; }
  }
  if (tid == 0) { s_n[0] = 0; s_n[1] = 0; s_over = 0; }
  __syncthreads();
  bool chg = false, fail = false, settled = false;
  int nt = 0, ne = 0, it = 0;
  while (it < budget) {
    if (marked) {
      // One round: pop up to kBurstPop entries, clear their bits, evaluate
      // them; after the barrier, mark the adjacency of what they tightened.
      const int par = it & 1;
      Marker mk{};
      mk.dirty = cur;                              // apply() lists tightened variables
      mk.list = s_list[par];
      mk.list_n = &s_n[par];
      mk.overflow = &s_over;
      const int32_t head = s_head, n = min(s_tail - head, kBurstPop);
      if (n == 0) { settled = true; ++it; break; }         // the quiet round
      const int32_t p = tid < n ? s_q[(head + tid) & (kBurstQ - 1)] : -1;
      if (p >= 0) {
        __hip_atomic_fetch_and(&cur[p >> 5], ~(1u << (p & 31)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p < t.n) { eval_prop<D, false, WG>(t, sd.store, p, chg, fail, nt, mk); ++ne; }
      }
      // The clears are non-returning atomics: wait for them explicitly, so a
      // re-mark after the barrier finds the bit clear and re-queues.
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();             // bits cleared, variables listed, head and tail read
      if (tid == 0) { s_head = head + n; s_n[par ^ 1] = 0; }
      const int32_t limit = head + n + kBurstQ;    // ring capacity after this pop
      const int nl = s_n[par];
      for (int i = tid; i < nl; i += BT) {
        const int32_t v = s_list[par][i];
        const int32_t e = sd.adj_off[v + 1];
        for (int32_t a = sd.adj_off[v]; a < e; a += 8) {
          int32_t q[8];
          uint32_t old[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) q[j] = a + j < e ? sd.adj[a + j] : -1;
#pragma unroll
          for (int j = 0; j < 8; ++j)
            old[j] = q[j] >= 0 ? atomicOr(&cur[q[j] >> 5], 1u << (q[j] & 31)) : ~0u;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            if (q[j] < 0 || (old[j] & (1u << (q[j] & 31)))) continue;
            const int32_t at = atomicAdd(&s_tail, 1);
            if (at < limit) s_q[at & (kBurstQ - 1)] = q[j];
            else s_over = 1;                               // the bit stays set; the summary covers it
          }
        }
      }
      ++it;
      const bool any_fail = __syncthreads_or(fail);
      if (any_fail || s_over) break;
    } else {
      bool c = false;
      for (int64_t q = tid; q < sd.m; q += BT) eval_prop<D, false, WG>(t, sd.store, q, c, fail, nt, Marker{});
      chg |= c;
      ++it;
      const bool any_fail = __syncthreads_or(fail);
      const bool any_chg = __syncthreads_or(c);
      if (any_fail) break;
      if (!any_chg) { settled = true; break; }             // a full sweep that changed nothing
    }
  }
  static_assert(3 * kBurstPop <= kMarkCap, "a round's tightenings fit the mark list");
  if (marked && !settled && !fail) {
    // The summary row again covers exactly the marked chunks: those of the
    // ring's entries, or every chunk when marks overflowed the ring.
    for (int j = tid; j < sd.swords; j += BT) s_sum[j] = 0;
    __syncthreads();
    const int64_t chunks = (sd.words + kChunkWords - 1) / kChunkWords;
    if (s_over) {
      for (int64_t j = tid; j < sd.swords; j += BT) {
        const int64_t left = chunks - j * 32;
        s_sum[j] = left >= 32 ? ~0u : ((1u << left) - 1u);
      }
    } else {
      const int32_t head = s_head, n = s_tail - head;
      for (int i = tid; i < n; i += BT) {
        const int32_t c = s_q[(head + i) & (kBurstQ - 1)] >> kChunkShift;
        atomicOr(&s_sum[c >> 5], 1u << (c & 31));
      }
    }
    __syncthreads();
    for (int j = tid; j < sd.swords; j += BT) srow[j] = s_sum[j];
  }
  publish(f, S, chg, fail, nt, tight);
  count_evals(f, ne);
  if (tid == 0) {
    f->extra_sweeps += it;
    if (settled && !fail) f->settled_at = S;
    if (marked && s_over && !fail) f->done_at = kSideBail;
  }
}

hipError_t launch_burst_slots(const SlotDesc* slots, InstFlags* flags, int32_t n_slots,
                              unsigned long long* tight, int32_t budget, int32_t lane, int dom,
                              hipStream_t stream) {
  if (n_slots <= 0 || budget <= 0) return hipSuccess;
  if (dom == kItvDiv)
    k_burst_slots<kItvDiv><<<n_slots, kBurstBlock, 0, stream>>>(slots, flags, n_slots, tight, budget, lane);
  else
    k_burst_slots<kItv><<<n_slots, kBurstBlock, 0, stream>>>(slots, flags, n_slots, tight, budget, lane);
  return hipGetLastError();
}

// ---- persistent mode: one launch, a device queue of instances -------------
// Each workgroup takes the next instance from the queue (largest first) and
// runs all of its sweeps itself, separated by workgroup barriers.  Store
// accesses are workgroup-scoped atomics; the barrier between sweeps orders
// one sweep's tightenings before the next sweep's reads, which is what the
// termination argument needs.
constexpr int kPBlock = 512;     // 8 waves per workgroup

// LDS: the instance's store (nvars 8-byte entries, at most lds_store_cap())
// is copied into the workgroup's LDS, swept there with LDS atomics, and
// written back when the instance is done.
template <int D, bool LDS>
__global__ __launch_bounds__(kPBlock) void k_persistent(Table t, int2* __restrict__ store,
                                                        int64_t nvars, Subsets sub, int64_t n_props,
                                                        InstFlags* flags, const int32_t* order,
                                                        int32_t n_inst, int32_t* queue,
                                                        unsigned long long* tight,
                                                        int32_t max_sweeps) {
  constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
  __shared__ int32_t s_slot, s_chg, s_fail;
  extern __shared__ u64 s_lds[];
  const int lane = lane_id();
  for (;;) {
    if (threadIdx.x == 0) s_slot = atomicAdd(queue, 1);
    __syncthreads();
    const int32_t slot = s_slot;
    if (slot >= n_inst) return;            // same value in every wave: uniform exit
    const int32_t k = order[slot];
    InstFlags* f = flags + k;
    const InstTable it = inst_table(t, sub, k, n_props);
    const Table& tk = it.t;
    const int32_t* idx = it.idx;
    const int64_t m = it.m;
    int2* gst = store + (int64_t)k * nvars;
    int2* st = LDS ? reinterpret_cast<int2*>(s_lds) : gst;
    const bool start_failed = f->failed_at != kNoFail;
    int32_t last_changed = -1, failed_at = kNoFail, s = 0;
    int nt = 0;
    if (LDS && !start_failed) {
      const u64* g = reinterpret_cast<const u64*>(gst);
      for (int64_t v = threadIdx.x; v < nvars; v += kPBlock) s_lds[v] = g[v];
      __syncthreads();
    }
    if (!start_failed) {
      for (s = 0; max_sweeps <= 0 || s < max_sweeps; ++s) {
        if (threadIdx.x == 0) { s_chg = 0; s_fail = 0; }
        __syncthreads();
        bool chg = false, fail = false;
        // Two propagators per lane per step: both sets of loads are issued
        // before either set of atomics (reading a bound another lane is
        // about to tighten is allowed: bounds only shrink).
        for (int64_t q = threadIdx.x; q < m; q += 2 * kPBlock) {
          const int64_t q2 = q + kPBlock;
          const int64_t i1 = prop_at(idx, q);
          const int64_t i2 = q2 < m ? prop_at(idx, q2) : -1;
          const uint32_t xo1 = tk.xo[i1];
          const int32_t y1 = tk.y[i1], z1 = tk.z[i1];
          uint32_t xo2 = 0;
          int32_t y2 = 0, z2 = 0;
          if (i2 >= 0) { xo2 = tk.xo[i2]; y2 = tk.y[i2]; z2 = tk.z[i2]; }
          const int32_t x1 = (int32_t)(xo1 & kVarMask), x2 = (int32_t)(xo2 & kVarMask);
          const u64 a1 = load_word<WG>(st + x1), b1 = load_word<WG>(st + y1), c1 = load_word<WG>(st + z1);
          u64 a2 = 0, b2 = 0, c2 = 0;
          if (i2 >= 0) { a2 = load_word<WG>(st + x2); b2 = load_word<WG>(st + y2); c2 = load_word<WG>(st + z2); }
          const Marker none{};
          apply<D, WG>(st, (int)(xo1 >> kOpShift), x1, y1, z1, a1, b1, c1, chg, fail, nt, none);
          if (i2 >= 0) apply<D, WG>(st, (int)(xo2 >> kOpShift), x2, y2, z2, a2, b2, c2, chg, fail, nt, none);
        }
        const bool wc = __ballot(chg) != 0, wf = __ballot(fail) != 0;
        if (lane == 0 && wc) s_chg = 1;
        if (lane == 0 && wf) s_fail = 1;
        __syncthreads();
        const bool any_chg = s_chg != 0, any_fail = s_fail != 0;
        __syncthreads();                     // all read the verdict before it is reset
        if (any_chg) last_changed = s;
        if (any_fail) { failed_at = s; break; }
        if (!any_chg) break;
      }
    }
    if (LDS && !start_failed) {
      __syncthreads();                       // every tightening is in LDS
      u64* g = reinterpret_cast<u64*>(gst);
      for (int64_t v = threadIdx.x; v < nvars; v += kPBlock) g[v] = s_lds[v];
    }
    const int sum = wave_sum(nt);
    if (lane == 0 && sum) atomicAdd(&tight[(k * 16 + (threadIdx.x >> 6)) & (kTightSlots - 1)],
                                    (unsigned long long)sum);
    if (threadIdx.x == 0 && !start_failed) {
      if (last_changed >= 0) f->changed[last_changed & 1] = last_changed;
      if (failed_at != kNoFail) f->failed_at = failed_at;
      f->done_at = s;
    }
    __syncthreads();                         // s_slot is rewritten next round
  }
}

int persistent_blocks_per_cu(int dom, int64_t lds_bytes) {
  int n = 0;
  hipError_t e;
  const size_t sm = (size_t)lds_bytes;
#define PCX_OCC(D, L) hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_persistent<D, L>, kPBlock, sm)
  switch (dom) {
    case kItvDiv: e = lds_bytes ? PCX_OCC(kItvDiv, true) : PCX_OCC(kItvDiv, false); break;
    case kBits: e = lds_bytes ? PCX_OCC(kBits, true) : PCX_OCC(kBits, false); break;
    default: e = lds_bytes ? PCX_OCC(kItv, true) : PCX_OCC(kItv, false); break;
  }
#undef PCX_OCC
  return (e == hipSuccess && n > 0) ? n : 1;
}

int64_t lds_store_cap(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess ||
      v <= kLdsStaticReserve)
    return kLdsStoreBytes;
  return (int64_t)v - kLdsStaticReserve;
}

hipError_t launch_persistent(const Table& t, int2* store, int64_t nvars, const Subsets& sub,
                             int64_t n_props, InstFlags* flags, const int32_t* order,
                             int32_t n_inst, int32_t* queue, unsigned long long* tight,
                             int32_t max_sweeps, int dom, int grid, bool lds, int64_t lds_cap,
                             hipStream_t stream) {
  if (lds && nvars * 8 > lds_cap) return hipErrorInvalidValue;
  const size_t sm = lds ? (size_t)nvars * 8 : 0;
  // Dynamic LDS beyond 64 KB (gfx950 has 160 KB per workgroup) is requested
  // from the runtime first.
#define PCX_PERSIST(D, L)                                                                        \
  do {                                                                                         \
    if (sm > 65536) {                                                                          \
      const hipError_t fe = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_persistent<D, L>), \
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm); \
      if (fe != hipSuccess) return fe;                                                         \
    }                                                                                          \
    k_persistent<D, L><<<grid, kPBlock, sm, stream>>>(t, store, nvars, sub, n_props, flags, order, \
                                                      n_inst, queue, tight, max_sweeps);       \
  } while (0)
  switch (dom) {
    case kItv: if (lds) PCX_PERSIST(kItv, true); else PCX_PERSIST(kItv, false); break;
    case kItvDiv: if (lds) PCX_PERSIST(kItvDiv, true); else PCX_PERSIST(kItvDiv, false); break;
    case kBits: if (lds) PCX_PERSIST(kBits, true); else PCX_PERSIST(kBits, false); break;
    default: return hipErrorInvalidValue;
  }
#undef PCX_PERSIST
  return hipGetLastError();
}

// ---- bookkeeping kernels --------------------------------------------------
__global__ void k_init_flags(InstFlags* flags, int32_t n_inst, WorkList wl,
                             const uint8_t* __restrict__ empty) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_inst) {
    flags[k].changed[0] = -2;
    flags[k].changed[1] = -1;      // "sweep -1 changed" lets sweep 0 run
    flags[k].failed_at = (empty && empty[k]) ? -1 : kNoFail;
    flags[k].done_at = -1;
    flags[k].marked_evals = 0;
    flags[k].extra_sweeps = 0;
    flags[k].settled_at = -1;
    flags[k].all_marked[0] = -1;
    flags[k].all_marked[1] = -1;
  }
  if (k == 0 && wl.sweep) {
    *wl.sweep = -1;
    *wl.n_act = 0;
    *wl.live = n_inst;
    wl.live[1] = 0;                // persistent-mode queue head (ctl[3])
  }
}

hipError_t launch_init_flags(InstFlags* flags, int32_t n_inst, WorkList wl, hipStream_t stream,
                             const uint8_t* empty) {
  const int grid = (int)((n_inst + 255) / 256);
  k_init_flags<<<grid > 0 ? grid : 1, 256, 0, stream>>>(flags, n_inst, wl, empty);
  return hipGetLastError();
}

// empty[k] = 1 when instance k's slice of `store` holds an empty domain
// (empty must be zeroed first).
template <bool BITS>
__global__ void k_mark_empty(const int2* __restrict__ store, int64_t nvars, int64_t total,
                             uint8_t* empty) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < total;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int2 b = store[v];
    if (BITS ? (b.x == 0 && b.y == 0) : (b.x > b.y)) empty[v / nvars] = 1;
  }
}

hipError_t launch_store_empty_map(const int2* store, int64_t nvars, int32_t n_inst, uint8_t* empty,
                                  int dom, hipStream_t stream) {
  const int64_t total = nvars * (int64_t)n_inst;
  int64_t grid = (total + 255) / 256;
  if (grid > 16384) grid = 16384;
  if (grid <= 0) return hipSuccess;
  if (dom == kBits)
    k_mark_empty<true><<<dim3((unsigned)grid), 256, 0, stream>>>(store, nvars, total, empty);
  else
    k_mark_empty<false><<<dim3((unsigned)grid), 256, 0, stream>>>(store, nvars, total, empty);
  return hipGetLastError();
}

// An instance whose store already holds an empty domain is failed before
// any sweep (the reference's is_bot() is true before the fixpoint).
template <bool BITS>
__global__ void k_check_empty(const int2* __restrict__ store, int64_t nvars, int64_t total,
                              InstFlags* flags) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < total;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int2 b = store[v];
    const bool empty = BITS ? (b.x == 0 && b.y == 0) : (b.x > b.y);
    if (empty) flags[v / nvars].failed_at = -1;
  }
}

hipError_t launch_store_check_empty(const int2* store, int64_t nvars, int32_t n_inst,
                                    InstFlags* flags, int dom, hipStream_t stream) {
  const int64_t total = nvars * (int64_t)n_inst;
  int64_t grid = (total + 255) / 256;
  if (grid > 16384) grid = 16384;
  if (grid <= 0) return hipSuccess;
  if (dom == kBits)
    k_check_empty<true><<<dim3((unsigned)grid), 256, 0, stream>>>(store, nvars, total, flags);
  else
    k_check_empty<false><<<dim3((unsigned)grid), 256, 0, stream>>>(store, nvars, total, flags);
  return hipGetLastError();
}

// Per-instance copies of the subsets: out[i] = table[idx[i]] for all i.
__global__ void k_compact(Table t, const int32_t* __restrict__ idx, int64_t n,
                          uint32_t* __restrict__ xo, int32_t* __restrict__ y,
                          int32_t* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t j = idx[i];
    xo[i] = t.xo[j];
    y[i] = t.y[j];
    z[i] = t.z[j];
  }
}

hipError_t launch_compact_table(const Table& t, const int32_t* idx, int64_t n, uint32_t* xo,
                                int32_t* y, int32_t* z, hipStream_t stream) {
  int64_t grid = (n + 255) / 256;
  if (grid > 65536) grid = 65536;
  if (grid > 0) k_compact<<<dim3((unsigned)grid), 256, 0, stream>>>(t, idx, n, xo, y, z);
  return hipGetLastError();
}

// ---- entailment over a batch ----------------------------------------------
// One workgroup per instance (grid-stride over instances).  count: the
// instance's propagators that are not entailed on its store, and whether the
// store holds an empty interval.  write: those propagators' table positions,
// ascending, at out_off[k] (a block scan per 256 propagators keeps the order).
__device__ __forceinline__ bool inst_prop_entailed(const Table& t, const int2* st,
                                                   const Subsets& sub, int64_t k, int64_t q,
                                                   int64_t& p) {
  p = sub.off ? (int64_t)sub.idx[sub.off[k] + q] : q;
  const uint32_t xo = t.xo[p];
  return entailed((int)(xo >> kOpShift), st[xo & kVarMask], st[t.y[p]], st[t.z[p]]);
}

__global__ __launch_bounds__(kBlock) void k_batch_ask_count(Table t, const int2* __restrict__ store,
                                                            int64_t nvars, Subsets sub,
                                                            int64_t n_props, int64_t n_inst,
                                                            int64_t* live, int32_t* empty) {
  __shared__ int s_cnt, s_empty;
  for (int64_t k = blockIdx.x; k < n_inst; k += gridDim.x) {
    if (threadIdx.x == 0) { s_cnt = 0; s_empty = 0; }
    __syncthreads();
    const int2* st = store + k * nvars;
    const int64_t m = sub.off ? sub.off[k + 1] - sub.off[k] : n_props;
    int c = 0;
    for (int64_t q = threadIdx.x; q < m; q += kBlock) {
      int64_t p;
      c += !inst_prop_entailed(t, st, sub, k, q, p);
    }
    bool e = false;
    for (int64_t v = threadIdx.x; v < nvars; v += kBlock) e |= st[v].x > st[v].y;
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(&s_cnt, c);
    if (__ballot(e) && lane_id() == 0) s_empty = 1;
    __syncthreads();
    if (threadIdx.x == 0) { live[k] = s_cnt; empty[k] = s_empty; }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_batch_ask_write(Table t, const int2* __restrict__ store,
                                                            int64_t nvars, Subsets sub,
                                                            int64_t n_props, int64_t n_inst,
                                                            const int64_t* out_off,
                                                            int32_t* out_idx) {
  __shared__ int s_scan[kBlock / kWave];
  for (int64_t k = blockIdx.x; k < n_inst; k += gridDim.x) {
    const int2* st = store + k * nvars;
    const int64_t m = sub.off ? sub.off[k + 1] - sub.off[k] : n_props;
    int64_t at = out_off[k];
    for (int64_t q0 = 0; q0 < m; q0 += kBlock) {               // same trip count in every thread
      const int64_t q = q0 + threadIdx.x;
      int64_t p = -1;
      const int keep = q < m && !inst_prop_entailed(t, st, sub, k, q, p);
      const int lane = lane_id(), wv = threadIdx.x / kWave;
      int incl = keep;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const int v = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += v;
      }
      if (lane == kWave - 1) s_scan[wv] = incl;
      __syncthreads();
      int off = incl - keep, total = 0;
#pragma unroll
      for (int i = 0; i < kBlock / kWave; ++i) {
        if (i < wv) off += s_scan[i];
        total += s_scan[i];
      }
      if (keep) out_idx[at + off] = (int32_t)p;
      at += total;
      __syncthreads();                                          // s_scan is rewritten next chunk
    }
  }
}

hipError_t launch_batch_ask(const Table& t, const int2* store, int64_t nvars, const Subsets& sub,
                            int64_t n_props, int64_t n_inst, int64_t* live, int32_t* empty,
                            const int64_t* out_off, int32_t* out_idx, hipStream_t stream) {
  const int64_t grid = n_inst < 65536 ? n_inst : 65536;
  if (grid <= 0) return hipSuccess;
  if (!out_idx)
    k_batch_ask_count<<<(unsigned)grid, kBlock, 0, stream>>>(t, store, nvars, sub, n_props, n_inst,
                                                            live, empty);
  else
    k_batch_ask_write<<<(unsigned)grid, kBlock, 0, stream>>>(t, store, nvars, sub, n_props, n_inst,
                                                            out_off, out_idx);
  return hipGetLastError();
}

// ---- branching: two children per instance ---------------------------------
// One pass per parent k (a workgroup each): children 2k and 2k+1 (stores
// and image) are written as copies of the parent while the workgroup finds
// var[k], the first variable (input order) that is not fixed (-1 when the
// store is empty or every variable is fixed), and whether the store is
// empty; then that variable is split in both children: intervals
// [lb, mid] / [mid+1, ub] (mid = lb + (ub - lb) / 2, floor), bitsets {min} /
// the rest.  The split halves are not empty, so each child is empty exactly
// when its parent is (child_empty, may be nullptr).
template <bool BITS>
__device__ __forceinline__ void branch_scan(int2 b, int32_t v, int32_t& first, bool& empty) {
  if (BITS) {
    const u64 wv = ((u64)(uint32_t)b.y << 32) | (uint32_t)b.x;
    empty |= wv == 0;
    if (wv & (wv - 1)) first = min(first, v);
  } else {
    empty |= b.x > b.y;
    if (b.x < b.y) first = min(first, v);
  }
}

// V2: two variables per 16-byte access (nvars even, rows 16-byte aligned).
template <bool BITS, bool V2>
__global__ __launch_bounds__(kBlock) void k_branch(const int2* __restrict__ parent, int64_t nvars,
                                                   int64_t n_inst, int32_t* __restrict__ var, int2* child,
                                                   int2* child_image, uint8_t* child_empty) {
  __shared__ int32_t s_first, s_empty;
  for (int64_t k = blockIdx.x; k < n_inst; k += gridDim.x) {
    if (threadIdx.x == 0) { s_first = 0x7fffffff; s_empty = 0; }
    __syncthreads();
    const int2* p = parent + k * nvars;
    int2* c0 = child + 2 * k * nvars;
    int2* c1 = c0 + nvars;
    int2* i0 = child_image + 2 * k * nvars;
    int2* i1 = i0 + nvars;
    int32_t first = 0x7fffffff;
    bool empty = false;
    if constexpr (V2) {
      const int4* p4 = reinterpret_cast<const int4*>(p);
      int4 *c04 = reinterpret_cast<int4*>(c0), *c14 = reinterpret_cast<int4*>(c1);
      int4 *i04 = reinterpret_cast<int4*>(i0), *i14 = reinterpret_cast<int4*>(i1);
      for (int64_t h = threadIdx.x; h < nvars / 2; h += kBlock) {
        const int4 b = p4[h];
        c04[h] = b;
        c14[h] = b;
        i04[h] = b;
        i14[h] = b;
        branch_scan<BITS>(make_int2(b.x, b.y), (int32_t)(2 * h), first, empty);
        branch_scan<BITS>(make_int2(b.z, b.w), (int32_t)(2 * h + 1), first, empty);
      }
    } else {
      for (int64_t v = threadIdx.x; v < nvars; v += kBlock) {
        const int2 b = p[v];
        c0[v] = b;
        c1[v] = b;
        i0[v] = b;
        i1[v] = b;
        branch_scan<BITS>(b, (int32_t)v, first, empty);
      }
    }
    if (first != 0x7fffffff) atomicMin(&s_first, first);
    if (__ballot(empty) && lane_id() == 0) s_empty = 1;
    __syncthreads();                             // the copies are written: the split overwrites them
    if (threadIdx.x == 0) {
      const int32_t sv = (s_empty || s_first == 0x7fffffff) ? -1 : s_first;
      var[k] = sv;
      if (child_empty) child_empty[2 * k] = child_empty[2 * k + 1] = (uint8_t)s_empty;
      if (sv >= 0) {
        const int2 b = p[sv];
        int2 left = b, right = b;
        if (BITS) {
          const u64 wv = ((u64)(uint32_t)b.y << 32) | (uint32_t)b.x;
          const u64 lo = wv & (~wv + 1), hi = wv & ~lo;
          left = make_int2((int32_t)(uint32_t)lo, (int32_t)(uint32_t)(lo >> 32));
          right = make_int2((int32_t)(uint32_t)hi, (int32_t)(uint32_t)(hi >> 32));
        } else {
          const int64_t mid = (int64_t)b.x + (((int64_t)b.y - (int64_t)b.x) >> 1);
          left.y = (int32_t)mid;
          right.x = (int32_t)(mid + 1);
        }
        c0[sv] = left;
        c1[sv] = right;
        i0[sv] = left;
        i1[sv] = right;
      }
    }
    __syncthreads();                             // s_first / s_empty read before the next parent
  }
}

// Seeds for children that start from their parent's fixpoint: child c's
// marks (plane 0, the first sweep's) are the member propagators adjacent to
// its parent's split variable.  One workgroup per child.
__global__ __launch_bounds__(kBlock) void k_branch_seed(EventSet ev, const int32_t* __restrict__ var,
                                                        int64_t n_child) {
  for (int64_t c = blockIdx.x; c < n_child; c += gridDim.x) {
    const int32_t v = var[c >> 1];
    if (v < 0) continue;                          // uniform: no split, no marks
    Marker mk = marker_for(&ev, c, -1);           // sweep -1 marks plane 0 (sweep 0's)
    for (int32_t i = ev.adj_off[v] + threadIdx.x; i < ev.adj_off[v + 1]; i += kBlock) {
      const int32_t p = ev.adj[i];
      const uint32_t bit = 1u << (p & 31);
      if (mk.member && !(mk.member[p >> 5] & bit)) continue;
      atomicOr(&mk.dirty[p >> 5], bit);
      if (mk.sum) set_summary(nullptr, mk.sum, p >> kChunkShift);
    }
  }
}

hipError_t launch_branch_seed(const EventSet& ev, const int32_t* var, int64_t n_child, hipStream_t stream) {
  const int64_t grid = n_child < 65536 ? n_child : 65536;
  if (grid > 0) k_branch_seed<<<(unsigned)grid, kBlock, 0, stream>>>(ev, var, n_child);
  return hipGetLastError();
}

// ---- first sweeps seeded from a base store --------------------------------
// Instances loaded together usually share most of their values (search nodes
// of one root).  A propagator whose three operands hold the same values in
// instance k as in a base store computes the same thing in both; if it does
// not tighten the base store, it does not tighten k's.  So a batch's first
// sweep needs only the propagators adjacent to a value that differs from the
// base, and those the base store does not satisfy.  The base holds, per
// variable, the value two of instances 0-2 agree on: the root's value in a
// batch of splits of one root.

// The base store (base_value per variable) from rows 0-2 of lb / ub, or of
// packed stores (lb == nullptr: rows of `store`).
__global__ __launch_bounds__(kBlock) void k_make_base(const int32_t* __restrict__ lb, const int32_t* __restrict__ ub,
                                                      const u64* __restrict__ store, int64_t nvars,
                                                      int64_t n_inst, u64* __restrict__ base) {
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < nvars; v += (int64_t)gridDim.x * kBlock) {
    auto row = [&](int64_t k) { return lb ? pack_bounds(lb[k * nvars + v], ub[k * nvars + v]) : store[k * nvars + v]; };
    const u64 a = row(0);
    base[v] = base_value(a, n_inst > 1 ? row(1) : a, n_inst > 2 ? row(2) : a, n_inst);
  }
}
hipError_t launch_make_base(const int32_t* lb, const int32_t* ub, const u64* store, int64_t nvars,
                            int64_t n_inst, u64* base, hipStream_t stream) {
  int64_t grid = (nvars + kBlock - 1) / kBlock;
  if (grid > 8192) grid = 8192;
  k_make_base<<<(unsigned)grid, kBlock, 0, stream>>>(lb, ub, store, nvars, n_inst, base);
  return hipGetLastError();
}

// The differences from the base store (made first) in stores already in
// device memory: one workgroup per (instance, 2048-variable block).
__global__ __launch_bounds__(kBlock) void k_list_diffs(const u64* __restrict__ store, int64_t nvars,
                                                       int64_t n_inst, int64_t vblocks, DiffSink ds) {
  constexpr int kPer = 8;
  const bool listing = diff_listing(ds);
  if (!listing) return;
  const int64_t k = blockIdx.x / vblocks;
  const int64_t v0 = (blockIdx.x % vblocks) * (kBlock * kPer);
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int64_t v = v0 + q * kBlock + threadIdx.x;
    bool d = false;
    u64 val = 0;
    if (v < nvars) {
      val = store[k * nvars + v];
      d = val != reinterpret_cast<const u64*>(ds.base)[v];
    }
    diff_append(ds, d, k, v, val);
  }
}

hipError_t launch_list_diffs(const int2* store, int64_t nvars, int64_t n_inst, const DiffSink& ds,
                             hipStream_t stream) {
  if (!ds.list || nvars <= 0 || n_inst < 2) return hipSuccess;
  const int64_t vblocks = (nvars + kBlock * 8 - 1) / (kBlock * 8);
  const int64_t grid = n_inst * vblocks;
  if (grid > 0x7fffffff) return hipErrorInvalidValue;
  const hipError_t e = launch_make_base(nullptr, nullptr, reinterpret_cast<const u64*>(store), nvars, n_inst,
                                        reinterpret_cast<u64*>(ds.base), stream);
  if (e != hipSuccess) return e;
  k_list_diffs<<<(unsigned)grid, kBlock, 0, stream>>>(reinterpret_cast<const u64*>(store), nvars, n_inst,
                                                      vblocks, ds);
  return hipGetLastError();
}

// Rows of the stores straight from lb / ub (no image copy), streamed like a
// copy: a grid-stride loop over groups of four values (nvars % 4 == 0,
// 16-byte aligned) or single values (a grid as large as the data measured
// fastest: 0.57 ms for 4,096 x 50,000 values against 0.67 ms with 4,096
// workgroups); each thread carries its (instance, variable) position from
// step to step.  The differences from the base store (made first) are listed
// with their values; a list that overflowed is not worth a check per
// workgroup here (the load then copies the image and later loads take the
// fused ingest).  With `held` (stores that hold these base values in every
// row) only the values that differ from it are written: the pass then reads
// the rows and writes a few values (single values only: k_ingest_held takes
// held loads four values at a time).
template <bool V4, bool HELD>
__global__ __launch_bounds__(kBlock) void k_ingest_rows(const int32_t* __restrict__ lb,
                                                        const int32_t* __restrict__ ub, u64* __restrict__ store,
                                                        int64_t nvars, int64_t n_inst,
                                                        uint8_t* __restrict__ empty, DiffSink ds,
                                                        const u64* __restrict__ held) {
  static_assert(!(V4 && HELD), "held loads of four values take k_ingest_held");
  constexpr int kPer = V4 ? 4 : 1;
  const bool listing = ds.list != nullptr;
  const u64* base = reinterpret_cast<const u64*>(ds.base);
  const int64_t n = n_inst * nvars;
  const int64_t stride = (int64_t)gridDim.x * kBlock * kPer;
  const int64_t dk = stride / nvars, dv = stride % nvars;
  int64_t i = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kPer;
  int64_t k = i / nvars, v = i % nvars;
  for (int64_t i0 = (int64_t)blockIdx.x * kBlock * kPer; i0 < n; i0 += stride) {
    const bool in = i < n;
    u64 w[kPer];
    bool bad = false;
    if (in) {
      if constexpr (V4) {
        const v4i a = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(lb + i));
        const v4i b = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(ub + i));
        w[0] = pack_bounds(a.x, b.x); w[1] = pack_bounds(a.y, b.y);
        w[2] = pack_bounds(a.z, b.z); w[3] = pack_bounds(a.w, b.w);
        const v4i p0 = {a.x, b.x, a.y, b.y}, p1 = {a.z, b.z, a.w, b.w};
        reinterpret_cast<v4i*>(store + i)[0] = p0;
        reinterpret_cast<v4i*>(store + i)[1] = p1;
        bad = (a.x > b.x) | (a.y > b.y) | (a.z > b.z) | (a.w > b.w);
      } else {
        const int32_t a = __builtin_nontemporal_load(lb + i), b = __builtin_nontemporal_load(ub + i);
        w[0] = pack_bounds(a, b);
        if (!HELD || w[0] != held[v]) store[i] = w[0];
        bad = a > b;
      }
      if (bad) empty[k] = 1;
    }
    if (listing) {
      // the base values in 16-byte loads (four scalar loads cost the
      // ingest 0.15 ms at 4,096 x 50,000 values, these nothing measurable)
      u64 bv[kPer];
      if constexpr (V4) {
        bv[0] = bv[1] = bv[2] = bv[3] = 0;
        if (in) {
          const ulonglong2 b0 = reinterpret_cast<const ulonglong2*>(base + v)[0];
          const ulonglong2 b1 = reinterpret_cast<const ulonglong2*>(base + v)[1];
          bv[0] = b0.x; bv[1] = b0.y; bv[2] = b1.x; bv[3] = b1.y;
        }
      } else {
        bv[0] = in ? base[v] : 0;
      }
      bool dif[kPer], any = false;
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        dif[j] = in && w[j] != bv[j];
        any |= dif[j];
      }
      if (__ballot(any)) {
#pragma unroll
        for (int j = 0; j < kPer; ++j) diff_append(ds, dif[j], k, v + j, in ? w[j] : 0);
      }
    }
    i += stride;
    k += dk;
    v += dv;
    if (v >= nvars) { v -= nvars; ++k; }
  }
}

// The same load into stores that hold the base store (`held`): one
// workgroup per (tile of 1,024 variables, chunk of kHeldChunk instances),
// each thread four variables whose held and base values it loads once, then
// the chunk's rows (4 KB of each row per array and workgroup).  Only the
// values that differ from `held` are written.  At 4,096 x 50,000 values a
// chunk of 4 rows measured 0.28 ms, of 2 0.29, 8 0.29, 16 0.30, 32 0.33,
// and one row per workgroup 0.36.
constexpr int kHeldChunk = 4;
__global__ __launch_bounds__(kBlock) void k_ingest_held(const int32_t* __restrict__ lb,
                                                        const int32_t* __restrict__ ub, u64* __restrict__ store,
                                                        int64_t nvars, int64_t n_inst, int64_t vtiles,
                                                        uint8_t* __restrict__ empty, DiffSink ds,
                                                        const u64* __restrict__ held) {
  const bool listing = ds.list != nullptr;
  const u64* base = reinterpret_cast<const u64*>(ds.base);
  const int64_t v = (blockIdx.x % vtiles) * (kBlock * 4) + threadIdx.x * 4;
  const int64_t k0 = (blockIdx.x / vtiles) * kHeldChunk;
  const int64_t k1 = min(n_inst, k0 + kHeldChunk);
  const bool in = v < nvars;                           // nvars % 4 == 0: then v + 3 < nvars
  u64 h[4] = {0, 0, 0, 0}, bv[4] = {0, 0, 0, 0};
  if (in) {
    const ulonglong2 h0 = reinterpret_cast<const ulonglong2*>(held + v)[0];
    const ulonglong2 h1 = reinterpret_cast<const ulonglong2*>(held + v)[1];
    h[0] = h0.x; h[1] = h0.y; h[2] = h1.x; h[3] = h1.y;
    if (listing) {
      const ulonglong2 b0 = reinterpret_cast<const ulonglong2*>(base + v)[0];
      const ulonglong2 b1 = reinterpret_cast<const ulonglong2*>(base + v)[1];
      bv[0] = b0.x; bv[1] = b0.y; bv[2] = b1.x; bv[3] = b1.y;
    }
  }
#pragma unroll 4
  for (int64_t k = k0; k < k1; ++k) {                  // the same trip count for the workgroup
    const int64_t i = k * nvars + v;
    u64 w[4] = {0, 0, 0, 0};
    if (in) {
      const v4i a = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(lb + i));
      const v4i b = __builtin_nontemporal_load(reinterpret_cast<const v4i*>(ub + i));
      w[0] = pack_bounds(a.x, b.x); w[1] = pack_bounds(a.y, b.y);
      w[2] = pack_bounds(a.z, b.z); w[3] = pack_bounds(a.w, b.w);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (w[j] != h[j]) store[i + j] = w[j];
      if ((a.x > b.x) | (a.y > b.y) | (a.z > b.z) | (a.w > b.w)) empty[k] = 1;
    }
    if (listing) {
      bool dif[4], any = false;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        dif[j] = in && w[j] != bv[j];
        any |= dif[j];
      }
      if (__ballot(any)) {                             // a sub-list per (workgroup, row)
        const int c = (int)((blockIdx.x * kHeldChunk + (k - k0)) % kDiffLists);
#pragma unroll
        for (int j = 0; j < 4; ++j) diff_append_to(ds, c, dif[j], k, v + j, w[j]);
      }
    }
  }
}

hipError_t launch_make_base_rows(const int32_t* lb, const int32_t* ub, int64_t nvars, int64_t n_inst,
                                 const DiffSink& ds, hipStream_t stream) {
  if (!ds.list || nvars <= 0 || n_inst <= 0) return hipSuccess;
  return launch_make_base(lb, ub, nullptr, nvars, n_inst, reinterpret_cast<u64*>(ds.base), stream);
}

hipError_t launch_ingest_rows(const int32_t* lb, const int32_t* ub, int2* store, int64_t nvars,
                              int64_t n_inst, uint8_t* empty, const DiffSink& ds, const int2* held,
                              hipStream_t stream) {
  if (nvars <= 0 || n_inst <= 0) return hipSuccess;
  const bool v4 = nvars % 4 == 0 &&
                  ((reinterpret_cast<uintptr_t>(lb) | reinterpret_cast<uintptr_t>(ub) |
                    reinterpret_cast<uintptr_t>(store) | reinterpret_cast<uintptr_t>(held)) & 15) == 0;
  const int64_t n = n_inst * nvars;
  int64_t grid = (n / (v4 ? 4 : 1) + kBlock - 1) / kBlock;
  if (grid > (1 << 20)) grid = 1 << 20;
  u64* st = reinterpret_cast<u64*>(store);
  const u64* h = reinterpret_cast<const u64*>(held);
  if (v4 && h) {
    const int64_t vtiles = (nvars + kBlock * 4 - 1) / (kBlock * 4);
    const int64_t g = vtiles * ((n_inst + kHeldChunk - 1) / kHeldChunk);
    if (g > 0x7fffffff) return hipErrorInvalidValue;
    k_ingest_held<<<(unsigned)g, kBlock, 0, stream>>>(lb, ub, st, nvars, n_inst, vtiles, empty, ds, h);
  } else if (v4) k_ingest_rows<true, false><<<(unsigned)grid, kBlock, 0, stream>>>(lb, ub, st, nvars, n_inst, empty, ds, h);
  else if (h) k_ingest_rows<false, true><<<(unsigned)grid, kBlock, 0, stream>>>(lb, ub, st, nvars, n_inst, empty, ds, h);
  else k_ingest_rows<false, false><<<(unsigned)grid, kBlock, 0, stream>>>(lb, ub, st, nvars, n_inst, empty, ds, h);
  return hipGetLastError();
}

// The image from a complete difference list: every row the base, then the
// listed values (a second launch, so that they land after the base).
__global__ __launch_bounds__(kBlock) void k_fill_base(const u64* __restrict__ base, u64* __restrict__ store,
                                                      int64_t nvars, int64_t vblocks) {
  const int64_t k = blockIdx.x / vblocks;
  const int64_t v0 = (blockIdx.x % vblocks) * (kBlock * 8);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int64_t v = v0 + q * kBlock + threadIdx.x;
    if (v < nvars) store[k * nvars + v] = base[v];
  }
}
// The listed differences, a group of workgroups per sub-list: workgroup b
// walks sub-list b % kDiffLists from chunk b / kDiffLists on, kBlock entries
// a chunk, up to the sub-list's count (read once), so the slots past the
// counts cost nothing.  The grid is a multiple of kDiffLists (diff_grid).
template <typename F>
__device__ __forceinline__ void for_listed_diffs(const DiffSink& ds, F f) {
  const int c = blockIdx.x % kDiffLists;
  const int64_t step = (int64_t)(gridDim.x / kDiffLists) * kBlock;
  const int64_t n = min<int64_t>(ds.count[c], ds.cap);
  for (int64_t at0 = (int64_t)(blockIdx.x / kDiffLists) * kBlock; at0 < n; at0 += step) {
    const int64_t at = at0 + threadIdx.x;
    if (at < n) {
      const int64_t slot = c * ds.cap + at;
      const u64 d = ds.list[slot];
      f(slot, (int64_t)(d >> 32), (uint32_t)d);
    }
  }
}
unsigned diff_grid(const DiffSink& ds) {
  const int64_t chunks = (ds.cap + kBlock - 1) / kBlock;
  return (unsigned)(kDiffLists * (chunks < 16 ? chunks : 16));
}

// Every listed difference's value into the stores.
__global__ __launch_bounds__(kBlock) void k_scatter_diffs(DiffSink ds, u64* __restrict__ store, int64_t nvars) {
  const u64* vals = reinterpret_cast<const u64*>(ds.vals);
  for_listed_diffs(ds, [&](int64_t slot, int64_t k, uint32_t v) { store[k * nvars + v] = vals[slot]; });
}

hipError_t launch_image_from_diffs(const DiffSink& ds, int2* store, int64_t nvars, int64_t n_inst,
                                   hipStream_t stream) {
  if (nvars <= 0 || n_inst <= 0) return hipSuccess;
  const int64_t vblocks = (nvars + kBlock * 8 - 1) / (kBlock * 8);
  const int64_t grid = n_inst * vblocks;
  if (grid > 0x7fffffff || !ds.vals) return hipErrorInvalidValue;
  u64* st = reinterpret_cast<u64*>(store);
  k_fill_base<<<(unsigned)grid, kBlock, 0, stream>>>(reinterpret_cast<const u64*>(ds.base), st, nvars, vblocks);
  k_scatter_diffs<<<diff_grid(ds), kBlock, 0, stream>>>(ds, st, nvars);
  return hipGetLastError();
}

// One workgroup per instance: its trail's variables (or, past the cap, its
// whole row) rewritten from the image: d_image, or the base store when the
// listed differences are scattered afterwards.
__global__ __launch_bounds__(kBlock) void k_restore_trail(InstLists tr, const u64* __restrict__ base,
                                                          const u64* __restrict__ image, u64* __restrict__ store,
                                                          int64_t nvars, int64_t n_inst) {
  for (int64_t k = blockIdx.x; k < n_inst; k += gridDim.x) {
    const int32_t n = tr.n[k];
    u64* row = store + k * nvars;
    const u64* irow = image ? image + k * nvars : nullptr;
    if (n <= tr.cap) {
      const int32_t* list = tr.list + k * tr.cap;
      for (int i = threadIdx.x; i < n; i += kBlock) {
        const int32_t v = list[i];
        row[v] = irow ? irow[v] : base[v];
      }
    } else {
      for (int64_t v = threadIdx.x; v < nvars; v += kBlock) row[v] = irow ? irow[v] : base[v];
    }
  }
}

hipError_t launch_restore_trail(const InstLists& tr, const DiffSink& ds, const int2* image, int2* store,
                                int64_t nvars, int64_t n_inst, hipStream_t stream) {
  if (nvars <= 0 || n_inst <= 0) return hipSuccess;
  if (!tr.list || (!image && !ds.vals)) return hipErrorInvalidValue;
  const int64_t grid = n_inst < 65536 ? n_inst : 65536;
  u64* st = reinterpret_cast<u64*>(store);
  k_restore_trail<<<(unsigned)grid, kBlock, 0, stream>>>(tr, reinterpret_cast<const u64*>(ds.base),
                                                         reinterpret_cast<const u64*>(image), st, nvars, n_inst);
  if (!image) k_scatter_diffs<<<diff_grid(ds), kBlock, 0, stream>>>(ds, st, nvars);
  return hipGetLastError();
}

// The base value at every listed difference's position.
__global__ __launch_bounds__(kBlock) void k_unscatter_diffs(DiffSink ds, u64* __restrict__ store, int64_t nvars) {
  const u64* base = reinterpret_cast<const u64*>(ds.base);
  for_listed_diffs(ds, [&](int64_t, int64_t k, uint32_t v) { store[k * nvars + v] = base[v]; });
}

hipError_t launch_stores_to_base(const InstLists* trail, const DiffSink& ds, int2* store, int64_t nvars,
                                 int64_t n_inst, hipStream_t stream) {
  if (nvars <= 0 || n_inst <= 0) return hipSuccess;
  if (!ds.list || !ds.count || !ds.base || (trail && !trail->list)) return hipErrorInvalidValue;
  u64* st = reinterpret_cast<u64*>(store);
  if (trail) {
    const int64_t grid = n_inst < 65536 ? n_inst : 65536;
    k_restore_trail<<<(unsigned)grid, kBlock, 0, stream>>>(*trail, reinterpret_cast<const u64*>(ds.base),
                                                           nullptr, st, nvars, n_inst);
  }
  k_unscatter_diffs<<<diff_grid(ds), kBlock, 0, stream>>>(ds, st, nvars);
  return hipGetLastError();
}

// One thread per propagator, on the base store alone (nothing is written
// back): positions of the propagators that would tighten it or read an
// empty domain.
template <int D>
__global__ __launch_bounds__(kBlock) void k_base_probe(Table t, const int2* __restrict__ base,
                                                       int32_t* __restrict__ list, int32_t* count) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t p0 = (int64_t)blockIdx.x * kBlock; p0 < t.n; p0 += stride) {
    const int64_t p = p0 + threadIdx.x;
    bool need = false;
    if (p < t.n) {
      const uint32_t xo = t.xo[p];
      const int op = (int)(xo >> kOpShift);
      if constexpr (D == kBits) {
        const u64* bw = reinterpret_cast<const u64*>(base);
        const u64 X = bw[xo & kVarMask], Y = bw[t.y[p]], Z = bw[t.z[p]];
        if (!X || !Y || !Z) {
          need = true;
        } else {
          u64 nx, ny, nz;
          deduce_bits(op, X, Y, Z, nx, ny, nz);
          need = (X & nx) != X || (Y & ny) != Y || (Z & nz) != Z;
        }
      } else {
        const int2 X = base[xo & kVarMask], Y = base[t.y[p]], Z = base[t.z[p]];
        if (X.x > X.y || Y.x > Y.y || Z.x > Z.y) {
          need = true;
        } else {
          int2 nx, ny, nz;
          deduce<true>(op, X, Y, Z, nx, ny, nz);
          need = nx.x > X.x || nx.y < X.y || ny.x > Y.x || ny.y < Y.y || nz.x > Z.x || nz.y < Z.y;
        }
      }
    }
    // the same wave-aggregated append, of table positions
    const u64 mask = __ballot(need);
    if (mask) {
      const int leader = __builtin_ctzll(mask);
      int b = 0;
      if (lane_id() == leader) b = atomicAdd(count, (int)__popcll(mask));
      b = __builtin_amdgcn_readlane(b, leader);
      const int at = b + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
      if (need && at < t.n) list[at] = (int32_t)p;   // each position is listed once: at < t.n
    }
  }
}

hipError_t launch_base_probe(const Table& t, const int2* base, int32_t* list, int32_t* count, int dom,
                             hipStream_t stream) {
  if (t.n <= 0) return hipSuccess;
  int64_t grid = (t.n + kBlock - 1) / kBlock;
  if (grid > 65536) grid = 65536;
  switch (dom) {
    case kItv: case kItvDiv: k_base_probe<kItv><<<(unsigned)grid, kBlock, 0, stream>>>(t, base, list, count); break;
    case kBits: k_base_probe<kBits><<<(unsigned)grid, kBlock, 0, stream>>>(t, base, list, count); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// One seed: propagator p of instance k marked for sweep 0 (plane 0), when
// its subset holds it; a bit this call sets is also staged for queue plane 0
// when the run sweeps from queues.
__device__ __forceinline__ void seed_prop(const Marker& mk, const EventSet& ev, const QueueStage& qs,
                                          const InstLists& sl, int64_t k, int32_t p) {
  const uint32_t bit = 1u << (p & 31);
  if (mk.member && !(mk.member[p >> 5] & bit)) return;
  if (sl.list) {                           // listed for the instance; past the cap, marked
    const int at = atomicAdd(&sl.n[k], 1);
    if (at < sl.cap) {
      sl.list[k * sl.cap + at] = p;
      return;
    }
  }
  if (!mk.sum && !ev.q) {                  // bits only: a non-returning atomic
    atomicOr(&mk.dirty[p >> 5], bit);
    return;
  }
  if (atomicOr(&mk.dirty[p >> 5], bit) & bit) return;
  if (mk.sum) set_summary(nullptr, mk.sum, p >> kChunkShift);
  if (ev.q) stage_entry(qs, ev, 0, ((unsigned long long)k << 32) | (uint32_t)p);
}

// Plane-0 marks, one block-wide step at a time (each step ends with a flush
// of the staged queue entries): steps over the difference sub-lists take up
// to 64 listed entries of one sub-list, four threads walking an entry's
// adjacency; the steps after them take kBlock (propagator of `unsat`,
// instance) pairs.
constexpr int kSeedSlotsPerStep = kBlock / 4;
__global__ __launch_bounds__(kBlock) void k_seed_from_base(EventSet ev, DiffSink ds,
                                                           const int32_t* __restrict__ unsat,
                                                           int64_t n_unsat, InstLists sl) {
  static_assert(kDiffLists == kBlock, "one thread per difference sub-list");
  __shared__ unsigned long long s_qbuf[kQStage];
  __shared__ int32_t s_qn;
  __shared__ int32_t s_scan[kBlock / kWave];
  __shared__ int32_t s_first[kDiffLists + 1];           // first step of each sub-list
  const QueueStage qs{s_qbuf, &s_qn};
  if (threadIdx.x == 0) s_qn = 0;
  const int own = (int)min<int64_t>(ds.count[threadIdx.x], ds.cap);
  int diff_steps;
  const int first = block_scan<kBlock>((own + kSeedSlotsPerStep - 1) / kSeedSlotsPerStep, s_scan, &diff_steps);
  s_first[threadIdx.x] = first;
  if (threadIdx.x == 0) s_first[kDiffLists] = diff_steps;
  __syncthreads();
  const int64_t n_pairs = n_unsat * ev.n_inst;
  const int64_t steps = diff_steps + (n_pairs + kBlock - 1) / kBlock;
  for (int64_t it = blockIdx.x; it < steps; it += gridDim.x) {
    if (it < diff_steps) {
      int lo = 0, hi = kDiffLists;                          // the sub-list c with s_first[c] <= it < s_first[c+1]
      while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (s_first[mid] <= it) lo = mid; else hi = mid;
      }
      const int c = lo;
      const int64_t at = (it - s_first[c]) * kSeedSlotsPerStep + threadIdx.x / 4;
      if (at < min<int64_t>(ds.count[c], ds.cap)) {
        const u64 d = ds.list[c * ds.cap + at];
        const int64_t k = (int64_t)(d >> 32);
        const int32_t v = (int32_t)(uint32_t)d;
        const Marker mk = marker_for(&ev, k, -1);          // sweep -1 marks plane 0 (sweep 0's)
        for (int32_t a = ev.adj_off[v] + (threadIdx.x & 3); a < ev.adj_off[v + 1]; a += 4)
          seed_prop(mk, ev, qs, sl, k, ev.adj[a]);
      }
    } else {
      const int64_t j = (it - diff_steps) * kBlock + threadIdx.x;
      if (j < n_pairs) {
        const int64_t k = j % ev.n_inst;
        seed_prop(marker_for(&ev, k, -1), ev, qs, sl, k, unsat[j / ev.n_inst]);
      }
    }
    if (ev.q) {
      __syncthreads();                                   // the step's entries are staged
      flush_stage(qs, ev, 0);
    }
  }
}

// The marked phase of a seeded run, one instance per workgroup of B threads:
// the workgroup owns the instance (its store, its bitmap row of plane 0 and
// its flags), so instead of device-wide sweeps it drains the instance's
// marks in rounds separated by workgroup barriers, and the whole phase is
// one launch.  Round r evaluates the list L_r; each variable it tightens
// puts its adjacent member propagators in L_{r+1}, once each (an LDS hash
// set of L_{r+1}).  A propagator marked while it is evaluated is listed
// again, so the rounds keep the marked-sweep invariant: at the end of round
// r every propagator not in L_{r+1} (or set in the row) is at its fixpoint
// on the store.  The bitmap row holds the seeds (round 0 scans it) and the
// marks that find L_{r+1} full (the next round also scans the row, for up to
// the list's capacity, and leaves the rest for a later scan).  The marks
// themselves are LDS operations: a round waits on memory for the table,
// the operands and the meets (phase B), then the adjacency (phase C).
// Workgroup-scope atomics and loads, as in the persistent sweeps.
constexpr int kWorkCap = 512;
constexpr int kDrainVars = 256;           // tightened variables listed per round
constexpr int kDrainHash = 1024;          // LDS hash slots of the next list

// List q for the next round: into the LDS hash set of the list (each
// propagator once) and at the end of the list; a propagator that finds the
// list full sets its bit in the row instead, for a scan.  Insertions start
// only while the list has room, so the set holds at most kWorkCap plus one
// per thread of the workgroup (< kDrainHash): a probe always ends.
__device__ __forceinline__ void drain_list(int32_t q, uint32_t* row, int32_t* hs, int32_t* list, int32_t* n,
                                           int32_t* over) {
  if (__hip_atomic_load(n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < kWorkCap) {
    uint32_t h = ((uint32_t)q * 2654435761u) >> 22;   // Fibonacci hashing: the top 10 bits
    for (;;) {
      const int32_t old = atomicCAS(&hs[h], -1, q);
      if (old == q) return;                            // listed already
      if (old == -1) break;
      h = (h + 1) & (kDrainHash - 1);
    }
    const int at = atomicAdd(n, 1);
    if (at < kWorkCap) {
      list[at] = q;
      return;
    }
  }
  __hip_atomic_fetch_or(&row[q >> 5], 1u << (q & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  *over = 1;
}

// The member propagators adjacent to v, listed for the next round; eight
// adjacency entries (and their member words) in flight at a time.
__device__ __forceinline__ void drain_mark(uint32_t* row, const uint32_t* mrow, const EventSet& ev, int32_t v,
                                           int32_t* hs, int32_t* list, int32_t* n, int32_t* over) {
  constexpr int kB = 8;
  const int32_t e = ev.adj_off[v + 1];
  for (int32_t i = ev.adj_off[v]; i < e; i += kB) {
    int32_t q[kB];
#pragma unroll
    for (int j = 0; j < kB; ++j) q[j] = i + j < e ? ev.adj[i + j] : -1;
    uint32_t in[kB];
#pragma unroll
    for (int j = 0; j < kB; ++j) in[j] = (q[j] >= 0 && mrow) ? mrow[q[j] >> 5] : ~0u;
#pragma unroll
    for (int j = 0; j < kB; ++j)
      if (q[j] >= 0 && ((in[j] >> (q[j] & 31)) & 1u)) drain_list(q[j], row, hs, list, n, over);
  }
}

// Round 0 reads the seeds with plain 16-byte loads, four in flight per
// thread (the row was written by earlier launches; the 16-byte chunks start
// at the aligned address below the row and stay inside the bitmap planes,
// the second of which follows this one); later scans read this workgroup's
// own atomics with workgroup-scope atomic loads.  Taken bits are cleared;
// bits beyond the list's room stay, and set *over.
template <int B>
__device__ __forceinline__ void drain_scan(uint32_t* row, int64_t words, bool first, int32_t* list,
                                           int32_t* n, int32_t* over) {
  constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
  const int tid = threadIdx.x;
  auto take = [&](int64_t w, uint32_t bits) {
    const int c = __builtin_popcount(bits);
    const int at = atomicAdd(n, c);
    uint32_t t = 0;
    int room = kWorkCap - at;
    for (uint32_t b = bits; b && room > 0; b &= b - 1, --room) t |= b & (~b + 1u);
    int j = at;
    for (uint32_t b = t; b; b &= b - 1) list[j++] = (int32_t)(w * 32 + __builtin_ctz(b));
    if (t) __hip_atomic_fetch_and(&row[w], ~t, __ATOMIC_RELAXED, WG);
    if (t != bits) *over = 1;                          // marks left for a later scan
  };
  const int o = (int)((reinterpret_cast<uintptr_t>(row) & 15) / 4);
  const uint4* row4 = reinterpret_cast<const uint4*>(row - o);
  const int64_t w4n = first ? (words + o + 3) / 4 : 0;
  for (int64_t g = tid; g < w4n; g += 4 * B) {
    uint4 v4[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v4[u] = g + u * B < w4n ? row4[g + u * B] : make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t ws[4] = {v4[u].x, v4[u].y, v4[u].z, v4[u].w};
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int64_t w = 4 * (g + u * B) + h - o;
        const uint32_t bits = (w >= 0 && w < words) ? ws[h] : 0u;
        if (bits) take(w, bits);
      }
    }
  }
  for (int64_t w = w4n ? words : tid; w < words; w += B) {
    const uint32_t bits = __hip_atomic_load(&row[w], __ATOMIC_RELAXED, WG);
    if (bits) take(w, bits);
  }
}

template <int D, int B>
__global__ __launch_bounds__(B) void k_drain_marks(Table t, int2* __restrict__ store, int64_t nvars,
                                                   InstFlags* flags, EventSet ev, int32_t* next_inst,
                                                   unsigned long long* tight, InstLists seeds,
                                                   InstLists trail) {
  constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
  __shared__ int32_t s_list[2][kWorkCap];
  __shared__ int32_t s_hash[kDrainHash];
  __shared__ int32_t s_tv[kDrainVars];
  __shared__ int32_t s_n[2], s_over[2];
  __shared__ int32_t s_k, s_chg, s_fail, s_ntv, s_tn;
  const int tid = threadIdx.x;
  int nt_all = 0;
  for (;;) {
    if (tid == 0) s_k = atomicAdd(next_inst, 1);
    __syncthreads();
    const int64_t k = s_k;
    __syncthreads();                                   // every thread has read s_k
    if (k >= ev.n_inst) break;
    InstFlags* f = flags + k;
    uint32_t* row = ev.dirty + k * ev.words;           // plane 0: the seeds, and marks that overflow
    if (f->failed_at != kNoFail) {                     // the same word for every thread: uniform
      for (int64_t w = tid; w < ev.words; w += B) row[w] = 0;   // its seeds: the row ends clear
      if (tid == 0 && trail.list) trail.n[k] = 0;
      continue;
    }
    int32_t* tlist = trail.list ? trail.list + k * trail.cap : nullptr;
    // a tightened variable on the trail (past the cap only counted)
    auto on_trail = [&](int32_t v) {
      if (!tlist) return;
      const int at = atomicAdd(&s_tn, 1);
      if (at < trail.cap) tlist[at] = v;
    };
    const uint32_t* mrow = ev.member ? ev.member + k * ev.words : nullptr;
    int2* st = store + k * nvars;
    // Round 0 takes the seed list (and scans the row when the list is
    // incomplete: the seeds past its cap were marked there), or scans the row.
    const int32_t m = seeds.list ? seeds.n[k] : 0;
    if (tid == 0) { s_n[0] = 0; s_over[0] = seeds.list ? m > seeds.cap : 1; s_tn = 0; }
    if (seeds.list) {
      for (int i = tid; i < kDrainHash; i += B) s_hash[i] = -1;
      __syncthreads();
      const int32_t* sl = seeds.list + k * seeds.cap;
      for (int i = tid; i < min(m, seeds.cap); i += B) drain_list(sl[i], row, s_hash, s_list[0], &s_n[0], &s_over[0]);
    }
    int cur = 0;
    int64_t evals = 0;
    for (int32_t r = 0;; ++r) {
      __syncthreads();
      const int nx = cur ^ 1;
      // (A) a round after an overflow (and round 0) also takes marks from the row
      if (s_over[cur]) {
        __syncthreads();                               // every thread has read s_over
        if (tid == 0) s_over[cur] = 0;
        __syncthreads();
        drain_scan<B>(row, ev.words, r == 0, s_list[cur], &s_n[cur], &s_over[cur]);
        __syncthreads();                               // the round's list is complete
      }
      const int n = min(s_n[cur], kWorkCap);
      if (n == 0) break;                               // nothing marked: the fixpoint
      if (tid == 0) { s_n[nx] = 0; s_over[nx] = s_over[cur]; s_chg = 0; s_fail = 0; s_ntv = 0; }
      for (int i = tid; i < kDrainHash; i += B) s_hash[i] = -1;
      __syncthreads();
      // (B) evaluate the list, listing the variables it tightens
      bool chg = false, fail = false;
      for (int i = tid; i < n; i += B) {
        const int32_t p = s_list[cur][i];
        const uint32_t xo = t.xo[p];
        const int32_t y = t.y[p], z = t.z[p], x = (int32_t)(xo & kVarMask);
        const u64 wx = load_word<WG>(st + x), wy = load_word<WG>(st + y), wz = load_word<WG>(st + z);
        bool c = false, fl = false, tx = false, ty = false, tz = false;
        int nt = 0;
        if constexpr (D == kBits) {
          if (!wx || !wy || !wz) {
            fl = true;
          } else {
            u64 nxw, nyw, nzw;
            deduce_bits((int)(xo >> kOpShift), wx, wy, wz, nxw, nyw, nzw);
            u64* sw = reinterpret_cast<u64*>(st);
            tx = meet_word<WG>(sw, x, wx, nxw, c, fl, nt);
            ty = meet_word<WG>(sw, y, wy, nyw, c, fl, nt);
            tz = meet_word<WG>(sw, z, wz, nzw, c, fl, nt);
          }
        } else {
          const int2 X = as_itv(wx), Y = as_itv(wy), Z = as_itv(wz);
          if (X.x > X.y || Y.x > Y.y || Z.x > Z.y) {
            fl = true;
          } else {
            int2 nxv, nyv, nzv;
            deduce<D == kItvDiv>((int)(xo >> kOpShift), X, Y, Z, nxv, nyv, nzv);
            if (nxv.x > nxv.y || nyv.x > nyv.y || nzv.x > nzv.y) fl = true;
            tx = meet_itv<WG>(st, x, X, nxv, c, fl, nt);
            ty = meet_itv<WG>(st, y, Y, nyv, c, fl, nt);
            tz = meet_itv<WG>(st, z, Z, nzv, c, fl, nt);
          }
        }
        ++evals;
        nt_all += nt;
        chg |= c;
        fail |= fl;
        const int32_t vs[3] = {x, y, z};
        const bool ts[3] = {tx, ty, tz};
        for (int j = 0; j < 3; ++j) {
          if (!ts[j]) continue;
          const int at = atomicAdd(&s_ntv, 1);
          if (at < kDrainVars) {
            s_tv[at] = vs[j];
          } else {
            drain_mark(row, mrow, ev, vs[j], s_hash, s_list[nx], &s_n[nx], &s_over[nx]);
            on_trail(vs[j]);
          }
        }
      }
      __syncthreads();                                 // the round's tightened variables are listed
      // (C) list their adjacency, a variable per thread
      const int ntv = min(s_ntv, kDrainVars);
      for (int i = tid; i < ntv; i += B) {
        drain_mark(row, mrow, ev, s_tv[i], s_hash, s_list[nx], &s_n[nx], &s_over[nx]);
        on_trail(s_tv[i]);
      }
      if (__ballot(chg) && lane_id() == 0) s_chg = 1;
      if (__ballot(fail) && lane_id() == 0) s_fail = 1;
      __syncthreads();                                 // the next list and the flags are in
      if (tid == 0) {
        if (s_chg) f->changed[r & 1] = r;
        if (s_fail) f->failed_at = r;
      }
      if (s_fail) {                                    // marks may be left: the row ends clear
        for (int64_t w = tid; w < ev.words; w += B) __hip_atomic_store(&row[w], 0u, __ATOMIC_RELAXED, WG);
        break;
      }
      cur = nx;
    }
    const int64_t we = wave_sum((int)evals);
    if (lane_id() == 0 && we) atomicAdd((unsigned long long*)&f->marked_evals, (unsigned long long)we);
    if (tid == 0 && tlist) trail.n[k] = s_tn;          // the rounds' writes of s_tn ended at a barrier
    __syncthreads();
  }
  const int sum = wave_sum(nt_all);
  if (lane_id() == 0 && sum)
    atomicAdd(&tight[blockIdx.x & (kTightSlots - 1)], (unsigned long long)sum);
}

hipError_t launch_drain_marks(const Table& t, int2* store, int64_t nvars, InstFlags* flags, const EventSet& ev,
                              int32_t* next_inst, unsigned long long* tight, const InstLists& seeds,
                              const InstLists& trail, int dom, int grid, hipStream_t stream) {
  if (ev.n_inst <= 0 || !ev.dirty) return hipSuccess;
  // Workgroups of two waves: at 4,096 interval instances of 50,000 variables
  // 0.228 ms against 0.236 with one wave and 0.255 with four (the replayed
  // step 0.33 ms against 0.35-0.36); at 1,024 bitset instances of 200,000
  // words 98 us against 105 with one wave.
  constexpr int kDrainB = 2 * kWave;
  switch (dom) {
    case kItv: k_drain_marks<kItv, kDrainB><<<grid, kDrainB, 0, stream>>>(t, store, nvars, flags, ev, next_inst, tight, seeds, trail); break;
    case kItvDiv: k_drain_marks<kItvDiv, kDrainB><<<grid, kDrainB, 0, stream>>>(t, store, nvars, flags, ev, next_inst, tight, seeds, trail); break;
    case kBits: k_drain_marks<kBits, kDrainB><<<grid, kDrainB, 0, stream>>>(t, store, nvars, flags, ev, next_inst, tight, seeds, trail); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_seed_from_base(const EventSet& ev, const DiffSink& diffs, int64_t n_diffs,
                                 const int32_t* unsat, int64_t n_unsat, const InstLists& seeds,
                                 hipStream_t stream) {
  // steps: at most one partial step per sub-list beyond the full ones
  const int64_t steps = n_diffs / kSeedSlotsPerStep + kDiffLists + (n_unsat * ev.n_inst + kBlock - 1) / kBlock;
  if (!ev.dirty || (n_diffs == 0 && n_unsat == 0)) return hipSuccess;
  const int64_t grid = steps > 4096 ? 4096 : steps;
  if (seeds.list && (ev.q || ev.summary)) return hipErrorInvalidValue;
  k_seed_from_base<<<(unsigned)grid, kBlock, 0, stream>>>(ev, diffs, unsat, n_unsat, seeds);
  return hipGetLastError();
}

hipError_t launch_branch(const int2* parent, int64_t nvars, int64_t n_inst, int32_t* var, int2* child,
                         int2* child_image, uint8_t* child_empty, int dom, hipStream_t stream) {
  if (n_inst <= 0) return hipSuccess;
  const int64_t grid = n_inst < 65536 ? n_inst : 65536;
  const bool v2 = nvars % 2 == 0 && ((reinterpret_cast<uintptr_t>(parent) | reinterpret_cast<uintptr_t>(child) |
                                      reinterpret_cast<uintptr_t>(child_image)) & 15) == 0;
  const unsigned g = (unsigned)grid;
  if (dom == kBits) {
    if (v2) k_branch<true, true><<<g, kBlock, 0, stream>>>(parent, nvars, n_inst, var, child, child_image, child_empty);
    else k_branch<true, false><<<g, kBlock, 0, stream>>>(parent, nvars, n_inst, var, child, child_image, child_empty);
  } else {
    if (v2) k_branch<false, true><<<g, kBlock, 0, stream>>>(parent, nvars, n_inst, var, child, child_image, child_empty);
    else k_branch<false, false><<<g, kBlock, 0, stream>>>(parent, nvars, n_inst, var, child, child_image, child_empty);
  }
  return hipGetLastError();
}

// ---- PC's formula trees: is_extractable --------------------------------------
// One thread per (instance, formula) interprets the formula's postorder
// nodes with a stack of int2: intervals for terms, {ask, nask} for formulas
// (PC's Formula::ask / nask, include/lala/formula.hpp; the arithmetic of
// the terms' projections over int32 bounds whose extremes are the
// infinities).  A formula not asked clears its instance's flag.
namespace tree {
constexpr int32_t kMinf = INT32_MIN, kInf = INT32_MAX;
__device__ __forceinline__ bool inf(i64 v) { return v == kMinf || v == kInf; }
__device__ __forceinline__ int32_t sat(i64 v) { return v <= kMinf ? kMinf : v >= kInf ? kInf : (int32_t)v; }
// a bound of a sum: the infinity it leans to wins over the other
__device__ __forceinline__ int32_t add(i64 a, i64 b, i64 first) {
  if (a == first || b == first) return (int32_t)first;
  if (inf(a)) return (int32_t)a;
  if (inf(b)) return (int32_t)b;
  return sat(a + b);
}
__device__ __forceinline__ int32_t neg(i64 a) { return a == kMinf ? kInf : a == kInf ? kMinf : (int32_t)-a; }
__device__ __forceinline__ int sgn(i64 a) { return (a > 0) - (a < 0); }
__device__ __forceinline__ i64 mul(i64 a, i64 b) {
  if (a == 0 || b == 0) return 0;
  if (inf(a) || inf(b)) return sgn(a) * sgn(b) > 0 ? kInf : kMinf;
  return a * b;
}
__device__ __forceinline__ i64 div(int kind, i64 a, i64 b) {
  if (inf(a)) return sgn(a) * sgn(b) > 0 ? kInf : kMinf;
  if (inf(b)) return 0;
  i64 q = a / b;                                   // truncated
  const i64 r = a - q * b;
  if (kind == PCX_FT_FDIV) { if (r != 0 && ((r < 0) != (b < 0))) --q; }
  else if (kind == PCX_FT_CDIV) { if (r != 0 && ((r < 0) == (b < 0))) ++q; }
  else if (kind == PCX_FT_EDIV) { if (r < 0) q += b > 0 ? -1 : 1; }
  return q;
}
}  // namespace tree

constexpr int kTreeDepth = 32;

__global__ __launch_bounds__(kBlock) void k_formula_ask(const int4* __restrict__ nodes, const int64_t* __restrict__ off,
                                                        int64_t n_formulas, const int2* __restrict__ store,
                                                        int64_t nvars, int64_t n_inst, int32_t* ok) {
  using tree::add;
  using tree::div;
  using tree::mul;
  using tree::neg;
  using tree::sat;
  constexpr int32_t kTMinf = tree::kMinf, kTInf = tree::kInf;
  // Threads take chunks of kBlock formulas of one instance, chunk by chunk
  // over all instances, so that once a formula of an instance is not asked
  // (is_extractable is false, as PC's loop returns at the first one) its
  // later chunks only read the flag.
  const int64_t chunks = (n_formulas + kBlock - 1) / kBlock;
  const int64_t total = chunks * n_inst * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock) {
    const int64_t row = i / kBlock, k = row % n_inst;
    const int64_t f = (row / n_inst) * kBlock + (i - row * kBlock);
    if (f >= n_formulas || !__hip_atomic_load(&ok[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;
    const int2* st = store + k * nvars;
    int2 s[kTreeDepth];
    int sp = 0;
    for (int64_t j = off[f]; j < off[f + 1]; ++j) {
      const int4 nd = nodes[j];
      switch (nd.x) {
        case PCX_FT_VAR: s[sp++] = st[nd.y]; break;
        case PCX_FT_CONST: s[sp++] = make_int2(nd.y, nd.y); break;
        case PCX_FT_TRUE: s[sp++] = make_int2(1, 0); break;
        case PCX_FT_FALSE: s[sp++] = make_int2(0, 1); break;
        case PCX_FT_NEG: { const int2 a = s[sp - 1]; s[sp - 1] = make_int2(neg(a.y), neg(a.x)); break; }
        case PCX_FT_ABS: {
          const int2 a = s[sp - 1];
          if (a.x >= 0) break;
          s[sp - 1] = a.y <= 0 ? make_int2(neg(a.y), neg(a.x)) : make_int2(0, max((int32_t)neg(a.x), a.y));
          break;
        }
        case PCX_FT_ADD: case PCX_FT_SUB: {
          const int2 a = s[sp - 2];
          int2 b = s[sp - 1];
          if (nd.x == PCX_FT_SUB) b = make_int2(neg(b.y), neg(b.x));
          s[sp - 2] = make_int2(add(a.x, b.x, kTMinf), add(a.y, b.y, kTInf));
          --sp;
          break;
        }
        case PCX_FT_MUL: {
          const int2 a = s[sp - 2], b = s[sp - 1];
          const i64 c0 = mul(a.x, b.x), c1 = mul(a.x, b.y), c2 = mul(a.y, b.x), c3 = mul(a.y, b.y);
          s[sp - 2] = make_int2(sat(min(min(c0, c1), min(c2, c3))), sat(max(max(c0, c1), max(c2, c3))));
          --sp;
          break;
        }
        case PCX_FT_TDIV: case PCX_FT_FDIV: case PCX_FT_CDIV: case PCX_FT_EDIV: {
          const int2 a = s[sp - 2], b = s[sp - 1];
          if (b.x <= 0 && 0 <= b.y) {
            s[sp - 2] = make_int2(kTMinf, kTInf);
          } else {
            const i64 c0 = div(nd.x, a.x, b.x), c1 = div(nd.x, a.x, b.y), c2 = div(nd.x, a.y, b.x),
                      c3 = div(nd.x, a.y, b.y);
            s[sp - 2] = make_int2(sat(min(min(c0, c1), min(c2, c3))), sat(max(max(c0, c1), max(c2, c3))));
          }
          --sp;
          break;
        }
        case PCX_FT_MIN: case PCX_FT_MAX: {
          const int2 a = s[sp - 2], b = s[sp - 1];
          s[sp - 2] = nd.x == PCX_FT_MIN ? make_int2(min(a.x, b.x), min(a.y, b.y))
                                         : make_int2(max(a.x, b.x), max(a.y, b.y));
          --sp;
          break;
        }
        case PCX_FT_LE: case PCX_FT_LT: case PCX_FT_EQ: case PCX_FT_NE: {
          const int2 l = s[sp - 2], r = s[sp - 1];
          int2 t;
          if (nd.x == PCX_FT_LE) t = make_int2(l.y <= r.x, l.x > r.y);
          else if (nd.x == PCX_FT_LT) t = make_int2(l.y < r.x, l.x >= r.y);
          else {
            const int same = l.x == l.y && r.x == r.y && l.x == r.x;
            const int apart = l.y < r.x || r.y < l.x;
            t = nd.x == PCX_FT_EQ ? make_int2(same, apart) : make_int2(apart, same);
          }
          s[sp - 2] = t;
          --sp;
          break;
        }
        case PCX_FT_AND: case PCX_FT_OR: {
          const int n = nd.y;
          int all_a = 1, any_a = 0, all_n = 1, any_n = 0;
          for (int q = sp - n; q < sp; ++q) {
            all_a &= s[q].x; any_a |= s[q].x; all_n &= s[q].y; any_n |= s[q].y;
          }
          sp -= n;
          s[sp++] = nd.x == PCX_FT_AND ? make_int2(all_a, any_n) : make_int2(any_a, all_n);
          break;
        }
        case PCX_FT_NOT: { const int2 a = s[sp - 1]; s[sp - 1] = make_int2(a.y, a.x); break; }
        case PCX_FT_EQUIV: case PCX_FT_XOR: {
          const int2 fm = s[sp - 2], g = s[sp - 1];
          const int same = (fm.x & g.x) | (fm.y & g.y), differ = (fm.x & g.y) | (fm.y & g.x);
          s[sp - 2] = nd.x == PCX_FT_EQUIV ? make_int2(same, differ) : make_int2(differ, same);
          --sp;
          break;
        }
        case PCX_FT_LIT: {
          const int2 a = s[sp - 1];
          s[sp - 1] = make_int2(!(a.x <= 0 && 0 <= a.y), a.x >= 0 && a.y <= 0);
          break;
        }
        case PCX_FT_PROP: {
          const int4 zn = nodes[j + 1];
          s[sp++] = make_int2(entailed(nd.y, st[nd.z], st[nd.w], st[zn.y]) ? 1 : 0, 0);
          ++j;                                       // the PROP_Z node
          break;
        }
        default: break;                              // (validated on the host)
      }
    }
    if (!s[0].x) __hip_atomic_store(&ok[k], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

hipError_t launch_formula_ask(const int4* nodes, const int64_t* off, int64_t n_formulas, const int2* store,
                              int64_t nvars, int64_t n_inst, int32_t* ok, hipStream_t stream) {
  if (n_inst <= 0 || n_formulas <= 0) return hipSuccess;
  int64_t grid = ((n_formulas + kBlock - 1) / kBlock) * n_inst;     // one workgroup per (chunk, instance)
  if (grid > 8192) grid = 8192;
  k_formula_ask<<<(unsigned)grid, kBlock, 0, stream>>>(nodes, off, n_formulas, store, nvars, n_inst, ok);
  return hipGetLastError();
}

__global__ void k_ask(Table t, const int2* __restrict__ st, uint8_t* ent, int32_t* all_flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = true;
  if (i < t.n) {
    const uint32_t xo = t.xo[i];
    ok = entailed((int)(xo >> kOpShift), st[xo & kVarMask], st[t.y[i]], st[t.z[i]]);
    if (ent) ent[i] = ok ? 1 : 0;
  }
  if (__ballot(!ok) && lane_id() == 0) *all_flag = 0;
}

hipError_t launch_ask(const Table& t, const int2* store, uint8_t* ent, int32_t* all_flag,
                      hipStream_t stream) {
  const int64_t grid = (t.n + 255) / 256;
  if (grid > 0) k_ask<<<dim3((unsigned)grid), 256, 0, stream>>>(t, store, ent, all_flag);
  return hipGetLastError();
}

__global__ void k_pack(const int32_t* __restrict__ lb, const int32_t* __restrict__ ub,
                       int2* __restrict__ st, int64_t n) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    st[v] = make_int2(lb[v], ub[v]);
}
__global__ void k_unpack(const int2* __restrict__ st, int32_t* __restrict__ lb,
                         int32_t* __restrict__ ub, int64_t n) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int2 b = st[v];
    lb[v] = b.x;
    ub[v] = b.y;
  }
}

// Device copy of n 8-byte words (batch reset: stores <- image): one
// non-temporal 16-byte load and store per thread per step, a large grid.
// Measured on the 1.64 GB config-2 reset: 562 us, against 657 us for
// hipMemcpyAsync (and 600-750 us with 2-8 loads in flight per thread).
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy_words(const v4u* __restrict__ src,
                                                    v4u* __restrict__ dst, int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

hipError_t launch_copy_words(const void* src, void* dst, int64_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int64_t n16 = n / 2;
  if (n16 > 0) {
    int64_t grid = (n16 + 255) / 256;
    if (grid > 131072) grid = 131072;
    k_copy_words<<<dim3((unsigned)grid), 256, 0, stream>>>(static_cast<const v4u*>(src),
                                                           static_cast<v4u*>(dst), n16);
  }
  if (n & 1) {                    // odd word count: the last word
    const hipError_t e = hipMemcpyAsync(static_cast<uint64_t*>(dst) + n - 1,
                                        static_cast<const uint64_t*>(src) + n - 1, 8,
                                        hipMemcpyDeviceToDevice, stream);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

// Stores loaded from device memory: image[v] = store[v] = {lb[v], ub[v]} and
// empty[k] = 1 when instance k holds some lb > ub (empty zeroed first), in
// one pass.  Four variables per thread and step: two 16-byte loads, four
// 16-byte stores; the source is read once, so its loads are non-temporal.
__global__ __launch_bounds__(256) void k_ingest4(const v4i* __restrict__ lb, const v4i* __restrict__ ub,
                                                 v4i* __restrict__ image, v4i* __restrict__ store,
                                                 int64_t n4, int64_t nvars, uint8_t* __restrict__ empty) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += stride) {
    const v4i a = __builtin_nontemporal_load(lb + i);
    const v4i b = __builtin_nontemporal_load(ub + i);
    const v4i p0 = {a.x, b.x, a.y, b.y};
    const v4i p1 = {a.z, b.z, a.w, b.w};
    image[2 * i] = p0;
    image[2 * i + 1] = p1;
    store[2 * i] = p0;
    store[2 * i + 1] = p1;
    const int bad = (a.x > b.x) | ((a.y > b.y) << 1) | ((a.z > b.z) << 2) | ((a.w > b.w) << 3);
    if (bad)
      for (int j = 0; j < 4; ++j)
        if (bad >> j & 1) empty[(4 * i + j) / nvars] = 1;
  }
}
// The same one variable at a time: sources that are not 16-byte aligned, and
// the tail of n % 4 variables.
__global__ void k_ingest1(const int32_t* __restrict__ lb, const int32_t* __restrict__ ub,
                          int2* __restrict__ image, int2* __restrict__ store, int64_t from, int64_t n,
                          int64_t nvars, uint8_t* __restrict__ empty) {
  for (int64_t v = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int2 p = make_int2(lb[v], ub[v]);
    image[v] = p;
    store[v] = p;
    if (p.x > p.y) empty[v / nvars] = 1;
  }
}

hipError_t launch_ingest(const int32_t* lb, const int32_t* ub, int2* image, int2* store, int64_t n,
                         int64_t nvars, uint8_t* empty, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const bool aligned = ((reinterpret_cast<uintptr_t>(lb) | reinterpret_cast<uintptr_t>(ub) |
                         reinterpret_cast<uintptr_t>(image) | reinterpret_cast<uintptr_t>(store)) & 15) == 0;
  int64_t done = 0;
  if (aligned && n >= 4) {
    const int64_t n4 = n / 4;
    int64_t grid = (n4 + 255) / 256;
    if (grid > 65536) grid = 65536;
    k_ingest4<<<dim3((unsigned)grid), 256, 0, stream>>>(reinterpret_cast<const v4i*>(lb),
                                                        reinterpret_cast<const v4i*>(ub),
                                                        reinterpret_cast<v4i*>(image),
                                                        reinterpret_cast<v4i*>(store), n4, nvars, empty);
    done = 4 * n4;
  }
  if (done < n) {
    int64_t grid = (n - done + 255) / 256;
    if (grid > 65536) grid = 65536;
    k_ingest1<<<dim3((unsigned)grid), 256, 0, stream>>>(lb, ub, image, store, done, n, nvars, empty);
  }
  return hipGetLastError();
}

hipError_t launch_store_pack(const int32_t* lb, const int32_t* ub, int2* store, int64_t n,
                             hipStream_t stream) {
  int64_t grid = (n + 255) / 256;
  if (grid > 65536) grid = 65536;
  if (grid > 0) k_pack<<<dim3((unsigned)grid), 256, 0, stream>>>(lb, ub, store, n);
  return hipGetLastError();
}
hipError_t launch_store_unpack(const int2* store, int32_t* lb, int32_t* ub, int64_t n,
                               hipStream_t stream) {
  int64_t grid = (n + 255) / 256;
  if (grid > 65536) grid = 65536;
  if (grid > 0) k_unpack<<<dim3((unsigned)grid), 256, 0, stream>>>(store, lb, ub, n);
  return hipGetLastError();
}

}  // namespace pcx
