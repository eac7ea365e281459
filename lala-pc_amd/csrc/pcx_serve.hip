// pcx_serve.hip -- continuous-batching server for propagation requests.
//
// Requests (a store over one of the registered networks) wait in a host
// queue.  Each tick admits queued requests into free device slots while the
// store arena has room (admission by HBM footprint), launches a few sweeps
// of the heterogeneous slot batch -- every live request advances by one sweep
// per launch, whatever its network -- and hands back the requests that
// reached their fixpoint, which frees their slot and arena block for the
// next tick.  Requests that arrive while others are on the GPU thus join the
// running batch instead of waiting for it to drain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "pcx_internal.h"

using namespace pcx;
using pcx_internal::fail_with;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over the store arena (units: 8-byte store entries).
class Arena {
 public:
  void init(int64_t words) { free_.clear(); if (words > 0) free_[0] = words; }
  int64_t alloc(int64_t n) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const int64_t off = it->first, rest = it->second - n;
        free_.erase(it);
        if (rest > 0) free_[off + n] = rest;
        return off;
      }
    }
    return -1;
  }
  void release(int64_t off, int64_t n) {
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
 private:
  std::map<int64_t, int64_t> free_;
};

// Units of 8 bytes an admitted request takes in the arena: its store, then
// (marked sweeps) two mark bitmaps, two summaries and a snapshot.
struct Footprint {
  int64_t words = 0, swords = 0, units = 0;
};

struct Request {
  int32_t net;
  int32_t* lb;
  int32_t* ub;
  int64_t nvars;
  int64_t tag;
  double t0;            // submit or arrival time (latency origin)
};

struct Slot {
  bool busy = false;
  bool side = false;    // owned by a side lane (bursts) instead of the slot batch
  int32_t lane = 0;     // that side lane
  Request req{};
  int64_t off = 0;      // arena offset
  int64_t units = 0;    // arena units held
  int32_t start = 0;    // first sweep number
  int32_t skipped = 0;  // slot-batch sweeps run while the slot was in the side lane
  int32_t retry_at = 0; // earliest sweep to try the side lane again after a bail
};

constexpr int32_t kFreeFailed = -2;   // failed_at marker of a free slot

// A network's told store and root fixpoint (pcx_server_set_root).
struct NetRoot {
  int2* told = nullptr;
  int2* root = nullptr;
  int64_t nvars = 0;
};

// out[0] = 1 when the store leaves the told store somewhere.
__global__ __launch_bounds__(256) void k_root_within(const int2* __restrict__ st, const int2* __restrict__ told,
                                                     int64_t nvars, int32_t* out) {
  bool outside = false;
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < nvars; v += (int64_t)gridDim.x * 256) {
    const int2 a = st[v], t = told[v];
    outside |= a.x < t.x || a.y > t.y;
  }
  if (__ballot(outside) && (threadIdx.x & 63) == 0) *out = 1;
}

// Unless out[0] is set: the store meets the root fixpoint.
__global__ __launch_bounds__(256) void k_root_meet(int2* __restrict__ st, const int2* __restrict__ root,
                                                   int64_t nvars, const int32_t* out) {
  if (*out) return;
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < nvars; v += (int64_t)gridDim.x * 256) {
    const int2 a = st[v], r = root[v];
    st[v] = make_int2(max(a.x, r.x), min(a.y, r.y));
  }
}

// Side lanes: concurrent burst launches, each on a stream of its own (with
// the slot batch's stream, four streams: the hardware queues of a process).
constexpr int kSideLanes = 3;
struct SideLane {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;        // burst and flag read-back done
  InstFlags* h_flags = nullptr;     // pinned; valid for the lane's slots once done
  bool busy = false;
};

}  // namespace

struct pcx_server {
  pcx_server_config cfg{};
  hipStream_t stream = nullptr;
  std::vector<pcx_engine*> nets;
  int32_t n_slots = 0;
  SlotDesc* d_slots = nullptr;
  InstFlags* d_flags = nullptr;
  int32_t* d_act = nullptr;
  int64_t* d_item_base = nullptr;
  int32_t* d_ctl = nullptr;
  unsigned long long* d_tight = nullptr;
  int2* d_arena = nullptr;
  Arena arena;
  std::vector<Slot> slots;
  SlotDesc* h_slots = nullptr;        // pinned mirrors
  InstFlags* h_flags = nullptr;
  int2* h_stage = nullptr;            // pinned staging ring for store uploads
  int64_t stage_words = 0;
  std::deque<Request> pending;
  std::vector<pcx_served> done;
  int32_t next_sweep = 0;
  pcx_server_stats stats{};
  // a tick's sweeps, per kernel domain: [dom][0] sweeps_per_tick sweeps,
  // [dom][1] one sweep (ticks that admitted requests)
  hipGraphExec_t tick_graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  // Side lanes: bursts of long-running slots beside the slot batch.
  SideLane lanes[kSideLanes];
  hipEvent_t handed = nullptr;        // slot batch: hand-over flags uploaded
  std::vector<NetRoot> roots;         // per network (pcx_server_set_root)
  int32_t* d_root_out = nullptr;      // per slot: its store left the told store
};

namespace {

// Marked sweeps per request.  A small request is done in a few cheap full
// sweeps, and full sweeps keep the propagation within a sweep that marked
// sweeps give up; a large one re-evaluates its whole table every sweep for
// ~100 sweeps, which is most of the server's work.  Automatic mode marks
// requests of at least kSlotMarkedMin propagators.
constexpr int64_t kSlotMarkedMin = (int64_t)1 << 17;

bool events_for(const pcx_server* s, int32_t net) {
  if (s->cfg.event_driven == 0) return false;
  const int64_t m = (int64_t)s->nets[net]->op.size();
  return m > 0 && (s->cfg.event_driven > 0 || m >= kSlotMarkedMin);
}

Footprint footprint(const pcx_server* s, const Request& rq) {
  Footprint fp;
  fp.units = rq.nvars;
  if (events_for(s, rq.net)) {
    const int64_t m = (int64_t)s->nets[rq.net]->op.size();
    fp.words = (m + 31) / 32;
    const int64_t chunks = (fp.words + kChunkWords - 1) / kChunkWords;
    fp.swords = (chunks + 31) / 32;
    fp.units += (2 * fp.words + 3 * fp.swords + 1) / 2;
  }
  return fp;
}

WorkList worklist(pcx_server* s) {
  return worklist_of(s->d_act, s->d_item_base, s->d_ctl);
}

pcx_status upload_slot_meta(pcx_server* s, int32_t k) {
  PCX_HIP(hipMemcpyAsync(s->d_slots + k, s->h_slots + k, sizeof(SlotDesc), hipMemcpyHostToDevice, s->stream));
  PCX_HIP(hipMemcpyAsync(s->d_flags + k, s->h_flags + k, sizeof(InstFlags), hipMemcpyHostToDevice, s->stream));
  return PCX_OK;
}

void finish(pcx_server* s, const Request& rq, const pcx_result& r) {
  pcx_served out{};
  out.tag = rq.tag;
  out.result = r;
  out.latency_ms = (now_s() - rq.t0) * 1e3;
  s->done.push_back(out);
  ++s->stats.completed;
}

// Move queued requests into free slots while the arena and staging ring
// have room.  Stores with an empty domain finish at once (failed, 0 sweeps).
pcx_status admit(pcx_server* s) {
  int64_t staged = 0;
  while (!s->pending.empty()) {
    Request& rq = s->pending.front();
    bool empty = false;
    for (int64_t v = 0; v < rq.nvars && !empty; ++v) empty = rq.lb[v] > rq.ub[v];
    if (empty) {
      pcx_result r{};
      r.failed = 1;
      finish(s, rq, r);
      s->pending.pop_front();
      continue;
    }
    int32_t k = -1;
    for (int32_t i = 0; i < s->n_slots; ++i)
      if (!s->slots[i].busy) { k = i; break; }
    if (k < 0) { ++s->stats.waits_slot; break; }
    if (staged + rq.nvars > s->stage_words) { ++s->stats.waits_stage_admit; break; }
    const Footprint fp = footprint(s, rq);
    const int64_t off = s->arena.alloc(fp.units);
    if (off < 0) { ++s->stats.waits_arena; break; }   // no HBM room: wait for a slot to drain
    int2* stage = s->h_stage + staged;
    for (int64_t v = 0; v < rq.nvars; ++v) stage[v] = make_int2(rq.lb[v], rq.ub[v]);
    PCX_HIP(hipMemcpyAsync(s->d_arena + off, stage, rq.nvars * sizeof(int2), hipMemcpyHostToDevice, s->stream));
    staged += rq.nvars;
    if (rq.net < (int32_t)s->roots.size() && s->roots[rq.net].root && s->roots[rq.net].nvars == rq.nvars) {
      // inside the told store: start from the meet with the root fixpoint
      const NetRoot& nr = s->roots[rq.net];
      const unsigned g = (unsigned)std::min<int64_t>((rq.nvars + 255) / 256, 2048);
      PCX_HIP(hipMemsetAsync(s->d_root_out + k, 0, sizeof(int32_t), s->stream));
      k_root_within<<<g, 256, 0, s->stream>>>(s->d_arena + off, nr.told, rq.nvars, s->d_root_out + k);
      k_root_meet<<<g, 256, 0, s->stream>>>(s->d_arena + off, nr.root, rq.nvars, s->d_root_out + k);
      PCX_HIP(hipGetLastError());
    }
    pcx_engine* e = s->nets[rq.net];
    const int32_t S = s->next_sweep;
    SlotDesc sd{};
    sd.xo = e->d_xo;
    sd.y = e->d_y;
    sd.z = e->d_z;
    sd.m = (int64_t)e->op.size();
    sd.store = s->d_arena + off;
    sd.start = S;
    if (fp.words > 0) {
      sd.adj_off = e->d_adj_off;
      sd.adj = e->d_adj;
      sd.marks = reinterpret_cast<uint32_t*>(s->d_arena + off + rq.nvars);
      sd.words = fp.words;
      sd.swords = fp.swords;
      PCX_HIP(hipMemsetAsync(sd.marks, 0, (2 * fp.words + 3 * fp.swords) * sizeof(uint32_t), s->stream));
    }
    s->h_slots[k] = sd;
    InstFlags f{};
    f.changed[(S - 1) & 1] = S - 1;            // primes sweep S
    f.changed[S & 1] = S - 2;
    f.failed_at = kNoFail;
    f.done_at = -1;
    f.settled_at = -1;
    f.all_marked[0] = f.all_marked[1] = -1;
    s->h_flags[k] = f;
    pcx_status st = upload_slot_meta(s, k);
    if (st != PCX_OK) return st;
    Slot& sl = s->slots[k];
    sl.busy = true;
    sl.req = rq;
    sl.off = off;
    sl.units = fp.units;
    sl.start = S;
    sl.side = false;
    sl.skipped = 0;
    sl.retry_at = 0;
    s->pending.pop_front();
    ++s->stats.admitted;
    s->stats.arena_peak_bytes = std::max(s->stats.arena_peak_bytes,
                                         (int64_t)((off + fp.units) * sizeof(int2)));
  }
  return PCX_OK;
}

// A slot's result from its flags: slot-batch sweeps up to the failing /
// settling / last changing one (plus the quiet sweep that found nothing more
// to do), plus every burst round, minus the slot-batch sweeps run while the
// slot was in the side lane.
pcx_result result_of(const pcx_server* s, int32_t k, const InstFlags& f) {
  const Slot& sl = s->slots[k];
  const int32_t last = std::max(f.changed[0], f.changed[1]);
  pcx_result r{};
  r.failed = f.failed_at != kNoFail;
  r.has_changed = last >= sl.start;
  if (r.failed)
    r.sweeps = (int64_t)f.failed_at - sl.start + 1;
  else if (f.settled_at >= 0)
    r.sweeps = (int64_t)f.settled_at - sl.start + 1;
  else
    r.sweeps = (int64_t)std::max(last, sl.start - 1) - sl.start + 2;
  r.sweeps += f.extra_sweeps - sl.skipped;
  if (s->h_slots[k].adj)          // one full sweep + the marked evaluations
    r.evals = (r.sweeps > 0 ? s->h_slots[k].m : 0) + f.marked_evals;
  else
    r.evals = r.sweeps * s->h_slots[k].m;
  return r;
}

// Hand back finished slots (flags in fl): their stores are copied back
// together (one synchronization), then their slot and arena block are freed.
pcx_status complete(pcx_server* s, const std::vector<int32_t>& ks, const InstFlags* fl) {
  if (ks.empty()) return PCX_OK;
  std::vector<int64_t> at;
  int64_t staged = 0;
  for (int32_t k : ks) {
    const Slot& sl = s->slots[k];
    PCX_HIP(hipMemcpyAsync(s->h_stage + staged, s->d_arena + sl.off, sl.req.nvars * sizeof(int2),
                           hipMemcpyDeviceToHost, s->stream));
    at.push_back(staged);
    staged += sl.req.nvars;
  }
  PCX_HIP(hipStreamSynchronize(s->stream));
  for (size_t i = 0; i < ks.size(); ++i) {
    const int32_t k = ks[i];
    Slot& sl = s->slots[k];
    pcx_result r = result_of(s, k, fl[k]);
    const int2* stage = s->h_stage + at[i];
    bool moved = false;                          // the fixpoint differs from the request's store
    for (int64_t v = 0; v < sl.req.nvars; ++v) {
      moved |= sl.req.lb[v] != stage[v].x || sl.req.ub[v] != stage[v].y;
      sl.req.lb[v] = stage[v].x;
      sl.req.ub[v] = stage[v].y;
    }
    // (a request started from its meet with a root fixpoint may change
    // nothing on the device and still differ from its own store)
    if (!r.failed) r.has_changed = moved;
    finish(s, sl.req, r);
    s->arena.release(sl.off, sl.units);
    sl.busy = false;
    sl.side = false;
    // park the slot: never scheduled until re-admitted
    s->h_flags[k].failed_at = kFreeFailed;
    s->h_flags[k].done_at = 0;
    pcx_status st = upload_slot_meta(s, k);
    if (st != PCX_OK) return st;
  }
  return PCX_OK;
}

// Slots of the slot batch whose request reached its fixpoint or failed
// (flags of this tick's read-back), as many as the staging buffer holds.
pcx_status harvest(pcx_server* s) {
  std::vector<int32_t> done;
  int64_t staged = 0;
  for (int32_t k = 0; k < s->n_slots; ++k) {
    const Slot& sl = s->slots[k];
    if (!sl.busy || sl.side) continue;
    const InstFlags& f = s->h_flags[k];
    const int32_t last = std::max(f.changed[0], f.changed[1]);
    const bool failed = f.failed_at != kNoFail;
    const bool quiet = last < s->next_sweep - 1;     // the latest sweep left it unchanged
    if (!failed && !quiet) continue;
    if (staged + sl.req.nvars > s->stage_words) {        // the rest next tick
      ++s->stats.waits_stage_harvest;
      break;
    }
    done.push_back(k);
    staged += sl.req.nvars;
  }
  return complete(s, done, s->h_flags);
}

int32_t busy_slots(const pcx_server* s) {
  int32_t n = 0;
  for (const Slot& sl : s->slots) n += sl.busy;
  return n;
}

// Side lanes.  A slot that has run burst_after sweeps and still changes
// leaves the slot batch (done_at = kSideLane + lane: the scheduler skips it)
// and is propagated by k_burst_slots on that lane's stream, concurrently
// with the slot batch's ticks, up to burst_sweeps rounds per launch.  Each
// lane has at most one launch in flight; the host looks at the lanes at each
// tick.  For each slot of a finished launch:
//  - settled or failed: the request completes;
//  - budget spent: the slot stays in the lane for its next launch;
//  - the burst stopped (its marks do not fit a workgroup): the slot rejoins
//    the slot batch at the batch's next sweep.  Its marks are for the sweep
//    after its hand-over; when the parities differ they move to the other
//    bitmap plane.  The batch sweeps it skipped are not its sweeps.
pcx_status side_collect(pcx_server* s) {
  std::vector<int32_t> done;
  int64_t staged = 0;
  const int32_t S1 = s->next_sweep - 1;          // the slot batch's last sweep
  for (int L = 0; L < kSideLanes; ++L) {
    SideLane& ln = s->lanes[L];
    if (!ln.busy) continue;
    const hipError_t q = hipEventQuery(ln.done);
    if (q == hipErrorNotReady) continue;
    PCX_HIP(q);
    ln.busy = false;
    for (int32_t k = 0; k < s->n_slots; ++k) {
      Slot& sl = s->slots[k];
      if (!sl.busy || !sl.side || sl.lane != L) continue;
      InstFlags f = ln.h_flags[k];
      if (f.failed_at != kNoFail || f.settled_at >= 0) {
        if (staged + sl.req.nvars <= s->stage_words) {
          s->h_flags[k] = f;                     // the flags its result is read from
          done.push_back(k);
          staged += sl.req.nvars;
        } else {
          ++s->stats.waits_stage_side;           // a full staging buffer: the lane's next
        }                                        // launch finds it settled again
        continue;
      }
      if (f.done_at != kSideBail) continue;      // budget spent: next launch
      const SlotDesc& sd = s->h_slots[k];
      const int32_t S0 = sd.handover;
      if (sd.adj && ((S0 + 1) & 1) != ((S1 + 1) & 1)) {
        const int64_t from = (S0 + 1) & 1, to = (S1 + 1) & 1;
        uint32_t* mk = sd.marks;
        PCX_HIP(hipMemcpyAsync(mk + to * sd.words, mk + from * sd.words, sd.words * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, s->stream));
        PCX_HIP(hipMemcpyAsync(mk + 2 * sd.words + to * sd.swords, mk + 2 * sd.words + from * sd.swords,
                               sd.swords * sizeof(uint32_t), hipMemcpyDeviceToDevice, s->stream));
        PCX_HIP(hipMemsetAsync(mk + from * sd.words, 0, sd.words * sizeof(uint32_t), s->stream));
        PCX_HIP(hipMemsetAsync(mk + 2 * sd.words + from * sd.swords, 0, sd.swords * sizeof(uint32_t),
                               s->stream));
      }
      const bool all = f.all_marked[(S0 + 1) & 1] == S0 + 1;
      f.changed[S1 & 1] = S1;                    // primes sweep S1 + 1
      f.changed[(S1 + 1) & 1] = S1 - 1;
      f.all_marked[0] = f.all_marked[1] = -1;
      if (all) f.all_marked[(S1 + 1) & 1] = S1 + 1;   // sweep S1 + 1 is then a full sweep
      f.done_at = -1;
      s->h_flags[k] = f;
      sl.side = false;
      sl.skipped += S1 - S0;
      sl.retry_at = S1 + std::max(s->cfg.burst_after, 4 * s->cfg.sweeps_per_tick);
      pcx_status st = upload_slot_meta(s, k);
      if (st != PCX_OK) return st;
    }
  }
  return complete(s, done, s->h_flags);
}

// Hand newly eligible slots of the slot batch to a free lane, and relaunch
// free lanes whose slots spent their budget.
pcx_status side_launch(pcx_server* s) {
  if (s->cfg.burst_sweeps <= 0) return PCX_OK;
  const int32_t S = s->next_sweep - 1;
  std::vector<int32_t> fresh;
  for (int32_t k = 0; k < s->n_slots; ++k) {
    const Slot& sl = s->slots[k];
    if (!sl.busy || sl.side) continue;
    const InstFlags& f = s->h_flags[k];
    const SlotDesc& sd = s->h_slots[k];
    const bool running = f.failed_at == kNoFail && f.done_at < 0 && f.changed[S & 1] == S &&
                         S - sl.start >= s->cfg.burst_after && S >= sl.retry_at;
    const bool fits = sd.adj ? sd.swords <= kBurstSummaryMax : sd.m <= kBurstFullMax;
    if (running && fits) fresh.push_back(k);
  }
  bool uploaded = false;
  for (int L = 0; L < kSideLanes; ++L) {
    SideLane& ln = s->lanes[L];
    if (ln.busy) continue;
    if (!fresh.empty()) {
      for (int32_t k : fresh) {
        s->h_flags[k].done_at = kSideLane + L;
        s->h_slots[k].handover = S;
        s->slots[k].side = true;
        s->slots[k].lane = L;
        pcx_status st = upload_slot_meta(s, k);
        if (st != PCX_OK) return st;
      }
      fresh.clear();
      uploaded = true;
    }
    bool any = false, div = false;
    for (const Slot& sl : s->slots)
      if (sl.busy && sl.side && sl.lane == L) {
        any = true;
        div |= s->nets[sl.req.net]->has_div;
      }
    if (!any) continue;
    if (uploaded) {
      PCX_HIP(hipEventRecord(s->handed, s->stream));
      uploaded = false;
    }
    PCX_HIP(hipStreamWaitEvent(ln.stream, s->handed, 0));
    PCX_HIP(launch_burst_slots(s->d_slots, s->d_flags, s->n_slots, s->d_tight, s->cfg.burst_sweeps,
                               L, div ? kItvDiv : kItv, ln.stream));
    PCX_HIP(hipMemcpyAsync(ln.h_flags, s->d_flags, s->n_slots * sizeof(InstFlags),
                           hipMemcpyDeviceToHost, ln.stream));
    PCX_HIP(hipEventRecord(ln.done, ln.stream));
    ln.busy = true;
  }
  return PCX_OK;
}

pcx_status tick(pcx_server* s) {
  pcx_status st = side_collect(s);
  if (st != PCX_OK) return st;
  const int64_t admitted = s->stats.admitted;
  st = admit(s);
  if (st != PCX_OK) return st;
  // A tick that admitted requests runs one sweep: most requests settle or
  // fail in their first sweep and are handed back at the end of the tick.
  // Other ticks run sweeps_per_tick sweeps for the requests still going.
  const bool fresh = s->stats.admitted > admitted;
  const int n_sweeps = fresh ? 1 : s->cfg.sweeps_per_tick;
  int32_t running = 0, in_batch = 0;
  bool div = false;
  for (const Slot& sl : s->slots) {
    if (!sl.busy) continue;
    ++running;
    if (sl.side) continue;
    ++in_batch;
    div |= s->nets[sl.req.net]->has_div;
  }
  s->stats.max_running = std::max(s->stats.max_running, running);
  if (in_batch > 0) {
    const int ppw = s->cfg.props_per_wave ? s->cfg.props_per_wave : 512;
    const int grid = pcx_internal::device_cus(s->cfg.device) * 8;
    // A tick's sweeps read everything (slots, flags, sweep number) from
    // device memory, so they are captured once per kernel domain and replayed.
    hipGraphExec_t& gx = s->tick_graph[div ? 1 : 0][fresh ? 1 : 0];
    if (!gx) {
      const WorkList wl = worklist(s);
      hipGraph_t graph = nullptr;
      PCX_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      hipError_t ce = hipSuccess;
      for (int g = 0; g < n_sweeps && ce == hipSuccess; ++g) {
        ce = launch_schedule_slots(s->d_slots, s->d_flags, s->n_slots, wl, ppw, s->stream);
        if (ce == hipSuccess)
          ce = launch_sweep_slots(s->d_slots, s->d_flags, wl, s->d_tight, ppw, div ? kItvDiv : kItv,
                                  grid, s->stream);
      }
      const hipError_t ee = hipStreamEndCapture(s->stream, &graph);
      if (ce != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        PCX_HIP(ce != hipSuccess ? ce : ee);
      }
      const hipError_t ie = hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      PCX_HIP(ie);
    }
    PCX_HIP(hipGraphLaunch(gx, s->stream));
    s->next_sweep += n_sweeps;
    s->stats.sweeps += n_sweeps;
    PCX_HIP(hipMemcpyAsync(s->h_flags, s->d_flags, s->n_slots * sizeof(InstFlags), hipMemcpyDeviceToHost, s->stream));
    PCX_HIP(hipStreamSynchronize(s->stream));
    st = harvest(s);
    if (st != PCX_OK) return st;
  } else {
    PCX_HIP(hipStreamSynchronize(s->stream));
  }
  st = side_launch(s);
  if (st != PCX_OK) return st;
  ++s->stats.ticks;
  return PCX_OK;
}

// One request on its own, as a caller without the server would run it: the
// drop-in fixpoint call on the request's network.
pcx_status run_alone(pcx_server* s, const Request& rq, pcx_result* out) {
  return pcx_fixpoint(s->nets[rq.net], rq.lb, rq.ub, rq.nvars, out);
}

}  // namespace

extern "C" {

void pcx_server_config_default(pcx_server_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->device = 0;
  c->max_slots = 256;
  c->arena_bytes = (int64_t)4 << 30;
  c->sweeps_per_tick = 4;
  c->props_per_wave = 0;
  c->stage_bytes = (int64_t)64 << 20;
  c->event_driven = -1;
  c->burst_sweeps = -1;
  c->burst_after = -1;
}

void pcx_server_destroy(pcx_server* s) {
  if (!s) return;
  (void)hipSetDevice(s->cfg.device);
  for (SideLane& ln : s->lanes)
    if (ln.stream) (void)hipStreamSynchronize(ln.stream);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& row : s->tick_graph)
    for (hipGraphExec_t g : row)
      if (g) (void)hipGraphExecDestroy(g);
  for (void* p : {(void*)s->d_slots, (void*)s->d_flags, (void*)s->d_act, (void*)s->d_item_base,
                  (void*)s->d_ctl, (void*)s->d_tight, (void*)s->d_arena})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)s->h_slots, (void*)s->h_flags, (void*)s->h_stage})
    if (p) (void)hipHostFree(p);
  for (SideLane& ln : s->lanes) {
    if (ln.h_flags) (void)hipHostFree(ln.h_flags);
    if (ln.done) (void)hipEventDestroy(ln.done);
    if (ln.stream) (void)hipStreamDestroy(ln.stream);
  }
  if (s->handed) (void)hipEventDestroy(s->handed);
  for (NetRoot& nr : s->roots) {
    if (nr.told) (void)hipFree(nr.told);
    if (nr.root) (void)hipFree(nr.root);
  }
  if (s->d_root_out) (void)hipFree(s->d_root_out);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

pcx_status pcx_server_create(const pcx_server_config* cfg, pcx_server** out) {
  if (!out) return fail_with(PCX_EINVAL, "out is NULL");
  *out = nullptr;
  pcx_server_config c;
  if (cfg) c = *cfg; else pcx_server_config_default(&c);
  if (c.max_slots <= 0 || c.max_slots > (1 << 20) || c.arena_bytes < 8 || c.sweeps_per_tick <= 0 ||
      c.stage_bytes < 8 || c.burst_sweeps < -1 || c.burst_after < -1)
    return fail_with(PCX_EINVAL, "bad server configuration");
  if (c.burst_sweeps == -1) c.burst_sweeps = 1024;
  if (c.burst_after == -1) c.burst_after = 256;
  if (c.props_per_wave && c.props_per_wave != 64 && c.props_per_wave != 128 &&
      c.props_per_wave != 256 && c.props_per_wave != 512)
    return fail_with(PCX_EINVAL, "props_per_wave must be 0, 64, 128, 256 or 512");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail_with(PCX_EDEVICE, "no HIP device");
  if (c.device < 0 || c.device >= n) return fail_with(PCX_EINVAL, "device ordinal out of range");
  PCX_HIP(hipSetDevice(c.device));
  auto* s = new (std::nothrow) pcx_server();
  if (!s) return fail_with(PCX_ENOMEM, "host allocation failed");
  s->cfg = c;
  s->n_slots = c.max_slots;
  s->slots.resize(c.max_slots);
  const int64_t words = c.arena_bytes / (int64_t)sizeof(int2);
  s->stage_words = c.stage_bytes / (int64_t)sizeof(int2);
  hipError_t he = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  auto chk = [&](hipError_t r) { if (he == hipSuccess) he = r; };
  chk(hipEventCreateWithFlags(&s->handed, hipEventDisableTiming));
  for (SideLane& ln : s->lanes) {
    chk(hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking));
    chk(hipEventCreateWithFlags(&ln.done, hipEventDisableTiming));
    chk(hipHostMalloc(&ln.h_flags, c.max_slots * sizeof(InstFlags), hipHostMallocDefault));
  }
  chk(hipMalloc(&s->d_slots, c.max_slots * sizeof(SlotDesc)));
  chk(hipMalloc(&s->d_flags, c.max_slots * sizeof(InstFlags)));
  chk(hipMalloc(&s->d_act, c.max_slots * sizeof(int32_t)));
  chk(hipMalloc(&s->d_item_base, (c.max_slots + 1) * sizeof(int64_t)));
  chk(hipMalloc(&s->d_ctl, kCtlWords * sizeof(int32_t)));
  chk(hipMalloc(&s->d_tight, kTightSlots * sizeof(unsigned long long)));
  chk(hipMalloc(&s->d_arena, words * sizeof(int2)));
  chk(hipHostMalloc(&s->h_slots, c.max_slots * sizeof(SlotDesc), hipHostMallocDefault));
  chk(hipHostMalloc(&s->h_flags, c.max_slots * sizeof(InstFlags), hipHostMallocDefault));
  chk(hipHostMalloc(&s->h_stage, s->stage_words * sizeof(int2), hipHostMallocDefault));
  if (he != hipSuccess) {
    pcx_server_destroy(s);
    return fail_with(he == hipErrorOutOfMemory ? PCX_ENOMEM : PCX_EDEVICE,
                     std::string("server allocation: ") + hipGetErrorString(he));
  }
  s->arena.init(words);
  for (int32_t k = 0; k < c.max_slots; ++k) {
    s->h_slots[k] = SlotDesc{};
    InstFlags f{};
    f.changed[0] = -2; f.changed[1] = -2; f.failed_at = kFreeFailed; f.done_at = 0;
    s->h_flags[k] = f;
  }
  he = hipMemcpy(s->d_slots, s->h_slots, c.max_slots * sizeof(SlotDesc), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(s->d_flags, s->h_flags, c.max_slots * sizeof(InstFlags), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemset(s->d_tight, 0, kTightSlots * sizeof(unsigned long long));
  int32_t ctl[kCtlWords] = {0};
  ctl[1] = -1;                                   // no sweep run yet
  if (he == hipSuccess) he = hipMemcpy(s->d_ctl, ctl, sizeof(ctl), hipMemcpyHostToDevice);
  if (he != hipSuccess) {
    pcx_server_destroy(s);
    return fail_with(PCX_EDEVICE, std::string("server init: ") + hipGetErrorString(he));
  }
  *out = s;
  return PCX_OK;
}

pcx_status pcx_server_add_network(pcx_server* s, pcx_engine* e, int32_t* net) {
  if (!s || !e || !net) return fail_with(PCX_EINVAL, "NULL argument");
  if (e->cfg.domain != PCX_DOMAIN_INTERVAL) return fail_with(PCX_EINVAL, "serving takes interval engines");
  if (e->cfg.device != s->cfg.device) return fail_with(PCX_EINVAL, "engine and server are on different devices");
  pcx_status st = pcx_internal::sync_table(e);
  if (st != PCX_OK) return st;
  s->nets.push_back(e);
  if (events_for(s, (int32_t)s->nets.size() - 1)) {
    if (3 * (int64_t)e->op.size() >= ((int64_t)1 << 31)) {
      s->nets.pop_back();
      return fail_with(PCX_EINVAL, "network too large for marked sweeps; set event_driven = 0");
    }
    st = pcx_internal::ensure_adjacency(e);
    if (st != PCX_OK) { s->nets.pop_back(); return st; }
  }
  *net = (int32_t)s->nets.size() - 1;
  return PCX_OK;
}

pcx_status pcx_server_set_root(pcx_server* s, int32_t net, const int32_t* lb, const int32_t* ub,
                               int64_t nvars) {
  if (!s || !lb || !ub) return fail_with(PCX_EINVAL, "NULL argument");
  if (net < 0 || net >= (int32_t)s->nets.size()) return fail_with(PCX_EINVAL, "unknown network");
  pcx_engine* e = s->nets[net];
  if (e->dirty) return fail_with(PCX_EINVAL, "network told more propagators after registration");
  if (nvars <= e->max_var) return fail_with(PCX_EINVAL, "store smaller than the network's variables");
  PCX_HIP(hipSetDevice(s->cfg.device));
  // the root fixpoint, by the engine's own fixpoint call
  std::vector<int32_t> rl(lb, lb + nvars), ru(ub, ub + nvars);
  pcx_result r{};
  pcx_status st = pcx_fixpoint(e, rl.data(), ru.data(), nvars, &r);
  if (st != PCX_OK) return st;
  std::vector<int2> told(nvars), root(nvars);
  for (int64_t v = 0; v < nvars; ++v) {
    told[v] = make_int2(lb[v], ub[v]);
    root[v] = make_int2(rl[v], ru[v]);
  }
  if (!s->d_root_out) PCX_HIP(hipMalloc(&s->d_root_out, s->n_slots * sizeof(int32_t)));
  if ((int32_t)s->roots.size() <= net) s->roots.resize(net + 1);
  NetRoot& nr = s->roots[net];
  PCX_HIP(hipStreamSynchronize(s->stream));        // no admission still reads the old root
  if (nr.told) (void)hipFree(nr.told);
  if (nr.root) (void)hipFree(nr.root);
  nr = NetRoot{};
  PCX_HIP(hipMalloc(&nr.told, nvars * sizeof(int2)));
  PCX_HIP(hipMalloc(&nr.root, nvars * sizeof(int2)));
  PCX_HIP(hipMemcpy(nr.told, told.data(), nvars * sizeof(int2), hipMemcpyHostToDevice));
  PCX_HIP(hipMemcpy(nr.root, root.data(), nvars * sizeof(int2), hipMemcpyHostToDevice));
  nr.nvars = nvars;
  return PCX_OK;
}

pcx_status pcx_server_submit(pcx_server* s, int32_t net, int32_t* lb, int32_t* ub, int64_t nvars,
                             int64_t tag) {
  if (!s || !lb || !ub) return fail_with(PCX_EINVAL, "NULL argument");
  if (net < 0 || net >= (int32_t)s->nets.size()) return fail_with(PCX_EINVAL, "unknown network");
  pcx_engine* e = s->nets[net];
  if (e->dirty) return fail_with(PCX_EINVAL, "network told more propagators after registration");
  if (nvars <= e->max_var) return fail_with(PCX_EINVAL, "store smaller than the network's variables");
  Request rq{net, lb, ub, nvars, tag, now_s()};
  if (nvars > s->stage_words || footprint(s, rq).units * (int64_t)sizeof(int2) > s->cfg.arena_bytes)
    return fail_with(PCX_EINVAL, "store larger than the server's arena or staging buffer");
  s->pending.push_back(rq);
  return PCX_OK;
}

pcx_status pcx_server_step(pcx_server* s, int32_t* in_flight) {
  if (!s) return fail_with(PCX_EINVAL, "NULL server");
  PCX_HIP(hipSetDevice(s->cfg.device));
  pcx_status st = tick(s);
  if (in_flight) *in_flight = busy_slots(s) + (int32_t)s->pending.size();
  return st;
}

pcx_status pcx_server_get_stats(const pcx_server* s, pcx_server_stats* out) {
  if (!s || !out) return fail_with(PCX_EINVAL, "NULL argument");
  *out = s->stats;
  return PCX_OK;
}

int64_t pcx_server_poll(pcx_server* s, pcx_served* out, int64_t max) {
  if (!s || !out || max <= 0) return 0;
  const int64_t n = std::min<int64_t>(max, (int64_t)s->done.size());
  std::copy(s->done.begin(), s->done.begin() + n, out);
  s->done.erase(s->done.begin(), s->done.begin() + n);
  return n;
}

pcx_status pcx_server_replay(pcx_server* s, int64_t n, const double* arrival_s, const int32_t* net,
                             int32_t* lb_all, int32_t* ub_all, const int64_t* store_off,
                             int32_t plain, pcx_served* out, double* wall_s) {
  if (!s || n < 0 || (n > 0 && (!arrival_s || !net || !lb_all || !ub_all || !store_off || !out)))
    return fail_with(PCX_EINVAL, "NULL argument");
  if (!s->pending.empty() || busy_slots(s) > 0 || !s->done.empty())
    return fail_with(PCX_EBUSY, "replay needs an idle server (no queued, running or unpolled requests)");
  PCX_HIP(hipSetDevice(s->cfg.device));
  for (int64_t i = 0; i < n; ++i) {
    if (net[i] < 0 || net[i] >= (int32_t)s->nets.size()) return fail_with(PCX_EINVAL, "unknown network");
    if (s->nets[net[i]]->dirty) return fail_with(PCX_EINVAL, "network told more propagators after registration");
    const int64_t nv = store_off[i + 1] - store_off[i];
    if (nv <= s->nets[net[i]]->max_var) return fail_with(PCX_EINVAL, "store smaller than the network's variables");
    const Request rq{net[i], nullptr, nullptr, nv, i, 0.0};
    if (nv > s->stage_words || footprint(s, rq).units * (int64_t)sizeof(int2) > s->cfg.arena_bytes)
      return fail_with(PCX_EINVAL, "store larger than the server's arena or staging buffer");
    if (i && arrival_s[i] < arrival_s[i - 1]) return fail_with(PCX_EINVAL, "arrivals must be sorted");
  }
  const double t_start = now_s();
  int64_t next = 0, completed = 0;
  std::vector<double> arrive(n);
  for (int64_t i = 0; i < n; ++i) arrive[i] = t_start + arrival_s[i];
  if (plain) {
    for (int64_t i = 0; i < n; ++i) {
      while (now_s() < arrive[i]) std::this_thread::yield();
      Request rq{net[i], lb_all + store_off[i], ub_all + store_off[i], store_off[i + 1] - store_off[i], i, arrive[i]};
      pcx_result r{};
      pcx_status st = run_alone(s, rq, &r);
      if (st != PCX_OK) return st;
      out[i].tag = i;
      out[i].result = r;
      out[i].latency_ms = (now_s() - arrive[i]) * 1e3;
    }
  } else {
    std::vector<pcx_served> buf(256);
    while (completed < n) {
      const double now = now_s();
      while (next < n && arrive[next] <= now) {
        Request rq{net[next], lb_all + store_off[next], ub_all + store_off[next],
                   store_off[next + 1] - store_off[next], next, arrive[next]};
        s->pending.push_back(rq);
        ++next;
      }
      if (s->pending.empty() && busy_slots(s) == 0) {
        if (next < n) std::this_thread::yield();
        continue;
      }
      pcx_status st = tick(s);
      if (st != PCX_OK) return st;
      int64_t got;
      while ((got = pcx_server_poll(s, buf.data(), (int64_t)buf.size())) > 0)
        for (int64_t j = 0; j < got; ++j) {
          out[buf[j].tag] = buf[j];
          ++completed;
        }
    }
  }
  if (wall_s) *wall_s = now_s() - t_start;
  return PCX_OK;
}

}  // extern "C"
