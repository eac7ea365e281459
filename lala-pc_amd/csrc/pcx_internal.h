// pcx_internal.h -- engine and batch objects shared by the host sources.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/pcx.h"
#include "pcx_kernels.h"

struct pcx_engine {
  pcx_config cfg{};
  std::vector<int32_t> op, x, y, z;   // table in engine order
  int32_t max_var = -1;
  bool has_div = false;               // table holds a division propagator
  bool dirty = true;
  uint64_t version = 0;               // bumped on every upload of the table
  // variable -> table positions adjacency (event-driven sweeps), built on demand
  int32_t* d_adj_off = nullptr;
  int32_t* d_adj = nullptr;
  uint64_t adj_version = 0;
  int64_t adj_vars = 0;
  uint32_t* d_xo = nullptr;
  int32_t* d_y = nullptr;
  int32_t* d_z = nullptr;
  int64_t d_cap = 0;
  hipStream_t stream = nullptr;
  pcx_batch* scratch = nullptr;       // batch of one for pcx_fixpoint / pcx_ask_all
  int64_t scratch_nvars = -1;
  int64_t scratch_table = -1;
};

struct pcx_batch {
  pcx_engine* e = nullptr;
  int64_t n_inst = 0, nvars = 0;
  int64_t table_n = 0;                // table size the subsets were built for
  int2* d_store = nullptr;
  int2* d_image = nullptr;
  int32_t* d_unpack = nullptr;        // staging of pcx_batch_read
  uint8_t* d_image_empty = nullptr;   // per instance: the image holds an empty domain
  bool store_is_image = false;        // the stores equal the image (after load / reset)
  int64_t* d_off = nullptr;
  int32_t* d_idx = nullptr;
  std::vector<int64_t> h_off;         // empty = whole table
  // Per-instance copies of the subsets (instance k's propagators at
  // [h_off[k], h_off[k+1]) of these arrays), built on first run.
  uint32_t* d_cxo = nullptr;
  int32_t* d_cy = nullptr;
  int32_t* d_cz = nullptr;
  uint64_t ctab_version = 0;          // engine table version they were built from
  bool compact = false;               // this run reads the per-instance copies
  bool stream = false;                // this run streams table loads
  // event-driven sweeps: member bitmaps (subsets only) and two dirty bitmaps
  uint32_t* d_member = nullptr;
  uint32_t* d_dirty = nullptr;
  uint32_t* d_summary = nullptr;      // chunk summaries (2 planes) + scheduler snapshot
  uint64_t member_version = 0;
  // lockstep full sweeps: instance-interleaved copy of the stores
  int2* d_shadow = nullptr;
  int64_t shadow_words = 0;
  int shadow_gi = 0;                  // group size of its layout
  bool shadow_is_image = false;       // it holds the image (interleaved since the last load)
  bool image_in_shadow = false;       // the reset image exists only as the interleaved copy
                                      // (a device load); d_image is written when needed
  bool run_from_image = false;        // this run started from stores equal to the image
  bool events = false;                // this run uses marked-only sweeps
  bool seeded = false;                // the next run may start from seeded marks
  bool seeded_run = false;            // this run starts with a marked sweep
  bool marks_clean = false;           // both mark planes and the summaries are zero
  // first sweeps seeded from a base store (per variable the value instances
  // 0-2 agree on): the base, the values of the image that differ from it
  // (both written by the load), and the propagators the base does not
  // satisfy (found by the run)
  int2* d_base = nullptr;             // [nvars]
  int2* d_base_alt = nullptr;         // [nvars]: the previous load's base store (device loads)
  unsigned long long* d_diff = nullptr;   // kDiffLists sub-lists of diff_cap entries
  int2* d_diff_val = nullptr;         // the listed values (device loads)
  bool image_in_diffs = false;        // the image is the base store + the listed values
  bool shared_loads = true;           // the last device load's stores shared most values
  double load_ms = 0.0;               // device time of the last device load (timing), until reported
  hipEvent_t load_ev[2] = {nullptr, nullptr};
  int64_t diff_cap = 0;
  int32_t* d_unsat = nullptr;         // [table_n]
  int32_t* d_trail = nullptr;         // [n_inst * kTrailCap] variables a drained run tightened,
                                      // then [n_inst * kSeedCap] the drain's seeds
  int32_t* d_trail_n = nullptr;       // [2 * n_inst]: the two lists' counts
  bool trail_valid = false;           // the stores are the image but for the trail's variables
  int32_t device = 0;                 // the engine's device (destroy reads it, not the engine)
  int32_t* d_bvar = nullptr;          // [n_inst] split variables of this batch's last branching
  bool probe_valid = false;           // d_unsat and the counts below hold this image's probe
  uint64_t probe_version = 0;         // of this table version
  int64_t probe_diff = 0, probe_unsat = 0;
  int32_t* d_seedctl = nullptr;       // differences per sub-list, then propagators in d_unsat
  bool diffs_listed = false;          // d_diff describes the current image
  // captured group of marked sweeps (batched mode), rebuilt when any launch
  // argument it was captured with differs
  struct MarkedGraphArgs {
    uint64_t version;                 // engine table version
    int32_t group, dom, max_sweeps, ppw;
    const void *xo, *y, *z, *adj_off, *adj, *q;
    bool operator==(const MarkedGraphArgs& o) const {
      return version == o.version && group == o.group && dom == o.dom && max_sweeps == o.max_sweeps &&
             ppw == o.ppw && xo == o.xo && y == o.y && z == o.z && adj_off == o.adj_off && adj == o.adj &&
             q == o.q;
    }
  };
  hipGraphExec_t marked_graph = nullptr;
  MarkedGraphArgs marked_graph_args{};
  // queued marked sweeps: two planes of kQLists sub-lists of qcap entries,
  // control words, and the captured group of (turn, queue pass) launches
  unsigned long long* d_q = nullptr;
  int32_t* d_qctl = nullptr;
  int64_t qcap = 0;
  hipGraphExec_t queue_graph = nullptr;
  MarkedGraphArgs queue_graph_args{};
  hipGraphExec_t plain_queue_graph = nullptr;     // the same for plain-mode runs
  MarkedGraphArgs plain_queue_graph_args{};
  pcx::InstFlags* d_flags = nullptr;
  unsigned long long* d_tight = nullptr;
  int32_t* d_act = nullptr;
  int64_t* d_item_base = nullptr;
  int32_t* d_ctl = nullptr;           // kCtlWords control words (see WorkList)
  int32_t* d_order = nullptr;         // instances by decreasing size (persistent queue)
  std::vector<hipEvent_t> ev;         // timing pool
  // pinned host memory for the readbacks of a run: kPinScratch bytes of small
  // values (a flags record, the live count), then the flags of every
  // instance, then the tightening counters
  uint8_t* h_pin = nullptr;
};

namespace pcx_internal {
pcx_status fail_with(pcx_status st, const std::string& msg);
pcx_status sync_table(pcx_engine* e);   // upload the table if it changed
int dom_of(const pcx_engine* e);        // kernel domain for this engine
int device_cus(int device);             // compute units of a device
pcx_status ensure_adjacency(pcx_engine* e);   // variable -> table positions, on the device
}  // namespace pcx_internal

#define PCX_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return pcx_internal::fail_with(e_ == hipErrorOutOfMemory ? PCX_ENOMEM : PCX_EDEVICE, \
                                     std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
