// pcx_noalloc.hip -- pcx_fixpoint_noalloc: the single-store fixpoint for
// callers that cannot allocate (tight serving loops, harnesses that forbid
// allocation).  Its working state -- the packed store, the instance flags,
// the tightening counters and the control words -- lives in one fixed
// device array of this file, `scratch`, instead of a batch the library
// allocates.  Each GPU holds its own copy of the array until the process
// ends; its device address is looked up once per device and kept.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "pcx_internal.h"

using namespace pcx;
using pcx_internal::fail_with;

namespace {

// Layout of the scratch array (byte offsets): flags, tightening counters and
// control words in the first 4 KiB, then the packed store.
constexpr int64_t kFlagsAt = 0;
constexpr int64_t kTightAt = 256;
constexpr int64_t kCtlAt = kTightAt + 2 * kTightSlots * 8;     // 1280
constexpr int64_t kActAt = kCtlAt + kCtlWords * 4;             // 1344
constexpr int64_t kBaseAt = kActAt + 64;                       // 1408: item_base (2 x int64)
constexpr int64_t kStoreAt = 4096;
constexpr int64_t kScratchBytes = kStoreAt + (int64_t)PCX_NOALLOC_MAX_VARS * 8;
static_assert(sizeof(InstFlags) <= kTightAt, "flags fit their slot");
static_assert(kBaseAt + 16 <= kStoreAt, "control words fit the header");

constexpr int kMaxDevices = 64;

}  // namespace

// The working state of every pcx_fixpoint_noalloc call on a device.  Host
// code never uses this name as an address: it uses the per-device pointer
// that hipGetSymbolAddress returns (scratch_of below).
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
uint8_t* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call on that device.
pcx_status scratch_of(int device, uint8_t** out) {
  if (device < 0 || device >= kMaxDevices) return fail_with(PCX_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (he != hipSuccess || !p)
      return fail_with(PCX_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(he));
    g_scratch[device] = static_cast<uint8_t*>(p);
  }
  *out = g_scratch[device];
  return PCX_OK;
}

// Zeroes n 8-byte words (the tightening counters): a kernel, so the call
// records no memset node if a caller captures its stream.
__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace

extern "C" {

int64_t pcx_noalloc_max_vars(void) { return PCX_NOALLOC_MAX_VARS; }

pcx_status pcx_fixpoint_noalloc(pcx_engine* e, int32_t* d_lb, int32_t* d_ub, int64_t nvars,
                                pcx_result* out) {
  if (!e || !d_lb || !d_ub) return fail_with(PCX_EINVAL, "NULL argument");
  if (e->cfg.domain != PCX_DOMAIN_INTERVAL)
    return fail_with(PCX_EINVAL, "pcx_fixpoint_noalloc needs an interval engine");
  if (nvars <= 0) return fail_with(PCX_EINVAL, "nvars must be positive");
  if (nvars > PCX_NOALLOC_MAX_VARS)
    return fail_with(PCX_EINVAL, "pcx_fixpoint_noalloc: " + std::to_string(nvars) +
                                     " variables exceed its fixed working store of " +
                                     std::to_string((long long)PCX_NOALLOC_MAX_VARS));
  if (e->max_var >= nvars) return fail_with(PCX_EINVAL, "a propagator names a variable outside the store");
  const int64_t m = (int64_t)e->op.size();
  // The table's device copy is refreshed in place; growing it would allocate.
  if (e->dirty && m > e->d_cap)
    return fail_with(PCX_EINVAL, "pcx_fixpoint_noalloc: the table outgrew its device copy; one "
                                 "pcx_fixpoint call (or a batch run) on this engine sizes it");
  PCX_HIP(hipSetDevice(e->cfg.device));
  pcx_status st = pcx_internal::sync_table(e);
  if (st != PCX_OK) return st;
  uint8_t* base = nullptr;
  st = scratch_of(e->cfg.device, &base);
  if (st != PCX_OK) return st;
  InstFlags* d_flags = reinterpret_cast<InstFlags*>(base + kFlagsAt);
  unsigned long long* d_tight = reinterpret_cast<unsigned long long*>(base + kTightAt);
  int32_t* d_ctl = reinterpret_cast<int32_t*>(base + kCtlAt);
  int32_t* d_act = reinterpret_cast<int32_t*>(base + kActAt);
  int64_t* d_base = reinterpret_cast<int64_t*>(base + kBaseAt);
  int2* d_store = reinterpret_cast<int2*>(base + kStoreAt);
  hipStream_t s = e->stream;
  const int dom = pcx_internal::dom_of(e);

  // Every value the sweeps read is written first: flags, counters, store.
  const WorkList wl = worklist_of(d_act, d_base, d_ctl);
  PCX_HIP(launch_init_flags(d_flags, 1, wl, s));
  k_zero_u64<<<1, 2 * kTightSlots, 0, s>>>(d_tight, 2 * kTightSlots);
  PCX_HIP(hipGetLastError());
  PCX_HIP(launch_store_pack(d_lb, d_ub, d_store, nvars, s));
  PCX_HIP(launch_store_check_empty(d_store, nvars, 1, d_flags, dom, s));

  // Full sweeps, the flags read back every `group` sweeps (as the plain mode
  // does), until a sweep changed nothing or the store failed.
  const Table t{e->d_xo, e->d_y, e->d_z, m};
  int ppw = e->cfg.props_per_wave ? std::min(e->cfg.props_per_wave, 512) : 0;
  if (!ppw) {
    const int cus = pcx_internal::device_cus(e->cfg.device);
    ppw = 64;
    for (int p = 512; p > 64; p /= 2)
      if (m / ((int64_t)kBlock * (p / kWave)) >= 2LL * cus) { ppw = p; break; }
  }
  const int group = e->cfg.sweeps_per_check > 0 ? e->cfg.sweeps_per_check : 8;
  const int64_t maxs = e->cfg.max_sweeps;
  InstFlags hf{};
  int32_t sw = 0;
  for (;;) {
    for (int g = 0; g < group; ++g, ++sw) {
      if (maxs > 0 && sw >= maxs) break;
      PCX_HIP(launch_sweep_plain(t, d_store, nullptr, m, sw, d_flags, d_tight, ppw, dom, false, s));
    }
    PCX_HIP(hipMemcpyAsync(&hf, d_flags, sizeof(hf), hipMemcpyDeviceToHost, s));
    PCX_HIP(hipStreamSynchronize(s));
    const int32_t last = std::max(hf.changed[0], hf.changed[1]);
    if (hf.failed_at != kNoFail || last + 1 < sw) break;     // the quiet sweep already ran
    if (maxs > 0 && sw >= maxs) break;
  }
  PCX_HIP(launch_store_unpack(d_store, d_lb, d_ub, nvars, s));
  PCX_HIP(hipStreamSynchronize(s));
  if (out) {
    const int32_t last = std::max(hf.changed[0], hf.changed[1]);
    pcx_result r{};
    r.failed = hf.failed_at != kNoFail;
    r.has_changed = last >= 0;
    r.sweeps = r.failed ? (int64_t)hf.failed_at + 1 : (int64_t)last + 2;
    if (maxs > 0 && r.sweeps > maxs) r.sweeps = maxs;
    r.evals = r.sweeps * m;
    *out = r;
  }
  return PCX_OK;
}

}  // extern "C"
