// pcx_gen.cpp -- seeded synthetic workloads (include/pcx_gen.h).
//
// Random source: splitmix64 (published by S. Vigna; constants of the
// reference implementation at prng.di.unimi.it), seeded per stream with a
// hash of (seed, stream id) so instances can be generated in parallel and
// still be reproducible.
#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pcx_gen.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  // uniform in [0, n)
  uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
  // uniform in [lo, hi]
  int64_t range(int64_t lo, int64_t hi) { return lo + (int64_t)below((uint64_t)(hi - lo + 1)); }
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

uint64_t stream_seed(uint64_t seed, uint64_t id) {
  SplitMix64 r(seed ^ (id * 0xd1b54a32d192ed03ULL));
  r.next();
  return r.next();
}

enum Kind { kAdd, kSub, kMul, kMin, kMax, kEqr, kLeqr, kKinds };

constexpr int32_t OP_ADD = 0, OP_MUL = 1, OP_MIN = 6, OP_MAX = 7, OP_EQ = 8, OP_LEQ = 9;

int threads_for(int64_t work) {
  unsigned hc = std::thread::hardware_concurrency();
  int t = (int)std::min<unsigned>(hc ? hc : 1, 16);
  if (work < 4096) t = 1;
  return t;
}

template <class F>
void parallel_for(int64_t n, F f) {
  const int t = threads_for(n);
  if (t <= 1) { for (int64_t i = 0; i < n; ++i) f(i); return; }
  std::vector<std::thread> th;
  for (int w = 0; w < t; ++w)
    th.emplace_back([=]() { for (int64_t i = w; i < n; i += t) f(i); });
  for (auto& x : th) x.join();
}

struct Emit {
  int32_t *op, *x, *y, *z;
  int64_t n = 0;
  void put(int kind, int32_t a, int32_t b, int32_t c) {
    // a = b <kind> c
    switch (kind) {
      case kAdd:  op[n] = OP_ADD; x[n] = a; y[n] = b; z[n] = c; break;
      case kSub:  op[n] = OP_ADD; x[n] = b; y[n] = a; z[n] = c; break;   // b = a + c
      case kMul:  op[n] = OP_MUL; x[n] = a; y[n] = b; z[n] = c; break;
      case kMin:  op[n] = OP_MIN; x[n] = a; y[n] = b; z[n] = c; break;
      case kMax:  op[n] = OP_MAX; x[n] = a; y[n] = b; z[n] = c; break;
      case kEqr:  op[n] = OP_EQ;  x[n] = a; y[n] = b; z[n] = c; break;
      default:    op[n] = OP_LEQ; x[n] = a; y[n] = b; z[n] = c; break;
    }
    ++n;
  }
};

int draw_kind(SplitMix64& r, const std::vector<int>& kinds) {
  return kinds[r.below(kinds.size())];
}

bool kind_value(int kind, int64_t a, int64_t b, int64_t* out) {
  switch (kind) {
    case kAdd:  *out = a + b; return true;
    case kSub:  *out = a - b; return true;
    case kMul:  *out = a * b; return true;
    case kMin:  *out = std::min(a, b); return true;
    case kMax:  *out = std::max(a, b); return true;
    case kEqr:  *out = a == b; return true;
    default:    *out = a <= b; return true;
  }
}

pcx_status literal_network(const pcxg_params& p, const std::vector<int>& kinds, Emit& em,
                           int32_t* lb, int32_t* ub) {
  SplitMix64 r(stream_seed(p.seed, 1));
  const int64_t V = p.nvars;
  for (int64_t v = 0; v < V; ++v) {
    const int64_t l = r.range(0, p.lb_max);
    lb[v] = (int32_t)l;
    ub[v] = (int32_t)std::min<int64_t>(l + r.range(p.width_min, p.width_max), PCX_INF - 1);
  }
  const int64_t chained = (int64_t)(p.chain_frac * (double)p.nprops);
  int64_t in_chain = 0;
  int32_t prev = -1;
  for (int64_t i = 0; i < p.nprops; ++i) {
    const int kind = draw_kind(r, kinds);
    int32_t a, b, c;
    const bool chain = i < chained;
    if (chain && prev >= 0 && in_chain < p.chain_len) {
      b = prev;
    } else {
      b = (int32_t)r.below(V);
      in_chain = 0;
    }
    do { a = (int32_t)r.below(V); } while (a == b);
    do { c = (int32_t)r.below(V); } while (c == a || c == b);
    em.put(kind, a, b, c);
    prev = chain ? a : -1;
    ++in_chain;
  }
  return PCX_OK;
}

struct Bucket {
  std::vector<int64_t> off;   // value -> [off[v], off[v+1]) into var
  std::vector<int32_t> var;

  void build(const std::vector<int32_t>& w, const std::vector<int32_t>& members, int64_t range) {
    off.assign(range + 1, 0);
    for (int32_t v : members) off[w[v] + 1]++;
    for (int64_t i = 0; i < range; ++i) off[i + 1] += off[i];
    var.resize(members.size());
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int32_t v : members) var[fill[w[v]]++] = v;
  }
  int64_t size(int64_t val) const {
    return (val < 0 || val + 1 >= (int64_t)off.size()) ? 0 : off[val + 1] - off[val];
  }
  // A member holding value `val`, different from a and b, or -1.
  int32_t pick(SplitMix64& r, int64_t val, int32_t a, int32_t b) const {
    const int64_t n = size(val);
    if (n <= 0) return -1;
    for (int t = 0; t < 8; ++t) {
      const int32_t v = var[off[val] + (int64_t)r.below((uint64_t)n)];
      if (v != a && v != b) return v;
    }
    return -1;
  }
};

// Domain from the [lb, ub] distribution, conditioned on containing w.
void planted_domain(SplitMix64& r, const pcxg_params& p, int64_t w, int32_t* lb, int32_t* ub) {
  const int64_t l = r.range(0, std::min<int64_t>(w, p.lb_max));
  const int64_t wmin = std::max<int64_t>(p.width_min, w - l);
  const int64_t width = wmin >= p.width_max ? wmin : r.range(wmin, p.width_max);
  *lb = (int32_t)l;
  *ub = (int32_t)std::min<int64_t>(l + width, PCX_INF - 1);
}

// The constant value c such that w_out = w_in <kind> c, or -1 when no single
// value does it (then `any_ge` / `any_ne` say which constants do).
struct NeedC { int64_t exact = -1; int64_t at_least = -1; int64_t at_most = -1; int64_t differ_from = -1; };
NeedC needed_constant(int kind, int64_t win, int64_t wout) {
  NeedC n;
  switch (kind) {
    case kAdd: n.exact = wout - win; break;
    case kSub: n.exact = win - wout; break;
    case kMul:
      if (win == 0) { if (wout == 0) n.at_least = 0; }
      else if (wout % win == 0) n.exact = wout / win;
      break;
    case kMin:
      if (wout < win) n.exact = wout; else if (wout == win) n.at_least = win;
      break;
    case kMax:
      if (wout > win) n.exact = wout; else if (wout == win) n.at_most = win;
      break;
    case kEqr:
      if (wout == 1) n.exact = win; else if (wout == 0) n.differ_from = win;
      break;
    default:   // win <= c  <=>  wout
      if (wout == 1) n.at_least = win; else if (wout == 0) n.at_most = win - 1;
      break;
  }
  return n;
}

pcx_status planted_network(const pcxg_params& p, const std::vector<int>& kinds, Emit& em,
                           int32_t* lb, int32_t* ub, int32_t* witness) {
  SplitMix64 r(stream_seed(p.seed, 2));
  const int64_t V = p.nvars, P = p.nprops;
  const int64_t range = std::max<int64_t>(2, p.value_range);
  constexpr int kBundle = 3;                 // links writing each chain variable
  const int depth = std::max(1, p.chain_len);
  int64_t chain_links = (int64_t)(p.chain_frac * (double)P);
  int64_t chain_vars = chain_links / kBundle;
  chain_vars = std::min<int64_t>(chain_vars, V / 2);

  // Variable roles on a random permutation: chain | constant | Boolean | int.
  std::vector<int32_t> perm(V);
  for (int64_t v = 0; v < V; ++v) perm[v] = (int32_t)v;
  for (int64_t i = V - 1; i > 0; --i) std::swap(perm[i], perm[(int64_t)r.below((uint64_t)i + 1)]);
  std::vector<int32_t> w(V, 0), chain_pool(perm.begin(), perm.begin() + chain_vars);
  std::vector<int32_t> consts, outs;         // outs: outputs of non-chain links
  for (int64_t i = chain_vars; i < V; ++i) {
    const int32_t v = perm[i];
    const double u = r.unit();
    if (u < p.const_frac) {
      w[v] = (int32_t)r.below((uint64_t)range);
      lb[v] = ub[v] = w[v];
      consts.push_back(v);
    } else if (u < p.const_frac + p.bool_frac) {
      w[v] = (int32_t)r.below(2);
      lb[v] = 0; ub[v] = 1;
      outs.push_back(v);
    } else {
      w[v] = (int32_t)r.below((uint64_t)range);
      planted_domain(r, p, w[v], &lb[v], &ub[v]);
      outs.push_back(v);
    }
  }
  if (consts.size() < 2 || outs.size() < 3) return PCX_EINVAL;
  Bucket cb;
  cb.build(w, consts, range);

  // ---- chains: a constant start, then `depth` exclusive variables, each
  // written by up to kBundle links x_k = x_{k-1} +/- constant.  Only ADD/SUB
  // links carry a bound along the whole chain (a MIN, MAX, MUL or reified
  // link against a constant can settle its output early), so chains use
  // those two kinds and the other propagators draw the remaining kinds.
  std::vector<int> chain_kinds, rest_kinds;
  for (int k : kinds) (k == kAdd || k == kSub ? chain_kinds : rest_kinds).push_back(k);
  if (chain_kinds.empty()) chain_kinds = kinds;
  if (rest_kinds.empty() || chain_links == 0) rest_kinds = kinds;
  int64_t used = 0, cv = 0;
  while (cv < chain_vars && used < chain_links && used < P) {
    int32_t prev = consts[r.below(consts.size())];
    for (int k = 0; k < depth && cv < chain_vars && used < P; ++k) {
      const int32_t x = chain_pool[cv++];
      int links = 0;
      for (int attempt = 0; attempt < 64 && links == 0; ++attempt) {
        const int kind = draw_kind(r, chain_kinds);
        const int32_t c = consts[r.below(consts.size())];
        int64_t val;
        kind_value(kind, w[prev], w[c], &val);
        if (val < 0 || val >= range) continue;
        w[x] = (int32_t)val;
        em.put(kind, x, prev, c);
        ++links;
      }
      if (links == 0) {                        // x = min(prev, c) always fits
        const int32_t c = consts[r.below(consts.size())];
        w[x] = std::min(w[prev], w[c]);
        em.put(kMin, x, prev, c);
        ++links;
      }
      for (int attempt = 0; attempt < 32 && links < kBundle && used + links < P; ++attempt) {
        const int kind = draw_kind(r, chain_kinds);
        const NeedC need = needed_constant(kind, w[prev], w[x]);
        int32_t c = -1;
        if (need.exact >= 0) c = cb.pick(r, need.exact, prev, x);
        else {
          for (int t = 0; t < 8 && c < 0; ++t) {
            const int32_t cand = consts[r.below(consts.size())];
            const int64_t wc = w[cand];
            if ((need.at_least >= 0 && wc >= need.at_least) ||
                (need.at_most >= 0 && wc <= need.at_most) ||
                (need.differ_from >= 0 && wc != need.differ_from))
              c = cand;
          }
        }
        if (c < 0 || c == prev || c == x) continue;
        int64_t val;
        kind_value(kind, w[prev], w[c], &val);
        if (val != w[x]) continue;
        em.put(kind, x, prev, c);
        ++links;
      }
      used += links;
      planted_domain(r, p, w[x], &lb[x], &ub[x]);
      prev = x;
    }
  }
  // unused chain variables become plain integers
  for (; cv < chain_vars; ++cv) {
    const int32_t v = chain_pool[cv];
    w[v] = (int32_t)r.below((uint64_t)range);
    planted_domain(r, p, w[v], &lb[v], &ub[v]);
    outs.push_back(v);
  }
  Bucket ob;
  ob.build(w, outs, range);

  // ---- the rest: operands uniform over the variables outside the chains
  // (so chains keep their depth), output a variable that holds the result
  // under the hidden assignment.  The kind is drawn once per propagator and
  // only the operands are redrawn, so the kind mix is what was asked for.
  std::vector<int32_t> pool(consts);
  pool.insert(pool.end(), outs.begin(), outs.end());
  const uint64_t np = pool.size();
  while (em.n < P) {
    const int kind = draw_kind(r, rest_kinds);
    bool done = false;
    for (int attempt = 0; attempt < 1024 && !done; ++attempt) {
      const int32_t b = pool[r.below(np)];
      int32_t c;
      do { c = pool[r.below(np)]; } while (c == b);
      int64_t val;
      kind_value(kind, w[b], w[c], &val);
      const int32_t a = ob.pick(r, val, b, c);
      if (a < 0) continue;
      em.put(kind, a, b, c);
      done = true;
    }
    if (!done) {
      for (;;) {
        const int32_t b = pool[r.below(np)];
        int32_t c;
        do { c = pool[r.below(np)]; } while (c == b);
        const int32_t a = ob.pick(r, std::min(w[b], w[c]), b, c);
        if (a < 0) continue;
        em.put(kMin, a, b, c);
        break;
      }
    }
  }
  if (witness) std::copy(w.begin(), w.end(), witness);
  return PCX_OK;
}

}  // namespace

extern "C" {

void pcxg_params_default(pcxg_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->seed = 1;
  p->nvars = 1000000;
  p->nprops = 4000000;
  p->kinds = PCXG_ALL;
  p->planted = 1;
  p->lb_max = 500000;
  p->width_min = 1;
  p->width_max = 500000;
  p->value_range = 125000;
  p->bool_frac = 0.25;
  p->const_frac = 0.10;
  p->chain_frac = 0.30;
  p->chain_len = 200;
}

pcx_status pcxg_network(const pcxg_params* p, int32_t* op, int32_t* x, int32_t* y, int32_t* z,
                        int32_t* lb, int32_t* ub, int32_t* witness) {
  if (!p || !op || !x || !y || !z || !lb || !ub) return PCX_EINVAL;
  if (p->nvars < 3 || p->nvars > PCX_MAX_VARS || p->nprops < 0) return PCX_EINVAL;
  if (p->lb_max < 0 || p->width_min < 0 || p->width_max < p->width_min) return PCX_EINVAL;
  std::vector<int> kinds;
  const uint32_t mask = p->kinds ? p->kinds : PCXG_ALL;
  for (int k = 0; k < kKinds; ++k)
    if (mask & (1u << k)) kinds.push_back(k);
  if (kinds.empty()) return PCX_EINVAL;
  Emit em{op, x, y, z};
  if (p->planted) return planted_network(*p, kinds, em, lb, ub, witness);
  if (witness) std::fill(witness, witness + p->nvars, 0);
  return literal_network(*p, kinds, em, lb, ub);
}

pcx_status pcxg_splits(uint64_t seed, const int32_t* lb, const int32_t* ub, int64_t nvars,
                       int64_t n_inst, int32_t kmin, int32_t kmax, int32_t* lb_out,
                       int32_t* ub_out) {
  if (!lb || !ub || !lb_out || !ub_out || nvars <= 0 || n_inst < 0 || kmin < 0 || kmax < kmin)
    return PCX_EINVAL;
  parallel_for(n_inst, [&](int64_t k) {
    SplitMix64 r(stream_seed(seed, 1000003ULL + (uint64_t)k));
    int32_t* l = lb_out + k * nvars;
    int32_t* u = ub_out + k * nvars;
    std::memcpy(l, lb, nvars * sizeof(int32_t));
    std::memcpy(u, ub, nvars * sizeof(int32_t));
    const int64_t splits = r.range(kmin, kmax);
    for (int64_t s = 0; s < splits; ++s) {
      for (int t = 0; t < 64; ++t) {
        const int64_t v = (int64_t)r.below((uint64_t)nvars);
        if (l[v] >= u[v]) continue;
        const int64_t c = r.range(l[v], (int64_t)u[v] - 1);   // v <= c  or  v > c
        if (r.below(2)) u[v] = (int32_t)c; else l[v] = (int32_t)(c + 1);
        break;
      }
    }
  });
  return PCX_OK;
}

pcx_status pcxg_subsets(uint64_t seed, int64_t nprops, int64_t n_inst, int64_t mmin, int64_t mmax,
                        int64_t* off, int32_t* idx) {
  if (!off || nprops < 0 || n_inst < 0 || mmin < 0 || mmax < mmin) return PCX_EINVAL;
  std::vector<int64_t> m(n_inst);
  off[0] = 0;
  for (int64_t k = 0; k < n_inst; ++k) {
    SplitMix64 r(stream_seed(seed, 7000001ULL + (uint64_t)k));
    m[k] = std::min<int64_t>(r.range(mmin, mmax), nprops);
    off[k + 1] = off[k] + m[k];
  }
  if (!idx) return PCX_OK;
  parallel_for(n_inst, [&](int64_t k) {
    SplitMix64 r(stream_seed(seed, 9000011ULL + (uint64_t)k));
    // selection sampling (Knuth, TAOCP vol. 2, Algorithm S): ascending output
    int64_t need = m[k], left = nprops;
    int32_t* out = idx + off[k];
    for (int64_t i = 0; i < nprops && need > 0; ++i, --left) {
      if ((int64_t)r.below((uint64_t)left) < need) {
        *out++ = (int32_t)i;
        --need;
      }
    }
  });
  return PCX_OK;
}

pcx_status pcxg_bits_network(uint64_t seed, int64_t nvars, int64_t nprops, uint32_t kinds,
                             double dens_lo, double dens_hi, int32_t planted,
                             int32_t* op, int32_t* x, int32_t* y, int32_t* z, uint64_t* words,
                             int32_t* witness) {
  if (!op || !x || !y || !z || !words || nvars < 3 || nvars > PCX_MAX_VARS || nprops < 0 ||
      dens_lo < 0 || dens_hi > 1 || dens_lo > dens_hi)
    return PCX_EINVAL;
  const uint32_t mask = kinds ? kinds : 0xfu;
  std::vector<int> ks;
  for (int k = 0; k < PCX_BITS_NUM_OPS; ++k)
    if (mask & (1u << k)) ks.push_back(k);
  if (ks.empty()) return PCX_EINVAL;
  SplitMix64 r(stream_seed(seed, 31));
  std::vector<int32_t> w(nvars, 0);
  for (int64_t v = 0; v < nvars; ++v) {
    const double d = dens_lo + (dens_hi - dens_lo) * r.unit();
    uint64_t word = 0;
    for (int b = 0; b < PCXG_BITS_VALUES; ++b)
      if (r.unit() < d) word |= 1ull << b;
    w[v] = (int32_t)r.below(PCXG_BITS_VALUES);
    if (planted) word |= 1ull << w[v];
    if (!word) word = 1ull << w[v];
    words[v] = word;
  }
  if (witness) std::copy(w.begin(), w.end(), witness);
  // value -> variables holding it as hidden value
  std::vector<int64_t> off(PCXG_BITS_VALUES + 1, 0);
  std::vector<int32_t> by;
  if (planted) {
    for (int64_t v = 0; v < nvars; ++v) off[w[v] + 1]++;
    for (int i = 0; i < PCXG_BITS_VALUES; ++i) off[i + 1] += off[i];
    by.resize(nvars);
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int64_t v = 0; v < nvars; ++v) by[fill[w[v]]++] = (int32_t)v;
  }
  auto holding = [&](int64_t val, int32_t a, int32_t b) -> int32_t {
    if (val < 0 || val >= PCXG_BITS_VALUES || off[val + 1] == off[val]) return -1;
    for (int t = 0; t < 8; ++t) {
      const int32_t v = by[off[val] + (int64_t)r.below((uint64_t)(off[val + 1] - off[val]))];
      if (v != a && v != b) return v;
    }
    return -1;
  };
  for (int64_t i = 0; i < nprops; ++i) {
    const int k = ks[r.below(ks.size())];
    int32_t a = -1, b = -1, c = -1;
    for (int attempt = 0; attempt < 1024; ++attempt) {
      b = (int32_t)r.below((uint64_t)nvars);
      do { c = (int32_t)r.below((uint64_t)nvars); } while (c == b);
      if (!planted) {
        do { a = (int32_t)r.below((uint64_t)nvars); } while (a == b || a == c);
        break;
      }
      if (k == PCX_BADD) { a = holding((int64_t)w[b] + w[c], b, c); if (a >= 0) break; continue; }
      if (k == PCX_BEQR) { a = holding(w[b] == w[c] ? 1 : 0, b, c); if (a >= 0) break; continue; }
      if (k == PCX_BNEQ && w[b] == w[c]) continue;
      if (k == PCX_BLEQ && w[b] > w[c]) std::swap(b, c);
      do { a = (int32_t)r.below((uint64_t)nvars); } while (a == b || a == c);
      break;
    }
    if (a < 0) return PCX_EINVAL;   // no variable holds the needed value
    op[i] = k; x[i] = a; y[i] = b; z[i] = c;
  }
  return PCX_OK;
}

pcx_status pcxg_bits_splits(uint64_t seed, const uint64_t* words, int64_t nvars, int64_t n_inst,
                            int32_t kmin, int32_t kmax, uint64_t* out) {
  if (!words || !out || nvars <= 0 || n_inst < 0 || kmin < 0 || kmax < kmin) return PCX_EINVAL;
  parallel_for(n_inst, [&](int64_t k) {
    SplitMix64 r(stream_seed(seed, 5000011ULL + (uint64_t)k));
    uint64_t* w = out + k * nvars;
    std::memcpy(w, words, nvars * sizeof(uint64_t));
    const int64_t n = r.range(kmin, kmax);
    for (int64_t s = 0; s < n; ++s) {
      for (int t = 0; t < 64; ++t) {
        const int64_t v = (int64_t)r.below((uint64_t)nvars);
        const int pc = __builtin_popcountll(w[v]);
        if (pc < 2) continue;
        int pick = (int)r.below((uint64_t)pc);
        uint64_t m = w[v];
        while (pick--) m &= m - 1;           // drop the lowest set bits
        w[v] &= ~(m & (~m + 1));             // remove the chosen value
        break;
      }
    }
  });
  return PCX_OK;
}

}  // extern "C"
