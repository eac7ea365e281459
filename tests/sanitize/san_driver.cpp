// san_driver.cpp -- the host code of the build under AddressSanitizer and
// UndefinedBehaviorSanitizer: the FlatZinc front-end (pcx_model.cpp), the
// workload generators (pcx_gen.cpp) and the CPU oracle (oracle/*.c).  Test
// infrastructure; no GPU code is built or run.
//
//   san_driver models   reads models separated by lines holding only "%%"
//                       from stdin; each is parsed (errors are expected for
//                       malformed text), queried, extended and destroyed
//   san_driver gen      generators and oracle on small seeded workloads
// Exit status 0 when every call returned; a sanitizer report aborts.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "pcx_gen.h"
#include "pcx_model.h"

// pcx_model.cpp reports errors through the engine's error slot; the engine
// (HIP code) is not part of this build.
namespace pcx_internal {
pcx_status fail_with(pcx_status st, const std::string& msg) {
  static std::string last;
  last = msg;
  return st;
}
}  // namespace pcx_internal

extern "C" {
int pcxo_tell_embed(int64_t, const int32_t*, const int32_t*, int32_t*, int32_t*);
int pcxo_sort_table(int64_t, int32_t*, int32_t*, int32_t*, int32_t*);
int pcxo_fixpoint(int64_t, const int32_t*, const int32_t*, const int32_t*, const int32_t*,
                  const int64_t*, int64_t, int32_t*, int32_t*, int64_t, int64_t, void*);
int pcxo_ask(int32_t, int32_t, int32_t, int32_t, const int32_t*, const int32_t*);
int pcxo_bits_tell_embed(int64_t, const int32_t*, const int32_t*, uint64_t*);
int pcxo_bits_sort(int64_t, int32_t*, int32_t*, int32_t*, int32_t*);
int pcxo_bits_fixpoint(int64_t, const int32_t*, const int32_t*, const int32_t*, const int32_t*,
                       uint64_t*, int64_t, int32_t, int64_t, void*);
int pcxo_check_exhaustive(int op, int lo, int hi, int complete, char* err, int errlen);
int pcxo_bits_check(int op, int U, int complete, char* err, int errlen);
}

struct Result { int32_t has_changed, failed; int64_t sweeps, evals; };

static int run_model(const std::string& text) {
  pcx_model* m = nullptr;
  if (pcx_model_parse(text.c_str(), &m) != PCX_OK) return 0;
  const int64_t nv = pcx_model_num_vars(m), np = pcx_model_num_props(m);
  std::vector<int32_t> op(np), x(np), y(np), z(np), lb(nv), ub(nv);
  if (pcx_model_table(m, 0, op.data(), x.data(), y.data(), z.data()) != PCX_OK) return 1;
  if (pcx_model_store(m, lb.data(), ub.data()) != PCX_OK) return 1;
  for (int64_t v = 0; v < nv; ++v) {
    const char* n = pcx_model_var_name(m, (int32_t)v);
    if (n && pcx_model_var(m, n) < 0) return 1;
  }
  // the oracle on the decomposition, as the tests run it
  pcxo_tell_embed(np, op.data(), x.data(), lb.data(), ub.data());
  pcxo_sort_table(np, op.data(), x.data(), y.data(), z.data());
  Result r{};
  pcxo_fixpoint(np, op.data(), x.data(), y.data(), z.data(), nullptr, 0, lb.data(), ub.data(), nv, 1000, &r);
  // a second tell on the same model (may fail: names may be missing)
  (void)pcx_model_add(m, "var 0..3: extra_v; constraint int_le(extra_v, 2);");
  (void)pcx_model_add(m, text.c_str());
  pcx_model_destroy(m);
  return 0;
}

static int gen_and_oracle() {
  for (uint64_t seed = 1; seed <= 6; ++seed) {
    pcxg_params p;
    pcxg_params_default(&p);
    p.seed = seed;
    p.nvars = 300 + 50 * seed;
    p.nprops = 4 * p.nvars;
    p.planted = seed & 1;
    p.chain_len = 5 + (int)seed;
    std::vector<int32_t> op(p.nprops), x(p.nprops), y(p.nprops), z(p.nprops), lb(p.nvars), ub(p.nvars),
        w(p.nvars);
    if (pcxg_network(&p, op.data(), x.data(), y.data(), z.data(), lb.data(), ub.data(), w.data()) != PCX_OK)
      return 1;
    pcxo_tell_embed(p.nprops, op.data(), x.data(), lb.data(), ub.data());
    pcxo_sort_table(p.nprops, op.data(), x.data(), y.data(), z.data());
    const int n_inst = 5;
    std::vector<int32_t> lo(n_inst * p.nvars), hi(n_inst * p.nvars);
    if (pcxg_splits(seed, lb.data(), ub.data(), p.nvars, n_inst, 1, 20, lo.data(), hi.data()) != PCX_OK) return 1;
    for (int k = 0; k < n_inst; ++k) {
      Result r{};
      pcxo_fixpoint(p.nprops, op.data(), x.data(), y.data(), z.data(), nullptr, 0, lo.data() + k * p.nvars,
                    hi.data() + k * p.nvars, p.nvars, 0, &r);
    }
    for (int64_t i = 0; i < p.nprops; i += 7)
      (void)pcxo_ask(op[i], x[i], y[i], z[i], lb.data(), ub.data());
    std::vector<int64_t> off(n_inst + 1);
    if (pcxg_subsets(seed, p.nprops, n_inst, 10, p.nprops, off.data(), nullptr) != PCX_OK) return 1;
    std::vector<int32_t> idx(off[n_inst]);
    if (pcxg_subsets(seed, p.nprops, n_inst, 10, p.nprops, off.data(), idx.data()) != PCX_OK) return 1;
    const int64_t bv = 200, bp = 800;
    std::vector<int32_t> bo(bp), bx(bp), by(bp), bz(bp), wit(bv);
    std::vector<uint64_t> words(bv), inst(3 * bv);
    if (pcxg_bits_network(seed, bv, bp, 0, 0.3, 1.0, (int32_t)(seed & 1), bo.data(), bx.data(), by.data(),
                          bz.data(), words.data(), wit.data()) != PCX_OK)
      return 1;
    if (pcxg_bits_splits(seed, words.data(), bv, 3, 1, 10, inst.data()) != PCX_OK) return 1;
    pcxo_bits_tell_embed(bp, bo.data(), bx.data(), inst.data());
    pcxo_bits_sort(bp, bo.data(), bx.data(), by.data(), bz.data());
    Result r{};
    pcxo_bits_fixpoint(bp, bo.data(), bx.data(), by.data(), bz.data(), inst.data(), bv, 0, 0, &r);
  }
  char err[256];
  for (int op = 0; op < 10; ++op)
    if (pcxo_check_exhaustive(op, -2, 2, 0, err, sizeof err) != 0) return 1;
  for (int op = 0; op < 4; ++op)
    if (pcxo_bits_check(op, 4, 1, err, sizeof err) != 0) return 1;
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  if (!std::strcmp(argv[1], "gen")) return gen_and_oracle();
  if (std::strcmp(argv[1], "models")) return 2;
  std::string line, cur;
  long n = 0;
  while (std::getline(std::cin, line)) {
    if (line == "%%") {
      if (run_model(cur)) return 1;
      cur.clear();
      ++n;
    } else {
      cur += line;
      cur += '\n';
    }
  }
  if (!cur.empty() && run_model(cur)) return 1;
  std::printf("%ld models\n", n);
  return 0;
}
