/*
 * lala_pcx_adapter.hpp -- the adapter a lala-pc maintainer adds on the
 * reference side to run PIR's fixpoint through pcx.
 *
 * It replaces, for one PIR over a VStore of integer intervals,
 *     GaussSeidelIteration{}.fixpoint(pir.num_deductions(),
 *                                     [&](size_t i) { return pir.deduce(i); });
 * (tests/pir_test.cpp:60-62) with the same fixpoint computed on the GPU.
 * Only PIR's public surface is used:
 *   num_deductions()   pir.hpp:382-384
 *   load_deduce(i)     pir.hpp:358-366   (bytecode_type {op, x, y, z}, pir.hpp:36-47)
 *   vars()             pir.hpp:853-855
 *   operator[](v)      pir.hpp:840-842   (an interval with lb().value() / ub().value())
 *   embed(AVar, itv)   pir.hpp:354-356
 * and AVar's aty() / vid() and Sig's names from lala-core.  The header
 * includes nothing of lala: the caller includes PIR first.
 */
#ifndef LALA_PCX_ADAPTER_HPP
#define LALA_PCX_ADAPTER_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pcx.h"

namespace lala {

/* PIR's Sig to pcx's opcode numbering (the numbering only fixes the table's
 * sort order).  PIR admits no other Sig (pir.hpp:811). */
template <class SigT>
inline int32_t pcx_opcode(SigT s) {
  switch (s) {
    case ADD: return PCX_ADD;   case MUL: return PCX_MUL;
    case TDIV: return PCX_TDIV; case FDIV: return PCX_FDIV;
    case CDIV: return PCX_CDIV; case EDIV: return PCX_EDIV;
    case MIN: return PCX_MIN;   case MAX: return PCX_MAX;
    case EQ: return PCX_EQ;     case LEQ: return PCX_LEQ;
    default: return -1;
  }
}

/* One engine per PIR table.  The engine is (re)built from the PIR's
 * bytecodes whenever num_deductions() changed since the last call, so a PIR
 * that is told more constraints between fixpoints stays in step with it. */
template <class PIR>
class PcxFixpoint {
 public:
  explicit PcxFixpoint(int device = 0) : device_(device) {}
  PcxFixpoint(const PcxFixpoint&) = delete;
  PcxFixpoint& operator=(const PcxFixpoint&) = delete;
  ~PcxFixpoint() { if (e_) pcx_destroy(e_); }

  /* The drop-in call: propagates pir's store to its fixpoint and returns
   * has_changed as GaussSeidelIteration::fixpoint does.  Afterwards
   * pir.is_bot() holds exactly when pcx reported a failure. */
  bool operator()(PIR& pir) {
    sync(pir);
    const int64_t nv = pir.vars();
    std::vector<int32_t> lb(nv), ub(nv);
    for (int64_t v = 0; v < nv; ++v) {
      lb[v] = pir[(int)v].lb().value();
      ub[v] = pir[(int)v].ub().value();
    }
    pcx_result r{};
    check(pcx_fixpoint(e_, lb.data(), ub.data(), nv, &r));
    for (int64_t v = 0; v < nv; ++v)
      pir.embed(AVar(sty_, (int)v), typename PIR::universe_type(lb[v], ub[v]));
    return r.has_changed != 0;
  }

  pcx_engine* engine() { return e_; }

 private:
  static void check(pcx_status st) {
    if (st != PCX_OK) throw std::runtime_error(std::string("pcx: ") + pcx_last_error());
  }

  void sync(const PIR& pir) {
    const int64_t n = pir.num_deductions();
    if (e_ && n == told_) return;
    if (e_) pcx_destroy(e_);
    e_ = nullptr;
    pcx_config c;
    pcx_config_default(&c);
    c.device = device_;
    check(pcx_create(&c, &e_));
    std::vector<int32_t> op(n), x(n), y(n), z(n);
    for (int64_t i = 0; i < n; ++i) {
      const auto b = pir.load_deduce((int)i);
      op[i] = pcx_opcode(b.op);
      x[i] = b.x.vid(); y[i] = b.y.vid(); z[i] = b.z.vid();
      sty_ = b.x.aty();             // the store's type: the bytecodes name its variables
    }
    // The PIR already embedded its reified outputs in [0,1] when told
    // (pir.hpp:333-335); pcx_tell sorts as PIR does (pir.hpp:343-347).
    check(pcx_tell(e_, op.data(), x.data(), y.data(), z.data(), n, nullptr, nullptr, 0));
    told_ = n;
  }

  int device_;
  pcx_engine* e_ = nullptr;
  int64_t told_ = -1;
  decltype(AVar().aty()) sty_{};
};

}  // namespace lala

#endif  /* LALA_PCX_ADAPTER_HPP */
