// pir_stub.hpp -- test stand-in for the public surface of lala-pc's PIR over
// a VStore of int intervals, as include/lala_pcx_adapter.hpp uses it.  Test
// infrastructure: it lets the adapter compile and run without lala-core,
// which is not part of the reference tree.
//
// Surface mirrored (file:line of the reference):
//   bytecode_type {Sig op; AVar x, y, z}          include/lala/pir.hpp:36-47
//   num_deductions()                              pir.hpp:382-384
//   load_deduce(i)                                pir.hpp:358-366
//   deduce(tell): append, embed reified x in [0,1], stable sort by
//                 (op, y, x, z)                   pir.hpp:327-352
//   embed(AVar, itv) = meet, true when tightened  pir.hpp:354-356
//   operator[](v), vars(), is_bot()               pir.hpp:831-855
// Sig names are lala-core's; only their identity matters here.
#pragma once
#include <algorithm>
#include <climits>
#include <cstdint>
#include <vector>

namespace lala {

enum Sig { ADD, SUB, MUL, TDIV, FDIV, CDIV, EDIV, MIN, MAX, EQ, NEQ, LEQ, GEQ };
using AType = int;

struct AVar {
  AType ty = -1;
  int v = -1;
  AVar() = default;
  AVar(AType t, int vid) : ty(t), v(vid) {}
  AType aty() const { return ty; }
  int vid() const { return v; }
};

struct Bound {
  int v;
  int value() const { return v; }
};

struct Itv {
  Bound l{INT_MIN}, u{INT_MAX};
  Itv() = default;
  Itv(int lo, int hi) : l{lo}, u{hi} {}
  Bound lb() const { return l; }
  Bound ub() const { return u; }
};

struct bytecode_type {
  Sig op;
  AVar x, y, z;
};

class StubPIR {
 public:
  using universe_type = Itv;
  static constexpr AType kStore = 0;     // the store's abstract type (sty in the tests)

  explicit StubPIR(int nvars) : store_(nvars) {}

  int vars() const { return (int)store_.size(); }
  Itv operator[](int v) const { return store_[v]; }
  bool is_bot() const {
    for (const Itv& d : store_) if (d.l.v > d.u.v) return true;
    return false;
  }
  bool embed(AVar x, const Itv& d) {
    Itv& s = store_[x.vid()];
    const Itv old = s;
    s.l.v = std::max(s.l.v, d.l.v);
    s.u.v = std::min(s.u.v, d.u.v);
    return s.l.v != old.l.v || s.u.v != old.u.v;
  }
  size_t num_deductions() const { return code_.size(); }
  bytecode_type load_deduce(int i) const { return code_[i]; }

  void tell(const std::vector<bytecode_type>& t) {
    for (const bytecode_type& b : t) {
      code_.push_back(b);
      if (b.op == EQ || b.op == LEQ) embed(b.x, Itv(0, 1));
    }
    std::stable_sort(code_.begin(), code_.end(), [](const bytecode_type& a, const bytecode_type& b) {
      if (a.op != b.op) return a.op < b.op;
      if (a.y.vid() != b.y.vid()) return a.y.vid() < b.y.vid();
      if (a.x.vid() != b.x.vid()) return a.x.vid() < b.x.vid();
      return a.z.vid() < b.z.vid();
    });
  }

 private:
  std::vector<Itv> store_;
  std::vector<bytecode_type> code_;
};

}  // namespace lala
