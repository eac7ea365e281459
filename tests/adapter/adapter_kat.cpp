// adapter_kat.cpp -- runs known-answer tests through include/lala_pcx_adapter.hpp
// over the PIR stand-in of pir_stub.hpp, as a reference-side caller would.
//
// Reads a script on stdin:
//   vars N                 then N lines "lb ub" (the initial store)
//   tell K                 then K lines "OP x y z" (OP in ADD MUL TDIV FDIV
//                          CDIV EDIV MIN MAX EQ LEQ)
//   embed v lb ub          meet one variable's domain
//   fix                    run the adapter's fixpoint; prints
//                          "fix <num_deductions> <has_changed> <is_bot>" and
//                          then N lines "lb ub"
// Exit status 0, or 2 with a message when pcx reports an error.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <map>
#include <string>

#include "pir_stub.hpp"
#include "lala_pcx_adapter.hpp"

using namespace lala;

int main() {
  const std::map<std::string, Sig> ops{{"ADD", ADD}, {"MUL", MUL}, {"TDIV", TDIV}, {"FDIV", FDIV},
                                       {"CDIV", CDIV}, {"EDIV", EDIV}, {"MIN", MIN}, {"MAX", MAX},
                                       {"EQ", EQ}, {"LEQ", LEQ}};
  StubPIR* pir = nullptr;
  PcxFixpoint<StubPIR> fix;
  std::string cmd;
  try {
    while (std::cin >> cmd) {
      if (cmd == "vars") {
        int n;
        std::cin >> n;
        delete pir;
        pir = new StubPIR(n);
        for (int v = 0; v < n; ++v) {
          long long a, b;
          std::cin >> a >> b;
          pir->embed(AVar(StubPIR::kStore, v), Itv((int)a, (int)b));
        }
      } else if (cmd == "tell") {
        int k;
        std::cin >> k;
        std::vector<bytecode_type> t(k);
        for (int i = 0; i < k; ++i) {
          std::string op;
          int x, y, z;
          std::cin >> op >> x >> y >> z;
          t[i] = bytecode_type{ops.at(op), AVar(StubPIR::kStore, x), AVar(StubPIR::kStore, y),
                               AVar(StubPIR::kStore, z)};
        }
        pir->tell(t);
      } else if (cmd == "embed") {
        int v;
        long long a, b;
        std::cin >> v >> a >> b;
        pir->embed(AVar(StubPIR::kStore, v), Itv((int)a, (int)b));
      } else if (cmd == "fix") {
        const bool changed = fix(*pir);
        std::printf("fix %zu %d %d\n", pir->num_deductions(), (int)changed, (int)pir->is_bot());
        for (int v = 0; v < pir->vars(); ++v) std::printf("%d %d\n", (*pir)[v].lb().value(), (*pir)[v].ub().value());
        std::fflush(stdout);
      } else {
        std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
        return 1;
      }
    }
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "%s\n", ex.what());
    return 2;
  }
  delete pir;
  return 0;
}
