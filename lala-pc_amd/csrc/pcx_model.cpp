// pcx_model.cpp -- front-end: FlatZinc-style models to PIR ternary form
// (include/pcx_model.h).
//
// The text is tokenized and parsed into expression trees item by item; each
// item is decomposed at once into ternary propagators `x = y op z` over the
// model's variables (declared, auxiliary, constant) and into domain
// restrictions.  The decomposition follows the shape the reference's
// ternarize gives (pir_test.cpp:193-206 lists its generated variables):
//   - a term (int_plus(x, y), int_eq(y, 2), ...) used as a value becomes an
//     auxiliary variable defined by one propagator (two for the negated
//     comparisons, three for int_abs: -a, max(a, -a), and 0 <= the result);
//   - an integer or boolean constant becomes a shared fixed singleton
//     variable (__CONSTANT_c);
//   - a comparison between a variable (or term) and a constant restricts that
//     variable's domain and adds no propagator, as the store keeps unary
//     constraints (pir_test.cpp:214-221: x + y = 5 is one propagator);
//   - `t = term` with a variable t puts the term's result directly in t
//     (pir_test.cpp:184-188: x = (y == 2) is one propagator);
//   - a predicate posted as true or false is the reified form with the
//     constant 1 or 0 as its result: y <= z is LEQ(1, y, z), y < z is
//     LEQ(0, z, y) (not z <= y), x != y is EQ(0, x, y), as in pir.hpp:270-273.
#include <algorithm>
#include <cctype>
#include <climits>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "../../include/pcx_model.h"

namespace pcx_internal {
pcx_status fail_with(pcx_status st, const std::string& msg);
}

namespace {

// ---- tokens -----------------------------------------------------------------
enum Tok { tEnd, tIdent, tInt, tPunct };

struct Token {
  Tok kind = tEnd;
  std::string_view text;  // identifier or punctuation (a view into the model text)
  long long value = 0;    // integer literal (no sign)
  int line = 1;
};

struct ParseError {
  pcx_status status;
  std::string msg;
};

// Tokens on demand: a model of millions of constraints is never held as a
// token list.
class Lexer {
 public:
  explicit Lexer(const char* s) : p_(s) {}
  Token next() {
    for (;;) {
      const char c = *p_;
      if (c == '\n') { ++line_; ++p_; continue; }
      if (c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v') { ++p_; continue; }
      if (c == '%') { while (*p_ && *p_ != '\n') ++p_; continue; }
      break;
    }
    Token t;
    t.line = line_;
    const char c = *p_;
    if (!c) return t;
    const char* b = p_;
    if (std::isalpha((unsigned char)c) || c == '_') {
      while (std::isalnum((unsigned char)*p_) || *p_ == '_') ++p_;
      t.kind = tIdent;
    } else if (c >= '0' && c <= '9') {
      long long v = 0;
      while (*p_ >= '0' && *p_ <= '9') {
        v = v * 10 + (*p_ - '0');
        if (v > (long long)INT32_MAX + 1)
          throw ParseError{PCX_EINVAL, "line " + std::to_string(line_) + ": integer out of int32 range"};
        ++p_;
      }
      t.kind = tInt;
      t.value = v;
    } else if (c == '"') {                       // strings only occur in annotations
      ++p_;
      while (*p_ && *p_ != '"') { if (*p_ == '\n') ++line_; ++p_; }
      if (*p_) ++p_;
      t.kind = tPunct;
      b = "\"";
      t.text = std::string_view(b, 1);
      return t;
    } else {
      t.kind = tPunct;
      p_ += ((c == '.' && p_[1] == '.') || (c == ':' && p_[1] == ':')) ? 2 : 1;
    }
    t.text = std::string_view(b, (size_t)(p_ - b));
    return t;
  }
 private:
  const char* p_;
  int line_ = 1;
};

// ---- expressions --------------------------------------------------------------
enum Ek { eInt, eBool, eIdent, eIndex, eCall, eArray, eRange, eSet };

struct Expr {
  Ek kind;
  long long v = 0;            // eInt / eBool value; eRange low
  long long hi = 0;           // eRange high
  std::string_view name;      // eIdent / eIndex / eCall (a view into the text or a literal)
  std::vector<Expr> args;     // eCall arguments, eArray / eSet elements, eIndex index
  int line = 0;
};

}  // namespace

// ---- the model ----------------------------------------------------------------
struct pcx_model {
  struct Var { std::string_view name; int32_t lb, ub; };
  std::deque<std::string> names;                                  // owned names (stable addresses)
  std::vector<Var> vars;
  std::unordered_map<std::string_view, int32_t> scalars;              // variable name -> index
  std::unordered_map<std::string_view, std::vector<int32_t>> arrays;  // var array -> element indices
  std::unordered_map<std::string_view, long long> par_ints;
  std::unordered_map<std::string_view, std::vector<long long>> par_arrays;
  std::unordered_map<long long, int32_t> consts;
  std::vector<int32_t> op, x, y, z;
  int32_t n_aux = 0;
  std::vector<int32_t> tnodes;                 // PC's formula trees, 4 int32 per node
  std::vector<int64_t> toff{0};                // node offset of each formula
  std::string_view own(std::string_view s) { names.emplace_back(s); return names.back(); }
};

namespace {

struct Builder {
  pcx_model& m;
  int line = 0;

  [[noreturn]] void fail(pcx_status st, const std::string& msg) const {
    throw ParseError{st, "line " + std::to_string(line) + ": " + msg};
  }

  int32_t new_var(std::string_view name, long long lb, long long ub) {
    if ((long long)m.vars.size() >= PCX_MAX_VARS) fail(PCX_EINVAL, "too many variables");
    m.vars.push_back({m.own(name), (int32_t)lb, (int32_t)ub});
    return (int32_t)m.vars.size() - 1;
  }
  int32_t aux(long long lb = PCX_MINF, long long ub = PCX_INF) {
    return new_var("__AUX_" + std::to_string(m.n_aux++), lb, ub);
  }
  int32_t aux01() { return aux(0, 1); }
  int32_t constant(long long c) {
    if (c < INT32_MIN || c > INT32_MAX) fail(PCX_EINVAL, "constant out of int32 range");
    auto it = m.consts.find(c);
    if (it != m.consts.end()) return it->second;
    const int32_t v = new_var("__CONSTANT_" + std::to_string(c), c, c);
    m.consts[c] = v;
    return v;
  }
  void emit(int o, int32_t vx, int32_t vy, int32_t vz) {
    m.op.push_back(o);
    m.x.push_back(vx);
    m.y.push_back(vy);
    m.z.push_back(vz);
  }
  void restrict_to(int32_t v, long long lo, long long hi) {
    auto& d = m.vars[v];
    if (lo > d.lb) d.lb = (int32_t)std::min<long long>(lo, INT32_MAX);
    if (hi < d.ub) d.ub = (int32_t)std::max<long long>(hi, INT32_MIN);
  }
  // An empty domain: the store is bot from the start.
  void contradiction() { restrict_to(constant(0), 1, 0); }

  // Is e a constant (literal, or a parameter)?  Sets *c.
  bool const_of(const Expr& e, long long* c) const {
    switch (e.kind) {
      case eInt: case eBool: *c = e.v; return true;
      case eIdent: {
        auto it = m.par_ints.find(e.name);
        if (it != m.par_ints.end()) { *c = it->second; return true; }
        return false;
      }
      case eIndex: {
        auto it = m.par_arrays.find(e.name);
        long long i;
        if (it == m.par_arrays.end() || !const_of(e.args[0], &i)) return false;
        if (i < 1 || i > (long long)it->second.size()) return false;
        *c = it->second[i - 1];
        return true;
      }
      default: return false;
    }
  }
  // A plain model variable (not a term)?  Sets *v.
  bool var_of(const Expr& e, int32_t* v) {
    if (e.kind == eIdent) {
      auto it = m.scalars.find(e.name);
      if (it == m.scalars.end()) return false;
      *v = it->second;
      return true;
    }
    if (e.kind == eIndex) {
      auto it = m.arrays.find(e.name);
      long long i;
      if (it == m.arrays.end() || !const_of(e.args[0], &i)) return false;
      if (i < 1 || i > (long long)it->second.size()) fail(PCX_EINVAL, "index out of range: " + std::string(e.name));
      *v = it->second[i - 1];
      return true;
    }
    return false;
  }

  // The elements of an array argument (literal, var array or par array).
  std::vector<Expr> elements(const Expr& e) {
    if (e.kind == eArray) return e.args;
    if (e.kind == eIdent) {
      std::vector<Expr> out;
      auto va = m.arrays.find(e.name);
      if (va != m.arrays.end()) {
        for (size_t i = 0; i < va->second.size(); ++i) {
          Expr x;
          x.kind = eIndex;
          x.name = e.name;
          Expr idx;
          idx.kind = eInt;
          idx.v = (long long)i + 1;
          x.args.push_back(idx);
          out.push_back(x);
        }
        return out;
      }
      auto pa = m.par_arrays.find(e.name);
      if (pa != m.par_arrays.end()) {
        for (long long c : pa->second) {
          Expr x;
          x.kind = eInt;
          x.v = c;
          out.push_back(x);
        }
        return out;
      }
    }
    fail(PCX_EINVAL, "expected an array");
  }

  // The variable holding the value of e (a term is materialized).
  int32_t value(const Expr& e) {
    long long c;
    int32_t v;
    if (const_of(e, &c)) return constant(c);
    if (var_of(e, &v)) return v;
    if (e.kind == eIdent || e.kind == eIndex) fail(PCX_EINVAL, "unknown name: " + std::string(e.name));
    if (e.kind == eCall) return term(e, -1);
    fail(PCX_EINVAL, "expected an integer or boolean expression");
  }

  static bool is_arith(std::string_view f) {
    static const char* names[] = {"int_plus", "int_minus", "int_times", "int_div", "int_tdiv",
                                  "int_fdiv", "int_cdiv", "int_ediv", "int_min", "int_max",
                                  "int_neg", "int_abs"};
    for (const char* n : names) if (f == n) return true;
    return false;
  }

  // Materialize call e with its result in `out` (a fresh auxiliary if < 0).
  int32_t term(const Expr& e, int32_t out) {
    const std::string_view f = e.name;
    if (!is_arith(f)) return reify(e, out);
    const size_t want = (f == "int_neg" || f == "int_abs") ? 1 : 2;
    if (e.args.size() != want) fail(PCX_EINVAL, std::string(f) + ": expected " + std::to_string(want) + " arguments");
    const int32_t A = value(e.args[0]);
    const int32_t B = want == 2 ? value(e.args[1]) : -1;
    const int32_t R = out >= 0 ? out : aux();
    if (f == "int_plus") emit(PCX_ADD, R, A, B);
    else if (f == "int_minus") emit(PCX_ADD, A, R, B);            // a = r + b
    else if (f == "int_times") emit(PCX_MUL, R, A, B);
    else if (f == "int_div" || f == "int_tdiv") emit(PCX_TDIV, R, A, B);
    else if (f == "int_fdiv") emit(PCX_FDIV, R, A, B);
    else if (f == "int_cdiv") emit(PCX_CDIV, R, A, B);
    else if (f == "int_ediv") emit(PCX_EDIV, R, A, B);
    else if (f == "int_min") emit(PCX_MIN, R, A, B);
    else if (f == "int_max") emit(PCX_MAX, R, A, B);
    else if (f == "int_neg") emit(PCX_ADD, constant(0), A, R);    // 0 = a + r
    else {                                                         // int_abs: r = max(a, -a), r >= 0
      const int32_t N = aux();
      emit(PCX_ADD, constant(0), A, N);
      emit(PCX_MAX, R, A, N);
      emit(PCX_LEQ, constant(1), constant(0), R);
    }
    return R;
  }

  // r = 1 - s
  void negate(int32_t r, int32_t s) { emit(PCX_ADD, constant(1), r, s); }

  // Is v one of the shared constants?  Sets *k.
  bool fixed(int32_t v, long long* k) const {
    if (v < 0) return false;
    const auto& d = m.vars[v];
    if (d.lb != d.ub) return false;
    auto it = m.consts.find(d.lb);
    if (it == m.consts.end() || it->second != v) return false;
    *k = d.lb;
    return true;
  }

  // max / min over a list of booleans into r (chains of binary propagators)
  int32_t fold(int o, const std::vector<int32_t>& v, int32_t r, long long empty) {
    if (v.empty()) {
      if (r < 0) return constant(empty);
      restrict_to(r, empty, empty);
      return r;
    }
    if (v.size() == 1) {
      if (r < 0) return v[0];
      emit(PCX_EQ, constant(1), r, v[0]);
      return r;
    }
    int32_t acc = v[0];
    for (size_t i = 1; i < v.size(); ++i) {
      const int32_t t = (i + 1 == v.size() && r >= 0) ? r : aux01();
      emit(o, t, acc, v[i]);
      acc = t;
    }
    return acc;
  }

  std::vector<int32_t> values(const Expr& arr) {
    std::vector<int32_t> out;
    for (const Expr& e : elements(arr)) out.push_back(value(e));
    return out;
  }

  // Σ a_i x_i into a fresh auxiliary
  int32_t linear(const Expr& as, const Expr& xs) {
    const std::vector<Expr> a = elements(as), xv = elements(xs);
    if (a.size() != xv.size()) fail(PCX_EINVAL, "int_lin_*: coefficient and variable arrays differ in length");
    int32_t acc = -1;
    for (size_t i = 0; i < a.size(); ++i) {
      long long c;
      if (!const_of(a[i], &c)) fail(PCX_EINVAL, "int_lin_*: coefficients must be constants");
      int32_t t = value(xv[i]);
      if (c != 1) {
        const int32_t p = aux();
        emit(PCX_MUL, p, constant(c), t);
        t = p;
      }
      if (acc < 0) acc = t;
      else {
        const int32_t s = aux();
        emit(PCX_ADD, s, acc, t);
        acc = s;
      }
    }
    return acc < 0 ? constant(0) : acc;
  }

  // The set argument of set_in: a range or a set literal (or a par int set).
  std::vector<long long> set_values(const Expr& s, bool* is_range, long long* lo, long long* hi) {
    std::vector<long long> vals;
    if (s.kind == eRange) {
      *is_range = true;
      *lo = s.v;
      *hi = s.hi;
      return vals;
    }
    if (s.kind != eSet) fail(PCX_EINVAL, "set_in: expected a range or a set literal");
    for (const Expr& e : s.args) {
      long long c;
      if (!const_of(e, &c)) fail(PCX_EINVAL, "set literal: expected constants");
      vals.push_back(c);
    }
    std::sort(vals.begin(), vals.end());
    vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
    *is_range = !vals.empty() && vals.back() - vals.front() + 1 == (long long)vals.size();
    if (*is_range) { *lo = vals.front(); *hi = vals.back(); }
    return vals;
  }

  // r <=> x in S
  int32_t set_in(int32_t X, const Expr& s, int32_t r) {
    bool rng;
    long long lo = 0, hi = 0;
    const std::vector<long long> vals = set_values(s, &rng, &lo, &hi);
    if (rng) {
      const int32_t s1 = aux01(), s2 = aux01();
      emit(PCX_LEQ, s1, constant(lo), X);
      emit(PCX_LEQ, s2, X, constant(hi));
      const int32_t R = r >= 0 ? r : aux01();
      emit(PCX_MIN, R, s1, s2);
      return R;
    }
    std::vector<int32_t> eqs;
    for (long long c : vals) {
      const int32_t e = aux01();
      emit(PCX_EQ, e, X, constant(c));
      eqs.push_back(e);
    }
    return fold(PCX_MAX, eqs, r, 0);
  }

  // Post x in S as a constraint.
  void post_set_in(int32_t X, const Expr& s) {
    bool rng;
    long long lo = 0, hi = 0;
    const std::vector<long long> vals = set_values(s, &rng, &lo, &hi);
    if (rng) { restrict_to(X, lo, hi); return; }
    if (vals.empty()) { contradiction(); return; }
    restrict_to(X, vals.front(), vals.back());
    std::vector<int32_t> eqs;
    for (long long c : vals) {
      const int32_t e = aux01();
      emit(PCX_EQ, e, X, constant(c));
      eqs.push_back(e);
    }
    const int32_t t = fold(PCX_MAX, eqs, -1, 0);
    restrict_to(t, 1, 1);
  }

  // element: c = as[b] (1-based): b = i implies c = as[i]
  void element(const Expr& b, const Expr& as, const Expr& c) {
    const int32_t B = value(b), C = value(c);
    const std::vector<Expr> a = elements(as);
    if (a.empty()) { contradiction(); return; }
    restrict_to(B, 1, (long long)a.size());
    for (size_t i = 0; i < a.size(); ++i) {
      const int32_t e = aux01(), g = aux01();
      emit(PCX_EQ, e, B, constant((long long)i + 1));
      emit(PCX_EQ, g, C, value(a[i]));
      emit(PCX_LEQ, constant(1), e, g);
    }
  }

  void arity(const Expr& e, size_t n) const {
    if (e.args.size() != n)
      throw ParseError{PCX_EINVAL, "line " + std::to_string(line) + ": " + std::string(e.name) + ": expected " +
                                       std::to_string(n) + " arguments"};
  }

  // The 0/1 variable `out` (fresh if < 0) equivalent to predicate e.
  int32_t reify(const Expr& e, int32_t out) {
    const std::string_view f = e.name;
    if (out >= 0) restrict_to(out, 0, 1);
    auto R = [&]() { return out >= 0 ? out : aux01(); };
    auto cmp = [&](int o, bool swap, bool neg) {
      arity(e, 2);
      const int32_t A = value(e.args[swap ? 1 : 0]), B = value(e.args[swap ? 0 : 1]);
      long long k;
      if (neg && fixed(out, &k) && (k == 0 || k == 1)) {   // posted: the opposite constant
        emit(o, constant(1 - k), A, B);
        return out;
      }
      const int32_t r = R();
      if (!neg) { emit(o, r, A, B); return r; }
      const int32_t s = aux01();
      emit(o, s, A, B);
      negate(r, s);
      return r;
    };
    if (f == "int_eq" || f == "bool_eq" || f == "nbool_equiv" || f == "bool_equiv") return cmp(PCX_EQ, false, false);
    if (f == "int_ne" || f == "bool_ne" || f == "bool_xor" || f == "nbool_xor") return cmp(PCX_EQ, false, true);
    if (f == "int_le" || f == "bool_le") return cmp(PCX_LEQ, false, false);
    if (f == "int_ge" || f == "bool_ge") return cmp(PCX_LEQ, true, false);
    if (f == "int_lt" || f == "bool_lt") return cmp(PCX_LEQ, true, true);    // a < b: not b <= a
    if (f == "int_gt" || f == "bool_gt") return cmp(PCX_LEQ, false, true);   // a > b: not a <= b
    if (f == "bool_not") {
      arity(e, 1);
      const int32_t A = value(e.args[0]);
      const int32_t r = R();
      negate(r, A);
      return r;
    }
    if (f == "bool_and" || f == "nbool_and" || f == "bool_or" || f == "nbool_or") {
      std::vector<int32_t> v;
      for (const Expr& a : e.args) v.push_back(value(a));
      const bool conj = f == "bool_and" || f == "nbool_and";
      return fold(conj ? PCX_MIN : PCX_MAX, v, out, conj ? 1 : 0);
    }
    if (f == "array_bool_and" || f == "array_bool_or") {
      arity(e, 1);
      const bool conj = f == "array_bool_and";
      return fold(conj ? PCX_MIN : PCX_MAX, values(e.args[0]), out, conj ? 1 : 0);
    }
    if (f == "bool_clause") {                  // OR(as) or OR(not bs): max(as) >= min(bs)
      arity(e, 2);
      const int32_t t = fold(PCX_MAX, values(e.args[0]), -1, 0);
      const int32_t u = fold(PCX_MIN, values(e.args[1]), -1, 1);
      const int32_t r = R();
      emit(PCX_LEQ, r, u, t);
      return r;
    }
    if (f == "set_in") {
      arity(e, 2);
      return set_in(value(e.args[0]), e.args[1], out);
    }
    if (f == "int_lin_le" || f == "int_lin_eq" || f == "int_lin_ne") {
      arity(e, 3);
      const int32_t s = linear(e.args[0], e.args[1]);
      const int32_t k = value(e.args[2]);
      const int32_t r = R();
      if (f == "int_lin_le") emit(PCX_LEQ, r, s, k);
      else if (f == "int_lin_eq") emit(PCX_EQ, r, s, k);
      else { const int32_t q = aux01(); emit(PCX_EQ, q, s, k); negate(r, q); }
      return r;
    }
    if (is_arith(f)) fail(PCX_EINVAL, std::string(f) + " is not a predicate");
    fail(PCX_ENOTSUP, "unsupported predicate: " + std::string(f));
  }

  // Post e as a constraint that holds.
  void post(const Expr& e) {
    long long c;
    int32_t v;
    if (const_of(e, &c)) { if (!c) contradiction(); return; }
    if (var_of(e, &v)) { restrict_to(v, 1, 1); return; }
    if (e.kind != eCall) fail(PCX_EINVAL, "expected a constraint");
    const std::string_view f = e.name;
    // reified forms: p_reif(args..., r)
    if (f.size() > 5 && f.compare(f.size() - 5, 5, "_reif") == 0) {
      if (e.args.empty()) fail(PCX_EINVAL, std::string(f) + ": missing arguments");
      Expr p = e;
      p.name = f.substr(0, f.size() - 5);
      p.args.pop_back();
      const Expr& r = e.args.back();
      int32_t rv;
      if (const_of(r, &c)) { reify(p, constant(c)); return; }
      if (var_of(r, &rv)) { restrict_to(rv, 0, 1); reify(p, rv); return; }
      fail(PCX_EINVAL, std::string(f) + ": the last argument must be a boolean variable");
    }
    // comparisons: a domain restriction when one side is a constant
    int kind = -1;                     // 0 le, 1 lt, 2 eq, 3 ne
    bool swap = false;
    if (f == "int_le" || f == "bool_le") kind = 0;
    else if (f == "int_ge" || f == "bool_ge") { kind = 0; swap = true; }
    else if (f == "int_lt" || f == "bool_lt") kind = 1;
    else if (f == "int_gt" || f == "bool_gt") { kind = 1; swap = true; }
    else if (f == "int_eq" || f == "bool_eq" || f == "bool2int" || f == "nbool_equiv" || f == "bool_equiv") kind = 2;
    else if (f == "int_ne" || f == "bool_ne") kind = 3;
    if (kind >= 0) {
      arity(e, 2);
      const Expr& a = e.args[swap ? 1 : 0];
      const Expr& b = e.args[swap ? 0 : 1];
      long long ca, cb;
      const bool ka = const_of(a, &ca), kb = const_of(b, &cb);
      if (ka && kb) {
        const bool ok = kind == 0 ? ca <= cb : kind == 1 ? ca < cb : kind == 2 ? ca == cb : ca != cb;
        if (!ok) contradiction();
        return;
      }
      if (kind == 2) {                 // t = term: the term's result goes into t
        int32_t t;
        if (b.kind == eCall && var_of(a, &t)) { term(b, t); return; }
        if (a.kind == eCall && var_of(b, &t)) { term(a, t); return; }
      }
      if (kind != 3 && (ka || kb)) {
        const int32_t X = value(ka ? b : a);
        const long long k = ka ? ca : cb;
        if (kind == 2) restrict_to(X, k, k);
        else if (kb) restrict_to(X, PCX_MINF, kind == 0 ? k : k - 1);     // x <= k, x < k
        else restrict_to(X, kind == 0 ? k : k + 1, PCX_INF);              // k <= x, k < x
        return;
      }
      const int32_t A = value(a), B = value(b);
      if (kind == 0) emit(PCX_LEQ, constant(1), A, B);
      else if (kind == 1) emit(PCX_LEQ, constant(0), B, A);
      else if (kind == 2) emit(PCX_EQ, constant(1), A, B);
      else emit(PCX_EQ, constant(0), A, B);
      return;
    }
    // c = a op b
    if (is_arith(f)) {
      const size_t in = (f == "int_neg" || f == "int_abs") ? 1 : 2;
      arity(e, in + 1);
      Expr t = e;
      t.args.pop_back();
      term(t, value(e.args.back()));
      return;
    }
    if (f == "bool_not") {
      if (e.args.size() == 1) {                  // not a: a is false
        const Expr& a = e.args[0];
        if (const_of(a, &c)) { if (c) contradiction(); return; }
        if (var_of(a, &v)) { restrict_to(v, 0, 0); return; }
        if (a.kind != eCall) fail(PCX_EINVAL, "bool_not: expected a boolean");
        reify(a, constant(0));
        return;
      }
      arity(e, 2);
      negate(value(e.args[1]), value(e.args[0]));
      return;
    }
    if (f == "bool_and" || f == "nbool_and") {
      if (e.args.size() == 3 && f == "bool_and") {
        emit(PCX_MIN, value(e.args[2]), value(e.args[0]), value(e.args[1]));
        return;
      }
      for (const Expr& a : e.args) post(a);
      return;
    }
    if ((f == "bool_or" || f == "bool_xor") && e.args.size() == 3) {
      Expr p = e;
      p.args.pop_back();
      const int32_t r = value(e.args[2]);
      restrict_to(r, 0, 1);
      reify(p, r);
      return;
    }
    if (f == "array_bool_and" || f == "array_bool_or") {
      if (e.args.size() == 2) {
        Expr p = e;
        p.args.pop_back();
        const int32_t r = value(e.args[1]);
        restrict_to(r, 0, 1);
        reify(p, r);
        return;
      }
      if (f == "array_bool_and") { for (const Expr& a : elements(e.args[0])) post(a); return; }
    }
    if (f == "array_int_element" || f == "array_var_int_element" || f == "array_bool_element" ||
        f == "array_var_bool_element") {
      arity(e, 3);
      element(e.args[0], e.args[1], e.args[2]);
      return;
    }
    if (f == "set_in") {
      arity(e, 2);
      post_set_in(value(e.args[0]), e.args[1]);
      return;
    }
    if (f == "int_lin_le" || f == "int_lin_eq" || f == "int_lin_ne") {
      arity(e, 3);
      long long k;
      const int32_t s = linear(e.args[0], e.args[1]);
      if (const_of(e.args[2], &k) && f != "int_lin_ne") {
        if (f == "int_lin_le") restrict_to(s, PCX_MINF, k); else restrict_to(s, k, k);
        return;
      }
      const int32_t K = value(e.args[2]);
      if (f == "int_lin_le") emit(PCX_LEQ, constant(1), s, K);
      else emit(PCX_EQ, constant(f == "int_lin_eq" ? 1 : 0), s, K);
      return;
    }
    reify(e, constant(1));             // any other predicate: its reification is true
  }
};

// ---- PC's formula trees ------------------------------------------------------
// Each constraint as the tree PC asks (formula.hpp: Inequality 760-770,
// Equality / Disequality 617-628, VariableLiteral 101-110, Conjunction,
// Disjunction, Biconditional, Xor, Implication 268-555), over the model's
// variables and constants.  Built after the constraint was decomposed, so a
// construct the tree form does not cover takes the conjunction of the
// ternary propagators the constraint was decomposed into.
struct TreeBuilder {
  Builder& b;
  std::vector<int32_t> out;
  bool ok = true;

  void node(int k, int32_t a = 0, int32_t bb = 0, int32_t c = 0) {
    out.push_back(k); out.push_back(a); out.push_back(bb); out.push_back(c);
  }
  static bool arith2(std::string_view f, int* k) {
    if (f == "int_plus") *k = PCX_FT_ADD;
    else if (f == "int_minus") *k = PCX_FT_SUB;
    else if (f == "int_times") *k = PCX_FT_MUL;
    else if (f == "int_div" || f == "int_tdiv") *k = PCX_FT_TDIV;
    else if (f == "int_fdiv") *k = PCX_FT_FDIV;
    else if (f == "int_cdiv") *k = PCX_FT_CDIV;
    else if (f == "int_ediv") *k = PCX_FT_EDIV;
    else if (f == "int_min") *k = PCX_FT_MIN;
    else if (f == "int_max") *k = PCX_FT_MAX;
    else return false;
    return true;
  }
  void term(const Expr& e) {
    long long c;
    int32_t v;
    if (!ok) return;
    if (b.const_of(e, &c)) {
      if (c < INT32_MIN || c > INT32_MAX) { ok = false; return; }
      node(PCX_FT_CONST, (int32_t)c);
      return;
    }
    if (b.var_of(e, &v)) { node(PCX_FT_VAR, v); return; }
    if (e.kind != eCall) { ok = false; return; }
    int k;
    if (arith2(e.name, &k) && e.args.size() == 2) {
      term(e.args[0]);
      term(e.args[1]);
      node(k);
      return;
    }
    if ((e.name == "int_neg" || e.name == "int_abs") && e.args.size() == 1) {
      term(e.args[0]);
      node(e.name == "int_neg" ? PCX_FT_NEG : PCX_FT_ABS);
      return;
    }
    if (e.name == "bool2int" && e.args.size() == 1) { term(e.args[0]); return; }
    ok = false;                                  // a formula used as a term
  }
  void lit(const Expr& e) { term(e); node(PCX_FT_LIT); }
  void lits(const Expr& arr, bool neg) {
    for (const Expr& a : b.elements(arr)) {
      lit(a);
      if (neg) node(PCX_FT_NOT);
    }
  }
  void formula(const Expr& e) {
    long long c;
    int32_t v;
    if (!ok) return;
    if (b.const_of(e, &c)) { node(c ? PCX_FT_TRUE : PCX_FT_FALSE); return; }
    if (b.var_of(e, &v)) { node(PCX_FT_VAR, v); node(PCX_FT_LIT); return; }
    if (e.kind != eCall) { ok = false; return; }
    const std::string_view f = e.name;
    const size_t n = e.args.size();
    if (f.size() > 5 && f.compare(f.size() - 5, 5, "_reif") == 0 && n >= 1) {
      Expr p = e;
      p.name = f.substr(0, f.size() - 5);
      p.args.pop_back();
      formula(p);
      lit(e.args.back());
      node(PCX_FT_EQUIV);
      return;
    }
    int cmp = -1;
    bool swap = false;
    if (f == "int_le" || f == "bool_le") cmp = PCX_FT_LE;
    else if (f == "int_ge" || f == "bool_ge") { cmp = PCX_FT_LE; swap = true; }
    else if (f == "int_lt" || f == "bool_lt") cmp = PCX_FT_LT;
    else if (f == "int_gt" || f == "bool_gt") { cmp = PCX_FT_LT; swap = true; }
    else if (f == "int_eq" || f == "bool_eq" || f == "bool2int") cmp = PCX_FT_EQ;
    else if (f == "int_ne" || f == "bool_ne") cmp = PCX_FT_NE;
    if (cmp >= 0 && n == 2) {
      term(e.args[swap ? 1 : 0]);
      term(e.args[swap ? 0 : 1]);
      node(cmp);
      return;
    }
    if ((f == "bool_equiv" || f == "nbool_equiv") && n == 2) {
      lit(e.args[0]);
      lit(e.args[1]);
      node(f == "bool_equiv" ? PCX_FT_EQUIV : PCX_FT_XOR);
      return;
    }
    int k;
    if (arith2(f, &k) && n == 3) {                 // a op b = r
      Expr t = e;
      t.args.pop_back();
      term(t);
      term(e.args[2]);
      node(PCX_FT_EQ);
      return;
    }
    if ((f == "int_neg" || f == "int_abs") && n == 2) {
      Expr t = e;
      t.args.pop_back();
      term(t);
      term(e.args[1]);
      node(PCX_FT_EQ);
      return;
    }
    if (f == "int_lin_le" || f == "int_lin_eq" || f == "int_lin_ne") {
      if (n != 3) { ok = false; return; }
      const std::vector<Expr> as = b.elements(e.args[0]), xs = b.elements(e.args[1]);
      if (as.size() != xs.size()) { ok = false; return; }
      if (as.empty()) node(PCX_FT_CONST, 0);
      for (size_t i = 0; i < as.size(); ++i) {
        term(as[i]);
        term(xs[i]);
        node(PCX_FT_MUL);
        if (i) node(PCX_FT_ADD);
      }
      term(e.args[2]);
      node(f == "int_lin_le" ? PCX_FT_LE : f == "int_lin_eq" ? PCX_FT_EQ : PCX_FT_NE);
      return;
    }
    if (f == "bool_clause" && n == 2) {
      const size_t k2 = b.elements(e.args[0]).size() + b.elements(e.args[1]).size();
      lits(e.args[0], false);
      lits(e.args[1], true);
      node(PCX_FT_OR, (int32_t)k2);
      return;
    }
    if (f == "bool_not") {
      if (n == 1) { formula(e.args[0]); node(PCX_FT_NOT); return; }
      if (n == 2) { lit(e.args[0]); lit(e.args[1]); node(PCX_FT_XOR); return; }
    }
    if ((f == "bool_and" || f == "bool_or" || f == "bool_xor") && (n == 2 || n == 3)) {
      lit(e.args[0]);
      lit(e.args[1]);
      if (f == "bool_xor") node(PCX_FT_XOR);
      else node(f == "bool_and" ? PCX_FT_AND : PCX_FT_OR, 2);
      if (n == 3) { lit(e.args[2]); node(PCX_FT_EQUIV); }
      return;
    }
    if ((f == "array_bool_and" || f == "array_bool_or") && (n == 1 || n == 2)) {
      const size_t k2 = b.elements(e.args[0]).size();
      lits(e.args[0], false);
      node(f == "array_bool_and" ? PCX_FT_AND : PCX_FT_OR, (int32_t)k2);
      if (n == 2) { lit(e.args[1]); node(PCX_FT_EQUIV); }
      return;
    }
    ok = false;
  }
};

// A declared domain with holes ({1, 3}) is a propagator of PC's (the
// disjunction of the equalities); a range is the store's own domain.
void record_domain_tree(Builder& b, int32_t v, const Expr& set) {
  bool rng;
  long long lo = 0, hi = 0;
  const std::vector<long long> vals = b.set_values(set, &rng, &lo, &hi);
  if (rng || vals.empty()) return;
  TreeBuilder t{b};
  for (long long c : vals) {
    t.node(PCX_FT_VAR, v);
    t.node(PCX_FT_CONST, (int32_t)c);
    t.node(PCX_FT_EQ);
  }
  t.node(PCX_FT_OR, (int32_t)vals.size());
  b.m.tnodes.insert(b.m.tnodes.end(), t.out.begin(), t.out.end());
  b.m.toff.push_back((int64_t)(b.m.tnodes.size() / 4));
}

// Record constraint e's tree (props [p0, end) are what it was decomposed into).
void record_tree(Builder& b, const Expr& e, size_t p0) {
  pcx_model& m = b.m;
  TreeBuilder t{b};
  t.formula(e);
  if (!t.ok) {
    t.out.clear();
    const size_t p1 = m.op.size();
    for (size_t p = p0; p < p1; ++p) {
      t.node(PCX_FT_PROP, m.op[p], m.x[p], m.y[p]);
      t.node(PCX_FT_PROP_Z, m.z[p]);
    }
    t.node(PCX_FT_AND, (int32_t)(p1 - p0));
  }
  m.tnodes.insert(m.tnodes.end(), t.out.begin(), t.out.end());
  m.toff.push_back((int64_t)(m.tnodes.size() / 4));
}

// ---- parser -------------------------------------------------------------------
struct Parser {
  Lexer lx;
  Builder& b;
  Token buf[2];                 // lookahead
  int nbuf = 0;

  Parser(const char* text, Builder& bb) : lx(text), b(bb) {}

  const Token& peek(int k = 0) {
    while (nbuf <= k) buf[nbuf++] = lx.next();
    return buf[k];
  }
  Token take() {
    const Token t = peek();
    buf[0] = buf[1];
    --nbuf;
    return t;
  }
  void skip() { take(); }
  bool is(const char* s) { const Token& t = peek(); return t.kind == tIdent || t.kind == tPunct ? t.text == s : false; }
  [[noreturn]] void fail(const std::string& msg) {
    throw ParseError{PCX_EINVAL, "line " + std::to_string(peek().line) + ": " + msg};
  }
  void expect(const char* s) {
    if (!is(s)) fail(std::string("expected '") + s + "'");
    skip();
  }
  std::string_view ident() {
    if (peek().kind != tIdent) fail("expected a name");
    return take().text;
  }
  long long integer() {
    bool neg = false;
    if (is("-")) { neg = true; skip(); }
    if (peek().kind != tInt) fail("expected an integer");
    const long long v = take().value;
    return neg ? -v : v;
  }
  void skip_annotations() {
    while (is("::")) {
      skip();
      int depth = 0;
      while (peek().kind != tEnd) {
        if (is("(") || is("[")) ++depth;
        else if (is(")") || is("]")) { if (depth == 0) break; --depth; }
        else if (depth == 0 && (is(";") || is("::") || is("="))) break;
        skip();
      }
    }
  }

  Expr expr() {
    Expr e;
    e.line = peek().line;
    if (peek().kind == tInt || is("-")) {
      e.kind = eInt;
      e.v = integer();
      if (is("..")) { skip(); e.kind = eRange; e.hi = integer(); }
      return e;
    }
    if (is("[")) {
      skip();
      e.kind = eArray;
      if (!is("]")) {
        e.args.push_back(expr());
        while (is(",")) { skip(); e.args.push_back(expr()); }
      }
      expect("]");
      return e;
    }
    if (is("{")) {
      skip();
      e.kind = eSet;
      if (!is("}")) {
        e.args.push_back(expr());
        while (is(",")) { skip(); e.args.push_back(expr()); }
      }
      expect("}");
      return e;
    }
    const std::string_view name = ident();
    if (name == "true" || name == "false") { e.kind = eBool; e.v = name == "true"; return e; }
    e.name = name;
    if (is("(")) {
      skip();
      e.kind = eCall;
      if (!is(")")) {
        e.args.push_back(expr());
        while (is(",")) { skip(); e.args.push_back(expr()); }
      }
      expect(")");
      return e;
    }
    if (is("[")) {
      skip();
      e.kind = eIndex;
      e.args.push_back(expr());
      expect("]");
      return e;
    }
    e.kind = eIdent;
    return e;
  }

  // dom := int | bool | L..U | {v, ...}; returns the hull, *set the literal (holes)
  void domain(long long* lo, long long* hi, Expr* set) {
    set->kind = eInt;
    if (is("int")) { skip(); *lo = PCX_MINF; *hi = PCX_INF; return; }
    if (is("bool")) { skip(); *lo = 0; *hi = 1; return; }
    if (is("{")) {
      *set = expr();
      if (set->args.empty()) { *lo = 1; *hi = 0; return; }
      *lo = LLONG_MAX;
      *hi = LLONG_MIN;
      for (const Expr& v : set->args) {
        long long c;
        if (!b.const_of(v, &c)) fail("set domain: expected constants");
        *lo = std::min(*lo, c);
        *hi = std::max(*hi, c);
      }
      return;
    }
    *lo = integer();
    expect("..");
    *hi = integer();
  }

  void check_int(long long v) {
    if (v < INT32_MIN || v > INT32_MAX) fail("bound out of int32 range");
  }

  void var_decl() {                          // after 'var'
    long long lo, hi;
    Expr set;
    domain(&lo, &hi, &set);
    check_int(lo);
    check_int(hi);
    expect(":");
    const std::string_view name = ident();
    skip_annotations();
    if (b.m.scalars.count(name) || b.m.arrays.count(name) || b.m.par_ints.count(name))
      fail("name declared twice: " + std::string(name));
    const int32_t v = b.new_var(name, lo, hi);
    b.m.scalars[b.m.vars[v].name] = v;
    if (set.kind == eSet && !set.args.empty()) { b.post_set_in(v, set); record_domain_tree(b, v, set); }
    if (is("=")) {
      skip();
      Expr eq;
      eq.kind = eCall;
      eq.name = "int_eq";
      Expr lhs;
      lhs.kind = eIdent;
      lhs.name = name;
      eq.args = {lhs, expr()};
      b.post(eq);
    }
  }

  void array_decl() {                        // after 'array'
    expect("[");
    const long long lo = integer();
    expect("..");
    const long long hi = integer();
    expect("]");
    if (lo != 1 || hi < 0) fail("arrays are indexed 1..n");
    // every element of a var array becomes a variable: refuse a size the
    // store cannot index before creating any of them
    if (hi > (long long)PCX_MAX_VARS - (long long)b.m.vars.size())
      fail("array of " + std::to_string(hi) + " elements: more variables than a store indexes");
    if (!is("of")) fail("expected 'of'");
    skip();
    const bool is_var = is("var");
    if (is_var) skip();
    long long dlo, dhi;
    Expr set;
    domain(&dlo, &dhi, &set);
    expect(":");
    const std::string_view name = ident();
    skip_annotations();
    if (b.m.scalars.count(name) || b.m.arrays.count(name) || b.m.par_arrays.count(name))
      fail("name declared twice: " + std::string(name));
    Expr init;
    const bool has_init = is("=");
    if (has_init) { skip(); init = expr(); }
    if (!is_var) {
      if (!has_init || init.kind != eArray || (long long)init.args.size() != hi)
        fail("a parameter array needs " + std::to_string(hi) + " values");
      std::vector<long long> vals;
      for (const Expr& e : init.args) {
        long long c;
        if (!b.const_of(e, &c)) fail("parameter array: expected constants");
        vals.push_back(c);
      }
      b.m.par_arrays[b.m.own(name)] = vals;
      return;
    }
    std::vector<int32_t> elems;
    if (has_init) {
      if (init.kind != eArray || (long long)init.args.size() != hi)
        fail("array initializer needs " + std::to_string(hi) + " elements");
      for (const Expr& e : init.args) {
        const int32_t v = b.value(e);
        b.restrict_to(v, dlo, dhi);
        elems.push_back(v);
      }
    } else {
      check_int(dlo);
      check_int(dhi);
      for (long long k = 1; k <= hi; ++k) {
        const int32_t v = b.new_var(std::string(name) + "[" + std::to_string(k) + "]", dlo, dhi);
        if (set.kind == eSet && !set.args.empty()) { b.post_set_in(v, set); record_domain_tree(b, v, set); }
        elems.push_back(v);
      }
    }
    b.m.arrays[b.m.own(name)] = elems;
  }

  void par_decl() {                          // int: n = 3;  bool: b = true;
    const bool is_bool = is("bool");
    skip();
    expect(":");
    const std::string_view name = ident();
    skip_annotations();
    expect("=");
    const Expr e = expr();
    long long c;
    if (!b.const_of(e, &c)) fail("a parameter needs a constant value");
    if (is_bool && c != 0 && c != 1) fail("a boolean parameter must be true or false");
    b.m.par_ints[b.m.own(name)] = c;
  }

  void run() {
    while (peek().kind != tEnd) {
      b.line = peek().line;
      if (is(";")) { skip(); continue; }
      if (is("var")) { skip(); var_decl(); }
      else if (is("array")) { skip(); array_decl(); }
      else if ((is("int") || is("bool")) && peek(1).kind == tPunct && peek(1).text == ":") par_decl();
      else if (is("constraint")) {
        skip();
        const Expr e = expr();
        skip_annotations();
        const size_t p0 = b.m.op.size();
        b.post(e);
        record_tree(b, e, p0);
      } else if (is("solve") || is("output")) {
        while (peek().kind != tEnd && !is(";")) skip();
      } else if (is("predicate") || is("function")) {
        fail("predicate definitions are not supported");
      } else {
        fail("expected 'var', 'array', 'constraint' or 'solve'");
      }
      if (peek().kind != tEnd) expect(";");
    }
  }
};

pcx_status run_text(pcx_model* m, const char* text) {
  if (!m || !text) return pcx_internal::fail_with(PCX_EINVAL, "NULL argument");
  // Items are applied as they are parsed; on an error the model keeps the
  // items before it.
  try {
    Builder b{*m};
    Parser p(text, b);
    p.run();
  } catch (const ParseError& e) {
    return pcx_internal::fail_with(e.status, e.msg);
  } catch (const std::bad_alloc&) {
    return pcx_internal::fail_with(PCX_ENOMEM, "host allocation failed");
  }
  return PCX_OK;
}

}  // namespace

extern "C" {

pcx_status pcx_model_parse(const char* text, pcx_model** out) {
  if (!out) return pcx_internal::fail_with(PCX_EINVAL, "out is NULL");
  *out = nullptr;
  auto* m = new (std::nothrow) pcx_model();
  if (!m) return pcx_internal::fail_with(PCX_ENOMEM, "host allocation failed");
  const pcx_status st = run_text(m, text);
  if (st != PCX_OK) { delete m; return st; }
  *out = m;
  return PCX_OK;
}

pcx_status pcx_model_add(pcx_model* m, const char* text) { return run_text(m, text); }

void pcx_model_destroy(pcx_model* m) { delete m; }

int64_t pcx_model_num_vars(const pcx_model* m) { return m ? (int64_t)m->vars.size() : 0; }

int64_t pcx_model_num_props(const pcx_model* m) { return m ? (int64_t)m->op.size() : 0; }

pcx_status pcx_model_table(const pcx_model* m, int64_t from, int32_t* op, int32_t* x, int32_t* y,
                           int32_t* z) {
  if (!m) return pcx_internal::fail_with(PCX_EINVAL, "NULL model");
  const int64_t n = (int64_t)m->op.size();
  if (from < 0 || from > n) return pcx_internal::fail_with(PCX_EINVAL, "from out of range");
  if (from < n && (!op || !x || !y || !z)) return pcx_internal::fail_with(PCX_EINVAL, "NULL argument");
  std::copy(m->op.begin() + from, m->op.end(), op);
  std::copy(m->x.begin() + from, m->x.end(), x);
  std::copy(m->y.begin() + from, m->y.end(), y);
  std::copy(m->z.begin() + from, m->z.end(), z);
  return PCX_OK;
}

pcx_status pcx_model_store(const pcx_model* m, int32_t* lb, int32_t* ub) {
  if (!m || ((!lb || !ub) && !m->vars.empty())) return pcx_internal::fail_with(PCX_EINVAL, "NULL argument");
  for (size_t v = 0; v < m->vars.size(); ++v) { lb[v] = m->vars[v].lb; ub[v] = m->vars[v].ub; }
  return PCX_OK;
}

int32_t pcx_model_var(const pcx_model* m, const char* name) {
  if (!m || !name) return -1;
  const std::string s(name);
  auto it = m->scalars.find(s);
  if (it != m->scalars.end()) return it->second;
  const size_t lb = s.find('[');
  if (lb != std::string::npos && s.back() == ']') {
    auto a = m->arrays.find(s.substr(0, lb));
    if (a == m->arrays.end()) return -1;
    char* end = nullptr;
    const long k = std::strtol(s.c_str() + lb + 1, &end, 10);
    if (end != s.c_str() + s.size() - 1 || k < 1 || k > (long)a->second.size()) return -1;
    return a->second[k - 1];
  }
  for (size_t v = 0; v < m->vars.size(); ++v)
    if (m->vars[v].name == std::string_view(s)) return (int32_t)v;
  return -1;
}

int64_t pcx_model_num_formulas(const pcx_model* m) { return m ? (int64_t)m->toff.size() - 1 : 0; }

int64_t pcx_model_num_formula_nodes(const pcx_model* m) { return m ? (int64_t)m->tnodes.size() / 4 : 0; }

pcx_status pcx_model_formulas(const pcx_model* m, int32_t* nodes, int64_t* off) {
  if (!m || !nodes || !off) return pcx_internal::fail_with(PCX_EINVAL, "NULL argument");
  std::copy(m->tnodes.begin(), m->tnodes.end(), nodes);
  std::copy(m->toff.begin(), m->toff.end(), off);
  return PCX_OK;
}

const char* pcx_model_var_name(const pcx_model* m, int32_t v) {
  if (!m || v < 0 || v >= (int32_t)m->vars.size()) return nullptr;
  return m->vars[v].name.data();   // an owned std::string: NUL-terminated
}

}  // extern "C"
