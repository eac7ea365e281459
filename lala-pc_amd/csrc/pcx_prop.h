// pcx_prop.h -- device-side bounds propagators for `x = y <op> z`.
//
// Semantics follow lattice-land/lala-pc include/lala/pir.hpp:721-817 (deduce)
// and 417-438 (ask).  A propagator reads the three intervals of its operands
// and produces, for each operand, an interval to meet into the store; an
// operand it does not narrow gets [-oo, +oo].  Within one evaluation the
// updates are sequential: a later line reads the bounds produced by earlier
// lines (pir.hpp:396-401 macros read the live r1/r2/r3), and that order is
// kept here.
//
// Arithmetic on bounds is done in 64 bits and saturated into the int32 range
// when it becomes a bound; the reference uses `int` and overflows for large
// products.  Where the reference does not overflow the results are equal.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcx {

constexpr int32_t kInf = 2147483647;
constexpr int32_t kMinf = -2147483647 - 1;

enum Op : int32_t {
  kAdd = 0, kMul = 1, kTdiv = 2, kFdiv = 3, kCdiv = 4, kEdiv = 5,
  kMin = 6, kMax = 7, kEq = 8, kLeq = 9, kNumOps = 10
};

typedef long long i64;

struct Bnd { i64 l, u; };   // working interval (values always in int32 range)

__device__ __forceinline__ i64 clamp32(i64 v) {
  return v > kInf ? (i64)kInf : (v < kMinf ? (i64)kMinf : v);
}
__device__ __forceinline__ i64 mn(i64 a, i64 b) { return a < b ? a : b; }
__device__ __forceinline__ i64 mx(i64 a, i64 b) { return a > b ? a : b; }
__device__ __forceinline__ Bnd bnd(i64 l, i64 u) { return Bnd{clamp32(l), clamp32(u)}; }
__device__ __forceinline__ Bnd top() { return Bnd{kMinf, kInf}; }
__device__ __forceinline__ Bnd empty() { return Bnd{kInf, kMinf}; }
__device__ __forceinline__ bool is_empty(Bnd a) { return a.l > a.u; }

// Integer divisions (battery::tdiv/fdiv/cdiv/ediv semantics). Operands fit in
// int32 except where a bound was shifted by one; 64-bit keeps those exact.
__device__ __forceinline__ i64 q_trunc(i64 a, i64 b) { return a / b; }
__device__ __forceinline__ i64 q_floor(i64 a, i64 b) {
  i64 q = a / b;
  return (q * b != a && ((a < 0) != (b < 0))) ? q - 1 : q;
}
__device__ __forceinline__ i64 q_ceil(i64 a, i64 b) {
  i64 q = a / b;
  return (q * b != a && ((a < 0) == (b < 0))) ? q + 1 : q;
}
__device__ __forceinline__ i64 q_euclid(i64 a, i64 b) {
  i64 q = a / b, r = a - q * b;
  return r < 0 ? (b > 0 ? q - 1 : q + 1) : q;
}
__device__ __forceinline__ i64 q_op(i64 a, int op, i64 b) {
  switch (op) {
    case kTdiv: return q_trunc(a, b);
    case kFdiv: return q_floor(a, b);
    case kCdiv: return q_ceil(a, b);
    default:    return q_euclid(a, b);
  }
}

// Division fast path for the hot MUL inverse: both operands are int32 and the
// divisor is non-zero, so 32-bit division is exact.
__device__ __forceinline__ int32_t d_floor32(int32_t a, int32_t b) {
  int32_t q = a / b;
  return (q * b != a && ((a ^ b) < 0)) ? q - 1 : q;
}
__device__ __forceinline__ int32_t d_ceil32(int32_t a, int32_t b) {
  int32_t q = a / b;
  return (q * b != a && ((a ^ b) >= 0)) ? q + 1 : q;
}

// Bot-aware join: an empty operand is the identity.
__device__ __forceinline__ Bnd hull(Bnd a, Bnd b) {
  if (is_empty(b)) return a;
  if (is_empty(a)) return b;
  return Bnd{mn(a.l, b.l), mx(a.u, b.u)};
}
__device__ __forceinline__ void narrow(Bnd& a, Bnd b) { a.l = mx(a.l, b.l); a.u = mn(a.u, b.u); }

// ---- division (pir.hpp:449-700), written without recursion ----------------
// x = y / d, forward.
__device__ __forceinline__ void div_forward(int op, Bnd& x, Bnd y, Bnd& d) {
  if (d.l < 0 && d.u > 0) {
    x.l = mx(x.l, mn(y.l, y.u == kMinf ? (i64)kInf : -y.u));
    x.u = mn(x.u, mx(y.l == kInf ? (i64)kMinf : -y.l, y.u));
    return;
  }
  if (d.l == 0) d.l = 1;
  if (d.u == 0) d.u = -1;
  if (y.l == kMinf || y.u == kInf || d.l == kMinf || d.u == kInf) return;
  if (is_empty(d)) return;
  i64 t1 = q_op(y.l, op, d.l), t2 = q_op(y.l, op, d.u), t3 = q_op(y.u, op, d.l), t4 = q_op(y.u, op, d.u);
  x.l = mx(x.l, mn(mn(t1, t2), mn(t3, t4)));
  x.u = mn(x.u, mx(mx(t1, t2), mx(t3, t4)));
}

// numerator bound for floor / ceil division
__device__ Bnd num_floor(Bnd x, Bnd d) {
  if (d.l < 0 && d.u > 0)
    return bnd(mn(mn(x.l, -x.u), mn(x.l * d.u, (x.u + 1) * d.l + 1)),
               mx(mx(-x.l, x.u), mx(x.l * d.l, (x.u + 1) * d.u - 1)));
  if (d.l > 0 || d.u < 0)
    return bnd(mn(mn(x.l * d.l, x.l * d.u), mn((x.u + 1) * d.l + 1, (x.u + 1) * d.u + 1)),
               mx(mx(x.l * d.l, x.l * d.u), mx((x.u + 1) * d.l - 1, (x.u + 1) * d.u - 1)));
  return top();
}
__device__ Bnd num_ceil(Bnd x, Bnd d) {
  if (d.l < 0 && d.u > 0)
    return bnd(mn(mn(x.l, -x.u), mn(x.u * d.l, (x.l - 1) * d.u + 1)),
               mx(mx(-x.l, x.u), mx(x.u * d.u, (x.l - 1) * d.l - 1)));
  if (d.l > 0 || d.u < 0)
    return bnd(mn(mn(x.u * d.l, x.u * d.u), mn((x.l - 1) * d.l + 1, (x.l - 1) * d.u + 1)),
               mx(mx(x.u * d.l, x.u * d.u), mx((x.l - 1) * d.l - 1, (x.l - 1) * d.u - 1)));
  return top();
}
__device__ Bnd num_trunc(Bnd x, Bnd d) {
  if (x.l > 0) return num_floor(x, d);
  if (x.u < 0) return num_ceil(x, d);
  if (x.l <= 0 && 0 <= x.u) {
    Bnd r = bnd(mn(d.l, -d.u) + 1, mx(-d.l, d.u) - 1);
    if (x.l != 0) r = hull(r, num_ceil(bnd(x.l, -1), d));
    if (x.u != 0) r = hull(r, num_floor(bnd(1, x.u), d));
    return r;
  }
  return top();
}
__device__ Bnd num_euclid(Bnd x, Bnd d) {
  if (d.l > 0) return num_floor(x, d);
  if (d.u < 0) return num_ceil(x, d);
  if (d.l < 0 && d.u > 0) return hull(num_ceil(x, bnd(d.l, -1)), num_floor(x, bnd(1, d.u)));
  return top();
}

// Denominator bound for floor division when x is one of the "leaf" shapes:
// x > 0 or x < -1 (with the numerator split by sign), x == 0, x == -1.
// Returns `ok=false` for x shapes that need splitting (handled by the caller).
__device__ Bnd den_floor_signed(Bnd x, Bnd y) {   // x > 0 || x.u + 1 < 0; y has no 0
  if (y.l > 0)
    return bnd(mn(q_floor(y.l, x.u + 1), q_floor(y.u, x.u + 1)) + 1, mx(q_floor(y.l, x.l), q_floor(y.u, x.l)));
  // y.u < 0
  return bnd(mn(q_ceil(y.l, x.l), q_ceil(y.u, x.l)), mx(q_ceil(y.l, x.u + 1), q_ceil(y.u, x.u + 1)) - 1);
}
__device__ Bnd den_floor_leaf(Bnd x, Bnd y) {
  if (x.l > 0 || x.u + 1 < 0) {
    if (y.l > 0 || y.u < 0) return den_floor_signed(x, y);
    if (0 == y.l && y.l < y.u) return den_floor_signed(x, bnd(1, y.u));
    if (y.l < y.u && y.u == 0) return den_floor_signed(x, bnd(y.l, -1));
    if (y.l < 0 && 0 < y.u) return hull(den_floor_signed(x, bnd(y.l, -1)), den_floor_signed(x, bnd(1, y.u)));
    if (y.l == 0 && y.u == 0) return empty();
    return top();
  }
  if (x.l == 0 && x.u == 0) {
    if (y.l > 0) return bnd(y.l + 1, kInf);
    if (y.u < 0) return bnd(kMinf, y.u - 1);
    return top();
  }
  if (x.l == -1 && x.u == -1) {
    if (y.l > 0) return bnd(kMinf, -y.l);
    if (y.u < 0) return bnd(-y.u, kInf);
    if (0 == y.l && y.l < y.u) return bnd(kMinf, -1);
    if (y.l < y.u && y.u == 0) return bnd(1, kInf);
    if (y.l == 0 && y.u == 0) return empty();
    return top();
  }
  return top();
}
__device__ Bnd den_floor(Bnd x, Bnd y) {
  if ((x.l > 0 || x.u + 1 < 0) || (x.l == 0 && x.u == 0) || (x.l == -1 && x.u == -1))
    return den_floor_leaf(x, y);
  if (x.l == 0 && 0 < x.u) return hull(den_floor_leaf(bnd(0, 0), y), den_floor_leaf(bnd(1, x.u), y));
  if (x.l < -1 && x.u == -1) return hull(den_floor_leaf(bnd(x.l, -2), y), den_floor_leaf(bnd(-1, -1), y));
  if (x.l <= -1 && x.u >= 0) {
    Bnd r = den_floor_leaf(bnd(-1, -1), y);
    r = hull(r, den_floor_leaf(bnd(0, 0), y));
    if (x.l != -1) r = hull(r, den_floor_leaf(bnd(x.l, -2), y));
    if (x.u != 0) r = hull(r, den_floor_leaf(bnd(1, x.u), y));
    return r;
  }
  return top();
}

__device__ Bnd den_ceil_signed(Bnd x, Bnd y) {    // x.l - 1 > 0 || x.u < 0; y has no 0
  if (y.l > 0)
    return bnd(mn(q_ceil(y.l, x.u), q_ceil(y.u, x.u)), mx(q_ceil(y.l, x.l - 1), q_ceil(y.u, x.l - 1)) - 1);
  return bnd(mn(q_floor(y.l, x.l - 1), q_floor(y.u, x.l - 1)) + 1, mx(q_floor(y.l, x.u), q_floor(y.u, x.u)));
}
__device__ Bnd den_ceil_leaf(Bnd x, Bnd y) {
  if (x.l - 1 > 0 || x.u < 0) {
    if (y.l > 0 || y.u < 0) return den_ceil_signed(x, y);
    if (0 == y.l && y.l < y.u) return den_ceil_signed(x, bnd(1, y.u));
    if (y.l < y.u && y.u == 0) return den_ceil_signed(x, bnd(y.l, -1));
    if (y.l < 0 && 0 < y.u) return hull(den_ceil_signed(x, bnd(y.l, -1)), den_ceil_signed(x, bnd(1, y.u)));
    if (y.l == 0 && y.u == 0) return empty();
    return top();
  }
  if (x.l == 0 && x.u == 0) {
    if (y.l > 0) return bnd(kMinf, -y.l - 1);
    if (y.u < 0) return bnd(-y.u + 1, kInf);
    return top();
  }
  if (x.l == 1 && x.u == 1) {
    if (y.l > 0) return bnd(y.l, kInf);
    if (y.u < 0) return bnd(kMinf, y.u);
    if (0 == y.l && y.l < y.u) return bnd(1, kInf);
    if (y.l < y.u && y.u == 0) return bnd(kMinf, -1);
    if (y.l == 0 && y.u == 0) return empty();
    return top();
  }
  return top();
}
__device__ Bnd den_ceil(Bnd x, Bnd y) {
  if ((x.l - 1 > 0 || x.u < 0) || (x.l == 0 && x.u == 0) || (x.l == 1 && x.u == 1))
    return den_ceil_leaf(x, y);
  if (x.l < 0 && x.u == 0) return hull(den_ceil_leaf(bnd(x.l, -1), y), den_ceil_leaf(bnd(0, 0), y));
  if (x.l == 1 && 1 < x.u) return hull(den_ceil_leaf(bnd(1, 1), y), den_ceil_leaf(bnd(2, x.u), y));
  if (x.l <= 0 && x.u >= 1) {
    Bnd r = den_ceil_leaf(bnd(1, 1), y);
    r = hull(r, den_ceil_leaf(bnd(0, 0), y));
    if (x.l != 0) r = hull(r, den_ceil_leaf(bnd(x.l, -1), y));
    if (x.u != 1) r = hull(r, den_ceil_leaf(bnd(2, x.u), y));
    return r;
  }
  return top();
}

__device__ Bnd den_trunc_zero(Bnd y, Bnd d) {     // x == [0,0]
  if (y.l > 0 && d.l > 0) return bnd(y.l + 1, kInf);
  if (y.l > 0 && d.u < 0) return bnd(kMinf, -y.l - 1);
  if (y.u < 0 && d.l > 0) return bnd(-y.u + 1, kInf);
  if (y.u < 0 && d.u < 0) return bnd(kMinf, y.u - 1);
  return top();
}
__device__ Bnd den_trunc(Bnd x, Bnd y, Bnd d) {
  if (x.l > 0) return den_floor(x, y);
  if (x.u < 0) return den_ceil(x, y);
  if (x.l == 0 && x.u == 0) return den_trunc_zero(y, d);
  if (x.l <= 0 && 0 <= x.u) {
    Bnd r = den_trunc_zero(y, d);
    if (x.l != 0) r = hull(r, den_ceil(bnd(x.l, -1), y));
    if (x.u != 0) r = hull(r, den_floor(bnd(1, x.u), y));
    return r;
  }
  return top();
}
__device__ Bnd den_euclid(Bnd x, Bnd y, Bnd d) {
  if (d.l > 0) return den_floor(x, y);
  if (d.u < 0) return den_ceil(x, y);
  if (d.l < 0 && 0 < d.u) return hull(den_floor(x, y), den_ceil(x, y));
  return top();
}

__device__ __noinline__ void prop_div(int op, Bnd& x, Bnd& y, Bnd& d) {
  div_forward(op, x, y, d);
  x = bnd(x.l, x.u);
  if (is_empty(x) || is_empty(d)) return;
  switch (op) {
    case kFdiv: narrow(y, num_floor(x, d)); break;
    case kCdiv: narrow(y, num_ceil(x, d)); break;
    case kTdiv: narrow(y, num_trunc(x, d)); break;
    default:    narrow(y, num_euclid(x, d)); break;
  }
  if (is_empty(y)) return;
  switch (op) {
    case kFdiv: narrow(d, den_floor(x, y)); break;
    case kCdiv: narrow(d, den_ceil(x, y)); break;
    case kTdiv: narrow(d, den_trunc(x, y, d)); break;
    default:    narrow(d, den_euclid(x, y, d)); break;
  }
}

// ---- 32-bit helpers for the hot propagators ------------------------------
// Saturating int32 add/sub: exactly clamp32 of the exact sum (one VALU op
// with the clamp bit on CDNA).
__device__ __forceinline__ int32_t sadd(int32_t a, int32_t b) { return __builtin_elementwise_add_sat(a, b); }
__device__ __forceinline__ int32_t ssub(int32_t a, int32_t b) { return __builtin_elementwise_sub_sat(a, b); }
__device__ __forceinline__ int32_t sat32(i64 v) { return (int32_t)clamp32(v); }

// x = y * z, forward (pir.hpp:768-775) on int32 bounds.
__device__ __forceinline__ void mul_forward32(int32_t& xl, int32_t& xu, int32_t yl, int32_t yu,
                                              int32_t zl, int32_t zu) {
  if (yl == kMinf || yu == kInf || zl == kMinf || zu == kInf) return;
  i64 lo, hi;
  if (min(yl, yu) >= 0 && min(zl, zu) >= 0) {
    // all four products are non-negative and monotone in both factors
    lo = (i64)min(yl, yu) * (i64)min(zl, zu);
    hi = (i64)max(yl, yu) * (i64)max(zl, zu);
  } else {
    const i64 a = (i64)yl * zl, b = (i64)yl * zu, c = (i64)yu * zl, d = (i64)yu * zu;
    lo = mn(mn(a, b), mn(c, d));
    hi = mx(mx(a, b), mx(c, d));
  }
  xl = max(xl, sat32(lo));
  xu = min(xu, sat32(hi));
}

// Narrow t from x = t * d (pir.hpp:702-718) on int32 bounds; d may lose a
// zero bound.
__device__ __forceinline__ void mul_inverse32(int32_t xl, int32_t xu, int32_t& tl, int32_t& tu,
                                              int32_t& dl, int32_t& du) {
  const bool x_nonzero = xl > 0 || xu < 0;
  if (x_nonzero) {
    if (dl == 0) dl = 1;
    if (du == 0) du = -1;
  }
  if (x_nonzero && dl < 0 && du > 0) {
    tl = max(tl, min(xl, xu == kMinf ? kInf : -xu));
    tu = min(tu, max(xl == kInf ? kMinf : -xl, xu));
    return;
  }
  if (!(x_nonzero || dl > 0 || du < 0)) return;
  if (xl == kMinf || xu == kInf || dl == kMinf || du == kInf) return;
  if (dl > du) return;
  if (min(xl, xu) >= 0 && dl > 0) {
    // Non-negative numerators, positive divisors: ceil(a/b) and floor(a/b)
    // are increasing in a and decreasing in b, so the extreme of the four
    // quotients is at (min a, max b) resp. (max a, min b).
    const uint32_t a_lo = (uint32_t)min(xl, xu), a_hi = (uint32_t)max(xl, xu);
    const uint32_t b_lo = (uint32_t)dl, b_hi = (uint32_t)du;
    const int32_t c = (int32_t)((a_lo + b_hi - 1u) / b_hi);
    const int32_t f = (int32_t)(a_hi / b_lo);
    tl = max(tl, c);
    tu = min(tu, f);
    return;
  }
  const int32_t c1 = d_ceil32(xl, dl), c2 = d_ceil32(xl, du), c3 = d_ceil32(xu, dl), c4 = d_ceil32(xu, du);
  const int32_t f1 = d_floor32(xl, dl), f2 = d_floor32(xl, du), f3 = d_floor32(xu, dl), f4 = d_floor32(xu, du);
  tl = max(tl, min(min(c1, c2), min(c3, c4)));
  tu = min(tu, max(max(f1, f2), max(f3, f4)));
}

// ---- one propagator ------------------------------------------------------
// In: loaded intervals X, Y, Z.  Out: the intervals to meet into x, y, z.
// kDiv = false compiles the division propagators out (they need many more
// registers); the engine uses that variant for tables without divisions.
template <bool kDiv = true>
__device__ __forceinline__ void deduce(int op, int2 X, int2 Y, int2 Z, int2& nx, int2& ny, int2& nz) {
  int32_t xl = X.x, xu = X.y, yl = Y.x, yu = Y.y, zl = Z.x, zu = Z.y;
  const int2 kTop = make_int2(kMinf, kInf);
  switch (op) {
    case kEq: {
      nx = kTop; ny = kTop; nz = kTop;
      if (xl == 1 && xu == 1) {                    // x true: y = z
        ny = Z; nz = Y;
      } else if (xl == 0 && xu == 0 && (yl == yu || zl == zu)) {
        // x false and one side fixed: shave that value off the other side
        const int2 cut = make_int2(yl == zl ? sadd(yl, 1) : kMinf, yu == zu ? ssub(yu, 1) : kInf);
        if (zl == zu) ny = cut; else nz = cut;
      } else if (yu == zl && yl == zu) {           // same singleton
        nx = make_int2(1, 1);
      } else if (yl > zu || yu < zl) {             // disjoint
        nx = make_int2(0, 0);
      }
      return;
    }
    case kLeq: {
      nx = kTop; ny = kTop; nz = kTop;
      if (xl == 1 && xu == 1) {                    // y <= z
        ny = make_int2(yl, zu); nz = make_int2(yl, zu);
      } else if (xl == 0 && xu == 0) {             // y > z
        ny = make_int2(sadd(zl, 1), yu);
        nz = make_int2(zl, ssub(yu, 1));
      } else if (yu <= zl) {
        nx = make_int2(1, 1);
      } else if (yl > zu) {
        nx = make_int2(0, 0);
      }
      return;
    }
    case kAdd: {
      // A sum or difference involving an infinite bound is that bound: the
      // projection then leaves its target alone.  Each test "a is not -inf
      // and b is not +inf" is min(a, ~b) != kMinf (~kInf == kMinf), one
      // comparison instead of two.
      if (min(yl, zl) != kMinf)  xl = max(xl, sadd(yl, zl));
      if (max(yu, zu) != kInf)   xu = min(xu, sadd(yu, zu));
      if (min(xl, ~zu) != kMinf) yl = max(yl, ssub(xl, zu));
      if (max(xu, ~zl) != kInf)  yu = min(yu, ssub(xu, zl));
      if (min(xl, ~yu) != kMinf) zl = max(zl, ssub(xl, yu));
      if (max(xu, ~yl) != kInf)  zu = min(zu, ssub(xu, yl));
      break;
    }
    case kMul: {
      mul_forward32(xl, xu, yl, yu, zl, zu);
      mul_inverse32(xl, xu, yl, yu, zl, zu);
      mul_inverse32(xl, xu, zl, zu, yl, yu);
      break;
    }
    case kMin: {
      xl = max(xl, min(yl, zl));
      xu = min(xu, min(yu, zu));
      yl = max(yl, xl);
      if (xu < zl) yu = min(yu, xu);
      zl = max(zl, xl);
      if (xu < yl) zu = min(zu, xu);
      break;
    }
    case kMax: {
      xl = max(xl, max(yl, zl));
      xu = min(xu, max(yu, zu));
      yu = min(yu, xu);
      if (xl > zu) yl = max(yl, xl);
      zu = min(zu, xu);
      if (xl > yu) zl = max(zl, xl);
      break;
    }
    case kTdiv: case kFdiv: case kCdiv: case kEdiv:
      if constexpr (kDiv) {
        Bnd x{xl, xu}, y{yl, yu}, z{zl, zu};
        prop_div(op, x, y, z);
        xl = sat32(x.l); xu = sat32(x.u); yl = sat32(y.l); yu = sat32(y.u);
        zl = sat32(z.l); zu = sat32(z.u);
      }
      break;
    default:
      nx = kTop; ny = kTop; nz = kTop;
      return;
  }
  nx = make_int2(xl, xu);
  ny = make_int2(yl, yu);
  nz = make_int2(zl, zu);
}

// ---- bitset domains -------------------------------------------------------
// One 64-bit word per variable, bit v <=> value v in [0,63]; empty word =
// failed.  Propagator semantics follow the reference's formula propagators
// over NBitset (include/lala/formula.hpp, terms.hpp): see the opcode list.
enum BitsOp : int32_t { kBLeq = 0, kBAdd = 1, kBNeq = 2, kBEqr = 3, kBitsOps = 4 };

typedef unsigned long long u64;

__device__ __forceinline__ bool word_single(u64 a) { return a && !(a & (a - 1)); }
// values <= v / values >= v
__device__ __forceinline__ u64 word_upto(int v) { return v >= 63 ? ~0ull : (v < 0 ? 0ull : ((2ull << v) - 1ull)); }
__device__ __forceinline__ u64 word_from(int v) { return v <= 0 ? ~0ull : (v > 63 ? 0ull : (~0ull << v)); }

// Membership tests for the set arithmetic of ADD over [0,63]:
// add_residual(X, Z) = {v : exists z in Z, v + z in X} (= X - Z within [0,63]),
// bit v set iff (X >> v) & Z != 0;
// add_support(Y, Z) = {x : exists y in Y, x - y in Z} (= Y + Z within [0,63]),
// bit x set iff (bitreverse(Z) >> (63 - x)) & Y != 0 (bit y of that word is
// bit 63 - (x - y) of bitreverse(Z), i.e. x - y in Z, for y <= x).
// One test per candidate value instead of a shifted OR per member.
__device__ __forceinline__ u64 add_residual(u64 X, u64 Z) {
  uint32_t lo = 0, hi = 0;
#pragma unroll 8
  for (int v = 0; v < 32; ++v) lo |= ((X >> v) & Z) ? (1u << v) : 0u;
#pragma unroll 8
  for (int v = 32; v < 64; ++v) hi |= ((X >> v) & Z) ? (1u << (v - 32)) : 0u;
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 add_support(u64 Y, u64 Z) {
  const u64 rz = __builtin_bitreverse64(Z);
  uint32_t lo = 0, hi = 0;
#pragma unroll 8
  for (int x = 0; x < 32; ++x) lo |= ((rz >> (63 - x)) & Y) ? (1u << x) : 0u;
#pragma unroll 8
  for (int x = 32; x < 64; ++x) hi |= ((rz >> (63 - x)) & Y) ? (1u << (x - 32)) : 0u;
  return ((u64)hi << 32) | lo;
}

// In: words X, Y, Z.  Out: masks to AND into x, y, z (all-ones = untouched).
__device__ __forceinline__ void deduce_bits(int op, u64 X, u64 Y, u64 Z, u64& nx, u64& ny, u64& nz) {
  nx = ~0ull; ny = ~0ull; nz = ~0ull;
  switch (op) {
    case kBLeq: {                         // y <= z: y <= max z, then z >= min y
      if (Z) ny = word_upto(63 - __builtin_clzll(Z));
      const u64 y2 = Y & ny;
      if (y2) nz = word_from(__builtin_ctzll(y2));
      return;
    }
    case kBAdd: {                         // x = y + z: y, z from residuals, then x
      const u64 y2 = Y & add_residual(X, Z);
      const u64 z2 = Z & add_residual(X, y2);
      ny = y2; nz = z2;
      nx = X & add_support(y2, z2);
      return;
    }
    case kBNeq:                           // y != z: shave a fixed side off the other
      if (word_single(Y)) nz = ~Y;
      else if (word_single(Z)) ny = ~Z;
      return;
    case kBEqr: {                         // x <=> (y = z), x a Boolean literal
      if (!(X & 1ull)) { nz = Y; ny = Y & Z; }               // x true
      else if (X == 1ull) {                                  // x false
        if (word_single(Y)) nz = ~Y;
        else if (word_single(Z)) ny = ~Z;
      } else if (Y == Z && word_single(Y)) nx = 2ull;        // entailed: x = 1
      else if (!(Y & Z)) nx = 1ull;                          // disentailed: x = 0
      return;
    }
    default: return;
  }
}

// Entailment of one propagator on its operands' intervals (pir.hpp:417-438).
__device__ __forceinline__ bool entailed(int op, int2 X, int2 Y, int2 Z) {
  const i64 xl = X.x, xu = X.y, yl = Y.x, yu = Y.y, zl = Z.x, zu = Z.y;
  switch (op) {
    case kEq:  return (xl == 1 && yu == zl && yl == zu) || (xu == 0 && (yu < zl || yl > zu));
    case kLeq: return (xl == 1 && yu <= zl) || (xu == 0 && yl > zu);
    case kAdd: return xl == xu && yl == yu && zl == zu && xl == yl + zl;
    case kMul: return xl == xu && ((yl == yu && zl == zu && xl == yl * zl)
                                   || (xl == 0 && ((yl == 0 && yu == 0) || (zl == 0 && zu == 0))));
    case kTdiv: case kFdiv: case kCdiv: case kEdiv:
      return (xl == xu && yl == yu && zl == zu && zl != 0 && xl == q_op(yl, op, zl))
          || (xl == yu && xu == yl && xl == 0 && (zl > 0 || zu < 0));
    case kMin: return (xl == yu && xu == yl && yu <= zl) || (xl == zu && xu == zl && zu <= yl);
    case kMax: return (xl == yu && xu == yl && yl >= zu) || (xl == zu && xu == zl && zl >= yu);
    default:   return false;
  }
}

}  // namespace pcx
