/*
 * pir_checks.c -- brute-force checks that pin the oracle (TEST INFRASTRUCTURE).
 *
 * Re-implements the methodology of the reference's bound-consistency harness
 * (tests/bound_consistency_test.hpp) against pir_oracle.c:
 *   - soundness over interval shapes: the fixpoint of one propagator over a
 *     box contains the hull of all integer solutions in the box, and equals it
 *     when the propagator is bound-complete          (bound_consistency_test.hpp:55-113)
 *   - completeness on singletons: a singleton box fails iff the predicate is
 *     false, and ask() holds iff it is true            (bound_consistency_test.hpp:115-153)
 *   - exhaustive box sweep                             (bound_consistency_test.hpp:155-225)
 *   - entailment soundness: when every ask() holds, every point of the box
 *     satisfies the predicate                          (bound_consistency_test.hpp:89-109)
 * plus one check the reference does not have:
 *   - monotonicity: for boxes a <= b (a inside b), f(a) <= f(b), for one
 *     application of deduce and for the whole fixpoint.  Uniqueness of the
 *     fixpoint under any iteration order (what the GPU engine relies on)
 *     follows from it.
 *
 * The predicates below are written from the mathematical definitions; the
 * divisions use floating-point floor/ceil on small integers, not the oracle's
 * integer division helpers, so a mistake in those helpers shows up here.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

typedef struct { int32_t has_changed, failed; int64_t sweeps, evals; } pcxo_result;
int pcxo_deduce(int32_t, int32_t, int32_t, int32_t, int32_t*, int32_t*, int32_t*);
int pcxo_ask(int32_t, int32_t, int32_t, int32_t, const int32_t*, const int32_t*);
int pcxo_tell_embed(int64_t, const int32_t*, const int32_t*, int32_t*, int32_t*);
int pcxo_fixpoint(int64_t, const int32_t*, const int32_t*, const int32_t*, const int32_t*,
                  const int64_t*, int64_t, int32_t*, int32_t*, int64_t, int64_t, pcxo_result*);

enum { O_ADD = 0, O_MUL = 1, O_TDIV = 2, O_FDIV = 3, O_CDIV = 4, O_EDIV = 5,
       O_MIN = 6, O_MAX = 7, O_EQ = 8, O_LEQ = 9 };

/* x = y <op> z, from the definitions (pir_test.cpp:123-136 lists the same). */
static int pred(int op, long x, long y, long z) {
  switch (op) {
    case O_EQ:  return (x == 0 || x == 1) && x == (y == z);
    case O_LEQ: return (x == 0 || x == 1) && x == (y <= z);
    case O_ADD: return x == y + z;
    case O_MUL: return x == y * z;
    case O_MIN: return x == (y < z ? y : z);
    case O_MAX: return x == (y > z ? y : z);
    case O_TDIV: return z != 0 && x == (long)trunc((double)y / (double)z);
    case O_FDIV: return z != 0 && x == (long)floor((double)y / (double)z);
    case O_CDIV: return z != 0 && x == (long)ceil((double)y / (double)z);
    case O_EDIV: {
      if (z == 0) return 0;
      /* Euclidean: y = z*q + r with 0 <= r < |z| */
      long q = (z > 0) ? (long)floor((double)y / (double)z) : (long)ceil((double)y / (double)z);
      return x == q;
    }
  }
  return 0;
}

/* Run one propagator (vars 0=x, 1=y, 2=z) from box b to its fixpoint. */
static void run_box(int op, const int32_t b[6], int32_t lb[3], int32_t ub[3], pcxo_result* r) {
  int32_t o = op, x = 0, y = 1, z = 2;
  lb[0] = b[0]; ub[0] = b[1]; lb[1] = b[2]; ub[1] = b[3]; lb[2] = b[4]; ub[2] = b[5];
  pcxo_tell_embed(1, &o, &x, lb, ub);
  pcxo_fixpoint(1, &o, &x, &y, &z, NULL, 0, lb, ub, 3, 0, r);
}

/* Hull of the solutions inside the box; returns 0 if there is none. */
static int hull(int op, const int32_t b[6], long h[6]) {
  int any = 0;
  for (long a = b[0]; a <= b[1]; ++a)
    for (long c = b[2]; c <= b[3]; ++c)
      for (long d = b[4]; d <= b[5]; ++d)
        if (pred(op, a, c, d)) {
          if (!any) { h[0] = h[1] = a; h[2] = h[3] = c; h[4] = h[5] = d; any = 1; }
          else {
            if (a < h[0]) h[0] = a; if (a > h[1]) h[1] = a;
            if (c < h[2]) h[2] = c; if (c > h[3]) h[3] = c;
            if (d < h[4]) h[4] = d; if (d > h[5]) h[5] = d;
          }
        }
  return any;
}

static int nfail; static char* ebuf; static int elen;
static void fail(const char* what, int op, const int32_t b[6], const int32_t* lb, const int32_t* ub) {
  if (nfail++ == 0 && ebuf)
    snprintf(ebuf, (size_t)elen, "%s op=%d box x[%d,%d] y[%d,%d] z[%d,%d] -> x[%d,%d] y[%d,%d] z[%d,%d]",
             what, op, b[0], b[1], b[2], b[3], b[4], b[5], lb[0], ub[0], lb[1], ub[1], lb[2], ub[2]);
}

/* Soundness (+ completeness if `complete`) and entailment soundness on one box. */
static void check_box(int op, const int32_t b[6], int complete) {
  int32_t lb[3], ub[3]; pcxo_result r; long h[6];
  for (int i = 0; i < 6; i += 2) if (b[i] > b[i + 1]) return;   /* empty declared box */
  run_box(op, b, lb, ub, &r);
  int has = hull(op, b, h);
  if (r.failed) { if (has) fail("unsound failure", op, b, lb, ub); return; }
  if (!has) {
    /* propagation may not detect every empty box, but ask() must not claim entailment */
    if (pcxo_ask(op, 0, 1, 2, lb, ub)) fail("entailed but no solution", op, b, lb, ub);
    if (complete) fail("incomplete: empty hull not detected", op, b, lb, ub);
    return;
  }
  for (int v = 0; v < 3; ++v) {
    if (lb[v] > h[2 * v] || ub[v] < h[2 * v + 1]) { fail("unsound bounds", op, b, lb, ub); return; }
    if (complete && (lb[v] != h[2 * v] || ub[v] != h[2 * v + 1])) { fail("not bound-complete", op, b, lb, ub); return; }
  }
  if (pcxo_ask(op, 0, 1, 2, lb, ub)) {
    for (long a = h[0]; a <= h[1]; ++a)
      for (long c = h[2]; c <= h[3]; ++c)
        for (long d = h[4]; d <= h[5]; ++d)
          if (!pred(op, a, c, d)) { fail("entailed but a point violates", op, b, lb, ub); return; }
  }
}

/* The 23 interval shapes of bound_consistency_test.hpp:59 (minval -20, maxval 20). */
static const int32_t SHAPES[23][2] = {
  {-20,-20},{-19,19},{-20,-3},{-20,-2},{-20,0},{-20,20},{-2,-1},{-1,0},{-2,2},{-1,-1},
  {0,0},{1,1},{2,2},{0,1},{1,2},{0,20},{2,20},{3,20},{20,20},{1,0},{2,10},{-25,25},{-2,3}};

int pcxo_check_shapes(int op, int complete, char* err, int errlen) {
  nfail = 0; ebuf = err; elen = errlen; if (err && errlen) err[0] = 0;
  for (int i = 0; i < 23; ++i)
    for (int j = 0; j < 23; ++j)
      for (int k = 0; k < 23; ++k) {
        int32_t b[6] = { SHAPES[i][0], SHAPES[i][1], SHAPES[j][0], SHAPES[j][1], SHAPES[k][0], SHAPES[k][1] };
        check_box(op, b, complete);
      }
  return nfail;
}

int pcxo_check_exhaustive(int op, int lo, int hi, int complete, char* err, int errlen) {
  nfail = 0; ebuf = err; elen = errlen; if (err && errlen) err[0] = 0;
  for (int xl = lo; xl <= hi; ++xl) for (int xu = xl; xu <= hi; ++xu)
  for (int yl = lo; yl <= hi; ++yl) for (int yu = yl; yu <= hi; ++yu)
  for (int zl = lo; zl <= hi; ++zl) for (int zu = zl; zu <= hi; ++zu) {
    int32_t b[6] = { xl, xu, yl, yu, zl, zu };
    check_box(op, b, complete);
  }
  return nfail;
}

/* Completeness on singletons over [lo, hi) (bound_consistency_test.hpp:115-153). */
int pcxo_check_singletons(int op, int lo, int hi, char* err, int errlen) {
  nfail = 0; ebuf = err; elen = errlen; if (err && errlen) err[0] = 0;
  for (int i = lo; i < hi; ++i) for (int j = lo; j < hi; ++j) for (int k = lo; k < hi; ++k) {
    int32_t b[6] = { i, i, j, j, k, k }, lb[3], ub[3]; pcxo_result r;
    int32_t o = op, x = 0;
    lb[0] = i; ub[0] = i; lb[1] = j; ub[1] = j; lb[2] = k; ub[2] = k;
    pcxo_tell_embed(1, &o, &x, lb, ub);
    int told_bot = lb[0] > ub[0];
    int entailed = told_bot ? 0 : pcxo_ask(op, 0, 1, 2, lb, ub);
    run_box(op, b, lb, ub, &r);
    if (!pred(op, i, j, k)) {
      if (entailed) fail("entailed on a non-solution", op, b, lb, ub);
      if (!r.failed) fail("non-solution singleton not refuted", op, b, lb, ub);
    } else {
      if (!entailed) fail("solution singleton not entailed", op, b, lb, ub);
      if (r.failed || lb[0] != i || ub[0] != i || lb[1] != j || ub[1] != j || lb[2] != k || ub[2] != k)
        fail("solution singleton changed", op, b, lb, ub);
    }
  }
  return nfail;
}

/* Monotonicity: shrinking any single bound of the box by one never yields a
 * larger result, for one deduce() and for the fixpoint.  Results that are
 * empty on the smaller box are vacuously below anything. */
static int leq_store(const int32_t* la, const int32_t* ua, const int32_t* lb, const int32_t* ub) {
  for (int v = 0; v < 3; ++v) if (la[v] < lb[v] || ua[v] > ub[v]) return 0;
  return 1;
}
static int any_bot(const int32_t* l, const int32_t* u) {
  for (int v = 0; v < 3; ++v) if (l[v] > u[v]) return 1;
  return 0;
}
int pcxo_check_monotone(int op, int lo, int hi, char* err, int errlen) {
  nfail = 0; ebuf = err; elen = errlen; if (err && errlen) err[0] = 0;
  int32_t x = 0, y = 1, z = 2, o = op;
  for (int xl = lo; xl <= hi; ++xl) for (int xu = xl; xu <= hi; ++xu)
  for (int yl = lo; yl <= hi; ++yl) for (int yu = yl; yu <= hi; ++yu)
  for (int zl = lo; zl <= hi; ++zl) for (int zu = zl; zu <= hi; ++zu) {
    int32_t big[6] = { xl, xu, yl, yu, zl, zu };
    for (int s = 0; s < 6; ++s) {
      int32_t small[6]; memcpy(small, big, sizeof small);
      if (s % 2 == 0) small[s] += 1; else small[s] -= 1;
      if (small[s & ~1] > small[s | 1]) continue;
      /* one application */
      int32_t la[3] = { small[0], small[2], small[4] }, ua[3] = { small[1], small[3], small[5] };
      int32_t lB[3] = { big[0], big[2], big[4] }, uB[3] = { big[1], big[3], big[5] };
      int32_t fa = 0, fb = 0;
      pcxo_tell_embed(1, &o, &x, la, ua);
      pcxo_tell_embed(1, &o, &x, lB, uB);
      if (!any_bot(la, ua)) pcxo_deduce(op, x, y, z, la, ua, &fa);
      if (!any_bot(lB, uB)) pcxo_deduce(op, x, y, z, lB, uB, &fb);
      if (!any_bot(la, ua)) {
        if (any_bot(lB, uB) || !leq_store(la, ua, lB, uB)) { fail("deduce not monotone", op, small, la, ua); continue; }
      }
      /* fixpoint */
      pcxo_result ra, rb;
      run_box(op, small, la, ua, &ra);
      run_box(op, big, lB, uB, &rb);
      if (!ra.failed) {
        if (rb.failed || !leq_store(la, ua, lB, uB)) fail("fixpoint not monotone", op, small, la, ua);
      }
    }
  }
  return nfail;
}

/* Multiplication by a non-zero constant is bound-complete at the fixpoint:
 * for x = y * c the solutions' hull is y in [ceil(xl/c), floor(xu/c)] (c > 0,
 * mirrored for c < 0) and x in c*y, and MUL's forward product plus mul_inv
 * reach exactly that.  The reference only asserts soundness for MUL
 * (pir_test.cpp:128); this sharper check pins the rounding of mul_inv. */
int pcxo_check_mul_const(int lo, int hi, char* err, int errlen) {
  nfail = 0; ebuf = err; elen = errlen; if (err && errlen) err[0] = 0;
  for (int xl = lo; xl <= hi; ++xl) for (int xu = xl; xu <= hi; ++xu)
  for (int yl = lo; yl <= hi; ++yl) for (int yu = yl; yu <= hi; ++yu)
  for (int c = lo; c <= hi; ++c) {
    if (c == 0) continue;
    int32_t b[6] = { xl, xu, yl, yu, c, c };
    check_box(O_MUL, b, 1);
    int32_t b2[6] = { xl, xu, c, c, yl, yu };
    check_box(O_MUL, b2, 1);
  }
  return nfail;
}
