/*
 * pir_oracle.c -- CPU restatement of the reference PIR interval propagators.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load it.  The
 * product library (lala-pc_amd/) never links, loads or calls it.
 *
 * What it restates (reference = lattice-land/lala-pc, include/lala/pir.hpp):
 *   - bytecode semantics `x = y <op> z`          pir.hpp:36-47
 *   - deduce(bytecode), statement by statement    pir.hpp:721-817
 *       EQ 730-744, LEQ 745-757, ADD 758-766, MUL 767-779 (+mul_inv 702-718),
 *       T/C/F/EDIV 780-792 (+itv_div 449-467, num_* 469-517, den_* 519-700),
 *       MIN 793-801, MAX 802-810, embeds 813-815
 *   - ask(bytecode)                               pir.hpp:417-438
 *   - tell: [0,1] embed of reified outputs        pir.hpp:333-335
 *           stable sort by (op, y, x, z)          pir.hpp:343-347
 *   - Gauss-Seidel fixpoint as called by the tests pir_test.cpp:60-62
 *     (sweep i = 0..n-1, repeat while some deduce(i) returned true).
 * External semantics restated from their call sites (lala-core v1.2.8 and
 * cuda-battery, not vendored in the reference):
 *   - VStore::embed = interval meet, returns true iff a bound strictly
 *     tightened; an interval is bot when lb > ub          (pir.hpp:813-815)
 *   - +/-infinity are INT32_MAX / INT32_MIN               (pir.hpp:403-404)
 *   - battery::tdiv/cdiv/fdiv/ediv = truncating, ceiling, floor and
 *     Euclidean division                                  (pir_test.cpp:129-136)
 *   - Interval join is bot-aware: joining with an empty interval returns the
 *     other operand (fjoin in den_fdiv / den_cdiv, pir.hpp:542, 561-571).
 *
 * One deliberate, documented divergence: the reference computes bound
 * arithmetic in `int` and overflows (undefined behaviour) for products
 * beyond 2^31.  Here every +, -, * on bounds is done in int64 and saturated
 * into [INT32_MIN, INT32_MAX] when stored back into an interval.  Wherever the
 * reference does not overflow the two agree bit for bit.
 *
 * Failure: after a fixpoint the store is failed iff some variable is empty.
 * The Gauss-Seidel loop stops as soon as a variable becomes empty (the
 * reference's test engine keeps sweeping on empty intervals; bound values
 * after failure are order-dependent and are never compared, only the flag).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef PCXO_MUTANT
#define PCXO_MUTANT 0   /* tests/test_oracle.py builds copies with planted slips */
#endif
#define MUT(k) (PCXO_MUTANT == (k))

#define INF  INT32_MAX
#define MINF INT32_MIN

/* Opcode numbering shared with include/pcx.h (checked by a test). */
enum { O_ADD = 0, O_MUL = 1, O_TDIV = 2, O_FDIV = 3, O_CDIV = 4, O_EDIV = 5,
       O_MIN = 6, O_MAX = 7, O_EQ = 8, O_LEQ = 9, O_NUM = 10 };

typedef struct { int64_t l, u; } itv;   /* local interval, int64 working copy */

static int32_t sat(int64_t v) {
  return v > INF ? INF : (v < MINF ? MINF : (int32_t)v);
}
static itv mk(int64_t l, int64_t u) { itv r; r.l = sat(l); r.u = sat(u); return r; }
static int64_t mn(int64_t a, int64_t b) { return a < b ? a : b; }
static int64_t mx(int64_t a, int64_t b) { return a > b ? a : b; }
static int is_bot(itv a) { return a.l > a.u; }
static itv top(void) { return mk(MINF, INF); }
static itv bot(void) { return mk(INF, MINF); }

/* battery division helpers, restated from their definitions. */
static int64_t tdiv(int64_t a, int64_t b) { return a / b; }
static int64_t fdiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
  return q;
}
static int64_t cdiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) == (b < 0))) q++;
  return q;
}
static int64_t ediv(int64_t a, int64_t b) {
  int64_t q = a / b, r = a % b;
  if (r < 0) q = (b > 0) ? q - 1 : q + 1;
  return q;
}
static int64_t divop(int64_t a, int op, int64_t b) {
  switch (op) {
    case O_TDIV: return tdiv(a, b);
    case O_CDIV: return cdiv(a, b);
    case O_FDIV: return fdiv(a, b);
    default:     return MUT(7) ? fdiv(a, b) : ediv(a, b);
  }
}

/* Interval meet (in place, local copy) and bot-aware joins. */
static void meet(itv* a, itv b) { a->l = mx(a->l, b.l); a->u = mn(a->u, b.u); }
static void join(itv* a, itv b) {
  if (is_bot(b)) return;
  if (is_bot(*a)) { *a = b; return; }
  a->l = mn(a->l, b.l); a->u = mx(a->u, b.u);
}
static itv fjoin(itv a, itv b) { join(&a, b); return a; }

/* ---- store ------------------------------------------------------------- */
typedef struct {
  int32_t* lb; int32_t* ub; int failed;
} store_t;

/* VStore::embed restated: meet, report strict tightening, flag emptiness. */
static int embed(store_t* s, int32_t v, itv d) {
  int ch = 0;
  int32_t l = sat(d.l), u = sat(d.u);
  if (l > s->lb[v]) { s->lb[v] = l; ch = 1; }
  if (u < s->ub[v]) { s->ub[v] = u; ch = 1; }
  if (ch && s->lb[v] > s->ub[v]) s->failed = 1;
  return ch;
}
static itv load(const store_t* s, int32_t v) { itv r; r.l = s->lb[v]; r.u = s->ub[v]; return r; }

/* ---- division sub-propagators (pir.hpp:449-700) ------------------------- */
/* r1 = r2 / r3 */
static void itv_div(int op, itv* r1, itv* r2, itv* r3) {
  if (r3->l < 0 && r3->u > 0) {
    r1->l = mx(r1->l, mn(r2->l, r2->u == MINF ? INF : -r2->u));
    r1->u = mn(r1->u, mx(r2->l == INF ? MINF : -r2->l, r2->u));
  } else {
    if (r3->l == 0) r3->l = 1;
    if (r3->u == 0) r3->u = -1;
    if (r2->l == MINF || r2->u == INF || r3->l == MINF || r3->u == INF) return;
    if (is_bot(*r3)) return;
    int64_t t1 = divop(r2->l, op, r3->l), t2 = divop(r2->l, op, r3->u);
    int64_t t3 = divop(r2->u, op, r3->l), t4 = divop(r2->u, op, r3->u);
    r1->l = mx(r1->l, mn(mn(t1, t2), mn(t3, t4)));
    r1->u = mn(r1->u, mx(mx(t1, t2), mx(t3, t4)));
  }
}

static itv num_fdiv(itv r1, itv r3) {
  int64_t xl = r1.l, xu = r1.u, zl = r3.l, zu = r3.u;
  if (zl < 0 && zu > 0)
    return mk(mn(mn(xl, -xu), mn(xl * zu, (xu + 1) * zl + 1)),
              mx(mx(-xl, xu), mx(xl * zl, (xu + 1) * zu - 1)));
  if (zl > 0 || zu < 0)
    return mk(mn(mn(xl * zl, xl * zu), mn((xu + 1) * zl + 1, (xu + 1) * zu + 1)),
              mx(mx(xl * zl, xl * zu), mx((xu + 1) * zl - 1, (xu + 1) * zu - 1)));
  return top();
}

static itv num_cdiv(itv r1, itv r3) {
  int64_t xl = r1.l, xu = r1.u, zl = r3.l, zu = r3.u;
  if (zl < 0 && zu > 0)
    return mk(mn(mn(xl, -xu), mn(xu * zl, (xl - 1) * zu + 1)),
              mx(mx(-xl, xu), mx(xu * zu, (xl - 1) * zl - 1)));
  if (zl > 0 || zu < 0)
    return mk(mn(mn(xu * zl, xu * zu), mn((xl - 1) * zl + 1, (xl - 1) * zu + 1)),
              mx(mx(xu * zl, xu * zu), mx((xl - 1) * zl - 1, (xl - 1) * zu - 1)));
  return top();
}

static itv num_tdiv(itv r1, itv r3) {
  int64_t xl = r1.l, xu = r1.u, zl = r3.l, zu = r3.u;
  if (xl > 0) return num_fdiv(r1, r3);
  if (xu < 0) return num_cdiv(r1, r3);
  if (xl <= 0 && 0 <= xu) {
    itv r = mk(mn(zl, -zu) + 1, mx(-zl, zu) - 1);
    if (xl != 0) join(&r, num_cdiv(mk(xl, -1), r3));
    if (xu != 0) join(&r, num_fdiv(mk(1, xu), r3));
    return r;
  }
  return top();
}

static itv num_ediv(itv r1, itv r3) {
  if (r3.l > 0) return num_fdiv(r1, r3);
  if (r3.u < 0) return num_cdiv(r1, r3);
  if (r3.l < 0 && r3.u > 0)
    return fjoin(num_cdiv(r1, mk(r3.l, -1)), num_fdiv(r1, mk(1, r3.u)));
  return top();
}

static itv den_fdiv(itv r1, itv r2) {
  int64_t xl = r1.l, xu = r1.u, yl = r2.l, yu = r2.u;
  if (xl > 0 || xu + 1 < 0) {
    if (yl > 0)
      return mk(mn(fdiv(yl, xu + 1), fdiv(yu, xu + 1)) + 1, mx(fdiv(yl, xl), fdiv(yu, xl)));
    else if (yu < 0)
      return mk(mn(cdiv(yl, xl), cdiv(yu, xl)), mx(cdiv(yl, xu + 1), cdiv(yu, xu + 1)) - 1);
    else if (0 == yl && yl < yu) return den_fdiv(r1, mk(1, yu));
    else if (yl < yu && yu == 0) return den_fdiv(r1, mk(yl, -1));
    else if (yl < 0 && 0 < yu) return fjoin(den_fdiv(r1, mk(yl, -1)), den_fdiv(r1, mk(1, yu)));
    else if (yl == 0 && yu == 0) return bot();
  } else if (xl == 0 && xu == 0) {
    if (yl > 0) return mk(yl + 1, INF);
    else if (yu < 0) return mk(MINF, yu - 1);
  } else if (xl == -1 && xu == -1) {
    if (yl > 0) return mk(MINF, -yl);
    else if (yu < 0) return mk(-yu, INF);
    else if (0 == yl && yl < yu) return mk(MINF, -1);
    else if (yl < yu && yu == 0) return mk(1, INF);
    else if (yl == 0 && yu == 0) return bot();
  } else if (xl == 0 && 0 < xu) {
    return fjoin(den_fdiv(mk(0, 0), r2), den_fdiv(mk(1, xu), r2));
  } else if (xl < -1 && xu == -1) {
    return fjoin(den_fdiv(mk(xl, -2), r2), den_fdiv(mk(-1, -1), r2));
  } else if (xl <= -1 && xu >= 0) {
    itv r = den_fdiv(mk(-1, -1), r2);
    join(&r, den_fdiv(mk(0, 0), r2));
    if (xl != -1) join(&r, den_fdiv(mk(xl, -2), r2));
    if (xu != 0) join(&r, den_fdiv(mk(1, xu), r2));
    return r;
  }
  return top();
}

static itv den_cdiv(itv r1, itv r2) {
  int64_t xl = r1.l, xu = r1.u, yl = r2.l, yu = r2.u;
  if (xl - 1 > 0 || xu < 0) {
    if (yl > 0)
      return mk(mn(cdiv(yl, xu), cdiv(yu, xu)), mx(cdiv(yl, xl - 1), cdiv(yu, xl - 1)) - 1);
    else if (yu < 0)
      return mk(mn(fdiv(yl, xl - 1), fdiv(yu, xl - 1)) + 1, mx(fdiv(yl, xu), fdiv(yu, xu)));
    else if (0 == yl && yl < yu) return den_cdiv(r1, mk(1, yu));
    else if (yl < yu && yu == 0) return den_cdiv(r1, mk(yl, -1));
    else if (yl < 0 && 0 < yu) return fjoin(den_cdiv(r1, mk(yl, -1)), den_cdiv(r1, mk(1, yu)));
    else if (yl == 0 && yu == 0) return bot();
  } else if (xl == 0 && xu == 0) {
    if (yl > 0) return mk(MINF, -yl - 1);
    else if (yu < 0) return mk(-yu + 1, INF);
  } else if (xl == 1 && xu == 1) {
    if (yl > 0) return mk(yl, INF);
    else if (yu < 0) return mk(MINF, yu);
    else if (0 == yl && yl < yu) return mk(1, INF);
    else if (yl < yu && yu == 0) return mk(MINF, -1);
    else if (yl == 0 && yu == 0) return bot();
  } else if (xl < 0 && xu == 0) {
    return fjoin(den_cdiv(mk(xl, -1), r2), den_cdiv(mk(0, 0), r2));
  } else if (xl == 1 && 1 < xu) {
    return fjoin(den_cdiv(mk(1, 1), r2), den_cdiv(mk(2, xu), r2));
  } else if (xl <= 0 && xu >= 1) {
    itv r = den_cdiv(mk(1, 1), r2);
    join(&r, den_cdiv(mk(0, 0), r2));
    if (xl != 0) join(&r, den_cdiv(mk(xl, -1), r2));
    if (xu != 1) join(&r, den_cdiv(mk(2, xu), r2));
    return r;
  }
  return top();
}

static itv den_tdiv(itv r1, itv r2, itv r3) {
  int64_t xl = r1.l, xu = r1.u, yl = r2.l, yu = r2.u, zl = r3.l, zu = r3.u;
  if (xl > 0) return den_fdiv(r1, r2);
  if (xu < 0) return den_cdiv(r1, r2);
  if (xl == 0 && xu == 0) {
    if (yl > 0 && zl > 0) return mk(yl + 1, INF);
    if (yl > 0 && zu < 0) return mk(MINF, -yl - 1);
    if (yu < 0 && zl > 0) return mk(-yu + 1, INF);
    if (yu < 0 && zu < 0) return mk(MINF, yu - 1);
  } else if (xl <= 0 && 0 <= xu) {
    itv r = den_tdiv(mk(0, 0), r2, r3);
    if (xl != 0) join(&r, den_cdiv(mk(xl, -1), r2));
    if (xu != 0) join(&r, den_fdiv(mk(1, xu), r2));
    return r;
  }
  return top();
}

static itv den_ediv(itv r1, itv r2, itv r3) {
  if (r3.l > 0) return den_fdiv(r1, r2);
  if (r3.u < 0) return den_cdiv(r1, r2);
  if (r3.l < 0 && 0 < r3.u) return fjoin(den_fdiv(r1, r2), den_cdiv(r1, r2));
  return top();
}

/* mul_inv(r1, r2, r3): narrow r2 from r1 = r2 * r3 (pir.hpp:702-718). */
static void mul_inv(itv r1, itv* r2, itv* r3) {
  int64_t xl = r1.l, xu = r1.u;
  if (xl > 0 || xu < 0) {
    if (r3->l == 0) r3->l = 1;
    if (r3->u == 0) r3->u = -1;
  }
  if ((xl > 0 || xu < 0) && r3->l < 0 && r3->u > 0) {
    r2->l = mx(r2->l, mn(xl, xu == MINF ? INF : -xu));
    r2->u = mn(r2->u, mx(xl == INF ? MINF : -xl, xu));
  } else if (xl > 0 || xu < 0 || r3->l > 0 || r3->u < 0) {
    if (xl == MINF || xu == INF || r3->l == MINF || r3->u == INF) return;
    if (is_bot(*r3)) return;
    int64_t zl = r3->l, zu = r3->u;
    if (MUT(2)) r2->l = mx(r2->l, mn(mn(fdiv(xl, zl), fdiv(xl, zu)), mn(fdiv(xu, zl), fdiv(xu, zu))));
    else r2->l = mx(r2->l, mn(mn(cdiv(xl, zl), cdiv(xl, zu)), mn(cdiv(xu, zl), cdiv(xu, zu))));
    r2->u = mn(r2->u, mx(mx(fdiv(xl, zl), fdiv(xl, zu)), mx(fdiv(xu, zl), fdiv(xu, zu))));
  }
}

/* Saturate a working interval into the int32 store range. */
static itv clip(itv a) { return mk(a.l, a.u); }

/* ---- deduce(bytecode) (pir.hpp:721-817) -------------------------------- */
static int deduce_one(store_t* s, int op, int32_t x, int32_t y, int32_t z) {
  int ch = 0;
  itv r1 = load(s, x), r2 = load(s, y), r3 = load(s, z);
  switch (op) {
    case O_EQ: {
      int one = (r1.l == 1 && r1.u == 1), zero = (r1.l == 0 && r1.u == 0);
      if (one) {
        ch |= embed(s, y, r3);
        ch |= embed(s, z, r2);
      } else if (zero && (r2.l == r2.u || r3.l == r3.u)) {
        ch |= embed(s, r3.l == r3.u ? y : z,
                    mk(r2.l == r3.l ? r2.l + (MUT(4) ? 0 : 1) : MINF, r2.u == r3.u ? r2.u - 1 : INF));
      } else if (r2.u == r3.l && r2.l == r3.u) {
        ch |= embed(s, x, mk(1, 1));
      } else if (r2.l > r3.u || r2.u < r3.l) {
        ch |= embed(s, x, mk(0, 0));
      }
      return ch;
    }
    case O_LEQ: {
      int one = (r1.l == 1 && r1.u == 1), zero = (r1.l == 0 && r1.u == 0);
      if (one) {
        ch |= embed(s, y, mk(r2.l, r3.u));
        ch |= embed(s, z, mk(r2.l, r3.u));
      } else if (zero) {
        ch |= embed(s, y, mk(r3.l + (MUT(5) ? 0 : 1), r2.u));
        ch |= embed(s, z, mk(r3.l, r2.u - 1));
      } else if (r2.u <= r3.l) {
        ch |= embed(s, x, mk(1, 1));
      } else if (r2.l > r3.u) {
        ch |= embed(s, x, mk(0, 0));
      }
      return ch;
    }
    case O_ADD: {
      /* Each line reads the values updated by the lines above it. */
      r1.l = (r2.l == MINF || r3.l == MINF) ? r1.l : mx(r1.l, r2.l + r3.l);
      r1.u = (r2.u == INF || r3.u == INF) ? r1.u : mn(r1.u, r2.u + r3.u);
      r1 = clip(r1);
      r2.l = (r1.l == MINF || r3.u == INF) ? r2.l : mx(r2.l, r1.l - (MUT(1) ? r3.l : r3.u));
      r2.u = (r1.u == INF || r3.l == MINF) ? r2.u : mn(r2.u, r1.u - r3.l);
      r2 = clip(r2);
      r3.l = (r1.l == MINF || r2.u == INF) ? r3.l : mx(r3.l, r1.l - r2.u);
      r3.u = (r1.u == INF || r2.l == MINF) ? r3.u : mn(r3.u, r1.u - r2.l);
      r3 = clip(r3);
      break;
    }
    case O_MUL: {
      if (r2.l != MINF && r2.u != INF && r3.l != MINF && r3.u != INF) {
        int64_t t1 = r2.l * r3.l, t2 = r2.l * r3.u, t3 = r2.u * r3.l, t4 = r2.u * r3.u;
        r1.l = mx(r1.l, mn(mn(t1, t2), mn(t3, t4)));
        r1.u = mn(r1.u, mx(mx(t1, t2), mx(t3, t4)));
        r1 = clip(r1);
      }
      mul_inv(r1, &r2, &r3);
      mul_inv(r1, &r3, &r2);
      break;
    }
    case O_TDIV: case O_CDIV: case O_FDIV: case O_EDIV: {
      itv_div(op, &r1, &r2, &r3);
      r1 = clip(r1);
      if (!is_bot(r1) && !is_bot(r3)) {
        switch (op) {
          case O_FDIV: meet(&r2, num_fdiv(r1, r3)); break;
          case O_CDIV: meet(&r2, num_cdiv(r1, r3)); break;
          case O_TDIV: meet(&r2, num_tdiv(r1, r3)); break;
          default:     meet(&r2, num_ediv(r1, r3)); break;
        }
        if (!is_bot(r2)) {
          switch (op) {
            case O_FDIV: meet(&r3, den_fdiv(r1, r2)); break;
            case O_CDIV: meet(&r3, den_cdiv(r1, r2)); break;
            case O_TDIV: meet(&r3, den_tdiv(r1, r2, r3)); break;
            default:     meet(&r3, den_ediv(r1, r2, r3)); break;
          }
        }
      }
      break;
    }
    case O_MIN: {
      r1.l = mx(r1.l, mn(r2.l, r3.l));
      r1.u = mn(r1.u, mn(r2.u, r3.u));
      r2.l = mx(r2.l, r1.l);
      if (MUT(3) ? r1.u <= r3.l : r1.u < r3.l) r2.u = mn(r2.u, r1.u);
      r3.l = mx(r3.l, r1.l);
      if (r1.u < r2.l) r3.u = mn(r3.u, r1.u);
      break;
    }
    case O_MAX: {
      r1.l = mx(r1.l, mx(r2.l, r3.l));
      r1.u = mn(r1.u, mx(r2.u, r3.u));
      if (!MUT(6)) r2.u = mn(r2.u, r1.u);
      if (r1.l > r3.u) r2.l = mx(r2.l, r1.l);
      r3.u = mn(r3.u, r1.u);
      if (r1.l > r2.u) r3.l = mx(r3.l, r1.l);
      break;
    }
    default: return 0;
  }
  ch |= embed(s, x, r1);
  ch |= embed(s, y, r2);
  ch |= embed(s, z, r3);
  return ch;
}

/* ---- ask(bytecode) (pir.hpp:417-438) ----------------------------------- */
static int ask_one(const store_t* s, int op, int32_t x, int32_t y, int32_t z) {
  itv r1 = load(s, x), r2 = load(s, y), r3 = load(s, z);
  int64_t xl = r1.l, xu = r1.u, yl = r2.l, yu = r2.u, zl = r3.l, zu = r3.u;
  switch (op) {
    case O_EQ:  return (xl == 1 && yu == zl && yl == zu) || (xu == 0 && (yu < zl || yl > zu));
    case O_LEQ: return (xl == 1 && yu <= zl) || (xu == 0 && yl > zu);
    case O_ADD: return xl == xu && yl == yu && zl == zu && xl == yl + zl;
    case O_MUL: return xl == xu && ((yl == yu && zl == zu && xl == yl * zl)
                                    || (xl == 0 && ((yl == 0 && yu == 0) || (zl == 0 && zu == 0))));
    case O_TDIV: case O_CDIV: case O_FDIV: case O_EDIV:
      return (xl == xu && yl == yu && zl == zu && zl != 0 && xl == divop(yl, op, zl))
          || (xl == yu && xu == yl && xl == 0 && (zl > 0 || zu < 0));
    case O_MIN: return (xl == yu && xu == yl && yu <= zl) || (xl == zu && xu == zl && zu <= yl);
    case O_MAX: return (xl == yu && xu == yl && yl >= zu) || (xl == zu && xu == zl && zl >= yu);
    default: return 0;
  }
}

/* ======================= exported oracle interface ======================= */

typedef struct {
  int32_t has_changed;
  int32_t failed;
  int64_t sweeps;
  int64_t evals;
} pcxo_result;

/* One propagator application; returns has_changed, sets *failed on bot. */
int pcxo_deduce(int32_t op, int32_t x, int32_t y, int32_t z,
                int32_t* lb, int32_t* ub, int32_t* failed) {
  store_t s = { lb, ub, 0 };
  int ch = deduce_one(&s, op, x, y, z);
  if (failed) *failed = s.failed;
  return ch;
}

int pcxo_ask(int32_t op, int32_t x, int32_t y, int32_t z,
             const int32_t* lb, const int32_t* ub) {
  store_t s = { (int32_t*)lb, (int32_t*)ub, 0 };
  return ask_one(&s, op, x, y, z);
}

/* Reified outputs are embedded into [0,1] when told (pir.hpp:333-335). */
int pcxo_tell_embed(int64_t n, const int32_t* op, const int32_t* x,
                    int32_t* lb, int32_t* ub) {
  store_t s = { lb, ub, 0 };
  int ch = 0;
  for (int64_t i = 0; i < n; ++i)
    if (!MUT(8) && (op[i] == O_EQ || op[i] == O_LEQ)) ch |= embed(&s, x[i], mk(0, 1));
  return ch;
}

/* Stable sort of the table by (op, y, x, z) (pir.hpp:343-347). */
typedef struct { int32_t op, x, y, z; int64_t pos; } bc_t;
static int bc_cmp(const void* pa, const void* pb) {
  const bc_t* a = (const bc_t*)pa; const bc_t* b = (const bc_t*)pb;
  if (a->op != b->op) return a->op < b->op ? -1 : 1;
  if (a->y != b->y) return a->y < b->y ? -1 : 1;
  if (a->x != b->x) return a->x < b->x ? -1 : 1;
  if (a->z != b->z) return a->z < b->z ? -1 : 1;
  return a->pos < b->pos ? -1 : (a->pos > b->pos);   /* stability */
}
int pcxo_sort_table(int64_t n, int32_t* op, int32_t* x, int32_t* y, int32_t* z) {
  bc_t* t = (bc_t*)malloc(sizeof(bc_t) * (size_t)(n ? n : 1));
  if (!t) return -1;
  for (int64_t i = 0; i < n; ++i) { t[i].op = op[i]; t[i].x = x[i]; t[i].y = y[i]; t[i].z = z[i]; t[i].pos = i; }
  qsort(t, (size_t)n, sizeof(bc_t), bc_cmp);
  for (int64_t i = 0; i < n; ++i) { op[i] = t[i].op; x[i] = t[i].x; y[i] = t[i].y; z[i] = t[i].z; }
  free(t);
  return 0;
}

static int store_is_bot(const int32_t* lb, const int32_t* ub, int64_t nvars) {
  for (int64_t v = 0; v < nvars; ++v) if (lb[v] > ub[v]) return 1;
  return 0;
}

/* Gauss-Seidel fixpoint over the table in the given order.
 * Table `idx` (may be NULL) selects a subset of propagators, in that order. */
int pcxo_fixpoint(int64_t n, const int32_t* op, const int32_t* x, const int32_t* y,
                  const int32_t* z, const int64_t* idx, int64_t nidx,
                  int32_t* lb, int32_t* ub, int64_t nvars, int64_t max_sweeps,
                  pcxo_result* out) {
  store_t s = { lb, ub, 0 };
  pcxo_result r = { 0, 0, 0, 0 };
  int64_t m = idx ? nidx : n;
  if (store_is_bot(lb, ub, nvars)) { r.failed = 1; if (out) *out = r; return 0; }
  int changed = 1;
  while (changed && !s.failed && (max_sweeps <= 0 || r.sweeps < max_sweeps)) {
    changed = 0;
    r.sweeps++;
    for (int64_t k = 0; k < m; ++k) {
      int64_t i = idx ? idx[k] : k;
      changed |= deduce_one(&s, op[i], x[i], y[i], z[i]);
      r.evals++;
      if (s.failed) break;
    }
    r.has_changed |= changed;
  }
  r.failed = s.failed;
  if (out) *out = r;
  return 0;
}

/* ---- batch of independent instances, spread over host threads ----------
 * Instance k owns the store slice lb/ub[k*nvars .. (k+1)*nvars) and runs the
 * Gauss-Seidel loop above over the propagators idx[off[k] .. off[k+1]) (all
 * propagators when idx == NULL).  This is the multi-core CPU baseline: the
 * reference algorithm is sequential per instance (pir_test.cpp:60-62), so
 * instances are the unit of host parallelism. */
#include <pthread.h>

typedef struct {
  int64_t n; const int32_t *op, *x, *y, *z;
  const int64_t* idx; const int64_t* off;
  int32_t *lb, *ub; int64_t nvars, n_inst, max_sweeps;
  pcxo_result* res;
  int64_t next;            /* shared work counter */
  pthread_mutex_t mu;
} batch_job;

static void* batch_worker(void* p) {
  batch_job* j = (batch_job*)p;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    int64_t k = j->next++;
    pthread_mutex_unlock(&j->mu);
    if (k >= j->n_inst) break;
    const int64_t* ix = j->idx ? j->idx + j->off[k] : NULL;
    int64_t nix = j->idx ? j->off[k + 1] - j->off[k] : 0;
    pcxo_fixpoint(j->n, j->op, j->x, j->y, j->z, ix, nix,
                  j->lb + k * j->nvars, j->ub + k * j->nvars, j->nvars,
                  j->max_sweeps, &j->res[k]);
  }
  return NULL;
}

int pcxo_fixpoint_batch(int64_t n, const int32_t* op, const int32_t* x, const int32_t* y,
                        const int32_t* z, const int64_t* idx, const int64_t* off,
                        int64_t n_inst, int32_t* lb, int32_t* ub, int64_t nvars,
                        int64_t max_sweeps, int32_t threads, pcxo_result* res) {
  batch_job j;
  j.n = n; j.op = op; j.x = x; j.y = y; j.z = z; j.idx = idx; j.off = off;
  j.lb = lb; j.ub = ub; j.nvars = nvars; j.n_inst = n_inst; j.max_sweeps = max_sweeps;
  j.res = res; j.next = 0;
  pthread_mutex_init(&j.mu, NULL);
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  if (!th) return -1;
  for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, batch_worker, &j);
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&j.mu);
  return 0;
}
