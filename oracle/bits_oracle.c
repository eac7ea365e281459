/*
 * bits_oracle.c -- CPU restatement of the reference's bitset propagation
 * (PC<VStore<NBitset<64>>>, tests/pc_bitset_test.cpp:23-25) for the four
 * ternary propagator kinds of the bitset workload.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg,
 * never by the product library.
 *
 * Domains are 64-bit words, bit v <=> value v in [0, 63]; a variable is empty
 * (failed) when its word is 0.  The reference's NBitset<64> holds the values
 * [0, 61] exactly and keeps one bucket bit for all negative values and one for
 * every value past 61 (pc_bitset_test.cpp:150-157: `var -15..5` is
 * NBit(-1, 5)).  This restatement uses plain words and drops results outside
 * [0, 63].  The two agree on stores inside [0, 61], and every generated
 * workload stays there (pcxg_bits_network / pcxg_bits_splits draw values below
 * PCXG_BITS_VALUES = 62 and propagation only removes values).
 * Inside that range LEQ, NEQ and EQR follow formula.hpp (lines cited below)
 * and are pinned by the pc_bitset_test.cpp cases and the brute-force checks.
 * ADD is "parity unpinned": its projections are NBitset::project(ADD / SUB)
 * of lala-core v1.2.8 (not in the reference tree), which no reference test
 * exercises; this restatement takes the exact set sum / difference.
 *
 * Propagators (restated from include/lala/formula.hpp and terms.hpp):
 *   LEQ  y <= z          Inequality<false>::deduce_impl  formula.hpp:796-805
 *        y <- y & [.., max z];  z <- z & [min y, ..]   (y already narrowed)
 *        has_changed is ASSIGNED by the second embed (formula.hpp:803), so the
 *        first embed's change is dropped: kept in "faithful" mode.
 *   ADD  x = y + z       Equality<false>::deduce_impl    formula.hpp:672-681
 *        with the residual embeds of Binary<GroupAdd>    terms.hpp:376-397
 *        y <- y & (x - z);  z <- z & (x - y);  x <- x & (y + z)
 *   NEQ  y != z          Disequality::deduce_impl         formula.hpp:638-662
 *        if y = {v}: z <- z \ {v} and return; else if z = {w}: y <- y \ {w}
 *   EQR  x <=> (y = z)   Biconditional::deduce            formula.hpp:421-427
 *        0 not in x: y = z (z <- z & y; y <- y & z);  x in {0}: NEQ on (y, z);
 *        else y = z = {v}: x <- x & {1};  else y & z empty: x <- x & {0}
 *        (x is a VariableLiteral: true iff 0 not in x, formula.hpp:98-110)
 *        Telling an EQR embeds x into {0,1}, as PIR's tell does for its
 *        reified outputs (pir.hpp:333-335): with other values in x the
 *        propagator is not monotone (x = {0,2} can fail where x = {2} does
 *        not), and the fixpoint would depend on the iteration order.
 * Table order: PC's tell sorts propagators by formula kind (pc.hpp:634-641);
 * the kinds above are, in variant order (formula.hpp:905-920), Leq(4) <
 * Eq(6) < Neq(7) < Bicond(10); the opcodes below are numbered in that order.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>

#ifndef PCXO_MUTANT
#define PCXO_MUTANT 0   /* tests/test_oracle_bits.py builds copies with planted slips */
#endif
#define MUT(k) (PCXO_MUTANT == (k))

enum { B_LEQ = 0, B_ADD = 1, B_NEQ = 2, B_EQR = 3 };

typedef uint64_t w64;

static w64 set_sum(w64 a, w64 b) {          /* {i + j} within [0,63] */
  w64 r = 0;
  for (int j = 0; j < 64; ++j) if (b >> j & 1) r |= a << j;
  return r;
}
static w64 set_diff(w64 a, w64 b) {         /* {i - j : i in a, j in b} within [0,63] */
  w64 r = 0;
  for (int j = 0; j < 64; ++j) if (b >> j & 1) r |= MUT(11) ? a << j : a >> j;
  return r;
}
static int wmin(w64 a) { return __builtin_ctzll(a); }
static int wmax(w64 a) { return 63 - __builtin_clzll(a); }
static w64 at_most(int v) { return v >= 63 ? ~0ULL : (v < 0 ? 0 : ((2ULL << v) - 1)); }
static w64 at_least(int v) { return v <= 0 ? ~0ULL : (v > 63 ? 0 : (~0ULL << v)); }
static int singleton(w64 a) { return a && !(a & (a - 1)); }

typedef struct { w64* w; int failed; } bstore;

static int embed(bstore* s, int32_t v, w64 d) {
  w64 n = s->w[v] & d;
  if (n == s->w[v]) return 0;
  s->w[v] = n;
  if (!n) s->failed = 1;
  return 1;
}

static int neq(bstore* s, int32_t y, int32_t z) {
  w64 l = s->w[y];
  if (MUT(13) && l) return embed(s, z, ~(l & (~l + 1)));
  if (singleton(l)) return embed(s, z, ~l);
  w64 r = s->w[z];
  if (singleton(r)) return embed(s, y, ~r);
  return 0;
}

/* one propagator; `faithful` reproduces the LEQ has_changed assignment */
static int deduce_bits(bstore* s, int op, int32_t x, int32_t y, int32_t z, int faithful) {
  switch (op) {
    case B_LEQ: {
      int c1 = 0, c2;
      if (s->w[z]) c1 = embed(s, y, at_most(MUT(12) ? wmin(s->w[z]) : wmax(s->w[z])));
      if (s->failed) return 1;
      c2 = embed(s, z, at_least(wmin(s->w[y])));
      return faithful ? c2 : (c1 | c2);
    }
    case B_ADD: {
      int ch = 0;
      w64 u = s->w[x];
      ch |= embed(s, y, set_diff(u, s->w[z]));
      ch |= embed(s, z, set_diff(u, s->w[y]));
      ch |= embed(s, x, set_sum(s->w[y], s->w[z]));
      return ch;
    }
    case B_NEQ:
      return neq(s, y, z);
    case B_EQR: {
      w64 wx = s->w[x], wy = s->w[y], wz = s->w[z];
      if (!(wx & 1) || (MUT(14) && wx == 1)) {   /* x true: y = z */
        int ch = embed(s, z, wy);
        ch |= embed(s, y, s->w[z]);
        return ch;
      }
      if (wx == 1) return neq(s, y, z);      /* x false: y != z */
      if (wy == wz && singleton(wy)) return embed(s, x, 2);   /* x <- {1} */
      if (!(wy & wz)) return embed(s, x, 1);                   /* x <- {0} */
      return 0;
    }
  }
  return 0;
}

typedef struct { int32_t has_changed, failed; int64_t sweeps, evals; } pcxo_result;

int pcxo_bits_deduce(int32_t op, int32_t x, int32_t y, int32_t z, uint64_t* w, int32_t faithful,
                     int32_t* failed) {
  bstore s = { w, 0 };
  int ch = deduce_bits(&s, op, x, y, z, faithful);
  if (failed) *failed = s.failed;
  return ch;
}

/* Reified outputs are Boolean: x <- x & {0,1} for every EQR told. */
int pcxo_bits_tell_embed(int64_t n, const int32_t* op, const int32_t* x, uint64_t* w) {
  int ch = 0;
  for (int64_t i = 0; i < n; ++i)
    if (op[i] == B_EQR && (w[x[i]] & ~3ULL)) { w[x[i]] &= 3ULL; ch = 1; }
  return ch;
}

/* Stable sort by opcode (= PC's kind order). */
int pcxo_bits_sort(int64_t n, int32_t* op, int32_t* x, int32_t* y, int32_t* z) {
  int64_t cnt[4] = {0, 0, 0, 0};
  for (int64_t i = 0; i < n; ++i) { if (op[i] < 0 || op[i] > 3) return -1; cnt[op[i]]++; }
  int64_t base[4] = {0, cnt[0], cnt[0] + cnt[1], cnt[0] + cnt[1] + cnt[2]};
  int32_t* t = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)(n ? n : 1));
  if (!t) return -1;
  for (int64_t i = 0; i < n; ++i) {
    int64_t j = base[op[i]]++;
    t[4 * j] = op[i]; t[4 * j + 1] = x[i]; t[4 * j + 2] = y[i]; t[4 * j + 3] = z[i];
  }
  for (int64_t i = 0; i < n; ++i) { op[i] = t[4 * i]; x[i] = t[4 * i + 1]; y[i] = t[4 * i + 2]; z[i] = t[4 * i + 3]; }
  free(t);
  return 0;
}

int pcxo_bits_fixpoint(int64_t n, const int32_t* op, const int32_t* x, const int32_t* y,
                       const int32_t* z, uint64_t* w, int64_t nvars, int32_t faithful,
                       int64_t max_sweeps, pcxo_result* out) {
  bstore s = { w, 0 };
  pcxo_result r = { 0, 0, 0, 0 };
  for (int64_t v = 0; v < nvars; ++v) if (!w[v]) { r.failed = 1; if (out) *out = r; return 0; }
  int changed = 1;
  while (changed && !s.failed && (max_sweeps <= 0 || r.sweeps < max_sweeps)) {
    changed = 0;
    r.sweeps++;
    for (int64_t i = 0; i < n; ++i) {
      changed |= deduce_bits(&s, op[i], x[i], y[i], z[i], faithful);
      r.evals++;
      if (s.failed) break;
    }
    r.has_changed |= changed;
  }
  r.failed = s.failed;
  if (out) *out = r;
  return 0;
}

/* ---- batch over host threads (CPU baseline) ---------------------------- */
typedef struct {
  int64_t n; const int32_t *op, *x, *y, *z;
  uint64_t* w; int64_t nvars, n_inst, max_sweeps; int32_t faithful;
  pcxo_result* res; int64_t next; pthread_mutex_t mu;
} bjob;

static void* bworker(void* p) {
  bjob* j = (bjob*)p;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    int64_t k = j->next++;
    pthread_mutex_unlock(&j->mu);
    if (k >= j->n_inst) break;
    pcxo_bits_fixpoint(j->n, j->op, j->x, j->y, j->z, j->w + k * j->nvars, j->nvars,
                       j->faithful, j->max_sweeps, &j->res[k]);
  }
  return NULL;
}

int pcxo_bits_fixpoint_batch(int64_t n, const int32_t* op, const int32_t* x, const int32_t* y,
                             const int32_t* z, uint64_t* w, int64_t nvars, int64_t n_inst,
                             int32_t faithful, int64_t max_sweeps, int32_t threads,
                             pcxo_result* res) {
  bjob j;
  j.n = n; j.op = op; j.x = x; j.y = y; j.z = z; j.w = w; j.nvars = nvars; j.n_inst = n_inst;
  j.max_sweeps = max_sweeps; j.faithful = faithful; j.res = res; j.next = 0;
  pthread_mutex_init(&j.mu, NULL);
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  if (!th) return -1;
  for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, bworker, &j);
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&j.mu);
  return 0;
}

/* ---- brute-force checks over small universes (values 0..U-1) ------------ */
#include <stdio.h>
#include <string.h>

static int bpred(int op, int a, int b, int c) {
  switch (op) {
    case B_LEQ: return b <= c;
    case B_ADD: return a == b + c;
    case B_NEQ: return b != c;
    default:    return (a == 0 || a == 1) && a == (b == c);
  }
}

/* For every triple of non-empty words over values [0,U): run one propagator
 * to its (true) fixpoint and compare with the set of solutions:
 *   sound  - every solution survives;
 *   GAC    - for add/leq/neq every surviving value has a support
 *            (the fixpoint equals the projections of the solutions);
 *   monotone - removing one value from one input never enlarges the result.
 * Returns the number of violations; the first is described in err. */
int pcxo_bits_check(int op, int U, int complete, char* err, int errlen) {
  int nf = 0;
  if (err && errlen) err[0] = 0;
  const w64 full = (U >= 64) ? ~0ULL : ((1ULL << U) - 1);
  const int xdom = (op == B_LEQ || op == B_NEQ) ? 1 : 0;   /* x unused: one fixed word */
  for (w64 a = 1; a <= full; ++a) {
    if (xdom && a != full) continue;
    if (op == B_EQR && a > 3) continue;     /* reified output told into {0,1} */
    for (w64 b = 1; b <= full; ++b)
      for (w64 c = 1; c <= full; ++c) {
        w64 w[3] = { a, b, c };
        int32_t o = op, x = 0, y = 1, z = 2;
        pcxo_result r;
        pcxo_bits_fixpoint(1, &o, &x, &y, &z, w, 3, 0, 0, &r);
        w64 pa = 0, pb = 0, pc = 0;
        for (int i = 0; i < U; ++i) if (a >> i & 1)
          for (int j = 0; j < U; ++j) if (b >> j & 1)
            for (int k = 0; k < U; ++k) if (c >> k & 1)
              if (bpred(op, i, j, k)) { pa |= 1ULL << i; pb |= 1ULL << j; pc |= 1ULL << k; }
        if (xdom) pa = a;
        int bad = 0;
        if (r.failed) bad = (pb != 0);
        else {
          if ((w[0] & pa) != pa || (w[1] & pb) != pb || (w[2] & pc) != pc) bad = 1;
          if (complete && (w[0] != pa || w[1] != pb || w[2] != pc) && pb) bad = 2;
          if (complete && !pb) bad = 3;
        }
        if (!bad) {
          /* monotone: drop one value of one input, result must not grow */
          for (int v = 0; v < 3 && !bad; ++v)
            for (int i = 0; i < U && !bad; ++i) {
              w64 s0[3] = { a, b, c };
              if (!(s0[v] >> i & 1) || s0[v] == (1ULL << i)) continue;
              if (xdom && v == 0) continue;
              s0[v] &= ~(1ULL << i);
              pcxo_result r2;
              pcxo_bits_fixpoint(1, &o, &x, &y, &z, s0, 3, 0, 0, &r2);
              if (r2.failed) continue;
              if (r.failed || (s0[0] & ~w[0]) || (s0[1] & ~w[1]) || (s0[2] & ~w[2])) bad = 4;
            }
        }
        if (bad && nf++ == 0 && err)
          snprintf(err, (size_t)errlen, "op %d kind %d: x=%llx y=%llx z=%llx -> %llx %llx %llx failed=%d",
                   op, bad, (unsigned long long)a, (unsigned long long)b, (unsigned long long)c,
                   (unsigned long long)w[0], (unsigned long long)w[1], (unsigned long long)w[2], r.failed);
      }
  }
  return nf;
}
