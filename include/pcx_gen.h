/*
 * pcx_gen.h -- seeded synthetic workloads for the fixpoint engine.
 *
 * No counterpart in the reference, which generates no instances: its tests
 * build models from FlatZinc (tests/pir_test.cpp).  These generators produce
 * the ternary networks, stores and search-node splits of the benchmark
 * workloads, bit-for-bit reproducibly from a 64-bit seed (splitmix64, Steele,
 * Lea & Flood 2014, as published by Vigna).  Host-only code; no GPU needed.
 */
#ifndef PCX_GEN_H
#define PCX_GEN_H

#include <stdint.h>
#include "pcx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Generator operation kinds (what the workloads draw uniformly from). */
#define PCXG_ADD  (1u << 0)   /* x = y + z                        -> ADD          */
#define PCXG_SUB  (1u << 1)   /* x = y - z, told as y = x + z     -> ADD          */
#define PCXG_MUL  (1u << 2)   /* x = y * z                        -> MUL          */
#define PCXG_MIN  (1u << 3)   /* x = min(y, z)                    -> MIN          */
#define PCXG_MAX  (1u << 4)   /* x = max(y, z)                    -> MAX          */
#define PCXG_EQR  (1u << 5)   /* x <=> (y = z)                    -> EQ           */
#define PCXG_LEQR (1u << 6)   /* x <=> (y <= z)                   -> LEQ          */
#define PCXG_ALL  0x7fu

typedef struct {
  uint64_t seed;
  int64_t  nvars;
  int64_t  nprops;
  uint32_t kinds;          /* PCXG_* mask; 0 = PCXG_ALL                         */
  int32_t  planted;        /* 0: operands uniform, bounds as given (literal     */
                           /*    random network);                               */
                           /* 1: every propagator is satisfied by a hidden      */
                           /*    assignment, so the network cannot fail         */
  int32_t  lb_max;         /* lb ~ U[0, lb_max]                                 */
  int32_t  width_min;      /* ub = lb + U[width_min, width_max] (planted:       */
  int32_t  width_max;      /*    conditioned on containing the hidden value)    */
  int32_t  value_range;    /* planted: hidden values in [0, value_range)        */
  double   bool_frac;      /* planted: share of variables with values in {0,1} */
  double   const_frac;     /* planted: share of variables fixed to their value  */
                           /*    (the constants a ternarizer introduces)        */
  double   chain_frac;     /* share of propagators arranged in chains, each     */
                           /* link's y = the previous link's x; planted chains  */
                           /* start at a constant and take z among constants    */
  int32_t  chain_len;      /* links per chain                                   */
  int32_t  reserved[3];
} pcxg_params;

/* Defaults of the 1e6-variable / 4e6-propagator workload. */
void pcxg_params_default(pcxg_params* p);

/* Fill op/x/y/z[nprops] (told order, not sorted) and lb/ub[nvars].  The [0,1]
 * embedding of reified outputs is left to pcx_tell.  Planted networks also
 * return the hidden assignment in `witness` (may be NULL). */
pcx_status pcxg_network(const pcxg_params* p, int32_t* op, int32_t* x, int32_t* y, int32_t* z,
                        int32_t* lb, int32_t* ub, int32_t* witness);

/* Search-node splits: for each of n_inst copies of the store lb/ub[nvars],
 * draw k ~ U[kmin, kmax] splits "v <= c" or "v > c" with c uniform inside the
 * current bounds of a random variable v with a non-singleton domain.
 * Writes the n_inst*nvars bounds to lb_out/ub_out. */
pcx_status pcxg_splits(uint64_t seed, const int32_t* lb, const int32_t* ub, int64_t nvars,
                       int64_t n_inst, int32_t kmin, int32_t kmax,
                       int32_t* lb_out, int32_t* ub_out);

/* Propagator subsets: instance k keeps m_k ~ U[mmin, mmax] (clamped to
 * nprops) distinct propagators of a table of nprops, chosen uniformly and
 * listed in ascending order.  off[n_inst+1] is written first; call with
 * idx == NULL to learn the total (off[n_inst]), then again with idx. */
pcx_status pcxg_subsets(uint64_t seed, int64_t nprops, int64_t n_inst, int64_t mmin, int64_t mmax,
                        int64_t* off, int32_t* idx);

/* Values of generated bitset domains: [0, PCXG_BITS_VALUES).  These are the
 * values NBitset<64> of the reference's bitset store holds exactly
 * (tests/pc_bitset_test.cpp:23-25): its first and last bits stand for all
 * negative values and for every value past 61 (pc_bitset_test.cpp:150-157,
 * `var -15..5` is NBit(-1, 5)).  Propagation only removes values, so every
 * store of a generated workload stays inside [0, 61], where a plain word and
 * NBitset<64> describe the same set. */
#define PCXG_BITS_VALUES 62

/* Bitset network: nvars words, each value in [0, PCXG_BITS_VALUES) present
 * with a density drawn per variable from U[dens_lo, dens_hi] (never empty); nprops
 * propagators with opcodes drawn uniformly from `kinds` (bit k = opcode k of
 * the bitset domain, 0 = all four), operands distinct.
 *   planted = 0: operands uniform (a random network; it usually fails);
 *   planted = 1: a hidden value per variable lies in its domain and every
 *     propagator holds under it (x = y + z picks x holding w_y + w_z, y <= z
 *     and y != z are oriented/redrawn to hold, an EQR output holds
 *     [w_y = w_z]), so the root store cannot fail; `witness` (may be NULL)
 *     receives the hidden values.
 * Told order, not sorted. */
pcx_status pcxg_bits_network(uint64_t seed, int64_t nvars, int64_t nprops, uint32_t kinds,
                             double dens_lo, double dens_hi, int32_t planted,
                             int32_t* op, int32_t* x, int32_t* y, int32_t* z, uint64_t* words,
                             int32_t* witness);

/* Bit-removal splits: n_inst copies of words[nvars], each with k ~
 * U[kmin, kmax] removals of one present value from a random variable that
 * holds at least two values. */
pcx_status pcxg_bits_splits(uint64_t seed, const uint64_t* words, int64_t nvars, int64_t n_inst,
                            int32_t kmin, int32_t kmax, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif /* PCX_GEN_H */
