/* pcx_model.h -- front-end: FlatZinc-style models to the PIR ternary form.
 *
 * The reference's PIR accepts only ternary constraints `X = Y op Z` and
 * `X <=> Y op Z` over three variables (PIR::interpret_formula,
 * include/lala/pir.hpp:254-318).  Its tests state models as FlatZinc text
 * (tests/pir_test.cpp, e.g. 175-182, 214-221, 362-369), which lala-core's
 * ternarize decomposes into that form before the PIR sees them: nested
 * terms become auxiliary variables, constants become fixed singleton
 * variables (the generated variables listed at pir_test.cpp:193-206), and a
 * unary constraint (a variable against a constant) is kept by the store as
 * a domain.  A pcx_model does that decomposition on the host and hands the
 * result to pcx_tell / pcx_fixpoint as plain arrays.
 *
 * Accepted items (one per `;`):
 *   var <dom>: name [= expr];          dom: int | bool | L..U | {v, ...}
 *   array[1..n] of var <dom>: name [= [e, ...]];
 *   array[1..n] of int|bool: name = [c, ...];   int: name = c;   bool: name = b;
 *   constraint expr;                    solve ...;   (ignored)
 * Annotations (`:: ...`) and `%` comments are skipped.  Predicates: the
 * int_/bool_ comparisons and their _reif forms, int_plus/minus/times/div/
 * tdiv/fdiv/cdiv/ediv/min/max/neg/abs (as 3-argument constraints or nested
 * terms), int_lin_le/eq/ne, bool_not/and/or/xor/clause, nbool_or/and/equiv,
 * array_bool_and/or, bool2int, array_int_element, array_var_int_element,
 * array_bool_element, set_in, set_in_reif.  Domains with holes ({1, 3}) are
 * kept as their hull plus a disjunction of equalities.
 *
 * Variables are numbered in declaration order (array elements in index
 * order), auxiliary and constant variables as they are created.  Unbounded
 * domains are [PCX_MINF, PCX_INF], the store's infinities.
 */
#ifndef PCX_MODEL_H
#define PCX_MODEL_H

#include "pcx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pcx_model pcx_model;

/* Parse a model.  Errors: PCX_EINVAL (syntax, unknown name, constant out of
 * int32 range) or PCX_ENOTSUP (a predicate this front-end does not
 * decompose), with the line and the reason in pcx_last_error().
 *   replaces interpret_tell of a model's formula (pir.hpp:316-318) composed
 *   with lala-core's ternarize. */
pcx_status pcx_model_parse(const char* text, pcx_model** out);
/* More items on the same model, as the reference's tests tell further
 * constraints into an existing PIR (interpret_must_succeed<IKind::TELL>,
 * e.g. pir_test.cpp:577-581).  New propagators are appended; domains only
 * shrink. */
pcx_status pcx_model_add(pcx_model* m, const char* text);
void       pcx_model_destroy(pcx_model* m);

int64_t    pcx_model_num_vars(const pcx_model* m);
/* Propagators decomposed so far (PIR::num_deductions, pir.hpp:382-384,
 * before the tell). */
int64_t    pcx_model_num_props(const pcx_model* m);
/* Propagators [from, num_props) in the order created, opcodes as in pcx.h:
 * pass them to pcx_tell. */
pcx_status pcx_model_table(const pcx_model* m, int64_t from, int32_t* op, int32_t* x, int32_t* y,
                           int32_t* z);
/* The declared and restricted domains of all variables (num_vars entries).
 * After pcx_model_add, meet them into the current store (lb = max, ub = min;
 * new variables take them as they are). */
pcx_status pcx_model_store(const pcx_model* m, int32_t* lb, int32_t* ub);
/* Index of a variable (`x`, `xs[2]`), or -1. */
int32_t    pcx_model_var(const pcx_model* m, const char* name);
/* Name of variable v (auxiliary: __AUX_k, constants: __CONSTANT_c), or NULL. */
const char* pcx_model_var_name(const pcx_model* m, int32_t v);

/* PC's formula trees.  The reference's PC keeps each constraint as a tree
 * of formulas over terms (include/lala/formula.hpp, terms.hpp) and its
 * is_extractable asks every tree (pc.hpp:727-738).  The model keeps each
 * constraint told so far as such a tree, in postorder: nodes of four
 * int32 {kind, a, b, c}.  Terms push an interval: VAR (a = variable),
 * CONST (a = value), NEG, ABS (one operand), ADD, SUB, MUL, TDIV, FDIV,
 * CDIV, EDIV, MIN, MAX (two).  Formulas push a truth value (entailed,
 * disentailed, or neither): LE, LT, EQ, NE (two terms), AND, OR (a =
 * number of operands), NOT, EQUIV, XOR (formulas), LIT (a term read as a
 * boolean: true when 0 is not in it, false when it is {0}), PROP (one
 * ternary propagator of pcx.h: a = op, b = x, c = y, and the next node's a =
 * z; a constraint this tree form does not cover is the conjunction of the
 * propagators it was decomposed into), TRUE, FALSE. */
enum {
  PCX_FT_VAR = 1, PCX_FT_CONST, PCX_FT_NEG, PCX_FT_ABS, PCX_FT_ADD, PCX_FT_SUB, PCX_FT_MUL,
  PCX_FT_TDIV, PCX_FT_FDIV, PCX_FT_CDIV, PCX_FT_EDIV, PCX_FT_MIN, PCX_FT_MAX,
  PCX_FT_LE, PCX_FT_LT, PCX_FT_EQ, PCX_FT_NE, PCX_FT_AND, PCX_FT_OR, PCX_FT_NOT, PCX_FT_EQUIV,
  PCX_FT_XOR, PCX_FT_LIT, PCX_FT_PROP, PCX_FT_PROP_Z, PCX_FT_TRUE, PCX_FT_FALSE
};
/* Formulas (constraints told so far) and their nodes in total. */
int64_t    pcx_model_num_formulas(const pcx_model* m);
int64_t    pcx_model_num_formula_nodes(const pcx_model* m);
/* nodes: 4 * num_formula_nodes int32; off: num_formulas + 1 offsets (in
 * nodes) of each formula's postorder list. */
pcx_status pcx_model_formulas(const pcx_model* m, int32_t* nodes, int64_t* off);

#ifdef __cplusplus
}
#endif
#endif
