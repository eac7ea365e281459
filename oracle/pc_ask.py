"""Test infrastructure (the oracle): PC's is_extractable over formula trees,
restated on the CPU in plain Python for the small stores of the reference's
model tests.  Only tests import it, as the checker.

A store is (lb, ub) int32 arrays with -2^31 / 2^31-1 as the infinities.  A
formula is a postorder list of nodes {kind, a, b, c} (kinds as in
include/pcx_model.h).  Terms evaluate to intervals (PC's Term::project:
include/lala/terms.hpp -- Variable 70-72, Constant 35, unary and binary
terms over lala-core's interval arithmetic, restated here: sums and
products over the bounds, division by an interval that excludes zero over
its corners, with the rounding of the division kind).  Formulas evaluate to
(ask, nask) pairs as PC's formulas do (include/lala/formula.hpp): an
inequality l <= r is asked as ub(l) <= lb(r) and disproved as lb(l) > ub(r)
(760-770), an equality as two equal singletons / disjoint intervals
(617-628), a variable literal as 0 not in / the value {0} (101-110), and
conjunction, disjunction, negation, equivalence and exclusive or from their
parts (268-273, 338-343, 408-418, 546-555).  is_extractable: the store is
not empty and every formula is asked (pc.hpp:727-738).  A PROP node (a
ternary propagator that stands for a constraint the trees do not cover) is
asked with the caller's ternary ask (PIR's, pir.hpp:417-438) and never
disproved.
"""
import math

(VAR, CONST, NEG, ABS, ADD, SUB, MUL, TDIV, FDIV, CDIV, EDIV, MIN, MAX, LE, LT, EQ, NE, AND, OR,
 NOT, EQUIV, XOR, LIT, PROP, PROP_Z, TRUE, FALSE) = range(1, 28)

MINF, INF = -2 ** 31, 2 ** 31 - 1


def _in(v):
    return -math.inf if v == MINF else math.inf if v == INF else v


def _sat(v):
    if v == -math.inf or v <= MINF:
        return -math.inf
    if v == math.inf or v >= INF:
        return math.inf
    return v


def _add(a, b, first):
    """A bound of a sum: the infinity `first` (the one the bound leans to)
    wins over the other."""
    if a == first or b == first:
        return first
    if math.isinf(a):
        return a
    if math.isinf(b):
        return b
    return a + b


def _mul(a, b):
    if a == 0 or b == 0:
        return 0
    return a * b


def _div(kind, a, b):
    if math.isinf(a):
        return a if b > 0 else -a
    if math.isinf(b):
        return 0
    q, r = divmod(a, b)                    # floor division
    if kind == FDIV:
        return q
    if kind == CDIV:
        return q + (1 if r else 0)
    if kind == TDIV:
        return q + (1 if r and (a < 0) != (b < 0) else 0)
    # Euclidean: the remainder is never negative
    q = a // b if b > 0 else -(a // -b)
    if a - q * b < 0:
        q += 1 if b < 0 else -1
    return q


def _term(kind, stack):
    if kind in (NEG, ABS):
        l, u = stack.pop()
        if kind == NEG:
            stack.append((-u, -l))
        elif l >= 0:
            stack.append((l, u))
        elif u <= 0:
            stack.append((-u, -l))
        else:
            stack.append((0, max(-l, u)))
        return
    (bl, bu) = stack.pop()
    (al, au) = stack.pop()
    if kind in (ADD, SUB):
        if kind == SUB:                    # a - b = a + (-b)
            bl, bu = -bu, -bl
        r = (_add(al, bl, -math.inf), _add(au, bu, math.inf))
    elif kind == MUL:
        c = [_mul(x, y) for x in (al, au) for y in (bl, bu)]
        r = (min(c), max(c))
    elif kind in (TDIV, FDIV, CDIV, EDIV):
        if bl <= 0 <= bu:
            r = (-math.inf, math.inf)
        else:
            c = [_div(kind, x, y) for x in (al, au) for y in (bl, bu)]
            r = (min(c), max(c))
    elif kind == MIN:
        r = (min(al, bl), min(au, bu))
    else:
        r = (max(al, bl), max(au, bu))
    stack.append((_sat(r[0]), _sat(r[1])))


def ask_formula(nodes, lb, ub, ask3):
    """(ask, nask) of one formula's postorder nodes on the store."""
    stack = []
    i = 0
    while i < len(nodes):
        k, a, b, c = (int(v) for v in nodes[i])
        if k == VAR:
            stack.append((_in(int(lb[a])), _in(int(ub[a]))))
        elif k == CONST:
            stack.append((_in(a), _in(a)))
        elif k in (NEG, ABS, ADD, SUB, MUL, TDIV, FDIV, CDIV, EDIV, MIN, MAX):
            _term(k, stack)
        elif k in (LE, LT, EQ, NE):
            (rl, ru) = stack.pop()
            (ll, lu) = stack.pop()
            if k == LE:
                t = (lu <= rl, ll > ru)
            elif k == LT:
                t = (lu < rl, ll >= ru)
            else:
                same = ll == lu == rl == ru
                apart = lu < rl or ru < ll
                t = (same, apart) if k == EQ else (apart, same)
            stack.append(t)
        elif k in (AND, OR):
            parts = stack[len(stack) - a:]
            del stack[len(stack) - a:]
            if k == AND:
                stack.append((all(p[0] for p in parts), any(p[1] for p in parts)))
            else:
                stack.append((any(p[0] for p in parts), all(p[1] for p in parts)))
        elif k == NOT:
            (p, n) = stack.pop()
            stack.append((n, p))
        elif k in (EQUIV, XOR):
            (gp, gn) = stack.pop()
            (fp, fn) = stack.pop()
            same = (fp and gp) or (fn and gn)
            differ = (fp and gn) or (fn and gp)
            stack.append((same, differ) if k == EQUIV else (differ, same))
        elif k == LIT:
            (l, u) = stack.pop()
            stack.append((not (l <= 0 <= u), l >= 0 and u <= 0))
        elif k == PROP:
            z = int(nodes[i + 1][1])
            stack.append((bool(ask3(a, b, c, z, lb, ub)), False))
            i += 1
        elif k == TRUE:
            stack.append((True, False))
        elif k == FALSE:
            stack.append((False, True))
        else:
            raise ValueError(f"unknown node kind {k}")
        i += 1
    assert len(stack) == 1
    return stack[0]


def is_extractable(nodes, off, lb, ub, ask3):
    """PC's is_extractable: not empty, and every formula asked."""
    if any(int(l) > int(u) for l, u in zip(lb, ub)):
        return False
    return all(ask_formula(nodes[off[f]:off[f + 1]], lb, ub, ask3)[0] for f in range(len(off) - 1))
