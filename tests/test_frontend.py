"""The front-end (include/pcx_model.h) on the reference's own test models.

tests/golden/fzn_kats.json holds the FlatZinc models of tests/pir_test.cpp
and the stores that test file expects before and after each Gauss-Seidel
fixpoint.  Here the front-end's decomposition is run through the oracle's
fixpoint (the GPU engine does the same in test_gpu_frontend.py): every
variable of the original model must end as the reference expects, failures
included.  The propagator counts are lala-core ternarize's, which the
front-end matches only where the decomposition is the obvious one; they are
reported, not pinned.  Parse errors and unsupported predicates are errors.
"""
import numpy as np
import pytest

import os

import oracle_lib as ol
import pc_ask
import pcx
from frontend_cases import load_cases, load_pc_cases, run_case

CASES = load_cases()
PC_CASES = load_pc_cases()


class Oracle:
    """The oracle as the fixpoint: keeps the whole table, sorted as tell does."""

    def __init__(self):
        self.t = [np.zeros(0, np.int32) for _ in range(4)]

    def __call__(self, op, x, y, z, lb, ub, before):
        op, x = op.copy(), x.copy()
        ol.tell_embed(op, x, lb, ub)
        self.t = [np.concatenate([a, b]) for a, b in zip(self.t, (op, x, y, z))]
        t = [a.copy() for a in self.t]
        ol.sort_table(*t)
        before()
        r = ol.fixpoint(*t, lb, ub)
        return bool(r.failed)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_reference_models_through_the_oracle(case):
    errs, _ = run_case(case, Oracle())
    assert not errs, f"{case['name']} ({case['ref']}): " + "; ".join(errs)


@pytest.mark.parametrize("case", PC_CASES, ids=lambda c: c["name"])
def test_reference_pc_models_through_the_oracle(case):
    """tests/pc_test.cpp's models (n-ary linear, XOR, clauses, element, abs,
    reified and nested terms): PC's expected stores on every original
    variable, failures included."""
    errs, _ = run_case(case, Oracle())
    assert not errs, f"{case['name']} ({case['ref']}): " + "; ".join(errs)


class AskingOracle(Oracle):
    """The oracle, also recording per step whether every ternary propagator is
    entailed (PIR's is_extractable) and whether the fixpoint changed the store."""

    def __init__(self):
        super().__init__()
        self.steps = []

    def __call__(self, op, x, y, z, lb, ub, before):
        op, x = op.copy(), x.copy()
        ol.tell_embed(op, x, lb, ub)
        self.t = [np.concatenate([a, b]) for a, b in zip(self.t, (op, x, y, z))]
        t = [a.copy() for a in self.t]
        ol.sort_table(*t)
        before()
        r = ol.fixpoint(*t, lb, ub)
        ua = not r.failed and all(ol.ask(int(t[0][i]), int(t[1][i]), int(t[2][i]), int(t[3][i]), lb, ub)
                                  for i in range(len(t[0])))
        self.steps.append((ua, bool(r.has_changed)))
        return bool(r.failed)


def _flag_agreement(cases):
    n = ua = ch = 0
    for c in cases:
        o = AskingOracle()
        errs, _ = run_case(c, o)
        assert not errs
        for st, (u, h) in zip(c["steps"], o.steps):
            if st["failed"]:
                continue
            n += 1
            ua += st["ua"] == u
            ch += st["changed"] == h
    return n, ua, ch


def test_extractable_and_changed_flags():
    """is_extractable and has_changed through the decomposition.  For PIR's
    tests every is_extractable flag is the reference's; has_changed differs
    where the front-end's auxiliary variables start wider than ternarize's
    (their first fixpoint tightens them).  PC asks its formula trees
    directly (`-x <= 2` is entailed once x >= -2, with no defining `0 = x + t`
    left to entail), so for PC's tests the flags are counted, not pinned."""
    n, ua, ch = _flag_agreement(CASES)
    print(f"PIR tests: is_extractable {ua}/{n}, has_changed {ch}/{n}")
    assert ua == n and ch >= n - 11
    n, ua, ch = _flag_agreement(PC_CASES)
    print(f"PC tests: is_extractable {ua}/{n}, has_changed {ch}/{n}")
    assert ua >= n - 10 and ch >= n - 11


class TreeAskingOracle(Oracle):
    """The oracle's fixpoint, then PC's is_extractable over the model's
    formula trees (oracle/pc_ask.py) per step."""

    def __init__(self):
        super().__init__()
        self.steps = []

    def __call__(self, op, x, y, z, lb, ub, before):
        failed = super().__call__(op, x, y, z, lb, ub, before)
        nodes, off = self.model.formulas()
        self.steps.append(pc_ask.is_extractable(nodes, off, lb, ub, ol.ask))
        return failed


def test_pc_formula_trees_give_pc_is_extractable():
    """PC asks each constraint as a formula tree.  With the model's trees
    (pcx_model_formulas), every is_extractable flag of tests/pc_test.cpp's
    model tests is the reference's (through the ternary propagators, ten of
    them differ: test_extractable_and_changed_flags): 98 of 98 steps."""
    n = 0
    for c in PC_CASES:
        o = TreeAskingOracle()
        errs, _ = run_case(c, o)
        assert not errs
        for k, (st, ua) in enumerate(zip(c["steps"], o.steps)):
            if st["failed"]:
                continue
            n += 1
            assert ua == st["ua"], f"{c['name']} ({c['ref']}) step {k}: is_extractable {ua}, expected {st['ua']}"
    assert n >= 90


def test_formula_tree_kinds_match_the_header():
    import re
    text = open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "include", "pcx_model.h")).read()
    body = text[text.index("PCX_FT_VAR = 1"):text.index("};", text.index("PCX_FT_VAR = 1"))]
    names = re.findall(r"PCX_FT_(\w+)", body)
    assert [getattr(pc_ask, n) for n in names] == list(range(1, len(names) + 1))


def test_fixture_covers_the_model_tests():
    assert len(PC_CASES) >= 60
    assert len(CASES) >= 70
    assert sum(len(c["steps"]) for c in CASES) >= 100
    assert any(s["failed"] for c in CASES for s in c["steps"])


def test_propagator_counts_reported():
    same = total = 0
    for c in CASES:
        _, counts = run_case(c, Oracle())
        for s, n in zip(c["steps"], counts):
            total += 1
            same += s["nd"] == n
    print(f"front-end propagator counts equal to the reference's in {same} of {total} steps")
    assert total > 0


@pytest.mark.parametrize("text,status", [
    ("var int: x; constraint int_le(x, ;", pcx.EINVAL),
    ("var int: x; constraint foo_bar(x, 3);", pcx.ENOTSUP),
    ("var int: x; var int: x;", pcx.EINVAL),
    ("var int: x; constraint int_le(y, 3);", pcx.EINVAL),
    ("var 1..99999999999: x;", pcx.EINVAL),
    ("predicate p(var int: x);", pcx.EINVAL),
])
def test_errors(text, status):
    with pytest.raises(pcx.PcxError) as e:
        pcx.Model(text)
    assert e.value.status == status
    assert "line 1" in str(e.value)


def test_decomposition_shapes():
    """The decomposition's shape on small models: unary constraints are
    domains, x = term puts the term in x, x < y is one LEQ with 0."""
    m = pcx.Model("var 0..10: x; var 0..10: y; var int: z; constraint int_le(x, 7);"
                  "constraint int_eq(z, int_plus(x, y)); constraint int_lt(x, y);"
                  "constraint int_ne(x, 3);")
    lb, ub = m.store()
    assert (lb[m.var("x")], ub[m.var("x")]) == (0, 7)
    op, x, y, z = m.table()
    X, Y, Z = m.var("x"), m.var("y"), m.var("z")
    c0, c3 = m.var("__CONSTANT_0"), m.var("__CONSTANT_3")
    rows = sorted(zip(op.tolist(), x.tolist(), y.tolist(), z.tolist()))
    assert rows == sorted([(pcx.ADD, Z, X, Y), (pcx.LEQ, c0, Y, X), (pcx.EQ, c0, X, c3)])
    # incremental: more constraints, more variables, domains only shrink
    n0, v0 = m.num_props(), m.num_vars()
    m.add("var bool: b; constraint int_le_reif(x, y, b); constraint int_ge(x, 2);")
    assert m.num_props() == n0 + 1 and m.num_vars() == v0 + 1
    lb2, ub2 = m.meet_store(lb, ub)
    assert (lb2[X], ub2[X]) == (2, 7) and (lb2[m.var("b")], ub2[m.var("b")]) == (0, 1)
    assert m.var("nope") == -1 and m.var_name(10**6) is None


@pytest.mark.parametrize("seed", [1, 2])
def test_flatzinc_round_trip(seed):
    """A generated network written as FlatZinc decomposes back into exactly
    its table and store, and the fixpoint is the oracle's on the table."""
    p = pcx.gen_params(nvars=1000, nprops=3000, seed=seed)
    op, x, y, z, lb, ub, _ = pcx.gen_network(p)
    m = pcx.Model(pcx.flatzinc(op, x, y, z, lb, ub))
    assert m.num_vars() == len(lb) and m.num_props() == len(op)
    mop, mx, my, mz = m.table()
    assert np.array_equal(mop, op) and np.array_equal(mx, x)
    assert np.array_equal(my, y) and np.array_equal(mz, z)
    mlb, mub = m.store()
    assert np.array_equal(mlb, lb) and np.array_equal(mub, ub)


def test_oversized_array_is_refused_before_creating_variables():
    """A var array larger than a store can index is an error at once (the
    sanitizer corpus found the front-end creating 2^28 variables first)."""
    import time
    t0 = time.time()
    with pytest.raises(pcx.PcxError) as e:
        pcx.Model("array[1..2147483647] of var int: q;")
    assert e.value.status == pcx.EINVAL and time.time() - t0 < 5
    m = pcx.Model("array[1..3] of var 0..1: q; constraint bool_clause([q[1]], [q[2]]);")
    assert m.num_vars() >= 3
