"""Pins the CPU oracle before anything is compared against it.

1. Known-answer tests transcribed from the reference (tests/pir_test.cpp).
2. The reference's brute-force bound-consistency methodology
   (tests/bound_consistency_test.hpp): soundness over interval shapes,
   completeness on singletons, exhaustive small boxes, entailment soundness.
3. Monotonicity of every propagator (not tested by the reference; it is what
   makes the fixpoint unique under any iteration order).
4. Mutation pinning: plausible mistakes planted in copies of the oracle (one
   per build flag) must each be caught by (1)-(3).
5. Order independence of the Gauss-Seidel fixpoint on random networks.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import kats
import oracle_lib as ol

COMPLETE = {"ADD": 1, "MUL": 0, "TDIV": 0, "FDIV": 0, "CDIV": 0, "EDIV": 0,
            "MIN": 1, "MAX": 1, "EQ": 1, "LEQ": 1}


class OracleEngine:
    """PIR over the oracle: tell appends + embeds [0,1] + stable-sorts."""

    def __init__(self):
        self.op = np.zeros(0, np.int32)
        self.x = np.zeros(0, np.int32)
        self.y = np.zeros(0, np.int32)
        self.z = np.zeros(0, np.int32)

    def tell(self, op, x, y, z, lb, ub):
        ol.tell_embed(op, x, lb, ub)
        self.op = np.concatenate([self.op, op]).astype(np.int32)
        self.x = np.concatenate([self.x, x]).astype(np.int32)
        self.y = np.concatenate([self.y, y]).astype(np.int32)
        self.z = np.concatenate([self.z, z]).astype(np.int32)
        ol.sort_table(self.op, self.x, self.y, self.z)

    def num_deductions(self):
        return len(self.op)

    def fixpoint(self, lb, ub):
        return bool(ol.fixpoint(self.op, self.x, self.y, self.z, lb, ub).failed)

    def all_entailed(self, lb, ub):
        return all(ol.ask(self.op[i], self.x[i], self.y[i], self.z[i], lb, ub)
                   for i in range(len(self.op)))


@pytest.mark.parametrize("case", kats.load_cases(), ids=lambda c: c["name"])
def test_kat(case):
    errs = kats.run_case(case, OracleEngine)
    assert not errs, f"{case['name']} ({case['ref']}): " + "; ".join(errs)


def _checks(L, op, buf):
    n = L.pcxo_check_shapes(op, COMPLETE_BY_CODE[op], buf, len(buf))
    if n:
        return "shapes: " + buf.value.decode()
    n = L.pcxo_check_singletons(op, -5, 5, buf, len(buf))
    if n:
        return "singletons: " + buf.value.decode()
    n = L.pcxo_check_exhaustive(op, -4, 4, COMPLETE_BY_CODE[op], buf, len(buf))
    if n:
        return "exhaustive: " + buf.value.decode()
    n = L.pcxo_check_monotone(op, -4, 4, buf, len(buf))
    if n:
        return "monotone: " + buf.value.decode()
    if op == ol.OPS["MUL"]:
        n = L.pcxo_check_mul_const(-7, 7, buf, len(buf))
        if n:
            return "mul by constant: " + buf.value.decode()
    return ""


COMPLETE_BY_CODE = {ol.OPS[k]: v for k, v in COMPLETE.items()}


@pytest.mark.parametrize("name", list(ol.OPS))
def test_bound_consistency(name):
    """bound_consistency_test.hpp:55-225 methodology + monotonicity."""
    L = ol.lib()
    buf = ctypes.create_string_buffer(512)
    err = _checks(L, ol.OPS[name], buf)
    assert not err, f"{name}: {err}"


# Each mutant is a plausible slip in the restatement (wrong operand, wrong
# rounding, off-by-one, dropped statement).  Every one must be caught by the
# KATs or the brute-force checks, otherwise those checks pin nothing.
MUTANTS = {
    1: "ADD: y.lb uses z.lb instead of z.ub",
    2: "MUL: mul_inv lower bound rounds down (fdiv) instead of up",
    3: "MIN: strict test xu < zl relaxed to <=",
    4: "EQ: x=0 shave keeps the equal bound (yl instead of yl+1)",
    5: "LEQ: x=0 branch uses zl instead of zl+1",
    6: "MAX: y.ub = min(yu, xu) dropped",
    7: "EDIV: Euclidean division replaced by floor division",
    8: "tell: reified outputs not embedded into [0,1]",
}


@pytest.fixture(scope="module")
def mutant_dir():
    d = tempfile.mkdtemp(prefix="pcx_mut_")
    yield d


def _build_mutant(d, k):
    out = os.path.join(d, f"libmut{k}.so")
    src = [os.path.join(ol.ORACLE_DIR, f) for f in ("pir_oracle.c", "pir_checks.c", "bits_oracle.c")]
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-shared", "-std=c11", f"-DPCXO_MUTANT={k}",
                           "-o", out] + src + ["-lm", "-pthread"])
    return out


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_mutant_is_caught(k, mutant_dir, monkeypatch):
    path = _build_mutant(mutant_dir, k)
    L = ctypes.CDLL(path)
    # run the KATs through the mutant
    monkeypatch.setattr(ol, "_lib", None)
    monkeypatch.setattr(ol, "ORACLE_LIB", path)
    caught = []
    for case in kats.load_cases():
        if kats.run_case(case, OracleEngine):
            caught.append(case["name"])
            break
    if not caught:
        ol._lib = None
        L = ol.lib()
        buf = ctypes.create_string_buffer(512)
        for op in range(10):
            if _checks(L, op, buf):
                caught.append(op)
                break
    assert caught, f"mutant {k} ({MUTANTS[k]}) passed every check"


def _random_net(rng, V, P, lo=0, hi=40):
    op = rng.integers(0, 10, P).astype(np.int32)
    op[op == 2] = 0   # keep divisions rare but present
    xyz = np.stack([rng.choice(V, 3, replace=False) for _ in range(P)]).astype(np.int32)
    lbv = rng.integers(lo, hi, V).astype(np.int32)
    ubv = (lbv + rng.integers(0, hi - lo, V)).astype(np.int32)
    return op, xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy(), lbv, ubv


@pytest.mark.parametrize("seed", range(12))
def test_order_independence(seed):
    """Any permutation of the table gives the same fixpoint (or both fail)."""
    rng = np.random.default_rng(seed)
    op, x, y, z, lb0, ub0 = _random_net(rng, 30, 45)
    lb0, ub0 = lb0.copy(), ub0.copy()
    ol.tell_embed(op, x, lb0, ub0)
    ref = None
    for trial in range(6):
        perm = rng.permutation(len(op))
        lb, ub = lb0.copy(), ub0.copy()
        r = ol.fixpoint(op[perm].copy(), x[perm].copy(), y[perm].copy(), z[perm].copy(), lb, ub)
        got = (bool(r.failed), None if r.failed else (lb.tolist(), ub.tolist()))
        if ref is None:
            ref = got
        assert got == ref, f"seed {seed} permutation {trial} differs"


def test_idempotent_at_fixpoint():
    rng = np.random.default_rng(7)
    for _ in range(20):
        op, x, y, z, lb, ub = _random_net(rng, 25, 30)
        ol.tell_embed(op, x, lb, ub)
        r = ol.fixpoint(op, x, y, z, lb, ub)
        if r.failed:
            continue
        lb2, ub2 = lb.copy(), ub.copy()
        r2 = ol.fixpoint(op, x, y, z, lb2, ub2)
        assert r2.sweeps == 1 and not r2.has_changed
        assert (lb2 == lb).all() and (ub2 == ub).all()
