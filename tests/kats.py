"""Known-answer-test runner shared by the oracle tests and the GPU parity tests.

A KAT (tests/golden/pir_kats.json) is a sequence of steps, each of which may
tell new ternary propagators and/or embed unary bounds, then runs a fixpoint.
The runner mirrors the reference test helpers `deduce_and_test` /
`deduce_and_test_bot` (tests/pir_test.cpp:54-89): check num_deductions and the
store before the fixpoint, run the fixpoint, check the store after it (or
failure), then check the is_ua flag (not failed and every propagator
entailed, tests/pir_test.cpp:32-52).

An engine object provides:
    tell(op, x, y, z, lb, ub)     append propagators, embed reified outputs
    num_deductions() -> int
    fixpoint(lb, ub) -> failed     mutate lb/ub in place
    all_entailed(lb, ub) -> bool
"""
import json
import os

import numpy as np

from oracle_lib import INF, MINF, OPS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pir_kats.json")


def _bound(v):
    if v == "INF":
        return INF
    if v == "MINF":
        return MINF
    return int(v)


def _itv(e):
    if e == "TOP":
        return (MINF, INF)
    return (_bound(e[0]), _bound(e[1]))


def load_cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def run_case(case, make_engine):
    """Run one KAT; returns a list of mismatch strings (empty = pass)."""
    names = {n: i for i, n in enumerate(case["vars"])}
    init = [_itv(e) for e in case["init"]]
    lb = np.array([a for a, _ in init], dtype=np.int32)
    ub = np.array([b for _, b in init], dtype=np.int32)
    eng = make_engine()
    errs = []
    for s, step in enumerate(case["steps"]):
        if "tell" in step:
            t = step["tell"]
            op = np.array([OPS[b[0]] for b in t], dtype=np.int32)
            x = np.array([names[b[1]] for b in t], dtype=np.int32)
            y = np.array([names[b[2]] for b in t], dtype=np.int32)
            z = np.array([names[b[3]] for b in t], dtype=np.int32)
            eng.tell(op, x, y, z, lb, ub)
        for v, lo, hi in step.get("embed", []):
            i = names[v]
            lb[i] = max(lb[i], _bound(lo))
            ub[i] = min(ub[i], _bound(hi))
        if eng.num_deductions() != step["nd"]:
            errs.append(f"step {s}: num_deductions {eng.num_deductions()} != {step['nd']}")
        for i, e in enumerate(step["before"]):
            if (int(lb[i]), int(ub[i])) != _itv(e):
                errs.append(f"step {s}: before var {case['vars'][i]} = [{lb[i]},{ub[i]}] != {_itv(e)}")
        failed = eng.fixpoint(lb, ub)
        if step["after"] == "bot":
            if not failed:
                errs.append(f"step {s}: expected failure, got none")
            continue
        if failed:
            errs.append(f"step {s}: unexpected failure")
            continue
        for i, e in enumerate(step["after"]):
            if (int(lb[i]), int(ub[i])) != _itv(e):
                errs.append(f"step {s}: after var {case['vars'][i]} = [{lb[i]},{ub[i]}] != {_itv(e)}")
        if "ua" in step:
            ua = eng.all_entailed(lb, ub)
            if ua != step["ua"]:
                errs.append(f"step {s}: is_ua {ua} != {step['ua']}")
    return errs
