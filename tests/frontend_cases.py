"""Runs the reference's FlatZinc test models (tests/golden/fzn_kats.json from
tests/pir_test.cpp, tests/golden/pc_kats.json from tests/pc_test.cpp)
through the front-end and a fixpoint, step by step, as the reference's
tests do: parse and tell, check the store before the
fixpoint, run Gauss-Seidel (here: the oracle or the GPU engine), check the
store after it, tell the next constraints into the same store.  Only the
variables of the original model are compared (by name); the generated ones
are numbered differently from lala-core's ternarize."""
import json
import os

import numpy as np

import pcx

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fzn_kats.json")
GOLDEN_PC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pc_kats.json")


def load_cases(path=GOLDEN):
    with open(path) as f:
        return json.load(f)["cases"]


def load_pc_cases():
    """The model tests of the reference's PC (tests/pc_test.cpp).  PC keeps
    formulas as trees and interprets them directly; here they go through the
    same ternary decomposition, and the stores of the original variables must
    still be PC's (the fixpoint of the decomposition projected on them)."""
    return load_cases(GOLDEN_PC)


def run_case(case, backend):
    """backend(new_op, new_x, new_y, new_z, lb, ub) -> failed: tells the new
    propagators (embedding into lb/ub) and runs the fixpoint in place.
    Returns (errors, propagator counts of the front-end per step)."""
    errs, counts = [], []
    model = None
    lb = ub = np.zeros(0, np.int32)
    told = 0
    for k, step in enumerate(case["steps"]):
        if model is None:
            model = pcx.Model(step["text"])
        elif step["text"]:
            model.add(step["text"])
        lb, ub = model.meet_store(lb, ub)
        backend.model = model                    # a backend may read the model (its formulas)
        op, x, y, z = model.table(told)
        told = model.num_props()
        counts.append(told)
        failed = backend(op, x, y, z, lb, ub, before=lambda: check(case, model, lb, ub, step["before"],
                                                                     f"step {k} before", errs))
        if step["failed"]:
            if not failed:
                errs.append(f"step {k}: expected failure")
            return errs, counts
        if failed:
            errs.append(f"step {k}: unexpected failure")
            return errs, counts
        check(case, model, lb, ub, step["after"], f"step {k} after", errs)
    return errs, counts


def check(case, model, lb, ub, want, what, errs):
    for i, name in enumerate(case["names"]):
        if name.startswith("__") or i >= len(want) or want[i] == "bot":
            continue
        v = model.var(name)
        if v < 0:
            errs.append(f"{what}: no variable {name}")
            continue
        got = [int(lb[v]), int(ub[v])]
        if got != want[i]:
            errs.append(f"{what}: {name} = {got}, expected {want[i]}")
