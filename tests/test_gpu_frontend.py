"""The reference's FlatZinc test models (tests/golden/fzn_kats.json) through
the front-end and the GPU engine: told step by step into one PIR (pcx_tell
appends and embeds, as PIR::deduce(tell_type) does), every original variable
must end as tests/pir_test.cpp expects, failures included.  A generated
network written as FlatZinc gives, on the GPU, the oracle's fixpoint."""
import numpy as np
import pytest

import oracle_lib as ol
import pcx
from frontend_cases import load_cases, load_pc_cases, run_case

pytestmark = pytest.mark.gpu

CASES = load_cases()
PC_CASES = load_pc_cases()


@pytest.fixture(scope="module")
def gpu():
    if pcx.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True


class Engine:
    def __init__(self):
        self.pir = pcx.PIR()

    def __call__(self, op, x, y, z, lb, ub, before):
        if len(op):
            self.pir.tell(op, x, y, z, lb, ub)
        before()
        return bool(self.pir.fixpoint(lb, ub).failed)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_reference_models_on_the_gpu(gpu, case):
    errs, _ = run_case(case, Engine())
    assert not errs, f"{case['name']} ({case['ref']}): " + "; ".join(errs)


@pytest.mark.parametrize("case", PC_CASES, ids=lambda c: c["name"])
def test_reference_pc_models_on_the_gpu(gpu, case):
    """tests/pc_test.cpp's models, decomposed and propagated on the GPU: PC's
    stores on every original variable, failures included."""
    errs, _ = run_case(case, Engine())
    assert not errs, f"{case['name']} ({case['ref']}): " + "; ".join(errs)


class TreeEngine(Engine):
    """The GPU fixpoint, then PC's is_extractable of the model's formula trees
    asked on the GPU (pcx_batch_ask_formulas) per step."""

    def __init__(self):
        super().__init__()
        self.steps = []

    def __call__(self, op, x, y, z, lb, ub, before):
        failed = super().__call__(op, x, y, z, lb, ub, before)
        nodes, off = self.model.formulas()
        b = pcx.Batch(self.pir, 1, len(lb))
        b.load(lb, ub)
        self.steps.append(bool(b.ask_formulas(nodes, off)[0]))
        b.close()
        return failed


def test_pc_is_extractable_on_the_gpu(gpu):
    """tests/pc_test.cpp's is_extractable flags (pc.hpp:727-738) from the
    model's formula trees, each tree interpreted by a GPU thread: every step
    of every PC model test that does not fail."""
    n = 0
    for c in PC_CASES:
        eng = TreeEngine()
        errs, _ = run_case(c, eng)
        assert not errs, c["name"]
        for k, (st, ua) in enumerate(zip(c["steps"], eng.steps)):
            if st["failed"]:
                continue
            n += 1
            assert ua == st["ua"], f"{c['name']} ({c['ref']}) step {k}: is_extractable {ua}, expected {st['ua']}"
    assert n >= 90


@pytest.mark.parametrize("seed", [3, 4])
def test_generated_model_matches_oracle(gpu, seed):
    p = pcx.gen_params(nvars=1000, nprops=3000, seed=seed)
    op, x, y, z, lb, ub, _ = pcx.gen_network(p)
    m = pcx.Model(pcx.flatzinc(op, x, y, z, lb, ub))
    mop, mx, my, mz = m.table()
    glb, gub = m.store()
    pir = pcx.PIR()
    pir.tell(mop, mx, my, mz, glb, gub)
    r = pir.fixpoint(glb, gub)
    top, tx, ty, tz = pir.table()
    olb, oub = m.store()
    ol.tell_embed(mop.copy(), mx.copy(), olb, oub)
    o = ol.fixpoint(top, tx, ty, tz, olb, oub)
    assert bool(r.failed) == bool(o.failed)
    if not o.failed:
        assert np.array_equal(glb, olb) and np.array_equal(gub, oub)
