"""PC's formula trees on the GPU (pcx_batch_ask_formulas): random trees over
random stores, each interpreted by a GPU thread, against the oracle's
restatement of PC's ask (oracle/pc_ask.py); malformed trees are refused."""
import numpy as np
import pytest

import oracle_lib as ol
import pc_ask as pa
import pcx

pytestmark = pytest.mark.gpu

V = 8
BOOLS = (6, 7)                       # 0/1 variables, read as literals
MINF, INF = -2 ** 31, 2 ** 31 - 1


@pytest.fixture(scope="module")
def gpu():
    if pcx.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True


def _term(rng, depth, out):
    if depth == 0 or rng.random() < 0.3:
        if rng.random() < 0.7:
            out.append((pa.VAR, int(rng.integers(0, 6)), 0, 0))
        else:
            out.append((pa.CONST, int(rng.integers(-4, 5)), 0, 0))
        return
    k = int(rng.choice([pa.NEG, pa.ABS, pa.ADD, pa.SUB, pa.MUL, pa.TDIV, pa.FDIV, pa.CDIV, pa.EDIV,
                        pa.MIN, pa.MAX]))
    _term(rng, depth - 1, out)
    if k not in (pa.NEG, pa.ABS):
        _term(rng, depth - 1, out)
    out.append((k, 0, 0, 0))


def _formula(rng, depth, out):
    r = rng.random()
    if depth == 0 or r < 0.45:
        c = rng.random()
        if c < 0.6:
            _term(rng, 2, out)
            _term(rng, 2, out)
            out.append((int(rng.choice([pa.LE, pa.LT, pa.EQ, pa.NE])), 0, 0, 0))
        elif c < 0.75:
            out.append((pa.VAR, int(rng.choice(BOOLS)), 0, 0))
            out.append((pa.LIT, 0, 0, 0))
        elif c < 0.9:
            op = int(rng.integers(0, 10))
            x, y, z = (int(v) for v in rng.integers(0, V, 3))
            out.append((pa.PROP, op, x, y))
            out.append((pa.PROP_Z, z, 0, 0))
        else:
            out.append((int(rng.choice([pa.TRUE, pa.FALSE])), 0, 0, 0))
        return
    k = int(rng.choice([pa.AND, pa.OR, pa.NOT, pa.EQUIV, pa.XOR]))
    if k in (pa.AND, pa.OR):
        n = int(rng.integers(0, 4))
        for _ in range(n):
            _formula(rng, depth - 1, out)
        out.append((k, n, 0, 0))
    elif k == pa.NOT:
        _formula(rng, depth - 1, out)
        out.append((k, 0, 0, 0))
    else:
        _formula(rng, depth - 1, out)
        _formula(rng, depth - 1, out)
        out.append((k, 0, 0, 0))


def _stores(rng, n):
    lo = rng.integers(-6, 7, (n, V))
    hi = lo + rng.integers(0, 5, (n, V))
    lo[:, list(BOOLS)] = rng.integers(0, 2, (n, 2))
    hi[:, list(BOOLS)] = np.maximum(lo[:, list(BOOLS)], rng.integers(0, 2, (n, 2)))
    lo, hi = lo.astype(np.int64), hi.astype(np.int64)
    inf_lo = rng.random((n, V)) < 0.05
    inf_hi = rng.random((n, V)) < 0.05
    lo[inf_lo] = MINF
    hi[inf_hi] = INF
    lo[:, list(BOOLS)] = np.maximum(lo[:, list(BOOLS)], 0)
    hi[:, list(BOOLS)] = np.minimum(hi[:, list(BOOLS)], 1)
    hi = np.maximum(hi, lo)
    return lo.astype(np.int32), hi.astype(np.int32)


def test_random_trees_match_the_oracle(gpu):
    rng = np.random.default_rng(5)
    pir = pcx.PIR()
    n = 96
    lo, hi = _stores(rng, n)
    lo[3, 2], hi[3, 2] = 4, 2                     # an empty store: never extractable
    b = pcx.Batch(pir, n, V)
    b.load(lo.ravel(), hi.ravel())
    asked = 0
    for t in range(300):
        out = []
        _formula(rng, 3, out)
        nodes = np.array(out, np.int32).reshape(-1, 4)
        off = np.array([0, len(nodes)], np.int64)
        got = b.ask_formulas(nodes, off)
        want = np.array([pa.is_extractable(nodes, off, lo[k], hi[k], ol.ask) for k in range(n)])
        bad = np.nonzero(got != want)[0]
        assert len(bad) == 0, f"tree {t}: {out}; instances {bad[:5]}: GPU {got[bad[:5]]}, oracle {want[bad[:5]]}"
        asked += int(want.sum())
    assert asked > 1000                           # both answers occur often
    # several formulas together: extractable iff every one is asked
    forms = []
    for _ in range(5):
        out = []
        _formula(rng, 2, out)
        forms.append(out)
    nodes = np.array([nd for f in forms for nd in f], np.int32).reshape(-1, 4)
    off = np.cumsum([0] + [len(f) for f in forms]).astype(np.int64)
    want = np.array([pa.is_extractable(nodes, off, lo[k], hi[k], ol.ask) for k in range(n)])
    assert np.array_equal(b.ask_formulas(nodes, off), want)


def test_malformed_trees_are_refused(gpu):
    pir = pcx.PIR()
    b = pcx.Batch(pir, 2, 4)
    b.load(np.zeros(8, np.int32), np.ones(8, np.int32))
    bad = [
        [(99, 0, 0, 0)],                                        # unknown kind
        [(pa.VAR, 4, 0, 0), (pa.LIT, 0, 0, 0)],                 # variable outside the store
        [(pa.ADD, 0, 0, 0)],                                    # no operands
        [(pa.VAR, 0, 0, 0), (pa.VAR, 1, 0, 0)],                 # two values left
        [(pa.PROP, 0, 0, 1)],                                   # PROP without PROP_Z
        [(pa.PROP, 12, 0, 1), (pa.PROP_Z, 2, 0, 0)],            # unknown opcode
        [(pa.VAR, 0, 0, 0)] * 33 + [(pa.AND, 33, 0, 0)],         # deeper than 32
    ]
    for nodes in bad:
        arr = np.array(nodes, np.int32).reshape(-1, 4)
        with pytest.raises(pcx.PcxError) as e:
            b.ask_formulas(arr, np.array([0, len(arr)], np.int64))
        assert e.value.status == pcx.EINVAL, nodes
    ok = np.array([(pa.VAR, 0, 0, 0), (pa.CONST, 1, 0, 0), (pa.LE, 0, 0, 0)], np.int32)
    assert b.ask_formulas(ok, np.array([0, 3], np.int64)).all()
    assert b.ask_formulas(np.zeros((0, 4), np.int32), np.array([0], np.int64)).all()   # no formula


def test_planted_solutions_are_extractable(gpu):
    """A generated network as FlatZinc, read back as PC's trees: stores
    holding its planted solution are extractable on the GPU, and a store
    with one value moved off the solution is not (a tree it breaks is
    disproved); the oracle agrees on every store."""
    p = pcx.gen_params(seed=9, nvars=5000, nprops=20000)
    op, x, y, z, lb, ub, w = pcx.gen_network(p)
    pir = pcx.PIR()
    pir.tell(op, x, y, z, lb, ub)
    m = pcx.Model(pcx.flatzinc(op, x, y, z, lb, ub))
    nodes, off = m.formulas()
    n, V = 6, len(w)
    stores = np.tile(w.astype(np.int32), (n, 1))
    free = np.nonzero(lb != ub)[0]
    for k in (2, 5):
        v = int(free[k])
        stores[k, v] = lb[v] if w[v] != lb[v] else ub[v]     # another value of its domain
    b = pcx.Batch(pir, n, V)
    b.load(stores.ravel(), stores.ravel())
    got = b.ask_formulas(nodes, off)
    want = np.array([pa.is_extractable(nodes, off, stores[k], stores[k], ol.ask) for k in range(n)])
    assert np.array_equal(got, want)
    assert got[[0, 1, 3, 4]].all()
