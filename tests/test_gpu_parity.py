"""GPU parity: every result of the HIP engine equals the CPU oracle's.

Monotone, extensive propagators have one fixpoint whatever the iteration
order, so the asynchronous GPU sweeps must reproduce the oracle's
Gauss-Seidel result bit for bit in every bound and in the failure flag (after
a failure only the flag is compared: bounds past a failure depend on order).
All calls go through the C ABI (lala-pc_amd/pcx/libpcx.so).
"""
import numpy as np
import pytest

import kats
import oracle_lib as ol
import pcx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    if pcx.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True


class GpuEngine:
    """KAT engine over pcx.PIR (same interface as the oracle's)."""

    def __init__(self):
        self.pir = pcx.PIR()

    def tell(self, op, x, y, z, lb, ub):
        self.pir.tell(op, x, y, z, lb, ub)

    def num_deductions(self):
        return self.pir.num_deductions()

    def fixpoint(self, lb, ub):
        return bool(self.pir.fixpoint(lb, ub).failed)

    def all_entailed(self, lb, ub):
        return self.pir.is_extractable(lb, ub)


@pytest.mark.parametrize("case", kats.load_cases(), ids=lambda c: c["name"])
def test_kat_on_gpu(gpu, case):
    errs = kats.run_case(case, GpuEngine)
    assert not errs, f"{case['name']} ({case['ref']}): " + "; ".join(errs)


def _boxes(lo, hi):
    r = np.arange(lo, hi + 1)
    iv = np.array([(a, b) for a in r for b in r if a <= b], dtype=np.int32)
    n = len(iv)
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return iv[i.ravel()], iv[j.ravel()], iv[k.ravel()]


def _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, nvars, what):
    n_inst = len(res_o)
    lb_g, ub_g = lb_g.reshape(n_inst, nvars), ub_g.reshape(n_inst, nvars)
    lb_o, ub_o = lb_o.reshape(n_inst, nvars), ub_o.reshape(n_inst, nvars)
    fg = np.array([r.failed for r in res_g], bool)
    fo = np.array([r.failed for r in res_o], bool)
    bad = np.nonzero(fg != fo)[0]
    assert len(bad) == 0, f"{what}: failure flag differs on {len(bad)} instances, first {bad[:5]}"
    ok = ~fo
    diff = ((lb_g != lb_o) | (ub_g != ub_o)).any(axis=1) & ok
    bad = np.nonzero(diff)[0]
    assert len(bad) == 0, (f"{what}: bounds differ on {len(bad)} instances, first {bad[0]}: "
                           f"gpu {lb_g[bad[0]]}/{ub_g[bad[0]]} oracle {lb_o[bad[0]]}/{ub_o[bad[0]]}")
    hg = np.array([r.has_changed for r in res_g], bool)
    ho = np.array([r.has_changed for r in res_o], bool)
    assert (hg[ok] == ho[ok]).all(), f"{what}: has_changed differs"


@pytest.mark.parametrize("name", list(ol.OPS))
@pytest.mark.parametrize("mode", [pcx.MODE_BATCHED, pcx.MODE_PERSISTENT, pcx.MODE_PLAIN])
def test_every_box_one_propagator(gpu, name, mode):
    """All boxes x,y,z in [-4,4] (91,125 instances) for one propagator."""
    lo, hi = (-2, 2) if mode == pcx.MODE_PLAIN else (-4, 4)
    X, Y, Z = _boxes(lo, hi)
    n = len(X)
    lb = np.stack([X[:, 0], Y[:, 0], Z[:, 0]], 1).astype(np.int32).ravel()
    ub = np.stack([X[:, 1], Y[:, 1], Z[:, 1]], 1).astype(np.int32).ravel()
    op = np.array([ol.OPS[name]], np.int32)
    x, y, z = (np.array([v], np.int32) for v in (0, 1, 2))
    # tell embeds [0,1] for reified outputs, instance by instance
    if name in ("EQ", "LEQ"):
        lb[0::3] = np.maximum(lb[0::3], 0)
        ub[0::3] = np.minimum(ub[0::3], 1)
    lb_o, ub_o = lb.copy(), ub.copy()
    res_o = ol.fixpoint_batch(op, x, y, z, lb_o, ub_o, 3, threads=8)
    pir = pcx.PIR()
    pir.tell(op, x, y, z)
    b = pcx.Batch(pir, n, 3)
    b.load(lb, ub)
    res_g, st = b.run(mode)
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, 3, f"{name} mode {mode}")


def _network(seed, V, P, **kw):
    p = pcx.gen_params(seed=seed, nvars=V, nprops=P, **kw)
    op, x, y, z, lb, ub, w = pcx.gen_network(p)
    return op, x, y, z, lb, ub, w


def _oracle_fix(op, x, y, z, lb, ub):
    op, x, y, z = op.copy(), x.copy(), y.copy(), z.copy()
    lb, ub = lb.copy(), ub.copy()
    ol.tell_embed(op, x, lb, ub)
    ol.sort_table(op, x, y, z)
    r = ol.fixpoint(op, x, y, z, lb, ub)
    return r, lb, ub


def _gpu_fix(op, x, y, z, lb, ub, **cfg):
    pir = pcx.PIR(**cfg)
    lb, ub = lb.copy(), ub.copy()
    pir.tell(op, x, y, z, lb, ub)
    r = pir.fixpoint(lb, ub)
    return r, lb, ub, pir


@pytest.mark.parametrize("seed", range(8))
def test_literal_random_network(gpu, seed):
    """1,000 vars in [0,1000], 3,000 propagators over the 7 kinds, operands uniform."""
    net = _network(seed, 1000, 3000, planted=0, lb_max=0, width_min=1000, width_max=1000,
                   chain_frac=0.0)
    ro, lbo, ubo = _oracle_fix(*net[:6])
    rg, lbg, ubg, _ = _gpu_fix(*net[:6])
    assert bool(rg.failed) == bool(ro.failed)
    if not ro.failed:
        assert (lbg == lbo).all() and (ubg == ubo).all()


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("ppw", [64, 512])
def test_literal_network_wide_bounds_fails_like_oracle(gpu, seed, ppw):
    """Random bounds as in the 1e6-variable workload but unplanted: fails."""
    net = _network(seed, 20000, 80000, planted=0, chain_frac=0.3)
    ro, _, _ = _oracle_fix(*net[:6])
    rg, _, _, _ = _gpu_fix(*net[:6], props_per_wave=ppw)
    assert bool(rg.failed) == bool(ro.failed)


@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
@pytest.mark.parametrize("seed,V,P,ppw", [(1, 20000, 80000, 256), (2, 20000, 80000, 64),
                                          (3, 100000, 400000, 128), (4, 100000, 400000, 512)])
def test_planted_network(gpu, seed, V, P, ppw, events):
    op, x, y, z, lb, ub, w = _network(seed, V, P)
    ro, lbo, ubo = _oracle_fix(op, x, y, z, lb, ub)
    rg, lbg, ubg, pir = _gpu_fix(op, x, y, z, lb, ub, props_per_wave=ppw, event_driven=events)
    assert not ro.failed and not rg.failed
    assert (lbg == lbo).all() and (ubg == ubo).all()
    assert ((lbg <= w) & (w <= ubg)).all()
    # idempotence: one more fixpoint changes nothing and takes one sweep
    r2 = pir.fixpoint(lbg, ubg)
    assert not r2.has_changed and r2.sweeps == 1


def test_full_size_network(gpu):
    """1,000,000 variables / 4,000,000 propagators: bit-exact with the oracle.
    The planted chains make the reference's Gauss-Seidel loop take at least
    50 sweeps; the GPU's marked sweeps advance a chain several links per
    sweep (repeated passes over a chunk), so they may take fewer."""
    op, x, y, z, lb, ub, w = _network(1, 1000000, 4000000)
    rg, lbg, ubg, pir = _gpu_fix(op, x, y, z, lb, ub)
    assert not rg.failed
    ro, lbo, ubo = _oracle_fix(op, x, y, z, lb, ub)
    assert ro.sweeps >= 50
    assert (lbg == lbo).all() and (ubg == ubo).all()
    assert ((lbg <= w) & (w <= ubg)).all()


@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
@pytest.mark.parametrize("tables", [(0, 0), (0, 1), (1, -1), (1, 0)],
                         ids=["indexed", "indexed-nt", "copies-nt", "copies-cached"])
@pytest.mark.parametrize("mode", [pcx.MODE_BATCHED, pcx.MODE_PERSISTENT, pcx.MODE_PLAIN])
def test_batch_of_search_nodes(gpu, mode, tables, events):
    """Shared table, per-instance subsets and splits (scaled-down batch), with
    the subsets read through indices or from per-instance copies, and later
    sweeps evaluating only marked propagators or every propagator."""
    V, P, n_inst = 5000, 20000, 96
    op, x, y, z, lb, ub, w = _network(11, V, P)
    pir = pcx.PIR(compact_tables=tables[0], stream_tables=tables[1], event_driven=events)
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_lb, root_ub).failed
    slb, sub = pcx.gen_splits(12, root_lb, root_ub, n_inst, 1, 20)
    off, idx = pcx.gen_subsets(13, P, n_inst, 200, 20000)
    top, tx, ty, tz = pir.table()
    lb_o, ub_o = slb.copy(), sub.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lb_o, ub_o, V, idx=idx.astype(np.int64), off=off, threads=8)
    b = pcx.Batch(pir, n_inst, V, off, idx)
    b.load(slb, sub)
    res_g, st = b.run(mode)
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, V, f"batch mode {mode}")
    # reset restores the loaded stores and a rerun gives the same answer
    b.reset()
    res_g2, _ = b.run(mode)
    lb_g2, ub_g2 = b.read()
    assert (lb_g2 == lb_g).all() and (ub_g2 == ub_g).all()


def test_marked_sweeps_grouped_items(gpu):
    """A batch whose marked sweeps group several marked chunks per work item:
    2,048 instances of a 200,000-propagator network with 150k-200k active
    propagators each leave more than 8,192 chunks holding marks after the
    first sweep.  Every instance equals the oracle and the same batch run
    with full sweeps only."""
    V, P, n_inst = 50000, 200000, 2048
    op, x, y, z, lb, ub, w = _network(21, V, P)
    pir = pcx.PIR(event_driven=1, base_seed=0)     # a full first sweep, then marked ones
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_lb, root_ub).failed
    slb, sub = pcx.gen_splits(22, root_lb, root_ub, n_inst, 1, 20)
    off, idx = pcx.gen_subsets(23, P, n_inst, 150000, 200000)
    top, tx, ty, tz = pir.table()
    lb_o, ub_o = slb.copy(), sub.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lb_o, ub_o, V, idx=idx.astype(np.int64), off=off,
                              threads=16)
    b = pcx.Batch(pir, n_inst, V, off, idx)
    b.load(slb, sub)
    res_g, st = b.run(pcx.MODE_BATCHED)
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, V, "grouped marked items")
    assert st.evals > st.full_evals > 0           # marked sweeps ran after the full one
    pir.set_config(event_driven=0)
    b.reset()
    res_f, _ = b.run(pcx.MODE_BATCHED)
    lb_f, ub_f = b.read()
    _compare(lb_g, ub_g, res_g, lb_f, ub_f, res_f, V, "grouped marked items vs full sweeps")


@pytest.mark.parametrize("mode", [pcx.MODE_BATCHED, pcx.MODE_PLAIN])
def test_dense_first_sweep_marks_everything(gpu, mode):
    """Batches of unpropagated stores (splits of the told store, not of its
    fixpoint): the first sweep tightens most bounds, work items overflow their
    mark lists and mark every propagator for the next sweep by one word.
    Every instance equals the oracle."""
    V, P, n_inst = 20000, 80000, 48
    op, x, y, z, lb, ub, w = _network(31, V, P)
    pir = pcx.PIR(event_driven=1, props_per_wave=2048 if mode == pcx.MODE_BATCHED else 512)
    pir.tell(op, x, y, z, lb, ub)
    slb, sub = pcx.gen_splits(32, lb, ub, n_inst, 1, 20)
    off, idx = pcx.gen_subsets(33, P, n_inst, 40000, 80000)
    top, tx, ty, tz = pir.table()
    lb_o, ub_o = slb.copy(), sub.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lb_o, ub_o, V, idx=idx.astype(np.int64), off=off,
                              threads=16)
    b = pcx.Batch(pir, n_inst, V, off, idx)
    b.load(slb, sub)
    res_g, st = b.run(mode)
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, V, f"dense first sweep, mode {mode}")
    assert st.full_tightened > n_inst * 1000     # dense: far more than one mark list per item


@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
@pytest.mark.parametrize("gi", [16, 32, 64])
def test_lockstep_full_sweeps(gpu, gi, events):
    """Full sweeps in lockstep (gi instances per step over the
    instance-interleaved copy of the stores): 100 instances (not a multiple
    of gi), per-instance subsets, splits of the root fixpoint.  Every instance
    equals the oracle, and a rerun after reset gives the same stores."""
    V, P, n_inst = 5000, 20000, 100
    op, x, y, z, lb, ub, w = _network(41, V, P)
    pir = pcx.PIR(event_driven=events, lockstep=gi, base_seed=0)
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_lb, root_ub).failed
    slb, sub = pcx.gen_splits(42, root_lb, root_ub, n_inst, 1, 20)
    off, idx = pcx.gen_subsets(43, P, n_inst, 200, 20000)
    top, tx, ty, tz = pir.table()
    lb_o, ub_o = slb.copy(), sub.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lb_o, ub_o, V, idx=idx.astype(np.int64), off=off, threads=8)
    b = pcx.Batch(pir, n_inst, V, off, idx)
    b.load(slb, sub)
    res_g, st = b.run(pcx.MODE_BATCHED)
    assert st.lockstep == gi
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, V, f"lockstep {gi} events {events}")
    # a rerun from the same image reuses its interleaved copy
    b.reset()
    res_g2, _ = b.run(pcx.MODE_BATCHED)
    lb_g2, ub_g2 = b.read()
    assert (lb_g2 == lb_g).all() and (ub_g2 == ub_g).all()
    # a new load replaces it
    slb3, sub3 = pcx.gen_splits(44, root_lb, root_ub, n_inst, 1, 20)
    lb_o3, ub_o3 = slb3.copy(), sub3.copy()
    res_o3 = ol.fixpoint_batch(top, tx, ty, tz, lb_o3, ub_o3, V, idx=idx.astype(np.int64), off=off,
                               threads=8)
    b.load(slb3, sub3)
    res_g3, _ = b.run(pcx.MODE_BATCHED)
    lb_g3, ub_g3 = b.read()
    _compare(lb_g3, ub_g3, res_g3, lb_o3, ub_o3, res_o3, V, f"lockstep {gi} after a new load")


@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
@pytest.mark.parametrize("gi", [16, 64])
def test_lockstep_dense_first_sweep(gpu, gi, events):
    """Unpropagated stores and no subsets: the lockstep first sweep tightens
    most bounds, overflows its workgroups' mark lists (whole instances marked
    for the next sweep) and, with full sweeps only, every later sweep is a
    lockstep sweep over the same interleaved copy.  48 instances: one
    partial group at gi = 64.  Every instance equals the oracle."""
    V, P, n_inst = 20000, 80000, 48
    op, x, y, z, lb, ub, w = _network(51, V, P)
    pir = pcx.PIR(event_driven=events, lockstep=gi)
    pir.tell(op, x, y, z, lb, ub)
    slb, sub = pcx.gen_splits(52, lb, ub, n_inst, 1, 20)
    top, tx, ty, tz = pir.table()
    lb_o, ub_o = slb.copy(), sub.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lb_o, ub_o, V, threads=16)
    b = pcx.Batch(pir, n_inst, V)
    b.load(slb, sub)
    res_g, st = b.run(pcx.MODE_BATCHED)
    assert st.lockstep == gi
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, V, f"lockstep {gi} dense first sweep")
    assert st.full_tightened > n_inst * 1000


@pytest.mark.parametrize("gi", [16, 64])
def test_lockstep_all_opcodes_mixed(gpu, gi):
    """A random table over all ten opcodes, divisions included (words of
    mixed opcodes take the per-propagator path, words of one opcode the
    specialized one), on 200 random stores with bounds in [-30, 30]."""
    rng = np.random.default_rng(71)
    V, P, n_inst = 60, 400, 200
    op = rng.integers(0, len(ol.OPS), P).astype(np.int32)
    x = rng.integers(0, V, P).astype(np.int32)
    y = rng.integers(0, V, P).astype(np.int32)
    z = rng.integers(0, V, P).astype(np.int32)
    a = rng.integers(-30, 31, (n_inst, V))
    b = rng.integers(-30, 31, (n_inst, V))
    lb = np.minimum(a, b).astype(np.int32).ravel()
    ub = np.maximum(a, b).astype(np.int32).ravel()
    pir = pcx.PIR(lockstep=gi)
    lb1, ub1 = lb[:V].copy(), ub[:V].copy()
    pir.tell(op, x, y, z, lb1, ub1)           # embeds reified outputs in [0, 1] (first store)
    top, tx, ty, tz = pir.table()
    for k in range(n_inst):                   # the same embedding in every store
        s = slice(k * V, (k + 1) * V)
        r = top >= ol.OPS["EQ"]
        lb[s][tx[r]] = np.maximum(lb[s][tx[r]], 0)
        ub[s][tx[r]] = np.minimum(ub[s][tx[r]], 1)
    lb_o, ub_o = lb.copy(), ub.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lb_o, ub_o, V, threads=8)
    b = pcx.Batch(pir, n_inst, V)
    b.load(lb, ub)
    res_g, st = b.run(pcx.MODE_BATCHED)
    assert st.lockstep == gi
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, V, f"all opcodes, lockstep {gi}")


@pytest.mark.parametrize("name", list(ol.OPS))
@pytest.mark.parametrize("gi", [32, 64])
def test_lockstep_every_box(gpu, name, gi):
    """All boxes x,y,z in [-3,3] (21,952 instances) for one propagator, swept
    in lockstep groups of gi."""
    X, Y, Z = _boxes(-3, 3)
    n = len(X)
    lb = np.stack([X[:, 0], Y[:, 0], Z[:, 0]], 1).astype(np.int32).ravel()
    ub = np.stack([X[:, 1], Y[:, 1], Z[:, 1]], 1).astype(np.int32).ravel()
    op = np.array([ol.OPS[name]], np.int32)
    x, y, z = (np.array([v], np.int32) for v in (0, 1, 2))
    if name in ("EQ", "LEQ"):
        lb[0::3] = np.maximum(lb[0::3], 0)
        ub[0::3] = np.minimum(ub[0::3], 1)
    lb_o, ub_o = lb.copy(), ub.copy()
    res_o = ol.fixpoint_batch(op, x, y, z, lb_o, ub_o, 3, threads=8)
    pir = pcx.PIR(lockstep=gi)
    pir.tell(op, x, y, z)
    b = pcx.Batch(pir, n, 3)
    b.load(lb, ub)
    res_g, st = b.run(pcx.MODE_BATCHED)
    assert st.lockstep == gi
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, 3, f"{name} lockstep {gi}")


def test_reset_restores_the_loaded_stores(gpu):
    """pcx_batch_reset copies the image back over the stores, here with an odd
    number of 8-byte words (3 instances x 5 variables: the copy's last word
    is its tail), after a run that tightened them."""
    op = np.array([pcx.ADD, pcx.LEQ], np.int32)
    x, y, z = np.array([0, 4], np.int32), np.array([1, 2], np.int32), np.array([2, 3], np.int32)
    lb = np.array([0, 1, 2, 0, 0] * 3, np.int32)
    ub = np.array([9, 3, 4, 9, 1] * 3, np.int32)
    lb[5 + 2], ub[5 + 2] = 4, 4            # instance 1: z fixed
    pir = pcx.PIR()
    pir.tell(op, x, y, z)
    b = pcx.Batch(pir, 3, 5)
    b.load(lb, ub)
    res1, _ = b.run(pcx.MODE_BATCHED)
    lb1, ub1 = b.read()
    assert (lb1 != lb).any() or (ub1 != ub).any()      # the run tightened something
    b.reset()
    lb0, ub0 = b.read()
    assert (lb0 == lb).all() and (ub0 == ub).all()
    res2, _ = b.run(pcx.MODE_BATCHED)
    lb2, ub2 = b.read()
    assert (lb2 == lb1).all() and (ub2 == ub1).all()
    assert [r.failed for r in res1] == [r.failed for r in res2]


def test_edge_cases(gpu):
    pir = pcx.PIR()
    # empty table: fixpoint is the store itself, one quiet sweep
    lb = np.array([0, 5], np.int32)
    ub = np.array([3, 4], np.int32)
    r = pir.fixpoint(lb, ub)
    assert r.failed        # var 1 is empty from the start
    lb = np.array([0, 1], np.int32)
    ub = np.array([3, 4], np.int32)
    r = pir.fixpoint(lb, ub)
    assert not r.failed and not r.has_changed
    # aliased operands: x = x + z with z in [1,1] has no solution
    pir2 = pcx.PIR()
    lb = np.array([0, 1], np.int32)
    ub = np.array([10, 1], np.int32)
    pir2.tell([pcx.ADD], [0], [0], [1], lb, ub)
    lbo, ubo = lb.copy(), ub.copy()
    ro = ol.fixpoint(np.array([0], np.int32), np.array([0], np.int32), np.array([0], np.int32),
                     np.array([1], np.int32), lbo, ubo)
    rg = pir2.fixpoint(lb, ub)
    assert bool(rg.failed) == bool(ro.failed)
    # a variable index outside the store is rejected, not read
    pir3 = pcx.PIR()
    pir3.tell([pcx.ADD], [0], [1], [5])
    with pytest.raises(pcx.PcxError) as e:
        pir3.fixpoint(np.zeros(3, np.int32), np.ones(3, np.int32))
    assert e.value.status == pcx.EINVAL
    with pytest.raises(pcx.PcxError):
        pir3.tell([42], [0], [1], [2])
    with pytest.raises(pcx.PcxError) as e:
        pir3.set_config(lockstep=8)
    assert e.value.status == pcx.EINVAL


@pytest.mark.parametrize("lds", [1, 0, -1], ids=["lds", "global", "auto"])
def test_persistent_lds_resident_stores(gpu, lds):
    """Small instances (3,000 variables = 24 KB stores) in the persistent mode
    with the stores staged in LDS or left in global memory."""
    V, P, n_inst = 3000, 12000, 64
    op, x, y, z, lb, ub, w = _network(21, V, P)
    pir = pcx.PIR(lds_mode=lds)
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_lb, root_ub).failed
    slb, sub = pcx.gen_splits(22, root_lb, root_ub, n_inst, 1, 20)
    off, idx = pcx.gen_subsets(23, P, n_inst, 100, P)
    top, tx, ty, tz = pir.table()
    lb_o, ub_o = slb.copy(), sub.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lb_o, ub_o, V, idx=idx.astype(np.int64), off=off, threads=8)
    b = pcx.Batch(pir, n_inst, V, off, idx)
    b.load(slb, sub)
    res_g, _ = b.run(pcx.MODE_PERSISTENT)
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, V, f"persistent lds_mode {lds}")


@pytest.mark.parametrize("V", [12000, 19000])
def test_persistent_lds_stores_beyond_64k(gpu, V):
    """Persistent mode with stores of 96 KB and 152 KB staged in LDS: beyond
    the 64 KB of a workgroup on earlier CDNA, within gfx950's 160 KB."""
    P, n_inst = 4 * V, 48
    op, x, y, z, lb, ub, w = _network(25, V, P)
    pir = pcx.PIR(lds_mode=1)
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_lb, root_ub).failed
    slb, sub = pcx.gen_splits(26, root_lb, root_ub, n_inst, 1, 20)
    top, tx, ty, tz = pir.table()
    lb_o, ub_o = slb.copy(), sub.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lb_o, ub_o, V, threads=8)
    b = pcx.Batch(pir, n_inst, V)
    b.load(slb, sub)
    res_g, _ = b.run(pcx.MODE_PERSISTENT)
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, V, f"persistent LDS store of {V * 8} bytes")


def test_autotuner_choice_is_bit_exact(gpu, tmp_path):
    """The autotuner tries every sweep configuration on the loaded batch and
    leaves the engine configured with the fastest; results stay bit-exact and
    the choice round-trips through the JSON cache."""
    V, P, n_inst = 3000, 12000, 48
    op, x, y, z, lb, ub, w = _network(41, V, P)
    pir = pcx.PIR()
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_lb, root_ub).failed
    slb, sub = pcx.gen_splits(42, root_lb, root_ub, n_inst, 1, 20)
    off, idx = pcx.gen_subsets(43, P, n_inst, 100, P)
    top, tx, ty, tz = pir.table()
    lb_o, ub_o = slb.copy(), sub.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lb_o, ub_o, V, idx=idx.astype(np.int64), off=off, threads=8)
    b = pcx.Batch(pir, n_inst, V, off, idx)
    b.load(slb, sub)
    best, points = b.autotune(reps=1)
    # batched grid + lockstep {16, 32, 64} + persistent x {LDS, global} x {index, copy}
    assert len(points) == 48 + 3 + 4
    assert sorted(p["lockstep"] for p in points if p["lockstep"]) == [16, 32, 64]
    assert all(p["ms"] > 0 for p in points)
    assert best["ms"] == min(p["ms"] for p in points)
    c = pir.get_config()
    assert (c.props_per_wave, c.lds_mode, c.compact_tables, c.lockstep) == \
        (best["props_per_wave"], best["lds_mode"], best["compact_tables"], best["lockstep"])
    res_g, _ = b.run(best["mode"])
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, V, f"autotuned {best}")
    path = str(tmp_path / "tuning.json")
    pcx.save_tuning(path, "k", best, points)
    assert pcx.load_tuning(path, "k")["best"] == best
    pir2 = pcx.PIR()
    assert pcx.apply_tuning(pir2, best) == best["mode"]
    assert pir2.get_config().event_driven == best["event_driven"]
    assert pir2.get_config().lockstep == best["lockstep"]


@pytest.mark.parametrize("mode", [pcx.MODE_PLAIN, pcx.MODE_BATCHED])
def test_marked_sweeps_when_a_work_item_tightens_many_variables(gpu, mode):
    """2,048 independent x_i = y_i + z_i whose x_i all tighten in the first
    sweep: one 2,048-propagator work item tightens more variables than its
    workgroup's mark list holds, so the whole instance is marked for the
    next sweep.  A chain x_0 <= x_1 <= ... reads them afterwards."""
    n = 2048
    V = 3 * n
    op = np.full(n, pcx.ADD, np.int32)
    x = np.arange(n, dtype=np.int32)
    y = (n + np.arange(n)).astype(np.int32)
    z = (2 * n + np.arange(n)).astype(np.int32)
    lb = np.zeros(V, np.int32)
    ub = np.full(V, 1000, np.int32)
    ub[n:] = np.arange(2 * n) % 7             # y, z small: every x_i tightens
    # a LEQ chain x_i <= x_{i+1} over the x's (reified output fixed to 1)
    b = np.arange(V, V + n - 1, dtype=np.int32)
    op2 = np.concatenate([op, np.full(n - 1, pcx.LEQ, np.int32)])
    x2 = np.concatenate([x, b])
    y2 = np.concatenate([y, np.arange(n - 1, dtype=np.int32)])
    z2 = np.concatenate([z, np.arange(1, n, dtype=np.int32)])
    lb = np.concatenate([lb, np.ones(n - 1, np.int32)])
    ub = np.concatenate([ub, np.ones(n - 1, np.int32)])
    tables = (op2, x2, y2, z2)
    lo, hi = lb.copy(), ub.copy()
    top, tx, ty, tz = (a.copy() for a in tables)
    ol.tell_embed(top, tx, lo, hi)
    ol.sort_table(top, tx, ty, tz)
    ro = ol.fixpoint(top, tx, ty, tz, lo, hi)
    pir = pcx.PIR(props_per_wave=512, event_driven=1)
    glb, gub = lb.copy(), ub.copy()
    pir.tell(*tables, glb, gub)
    if mode == pcx.MODE_PLAIN:
        rg = pir.fixpoint(glb, gub)
    else:
        bt = pcx.Batch(pir, 2, len(glb))
        bt.load(np.concatenate([glb, glb]), np.concatenate([gub, gub]))
        res, _ = bt.run(mode)
        rg = res[0]
        l2, u2 = bt.read()
        assert (l2[:len(glb)] == l2[len(glb):]).all() and (u2[:len(glb)] == u2[len(glb):]).all()
        glb, gub = l2[:len(glb)], u2[:len(glb)]
    assert bool(rg.failed) == bool(ro.failed)
    if not ro.failed:
        assert (glb == lo).all() and (gub == hi).all()


@pytest.mark.parametrize("ppw", [1024, 2048])
def test_batched_long_work_items(gpu, ppw):
    """Work items of 2 and 4 rounds of 512 propagators per wave."""
    V, P, n_inst = 5000, 20000, 64
    op, x, y, z, lb, ub, w = _network(51, V, P)
    pir = pcx.PIR(props_per_wave=ppw)
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_lb, root_ub).failed       # plain sweeps cap at 512 per wave
    slb, sub = pcx.gen_splits(52, root_lb, root_ub, n_inst, 1, 20)
    off, idx = pcx.gen_subsets(53, P, n_inst, 200, P)
    top, tx, ty, tz = pir.table()
    lb_o, ub_o = slb.copy(), sub.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lb_o, ub_o, V, idx=idx.astype(np.int64), off=off, threads=8)
    b = pcx.Batch(pir, n_inst, V, off, idx)
    b.load(slb, sub)
    res_g, _ = b.run(pcx.MODE_BATCHED)
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lb_o, ub_o, res_o, V, f"batched ppw {ppw}")


def test_batch_entailment_matches_oracle(gpu):
    """pcx_batch_ask on propagated search nodes: the live (not entailed)
    propagators and the extractable flag of every instance equal the
    oracle's ask (pir.hpp:417-438, 873-884); the live subsets carry the
    whole fixpoint (re-running only them changes nothing); a store fixed to
    the planted witness is extractable."""
    V, P, n_inst = 2000, 8000, 40
    op, x, y, z, lb, ub, w = _network(61, V, P)
    pir = pcx.PIR()
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_lb, root_ub).failed
    slb, sub = pcx.gen_splits(62, root_lb, root_ub, n_inst, 1, 10)
    slb = slb.reshape(n_inst, V)
    sub = sub.reshape(n_inst, V)
    slb[0] = w                                # instance 0: the witness, every variable fixed
    sub[0] = w
    off, idx = pcx.gen_subsets(63, P, n_inst, 100, P)
    b = pcx.Batch(pir, n_inst, V, off, idx)
    b.load(slb.reshape(-1), sub.reshape(-1))
    res, _ = b.run(pcx.MODE_BATCHED)
    lb_g, ub_g = b.read()
    ext, loff, lidx = b.ask()
    top, tx, ty, tz = pir.table()
    for k in range(n_inst):
        lo, hi = lb_g[k * V:(k + 1) * V].copy(), ub_g[k * V:(k + 1) * V].copy()
        props = idx[off[k]:off[k + 1]]
        live = [int(p) for p in props if not ol.ask(top[p], tx[p], ty[p], tz[p], lo, hi)]
        assert list(lidx[loff[k]:loff[k + 1]]) == live, f"instance {k}"
        assert ext[k] == (not res[k].failed and not live), f"instance {k}"
    assert ext[0]
    # the live subsets alone keep every fixpoint: nothing changes when they run
    b2 = pcx.Batch(pir, n_inst, V, loff, lidx)
    b2.load(lb_g, ub_g)
    res2, _ = b2.run(pcx.MODE_BATCHED)
    lb2, ub2 = b2.read()
    for k in range(n_inst):
        if not res[k].failed:
            assert not res2[k].failed and not res2[k].has_changed, f"instance {k}"
    ok = np.repeat(~np.array([bool(r.failed) for r in res]), V)
    assert (lb2[ok] == lb_g[ok]).all() and (ub2[ok] == ub_g[ok]).all()


def _host_children(lo, hi, n_inst, V):
    """Branching restated on the host: first unfixed variable, bisected."""
    lo, hi = lo.reshape(n_inst, V), hi.reshape(n_inst, V)
    var = np.full(n_inst, -1, np.int32)
    clo = np.repeat(lo, 2, axis=0).copy()
    chi = np.repeat(hi, 2, axis=0).copy()
    for k in range(n_inst):
        if (lo[k] > hi[k]).any():
            continue
        free = np.nonzero(lo[k] < hi[k])[0]
        if len(free) == 0:
            continue
        v = int(free[0])
        var[k] = v
        mid = int(lo[k, v]) + (int(hi[k, v]) - int(lo[k, v])) // 2
        chi[2 * k, v] = mid
        clo[2 * k + 1, v] = mid + 1
    return var, clo.reshape(-1), chi.reshape(-1)


@pytest.mark.parametrize("mode", [pcx.MODE_BATCHED, pcx.MODE_PLAIN])
@pytest.mark.parametrize("incremental", [False, True], ids=["full", "from-fixpoint"])
@pytest.mark.parametrize("V", [3000, 2999], ids=["even", "odd"])
def test_branching_on_device(gpu, incremental, mode, V):
    """pcx_batch_branch splits each propagated node on its first unfixed
    variable; the children equal the host restatement and propagate to the
    oracle's fixpoints, also when their run starts from the split variable's
    propagators only (the parents are at their fixpoints).  Stores of an
    even number of variables are copied two variables per access.  Parent 3
    is empty: so are both its children."""
    P, n_inst = 12000, 24
    op, x, y, z, lb, ub, w = _network(71, V, P)
    pir = pcx.PIR()
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_lb, root_ub).failed
    slb, sub = pcx.gen_splits(72, root_lb, root_ub, n_inst, 1, 6)
    slb[3 * V + 5], sub[3 * V + 5] = 4, 3
    off, idx = pcx.gen_subsets(73, P, n_inst, 1000, P)
    parent = pcx.Batch(pir, n_inst, V, off, idx)
    parent.load(slb, sub)
    parent.run(pcx.MODE_BATCHED)
    plo, phi = parent.read()
    # children inherit their parent's subset
    sizes = np.diff(off)
    coff = np.concatenate([[0], np.cumsum(np.repeat(sizes, 2))]).astype(np.int64)
    cidx = np.concatenate([np.concatenate([idx[off[k]:off[k + 1]]] * 2) for k in range(n_inst)])
    child = pcx.Batch(pir, 2 * n_inst, V, coff, cidx)
    var = parent.branch(child, from_fixpoint=incremental)
    hvar, clo, chi = _host_children(plo, phi, n_inst, V)
    assert (var == hvar).all()
    glo, ghi = child.read()
    assert (glo == clo).all() and (ghi == chi).all()
    res, st = child.run(mode)
    lb_g, ub_g = child.read()
    if incremental:                           # no full first sweep
        assert st.full_evals == 0 and st.evals < int(np.diff(coff).sum())
    top, tx, ty, tz = pir.table()
    lo_o, hi_o = clo.copy(), chi.copy()
    res_o = ol.fixpoint_batch(top, tx, ty, tz, lo_o, hi_o, V, idx=cidx.astype(np.int64), off=coff, threads=8)
    _compare(lb_g, ub_g, res, lo_o, hi_o, res_o, V, "children")
    assert res[6].failed and res[7].failed
    child.reset()                             # the children are the child batch's image
    glo2, ghi2 = child.read()
    assert (glo2 == clo).all() and (ghi2 == chi).all()


@pytest.mark.parametrize("queue", [-1, 0, 10], ids=["queue", "bitmaps", "overflow"])
@pytest.mark.parametrize("gi", [16, 64])
def test_queued_marked_sweeps(gpu, gi, queue):
    """Marked sweeps after a lockstep full sweep, from queue planes (the
    default), from the mark bitmaps, or from a queue of 1,024 entries that
    overflows (the run falls back to the bitmaps mid-way): every instance
    equals the oracle, subsets and failures included."""
    V, P, n = 3000, 12000, 200
    op, x, y, z, lb, ub, _ = _network(41 + gi, V, P)
    pir = pcx.PIR(lockstep=gi, event_driven=1, mark_queue=queue, base_seed=0)
    pir.tell(op, x, y, z, lb, ub)
    root_l, root_u = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_l, root_u).failed
    lo, hi = pcx.gen_splits(gi + 7, root_l, root_u, n, 1, 20)
    off, idx = pcx.gen_subsets(gi + 8, P, n, 500, P)
    top, tx, ty, tz = pir.table()
    lo_o, hi_o = lo.copy(), hi.copy()
    ro = ol.fixpoint_batch(top, tx, ty, tz, lo_o, hi_o, V, idx=idx.astype(np.int64), off=off, threads=8)
    b = pcx.Batch(pir, n, V, off, idx)
    b.load(lo, hi)
    for _ in range(2):
        rg, st = b.run(pcx.MODE_BATCHED)
        assert st.lockstep == gi
        lg, ug = b.read()
        _compare(lg, ug, rg, lo_o, hi_o, ro, V, f"queue {queue} lockstep {gi}")
        b.reset()
    with pytest.raises(pcx.PcxError):
        pir.set_config(mark_queue=5)


@pytest.mark.parametrize("queue", [23, 0, 12], ids=["queue", "bitmaps", "overflow"])
def test_queued_marked_sweeps_plain(gpu, queue):
    """Single instances with marked sweeps (plain mode): from queue planes
    (asked for with a capacity of 2^23), from the bitmaps (the default), or
    from a queue of 4,096 entries that overflows; chained networks (long
    chains of tightenings) and a batch of search nodes, every one equal to
    the oracle."""
    V, P = 20000, 80000
    op, x, y, z, lb, ub, _ = _network(77, V, P, chain_frac=0.3, chain_len=200)
    pir = pcx.PIR(event_driven=1, mark_queue=queue)
    g_lb, g_ub = lb.copy(), ub.copy()
    pir.tell(op, x, y, z, g_lb, g_ub)
    root_l, root_u = g_lb.copy(), g_ub.copy()
    r = pir.fixpoint(root_l, root_u)
    ro, o_lb, o_ub = _oracle_fix(op, x, y, z, lb, ub)
    assert bool(r.failed) == bool(ro.failed)
    if not ro.failed:
        assert np.array_equal(root_l, o_lb) and np.array_equal(root_u, o_ub)
    lo, hi = pcx.gen_splits(78, root_l, root_u, 12, 1, 20)
    top, tx, ty, tz = pir.table()
    lo_o, hi_o = lo.copy(), hi.copy()
    ros = ol.fixpoint_batch(top, tx, ty, tz, lo_o, hi_o, V, threads=8)
    b = pcx.Batch(pir, 12, V)
    b.load(lo, hi)
    rg, _ = b.run(pcx.MODE_PLAIN)
    lg, ug = b.read()
    _compare(lg, ug, rg, lo_o, hi_o, ros, V, f"plain queue {queue}")


def _seed_case(seed, V, P, n_inst, subsets):
    op, x, y, z, lb, ub, w = _network(seed, V, P)
    pir = pcx.PIR(event_driven=1)
    pir.tell(op, x, y, z, lb, ub)
    root_lb, root_ub = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_lb, root_ub).failed
    slb, sub = pcx.gen_splits(seed + 1, root_lb, root_ub, n_inst, 1, 20)
    off = idx = None
    if subsets:
        off, idx = pcx.gen_subsets(seed + 2, P, n_inst, 500, P)
    return pir, lb, ub, root_lb, root_ub, slb, sub, off, idx


def _oracle_batch(pir, slb, sub, V, off, idx):
    top, tx, ty, tz = pir.table()
    lo, hi = slb.copy(), sub.copy()
    kw = {} if off is None else {"idx": idx.astype(np.int64), "off": off}
    res = ol.fixpoint_batch(top, tx, ty, tz, lo, hi, V, threads=8, **kw)
    return lo, hi, res


@pytest.mark.parametrize("subsets", [True, False], ids=["subsets", "whole"])
@pytest.mark.parametrize("load", ["host", "device"])
@pytest.mark.parametrize("gi", [0, 64])
def test_first_sweep_seeded_from_instance_zero(gpu, gi, load, subsets):
    """Splits of the root fixpoint share all but a few values with instance
    0: the run evaluates first only the propagators adjacent to a value that
    differs from instance 0's and those instance 0's store does not satisfy
    (no full sweep), and every instance equals the oracle and the same run
    with a full first sweep; after a reset the seeded run repeats.  Loads from
    host and device memory, with and without lockstep copies and subsets;
    instance 5 holds an empty interval."""
    from pcx.device import DeviceArray
    V, P, n_inst = 4000, 16000, 150
    pir, lb, ub, rl, ru, slb, sub, off, idx = _seed_case(81, V, P, n_inst, subsets)
    slb[5 * V + 17], sub[5 * V + 17] = 9, 3                 # instance 5 starts failed
    pir.set_config(lockstep=gi)
    lo_o, hi_o, res_o = _oracle_batch(pir, slb, sub, V, off, idx)
    b = pcx.Batch(pir, n_inst, V, off, idx) if subsets else pcx.Batch(pir, n_inst, V)
    dev = []
    if load == "host":
        b.load(slb, sub)
    else:
        dev = [DeviceArray(slb), DeviceArray(sub)]
        b.load_device(dev[0].ptr, dev[1].ptr)
    pairs = int(off[-1]) if subsets else n_inst * P
    res_g, st = b.run(pcx.MODE_BATCHED)
    assert st.full_evals == 0 and st.evals < pairs // 4, (st.full_evals, st.evals, pairs)
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lo_o, hi_o, res_o, V, f"seeded, lockstep {gi}, {load} load")
    assert res_g[5].failed
    b.reset()
    lb_r, ub_r = b.read()
    assert np.array_equal(lb_r, slb) and np.array_equal(ub_r, sub), "reset after a seeded run"
    res_2, st2 = b.run(pcx.MODE_BATCHED)
    assert st2.full_evals == 0
    lb_2, ub_2 = b.read()
    _compare(lb_2, ub_2, res_2, lo_o, hi_o, res_o, V, "seeded after reset")
    pir.set_config(base_seed=0)
    b.reset()
    res_f, stf = b.run(pcx.MODE_BATCHED)
    assert stf.full_evals > 0
    lb_f, ub_f = b.read()
    _compare(lb_f, ub_f, res_f, lo_o, hi_o, res_o, V, "full first sweep")
    for d in dev:
        d.close()


@pytest.mark.parametrize("case", ["widened", "empty", "base-unsatisfied", "unpropagated"])
def test_seeded_first_sweep_around_the_base_store(gpu, case):
    """The base store holds per variable the value two of instances 0-2
    agree on.  One instance with a variable widened back to its told bounds,
    or holding an empty interval, differs from it there; with that variable
    widened in instances 0-2 the base itself is no fixpoint, and the
    propagators it does not satisfy are seeded in every instance; with every
    instance unpropagated the base leaves too many of them and the run takes
    a full first sweep.  Every instance equals the oracle."""
    V, P, n_inst = 4000, 16000, 120
    pir, lb, ub, rl, ru, slb, sub, off, idx = _seed_case(91, V, P, n_inst, True)
    v = int(np.nonzero((lb != rl) | (ub != ru))[0][0])         # a variable the root tightened
    if case == "widened":
        slb[v], sub[v] = lb[v], ub[v]
    elif case == "empty":
        slb[33], sub[33] = 5, 4
    elif case == "base-unsatisfied":
        for k in range(3):
            slb[k * V + v], sub[k * V + v] = lb[v], ub[v]
    else:
        slb, sub = pcx.gen_splits(92, lb, ub, n_inst, 1, 20)
    lo_o, hi_o, res_o = _oracle_batch(pir, slb, sub, V, off, idx)
    b = pcx.Batch(pir, n_inst, V, off, idx)
    b.load(slb, sub)
    res_g, st = b.run(pcx.MODE_BATCHED)
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lo_o, hi_o, res_o, V, f"base store, {case}")
    if case == "unpropagated":
        assert st.full_evals > 0
    else:
        assert st.full_evals == 0
    if case == "empty":
        assert res_g[0].failed


def test_drained_runs_leave_the_marks_clear(gpu):
    """A drained run takes and clears every mark it evaluates, and clears the
    rows of instances that start failed or fail while draining, so the next
    run on the batch skips clearing the mark planes; its trail of tightened
    variables lets a reset rewrite only those (every reset gives the loaded
    stores back).  Instances 7 and 20 fail
    in their first round (an add whose operands are fixed, and its result
    fixed to another value than their sum), instance 3 starts failed.  Seeded runs after a
    reset, then runs with a full first sweep (and marked sweeps after it) and
    with a sweep limit: every instance equals the oracle each time."""
    V, P, n_inst = 4000, 16000, 60
    pir, lb, ub, rl, ru, slb, sub, off, idx = _seed_case(111, V, P, n_inst, False)
    top, tx, ty, tz = pir.table()
    made = 0
    for p in np.nonzero(top == pcx.ADD)[0]:
        x, y, z = int(tx[p]), int(ty[p]), int(tz[p])
        if len({x, y, z}) < 3:
            continue
        v = int(rl[y]) + int(rl[z]) + 1                       # not the sum
        for k in (7, 20):
            slb[k * V + y] = sub[k * V + y] = rl[y]
            slb[k * V + z] = sub[k * V + z] = rl[z]
            slb[k * V + x] = sub[k * V + x] = v
        made = 1
        break
    assert made
    slb[3 * V + 11], sub[3 * V + 11] = 7, 6
    # instance 12: 1,500 variables back at their told bounds (more tightenings
    # than its trail holds: its restore rewrites the whole row)
    wide = np.nonzero((lb != rl) | (ub != ru))[0][:1500]
    assert len(wide) == 1500
    slb[12 * V + wide], sub[12 * V + wide] = lb[wide], ub[wide]
    lo_o, hi_o, res_o = _oracle_batch(pir, slb, sub, V, None, None)
    assert res_o[7].failed and res_o[20].failed and res_o[3].failed
    b = pcx.Batch(pir, n_inst, V)
    b.load(slb, sub)
    for what, cfg in [("seeded", {}), ("seeded again", {}), ("full first sweep", {"base_seed": 0}),
                      ("swept", {"base_seed": -1, "max_sweeps": 1000}), ("seeded last", {"max_sweeps": 0})]:
        if cfg:
            pir.set_config(**cfg)
        b.reset()
        lb_r, ub_r = b.read()                                  # the reset gives the loaded stores back
        assert np.array_equal(lb_r, slb) and np.array_equal(ub_r, sub), f"reset before {what}"
        res_g, st = b.run(pcx.MODE_BATCHED)
        assert (st.full_evals == 0) == (what != "full first sweep"), what
        lb_g, ub_g = b.read()
        _compare(lb_g, ub_g, res_g, lo_o, hi_o, res_o, V, what)


def _failing_set(pir, lb, ub, rl, ru, slb, sub, V):
    """Instances 7 and 20 fail in their first round (an add whose operands
    are fixed and whose result is fixed to another value), instance 3 starts
    failed, and instance 12 has 1,500 variables back at their told bounds
    (more tightenings than a drained run's trail holds)."""
    top, tx, ty, tz = pir.table()
    slb, sub = slb.copy(), sub.copy()
    for p in np.nonzero(top == pcx.ADD)[0]:
        x, y, z = int(tx[p]), int(ty[p]), int(tz[p])
        if len({x, y, z}) < 3:
            continue
        for k in (7, 20):
            slb[k * V + y] = sub[k * V + y] = rl[y]
            slb[k * V + z] = sub[k * V + z] = rl[z]
            slb[k * V + x] = sub[k * V + x] = int(rl[y]) + int(rl[z]) + 1
        break
    slb[3 * V + 11], sub[3 * V + 11] = 7, 6
    wide = np.nonzero((lb != rl) | (ub != ru))[0][:1500]
    assert len(wide) == 1500
    slb[12 * V + wide], sub[12 * V + wide] = lb[wide], ub[wide]
    return slb, sub


@pytest.mark.parametrize("V", [4000, 3999], ids=["v4", "v1"])
def test_device_loads_write_only_what_differs(gpu, V):
    """A device load into a batch whose stores hold the last image (base
    store plus its complete difference list), or that image but for the
    trail of a drained run from it, puts the stores back to the base and then
    writes only the loaded values that differ from it.  Loads in a row: after
    drained runs (trails with failed instances and one past its cap), after
    a reset, with a new base store (splits of the told store instead of the
    root fixpoint), with random stores (whose differences overflow the list
    when a row's land in one sub-list: the image is then copied and the next
    load writes every value), after a swept run (no trail: every value
    written) and after a host load.  After each load the stores read back
    equal the loaded ones, and every run equals the oracle."""
    from pcx.device import DeviceArray
    n_inst = 60
    pir, lb, ub, rl, ru, slb, sub, off, idx = _seed_case(121, V, 4 * V, n_inst, False)
    rng = np.random.default_rng(125)
    lo_t, hi_t = np.tile(lb, n_inst).astype(np.int64), np.tile(ub, n_inst).astype(np.int64)
    ra, rb = rng.integers(lo_t, hi_t + 1), rng.integers(lo_t, hi_t + 1)
    sets = {"A": _failing_set(pir, lb, ub, rl, ru, slb, sub, V),
            "B": pcx.gen_splits(123, rl, ru, n_inst, 1, 20),
            "C": pcx.gen_splits(124, lb, ub, n_inst, 1, 20),
            "D": (np.minimum(ra, rb).astype(np.int32), np.maximum(ra, rb).astype(np.int32))}
    oracle = {k: _oracle_batch(pir, lo, hi, V, None, None) for k, (lo, hi) in sets.items()}
    assert oracle["A"][2][7].failed and oracle["A"][2][3].failed
    dev = {k: (DeviceArray(lo), DeviceArray(hi)) for k, (lo, hi) in sets.items()}
    b = pcx.Batch(pir, n_inst, V)
    steps = [("A", "drain"), ("B", "drain"), ("A", "drain"), ("B", "reset"), ("C", "drain"),
             ("D", "drain"), ("B", "drain"), ("B", "drain"), ("B", "swept"), ("B", "drain"),
             ("A", "host"), ("B", "drain"), ("C", "drain")]
    try:
        for i, (k, how) in enumerate(steps):
            what = f"step {i}: load {k} ({how})"
            if how == "host":
                b.load(*sets[k])
                b.run(pcx.MODE_BATCHED)
                b.reset()
                continue
            b.load_device(dev[k][0].ptr, dev[k][1].ptr)
            lb_r, ub_r = b.read()
            assert np.array_equal(lb_r, sets[k][0]) and np.array_equal(ub_r, sets[k][1]), what
            pir.set_config(max_sweeps=1000 if how == "swept" else 0)
            res_g, st = b.run(pcx.MODE_BATCHED)
            lb_g, ub_g = b.read()
            lo_o, hi_o, res_o = oracle[k]
            _compare(lb_g, ub_g, res_g, lo_o, hi_o, res_o, V, what)
            if how == "reset":
                b.reset()
                lb_r, ub_r = b.read()
                assert np.array_equal(lb_r, sets[k][0]) and np.array_equal(ub_r, sets[k][1]), what
    finally:
        pir.set_config(max_sweeps=0)
        for d in dev.values():
            d[0].close()
            d[1].close()


@pytest.mark.parametrize("queue", [-1, 0, 10], ids=["queue", "bitmaps", "overflow"])
@pytest.mark.parametrize("limit", [0, 1000], ids=["drain", "swept"])
def test_seeded_first_sweep_marks_from_bitmaps(gpu, queue, limit):
    """A seeded run to the fixpoint drains each instance's marks in one
    launch; with a sweep limit (here one the run never reaches) it sweeps:
    from queue planes, without queues (mark bitmaps only), or from a queue of
    1,024 entries that the seeds overflow (bitmap sweeps from sweep 0 on).
    Every instance equals the oracle."""
    V, P, n_inst = 4000, 16000, 150
    pir, lb, ub, rl, ru, slb, sub, off, idx = _seed_case(101, V, P, n_inst, True)
    pir.set_config(mark_queue=queue, max_sweeps=limit)
    lo_o, hi_o, res_o = _oracle_batch(pir, slb, sub, V, off, idx)
    b = pcx.Batch(pir, n_inst, V, off, idx)
    b.load(slb, sub)
    res_g, st = b.run(pcx.MODE_BATCHED)
    assert st.full_evals == 0
    lb_g, ub_g = b.read()
    _compare(lb_g, ub_g, res_g, lo_o, hi_o, res_o, V, f"seeded, mark_queue {queue}")
