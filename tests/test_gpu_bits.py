"""GPU parity for the bitset domain: every word equals the oracle's true
fixpoint (oracle/bits_oracle.c, restating PC over NBitset), and the failure
flag matches.  All calls go through the C ABI."""
import numpy as np
import pytest

import oracle_lib as ol
import pcx
from test_oracle_bits import load_bits_cases, run_bits_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    if pcx.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True


@pytest.mark.parametrize("case", load_bits_cases(), ids=lambda c: c["name"])
def test_bits_kat_on_gpu(gpu, case):
    def fx(op, x, y, z, w):
        pir = pcx.BitsPIR()
        pir.tell(op, x, y, z)
        r = pir.fixpoint(w)
        return r.failed, r.has_changed
    errs = run_bits_case(case, fx)
    assert not errs, f"{case['name']} ({case['ref']}): " + "; ".join(errs)


def _compare(wg, rg, wo, ro, nvars, what):
    n = len(ro)
    wg, wo = wg.reshape(n, nvars), wo.reshape(n, nvars)
    fg = np.array([r.failed for r in rg], bool)
    fo = np.array([r.failed for r in ro], bool)
    bad = np.nonzero(fg != fo)[0]
    assert len(bad) == 0, f"{what}: failure flag differs on {len(bad)} instances, first {bad[:5]}"
    diff = (wg != wo).any(axis=1) & ~fo
    bad = np.nonzero(diff)[0]
    assert len(bad) == 0, f"{what}: words differ on {len(bad)} instances, first {bad[0]}: {wg[bad[0]]} vs {wo[bad[0]]}"


@pytest.mark.parametrize("name", list(ol.BITS_OPS))
@pytest.mark.parametrize("mode", [pcx.MODE_BATCHED, pcx.MODE_PERSISTENT, pcx.MODE_PLAIN])
def test_every_domain_one_propagator(gpu, name, mode):
    U = 5 if mode != pcx.MODE_PLAIN else 3
    doms = np.arange(1, 1 << U, dtype=np.uint64)
    xs = np.array([1, 2, 3], np.uint64) if name == "EQR" else doms
    X, Y, Z = np.meshgrid(xs, doms, doms, indexing="ij")
    words = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.uint64).ravel()
    n = len(words) // 3
    op = np.array([ol.BITS_OPS[name]], np.int32)
    x, y, z = (np.array([v], np.int32) for v in (0, 1, 2))
    wo = words.copy()
    ro = ol.bits_fixpoint_batch(op, x, y, z, wo, 3, threads=8)
    pir = pcx.BitsPIR()
    pir.tell(op, x, y, z)
    b = pcx.Batch(pir, n, 3)
    b.load_words(words)
    rg, _ = b.run(mode)
    _compare(b.read_words(), rg, wo, ro, 3, f"{name} mode {mode}")


@pytest.mark.parametrize("seed,V,P", [(1, 5000, 20000), (2, 20000, 80000), (3, 200000, 800000)])
def test_bits_network(gpu, seed, V, P):
    op, x, y, z, w = pcx.gen_bits_network(seed, V, P)
    pir = pcx.BitsPIR()
    wg = w.copy()
    pir.tell(op, x, y, z, wg)
    rg = pir.fixpoint(wg)
    top, tx, ty, tz = pir.table()
    wo = w.copy()
    ol.bits_tell_embed(op, x, wo)
    ro = ol.bits_fixpoint(top, tx, ty, tz, wo)
    assert bool(rg.failed) == bool(ro.failed)
    if not ro.failed:
        assert (wg == wo).all()


@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
@pytest.mark.parametrize("stream", [0, 1])
@pytest.mark.parametrize("mode", [pcx.MODE_BATCHED, pcx.MODE_PERSISTENT, pcx.MODE_PLAIN])
def test_bits_batch_of_splits(gpu, mode, stream, events):
    V, P, n = 4000, 16000, 64
    op, x, y, z, w = pcx.gen_bits_network(9, V, P)
    pir = pcx.BitsPIR(stream_tables=stream, event_driven=events)
    pir.tell(op, x, y, z, w)
    root = w.copy()
    r = pir.fixpoint(root)
    assert not r.failed
    inst = pcx.gen_bits_splits(10, root, n, 1, 40)
    top, tx, ty, tz = pir.table()
    wo = inst.copy()
    ro = ol.bits_fixpoint_batch(top, tx, ty, tz, wo, V, threads=8)
    b = pcx.Batch(pir, n, V)
    b.load_words(inst)
    rg, _ = b.run(mode)
    _compare(b.read_words(), rg, wo, ro, V, f"bits batch mode {mode}")


@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
@pytest.mark.parametrize("gi", [16, 32, 64])
def test_bits_lockstep(gpu, gi, events):
    """Bitset full sweeps in lockstep groups of gi over 100 instances."""
    V, P, n = 4000, 16000, 100
    op, x, y, z, w = pcx.gen_bits_network(61, V, P)
    pir = pcx.BitsPIR(event_driven=events, lockstep=gi, base_seed=0)
    pir.tell(op, x, y, z, w)
    root = w.copy()
    assert not pir.fixpoint(root).failed
    inst = pcx.gen_bits_splits(62, root, n, 1, 40)
    top, tx, ty, tz = pir.table()
    wo = inst.copy()
    ro = ol.bits_fixpoint_batch(top, tx, ty, tz, wo, V, threads=8)
    b = pcx.Batch(pir, n, V)
    b.load_words(inst)
    rg, st = b.run(pcx.MODE_BATCHED)
    assert st.lockstep == gi
    _compare(b.read_words(), rg, wo, ro, V, f"bits lockstep {gi}")


def _bits_seed_case(seed, V, P, n):
    op, x, y, z, w = pcx.gen_bits_network(seed, V, P)
    pir = pcx.BitsPIR()
    pir.tell(op, x, y, z, w)
    root = w.copy()
    assert not pir.fixpoint(root).failed
    return pir, w, root, pcx.gen_bits_splits(seed + 1, root, n, 1, 40)


@pytest.mark.parametrize("case", ["splits", "empty", "base-unsatisfied", "unpropagated"])
@pytest.mark.parametrize("gi", [0, 64])
def test_bits_seeded_first_sweep(gpu, gi, case):
    """Bit-removal splits of the root fixpoint share all but a few words with
    the base store (per word the value two of instances 0-2 hold): the first
    sweep evaluates only the propagators adjacent to a differing word and
    those the base does not satisfy, and the marks drain in one launch.  One
    instance holding an empty word, the base widened back to a told word in
    instances 0-2 (not a fixpoint), and unpropagated stores (too many
    unsatisfied propagators: a full first sweep).  Every instance equals the
    oracle, after a reset too, and the run with a full first sweep agrees."""
    V, P, n = 4000, 16000, 150
    pir, w, root, inst = _bits_seed_case(71, V, P, n)
    v = int(np.nonzero(w != root)[0][0])                   # a word the root narrowed
    if case == "empty":
        inst[33] = 0
    elif case == "base-unsatisfied":
        for k in range(3):
            inst[k * V + v] = w[v]
    elif case == "unpropagated":
        inst = pcx.gen_bits_splits(72, w, n, 1, 40)
    pir.set_config(lockstep=gi)
    top, tx, ty, tz = pir.table()
    wo = inst.copy()
    ro = ol.bits_fixpoint_batch(top, tx, ty, tz, wo, V, threads=8)
    b = pcx.Batch(pir, n, V)
    b.load_words(inst)
    rg, st = b.run(pcx.MODE_BATCHED)
    _compare(b.read_words(), rg, wo, ro, V, f"bits seeded, {case}, lockstep {gi}")
    if case == "unpropagated":
        assert st.full_evals > 0
        return
    assert st.full_evals == 0 and st.evals < n * P // 4, (st.full_evals, st.evals)
    if case == "empty":
        assert rg[0].failed
    b.reset()
    assert np.array_equal(b.read_words(), inst), "reset after a seeded run"
    r2, st2 = b.run(pcx.MODE_BATCHED)
    assert st2.full_evals == 0
    _compare(b.read_words(), r2, wo, ro, V, "bits seeded after reset")
    pir.set_config(base_seed=0)
    b.reset()
    rf, stf = b.run(pcx.MODE_BATCHED)
    assert stf.full_evals > 0
    _compare(b.read_words(), rf, wo, ro, V, "bits full first sweep")


@pytest.mark.parametrize("queue", [-1, 0, 10], ids=["queue", "bitmaps", "overflow"])
def test_bits_seeded_swept(gpu, queue):
    """With a sweep limit (one the run never reaches) a seeded bitset run
    sweeps instead of draining: from queue planes, from the mark bitmaps
    alone, or from a queue the seeds overflow.  Every instance equals the
    oracle."""
    V, P, n = 4000, 16000, 150
    pir, w, root, inst = _bits_seed_case(75, V, P, n)
    pir.set_config(mark_queue=queue, max_sweeps=1000)
    top, tx, ty, tz = pir.table()
    wo = inst.copy()
    ro = ol.bits_fixpoint_batch(top, tx, ty, tz, wo, V, threads=8)
    b = pcx.Batch(pir, n, V)
    b.load_words(inst)
    rg, st = b.run(pcx.MODE_BATCHED)
    assert st.full_evals == 0
    _compare(b.read_words(), rg, wo, ro, V, f"bits seeded, mark_queue {queue}")


@pytest.mark.parametrize("lds", [1, 0])
def test_bits_persistent_lds(gpu, lds):
    V, P, n = 2000, 8000, 48
    op, x, y, z, w = pcx.gen_bits_network(31, V, P)
    pir = pcx.BitsPIR(lds_mode=lds)
    pir.tell(op, x, y, z, w)
    root = w.copy()
    assert not pir.fixpoint(root).failed
    inst = pcx.gen_bits_splits(32, root, n, 1, 40)
    top, tx, ty, tz = pir.table()
    wo = inst.copy()
    ro = ol.bits_fixpoint_batch(top, tx, ty, tz, wo, V, threads=8)
    b = pcx.Batch(pir, n, V)
    b.load_words(inst)
    rg, _ = b.run(pcx.MODE_PERSISTENT)
    _compare(b.read_words(), rg, wo, ro, V, f"bits persistent lds {lds}")


@pytest.mark.parametrize("V", [1000, 999], ids=["even", "odd"])
def test_bits_branching(gpu, V):
    P, n = 4000, 16
    op, x, y, z, w = pcx.gen_bits_network(81, V, P)
    pir = pcx.BitsPIR()
    pir.tell(op, x, y, z, w)
    root = w.copy()
    assert not pir.fixpoint(root).failed
    inst = pcx.gen_bits_splits(82, root, n, 1, 10)
    parent = pcx.Batch(pir, n, V)
    parent.load_words(inst)
    parent.run(pcx.MODE_BATCHED)
    pw = parent.read_words().reshape(n, V)
    child = pcx.Batch(pir, 2 * n, V)
    var = parent.branch(child)
    cw = child.read_words().reshape(2 * n, V)
    for k in range(n):
        words = pw[k]
        multi = [v for v in range(V) if int(words[v]) & (int(words[v]) - 1)]
        expect = -1 if (words == 0).any() or not multi else multi[0]
        assert var[k] == expect
        left, right = words.copy(), words.copy()
        if expect >= 0:
            wv = int(words[expect])
            lowbit = wv & -wv
            left[expect] = np.uint64(lowbit)
            right[expect] = np.uint64(wv & ~lowbit)
        assert (cw[2 * k] == left).all() and (cw[2 * k + 1] == right).all()
    rg, _ = child.run(pcx.MODE_BATCHED)
    top, tx, ty, tz = pir.table()
    wo = cw.reshape(-1).copy()
    ro = ol.bits_fixpoint_batch(top, tx, ty, tz, wo, V, threads=8)
    _compare(child.read_words(), rg, wo, ro, V, "bits children")


def test_bits_full_size_batch_of_1024(gpu):
    """The bitset workload at its full size: 1,024 bit-removal subproblems of
    one 200,000-word / 800,000-propagator network (values inside [0, 61]),
    run batched with the automatic lockstep full sweep; every instance equals
    the oracle."""
    V, P, n = 200000, 800000, 1024
    op, x, y, z, w = pcx.gen_bits_network(1, V, P)
    pir = pcx.BitsPIR()
    pir.tell(op, x, y, z, w)
    root = w.copy()
    assert not pir.fixpoint(root).failed
    inst = pcx.gen_bits_splits(2, root, n, 1, 20)
    assert not (inst & np.uint64((1 << 62) | (1 << 63))).any()
    b = pcx.Batch(pir, n, V)
    b.load_words(inst)
    rg, st = b.run(pcx.MODE_BATCHED)
    wg = b.read_words()
    b.close()
    top, tx, ty, tz = pir.table()
    ro = ol.bits_fixpoint_batch(top, tx, ty, tz, inst, V, threads=16)
    _compare(wg, rg, inst, ro, V, "bits 1,024 x 200k/800k")
