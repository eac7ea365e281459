"""GPU tests of the engine's API surface beyond the fixpoint itself:

- the table order after incremental tells equals the reference's stable sort
  by (op, y, x, z) of everything told (pir.hpp:327-352), as the oracle
  restates it;
- snapshot / restore pop the last entries of the sorted table, as the
  reference does (pir.hpp:857-870), and the fixpoint over the table left
  behind equals the oracle's;
- a configuration change between runs (max_sweeps) is obeyed by the
  captured graph of marked sweeps.

All calls go through the C ABI (lala-pc_amd/pcx/libpcx.so).
"""
import numpy as np
import pytest

import oracle_lib as ol
import pcx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    if pcx.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True


def _blocks(seed, V, sizes):
    """Blocks of random propagators over V variables (all ten opcodes,
    repeated and aliased operands included), told one block at a time."""
    rng = np.random.default_rng(seed)
    out = []
    for n in sizes:
        op = rng.integers(0, len(pcx.OPS), n).astype(np.int32)
        x, y, z = (rng.integers(0, V, n).astype(np.int32) for _ in range(3))
        dup = rng.integers(0, n, n // 8)          # repeated bytecodes (none in tiny blocks)
        if len(dup) > 1:
            op[dup[1:]], x[dup[1:]], y[dup[1:]], z[dup[1:]] = op[dup[0]], x[dup[0]], y[dup[0]], z[dup[0]]
        out.append((op, x, y, z))
    return out


def _oracle_order(blocks):
    """The reference's table after telling the blocks in turn: each tell
    appends and stable-sorts the whole table (pir.hpp:330-347)."""
    t = [np.zeros(0, np.int32) for _ in range(4)]
    for b in blocks:
        t = [np.concatenate([t[i], b[i]]) for i in range(4)]
        ol.sort_table(*t)
    return t


def _store(seed, V):
    rng = np.random.default_rng(seed)
    lb = rng.integers(-50, 50, V).astype(np.int32)
    ub = (lb + rng.integers(0, 60, V)).astype(np.int32)
    return lb, ub


@pytest.mark.parametrize("sizes", [[300], [200, 150, 1], [1, 1, 1, 400]])
def test_table_order_matches_reference_sort(gpu, sizes):
    blocks = _blocks(sum(sizes), 40, sizes)
    pir = pcx.PIR()
    for b in blocks:
        pir.tell(*b)
    got = pir.table()
    want = _oracle_order(blocks)
    for g, w, name in zip(got, want, "op x y z".split()):
        assert np.array_equal(g, w), f"table column {name} differs from the reference order"
    assert pir.num_deductions() == sum(sizes)


def test_snapshot_restore_pops_the_sorted_end(gpu):
    """Tell A, snapshot, tell B, propagate; restore; propagate again.  After
    the restore the table is the first |A| entries of sort(A + B) -- the
    reference pops from the end of its sorted vector -- and each fixpoint
    equals the oracle's over the same table."""
    V = 60
    A, B = _blocks(5, V, [250, 120])
    lb0, ub0 = _store(6, V)
    pir = pcx.PIR()
    lb, ub = lb0.copy(), ub0.copy()
    pir.tell(*A, lb, ub)
    snap = pir.snapshot()
    assert snap == len(A[0])
    pir.tell(*B, lb, ub)
    full = _oracle_order([A, B])
    assert all(np.array_equal(g, w) for g, w in zip(pir.table(), full))

    def both(table, lb_in, ub_in):
        lg, ug = lb_in.copy(), ub_in.copy()
        rg = pir.fixpoint(lg, ug)
        lo, uo = lb_in.copy(), ub_in.copy()
        ro = ol.fixpoint(*[c.copy() for c in table], lo, uo)
        assert bool(rg.failed) == bool(ro.failed)
        if not ro.failed:
            assert np.array_equal(lg, lo) and np.array_equal(ug, uo)
            assert bool(rg.has_changed) == bool(ro.has_changed)
        return ro

    # the embedded store (reified outputs in [0,1]) with wide domains so that
    # propagation does not fail at once
    lbw, ubw = np.full(V, -1000, np.int32), np.full(V, 1000, np.int32)
    ol.tell_embed(full[0], full[1], lbw, ubw)
    both(full, lbw, ubw)
    pir.restore(snap)
    kept = [c[:snap] for c in full]
    got = pir.table()
    assert pir.num_deductions() == snap
    assert all(np.array_equal(g, w) for g, w in zip(got, kept))
    both(kept, lbw, ubw)
    # a tell after the restore sorts into the remaining table
    C = _blocks(8, V, [40])[0]
    pir.tell(*C)
    merged = [np.concatenate([kept[i], C[i]]) for i in range(4)]
    ol.sort_table(*merged)
    assert all(np.array_equal(g, w) for g, w in zip(pir.table(), merged))
    both(merged, lbw, ubw)
    with pytest.raises(pcx.PcxError) as e:
        pir.restore(pir.num_deductions() + 1)
    assert e.value.status == pcx.EINVAL


def _climbers(n, seed, cap_lo, cap_hi):
    """n gadgets a = b + 1, t = a + 1, b = min(t, cap): the lower bounds
    climb until b reaches cap, a few links per sweep."""
    rng = np.random.default_rng(seed)
    op, x, y, z, lb, ub = [], [], [], [], [], []
    for g in range(n):
        a, b, t, one, cap = range(5 * g, 5 * g + 5)
        K = int(rng.integers(cap_lo, cap_hi))
        lb += [0, 0, 0, 1, K]
        ub += [10**6] * 3 + [1, K]
        op += [pcx.ADD, pcx.ADD, pcx.MIN]
        x += [a, t, b]
        y += [b, a, t]
        z += [one, one, cap]
    return [np.array(v, np.int32) for v in (op, x, y, z, lb, ub)]


def test_max_sweeps_change_between_runs(gpu):
    """Batched mode replays a captured graph for the marked sweeps after the
    first group of 8.  A run with max_sweeps = 0 (capture), then 16, then 0
    again on the same batch: the limited run stops every instance within 16
    sweeps, short of the fixpoint (the caps need far more), with bounds
    between the start and the fixpoint; the unlimited runs equal the
    oracle."""
    op, x, y, z, lb, ub = _climbers(50, 3, 1000, 1500)
    V, n = len(lb), 64
    pir = pcx.PIR(event_driven=1, lockstep=0)
    pir.tell(op, x, y, z, lb, ub)
    rng = np.random.default_rng(4)
    lo = np.tile(lb, n)
    hi = np.tile(ub, n)
    caps = rng.integers(1000, 1500, (n, V // 5)).astype(np.int32)
    lo.reshape(n, V)[:, 4::5] = caps
    hi.reshape(n, V)[:, 4::5] = caps
    top, tx, ty, tz = pir.table()
    lo_o, hi_o = lo.copy(), hi.copy()
    ro = ol.fixpoint_batch(top, tx, ty, tz, lo_o, hi_o, V, threads=8)
    assert not any(r.failed for r in ro)
    b = pcx.Batch(pir, n, V)
    b.load(lo, hi)

    def run_and_read(max_sweeps):
        pir.set_config(max_sweeps=max_sweeps)
        b.reset()
        res, _ = b.run(pcx.MODE_BATCHED)
        lg, ug = b.read()
        return res, lg, ug

    for max_sweeps in (0, 16, 0, 16):
        res, lg, ug = run_and_read(max_sweeps)
        if max_sweeps == 0:
            assert np.array_equal(lg, lo_o) and np.array_equal(ug, hi_o)
            assert max(r.sweeps for r in res) > 16
            continue
        assert max(r.sweeps for r in res) <= 16
        assert (lg >= lo).all() and (ug <= hi).all(), "bounds loosened"
        assert (lg <= lo_o).all() and (ug >= hi_o).all(), "tighter than the fixpoint"
        short = ((lg != lo_o) | (ug != hi_o)).reshape(n, V).any(axis=1)
        assert short.all(), f"{int((~short).sum())} instances reached the fixpoint within 16 sweeps"


@pytest.mark.parametrize("nvars", [50, 37])
def test_load_from_device_memory(gpu, nvars):
    """pcx_batch_load_device: stores staged in device memory (16-byte
    aligned and, with an odd variable count over an odd instance count, not)
    give the same results as a host load; the reset image is the device
    input; empty stores fail at once."""
    from pcx.device import DeviceArray
    rng = np.random.default_rng(nvars)
    P, n = 4 * nvars, 33
    op = rng.integers(0, 10, P).astype(np.int32)
    x, y, z = (rng.integers(0, nvars, P).astype(np.int32) for _ in range(3))
    pir = pcx.PIR()
    pir.tell(op, x, y, z)
    lo = rng.integers(-20, 20, n * nvars).astype(np.int32)
    hi = (lo + rng.integers(0, 30, n * nvars)).astype(np.int32)
    for b in pir.table()[1][pir.table()[0] >= pcx.EQ]:
        lo[b::nvars], hi[b::nvars] = np.maximum(lo[b::nvars], 0), np.minimum(hi[b::nvars], 1)
    lo[3 * nvars + 1], hi[3 * nvars + 1] = 5, 4          # instance 3 starts empty
    host = pcx.Batch(pir, n, nvars)
    host.load(lo, hi)
    r_h, _ = host.run(pcx.MODE_BATCHED)
    l_h, u_h = host.read()
    # one allocation holding lb then ub: ub starts 4 * n * nvars bytes in,
    # which is not 16-byte aligned when n * nvars is odd
    both = DeviceArray(np.concatenate([lo, hi]))
    dev = pcx.Batch(pir, n, nvars)
    dev.load_device(both.ptr, both.ptr + 4 * n * nvars)
    both.close()
    for _ in range(2):                                  # the run, then the reset image
        r_d, _ = dev.run(pcx.MODE_BATCHED)
        l_d, u_d = dev.read()
        assert [bool(r.failed) for r in r_d] == [bool(r.failed) for r in r_h]
        assert r_d[3].failed
        ok = np.repeat(~np.array([bool(r.failed) for r in r_h]), nvars)
        assert np.array_equal(l_d[ok], l_h[ok]) and np.array_equal(u_d[ok], u_h[ok])
        dev.reset()


@pytest.mark.parametrize("V", [700, 701], ids=["vector", "scalar"])
@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
@pytest.mark.parametrize("gi", [16, 32, 64])
def test_device_load_into_the_lockstep_copy(gpu, gi, events, V):
    """A device load into a lockstep batch writes the stores and the
    interleaved copy in one pass, and the reset image lives in that copy
    until a run would overwrite it.  Every sequence -- reset right after the
    load, run / run again without a reset / reset / run, a host load after
    it -- gives the host-loaded batch's results; 131 instances leave a
    partial last group; 700 variables take the 16-byte path, 701 the scalar
    one."""
    from pcx.device import DeviceArray
    P, n = 4 * V, 131
    op, x, y, z, lb, ub, _ = pcx.gen_network(pcx.gen_params(seed=gi + events, nvars=V, nprops=P))
    pir = pcx.PIR(lockstep=gi, event_driven=events)
    pir.tell(op, x, y, z, lb, ub)
    root_l, root_u = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_l, root_u).failed
    lo, hi = pcx.gen_splits(gi, root_l, root_u, n, 1, 20)
    off, idx = pcx.gen_subsets(gi + 1, P, n, 100, P)
    lo[5 * V + 3], hi[5 * V + 3] = 9, 8                        # instance 5 starts empty
    ref = pcx.Batch(pir, n, V, off, idx)
    ref.load(lo, hi)
    r_ref, st = ref.run(pcx.MODE_BATCHED)
    assert st.lockstep == gi
    l_ref, u_ref = ref.read()
    fail = np.array([bool(r.failed) for r in r_ref])
    ok = np.repeat(~fail, V)

    def same(b, what):
        res, _ = b.run(pcx.MODE_BATCHED)
        lg, ug = b.read()
        assert np.array_equal(np.array([bool(r.failed) for r in res]), fail), what
        assert np.array_equal(lg[ok], l_ref[ok]) and np.array_equal(ug[ok], u_ref[ok]), what

    d_lo, d_hi = DeviceArray(lo), DeviceArray(hi)
    b = pcx.Batch(pir, n, V, off, idx)
    b.load_device(d_lo.ptr, d_hi.ptr)
    b.reset()                                                  # image from the interleaved copy
    l0, u0 = b.read()
    assert np.array_equal(l0, lo) and np.array_equal(u0, hi)
    same(b, "after the load and a reset")
    b.run(pcx.MODE_BATCHED)                                    # from the fixpoint: rewrites the copy
    b.reset()
    l0, u0 = b.read()
    assert np.array_equal(l0, lo) and np.array_equal(u0, hi), "reset image lost"
    same(b, "after a run without reset and a reset")
    b.load_device(d_lo.ptr, d_hi.ptr)
    same(b, "straight after a second device load")
    b.load(lo, hi)                                             # host load: the image is d_image again
    b.run(pcx.MODE_PLAIN)
    b.reset()
    same(b, "after a host load")
    d_lo.close()
    d_hi.close()


@pytest.mark.parametrize("gi", [0, 64])
def test_device_loads_switch_between_image_forms(gpu, gi):
    """A device load whose stores share most values writes the stores only
    (its reset image is the base store plus the listed differences); one
    whose stores do not overflows the list and copies the image from the
    stores, and the loads after it write the image in the ingest pass until
    a load of shared stores again.  Along a sequence of unshared and shared
    loads, every run equals the same stores loaded from the host, and every
    reset gives back the loaded stores."""
    from pcx.device import DeviceArray
    V, P, n = 700, 2800, 800                 # unshared loads overflow sub-lists of 2,048 entries
    op, x, y, z, lb, ub, _ = pcx.gen_network(pcx.gen_params(seed=7 + gi, nvars=V, nprops=P))
    pir = pcx.PIR(lockstep=gi)
    pir.tell(op, x, y, z, lb, ub)
    root_l, root_u = lb.copy(), ub.copy()
    assert not pir.fixpoint(root_l, root_u).failed
    shared = pcx.gen_splits(gi + 3, root_l, root_u, n, 1, 20)
    rng = np.random.default_rng(gi)
    a = rng.integers(np.repeat(lb, n).reshape(V, n).T.ravel(), np.repeat(ub, n).reshape(V, n).T.ravel() + 1)
    b_ = rng.integers(np.repeat(lb, n).reshape(V, n).T.ravel(), np.repeat(ub, n).reshape(V, n).T.ravel() + 1)
    unshared = (np.minimum(a, b_).astype(np.int32), np.maximum(a, b_).astype(np.int32))

    def reference(lo, hi):
        r = pcx.Batch(pir, n, V)
        r.load(lo, hi)
        res, _ = r.run(pcx.MODE_BATCHED)
        lg, ug = r.read()
        return np.array([bool(x.failed) for x in res]), lg, ug

    refs = {"shared": reference(*shared), "unshared": reference(*unshared)}
    data = {"shared": shared, "unshared": unshared}
    batch = pcx.Batch(pir, n, V)
    for kind in ["unshared", "unshared", "shared", "shared", "unshared", "shared"]:
        lo, hi = data[kind]
        d_lo, d_hi = DeviceArray(lo), DeviceArray(hi)
        batch.load_device(d_lo.ptr, d_hi.ptr)
        d_lo.close()
        d_hi.close()
        for rep in range(2):                     # the run, then the run after a reset
            res, _ = batch.run(pcx.MODE_BATCHED)
            lg, ug = batch.read()
            fail, l_ref, u_ref = refs[kind]
            assert np.array_equal(np.array([bool(x.failed) for x in res]), fail), (kind, rep)
            ok = np.repeat(~fail, V)
            assert np.array_equal(lg[ok], l_ref[ok]) and np.array_equal(ug[ok], u_ref[ok]), (kind, rep)
            batch.reset()
            l0, u0 = batch.read()
            assert np.array_equal(l0, lo) and np.array_equal(u0, hi), ("reset", kind, rep)


def test_batch_outlives_its_engine(gpu):
    """A garbage collector may free an engine before a batch over it (a
    reference cycle through a test's traceback does): destroying the batch
    then reads nothing of the engine, and the device stays usable."""
    op = np.array([pcx.ADD], np.int32)
    x, y, z = (np.array([v], np.int32) for v in (0, 1, 2))
    lb = np.array([0, 0, 0], np.int32)
    ub = np.array([10, 3, 3], np.int32)
    pir = pcx.PIR()
    pir.tell(op, x, y, z, lb.copy(), ub.copy())
    b = pcx.Batch(pir, 2, 3)
    b.load(np.tile(lb, 2), np.tile(ub, 2))
    pir.close()                                  # the engine first
    b.close()
    p2 = pcx.PIR()
    lo, hi = lb.copy(), ub.copy()
    p2.tell(op, x, y, z, lo, hi)
    assert not p2.fixpoint(lo, hi).failed and list(hi) == [6, 3, 3]
