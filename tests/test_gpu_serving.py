"""Serving: each request's result is the same alone and in the mix.

Requests of several networks and sizes are served by the continuous-batching
server (many requests on the GPU at once, admitted as slots and arena memory
free up) and each result is compared bit for bit with the same store
propagated alone (pcx_fixpoint) and with the CPU oracle: every bound when the
store did not fail, the failure flag always.  Admission limits (few slots, an
arena that holds one large store) are forced from the host.
"""
import numpy as np
import pytest

import oracle_lib as ol
import pcx
from pcx.serving import make_trace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    if pcx.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True


def _trace(seed, n_req, **kw):
    """A trace and one engine per network; telling a network embeds its
    reified outputs before the requests' splits are drawn."""
    pirs = []

    def embed(op, x, y, z, lb, ub):
        p = pcx.PIR()
        p.tell(op, x, y, z, lb, ub)
        pirs.append(p)

    return make_trace(seed, n_req, embed=embed, **kw), pirs


@pytest.fixture(scope="module")
def small(gpu):
    tr, pirs = _trace(11, 90, rate=1e6, n_nets=6, pmin=1000, pmax=40000)
    alone = []
    for i in range(len(tr)):
        lo, hi = (a.copy() for a in tr.store(i))
        r = pirs[tr.net[i]].fixpoint(lo, hi)
        alone.append((lo, hi, r))
    return tr, pirs, alone


def _serve(tr, pirs, **cfg):
    srv = pcx.Server(**cfg)
    ids = [srv.add_network(p) for p in pirs]
    assert ids == list(range(len(pirs)))
    stores = [tuple(a.copy() for a in tr.store(i)) for i in range(len(tr))]
    for i, (lo, hi) in enumerate(stores):
        srv.submit(int(tr.net[i]), lo, hi, tag=i)
    got = {}
    steps = 0
    while len(got) < len(tr):
        srv.step()
        for tag, r, lat in srv.poll():
            assert tag not in got
            assert lat >= 0
            got[tag] = r
        steps += 1
        assert steps < 200000, "server made no progress"
    return stores, got, srv.stats()


def _footprint_bytes(nvars, nprops, events):
    """Arena bytes a request takes: its store, plus mark bitmaps with marked sweeps
    (events: 1 always, 0 never, -1 from 2^17 propagators)."""
    units = nvars
    if events == 1 or (events == -1 and nprops >= 1 << 17):
        words = (nprops + 31) // 32
        swords = ((words + 255) // 256 + 31) // 32
        units += (2 * words + 3 * swords + 1) // 2
    return 8 * units


def _same(a_lo, a_hi, a_res, b_lo, b_hi, b_res):
    if bool(a_res.failed) != bool(b_res.failed):
        return False
    if a_res.failed:
        return True
    return (np.array_equal(a_lo, b_lo) and np.array_equal(a_hi, b_hi)
            and bool(a_res.has_changed) == bool(b_res.has_changed))


def test_alone_matches_oracle(small):
    tr, pirs, alone = small
    for i in range(0, len(tr), 3):
        top, tx, ty, tz = pirs[tr.net[i]].table()
        lo, hi = (a.copy() for a in tr.store(i))
        r = ol.fixpoint(top, tx, ty, tz, lo, hi)
        g_lo, g_hi, g = alone[i]
        assert bool(r.failed) == bool(g.failed), f"request {i}"
        if not r.failed:
            assert np.array_equal(lo, g_lo) and np.array_equal(hi, g_hi), f"request {i}"


@pytest.mark.parametrize("events", [1, 0, -1], ids=["marked", "full", "auto"])
def test_mix_equals_alone(small, events):
    tr, pirs, alone = small
    stores, got, st = _serve(tr, pirs, max_slots=64, sweeps_per_tick=4, event_driven=events)
    assert st.max_running > 8, "requests never shared the GPU"
    assert st.completed == len(tr) and st.admitted == len(tr)
    for i in range(len(tr)):
        lo, hi = stores[i]
        assert _same(lo, hi, got[i], *alone[i]), f"request {i} differs in the mix"


@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
def test_few_slots_and_tiny_arena(small, events):
    tr, pirs, alone = small
    biggest = max(_footprint_bytes(len(tr.store(i)[0]), tr.nprops(i), events) for i in range(len(tr)))
    arena = biggest + 8 * 300
    stores, got, st = _serve(tr, pirs, max_slots=3, sweeps_per_tick=1, arena_bytes=arena,
                             event_driven=events)
    assert st.waits_slot > 0 and st.waits_arena > 0
    assert st.arena_peak_bytes <= arena
    for i in range(len(tr)):
        lo, hi = stores[i]
        assert _same(lo, hi, got[i], *alone[i]), f"request {i} differs under admission limits"


@pytest.mark.parametrize("ppw", [64, 512])
def test_ppw_and_many_sweeps_per_tick(small, ppw):
    tr, pirs, alone = small
    stores, got, _ = _serve(tr, pirs, props_per_wave=ppw, sweeps_per_tick=16)
    for i in range(len(tr)):
        assert _same(*stores[i], got[i], *alone[i]), f"request {i}"


def test_replay_continuous_and_plain_agree(small):
    tr, pirs, alone = small
    srv = pcx.Server(max_slots=32)
    for p in pirs:
        srv.add_network(p)
    # spread arrivals over ~20 ms so requests join a running batch
    arr = np.linspace(0.0, 0.02, len(tr))
    outs = {}
    for plain in (False, True):
        lo, hi = tr.lb_all.copy(), tr.ub_all.copy()
        out, wall = srv.replay(arr, tr.net, lo, hi, tr.store_off, plain=plain)
        assert wall > 0
        outs[plain] = (lo, hi, out)
    for i in range(len(tr)):
        s = slice(int(tr.store_off[i]), int(tr.store_off[i + 1]))
        for plain in (False, True):
            lo, hi, out = outs[plain]
            assert out[i].tag == i and out[i].latency_ms >= 0
            assert _same(lo[s], hi[s], out[i].result, *alone[i]), f"request {i} plain={plain}"


def test_empty_store_and_errors(gpu):
    op = np.array([pcx.ADD], np.int32)
    x, y, z = (np.array([v], np.int32) for v in (0, 1, 2))
    lb = np.array([0, 0, 0], np.int32)
    ub = np.array([10, 3, 3], np.int32)
    p = pcx.PIR()
    p.tell(op, x, y, z, lb, ub)
    srv = pcx.Server(max_slots=2, arena_bytes=64)      # full sweeps: 8 B per variable
    n = srv.add_network(p)
    e_lo, e_hi = lb.copy(), ub.copy()
    e_lo[1], e_hi[1] = 5, 4                      # empty interval: fails at once
    srv.submit(n, e_lo, e_hi, tag=7)
    ok_lo, ok_hi = lb.copy(), ub.copy()
    srv.submit(n, ok_lo, ok_hi, tag=8)
    done = {}
    while len(done) < 2:
        srv.step()
        done.update({t: r for t, r, _ in srv.poll()})
    assert done[7].failed and done[7].sweeps == 0
    assert not done[8].failed and list(ok_hi) == [6, 3, 3]
    with pytest.raises(pcx.PcxError) as e:
        srv.submit(n, np.zeros(9, np.int32), np.ones(9, np.int32), tag=9)   # 72 B > 64 B arena
    assert e.value.status == pcx.EINVAL
    with pytest.raises(pcx.PcxError) as e:
        srv.submit(n, lb[:2].copy(), ub[:2].copy(), tag=10)                 # too few variables
    assert e.value.status == pcx.EINVAL
    with pytest.raises(pcx.PcxError):
        srv.submit(5, lb.copy(), ub.copy(), tag=11)                         # unknown network


def test_auto_marks_large_requests_only(gpu):
    """Automatic mode: requests of at least 2^17 propagators run marked sweeps,
    smaller ones full sweeps, in the same slot batch; every result equals the
    request propagated alone."""
    tr, pirs = _trace(35, 30, rate=1e6, n_nets=4, pmin=2000, pmax=300000)
    sizes = [p.num_deductions() for p in pirs]
    used = set(tr.net.tolist())
    assert any(sizes[i] >= 1 << 17 for i in used) and any(sizes[i] < 1 << 17 for i in used)
    alone = []
    for i in range(len(tr)):
        lo, hi = (a.copy() for a in tr.store(i))
        alone.append((lo, hi, pirs[tr.net[i]].fixpoint(lo, hi)))
    stores, got, st = _serve(tr, pirs, max_slots=64, sweeps_per_tick=4)
    assert st.max_running > 4
    for i in range(len(tr)):
        assert _same(*stores[i], got[i], *alone[i]), f"request {i} ({sizes[tr.net[i]]} propagators)"
        if sizes[tr.net[i]] >= 1 << 17 and not got[i].failed:
            assert got[i].evals < got[i].sweeps * sizes[tr.net[i]]   # marked sweeps did less work
