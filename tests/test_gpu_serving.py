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


def _serve(tr, pirs, roots=False, extra=(), **cfg):
    """Serve every request of the trace (and `extra` (net, lb, ub) requests,
    tagged after them); roots: give each network its told store."""
    srv = pcx.Server(**cfg)
    ids = [srv.add_network(p) for p in pirs]
    assert ids == list(range(len(pirs)))
    if roots:
        for j, net in enumerate(tr.nets):
            srv.set_root(j, net[4], net[5])
    stores = [tuple(a.copy() for a in tr.store(i)) for i in range(len(tr))]
    for j, lo, hi in extra:
        stores.append((lo.copy(), hi.copy()))
    nets = list(tr.net) + [j for j, _, _ in extra]
    for i, (lo, hi) in enumerate(stores):
        srv.submit(int(nets[i]), lo, hi, tag=i)
    got = {}
    steps = 0
    while len(got) < len(stores):
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


@pytest.mark.parametrize("events", [1, 0, -1], ids=["marked", "full", "auto"])
def test_requests_start_from_the_root_fixpoint(small, events):
    """With each network's told store given (pcx_server_set_root), a request
    inside it starts from its meet with the root fixpoint: the results are
    the same as alone (stores, failure, has_changed), in fewer sweeps.
    Requests that leave the told store (a bound widened by one) start from
    their own store and match too."""
    tr, pirs, alone = small
    extra, extra_alone = [], []
    for i in range(0, len(tr), 9):
        j = int(tr.net[i])
        lo, hi = (a.copy() for a in tr.store(i))
        v = int(np.nonzero(lo == tr.nets[j][4])[0][0])
        lo[v] -= 1                                        # below the told lower bound
        extra.append((j, lo, hi))
        a_lo, a_hi = lo.copy(), hi.copy()
        extra_alone.append((a_lo, a_hi, pirs[j].fixpoint(a_lo, a_hi)))
    stores, got, st = _serve(tr, pirs, roots=True, extra=extra, max_slots=64, sweeps_per_tick=4,
                             event_driven=events)
    assert st.completed == len(tr) + len(extra)
    sweeps_root = 0
    for i in range(len(tr)):
        lo, hi = stores[i]
        assert _same(lo, hi, got[i], *alone[i]), f"request {i} differs when started from the root"
        sweeps_root += int(got[i].sweeps)
    for e, (a_lo, a_hi, a_r) in enumerate(extra_alone):
        lo, hi = stores[len(tr) + e]
        assert _same(lo, hi, got[len(tr) + e], a_lo, a_hi, a_r), f"request outside the told store {e}"
    assert sweeps_root < sum(int(alone[i][2].sweeps) for i in range(len(tr)))


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
    with pytest.raises(pcx.PcxError) as e:
        srv.set_root(5, lb, ub)                                             # unknown network
    assert e.value.status == pcx.EINVAL
    with pytest.raises(pcx.PcxError) as e:
        srv.set_root(n, lb[:2].copy(), ub[:2].copy())                       # too few variables
    assert e.value.status == pcx.EINVAL


def test_root_without_a_fixpoint(gpu):
    """A told store whose propagation fails: its root fixpoint is empty, so
    every request inside it fails (as it does alone); a request outside it
    (y widened past the told bound) runs from its own store and does not."""
    op = np.array([pcx.ADD], np.int32)
    x, y, z = (np.array([v], np.int32) for v in (0, 1, 2))
    lb = np.array([20, 0, 0], np.int32)
    ub = np.array([30, 3, 3], np.int32)                 # x = y + z <= 6 < 20: no fixpoint
    p = pcx.PIR()
    p.tell(op, x, y, z, lb, ub)
    srv = pcx.Server(max_slots=4)
    n = srv.add_network(p)
    srv.set_root(n, lb, ub)
    inside = (np.array([20, 1, 1], np.int32), np.array([25, 3, 3], np.int32))
    outside = (np.array([20, 0, 0], np.int32), np.array([30, 30, 3], np.int32))
    for tag, (lo, hi) in enumerate([inside, outside]):
        srv.submit(n, lo, hi, tag=tag)
    done = {}
    while len(done) < 2:
        srv.step()
        done.update({t: r for t, r, _ in srv.poll()})
    assert done[0].failed
    assert not done[1].failed and list(outside[0]) == [20, 17, 0] and list(outside[1]) == [30, 30, 3]
    a_lo, a_hi = (v.copy() for v in (np.array([20, 0, 0], np.int32), np.array([30, 30, 3], np.int32)))
    r = p.fixpoint(a_lo, a_hi)
    assert not r.failed and np.array_equal(a_lo, outside[0]) and np.array_equal(a_hi, outside[1])


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


def _climbers(n, seed, big=10**6):
    """n copies of a = b + 1, t = a + 1, b = min(t, cap): the lower bounds of
    a, b, t climb by about one per sweep until b reaches cap, so a request
    needs about max(cap) sweeps.  Variable 5g+4 is gadget g's cap and 5g its
    a (an upper bound on a below cap + 1 makes the request fail after about
    that many sweeps)."""
    rng = np.random.default_rng(seed)
    op, x, y, z, lb, ub = [], [], [], [], [], []
    for g in range(n):
        a, b, t, one, cap = range(5 * g, 5 * g + 5)
        K = int(rng.integers(10, 1200))
        lb += [0, 0, 0, 1, K]
        ub += [big, big, big, 1, K]
        op += [pcx.ADD, pcx.ADD, pcx.MIN]
        x += [a, t, b]
        y += [b, a, t]
        z += [one, one, cap]
    return [np.array(v, np.int32) for v in (op, x, y, z, lb, ub)]


@pytest.fixture(scope="module")
def climbers(gpu):
    """Two networks (50 and 4,000 gadgets) and 24 requests with random caps,
    a quarter of them made to fail; the oracle's fixpoint and the request
    alone for each."""
    nets, pirs = [], []
    for n, seed in ((50, 1), (4000, 2)):
        op, x, y, z, lb, ub = _climbers(n, seed)
        p = pcx.PIR()
        p.tell(op, x, y, z, lb, ub)
        nets.append((lb, ub))
        pirs.append(p)
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(24):
        j = i % 2
        lo, hi = (a.copy() for a in nets[j])
        G = len(lo) // 5
        caps = rng.integers(10, 1500, G)
        lo[4::5] = caps
        hi[4::5] = caps
        if i % 4 == 3:                           # one gadget's a cannot reach cap + 1
            g = int(rng.integers(G))
            hi[5 * g] = int(caps[g]) // 2 + 5
        reqs.append((j, lo, hi))
    oracle, alone = [], []
    for j, lo, hi in reqs:
        top, tx, ty, tz = pirs[j].table()
        o_lo, o_hi = lo.copy(), hi.copy()
        oracle.append((o_lo, o_hi, ol.fixpoint(top, tx, ty, tz, o_lo, o_hi)))
        a_lo, a_hi = lo.copy(), hi.copy()
        alone.append((a_lo, a_hi, pirs[j].fixpoint(a_lo, a_hi)))
    return pirs, reqs, oracle, alone


@pytest.mark.parametrize("spt", [2, 3])
@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
@pytest.mark.parametrize("burst", [(-1, -1), (5, 0), (0, 0)], ids=["auto", "short", "off"])
def test_bursts_slow_convergence(climbers, events, burst, spt):
    """Requests needing up to ~1,500 sweeps, served with the side lane (long
    bursts, bursts of 5 rounds relaunched again and again, or none): every
    result equals the oracle's and the request alone.  The 4,000-gadget
    network starts with more marks than a burst holds, so its requests go
    back to the slot batch (with 3 sweeps per tick, at a sweep of the other
    parity) and return to the side lane later."""
    pirs, reqs, oracle, alone = climbers
    for i, (o, a) in enumerate(zip(oracle, alone)):
        assert _same(*o, *a), f"request {i}: alone differs from the oracle"
    assert sum(bool(o[2].failed) for o in oracle) == len(reqs) // 4
    got, st, stores = _serve_climbers(pirs, reqs, range(len(reqs)), events, *burst, spt=spt)
    for i in range(len(reqs)):
        assert _same(*stores[i], got[i], *oracle[i]), f"request {i} (net {reqs[i][0]})"
        if not got[i].failed:
            assert got[i].sweeps >= 10
    if burst[0] == 0:
        assert st.sweeps >= max(r.sweeps for r in got.values())


@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
def test_bursts_take_the_long_requests(climbers, events):
    """The small network's requests (150 propagators: bursts apply with or
    without marks) do most of their sweeps in the side lane: the slot batch
    runs far fewer sweeps than they need."""
    pirs, reqs, oracle, _ = climbers
    sel = [i for i in range(len(reqs)) if reqs[i][0] == 0]
    got, st, stores = _serve_climbers(pirs, reqs, sel, events, -1, 8)
    for i in sel:
        assert _same(*stores[i], got[i], *oracle[i]), f"request {i}"
    longest = max(r.sweeps for r in got.values())
    assert longest > 400 and st.sweeps < longest // 4, (st.sweeps, longest)


def _serve_climbers(pirs, reqs, sel, events, burst_sweeps, burst_after, spt=2):
    srv = pcx.Server(max_slots=8, sweeps_per_tick=spt, event_driven=events,
                     burst_sweeps=burst_sweeps, burst_after=burst_after)
    for p in pirs:
        srv.add_network(p)
    stores = {i: (reqs[i][1].copy(), reqs[i][2].copy()) for i in sel}
    for i in sel:
        srv.submit(reqs[i][0], *stores[i], tag=i)
    got = {}
    while len(got) < len(stores):
        srv.step()
        for tag, r, _ in srv.poll():
            got[tag] = r
        assert srv.stats().ticks < 10000000, "server made no progress"
    return got, srv.stats(), stores


@pytest.mark.parametrize("burst", [(-1, 8), (0, 0)], ids=["side-lanes", "batch-only"])
@pytest.mark.parametrize("events", [1, 0], ids=["marked", "full"])
def test_staging_buffer_holds_one_store(climbers, events, burst):
    """Regression test for the pinned staging buffer (stage_bytes), the
    buffer every store passes through on its way to and from the device.  It
    holds one and a half of these 250-variable stores, so admissions,
    hand-backs from the slot batch and hand-backs from the side lanes each
    find it full and defer the rest to later ticks.

    Eight requests share one store; admitted one per tick (one sweep each),
    they start on eight consecutive sweeps and need the same ~hundreds of
    sweeps, so with 16 sweeps per tick several settle in one tick (slot
    batch), or are handed to one side lane together and settle in one launch.
    Four requests with other caps follow.  Every result equals the oracle's
    and the request alone."""
    pirs, reqs, oracle, alone = climbers
    sel = [i for i in range(len(reqs)) if reqs[i][0] == 0]
    twin = sel[0]
    nv = len(reqs[twin][1])
    cases = [twin] * 8 + sel[1:5]
    srv = pcx.Server(max_slots=8, sweeps_per_tick=16, event_driven=events, burst_sweeps=burst[0],
                     burst_after=burst[1], stage_bytes=nv * 8 * 3 // 2)
    for p in pirs:
        srv.add_network(p)
    stores = []
    for tag, i in enumerate(cases):
        lo, hi = reqs[i][1].copy(), reqs[i][2].copy()
        stores.append((lo, hi))
        srv.submit(0, lo, hi, tag=tag)
    got = {}
    while len(got) < len(cases):
        srv.step()
        for tag, r, _ in srv.poll():
            assert tag not in got
            got[tag] = r
        assert srv.stats().ticks < 10000000, "server made no progress"
    st = srv.stats()
    for tag, i in enumerate(cases):
        assert _same(*stores[tag], got[tag], *oracle[i]), f"request {tag} (trace request {i}) vs oracle"
        assert _same(*stores[tag], got[tag], *alone[i]), f"request {tag} (trace request {i}) vs alone"
    assert not oracle[twin][2].failed and got[0].sweeps > 64
    assert st.waits_stage_admit > 0
    if burst[0] == 0:
        assert st.waits_stage_harvest > 0
    else:
        assert st.waits_stage_side > 0
