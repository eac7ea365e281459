"""pcx_fixpoint_noalloc (the allocation-free fixpoint) against pcx_fixpoint.

Two calls in a row on one GPU, the second with a smaller store, each
compared bit for bit with the existing entry point on the same network and
store; a store beyond the fixed working array and a table that outgrew its
device copy are refused; a store empty at the start fails.  All calls go
through the C ABI."""
import numpy as np
import pytest

import pcx
from pcx.device import DeviceArray

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    if pcx.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True


def _case(seed, V, P):
    p = pcx.gen_params(seed=seed, nvars=V, nprops=P)
    op, x, y, z, lb, ub, _ = pcx.gen_network(p)
    pir = pcx.PIR()
    pir.tell(op, x, y, z, lb, ub)
    return pir, lb, ub


def _both(pir, lb, ub):
    lo_r, hi_r = lb.copy(), ub.copy()
    r_ref = pir.fixpoint(lo_r, hi_r)                 # the existing entry point
    d_lb, d_ub = DeviceArray(lb), DeviceArray(ub)
    r = pir.fixpoint_noalloc(d_lb.ptr, d_ub.ptr, len(lb))
    lo, hi = d_lb.read(), d_ub.read()
    d_lb.close()
    d_ub.close()
    assert bool(r.failed) == bool(r_ref.failed)
    if not r_ref.failed:
        assert np.array_equal(lo, lo_r) and np.array_equal(hi, hi_r)
        assert bool(r.has_changed) == bool(r_ref.has_changed)
    return r


def test_two_calls_second_smaller(gpu):
    assert pcx.lib().pcx_noalloc_max_vars() == 1 << 21
    big = _case(5, 200000, 800000)
    r1 = _both(*big)
    assert not r1.failed and r1.sweeps > 2
    small = _case(6, 3000, 12000)
    r2 = _both(*small)
    assert not r2.failed
    r3 = _both(*big)                                 # and back to the large one
    assert not r3.failed


def test_limits_and_failure(gpu):
    pir, lb, ub = _case(7, 2000, 8000)
    d = DeviceArray(np.zeros(16, np.int32))
    with pytest.raises(pcx.PcxError) as e:
        pir.fixpoint_noalloc(d.ptr, d.ptr, (1 << 21) + 1)
    assert e.value.status == pcx.EINVAL and "exceed" in str(e.value)
    d_lb, d_ub = DeviceArray(lb), DeviceArray(ub)
    with pytest.raises(pcx.PcxError) as e:           # a new engine's table is not on the device yet
        pir.fixpoint_noalloc(d_lb.ptr, d_ub.ptr, len(lb))
    assert e.value.status == pcx.EINVAL and "outgrew" in str(e.value)
    d_lb.close()
    d_ub.close()
    pir.fixpoint(lb.copy(), ub.copy())               # one allocating call uploads it
    lo, hi = lb.copy(), ub.copy()
    lo[17], hi[17] = 5, 4                            # empty from the start
    d_lb, d_ub = DeviceArray(lo), DeviceArray(hi)
    r = pir.fixpoint_noalloc(d_lb.ptr, d_ub.ptr, len(lo))
    assert r.failed and r.sweeps == 0
    d.close()
    d_lb.close()
    d_ub.close()
    # a tell that grows the table past its device copy needs one sizing call
    more = pcx.gen_network(pcx.gen_params(seed=8, nvars=2000, nprops=500))
    pir.tell(*more[:4])
    d_lb, d_ub = DeviceArray(lb), DeviceArray(ub)
    with pytest.raises(pcx.PcxError) as e:
        pir.fixpoint_noalloc(d_lb.ptr, d_ub.ptr, len(lb))
    assert e.value.status == pcx.EINVAL
    d_lb.close()
    d_ub.close()
    _both(pir, lb, ub)                               # pcx_fixpoint sizes it; both agree
