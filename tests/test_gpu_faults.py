"""Error paths under refused device allocations (exm_debug_fail_alloc makes the n-th next allocation fail
as a refused hipMalloc would): every call either succeeds or raises MemoryError, leaks no device memory,
and leaves the library able to repeat the same analysis with the same results.  The failure is forced
from the host; every kernel still runs to its normal end."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu
# repeated runs of one analysis: the band kernel's default form adds each read's pieces by float64
# atomic adds in LDS, whose order varies from run to run, so repeats agree to float64 rounding
REPEAT_RTOL = 1e-13


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def small(hip):
    from exmarg_amd.search import mass_bank, store as stm
    seg, _, _ = synth.simulate_at_snr(2, 41, (30.0, 0.8, 0.1), 12.0, 2.0, 20.0, 256.0)
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[::4])
    bank = mass_bank(synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 6), seg, "hip")
    yield seg, bank, st
    hip.call("exm_debug_fail_alloc", -1)
    stm.free(st)


def _in_use(hip):
    return hip.mem_in_use()


def test_refused_allocations_in_pe_intrinsic(hip, small):
    from exmarg_amd.search import pe_intrinsic
    seg, bank, st = small
    kw = dict(n_loud=64, n_random=64, seed=2, resident=True)
    # the two segments share a gather launch; a refused batch allocation falls back to one launch per
    # segment, whose work memory the library then keeps: run that path once before taking the baseline
    pe_intrinsic(seg, bank, st, chunk=1, **kw)
    ref = pe_intrinsic(seg, bank, st, **kw)
    base = _in_use(hip)
    failed = 0
    for k in range(24):
        hip.call("exm_debug_fail_alloc", k)
        try:
            pe_intrinsic(seg, bank, st, **kw)
        except MemoryError as e:
            assert "injected" in str(e)
            failed += 1
        finally:
            hip.call("exm_debug_fail_alloc", -1)
        assert _in_use(hip) == base, "device memory held after a refused allocation (k = %d)" % k
        again = pe_intrinsic(seg, bank, st, **kw)
        np.testing.assert_allclose(again["ln_lambda"], ref["ln_lambda"], rtol=REPEAT_RTOL, atol=0)
    say("pe_intrinsic: %d of 24 injection points hit an allocation" % failed)
    assert failed > 0


def test_refused_allocations_in_pe_redshift(hip, small):
    """pe_redshift of two segments sharing a gather launch, with the source-frame grid kept on the device"""
    from exmarg_amd.search import mass_bank, pe_redshift
    seg, _, st = small
    rbank = mass_bank(np.geomspace(20.0, 60.0, 6), seg, "hip")
    kw = dict(n_loud=64, n_random=64, seed=2, resident=True, zmax=3.0, nz=60, m1_range=(5.0, 40.0), nm1=50,
              mdet_range=(20.0, 60.0))
    pe_redshift(seg, rbank, st, chunk=1, **kw)          # the fallback path's kept work memory, as above
    ref = pe_redshift(seg, rbank, st, **kw)
    base = _in_use(hip)
    failed = 0
    for k in range(24):
        hip.call("exm_debug_fail_alloc", k)
        try:
            pe_redshift(seg, rbank, st, **kw)
        except MemoryError as e:
            assert "injected" in str(e)
            failed += 1
        finally:
            hip.call("exm_debug_fail_alloc", -1)
        assert _in_use(hip) == base, "device memory held after a refused allocation (k = %d)" % k
    again = pe_redshift(seg, rbank, st, **kw)
    fin = np.isfinite(ref["ln_lambda"])
    np.testing.assert_array_equal(np.isfinite(again["ln_lambda"]), fin)
    np.testing.assert_allclose(again["ln_lambda"][fin], ref["ln_lambda"][fin], rtol=REPEAT_RTOL, atol=1e-12)
    np.testing.assert_allclose(again["ln_b"], ref["ln_b"], rtol=REPEAT_RTOL, atol=1e-12)
    say("pe_redshift: %d of 24 injection points hit an allocation" % failed)
    assert failed > 0


def test_refused_allocations_in_batches(hip, small):
    from exmarg_amd.search import pe_intrinsic_batch
    seg, bank, st = small
    kw = dict(n_loud=64, n_random=64, seed=2, max_templates=24)
    ref = pe_intrinsic_batch([(seg, bank)], st, **kw)
    base = _in_use(hip)
    failed = 0
    for k in range(24):
        hip.call("exm_debug_fail_alloc", k)
        try:
            pe_intrinsic_batch([(seg, bank)], st, **kw)
        except MemoryError:
            failed += 1
        finally:
            hip.call("exm_debug_fail_alloc", -1)
        assert _in_use(hip) == base, "device memory held after a refused allocation (k = %d)" % k
    again = pe_intrinsic_batch([(seg, bank)], st, **kw)
    np.testing.assert_allclose(again[0]["ln_lambda"], ref[0]["ln_lambda"], rtol=REPEAT_RTOL, atol=0)
    say("pe_intrinsic_batch: %d of 24 injection points hit an allocation" % failed)
    assert failed > 0


def test_refused_allocation_fails_one_request_only(hip, small):
    """a request whose admission meets a refused allocation is refused with that error; the others are
    served as without it"""
    from exmarg_amd.search import Server
    seg, bank, st = small
    rows = synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 6)
    srv = Server(st, n_loud=64, n_random=64, max_request=(64, 8 * 512, 1024))
    try:
        hs = [srv.submit(seg, rows, seed=0)]
        srv.step()
        hip.call("exm_debug_fail_alloc", 0)
        try:
            hs.append(srv.submit(seg, rows * np.array([1.01, 1.0, 1.0]), seed=1))
            srv.step()
        finally:
            hip.call("exm_debug_fail_alloc", -1)
        hs.append(srv.submit(seg, rows, seed=2))
        srv.run()
    finally:
        srv.close()
    assert hs[0].error is None and hs[2].error is None
    assert isinstance(hs[1].error, MemoryError)
    assert np.all(np.isfinite(hs[0].ln_lambda)) and np.all(np.isfinite(hs[2].ln_lambda))
