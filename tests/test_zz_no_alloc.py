"""exm_plan_gather_noalloc (the plan gather with every scratch buffer in the library's static per-device
region) against exm_plan_gather through the C ABI: the same results on two calls in a row (the second
with fewer templates, so that it runs over what the first left in the region), for a subsampled walk and
an every-time walk (band kernel, in slices); a gather larger than the region is refused before anything
runs.  The band kernel's default (scheduled) form adds each read's pieces with float64 atomic adds in
LDS, whose order varies from run to run, so two runs agree to float64 rounding (relative 1e-13 here),
not bit for bit."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def setup(hip):
    from exmarg_amd.search import store as stm
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[::3])
    yield st
    stm.free(st)


def analysis(st, n_templates, T=2.0, fmax=256.0, seed=3):
    from exmarg_amd.search import mass_bank, pe as _pe
    seg, _, _ = synth.simulate_at_snr(1, seed, (30.0, 0.8, 0.1), 12.0, T, 20.0, fmax)
    bank = mass_bank(synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, n_templates), seg, "hip")
    an = _pe._Analysis(bank, st, 0.1267, True).start()
    from exmarg_amd.search.segments import read_segments
    r = read_segments(seg)
    traj = bank.trajectories({I: r["d_" + I][0] for I in "HL"}, {I: r["sig2_" + I] for I in "HL"})
    return bank, an, traj


def both(hip, an, traj, st, times):
    n = an.plan.job_m.size
    a = np.full(n, np.nan)
    b = np.full(n, np.nan)
    th = int(times.handle) if times is not None else 0
    hip.call("exm_plan_gather", int(an.plan.handle), int(traj.handle), int(st.device.handle), th, hip.ptr(a))
    hip.call("exm_plan_gather_noalloc", int(an.plan.handle), int(traj.handle), int(st.device.handle), th, hip.ptr(b))
    return a, b


@pytest.mark.parametrize("n_loud", [256, 0])
def test_noalloc_equals_plan_gather(hip, setup, n_loud):
    from exmarg_amd.search import gather as g
    st = setup
    for n_templates in (24, 9):                     # the second call smaller than the first
        bank, an, traj = analysis(st, n_templates)
        times = g.select_times(traj, st.tau_idx, n_loud, 256, seed=1) if n_loud else None
        say("no-alloc gather: %d templates, %d reads, n_loud %d" % (n_templates, an.plan.job_m.size, n_loud))
        a, b = both(hip, an, traj, st, times)
        assert np.isfinite(a).all()
        np.testing.assert_allclose(b, a, rtol=1e-13, atol=0)
        # the no-allocation call itself takes no pooled memory
        c = np.empty_like(b)
        before = hip.mem_in_use()
        hip.call("exm_plan_gather_noalloc", int(an.plan.handle), int(traj.handle), int(st.device.handle),
                 int(times.handle) if times is not None else 0, hip.ptr(c))
        assert hip.mem_in_use() == before
        np.testing.assert_allclose(c, a, rtol=1e-13, atol=0)
        if times is not None:
            times.free()
        traj.free()


def test_noalloc_refuses_a_gather_beyond_its_region(hip, setup):
    from exmarg_amd.search import gather as g
    st = setup
    bank, an, traj = analysis(st, 120)
    times = g.select_times(traj, st.tau_idx, 1024, 1024, seed=1)
    say("no-alloc gather beyond the region: %d templates x 2048 walked times" % bank.u.size)
    out = np.empty(an.plan.job_m.size)
    with pytest.raises(MemoryError, match="no-allocation maximum"):
        hip.call("exm_plan_gather_noalloc", int(an.plan.handle), int(traj.handle), int(st.device.handle),
                 int(times.handle), hip.ptr(out))
    # the allocating entry point still takes it
    hip.call("exm_plan_gather", int(an.plan.handle), int(traj.handle), int(st.device.handle), int(times.handle),
             hip.ptr(out))
    assert np.isfinite(out).all()
    times.free()
    traj.free()
