"""The autotuner: every (piece, tile) candidate gives the same ln Lambda to the band kernel's rounding
(float32 exp of non-positive differences, grouped per piece: 1e-7), the fastest is kept and
recorded, and apply() restores it."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_tuner_candidates_agree_and_choice_is_recorded(hip, tmp_path, monkeypatch):
    from exmarg_amd import tune
    from exmarg_amd.search import SegmentBatch, mass_bank, read_segments, store as stm
    from exmarg_amd.search import pe as _pe
    monkeypatch.setenv("EXMARG_TUNE_CACHE", str(tmp_path / "tuning.json"))
    say("tune: 3 segments, forms (2, 3) x pieces (4, 8, 16) x tiles (256, 1024)")
    st = stm.synthetic_store(seed=0, s_nodes=synth.S_NODES[::4])
    seg, _, _ = synth.simulate_at_snr(3, 71, (30.0, 0.8, 0.1), 14.0, 4.0, 20.0, 256.0)
    bank = mass_bank(synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 40), seg, "hip")
    rseg = read_segments(seg)
    an = _pe._Analysis(bank, st, 0.1267, True).start()
    sb = SegmentBatch(st, 128, 128, max_templates=3 * 40)
    for j in range(3):
        sb.add(an, {I: rseg["d_" + I][j] for I in "HL"}, {I: rseg["sig2_" + I] for I in "HL"}, j)
    try:
        ref = sb.run()
        outs = {}
        for f in (2, 3):
            for p in ((4, 8, 16) if f == 2 else (4, 8)):
                for t in (256, 1024):
                    tune.set_knobs(p, t, form=f)
                    outs[(f, p, t)] = sb.run()
        for k, v in outs.items():
            for a, b in zip(v, ref):
                fin = np.isfinite(b)
                np.testing.assert_array_equal(np.isfinite(a), fin)
                np.testing.assert_allclose(a[fin], b[fin], rtol=0, atol=1e-7, err_msg=str(k))
        best = tune.tune_batch(sb, repeats=1, forms=(2, 3), pieces=(4, 8), tiles=(256, 1024))
        say("best", best)
        assert best["form"] in (2, 3) and best["piece"] in (4, 8) and best["tile"] in (256, 1024)
        assert tune.current()[:3] == (best["form"], best["piece"], best["tile"])
        tune.set_knobs(0, 0, form=0)
        rec = tune.apply(st, 256)
        assert rec is not None and tune.current()[:3] == (best["form"], best["piece"], best["tile"])
    finally:
        tune.set_knobs(0, 0, form=0)
        sb.free()
        stm.free(st)
