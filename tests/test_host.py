"""Host-side logic of the package against the oracle, the reference's file formats and the C ABI
(no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from exmarg_amd import cosmology, synth
from exmarg_amd.search import pe, segments, store as stm

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class FakeBank:
    def __init__(self, u, M=None):
        self.u = np.asarray(u, float)
        self.M = np.arange(self.u.size, dtype=float) + 10 if M is None else M
        self.q = np.ones(self.u.size)
        self.chi = np.zeros(self.u.size)


def fake_store(s):
    return stm.RedshiftStore(s=np.asarray(s), alpha_lo=np.full(len(s), 0.16), alpha_hi=np.full(len(s), 4e4),
                             beta=[], delta=[], tau_idx=np.arange(3), tables=None, n_max=120)


@pytest.mark.parametrize("ymin", [0.1267, 0.0982, 0.5])
def test_node_geometry_matches_oracle(oracle, ymin):
    u = np.geomspace(0.05, 500.0, 37)
    st = fake_store(synth.S_NODES)
    got = pe._node_geometry(FakeBank(u), st, ymin)
    ref = oracle.node_geometry(st.s, st.alpha_lo, st.alpha_hi, u, ymin)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)


def test_bank_weights_reject_a_template_with_one_point():
    st = fake_store(synth.S_NODES[:3])
    u = np.array([1.0, 1e6])
    yN = pe._node_geometry(FakeBank(u), st, 0.1267)[2]
    with pytest.raises(ValueError, match="fewer than two"):
        pe._bank_weights(yN, "volume", 0.1267, FakeBank(u))


def test_prior_name_checked():
    with pytest.raises(ValueError):
        pe.distance_weights(np.array([0.2, 0.3]), "uniform", 0.1)


def test_luminosity_distance_golden_pairs(oracle):
    z = np.load(os.path.join(GOLDEN, "planck18_pairs.npz"))
    np.testing.assert_allclose(cosmology.luminosity_distance(z["z"]), z["dl"], rtol=1e-7)
    zz = np.linspace(0.01, 8.0, 57)
    np.testing.assert_allclose(cosmology.luminosity_distance(zz), oracle.dl_of_z(zz), rtol=1e-12)


def test_cells_match_oracle(oracle):
    rows = synth.grid_bank(shape=(4, 3, 2))
    b = FakeBank(np.ones(24), M=rows[:, 0])
    b.q, b.chi = rows[:, 1], rows[:, 2]
    np.testing.assert_allclose(pe._cell_weights(b), oracle.cell_weights(rows[:, 0], rows[:, 1], rows[:, 2]))


def test_read_segments_file_format():
    path = os.path.join(GOLDEN, "segments_2p.npz")
    seg = segments.read_segments(path, rows=[1])
    raw = np.load(path)
    assert seg["d_H"].dtype == np.complex128 and seg["d_H"].shape == (1, raw["nu"].size)
    np.testing.assert_array_equal(seg["d_L"][0], raw["d_L"][1].astype(complex))
    np.testing.assert_array_equal(seg["sig2_H"], raw["sig2"])
    assert seg["T"] == 8.0 and seg["fmin"] == 20.0 and seg["fmax"] == 512.0
    seg2 = segments.read_segments(dict(raw))
    assert seg2["rows"].tolist() == [0, 1]


def test_store_files_roundtrip(tmp_path):
    arr = synth.synthetic_store_arrays(n_nodes=3, seed=4, kind="rough", shape=(3, 49, 5, 7))
    arr["s"] = arr["s"][::-1].copy()
    arr["tables"] = arr["tables"][::-1]
    synth.write_store(str(tmp_path), arr)
    for mmap in (False, True):
        st = stm.load_redshift_store(str(tmp_path), mmap=mmap)
        assert np.all(np.diff(st.s) > 0) and st.n_max == 120
        order = np.argsort(arr["s"])
        for k, o in enumerate(order):
            np.testing.assert_array_equal(np.asarray(st.tables[k]), arr["tables"][o])
        np.testing.assert_array_equal(st.tau_idx, arr["tau_idx"])
        assert st.alpha_lo[0] == pytest.approx(0.16) and st.alpha_hi[0] == pytest.approx(4e4)


def test_store_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        stm.load_redshift_store(str(tmp_path))
    p = tmp_path / "bd_n120_s0001.000.npz"
    np.savez_compressed(p, T=np.zeros((1, 49, 2, 2), np.float32), beta=np.zeros(2), delta=np.zeros(2),
                        tau_idx=np.zeros(1, int), alpha=np.ones(2), s_node=1.0)
    with pytest.raises(ValueError, match="compressed"):
        stm.load_redshift_store(str(tmp_path))


def test_device_names_checked():
    from exmarg_amd.search import mass_bank
    seg = os.path.join(GOLDEN, "segments_2p.npz")
    for dev in ("cpu", "gpu", "cuda"):
        with pytest.raises(ValueError, match="hip"):
            mass_bank(np.array([30.0]), seg, dev)


def test_synthetic_tables_numpy_equals_c(oracle):
    shape = (87, 49, 200, 219)
    tau = synth.lag_indices()
    idx = np.random.default_rng(0).integers(0, int(np.prod(shape)), 5000)
    for kind in ("smooth", "rough"):
        full = oracle.synth_table(7, 3, (5, 49, 8, 11), synth.lag_indices(2), kind)
        np.testing.assert_array_equal(full, synth.table_values(7, 3, (5, 49, 8, 11), synth.lag_indices(2),
                                                              kind).reshape(5, 49, 8, 11))
        lib = oracle.lib()
        lib.oracle_synth_value.restype = ctypes.c_float
        lib.oracle_synth_value.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int, ctypes.c_int64,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        tau64 = np.ascontiguousarray(tau, np.int64)
        c = np.array([lib.oracle_synth_value(0, 9, 1 if kind == "rough" else 0, int(i), 49, 200, 219,
                                             tau64.ctypes.data) for i in idx], np.float32)
        np.testing.assert_array_equal(c, synth.table_values(0, 9, shape, tau, kind, flat=idx))


def test_store_grids_reproduce_the_reference_constants():
    beta, delta, lnq = synth.store_grids()
    assert beta.size == 200 and delta.size == 219
    assert beta[0] == pytest.approx(-7.824046, abs=1e-6) and beta[1] - beta[0] == pytest.approx(0.139466, abs=1e-6)
    assert delta[0] == pytest.approx(-7.600902, abs=1e-6) and np.allclose(np.diff(delta), 0.0697330500875406)
    g = np.load(os.path.join(GOLDEN, "s_grid.npz"))["s_nodes"]
    np.testing.assert_allclose(synth.S_NODES, g, rtol=1e-14)
    tau = synth.lag_indices()
    assert tau.size == 87 and tau[43] == 43 and tau[44] == -43


# ---------------------------------------------------------------- C ABI
def header_symbols():
    with open(os.path.join(ROOT, "include", "exmarg_hip.h")) as f:
        return re.findall(r"^int (exm_\w+)\(", f.read(), flags=re.M)


def test_library_exports_every_declared_symbol():
    from exmarg_amd import _hip
    L = _hip.load()
    names = header_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
        assert n in _hip.SIGNATURES, n
    assert set(_hip.SIGNATURES) == set(names)


def test_library_fails_loudly_without_a_device():
    from exmarg_amd import _hip
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a device is visible")
    except ImportError:
        pass
    rc = _hip.load().exm_store_create(1, 1, 49, 2, 2, None, 0.0, 1.0, 0.0, 1.0, None)
    assert rc == _hip.EXM_ERR_STATE
    assert "exm_init" in _hip.last_error()


def test_rng_state_is_numpys_pcg64_state():
    """the four words handed to the device draw are the PCG64 state and increment of default_rng(seed)"""
    from exmarg_amd.search.gather import rng_state
    for seed in (0, 5, 2**40 + 3):
        st = np.random.default_rng(seed).bit_generator.state
        assert st["bit_generator"] == "PCG64" and st["has_uint32"] == 0
        w = [int(x) for x in rng_state(seed)]
        assert (w[0] << 64) | w[1] == st["state"]["state"]
        assert (w[2] << 64) | w[3] == st["state"]["inc"]


def test_template_bytes_covers_records_and_launch_work():
    from exmarg_amd.search.batch import template_bytes
    per = template_bytes(87, 2048)
    assert per >= 87 * 2048 * 8 + 2048 * 4          # walk records and walk
    assert per - template_bytes(87, 2048, reads=0) == 54 * (16 * 87 + 2 * 201 + 32)


def test_tune_cache_roundtrip(tmp_path, monkeypatch):
    from exmarg_amd import tune
    monkeypatch.setenv("EXMARG_TUNE_CACHE", str(tmp_path / "t.json"))
    assert tune._load() == {}
    tune._save({"k": {"form": 3, "piece": 4, "tile": 1024, "ms": 1.0}})
    assert tune._load()["k"]["piece"] == 4
