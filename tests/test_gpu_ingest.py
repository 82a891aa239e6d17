"""Tables streamed from the store files into HBM (exm_store_load_files) equal the tables read on the
host, with tables larger than one staging chunk and more tables than staging slabs."""
import numpy as np
import pytest

from exmarg_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("readers", [1, 6])
def test_file_ingest_equals_host_tables(hip, oracle, tmp_path, readers):
    from exmarg_amd.search import gather as g, store as stm
    say("ingest: 3 tables of (31, 49, 60, 219) float32 (80 MB each), %d readers" % readers)
    arrays = synth.synthetic_store_arrays(n_nodes=3, seed=2, kind="rough", shape=(31, 49, 60, 219),
                                          s_nodes=synth.S_NODES[20:23])
    synth.write_store(str(tmp_path), arrays)
    direct = stm.load_redshift_store(str(tmp_path), readers=readers, host=False)
    assert direct.tables is None and direct.files is not None
    stm.upload(direct)
    say("ingest seconds %.3f" % direct.ingest_seconds)
    for k, T in enumerate(arrays["tables"]):
        np.testing.assert_array_equal(direct.device.read_table(k), T)
    # mapped tables take the same path and give the same reads as tables read into host memory
    mapped = stm.load_redshift_store(str(tmp_path), mmap=True)
    stm.upload(mapped)
    np.testing.assert_array_equal(mapped.device.read_table(2), arrays["tables"][2])
    stm.free(direct)
    stm.free(mapped)


def test_file_ingest_rejects_a_missing_file(hip, tmp_path):
    from exmarg_amd.search import store as stm
    arrays = synth.synthetic_store_arrays(n_nodes=2, seed=1, shape=(5, 49, 8, 9), s_nodes=synth.S_NODES[:2])
    paths = synth.write_store(str(tmp_path), arrays)
    st = stm.load_redshift_store(str(tmp_path), host=False)
    import os
    os.remove(paths[1])
    with pytest.raises(ValueError, match="cannot open"):
        stm.upload(st)
