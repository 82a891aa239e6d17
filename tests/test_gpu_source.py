"""The device source-frame step of pe_redshift (exm_source_frame) against the oracle's restatement
of src/exmarg/search/pe.py:87-142, with its edge cases."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def case(seed, nn=9, nm=14, nz=40, nm1=33):
    from exmarg_amd import cosmology
    from exmarg_amd.search import pe
    rng = np.random.default_rng(seed)
    yN = np.where(rng.uniform(size=(nn, nm)) < 0.8, rng.uniform(0.1, 3.0, (nn, nm)), 0.0)
    lamN = np.where(rng.uniform(size=(nn, nm)) < 0.9, rng.standard_normal((nn, nm)) * 3, -np.inf)
    lamN[:, 0] = -np.inf                 # a template with no finite point
    lamN[1:, 1] = -np.inf                # and one with a single point
    M = np.sort(rng.uniform(20, 120, nm))
    z = np.linspace(0.02, 2.0, nz)
    yq = pe.D0 / cosmology.luminosity_distance(z)
    m1 = np.geomspace(3.0, 60.0, nm1)
    m1[5] = M[4] / (2.0 * (1.0 + z[7]))  # a grid point exactly on a template mass
    return lamN, yN, M, yq, m1, z


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_source_frame_matches_oracle(hip, oracle, seed):
    from exmarg_amd.search import pe
    from exmarg_amd import cosmology
    say("source frame: case seed %d" % seed)
    lamN, yN, M, yq, m1, z = case(seed)
    dl = cosmology.luminosity_distance(z)
    w = pe.redshift_prior_weights(m1, z, dl, (10.0, 500.0))
    out = np.empty((m1.size, z.size))
    ln_b = pe.source_frame(lamN, yN, M, yq, m1, z, w, out=out)
    ref = oracle.to_source_frame(lamN, yN, M, yq, m1, z)
    np.testing.assert_array_equal(np.isfinite(out), np.isfinite(ref))
    assert np.isfinite(ref).sum() > 100
    np.testing.assert_allclose(out[np.isfinite(out)], ref[np.isfinite(ref)], rtol=1e-13, atol=1e-13)
    assert abs(ln_b - oracle.log_mean(ref, w)) < 1e-12


def test_source_frame_at_full_size(hip, oracle):
    """the pe_redshift default grid (1600 x 1600) over 3517 masses and 54 rows"""
    from exmarg_amd.search import pe
    from exmarg_amd import cosmology
    rng = np.random.default_rng(9)
    nn, nm = 54, 3517
    yN = np.sort(rng.uniform(0.05, 20.0, (nn, nm)), axis=0)
    lamN = np.where(rng.uniform(size=(nn, nm)) < 0.95, rng.standard_normal((nn, nm)), -np.inf)
    M = np.geomspace(10.0, 500.0, nm)
    z = np.linspace(0.02, 8.0, 1600)
    m1 = np.geomspace(3.0, 150.0, 1600)
    yq = pe.D0 / cosmology.luminosity_distance(z)
    w = pe.redshift_prior_weights(m1, z, cosmology.luminosity_distance(z), (10.0, 500.0))
    out = np.empty((m1.size, z.size))
    ln_b = pe.source_frame(lamN, yN, M, yq, m1, z, w, out=out)
    # columns checked against the oracle on a sample of z (the oracle loops in Python)
    cols = np.arange(0, 1600, 97)
    ref = oracle.to_source_frame(lamN, yN, M, yq[cols], m1, z[cols])
    got = out[:, cols]
    np.testing.assert_array_equal(np.isfinite(got), np.isfinite(ref))
    np.testing.assert_allclose(got[np.isfinite(got)], ref[np.isfinite(ref)], rtol=1e-12, atol=1e-12)
    fin = np.isfinite(out) & (w > 0)
    mx = out[fin].max()
    assert abs(ln_b - (mx + np.log(np.sum(w[fin] * np.exp(out[fin] - mx))))) < 1e-12


def test_source_frame_rejects_unsorted_masses(hip):
    from exmarg_amd.search import pe
    lamN, yN, M, yq, m1, z = case(3)
    with pytest.raises(ValueError, match="ascend"):
        pe.source_frame(lamN, yN, M[::-1].copy(), yq, m1, z)


def test_source_grid_equals_source_frame(hip, oracle):
    """exm_source_grid_create / _run (the call's constants kept on the device) give the grid and ln B of
    exm_source_frame bit for bit, segment after segment on one grid; a grid without w returns no ln B;
    the full-size ln B against a numpy sum over the 2.56e6 cells (the two-stage reduction)"""
    from exmarg_amd.search import pe
    from exmarg_amd import cosmology
    say("source grid: three segments on one grid, then the 1600 x 1600 grid")
    lamN, yN, M, yq, m1, z = case(4)
    w = pe.redshift_prior_weights(m1, z, cosmology.luminosity_distance(z), (10.0, 500.0))
    grid = pe.SourceGrid(yN, M, yq, m1, z, w)
    rng = np.random.default_rng(21)
    for k in range(3):
        L = np.where(np.isfinite(lamN), lamN + rng.standard_normal(lamN.shape), -np.inf)
        a, b = np.empty((m1.size, z.size)), np.empty((m1.size, z.size))
        lb_a = grid.run(L, out=a)
        lb_b = pe.source_frame(L, yN, M, yq, m1, z, w, out=b)
        np.testing.assert_array_equal(a, b)
        assert lb_a == lb_b
        ref = oracle.to_source_frame(L, yN, M, yq, m1, z)
        assert abs(lb_a - oracle.log_mean(ref, w)) < 1e-12
    # run_many over five segments (the staging slots reused twice over) equals run one at a time, bit for bit
    Ls = np.stack([np.where(np.isfinite(lamN), lamN + rng.standard_normal(lamN.shape), -np.inf) for _ in range(5)])
    many = np.empty((5, m1.size, z.size))
    lb_many = grid.run_many(Ls, out=many)
    for k in range(5):
        one = np.empty((m1.size, z.size))
        assert grid.run(Ls[k], out=one) == lb_many[k]
        np.testing.assert_array_equal(many[k], one)
    np.testing.assert_array_equal(grid.run_many(Ls[:2]), lb_many[:2])
    assert grid.run_many(Ls[:0]).size == 0
    with pytest.raises(ValueError, match="shape"):
        grid.run(lamN[:-1])
    with pytest.raises(ValueError, match="C-contiguous"):
        grid.run_many(Ls, out=many[:4])
    grid.free()
    nw = pe.SourceGrid(yN, M, yq, m1, z)
    assert nw.run(lamN, out=None) is None
    nw.free()
    nn, nm = 54, 3517
    yN = np.sort(rng.uniform(0.05, 20.0, (nn, nm)), axis=0)
    lamN = np.where(rng.uniform(size=(nn, nm)) < 0.95, rng.standard_normal((nn, nm)), -np.inf)
    M = np.geomspace(10.0, 500.0, nm)
    z = np.linspace(0.02, 8.0, 1600)
    m1 = np.geomspace(3.0, 150.0, 1600)
    yq = pe.D0 / cosmology.luminosity_distance(z)
    w = pe.redshift_prior_weights(m1, z, cosmology.luminosity_distance(z), (10.0, 500.0))
    grid = pe.SourceGrid(yN, M, yq, m1, z, w)
    out = np.empty((m1.size, z.size))
    ln_b = grid.run(lamN, out=out)
    grid.free()
    fin = np.isfinite(out) & (w > 0)
    mx = out[fin].max()
    assert abs(ln_b - (mx + np.log(np.sum(w[fin] * np.exp(out[fin] - mx))))) < 1e-12
