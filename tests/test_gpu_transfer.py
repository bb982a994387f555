"""Mesh-field transfer (src/py/interpolate_mesh_field.py,
src/py/build_mesh_transfer_tables.py): the GPU k-NN against the scipy
cKDTree the scripts query, the three interpolation modes and the two lookup
tables against the scripts' loops (oracle.transfer_field / transfer_tables),
and the files in the scripts' own format.  Indices must match exactly where
the neighbour distances are distinct; values within 1e-12 relative."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tr():
    from illum_amd import api, transfer
    api.default_context()
    return transfer


def _write_obj(path, V, F):
    with open(path, "w") as f:
        for v in V:
            f.write("v %r %r %r\n" % tuple(float(x) for x in v))
        for t in F + 1:
            f.write("f %d %d %d\n" % tuple(t))


def _mesh(sub, seed):
    from illum_amd import api
    V, F = api.synth_ellipsoid(sub, ncraters=30, seed=seed)
    return V * 100.0, F  # hundreds of units, as shape-model files hold them


def _check_knn(idx, dist, Dr, Ir):
    np.testing.assert_allclose(dist, Dr, rtol=1e-13, atol=1e-13)
    # where the reference's neighbour distances are distinct, the order is fixed
    gap = np.diff(Dr, axis=1) > 1e-9 * np.maximum(1.0, Dr[:, 1:])
    ok = np.concatenate([np.ones((len(Dr), 1), bool), gap], 1) & np.concatenate(
        [gap, np.ones((len(Dr), 1), bool)], 1)
    assert np.array_equal(idx[ok], Ir[ok])


@pytest.mark.parametrize("k", [1, 5, 12])
def test_knn_matches_ckdtree(tr, k):
    import scipy.spatial
    rng = np.random.default_rng(k)
    src = rng.normal(size=(20000, 3))
    tgt = rng.normal(size=(3000, 3)) * 1.2
    idx, dist = tr.knn(src, tgt, k)
    Dr, Ir = scipy.spatial.cKDTree(src).query(tgt, k)
    Dr, Ir = Dr.reshape(len(tgt), k), Ir.reshape(len(tgt), k)
    _check_knn(idx, dist, Dr, Ir)
    # mesh centroids across resolutions
    V4, F4 = _mesh(4, 3)
    V5, F5 = _mesh(5, 3)
    P4, P5 = tr.centroids(V4, F4), tr.centroids(V5, F5)
    idx, dist = tr.knn(P4, P5, k)
    Dr, Ir = scipy.spatial.cKDTree(P4).query(P5, k)
    _check_knn(idx, dist, Dr.reshape(len(P5), k), Ir.reshape(len(P5), k))


@pytest.mark.parametrize("mode", ["nearest", "weighted", "area_weighted"])
def test_interpolate_mesh_field(tr, tmp_path, mode):
    V3, F3 = _mesh(3, 5)
    V4, F4 = _mesh(4, 6)
    src, tgt = tmp_path / "src.obj", tmp_path / "tgt.obj"
    _write_obj(src, V3, F3)
    _write_obj(tgt, V4, F4)
    X = np.random.default_rng(2).uniform(100, 400, (len(F3), 3))  # facets x layers
    tr.arma_tofile(X, tmp_path / "x.bin")
    Y = tr.interpolate_mesh_field(str(src), str(tmp_path / "x.bin"), str(tgt),
                                  str(tmp_path / "y.bin"), mode=mode, neighbors=5)
    assert Y.shape == (len(F4), 3)
    assert np.array_equal(tr.arma_fromfile(tmp_path / "y.bin"), Y)
    Vs, Fs = tr.load_mesh(src)
    Vt, Ft = tr.load_mesh(tgt)
    Ps, Pt = tr.centroids(Vs, Fs), tr.centroids(Vt, Ft)
    As, At = tr.areas(Vs, Fs), tr.areas(Vt, Ft)
    Yo = O.transfer_field(Ps, X.T, Pt, mode, 5, As, At).T
    np.testing.assert_allclose(Y, Yo, rtol=1e-12, atol=0)


def test_build_mesh_transfer_tables(tr, tmp_path):
    paths = []
    for sub in (3, 4, 5):
        V, F = _mesh(sub, 9)
        p = tmp_path / f"m{sub}.obj"
        _write_obj(p, V, F)
        paths.append(str(p))
    down, neigh = tr.build_mesh_transfer_tables(paths, 5, str(tmp_path / "out"))
    P = [tr.centroids(*tr.load_mesh(p)) for p in paths]
    do, no = O.transfer_tables(P, 5)
    for i in (1, 2):
        assert np.array_equal(down[i], do[i])
        # the facet itself first (distance 0), then its neighbours; the three
        # edge neighbours of an icosphere facet can tie, so compare as sets
        assert np.array_equal(neigh[i][:, 0], no[i][:, 0])
        assert all(set(a) == set(b) for a, b in zip(neigh[i], no[i]))
        assert np.array_equal(tr.arma_fromfile(tmp_path / "out" / f"downsample_lut_{i}.bin",
                                               np.uint32), do[i])
    import yaml
    info = yaml.safe_load(open(tmp_path / "out" / "info.yaml"))
    assert info["num_neighbors"] == 5 and info["paths"] == paths
