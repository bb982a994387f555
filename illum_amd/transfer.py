"""Mesh-field transfer between shape models on the GPU: the reference's
src/py/interpolate_mesh_field.py and src/py/build_mesh_transfer_tables.py
with their nearest-neighbour queries (scipy cKDTree) done by illum_knn.

Meshes are read as the reference reads them (trimesh.load of an OBJ): every
face of the file, FP64 vertices as written, centroids mean(V[F], 1).  Fields
and tables are read and written as the scripts' src/py/arma.py does: an
Armadillo-style header ("ARMA_MAT_BIN_<type>", then the dimensions, one
number for a vector) followed by the values in C (row-major) order -- not
Armadillo's own column-major order, which is why the interpolation script
has its --fix flag.  Run as
  python -m illum_amd.transfer interpolate SRC_MESH SRC_FIELD TGT_MESH TGT_FIELD
      [--mode nearest|weighted|area_weighted] [--neighbors 5] [--fix] [--save_layers]
  python -m illum_amd.transfer tables MESH_LIST [-n 5] [-o OUT]
"""
import argparse
import os

import numpy as np

from . import api
from ._lib import call

MODES = {"nearest": 0, "weighted": 1, "area_weighted": 2}


_CODES = {np.dtype(np.float64): b"FN008", np.dtype(np.uint32): b"IU004",
          np.dtype(np.uint64): b"IU008", np.dtype(np.int64): b"IS008",
          np.dtype(np.complex128): b"FC016"}


def arma_fromfile(path, dtype=np.float64):
    """The scripts' reader (arma.py:4-11): dims from the header, values in C
    order read as `dtype`; a file without the header is raw values."""
    with open(path, "rb") as f:
        line = f.readline()
        if line[:4] != b"ARMA":
            return np.fromfile(path, dtype)
        dims = tuple(map(int, f.readline().split()))
        return np.fromfile(f, dtype=dtype).reshape(*dims)


def arma_tofile(X, path):
    """The scripts' writer (arma.py:23-32): type code, then "n" for a vector
    or "rows cols", then the values in C order."""
    X = np.asarray(X)
    if X.dtype not in _CODES:
        raise ValueError("invalid datatype")
    with open(path, "wb") as f:
        f.write(b"ARMA_MAT_BIN_%s\n" % _CODES[X.dtype])
        if X.ndim == 1:
            f.write(b"%d\n" % X.size)
        elif X.ndim == 2:
            f.write(b"%d %d\n" % X.shape)
        else:
            raise ValueError("only 1-D and 2-D arrays")
        np.ascontiguousarray(X).tofile(f)


def load_mesh(path):
    """(V, F) of every face in an OBJ file, polygons fan-triangulated."""
    V, F = [], []
    for line in open(path):
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "v":
            V.append([float(x) for x in tok[1:4]])
        elif tok[0] == "f":
            idx = [int(t.split("/")[0]) for t in tok[1:]]
            idx = [i - 1 if i > 0 else len(V) + i for i in idx]
            F += [[idx[0], idx[t], idx[t + 1]] for t in range(1, len(idx) - 1)]
    return np.asarray(V, np.float64), np.asarray(F, np.int64)


def centroids(V, F):
    return np.ascontiguousarray(np.mean(V[F], 1))


def areas(V, F):
    v0, v1, v2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    return np.linalg.norm(np.cross(v1 - v0, v2 - v0), axis=1) / 2


def knn(src, tgt, k=1, ctx=None):
    """(idx nt x k int32, dist nt x k): the k nearest source points of every
    target point, nearest first (ties to the smaller index)."""
    ctx = ctx or api.default_context()
    src = np.ascontiguousarray(src, np.float64)
    tgt = np.ascontiguousarray(tgt, np.float64)
    nt = tgt.shape[0]
    idx = np.zeros((nt, k), np.int32)
    dist = np.zeros((nt, k))
    call("illum_knn", ctx.h, api._ptr(src), src.shape[0], api._ptr(tgt), nt, k, api._ptr(idx),
         api._ptr(dist), 0)
    return idx, dist


def transfer(P_src, X_src, P_tgt, mode="nearest", neighbors=5, A_src=None, A_tgt=None, ctx=None):
    """X_tgt (nlayers x nt) from X_src (nlayers x ns) on the GPU."""
    ctx = ctx or api.default_context()
    P_src = np.ascontiguousarray(P_src, np.float64)
    P_tgt = np.ascontiguousarray(P_tgt, np.float64)
    X_src = np.ascontiguousarray(np.atleast_2d(X_src), np.float64)
    nl, ns = X_src.shape
    if ns != P_src.shape[0]:
        raise api._shape_error("mismatched sizes")
    nt = P_tgt.shape[0]
    X_tgt = np.zeros((nl, nt))
    m = MODES[mode]
    As = None if A_src is None else np.ascontiguousarray(A_src, np.float64)
    At = None if A_tgt is None else np.ascontiguousarray(A_tgt, np.float64)
    call("illum_mesh_transfer", ctx.h, api._ptr(P_src), ns, api._ptr(P_tgt), nt, api._ptr(X_src), nl,
         m, neighbors, api._ptr(As), api._ptr(At), api._ptr(X_tgt))
    return X_tgt


def interpolate_mesh_field(source_mesh_path, source_field_path, target_mesh_path,
                           target_field_path, mode="nearest", fix=False, save_layers=False,
                           neighbors=5):
    """interpolate_mesh_field.py: the same inputs, transposition rules and
    outputs, the nearest-neighbour work on the GPU."""
    if mode not in MODES:
        raise ValueError('invalid interpolation mode "%s"' % mode)
    V_src, F_src = load_mesh(source_mesh_path)
    V_tgt, F_tgt = load_mesh(target_mesh_path)
    P_src, P_tgt = centroids(V_src, F_src), centroids(V_tgt, F_tgt)
    X_src = arma_fromfile(source_field_path)
    if fix:
        m, n = X_src.shape
        X_src = X_src.reshape(n, m).T
    transposed = P_src.shape[0] == X_src.shape[0]
    if transposed:
        X_src = X_src.T
    if P_src.shape[0] != X_src.shape[1]:
        raise api._shape_error("mismatched sizes")
    A_src = A_tgt = None
    if mode == "area_weighted":
        A_src, A_tgt = areas(V_src, F_src), areas(V_tgt, F_tgt)
    X_tgt = transfer(P_src, X_src, P_tgt, mode, neighbors, A_src, A_tgt)
    if transposed:
        X_tgt = X_tgt.T
    if fix:
        m, n = X_tgt.shape
        X_tgt = X_tgt.T.reshape(m, n)
    if save_layers:
        for layer in range(X_tgt.shape[0]):
            arma_tofile(X_tgt[layer, :].reshape(X_tgt.shape[1], 1), target_field_path % layer)
    else:
        arma_tofile(X_tgt, target_field_path)
    return X_tgt


def build_mesh_transfer_tables(paths, num_neighbors=5, output_path="."):
    """build_mesh_transfer_tables.py: for consecutive meshes of the list, the
    nearest previous-mesh facet of every facet (downsample_lut_i) and each
    facet's num_neighbors nearest facets of its own mesh (neighbor_lut_i),
    uint32, plus info.yaml.  Returns the two dicts of tables."""
    P_src = centroids(*load_mesh(paths[0]))
    down, neigh = {}, {}
    for i in range(1, len(paths)):
        P_tgt = centroids(*load_mesh(paths[i]))
        down[i] = knn(P_src, P_tgt, 1)[0][:, 0].astype(np.uint32)
        neigh[i] = knn(P_tgt, P_tgt, num_neighbors)[0].astype(np.uint32)
        P_src = P_tgt
    os.makedirs(output_path, exist_ok=True)
    for i in range(1, len(paths)):
        arma_tofile(down[i], os.path.join(output_path, "downsample_lut_%d.bin" % i))
        arma_tofile(neigh[i], os.path.join(output_path, "neighbor_lut_%d.bin" % i))
    import yaml
    with open(os.path.join(output_path, "info.yaml"), "w") as f:
        print(yaml.dump({
            "num_neighbors": num_neighbors,
            "paths": list(paths),
            "downsample_lut": ["downsample_lut_%d.bin" % i for i in range(1, len(paths))],
            "neighbor_lut": ["neighbor_lut_%d.bin" % i for i in range(1, len(paths))]}), file=f)
    return down, neigh


def main(argv=None):
    p = argparse.ArgumentParser(description="mesh-field transfer on the GPU")
    sub = p.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("interpolate")
    a.add_argument("source_mesh_path")
    a.add_argument("source_field_path")
    a.add_argument("target_mesh_path")
    a.add_argument("target_field_path")
    a.add_argument("--mode", default="nearest")
    a.add_argument("--fix", action="store_true")
    a.add_argument("--save_layers", action="store_true")
    a.add_argument("--neighbors", type=int, default=5)
    b = sub.add_parser("tables")
    b.add_argument("mesh_list_path")
    b.add_argument("-n", "--num_neighbors", type=int, default=5)
    b.add_argument("-o", "--output_path", default=".")
    args = p.parse_args(argv)
    if args.cmd == "interpolate":
        interpolate_mesh_field(args.source_mesh_path, args.source_field_path,
                               args.target_mesh_path, args.target_field_path, args.mode, args.fix,
                               args.save_layers, args.neighbors)
    else:
        with open(args.mesh_list_path) as f:
            paths = [line.strip() for line in f if line.strip()]
        build_mesh_transfer_tables(paths, args.num_neighbors, args.output_path)


if __name__ == "__main__":
    main()
