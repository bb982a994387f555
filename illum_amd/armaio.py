"""Armadillo binary matrix files (the reference's on-disk format).

Dense files: text header "ARMA_MAT_BIN_FN008\\n" then "<n_rows> <n_cols>\\n", then
n_rows*n_cols FP64 values in column-major order -- what arma::mat::save writes
(arma_util.hpp:10-13) and what src/py/arma.py:4-35 reads.
"""
import numpy as np

_DT = {b"FN008": np.float64, b"IU004": np.uint32, b"IU008": np.uint64, b"IS008": np.int64}


def save_mat(X, path):
    X = np.asarray(X)
    if X.ndim == 1:
        X = X.reshape(-1, 1)
    code = {np.dtype(v): k for k, v in _DT.items()}[X.dtype]
    with open(path, "wb") as f:
        f.write(b"ARMA_MAT_BIN_" + code + b"\n")
        f.write(b"%d %d\n" % X.shape)
        f.write(np.asfortranarray(X).tobytes(order="F"))


def load_mat(path):
    with open(path, "rb") as f:
        head = f.readline().strip()
        if not head.startswith(b"ARMA_MAT_BIN_"):
            raise ValueError(f"{path}: not an Armadillo dense binary matrix")
        dt = _DT[head[len(b"ARMA_MAT_BIN_"):]]
        r, c = map(int, f.readline().split())
        data = np.frombuffer(f.read(), dtype=dt, count=r * c)
    return data.reshape((r, c), order="F").copy()
