"""Binary tensor (``.bt``) reader/writer for the problem file set.

Format (documented by the reference's Julia reader,
``tests/julia/modelFactory/src/modelFactory.jl:29-79``): three little-endian
uint64 dimensions (rows, cols, mats) followed by rows*cols*mats elements, column
major inside each matrix and matrices stored one after another, so element
(i, j, k) sits at ``i + j*rows + k*rows*cols`` (``src/definitions.cu:11-14``).
The suffix selects the element type: ``_u`` uint64, ``_d`` float64, ``_f``
float32 (``api/python/spock/tree.py:231-238``).  Negative integers (the root's
ancestor and event) wrap to 2**64-1 (``src/tree.cuh:129``).
"""
import os
import numpy as np

_DTYPES = {"u": np.uint64, "d": np.float64, "f": np.float32}


def as_tensor3(arr):
    """View any 0/1/2/3-D array as (rows, cols, mats) the way the writer lays it out."""
    a = np.asarray(arr)
    if a.ndim == 0:
        a = a.reshape(1, 1, 1)
    elif a.ndim == 1:
        a = a.reshape(-1, 1, 1)
    elif a.ndim == 2:
        a = a.reshape(a.shape[0], a.shape[1], 1)
    elif a.ndim != 3:
        raise ValueError(f"tensor rank {a.ndim} not supported")
    return a


def write_bt(path, arr, kind):
    """Write ``arr`` to ``path`` with element kind 'u', 'd' or 'f'."""
    dt = _DTYPES[kind]
    a = np.asarray(arr)
    if kind == "u":
        a = np.asarray(a, dtype=np.int64).astype(np.uint64)
    a = as_tensor3(a).astype(dt)
    r, c, m = a.shape
    with open(path, "wb") as fh:
        np.array([r, c, m], dtype="<u8").tofile(fh)
        a.reshape(-1, order="F").astype(np.dtype(dt).newbyteorder("<")).tofile(fh)


def read_bt(path):
    """Read a ``.bt`` file; returns a (rows, cols, mats) array (Fortran order semantics)."""
    kind = os.path.basename(path).rsplit(".", 1)[0][-1]
    dt = _DTYPES[kind]
    with open(path, "rb") as fh:
        dims = np.fromfile(fh, dtype="<u8", count=3)
        r, c, m = (int(x) for x in dims)
        data = np.fromfile(fh, dtype=dt, count=r * c * m)
    if data.size != r * c * m:
        raise ValueError(f"{path}: expected {r*c*m} elements, found {data.size}")
    return data.reshape((r, c, m), order="F")


def read_flat(path):
    """Read a ``.bt`` file as the flat element sequence it stores."""
    return read_bt(path).reshape(-1, order="F")
