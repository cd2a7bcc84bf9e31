"""Host-side packing of a hierarchy into the device layout.

Sparse matrices become sliced ELLPACK with 64-row slices (one slice per
wavefront): slice ``s`` covers rows ``64s .. 64s+63`` and has width
``w_s = max row length in the slice``; entry ``j`` of row ``64s + r`` is stored
at ``sptr[s] + 64*j + r``, unused slots hold ``col = -1, val = 0``.  Rows keep
their CSR (ascending column) order, so a row sum visits entries in the same
order as a CSR product.  All integer arrays of a hierarchy share one int32
buffer and all values one value buffer ("level-packed").
"""

import numpy as np
import scipy.sparse as sp

SLICE = 64


def sell_arrays(M, return_dest=False):
    """Return (sptr, col, val, nnz) of the 64-row sliced-ELL form of CSR ``M``
    (and, with ``return_dest``, the slot of every CSR entry)."""
    M = sp.csr_matrix(M)
    M.sum_duplicates()
    M.sort_indices()
    n = M.shape[0]
    nsl = (n + SLICE - 1) // SLICE
    lens = np.diff(M.indptr).astype(np.int64)
    lens_p = np.zeros(nsl * SLICE, dtype=np.int64)
    lens_p[:n] = lens
    width = lens_p.reshape(nsl, SLICE).max(axis=1) if nsl else np.zeros(0, dtype=np.int64)
    sptr = np.zeros(nsl + 1, dtype=np.int64)
    sptr[1:] = np.cumsum(width * SLICE)
    total = int(sptr[-1])
    if total >= 2 ** 31:
        raise ValueError("matrix too large for int32 sliced-ELL offsets")
    col = np.full(total, -1, dtype=np.int32)
    val = np.zeros(total, dtype=M.data.dtype)
    rows = np.repeat(np.arange(n, dtype=np.int64), lens)
    j = np.arange(M.nnz, dtype=np.int64) - M.indptr[rows].astype(np.int64)
    dest = sptr[rows // SLICE] + SLICE * j + (rows % SLICE)
    col[dest] = M.indices
    val[dest] = M.data
    if return_dest:
        return sptr.astype(np.int32), col, val, int(M.nnz), dest.astype(np.int32)
    return sptr.astype(np.int32), col, val, int(M.nnz)


def max_width(sptr):
    """Widest slice of a sliced-ELL matrix (slots per row)."""
    return int(np.diff(sptr).max() // SLICE) if sptr.size > 1 else 0


class Packer:
    """Accumulates int32 and value arrays and hands out offsets."""

    def __init__(self, dtype):
        self.dtype = np.dtype(dtype)
        self.ints = []
        self.vals = []
        self.ni = 0
        self.nv = 0

    def add_int(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        off = self.ni
        self.ints.append(a)
        # keep every array 16-byte aligned
        self.ni += a.size + (-a.size) % 4
        if (-a.size) % 4:
            self.ints.append(np.zeros((-a.size) % 4, dtype=np.int32))
        return off

    def add_val(self, a):
        a = np.ascontiguousarray(np.ravel(a), dtype=self.dtype)
        off = self.nv
        self.vals.append(a)
        pad = (-a.size) % 4
        self.nv += a.size + pad
        if pad:
            self.vals.append(np.zeros(pad, dtype=self.dtype))
        return off

    def add_sell(self, M):
        sptr, col, val, nnz = sell_arrays(sp.csr_matrix(M).astype(self.dtype))
        return {
            "nrows": int(M.shape[0]), "ncols": int(M.shape[1]),
            "sptr_off": self.add_int(sptr), "col_off": self.add_int(col),
            "val_off": self.add_val(val), "nnz": nnz, "padded": int(col.size),
            "max_width": max_width(sptr),
        }

    def int_buffer(self):
        return np.concatenate(self.ints) if self.ints else np.zeros(4, dtype=np.int32)

    def val_buffer(self):
        return np.concatenate(self.vals) if self.vals else np.zeros(4, dtype=self.dtype)
