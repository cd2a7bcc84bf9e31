"""Device sparse matrix: 64-row sliced ELLPACK on a PyTorch-ROCm device.

Stands in for the reference's ``jax.experimental.sparse.BCOO`` operands
(amjax/multilevel.py:548-554 converts every PyAMG matrix with
``BCOO.from_scipy_sparse``).  ``A @ x`` runs the gfx950 SpMV kernel through the
C-ABI; there is no CPU path.
"""

import ctypes

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from ._pack import Packer

_NP_OF = {torch.float64: np.float64, torch.float32: np.float32}
_TORCH_OF = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32}


def as_torch_dtype(dtype):
    """Accept torch / numpy / string dtypes; return torch.float32 or torch.float64."""
    if dtype is None:
        return None
    if isinstance(dtype, torch.dtype):
        td = dtype
    else:
        td = _TORCH_OF.get(np.dtype(dtype))
    if td not in (torch.float32, torch.float64):
        raise ValueError(f"unsupported dtype {dtype!r}: use float32 or float64")
    return td


def np_dtype(td):
    return _NP_OF[td]


def amg_dtype(td):
    return _lib.AMG_F64 if td == torch.float64 else _lib.AMG_F32


def default_device():
    if not torch.cuda.is_available():
        raise RuntimeError("amjax_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU execution path")
    return torch.device("cuda", torch.cuda.current_device())


class SparseMatrix:
    """A sliced-ELL matrix resident on the GPU, plus its host CSR copy.

    Instances are either standalone (own buffers) or views into the
    level-packed buffers of a ``MultilevelSolver``.
    """

    def __init__(self, host, ibuf, vbuf, desc):
        self.host = host            # scipy CSR in the device dtype
        self._ibuf = ibuf           # torch int32 (device)
        self._vbuf = vbuf           # torch float (device)
        self._desc = desc           # _lib.SellDesc

    @classmethod
    def from_scipy(cls, M, dtype=None, device=None):
        M = sp.csr_matrix(M)
        td = as_torch_dtype(dtype) or as_torch_dtype(M.dtype)
        host = M.astype(np_dtype(td))
        host.sum_duplicates()
        host.sort_indices()
        pk = Packer(np_dtype(td))
        d = pk.add_sell(host)
        device = torch.device(device) if device is not None else default_device()
        ibuf = torch.from_numpy(pk.int_buffer()).to(device)
        vbuf = torch.from_numpy(pk.val_buffer()).to(device)
        return cls(host, ibuf, vbuf, _lib.sell_desc(d, ibuf, vbuf))

    # -- array-like surface used by the reference (shape, dtype, nnz / nse, todense, T)
    @property
    def shape(self):
        return (int(self._desc.nrows), int(self._desc.ncols))

    @property
    def dtype(self):
        return self._vbuf.dtype

    @property
    def device(self):
        return self._vbuf.device

    @property
    def nnz(self):
        return int(self._desc.nnz)

    nse = nnz

    def todense(self):
        return torch.as_tensor(self.host.toarray(), device=self.device)

    @property
    def T(self):
        return SparseMatrix.from_scipy(self.host.T, self.dtype, self.device)

    def astype(self, dtype):
        td = as_torch_dtype(dtype)
        if td == self.dtype:
            return self
        return SparseMatrix.from_scipy(self.host, td, self.device)

    def with_values(self, values):
        """Same pattern (host CSR order), new values (numpy or torch, length nnz)."""
        vals = values.detach().cpu().numpy() if torch.is_tensor(values) else np.asarray(values)
        M = self.host.copy()
        M.data = vals.astype(M.data.dtype)
        return SparseMatrix.from_scipy(M, self.dtype, self.device)

    def matmul(self, x, out=None, add=False):
        x = _prep(x, self.dtype, self.device)
        if x.shape[0] != self.shape[1]:
            raise ValueError(f"shape mismatch: {self.shape} @ {tuple(x.shape)}")
        k = 1 if x.dim() == 1 else x.shape[1]
        shape = (self.shape[0],) if x.dim() == 1 else (self.shape[0], k)
        if out is None:
            out = torch.empty(shape, dtype=self.dtype, device=self.device)
        lib = _lib.load()
        if out.data_ptr() == x.data_ptr():
            raise ValueError("out may not alias x")
        _lib.check(lib.amg_sell_spmv(ctypes.byref(self._desc), amg_dtype(self.dtype), x.data_ptr(),
                                     out.data_ptr(), k, int(add), _lib.stream_ptr(self.device)),
                   "amg_sell_spmv")
        return out

    def __matmul__(self, x):
        return self.matmul(x)

    def __repr__(self):
        return f"SparseMatrix(shape={self.shape}, nnz={self.nnz}, dtype={self.dtype}, device={self.device})"


def _prep(x, dtype, device):
    """Contiguous device tensor of `dtype` ((n,) or (n, k))."""
    if not torch.is_tensor(x):
        x = torch.as_tensor(np.asarray(x))
    if x.device != device or x.dtype != dtype:
        x = x.to(device=device, dtype=dtype)
    if x.dim() not in (1, 2):
        raise ValueError("expected a vector (n,) or a block (n, k)")
    return x.contiguous()
