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
        self._pattern = None        # (rows, cols, dest) int32 device arrays, CSR order
        self.values_tensor = None   # values in CSR order when built by with_values()

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

    def pattern(self):
        """(rows, cols, dest): row and column of every stored entry in CSR order and
        its sliced-ELL slot, as int32 device tensors (built once)."""
        if self._pattern is None:
            from ._pack import sell_arrays
            H = sp.csr_matrix(self.host)
            _, _, _, _, dest = sell_arrays(H, return_dest=True)
            rows = np.repeat(np.arange(H.shape[0], dtype=np.int32), np.diff(H.indptr))
            dev = self.device
            self._pattern = tuple(torch.as_tensor(a.astype(np.int32), device=dev)
                                  for a in (rows, H.indices, dest))
        return self._pattern

    def with_values(self, values):
        """Same pattern, new stored values (length nnz, CSR order).

        A torch tensor is scattered on the device into a copy of the value
        buffer and kept as ``values_tensor`` (gradients with respect to it are
        available through ``MultilevelSolver.solve(b, A=...)`` and
        ``amjax_amd.cg``); other array-likes are repacked on the host.
        """
        if not torch.is_tensor(values):
            M = self.host.copy()
            M.data = np.asarray(values).astype(M.data.dtype)
            return SparseMatrix.from_scipy(M, self.dtype, self.device)
        if values.numel() != self.nnz:
            raise ValueError(f"expected {self.nnz} values, got {values.numel()}")
        v = values.detach().to(device=self.device, dtype=self.dtype).contiguous()
        rows, cols, dest = self.pattern()
        vbuf = self._vbuf.clone()
        _lib.check(_lib.load().amg_scatter_values(dest.data_ptr(), self.nnz, amg_dtype(self.dtype), v.data_ptr(),
                                                  vbuf.data_ptr() + self._val_offset_bytes(),
                                                  _lib.stream_ptr(self.device)), "amg_scatter_values")
        d = _lib.SellDesc()
        ctypes.memmove(ctypes.byref(d), ctypes.byref(self._desc), ctypes.sizeof(d))
        d.val = vbuf.data_ptr() + self._val_offset_bytes()
        host = self.host.copy()
        host.data = v.cpu().numpy().astype(host.data.dtype)
        out = SparseMatrix(host, self._ibuf, vbuf, d)
        out._pattern = self._pattern
        out.values_tensor = values
        return out

    def _val_offset_bytes(self):
        return int(self._desc.val or 0) - self._vbuf.data_ptr()

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
    x = x.contiguous()
    if x.data_ptr() % 16:
        # the kernels read rows of k columns as 16-byte vectors
        x = x.clone()
    return x
