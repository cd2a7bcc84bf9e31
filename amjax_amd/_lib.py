"""ctypes binding of ``libamjax_amd.so`` (C-ABI declared in include/amjax_amd.h).

The library must be built (``python -m amjax_amd.build`` or
``__graft_entry__.build()``); there is no fallback path: a missing library or
a missing GPU raises.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamjax_amd.so")

AMG_F64, AMG_F32 = 0, 1
AMG_SMOOTH_NONE, AMG_SMOOTH_JACOBI, AMG_SMOOTH_POLY = 0, 1, 2
COARSE_KIND = {"pinv": 0, "lu": 1, "qr": 2, "jacobi": 3}
AMG_OPT_FUSE = 1
AMG_OPT_AUTOTUNE = 2
AMG_OPT_VECTOR = 3
AMG_OPT_TAIL = 4
AMG_BATCH_OPT_GRID = 1
AMG_BATCH_OPT_SPLIT_KP = 2
AMG_BATCH_OPT_AUTOTUNE = 3
AMG_BATCH_OPT_CYCLE = 4
AMG_BATCH_OPT_CPL = 5
AMG_TIME_COLD = 256
AMG_NOALLOC_SCRATCH_MB = 64

# every symbol include/amjax_amd.h declares
EXPORTS = ("amg_create", "amg_destroy", "amg_set_level0", "amg_set_option", "amg_cycle", "amg_solve", "amg_pcg",
           "amg_sell_spmv", "amg_sell_jacobi", "amg_dense_solve", "amg_grad_values", "amg_scatter_values",
           "amg_time_kernel", "amg_tail_level", "amg_info", "amg_last_error", "amg_batch_create", "amg_batch_pcg",
           "amg_batch_set_option", "amg_batch_info", "amg_batch_destroy", "amg_tune_report", "amg_sell_poly", "amg_bicgstab",
           "amg_pcg_ir", "amg_gmres", "amg_cycle_noalloc", "amg_pcg_ru")


class SellDesc(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int32), ("ncols", ctypes.c_int32),
                ("sptr", ctypes.c_void_p), ("col", ctypes.c_void_p), ("val", ctypes.c_void_p),
                ("nnz", ctypes.c_int64), ("padded", ctypes.c_int64), ("max_width", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class LevelDesc(ctypes.Structure):
    _fields_ = [("A", SellDesc), ("P", SellDesc), ("R", SellDesc), ("D", SellDesc), ("dinv", ctypes.c_void_p),
                ("pre_kind", ctypes.c_int32), ("pre_iters", ctypes.c_int32), ("pre_omega", ctypes.c_double),
                ("post_kind", ctypes.c_int32), ("post_iters", ctypes.c_int32),
                ("post_omega", ctypes.c_double), ("pre_coef", ctypes.c_void_p), ("pre_ncoef", ctypes.c_int32),
                ("reserved0", ctypes.c_int32), ("post_coef", ctypes.c_void_p), ("post_ncoef", ctypes.c_int32),
                ("reserved1", ctypes.c_int32)]


class CoarseDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("n", ctypes.c_int32), ("f0", ctypes.c_void_p),
                ("f1", ctypes.c_void_p), ("iters", ctypes.c_int32), ("omega", ctypes.c_double),
                ("factor_f64", ctypes.c_int32), ("reserved", ctypes.c_int32)]


_lib = None


class AmgError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AmgError(f"{LIB_PATH} is missing: build it with `python -m amjax_amd.build` "
                       "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    c_int, c_char, c_double = ctypes.c_int, ctypes.c_char, ctypes.c_double
    sig = {
        "amg_create": [ctypes.POINTER(LevelDesc), c_int, ctypes.POINTER(CoarseDesc), c_int,
                       ctypes.POINTER(SellDesc), ctypes.POINTER(vp)],
        "amg_destroy": [vp],
        "amg_set_level0": [vp, ctypes.POINTER(SellDesc)],
        "amg_set_option": [vp, c_int, ctypes.c_longlong],
        "amg_cycle": [vp, c_char, c_int, vp, vp, c_int, c_int, vp],
        "amg_cycle_noalloc": [vp, c_char, c_int, vp, vp, c_int, c_int, vp],
        "amg_solve": [vp, c_char, c_int, vp, vp, c_int, c_double, c_int, c_int, ip, dp, vp],
        "amg_pcg": [vp, c_char, c_int, vp, vp, c_int, c_double, c_double, c_int, c_int, c_int, ip, dp, vp],
        "amg_bicgstab": [vp, c_char, c_int, vp, vp, c_int, c_double, c_double, c_int, c_int, ip, dp, vp],
        "amg_pcg_ir": [vp, c_char, c_int, vp, vp, c_int, c_double, c_double, c_int, c_int, c_int, ip, ip, dp, vp],
        "amg_pcg_ru": [vp, c_char, c_int, vp, vp, c_int, c_double, c_double, c_double, c_int, c_int, ip, ip, dp, vp],
        "amg_gmres": [vp, c_char, c_int, vp, vp, c_int, c_double, c_double, c_int, c_int, c_int, c_int, ip, dp, vp],
        "amg_sell_spmv": [ctypes.POINTER(SellDesc), c_int, vp, vp, c_int, c_int, vp],
        "amg_sell_jacobi": [ctypes.POINTER(SellDesc), vp, c_int, vp, vp, vp, c_int, c_double, vp],
        "amg_dense_solve": [ctypes.POINTER(CoarseDesc), c_int, vp, vp, c_int, vp],
        "amg_grad_values": [vp, vp, ctypes.c_longlong, c_int, vp, vp, c_int, c_double, vp, vp],
        "amg_scatter_values": [vp, ctypes.c_longlong, c_int, vp, vp, vp],
        "amg_time_kernel": [vp, c_int, c_int, c_int, c_int, ctypes.POINTER(ctypes.c_float), vp],
        "amg_info": [vp, ip, ip],
        "amg_tail_level": [vp, ip],
        "amg_last_error": [ctypes.c_char_p, c_int],
        "amg_batch_create": [ctypes.POINTER(vp), ip, c_int, ctypes.POINTER(vp)],
        "amg_batch_pcg": [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), c_double, c_double, c_int, c_int, ip, dp, vp],
        "amg_batch_set_option": [vp, c_int, ctypes.c_longlong],
        "amg_batch_info": [vp, ip, ip],
        "amg_batch_destroy": [vp],
        "amg_tune_report": [ctypes.c_char_p, c_int],
        "amg_sell_poly": [ctypes.POINTER(SellDesc), c_int, dp, c_int, c_int, vp, vp, vp, vp, c_int, c_int, vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = c_int
    _lib = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().amg_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(code, what):
    if code == 0:
        return
    msg = last_error()
    if code == -1:
        raise ValueError(f"{what}: {msg}")
    raise AmgError(f"{what} failed ({code}): {msg}")


def tune_report():
    """The autotuner's cache: a list of (key, plan, us_default, us_best) tuples with
    key = (kind, dtype, kp(k), entries/row, narrow, size class) and
    plan = (kp, wide, log2 rows per unit, grid)."""
    lib = load()
    n = lib.amg_tune_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib.amg_tune_report(buf, n + 1)
    out = []
    for line in buf.value.decode().splitlines():
        lhs, rhs = line.split(" -> ")
        kind, dt, kp, npr, narrow, size = lhs.split()
        pk, wide, subl, grid, ud, ub = rhs.split()
        out.append(((kind, dt, int(kp), int(npr), int(narrow), int(size)), (int(pk), int(wide), int(subl), int(grid)),
                    float(ud), float(ub)))
    return out


def stream_ptr(device=None):
    """The current torch stream of ``device`` as a ``hipStream_t``."""
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def sell_desc(d, ibuf, vbuf):
    """SellDesc with absolute device pointers from a Packer record and the uploaded buffers."""
    ib = ibuf.data_ptr()
    vb = vbuf.data_ptr()
    es = vbuf.element_size()
    return SellDesc(nrows=d["nrows"], ncols=d["ncols"], sptr=ib + 4 * d["sptr_off"],
                    col=ib + 4 * d["col_off"], val=vb + es * d["val_off"], nnz=d["nnz"],
                    padded=d["padded"], max_width=d["max_width"], reserved=0)
