"""ctypes binding of libfricc_hip.so (C ABI declared in include/fricc_capi.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C fri-cc_amd/csrc``).  There is no fallback: if the library is missing,
importing anything that needs it raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfricc_hip.so")

FRICC_OK, FRICC_EINVAL, FRICC_EDEVICE, FRICC_ENOMEM, FRICC_EPROB, FRICC_ENONFINITE, FRICC_ESAMPLE = range(7)
LARGEST, FRI_SYSTEMATIC, FRI_PIVOTAL = 0, 1, 2
FP64, FP32, BF16, BF16X1 = 0, 1, 2, 3
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16": BF16, "bf16x1": BF16X1}

# every symbol include/fricc_capi.h declares
EXPORTS = (
    "fricc_abi_version", "fricc_last_error", "fricc_compress", "fricc_sample_systematic", "fricc_sparse_create", "fricc_sparse_destroy",
    "fricc_sparse_compress", "fricc_sparse_set", "fricc_sparse_info", "fricc_sparse_get", "fricc_sparse_device",
    "fricc_contract", "fricc_contract_prepared", "fricc_contract_spec", "fricc_eri_build", "fricc_set_gemm_order", "fricc_set_gemm_split", "fricc_dense_zpair",
    "fricc_sparse_prepare", "fricc_split_bf16", "fricc_set_bf16_passes", "fricc_t2_finalize",
    "fricc_set_bf16_view_width", "fricc_stream_copy", "fricc_sparse_compress_async", "fricc_sparse_status",
    "fricc_workspace_select", "fricc_workspace_release", "fricc_split_bf16_lincomb",
    "fricc_sparse_compress_pair_async", "fricc_sparse_compress_multi_async", "fricc_workspace_pin",
    "fricc_compress_noalloc", "fricc_compress_noalloc_max_n", "fricc_dense_contract",
    "fricc_split_bf16_hi", "fricc_sample_pivotal", "fricc_parallel_sample", "fricc_dense_contract_from",
    "fricc_split_bf16_lincomb3",
)


class CompressInfo(ctypes.Structure):
    _fields_ = [
        ("n_D", ctypes.c_int64),
        ("n_sampled", ctypes.c_int64),
        ("n_notfound", ctypes.c_int64),
        ("n_candidates", ctypes.c_int64),
        ("norm0", ctypes.c_double),
        ("remaining_norm", ctypes.c_double),
        ("p_total", ctypes.c_double),
        ("nonfinite", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class FriccError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"libfricc_hip status {status}: {msg}")
        self.status = status


class FriccValueError(FriccError, ValueError):
    pass


_LIB = None

I64x4 = ctypes.c_int64 * 4
I64x8 = ctypes.c_int64 * 8
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64


def lib():
    """Load libfricc_hip.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libfricc_hip.so not built: {LIB_PATH} missing (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    L.fricc_abi_version.restype = ctypes.c_int
    L.fricc_last_error.restype = ctypes.c_char_p
    L.fricc_compress.argtypes = [_vp, _i64, _i64, ctypes.c_int, ctypes.c_double, _vp, _vp,
                                 ctypes.POINTER(CompressInfo), _vp]
    L.fricc_compress_noalloc.argtypes = [_vp, _i64, _i64, ctypes.c_int, ctypes.c_double, _vp, _vp,
                                         ctypes.POINTER(CompressInfo), _vp]
    L.fricc_compress_noalloc_max_n.restype = _i64
    L.fricc_compress_noalloc_max_n.argtypes = []
    L.fricc_sample_systematic.argtypes = [_vp, _i64, _i64, ctypes.c_double, _vp, ctypes.POINTER(_i64), _vp]
    L.fricc_sample_pivotal.argtypes = [_vp, _i64, _i64, _vp, ctypes.c_double, _vp, _vp]
    L.fricc_parallel_sample.argtypes = [_vp, _i64, _i64, _i64, _vp, ctypes.c_double, _vp, _vp]
    L.fricc_sparse_create.argtypes = [I64x4, ctypes.POINTER(_vp)]
    L.fricc_sparse_destroy.argtypes = [_vp]
    L.fricc_sparse_compress.argtypes = [_vp, _vp, _i64, ctypes.c_int, ctypes.c_double, _vp]
    L.fricc_sparse_set.argtypes = [_vp, _vp, _vp, _i64, _vp]
    L.fricc_sparse_info.argtypes = [_vp, _vp, _vp, _vp]
    L.fricc_sparse_get.argtypes = [_vp, _vp, _vp, _vp]
    L.fricc_sparse_device.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]
    L.fricc_contract.argtypes = [ctypes.c_char_p, _vp, I64x4, I64x4, _vp, _vp, I64x4, I64x4, ctypes.c_int, _vp]
    L.fricc_contract_prepared.argtypes = [ctypes.c_char_p, _vp, I64x4, _vp, _vp, I64x4, I64x4, ctypes.c_int, _vp]
    L.fricc_contract_spec.argtypes = [ctypes.c_char_p, ctypes.c_double, _vp, ctypes.c_int, I64x4, I64x4, _vp, _vp,
                                      ctypes.c_int, I64x4, I64x4, ctypes.c_int, _vp]
    _pi = ctypes.POINTER(_i64)
    L.fricc_dense_contract.argtypes = [ctypes.c_char_p, ctypes.c_double, _vp, ctypes.c_int, ctypes.c_int, _pi, _pi,
                                       _vp, ctypes.c_int, ctypes.c_int, _pi, _pi, ctypes.c_double, _vp, ctypes.c_int,
                                       ctypes.c_int, _pi, _pi, ctypes.c_int, _vp]
    L.fricc_dense_contract_from.argtypes = [ctypes.c_char_p, ctypes.c_double, _vp, ctypes.c_int, ctypes.c_int, _pi,
                                            _pi, _vp, ctypes.c_int, ctypes.c_int, _pi, _pi, ctypes.c_double, _vp, _vp,
                                            ctypes.c_int, ctypes.c_int, _pi, _pi, ctypes.c_int, _vp]
    L.fricc_eri_build.argtypes = [_vp, _i64, _i64, I64x8, I64x4, _vp, _vp]
    L.fricc_set_gemm_order.argtypes = [ctypes.c_int]
    L.fricc_set_gemm_split.argtypes = [ctypes.c_int]
    L.fricc_dense_zpair.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 6
    L.fricc_split_bf16.argtypes = [_vp, ctypes.c_int, _i64, _i64, _vp, _vp]
    L.fricc_split_bf16_hi.argtypes = [_vp, ctypes.c_int, _i64, _i64, _vp, _vp]
    L.fricc_split_bf16_lincomb.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _i64, _i64, _vp,
                                           _vp]
    L.fricc_split_bf16_lincomb3.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _i64, _i64,
                                            _vp, _vp, _vp, _vp]
    L.fricc_set_bf16_passes.argtypes = [ctypes.c_int]
    L.fricc_set_bf16_view_width.argtypes = [ctypes.c_int]
    L.fricc_stream_copy.argtypes = [_vp, _vp, _i64, _vp]
    L.fricc_sparse_compress_async.argtypes = [_vp, _vp, _i64, ctypes.c_int, _vp, _vp]
    L.fricc_sparse_compress_pair_async.argtypes = [_vp, _vp, _vp, _i64, ctypes.c_int, _vp, _vp, _vp]
    L.fricc_sparse_compress_multi_async.argtypes = [_vp, ctypes.c_int, _vp, _i64, ctypes.c_int, _vp, _vp]
    L.fricc_sparse_status.argtypes = [_vp, _vp]
    L.fricc_workspace_select.argtypes = [ctypes.c_int]
    L.fricc_workspace_release.argtypes = [ctypes.c_int]
    L.fricc_workspace_pin.argtypes = [ctypes.c_int, ctypes.c_int]
    L.fricc_t2_finalize.argtypes = [ctypes.c_int, ctypes.c_int] + [_vp] * 3 + [ctypes.c_int] + [_vp] * 8 + [_vp]
    L.fricc_sparse_prepare.argtypes = [_vp, ctypes.c_char_p, I64x4, I64x4, ctypes.c_int, _vp]
    for name in EXPORTS:
        getattr(L, name)  # every declared symbol must resolve
    if L.fricc_abi_version() != 1:
        raise ImportError("libfricc_hip.so ABI version mismatch")
    _LIB = L
    return L


def check(status):
    if status != FRICC_OK:
        msg = lib().fricc_last_error().decode(errors="replace")
        if status == FRICC_EINVAL:
            raise FriccValueError(status, msg)
        raise FriccError(status, msg)


def stream_ptr(stream=None):
    """hipStream_t of the given torch stream (default: torch's current stream)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def i64x4(seq):
    seq = [int(v) for v in seq]
    if len(seq) != 4:
        raise ValueError("expected 4 values")
    return I64x4(*seq)
