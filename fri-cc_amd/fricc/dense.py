"""Dense two-operand contractions on the GPU's matrix cores
(``fricc_dense_contract``, fri-cc_amd/csrc/dense.hip).

``einsum(spec, A, B, out=None, alpha=1, beta=0, compute=None, c_src=None)``
computes ``out = alpha * einsum(spec, A, B) + beta * C`` with C = ``c_src``
(a tensor of out's shape, strides and dtype that does not overlap it) or
``out`` itself, for torch CUDA tensors of
any strides (permuted views and broadcasts need no copies), FP32 or FP64
elements (converted to the compute type on load).  ``compute``: torch.float32
(FP32 MFMA) or torch.float64 (FP64 MFMA); default FP64 when both inputs are
FP64, else FP32.  This replaces the NumPy/PySCF einsums of the reference's
dense update terms (fricc/friccsd.py:144-207, 309-376)."""
import ctypes

from . import _capi


def _torch():
    import torch

    return torch


def _prec(t):
    torch = _torch()
    if t.dtype == torch.float64:
        return _capi.FP64
    if t.dtype == torch.float32:
        return _capi.FP32
    raise TypeError(f"dense einsum operands must be float32 or float64, not {t.dtype}")


def _arr(vals):
    vals = [int(v) for v in vals]
    return (ctypes.c_int64 * max(1, len(vals)))(*vals)


def output_shape(spec, A, B):
    ins, outs = spec.split("->")
    sa, sb = ins.split(",")
    ext = {}
    for s, t in ((sa, A), (sb, B)):
        for c, n in zip(s, t.shape):
            ext[c] = int(n)
    return tuple(ext[c] for c in outs)


# When a dict: every call adds its algorithmic work to it, keyed by spec
# ("flops": 2 * batch * M * N * K; "bytes": A and B read once, C written once
# and read once more when beta != 0) -- the figures a roofline is built from.
COUNTER = None


def _count(spec, A, B, out, beta, compute):
    ins, outs = spec.split("->")
    sa, sb = ins.split(",")
    ext = {}
    for s, t in ((sa, A), (sb, B), (outs, out)):
        for c, n in zip(s, t.shape):
            ext[c] = int(n)
    n_out = 1
    for c in outs:
        n_out *= ext[c]
    k = 1
    for c in set(sa) & set(sb):
        if c not in outs:
            k *= ext[c]
    rec = COUNTER.setdefault(spec, {"calls": 0, "flops": 0.0, "bytes": 0.0, "fp64": compute == _torch().float64})
    rec["calls"] += 1
    rec["flops"] += 2.0 * n_out * k
    rec["bytes"] += (A.numel() * A.element_size() + B.numel() * B.element_size()
                     + out.numel() * out.element_size() * (2 if beta != 0.0 else 1))


def einsum(spec, A, B, out=None, alpha=1.0, beta=0.0, compute=None, stream=None, c_src=None):
    torch = _torch()
    if compute is None:
        compute = torch.float64 if (A.dtype == torch.float64 and B.dtype == torch.float64) else torch.float32
    if out is None:
        if beta != 0.0:
            raise ValueError("beta != 0 needs an output tensor")
        out = torch.empty(output_shape(spec, A, B), dtype=compute, device=A.device)
    for t in (A, B, out):
        if not t.is_cuda:
            raise ValueError("dense einsum operands must be CUDA tensors")
    cp = _capi.FP64 if compute == torch.float64 else _capi.FP32
    a_args = (ctypes.c_void_p(A.data_ptr()), _prec(A), A.dim(), _arr(A.shape), _arr(A.stride()),
              ctypes.c_void_p(B.data_ptr()), _prec(B), B.dim(), _arr(B.shape), _arr(B.stride()), float(beta))
    c_args = (ctypes.c_void_p(out.data_ptr()), _prec(out), out.dim(), _arr(out.shape), _arr(out.stride()),
              cp, _capi.stream_ptr(stream))
    if c_src is None:
        _capi.check(_capi.lib().fricc_dense_contract(spec.encode(), float(alpha), *a_args, *c_args))
    else:
        if (not c_src.is_cuda or c_src.shape != out.shape or c_src.stride() != out.stride()
                or c_src.dtype != out.dtype):
            raise ValueError("c_src must be a CUDA tensor of out's shape, strides and dtype")
        _capi.check(_capi.lib().fricc_dense_contract_from(spec.encode(), float(alpha), *a_args,
                                                          ctypes.c_void_p(c_src.data_ptr()), *c_args))
    if COUNTER is not None:
        _count(spec, A, B, out, beta, compute)
    return out
