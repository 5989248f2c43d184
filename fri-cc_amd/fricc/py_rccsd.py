"""Drop-in replacement for the reference's pybind11 module ``fricc.py_rccsd``.

Mirrors ``PYBIND11_MODULE(py_rccsd, m)`` (src/py_rccsd.cpp:16-55): same names,
positional order, defaults and keyword spellings (including ``ouput`` and the
swapped ``get_d_largest`` keywords), backed by the gfx950 kernels of
``libfricc_hip.so`` through its C ABI (include/fricc_capi.h).

Inputs may be NumPy float64 arrays (copied to the GPU and, for outputs, back
in place) or torch CUDA float64 tensors (used in place, no copies).

Deviations, all deliberate:
* The reference terminates the process (``exit(EXIT_FAILURE)``) on a bad
  compression string, an unknown term, p > 1 or an out-of-range index; here
  those raise ``ValueError`` / ``RuntimeError``.
* ``sampling_method="pivotal"`` (the pybind default, src/py_rccsd.cpp:30,39)
  runs the pivotal chain (src/fri_utils.cpp:63-127) on the GPU, all steps in
  parallel on the exact prefix of p.  The reference draws rand()/RAND_MAX at
  every step of the chain; here one uniform per call (drawn like the
  systematic one, or given as ``u``) keys a SplitMix64 stream of step draws.
  ``sample_pivotal`` also takes the step uniforms as an array.
* An unknown sampling method raises ``ValueError`` (the reference silently
  emits index 0, src/fri_utils.cpp:443-444).
"""
import ctypes

import numpy as np

from . import _capi
from .uniforms import draw_libc

__all__ = [
    "SparseTensor4d", "contract_DTSpT", "contract_DTSpT_experimental", "argsort", "sample_pivotal",
    "sample_systematic", "parallel_sample", "make_probability_vector", "get_d_largest", "fri_compression",
]

_TERMS = ("0101", "2323", "1302", "1202", "1303", "1203", "1323", "0112", "0113", "0313", "0312")


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("fricc.py_rccsd needs a GPU (torch.cuda is not available)")
    return torch


def _is_torch(a):
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(a, torch.Tensor)


def _as_device_vector(x, name="x"):
    """float64 1-D contiguous CUDA tensor for x (numpy float64 or torch float64)."""
    torch = _torch()
    if _is_torch(x):
        if x.dtype != torch.float64:
            raise TypeError(f"{name} must be float64")
        return x.reshape(-1).contiguous().cuda()
    if not isinstance(x, np.ndarray) or x.dtype != np.float64:
        raise TypeError(f"{name} must be a float64 numpy array (noconvert)")
    return torch.from_numpy(np.ascontiguousarray(x).reshape(-1)).cuda()


def _draw(u):
    return draw_libc() if u is None else float(u)


def _method(compression, sampling_method="systematic"):
    if compression == "largest":
        return _capi.LARGEST
    if compression != "fri":
        raise ValueError("compression must be 'largest' or 'fri'")
    if sampling_method == "systematic":
        return _capi.FRI_SYSTEMATIC
    if sampling_method == "pivotal":
        return _capi.FRI_PIVOTAL
    raise ValueError("sampling_method must be 'pivotal' or 'systematic'")


class SparseTensor4d:
    """Device-resident SparseTensor4d (include/sparse_tensor.hpp:13-152).

    ``SparseTensor4d(tensor_arr, dims, m, compression="fri",
    sampling_method="pivotal", verbose=False)`` compresses a dense (o,o,v,v)
    tensor; ``SparseTensor4d(dims, nnz)`` makes an empty one.  Entries are kept
    on the GPU in compression order: D (|x| descending) then the samples
    (systematic: ascending flat index; pivotal: the chain's emission order).
    ``u`` (keyword only, not in the reference) supplies the uniform instead of
    drawing ``rand()/RAND_MAX``.
    """

    def __init__(self, *args, **kwargs):
        self._h = ctypes.c_void_p()
        self._host = None
        self.info = {}
        if len(args) == 2 and not kwargs and not isinstance(args[0], np.ndarray) and not _is_torch(args[0]):
            # SparseTensor4d(dims, nnz): nnz zero entries (include/sparse_tensor.hpp:27-31)
            dims, nnz = args
            self._create(dims)
            nnz = int(nnz)
            if nnz > 0:
                self.set_entries(np.zeros(nnz, dtype=np.int64), np.zeros(nnz))
            return
        self._init_compress(*args, **kwargs)

    @classmethod
    def empty(cls, dims):
        """A tensor with no entries, to be filled by ``recompress``."""
        obj = cls.__new__(cls)
        obj._h = ctypes.c_void_p()
        obj._host = None
        obj.info = {}
        obj.u = None
        obj._create(dims)
        return obj

    def _create(self, dims):
        dims = [int(d) for d in dims]
        if len(dims) != 4:
            raise ValueError("dims must have 4 entries")
        self.dims = tuple(dims)
        _capi.check(_capi.lib().fricc_sparse_create(_capi.i64x4(dims), ctypes.byref(self._h)))

    def _init_compress(self, tensor_arr, dims, m, compression="fri", sampling_method="pivotal", verbose=False,
                       *, u=None, stream=None):
        import time

        if _is_torch(tensor_arr):
            x = _as_device_vector(tensor_arr, "tensor_arr")
            if not tensor_arr.is_contiguous():
                raise TypeError("tensor_arr must be C-contiguous")
        else:
            if not isinstance(tensor_arr, np.ndarray) or tensor_arr.dtype != np.float64:
                raise TypeError("tensor_arr must be a float64 numpy array (noconvert)")
            if not tensor_arr.flags.c_contiguous:
                raise TypeError("tensor_arr must be C-contiguous (py::array::c_style)")
            x = _as_device_vector(tensor_arr, "tensor_arr")
        self._create(dims)
        n = int(np.prod(self.dims))
        if x.numel() != n:
            raise ValueError(f"tensor has {x.numel()} elements but dims {self.dims} give {n}")
        m = int(m)
        method = _method(compression, sampling_method)
        # one rand() per "fri" call, also when nothing is sampled
        uu = 0.0 if method == _capi.LARGEST else _draw(u)
        if m < 0 or m > n:
            raise ValueError(f"m must lie in [0, {n}]")
        self.u = None if method == _capi.LARGEST else uu
        if m == 0:
            return
        t0 = time.perf_counter()
        st = _capi.stream_ptr(stream)
        _capi.check(_capi.lib().fricc_sparse_compress(self._h, ctypes.c_void_p(x.data_ptr()), m, method, uu, st))
        self._refresh_info()
        if verbose:
            _torch().cuda.synchronize()
            dt = time.perf_counter() - t0
            print("Subsection    Fraction of Total    Total Time")
            print("==========    =================    ==========")
            print(f"GPU compress  {1.0:6.4f}               {dt:6.4f} (s)")
            print(f"|D| = {self.info['n_D']}  n' = {self.info['n_sampled']}")

    def _refresh_info(self):
        info = _capi.CompressInfo()
        nnz = ctypes.c_int64()
        dims = _capi.I64x4()
        _capi.check(_capi.lib().fricc_sparse_info(self._h, ctypes.byref(dims), ctypes.byref(nnz), ctypes.byref(info)))
        self.info = info.as_dict()
        self._host = None

    # ---- reference API
    def get_element(self, mi):
        """([i, j, a, b], value) of entry mi (include/sparse_tensor.hpp:138-143)."""
        idx, val = self._host_entries()
        mi = int(mi)
        if mi < 0 or mi >= idx.size:
            raise IndexError("entry index out of range")
        f = int(idx[mi])
        d0, d1, d2, d3 = self.dims
        return [f // (d1 * d2 * d3), f % (d1 * d2 * d3) // (d2 * d3), f % (d2 * d3) // d3, f % d3], float(val[mi])

    def print(self):
        idx, val = self._host_entries()
        print("i   j   k   l        val")
        for mi in range(idx.size):
            i, j, a, b = self.get_element(mi)[0]
            print(f"{i}   {j}   {a}   {b}        {val[mi]}")
        print()

    # ---- extensions
    @property
    def nnz(self):
        n = ctypes.c_int64()
        _capi.check(_capi.lib().fricc_sparse_info(self._h, None, ctypes.byref(n), None))  # NULL: not wanted
        return int(n.value)

    def size(self):
        return self.nnz

    def handle(self):
        return self._h

    def _host_entries(self):
        if self._host is None:
            n = self.nnz
            idx = np.zeros(n, dtype=np.int64)
            val = np.zeros(n, dtype=np.float64)
            if n:
                _capi.check(_capi.lib().fricc_sparse_get(self._h, idx.ctypes.data_as(ctypes.c_void_p),
                                                         val.ctypes.data_as(ctypes.c_void_p), _capi.stream_ptr()))
            self._host = (idx, val)
        return self._host

    def indices(self):
        """Flat indices of all entries (host copy, compression order)."""
        return self._host_entries()[0].copy()

    def values(self):
        return self._host_entries()[1].copy()

    def set_entries(self, idx, val):
        """set_element for every entry (include/sparse_tensor.hpp:127-136)."""
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        val = np.ascontiguousarray(val, dtype=np.float64)
        if idx.shape != val.shape:
            raise ValueError("idx and val differ in length")
        _capi.check(_capi.lib().fricc_sparse_set(self._h, idx.ctypes.data_as(ctypes.c_void_p),
                                                 val.ctypes.data_as(ctypes.c_void_p), idx.size, _capi.stream_ptr()))
        self._refresh_info()

    def recompress(self, x, m, compression="fri", u=None, stream=None, sampling_method="systematic"):
        """Compress a new device tensor into this object's storage (no reallocation
        when m does not grow).  Consumes one uniform for "fri"."""
        x = _as_device_vector(x)
        method = _method(compression, sampling_method)
        uu = _draw(u) if method != _capi.LARGEST else 0.0
        self.u = uu if method != _capi.LARGEST else None
        _capi.check(_capi.lib().fricc_sparse_compress(self._h, ctypes.c_void_p(x.data_ptr()), int(m), method, uu,
                                                      _capi.stream_ptr(stream)))
        self._async = False
        self._refresh_info()
        return self

    def recompress_async(self, x, m, compression, u_dev, stream=None, sampling_method="systematic"):
        """``recompress`` with no host round trip (device-resident loop, graph
        capture): the uniform is read from ``u_dev`` (a one-element CUDA float64
        tensor, unused for "largest"); statistics and errors wait for
        ``sync_status``."""
        x = _as_device_vector(x)
        method = _method(compression, sampling_method)
        _capi.check(_capi.lib().fricc_sparse_compress_async(self._h, ctypes.c_void_p(x.data_ptr()), int(m), method,
                                                            ctypes.c_void_p(u_dev.data_ptr()),
                                                            _capi.stream_ptr(stream)))
        self._host = None
        self._async = True
        return self

    def recompress_pair_async(self, other, x, m, compression, u_dev, u_dev2, stream=None,
                              sampling_method="systematic"):
        """``recompress_async`` of x into this tensor (uniform u_dev) and into
        ``other`` (uniform u_dev2), sharing everything but the sampling: each
        tensor ends bit for bit as its own ``recompress_async`` would leave it.
        Returns (self, other)."""
        x = _as_device_vector(x)
        method = _method(compression, sampling_method)
        _capi.check(_capi.lib().fricc_sparse_compress_pair_async(
            self._h, other._h, ctypes.c_void_p(x.data_ptr()), int(m), method, ctypes.c_void_p(u_dev.data_ptr()),
            ctypes.c_void_p(u_dev2.data_ptr()), _capi.stream_ptr(stream)))
        for t in (self, other):
            t._host = None
            t._async = True
        return self, other

    @staticmethod
    def compress_multi_async(tensors, x, m, compression, u_dev, stream=None, sampling_method="systematic"):
        """``recompress_async`` of x into every tensor of ``tensors``, tensor k
        with the uniform u_dev[k] (a CUDA float64 tensor of len(tensors)
        consecutive values), sharing everything but the sampling."""
        x = _as_device_vector(x)
        method = _method(compression, sampling_method)
        if u_dev.numel() < len(tensors) or not u_dev.is_contiguous():
            raise ValueError("u_dev must hold one contiguous uniform per tensor")
        handles = (ctypes.c_void_p * len(tensors))(*[t._h.value for t in tensors])
        _capi.check(_capi.lib().fricc_sparse_compress_multi_async(
            handles, len(tensors), ctypes.c_void_p(x.data_ptr()), int(m), method, ctypes.c_void_p(u_dev.data_ptr()),
            _capi.stream_ptr(stream)))
        for t in tensors:
            t._host = None
            t._async = True
        return list(tensors)

    def sync_status(self, stream=None):
        """Statistics and status of the last asynchronous compression (raises
        on non-finite input, p > 1 or a bad uniform)."""
        if getattr(self, "_async", False):
            st = _capi.lib().fricc_sparse_status(self._h, _capi.stream_ptr(stream))
            self._refresh_info()
            _capi.check(st)
        return self.info

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.lib().fricc_sparse_destroy(h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self._h = ctypes.c_void_p()


def _dev4(a, name, writeable=False):
    """(device tensor, strides, keep_numpy) for a 4-D float64 operand."""
    torch = _torch()
    if _is_torch(a):
        if a.dtype != torch.float64:
            raise TypeError(f"{name} must be float64")
        if a.dim() != 4:
            raise ValueError(f"{name} must have ndim 4")
        if not a.is_cuda:
            raise TypeError(f"{name} must be a CUDA tensor")
        return a, list(a.stride()), None
    if not isinstance(a, np.ndarray) or a.dtype != np.float64:
        raise TypeError(f"{name} must be a float64 numpy array (noconvert)")
    if a.ndim != 4:
        raise ValueError(f"{name} must have ndim 4 (unchecked<4>)")
    if writeable and not a.flags.writeable:
        raise ValueError(f"{name} must be writeable")
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t, list(t.stride()), a


def contract_DTSpT(W, T, ouput, term, precision="fp64"):
    """ouput += term(W, T), in place (contract_SparseTensor4d_wrapper,
    src/sparse_tensor.cpp:349-382).  ``precision`` (not in the reference):
    "fp64", "fp32" (FP32 MFMA) or "bf16" (BF16 MFMA, hi/lo split operands,
    FP32 accumulate)."""
    if term not in _TERMS:
        raise ValueError(f"contraction term {term!r} not implemented")
    if not isinstance(T, SparseTensor4d):
        raise TypeError("T must be a SparseTensor4d")
    if precision not in ("fp64", "fp32", "bf16", "bf16x1"):
        raise ValueError("precision must be 'fp64', 'fp32', 'bf16' or 'bf16x1'")
    Wd, ws, _ = _dev4(W, "W")
    Od, os_, host = _dev4(ouput, "ouput", writeable=True)
    if T.nnz == 0:
        return None
    st = _capi.stream_ptr()
    _capi.check(_capi.lib().fricc_contract(term.encode(), ctypes.c_void_p(Wd.data_ptr()), _capi.i64x4(Wd.shape),
                                           _capi.i64x4(ws), T.handle(), ctypes.c_void_p(Od.data_ptr()),
                                           _capi.i64x4(Od.shape), _capi.i64x4(os_), _capi.PRECISIONS[precision], st))
    if host is not None:
        host[...] = Od.cpu().numpy()
    return None


def contract_DTSpT_experimental(W, T, ouput, term):
    """The reference's alternative OpenMP schedules of the ladder
    (src/sparse_tensor.cpp:389-459); both compute the "2323" term."""
    if term not in ("taskloop", "atomic"):
        raise ValueError(f"experimental term {term!r} not implemented")
    return contract_DTSpT(W, T, ouput, "2323")


# ---------------------------------------------------------------- free functions
def _compress_raw(x, m, method, u):
    torch = _torch()
    xd = _as_device_vector(x)
    m = int(m)
    if m <= 0 or m > xd.numel():
        raise ValueError(f"m must lie in [1, {xd.numel()}]")
    idx = torch.empty(m, dtype=torch.int32, device="cuda")
    val = torch.empty(m, dtype=torch.float64, device="cuda")
    info = _capi.CompressInfo()
    _capi.check(_capi.lib().fricc_compress(ctypes.c_void_p(xd.data_ptr()), xd.numel(), m, method, float(u),
                                           ctypes.c_void_p(idx.data_ptr()), ctypes.c_void_p(val.data_ptr()),
                                           ctypes.byref(info), _capi.stream_ptr()))
    return idx.cpu().numpy().astype(np.uint32).astype(np.int64), val.cpu().numpy(), info.as_dict()


def argsort(array, m):
    """Indices of the m largest |x|, ties by index (src/fri_utils.cpp:11-36)."""
    idx, _, _ = _compress_raw(array, m, _capi.LARGEST, 0.0)
    return idx.tolist()


def fri_compression(x, n_sample, sampling_method="pivotal", verbose=False, *, u=None):
    """(indices, values) of the FRI compression (src/fri_utils.cpp:439-522)."""
    idx, val, info = _compress_raw(x, n_sample, _method("fri", sampling_method), _draw(u))
    if verbose:
        print(f"|D| = {info['n_D']}  n' = {info['n_sampled']}  remaining norm = {info['remaining_norm']}")
    return idx.tolist(), val.tolist()


def get_d_largest(n_sample, x):
    """(D, remaining_norm) (src/fri_utils.cpp:400-437).  Keyword names follow
    the reference binding, which swaps them (src/py_rccsd.cpp:28): the first
    positional argument is the vector, the second the budget."""
    vec, m = n_sample, x
    idx, _, info = _compress_raw(vec, m, _capi.FRI_SYSTEMATIC, 0.0)
    return idx[: info["n_D"]].tolist(), info["remaining_norm"]


def make_probability_vector(x, n_sample, D, remaining_norm):
    """p_i = |x_i| / R * n_sample, p[D] = 0 (src/fri_utils.cpp:367-398);
    raises ValueError where the reference exits on p > 1."""
    torch = _torch()
    xd = _as_device_vector(x)
    # elementwise tensor division (IEEE division, as the reference's abs(x)/R*n)
    R = torch.full_like(xd, float(remaining_norm))
    p = torch.abs(xd) / R * torch.full_like(xd, float(n_sample))
    if len(D):
        p[torch.as_tensor(np.asarray(D, dtype=np.int64), device="cuda")] = 0.0
    if bool((p > 1.0).any()):
        raise ValueError("probability vector has an entry > 1 (src/fri_utils.cpp:383-395)")
    return p.cpu().numpy().tolist()


def sample_systematic(n_sample, probs, *, u=None):
    """Systematic sampling (src/fri_utils.cpp:129-159) with the canonical
    interval ends; draws one uniform."""
    torch = _torch()
    pd = _as_device_vector(np.asarray(probs, dtype=np.float64) if not _is_torch(probs) else probs, "probs")
    n_sample = int(n_sample)
    uu = _draw(u)
    out = torch.zeros(max(n_sample, 1), dtype=torch.int32, device="cuda")
    nf = ctypes.c_int64()
    _capi.check(_capi.lib().fricc_sample_systematic(ctypes.c_void_p(pd.data_ptr()), pd.numel(), n_sample, uu,
                                                    ctypes.c_void_p(out.data_ptr()), ctypes.byref(nf),
                                                    _capi.stream_ptr()))
    return out[:n_sample].cpu().numpy().astype(np.int64).tolist()


def sample_pivotal(n_sample, probs, *, u=None, uniforms=None):
    """Pivotal sampling (src/fri_utils.cpp:63-127) of n_sample indices, in the
    order the chain emits them.  Step j draws ``uniforms[j]`` when given (n
    values; uniforms[0] is not read -- e.g. the reference's rand()/RAND_MAX
    stream shifted by one), else the SplitMix64 stream keyed by one uniform
    (``u`` or one rand()/RAND_MAX draw).  Raises where the reference exits."""
    torch = _torch()
    pd = _as_device_vector(np.asarray(probs, dtype=np.float64) if not _is_torch(probs) else probs, "probs")
    n_sample = int(n_sample)
    Ud = None
    uu = 0.0
    if uniforms is not None:
        Ud = _as_device_vector(np.asarray(uniforms, dtype=np.float64) if not _is_torch(uniforms) else uniforms,
                               "uniforms")
        if Ud.numel() < pd.numel():
            raise ValueError("uniforms must hold one value per probability")
    else:
        uu = _draw(u)
    out = torch.zeros(max(n_sample, 1), dtype=torch.int32, device="cuda")
    _capi.check(_capi.lib().fricc_sample_pivotal(ctypes.c_void_p(pd.data_ptr()), pd.numel(), n_sample,
                                                 ctypes.c_void_p(Ud.data_ptr() if Ud is not None else None), uu,
                                                 ctypes.c_void_p(out.data_ptr()), _capi.stream_ptr()))
    return out[:n_sample].cpu().numpy().astype(np.int64).tolist()


PARALLEL_CHUNKS = 256


def parallel_sample(n_sample, probs, *, n_chunks=None, u=None, uniforms=None):
    """Thread-partitioned pivotal sampling (src/fri_utils.cpp:176-362) with
    ``n_chunks`` chunks (default min(256, len(probs))) in place of the
    reference's OpenMP threads, whose count its result depends on.  Draws as
    ``sample_pivotal``, by position.  Raises where the reference exits."""
    torch = _torch()
    pd = _as_device_vector(np.asarray(probs, dtype=np.float64) if not _is_torch(probs) else probs, "probs")
    n_sample = int(n_sample)
    T = int(min(PARALLEL_CHUNKS, pd.numel()) if n_chunks is None else n_chunks)
    Ud = None
    uu = 0.0
    if uniforms is not None:
        Ud = _as_device_vector(np.asarray(uniforms, dtype=np.float64) if not _is_torch(uniforms) else uniforms,
                               "uniforms")
        if Ud.numel() < pd.numel():
            raise ValueError("uniforms must hold one value per probability")
    else:
        uu = _draw(u)
    out = torch.zeros(max(n_sample, 1), dtype=torch.int32, device="cuda")
    _capi.check(_capi.lib().fricc_parallel_sample(ctypes.c_void_p(pd.data_ptr()), pd.numel(), n_sample, T,
                                                  ctypes.c_void_p(Ud.data_ptr() if Ud is not None else None), uu,
                                                  ctypes.c_void_p(out.data_ptr()), _capi.stream_ptr()))
    return out[:n_sample].cpu().numpy().astype(np.int64).tolist()
