"""ctypes bindings to the CPU oracle library (test infrastructure only).

Each wrapper cites the reference function it restates; the C source is
``oracle/fricc_oracle.cpp``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

CANONICAL = 0
FAITHFUL = 1
LARGEST = 0
FRI_SYSTEMATIC = 1
FRI_PIVOTAL = 2


class OrcInfo(ctypes.Structure):
    _fields_ = [
        ("n_D", ctypes.c_int64),
        ("n_sampled", ctypes.c_int64),
        ("n_notfound", ctypes.c_int64),
        ("norm0", ctypes.c_double),
        ("remaining_norm", ctypes.c_double),
        ("p_total", ctypes.c_double),
        ("status", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libfricc_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int64)
        L.orc_compress.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ip, dp, ctypes.POINTER(OrcInfo)]
        L.orc_compress.restype = ctypes.c_int
        L.orc_norm_canonical.argtypes = [dp, ctypes.c_int64]
        L.orc_norm_canonical.restype = ctypes.c_double
        L.orc_prefix_canonical.argtypes = [dp, ctypes.c_int64, dp]
        L.orc_prefix_canonical.restype = None
        L.orc_contract_ref.argtypes = [ctypes.c_char_p, dp, ip, ip, ip, dp, ctypes.c_int64, ip, dp, ip,
                                       ctypes.c_int64, ctypes.c_int64]
        L.orc_contract_ref.restype = ctypes.c_int
        L.orc_eri_block.argtypes = [dp, ctypes.c_int64, ctypes.c_int64] + [ctypes.c_int64] * 8 + [dp]
        L.orc_eri_block.restype = None
        L.orc_eri_block_plain.argtypes = L.orc_eri_block.argtypes
        L.orc_eri_block_plain.restype = None
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_pivotal.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, dp, ctypes.c_uint64, ctypes.c_int, ip, ip]
        L.orc_pivotal.restype = ctypes.c_int
        L.orc_parallel_sample.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, dp, ctypes.c_uint64,
                                          ctypes.c_int, ip]
        L.orc_parallel_sample.restype = ctypes.c_int
        L.orc_splitmix_uniforms.argtypes = [ctypes.c_uint64, ctypes.c_int64, dp]
        L.orc_splitmix_uniforms.restype = None
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def compress(x, m, method=FRI_SYSTEMATIC, u=0.0, mode=CANONICAL):
    """fri_compression / "largest" argsort (src/fri_utils.cpp:439-522,
    include/sparse_tensor.hpp:46-67).  Returns (idx int64[m], val[m], info)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    m = int(m)
    idx = np.zeros(max(m, 1), dtype=np.int64)
    val = np.zeros(max(m, 1), dtype=np.float64)
    info = OrcInfo()
    st = lib().orc_compress(_dp(x), x.size, m, int(method), int(mode), float(u), _ip(idx), _dp(val),
                            ctypes.byref(info))
    d = info.as_dict()
    d["status"] = st
    return idx[:m], val[:m], d


def key_of(u):
    """The 64-bit SplitMix64 key of a call's uniform u: its IEEE-754 bits."""
    return int(np.array([u], dtype="<f8").view("<u8")[0])


def splitmix_uniforms(key, n):
    """The step uniforms of pivotal sampling when no array is given: element j
    is the j-th SplitMix64 output seeded with key, top 53 bits times 2^-53."""
    out = np.zeros(max(int(n), 1))
    lib().orc_splitmix_uniforms(ctypes.c_uint64(int(key)), int(n), _dp(out))
    return out[:int(n)]


def pivotal(p, n_sample, U=None, key=0, mode=CANONICAL):
    """sample_pivotal (src/fri_utils.cpp:63-127).  Step j draws U[j] (U[0]
    unused) or, with U None, the SplitMix64 stream of key.  Returns
    (S int64[n_sample], count emitted, status): status 0, 1 (p outside [0, 1])
    or 5 (a reference exit: wrong number of samples)."""
    p = np.ascontiguousarray(p, dtype=np.float64).ravel()
    ns = int(n_sample)
    S = np.zeros(max(ns, 1), dtype=np.int64)
    cnt = ctypes.c_int64(0)
    Up = None
    if U is not None:
        U = np.ascontiguousarray(U, dtype=np.float64).ravel()
        assert U.size >= p.size
        Up = _dp(U)
    st = lib().orc_pivotal(_dp(p), p.size, ns, Up, ctypes.c_uint64(int(key)), int(mode), _ip(S), ctypes.byref(cnt))
    return S[:ns], cnt.value, st


def parallel_sample(q, n_sample, n_chunks, U=None, key=0, mode=CANONICAL):
    """parallel_sample (src/fri_utils.cpp:176-362) with n_chunks chunks in place
    of the reference's OpenMP threads.  Returns (S int64[n_sample], status):
    0, 1 (q outside [0, 1]), 2 (n_chunks outside [1, n]) or 5 (a reference
    exit)."""
    q = np.ascontiguousarray(q, dtype=np.float64).ravel()
    ns = int(n_sample)
    S = np.zeros(max(ns, 1), dtype=np.int64)
    Up = None
    if U is not None:
        U = np.ascontiguousarray(U, dtype=np.float64).ravel()
        assert U.size >= q.size
        Up = _dp(U)
    st = lib().orc_parallel_sample(_dp(q), q.size, ns, int(n_chunks), Up, ctypes.c_uint64(int(key)), int(mode),
                                   _ip(S))
    return S[:ns], st


def norm_canonical(x):
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    return lib().orc_norm_canonical(_dp(x), x.size)


def prefix_canonical(p):
    p = np.ascontiguousarray(p, dtype=np.float64).ravel()
    P = np.zeros(p.size + 1)
    lib().orc_prefix_canonical(_dp(p), p.size, _dp(P))
    return P


def unpack(idx, dims):
    """SparseTensor4d::unpack_idx (include/sparse_tensor.hpp:108-125)."""
    idx = np.asarray(idx, dtype=np.int64)
    d0, d1, d2, d3 = [int(d) for d in dims]
    out = np.empty((idx.size, 4), dtype=np.int64)
    out[:, 0] = idx // (d1 * d2 * d3)
    out[:, 1] = idx % (d1 * d2 * d3) // (d2 * d3)
    out[:, 2] = idx % (d2 * d3) // d3
    out[:, 3] = idx % d3
    return out


def contract_ref(term, W, idx, val, tdims, out, block_lo=0, block_hi=-1, idx4=None):
    """The reference's loop-nest kernels (src/sparse_tensor.cpp:11-346): rescans
    every sparse entry per output block.  ``out`` is accumulated in place.
    ``idx4`` (optional): the entries already unpacked to [nnz][4] int64, the
    reference's AoS index layout (include/sparse_tensor.hpp:15-18)."""
    assert W.dtype == np.float64 and out.dtype == np.float64
    idx4 = np.ascontiguousarray(unpack(idx, tdims) if idx4 is None else idx4, dtype=np.int64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    wshape = np.array(W.shape, dtype=np.int64)
    ws = np.array([s // 8 for s in W.strides], dtype=np.int64)
    os_ = np.array([s // 8 for s in out.strides], dtype=np.int64)
    td = np.array(tdims, dtype=np.int64)
    st = lib().orc_contract_ref(term.encode(), _dp(W), _ip(wshape), _ip(ws), _ip(idx4), _dp(val),
                                val.size, _ip(td), _dp(out), _ip(os_), int(block_lo), int(block_hi))
    if st != 0:
        raise ValueError(f"unknown contraction term {term!r}")
    return out


def eri_block(L, p, q, r, s, plain=False):
    """(pq|rs) as the fma chain over Q (bit-identical to the GPU builder).
    ``plain``: the unblocked loop (the packed, tiled one computes the same
    chains; the two are compared in tests)."""
    L = np.ascontiguousarray(L, dtype=np.float64)
    naux, nmo, _ = L.shape
    out = np.empty((p[1] - p[0], q[1] - q[0], r[1] - r[0], s[1] - s[0]))
    f = lib().orc_eri_block_plain if plain else lib().orc_eri_block
    f(_dp(L), naux, nmo, p[0], p[1], q[0], q[1], r[0], r[1], s[0], s[1], _dp(out))
    return out


def num_threads():
    return lib().orc_num_threads()
