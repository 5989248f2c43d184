"""Synthetic density-fitted molecular systems and GPU integral containers.

The reference takes integrals from PySCF (``self.ao2mo``, fricc/friccsd.py:503-504),
which is not available here.  The benchmark systems are synthetic:

* orbital energies: occupied ~ U[-1.5,-0.3], virtual ~ U[0.1,3.0], each
  sorted (``numpy.random.default_rng(seed)``, drawn in that order);
* factors L[Q,p,q] = N(0,1) * exp(-|p-q|/10) / sqrt(naux), then
  L <- (L + L^T_pq)/2;
* (pq|rs) = sum_Q L[Q,p,q] L[Q,r,s], built on the GPU as a sequential fma chain
  over Q (bitwise 8-fold symmetric; libfricc_hip ``fricc_eri_build``);
* Fock = diag(orbital energies), so f_ov = 0 and the MP2 guess has t1 = 0.
"""
import ctypes
import weakref

import numpy as np

from . import _capi

CONFIGS = {
    # name: (nocc, nvir, naux, seed, m_keep, iterations)
    "c2": dict(nocc=20, nvir=180, naux=600, seed=1, m_keep=1_000_000, iterations=30),
    "c3": dict(nocc=32, nvir=256, naux=864, seed=1, m_keep=4_000_000, iterations=40),
}


def synthetic_factors(nocc, nvir, naux, seed):
    """(L[naux,nmo,nmo], mo_energy[nmo]) of the synthetic recipe."""
    nmo = nocc + nvir
    rng = np.random.default_rng(seed)
    eo = np.sort(rng.uniform(-1.5, -0.3, nocc))
    ev = np.sort(rng.uniform(0.1, 3.0, nvir))
    p = np.arange(nmo)
    decay = np.exp(-np.abs(p[:, None] - p[None, :]) / 10.0) / np.sqrt(naux)
    L = rng.standard_normal((naux, nmo, nmo)) * decay
    L = 0.5 * (L + L.transpose(0, 2, 1))
    return np.ascontiguousarray(L), np.concatenate([eo, ev])


class SyntheticSystem:
    """A synthetic closed-shell system (nocc doubly occupied orbitals)."""

    def __init__(self, nocc, nvir, naux, seed):
        self.nocc, self.nvir, self.naux, self.seed = int(nocc), int(nvir), int(naux), int(seed)
        self.nmo = self.nocc + self.nvir
        self.L, self.mo_energy = synthetic_factors(self.nocc, self.nvir, self.naux, self.seed)
        self.fock = np.diag(self.mo_energy)
        self.e_hf = 0.0

    @classmethod
    def from_config(cls, name):
        c = CONFIGS[name]
        return cls(c["nocc"], c["nvir"], c["naux"], c["seed"])


def eri_block(Ld, naux, nmo, ranges, out, strides, stream=None):
    """(pq|rs) for the four [lo, hi) ranges into ``out`` at the given element strides."""
    import torch

    assert out.dtype == torch.float64 and out.is_cuda
    rng = _capi.I64x8(*[int(v) for r in ranges for v in r])
    _capi.check(_capi.lib().fricc_eri_build(ctypes.c_void_p(Ld.data_ptr()), int(naux), int(nmo), rng,
                                            _capi.i64x4(strides), ctypes.c_void_p(out.data_ptr()),
                                            _capi.stream_ptr(stream)))
    return out


def bf16_xblocks(X):
    """256-column blocks of a BF16 operand."""
    return (X + 255) // 256


def bf16_planes(A, R, X, stream=None):
    """BF16 hi/lo operand of A viewed as (R, X): 1 KB records
    [X/256][R][hi 256 | lo 256] (fricc_split_bf16 layout)."""
    import torch

    A = A.contiguous()
    assert A.numel() == R * X and A.dtype in (torch.float64, torch.float32)
    src = _capi.FP64 if A.dtype == torch.float64 else _capi.FP32
    out = torch.empty(R * bf16_xblocks(X) * 512, dtype=torch.int16, device=A.device)
    _capi.check(_capi.lib().fricc_split_bf16(ctypes.c_void_p(A.data_ptr()), src, int(R), int(X),
                                             ctypes.c_void_p(out.data_ptr()), _capi.stream_ptr(stream)))
    return out


def bf16_planes_hi(A, R, X, stream=None):
    """BF16 hi-only operand of A viewed as (R, X) for single-pass products:
    512-byte records [X/256][R][hi 256] (fricc_split_bf16_hi layout)."""
    import torch

    A = A.contiguous()
    assert A.numel() == R * X and A.dtype in (torch.float64, torch.float32)
    src = _capi.FP64 if A.dtype == torch.float64 else _capi.FP32
    out = torch.empty(R * bf16_xblocks(X) * 256, dtype=torch.int16, device=A.device)
    _capi.check(_capi.lib().fricc_split_bf16_hi(ctypes.c_void_p(A.data_ptr()), src, int(R), int(X),
                                                ctypes.c_void_p(out.data_ptr()), _capi.stream_ptr(stream)))
    return out


def bf16_planes_lincomb(A, B, alpha, beta, R, X, stream=None):
    """bf16_planes of alpha*A + beta*B without forming the sum (fricc_split_bf16_lincomb)."""
    import torch

    A, B = A.contiguous(), B.contiguous()
    assert A.shape == B.shape and A.dtype == B.dtype and A.numel() == R * X
    assert A.dtype in (torch.float64, torch.float32)
    src = _capi.FP64 if A.dtype == torch.float64 else _capi.FP32
    out = torch.empty(R * bf16_xblocks(X) * 512, dtype=torch.int16, device=A.device)
    _capi.check(_capi.lib().fricc_split_bf16_lincomb(ctypes.c_void_p(A.data_ptr()), ctypes.c_void_p(B.data_ptr()),
                                                     src, float(alpha), float(beta), int(R), int(X),
                                                     ctypes.c_void_p(out.data_ptr()), _capi.stream_ptr(stream)))
    return out


def bf16_planes_lincomb3(A, B, alpha, beta, R, X, stream=None):
    """(bf16_planes_lincomb(A, B, alpha, beta), bf16_planes(A), bf16_planes(B))
    with A and B read once (fricc_split_bf16_lincomb3)."""
    import torch

    A, B = A.contiguous(), B.contiguous()
    assert A.shape == B.shape and A.dtype == B.dtype and A.numel() == R * X
    assert A.dtype in (torch.float64, torch.float32)
    src = _capi.FP64 if A.dtype == torch.float64 else _capi.FP32
    outs = [torch.empty(R * bf16_xblocks(X) * 512, dtype=torch.int16, device=A.device) for _ in range(3)]
    _capi.check(_capi.lib().fricc_split_bf16_lincomb3(ctypes.c_void_p(A.data_ptr()), ctypes.c_void_p(B.data_ptr()),
                                                      src, float(alpha), float(beta), int(R), int(X),
                                                      *(ctypes.c_void_p(o.data_ptr()) for o in outs),
                                                      _capi.stream_ptr(stream)))
    return tuple(outs)


class _LazyOperands(dict):
    """Operands of one dtype, built on first access.  Holds its integrals
    weakly: the integrals own this cache, and a reference cycle would keep
    every device buffer of a dropped system alive until a cyclic collection."""

    def __init__(self, eris, dtype):
        super().__init__()
        self._eris, self.dtype = weakref.ref(eris), dtype

    def __missing__(self, name):
        alias = {"Lovov": "Lov2", "Lov2": "Lovov"}.get(name)
        if alias is not None and dict.__contains__(self, alias):
            val = dict.__getitem__(self, alias)
        else:
            val = self._eris()._build(name, self.dtype)
        self[name] = val
        return val


def _cstrides(shape):
    s = [1, 1, 1, 1]
    for i in (2, 1, 0):
        s[i] = s[i + 1] * shape[i + 1]
    return s


class DeviceERIs:
    """GPU-resident chemists'-notation integrals with the attributes
    ``ccsd._ChemistsERIs`` exposes to update_amps (fricc/friccsd.py:80-86,
    164-187, 310-330): fock, mo_energy, oooo, ovoo, ovov, oovv, ovvo,
    get_ovvv(), plus ``ladder`` = W[c,d,a,b] = (ac|bd), the 2323 operand."""

    def __init__(self):
        self._prep = {}
        self._s_ab = None
        self.ladder = None
        self.L = None

    @classmethod
    def from_factors(cls, L, mo_energy, nocc, fock=None, build_ladder=True):
        import torch

        e = cls()
        L = np.ascontiguousarray(L, dtype=np.float64)
        naux, nmo, _ = L.shape
        o, v = int(nocc), nmo - int(nocc)
        e.nocc, e.nvir, e.nmo, e.naux = o, v, nmo, naux
        dev = torch.device("cuda")
        e.mo_energy = torch.as_tensor(np.asarray(mo_energy, dtype=np.float64), device=dev)
        e.fock = torch.as_tensor(np.diag(mo_energy) if fock is None else np.asarray(fock, np.float64), device=dev)
        Ld = torch.as_tensor(L, device=dev)
        O, V = (0, o), (o, nmo)

        def blk(r0, r1, r2, r3):
            shape = [r[1] - r[0] for r in (r0, r1, r2, r3)]
            t = torch.empty(shape, dtype=torch.float64, device=dev)
            return eri_block(Ld, naux, nmo, (r0, r1, r2, r3), t, _cstrides(shape))

        e.oooo = blk(O, O, O, O)
        e.ovoo = blk(O, V, O, O)
        e.ovov = blk(O, V, O, V)
        e.oovv = blk(O, O, V, V)
        e.ovvo = blk(O, V, V, O)
        e.ovvv = blk(O, V, V, V)
        if build_ladder:
            # W[c,d,a,b] = (ac|bd): p=a, q=c, r=b, s=d
            W = torch.empty((v, v, v, v), dtype=torch.float64, device=dev)
            e.ladder = eri_block(Ld, naux, nmo, (V, V, V, V), W, [v, v ** 3, 1, v * v])
        else:
            e.ladder = None
        e.L = Ld
        return e

    @classmethod
    def from_blocks(cls, fock, mo_energy, nocc, oooo, ovoo, ovov, oovv, ovvo, ovvv, vvvv):
        """From host chemists' blocks (e.g. a PySCF ``_ChemistsERIs`` with vvvv
        unpacked to (v,v,v,v))."""
        import torch

        e = cls()
        dev = torch.device("cuda")
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
        e.nocc = int(nocc)
        e.mo_energy, e.fock = t(mo_energy), t(fock)
        e.oooo, e.ovoo, e.ovov, e.oovv, e.ovvo, e.ovvv = t(oooo), t(ovoo), t(ovov), t(oovv), t(ovvo), t(ovvv)
        e.nvir = e.ovov.shape[1]
        e.nmo = e.nocc + e.nvir
        # W[c,d,a,b] = (ac|bd) = vvvv[a,c,b,d]
        e.ladder = t(vvvv).permute(1, 3, 0, 2).contiguous()
        e.L = None
        return e

    def get_ovvv(self):
        return self.ovvv

    def sparse_operand(self, name, precision):
        """Static A operand of a sparse contraction ("A1323", "A0112", "A0313",
        "Gx", "Gy", "ladder") in the kernel layout for ``precision``
        ("fp64" | "fp32" | "bf16" hi | lo records | "bf16x1" hi-only records),
        cached.  For FP32/BF16 the FP64 layout is
        built only transiently."""
        import torch

        key = ("sparse", name, precision)
        if key in self._prep:
            return self._prep[key]
        if precision == "fp64":
            out = self.ladder if name == "ladder" else self.prepared(torch.float64)[name]
        elif name == "ladder" and self.ladder is None and self.L is not None:
            out = self._reduced_ladder(precision)
        else:
            A = self.ladder if name == "ladder" else self._build(name, torch.float64)
            if A is None:
                raise RuntimeError("the FP64 ladder was released; its BF16/FP32 operand must be built before")
            if precision == "fp32":
                out = A.to(torch.float32).contiguous()
            elif precision == "bf16x1":
                out = bf16_planes_hi(A, A.shape[0] * A.shape[1], A.shape[2] * A.shape[3])
            else:
                out = bf16_planes(A, A.shape[0] * A.shape[1], A.shape[2] * A.shape[3])
            del A
        self._prep[key] = out
        return out

    def _build(self, name, dtype):
        """One static operand in the layout the update reads (see prepared)."""
        import torch

        o, v = self.nocc, self.nvir
        c = lambda a: a.to(dtype).contiguous()  # noqa: E731
        ovov, ovvv = self.ovov, self.ovvv
        if name == "A1323":
            return c(ovov.permute(1, 3, 0, 2))          # [c,d,k,l] = (kc|ld)
        if name in ("A0112", "ovov"):
            return c(ovov)                              # [l,d,k,c] = (ld|kc)
        if name == "A0313":
            return c(ovov.permute(0, 3, 2, 1))          # [l,d,k,c] = (lc|kd)
        if name == "A0113c":
            return c(ovov - 0.5 * ovov.permute(0, 3, 2, 1))  # [l,d,k,c] = (ld|kc) - 1/2 (lc|kd)
        if name == "A12b":
            # [k,c,i,a] = 2 (kc|ai) - (ki|ac): bare 2 Wvoov - Wvovo of the linearised update
            return c(2.0 * self.ovvo.permute(0, 1, 3, 2) - self.oovv.permute(0, 3, 1, 2))
        if name == "Gx":
            return c(ovvv.permute(3, 1, 0, 2))          # [c,d,k,a] = (kd|ac)
        if name == "Gy":
            return c(ovvv.permute(1, 3, 0, 2))          # [c,d,k,b] = (kc|bd)
        if name == "P1":
            return c(ovvv.permute(0, 1, 3, 2))          # [k,c,d,a] = (kc|ad)
        if name == "P2":
            return c(ovvv.permute(0, 3, 1, 2))          # [k,c,d,a] = (kd|ac)
        if name == "Z":
            return c(2 * ovvv.permute(0, 3, 1, 2) - ovvv.permute(0, 1, 3, 2))
        if name == "Qv":
            return c(ovvv.permute(0, 2, 1, 3))          # [i,c,a,b] = (ia|cb)
        if name == "Lov":
            return c(2 * ovov.permute(0, 2, 3, 1) - ovov.permute(0, 2, 1, 3))   # [k,l,d,c]
        if name in ("Lov2", "Lovov"):
            # [k,l,c,d] = 2 (kc|ld) - (kd|lc): one tensor serves both names
            return c(2 * ovov.permute(0, 2, 1, 3) - ovov.permute(0, 2, 3, 1))
        if name == "ovovT":
            return c(ovov.permute(0, 2, 1, 3))          # [i,j,a,b] = (ia|jb)
        if name == "ovvo_p":
            return c(self.ovvo.permute(0, 1, 3, 2))     # [k,c,i,a] = (kc|ai)
        if name == "oovv_p":
            return c(self.oovv.permute(0, 3, 1, 2))     # [k,c,i,a] = (ki|ac)
        if name == "ovoo_P":
            return c(self.ovoo.permute(2, 3, 0, 1))     # [k,i,l,c] = (lc|ki)
        if name == "ovoo_Q":
            return c(self.ovoo.permute(0, 3, 1, 2))     # [k,i,c,l] = (kc|li)
        if name == "ovoo_Y":
            return c(self.ovoo.permute(0, 2, 3, 1))     # [l,k,i,c] = (lc|ki)
        if name == "ovoo_ijak":
            return c(self.ovoo.permute(0, 2, 1, 3))     # [i,j,a,k] = (ia|jk)
        if name == "Aoo":
            # [k,l,i,c] = -2 (lc|ki) + (kc|li)
            Y = self.ovoo.permute(0, 2, 3, 1)
            return c(-2 * Y.permute(1, 0, 2, 3) + Y)
        if name in ("ovoo", "oovv", "ovvo", "oooo"):
            return c(getattr(self, name))
        if name == "Woooo_b":
            return c(self.oooo.permute(0, 2, 1, 3))     # [k,l,i,j] = (ki|lj)
        if name == "ladder":
            if dtype != torch.float64 or self.ladder is None:
                raise KeyError("the FP64 ladder operand is not available")
            return self.ladder
        if name == "s_ab":
            # sum over (c,d) of the bare part of Wvvvv, (ac|bd)
            if self.ladder is not None:
                return self.ladder.reshape(v * v, v * v).sum(0).reshape(v, v).to(dtype)
            if self._s_ab is None:
                if self.L is None:
                    raise KeyError("s_ab needs the ladder or the factors")
                # sum_cd sum_Q L[Q,a,c] L[Q,b,d] = sum_Q M[Q,a] M[Q,b], M = sum_c L[Q,a,c]
                from .dense import einsum

                M = self.L[:, o:, o:].sum(2)
                self._s_ab = einsum("qa,qb->ab", M, M, compute=torch.float64)
            return self._s_ab.to(dtype)
        if name == "g_ka":
            return ovvv.sum(dim=(1, 3)).to(dtype)       # sum_cd (kd|ac) = sum_cd (kc|ad)
        raise KeyError(name)

    def _reduced_ladder(self, precision, chunk_bytes=1 << 31):
        """BF16 planes (or FP32 copy) of the ladder W[c,d,a,b] = (ac|bd) built
        from the factors in slices of c, so the FP64 ladder (v^4 doubles) never
        exists whole: each slice's FP64 rows are evaluated, converted into
        their place in the operand and dropped.  The elements are those of the
        whole-ladder build (each integral is the same FMA chain, each
        conversion elementwise)."""
        import torch

        o, v, nmo = self.nocc, self.nvir, self.nmo
        X = v * v
        Xb = bf16_xblocks(X)
        cc = min(v, max(1, int(chunk_bytes // (8 * v ** 3))))
        dev = self.L.device
        rl = 512 if precision == "bf16" else 256  # record length: hi | lo, or hi only
        if precision == "fp32":
            out = torch.empty((v, v, v, v), dtype=torch.float32, device=dev)
        else:
            out = torch.empty(X * Xb * rl, dtype=torch.int16, device=dev)
            rec = out.view(Xb, X, rl)  # [column block][row (c,d)][hi | lo] or [hi]
        Wc = torch.empty((cc, v, v, v), dtype=torch.float64, device=dev)
        for c0 in range(0, v, cc):
            c1 = min(v, c0 + cc)
            n = c1 - c0
            W = Wc[:n]
            # W[c,d,a,b] = (ac|bd): p=a, q=c, r=b, s=d
            eri_block(self.L, self.naux, nmo, ((o, nmo), (o + c0, o + c1), (o, nmo), (o, nmo)), W,
                      [v, v ** 3, 1, v * v])
            rows = W.view(n * v, X)
            if precision == "fp32":
                out.view(X, X)[c0 * v:c1 * v].copy_(rows)
            else:
                part = (bf16_planes if precision == "bf16" else bf16_planes_hi)(rows, n * v, X)
                rec[:, c0 * v:c1 * v].copy_(part.view(Xb, n * v, rl))
                del part
        del Wc, W, rows
        return out

    def prepared(self, dtype):
        """Static operands in the layouts the update reads, per dtype, each built
        on first use and cached (Lov2 and Lovov share one tensor)."""
        key = str(dtype)
        if key not in self._prep:
            self._prep[key] = _LazyOperands(self, dtype)
        return self._prep[key]

    def release_fp64_ladder(self):
        """Free the FP64 ladder (v^4 doubles) once the reduced-precision operand
        and its column sums are built: a solve that runs in FP32/BF16 only
        needs those."""
        import torch

        if self.ladder is None:
            return
        v = self.nvir
        self._s_ab = self.ladder.reshape(v * v, v * v).sum(0).reshape(v, v)
        for d in list(self._prep):
            if not isinstance(d, tuple) and "ladder" in self._prep[d]:
                del self._prep[d]["ladder"]
        self._prep.pop(("sparse", "ladder", "fp64"), None)
        self.ladder = None
        torch.cuda.empty_cache()
