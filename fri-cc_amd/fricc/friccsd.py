"""Device-resident FRI-CCSD: ``FRICCSD``, ``update_amps``, ``kernel``.

Mirrors ``fricc/friccsd.py`` of the reference (same settings dict, same call
order, same compressed/uncompressed branches) with every array a torch CUDA
tensor.  The compressions and the 11 sparse contractions run in the gfx950
kernels of libfricc_hip.so (``py_rccsd``); the dense remainder (PySCF's
``imd.*`` formulas, fricc/friccsd.py:144-192) is one MFMA contraction per
product (``fricc.dense``, libfricc_hip's own kernel, no vendor BLAS) over
operands pre-arranged once per system (``DeviceERIs.prepared``) or read in
place through their strides.

Differences from the reference, each deliberate:
* ``cc_Wvvvv`` (fricc/friccsd.py:325-331) is never materialised: its bare part
  (ac|bd) is the static ladder operand and its t1 dressing is contracted with
  the sparse T2 directly (two extra gathered contractions), which is the same
  sum in a different order.
* precision (setting ``"precision"``, PRECISION_MODES): "fp64" (plain, the
  reference's arithmetic); "fp32": sparse contractions on FP32 MFMA and the
  O(o^2 v^3) dense products in FP32; "bf16-ladder": the ladder on BF16 MFMA
  (one pass, FP32 accumulate), every other contraction FP32; "mixed": sparse
  contractions on BF16 MFMA with hi/lo split operands (hi*hi + hi*lo +
  lo*hi, FP32 accumulate) and FP32 dense products.  Compression, the energy
  and the denominators stay FP64 in all modes.
* ``cc._iter`` is set by ``kernel`` (the reference never sets it, so
  ``unbiased=True`` raised AttributeError; quirk A.3), and both the
  ``"unbias"`` and ``"unbiased"`` keys are accepted (quirk A.2).
* The settings dict passed in is copied, not mutated (quirk A.12).
* ``vvvv_t1t1``: "reference" (default) keeps the t1*t1 term exactly as
  fricc/friccsd.py:222 writes it, t_ia t_jb sum_cd W_abcd; "rccsd" uses the
  CCSD tau form sum_cd W_abcd t_ic t_jd.
"""
import copy
import logging
import math
import time
import weakref

import numpy as np

from . import _capi
from .py_rccsd import SparseTensor4d
from .uniforms import GlibcUniforms

log = logging.getLogger("fricc")

default_fri_settings = {
    "m_keep": 1000,
    "compression": "fri",
    "sampling_method": "systematic",
    "verbose": False,
    "compressed_contractions": ["O^2V^4", "O^3V^3", "O^4V^2"],
    "burn_in": 100,
    "unbias": False,
}

ALLOWED_CONTRACTIONS = ["O^2V^4", "O^3V^3", "O^4V^2"]

# precision mode -> (precision of the sparse contractions, precision of the
# ladder 2323); the dense products run in FP64 for "fp64", else FP32, and the
# compressions, denominators, energy and norm are FP64 in every mode.
#   "fp64": plain; "fp32": every contraction FP32;
#   "bf16-ladder": the ladder on single-pass BF16 MFMA (FP32 accumulate),
#       every other contraction FP32;
#   "mixed": every sparse contraction on three-pass BF16 (hi*hi + hi*lo +
#       lo*hi, about 16 significant bits per product), FP32 accumulate.
PRECISION_MODES = {"fp64": ("fp64", "fp64"), "fp32": ("fp32", "fp32"), "bf16-ladder": ("fp32", "bf16x1"),
                   "mixed": ("bf16", "bf16")}


def _torch():
    import torch

    return torch


def _contract(term, A, wshape, T, out, precision):
    """out += term(A, T) with A already in the kernel layout (fricc_contract_prepared)."""
    _capi.check(_capi.lib().fricc_contract_prepared(
        term.encode(), _vp(A), _capi.i64x4(wshape), T.handle(), _vp(out), _capi.i64x4(out.shape),
        _capi.i64x4(out.stride()), _capi.PRECISIONS[precision], _capi.stream_ptr()))


def _contract_spec(spec, alpha, A, wshape, T, out, precision):
    _capi.check(_capi.lib().fricc_contract_spec(
        spec.encode(), float(alpha), _vp(A), 1, _capi.i64x4(wshape), _capi.i64x4([0, 0, 0, 0]), T.handle(),
        _vp(out), _capi.i64x4(out.shape), _capi.i64x4(out.stride()), _capi.PRECISIONS[precision],
        _capi.stream_ptr()))


def _vp(t):
    import ctypes

    return ctypes.c_void_p(t.data_ptr())


class SectionTimer:
    """HIP-event timing of update_amps sections on the current stream
    (enabled with ``cc.timer = SectionTimer()``)."""

    def __init__(self):
        self.events = {}

    def start(self, name):
        import torch

        e = torch.cuda.Event(enable_timing=True)
        e.record()
        self.events.setdefault(name, []).append([e, None])

    def stop(self, name):
        import torch

        e = torch.cuda.Event(enable_timing=True)
        e.record()
        self.events[name][-1][1] = e

    def ms(self):
        """{section: [ms per call]} (synchronizes)."""
        import torch

        torch.cuda.synchronize()
        return {k: [a.elapsed_time(b) for a, b in v if b is not None] for k, v in self.events.items()}


class _NoTimer:
    def start(self, name):
        pass

    def stop(self, name):
        pass


def _prepare(T, spec, wshape, oshape, precision):
    _capi.check(_capi.lib().fricc_sparse_prepare(T.handle(), spec.encode(), _capi.i64x4(wshape),
                                                 _capi.i64x4(oshape), _capi.PRECISIONS[precision],
                                                 _capi.stream_ptr()))


def _prepare_views(cc, ts, ts_alt, o, v, sp, ccs):
    """Build the gathered views of T that the step's compressed contractions
    read (each contraction finds its view built and reuses it)."""
    if ts_alt is not None and "O^4V^2" in ccs:
        _prepare(ts_alt, "kcld,ijcd->klij", (o, v, o, v), (o, o, o, o), sp)
    if ts_alt is not None and "O^3V^3" in ccs:
        _prepare(ts_alt, "ldkc,ilda->akic", (o, v, o, v), (v, o, o, v), sp)
        _prepare(ts_alt, "ldkc,ilad->akic", (o, v, o, v), (v, o, o, v), sp)
        _prepare(ts_alt, "lckd,ilda->akci", (o, v, o, v), (v, o, v, o), sp)
    if "O^2V^4" in ccs:
        _prepare(ts, "kdac,ijcd->ijka", (o, v, v, v), (o, o, o, v), sp)
        _prepare(ts, "abcd,ijcd->ijab", (v, v, v, v), (o, o, v, v), cc.ladder_precision)
    if "O^4V^2" in ccs:
        _prepare(ts, "klij,klab->ijab", (o, o, o, o), (o, o, v, v), sp)
    if "O^3V^3" in ccs:
        _prepare(ts, "akic,kjcb->ijab", (v, o, o, v), (o, o, v, v), sp)
        _prepare(ts, "akic,kjbc->ijab", (v, o, o, v), (o, o, v, v), sp)
        _prepare(ts, "bkci,kjac->jiba", (v, o, v, o), (o, o, v, v), sp)


def _dyn_operand(A, R, X, precision):
    """Per-iteration A operand (Woooo, Wvoov, Wvovo) in the kernel precision."""
    torch = _torch()
    if precision == "fp64":
        return A.contiguous()
    if precision == "fp32":
        return A.to(torch.float32).contiguous()
    from .systems import bf16_planes

    return bf16_planes(A, R, X)


def _tr(t):
    """tmp.transpose(1,0,3,2)"""
    return t.permute(1, 0, 3, 2)


# Compressed terms as index specs with the sign they carry into the
# symmetrised buffer S of update_amps (t2new += S + S.transpose(1,0,3,2)).
_O3V3 = (("1302", "akic,kjcb->ijab", 2.0), ("1202", "akci,kjcb->ijab", -1.0),
         ("1303", "akic,kjbc->ijab", -1.0), ("1203", "bkci,kjac->ijab", -1.0))
_PREC_ID = {"fp64": 0, "fp32": 1}


def _out_prec(t):
    return _capi.FP32 if t.dtype == _torch().float32 else _capi.FP64


# When a dict: every compressed contraction of a step is also evaluated alone
# into a zeroed FP64 buffer stored under its spec, so the contributions of
# two steps (e.g. a reduced precision mode against FP64 on the same sampled
# T2) can be compared term by term.  Off in normal runs.
TERM_PROBE = None


def _contract_into(spec, alpha, A, wshape, T, out, precision):
    """out += alpha * einsum(spec, W, T) with A the prepared operand of W."""
    outs = [out]
    if TERM_PROBE is not None:
        probe = _torch().zeros(out.shape, dtype=_torch().float64, device=out.device)
        TERM_PROBE[spec] = probe
        outs.append(probe)
    for o_ in outs:
        _capi.check(_capi.lib().fricc_contract_spec(
            spec.encode(), float(alpha), _vp(A), 1, _capi.i64x4(wshape), _capi.i64x4([0, 0, 0, 0]), T.handle(),
            _vp(o_), _out_prec(o_), _capi.i64x4(o_.shape), _capi.i64x4(o_.stride()), _capi.PRECISIONS[precision],
            _capi.stream_ptr()))


def _t2_reduced(cc, t2, cdt):
    """t2 in the products' reduced precision: the copy the previous step's
    finalize wrote when t2 is that step's output, unmodified (same tensor,
    same version counter) and no graph is being recorded; else a conversion."""
    torch = _torch()
    c = getattr(cc, "_t2_low", None)
    if (c is not None and c[2].dtype == cdt and c[0]() is t2 and t2._version == c[1]
            and not torch.cuda.is_current_stream_capturing()):
        return c[2]
    return t2.to(cdt)


def _check_m_keep(cc, t2):
    m_keep = cc.fri_settings["m_keep"]
    if m_keep > t2.numel() or m_keep == 0:
        raise ValueError("Bad m_keep value! The following condition wasn't met 0 < m_keep <= t2.size")


def _needs_alt(cc):
    ccs = cc.fri_settings["compressed_contractions"]
    return "O^3V^3" in ccs or "O^4V^2" in ccs


def _draw_uniforms(cc):
    """The uniforms of one step, in the reference's order (t2_sparse, then
    t2_sparse_alt, fricc/friccsd.py:99-138); none for "largest"."""
    if cc.fri_settings["compression"] != "fri":
        return []
    return [next(cc.uniforms) for _ in range(cc._n_compressions())]


def _set_uniforms(cc, us):
    """Write the step's uniforms into the device buffer the compressions read
    (fill kernels: stream-ordered, no host round trip)."""
    ud = cc._udev()
    for k, u in enumerate(us):
        ud[k].fill_(float(u))


def _finish_step(cc, t1new, t2new, red):
    """The one host read of a step: {E2, |dt2|^2, E1, |dt1|^2} and the
    compression statuses (raises on non-finite amplitudes, p > 1)."""
    e2, dn, e1, d1 = red.tolist()
    info = []
    for k in sorted(cc._slots):
        if k in cc._used_slots:
            info.append(cc._slots[k].sync_status())
    cc._last_step = (t1new, t2new, e1 + e2, math.sqrt(d1 + dn))
    cc.sparse_info = (info[0] if info else None, info[1] if len(info) > 1 else None)


def update_amps(cc, t1, t2, eris):
    """One FRI-CCSD Jacobi step (fricc/friccsd.py:29-301) on the GPU: draws the
    step's uniforms, runs the device part (no host round trip) and reads back
    the energy, the amplitude-change norm and the compression statuses once."""
    _check_m_keep(cc, t2)
    _set_uniforms(cc, _draw_uniforms(cc))
    t1new, t2new, red = cc._device_step(t1, t2, eris)
    _finish_step(cc, t1new, t2new, red)
    return t1new, t2new


def _update_amps_device(cc, t1, t2, eris):
    """One FRI-CCSD Jacobi step (fricc/friccsd.py:29-301) on the GPU.

    Hirata et al., J. Chem. Phys. 120, 2581 (2004), eqs. 35-36.  Same terms as
    the reference, accumulated differently: every contribution the reference
    adds as ``tmp + tmp.transpose(1,0,3,2)`` goes into one buffer S (GEMMs
    accumulate in place, compressed contractions add with their sign), every
    other T2 contribution into the output buffer, and one fused kernel adds
    S + S^T and the t1*t1 vvvv term, divides by the denominators and reduces
    the energy and the amplitude-change norm (fricc_t2_finalize).

    Nothing here waits for the GPU: the uniforms come from the device buffer
    cc._udev(), compression statistics and errors stay on the device, and the
    reductions are returned as a device tensor {E2, |dt2|^2, E1, |dt1|^2}, so
    the whole call can be recorded as one graph."""
    torch = _torch()
    st = cc.fri_settings
    m_keep = st["m_keep"]
    ccs = st["compressed_contractions"]
    _capi.check(_capi.lib().fricc_workspace_select(int(cc.arena)))
    sp = cc.sparse_precision
    timings = cc._timings = {}
    tm = cc.timer if getattr(cc, "timer", None) is not None else _NoTimer()
    nocc, nvir = t1.shape
    o, v = nocc, nvir
    ov = o * v
    f64 = torch.float64
    cdt = cc.dense_dtype  # dense GEMMs and the S / W buffers
    P = eris.prepared(cdt)
    P64 = eris.prepared(f64)
    fock = eris.fock
    mo_e_o = eris.mo_energy[:o]
    mo_e_v = eris.mo_energy[o:] + cc.level_shift
    fov = fock[:o, o:]
    foo = fock[:o, :o]
    fvv = fock[o:, o:]
    t1 = t1.contiguous()
    t2 = t2.contiguous()
    t2c = _t2_reduced(cc, t2, cdt) if cdt != f64 else t2
    C = lambda a: a.to(f64) if cdt != f64 else a  # noqa: E731

    # ---- compression (fricc/friccsd.py:99-138)
    # The two samples of T2 run on the arena's side stream while the dense
    # products that do not read the compressed T (Fock intermediates, the
    # t1 equation, the W builds) run on this one; the stream joins before
    # the first compressed contraction.  The unbiased branch (iteration
    # dependent) and recorded steps (cc._single_stream: a captured graph that
    # forked to a second stream kept its intermediate buffers allocated after
    # it was destroyed, on this HIP/PyTorch stack) stay on one stream.
    t0 = time.perf_counter()
    comp = st["compression"]
    smeth = st["sampling_method"]
    ud = cc._udev()
    ts_alt = None
    main = torch.cuda.current_stream()
    averaging = _needs_alt(cc) and st["unbiased"] and cc._iter >= st["burn_in"]
    single = averaging or getattr(cc, "_single_stream", False)
    side = None if single else arena_side_stream(cc.arena)
    if side is not None:
        side.wait_stream(main)
    with torch.cuda.stream(side if side is not None else main):
        if side is not None:
            _capi.check(_capi.lib().fricc_workspace_select(side_arena(cc.arena)))
        tm.start("compress")
        if _needs_alt(cc) and not averaging:
            # t2_sparse and t2_sparse_alt: two samples of the same T2 (the
            # second uniform for the second), which share everything but the
            # sampling
            ualt = ud[1:2] if comp == "fri" else ud[0:1]
            ts, ts_alt = cc._slot(0, t2.shape).recompress_pair_async(cc._slot(1, t2.shape), t2, int(m_keep), comp,
                                                                     ud[0:1], ualt, sampling_method=smeth)
            cc._used_slots = {0, 1}
        else:
            ts = cc._slot(0, t2.shape).recompress_async(t2, int(m_keep), comp, ud[0:1], sampling_method=smeth)
            cc._used_slots = {0}
        if averaging:
            # unbiased averaging after burn-in: t2_sparse_alt samples the running average
            ualt = ud[1:2] if comp == "fri" else ud[0:1]
            it = cc._iter - st["burn_in"]
            if it == 0:
                cc._t2_avg = t2.clone()
            else:
                cc._t2_avg = t2 / it + (it - 1) / it * cc._t2_avg
            ts_alt = cc._slot(1, t2.shape).recompress_async(cc._t2_avg, int(m_keep), comp, ualt,
                                                            sampling_method=smeth)
            cc._used_slots = {0, 1}
        tm.stop("compress")
        if side is not None:
            # the views every compressed contraction reads, built here too
            _prepare_views(cc, ts, ts_alt, o, v, sp, ccs)
        _capi.check(_capi.lib().fricc_workspace_select(int(cc.arena)))
    timings["compress"] = time.perf_counter() - t0

    def join():
        if side is not None:
            main.wait_stream(side)

    # every dense product below is one MFMA contraction (fricc.dense): E in
    # the dense precision (FP32 in the reduced modes), E64 in FP64; operands
    # are read in place through their strides, FP64 ones converted on load
    from .dense import einsum as _de

    def E(spec, A, B, out=None, alpha=1.0, beta=0.0, c_src=None):
        return _de(spec, A, B, out=out, alpha=alpha, beta=beta, compute=cdt, c_src=c_src)

    def E64(spec, A, B, out=None, alpha=1.0, beta=0.0):
        return _de(spec, A, B, out=out, alpha=alpha, beta=beta, compute=f64)

    # ---- Foo, Fvv, Fov (imd.cc_Foo / cc_Fvv / cc_Fov)
    # Fkc = sum_ld [2 (kc|ld) - (kd|lc)] t1[l,d]
    Fov_ = E("klcd,ld->kc", P["Lov2"], t1)
    # Fki = sum_lcd Lov2[k,l,c,d] t2[i,l,c,d] + Fkc t1[i,c];  Fac = -sum_kld t2[k,l,a,d] Lov[k,l,d,c] - t1[k,a] Fkc
    Fki = E("klcd,ilcd->ki", P["Lov2"], t2c)
    E("kc,ic->ki", Fov_, t1, out=Fki, beta=1.0)
    Fac = E("klad,kldc->ac", t2c, P["Lov"], alpha=-1.0)
    E("ka,kc->ac", t1, Fov_, out=Fac, alpha=-1.0, beta=1.0)
    Foo_u = C(Fki) + foo
    Fvv_u = C(Fac) + fvv
    Fov = C(Fov_) + fov
    Foo = Foo_u.clone()
    Fvv = Fvv_u.clone()
    Foo.diagonal().sub_(mo_e_o)
    Fvv.diagonal().sub_(mo_e_v)

    # ---- T1 equation (fricc/friccsd.py:152-176)
    t1new = E64("ik,ka->ia", E64("ic,kc->ik", t1, fov), t1, alpha=-2.0)
    E64("ic,ac->ia", t1, Fvv, out=t1new, beta=1.0)
    E64("ki,ka->ia", Foo, t1, out=t1new, alpha=-1.0, beta=1.0)
    # 2 sum_kc Fov[k,c] t2[k,i,c,a] - sum_kc Fov[k,c] t2[i,k,c,a]
    E("kc,kica->ia", Fov, t2c, out=t1new, alpha=2.0, beta=1.0)
    E("kc,ikca->ia", Fov, t2c, out=t1new, alpha=-1.0, beta=1.0)
    E64("ik,ka->ia", E64("ic,kc->ik", t1, Fov), t1, out=t1new, beta=1.0)
    t1new += fov
    # 2 sum_kc (kc|ai) t1[k,c] - sum_kc (ki|ac) t1[k,c]
    E("kcai,kc->ia", P["ovvo"], t1, out=t1new, alpha=2.0, beta=1.0)
    E("kiac,kc->ia", P["oovv"], t1, out=t1new, alpha=-1.0, beta=1.0)
    # O^2V^3: 2 kdac,ikcd - kcad,ikcd  ==  t2[i,(kcd)] Z[(kcd),a], and
    # LvvT1[a,c] = sum_kd Z[k,c,d,a] t1[k,d]  (== 2 kdac,kd - kcad,kd -> ac):
    # both read the o*v^3 operand Z, so one kernel streams it once
    Z = P["Z"]
    if o <= 32 and v % 128 == 0:
        LvvT1 = torch.empty((v, v), dtype=f64, device=t1.device)
        t1f = t1.to(cdt).contiguous()
        t2z = t2c.contiguous()
        _capi.check(_capi.lib().fricc_dense_zpair(o, v, _PREC_ID["fp64" if cdt == f64 else "fp32"], _vp(t2z),
                                                  _vp(t1f), _vp(Z), _vp(t1new), _vp(LvvT1), _capi.stream_ptr()))
    else:
        E("ikcd,kcda->ia", t2c, Z, out=t1new, beta=1.0)
        LvvT1 = C(E("kd,kcda->ac", t1, Z))
    E64("ic,ac->ia", t1, LvvT1, out=t1new, beta=1.0)
    # -2 lcki,klac + kcli,klac  ==  sum_klc Aoo[k,l,i,c] t2[k,l,a,c]
    E("klic,klac->ia", P["Aoo"], t2c, out=t1new, beta=1.0)
    # Pk[k,i] = sum_lc (lc|ki) t1[l,c] and Qk[k,i] = sum_lc (kc|li) t1[l,c], shared
    # by the t1 equation and Loo
    Pk = E64("kilc,lc->ki", P64["ovoo_P"], t1)
    Qk = E64("kicl,lc->ki", P64["ovoo_Q"], t1)
    E64("ki,ka->ia", Qk - 2 * Pk, t1, out=t1new, beta=1.0)  # -2 lcki,lc,ka + kcli,lc,ka

    # ---- Loo, Lvv (imd.Loo / imd.Lvv)
    Loo = Foo_u + 2 * Pk - Qk
    E64("kc,ic->ki", fov, t1, out=Loo, beta=1.0)
    Lvv = Fvv_u + LvvT1
    E64("ka,kc->ac", t1, fov, out=Lvv, alpha=-1.0, beta=1.0)
    Loo.diagonal().sub_(mo_e_o)
    Lvv.diagonal().sub_(mo_e_v)

    # ---- W intermediates (fricc/friccsd.py:194-207)
    t0 = time.perf_counter()
    # U[k,c,l,i] = sum_d (kc|ld) t1[i,d]: with (ld|kc) = (kc|ld) and (kc|ai) = (ia|kc)
    # it gives the t1 pieces of Woooo, of both A_vo intermediates and of T2p
    U = E("kcld,id->kcli", P["ovov"], t1)
    # lcki,jc->klij + kclj,ic->klij: with Y[l,k,i,c] = (lc|ki) the first is
    # sum_c Y[l,k,i,c] t1[j,c], the second sum_c Y[k,l,j,c] t1[i,c]
    Woooo = E64("lkic,jc->klij", P64["ovoo_Y"], t1)
    E64("kljc,ic->klij", P64["ovoo_Y"], t1, out=Woooo, beta=1.0)
    if "O^4V^2" not in ccs:
        E64("kcld,ijcd->klij", eris.ovov, t2, out=Woooo, beta=1.0)
    E("ic,kclj->klij", t1, U, out=Woooo, beta=1.0)  # sum_cd (kc|ld) t1[i,c] t1[j,d]
    Woooo += eris.oooo.permute(0, 2, 1, 3)
    # Wvoov as A_voov[k,c,i,a] = Wvoov[a,k,i,c];  Wvovo as A_vovo[k,c,i,a] = Wvovo[a,k,c,i]
    # (the integral part is read as the beta term's source: no copy of it first)
    A_voov = torch.empty_like(P["ovvo_p"])
    E("id,kcda->kcia", t1, P["P1"], out=A_voov, beta=1.0, c_src=P["ovvo_p"])
    A_vovo = torch.empty_like(P["oovv_p"])
    E("id,kcda->kcia", t1, P["P2"], out=A_vovo, beta=1.0, c_src=P["oovv_p"])
    # - sum_l (kcli + sum_d ldkc t1[i,d]) t1[l,a];  - sum_l (lcki + sum_d lckd t1[i,d]) t1[l,a]:
    # both read N1[p,c,q,i] = (pc|qi) + U[p,c,q,i], with (p,q) = (k,l) and (l,k)
    N1 = P["ovoo"] + U
    E("kcli,la->kcia", N1, t1, out=A_voov, alpha=-1.0, beta=1.0)
    E("lcki,la->kcia", N1, t1, out=A_vovo, alpha=-1.0, beta=1.0)
    del N1
    Wvoov = A_voov.permute(3, 0, 2, 1)  # [a,k,i,c] view
    Wvovo = A_vovo.permute(3, 0, 1, 2)  # [a,k,c,i] view
    join()  # the compressed contractions start here
    if "O^4V^2" in ccs:
        _contract_into("kcld,ijcd->klij", 1.0, eris.sparse_operand("A1323", sp), (o, v, o, v), ts_alt, Woooo, sp)
    if "O^3V^3" in ccs:
        A0112 = eris.sparse_operand("A0112", sp)
        A0313 = eris.sparse_operand("A0313", sp)
        _contract_into("ldkc,ilda->akic", -0.5, A0112, (o, v, o, v), ts_alt, Wvoov, sp)
        # 0313 and 0113 read the same view of T with operands in the same
        # [(l,d)][(k,c)] layout: one contraction with (ld|kc) - 1/2 (lc|kd)
        _contract_into("ldkc,ilad->akic", 1.0, eris.sparse_operand("A0113c", sp), (o, v, o, v), ts_alt, Wvoov, sp)
        _contract_into("lckd,ilda->akci", -0.5, A0313, (o, v, o, v), ts_alt, Wvovo, sp)
    else:
        ovov_c = P["ovov"]
        E("ldkc,ilda->akic", ovov_c, t2c, out=Wvoov, alpha=-0.5, beta=1.0)
        E("lckd,ilad->akic", ovov_c, t2c, out=Wvoov, alpha=-0.5, beta=1.0)
        E("ldkc,ilad->akic", ovov_c, t2c, out=Wvoov, beta=1.0)
        E("lckd,ilda->akci", ovov_c, t2c, out=Wvovo, alpha=-0.5, beta=1.0)
    timings["W"] = time.perf_counter() - t0

    # ---- direct T2 terms into the output buffer acc (fricc/friccsd.py:210-235)
    t0 = time.perf_counter()
    # the ladder's t1 dressing (Gx/Gy contractions and their t1 products) in
    # the dense-GEMM precision
    # X[ij,k,a] = sum_cd (kd|ac) T[ij,cd] and Y[ij,k,b] = sum_cd (kc|bd) T[ij,cd]
    # live in halves of two stacked buffers whose other halves hold t1, so that
    # acc = -(X.t1 + t1.Y) below is one batched product over ij with the sum
    # running over (h, k):  XT[ij,0] = X, XT[ij,1] = t1;  YT[ij,0] = t1, YT[ij,1] = Y
    XT = torch.empty((o, o, 2, o, v), dtype=cdt, device=t2.device)
    YT = torch.empty((o, o, 2, o, v), dtype=cdt, device=t2.device)
    X, Y = XT[:, :, 0], YT[:, :, 1]
    X.zero_()
    Y.zero_()
    t1c = t1.to(cdt)
    XT[:, :, 1].copy_(t1c.expand(o, o, o, v))
    YT[:, :, 0].copy_(t1c.expand(o, o, o, v))
    if "O^2V^4" in ccs:
        _contract_into("kdac,ijcd->ijka", 1.0, eris.sparse_operand("Gx", sp), (o, v, v, v), ts, X, sp)
        _contract_into("kcbd,ijcd->ijkb", 1.0, eris.sparse_operand("Gy", sp), (o, v, v, v), ts, Y, sp)
    else:
        E("kdac,ijcd->ijka", eris.ovvv, t2c, out=X, beta=1.0)
        E("kcbd,ijcd->ijkb", eris.ovvv, t2c, out=Y, beta=1.0)
    if "O^4V^2" in ccs:
        # klij,ka,lb -> ijab  ==  sum_l Q[ij,l,a] t1[l,b],  Q[ij,l,a] = sum_k Woooo[k,l,i,j] t1[k,a]
        E("klij,ka->ijla", Woooo, t1, out=X, alpha=-1.0, beta=1.0)
    # acc[ij,a,b] = -sum_k X[ij,k,a] t1[k,b] - sum_k t1[k,a] Y[ij,k,b]
    acc = torch.empty((o, o, v, v), dtype=f64, device=t2.device)
    E("ijhka,ijhkb->ijab", XT, YT, out=acc, alpha=-1.0)
    del XT, YT, X, Y
    if "O^4V^2" in ccs:
        _contract_into("klij,klab->ijab", 1.0, _dyn_operand(Woooo, o * o, o * o, sp), (o, o, o, o), ts, acc, sp)
    else:
        tau = t2 + t1[:, None, :, None] * t1[None, :, None, :]
        E64("klij,klab->ijab", Woooo, tau, out=acc, beta=1.0)
    if cc.vvvv_t1t1 == "reference":
        # t_ia t_jb sum_cd W_abcd (fricc/friccsd.py:222, quirk A.1), added in the finalize kernel
        G1 = E64("ka,kb->ab", P64["g_ka"], t1)
        wsum = (P64["s_ab"] - G1 - G1.T).contiguous()
    else:
        wsum = None
        # sum_cd W_abcd t_ic t_jd with W = (ac|bd) - sum_k (kd|ac) t_kb - sum_k (kc|bd) t_ka
        tt = t1[:, None, :, None] * t1[None, :, None, :]
        E64("ijcd,cdab->ijab", tt, P64["ladder"], out=acc, beta=1.0)
        Xd = E64("kdai,jd->ijka", E64("kdac,ic->kdai", eris.ovvv, t1), t1)
        Yd = E64("kbdi,jd->ijkb", E64("kcbd,ic->kbdi", eris.ovvv, t1), t1)
        E64("ijka,kb->ijab", Xd, t1, out=acc, alpha=-1.0, beta=1.0)
        E64("ijkb,ka->ijab", Yd, t1, out=acc, alpha=-1.0, beta=1.0)
    if "O^2V^4" in ccs:
        spl = cc.ladder_precision
        _prepare(ts, "2323", (v, v, v, v), (o, o, v, v), spl)
        tm.start("ladder")
        _contract_into("abcd,ijcd->ijab", 1.0, eris.sparse_operand("ladder", spl), (v, v, v, v), ts, acc, spl)
        tm.stop("ladder")
    else:
        E64("ijcd,cdab->ijab", t2, P64["ladder"], out=acc, beta=1.0)
    timings["ladder"] = time.perf_counter() - t0

    # ---- symmetrised T2 terms into S (fricc/friccsd.py:178-186, 237-279)
    t0 = time.perf_counter()
    # tmp = (ia|cb) t1[j,c] - sum_k oovv[k,i,b,c] t1[j,c] t1[k,a]
    S = E("jc,icab->ijab", t1, P["Qv"])
    # - sum_k t1[k,a] Mt[k,i,j,b] with Mt[k,i,j,b] = sum_c (ki|bc) t1[j,c], and
    # - sum_k T2p[i,a,k,j] t1[k,b] with T2p = sum_c (ia|kc) t1[j,c] + ovoo[i,a,j,k]
    # (sum_c (ia|kc) t1[j,c] = U[i,a,k,j]): one batched product over ij with
    # the sum over (h, k), from stacked buffers written in place:
    # AS[ij,0,k,a] = t1[k,a], AS[ij,1,k,a] = T2p[i,a,k,j];
    # BS[ij,0,k,b] = Mt[k,i,j,b], BS[ij,1,k,b] = t1[k,b]
    AS = torch.empty((o, o, 2, o, v), dtype=cdt, device=t2.device)
    BS = torch.empty((o, o, 2, o, v), dtype=cdt, device=t2.device)
    AS[:, :, 0].copy_(t1c.expand(o, o, o, v))
    BS[:, :, 1].copy_(t1c.expand(o, o, o, v))
    E("kibc,jc->kijb", P["oovv"], t1, out=BS[:, :, 0].permute(2, 0, 1, 3))
    torch.add(U, P["ovoo"].permute(0, 1, 3, 2), out=AS[:, :, 1].permute(0, 3, 2, 1))
    E("ijhka,ijhkb->ijab", AS, BS, out=S, alpha=-1.0, beta=1.0)
    del AS, BS
    # + Lvv[a,c] t2[i,j,c,b] - Loo[k,i] t2[k,j,a,b]
    E("ac,ijcb->ijab", Lvv, t2c, out=S, beta=1.0)
    E("ki,kjab->ijab", Loo, t2c, out=S, alpha=-1.0, beta=1.0)
    if "O^3V^3" in ccs:
        _prepare(ts, "1302", (v, o, o, v), (o, o, v, v), sp)
        _prepare(ts, "1303", (v, o, o, v), (o, o, v, v), sp)
        tm.start("O3V3")
        # 1302 and 1202 read the same view of T with operands in the same
        # [(k,c)][(i,a)] layout: one contraction with 2 Wvoov - Wvovo
        if sp == "bf16":
            from .systems import bf16_planes_lincomb3

            # the three BF16 operands of this stage from one read of Wvoov and Wvovo
            A12, Av, Avo = bf16_planes_lincomb3(A_voov, A_vovo, 2.0, -1.0, ov, ov)
            a12 = 1.0
        else:
            # 2 Wvoov - Wvovo as 2 (Wvoov - Wvovo / 2): one pass instead of
            # two, and scaling by 2 is exact, so the sum is bit for bit the same
            A12 = _dyn_operand(torch.add(A_voov, A_vovo, alpha=-0.5), ov, ov, sp)
            a12 = 2.0
            Av = _dyn_operand(A_voov, ov, ov, sp)
            Avo = _dyn_operand(A_vovo, ov, ov, sp)
        _contract_into("akic,kjcb->ijab", a12, A12, (v, o, o, v), ts, S, sp)
        del A12
        _contract_into("akic,kjbc->ijab", -1.0, Av, (v, o, o, v), ts, S, sp)    # 1303
        # 1203, added transposed (S[j,i,b,a]): S enters only as S + S^T, and
        # this way the gathered columns (j, a) run along S's unit-stride axis
        _contract_into("bkci,kjac->jiba", -1.0, Avo, (v, o, v, o), ts, S, sp)
        tm.stop("O3V3")
    else:
        E("akic,kjcb->ijab", Wvoov, t2c, out=S, alpha=2.0, beta=1.0)
        E("akci,kjcb->ijab", Wvovo, t2c, out=S, alpha=-1.0, beta=1.0)
        E("akic,kjbc->ijab", Wvoov, t2c, out=S, alpha=-1.0, beta=1.0)
        E("bkci,kjac->ijab", Wvovo, t2c, out=S, alpha=-1.0, beta=1.0)
    timings["O3V3"] = time.perf_counter() - t0

    # ---- denominators (fricc/friccsd.py:281-284), energy and norm: FP64, fused
    eia = (mo_e_o[:, None] - mo_e_v).contiguous()
    t1new = (t1new / eia).contiguous()
    red = torch.empty(4, dtype=f64, device=t2.device)
    # the next step's FP32 copy of T2 comes out of the finalize (not in a
    # recorded graph, whose replays convert their own input)
    t2low = None
    if cdt == torch.float32 and not torch.cuda.is_current_stream_capturing():
        t2low = torch.empty(acc.shape, dtype=torch.float32, device=acc.device)
    _capi.check(_capi.lib().fricc_t2_finalize(
        o, v, _vp(P64["ovovT"]), _vp(acc), _vp(S), _PREC_ID["fp32" if S.dtype == torch.float32 else "fp64"],
        _vp(t1), _vp(wsum) if wsum is not None else None, _vp(eia), _vp(t2), None, _vp(t1new),
        _vp(t2low) if t2low is not None else None, _vp(red), _capi.stream_ptr()))
    cc._t2_low = (weakref.ref(acc), acc._version, t2low) if t2low is not None else None
    red[2].copy_(2 * (fov * t1new).sum())
    red[3].copy_(((t1new - t1) ** 2).sum())
    return t1new, acc, red


def energy(cc, t1, t2, eris):
    """CCSD correlation energy (PySCF CCSD.energy, called at fricc/friccsd.py:402,427),
    FP64: 2 sum f_ia t_ia + sum tau_ijab [2(ia|jb) - (ib|ja)], tau = t2 + t1 t1."""
    o = t1.shape[0]
    P = eris.prepared(_torch().float64)
    fov = eris.fock[:o, o:]
    tau = t2 + t1[:, None, :, None] * t1[None, :, None, :]
    e = 2 * (fov * t1).sum() + (tau * P["Lovov"]).sum()
    return float(e.item())


def init_amps(eris):
    """MP2 guess (PySCF CCSD.init_amps, via get_init_guess at fricc/friccsd.py:396):
    t1 = f_ov / e_ia, t2 = (ia|jb) / D_ijab; returns (emp2, t1, t2)."""
    torch = _torch()
    o = eris.nocc
    mo_e = eris.mo_energy
    eia = mo_e[:o, None] - mo_e[None, o:]
    t1 = eris.fock[:o, o:] / eia
    t2 = (eris.prepared(torch.float64)["ovovT"] / (eia[:, None, :, None] + eia[None, :, None, :])).contiguous()
    emp2 = (t2 * eris.prepared(torch.float64)["Lovov"]).sum()
    return float(emp2.item()), t1, t2


def amplitudes_to_vector_norm(t1new, t2new, t1, t2):
    """|| vec(t1new,t2new) - vec(t1,t2) ||_2 with ccsd.amplitudes_to_vector's
    packing (t1, then the lower triangle of t2 as an (ia, jb) matrix),
    fricc/friccsd.py:418-420."""
    torch = _torch()
    o, v = t1.shape
    d2 = (t2new - t2).permute(0, 2, 1, 3).reshape(o * v, o * v)
    tri = torch.tril(d2)
    return float(torch.sqrt(((t1new - t1) ** 2).sum() + (tri ** 2).sum()).item())


_ARENA_STREAMS = {}


def side_arena(arena):
    """The scratch arena of a step's side-stream work (its compressions),
    so that nothing the side stream writes shares a slot with the kernels the
    main stream runs at the same time."""
    if not 0 <= int(arena) < 32:
        raise ValueError("solves use arenas 0..31 (their side arenas are 32..63)")
    return int(arena) + 32


def arena_side_stream(arena):
    """A second stream per arena and device (reused), on which a step runs
    its compressions while the dense products run on the step's own stream."""
    torch = _torch()
    key = (torch.cuda.current_device(), int(arena), "side")
    st = _ARENA_STREAMS.get(key)
    if st is None:
        st = _ARENA_STREAMS[key] = torch.cuda.Stream()
    return st


def arena_stream(arena):
    """The stream of a library scratch arena (one per arena and device, reused).

    A recorded step is warmed up and captured on its arena's stream.  The BLAS
    library keeps one workspace per stream and a captured GEMM keeps using the
    workspace of the stream it was captured on; steps recorded on one shared
    capture stream would all use the same workspace, and replaying them side
    by side (one solve per arena, several at once) would run GEMMs
    concurrently on that one workspace.  One stream per arena gives every
    concurrently replayed step its own; reusing the streams keeps the number
    of workspaces bounded."""
    torch = _torch()
    key = (torch.cuda.current_device(), int(arena))
    st = _ARENA_STREAMS.get(key)
    if st is None:
        st = _ARENA_STREAMS[key] = torch.cuda.Stream()
    return st


def _unpin(arena, device):
    torch = _torch()
    with torch.cuda.device(device):
        _capi.check(_capi.lib().fricc_workspace_pin(int(arena), -1))


def _pin_arena(arena, owner):
    """Pin the library arena ``arena`` on the current device for as long as
    ``owner`` (a recorded graph's holder) lives."""
    _capi.check(_capi.lib().fricc_workspace_pin(int(arena), 1))
    weakref.finalize(owner, _unpin, int(arena), _torch().cuda.current_device())


class GraphStep:
    """``update_amps`` recorded once as a graph for this (nocc, nvir, m) shape
    and replayed every iteration: one launch of the whole step instead of a
    few hundred.  Inputs live in static buffers; per iteration the host writes
    the step's uniforms (fill kernels, stream-ordered), replays, reads the
    step's reductions and statuses once (``_finish_step``) and copies the
    outputs into the inputs.  Needs the unbiased averaging off (its branch
    depends on the iteration)."""

    def __init__(self, cc, eris, t1, t2):
        torch = _torch()
        st = cc.fri_settings
        if cc._iteration_dependent():
            raise ValueError("graph mode needs unbiased averaging off")
        if getattr(cc, "timer", None) is not None:
            raise ValueError("graph mode cannot run with a section timer")
        _check_m_keep(cc, t2)
        self.cc, self.eris = cc, eris
        cc._single_stream = True  # recorded on one stream (see _update_amps_device)
        self.t1 = t1.detach().clone().contiguous()
        self.t2 = t2.detach().clone().contiguous()
        # warm-up step on the arena's stream: allocates every workspace, view
        # and operand the step needs (the BLAS workspace of that stream
        # included), so the recorded step allocates nothing outside the
        # graph's pool; then the capture on the same stream
        st = arena_stream(cc.arena)
        st.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(st):
            _set_uniforms(cc, [0.5] * cc._n_compressions())
            cc._device_step(self.t1, self.t2, eris)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=st):
            self.out = cc._device_step(self.t1, self.t2, eris)
        torch.cuda.synchronize()
        # the graph holds the addresses of the arena's slots: while it lives,
        # growing one of them (an eager call on a larger shape on this arena)
        # is an error instead of freeing memory the graph uses.  The recorded
        # step runs on one stream, so the side arena is not in the graph and
        # stays free for eager steps.
        _pin_arena(cc.arena, self)

    def issue(self):
        """Queue one iteration on the current stream (uniforms, replay); no wait."""
        cc = self.cc
        _set_uniforms(cc, _draw_uniforms(cc))
        self.graph.replay()

    def complete(self, before_copy=None):
        """Read the iteration's energy, norm and statuses (waits for the current
        stream), then queue the copy of its outputs into the inputs.
        ``before_copy`` (optional) is called between the two, with the current
        stream, so a scheduler can order the copies after other work."""
        cc = self.cc
        t1n, t2n, red = self.out
        _finish_step(cc, t1n, t2n, red)
        if before_copy is not None:
            before_copy()
        e, nt = cc._last_step[2], cc._last_step[3]
        self.t1.copy_(t1n)
        self.t2.copy_(t2n)
        cc._last_step = (self.t1, self.t2, e, nt)
        return self.t1, self.t2

    def step(self, t1, t2):
        if t1 is not self.t1:
            self.t1.copy_(t1)
        if t2 is not self.t2:
            self.t2.copy_(t2)
        self.issue()
        return self.complete()


def kernel(mycc, eris=None, t1=None, t2=None, max_cycle=50, tol=1e-8, tolnormt=1e-6, verbose=None):
    """The iteration loop (fricc/friccsd.py:379-456), device-resident: t1, t2
    stay on the GPU; the energy and the convergence/NaN guards are read back
    once per cycle."""
    max_cycle = mycc.max_cycle  # the reference ignores its argument (:390)
    if eris is None:
        eris = mycc.ao2mo()
    if t1 is None and t2 is None:
        t1, t2 = mycc.get_init_guess(eris)
    elif t2 is None:
        t2 = mycc.get_init_guess(eris)[1]
    t0 = time.perf_counter()
    eold = 0.0
    eccsd = mycc.energy(t1, t2, eris)
    log.info("Init E_corr(%s) = %.15g", mycc.__class__.__name__, eccsd)
    mycc.energies = []
    mycc.cycle_times = []
    conv = False
    runner = GraphStep(mycc, eris, t1, t2) if mycc.fri_settings.get("graph") else None
    for istep in range(max_cycle):
        tc = time.perf_counter()
        mycc._iter = istep
        if runner is not None:
            t1new, t2new = runner.step(t1, t2)
        else:
            t1new, t2new = mycc.update_amps(t1, t2, eris)
        last = getattr(mycc, "_last_step", None)
        if last is not None and last[0] is t1new and last[1] is t2new:
            enew, normt = last[2], last[3]  # computed in the fused finalize pass
        else:
            normt = amplitudes_to_vector_norm(t1new, t2new, t1, t2)
            enew = mycc.energy(t1new, t2new, eris)
        t1, t2 = t1new, t2new
        t1new = t2new = None
        eold, eccsd = eccsd, enew
        mycc.energies.append(eccsd)
        mycc.cycle_times.append(time.perf_counter() - tc)
        log.info("cycle = %d  E_corr(%s) = %.15g  dE = %.9g  norm(t1,t2) = %.6g", istep + 1,
                 mycc.__class__.__name__, eccsd, eccsd - eold, normt)
        if abs(eccsd - eold) < tol and normt < tolnormt:
            conv = True
            break
        if math.isnan(eccsd - eold):
            conv = False
            log.warning("%s FAILED THERE ARE NANS", mycc.__class__.__name__)
            break
        elif abs(eccsd - eold) > 1e5:
            conv = False
            log.warning("%s FAILED BECAUSE dE = %s", mycc.__class__.__name__, abs(eccsd - eold))
            break
    log.info("%s time %.3f s", mycc.__class__.__name__, time.perf_counter() - t0)
    return conv, eccsd, t1, t2


class FRICCSD:
    """FRI-CCSD driver (fricc/friccsd.py:462-524).

    ``mf`` is a PySCF SCF object (when PySCF is importable), a
    ``systems.SyntheticSystem`` or a ready ``systems.DeviceERIs``.
    Extra settings keys (not in the reference): ``"precision"`` ("fp64" |
    "fp32" | "bf16-ladder" | "mixed"), ``"seed"`` (srand seed for the uniform stream, default:
    none, as the reference), ``"vvvv_t1t1"`` ("reference" | "rccsd"),
    ``"graph"`` (True: ``kernel`` records the step once and replays it)."""

    def __init__(self, mf, fri_settings=default_fri_settings, frozen=None, mo_coeff=None, mo_occ=None,
                 uniforms=None):
        self.mf = mf
        settings = copy.deepcopy(fri_settings)
        for k, val in default_fri_settings.items():
            if k not in settings:
                log.debug("FRI: Setting %s to %s", k, val)
                settings[k] = val
        if "unbiased" not in settings:
            settings["unbiased"] = bool(settings.get("unbias", False))
        settings.setdefault("precision", "fp64")
        settings.setdefault("vvvv_t1t1", "reference")
        self._fri_settings = settings
        for c in settings["compressed_contractions"]:
            if c not in ALLOWED_CONTRACTIONS:
                raise ValueError(f"Contraction ({c}) is not supported!")
        if settings["compression"] not in ("fri", "largest"):
            raise ValueError("compression must be 'fri' or 'largest'")
        if settings["compression"] == "fri" and settings["sampling_method"] not in ("systematic", "pivotal"):
            raise ValueError("sampling_method must be 'systematic' or 'pivotal'")
        if settings["precision"] not in PRECISION_MODES:
            raise ValueError(f"precision must be one of {sorted(PRECISION_MODES)}")
        self.frozen = frozen
        self.max_cycle = 50
        self.conv_tol = 1e-7
        self.conv_tol_normt = 1e-5
        self.level_shift = 0.0
        self.verbose = 0
        self.energies = []
        self.e_corr = None
        self.converged = False
        self.t1 = self.t2 = None
        self._iter = 0
        self._slots = {}
        self._used_slots = set()
        self._u_dev = None
        self.arena = 0  # library scratch arena; solves that run concurrently need different ones
        self._eris = None
        self.uniforms = uniforms if uniforms is not None else GlibcUniforms(settings.get("seed"))
        self.vvvv_t1t1 = settings["vvvv_t1t1"]
        self.timer = None
        torch = _torch()
        prec = settings["precision"]
        self.sparse_precision, self.ladder_precision = PRECISION_MODES[prec]
        self.dense_dtype = torch.float64 if prec == "fp64" else torch.float32
        self.e_hf = getattr(mf, "e_hf", 0.0) if mf is not None else 0.0

    @property
    def fri_settings(self):
        return self._fri_settings

    # ---- integrals
    def ao2mo(self, mo_coeff=None):
        if self._eris is not None:
            return self._eris
        from .systems import DeviceERIs, SyntheticSystem

        mf = self.mf
        if isinstance(mf, DeviceERIs):
            self._eris = mf
        elif isinstance(mf, SyntheticSystem):
            self._eris = DeviceERIs.from_factors(mf.L, mf.mo_energy, mf.nocc)
        else:
            self._eris = _eris_from_pyscf(mf)
        return self._eris

    @property
    def nocc(self):
        return self.ao2mo().nocc

    def get_init_guess(self, eris=None):
        eris = eris or self.ao2mo()
        self.emp2, t1, t2 = init_amps(eris)
        return t1, t2

    def energy(self, t1, t2, eris=None):
        return energy(self, t1, t2, eris or self.ao2mo())

    def update_amps(self, t1, t2, eris):
        return update_amps(self, t1, t2, eris)

    # ---- step hooks shared with GraphStep, replica batches and the server
    def _device_step(self, t1, t2, eris):
        """The device part of update_amps: (t1new, t2new, reductions), no host wait."""
        return _update_amps_device(self, t1, t2, eris)

    def _n_compressions(self):
        """FRI compressions (uniforms) per step: t2_sparse, and t2_sparse_alt
        when an O^3V^3 / O^4V^2 intermediate is compressed."""
        return 2 if _needs_alt(self) else 1

    def _iteration_dependent(self):
        """True when a step's work depends on the iteration count (unbiased averaging)."""
        return bool(self.fri_settings.get("unbiased")) and _needs_alt(self)

    def amplitudes_to_vector_norm(self, t1new, t2new, t1, t2):
        return amplitudes_to_vector_norm(t1new, t2new, t1, t2)

    def ccsd(self, t1=None, t2=None, eris=None):
        if eris is None:
            eris = self.ao2mo()
        self.converged, self.e_corr, self.t1, self.t2 = kernel(
            self, eris, t1, t2, max_cycle=self.max_cycle, tol=self.conv_tol, tolnormt=self.conv_tol_normt,
            verbose=self.verbose)
        return self.e_corr, self.t1, self.t2

    def kernel(self, t1=None, t2=None, eris=None):
        return self.ccsd(t1, t2, eris)

    @property
    def e_tot(self):
        return (self.e_hf or 0.0) + (self.e_corr or 0.0)

    # ---- internals
    def _udev(self):
        """Device buffer of the step's uniforms (read by the compressions)."""
        if self._u_dev is None:
            self._u_dev = _torch().zeros(8, dtype=_torch().float64, device="cuda")
        return self._u_dev

    def _slot(self, k, dims):
        s = self._slots.get(k)
        if s is None or tuple(s.dims) != tuple(dims):
            s = SparseTensor4d.empty(dims)
            self._slots[k] = s
        return s


def _eris_from_pyscf(mf):  # pragma: no cover - PySCF is not installed here
    try:
        from pyscf import ao2mo
        from pyscf.cc import ccsd
    except ImportError as e:
        raise ImportError("FRICCSD needs PySCF for a mean-field object; pass a SyntheticSystem instead") from e
    from .systems import DeviceERIs

    mycc = ccsd.CCSD(mf)
    er = mycc.ao2mo()
    o = mycc.nocc
    v = er.ovov.shape[1]
    vvvv = ao2mo.restore(1, np.asarray(er.vvvv), v)
    return DeviceERIs.from_blocks(er.fock, er.mo_energy, o, er.oooo, er.ovoo, er.ovov, er.oovv, er.ovvo,
                                  np.asarray(er.get_ovvv()), vvvv)
