"""Device-resident FRI-LCCSD: ``FRILCCSD``, its ``update_amps`` and energy.

Mirrors ``fricc/frilccsd.py`` of the reference: the closed-shell linearised
CCSD of Taube & Bartlett, J. Chem. Phys. 130, 144112 (2009), eqs. 13 and 15,
with T2 FRI-compressed for the 2323, 0101 and four O^3V^3 contractions
(fricc/frilccsd.py:27-230), optionally one independent compression per
compressed term (``"Independent Compressions"``, :233-457, drawn in the order
2323, 0101, 1302, 1202, 1303, 1203) and optionally LCCD (``"LCCD"``: T1 held
at zero).  The energy is ``lin_cc_energy`` (:460-476).  The iteration loop is
FRICCSD's (the reference class inherits it, :479-519).

The intermediates are the bare integral blocks, so every compressed term
reads a static operand prepared once per system: the ladder (ac|bd) (the same
operand FRI-CCSD uses), (ki|lj), (kc|ai) and (ki|ac).  The device step has
FRI-CCSD's shape -- symmetrised terms into one buffer S, direct terms into
``acc``, one fused finalize that adds S + S^T, divides by the denominators and
reduces the energy and norm -- and records as one graph the same way.
Precision modes as in FRICCSD.
"""
import copy
import logging

from . import _capi
from .py_rccsd import SparseTensor4d
from .dense import einsum as _de
from .friccsd import (FRICCSD, _check_m_keep, _contract_into, _prepare, _set_uniforms, _torch, _vp,
                      default_fri_settings as _ccsd_defaults, _finish_step, _draw_uniforms, _O3V3, _PREC_ID)

log = logging.getLogger("fricc")

default_fri_settings = {
    "m_keep": 1000,
    "compression": "fri",
    "sampling_method": "systematic",
    "verbose": False,
    "compressed_contractions": ["O^2V^4", "O^3V^3", "O^4V^2"],
    "LCCD": False,
    "Independent Compressions": False,
}

# compressed terms in the order the independent variant draws their tensors
# (fricc/frilccsd.py:333-342 builds them as a list, used as [0] .. [5])
TERM_ORDER = ("2323", "0101", "1302", "1202", "1303", "1203")


def update_amps(cc, t1, t2, eris):
    """One FRI-LCCSD Jacobi step on the GPU (fricc/frilccsd.py:27-230 and
    :233-457): uniforms drawn, device step, one read of the energy, norm and
    compression statuses."""
    _check_m_keep(cc, t2)
    _set_uniforms(cc, _draw_uniforms(cc))
    t1new, t2new, red = cc._device_step(t1, t2, eris)
    _finish_step(cc, t1new, t2new, red)
    return t1new, t2new


def _update_amps_lccsd_device(cc, t1, t2, eris):
    """Device part of the step; returns (t1new, t2new, {E2, |dt2|^2, E1, |dt1|^2})
    without waiting for the GPU."""
    torch = _torch()
    st = cc.fri_settings
    m_keep = int(st["m_keep"])
    ccs = st["compressed_contractions"]
    _capi.check(_capi.lib().fricc_workspace_select(int(cc.arena)))
    sp = cc.sparse_precision
    o, v = t1.shape
    ov = o * v
    f64 = torch.float64
    cdt = cc.dense_dtype
    P = eris.prepared(cdt)
    P64 = eris.prepared(f64)
    fock = eris.fock
    mo_e_o = eris.mo_energy[:o]
    mo_e_v = eris.mo_energy[o:] + cc.level_shift
    fov = fock[:o, o:]
    t1_in = t1.contiguous()
    t2 = t2.contiguous()
    lccd = bool(st["LCCD"])
    t1 = torch.zeros_like(t1_in) if lccd else t1_in
    t2c = t2.to(cdt) if cdt != f64 else t2

    def E(spec, A, B, out=None, alpha=1.0, beta=0.0):  # MFMA contraction in the dense precision
        return _de(spec, A, B, out=out, alpha=alpha, beta=beta, compute=cdt)

    def E64(spec, A, B, out=None, alpha=1.0, beta=0.0):
        return _de(spec, A, B, out=out, alpha=alpha, beta=beta, compute=f64)

    # ---- compression: one tensor, or one per compressed term
    comp = st["compression"]
    ud = cc._udev()
    if st["Independent Compressions"] and comp == "fri":
        # one sample of the same T2 per term, uniforms in term order: one
        # selection shared by all samples
        slots = [cc._slot(k, t2.shape) for k in range(len(TERM_ORDER))]
        SparseTensor4d.compress_multi_async(slots, t2, m_keep, comp, ud[:len(TERM_ORDER)],
                                            sampling_method=st["sampling_method"])
        ts = dict(zip(TERM_ORDER, slots))
        cc._used_slots = set(range(len(TERM_ORDER)))
    else:
        one = cc._slot(0, t2.shape).recompress_async(t2, m_keep, comp, ud[0:1],
                                                     sampling_method=st["sampling_method"])
        ts = dict.fromkeys(TERM_ORDER, one)
        cc._used_slots = {0}

    Foo = fock[:o, :o].clone()
    Fvv = fock[o:, o:].clone()
    Foo.diagonal().sub_(mo_e_o)
    Fvv.diagonal().sub_(mo_e_v)

    # ---- T1 equation (fricc/frilccsd.py:126-154)
    if lccd:
        t1new = torch.zeros_like(t1)
    else:
        t1new = fov.clone()
        E64("ic,ac->ia", t1, Fvv, out=t1new, beta=1.0)
        E64("ki,ka->ia", Foo, t1, out=t1new, alpha=-1.0, beta=1.0)
        # 2 sum_kc (kc|ai) t1[k,c] - sum_kc (ki|ac) t1[k,c]
        E("kcai,kc->ia", P["ovvo"], t1, out=t1new, alpha=2.0, beta=1.0)
        E("kiac,kc->ia", P["oovv"], t1, out=t1new, alpha=-1.0, beta=1.0)
        # 2 kdac,ikcd - kcad,ikcd  ==  t2[i,(kcd)] Z[(kcd),a]
        E("ikcd,kcda->ia", t2c, P["Z"], out=t1new, beta=1.0)
        E("klic,klac->ia", P["Aoo"], t2c, out=t1new, beta=1.0)
        E("kc,kica->ia", fov, t2c, out=t1new, alpha=2.0, beta=1.0)
        E("kc,ikca->ia", fov, t2c, out=t1new, alpha=-1.0, beta=1.0)

    # ---- symmetrised T2 terms into S (fricc/frilccsd.py:168-224 "tmp + tmp^T")
    # Term 8: sum_c (ia|bc) t1[j,c];  Term 7: - sum_k (ia|jk) t1[k,b]
    S = E("jc,icab->ijab", t1, P["Qv"])
    E("ijak,kb->ijab", P["ovoo_ijak"], t1, out=S, alpha=-1.0, beta=1.0)
    # Term 3: Fvv[a,c] t2[i,j,c,b];  Term 2: - Foo[k,i] t2[k,j,a,b]
    E("ac,ijcb->ijab", Fvv, t2c, out=S, beta=1.0)
    E("ki,kjab->ijab", Foo, t2c, out=S, alpha=-1.0, beta=1.0)
    # Term 6: O^3V^3 with Wvoov[a,k,i,c] = (kc|ai), Wvovo[a,k,c,i] = (ki|ac)
    if "O^3V^3" in ccs:
        Av = eris.sparse_operand("ovvo_p", sp)
        Avo = eris.sparse_operand("oovv_p", sp)
        terms = list(zip(_O3V3, (Av, Avo, Av, Avo), ((v, o, o, v), (v, o, v, o), (v, o, o, v), (v, o, v, o))))
        if ts["1302"] is ts["1202"]:
            # one compressed tensor: 1302 and 1202 read the same view with
            # operands in the same layout, so they are one contraction with
            # the bare 2 Wvoov - Wvovo
            _prepare(ts["1302"], "1302", (v, o, o, v), (o, o, v, v), sp)
            _contract_into("akic,kjcb->ijab", 1.0, eris.sparse_operand("A12b", sp), (v, o, o, v), ts["1302"], S, sp)
            terms = terms[2:]
        for (term, spec, alpha), A, ws in terms:
            _prepare(ts[term], term, ws, (o, o, v, v), sp)
            _contract_into(spec, alpha, A, ws, ts[term], S, sp)
    else:
        Wvoov = P["ovvo_p"].permute(3, 0, 2, 1)   # [a,k,i,c]
        Wvovo = P["oovv_p"].permute(3, 0, 1, 2)   # [a,k,c,i]
        E("akic,kjcb->ijab", Wvoov, t2c, out=S, alpha=2.0, beta=1.0)
        E("akci,kjcb->ijab", Wvovo, t2c, out=S, alpha=-1.0, beta=1.0)
        E("akic,kjbc->ijab", Wvoov, t2c, out=S, alpha=-1.0, beta=1.0)
        E("bkci,kjac->ijab", Wvovo, t2c, out=S, alpha=-1.0, beta=1.0)

    # ---- direct T2 terms into acc (FP64): Term 4 (ladder) and Term 5 (0101)
    acc = torch.zeros((o, o, v, v), dtype=f64, device=t2.device)
    if "O^2V^4" in ccs:
        spl = cc.ladder_precision
        _prepare(ts["2323"], "2323", (v, v, v, v), (o, o, v, v), spl)
        _contract_into("abcd,ijcd->ijab", 1.0, eris.sparse_operand("ladder", spl), (v, v, v, v), ts["2323"], acc,
                       spl)
    else:
        if eris.ladder is None:
            raise RuntimeError("the uncompressed O^2V^4 term needs the FP64 ladder")
        E64("ijcd,cdab->ijab", t2, P64["ladder"], out=acc, beta=1.0)
    if "O^4V^2" in ccs:
        _contract_into("klij,klab->ijab", 1.0, eris.sparse_operand("Woooo_b", sp), (o, o, o, o), ts["0101"], acc,
                       sp)
    else:
        E64("klij,klab->ijab", P64["Woooo_b"], t2, out=acc, beta=1.0)

    # ---- denominators, linearised energy and norm: FP64, fused
    eia = (mo_e_o[:, None] - mo_e_v).contiguous()
    t1new = (t1new / eia).contiguous()
    red = torch.empty(4, dtype=f64, device=t2.device)
    zero = torch.zeros_like(t1new)  # the linearised energy has no t1*t1 part
    _capi.check(_capi.lib().fricc_t2_finalize(
        o, v, _vp(P64["ovovT"]), _vp(acc), _vp(S), _PREC_ID["fp32" if S.dtype == torch.float32 else "fp64"],
        _vp(t1), None, _vp(eia), _vp(t2), None, _vp(zero), None, _vp(red), _capi.stream_ptr()))
    red[2].copy_(2 * (fov * t1new).sum())
    red[3].copy_(((t1new - t1_in) ** 2).sum())
    return t1new, acc, red


def energy(cc, t1, t2, eris):
    """lin_cc_energy (fricc/frilccsd.py:460-476), FP64:
    2 sum f_ia t_ia + sum t2_ijab [2 (ia|jb) - (ib|ja)]."""
    o = t1.shape[0]
    P = eris.prepared(_torch().float64)
    e = 2 * (eris.fock[:o, o:] * t1).sum() + (t2 * P["Lovov"]).sum()
    return float(e.item())


class FRILCCSD(FRICCSD):
    """FRI-LCCSD driver (fricc/frilccsd.py:479-519).  Settings: the reference's
    keys ("LCCD", "Independent Compressions", ...) plus this package's
    "precision", "seed" and "graph"."""

    def __init__(self, mf, fri_settings=default_fri_settings, frozen=None, mo_coeff=None, mo_occ=None,
                 uniforms=None):
        settings = copy.deepcopy(fri_settings)
        for k, val in default_fri_settings.items():
            settings.setdefault(k, val)
        # keys FRICCSD reads that the linearised update does not use
        settings.setdefault("burn_in", _ccsd_defaults["burn_in"])
        settings.setdefault("unbiased", False)
        if settings["unbiased"]:
            raise ValueError("FRILCCSD has no unbiased averaging")
        super().__init__(mf, settings, frozen=frozen, mo_coeff=mo_coeff, mo_occ=mo_occ, uniforms=uniforms)

    def update_amps(self, t1, t2, eris):
        return update_amps(self, t1, t2, eris)

    def energy(self, t1, t2, eris=None):
        return energy(self, t1, t2, eris or self.ao2mo())

    def _device_step(self, t1, t2, eris):
        return _update_amps_lccsd_device(self, t1, t2, eris)

    def _n_compressions(self):
        return len(TERM_ORDER) if self.fri_settings["Independent Compressions"] else 1

    def _iteration_dependent(self):
        return False


__all__ = ["FRILCCSD", "update_amps", "energy", "default_fri_settings", "TERM_ORDER"]
