"""NumPy FP64 restatement of FRI-LCCSD's amplitude update -- TEST INFRASTRUCTURE.

Follows ``fricc/frilccsd.py`` of the reference:
  * ``update_amps``                          fricc/frilccsd.py:27-230
  * ``update_amps_independent_compressions`` fricc/frilccsd.py:233-457
    (one compression per compressed term, drawn in the order 2323, 0101,
    1302, 1202, 1303, 1203)
  * ``lin_cc_energy``                        fricc/frilccsd.py:460-476
  * settings defaults                        fricc/frilccsd.py:15-23
and the iteration loop the FRILCCSD class inherits from FRICCSD
(fricc/friccsd.py:379-456, with FRILCCSD.energy = lin_cc_energy,
fricc/frilccsd.py:510-519).

The terms are those of the closed-shell linearised CCSD of Taube & Bartlett,
J. Chem. Phys. 130, 144112 (2009), eqs. 13 and 15, with the bare integrals in
place of the dressed W intermediates (the W's are the integral blocks alone,
fricc/frilccsd.py:157-160).  ``Independent Compressions`` and ``LCCD`` follow
the reference's settings keys.

Parity: pinned by ``oracle/spinorb.update_linear`` (an independent
spin-orbital linearised CCSD) through tests/test_oracle_lccsd.py, and by the
"largest, m = N" identity.  The NWChem LCCD/LCCSD energies of
fricc/tests/test_lccsd.py:65-140 need PySCF integrals for H2O, which cannot be
generated here: those stay unchecked.
"""
import math

import numpy as np

from .rccsd import SparseT, amplitudes_to_vector, contract, init_amps

default_settings_ref = {
    "m_keep": 1000,
    "compression": "fri",
    "sampling_method": "systematic",
    "verbose": False,
    "compressed_contractions": ["O^2V^4", "O^3V^3", "O^4V^2"],
    "LCCD": False,
    "Independent Compressions": False,
}


def default_settings(**kw):
    from . import liboracle

    s = dict(default_settings_ref)
    s.update({"level_shift": 0.0, "mode": liboracle.CANONICAL})
    s.update(kw)
    return s


# compressed terms in the order the independent variant draws their tensors
TERM_ORDER = ("2323", "0101", "1302", "1202", "1303", "1203")


def update_amps(t1, t2, eris, settings, uniforms, return_sparse=False):
    """fricc/frilccsd.py:27-230 (and :233-457 when settings["Independent
    Compressions"]).  ``uniforms`` yields one float per FRI compression."""
    o, v = t1.shape
    m = int(settings["m_keep"])
    if m > t2.size or m == 0:
        raise ValueError("Bad m_keep value! The following condition wasn't met 0 < m_keep <= t2.size")
    cc = settings["compressed_contractions"]
    fock = eris.fock
    mo_e_o = eris.mo_energy[:o]
    mo_e_v = eris.mo_energy[o:] + settings["level_shift"]
    fov = fock[:o, o:].copy()
    t2 = np.ascontiguousarray(t2)
    comp = settings["compression"]
    mode = settings["mode"]

    def draw():
        return next(uniforms) if comp == "fri" else 0.0

    if settings["Independent Compressions"]:
        sparse = [SparseT(t2, m, comp, draw(), mode) for _ in TERM_ORDER]
    else:
        one = SparseT(t2, m, comp, draw(), mode)
        sparse = [one] * len(TERM_ORDER)
    ts = dict(zip(TERM_ORDER, sparse))

    Foo = fock[:o, :o].copy()
    Fvv = fock[o:, o:].copy()
    Fov = fock[:o, o:].copy()
    Foo[np.diag_indices(o)] -= mo_e_o
    Fvv[np.diag_indices(v)] -= mo_e_v

    ovoo = eris.ovoo
    ovvv = eris.ovvv
    if settings["LCCD"]:
        t1 = np.zeros_like(t1)
        t1new = np.zeros_like(t1)
    else:
        t1new = np.zeros_like(t1)
        t1new += fov
        t1new += np.einsum("ac,ic->ia", Fvv, t1)
        t1new += -np.einsum("ki,ka->ia", Foo, t1)
        t1new += 2 * np.einsum("kcai,kc->ia", eris.ovvo, t1)
        t1new += -np.einsum("kiac,kc->ia", eris.oovv, t1)
        t1new += 2 * np.einsum("kdac,ikcd->ia", ovvv, t2)
        t1new += -np.einsum("kcad,ikcd->ia", ovvv, t2)
        t1new += -2 * np.einsum("lcki,klac->ia", ovoo, t2)
        t1new += np.einsum("kcli,klac->ia", ovoo, t2)
        t1new += 2 * np.einsum("kc,kica->ia", Fov, t2)
        t1new += -np.einsum("kc,ikca->ia", Fov, t2)

    t2new = eris.ovov.transpose(0, 2, 1, 3).copy()
    tmp = np.einsum("ki,kjab->ijab", Foo, t2)
    t2new -= tmp + tmp.transpose(1, 0, 3, 2)
    tmp = np.einsum("ac,ijcb->ijab", Fvv, t2)
    t2new += tmp + tmp.transpose(1, 0, 3, 2)

    Woooo = eris.oooo.transpose(0, 2, 1, 3)
    Wvoov = eris.ovvo.transpose(2, 0, 3, 1)
    Wvovo = eris.oovv.transpose(2, 0, 3, 1)
    Wvvvv = eris.vvvv.transpose(0, 2, 1, 3)

    if "O^2V^4" in cc:
        contract(Wvvvv, ts["2323"], t2new, "2323")
    else:
        t2new += np.einsum("abcd,ijcd->ijab", Wvvvv, t2)
    if "O^4V^2" in cc:
        contract(Woooo, ts["0101"], t2new, "0101")
    else:
        t2new += np.einsum("klij,klab->ijab", Woooo, t2)
    if "O^3V^3" in cc:
        tmp = np.zeros_like(t2new)
        contract(Wvoov, ts["1302"], tmp, "1302")
        contract(Wvovo, ts["1202"], tmp, "1202")
        t2new += tmp + tmp.transpose(1, 0, 3, 2)
        tmp = np.zeros_like(t2new)
        contract(Wvoov, ts["1303"], tmp, "1303")
        t2new -= tmp + tmp.transpose(1, 0, 3, 2)
        tmp = np.zeros_like(t2new)
        contract(Wvovo, ts["1203"], tmp, "1203")
        t2new -= tmp + tmp.transpose(1, 0, 3, 2)
    else:
        tmp = 2 * np.einsum("akic,kjcb->ijab", Wvoov, t2)
        tmp -= np.einsum("akci,kjcb->ijab", Wvovo, t2)
        t2new += tmp + tmp.transpose(1, 0, 3, 2)
        tmp = np.einsum("akic,kjbc->ijab", Wvoov, t2)
        t2new -= tmp + tmp.transpose(1, 0, 3, 2)
        tmp = np.einsum("bkci,kjac->ijab", Wvovo, t2)
        t2new -= tmp + tmp.transpose(1, 0, 3, 2)

    tmp = np.einsum("akij,kb->ijab", ovoo.transpose(1, 3, 0, 2), t1)
    t2new -= tmp + tmp.transpose(1, 0, 3, 2)
    tmp = np.einsum("iabc,jc->ijab", ovvv, t1)
    t2new += tmp + tmp.transpose(1, 0, 3, 2)

    eia = mo_e_o[:, None] - mo_e_v
    eijab = eia[:, None, :, None] + eia[None, :, None, :]
    t1new /= eia
    t2new /= eijab
    if return_sparse:
        return t1new, t2new, sparse if settings["Independent Compressions"] else [sparse[0]]
    return t1new, t2new


def energy(t1, t2, eris):
    """lin_cc_energy, fricc/frilccsd.py:460-476: 2 f_ia t_ia + t2 (2 (ia|jb) - (ib|ja))."""
    o = t1.shape[0]
    e = 2 * np.einsum("ia,ia", eris.fock[:o, o:], t1)
    e += 2 * np.einsum("ijab,iajb", t2, eris.ovov)
    e += -np.einsum("ijab,ibja", t2, eris.ovov)
    return float(e)


def kernel(eris, settings, uniforms, max_cycle=50, tol=1e-8, tolnormt=1e-6, t1=None, t2=None):
    """fricc/friccsd.py:379-456 as FRILCCSD runs it (its update and energy)."""
    if t1 is None and t2 is None:
        _, t1, t2 = init_amps(eris)
    eold = 0.0
    ecc = energy(t1, t2, eris)
    energies = []
    conv = False
    for _ in range(max_cycle):
        t1new, t2new = update_amps(t1, t2, eris, settings, uniforms)
        normt = np.linalg.norm(amplitudes_to_vector(t1new, t2new) - amplitudes_to_vector(t1, t2))
        t1, t2 = t1new, t2new
        eold, ecc = ecc, energy(t1, t2, eris)
        energies.append(ecc)
        if abs(ecc - eold) < tol and normt < tolnormt:
            conv = True
            break
        if math.isnan(ecc - eold) or abs(ecc - eold) > 1e5:
            conv = False
            break
    return conv, ecc, t1, t2, energies
