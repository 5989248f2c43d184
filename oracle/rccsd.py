"""NumPy FP64 restatement of FRI-CCSD's amplitude update -- TEST INFRASTRUCTURE.

Follows ``fricc/friccsd.py`` of the reference line by line:
  * ``update_amps``        fricc/friccsd.py:29-301
  * ``sparse_cc_Woooo``    fricc/friccsd.py:309-322
  * ``cc_Wvvvv``           fricc/friccsd.py:325-331
  * ``sparse_cc_Wvoov``    fricc/friccsd.py:334-360
  * ``sparse_cc_Wvovo``    fricc/friccsd.py:363-376
  * ``kernel``             fricc/friccsd.py:379-456
and restates the PySCF formulas it calls (PySCF is not pinned by the
reference, ``setup.py:17``; these are the published ``pyscf.cc.rintermediates``
/ ``pyscf.cc.ccsd`` expressions):
  * ``cc_Foo/cc_Fvv/cc_Fov/Loo/Lvv`` (imd.*, called at friccsd.py:144-146,189-190)
  * ``energy``  (CCSD.energy, friccsd.py:402,427)
  * ``init_amps`` (CCSD.get_init_guess, friccsd.py:396)
  * ``amplitudes_to_vector`` (packed lower triangle of t2 as (ia,jb), friccsd.py:418-419)

Parity: pinned by ``oracle/spinorb.py`` (an independent spin-orbital CCSD) via
tests/test_oracle_rccsd.py, and by the "largest, m = N" identity.

The compressed contractions use the CPU oracle compression (canonical mode by
default) and evaluate each ``contract_DTSpT`` as ``np.einsum`` on the
decompressed tensor -- the identity the reference's own tests assert
(fricc/tests/test_sparse_tensors.py:48-269).
"""
import ctypes
import math

import numpy as np

from . import liboracle

ALLOWED_CONTRACTIONS = ("O^2V^4", "O^3V^3", "O^4V^2")


def einsum(spec, *ops):
    """np.einsum with its contraction-order search (pairwise BLAS
    contractions): the same sums as PySCF's lib.einsum, which the reference
    calls, in a BLAS order; FP64 throughout."""
    return np.einsum(spec, *ops, optimize=True)


class ERIs:
    """Chemists'-notation MO integral blocks, as ``ccsd._ChemistsERIs`` holds them
    (fricc/friccsd.py:80-86,164-187,310-330): oooo=(ij|kl), ovoo=(ia|jk),
    ovov=(ia|jb), oovv=(ij|ab), ovvo=(ia|bj), ovvv=(ia|bc), vvvv=(ab|cd)."""

    def __init__(self, eri_full, fock, nocc):
        o = nocc
        self.fock = fock
        self.mo_energy = fock.diagonal().copy()
        self.nocc = o
        e = eri_full
        self.oooo = e[:o, :o, :o, :o].copy()
        self.ovoo = e[:o, o:, :o, :o].copy()
        self.ovov = e[:o, o:, :o, o:].copy()
        self.oovv = e[:o, :o, o:, o:].copy()
        self.ovvo = e[:o, o:, o:, :o].copy()
        self.ovvv = e[:o, o:, o:, o:].copy()
        self.vvvv = e[o:, o:, o:, o:].copy()

    def get_ovvv(self):
        return self.ovvv

    @classmethod
    def from_factors(cls, L, mo_energy, nocc):
        """The blocks straight from the factors (no full nmo^4 tensor): each
        integral is the same fma chain as in the full build."""
        e = cls.__new__(cls)
        o, nmo = nocc, L.shape[1]
        O, V = (0, o), (o, nmo)
        e.fock = np.diag(mo_energy)
        e.mo_energy = np.array(mo_energy, dtype=np.float64)
        e.nocc = o
        blk = lambda a, b, c, d: liboracle.eri_block(L, a, b, c, d)  # noqa: E731
        e.oooo = blk(O, O, O, O)
        e.ovoo = blk(O, V, O, O)
        e.ovov = blk(O, V, O, V)
        e.oovv = blk(O, O, V, V)
        e.ovvo = blk(O, V, V, O)
        e.ovvv = blk(O, V, V, V)
        e.vvvv = blk(V, V, V, V)
        return e


def synthetic_L(nocc, nvir, naux, seed):
    """Synthetic density-fitting factors (the c2/c3 recipe): occupied energies
    U[-1.5,-0.3], virtual U[0.1,3.0], each sorted; L[Q,p,q] = N(0,1)
    exp(-|p-q|/10)/sqrt(naux), symmetrised in (p,q).  Returns (L, mo_energy)."""
    nmo = nocc + nvir
    rng = np.random.default_rng(seed)
    eo = np.sort(rng.uniform(-1.5, -0.3, nocc))
    ev = np.sort(rng.uniform(0.1, 3.0, nvir))
    p = np.arange(nmo)
    decay = np.exp(-np.abs(p[:, None] - p[None, :]) / 10.0) / np.sqrt(naux)
    L = rng.standard_normal((naux, nmo, nmo)) * decay
    L = 0.5 * (L + L.transpose(0, 2, 1))
    return np.ascontiguousarray(L), np.concatenate([eo, ev])


def synthetic_eris_blocks(nocc, nvir, naux, seed):
    """synthetic_eris without the full ERI tensor (for c2/c3-sized systems)."""
    L, e = synthetic_L(nocc, nvir, naux, seed)
    return ERIs.from_factors(L, e, nocc), L


def synthetic_eris(nocc, nvir, naux, seed):
    L, e = synthetic_L(nocc, nvir, naux, seed)
    nmo = nocc + nvir
    eri = liboracle.eri_block(L, (0, nmo), (0, nmo), (0, nmo), (0, nmo))
    return ERIs(eri, np.diag(e), nocc), L


# --- PySCF intermediates ------------------------------------------------------

def cc_Foo(t1, t2, eris):
    o = t1.shape[0]
    ovov = eris.ovov
    Fki = 2 * einsum("kcld,ilcd->ki", ovov, t2)
    Fki -= einsum("kdlc,ilcd->ki", ovov, t2)
    Fki += 2 * einsum("kcld,ic,ld->ki", ovov, t1, t1)
    Fki -= einsum("kdlc,ic,ld->ki", ovov, t1, t1)
    Fki += eris.fock[:o, :o]
    return Fki


def cc_Fvv(t1, t2, eris):
    o = t1.shape[0]
    ovov = eris.ovov
    Fac = -2 * einsum("kcld,klad->ac", ovov, t2)
    Fac += einsum("kdlc,klad->ac", ovov, t2)
    Fac -= 2 * einsum("kcld,ka,ld->ac", ovov, t1, t1)
    Fac += einsum("kdlc,ka,ld->ac", ovov, t1, t1)
    Fac += eris.fock[o:, o:]
    return Fac


def cc_Fov(t1, t2, eris):
    o = t1.shape[0]
    ovov = eris.ovov
    Fkc = 2 * einsum("kcld,ld->kc", ovov, t1)
    Fkc -= einsum("kdlc,ld->kc", ovov, t1)
    Fkc += eris.fock[:o, o:]
    return Fkc


def Loo(t1, t2, eris):
    o = t1.shape[0]
    fov = eris.fock[:o, o:]
    Lki = cc_Foo(t1, t2, eris) + einsum("kc,ic->ki", fov, t1)
    Lki += 2 * einsum("lcki,lc->ki", eris.ovoo, t1)
    Lki -= einsum("kcli,lc->ki", eris.ovoo, t1)
    return Lki


def Lvv(t1, t2, eris):
    o = t1.shape[0]
    fov = eris.fock[:o, o:]
    Lac = cc_Fvv(t1, t2, eris) - einsum("kc,ka->ac", fov, t1)
    Lac += 2 * einsum("kdac,kd->ac", eris.ovvv, t1)
    Lac -= einsum("kcad,kd->ac", eris.ovvv, t1)
    return Lac


def cc_Woooo(t1, t2, eris):
    Wklij = einsum("lcki,jc->klij", eris.ovoo, t1)
    Wklij += einsum("kclj,ic->klij", eris.ovoo, t1)
    Wklij += einsum("kcld,ijcd->klij", eris.ovov, t2)
    Wklij += einsum("kcld,ic,jd->klij", eris.ovov, t1, t1)
    Wklij += eris.oooo.transpose(0, 2, 1, 3)
    return Wklij


def cc_Wvoov(t1, t2, eris):
    Wakic = einsum("kcad,id->akic", eris.ovvv, t1)
    Wakic -= einsum("kcli,la->akic", eris.ovoo, t1)
    Wakic += eris.ovvo.transpose(2, 0, 3, 1)
    Wakic -= 0.5 * einsum("ldkc,ilda->akic", eris.ovov, t2)
    Wakic -= 0.5 * einsum("lckd,ilad->akic", eris.ovov, t2)
    Wakic -= einsum("ldkc,id,la->akic", eris.ovov, t1, t1)
    Wakic += einsum("ldkc,ilad->akic", eris.ovov, t2)
    return Wakic


def cc_Wvovo(t1, t2, eris):
    Wakci = einsum("kdac,id->akci", eris.ovvv, t1)
    Wakci -= einsum("lcki,la->akci", eris.ovoo, t1)
    Wakci += eris.oovv.transpose(2, 0, 3, 1)
    Wakci -= 0.5 * einsum("lckd,ilda->akci", eris.ovov, t2)
    Wakci -= einsum("lckd,id,la->akci", eris.ovov, t1, t1)
    return Wakci


def cc_Wvvvv(t1, t2, eris, chunk=16):
    """fricc/friccsd.py:325-331: W = -sum_k (kd|ac) t_kb - sum_k (kc|bd) t_ka
    + (ac|bd), in that order per element; built over blocks of a so that no
    second v^4 temporary exists."""
    v = t1.shape[1]
    ovvv = eris.ovvv
    Wabcd = np.empty((v, v, v, v))
    for a0 in range(0, v, chunk):
        a1 = min(v, a0 + chunk)
        w = einsum("kdac,kb->abcd", ovvv[:, :, a0:a1, :], -t1)
        w -= einsum("kcbd,ka->abcd", ovvv, t1[:, a0:a1])
        w += eris.vvvv[a0:a1].transpose(0, 2, 1, 3)
        Wabcd[a0:a1] = w
    return Wabcd


def energy(t1, t2, eris):
    o = t1.shape[0]
    e = 2 * einsum("ia,ia", eris.fock[:o, o:], t1)
    tau = einsum("ia,jb->ijab", t1, t1) + t2
    e += 2 * einsum("ijab,iajb", tau, eris.ovov)
    e -= einsum("ijab,ibja", tau, eris.ovov)
    return float(e)


def init_amps(eris):
    o = eris.nocc
    mo_e = eris.mo_energy
    eia = mo_e[:o, None] - mo_e[None, o:]
    t1 = eris.fock[:o, o:] / eia
    t2 = eris.ovov.transpose(0, 2, 1, 3) / (eia[:, None, :, None] + eia[None, :, None, :])
    emp2 = 2 * einsum("ijab,iajb", t2, eris.ovov) - einsum("jiab,iajb", t2, eris.ovov)
    return float(emp2), t1, np.ascontiguousarray(t2)


def amplitudes_to_vector(t1, t2):
    o, v = t1.shape
    nov = o * v
    m = t2.transpose(0, 2, 1, 3).reshape(nov, nov)
    return np.concatenate([t1.ravel(), m[np.tril_indices(nov)]])


# --- sparse pieces ------------------------------------------------------------

class SparseT:
    """Decompressed view of a compressed tensor (SparseTensor4d semantics)."""

    def __init__(self, t2, m, compression, u, mode, sampling="systematic"):
        self.dims = t2.shape
        if compression != "fri":
            method = liboracle.LARGEST
        else:
            method = liboracle.FRI_PIVOTAL if sampling == "pivotal" else liboracle.FRI_SYSTEMATIC
        self.idx, self.val, self.info = liboracle.compress(t2.ravel(), m, method, u, mode)
        if self.info["status"] != 0:
            raise RuntimeError(f"oracle compression failed: {self.info}")
        dense = np.zeros(t2.size)
        np.add.at(dense, self.idx, self.val)
        self.dense = dense.reshape(t2.shape)


def contract(W, T, out, term):
    """contract_DTSpT as the einsum identity of fricc/tests/test_sparse_tensors.py."""
    t = T.dense
    if term == "0101":
        out += einsum("klij,klab->ijab", W, t)
    elif term == "2323":
        out += einsum("abcd,ijcd->ijab", W, t)
    elif term == "1302":
        out += 2 * einsum("akic,kjcb->ijab", W, t)
    elif term == "1202":
        out -= einsum("akci,kjcb->ijab", W, t)
    elif term == "1303":
        out += einsum("akic,kjbc->ijab", W, t)
    elif term == "1203":
        out += einsum("bkci,kjac->ijab", W, t)
    elif term == "1323":
        out += einsum("kcld,ijcd->klij", W, t)
    elif term == "0112":
        out -= 0.5 * einsum("ldkc,ilda->akic", W, t)
    elif term == "0313":
        out -= 0.5 * einsum("lckd,ilad->akic", W, t)
    elif term == "0113":
        out += einsum("ldkc,ilad->akic", W, t)
    elif term == "0312":
        out -= 0.5 * einsum("lckd,ilda->akci", W, t)
    else:
        raise ValueError(term)
    return out


def sparse_cc_Woooo(t1, ts, eris):
    Wklij = einsum("lcki,jc->klij", eris.ovoo, t1)
    Wklij += einsum("kclj,ic->klij", eris.ovoo, t1)
    contract(eris.ovov, ts, Wklij, "1323")
    Wklij += einsum("kcld,ic,jd->klij", eris.ovov, t1, t1)
    Wklij += eris.oooo.transpose(0, 2, 1, 3)
    return Wklij


def sparse_cc_Wvoov(t1, ts, eris):
    Wakic = einsum("kcad,id->akic", eris.ovvv, t1)
    Wakic -= einsum("kcli,la->akic", eris.ovoo, t1)
    Wakic += eris.ovvo.transpose(2, 0, 3, 1)
    contract(eris.ovov, ts, Wakic, "0112")
    contract(eris.ovov, ts, Wakic, "0313")
    contract(eris.ovov, ts, Wakic, "0113")
    Wakic -= einsum("ldkc,id,la->akic", eris.ovov, t1, t1)
    return Wakic


def sparse_cc_Wvovo(t1, ts, eris):
    Wakci = einsum("kdac,id->akci", eris.ovvv, t1)
    Wakci -= einsum("lcki,la->akci", eris.ovoo, t1)
    Wakci += eris.oovv.transpose(2, 0, 3, 1)
    contract(eris.ovov, ts, Wakci, "0312")
    Wakci -= einsum("lckd,id,la->akci", eris.ovov, t1, t1)
    return Wakci


def default_settings(**kw):
    s = {
        "m_keep": 1000,
        "compression": "fri",
        "sampling_method": "systematic",
        "compressed_contractions": list(ALLOWED_CONTRACTIONS),
        "unbiased": False,
        "burn_in": 100,
        "vvvv_t1t1": "reference",
        "level_shift": 0.0,
        "mode": liboracle.CANONICAL,
    }
    s.update(kw)
    return s


def update_amps(t1, t2, eris, settings, uniforms, return_sparse=False, state=None):
    """fricc/friccsd.py:29-301.  ``uniforms`` is an iterator of floats in [0,1],
    one consumed per FRI compression (lin_space_add_const, src/fri_utils.cpp:47),
    in the order t2_sparse, t2_sparse_alt.

    ``state`` (a dict, kept across calls by the caller) carries what the
    reference keeps on the solver object for the unbiased estimator: "iter"
    (``cc._iter``, the iteration index) and "t2_avg" (``cc._t2_avg``)."""
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
    smeth = settings.get("sampling_method", "systematic")

    def draw():
        return next(uniforms) if comp == "fri" else 0.0

    ts = SparseT(t2, m, comp, draw(), mode, smeth)
    ts_alt = None
    if "O^3V^3" in cc or "O^4V^2" in cc:
        it_now = 0 if state is None else state.get("iter", 0)
        if not settings["unbiased"] or it_now < settings["burn_in"]:
            ts_alt = SparseT(t2, m, comp, draw(), mode, smeth)
        else:
            # fricc/friccsd.py:117-136: after the burn-in the second sample is
            # taken of the running average Y(t+1) = (1 - 1/t) Y(t) + X(t+1)/t,
            # with the reference's operation order t2/it + (it-1)/it * avg;
            # at it == 0 the average starts as t2 itself
            it = it_now - settings["burn_in"]
            if it == 0:
                state["t2_avg"] = t2.copy()
            else:
                state["t2_avg"] = t2 / it + (it - 1) / it * state["t2_avg"]
            ts_alt = SparseT(np.ascontiguousarray(state["t2_avg"]), m, comp, draw(), mode, smeth)

    Foo = cc_Foo(t1, t2, eris)
    Fvv = cc_Fvv(t1, t2, eris)
    Fov = cc_Fov(t1, t2, eris)
    Foo[np.diag_indices(o)] -= mo_e_o
    Fvv[np.diag_indices(v)] -= mo_e_v

    t1new = -2 * einsum("kc,ka,ic->ia", fov, t1, t1)
    t1new += einsum("ac,ic->ia", Fvv, t1)
    t1new += -einsum("ki,ka->ia", Foo, t1)
    t1new += 2 * einsum("kc,kica->ia", Fov, t2)
    t1new += -einsum("kc,ikca->ia", Fov, t2)
    t1new += einsum("kc,ic,ka->ia", Fov, t1, t1)
    t1new += fov
    t1new += 2 * einsum("kcai,kc->ia", eris.ovvo, t1)
    t1new += -einsum("kiac,kc->ia", eris.oovv, t1)
    ovvv = eris.ovvv
    t1new += 2 * einsum("kdac,ikcd->ia", ovvv, t2)
    t1new += -einsum("kcad,ikcd->ia", ovvv, t2)
    t1new += 2 * einsum("kdac,kd,ic->ia", ovvv, t1, t1)
    t1new += -einsum("kcad,kd,ic->ia", ovvv, t1, t1)
    ovoo = eris.ovoo
    t1new += -2 * einsum("lcki,klac->ia", ovoo, t2)
    t1new += einsum("kcli,klac->ia", ovoo, t2)
    t1new += -2 * einsum("lcki,lc,ka->ia", ovoo, t1, t1)
    t1new += einsum("kcli,lc,ka->ia", ovoo, t1, t1)

    tmp2 = einsum("kibc,ka->abic", eris.oovv, -t1)
    tmp2 += ovvv.transpose(1, 3, 0, 2)
    tmp = einsum("abic,jc->ijab", tmp2, t1)
    t2new = tmp + tmp.transpose(1, 0, 3, 2)
    tmp2 = einsum("kcai,jc->akij", eris.ovvo, t1)
    tmp2 += ovoo.transpose(1, 3, 0, 2)
    tmp = einsum("akij,kb->ijab", tmp2, t1)
    t2new -= tmp + tmp.transpose(1, 0, 3, 2)
    t2new += eris.ovov.transpose(0, 2, 1, 3)

    Loo_ = Loo(t1, t2, eris)
    Lvv_ = Lvv(t1, t2, eris)
    Loo_[np.diag_indices(o)] -= mo_e_o
    Lvv_[np.diag_indices(v)] -= mo_e_v

    Woooo = sparse_cc_Woooo(t1, ts_alt, eris) if "O^4V^2" in cc else cc_Woooo(t1, t2, eris)
    if "O^3V^3" in cc:
        Wvoov = sparse_cc_Wvoov(t1, ts_alt, eris)
        Wvovo = sparse_cc_Wvovo(t1, ts_alt, eris)
    else:
        Wvoov = cc_Wvoov(t1, t2, eris)
        Wvovo = cc_Wvovo(t1, t2, eris)
    Wvvvv = cc_Wvvvv(t1, t2, eris)

    if "O^4V^2" in cc:
        t2new += einsum("klij,ka,lb->ijab", Woooo, t1, t1)
        contract(Woooo, ts, t2new, "0101")
    else:
        tau = t2 + einsum("ia,jb->ijab", t1, t1)
        t2new += einsum("klij,klab->ijab", Woooo, tau)

    if settings["vvvv_t1t1"] == "reference":
        # friccsd.py:222 (quirk A.1): einsum("abcd,ia,jb->ijab") = t_ia t_jb sum_cd W_abcd
        Wsum = Wvvvv.reshape(Wvvvv.shape[0], Wvvvv.shape[1], -1).sum(axis=2)
        t2new += einsum("ab,ia,jb->ijab", Wsum, t1, t1)
    else:
        t2new += einsum("abcd,ic,jd->ijab", Wvvvv, t1, t1)  # RCCSD tau form
    if "O^2V^4" in cc:
        contract(Wvvvv, ts, t2new, "2323")
    else:
        t2new += einsum("abcd,ijcd->ijab", Wvvvv, t2)

    tmp = einsum("ac,ijcb->ijab", Lvv_, t2)
    t2new += tmp + tmp.transpose(1, 0, 3, 2)
    tmp = einsum("ki,kjab->ijab", Loo_, t2)
    t2new -= tmp + tmp.transpose(1, 0, 3, 2)

    if "O^3V^3" in cc:
        tmp = np.zeros_like(t2new)
        contract(Wvoov, ts, tmp, "1302")
        contract(Wvovo, ts, tmp, "1202")
        t2new += tmp + tmp.transpose(1, 0, 3, 2)
        tmp = np.zeros_like(t2new)
        contract(Wvoov, ts, tmp, "1303")
        t2new -= tmp + tmp.transpose(1, 0, 3, 2)
        tmp = np.zeros_like(t2new)
        contract(Wvovo, ts, tmp, "1203")
        t2new -= tmp + tmp.transpose(1, 0, 3, 2)
    else:
        tmp = 2 * einsum("akic,kjcb->ijab", Wvoov, t2)
        tmp -= einsum("akci,kjcb->ijab", Wvovo, t2)
        t2new += tmp + tmp.transpose(1, 0, 3, 2)
        tmp = einsum("akic,kjbc->ijab", Wvoov, t2)
        t2new -= tmp + tmp.transpose(1, 0, 3, 2)
        tmp = einsum("bkci,kjac->ijab", Wvovo, t2)
        t2new -= tmp + tmp.transpose(1, 0, 3, 2)

    eia = mo_e_o[:, None] - mo_e_v
    eijab = eia[:, None, :, None] + eia[None, :, None, :]
    t1new /= eia
    t2new /= eijab
    if return_sparse:
        return t1new, t2new, (ts, ts_alt)
    return t1new, t2new


class GlibcUniforms:
    """u = rand()/RAND_MAX from the C library's own generator after srand(seed)
    (lin_space_add_const, src/fri_utils.cpp:47).  Calls the function; does not
    reimplement it."""

    def __init__(self, seed):
        self._libc = ctypes.CDLL(None)
        self._libc.rand.restype = ctypes.c_int
        self._libc.srand(ctypes.c_uint(seed))

    def __iter__(self):
        return self

    def __next__(self):
        return self._libc.rand() / 2147483647.0


def kernel(eris, settings, uniforms, max_cycle=50, tol=1e-8, tolnormt=1e-6, t1=None, t2=None):
    """fricc/friccsd.py:379-456 (no DIIS)."""
    if t1 is None and t2 is None:
        _, t1, t2 = init_amps(eris)
    eold = 0.0
    ecc = energy(t1, t2, eris)
    energies = []
    conv = False
    state = {}
    for istep in range(max_cycle):
        state["iter"] = istep  # cc._iter (the reference never sets it, quirk A.3)
        t1new, t2new = update_amps(t1, t2, eris, settings, uniforms, state=state)
        normt = np.linalg.norm(amplitudes_to_vector(t1new, t2new) - amplitudes_to_vector(t1, t2))
        t1, t2 = t1new, t2new
        eold, ecc = ecc, energy(t1, t2, eris)
        energies.append(ecc)
        if abs(ecc - eold) < tol and normt < tolnormt:
            conv = True
            break
        if math.isnan(ecc - eold):
            conv = False
            break
        elif abs(ecc - eold) > 1e5:
            conv = False
            break
    return conv, ecc, t1, t2, energies
