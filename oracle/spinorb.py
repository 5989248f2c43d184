"""Independent spin-orbital CCSD Jacobi update -- TEST INFRASTRUCTURE ONLY.

Stanton, Gauss, Watts & Bartlett, J. Chem. Phys. 94, 4334 (1991), eqs. 1-13
(the form used in T. D. Crawford's CCSD programming notes).  It shares no code
and no derivation with ``oracle/rccsd.py`` (closed-shell spatial-orbital
PySCF formulas), so agreement of one Jacobi step pins those formulas:
the update t -> (residual without diagonal Fock terms) / D is unique.

Spin orbital P = 2*p + sigma (sigma 0 = alpha, 1 = beta); the first 2*nocc
are occupied.
"""
import numpy as np


def spin_integrals(eri, fock, nocc):
    """<PQ||RS> and the spin-orbital Fock matrix from chemists' (pq|rs)."""
    n = eri.shape[0]
    spin = np.arange(2 * n) % 2
    sp = np.arange(2 * n) // 2
    # <PQ|RS> = (pr|qs) delta(sP,sR) delta(sQ,sS)
    g = eri[np.ix_(sp, sp, sp, sp)].transpose(0, 2, 1, 3)
    same = (spin[:, None] == spin[None, :])
    g = g * same[:, None, :, None] * same[None, :, None, :]
    asym = g - g.transpose(0, 1, 3, 2)
    f = fock[np.ix_(sp, sp)] * same
    return asym, f


def spin_amplitudes(t1, t2):
    """Closed-shell spatial (t1, t2) -> spin-orbital amplitudes.  t2 must obey
    t2[i,j,a,b] == t2[j,i,b,a]."""
    o, v = t1.shape
    so, sv = 2 * o, 2 * v
    si = np.arange(so) % 2
    sa = np.arange(sv) % 2
    oi = np.arange(so) // 2
    va = np.arange(sv) // 2
    T1 = t1[np.ix_(oi, va)] * (si[:, None] == sa[None, :])
    d = lambda x, y: (x[:, None] == y[None, :]).astype(float)
    t_ijab = t2[np.ix_(oi, oi, va, va)]
    t_ijba = t2[np.ix_(oi, oi, va, va)].transpose(0, 1, 3, 2)
    t_ijba = t2.transpose(0, 1, 3, 2)[np.ix_(oi, oi, va, va)]
    T2 = (d(si, sa)[:, None, :, None] * d(si, sa)[None, :, None, :]) * t_ijab \
        - (d(si, sa)[:, None, None, :] * d(si, sa)[None, :, :, None]) * t_ijba
    return T1, T2


def update(T1, T2, asym, f, nocc):
    """One Jacobi step; returns (T1new, T2new) with D = f_ii - f_aa etc."""
    so = 2 * nocc
    o = slice(0, so)
    v = slice(so, None)
    fov = f[o, v]
    foo = f[o, o]
    fvv = f[v, v]
    nv = fvv.shape[0]
    g_oovv = asym[o, o, v, v]
    g_ovvv = asym[o, v, v, v]
    g_ooov = asym[o, o, o, v]
    g_oooo = asym[o, o, o, o]
    g_vvvv = asym[v, v, v, v]
    g_ovvo = asym[o, v, v, o]
    g_ovov = asym[o, v, o, v]
    g_vvvo = asym[v, v, v, o]
    g_ovoo = asym[o, v, o, o]
    g_vovo = asym[v, o, v, o]

    taut = T2 + 0.5 * (np.einsum("ia,jb->ijab", T1, T1) - np.einsum("ib,ja->ijab", T1, T1))
    tau = T2 + np.einsum("ia,jb->ijab", T1, T1) - np.einsum("ib,ja->ijab", T1, T1)

    Fae = fvv * (1 - np.eye(nv))
    Fae -= 0.5 * np.einsum("me,ma->ae", fov, T1)
    Fae += np.einsum("mf,mafe->ae", T1, g_ovvv)
    Fae -= 0.5 * np.einsum("mnaf,mnef->ae", taut, g_oovv)

    Fmi = foo * (1 - np.eye(so))
    Fmi += 0.5 * np.einsum("ie,me->mi", T1, fov)
    Fmi += np.einsum("ne,mnie->mi", T1, g_ooov)
    Fmi += 0.5 * np.einsum("inef,mnef->mi", taut, g_oovv)

    Fme = fov + np.einsum("nf,mnef->me", T1, g_oovv)

    Wmnij = g_oooo.copy()
    x = np.einsum("je,mnie->mnij", T1, g_ooov)
    Wmnij += x - x.transpose(0, 1, 3, 2)
    Wmnij += 0.25 * np.einsum("ijef,mnef->mnij", tau, g_oovv)

    Wabef = g_vvvv.copy()
    # <am||ef> = -<ma||ef>
    x = np.einsum("mb,maef->abef", T1, g_ovvv)
    Wabef += x - x.transpose(1, 0, 2, 3)
    Wabef += 0.25 * np.einsum("mnab,mnef->abef", tau, g_oovv)

    Wmbej = g_ovvo.copy()
    Wmbej += np.einsum("jf,mbef->mbej", T1, g_ovvv)
    # <mn||ej> = -<mn||je>
    Wmbej += np.einsum("nb,mnje->mbej", T1, g_ooov)
    Wmbej -= 0.5 * np.einsum("jnfb,mnef->mbej", T2, g_oovv)
    Wmbej -= np.einsum("jf,nb,mnef->mbej", T1, T1, g_oovv)

    # T1
    r1 = fov.copy()
    r1 += np.einsum("ie,ae->ia", T1, Fae)
    r1 -= np.einsum("ma,mi->ia", T1, Fmi)
    r1 += np.einsum("imae,me->ia", T2, Fme)
    r1 -= np.einsum("nf,naif->ia", T1, g_ovov)
    r1 -= 0.5 * np.einsum("imef,maef->ia", T2, g_ovvv)
    # <nm||ei> = <mn||ie>... use g_ooov: <nm||ei> = <mn||ie>
    r1 -= 0.5 * np.einsum("mnae,mnie->ia", T2, g_ooov)

    # T2
    r2 = g_oovv.copy()
    Ft = Fae - 0.5 * np.einsum("mb,me->be", T1, Fme)
    x = np.einsum("ijae,be->ijab", T2, Ft)
    r2 += x - x.transpose(0, 1, 3, 2)
    Ft = Fmi + 0.5 * np.einsum("je,me->mj", T1, Fme)
    x = np.einsum("imab,mj->ijab", T2, Ft)
    r2 -= x - x.transpose(1, 0, 2, 3)
    r2 += 0.5 * np.einsum("mnab,mnij->ijab", tau, Wmnij)
    r2 += 0.5 * np.einsum("ijef,abef->ijab", tau, Wabef)
    x = np.einsum("imae,mbej->ijab", T2, Wmbej)
    x -= np.einsum("ie,ma,mbej->ijab", T1, T1, g_ovvo)
    r2 += x - x.transpose(0, 1, 3, 2) - x.transpose(1, 0, 2, 3) + x.transpose(1, 0, 3, 2)
    x = np.einsum("ie,abej->ijab", T1, g_vvvo)
    r2 += x - x.transpose(1, 0, 2, 3)
    x = np.einsum("ma,mbij->ijab", T1, g_ovoo)
    r2 -= x - x.transpose(0, 1, 3, 2)

    eo = np.diag(foo)
    ev = np.diag(fvv)
    D1 = eo[:, None] - ev[None, :]
    D2 = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    return r1 / D1, r2 / D2


def energy(T1, T2, asym, f, nocc):
    so = 2 * nocc
    o = slice(0, so)
    v = slice(so, None)
    g = asym[o, o, v, v]
    return float(np.einsum("ia,ia", f[o, v], T1) + 0.25 * np.einsum("ijab,ijab", g, T2)
                 + 0.5 * np.einsum("ijab,ia,jb", g, T1, T1))


def update_linear(T1, T2, asym, f, nocc, lccd=False):
    """One Jacobi step of linearised CCSD (every term of ``update`` that is
    at most linear in T1 and T2; Taube & Bartlett, J. Chem. Phys. 130, 144112
    (2009), eqs. 13 and 15).  ``lccd``: T1 is held at zero (LCCD)."""
    so = 2 * nocc
    o = slice(0, so)
    v = slice(so, None)
    fov = f[o, v]
    foo = f[o, o]
    fvv = f[v, v]
    nv = fvv.shape[0]
    if lccd:
        T1 = np.zeros_like(T1)
    Fae = fvv * (1 - np.eye(nv))
    Fmi = foo * (1 - np.eye(so))

    r1 = fov.copy()
    r1 += np.einsum("ie,ae->ia", T1, Fae)
    r1 -= np.einsum("ma,mi->ia", T1, Fmi)
    r1 += np.einsum("imae,me->ia", T2, fov)
    r1 -= np.einsum("nf,naif->ia", T1, asym[o, v, o, v])
    r1 -= 0.5 * np.einsum("imef,maef->ia", T2, asym[o, v, v, v])
    r1 -= 0.5 * np.einsum("mnae,mnie->ia", T2, asym[o, o, o, v])

    r2 = asym[o, o, v, v].copy()
    x = np.einsum("ijae,be->ijab", T2, Fae)
    r2 += x - x.transpose(0, 1, 3, 2)
    x = np.einsum("imab,mj->ijab", T2, Fmi)
    r2 -= x - x.transpose(1, 0, 2, 3)
    r2 += 0.5 * np.einsum("mnab,mnij->ijab", T2, asym[o, o, o, o])
    r2 += 0.5 * np.einsum("ijef,abef->ijab", T2, asym[v, v, v, v])
    x = np.einsum("imae,mbej->ijab", T2, asym[o, v, v, o])
    r2 += x - x.transpose(0, 1, 3, 2) - x.transpose(1, 0, 2, 3) + x.transpose(1, 0, 3, 2)
    x = np.einsum("ie,abej->ijab", T1, asym[v, v, v, o])
    r2 += x - x.transpose(1, 0, 2, 3)
    x = np.einsum("ma,mbij->ijab", T1, asym[o, v, o, o])
    r2 -= x - x.transpose(0, 1, 3, 2)

    eo = np.diag(foo)
    ev = np.diag(fvv)
    D1 = eo[:, None] - ev[None, :]
    D2 = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    if lccd:
        return np.zeros_like(T1), r2 / D2
    return r1 / D1, r2 / D2


def energy_linear(T1, T2, asym, f, nocc):
    """Linearised CC energy: sum f_ia t_ia + 1/4 sum <ij||ab> t_ijab."""
    so = 2 * nocc
    o = slice(0, so)
    v = slice(so, None)
    return float(np.einsum("ia,ia", f[o, v], T1) + 0.25 * np.einsum("ijab,ijab", asym[o, o, v, v], T2))
