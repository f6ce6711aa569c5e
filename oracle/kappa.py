"""RTA, iterative-BTE and off-diagonal thermal conductivity, restated in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows src/thermal_conductivity/kappa.f90: get_kappa :28-75,
get_kappa_offdiag :77-306, symmetrize_kappa :308-324,
iterative_solution :326-495.
"""
import numpy as np

from . import units as U
from .phonons import dynamical_matrix, eigenvector_transformation


def get_kappa(p, mesh, disp, ops, Fn, temperature, classical=False, per_mode=False):
    """get_kappa (kappa.f90:28-75)."""
    kappa = np.zeros((3, 3))
    kq = np.zeros((mesh.nirr, disp.nb, 3, 3))
    M = np.array([o.m for o in ops])
    for q in range(mesh.nirr):
        for b in range(disp.nb):
            om = disp.om_irr[q, b]
            if om < U.FREQTOL:
                continue
            cv = U.KB_HARTREE if classical else float(U.harmonic_oscillator_cv(temperature, om))
            v0 = M @ Fn[q, b]
            v1 = M @ disp.vel_irr[q, b]
            v2 = np.einsum("ni,nj->ij", v0, v1) / len(ops)
            buf = cv * v2 / p.volume
            kq[q, b] = buf
            kappa += buf * mesh.weight_irr[q]
    kappa = U.chop(kappa, np.abs(kappa).sum() * 1e-6)
    if per_mode:
        return kappa, kq
    return kappa


def symmetrize_kappa(kappa, ops):
    """symmetrize_kappa (kappa.f90:308-324)."""
    tmp = np.zeros((3, 3))
    for o in ops:
        tmp += o.m @ kappa @ o.m.T
    k = tmp / len(ops)
    return U.chop(k, np.abs(k).sum() * 1e-6)


def iterative_solution(p, mesh, disp, ops, sr, temperature, niter=200, tol=1e-5,
                       classical=False, log=None):
    """iterative_solution (kappa.f90:326-495). Returns (Fn, scfcheck list, kappa history)."""
    nb, nirr, nfull = disp.nb, mesh.nirr, mesh.nfull
    Fn = sr.F0.copy()
    mix = 0.95
    hist = [get_kappa(p, mesh, disp, ops, Fn, temperature, classical)]
    scf = []
    rows = sr.rows
    for it in range(1, niter + 1):
        Fbb = np.zeros((nfull, nb, 3))
        for q in range(nfull):
            iop = mesh.op_of_full[q]
            k = mesh.irr_of_full[q]
            m = ops[abs(iop) - 1].m
            v = Fn[k] @ m.T
            Fbb[q] = v if iop > 0 else -v
        buf = sr.Xi @ Fbb.reshape(nfull * nb, 3)      # [nrow, 3]
        Fnb = np.zeros((nirr, nb, 3))
        for i, (q1, b1) in enumerate(rows):
            if sr.qs[q1, b1] < U.FREQTOL:
                Fnb[q1, b1] = 0.0
            else:
                Fnb[q1, b1] = -buf[i] / sr.qs[q1, b1]
        # degeneracy average, then add F0, sequentially per band (kappa.f90:429-448)
        for q1 in range(nirr):
            for b1 in range(nb):
                dm = disp.degen_irr[q1][b1]
                v0 = Fnb[q1, dm].sum(axis=0) / len(dm)
                Fnb[q1, dm] = v0
                Fnb[q1, b1] = sr.F0[q1, b1] + Fnb[q1, b1]
        g0 = 0.0
        for q in range(nirr):
            for b in range(nb):
                g1 = np.sum((Fn[q, b] - Fnb[q, b]) ** 2)
                g2 = np.sum(Fn[q, b] ** 2)
                if g2 > U.SQTOL:
                    g0 += g1 / g2
                Fn[q, b] = Fn[q, b] * (1.0 - mix) + mix * Fnb[q, b]
        scf.append(g0 / nirr / nb)
        kap = get_kappa(p, mesh, disp, ops, Fn, temperature, classical)
        hist.append(kap)
        if log is not None:
            log.append((it, kap, scf[-1]))
        if it >= 3 and sum(scf[-3:]) < tol:
            break
        if it > 15 and mix > 0.5:
            mix *= 0.98
    return Fn, scf, hist


def get_kappa_offdiag(p, fc2, mesh, disp, ops, linewidth, temperature, classical=False,
                      include_diagonal=False, terms=False):
    """get_kappa_offdiag (kappa.f90:77-306).  The reference forms a (nb^2 x nb^2)
    Kronecker matrix and multiplies it by the flattened gradient; that product
    equals, entry for entry, avg_op[ egw^H G egw ], which is what is formed here.

    The reference sums over mode pairs j != k (:246-285).  include_diagonal=True adds the
    j = k terms, where the coherence velocity is the group velocity, the Lorentzian is
    1/(2 Gamma) and the heat-capacity mean is c_v, so that each such term is the mode's
    single-mode-approximation kappa: a check on the velocities, the Lorentzian, the c_v
    mean and the prefactor.  terms=True also returns every pair's weighted contribution
    [q_irr, j, k, 3, 3] (before the chop)."""
    nb = disp.nb
    kod = np.zeros((3, 3))
    tq = np.zeros((mesh.nirr, nb, nb, 3, 3)) if terms else None
    ism = p.invsqrtmass
    mw = np.repeat(ism, 3)
    for iq in range(mesh.nirr):
        qf = mesh.irr_full_index[iq]
        q = mesh.qcart[qf]
        _, gD = dynamical_matrix(p, fc2, q)
        # buf_cm1: gradient with the masses divided out, indexed G[ia, ib] (kappa.f90:143-156)
        G = gD / (mw[:, None] * mw[None, :])[None]
        om = disp.om_irr[iq]
        lg = mesh.littlegroup_irr[iq]
        V = np.zeros((3, nb, nb), dtype=np.complex128)
        for k in lg:
            op = ops[k]
            gm = eigenvector_transformation(p, q, op)
            egw = gm @ disp.egv_irr[iq]
            scale = np.where(om > U.FREQTOL, 1.0 / np.sqrt(np.where(om > 0, om, 1.0) * 2.0), 0.0)
            egw = egw * mw[:, None] * scale[None, :]
            for z in range(3):
                V[z] += egw.conj().T @ G[z] @ egw
        V /= len(lg)
        V = U.chop_complex(V, 1e-10 / (U.GROUPVEL_HARTREEBOHR_TO_MS / 1000))
        vel = V.real  # [z, i, j]
        # rotate over the star (kappa.f90:221-233)
        sops = mesh.star_ops[iq]
        Ms = np.array([ops[abs(o) - 1].m for o in sops])
        velsq = np.zeros((nb, nb, 3, 3))
        for i in range(nb):
            for j in range(nb):
                v1 = Ms @ vel[:, i, j]
                velsq[i, j] = np.einsum("ni,nj->ij", v1, v1) / len(sops)
        pref = mesh.weight_irr[iq] / p.volume
        for j in range(nb):
            om1 = om[j]
            if om1 < U.FREQTOL:
                continue
            t1 = linewidth[iq, j]
            if t1 < U.FREQTOL:
                continue
            for k in range(nb):
                if j == k and not include_diagonal:
                    continue
                om2 = om[k]
                if om2 < U.FREQTOL:
                    continue
                t2 = linewidth[iq, k]
                if t2 < U.FREQTOL:
                    continue
                if classical:
                    f0 = U.KB_HARTREE
                else:
                    f0 = 0.5 * (float(U.harmonic_oscillator_cv(temperature, om1))
                                + float(U.harmonic_oscillator_cv(temperature, om2)))
                tau = (t1 + t2) / ((t1 + t2) ** 2 + (om1 - om2) ** 2)
                kod += velsq[j, k] * tau * f0 * pref
                if terms:
                    tq[iq, j, k] = velsq[j, k] * tau * f0 * pref
    kod = U.chop(kod, np.abs(kod).sum() * 1e-6)
    if terms:
        return kod, tq
    return kod
