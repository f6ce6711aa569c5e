"""Harmonic properties on the q-mesh, restated in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows type_forceconstant_secondorder_dynamicalmatrix.f90 (dynamical matrix
:484-788, eigen/velocities :13-481), type_phonon_dispersions.f90:1083-1267
(degeneracy table) and type_phonon_dispersions_generation.f90:8-241
(unfolding to the full grid).  numpy's LAPACK `eigh` plays the role of zheev.

Degenerate-subspace gauge: the reference keeps whatever basis zheev returns.
This build fixes the basis of every degenerate subspace deterministically
(probe diagonalisation + phase fix, `canonical_gauge`) and the HIP path does
the same, so the two agree to rounding.  All degeneracy-averaged outputs are
independent of this choice.
"""
import numpy as np

from . import units as U

GOLD = 0.6180339887498949
PLASTIC = 0.7548776662466927


def gauge_probe(nb):
    """Diagonal of the fixed Hermitian probe used to split degenerate subspaces."""
    k = np.arange(1, nb + 1, dtype=np.float64)
    return (k * GOLD) % 1.0


def gauge_weights(nb):
    """Fixed complex linear functional used to fix the phase of each eigenvector."""
    k = np.arange(1, nb + 1, dtype=np.float64)
    return (1.0 + (k * PLASTIC) % 1.0) * np.exp(1j * 0.9 * k)


def dynamical_matrix(p, fc2, q):
    """fc%dynamicalmatrix (type_forceconstant_secondorder_dynamicalmatrix.f90:484-788),
    non-polar: D[(a2,j),(a1,i)] = sum_pairs m_ij e^{i 2pi q.R} / sqrt(m1 m2)
    -- stored transposed, as the reference flattens it (:747-777).  Also the
    three Cartesian gradients (i R_alpha e^{iqR} m)."""
    na = p.na
    nb = 3 * na
    qv = U.chop(q * U.TWOPI, 1e-13)
    Dt = np.zeros((na, na, 3, 3), dtype=np.complex128)
    gDt = np.zeros((3, na, na, 3, 3), dtype=np.complex128)
    ph = np.exp(1j * (fc2["lv2"] @ qv))
    for n in range(len(fc2["a1"])):
        a1 = fc2["a1"][n]; a2 = fc2["a2"][n]
        Dt[a1, a2] += ph[n] * fc2["m"][n]
        for al in range(3):
            gDt[al, a1, a2] += 1j * fc2["lv2"][n, al] * ph[n] * fc2["m"][n]
    ism = p.invsqrtmass
    w = ism[:, None] * ism[None, :]
    Dt *= w[:, :, None, None]
    gDt *= w[None, :, :, None, None]
    # D[(a2,j),(a1,i)] = Dt[a1,a2,i,j]
    D = Dt.transpose(1, 3, 0, 2).reshape(nb, nb)
    gD = gDt.transpose(0, 2, 4, 1, 3).reshape(3, nb, nb)
    return D, gD


def degenerate_subspaces(omega):
    """return_degenerate_subspaces (…_dynamicalmatrix.f90:911-980)."""
    subs = []
    for i in range(len(omega)):
        placed = False
        for s in subs:
            if any(abs(omega[j] - omega[i]) < U.FREQTOL for j in s):
                s.append(i)
                placed = True
                break
        if not placed:
            subs.append([i])
    return subs


def canonical_gauge(omega, egv):
    """Fix the basis inside every degenerate subspace (consecutive sorted modes
    closer than lo_freqtol) by diagonalising a fixed diagonal probe, then fix
    each vector's phase so a fixed linear functional is real positive."""
    nb = len(omega)
    pr = gauge_probe(nb)
    w = gauge_weights(nb)
    egv = egv.copy()
    i = 0
    while i < nb:
        j = i + 1
        while j < nb and omega[j] - omega[j - 1] < U.FREQTOL:
            j += 1
        if j - i > 1:
            Us = egv[:, i:j]
            M = Us.conj().T @ (pr[:, None] * Us)
            M = 0.5 * (M + M.conj().T)
            ev, V = np.linalg.eigh(M)
            egv[:, i:j] = Us @ V
        i = j
    c = w @ egv
    egv = egv * (np.conj(c) / np.abs(c))[None, :]
    return egv


def harmonic_point(p, fc2, q, at_gamma):
    """frequencies, eigenvectors, group velocities and degeneracy for one q."""
    D, gD = dynamical_matrix(p, fc2, q)
    nb = D.shape[0]
    lam, egv = np.linalg.eigh(D)
    if at_gamma:
        idx = np.argsort(np.abs(lam), kind="stable")[:3]
        lam[idx] = 0.0
    om_s = U.negsqrt(lam)
    order = np.argsort(om_s, kind="stable")
    lam = lam[order]
    egv = egv[:, order]
    omega = U.negsqrt(lam)
    egv = canonical_gauge(omega, egv)
    egv = U.chop_complex(egv, 1e-13)
    # Hellmann-Feynman velocities, degenerate subspaces averaged (:198-215, :287-374)
    hf = np.real(np.einsum("ib,aij,jb->ab", egv.conj(), gD, egv))
    vel = np.zeros((3, nb))
    for s in degenerate_subspaces(omega):
        vel[:, s] = hf[:, s].mean(axis=1)[:, None]
    for b in range(nb):
        if omega[b] > U.FREQTOL:
            vel[:, b] /= 2.0 * omega[b]
        else:
            vel[:, b] = 0.0
    vel = U.chop(vel, 1e-9 / U.GROUPVEL_HARTREEBOHR_TO_MS)
    return omega, egv, vel.T.copy()


def degeneracy_table(omega):
    """type_phonon_dispersions.f90:1190-1214: degenmode lists per band."""
    nb = len(omega)
    return [np.nonzero(np.abs(omega - omega[b]) < U.FREQTOL)[0] for b in range(nb)]


def eigenvector_transformation(p, q, op):
    """lo_eigenvector_transformation_matrix (type_symmetryoperation.f90:243-293)."""
    na = p.na
    gm = np.zeros((3 * na, 3 * na), dtype=np.complex128)
    for a2 in range(na):
        a1 = op.fmap[a2]
        v0 = op.im @ (p.rcart[a1] - op.tr) - p.rcart[a2]
        iq = -(v0 @ q) * U.TWOPI
        gm[3 * a1:3 * a1 + 3, 3 * a2:3 * a2 + 3] = op.m * np.exp(1j * iq)
    return gm


class Dispersions:
    pass


def generate_dispersions(p, fc2, ops, mesh):
    """lo_phonon_dispersions%generate (type_phonon_dispersions_generation.f90:8-241)."""
    nb = 3 * p.na
    d = Dispersions()
    d.nb = nb
    d.om_irr = np.zeros((mesh.nirr, nb))
    d.egv_irr = np.zeros((mesh.nirr, nb, nb), dtype=np.complex128)
    d.vel_irr = np.zeros((mesh.nirr, nb, 3))
    d.degen_irr = []
    for k, qi in enumerate(mesh.irr_full_index):
        q = mesh.qcart[qi]
        om, e, v = harmonic_point(p, fc2, q, q @ q < U.SQTOL)
        d.om_irr[k] = om; d.egv_irr[k] = e; d.vel_irr[k] = v
        d.degen_irr.append(degeneracy_table(om))
    d.om_full = np.zeros((mesh.nfull, nb))
    d.egv_full = np.zeros((mesh.nfull, nb, nb), dtype=np.complex128)
    d.vel_full = np.zeros((mesh.nfull, nb, 3))
    for q in range(mesh.nfull):
        k = mesh.irr_of_full[q]
        iop = mesh.op_of_full[q]
        op = ops[abs(iop) - 1]
        gm = eigenvector_transformation(p, mesh.qcart[mesh.irr_full_index[k]], op)
        e = gm @ d.egv_irr[k]
        v = d.vel_irr[k] @ op.m.T
        if iop < 0:
            e = np.conj(e)
            v = -v
        d.egv_full[q] = e
        d.om_full[q] = d.om_irr[k]
        d.vel_full[q] = v
    d.omega_min = min([o for o in d.om_irr.ravel() if abs(o) > U.FREQTOL] + [np.inf])
    return d
