"""Adjoint of the three-phonon linewidths and of the RTA thermal conductivity with respect
to the third-order force constants, restated in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference has no adjoint; this differentiates its forward algorithm as the oracle
restates it in scattering.py and kappa.py:

  Gamma_raw(q1,b1) = sum_trip sum_{b2,b3} c(b1,b2,b3) |Psi(b1,b2,b3)|^2
      (compute_threephonon_scattering, scattering_threephonon.f90:88-135), with
      Psi = sum E1 E2 E3 Phi and Phi = sum_t mwm_t exp(-2 pi i (q2.lv2_t + q3.lv3_t))
      (pretransform_phi3 :275-317);
  linewidth = degeneracy average of Gamma_raw (scattering.f90:260-277);
  kappa_RTA = symmetrize_kappa(get_kappa(F0)), F0 = v / (2 linewidth)
      (scattering.f90:278-292, kappa.f90:28-75, :308-324).

Gamma is quadratic in Phi3, so for a cotangent g = dL/dGamma_raw

  dL/dmwm_t[i,j,k] = sum_trip 2 Re( ph_t * Y[3a1+i, 3a2+j, 3a3+k] ),
  Y[ia,ib,ic] = sum_b E1[ia,b1] E2[ib,b2] E3[ic,b3] H[b] conj(Psi[b]),
  H[b] = g(q1,b1) c(b1,b2,b3).
"""
import numpy as np

from . import units as U
from .scattering import (THREEPHONON_PREFACTOR, irreducible_triplets, occupations_and_smearing,
                         pretransform_phi3, psi_chain, rotation_table, velocity_difference_widths)


def raw_cotangent(disp, dL_dlw):
    """Cotangent of the raw linewidths from that of the degeneracy-averaged ones.  The
    forward averages band by band in place (scattering.f90:264-277); each step is a
    symmetric averaging map, so the adjoint applies the same steps in reverse order."""
    nirr, nb = disp.om_irr.shape
    g = np.where(disp.om_irr >= U.FREQTOL, np.asarray(dL_dlw, float), 0.0).copy()
    for q in range(nirr):
        for b in reversed(range(nb)):
            if disp.om_irr[q, b] < U.FREQTOL:
                continue
            dm = disp.degen_irr[q][b]
            g[q, dm] = g[q, dm].sum() / len(dm)
    return np.where(disp.om_irr >= U.FREQTOL, g, 0.0)


def gamma_adjoint(p, fc3, mesh, disp, ops, temperature, g_raw, classical=False,
                  integrationtype=2, sigma=1.0):
    """dL/dmwm [nfc3, 3, 3, 3] for L = sum g_raw * Gamma_raw (three-phonon part) at one
    temperature; the loop of scattering.scattering_rates run backwards."""
    nb = disp.nb
    nfull = mesh.nfull
    be, sigsq, sigavg = occupations_and_smearing(p, mesh, disp, temperature, classical)
    sigma_fixed = sigma * U.FREQ_THZ_TO_HARTREE
    thresh, reclat6 = np.max(sigavg) * 1e-4, p.reclat / np.array(mesh.dims)[None, :]
    rot = rotation_table(p, mesh, ops)
    mcw = 1.0 / nfull
    grad = np.zeros_like(fc3["mwm"])
    a1, a2, a3 = fc3["a1"], fc3["a2"], fc3["a3"]
    for q1 in range(mesh.nirr):
        rowmask = disp.om_irr[q1] >= U.FREQTOL
        g1 = np.where(rowmask, g_raw[q1], 0.0)
        if not np.any(g1):
            continue
        om1 = disp.om_irr[q1]
        E1 = disp.egv_irr[q1] / np.sqrt(np.where(om1 > 0, om1, 1.0))[None, :]
        for (q2, q3, perm, red) in irreducible_triplets(mesh, rot, q1):
            om2 = disp.om_full[q2]; om3 = disp.om_full[q3]
            E2 = disp.egv_full[q2] / np.sqrt(np.where(om2 > 0, om2, 1.0))[None, :]
            E3 = disp.egv_full[q3] / np.sqrt(np.where(om3 > 0, om3, 1.0))[None, :]
            ptf = pretransform_phi3(fc3, mesh.qcart[q2], mesh.qcart[q3], nb)
            psi = psi_chain(E1, E2, E3, ptf)
            i2, i3 = mesh.irr_of_full[q2], mesh.irr_of_full[q3]
            n2 = be[i2][None, :, None]; n3 = be[i3][None, None, :]
            w1 = om1[:, None, None]; w2 = om2[None, :, None]; w3 = om3[None, None, :]
            if integrationtype == 2:
                sig = np.sqrt(sigsq[q1][:, None, None] + sigsq[i2][None, :, None]
                              + sigsq[i3][None, None, :])
            elif integrationtype == 6:
                sig = velocity_difference_widths(disp.vel_irr[q1], disp.vel_full[q2], disp.vel_full[q3],
                                                 reclat6, thresh)
            else:
                sig = np.full((nb, nb, nb), sigma_fixed)
            d0 = U.gauss(w1, -w2 + w3, sig) - U.gauss(w1, w2 - w3, sig)
            d1 = U.gauss(w1, w2 + w3, sig) - U.gauss(w1, -w2 - w3, sig)
            c = perm * THREEPHONON_PREFACTOR * mcw * ((n2 - n3) * d0 + (n2 + n3 + 1.0) * d1)
            mask = ((w2 >= U.FREQTOL) & (np.abs(w1 - w2) >= U.FREQTOL)
                    & (w3 >= U.FREQTOL) & (np.abs(w1 - w3) >= U.FREQTOL)
                    & (np.abs(w2 - w3) >= U.FREQTOL))
            H = np.where(mask, c, 0.0) * g1[:, None, None] * len(red)
            Z = H * np.conj(psi)
            Y = np.einsum("ia,jb,kc,abc->ijk", E1, E2, E3, Z, optimize=True)
            ph = np.exp(-1j * U.TWOPI * (fc3["lv2"] @ mesh.qcart[q2] + fc3["lv3"] @ mesh.qcart[q3]))
            Yb = Y.reshape(len(Y) // 3, 3, len(Y) // 3, 3, len(Y) // 3, 3)
            blk = Yb[a1, :, a2, :, a3, :]             # [t, i, j, k]
            grad += 2.0 * np.real(ph[:, None, None, None] * blk)
    return grad


def mwm_to_m(p, fc3, grad_mwm):
    """Gradient with respect to the un-weighted tensors m (mwm = m / sqrt(m1 m2 m3),
    type_forceconstant_thirdorder.f90:277)."""
    ism = p.invsqrtmass
    w = ism[fc3["a1"]] * ism[fc3["a2"]] * ism[fc3["a3"]]
    return grad_mwm * w[:, None, None, None]


def kappa_rta_cotangent(p, mesh, disp, ops, sr, temperature, dL_dkappa, classical=False):
    """dL/dlinewidth [nirr, nb] for L = sum dL_dkappa * kappa_RTA, kappa_RTA the symmetrised
    single-mode kappa from F0 = v / (2 linewidth) (kappa.f90:28-75, :308-324).  lo_chop is
    taken as the identity."""
    from .kappa import get_kappa
    G = np.asarray(dL_dkappa, float)
    # symmetrize_kappa is an average of R k R^T; its adjoint averages R^T G R
    Gs = sum(o.m.T @ G @ o.m for o in ops) / len(ops)
    _, kq = get_kappa(p, mesh, disp, ops, sr.F0, temperature, classical, per_mode=True)
    out = np.zeros((mesh.nirr, disp.nb))
    for q in range(mesh.nirr):
        for b in range(disp.nb):
            lw = sr.linewidth[q, b]
            if disp.om_irr[q, b] < U.FREQTOL or not lw > U.FREQTOL:
                continue
            # kq is linear in F0 = v / (2 lw)
            out[q, b] = -mesh.weight_irr[q] * np.sum(Gs * kq[q, b]) / lw
    return out
