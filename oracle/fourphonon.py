"""Four-phonon scattering (test infrastructure, see oracle/__init__.py).

compute_fourphonon_scattering (scattering_fourphonon.f90:1-328) with qg4ph = qgrid (the
Monte-Carlo grids are then the identity with weight 1/N_q): for an irreducible q1, the
quartets q2 <= q3 <= q4 = -q1-q2-q3 kept by quartet_is_irreducible (:330-456), their images
under the little group of q1 with permutations removed, pretransform_phi4 (:459-506),
Psi = sum phi4 e1 e2 e3 e4 (unconjugated, e / sqrt(omega)), the four occupation and
smeared-delta channels with the permutation multiplicities, the linewidth and the
off-diagonal scattering-matrix terms averaged over the little group (:143-205).
"""
import math

import numpy as np

from . import units as U

FOURPHONON_PREFACTOR = math.pi / 96.0   # scattering.f90:31


def fourth_grid_index(i1, i2, i3, dims):
    """fft_fourth_grid_index: the grid point -q1-q2-q3, 0-based indices."""
    n1, n2, n3 = dims
    g = np.zeros(3, dtype=np.int64)
    for i in (i1, i2, i3):
        g += np.array([i // (n2 * n3), (i // n3) % n2, i % n3])
    g = (-g) % np.array(dims)
    return int(g[0] * n2 * n3 + g[1] * n3 + g[2])


def quartet_images(mesh, rot, q1, q2, q3, q4):
    """quartet_is_irreducible (scattering_fourphonon.f90:330-456): (reducible, images)."""
    lg = mesh.littlegroup_irr[q1]
    newqp, newsort = [], []
    isred = False
    for o in lg:
        a, b, c = int(rot[o, q2]), int(rot[o, q3]), int(rot[o, q4])
        if a < 0 or b < 0 or c < 0:
            newqp.append(None)
            newsort.append(None)
            continue
        s = sorted((a, b, c))
        newqp.append((a, b, c))
        newsort.append(tuple(s))
        if s[0] > q2 or s[1] > q3:
            isred = True
    if isred:
        return True, []
    newqp = [x if x is not None else (q2, q3, q4) for x in newqp]
    newsort = [x if x is not None else (q2, q3, q4) for x in newsort]
    red = []
    n = len(lg)
    for j in range(n):
        if not any(newsort[j] == newsort[k] for k in range(j + 1, n)):
            red.append(newqp[j])
    return False, red


def pretransform_phi4(fc4, q2, q3, q4, nb):
    """pretransform_phi4 (scattering_fourphonon.f90:459-506): [ia, ib, ic, id]."""
    ptf = np.zeros((nb, nb, nb, nb), dtype=np.complex128)
    iqr = -(fc4["lv2"] @ q2 + fc4["lv3"] @ q3 + fc4["lv4"] @ q4) * U.TWOPI
    ph = np.cos(iqr) + 1j * np.sin(iqr)
    for t in range(len(ph)):
        a1, a2, a3, a4 = fc4["a1"][t], fc4["a2"][t], fc4["a3"][t], fc4["a4"][t]
        ptf[3 * a1:3 * a1 + 3, 3 * a2:3 * a2 + 3, 3 * a3:3 * a3 + 3, 3 * a4:3 * a4 + 3] += fc4["mwm"][t] * ph[t]
    return ptf


def dirac4(om1, om2, om3, om4, sig):
    """get_dirac (scattering_fourphonon.f90:259-327) for broadcast frequency arrays."""
    g = lambda x, mu: U.gauss(x, mu, sig)
    d0 = g(om1, om2 + om3 + om4) - g(om1, -om2 - om3 - om4)
    d1 = g(om1, -om2 + om3 + om4) - g(om1, om2 - om3 - om4)
    d2 = g(om1, om2 - om3 + om4) - g(om1, -om2 + om3 - om4)
    d3 = g(om1, om2 + om3 - om4) - g(om1, -om2 - om3 + om4)
    return d0, d1, d2, d3


def widths6(v1, v2, v3, v4, reclat, thresh):
    """Integration type 6 of get_dirac: mean over the six pairs of sqrt(max((dv.reclat)^2)/2)."""
    vs = [v1[:, None, None, None, :], v2[None, :, None, None, :], v3[None, None, :, None, :],
          v4[None, None, None, :, :]]
    s = 0.0
    for i in range(4):
        for j in range(i + 1, 4):
            d = (vs[i] - vs[j]) @ reclat
            s = s + np.sqrt(np.max(d * d, axis=-1) * 0.5) / 6.0
    return s + thresh


def row_terms(p, fc4, mesh, disp, be, sigsq, q1, integrationtype, sigma_fixed, thresh, rot, reduce=True,
              pairs=None, npoints=None):
    """Linewidth contributions g[b1] and off-diagonal terms od[b1, q, b] of every row of q1
    (before the little-group average).  reduce=False sums every ordered (q2, q3) pair with
    channels 0 and 1 only (the brute-force form: channels 2 and 3 relabel channel 1), and
    channel 1 three times where q2 = q3 = q4 (the reference weights each of the three
    channels once there, :114-119).  pairs: the (q2, q3) draws of a Monte-Carlo row in loop
    order (every pair visited as drawn, duplicates included) and npoints the points of one
    grid; the terms then carry the weight 1 / npoints^2."""
    nb, nfull, dims = disp.nb, mesh.nfull, [int(x) for x in mesh.dims]
    om1 = disp.om_irr[q1]
    E = lambda e, w: e / np.sqrt(np.where(w > 0, w, 1.0))[None, :]
    E1 = E(disp.egv_irr[q1], om1)
    w = 1.0 / (nfull if npoints is None else npoints)
    q1f = int(mesh.irr_full_index[q1])
    reclat6 = p.reclat / np.array(dims)[None, :]
    g = np.zeros(nb)
    od = np.zeros((nb, nfull, nb))
    if pairs is None:
        pairs = [(q2, q3) for q2 in range(nfull) for q3 in range(q2 if reduce else 0, nfull)]
    for (q2, q3) in pairs:
        if reduce and q3 < q2:
            continue
        q4 = fourth_grid_index(q1f, q2, q3, dims)
        if reduce:
            if q4 < q3:
                continue
            red, imgs = quartet_images(mesh, rot, q1, q2, q3, q4)
            if red:
                continue
            if q2 == q3 == q4:
                mult = (1.0, 1.0, 1.0, 1.0)
            elif q3 == q4 or q2 == q4 or q2 == q3:
                mult = (3.0, 1.0, 1.0, 1.0)
            else:
                mult = (6.0, 2.0, 2.0, 2.0)
        else:
            imgs = [(q2, q3, q4)]
            mult = (1.0, 3.0, 0.0, 0.0) if q2 == q3 == q4 else (1.0, 1.0, 0.0, 0.0)
        oms = [disp.om_full[q] for q in (q2, q3, q4)]
        Es = [E(disp.egv_full[q], o) for q, o in zip((q2, q3, q4), oms)]
        ptf = pretransform_phi4(fc4, mesh.qcart[q2], mesh.qcart[q3], mesh.qcart[q4], nb)
        psi = np.einsum("abcd,ai,bj,ck,dl->ijkl", ptf, E1, Es[0], Es[1], Es[2], optimize=True)
        psisq = FOURPHONON_PREFACTOR * np.abs(psi) ** 2 * w * w
        n = [be[mesh.irr_of_full[q]] for q in (q2, q3, q4)]
        n2, n3, n4 = n[0][None, :, None, None], n[1][None, None, :, None], n[2][None, None, None, :]
        n2p, n3p, n4p = n2 + 1.0, n3 + 1.0, n4 + 1.0
        plf0 = n2p * n3p * n4p - n2 * n3 * n4
        plf1 = 3.0 * n2 * n3p * n4p - n2p * n3 * n4
        plf2 = 3.0 * n3 * n2p * n4p - n3p * n2 * n4
        plf3 = 3.0 * n4 * n3p * n2p - n4p * n3 * n2
        w1 = om1[:, None, None, None]
        w2, w3, w4 = oms[0][None, :, None, None], oms[1][None, None, :, None], oms[2][None, None, None, :]
        if integrationtype == 2:
            sg = [sigsq[q1]] + [sigsq[mesh.irr_of_full[q]] for q in (q2, q3, q4)]
            sig = np.sqrt(sg[0][:, None, None, None] + sg[1][None, :, None, None] + sg[2][None, None, :, None]
                          + sg[3][None, None, None, :])
        elif integrationtype == 1:
            sig = np.full((nb,) * 4, sigma_fixed)
        else:
            sig = widths6(disp.vel_irr[q1], disp.vel_full[q2], disp.vel_full[q3], disp.vel_full[q4],
                          reclat6, thresh)
        d0, d1, d2, d3 = dirac4(w1, w2, w3, w4, sig)
        f = psisq * (mult[0] * plf0 * d0 + mult[1] * plf1 * d1 + mult[2] * plf2 * d2 + mult[3] * plf3 * d3)
        mask = (w2 >= U.FREQTOL) & (w3 >= U.FREQTOL) & (w4 >= U.FREQTOL) & (w1 >= U.FREQTOL)
        f = np.where(mask, f, 0.0)
        tot = f.sum(axis=(1, 2, 3))
        s2, s3, s4 = f.sum(axis=(2, 3)), f.sum(axis=(1, 3)), f.sum(axis=(1, 2))
        inv1 = np.where(om1 > 0, 1.0 / np.where(om1 > 0, om1, 1.0), 0.0)
        for (a, b, c) in imgs:
            g += tot
            od[:, a, :] += 2.0 * s2 * oms[0][None, :] * inv1[:, None]
            od[:, b, :] += 2.0 * s3 * oms[1][None, :] * inv1[:, None]
            od[:, c, :] += 2.0 * s4 * oms[2][None, :] * inv1[:, None]
    return g, od


def symmetrize_od(mesh, rot, q1, od, grid1=None):
    """symmetrize_and_distribute (scattering_fourphonon.f90:143-205) for every b1 of q1, over
    the points of the first Monte-Carlo grid in drawn order (all points on the full mesh)."""
    lg = mesh.littlegroup_irr[q1]
    for q2 in (range(mesh.nfull) if grid1 is None else grid1):
        imgs = rot[lg, q2]
        imgs = imgs[imgs >= 0]
        if np.any(imgs < q2):
            continue
        vals = od[:, imgs, :]                     # [b1, n_img, b]
        pos = vals > 0.0
        nn = pos.sum(axis=1)
        s = np.where(pos, vals, 0.0).sum(axis=1)
        avg = np.where(nn > 0, s / np.maximum(nn, 1), 0.0)
        for x in imgs:
            od[:, x, :] = np.where(nn > 0, avg, od[:, x, :])
    return od
