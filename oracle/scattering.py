"""Three-phonon scattering rates and scattering matrix, restated in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows src/thermal_conductivity/scattering.f90 (lo_scattering_rates%generate
:58-358) and scattering_threephonon.f90 (compute_threephonon_scattering :2-217,
get_dirac :219-271, pretransform_phi3 :275-317, triplet_is_irreducible
:319-448).  The per-row loop of the reference is kept, except that the |Psi|^2
of one (q1, q2, q3) triplet is formed for all b1 at once (the triplet set of a
row depends on q1 only); `psi_reference_form` keeps the reference's own
zgeru/dot_product form for small cross-checks.
"""
import numpy as np

from . import units as U

THREEPHONON_PREFACTOR = U.PI / 16.0  # scattering.f90:30


def third_grid_index(i1, i2, dims):
    """fft_third_grid_index (lo_fftgrid_helper.f90:136-163), 0-based indices."""
    n1, n2, n3 = dims
    g1 = np.array([i1 // (n2 * n3), (i1 // n3) % n2, i1 % n3])
    g2 = np.array([i2 // (n2 * n3), (i2 // n3) % n2, i2 % n3])
    g3 = (-g1 - g2) % np.array(dims)
    return int(g3[0] * n2 * n3 + g3[1] * n3 + g3[2])


def rotation_table(p, mesh, ops):
    """rot[o, q] = grid index of R_o q (reciprocal fractional), -1 if off grid."""
    from .lattice import _on_grid, _index_from_coordinate
    dims = mesh.dims
    frac = mesh.qcart @ p.inv_reclat.T
    rot = np.full((len(ops), mesh.nfull), -1, dtype=np.int64)
    for o, op in enumerate(ops):
        for q in range(mesh.nfull):
            v1 = op.rfm @ frac[q]
            if _on_grid(v1, dims):
                rot[o, q] = mesh.gi2ind(_index_from_coordinate(v1, dims))
    return rot


def irreducible_triplets(mesh, rot, q1):
    """Visited triplets of irreducible point q1 (compute_threephonon_scattering :72-84 with
    triplet_is_irreducible :319-448): list of (q2, q3, perm, red_triplet)."""
    lg = mesh.littlegroup_irr[q1]
    q1f = mesh.irr_full_index[q1]
    out = []
    for q2 in range(mesh.nfull):
        q3 = third_grid_index(q1f, q2, mesh.dims)
        if q3 < q2:
            continue
        newqp = []
        isred = False
        for k in lg:
            a = rot[k, q2]
            if a < 0:
                newqp.append(None)
                continue
            b = rot[k, q3]
            if b < 0:
                newqp.append(None)
                continue
            if min(a, b) > q2:
                isred = True
            newqp.append((a, b))
        if isred:
            continue
        newqp = [(q2, q3) if x is None else x for x in newqp]
        srt = [tuple(sorted(x)) for x in newqp]
        red = []
        n = len(newqp)
        for j in range(n):
            if not any(srt[j] == srt[k] for k in range(j + 1, n)):
                red.append(newqp[j])
        perm = 1.0 if q2 == q3 else 2.0
        out.append((q2, q3, perm, red))
    return out


def pretransform_phi3(fc3, q2, q3, nb):
    """pretransform_phi3 (scattering_threephonon.f90:275-317) as a (nb, nb, nb) array
    indexed [ia, ib, ic]."""
    ph = np.exp(-1j * U.TWOPI * (fc3["lv2"] @ q2 + fc3["lv3"] @ q3))
    ptf = np.zeros((nb, nb, nb), dtype=np.complex128)
    for t in range(len(ph)):
        a1, a2, a3 = fc3["a1"][t], fc3["a2"][t], fc3["a3"][t]
        ptf[3 * a1:3 * a1 + 3, 3 * a2:3 * a2 + 3, 3 * a3:3 * a3 + 3] += fc3["mwm"][t] * ph[t]
    return ptf


def psi_reference_form(egv1, egv2, egv3, ptf):
    """One matrix element exactly as scattering_threephonon.f90:96-112 forms it:
    evp1 = zgeru(egv2, egv1); evp2 = conj(zgeru(egv3, evp1)); c0 = dot_product(evp2, ptf)."""
    nb = len(egv1)
    evp1 = np.zeros(nb * nb, dtype=np.complex128)
    for j in range(nb):            # zgeru: A(i,j) += x(i) y(j), column-major
        evp1[j * nb:(j + 1) * nb] += egv2 * egv1[j]
    evp2 = np.zeros(nb ** 3, dtype=np.complex128)
    for j in range(nb * nb):
        evp2[j * nb:(j + 1) * nb] += egv3 * evp1[j]
    evp2 = np.conj(evp2)
    # ptf flat index l = ia*nb^2 + ib*nb + ic ; evp2 flat index = ic + nb*(ib + nb*ia)
    flat = ptf.reshape(-1)
    return np.sum(np.conj(evp2) * flat)   # dot_product conjugates its first argument


def psi_chain(E1, E2, E3, ptf):
    """Psi[b1,b2,b3] = sum E1[ia,b1] E2[ib,b2] E3[ic,b3] ptf[ia,ib,ic] as three GEMM steps."""
    t1 = np.tensordot(E1, ptf, axes=([0], [0]))          # [b1, ib, ic]
    t2 = np.einsum("ib,aic->abc", E2, t1, optimize=True)  # [b1, b2, ic]
    return np.einsum("abc,cd->abd", t2, E3, optimize=True)


class ScatteringRates:
    pass


def occupations_and_smearing(p, mesh, disp, temperature, classical=False):
    """Bose-Einstein occupations and adaptive smearing widths (scattering.f90:131-167)."""
    nb, nirr, dims = disp.nb, mesh.nirr, mesh.dims
    be = np.zeros((nirr, nb))
    sigsq = np.zeros((nirr, nb))
    sigavg = np.zeros(nb)
    reclat = p.reclat / np.array(dims)[None, :]
    for q in range(nirr):
        for b in range(nb):
            om = disp.om_irr[q, b]
            if om < U.FREQTOL:
                continue
            if classical:
                be[q, b] = U.KB_HARTREE * temperature / om
            else:
                be[q, b] = U.planck(temperature, om)
            v0 = (np.abs(disp.vel_irr[q, b]) @ reclat) ** 2
            s = np.max(v0 * 0.5)
            sigsq[q, b] = s
            sigavg[b] += s * mesh.weight_irr[q]
    sigsq += np.max(sigavg) * 1e-4
    return be, sigsq, sigavg


def velocity_difference_widths(v1, v2, v3, reclat, thresh):
    """get_dirac integration type 6 (scattering_threephonon.f90:251-261): for the group
    velocities v1 [n1, 3], v2 [n2, 3], v3 [n3, 3] the widths [n1, n2, n3], the sum over the
    three pairs of sqrt(max_c ((v_a - v_b) reclat)_c^2 / 2) / 3, plus thresh; reclat is the
    reciprocal basis with column j divided by the mesh dimension j (scattering.f90:138-140)."""
    def pw(a, b):
        return np.sqrt(np.max(((a - b) @ reclat) ** 2, axis=-1) * 0.5)
    s = pw(v1[:, None, :], v2[None, :, :])[:, :, None] / 3.0
    s = s + pw(v1[:, None, :], v3[None, :, :])[:, None, :] / 3.0
    s = s + pw(v2[:, None, :], v3[None, :, :])[None, :, :] / 3.0
    return s + thresh


ISOTOPE_PREFACTOR = U.PI / 4.0  # scattering.f90:29


def isotope_strength(e1, e2, g):
    """isotope_scattering_strength (scattering_isotope.f90:72-92): sum over atoms of
    g_a (sum_x |conj(e1_ax) e2_ax|)^2.  e1: [component]; e2: [..., component]."""
    na = len(g)
    e2a = e2.reshape(e2.shape[:-1] + (na, 3))             # [..., atom, xyz]
    e1a = np.conj(e1).reshape(na, 3)
    f0 = np.abs(e1a * e2a).sum(axis=-1) ** 2              # [..., atom]
    return f0 @ np.asarray(g, float)


def isotope_row(mesh, disp, sigsq, q1, b1, g, integrationtype=2, sigma_fixed=0.0, reclat=None, thresh=0.0):
    """compute_isotope_scattering (scattering_isotope.f90:2-70) for one row (q1, b1):
    returns the linewidth term and the row of scattering-matrix terms [q2, b2].  Integration
    type 6 (:52-54) needs reclat (columns divided by the mesh dimensions) and thresh."""
    nb, nfull = disp.nb, mesh.nfull
    om1 = disp.om_irr[q1, b1]
    e1 = disp.egv_irr[q1][:, b1]
    gam = 0.0
    xi = np.zeros((nfull, nb))
    for q2 in range(nfull):
        om2 = disp.om_full[q2]
        e2 = disp.egv_full[q2]                      # [component, mode]
        st = isotope_strength(e1, e2.T, g)          # [mode]
        i2 = mesh.irr_of_full[q2]
        for b2 in range(nb):
            if om2[b2] < U.FREQTOL:
                continue
            if integrationtype == 2:
                sig = np.sqrt(sigsq[q1, b1] + sigsq[i2, b2])
            elif integrationtype == 6:
                sig = np.sqrt(np.max((disp.vel_full[q2, b2] @ reclat) ** 2) * 0.5) + thresh
            else:
                sig = sigma_fixed
            psisq = st[b2] * ISOTOPE_PREFACTOR / nfull
            f0 = psisq * om1 * om2[b2] * U.gauss(om1, om2[b2], sig)
            gam += f0
            xi[q2, b2] += f0 * om2[b2] / om1
    return gam, xi


def scattering_rates(p, fc3, mesh, disp, ops, temperature, classical=False,
                     integrationtype=2, sigma=1.0, rows_only=None, isotope_g=None, mfp_max=None,
                     qg3ph=None, seed=None, fc4=None, qg4ph=None, thirdorder=True):
    """lo_scattering_rates%generate (scattering.f90:58-358): three-phonon scattering with
    qg3ph = qgrid (the Monte-Carlo grid is then the identity with weight 1/N_q), plus
    isotope scattering (:229-232) when per-atom mass-disorder parameters are given, plus
    boundary scattering |v| / mfp_max (Bohr, :245-250) when mfp_max > 0.  With qg3ph smaller
    than the mesh, every row (q1, b1) visits the triplets of its own Monte-Carlo q2 draw
    (compute_threephonon_scattering :69-78, montecarlo.row_grids), each draw weighted
    1 / prod(qg3ph).  With fourth-order force constants fc4 (lattice.read_fc4) the
    four-phonon term of compute_fourphonon_scattering is added to every row (qg4ph = qgrid,
    oracle/fourphonon.py).  thirdorder=False leaves the three-phonon term out (--nothirdorder,
    scattering.f90:234-238; full q-grids only here)."""
    if not thirdorder and (qg3ph is not None or qg4ph is not None):
        raise NotImplementedError("the oracle leaves the three-phonon term out on full q-grids only")
    nb = disp.nb
    nirr, nfull = mesh.nirr, mesh.nfull
    sr = ScatteringRates()
    be, sigsq, sigavg = occupations_and_smearing(p, mesh, disp, temperature, classical)
    sr.be, sr.sigsq = be, sigsq
    sr.sigma_fixed = sigma * U.FREQ_THZ_TO_HARTREE
    sr.thresh_sigma = np.max(sigavg) * 1e-4
    reclat6 = p.reclat / np.array(mesh.dims)[None, :]
    rows = [(q, b) for q in range(nirr) for b in range(nb) if disp.om_irr[q, b] >= U.FREQTOL]
    sr.rows = rows
    row_of = {r: i for i, r in enumerate(rows)}
    nrow = len(rows)
    gamma_row = np.zeros(nrow)
    Xi = np.zeros((nrow, nfull * nb))
    rot = rotation_table(p, mesh, ops)
    sr.rot = rot
    mcw = 1.0 / nfull
    grids = None
    grids4 = None
    from . import montecarlo as MC
    mc3 = MC.mc_dims_of(list(mesh.dims), list(qg3ph)) if qg3ph is not None else list(mesh.dims)
    mc4 = MC.mc_dims_of(list(mesh.dims), list(qg4ph)) if qg4ph is not None else list(mesh.dims)
    if fc4 is not None and (mc3 != list(mesh.dims) or mc4 != list(mesh.dims)):
        # with four-phonon scattering every row draws its three grids from one generator
        allg, npts, npts4 = MC.row_grids_4ph(list(mesh.dims), rows, mc3, mc4, seed)
        grids = {r: g[0] for r, g in allg.items()}
        mcw = 1.0 / npts
        if mc4 != list(mesh.dims):
            grids4 = {r: (g[1], g[2], npts4) for r, g in allg.items()}
    elif mc3 != list(mesh.dims):
        grids, npts = MC.row_grids(list(mesh.dims), rows, mc3, seed)
        mcw = 1.0 / npts
    ntrip = 0
    for q1 in range(nirr):
        myb = [b for b in range(nb) if (q1, b) in row_of]
        if rows_only is not None:
            myb = [b for b in myb if row_of[(q1, b)] in rows_only]
        if not myb:
            continue
        om1 = disp.om_irr[q1]
        E1 = disp.egv_irr[q1] / np.sqrt(np.where(om1 > 0, om1, 1.0))[None, :]
        od = np.zeros((nb, nfull, nb))   # [b1, q, b]
        trips = irreducible_triplets(mesh, rot, q1) if thirdorder else []
        ntrip += len(trips)
        for (q2, q3, perm, red) in trips:
            if grids is not None:
                # draw count of q2 for every row of q1 (zero: the row does not visit it)
                cnt = np.array([np.count_nonzero(grids[(q1, b)] == q2) if (q1, b) in grids else 0
                                for b in range(nb)], dtype=float)
                if not np.any(cnt[myb]):
                    continue
            om2 = disp.om_full[q2]; om3 = disp.om_full[q3]
            E2 = disp.egv_full[q2] / np.sqrt(np.where(om2 > 0, om2, 1.0))[None, :]
            E3 = disp.egv_full[q3] / np.sqrt(np.where(om3 > 0, om3, 1.0))[None, :]
            ptf = pretransform_phi3(fc3, mesh.qcart[q2], mesh.qcart[q3], nb)
            psi = psi_chain(E1, E2, E3, ptf)
            psisq = THREEPHONON_PREFACTOR * np.abs(psi) ** 2 * mcw
            i2, i3 = mesh.irr_of_full[q2], mesh.irr_of_full[q3]
            n2 = be[i2][None, :, None]; n3 = be[i3][None, None, :]
            w1 = om1[:, None, None]; w2 = om2[None, :, None]; w3 = om3[None, None, :]
            if integrationtype == 2:
                sig = np.sqrt(sigsq[q1][:, None, None] + sigsq[i2][None, :, None]
                              + sigsq[i3][None, None, :])
            elif integrationtype == 1:
                sig = np.full((nb, nb, nb), sr.sigma_fixed)
            else:
                sig = velocity_difference_widths(disp.vel_irr[q1], disp.vel_full[q2], disp.vel_full[q3],
                                                 reclat6, sr.thresh_sigma)
            d0 = U.gauss(w1, -w2 + w3, sig) - U.gauss(w1, w2 - w3, sig)
            d1 = U.gauss(w1, w2 + w3, sig) - U.gauss(w1, -w2 - w3, sig)
            f = perm * psisq * ((n2 - n3) * d0 + (n2 + n3 + 1.0) * d1)
            if grids is not None:
                f = f * cnt[:, None, None]
            mask = ((w2 >= U.FREQTOL) & (np.abs(w1 - w2) >= U.FREQTOL)
                    & (w3 >= U.FREQTOL) & (np.abs(w1 - w3) >= U.FREQTOL)
                    & (np.abs(w2 - w3) >= U.FREQTOL))
            f = np.where(mask, f, 0.0)
            tot = f.sum(axis=(1, 2))
            s2 = f.sum(axis=2)   # [b1, b2]
            s3 = f.sum(axis=1)   # [b1, b3]
            for b1 in myb:
                gamma_row[row_of[(q1, b1)]] += tot[b1] * len(red)
                for (q2p, q3p) in red:
                    od[b1, q2p, :] += 2.0 * s2[b1] * om2 / om1[b1]
                    od[b1, q3p, :] += 2.0 * s3[b1] * om3 / om1[b1]
        # symmetrize_and_distribute (:143-205)
        lg = mesh.littlegroup_irr[q1]
        for q2 in range(nfull):
            imgs = rot[lg, q2]
            imgs = imgs[imgs >= 0]
            if np.any(imgs < q2):
                continue
            for b1 in myb:
                vals = od[b1][imgs]              # [n_img, nb]
                pos = vals > 0.0
                nn = pos.sum(axis=0)
                s = np.where(pos, vals, 0.0).sum(axis=0)
                upd = nn > 0
                avg = np.where(upd, s / np.maximum(nn, 1), 0.0)
                for x in imgs:
                    od[b1, x, upd] = avg[upd]
        for b1 in myb:
            Xi[row_of[(q1, b1)]] += od[b1].reshape(-1)
        if fc4 is not None:
            from . import fourphonon as F4
            if grids4 is None:
                g4, od4 = F4.row_terms(p, fc4, mesh, disp, be, sigsq, q1, integrationtype, sr.sigma_fixed,
                                       sr.thresh_sigma, rot)
                od4 = F4.symmetrize_od(mesh, rot, q1, od4)
                for b1 in myb:
                    gamma_row[row_of[(q1, b1)]] += g4[b1]
                    Xi[row_of[(q1, b1)]] += od4[b1].reshape(-1)
            else:
                for b1 in myb:
                    ga, gb, n4 = grids4[(q1, b1)]
                    pairs = [(int(x), int(y)) for x in ga for y in gb]
                    g4, od4 = F4.row_terms(p, fc4, mesh, disp, be, sigsq, q1, integrationtype, sr.sigma_fixed,
                                           sr.thresh_sigma, rot, pairs=pairs, npoints=n4)
                    od4 = F4.symmetrize_od(mesh, rot, q1, od4, grid1=[int(x) for x in ga])
                    gamma_row[row_of[(q1, b1)]] += g4[b1]
                    Xi[row_of[(q1, b1)]] += od4[b1].reshape(-1)
    sr.ntriplets = ntrip
    if isotope_g is not None:
        for i, (q1, b1) in enumerate(rows):
            if rows_only is not None and i not in rows_only:
                continue
            gi, xi = isotope_row(mesh, disp, sigsq, q1, b1, isotope_g, integrationtype, sr.sigma_fixed,
                                 reclat6, sr.thresh_sigma)
            gamma_row[i] += gi
            Xi[i] += xi.reshape(-1)
    if mfp_max is not None and mfp_max > 0.0:
        for i, (q1, b1) in enumerate(rows):
            v = disp.vel_irr[q1, b1]
            vn = np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
            if vn > U.GROUPVELTOL:
                gamma_row[i] += vn / mfp_max
    # reduce + degeneracy average + F0 (:260-292)
    lw = np.zeros((nirr, nb))
    for i, (q, b) in enumerate(rows):
        lw[q, b] = gamma_row[i]
    for q in range(nirr):
        for b in range(nb):
            if disp.om_irr[q, b] < U.FREQTOL:
                continue
            dm = disp.degen_irr[q][b]
            f0 = lw[q, dm].sum() / len(dm)
            lw[q, dm] = f0
    linewidth = np.zeros((nirr, nb)); qs = np.zeros((nirr, nb))
    F0 = np.zeros((nirr, nb, 3)); mfp = np.zeros((nirr, nb, 3)); smfp = np.zeros((nirr, nb))
    for q in range(nirr):
        for b in range(nb):
            if disp.om_irr[q, b] < U.FREQTOL:
                continue
            linewidth[q, b] = lw[q, b]
            qs[q, b] = 2.0 * linewidth[q, b]
            if linewidth[q, b] > U.FREQTOL:
                v = disp.vel_irr[q, b]
                mfp[q, b] = v / qs[q, b]
                smfp[q, b] = np.linalg.norm(v) / qs[q, b]
                F0[q, b] = mfp[q, b]
    # symmetrize_scatmat (:296-357)
    for i, (q1, b1) in enumerate(rows):
        lg = mesh.littlegroup_irr[q1]
        row = Xi[i].reshape(nfull, nb)
        for q2 in range(nfull):
            imgs = rot[lg, q2]
            imgs = imgs[imgs >= 0]
            if np.any(imgs < q2) or len(imgs) == 0:
                continue
            buf = row[imgs].sum(axis=0) / len(imgs)
            row[imgs] = buf[None, :]
    sr.Xi = Xi
    sr.linewidth, sr.qs, sr.F0, sr.mfp, sr.scalar_mfp = linewidth, qs, F0, mfp, smfp
    sr.gamma_row = gamma_row
    return sr
