"""Cumulative, boundary and spectral thermal conductivity and the phonon density of states of
outfile.cumulative_thermal_conductivity.hdf5 (test infrastructure, see oracle/__init__.py).

Restates, for one temperature:
  lo_cumulative_kappa%get_cumulative_kappa   cumulative_kappa.f90:58-414
  lo_cumulative_kappa%get_boundary_kappa     cumulative_kappa.f90:416-508
  lo_cumulative_kappa%get_spectral_kappa     cumulative_kappa.f90:510-561 and
      lo_phonon_dos%spectral_kappa           type_phonon_dos_integrations.f90:428-809
  lo_phonon_dos%generate                     type_phonon_dos.f90:110-297 with
      lo_get_phonon_dos_tetrahedron          type_phonon_dos_integrations.f90:812-924
      lo_get_phonon_dos_gaussian             type_phonon_dos_integrations.f90:927-1021
  the irreducible tetrahedra of the FFT mesh type_qpointmesh_gridgeneration.f90:700-1260,
      1342-1470
  lo_LV_tetrahedron_weights                  type_qpointmesh_integrationweights.f90:148-262
  phonon_angular_momentum_matrix             type_phonon_dispersions.f90:868-971
  spectral_angular_momentum                  type_phonon_dos_integrations.f90:8-425
and the helpers they call (gottochblandat_calculus.f90).  Inputs are the dispersions, the
final F vectors (before the omega/T prefactor of main.f90:323-328), the linewidths and the
symmetrised off-diagonal kappa, all in atomic units.  Two places where the reference reads a
variable before it is ever set are given a value here, and the same value in the engine: the
boundary-scattering rate of cumulative_kappa.f90:464-467 starts at zero, and the full-grid
per-mode kappa of modes below lo_freqtol (cumulative_kappa.f90:538-553) is zero.

spectral_angular_momentum: parity unpinned.  The tensor vanishes by symmetry in every test
crystal, so the spectrum's normalisation is a ratio of rounding noise and no closed form,
invariant or brute force checks it beyond what it shares with spectral_kappa.
"""
import math

import numpy as np
from scipy.special import erf

from . import units as U

KAPPA_SI_TO_AU = 1.0 / U.KAPPA_AU_TO_SI
MFPTOL = 1e-12 / U.GROUPVEL_HARTREEBOHR_TO_MS         # 1e-12 m/s in atomic units (cumulative_kappa.f90:167)


# ---------------------------------------------------------------- helpers (gottochblandat)

def gauss(x, mu, sigma):
    """lo_gauss (gottochblandat_calculus.f90:268-287)."""
    if abs(x - mu) > 8 * sigma:
        return 0.0
    inv = 1.0 / sigma
    return math.exp(-(x - mu) * (x - mu) * 0.5 * inv * inv) * 0.398942280401433 * inv


def linspace(a, b, n):
    """lo_linspace (gottochblandat_calculus.f90:188-214)."""
    if n == 1:
        return np.array([(a + b) * 0.5])
    f1 = b - a
    x = np.array([a + ((i + 1) * 1.0 - 1.0) / (n * 1.0 - 1.0) * f1 for i in range(n)])
    x[0], x[-1] = a, b
    return x


def logspace(a, b, n):
    """lo_logspace (gottochblandat_calculus.f90:217-249)."""
    if n == 1:
        return np.array([(a + b) * 0.5])
    shift = 1.0 - a if a < U.TOL else 0.0
    d = linspace(math.log(a + shift), math.log(b + shift), n)
    return np.array([math.exp(v) - shift for v in d])


def trapezoid(x, y):
    """lo_trapezoid_integration (gottochblandat_calculus.f90:252-265), summed in order."""
    f = 0.0
    for i in range(len(x) - 1):
        f += (y[i] + y[i + 1]) * 0.5 * (x[i + 1] - x[i])
    return f


def linear_interpolation(x, y, xi, thres=U.SQTOL):
    """lo_linear_interpolation (gottochblandat_calculus.f90:75-151)."""
    n = len(x)
    if abs(xi - x[0]) < thres:
        return y[0]
    if abs(xi - x[n - 1]) < thres:
        return y[n - 1]
    if np.min(x) - xi > 2 * thres or abs(np.max(x)) - xi < -2 * thres:
        return 0.0
    # first i (1-based, from 2) with x(i) - xi > thres; x is sorted, so the predicate is
    # monotone and a bisection finds what the reference's scan finds
    lo, hi = 2, n + 1
    while lo < hi:
        mid = (lo + hi) // 2
        if x[mid - 1] - xi > thres:
            hi = mid
        else:
            lo = mid
            lo += 1
    if lo <= n:
        ii, jj = lo, lo - 1
    else:
        ii, jj = n, 0
    if ii < 2:
        ii, jj = 2, 1
    if ii >= n:
        ii, jj = n, n - 1
    A = (x[ii - 1] - xi) / (x[ii - 1] - x[jj - 1])
    return A * y[jj - 1] + (1.0 - A) * y[ii - 1]


def atom_classes(ops, na):
    """sym%degeneracy and sym%degenerate_atom (type_symmetryoperation.f90:651-676): the atoms
    any operation maps onto each other, in increasing order."""
    di = np.zeros((na, na), dtype=bool)
    for o in ops:
        for i in range(na):
            j = o.fmap[i]
            di[i, j] = di[j, i] = True
    deg = [int(di[:, i].sum()) for i in range(na)]
    members = [[j for j in range(na) if di[j, i]] for i in range(na)]
    return deg, members


def default_smearing(om_irr):
    """dr%default_smearing, omega_max, omega_min (type_phonon_dispersions_generation.f90:
    195-233)."""
    nirr, nb = om_irr.shape
    ds = np.zeros(nb)
    for b in range(nb):
        x = np.sort(om_irr[:, b])
        f0 = 0.0
        for i in range(nirr - 1):
            f0 = max(f0, abs(x[i + 1] - x[i]))
        ds[b] = f0
    ds = np.maximum(ds, ds.max() / 5.0)
    omax = float(om_irr.max())
    nz = om_irr[np.abs(om_irr) > U.FREQTOL]
    omin = float(nz.min()) if nz.size else np.inf
    return ds, omax, omin


def site_projection(egv, na, mode):
    """|e_a|^2 of every atom a of one mode (egv[component, mode])."""
    return np.array([float(np.sum(np.abs(egv[3 * a:3 * a + 3, mode]) ** 2)) for a in range(na)])


def cv_of(temperature, om, classical):
    return U.KB_HARTREE if classical else float(U.harmonic_oscillator_cv(temperature, om))


def rotate_vec(ops, k, v):
    """lo_operate_on_vector of a Cartesian vector, k signed 1-based (negative: time reversal)."""
    m = ops[abs(k) - 1].m
    return (m @ v) if k > 0 else -(m @ v)


# ---------------------------------------------------------------- tetrahedra of the FFT mesh

def lv_tetrahedron_weights(ein, z, tol, sigma):
    """lo_LV_tetrahedron_weights (type_qpointmesh_integrationweights.f90:148-262): the
    Lambin-Vigneron (PRB 29, 3430 (1984)) weights of the four corners of a linear
    tetrahedron for the density of states at z."""
    ind = np.argsort(ein, kind="stable")
    ei = np.asarray(ein)[ind]
    w = np.zeros(4)
    if z <= ei[0] or z >= ei[3]:
        return w
    EZ1, EZ2, EZ3, EZ4 = z - ei[0], z - ei[1], z - ei[2], z - ei[3]
    wt = np.zeros(4)
    if ei[3] - ei[0] < tol:
        for i, ez in enumerate((EZ1, EZ2, EZ3, EZ4)):
            wt[i] = gauss(ez, 0.0, sigma) * 0.25
    elif z < ei[1]:
        a, b, c = 1.0 / (ei[1] - ei[0]), 1.0 / (ei[2] - ei[0]), 1.0 / (ei[3] - ei[0])
        wt[0] = a * b * c * EZ1 ** 2 * (-a * EZ2 - b * EZ3 - c * EZ4)
        wt[1] = a ** 2 * b * c * EZ1 ** 3
        wt[2] = a * b ** 2 * c * EZ1 ** 3
        wt[3] = a * b * c ** 2 * EZ1 ** 3
    elif z < ei[2]:
        b, c = 1.0 / (ei[2] - ei[0]), 1.0 / (ei[3] - ei[0])
        d, e = 1.0 / (ei[2] - ei[1]), 1.0 / (ei[3] - ei[1])
        ff0, ff2 = -b ** 2 * EZ3, -c ** 2 * EZ4
        gg0, gg2 = -d ** 2 * EZ3, -e ** 2 * EZ4
        hh0, hh2 = d ** 2 * EZ2, b ** 2 * EZ1
        ii0, ii2 = e ** 2 * EZ2, c ** 2 * EZ1
        ff1 = -c * d * EZ1 * EZ3 - d * e * EZ2 * EZ3 - c * e * EZ1 * EZ4
        ff3 = -b * d * EZ1 * EZ3 - b * e * EZ1 * EZ4 - d * e * EZ2 * EZ4
        gg1 = -b * c * EZ1 * EZ3 - b * e * EZ2 * EZ3 - c * e * EZ2 * EZ4
        gg3 = -b * d * EZ2 * EZ3 - b * c * EZ1 * EZ4 - c * d * EZ2 * EZ4
        hh1 = -b * c * EZ1 * EZ3 - b * e * EZ2 * EZ3 - c * e * EZ2 * EZ4
        hh3 = -c * d * EZ1 * EZ3 - d * e * EZ2 * EZ3 - c * e * EZ1 * EZ4
        ii1 = -b * d * EZ2 * EZ3 - b * c * EZ1 * EZ4 - c * d * EZ2 * EZ4
        ii3 = -b * d * EZ1 * EZ3 - b * e * EZ1 * EZ4 - d * e * EZ2 * EZ4
        wt[0] = 0.5 * (ff0 * ff1 + ff2 * ff3)
        wt[1] = 0.5 * (gg0 * gg1 + gg2 * gg3)
        wt[2] = 0.5 * (hh0 * hh1 + hh2 * hh3)
        wt[3] = 0.5 * (ii0 * ii1 + ii2 * ii3)
    else:
        c, e, f = 1.0 / (ei[3] - ei[0]), 1.0 / (ei[3] - ei[1]), 1.0 / (ei[3] - ei[2])
        wt[0] = -(c ** 2 * e * EZ4 ** 3 * f)
        wt[1] = -(c * e ** 2 * EZ4 ** 3 * f)
        wt[2] = -(c * e * EZ4 ** 3 * f ** 2)
        wt[3] = c * e * EZ4 ** 2 * f * (c * EZ1 + e * EZ2 + EZ3 * f)
    w[ind] = wt
    return w


def _clean_frac(x):
    """lo_clean_fractional_coordinates (gottochblandat.f90:1261-1283), default tolerance."""
    y = x
    while True:
        if abs(y) < U.SQTOL:
            y = 0.0
        if abs(y - 1.0) < U.SQTOL:
            y = 0.0
        if y > 1.0:
            y = y - 1.0
        if y < 0.0:
            y = y + 1.0
        if -U.SQTOL < y < 1.0:
            return y


def cell_center(gi, dims):
    """Centre of the box above grid point gi (0-based), type_qpointmesh_gridgeneration.f90:
    729-735 (chop, then clean, twice)."""
    out = []
    for d in range(3):
        v = (gi[d] + 1 - 0.5) / float(dims[d])
        for _ in range(2):
            v = _clean_frac(float(U.chop(np.array([v]), 1e-11)[0]))
        out.append(v)
    return out


def chop_box(reclat, dims, center):
    """neighbour_corners_from_center + chop_box_into_tetrahedrons (type_qpointmesh_
    gridgeneration.f90:1342-1470): the eight corners of the box (x slowest, z fastest) as
    grid offsets, and the six tetrahedra around the shortest main diagonal as corner
    numbers.  The diagonal is chosen with the reference's arithmetic (first of equal
    lengths in corner order); the six tetrahedra are consecutive pairs of the other corners
    taken around that diagonal."""
    half = [0.5 / dims[d] for d in range(3)]
    box, off = [], []
    for i in (0, 1):
        for j in (0, 1):
            for k in (0, 1):
                box.append([center[0] + (half[0] if i else -half[0]),
                            center[1] + (half[1] if j else -half[1]),
                            center[2] + (half[2] if k else -half[2])])
                off.append((i, j, k))
    B = reclat
    r = []
    for c in box:
        v1 = [c[d] - box[0][d] for d in range(3)]
        r.append([B[x, 0] * v1[0] + B[x, 1] * v1[1] + B[x, 2] * v1[2] for x in range(3)])
    com = [0.0, 0.0, 0.0]
    for c in r:
        for x in range(3):
            com[x] = com[x] + c[x] / 8.0
    f2 = np.inf
    diag = None
    for i in range(8):
        for j in range(i + 1, 8):
            v0 = [r[i][x] - com[x] for x in range(3)]
            v1 = [r[j][x] - com[x] for x in range(3)]
            d01 = v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2]
            d00 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]
            f0 = abs(d01 + d00) + abs(d01 + d00)
            if f0 < 1e-10:
                dv = [r[j][x] - r[i][x] for x in range(3)]
                f1 = math.sqrt(dv[0] ** 2 + dv[1] ** 2 + dv[2] ** 2)
                if f1 < f2:
                    f2 = f1
                    diag = (i, j)
    other = [i for i in range(8) if i not in diag]
    # angles of the other corners around the diagonal
    n = np.array(r[diag[0]]) - np.array(r[diag[1]])
    n = n / np.linalg.norm(n)
    u = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = u - (u @ n) * n
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(n, e1)
    ang = []
    for i in other:
        v = np.array(r[i]) - np.array(com)
        ang.append(math.atan2(v @ e2, v @ e1))
    order = [other[i] for i in np.argsort(ang, kind="stable")]
    tets = [(diag[0], diag[1], order[i], order[(i + 1) % 6]) for i in range(6)]
    return off, tets


def irreducible_tetrahedra(p, ops, mesh, timereversal=True):
    """The irreducible tetrahedra of the FFT mesh (type_qpointmesh_gridgeneration.f90:
    700-1260): boxes above grid points reduced by the point group (and time reversal) as
    the q-points are, the lowest-numbered box of each class as its prototype, whose six
    tetrahedra carry the number of boxes in the class as integration weight.  Returns
    (corners [ntet, 4] full-grid indices, weights [ntet])."""
    dims = [int(d) for d in mesh.dims]
    n1, n2, n3 = dims
    npt = n1 * n2 * n3
    gis = [(i, j, k) for i in range(n1) for j in range(n2) for k in range(n3)]
    centers = [cell_center(g, dims) for g in gis]
    nsym = len(ops)
    ncol = 2 * nsym if timereversal else nsym
    conn = np.zeros((npt, ncol), dtype=np.int64)
    for ipt in range(npt):
        v0 = np.array(centers[ipt])
        for tr in ((False, True) if timereversal else (False,)):
            for iop, o in enumerate(ops):
                v1 = (-1.0 if tr else 1.0) * (o.rfm @ v0)
                w0 = (1.0 + 2.0 * np.array(dims) * v1) * 0.5
                if np.abs(w0 - np.round(w0)).sum() > U.SQTOL:
                    continue
                g = [int(np.round(w0[d])) for d in range(3)]          # 1-based, periodic
                g = [(g[d] - 1) % dims[d] for d in range(3)]
                conn[ipt, iop + (nsym if tr else 0)] = (g[0] * n2 + g[1]) * n3 + g[2] + 1
    keep = np.ones(npt, dtype=bool)
    for ipt in range(npt):
        for c in conn[ipt]:
            j = abs(int(c)) - 1
            if j > ipt:
                keep[j] = False
    corners, weights = [], []
    for ipt in np.nonzero(keep)[0]:
        members = []
        for c in conn[ipt]:
            if c != 0 and int(c) not in members:
                members.append(int(c))
        off, tets = chop_box(p.reclat, dims, centers[ipt])
        g = gis[ipt]
        cidx = [(((g[0] + o[0]) % n1) * n2 + (g[1] + o[1]) % n2) * n3 + (g[2] + o[2]) % n3 for o in off]
        for t in tets:
            corners.append([cidx[c] for c in t])
            weights.append(float(len(members)))
    return np.array(corners, dtype=np.int64), np.array(weights)


# ---------------------------------------------------------------- density of states

class PhononDos:
    pass


def phonon_dos(p, ops, mesh, disp, n_dos_point=1000, integrationtype=3, sigma=1.0, tets=None):
    """lo_phonon_dos%generate (type_phonon_dos.f90:110-297)."""
    na, nb = p.na, disp.nb
    ds, omax, omin = default_smearing(disp.om_irr)
    d = PhononDos()
    d.integrationtype, d.n = integrationtype, n_dos_point
    dosmin = float(disp.om_irr.min())
    if dosmin > -U.FREQTOL:
        dosmin = 0.0
    dosmax = omax + 4.01 * ds.max()
    om = linspace(dosmin, dosmax, n_dos_point)
    pm = np.zeros((n_dos_point, nb))
    ps = np.zeros((n_dos_point, na))
    tmax, tmin = om.max(), om.min()
    invf = n_dos_point / (tmax - tmin)
    if integrationtype == 3:
        corners, weights = tets if tets is not None else irreducible_tetrahedra(p, ops, mesh)
        deltae = (om[1] - om[0]) * 0.5
        for it in range(len(weights)):
            irr = [int(mesh.irr_of_full[c]) for c in corners[it]]
            for j in range(nb):
                sg = ds[j]
                tete = np.array([disp.om_irr[q, j] for q in irr])
                proj = np.zeros(na)
                for q in irr:
                    proj += site_projection(disp.egv_irr[q], na, j)
                ii = max(math.floor((tete.min() - tmin) * invf), 1)
                jj = min(math.ceil((tete.max() - tmin) * invf), n_dos_point)
                for k in range(ii, jj + 1):
                    e = om[k - 1]
                    w1 = lv_tetrahedron_weights(tete, e - deltae, U.FREQTOL, sg)
                    w2 = lv_tetrahedron_weights(tete, e + deltae, U.FREQTOL, sg)
                    f0 = np.sum(w1 + w2) * weights[it] * 0.5
                    pm[k - 1, j] += f0
                    ps[k - 1] += proj * f0
    else:
        radius = (3.0 / p.volume / float(mesh.nfull) / 4.0 / math.pi) ** (1.0 / 3.0)
        for q in range(mesh.nirr):
            wq = mesh.weight_irr[q]
            for j in range(nb):
                f1 = disp.om_irr[q, j]
                if f1 < omin * 0.5:
                    continue
                if integrationtype == 1:
                    sg = ds[j] * sigma
                else:
                    sg = sigma * (2 * math.pi / math.sqrt(2.0)) * radius * np.linalg.norm(disp.vel_irr[q, j])
                    sg = max(ds[j] * 0.25 * sigma, sg)
                    sg = min(ds[j] * 4.0 * sigma, sg)
                proj = site_projection(disp.egv_irr[q], na, j)
                ii = max(math.floor((f1 - tmin - 4 * sg) * invf), 1)
                jj = min(math.floor((f1 - tmin + 4 * sg) * invf), n_dos_point)
                for k in range(ii, jj + 1):
                    dy = gauss(f1, om[k - 1], sg)
                    pm[k - 1, j] += dy * wq
                    ps[k - 1] += dy * proj
    # makenice
    for j in range(nb):
        f1 = pm[0, j]
        for i in range(1, n_dos_point + 1):
            f0 = ((n_dos_point - i) * 1.0 / ((n_dos_point - 1) * 1.0)) ** 2
            pm[i - 1, j] = pm[i - 1, j] - f0 * f1
            if pm[i - 1, j] < U.FREQTOL:
                pm[i - 1, j] = 0.0
        pm[0, j] = 0.0
    for j in range(na):
        f1 = ps[0, j]
        for i in range(1, n_dos_point + 1):
            f0 = ((n_dos_point - i) * 1.0 / ((n_dos_point - 1) * 1.0)) ** 2
            ps[i - 1, j] = ps[i - 1, j] - f0 * f1
            if ps[i - 1, j] < U.FREQTOL:
                ps[i - 1, j] = 0.0
        ps[0, j] = 0.0
    dos = np.zeros(n_dos_point)
    for j in range(nb):
        f0 = 1.0 / trapezoid(om, pm[:, j])
        dos = dos + pm[:, j]
        pm[:, j] = pm[:, j] * f0
    dos[0] = 0.0
    dostol = float(np.mean(dos)) * U.SQTOL
    dos = U.chop(dos, dostol)
    f0 = trapezoid(om, dos)
    dos = dos * nb / f0
    for i in range(n_dos_point):
        f0 = pm[i].sum()
        pm[i] = pm[i] * dos[i] / f0 if f0 > dostol else 0.0
    deg, mem_ = atom_classes(ops, na)
    sb = np.zeros((n_dos_point, na))
    for i in range(na):
        for j in range(deg[i]):
            sb[:, i] = sb[:, i] + ps[:, mem_[i][j]] / (1.0 * deg[j])
        f0 = trapezoid(om, sb[:, i])
        sb[:, i] = sb[:, i] * 3.0 / f0
    ps = sb
    for i in range(n_dos_point):
        f0 = ps[i].sum()
        ps[i] = ps[i] * dos[i] / f0 if f0 > dostol else 0.0
    d.omega, d.dos, d.pdos_mode, d.pdos_site = om, dos, pm, ps
    d.default_smearing, d.omega_max, d.omega_min = ds, omax, omin
    return d


# ---------------------------------------------------------------- cumulative kappa vs mfp

def scalar_mfp(disp, linewidth):
    """scattering.f90:284-289: |v| / (2 Gamma) where Gamma > lo_freqtol, else zero."""
    v = np.linalg.norm(disp.vel_irr, axis=2)
    out = np.zeros_like(linewidth)
    m = linewidth > U.FREQTOL
    out[m] = v[m] / (2.0 * linewidth[m])
    return out


class Cumulative:
    pass


def cumulative_kappa(p, ops, mesh, disp, Fn, mfp, temperature, classical=False, npts=1000, sigma=1.0):
    """get_cumulative_kappa (cumulative_kappa.f90:58-414)."""
    na, nb, nirr = p.na, disp.nb, mesh.nirr
    kappatol = 1e-5 * KAPPA_SI_TO_AU
    kflat = np.zeros((nirr, nb, 6))
    for fq in range(mesh.nfull):
        iq, k = int(mesh.irr_of_full[fq]), int(mesh.op_of_full[fq])
        for b in range(nb):
            om = disp.om_irr[iq, b]
            if om < U.FREQTOL:
                continue
            cv = cv_of(temperature, om, classical)
            v0 = rotate_vec(ops, k, Fn[iq, b])
            v1 = rotate_vec(ops, k, disp.vel_irr[iq, b])
            buf = cv * np.outer(v0, v1) / p.volume
            kflat[iq, b] += [buf[0, 0], buf[1, 1], buf[2, 2], buf[1, 2], buf[0, 2], buf[0, 1]]
    kchop = np.zeros((nirr, nb, na, 6))
    for iq in range(nirr):
        for b in range(nb):
            prj = site_projection(disp.egv_irr[iq], na, b)
            prj = U.chop(prj / prj.sum(), U.SQTOL)
            for a in range(na):
                kchop[iq, b, a] = kflat[iq, b] * prj[a] / mesh.nfull
    ktot = np.array([kchop[..., i].sum() for i in range(6)])
    ktot = U.chop(ktot, kappatol)
    live = ktot >= kappatol
    # smearing width in log(mfp)
    x = np.array([math.log(v) if v > MFPTOL else math.log(MFPTOL) for v in mfp.reshape(-1)])
    x = np.sort(x)
    a = x[math.ceil(len(x) * 0.5) - 1]
    b = x[min(len(x), math.ceil(len(x) * 0.98)) - 1]
    logsig = 0.0
    for i in range(len(x) - 1):
        if a < x[i] < b:
            logsig = max(logsig, x[i + 1] - x[i])
    logsig *= sigma
    nn = npts * 20
    xm = np.zeros((nn, nb, na))
    ym = np.zeros((nn, nb, na, 6))
    for at in range(na):
        for md in range(nb):
            xq = mfp[:, md].copy()
            ind = np.argsort(xq, kind="stable")
            xq = xq[ind]
            y = np.zeros((nirr, 6))
            for ii in range(6):
                if live[ii]:
                    y[:, ii] = kchop[ind, md, at, ii]
            minx = xq[xq > U.TOL].min() * 0.1
            maxx = xq.max() * 1.5
            xm[:, md, at] = linspace(minx, maxx, nn)
            for iq in range(nirr):
                xv = xq[iq]
                if xv < minx or xv > maxx:
                    continue
                la = math.log(xv)
                s1 = math.exp(la - 4 * logsig) - xv
                s2 = math.exp(la + 4 * logsig) - xv
                isig = 1.0 / logsig
                ii = max(1, math.floor(nn * (xv + s1 - minx) / maxx))
                jj = min(nn, math.ceil(nn * (xv + s2 - minx) / maxx))
                if jj <= ii:
                    continue
                f1 = (1.0 + erf((np.log(xm[ii - 1:jj, md, at]) - math.log(xv)) * isig)) * 0.5
                ym[ii - 1:jj, md, at] += f1[:, None] * y[iq][None, :]
                ym[jj:, md, at] += y[iq][None, :]
    axis = logspace(xm.min(), xm.max() * 2.0, npts)
    mk = np.zeros((npts, 6))
    mb = np.zeros((npts, nb, 6))
    ma = np.zeros((npts, na, 6))
    f0 = np.zeros(6)
    for md in range(nb):
        for at in range(na):
            for i in range(npts):
                if axis[i] < xm[0, md, at] * 1.01:
                    f0 = np.zeros(6)
                elif axis[i] > xm[nn - 1, md, at] * 0.999:
                    f0 = ym[nn - 1, md, at].copy()
                else:
                    f0 = f0.copy()
                    for ii in range(6):
                        if live[ii]:
                            f0[ii] = linear_interpolation(xm[:, md, at], ym[:, md, at, ii], axis[i])
                mk[i] += f0
                mb[i, md] += f0
                ma[i, at] += f0
    for i in range(npts - 1):
        for ii in range(6):
            if not live[ii]:
                continue
            if mk[i + 1, ii] < mk[i, ii]:
                mk[i + 1, ii] = mk[i, ii]
            for at in range(na):
                if ma[i + 1, at, ii] < ma[i, at, ii]:
                    ma[i + 1, at, ii] = ma[i, at, ii]
            for md in range(nb):
                if mb[i + 1, md, ii] < mb[i, md, ii]:
                    mb[i + 1, md, ii] = mb[i, md, ii]
    deg, mem_ = atom_classes(ops, na)
    for ii in range(6):
        if not live[ii]:
            continue
        sb = np.zeros((npts, na))
        for i in range(na):
            for j in range(deg[i]):
                sb[:, i] = sb[:, i] + ma[:, mem_[i][j], ii] / float(deg[j])
        ma[:, :, ii] = sb
    for ii in range(6):
        if live[ii]:
            mk[:, ii] = mk[:, ii] * ktot[ii] / mk[npts - 1, ii]
    for i in range(npts):
        for ii in range(6):
            if not live[ii]:
                continue
            f1, f2 = mb[i, :, ii].sum(), mk[i, ii]
            if f1 > kappatol and f2 > kappatol:
                mb[i, :, ii] = mb[i, :, ii] * mk[i, ii] / f1
            else:
                mb[i, :, ii] = 0.0
            f1 = ma[i, :, ii].sum()
            if f1 > kappatol and f2 > kappatol:
                ma[i, :, ii] = ma[i, :, ii] * mk[i, ii] / f1
            else:
                ma[i, :, ii] = 0.0
    c = Cumulative()
    c.kappa_total, c.mfaxis, c.mf_kappa, c.mf_kappa_band, c.mf_kappa_atom = ktot, axis, mk, mb, ma
    return c


# ---------------------------------------------------------------- boundary scattering

def boundary_kappa(p, ops, mesh, disp, Fn, linewidth, mfp, kod, temperature, classical=False, npts=1000):
    """get_boundary_kappa (cumulative_kappa.f90:416-508): kappa with a boundary rate |v|/L
    added by Matthiessen's rule, for L on a log axis.  Returns (axis [npts], kappa [6, npts])."""
    nb, nirr = disp.nb, mesh.nirr
    pos = mfp[mfp > U.TOL]
    minx = max(pos.min() * 1e-1, U.GROUPVELTOL)
    maxx = pos.max() * 1e6
    axis = logspace(minx, maxx, npts)
    out = np.zeros((6, npts))
    f0 = 0.0
    M = [o.m for o in ops]
    for ll in range(npts):
        kappa = np.array(kod, dtype=np.float64).copy()
        for iq in range(nirr):
            for b in range(nb):
                om = disp.om_irr[iq, b]
                if om < U.FREQTOL:
                    continue
                cv = cv_of(temperature, om, classical)
                vel, fn, lw = disp.vel_irr[iq, b], Fn[iq, b], linewidth[iq, b]
                vn = float(np.linalg.norm(vel))
                if vn > U.GROUPVELTOL:
                    f0 = vn / axis[ll]
                f1 = lw / (lw + f0) if lw + f0 > U.FREQTOL else 1.0
                v2 = np.zeros((3, 3))
                for m in M:
                    v2 = v2 + np.outer(m @ fn, m @ vel) / len(ops)
                kappa = kappa + cv * v2 * mesh.weight_irr[iq] * f1 / p.volume
        tmp = np.zeros((3, 3))
        for m in M:
            tmp += m @ kappa @ m.T
        kappa = U.chop(tmp / len(ops), np.abs(tmp / len(ops)).sum() * 1e-6)
        kappa = U.chop(kappa, np.abs(kappa).sum() * 1e-6)
        out[:, ll] = [kappa[0, 0], kappa[1, 1], kappa[2, 2], kappa[1, 2], kappa[0, 2], kappa[0, 1]]
    return axis, out


# ---------------------------------------------------------------- spectral kappa

def spectral_kappa(p, ops, mesh, disp, Fn, dos, temperature, classical=False, sigma=1.0, tets=None):
    """get_spectral_kappa (cumulative_kappa.f90:510-561) and lo_phonon_dos%spectral_kappa
    (type_phonon_dos_integrations.f90:428-809).  Returns (spec [n, 9], band [n, nb, 9],
    atom [n, na, 9]), the nine components in the order (xx, xy, xz, yx, ..., zz)."""
    na, nb, nirr, nf = p.na, disp.nb, mesh.nirr, mesh.nfull
    kappatol = 1e-6 * KAPPA_SI_TO_AU
    n, om = dos.n, dos.omega
    kf = np.zeros((9, nirr, nb))
    for aq in range(nf):
        iq, k = int(mesh.irr_of_full[aq]), int(mesh.op_of_full[aq])
        for b in range(nb):
            w = disp.om_irr[iq, b]
            if w < U.FREQTOL:
                continue
            cv = cv_of(temperature, w, classical)
            buf = cv * np.outer(rotate_vec(ops, k, Fn[iq, b]), rotate_vec(ops, k, disp.vel_irr[iq, b])) / p.volume
            kf[:, iq, b] += buf.reshape(-1)
    kflat = np.zeros((nf, nb, 9))
    for aq in range(nf):
        iq = int(mesh.irr_of_full[aq])
        f0 = 1.0 / mesh.weight_irr[iq]
        for b in range(nb):
            kflat[aq, b] = kf[:, iq, b] * f0
    kflat = U.chop(kflat / nf, kappatol)
    ktot = np.array([kflat[:, :, i].sum() / nf for i in range(9)])
    sb = np.zeros((n, nb, 9))
    sa = np.zeros((n, na, 9))
    ds = dos.default_smearing
    invf = n / om.max()
    if dos.integrationtype in (1, 2):
        scaled = p.reclat * 2 * math.pi / np.array(mesh.dims, dtype=np.float64)[None, :]
        for dr in range(9):
            if abs(ktot[dr]) < kappatol:
                continue
            for iq in range(nirr):
                wq = mesh.weight_irr[iq]
                fi = int(mesh.irr_full_index[iq])
                for b in range(nb):
                    f1 = disp.om_irr[iq, b]
                    if f1 < dos.omega_min * 0.5:
                        continue
                    if dos.integrationtype == 1:
                        sg = ds[b] * sigma
                    else:
                        g = np.abs(disp.vel_irr[iq, b])
                        s = 0.0
                        for i in range(3):
                            s += float(g @ scaled[:, i]) ** 2
                        sg = math.sqrt(s * 0.083333333333) * sigma
                        sg = min(sg, sigma * ds[b] * 4.0)
                        sg = max(sg, sigma * ds[b] * 0.25)
                    proj = site_projection(disp.egv_irr[iq], na, b)
                    qk = kflat[fi, b, dr]
                    ii = max(math.floor((f1 - 4 * sg) * invf), 1)
                    jj = min(math.floor((f1 + 4 * sg) * invf), n)
                    for x in range(ii, jj + 1):
                        f0 = gauss(f1, om[x - 1], sg) * wq
                        sb[x - 1, b, dr] += qk * f0
                        sa[x - 1, :, dr] += qk * proj * f0
    else:
        corners, weights = tets if tets is not None else irreducible_tetrahedra(p, ops, mesh)
        tps = [[np.array([site_projection(disp.egv_full[c], na, b) for c in corners[it]])
                for it in range(len(weights))] for b in range(nb)]                      # [4, na]
        for dr in range(9):
            for b in range(nb):
                sg = ds[b]
                for it in range(len(weights)):
                    cs = corners[it]
                    tete = np.array([disp.om_full[c, b] for c in cs])
                    tk = np.array([kflat[c, b, dr] for c in cs])
                    ii = max(math.floor(tete.min() * invf), 1)
                    jj = min(math.ceil(tete.max() * invf), n)
                    if jj < ii:
                        continue
                    tp = tps[b][it]
                    for x in range(ii, jj + 1):
                        w = lv_tetrahedron_weights(tete, om[x - 1], U.FREQTOL, sg) * weights[it]
                        sb[x - 1, b, dr] += np.sum(w * tk)
                        sa[x - 1, :, dr] += (w * tk) @ tp
    # makenice
    ker = np.array([gauss(0.0, i * 1.0, 1.5) for i in range(-6, 7)])
    ker = ker / ker.sum()
    sa = np.maximum(sa, 0.0)
    sb = np.maximum(sb, 0.0)
    spec = np.zeros((n, 9))
    deg, mem_ = atom_classes(ops, na)

    def smooth(y):
        ysm = np.zeros(n + 12)
        for i in range(n):
            for j in range(-6, 7):
                ysm[i + 6 + j] += ker[j + 6] * y[i]
        return ysm[6:6 + n].copy()

    for dr in range(9):
        if ktot[dr] < kappatol:
            continue
        for a in range(na):
            sa[:, a, dr] = smooth(sa[:, a, dr])
        for b in range(nb):
            sb[:, b, dr] = smooth(sb[:, b, dr])
        for i in range(na):
            ysm = np.zeros(n)
            for _ in range(deg[i]):
                ysm = ysm + sa[:, i, dr] / (1.0 * deg[i])
            for j in range(deg[i]):
                sa[:, mem_[i][j], dr] = ysm
        for a in range(na):
            f0 = sa[0, a, dr]
            for i in range(1, n + 1):
                f1 = 1.0 - (i - 1.0) / float(n - 1)
                sa[i - 1, a, dr] = max(sa[i - 1, a, dr] - f1 * f0, 0.0)
            sa[0, a, dr] = 0.0
        for b in range(nb):
            f0 = sb[0, b, dr]
            for i in range(1, n + 1):
                f1 = 1.0 - (i - 1.0) / float(n - 1)
                sb[i - 1, b, dr] = max(sb[i - 1, b, dr] - f1 * f0, 0.0)
            sb[0, b, dr] = 0.0
        spec[:, dr] = 0.0
        for b in range(nb):
            spec[:, dr] = spec[:, dr] + sb[:, b, dr]
        f0 = trapezoid(om, spec[:, dr])
        spec[:, dr] = spec[:, dr] * ktot[dr] / f0 if f0 > kappatol else 0.0
        for i in range(n):
            f0, f1, f2 = sa[i, :, dr].sum(), sb[i, :, dr].sum(), spec[i, dr]
            if f2 > kappatol:
                sa[i, :, dr] = sa[i, :, dr] * f2 / f0 if f0 > kappatol else 0.0
                sb[i, :, dr] = sb[i, :, dr] * f2 / f1 if f1 > kappatol else 0.0
            else:
                spec[i, dr] = 0.0
                sb[i, :, dr] = 0.0
                sa[i, :, dr] = 0.0
    return spec, sb, sa, ktot


# ---------------------------------------------------------------- angular momentum

def _small_group(p, ops, qc):
    """lo_get_small_group_of_qpoint (type_qpointmesh.f90:660-700), no time reversal."""
    out = []
    for i, o in enumerate(ops):
        v1 = p.inv_reclat @ (qc - o.m @ qc)
        v1 = v1 - np.round(v1)
        if v1 @ v1 < U.SQTOL:
            out.append(i)
    return out


def _mode_angular_momentum(egv, na, mode):
    """sum over atoms of e^H M_a e, M_a the spin-1 generators of cumulative_kappa's
    dispersions routine: 2 Im(conj(e_y) e_z), 2 Im(conj(e_z) e_x), 2 Im(conj(e_x) e_y)."""
    cv = np.zeros(3, dtype=np.complex128)
    M = np.zeros((3, 3, 3), dtype=np.complex128)
    M[0, 1, 2], M[0, 2, 1] = 1, -1
    M[1, 0, 2], M[1, 2, 0] = -1, 1
    M[2, 0, 1], M[2, 1, 0] = 1, -1
    M = -M * 1j
    for k in range(na):
        c1 = egv[3 * k:3 * k + 3, mode]
        for a in range(3):
            cv[a] += np.vdot(c1, M[a] @ c1)
    return cv.real


def angular_momentum_matrix(p, ops, mesh, disp, linewidth, temperature):
    """phonon_angular_momentum_matrix (type_phonon_dispersions.f90:868-971), with linewidths."""
    alpha = np.zeros((3, 3))
    for i in range(mesh.nfull):
        lg = _small_group(p, ops, mesh.qcart[i])
        k = int(mesh.irr_of_full[i])
        for j in range(disp.nb):
            om = disp.om_full[i, j]
            if om < U.FREQTOL:
                continue
            v1 = _mode_angular_momentum(disp.egv_full[i], p.na, j)
            w1 = disp.vel_full[i, j]
            v0 = np.zeros(3)
            w0 = np.zeros(3)
            for o in lg:
                v0 = v0 + ops[o].m @ v1
                w0 = w0 + ops[o].m @ w1
            v0, w0 = v0 / len(lg), w0 / len(lg)
            f0 = float(U.harmonic_oscillator_cv(temperature, om)) / om
            f0 = f0 * (1.0 / mesh.nfull)
            f0 = f0 / (2.0 * linewidth[k, j])
            alpha = alpha + np.outer(v0, w0) * f0
    alpha = alpha / p.volume
    return U.chop(alpha, np.linalg.norm(alpha) * 1e-10)


def spectral_angular_momentum(p, ops, mesh, disp, linewidth, dos, temperature, sigma=1.0, tets=None):
    """spectral_angular_momentum (type_phonon_dos_integrations.f90:8-425), as written: the
    mode's angular momentum (not its little-group average) times c_v/omega times twice the
    linewidth, outer the averaged velocity.  Returns (spec [n, 9], band [n, nb, 9],
    atom [n, na, 9], total [9]), components in the order of the outer product's rows."""
    na, nb, nirr, nf = p.na, disp.nb, mesh.nirr, mesh.nfull
    tol = 1e-30
    n, om = dos.n, dos.omega
    kf = np.zeros((9, nirr, nb))
    for iq in range(nf):
        ii = int(mesh.irr_of_full[iq])
        lg = _small_group(p, ops, mesh.qcart[iq])
        for b in range(nb):
            w = disp.om_full[iq, b]
            if w < U.FREQTOL:
                continue
            cv0 = _mode_angular_momentum(disp.egv_full[iq], na, b)
            w0 = np.zeros(3)
            for o in lg:
                w0 = w0 + ops[o].m @ disp.vel_full[iq, b]
            w0 = w0 / len(lg)
            f0 = float(U.harmonic_oscillator_cv(temperature, w))
            f0 = f0 / w
            f0 = f0 * (1.0 / nf)
            f0 = f0 * (2.0 * linewidth[ii, b])
            kf[:, ii, b] += np.outer(cv0 * f0, w0).reshape(-1)
    flat = np.zeros((nf, nb, 9))
    for iq in range(nf):
        ii = int(mesh.irr_of_full[iq])
        flat[iq] = (kf[:, ii, :] * (1.0 / mesh.weight_irr[ii])).T
    total = np.array([flat[:, :, i].sum() / nf for i in range(9)])
    total = U.chop(total, np.linalg.norm(total) * 1e-9)
    sb = np.zeros((n, nb, 9))
    sa = np.zeros((n, na, 9))
    ds = dos.default_smearing
    invf = n / om.max()
    if dos.integrationtype in (1, 2):
        scaled = p.reclat * 2 * math.pi / np.array(mesh.dims, dtype=np.float64)[None, :]
        for dr in range(9):
            if abs(total[dr]) < tol:
                continue
            for iq in range(nirr):
                wq = mesh.weight_irr[iq]
                fi = int(mesh.irr_full_index[iq])
                for b in range(nb):
                    f1 = disp.om_irr[iq, b]
                    if f1 < dos.omega_min * 0.5:
                        continue
                    if dos.integrationtype == 1:
                        sg = ds[b] * sigma
                    else:
                        g = np.abs(disp.vel_irr[iq, b])
                        s = 0.0
                        for i in range(3):
                            s += float(g @ scaled[:, i]) ** 2
                        sg = math.sqrt(s * 0.083333333333) * sigma
                        sg = min(sg, sigma * ds[b] * 4.0)
                        sg = max(sg, sigma * ds[b] * 0.25)
                    proj = site_projection(disp.egv_irr[iq], na, b)
                    qa = flat[fi, b, dr]
                    ii = max(math.floor((f1 - 4 * sg) * invf), 1)
                    jj = min(math.floor((f1 + 4 * sg) * invf), n)
                    for x in range(ii, jj + 1):
                        f0 = gauss(f1, om[x - 1], sg) * wq
                        sb[x - 1, b, dr] += qa * f0
                        sa[x - 1, :, dr] += qa * proj * f0
    else:
        corners, weights = tets if tets is not None else irreducible_tetrahedra(p, ops, mesh)
        tps = [[np.array([site_projection(disp.egv_full[c], na, b) for c in corners[it]])
                for it in range(len(weights))] for b in range(nb)]
        for dr in range(9):
            for b in range(nb):
                sg = ds[b]
                for it in range(len(weights)):
                    cs = corners[it]
                    tete = np.array([disp.om_full[c, b] for c in cs])
                    ta = np.array([flat[c, b, dr] for c in cs])
                    ii = max(math.floor(tete.min() * invf), 1)
                    jj = min(math.ceil(tete.max() * invf), n)
                    for x in range(ii, jj + 1):
                        w = lv_tetrahedron_weights(tete, om[x - 1], U.FREQTOL, sg) * weights[it]
                        sb[x - 1, b, dr] += np.sum(w * ta)
                        sa[x - 1, :, dr] += (w * ta) @ tps[b][it]
    ker = np.array([gauss(0.0, i * 1.0, 1.5) for i in range(-6, 7)])
    ker = ker / ker.sum()
    spec = np.zeros((n, 9))
    deg, mem_ = atom_classes(ops, na)

    def smooth(y):
        ysm = np.zeros(n + 12)
        for i in range(n):
            for j in range(-6, 7):
                ysm[i + 6 + j] += ker[j + 6] * y[i]
        return ysm[6:6 + n].copy()

    for dr in range(9):
        if abs(total[dr]) < tol:
            continue
        for a in range(na):
            sa[:, a, dr] = smooth(sa[:, a, dr])
        for b in range(nb):
            sb[:, b, dr] = smooth(sb[:, b, dr])
        for i in range(na):
            ysm = np.zeros(n)
            for _ in range(deg[i]):
                ysm = ysm + sa[:, i, dr] / (1.0 * deg[i])
            for j in range(deg[i]):
                sa[:, mem_[i][j], dr] = ysm
        spec[:, dr] = 0.0
        for b in range(nb):
            spec[:, dr] = spec[:, dr] + sb[:, b, dr]
        f0 = trapezoid(om, spec[:, dr])
        spec[:, dr] = spec[:, dr] * total[dr] / f0 if abs(f0) > tol else 0.0
        for i in range(n):
            f0, f1, f2 = sa[i, :, dr].sum(), sb[i, :, dr].sum(), spec[i, dr]
            if abs(f2) > tol:
                sa[i, :, dr] = sa[i, :, dr] * f2 / f0 if abs(f0) > tol else 0.0
                sb[i, :, dr] = sb[i, :, dr] * f2 / f1 if abs(f1) > tol else 0.0
            else:
                spec[i, dr] = 0.0
                sb[i, :, dr] = 0.0
                sa[i, :, dr] = 0.0
    return spec, sb, sa, total
