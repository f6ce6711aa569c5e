"""Readers, space group and FFT q-mesh of the reference, restated in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Arrays are 0-based and row-major; a Fortran (3, n) array of the reference is an
(n, 3) array here.  Lattice matrices keep the reference's column convention:
latvec[:, i] is lattice vector i (Bohr), reclat[:, i] reciprocal vector i
(no 2*pi).
"""
import numpy as np

from . import units as U


# ----------------------------------------------------------------------------
# Readers
# ----------------------------------------------------------------------------
class Structure:
    pass


def read_ucposcar(path):
    """lo_crystalstructure%readfromfile (type_crystalstructure_io.f90:9-350)
    for plain (non-alloy, non-magnetic) POSCARs, followed by the unit handling
    of lo_crystalstructure%generate with enhet=1
    (type_crystalstructure.f90:446-480)."""
    with open(path) as fh:
        lines = [ln for ln in fh.read().splitlines()]
    scale = float(lines[1].split()[0])
    m = np.array([[float(x) for x in lines[2 + i].split()[:3]] for i in range(3)]).T
    symbols = lines[5].split()
    counts = [int(x) for x in lines[6].split()]
    symbols = symbols[:len(counts)]
    li = 7
    if lines[li].strip()[:1] in ("s", "S"):
        li += 1
    mode = lines[li].strip()[:1]
    li += 1
    if mode in ("C", "c"):
        cartesian = True
    elif mode in ("D", "d"):
        cartesian = False
    else:
        raise ValueError("Specify either Cartesian or direct coordinates")
    if scale > 0:
        m = m * scale
    else:
        m = m * (abs(scale) / abs(np.linalg.det(m))) ** (1.0 / 3.0)
    im = np.linalg.inv(m)
    na = sum(counts)
    r = np.zeros((na, 3))
    species = []
    for s, c in zip(symbols, counts):
        species += [U.symbol_to_z(s)] * c
    for a in range(na):
        v = np.array([float(x) for x in lines[li + a].split()[:3]])
        if cartesian:
            v = im @ v
        r[a] = U.clean_fractional(v)
    p = Structure()
    p.na = na
    p.symbols = [U.SYMBOLS[z - 1] for z in species]
    p.atomic_number = np.array(species, dtype=np.int64)
    lat = m * U.A_TO_BOHR
    p.latvec = U.chop(lat, U.SQTOL)
    p.inv_latvec = U.chop(np.linalg.inv(lat), U.SQTOL)
    rec = np.linalg.inv(lat).T
    p.reclat = U.chop(rec, U.SQTOL)
    p.inv_reclat = U.chop(np.linalg.inv(rec), U.SQTOL)
    p.volume = abs(np.linalg.det(p.latvec))
    p.rfrac = U.chop(U.clean_fractional(r), U.SQTOL)
    p.rcart = U.chop(p.rfrac @ p.latvec.T, U.SQTOL)
    # species ids: first-seen order of distinct atomic numbers
    uniq = []
    for z in p.atomic_number:
        if z not in uniq:
            uniq.append(z)
    p.species = np.array([uniq.index(z) for z in p.atomic_number], dtype=np.int64)
    p.mass = None
    p.invsqrtmass = None
    return p


def read_isotopes(p, path):
    """lo_crystalstructure%readisotopefromfile (type_crystalstructure_io.f90:921-959)."""
    toks = open(path).read().split()
    pos = 0
    mass = np.zeros(p.na)
    disorder = np.zeros(p.na)
    for a in range(p.na):
        n = int(toks[pos]); pos += 1
        conc = np.zeros(n); mm = np.zeros(n)
        for j in range(n):
            conc[j] = float(toks[pos]); mm[j] = float(toks[pos + 1]); pos += 2
        mm = mm * U.AMU_TO_EMU
        conc = conc / conc.sum()
        mean = (conc * mm).sum()
        mass[a] = mean
        # mass_disorder_parameter (type_crystalstructure.f90:1505-1519)
        disorder[a] = (conc * (mm - mean) ** 2).sum() / mean ** 2
    p.mass = mass
    p.invsqrtmass = 1.0 / np.sqrt(mass)
    p.mass_disorder = disorder
    return p


def read_fc2(p, path):
    """lo_forceconstant_secondorder%readfromfile (type_forceconstant_secondorder_io.f90:109-247),
    non-polar branch. Returns flat pair arrays: a1, a2, lv2 (Cartesian Bohr), m (Ha/Bohr^2)."""
    toks = open(path).read().split("\n")
    it = iter(toks)

    def nxt():
        while True:
            ln = next(it)
            if ln.strip():
                return ln.split()
    na = int(nxt()[0])
    if na != p.na:
        raise ValueError("different number of atoms in forceconstant and structure")
    nxt()  # cutoff
    a1s, a2s, lvs, ms = [], [], [], []
    for a1 in range(na):
        n = int(nxt()[0])
        for _ in range(n):
            a2 = int(nxt()[0]) - 1
            lv = np.array([float(x) for x in nxt()[:3]])
            m = np.array([[float(x) for x in nxt()[:3]] for _ in range(3)])
            a1s.append(a1); a2s.append(a2); lvs.append(lv); ms.append(m * U.FC2_EVA_TO_HARTREEBOHR)
    polar = int(nxt()[0])
    if polar != 0:
        raise NotImplementedError("polar force constants are outside this path")
    lv = np.array(lvs)
    lvc = U.chop(lv @ p.latvec.T, 1e-13)
    return dict(a1=np.array(a1s), a2=np.array(a2s), lv2=lvc, m=np.array(ms))


def read_fc3(p, path):
    """lo_forceconstant_thirdorder%readfromfile (type_forceconstant_thirdorder.f90:243-311).
    Returns flat triplet arrays with Cartesian lattice vectors (Bohr) and the mass-weighted
    tensor mwm[t, i, j, k] (Ha/Bohr^3/emu^1.5)."""
    lines = [ln.split() for ln in open(path).read().split("\n") if ln.strip()]
    pos = 0
    na = int(lines[pos][0]); pos += 1
    pos += 1  # cutoff
    out = {k: [] for k in ("a1", "a2", "a3", "lv1", "lv2", "lv3", "m")}
    for a in range(na):
        n = int(lines[pos][0]); pos += 1
        for _ in range(n):
            i1 = int(lines[pos][0]) - 1; i2 = int(lines[pos + 1][0]) - 1; i3 = int(lines[pos + 2][0]) - 1
            lv1 = [float(x) for x in lines[pos + 3][:3]]
            lv2 = [float(x) for x in lines[pos + 4][:3]]
            lv3 = [float(x) for x in lines[pos + 5][:3]]
            pos += 6
            m = np.zeros((3, 3, 3))
            for ii in range(3):
                for jj in range(3):
                    m[ii, jj, :] = [float(x) for x in lines[pos][:3]]
                    pos += 1
            out["a1"].append(a); out["a2"].append(i2); out["a3"].append(i3)
            out["lv1"].append(lv1); out["lv2"].append(lv2); out["lv3"].append(lv3)
            out["m"].append(m * U.FC3_EVA_TO_HARTREEBOHR)
    fc = {k: np.array(v) for k, v in out.items()}
    ism = p.invsqrtmass
    fc["mwm"] = fc["m"] * (ism[fc["a1"]] * ism[fc["a2"]] * ism[fc["a3"]])[:, None, None, None]
    for k in ("lv1", "lv2", "lv3"):
        fc[k] = U.chop(fc[k] @ p.latvec.T, U.SQTOL)
    return fc


def read_fc4(p, path):
    """lo_forceconstant_fourthorder%readfromfile (type_forceconstant_fourthorder.f90:259-340).
    Flat quartet arrays with Cartesian lattice vectors (Bohr) and the mass-weighted tensor
    mwm[t, i, j, k, l] (Ha/Bohr^4/emu^2)."""
    lines = [ln.split() for ln in open(path).read().split("\n") if ln.strip()]
    pos = 0
    na = int(lines[pos][0]); pos += 1
    pos += 1  # cutoff
    out = {k: [] for k in ("a1", "a2", "a3", "a4", "lv2", "lv3", "lv4", "m")}
    for a in range(na):
        n = int(lines[pos][0]); pos += 1
        for _ in range(n):
            ii = [int(lines[pos + k][0]) - 1 for k in range(4)]
            lv = [[float(x) for x in lines[pos + 4 + k][:3]] for k in range(4)]
            pos += 8
            m = np.zeros((3, 3, 3, 3))
            for i in range(3):
                for j in range(3):
                    for k in range(3):
                        m[i, j, k, :] = [float(x) for x in lines[pos][:3]]
                        pos += 1
            out["a1"].append(a); out["a2"].append(ii[1]); out["a3"].append(ii[2]); out["a4"].append(ii[3])
            out["lv2"].append(lv[1]); out["lv3"].append(lv[2]); out["lv4"].append(lv[3])
            out["m"].append(m * U.FC4_EVA_TO_HARTREEBOHR)
    fc = {k: np.array(v) for k, v in out.items()}
    ism = p.invsqrtmass
    fc["mwm"] = fc["m"] * (ism[fc["a1"]] * ism[fc["a2"]] * ism[fc["a3"]] * ism[fc["a4"]])[:, None, None, None, None]
    for k in ("lv2", "lv3", "lv4"):
        fc[k] = U.chop(fc[k] @ p.latvec.T, U.SQTOL)
    return fc


# ----------------------------------------------------------------------------
# Space group
# ----------------------------------------------------------------------------
class SymOp:
    pass


def spacegroup(p, tol=U.TOL):
    """Space-group operations (R, t) of the crystal.

    Same contract as lo_symset%generate (type_symmetryoperation.f90:296-~700):
    every (R, t) that maps the crystal onto itself (species preserved), with the
    Cartesian rotation m, its transpose im, the reciprocal fractional rotation
    rfm = chop(B^-1 m B, 1e-10), the Cartesian translation and the atom map
    fmap[a] = image of atom a.  Enumeration order is this build's own and is
    fixed: identity first, then rotations in lexicographic order of their
    fractional matrix entries, translations in lexicographic order.
    """
    A = p.latvec
    G = A.T @ A
    vals = np.array([-1, 0, 1])
    grid = np.array(np.meshgrid(*([vals] * 9), indexing="ij")).reshape(9, -1).T
    W = grid.reshape(-1, 3, 3).astype(np.float64)
    GW = np.einsum("nki,kl,nlj->nij", W, G, W)
    ok = np.abs(GW - G[None]).sum(axis=(1, 2)) < 1e-6 * np.abs(G).sum()
    Ws = W[ok]
    ident = np.eye(3)
    order = sorted(range(len(Ws)), key=lambda i: (0 if np.allclose(Ws[i], ident) else 1, i))
    Ws = Ws[order]
    rf = p.rfrac
    ops = []
    for Wm in Ws:
        # candidate translations from atom 0
        rot = rf @ Wm.T
        cands = []
        for b in range(p.na):
            if p.species[b] != p.species[0]:
                continue
            t = U.clean_fractional(rf[b] - rot[0])
            cands.append(t)
        found = []
        for t in cands:
            fmap = np.full(p.na, -1)
            good = True
            for a in range(p.na):
                d = rf - (rot[a] + t)
                d = d - np.round(d)
                dc = np.linalg.norm(d @ A.T, axis=1)
                j = np.argmin(dc)
                if dc[j] > tol or p.species[j] != p.species[a]:
                    good = False
                    break
                fmap[a] = j
            if good:
                found.append((np.round(t, 8).tolist(), t, fmap))
        found.sort(key=lambda x: x[0])
        for _, t, fmap in found:
            o = SymOp()
            o.fm = Wm.copy()
            o.m = U.chop(A @ Wm @ np.linalg.inv(A), 1e-10)
            o.im = o.m.T.copy()
            o.rfm = U.chop(p.inv_reclat @ o.m @ p.reclat, 1e-10)
            o.ftr = t
            o.tr = U.chop(A @ t, 1e-11)
            o.fmap = fmap
            ops.append(o)
    return ops


# ----------------------------------------------------------------------------
# FFT q-point mesh
# ----------------------------------------------------------------------------
class Mesh:
    pass


def _bz_fold(p, v0):
    """Move a fractional q to the first Brillouin zone as
    type_qpointmesh_gridgeneration.f90:115-133 does: subtract the reciprocal lattice
    vector nearest to q + lo_degenvector (the shift breaks ties on the zone boundary).
    The reference skips points inside the inscribed sphere first; that shortcut only
    changes which of two boundary-equivalent images a point on the sphere gets, so it is
    left out to keep the choice independent of rounding."""
    B = p.reclat
    q = B @ v0
    v1 = q + U.DEGENVECTOR
    n2 = np.array(np.meshgrid(*([np.arange(-2, 3)] * 3), indexing="ij")).reshape(3, -1).T
    Gs = n2 @ B.T
    d = np.linalg.norm(v1[None, :] - Gs, axis=1)
    g = n2[np.argmin(d)]
    return v0 - g


def _on_grid(v, dims):
    w = v * dims + 1.0
    return np.abs(w - np.round(w)).sum() < U.TOL


def _index_from_coordinate(v, dims):
    w = U.clean_fractional(v) * dims + 1.0
    gi = np.round(w).astype(int)  # 1-based
    return gi


def generate_mesh(p, ops, dims, timereversal=True):
    """lo_generate_qmesh(... 'fft' ...) -> build_grid (type_qpointmesh_gridgeneration.f90:8-545)
    plus lo_get_small_group_of_qpoint (type_qpointmesh.f90:660-700)."""
    dims = np.array(dims, dtype=np.int64)
    n1, n2, n3 = dims
    npt = int(n1 * n2 * n3)
    nsym = len(ops)
    gridind = np.zeros((npt, 3), dtype=np.int64)
    pts = np.zeros((npt, 3))
    l = 0
    for i in range(n1):
        for j in range(n2):
            for k in range(n3):
                v0 = np.array([i / n1, j / n2, k / n3])
                gridind[l] = (i, j, k)
                pts[l] = _bz_fold(p, v0)
                l += 1

    def gi2ind(gi):
        return int((gi[0] - 1) * n2 * n3 + (gi[1] - 1) * n3 + (gi[2] - 1))

    ncol = 2 * nsym if timereversal else nsym
    conn = np.zeros((npt, ncol), dtype=np.int64)  # 1-based point index, sign = TR
    for i in range(npt):
        v0 = pts[i]
        for j, o in enumerate(ops):
            for tr in ((False, True) if timereversal else (False,)):
                s = -1.0 if tr else 1.0
                v1 = s * (o.rfm @ v0)
                if not _on_grid(v1, dims):
                    continue
                ii = gi2ind(_index_from_coordinate(v1, dims))
                v2 = s * (np.linalg.inv(o.rfm) @ pts[ii])
                if not _on_grid(v2, dims):
                    continue
                jj = gi2ind(_index_from_coordinate(v2, dims))
                if jj != i:
                    continue
                if tr:
                    conn[i, j + nsym] = -(ii + 1)
                else:
                    conn[i, j] = ii + 1
    ind = np.ones(npt, dtype=bool)
    for i in range(npt):
        for j in range(ncol):
            k = abs(conn[i, j]) - 1
            if k > i:
                ind[k] = False
    irr = np.nonzero(ind)[0]
    m = Mesh()
    m.dims = dims
    m.nfull = npt
    m.nirr = len(irr)
    m.irr_full_index = irr
    m.irr_of_full = np.full(npt, -1)
    m.op_of_full = np.zeros(npt, dtype=np.int64)  # 1-based op, negative = time reversal
    for k, qi in enumerate(irr):
        for j in range(nsym):
            i = conn[qi, j] - 1
            if i < 0 or m.irr_of_full[i] >= 0:
                continue
            m.irr_of_full[i] = k
            m.op_of_full[i] = j + 1
        if timereversal:
            for j in range(nsym):
                i = -conn[qi, j + nsym] - 1
                if i < 0 or m.irr_of_full[i] >= 0:
                    continue
                m.irr_of_full[i] = k
                m.op_of_full[i] = -(j + 1)
    m.weight_irr = np.bincount(m.irr_of_full, minlength=m.nirr) / float(npt)
    m.frac = pts
    m.qcart = U.chop(pts @ p.reclat.T, 1e-14)
    m.gridind = gridind

    def small_group(qc):
        out = []
        for i, o in enumerate(ops):
            v1 = qc - o.m @ qc
            v1 = p.inv_reclat @ v1
            v1 = v1 - np.round(v1)
            if v1 @ v1 < U.SQTOL:
                out.append(i)
        return np.array(out, dtype=np.int64)

    m.littlegroup_irr = [small_group(m.qcart[q]) for q in irr]
    # star of each irreducible point (index_full_point / operation_full_point)
    m.star_ops = []
    for k, qi in enumerate(irr):
        v0 = p.inv_reclat @ m.qcart[qi]
        seen, sops = [], []
        for tr in ((False, True) if timereversal else (False,)):
            for j, o in enumerate(ops):
                v1 = (-1.0 if tr else 1.0) * (o.rfm @ v0)
                if not _on_grid(v1, dims):
                    continue
                jj = gi2ind(_index_from_coordinate(v1, dims))
                if jj in seen:
                    continue
                seen.append(jj)
                sops.append(-(j + 1) if tr else (j + 1))
        m.star_ops.append(np.array(sops))
    m.gi2ind = gi2ind
    return m


def rotate_qindex(p, m, op, q):
    """Grid index of R q (reciprocal, fractional) or -1 if off grid
    (used by scattering_threephonon.f90:165-169, 377-388)."""
    v = p.inv_reclat @ m.qcart[q]
    v1 = op.rfm @ v
    if not _on_grid(v1, m.dims):
        return -1
    return m.gi2ind(_index_from_coordinate(v1, m.dims))
