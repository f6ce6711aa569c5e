"""Readers of the reference's input files, producing arrays in the internal
Hartree atomic units the C ABI takes.

  infile.ucposcar                 type_crystalstructure_io.f90:9-350 (plain POSCAR)
                                  + unit handling of type_crystalstructure.f90:446-480
  infile.isotopes                 type_crystalstructure_io.f90:921-959
  infile.forceconstant            type_forceconstant_secondorder_io.f90:109-247 (non-polar)
  infile.forceconstant_thirdorder type_forceconstant_thirdorder.f90:243-311

Numbers are read the way Fortran list-directed input reads them: the leading
tokens of a record, the rest of the line ignored.
"""
import os

import numpy as np

# konstanter.f90:36-91
BOHR_TO_A = 0.52917721067
A_TO_BOHR = 1.0 / BOHR_TO_A
HARTREE_TO_EV = 27.21138602
EV_TO_HARTREE = 1.0 / HARTREE_TO_EV
AMU_TO_EMU = 1.660539040e-27 / 9.10938356e-31
FC2_UNIT = EV_TO_HARTREE / A_TO_BOHR ** 2
FC3_UNIT = EV_TO_HARTREE / A_TO_BOHR ** 3
FC4_UNIT = EV_TO_HARTREE / (A_TO_BOHR ** 4)   # lo_forceconstant_4th_eVA_to_HartreeBohr
SQTOL = 1e-10

_ELEMENTS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn "
    "Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce "
    "Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn "
    "Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr").split()

# gottochblandat.f90:1286-1374: values that lo_chop snaps to, in order
_SNAP = np.array([0.0] + [s * v for v in (
    0.1, 0.1111111111111111, 0.125, 0.1428571428571428, 0.1666666666666667, 0.2,
    0.2222222222222222, 0.25, 0.2857142857142857, 0.3, 0.3333333333333333, 0.375, 0.4,
    0.4285714285714285, 0.4444444444444444, 0.5, 0.5555555555555556, 0.5714285714285714,
    0.6, 0.625, 0.6666666666666667, 0.7, 0.7071067811865475, 0.7142857142857143, 0.75,
    0.7777777777777778, 0.8, 0.8333333333333334, 0.8571428571428571, 0.8660254037844386,
    0.875, 0.8888888888888888, 0.9, 1.0) for s in (1.0, -1.0)])


def chop(a, tol):
    a = np.array(a, dtype=np.float64, copy=True)
    f = a.reshape(-1)
    for i in range(f.size):
        hit = np.nonzero(np.abs(f[i] - _SNAP) < tol)[0]
        if hit.size:
            f[i] = _SNAP[hit[0]]
    return a


def _clean_frac(v, tol=SQTOL):
    out = []
    for x in np.ravel(v):
        while True:
            if abs(x) < tol:
                x = 0.0
            if abs(x - 1.0) < tol:
                x = 0.0
            if x > 1.0:
                x -= 1.0
            if x < 0.0:
                x += 1.0
            if -tol < x < 1.0:
                break
        out.append(x)
    return np.array(out).reshape(np.shape(v))


def _records(path):
    with open(path) as fh:
        return [ln.split() for ln in fh if ln.strip()]


class Structure:
    """latvec: 3x3, columns are lattice vectors (Bohr); rfrac: (na,3); species: (na,)
    species ids; mass: (na,) emu."""


def read_ucposcar(path):
    lines = open(path).read().splitlines()
    scale = float(lines[1].split()[0])
    m = np.array([[float(x) for x in lines[2 + i].split()[:3]] for i in range(3)]).T
    counts = [int(x) for x in lines[6].split()]
    symbols = lines[5].split()[:len(counts)]
    li = 7
    if lines[li].strip()[:1] in "sS":
        li += 1
    mode = lines[li].strip()[:1]
    li += 1
    if mode not in "CcDd" or not mode:
        raise ValueError(f"{path}: specify either Cartesian or direct coordinates")
    m = m * scale if scale > 0 else m * (abs(scale) / abs(np.linalg.det(m))) ** (1.0 / 3.0)
    na = sum(counts)
    r = np.zeros((na, 3))
    z = []
    for s, c in zip(symbols, counts):
        if s not in _ELEMENTS:
            raise ValueError(f"{path}: unknown element '{s}'")
        z += [_ELEMENTS.index(s) + 1] * c
    im = np.linalg.inv(m)
    for a in range(na):
        v = np.array([float(x) for x in lines[li + a].split()[:3]])
        r[a] = _clean_frac(im @ v if mode in "Cc" else v)
    p = Structure()
    p.na = na
    p.symbols = [_ELEMENTS[x - 1] for x in z]
    uniq = list(dict.fromkeys(z))
    p.species = np.array([uniq.index(x) for x in z], dtype=np.int32)
    lat = m * A_TO_BOHR
    p.latvec = chop(lat, SQTOL)
    p.reclat = chop(np.linalg.inv(lat).T, SQTOL)
    p.volume = abs(np.linalg.det(p.latvec))
    p.rfrac = chop(_clean_frac(r), SQTOL)
    p.mass = None
    return p


def _distribution(conc, mass_amu):
    """Mean mass (emu) and mass-disorder parameter g = sum_i c_i (m_i - m)^2 / m^2 of an
    isotope distribution, abundances normalised first (type_crystalstructure_io.f90:953-957,
    type_crystalstructure.f90:1505-1519); sums run in order, as Fortran's sum does."""
    mm = [float(x) * AMU_TO_EMU for x in mass_amu]
    tot = 0.0
    for c in conc:
        tot += float(c)
    c = [float(x) / tot for x in conc]
    mean = 0.0
    for ci, mi in zip(c, mm):
        mean += ci * mi
    f = 0.0
    for ci, mi in zip(c, mm):
        f += ci * (mi - mean) ** 2
    return mean, f / mean ** 2


def read_isotopes(p, path):
    """infile.isotopes: per atom n, then n lines of abundance and mass (amu)."""
    tok = open(path).read().split()
    pos = 0
    p.mass = np.zeros(p.na)
    p.mass_disorder = np.zeros(p.na)
    for a in range(p.na):
        n = int(tok[pos]); pos += 1
        conc = [float(tok[pos + 2 * j]) for j in range(n)]
        mm = [float(tok[pos + 2 * j + 1]) for j in range(n)]
        pos += 2 * n
        p.mass[a], p.mass_disorder[a] = _distribution(conc, mm)
    return p


def natural_isotopes(symbol):
    """(abundances, masses in amu) of an element's natural isotope distribution, from the
    library's table (tc3_natural_isotopes; naturaldistribution, type_crystalstructure.f90:1521-2466)."""
    import ctypes
    from .lib import dptr, load
    lib = load()
    n = ctypes.c_int(0)
    conc = np.zeros(16)
    mass = np.zeros(16)
    r = lib.tc3_natural_isotopes(symbol.encode(), 16, dptr(conc), dptr(mass), ctypes.byref(n))
    if r:
        raise ValueError(f"no isotope distributions available for {symbol}, are you sure it is a stable element?")
    return conc[:n.value].copy(), mass[:n.value].copy()


def natural_masses(p):
    """Masses and mass-disorder parameters of every atom from the natural isotope
    distribution of its element (type_crystalstructure.f90:650-652, the default without
    --readiso)."""
    p.mass = np.zeros(p.na)
    p.mass_disorder = np.zeros(p.na)
    for a in range(p.na):
        conc, mm = natural_isotopes(p.symbols[a])
        p.mass[a], p.mass_disorder[a] = _distribution(conc, mm)
    return p


def read_fc2(p, path):
    rec = _records(path)
    na = int(rec[0][0])
    if na != p.na:
        raise ValueError(f"different number of atoms in \"{path}\" and the structure")
    pos = 2
    a1s, a2s, lvs, ms = [], [], [], []
    for a1 in range(na):
        n = int(rec[pos][0]); pos += 1
        for _ in range(n):
            a2s.append(int(rec[pos][0]) - 1)
            lvs.append([float(x) for x in rec[pos + 1][:3]])
            ms.append([[float(x) for x in rec[pos + 2 + j][:3]] for j in range(3)])
            a1s.append(a1)
            pos += 5
    if int(rec[pos][0]) != 0:
        raise NotImplementedError("polar force constants (LO-TO) are not part of this path")
    lv = chop(np.array(lvs) @ p.latvec.T, 1e-13)
    m = np.array(ms) * FC2_UNIT                       # [n, i, j]
    return dict(a1=np.array(a1s, np.int32), a2=np.array(a2s, np.int32), lv2=np.ascontiguousarray(lv),
                m=m)


def read_fc3(p, path):
    rec = _records(path)
    na = int(rec[0][0])
    pos = 2
    out = {k: [] for k in ("a1", "a2", "a3", "lv2", "lv3", "m")}
    for a1 in range(na):
        n = int(rec[pos][0]); pos += 1
        for _ in range(n):
            out["a1"].append(a1)
            out["a2"].append(int(rec[pos + 1][0]) - 1)
            out["a3"].append(int(rec[pos + 2][0]) - 1)
            out["lv2"].append([float(x) for x in rec[pos + 4][:3]])
            out["lv3"].append([float(x) for x in rec[pos + 5][:3]])
            m = np.array([[float(x) for x in rec[pos + 6 + r][:3]] for r in range(9)]).reshape(3, 3, 3)
            out["m"].append(m)
            pos += 15
    fc = dict(a1=np.array(out["a1"], np.int32), a2=np.array(out["a2"], np.int32),
              a3=np.array(out["a3"], np.int32))
    fc["lv2"] = np.ascontiguousarray(chop(np.array(out["lv2"]) @ p.latvec.T, SQTOL))
    fc["lv3"] = np.ascontiguousarray(chop(np.array(out["lv3"]) @ p.latvec.T, SQTOL))
    fc["m"] = np.array(out["m"]) * FC3_UNIT          # [n, i, j, k]
    return fc


def read_fc4(p, path):
    """infile.forceconstant_fourthorder (type_forceconstant_fourthorder.f90:259-340):
    quartet arrays with Cartesian lattice vectors (Bohr) and m[n, i, j, k, l] (Ha/Bohr^4)."""
    rec = _records(path)
    na = int(rec[0][0])
    pos = 2
    out = {k: [] for k in ("a1", "a2", "a3", "a4", "lv2", "lv3", "lv4", "m")}
    for a1 in range(na):
        n = int(rec[pos][0]); pos += 1
        for _ in range(n):
            out["a1"].append(a1)
            for k, key in ((1, "a2"), (2, "a3"), (3, "a4")):
                out[key].append(int(rec[pos + k][0]) - 1)
            for k, key in ((5, "lv2"), (6, "lv3"), (7, "lv4")):
                out[key].append([float(x) for x in rec[pos + k][:3]])
            m = np.array([[float(x) for x in rec[pos + 8 + r][:3]] for r in range(27)]).reshape(3, 3, 3, 3)
            out["m"].append(m)
            pos += 35
    fc = {k: np.array(out[k], np.int32) for k in ("a1", "a2", "a3", "a4")}
    for k in ("lv2", "lv3", "lv4"):
        fc[k] = np.ascontiguousarray(chop(np.array(out[k]).reshape(-1, 3) @ p.latvec.T, SQTOL))
    fc["m"] = np.array(out["m"]).reshape(-1, 3, 3, 3, 3) * FC4_UNIT   # [n, i, j, k, l]
    return fc


def read_directory(indir, readiso=True):
    """Structure and force constants of a run directory; masses from infile.isotopes with
    readiso (--readiso), else from the natural isotope distributions (the reference's default,
    type_crystalstructure.f90:650-652, readisotopefromfile only with --readiso, main.f90:100-104)."""
    p = read_ucposcar(os.path.join(indir, "infile.ucposcar"))
    if readiso:
        read_isotopes(p, os.path.join(indir, "infile.isotopes"))
    else:
        natural_masses(p)
    fc2 = read_fc2(p, os.path.join(indir, "infile.forceconstant"))
    fc3 = read_fc3(p, os.path.join(indir, "infile.forceconstant_thirdorder"))
    return p, fc2, fc3
