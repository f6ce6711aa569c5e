"""Synthetic TDEP inputs from Morse pair potentials.

Writes infile.ucposcar, infile.isotopes, infile.forceconstant and
infile.forceconstant_thirdorder in the formats the reference reads
(type_crystalstructure_io.f90:9-350, type_crystalstructure_io.f90:921-959,
type_forceconstant_secondorder_io.f90:109-247,
type_forceconstant_thirdorder.f90:243-311), with the force constants given by
the analytic second and third derivatives of a Morse pair potential

    phi(r) = D [exp(-2 a (r - r0)) - 2 exp(-a (r - r0))].

For a pair potential only clusters with a repeated atom carry third-order
terms: (a,a,a), (a,a,b), (a,b,a) and (a,b,b).  All of them are written, and
the self terms follow from the acoustic sum rule.  D carries a +-1% jitter
drawn once per symmetry shell (distance, unordered species pair) from the seed,
so the jitter never lowers the crystal symmetry.

The five workload families (FCC Al, rocksalt NaCl, 16- and 32-atom FCC mass
superlattices) are available through `make_config`, with the q-mesh and
temperatures of each family.
"""
import os
import numpy as np

A_TO_BOHR = 1.0 / 0.52917721067


class Morse:
    def __init__(self, D, alpha, r0):
        self.D, self.alpha, self.r0 = D, alpha, r0

    def derivs(self, r, D=None):
        """phi', phi'', phi''' at r (eV/A^n)."""
        D = self.D if D is None else D
        a, r0 = self.alpha, self.r0
        e1 = np.exp(-a * (r - r0))
        e2 = e1 * e1
        d1 = D * (-2 * a * e2 + 2 * a * e1)
        d2 = D * (4 * a * a * e2 - 2 * a * a * e1)
        d3 = D * (-8 * a ** 3 * e2 + 2 * a ** 3 * e1)
        return d1, d2, d3

    def derivs4(self, r, D=None):
        """Fourth derivative of phi at r (eV/A^4)."""
        D = self.D if D is None else D
        a, r0 = self.alpha, self.r0
        e1 = np.exp(-a * (r - r0))
        return D * (16 * a ** 4 * e1 * e1 - 2 * a ** 4 * e1)


class MorseDerivative(Morse):
    """Derivatives of (phi', phi'', phi''') of a Morse potential with respect to one of its
    parameters, "D", "alpha" or "r0" (the D given to derivs includes the shell's jitter
    factor, so the derivative with respect to D scales by it)."""

    def __init__(self, base, which):
        super().__init__(base.D, base.alpha, base.r0)
        self.which = which

    def derivs(self, r, D=None):
        D = self.D if D is None else D
        a, x = self.alpha, r - self.r0
        e1 = np.exp(-a * x)
        e2 = e1 * e1
        if self.which == "D":
            return tuple(v / self.D for v in Morse.derivs(self, r, D))
        if self.which == "alpha":
            return (D * (-2 * e2 + 2 * e1 + 4 * a * x * e2 - 2 * a * x * e1),
                    D * (8 * a * e2 - 4 * a * e1 - 8 * a * a * x * e2 + 2 * a * a * x * e1),
                    D * (-24 * a * a * e2 + 6 * a * a * e1 + 16 * a ** 3 * x * e2 - 2 * a ** 3 * x * e1))
        if self.which == "r0":
            return (D * (-4 * a * a * e2 + 2 * a * a * e1),
                    D * (8 * a ** 3 * e2 - 2 * a ** 3 * e1),
                    D * (-16 * a ** 4 * e2 + 2 * a ** 4 * e1))
        raise KeyError(self.which)


class ZeroPotential(Morse):
    """A pair whose potential does not depend on the parameter (the other pair classes)."""

    def __init__(self, base):
        super().__init__(base.D, base.alpha, base.r0)

    def derivs(self, r, D=None):
        return 0.0, 0.0, 0.0


def pair_tensors(R, d1, d2, d3):
    """Second (H_ij) and third (T_ijk) Cartesian derivatives of phi(|R|)."""
    r = np.linalg.norm(R)
    u = R / r
    I = np.eye(3)
    H = d2 * np.outer(u, u) + d1 / r * (I - np.outer(u, u))
    c = d2 / r - d1 / r ** 2
    uuu = np.einsum("i,j,k->ijk", u, u, u)
    sym = (np.einsum("ij,k->ijk", I, u) + np.einsum("ik,j->ijk", I, u)
           + np.einsum("jk,i->ijk", I, u))
    T = d3 * uuu + c * (sym - 3 * uuu)
    return H, T


def pair_tensor4(R, d1, d2, d3, d4):
    """Fourth Cartesian derivative of phi(|R|) from its radial derivatives d1..d4: with
    A_1 = phi'/r and A_(n+1) = (1/r) dA_n/dr,
    Q_abcd = (d_ab d_cd + 2 perms) A_2 + (d_ab R_c R_d + 5 perms) A_3 + R_a R_b R_c R_d A_4."""
    r = np.linalg.norm(R)
    A2 = d2 / r ** 2 - d1 / r ** 3
    A3 = d3 / r ** 3 - 3 * d2 / r ** 4 + 3 * d1 / r ** 5
    A4 = d4 / r ** 4 - 6 * d3 / r ** 5 + 15 * d2 / r ** 6 - 15 * d1 / r ** 7
    I = np.eye(3)
    dd = np.einsum("ab,cd->abcd", I, I)
    dd = dd + dd.transpose(0, 2, 1, 3) + dd.transpose(0, 3, 2, 1)
    dRR = np.einsum("ab,c,d->abcd", I, R, R)
    six = (dRR + dRR.transpose(0, 2, 1, 3) + dRR.transpose(0, 2, 3, 1) + dRR.transpose(2, 0, 1, 3)
           + dRR.transpose(2, 0, 3, 1) + dRR.transpose(2, 3, 0, 1))
    return dd * A2 + six * A3 + np.einsum("a,b,c,d->abcd", R, R, R, R) * A4


class Crystal:
    """Lattice (columns, Angstrom), fractional positions, element labels, masses (amu)."""

    def __init__(self, lat, frac, symbols, masses):
        self.lat = np.array(lat, dtype=np.float64)
        self.frac = np.array(frac, dtype=np.float64)
        self.symbols = list(symbols)
        self.masses = np.array(masses, dtype=np.float64)
        self.na = len(self.symbols)


def fcc_primitive(a, symbols, masses, basis):
    lat = a * np.array([[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]).T
    return Crystal(lat, basis, symbols, masses)


def fcc_superlattice(a, nrep, layer_mass):
    """Conventional 4-atom cubic FCC cell stacked nrep times along z.  Atomic layer l
    (2 per cubic cell) gets mass layer_mass[l].  Element labels encode the mass, the
    in-plane site in the layers through the origin, and whether the atom sits in the
    first cubic cell, so the labelled cell is primitive (no (1/2,1/2,0) centring, no
    translation along z) and keeps the 4mm point group."""
    lat = np.diag([a, a, a * nrep])
    frac, syms, masses = [], [], []
    labels = {}
    pool = ["Ar", "Ca", "Kr", "Br", "Si", "Al", "Fe", "Co", "Ni", "Cu", "Zn", "Ga",
            "Ge", "As", "Se", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Ru", "Rh", "Pd"]
    for c in range(nrep):
        for half, sites in ((0, [(0.0, 0.0), (0.5, 0.5)]), (1, [(0.5, 0.0), (0.0, 0.5)])):
            l = 2 * c + half
            z = (c + 0.5 * half) / nrep
            for s, (x, y) in enumerate(sites):
                # even layers: one label per in-plane site (breaks the (1/2,1/2,0)
                # centring); odd layers: one label for both sites (keeps the 4-fold
                # axis); the first cubic cell gets its own labels (breaks any pure
                # translation along z), so the labelled cell is primitive
                key = (layer_mass[l], s if half == 0 else 2, c == 0)
                if key not in labels:
                    labels[key] = pool[len(labels)]
                frac.append((x, y, z))
                syms.append(labels[key])
                masses.append(layer_mass[l])
    # POSCAR groups atoms by species: order atoms by first appearance of their label
    order = sorted(range(len(syms)), key=lambda i: (list(labels.values()).index(syms[i]), i))
    return Crystal(lat, np.array(frac)[order], [syms[i] for i in order],
                   np.array(masses)[order])


def _neighbours(cr, cutoff):
    """All (a1, a2, L) with 0 < |r2 + L - r1| < cutoff (Angstrom)."""
    n = [int(np.ceil(cutoff / np.linalg.norm(cr.lat[:, i]))) + 1 for i in range(3)]
    out = []
    cart = cr.frac @ cr.lat.T
    for a1 in range(cr.na):
        for a2 in range(cr.na):
            for i in range(-n[0], n[0] + 1):
                for j in range(-n[1], n[1] + 1):
                    for k in range(-n[2], n[2] + 1):
                        L = np.array([i, j, k], dtype=np.float64)
                        R = cart[a2] + cr.lat @ L - cart[a1]
                        d = np.linalg.norm(R)
                        if 1e-8 < d < cutoff:
                            out.append((a1, a2, L, R, d))
    return out


def morse_ifcs(cr, potential, cutoff, seed=0, jitter=0.01, fourth=False):
    """Second- and third-order IFC lists (eV/A^2, eV/A^3) for a pair potential, and with
    fourth=True the fourth-order list (eV/A^4) as a third value.
    potential(sym1, sym2) -> Morse."""
    nb = _neighbours(cr, cutoff)
    rng = np.random.default_rng(seed)
    shells = {}
    keys = []
    for (a1, a2, L, R, d) in nb:
        key = (round(d, 6), tuple(sorted((cr.symbols[a1], cr.symbols[a2]))))
        keys.append(key)
    for key in sorted(set(keys)):
        shells[key] = 1.0 + jitter * (2.0 * rng.random() - 1.0)
    fc2 = [[] for _ in range(cr.na)]
    fc3 = [[] for _ in range(cr.na)]
    self2 = [np.zeros((3, 3)) for _ in range(cr.na)]
    self3 = [np.zeros((3, 3, 3)) for _ in range(cr.na)]
    fc4 = [[] for _ in range(cr.na)]
    self4 = [np.zeros((3, 3, 3, 3)) for _ in range(cr.na)]
    z3 = np.zeros(3)
    for (a1, a2, L, R, d), key in zip(nb, keys):
        pot = potential(cr.symbols[a1], cr.symbols[a2])
        d1, d2, d3 = pot.derivs(d, pot.D * shells[key])
        H, T = pair_tensors(R, d1, d2, d3)
        fc2[a1].append((a2, L, -H))
        self2[a1] += H
        # clusters with a repeated atom: (a1,b,b) -T, (a1,a1,b) +T, (a1,b,a1) +T
        fc3[a1].append((a1, a2, a2, z3, L, L, -T))
        fc3[a1].append((a1, a1, a2, z3, z3, L, T))
        fc3[a1].append((a1, a2, a1, z3, L, z3, T))
        self3[a1] -= T
        if fourth:
            # clusters of a1 and b, a1 first: (-1)^(number of a1) Q
            Q = pair_tensor4(R, d1, d2, d3, pot.derivs4(d, pot.D * shells[key]))
            for s2 in (0, 1):
                for s3 in (0, 1):
                    for s4 in (0, 1):
                        if s2 + s3 + s4 == 0:
                            continue
                        n1 = 4 - (s2 + s3 + s4)
                        at = [a2 if s else a1 for s in (s2, s3, s4)]
                        lv = [L if s else z3 for s in (s2, s3, s4)]
                        fc4[a1].append((a1, at[0], at[1], at[2], z3, lv[0], lv[1], lv[2], Q * (-1.0) ** n1))
            self4[a1] += Q
    for a in range(cr.na):
        fc2[a].insert(0, (a, z3, self2[a]))
        fc3[a].insert(0, (a, a, a, z3, z3, z3, self3[a]))
        fc4[a].insert(0, (a, a, a, a, z3, z3, z3, z3, self4[a]))
    if fourth:
        return fc2, fc3, fc4
    return fc2, fc3


def write_fourthorder(cr, fc4, cutoff, path):
    """infile.forceconstant_fourthorder (type_forceconstant_fourthorder.f90:259-299)."""
    with open(os.path.join(path, "infile.forceconstant_fourthorder"), "w") as f:
        f.write(f"{cr.na:11d}               How many atoms per unit cell\n")
        f.write(f"{cutoff:21.15f}     Realspace cutoff (A)\n")
        for a in range(cr.na):
            f.write(f"{len(fc4[a]):11d}               How many quartets are atom {a + 1:3d} part of\n")
            for n, (i1, i2, i3, i4, l1, l2, l3, l4, m) in enumerate(fc4[a]):
                for k, ii in enumerate((i1, i2, i3, i4)):
                    f.write(f"{ii + 1:11d}                atom {k + 1} in quartet {n + 1}\n")
                for L in (l1, l2, l3, l4):
                    f.write(" ".join(f"{x:25.16e}" for x in L) + "\n")
                for ii in range(3):
                    for jj in range(3):
                        for kk in range(3):
                            f.write(" ".join(f"{x:25.16e}" for x in m[ii, jj, kk]) + "\n")


def write_inputs(cr, fc2, fc3, cutoff, path):
    os.makedirs(path, exist_ok=True)
    species = []
    for s in cr.symbols:
        if s not in species:
            species.append(s)
    counts = [cr.symbols.count(s) for s in species]
    with open(os.path.join(path, "infile.ucposcar"), "w") as f:
        f.write("generated Morse crystal\n1.0\n")
        for i in range(3):
            f.write(" ".join(f"{x:.16f}" for x in cr.lat[:, i]) + "\n")
        f.write(" ".join(species) + "\n")
        f.write(" ".join(str(c) for c in counts) + "\n")
        f.write("direct coordinates\n")
        for r in cr.frac:
            f.write(" ".join(f"{x:.16f}" for x in r) + "\n")
    with open(os.path.join(path, "infile.isotopes"), "w") as f:
        for m in cr.masses:
            f.write("1\n1.0 %.10f\n" % m)
    with open(os.path.join(path, "infile.forceconstant"), "w") as f:
        f.write(f"{cr.na:11d}               How many atoms per unit cell\n")
        f.write(f"{cutoff:21.15f}     Realspace cutoff (A)\n")
        for a in range(cr.na):
            f.write(f"{len(fc2[a]):11d}               How many neighbours does atom {a + 1:3d} have\n")
            for n, (a2, L, m) in enumerate(fc2[a]):
                f.write(f"{a2 + 1:11d}               In the unit cell, what is the index of neighbour {n + 1} of atom {a + 1}\n")
                f.write(" ".join(f"{x:25.16e}" for x in L) + "\n")
                for row in m:
                    f.write(" ".join(f"{x:25.16e}" for x in row) + "\n")
        f.write(" 0   # no Born charges or dielectric tensor\n")
    with open(os.path.join(path, "infile.forceconstant_thirdorder"), "w") as f:
        f.write(f"{cr.na:11d}               How many atoms per unit cell\n")
        f.write(f"{cutoff:21.15f}     Realspace cutoff (A)\n")
        for a in range(cr.na):
            f.write(f"{len(fc3[a]):11d}               How many triplets are atom {a + 1:3d} part of\n")
            for n, (i1, i2, i3, l1, l2, l3, m) in enumerate(fc3[a]):
                for k, ii in enumerate((i1, i2, i3)):
                    f.write(f"{ii + 1:11d}                atom {k + 1} in triplet {n + 1}\n")
                for L in (l1, l2, l3):
                    f.write(" ".join(f"{x:25.16e}" for x in L) + "\n")
                for ii in range(3):
                    for jj in range(3):
                        f.write(" ".join(f"{x:25.16e}" for x in m[ii, jj]) + "\n")


# ----------------------------------------------------------------------------
# The workload families
# ----------------------------------------------------------------------------
MORSE_AL = Morse(0.30, 1.40, 2.864)


def _rocksalt_potential(s1, s2):
    if s1 != s2:
        return Morse(0.45, 1.30, 2.82)
    return Morse(0.10, 1.10, 3.99)


CONFIGS = {
    # FCC monatomic, 1-atom primitive cell, 12x12x12, 300 K
    "fcc_al": dict(qgrid=(12, 12, 12), temperatures=(300.0,)),
    # rocksalt binary, 2-atom primitive cell, 20x20x20, 100/300/500 K
    "rocksalt": dict(qgrid=(20, 20, 20), temperatures=(100.0, 300.0, 500.0)),
    # 16-atom FCC mass superlattice, 10x10x4, 7 temperatures
    "sl16": dict(qgrid=(10, 10, 4), temperatures=(100.0, 200.0, 300.0, 400.0, 500.0, 600.0, 700.0)),
    # rocksalt adjoint workload, 16x16x16, 300 K
    "rocksalt_adjoint": dict(qgrid=(16, 16, 16), temperatures=(300.0,)),
    # 32-atom FCC superlattice, 8x8x2, 300 K
    "sl32": dict(qgrid=(8, 8, 2), temperatures=(300.0,)),
    # 4-atom cubic FCC cell, two atomic layers of 40 and 80 amu (12 bands): four-phonon tests
    # beyond two atoms
    "sl4": dict(qgrid=(3, 3, 3), temperatures=(300.0,)),
}


MORSE_FIELDS = ("D", "alpha", "r0")


def morse_parameters(name):
    """The Morse parameters of a family as a vector, (D, alpha, r0) per pair class (rocksalt:
    unlike pairs, then like pairs), their labels, potential(v) -> potential(s1, s2), and
    derivative(v, k) -> the potential(s1, s2) of the derivatives with respect to v[k]."""
    if name in ("rocksalt", "rocksalt_adjoint"):
        p = np.array([0.45, 1.30, 2.82, 0.10, 1.10, 3.99])
        labels = [f"{c}_{f}" for c in ("unlike", "like") for f in MORSE_FIELDS]
        cls = lambda s1, s2: 0 if s1 != s2 else 1
    else:
        p = np.array([MORSE_AL.D, MORSE_AL.alpha, MORSE_AL.r0])
        labels = list(MORSE_FIELDS)
        cls = lambda s1, s2: 0

    def potential(v):
        v = np.asarray(v, dtype=np.float64)
        return lambda s1, s2: Morse(*v[3 * cls(s1, s2):3 * cls(s1, s2) + 3])

    def derivative(v, k):
        base = potential(v)
        return lambda s1, s2: (MorseDerivative(base(s1, s2), MORSE_FIELDS[k % 3]) if cls(s1, s2) == k // 3
                               else ZeroPotential(base(s1, s2)))
    return p, labels, potential, derivative


def make_crystal(name, seed=0):
    """Crystal, pair potential, cutoff (A) of a workload family."""
    if name == "fcc_al":
        cr = fcc_primitive(4.05, ["Al"], [26.98], [[0.0, 0.0, 0.0]])
        return cr, (lambda s1, s2: MORSE_AL), 5.0
    if name in ("rocksalt", "rocksalt_adjoint"):
        cr = fcc_primitive(5.64, ["Na", "Cl"], [22.99, 35.45], [[0, 0, 0], [0.5, 0.5, 0.5]])
        return cr, _rocksalt_potential, 6.0
    if name == "sl16":
        lm = [40.0 if (l // 2) % 2 == 0 else 80.0 for l in range(8)]
        return fcc_superlattice(4.05, 4, lm), (lambda s1, s2: MORSE_AL), 5.0
    if name == "sl4":
        return fcc_superlattice(4.05, 1, [40.0, 80.0]), (lambda s1, s2: MORSE_AL), 5.0
    if name == "sl32":
        rng = np.random.default_rng(4)
        lm = [float(x) for x in rng.choice([28.0, 56.0, 84.0], size=16)]
        return fcc_superlattice(4.05, 8, lm), (lambda s1, s2: MORSE_AL), 6.0
    raise KeyError(name)


def make_config(name, path, seed=0, fourth=False, params=None):
    """Write the four input files of a workload family into `path` (and with fourth=True
    infile.forceconstant_fourthorder); returns its dict.  params: the family's Morse
    parameter vector (morse_parameters) in place of the default."""
    cr, pot, cutoff = make_crystal(name, seed)
    if params is not None:
        pot = morse_parameters(name)[2](params)
    if fourth:
        fc2, fc3, fc4 = morse_ifcs(cr, pot, cutoff, seed=seed, fourth=True)
    else:
        fc2, fc3 = morse_ifcs(cr, pot, cutoff, seed=seed)
    write_inputs(cr, fc2, fc3, cutoff, path)
    if fourth:
        write_fourthorder(cr, fc4, cutoff, path)
    return dict(CONFIGS[name], name=name, natoms=cr.na)


def morse_ifc_derivatives(name, params, k, seed=0):
    """Derivatives of the force-constant lists of a family with respect to params[k], lists as
    morse_ifcs returns them (eV/A^2, eV/A^3), entry for entry in the order of the files
    make_config writes."""
    cr, _, cutoff = make_crystal(name, seed)
    return morse_ifcs(cr, morse_parameters(name)[3](params, k), cutoff, seed=seed)
