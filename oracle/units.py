"""Unit constants and small math helpers of the reference, restated.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Constants follow src/libolle/konstanter.f90:26-125 (NIST 2017 values as the
reference states them); helpers follow the cited gottochblandat*.f90 lines.
"""
import numpy as np

# konstanter.f90:26-28
PI = 3.141592653589793
TWOPI = 6.283185307179586
# konstanter.f90:36-66
EV_TO_JOULE = 1.6021766208e-19
HARTREE_TO_EV = 27.21138602
EV_TO_HARTREE = 1.0 / HARTREE_TO_EV
HARTREE_TO_JOULE = HARTREE_TO_EV * EV_TO_JOULE
BOHR_TO_M = 0.52917721067e-10
M_TO_BOHR = 1.0 / BOHR_TO_M
BOHR_TO_A = 0.52917721067
A_TO_BOHR = 1.0 / BOHR_TO_A
HBAR_EV = 6.582119514e-16
HBAR_JOULE = HBAR_EV * EV_TO_JOULE
HBAR_HARTREE = HBAR_EV * EV_TO_HARTREE
KB_EV = 8.6173303e-5
KB_HARTREE = KB_EV * EV_TO_HARTREE
AMU_TO_KG = 1.660539040e-27
EMU_TO_KG = 9.10938356e-31
AMU_TO_EMU = AMU_TO_KG / EMU_TO_KG
EMU_TO_AMU = 1.0 / AMU_TO_EMU
# konstanter.f90:78-109
TIME_AU_TO_S = HBAR_JOULE / HARTREE_TO_JOULE
FC2_EVA_TO_HARTREEBOHR = EV_TO_HARTREE / (A_TO_BOHR ** 2)
FC3_EVA_TO_HARTREEBOHR = EV_TO_HARTREE / (A_TO_BOHR ** 3)
FC4_EVA_TO_HARTREEBOHR = EV_TO_HARTREE / (A_TO_BOHR ** 4)
FREQ_HARTREE_TO_HZ = 1.0 / HBAR_HARTREE
FREQ_HARTREE_TO_THZ = 1e-12 / TWOPI / HBAR_HARTREE
FREQ_THZ_TO_HARTREE = 1.0 / FREQ_HARTREE_TO_THZ
GROUPVEL_HARTREEBOHR_TO_MS = BOHR_TO_M / TIME_AU_TO_S
KAPPA_AU_TO_SI = HARTREE_TO_JOULE / TIME_AU_TO_S / BOHR_TO_M
# konstanter.f90:113-135
TOL = 1e-5
SQTOL = TOL ** 2
FREQTOL = TOL * 1e-4
GROUPVELTOL = TOL * 1e-5      # lo_phonongroupveltol
TEMPERATURETOL = 1e-3
DEGENVECTOR = np.array([1.0, 1.33, 1.45624623]) * 1e-8

# gottochblandat.f90:1286-1374 -- the reference's list of "well defined" values
_WELLDEFINED = np.array([
    0.0, 0.1, -0.1, 0.1111111111111111, -0.1111111111111111, 0.125, -0.125,
    0.1428571428571428, -0.1428571428571428, 0.1666666666666667, -0.1666666666666667,
    0.2, -0.2, 0.2222222222222222, -0.2222222222222222, 0.25, -0.25,
    0.2857142857142857, -0.2857142857142857, 0.3, -0.3, 0.3333333333333333,
    -0.3333333333333333, 0.375, -0.375, 0.4, -0.4, 0.4285714285714285,
    -0.4285714285714285, 0.4444444444444444, -0.4444444444444444, 0.5, -0.5,
    0.5555555555555556, -0.5555555555555556, 0.5714285714285714, -0.5714285714285714,
    0.6, -0.6, 0.625, -0.625, 0.6666666666666667, -0.6666666666666667, 0.7, -0.7,
    0.7071067811865475, -0.7071067811865475, 0.7142857142857143, -0.7142857142857143,
    0.75, -0.75, 0.7777777777777778, -0.7777777777777778, 0.8, -0.8,
    0.8333333333333334, -0.8333333333333334, 0.8571428571428571, -0.8571428571428571,
    0.8660254037844386, -0.8660254037844386, 0.875, -0.875, 0.8888888888888888,
    -0.8888888888888888, 0.9, -0.9, 1.0, -1.0])


def chop(x, tol):
    """lo_chop_real (gottochblandat.f90:1286-1374): snap to the first listed
    well-defined value within tol, in list order."""
    x = np.array(x, dtype=np.float64, copy=True)
    flat = x.reshape(-1)
    for n in range(flat.size):
        d = np.abs(flat[n] - _WELLDEFINED) < tol
        if d.any():
            flat[n] = _WELLDEFINED[np.argmax(d)]
    return x


def chop_complex(z, tol):
    """lo_chop_complex (gottochblandat.f90:1377-1392): zero tiny parts."""
    re = np.where(np.abs(z.real) < tol, 0.0, z.real)
    im = np.where(np.abs(z.imag) < tol, 0.0, z.imag)
    return re + 1j * im


def gauss(x, mu, sigma):
    """lo_gauss (gottochblandat_calculus.f90:268-288), 8-sigma cutoff."""
    x = np.asarray(x, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    sigma = np.asarray(sigma, dtype=np.float64)
    d = x - mu
    inv = 1.0 / sigma
    g = np.exp(-d * d * 0.5 * inv * inv) * 0.398942280401433 * inv
    return np.where(np.abs(d) > 8.0 * sigma, 0.0, g)


def planck(temperature, omega):
    """lo_planck (gottochblandat_physics.f90:160-188)."""
    omega = np.asarray(omega, dtype=np.float64)
    if temperature < TEMPERATURETOL:
        return np.zeros_like(omega)
    x = omega / (KB_HARTREE * temperature)
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        n = 1.0 / (np.exp(np.minimum(x, 700.0)) - 1.0)
    n = np.where(x > 100.0, 0.0, n)
    return np.where(omega < FREQTOL, 0.0, n)


def harmonic_oscillator_cv(temperature, omega):
    """lo_harmonic_oscillator_cv (gottochblandat_physics.f90:331-367)."""
    omega = np.asarray(omega, dtype=np.float64)
    if temperature < TEMPERATURETOL:
        return np.zeros_like(omega)
    x = omega / temperature / KB_HARTREE
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        xs = np.minimum(x, 700.0)
        mid = KB_HARTREE * xs * xs * np.exp(xs) / (np.exp(xs) - 1.0) ** 2
        low = KB_HARTREE * xs * xs * np.exp(-xs)
        high = KB_HARTREE * (1.0 - x * x / 12.0 + x ** 4 / 240.0)
    cv = np.where(x > 200.0, 0.0, np.where(x > 21.0, low, np.where(x < 2e-2, high, mid)))
    return np.where(omega < FREQTOL, 0.0, cv)


def negsqrt(x):
    """lo_negsqrt (gottochblandat_linalg.f90:7-11)."""
    x = np.asarray(x, dtype=np.float64)
    return np.sqrt(np.abs(x)) * np.where(x >= 0.0, 1.0, -1.0)


def clean_fractional(x, tol=SQTOL):
    """lo_clean_fractional_coordinates (gottochblandat.f90:1261-1283)."""
    y = np.array(x, dtype=np.float64, copy=True)
    flat = y.reshape(-1)
    for n in range(flat.size):
        v = flat[n]
        while True:
            if abs(v) < tol:
                v = 0.0
            if abs(v - 1.0) < tol:
                v = 0.0
            if v > 1.0:
                v -= 1.0
            if v < 0.0:
                v += 1.0
            if -tol < v < 1.0:
                break
        flat[n] = v
    return y


# Periodic table symbols, used only to turn POSCAR species labels into species ids
# (type_crystalstructure_atomdata.f90 symbol_to_z maps symbols to Z the same way).
SYMBOLS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn "
    "Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce "
    "Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn "
    "Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh Fl "
    "Mc Lv Ts Og").split()


def symbol_to_z(sym):
    s = sym.strip()
    if s not in SYMBOLS:
        raise ValueError(f"unknown element symbol '{sym}'")
    return SYMBOLS.index(s) + 1
