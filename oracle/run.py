"""End-to-end oracle driver: the reference's main.f90 flow, three-phonon only.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

run(indir, qgrid, temperature) mirrors src/thermal_conductivity/main.f90:42-296
(--readiso --noisotope): read inputs, space group, q-mesh, harmonic
dispersions, scattering rates, SMA / off-diagonal / iterative kappa,
symmetrisation, and the outfile.thermal_conductivity text.
"""
import os
import numpy as np

from . import units as U
from . import lattice as L
from . import phonons as P
from . import scattering as S
from . import kappa as K


class Result:
    pass


def setup(indir, qgrid, timereversal=True):
    p = L.read_ucposcar(os.path.join(indir, "infile.ucposcar"))
    L.read_isotopes(p, os.path.join(indir, "infile.isotopes"))
    fc2 = L.read_fc2(p, os.path.join(indir, "infile.forceconstant"))
    fc3 = L.read_fc3(p, os.path.join(indir, "infile.forceconstant_thirdorder"))
    ops = L.spacegroup(p)
    mesh = L.generate_mesh(p, ops, qgrid, timereversal=timereversal)
    disp = P.generate_dispersions(p, fc2, ops, mesh)
    return p, fc2, fc3, ops, mesh, disp


def run(indir, qgrid, temperature, maxiter=200, tol=1e-5, classical=False, offdiag=True,
        integrationtype=2, sigma=1.0, pre=None, max_mfp=None):
    """max_mfp: boundary scattering length in metres (--max_mfp), None for none."""
    p, fc2, fc3, ops, mesh, disp = pre if pre is not None else setup(indir, qgrid)
    if disp.omega_min < -U.FREQTOL:
        raise RuntimeError("FOUND UNSTABLE MODES")
    r = Result()
    r.p, r.fc2, r.fc3, r.ops, r.mesh, r.disp = p, fc2, fc3, ops, mesh, disp
    sr = S.scattering_rates(p, fc3, mesh, disp, ops, temperature, classical=classical,
                            integrationtype=integrationtype, sigma=sigma,
                            mfp_max=None if max_mfp is None else max_mfp * U.M_TO_BOHR)
    r.sr = sr
    r.kappa_sma = K.get_kappa(p, mesh, disp, ops, sr.F0, temperature, classical)
    r.kappa_offdiag = (K.get_kappa_offdiag(p, fc2, mesh, disp, ops, sr.linewidth,
                                           temperature, classical)
                       if offdiag else np.zeros((3, 3)))
    if maxiter > 0:
        Fn, scf, hist = K.iterative_solution(p, mesh, disp, ops, sr, temperature, maxiter,
                                             tol, classical)
    else:
        Fn, scf, hist = sr.F0.copy(), [], []
    r.Fn, r.scf, r.kappa_hist = Fn, scf, hist
    r.kappa_iter = K.get_kappa(p, mesh, disp, ops, Fn, temperature, classical)
    r.kappa_iter = K.symmetrize_kappa(r.kappa_iter, ops)
    r.kappa_offdiag = K.symmetrize_kappa(r.kappa_offdiag, ops)
    r.kappa_sma = K.symmetrize_kappa(r.kappa_sma, ops)
    r.temperature = temperature
    return r


def fortran_e(x, w=24, d=12):
    """Fortran Ew.d edit descriptor (gfortran style: 0.ddddE+xx)."""
    if x == 0.0 or not np.isfinite(x):
        s = "0." + "0" * d + "E+00"
    else:
        e = int(np.floor(np.log10(abs(x)))) + 1
        m = abs(x) / 10.0 ** e
        ms = f"{m:.{d}f}"
        if ms.startswith("1."):
            e += 1
            m = abs(x) / 10.0 ** e
            ms = f"{m:.{d}f}"
        s = ms + ("E+" if e >= 0 else "E-") + f"{abs(e):02d}"
        if x < 0:
            s = "-" + s
    return s.rjust(w)


def outfile_text(r):
    """main.f90:248-294."""
    def six(m):
        m = m * U.KAPPA_AU_TO_SI
        v = [m[0, 0], m[1, 1], m[2, 2], m[0, 1], m[0, 2], m[1, 2]]
        return " " + "".join(" " + fortran_e(x) for x in v) + "\n"
    hdr = "#" + "".join(" " + s.rjust(24) for s in ("kxx", "kyy", "kzz", "kxy", "kxz", "kyz")) + "\n"
    t = "# Unit:" + " " * 15 + "W/m/K\n"
    t += "# Temperature:" + " " * 8 + fortran_e(r.temperature, 20, 12) + "\n"
    t += "# Single mode approximation\n" + hdr + six(r.kappa_sma)
    t += "# Collective contribution\n" + hdr + six(r.kappa_iter - r.kappa_sma)
    t += "# Off diagonal (coherence) contribution\n" + hdr + six(r.kappa_offdiag)
    t += "# Total thermal conductivity\n" + hdr + six(r.kappa_iter + r.kappa_offdiag)
    return t
