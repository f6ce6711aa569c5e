"""Command-line host for the three-phonon thermal-conductivity path.

Run in a directory holding infile.ucposcar, infile.forceconstant,
infile.forceconstant_thirdorder (and infile.isotopes): reads them, runs the harmonic
step, three-phonon scattering, RTA + iterative BTE (+ off-diagonal) kappa on the GPU, and
writes outfile.thermal_conductivity, outfile.thermal_conductivity_grid.hdf5 and
outfile.cumulative_thermal_conductivity.hdf5 like the reference program
(src/thermal_conductivity/main.f90:246-340), for every requested temperature (one pass;
files of temperatures after the first get a suffix).

Options follow the reference's (src/thermal_conductivity/options.f90:44-237) for the
parts of the path built here:

    python -m tdep_amd.host -qg 12 12 12 --temperature 300 [--precision bf16x3]
"""
import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def parse(argv=None):
    ap = argparse.ArgumentParser(prog="thermal_conductivity (MI355X)", description=__doc__.split("\n\n")[0])
    ap.add_argument("-qg", "--qpoint_grid", type=int, nargs=3, default=[26, 26, 26])
    ap.add_argument("--temperature", type=float, nargs="+", default=[300.0],
                    help="one or more temperatures (K); all share one scattering pass")
    ap.add_argument("-it", "--integrationtype", type=int, default=2, choices=[1, 2, 6])
    ap.add_argument("--sigma", type=float, default=1.0,
                    help="fixed smearing (THz) for integrationtype 1; also scales the smoothing of "
                         "the cumulative kappa against mean free path")
    ap.add_argument("--max_mfp", type=float, default=-1.0, help="boundary scattering length (m)")
    ap.add_argument("-qg3ph", "--qpoint_grid3ph", type=int, nargs=3, default=[-1, -1, -1],
                    help="Monte-Carlo q-point grid of the three-phonon term")
    ap.add_argument("--seed", type=int, default=-1, help="seed of the Monte-Carlo grids")
    ap.add_argument("-qg4ph", "--qpoint_grid4ph", type=int, nargs=3, default=[-1, -1, -1],
                    help="Monte-Carlo q-point grid of the four-phonon term")
    ap.add_argument("--fourthorder", action="store_true",
                    help="add four-phonon scattering (reads infile.forceconstant_fourthorder; full q-grid)")
    ap.add_argument("--mfppts", type=int, default=1000, help="points of the mean-free-path axes")
    ap.add_argument("--freqpts", type=int, default=1000, help="points of the frequency axis")
    ap.add_argument("--dosintegrationtype", type=int, default=3, choices=[1, 2, 3],
                    help="spectral integration: 1 Gaussian, 2 adaptive Gaussian, 3 tetrahedron")
    ap.add_argument("--dossigma", type=float, default=1.0, help="smearing scale of the spectral integration")
    ap.add_argument("--unit", default="thz", choices=["thz", "mev", "icm"], help="frequency unit of the spectra")
    ap.add_argument("--iterative_tolerance", type=float, default=1e-5)
    ap.add_argument("--iterative_maxsteps", type=int, default=200)
    ap.add_argument("--classical", action="store_true")
    ap.add_argument("--readiso", action="store_true",
                    help="read the isotope distribution from infile.isotopes (default: natural isotopes)")
    ap.add_argument("--nothirdorder", action="store_true", help="do not consider three-phonon scattering")
    ap.add_argument("--nosym", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--noisotope", action="store_true", help="switch off isotope scattering")
    ap.add_argument("--notr", action="store_true", help="do not use time-reversal symmetry")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16x3"],
                    help="precision of the |Psi|^2 contraction")
    ap.add_argument("--no-offdiag", action="store_true", help="skip the off-diagonal (coherence) term")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--directory", default=".")
    return ap.parse_args(argv)


def main(argv=None):
    a = parse(argv)
    from tdep_amd.engine import Engine
    if max(a.qpoint_grid4ph) > 0 and not a.fourthorder:
        # options.f90:220-223
        print("You have to enable fourthorder to use a fourth order Monte-Carlo grid, stopping calculation.")
        return 0
    eng = Engine.from_directory(a.directory, tuple(a.qpoint_grid), device=a.device, timereversal=not a.notr,
                                readiso=a.readiso, qpsymmetry=not a.nosym)
    eng.set_isotopes(eng.structure.mass_disorder, enable=not a.noisotope)
    eng.set_thirdorder(not a.nothirdorder)
    eng.set_boundary(a.max_mfp)
    if max(a.qpoint_grid3ph) > 0 or max(a.qpoint_grid4ph) > 0:
        eng.set_montecarlo(tuple(a.qpoint_grid3ph) if max(a.qpoint_grid3ph) > 0 else None, a.seed)
    if max(a.qpoint_grid4ph) > 0:
        eng.set_montecarlo4(tuple(a.qpoint_grid4ph))
    if a.fourthorder:
        eng.set_fourthorder(a.directory)
    eng.dispersions()
    eng.scattering_rates(a.temperature, precision=a.precision, integrationtype=a.integrationtype,
                         sigma=a.sigma, classical=a.classical)
    eng.kappa(maxiter=a.iterative_maxsteps, tol=a.iterative_tolerance, offdiag=not a.no_offdiag)
    for iT, T in enumerate(a.temperature):
        sfx = "" if iT == 0 else f".T{T:g}"
        with open(os.path.join(a.directory, "outfile.thermal_conductivity" + sfx), "w") as f:
            f.write(eng.outfile_text(iT))
        eng.write_grid_hdf5(os.path.join(a.directory, "outfile.thermal_conductivity_grid" + sfx + ".hdf5"), iT)
        eng.write_cumulative_hdf5(os.path.join(a.directory, "outfile.cumulative_thermal_conductivity" + sfx + ".hdf5"),
                                  iT, mfppts=a.mfppts, freqpts=a.freqpts, dosintegrationtype=a.dosintegrationtype,
                                  sigma=a.sigma, dossigma=a.dossigma, unit=a.unit)
    t = eng.timings()
    print(f"harmonic {t['harmonic']:.3f} s, scattering {t['scatter']:.3f} s, kappa {t['kappa']:.3f} s "
          f"(BTE iterations {list(int(x) for x in eng.niter)})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
