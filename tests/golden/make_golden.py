"""Write the oracle's FP64 results for small workloads as committed fixtures.

Run from the repository root:  python tests/golden/make_golden.py
The inputs are regenerated from tdep_amd.inputs with the default seed; each fixture
stores the q-grid, temperature and the oracle's frequencies, velocities, linewidths
and kappa tensors (atomic units)."""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import run as R  # noqa: E402
from tdep_amd import inputs  # noqa: E402

CASES = [("fcc_al", (6, 6, 6), (300.0,)), ("rocksalt", (4, 4, 4), (100.0, 300.0)),
         ("sl16", (4, 4, 2), (300.0,)),
         # the monatomic FCC workload at its full 12x12x12 mesh, and the 32-atom superlattice
         # (96 bands) on a small in-plane mesh
         ("fcc_al", (12, 12, 12), (300.0,)), ("sl32", (4, 4, 1), (300.0,))]


# four-phonon scattering (--fourthorder, full grids) on the 16-atom superlattice (48 bands):
# linewidths with the three- and four-phonon terms, and twelve rows of the scattering matrix
FOUR_CASES = [("sl16", (3, 3, 1), (300.0,))]


def main_fourth(only=None):
    from oracle import lattice as L, scattering as S
    out = os.path.join(ROOT, "tests", "golden")
    for name, grid, temps in FOUR_CASES:
        tag = f"fourth_{name}_{'x'.join(map(str, grid))}"
        if only and tag not in only:
            continue
        with tempfile.TemporaryDirectory() as d:
            inputs.make_config(name, d, fourth=True)
            p, fc2, fc3, ops, mesh, disp = R.setup(d, grid)
            fc4 = L.read_fc4(p, os.path.join(d, "infile.forceconstant_fourthorder"))
            rec = dict(grid=np.array(grid), temperatures=np.array(temps), om_irr=disp.om_irr)
            for iT, T in enumerate(temps):
                sr = S.scattering_rates(p, fc3, mesh, disp, ops, T, fc4=fc4)
                rec[f"linewidth_{iT}"] = sr.linewidth
                pick = np.linspace(0, len(sr.rows) - 1, 12).astype(np.int64)
                rec[f"xi_rows_{iT}"] = pick
                rec[f"xi_{iT}"] = sr.Xi[pick]
                rec[f"rows_{iT}"] = np.array(sr.rows)
            fn = os.path.join(out, f"oracle_{tag}.npz")
            np.savez_compressed(fn, **rec)
            print(fn, {k: np.shape(v) for k, v in rec.items()})


def main(only=None):
    """only: fixture names (e.g. "sl32_2x2x2") to write; all when empty."""
    main_fourth(only)
    out = os.path.join(ROOT, "tests", "golden")
    for name, grid, temps in CASES:
        if only and f"{name}_{'x'.join(map(str, grid))}" not in only:
            continue
        with tempfile.TemporaryDirectory() as d:
            inputs.make_config(name, d)
            pre = R.setup(d, grid)
            rec = dict(grid=np.array(grid), temperatures=np.array(temps), om_irr=pre[5].om_irr,
                       vel_irr=pre[5].vel_irr, nirr=pre[4].nirr)
            for iT, T in enumerate(temps):
                r = R.run(d, grid, T, pre=pre)
                rec[f"linewidth_{iT}"] = r.sr.linewidth
                rec[f"kappa_sma_{iT}"] = r.kappa_sma
                rec[f"kappa_iter_{iT}"] = r.kappa_iter
                rec[f"kappa_offdiag_{iT}"] = r.kappa_offdiag
                rec[f"niter_{iT}"] = len(r.scf)
            fn = os.path.join(out, f"oracle_{name}_{'x'.join(map(str, grid))}.npz")
            np.savez_compressed(fn, **rec)
            print(fn, {k: np.shape(v) for k, v in rec.items()})


if __name__ == "__main__":
    main(sys.argv[1:])
