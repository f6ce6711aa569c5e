"""Finite-difference derivatives of the oracle's FP64 results with respect to the six Morse
parameters of the rocksalt family, as a committed fixture.

Run from the repository root:  python tests/golden/make_morse_golden.py
For each parameter p_k the inputs are regenerated at p_k (1 +- h) and the oracle's
frequencies, velocities, linewidths and single-mode kappa are differenced centrally at
h = 4e-6, 2e-6 and 1e-6 (relative); the fixture holds the Richardson combination of the two
smallest steps, (4 D(1e-6) - D(2e-6)) / 3, and a mask of the linewidths whose three central
differences disagree by more than 1e-3 of the largest one (a threshold of the forward, such as
the lo_freqtol tests of near-degenerate modes, crossed between the steps).
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import run as R  # noqa: E402
from tdep_amd import inputs  # noqa: E402

NAME, GRID, T = "rocksalt", (6, 6, 6), 300.0
STEPS = (4e-6, 2e-6, 1e-6)


def evaluate(params):
    with tempfile.TemporaryDirectory() as d:
        inputs.make_config(NAME, d, params=params)
        r = R.run(d, GRID, T, maxiter=0, offdiag=False)
        return dict(linewidth=r.sr.linewidth, kappa_sma=r.kappa_sma, omega=r.disp.om_irr, velocity=r.disp.vel_irr)


def main():
    p0, labels, _, _ = inputs.morse_parameters(NAME)
    keys = ("linewidth", "kappa_sma", "omega", "velocity")
    out = {k: [] for k in keys}
    mask = []
    for k in range(len(p0)):
        fd = {}
        for h in STEPS:
            pp, pm = p0.copy(), p0.copy()
            pp[k] *= 1 + h
            pm[k] *= 1 - h
            a, b = evaluate(pp), evaluate(pm)
            fd[h] = {key: (a[key] - b[key]) / (2 * h * p0[k]) for key in keys}
        for key in keys:
            out[key].append((4 * fd[STEPS[2]][key] - fd[STEPS[1]][key]) / 3)
        lw = np.stack([fd[h]["linewidth"] for h in STEPS])
        spread = lw.max(axis=0) - lw.min(axis=0)
        mask.append(spread <= 1e-3 * np.abs(lw).max())
        print(labels[k], "linewidths kept", int(mask[-1].sum()), "of", mask[-1].size, flush=True)
    rec = {f"d{key}": np.stack(v) for key, v in out.items()}
    rec.update(params=p0, grid=np.array(GRID), temperature=np.array(T), steps=np.array(STEPS),
               linewidth_consistent=np.stack(mask))
    fn = os.path.join(ROOT, "tests", "golden", f"oracle_morse_{NAME}_{'x'.join(map(str, GRID))}.npz")
    np.savez_compressed(fn, **rec)
    print(fn, {k: np.shape(v) for k, v in rec.items()})


if __name__ == "__main__":
    main()
