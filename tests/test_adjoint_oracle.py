"""Pins of the oracle's adjoint (oracle/adjoint.py) on the CPU.

The linewidths are quadratic in Phi3, so (i) Euler's theorem for homogeneous functions
fixes <mwm, dL/dmwm> = 2 L exactly, and (ii) a central difference is exact up to rounding
for any step; the RTA kappa goes as 1 / linewidth, so its central difference carries an
O(h^2) truncation error, which Richardson extrapolation removes.  A wrong phase sign, a
transposed eigenvector or a dropped equivalent-pair weight breaks (i) and (ii).
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def rocksalt333(oracle_setup):
    from oracle import scattering as S
    pre = oracle_setup("rocksalt", (3, 3, 3))
    p, fc2, fc3, ops, mesh, disp = pre
    sr = S.scattering_rates(p, fc3, mesh, disp, ops, 300.0)
    return pre, sr


def _with_mwm(fc3, x):
    f = dict(fc3)
    f["mwm"] = x
    return f


def test_gamma_adjoint_euler_identity_and_central_difference(rocksalt333):
    from oracle import adjoint as A, scattering as S
    (p, fc2, fc3, ops, mesh, disp), sr = rocksalt333
    rng = np.random.default_rng(7)
    dl = rng.standard_normal((mesh.nirr, disp.nb))
    g = A.raw_cotangent(disp, dl)
    grad = A.gamma_adjoint(p, fc3, mesh, disp, ops, 300.0, g)
    raw = np.zeros((mesh.nirr, disp.nb))
    for i, (q, b) in enumerate(sr.rows):
        raw[q, b] = sr.gamma_row[i]
    L0 = np.sum(dl * sr.linewidth)
    assert np.isclose(np.sum(g * raw), L0, rtol=1e-13)
    assert np.isclose(np.sum(grad * fc3["mwm"]), 2.0 * L0, rtol=1e-12)
    # a random direction and two single entries, at two step sizes each (exact for a
    # quadratic; the single entries take large steps to stay clear of rounding)
    dirs = [(rng.standard_normal(fc3["mwm"].shape) * fc3["mwm"], (1e-4, 1e-2))]
    flat = np.abs(fc3["mwm"]).reshape(-1)
    nz = np.flatnonzero(flat > 1e-3 * flat.max())
    for e in (int(np.argmax(flat)), int(nz[len(nz) // 3])):
        d = np.zeros(flat.size)
        d[e] = fc3["mwm"].reshape(-1)[e]
        dirs.append((d.reshape(fc3["mwm"].shape), (0.1, 1.0)))
    for d, steps in dirs:
        an = np.sum(grad * d)
        for h in steps:
            lp = np.sum(dl * S.scattering_rates(p, _with_mwm(fc3, fc3["mwm"] + h * d), mesh, disp, ops, 300.0).linewidth)
            lm = np.sum(dl * S.scattering_rates(p, _with_mwm(fc3, fc3["mwm"] - h * d), mesh, disp, ops, 300.0).linewidth)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - an) <= 1e-9 * abs(an) + 1e-12 * np.sum(np.abs(grad * d)), (h, fd, an)


def test_raw_cotangent_is_adjoint_of_degeneracy_average(rocksalt333):
    """<dl, A x> = <A^T dl, x> for the band-by-band in-place averaging of the forward."""
    from oracle import adjoint as A
    (p, fc2, fc3, ops, mesh, disp), sr = rocksalt333
    rng = np.random.default_rng(3)
    x = rng.standard_normal((mesh.nirr, disp.nb))
    dl = rng.standard_normal((mesh.nirr, disp.nb))
    ax = x.copy()
    for q in range(mesh.nirr):
        for b in range(disp.nb):
            if disp.om_irr[q, b] < 1e-9:
                ax[q, b] = 0.0
                continue
            dm = disp.degen_irr[q][b]
            ax[q, dm] = ax[q, dm].sum() / len(dm)
    ax = np.where(disp.om_irr >= 1e-9, ax, 0.0)
    assert np.isclose(np.sum(dl * ax), np.sum(A.raw_cotangent(disp, dl) * x), rtol=1e-13)


def test_kappa_rta_adjoint_richardson_difference(rocksalt333):
    from oracle import adjoint as A, kappa as K, scattering as S
    (p, fc2, fc3, ops, mesh, disp), sr = rocksalt333
    G = np.zeros((3, 3))
    G[0, 0] = 1.0
    G[1, 2] = 0.3
    cot = A.kappa_rta_cotangent(p, mesh, disp, ops, sr, 300.0, G)
    grad = A.gamma_adjoint(p, fc3, mesh, disp, ops, 300.0, A.raw_cotangent(disp, cot))
    rng = np.random.default_rng(11)
    d = rng.standard_normal(fc3["mwm"].shape) * fc3["mwm"]

    def L(x):
        s = S.scattering_rates(p, _with_mwm(fc3, x), mesh, disp, ops, 300.0)
        k = K.symmetrize_kappa(K.get_kappa(p, mesh, disp, ops, s.F0, 300.0), ops)
        return np.sum(G * k)
    h = 1e-4
    f1 = (L(fc3["mwm"] + h * d) - L(fc3["mwm"] - h * d)) / (2 * h)
    f2 = (L(fc3["mwm"] + 2 * h * d) - L(fc3["mwm"] - 2 * h * d)) / (4 * h)
    an = np.sum(grad * d)
    assert abs(f1 - an) < 1e-5 * abs(an)                 # central difference, step 1e-4 relative
    assert abs((4 * f1 - f2) / 3 - an) < 1e-8 * abs(an)  # Richardson: O(h^4)
