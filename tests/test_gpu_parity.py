"""HIP path against the FP64 oracle and the committed oracle fixtures, through the C ABI.

Tolerances:
  FP64 mode:      frequencies/eigenvectors 1e-10, linewidths and Xi 1e-9 relative,
                  kappa 1e-8 relative (summation order is the only difference).
  FP32 / BF16x3:  kappa within 1e-4 relative of the FP64 GPU result, linewidths of the
                  strongly scattered modes within 1e-3 relative.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _engine(inputs_dir, name, grid):
    from tdep_amd.engine import Engine
    print(f"[case] {name} {grid}", flush=True)
    e = Engine.from_directory(inputs_dir(name), grid, device=0)
    e.dispersions()
    return e


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name,grid", [("fcc_al", (6, 6, 6)), ("rocksalt", (4, 4, 4)), ("sl16", (4, 4, 2))])
def test_harmonic_matches_oracle(inputs_dir, oracle_setup, name, grid):
    p, fc2, fc3, ops, mesh, disp = oracle_setup(name, grid)
    e = _engine(inputs_dir, name, grid)
    om, vel, egv = e.harmonic()
    assert np.abs(om - disp.om_irr).max() < 1e-10 * np.abs(disp.om_irr).max()
    assert np.abs(vel - disp.vel_irr).max() < 1e-9 * np.abs(disp.vel_irr).max() + 1e-12
    assert np.abs(egv - disp.egv_irr).max() < 1e-9
    omf, velf, egvf = e.harmonic(full=True)
    assert np.abs(egvf - disp.egv_full).max() < 1e-9
    assert np.abs(velf - disp.vel_full).max() < 1e-9 * np.abs(disp.vel_full).max() + 1e-12


@pytest.mark.parametrize("name,grid", [("fcc_al", (6, 6, 6)), ("rocksalt", (4, 4, 4)), ("sl16", (4, 4, 2))])
def test_fp64_scattering_and_kappa_match_golden(inputs_dir, name, grid):
    g = np.load(os.path.join(GOLDEN, f"oracle_{name}_{'x'.join(map(str, grid))}.npz"))
    e = _engine(inputs_dir, name, grid)
    T = g["temperatures"]
    e.scattering_rates(T, precision="fp64")
    ks, ki, ko = e.kappa()
    for iT in range(len(T)):
        lw, F0, _ = e.linewidths(iT)
        assert _rel(lw, g[f"linewidth_{iT}"]) < 1e-9
        assert _rel(ks[iT], g[f"kappa_sma_{iT}"]) < 1e-9
        assert _rel(ki[iT], g[f"kappa_iter_{iT}"]) < 1e-7
        assert _rel(ko[iT], g[f"kappa_offdiag_{iT}"]) < 1e-8
        assert e.niter[iT] == int(g[f"niter_{iT}"])


@pytest.mark.parametrize("name,grid", [("rocksalt", (4, 4, 4)), ("sl16", (4, 4, 2)), ("sl32", (4, 4, 1))])
def test_offdiag_device_velocities_match_host(inputs_dir, monkeypatch, name, grid):
    """The off-diagonal kappa formed on the device (coherence velocities in LDS up to 48 modes,
    in global scratch at 96, pair sums per point) equals the host restatement's to summation
    order."""
    e = _engine(inputs_dir, name, grid)
    e.scattering_rates([100.0, 300.0], precision="fp64")
    _, _, kd = e.kappa()
    monkeypatch.setenv("TC3_OD_HOST", "1")
    _, _, kh = e.kappa()
    assert np.abs(kh).max() > 0
    assert _rel(kd, kh) < 1e-11


@pytest.mark.parametrize("name,grid,T", [("rocksalt", (4, 4, 4), 300.0), ("sl16", (4, 4, 2), 300.0)])
def test_fp64_scattering_matrix_matches_oracle(inputs_dir, oracle_setup, name, grid, T):
    from oracle import scattering as S
    p, fc2, fc3, ops, mesh, disp = oracle_setup(name, grid)
    sr = S.scattering_rates(p, fc3, mesh, disp, ops, T)
    e = _engine(inputs_dir, name, grid)
    e.scattering_rates([T], precision="fp64")
    xi, rq = e.scattering_matrix(0)
    assert [tuple(x) for x in rq] == list(sr.rows)
    assert _rel(xi, sr.Xi) < 1e-9


@pytest.mark.parametrize("name,grid,prec", [("sl16", (4, 4, 2), "fp32"), ("sl16", (4, 4, 2), "bf16x3"),
                                            ("rocksalt", (6, 6, 6), "fp32"), ("rocksalt", (6, 6, 6), "bf16x3"),
                                            ("sl32", (4, 4, 1), "bf16x3")])
def test_mixed_precision_kappa_within_1e4(inputs_dir, name, grid, prec):
    """Mixed-precision contraction against the FP64 GPU path: 2-atom rocksalt (small-band
    path), 16- and 32-atom superlattices (MFMA kernels for 48 and 96 bands)."""
    e = _engine(inputs_dir, name, grid)
    T = [100.0, 300.0, 700.0]
    e.scattering_rates(T, precision="fp64")
    k64 = e.kappa()
    lw64 = [e.linewidths(i)[0] for i in range(3)]
    e.scattering_rates(T, precision=prec)
    kmx = e.kappa()
    for a, b in zip(kmx, k64):
        assert _rel(a, b) < 1e-4
    for i in range(3):
        lw = e.linewidths(i)[0]
        big = lw64[i] > 1e-3 * lw64[i].max()
        assert np.abs(lw[big] / lw64[i][big] - 1).max() < 1e-3


def test_full_size_superlattice_properties(inputs_dir):
    """16-atom superlattice at its full 10x10x4 mesh: linewidths are non-negative and
    equal on degenerate sets, the scattering matrix is non-negative, kappa is tetragonal
    (xx = yy, no off-diagonal part) and mixed precision stays within 1e-4 of FP64."""
    e = _engine(inputs_dir, "sl16", (10, 10, 4))
    om, vel, egv = e.harmonic()
    e.scattering_rates([300.0], precision="fp64")
    k64 = e.kappa(offdiag=False)
    lw, F0, _ = e.linewidths(0)
    assert lw.min() >= 0
    for q in range(e.nirr):
        for b in range(e.nb):
            same = np.abs(om[q] - om[q, b]) < 1e-9
            assert np.ptp(lw[q, same]) <= 1e-12 * max(lw[q, same].max(), 1e-300)
    xi, _ = e.scattering_matrix(0, 0, 64)
    assert xi.min() >= 0
    ki = k64[1][0]
    assert abs(ki[0, 0] - ki[1, 1]) < 1e-10 * ki[0, 0] and np.abs(ki - np.diag(np.diag(ki))).max() == 0
    e.scattering_rates([300.0], precision="bf16x3")
    kb = e.kappa(offdiag=False)
    assert _rel(kb[1][0], ki) < 1e-4


@pytest.mark.parametrize("name,grid", [("rocksalt", (4, 4, 4)), ("sl16", (4, 4, 2))])
def test_isotope_scattering_matches_oracle(inputs_dir, oracle_setup, name, grid):
    """Isotope scattering (scattering_isotope.f90) with artificial mass disorder: FP64
    linewidths and symmetrised scattering matrix against the oracle, and the mixed mode
    within 1e-4 in kappa."""
    from oracle import scattering as S
    p, fc2, fc3, ops, mesh, disp = oracle_setup(name, grid)
    g = np.linspace(1e-3, 3e-3, p.na)
    T = 300.0
    sr = S.scattering_rates(p, fc3, mesh, disp, ops, T, isotope_g=g)
    e = _engine(inputs_dir, name, grid)
    e.set_isotopes(g, enable=True)
    e.scattering_rates([T], precision="fp64")
    lw, F0, _ = e.linewidths(0)
    assert _rel(lw, sr.linewidth) < 1e-9
    xi, rq = e.scattering_matrix(0)
    assert _rel(xi, sr.Xi) < 1e-9
    k64 = e.kappa(offdiag=False)
    e.scattering_rates([T], precision="bf16x3" if e.nb in (48, 96) else "fp32")
    kmx = e.kappa(offdiag=False)
    assert _rel(kmx[1], k64[1]) < 1e-4
    # switching the term off restores the three-phonon-only result
    e.set_isotopes(g, enable=False)
    e.scattering_rates([T], precision="fp64")
    sr0 = S.scattering_rates(p, fc3, mesh, disp, ops, T)
    assert _rel(e.linewidths(0)[0], sr0.linewidth) < 1e-9


@pytest.mark.parametrize("name,grid", [("fcc_al", (6, 6, 6)), ("sl16", (4, 4, 2))])
def test_boundary_scattering_matches_oracle(inputs_dir, oracle_setup, name, grid):
    """Boundary scattering |v| / L (scattering.f90:245-250, --max_mfp): FP64 linewidths and
    iterative kappa against the oracle with L = 20 nm (comparable to the three-phonon term),
    and switching it off restores the three-phonon result."""
    from oracle import kappa as K
    from oracle import scattering as S
    from oracle import units as U
    p, fc2, fc3, ops, mesh, disp = oracle_setup(name, grid)
    T, length = 300.0, 20e-9
    sr = S.scattering_rates(p, fc3, mesh, disp, ops, T, mfp_max=length * U.M_TO_BOHR)
    e = _engine(inputs_dir, name, grid)
    e.set_boundary(length)
    e.scattering_rates([T], precision="fp64")
    lw, F0, _ = e.linewidths(0)
    assert _rel(lw, sr.linewidth) < 1e-9
    ks = e.kappa(offdiag=False)
    Fn, _, _ = K.iterative_solution(p, mesh, disp, ops, sr, T)
    kit = K.symmetrize_kappa(K.get_kappa(p, mesh, disp, ops, Fn, T), ops)
    assert _rel(ks[1][0], kit) < 1e-8
    e.set_boundary(None)
    e.scattering_rates([T], precision="fp64")
    sr0 = S.scattering_rates(p, fc3, mesh, disp, ops, T)
    assert _rel(e.linewidths(0)[0], sr0.linewidth) < 1e-9
    assert e.kappa(offdiag=False)[1][0][0, 0] > ks[1][0][0, 0]


@pytest.mark.parametrize("name,grid", [("fcc_al", (6, 6, 6)), ("rocksalt", (4, 4, 4)), ("sl16", (4, 4, 2))])
def test_integration_type6_matches_oracle(inputs_dir, oracle_setup, name, grid):
    """Integration type 6 (velocity-difference widths, scattering_threephonon.f90:251-261):
    FP64 linewidths and scattering matrix against the oracle on the small-cell (fcc, rocksalt)
    and MFMA (16-atom) paths, the mixed mode within 1e-4 in kappa, and the isotope term with
    its own type-6 width (scattering_isotope.f90:52-54)."""
    from oracle import scattering as S
    p, fc2, fc3, ops, mesh, disp = oracle_setup(name, grid)
    T = 300.0
    sr = S.scattering_rates(p, fc3, mesh, disp, ops, T, integrationtype=6)
    e = _engine(inputs_dir, name, grid)
    e.scattering_rates([T], precision="fp64", integrationtype=6)
    lw, F0, _ = e.linewidths(0)
    assert _rel(lw, sr.linewidth) < 1e-9
    xi, rq = e.scattering_matrix(0)
    assert _rel(xi, sr.Xi) < 1e-9
    k64 = e.kappa(offdiag=False)
    e.scattering_rates([T], precision="bf16x3" if e.nb in (48, 96) else "fp32", integrationtype=6)
    assert _rel(e.kappa(offdiag=False)[1], k64[1]) < 1e-4
    # the widths differ from the adaptive ones, and so do the linewidths
    sr2 = S.scattering_rates(p, fc3, mesh, disp, ops, T)
    assert _rel(sr.linewidth, sr2.linewidth) > 1e-3
    if name == "rocksalt":
        g = np.linspace(1e-3, 3e-3, p.na)
        sri = S.scattering_rates(p, fc3, mesh, disp, ops, T, integrationtype=6, isotope_g=g)
        e.set_isotopes(g, enable=True)
        e.scattering_rates([T], precision="fp64", integrationtype=6)
        assert _rel(e.linewidths(0)[0], sri.linewidth) < 1e-9
        assert _rel(e.scattering_matrix(0)[0], sri.Xi) < 1e-9
        e.set_isotopes(g, enable=False)


@pytest.mark.parametrize("name,grid,qg3", [("fcc_al", (6, 6, 6), (3, 3, 3)), ("rocksalt", (4, 4, 4), (2, 2, 2)),
                                           ("sl16", (4, 4, 2), (2, 2, 1))])
def test_montecarlo_grid_matches_oracle(inputs_dir, oracle_setup, name, grid, qg3):
    """Monte-Carlo q2 grids (-qg3ph, --seed 7): the host's 64-bit Mersenne twister and
    stratified draws give the oracle's per-row triplets and weights, so FP64 linewidths and
    scattering matrix agree to 1e-9 on the small-cell and MFMA paths; the iterative solution
    (whose scattering matrix a Monte-Carlo grid leaves without time-reversal symmetry) gives
    the oracle's kappa to 1e-7 in the oracle's iteration count; mixed kappa within 1e-4;
    switching the grid off restores the full-mesh rates."""
    from oracle import kappa as K
    from oracle import scattering as S
    p, fc2, fc3, ops, mesh, disp = oracle_setup(name, grid)
    T = 300.0
    sr = S.scattering_rates(p, fc3, mesh, disp, ops, T, qg3ph=qg3, seed=7)
    e = _engine(inputs_dir, name, grid)
    e.set_montecarlo(qg3, 7)
    e.scattering_rates([T], precision="fp64")
    lw, F0, _ = e.linewidths(0)
    assert _rel(lw, sr.linewidth) < 1e-9
    assert _rel(e.scattering_matrix(0)[0], sr.Xi) < 1e-9
    k64 = e.kappa(offdiag=False)
    Fn, scf, _ = K.iterative_solution(p, mesh, disp, ops, sr, T)
    kit = K.symmetrize_kappa(K.get_kappa(p, mesh, disp, ops, Fn, T), ops)
    assert _rel(k64[1][0], kit) < 1e-7
    assert e.niter[0] == len(scf)
    e.scattering_rates([T], precision="bf16x3" if e.nb in (48, 96) else "fp32")
    assert _rel(e.kappa(offdiag=False)[1], k64[1]) < 1e-4
    e.set_montecarlo(None, 0)
    e.scattering_rates([T], precision="fp64")
    sr0 = S.scattering_rates(p, fc3, mesh, disp, ops, T)
    assert _rel(e.linewidths(0)[0], sr0.linewidth) < 1e-9


def test_fused_contraction_matches_two_kernel_path(inputs_dir, monkeypatch):
    """The opt-in fused 48-band contraction (TC3_FUSED=1: one persistent kernel, no W
    intermediate in HBM) against the two-kernel split-BF16x3 path and FP64 on the 16-atom
    superlattice with 7 temperatures: linewidths within 1e-4 of the two-kernel path, kappa
    within 1e-4 of FP64 (both are split-BF16x3; the summation orders differ)."""
    e = _engine(inputs_dir, "sl16", (4, 4, 2))
    T = [100.0 * i for i in range(1, 8)]
    e.scattering_rates(T, precision="fp64")
    k64 = e.kappa(offdiag=False)
    e.scattering_rates(T, precision="bf16x3")
    lwc = np.array([e.linewidths(i)[0] for i in range(7)])
    monkeypatch.setenv("TC3_FUSED", "1")
    e.scattering_rates(T, precision="bf16x3")
    lwf = np.array([e.linewidths(i)[0] for i in range(7)])
    kf = e.kappa(offdiag=False)
    assert _rel(lwf, lwc) < 1e-4
    for a, b in zip(kf, k64):
        assert _rel(a, b) < 1e-4


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_rows_stored_per_orbit_match_full_rows(inputs_dir, monkeypatch, prec):
    """The scattering matrix stored per orbit of the little group of each row's q-point (one
    value per orbit and band, the BTE product against the orbit sums of F) against the full
    rows (TC3_XI_FULL=1) on the 16-atom superlattice: fewer elements stored, every row the same
    bits once expanded, the same iteration counts, and kappa within 1e-12 relative in FP64 (the
    products are summed per orbit first) and 1e-9 in split-BF16x3 (the orbit sums of the FP32
    folded-out F are rounded to FP32)."""
    e = _engine(inputs_dir, "sl16", (4, 4, 2))
    T = [100.0, 300.0, 700.0]
    out = {}
    for mode in ("full", "orbit", "full again"):
        if mode != "orbit":
            monkeypatch.setenv("TC3_XI_FULL", "1")
        else:
            monkeypatch.delenv("TC3_XI_FULL", raising=False)
        e.scattering_rates(T, precision=prec)
        n, es, per_orbit = e.xi_storage()
        xi = np.array([e.scattering_matrix(i)[0] for i in range(len(T))])
        k = e.kappa(offdiag=False)
        out[mode] = (n, es, per_orbit, xi, k, np.array(e.niter[:len(T)]))
    nf, esf, pof, xif, kf, itf = out["full"]
    no, eso, poo, xio, ko, ito = out["orbit"]
    assert not pof and poo and esf == eso == (8 if prec == "fp64" else 4)
    assert nf == e.nrows * e.nfull * e.nb and no < nf
    assert np.array_equal(xif, xio)
    assert np.array_equal(itf, ito)
    # the full rows after the per-orbit ones: zeroed again where no triplet writes
    assert np.array_equal(out["full again"][3], xif)
    tol = 1e-12 if prec == "fp64" else 1e-9
    for a, b in zip(ko, kf):
        assert _rel(a, b) < tol


def test_contraction_graph_replay_matches_direct_launches(inputs_dir, monkeypatch):
    """The scattering pass replays its contraction launches from a HIP graph after the first
    call with the same arguments; the replays (three in a row, with other work between them)
    give the same linewidths and scattering-matrix rows, bit for bit, as direct launches
    (TC3_NO_GRAPH=1), in the split-BF16x3 and FP64 modes, and the graph is rebuilt when the
    temperatures change."""
    e = _engine(inputs_dir, "sl16", (4, 4, 2))
    for prec in ("bf16x3", "fp64"):
        for T in ([100.0, 300.0], [200.0, 400.0, 600.0]):
            monkeypatch.setenv("TC3_NO_GRAPH", "1")
            e.scattering_rates(T, precision=prec)
            lw0 = np.array([e.linewidths(i)[0] for i in range(len(T))])
            xi0 = np.array([e.scattering_matrix(i, 0, 32)[0] for i in range(len(T))])
            monkeypatch.delenv("TC3_NO_GRAPH")
            for _ in range(3):
                e.scattering_rates(T, precision=prec)
                e.kappa(offdiag=False)
                lw = np.array([e.linewidths(i)[0] for i in range(len(T))])
                xi = np.array([e.scattering_matrix(i, 0, 32)[0] for i in range(len(T))])
                assert np.array_equal(lw, lw0) and np.array_equal(xi, xi0)


def test_whole_triplet_psi_workgroup_matches(inputs_dir, monkeypatch):
    """The last contraction kernel with one workgroup per triplet (TC3_PSI_BB=3: the three
    blocks of 16 b1 in one workgroup, the shape the autotuner also times) gives the same
    linewidths and scattering-matrix rows, bit for bit, as one workgroup per block of 16 b1,
    on the 16-atom superlattice in split-BF16x3 with 7 temperatures."""
    e = _engine(inputs_dir, "sl16", (4, 4, 2))
    T = [100.0 * i for i in range(1, 8)]
    out = []
    for bb in ("1", "3"):
        monkeypatch.setenv("TC3_PSI_BB", bb)
        e.scattering_rates(T, precision="bf16x3")
        out.append((np.array([e.linewidths(i)[0] for i in range(7)]),
                    np.array([e.scattering_matrix(i, 0, 48)[0] for i in range(7)])))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
