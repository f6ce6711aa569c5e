"""Four-phonon scattering (--fourthorder) on the GPU against the oracle (oracle/fourphonon.py),
through the C ABI: linewidths and the symmetrised scattering matrix, FP64, within 1e-9 of
their largest value (1e-8 with velocity-difference widths); the four-phonon term changes the
three-phonon linewidths by more than that.  Cells of one and two atoms go through the
one-workgroup-per-quartet kernel, four and sixteen atoms (12 and 48 bands) through the batched
phi4 + MFMA contraction + epilogue kernels."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name,grid,it,temps", [("fcc_al", (4, 4, 4), 2, (100.0, 300.0)),
                                                ("fcc_al", (4, 4, 4), 1, (100.0, 300.0)),
                                                ("fcc_al", (4, 4, 4), 6, (100.0, 300.0)),
                                                ("rocksalt", (3, 3, 3), 2, (100.0, 300.0)),
                                                # four atoms (12 bands): the batched MFMA contraction
                                                ("sl4", (3, 3, 3), 2, (100.0, 300.0, 500.0, 700.0, 900.0)),
                                                ("sl4", (3, 3, 3), 6, (300.0,)),
                                                ("sl4", (3, 3, 3), 1, (300.0,))])
def test_fourphonon_matches_oracle(tmp_path, name, grid, it, temps):
    from oracle import lattice as L, run as R, scattering as S
    from tdep_amd import inputs
    from tdep_amd.engine import Engine
    print(f"[case] fourphonon {name} {grid} it={it} T={temps}", flush=True)
    d = str(tmp_path / name)
    inputs.make_config(name, d, fourth=True)
    temps = list(temps)
    e = Engine.from_directory(d, grid, device=0)
    e.dispersions()
    e.scattering_rates(temps, precision="fp64", integrationtype=it)
    lw3 = [e.linewidths(i)[0].copy() for i in range(len(temps))]
    e.set_fourthorder(d)
    e.scattering_rates(temps, precision="fp64", integrationtype=it)
    p, fc2, fc3, ops, mesh, disp = R.setup(d, grid)
    fc4 = L.read_fc4(p, os.path.join(d, "infile.forceconstant_fourthorder"))
    # integration type 6 takes its widths from group-velocity differences, which the engine
    # and the oracle agree on to 1e-9; a Gaussian out to its 8-sigma cut scales a relative
    # width change by up to 63, hence 1e-8 there
    tol = 1e-8 if it == 6 else 1e-9
    for iT, T in enumerate(temps):
        sr = S.scattering_rates(p, fc3, mesh, disp, ops, T, integrationtype=it, fc4=fc4)
        lw = e.linewidths(iT)[0]
        assert _rel(lw, sr.linewidth) < tol
        assert _rel(lw, lw3[iT]) > 1e-6
        xi, rq = e.scattering_matrix(iT)
        assert [tuple(x) for x in rq] == list(sr.rows)
        assert _rel(xi, sr.Xi) < tol


def test_hosts_with_fourthorder(tmp_path):
    """Both hosts run with --fourthorder on the same infiles write the same
    outfile.thermal_conductivity; the four-phonon term lowers kappa."""
    import subprocess
    from tdep_amd import host, inputs
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tdep_amd", "tc_host")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (no Fortran compiler)")
    print("[case] hosts --fourthorder fcc_al 4x4x4", flush=True)
    dirs = {k: tmp_path / k for k in ("f", "p", "p3")}
    for d in dirs.values():
        inputs.make_config("fcc_al", str(d), fourth=True)
    args = ["-qg", "4", "4", "4", "--temperature", "300", "--readiso"]
    r = subprocess.run([exe] + args + ["--fourthorder"], cwd=dirs["f"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert host.main(args + ["--fourthorder", "--directory", str(dirs["p"])]) == 0
    assert host.main(args + ["--directory", str(dirs["p3"])]) == 0
    a, b, c = (np.loadtxt(dirs[k] / "outfile.thermal_conductivity") for k in ("f", "p", "p3"))
    assert np.allclose(a, b, rtol=1e-10, atol=1e-14 * np.abs(b).max())
    # total kappa (last row): xx
    assert b[-1][0] < c[-1][0] * (1 - 1e-6)


@pytest.mark.parametrize("name,grid,qg3,qg4", [("fcc_al", (4, 4, 4), None, (2, 2, 2)),
                                               ("rocksalt", (3, 3, 3), (2, 2, 2), (2, 2, 2)),
                                               ("sl4", (3, 3, 3), (2, 2, 2), (2, 2, 2))])
def test_fourphonon_montecarlo_matches_oracle(tmp_path, name, grid, qg3, qg4):
    """-qg4ph (and -qg3ph) with --seed 7: every row draws its three-phonon grid and two
    four-phonon grids in turn; linewidths and scattering matrix as the oracle's."""
    from oracle import lattice as L, run as R, scattering as S
    from tdep_amd import inputs
    from tdep_amd.engine import Engine
    print(f"[case] fourphonon Monte-Carlo {name} {grid} qg3ph={qg3} qg4ph={qg4}", flush=True)
    d = str(tmp_path / name)
    inputs.make_config(name, d, fourth=True)
    e = Engine.from_directory(d, grid, device=0)
    e.set_montecarlo(qg3, 7)
    e.set_montecarlo4(qg4)
    e.set_fourthorder(d)
    e.dispersions()
    e.scattering_rates([300.0], precision="fp64")
    p, fc2, fc3, ops, mesh, disp = R.setup(d, grid)
    fc4 = L.read_fc4(p, os.path.join(d, "infile.forceconstant_fourthorder"))
    sr = S.scattering_rates(p, fc3, mesh, disp, ops, 300.0, fc4=fc4, qg3ph=qg3, qg4ph=qg4, seed=7)
    assert _rel(e.linewidths(0)[0], sr.linewidth) < 1e-9
    xi, rq = e.scattering_matrix(0)
    assert [tuple(x) for x in rq] == list(sr.rows)
    assert _rel(xi, sr.Xi) < 1e-9


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_fourphonon_48_bands_matches_golden(tmp_path, prec):
    """16-atom superlattice (48 bands) on a 3x3x1 mesh with --fourthorder: the three-phonon
    term on the MFMA chain and the four-phonon term through the batched kernels; linewidths
    and twelve scattering-matrix rows against the oracle's (tests/golden/make_golden.py,
    FOUR_CASES; the oracle takes about a minute here).  FP64 within 1e-9 of the largest value;
    with the split-BF16x3 three-phonon contraction (the four-phonon contraction then in FP32,
    the scattering matrix stored in FP32) within the mixed mode's 1e-4 budget."""
    from tdep_amd import inputs
    from tdep_amd.engine import Engine
    root = os.path.dirname(os.path.abspath(__file__))
    g = np.load(os.path.join(root, "golden", "oracle_fourth_sl16_3x3x1.npz"))
    grid, temps = tuple(int(x) for x in g["grid"]), [float(x) for x in g["temperatures"]]
    print(f"[case] fourphonon 48 bands sl16 {grid} {prec}", flush=True)
    d = str(tmp_path / "sl16")
    inputs.make_config("sl16", d, fourth=True)
    e = Engine.from_directory(d, grid, device=0)
    e.set_fourthorder(d)
    e.dispersions()
    e.scattering_rates(temps, precision=prec)
    tol = 1e-9 if prec == "fp64" else 1e-4
    for iT in range(len(temps)):
        assert _rel(e.linewidths(iT)[0], g[f"linewidth_{iT}"]) < tol
        xi, rq = e.scattering_matrix(iT)
        assert [tuple(x) for x in rq] == [tuple(x) for x in g[f"rows_{iT}"]]
        assert _rel(xi[g[f"xi_rows_{iT}"]], g[f"xi_{iT}"]) < tol


@pytest.mark.parametrize("name,grid", [("fcc_al", (4, 4, 4)), ("rocksalt", (3, 3, 3))])
def test_fourphonon_batched_kernels_on_small_cells(tmp_path, monkeypatch, name, grid):
    """The batched four-phonon kernels forced on cells of one and two atoms (TC3_F4_BATCHED),
    three quartets per batch (TC3_F4_NQB): linewidths and scattering matrix as the oracle's,
    like the one-workgroup-per-quartet kernel's."""
    from oracle import lattice as L, run as R, scattering as S
    from tdep_amd import inputs
    from tdep_amd.engine import Engine
    print(f"[case] fourphonon batched kernels {name} {grid}", flush=True)
    monkeypatch.setenv("TC3_F4_BATCHED", "1")
    monkeypatch.setenv("TC3_F4_NQB", "3")
    d = str(tmp_path / name)
    inputs.make_config(name, d, fourth=True)
    e = Engine.from_directory(d, grid, device=0)
    e.set_fourthorder(d)
    e.dispersions()
    e.scattering_rates([300.0], precision="fp64")
    p, fc2, fc3, ops, mesh, disp = R.setup(d, grid)
    fc4 = L.read_fc4(p, os.path.join(d, "infile.forceconstant_fourthorder"))
    sr = S.scattering_rates(p, fc3, mesh, disp, ops, 300.0, fc4=fc4)
    assert _rel(e.linewidths(0)[0], sr.linewidth) < 1e-9
    xi, rq = e.scattering_matrix(0)
    assert [tuple(x) for x in rq] == list(sr.rows)
    assert _rel(xi, sr.Xi) < 1e-9


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_fourphonon_mixed_precision_matches_oracle(tmp_path, prec):
    """The mixed modes on the 4-atom cell (12 bands): three-phonon contraction in FP32 or
    split-BF16x3, the batched four-phonon contraction in FP32 (FP64 epilogue); linewidths and
    scattering matrix within 1e-5 of the FP64 oracle's largest value (both contractions carry
    about 1e-6 relative error at this size)."""
    from oracle import lattice as L, run as R, scattering as S
    from tdep_amd import inputs
    from tdep_amd.engine import Engine
    print(f"[case] fourphonon mixed {prec} sl4 (3, 3, 3)", flush=True)
    d = str(tmp_path / "sl4")
    inputs.make_config("sl4", d, fourth=True)
    grid = (3, 3, 3)
    e = Engine.from_directory(d, grid, device=0)
    e.set_fourthorder(d)
    e.dispersions()
    e.scattering_rates([300.0], precision=prec)
    p, fc2, fc3, ops, mesh, disp = R.setup(d, grid)
    fc4 = L.read_fc4(p, os.path.join(d, "infile.forceconstant_fourthorder"))
    sr = S.scattering_rates(p, fc3, mesh, disp, ops, 300.0, fc4=fc4)
    assert _rel(e.linewidths(0)[0], sr.linewidth) < 1e-5
    xi, rq = e.scattering_matrix(0)
    assert _rel(xi, sr.Xi) < 1e-5
