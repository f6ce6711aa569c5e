"""outfile.cumulative_thermal_conductivity.hdf5 (tc3_write_cumulative_hdf5) against the oracle
(oracle/cumulative.py), through the C ABI.

The oracle runs on the engine's own harmonic modes, F vectors, linewidths and off-diagonal
kappa, each of which the other GPU tests pin against the oracle (1e-9 or better), so the
comparison isolates the DOS, tetrahedra, cumulative, boundary and spectral code.  Both sides
are FP64 with the same operations; datasets agree within 1e-9 of their largest value.  At
the default sizes (1000-point axes) the file is checked through properties that do not
depend on the axes: the DOS integrates to 3N, the cumulative and spectral kappa end at /
integrate to the total, per-mode and per-atom parts add up to the total, and kappa with
boundary scattering rises with the boundary length towards kappa_iter + kappa_offdiag.
The angular-momentum tensor vanishes by inversion symmetry in the test crystals: it is
compared with the oracle's against its no-cancellation size, and its spectrum (a ratio of
rounding noise here) only for shape and finiteness.
The file is read back with h5py from the image's conda interpreter.
"""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONDA_PY = "/opt/conda/bin/python3.9"

_DUMP = r"""
import json, sys, h5py, numpy as np
out = {"attrs": {}, "data": {}}
with h5py.File(sys.argv[1], "r") as f:
    for k, v in f.attrs.items():
        out["attrs"][k] = v.decode() if isinstance(v, bytes) else np.asarray(v).tolist()
    for k in f:
        d = f[k]
        out["data"][k] = {"values": d[()].tolist(), "unit": d.attrs["unit"].decode().strip(),
                          "dimensions": d.attrs["dimensions"].decode().strip()}
json.dump(out, sys.stdout)
"""


def _read(path):
    if not os.path.exists(CONDA_PY):
        pytest.skip("no interpreter with h5py in this image")
    r = subprocess.run([CONDA_PY, "-s", "-c", _DUMP, path], capture_output=True, text=True, timeout=120)
    if r.returncode != 0 and "No module named 'h5py'" in r.stderr:
        pytest.skip("h5py not importable")
    assert r.returncode == 0, r.stderr
    h = json.loads(r.stdout)
    return h["attrs"], {k: np.array(v["values"]) for k, v in h["data"].items()}, h["data"]


class _Disp:
    pass


def _engine_state(inputs_dir, name, grid, T):
    from tdep_amd.engine import Engine
    print(f"[case] cumulative {name} {grid}", flush=True)
    e = Engine.from_directory(inputs_dir(name), grid, device=0)
    e.dispersions()
    e.scattering_rates([T], precision="fp64")
    ks, ki, ko = e.kappa()
    d = _Disp()
    d.om_irr, d.vel_irr, d.egv_irr = e.harmonic()
    d.om_full, d.vel_full, d.egv_full = e.harmonic(full=True)
    d.nb = e.nb
    Fn = e.modes(0)[0]
    lw = e.linewidths(0)[0]
    return e, d, Fn, lw, ko[0], ki[0]   # kappa in atomic units


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _voigt33(v6):
    """[p][6] (xx yy zz yz xz xy) -> [3][3][p]"""
    idx = [[0, 5, 4], [5, 1, 3], [4, 3, 2]]
    return np.array([[v6[:, idx[i][j]] for i in range(3)] for j in range(3)])


@pytest.mark.parametrize("name,grid,dosint", [("fcc_al", (6, 6, 6), 3), ("fcc_al", (6, 6, 6), 1),
                                              ("rocksalt", (4, 4, 4), 3), ("rocksalt", (4, 4, 4), 2),
                                              ("sl16", (4, 4, 2), 3)])
def test_cumulative_file_matches_oracle(inputs_dir, oracle_setup, tmp_path, name, grid, dosint):
    from oracle import cumulative as C, units as U
    T, npts, nfreq = 300.0, 40, 120
    p, fc2, fc3, ops, mesh, _ = oracle_setup(name, grid)
    e, d, Fn, lw, kod, _ = _engine_state(inputs_dir, name, grid, T)
    path = str(tmp_path / "c.hdf5")
    e.write_cumulative_hdf5(path, 0, mfppts=npts, freqpts=nfreq, dosintegrationtype=dosint)
    attrs, h, raw = _read(path)
    tets = C.irreducible_tetrahedra(p, ops, mesh) if dosint == 3 else None
    dos = C.phonon_dos(p, ops, mesh, d, nfreq, dosint, 1.0, tets=tets)
    mfp = C.scalar_mfp(d, lw)
    cu = C.cumulative_kappa(p, ops, mesh, d, Fn, mfp, T, npts=npts)
    bax, bk = C.boundary_kappa(p, ops, mesh, d, Fn, lw, mfp, kod, T, npts=npts)
    sp, sb, sa, _ = C.spectral_kappa(p, ops, mesh, d, Fn, dos, T, tets=tets)
    uf, k = U.FREQ_HARTREE_TO_THZ, U.KAPPA_AU_TO_SI
    assert attrs["temperature"].strip() == "temperature"
    assert attrs["atomic_species"].split() == list(dict.fromkeys(p.symbols))
    assert raw["spectral_kappa_vs_frequency"]["dimensions"] == "xyz,xyz,frequency"
    assert raw["cumulative_kappa_vs_mean_free_path_per_mode"]["unit"] == "W/m/K"
    for key in ("dos", "cumulative_kappa_vs_mean_free_path", "spectral_kappa_vs_frequency",
                "boundary_scattering_kappa", "spectral_kappa_vs_frequency_per_atom"):
        assert np.abs(h[key]).max() > 0, key
    assert _rel(h["frequencies"], dos.omega * uf) < 1e-12
    assert _rel(h["dos"], dos.dos / uf) < 1e-9
    assert _rel(h["mean_free_path_axis"], cu.mfaxis * U.BOHR_TO_M) < 1e-9
    assert _rel(h["cumulative_kappa_vs_mean_free_path"], _voigt33(cu.mf_kappa) * k) < 1e-9
    band = np.array([[_voigt33(cu.mf_kappa_band[:, m, :])[j, i] for m in range(d.nb)] for j in range(3) for i in range(3)])
    assert _rel(h["cumulative_kappa_vs_mean_free_path_per_mode"].reshape(9, d.nb, npts), band * k) < 1e-9
    atom = np.array([[_voigt33(cu.mf_kappa_atom[:, a, :])[j, i] for a in range(p.na)] for j in range(3) for i in range(3)])
    assert _rel(h["cumulative_kappa_vs_mean_free_path_per_atom"].reshape(9, p.na, npts), atom * k) < 1e-9
    assert _rel(h["boundary_scattering_lengths"], bax * U.BOHR_TO_M) < 1e-9
    assert _rel(h["boundary_scattering_kappa"], _voigt33(bk.T) * k) < 1e-9
    spec = np.array([[sp[:, 3 * i + j] for i in range(3)] for j in range(3)])
    assert _rel(h["spectral_kappa_vs_frequency"], spec * k / uf) < 1e-9
    specb = np.array([[sb[:, :, 3 * i + j].T for i in range(3)] for j in range(3)])
    assert _rel(h["spectral_kappa_vs_frequency_per_mode"], specb * k / uf) < 1e-9
    speca = np.array([[sa[:, :, 3 * i + j].T for i in range(3)] for j in range(3)])
    assert _rel(h["spectral_kappa_vs_frequency_per_atom"], speca * k / uf) < 1e-9
    # angular momentum: zero by inversion symmetry in these crystals, so the tensor is
    # rounding noise; it matches the oracle's against the size it would have if every mode
    # were circularly polarised, sum |v| c_v / omega / (2 Gamma) / (N V)
    alpha = C.angular_momentum_matrix(p, ops, mesh, d, lw, T)
    scale = 0.0
    for q in range(mesh.nfull):
        for b in range(d.nb):
            w = d.om_full[q, b]
            if w >= U.FREQTOL:
                scale += np.linalg.norm(d.vel_full[q, b]) * float(U.harmonic_oscillator_cv(T, w)) / w / (
                    2 * lw[mesh.irr_of_full[q], b]) / mesh.nfull / p.volume
    assert np.abs(h["generating_angular_momentum_tensor"] - alpha.T).max() < 1e-12 * scale
    assert np.abs(alpha).max() < 1e-9 * scale
    am = h["generating_angular_momentum_tensor_vs_frequency"]
    assert am.shape == (3, 3, nfreq) and np.all(np.isfinite(am))
    assert raw["generating_angular_momentum_tensor"]["unit"] == "hbar/m/K"


def test_cumulative_file_properties_at_default_sizes(inputs_dir, tmp_path):
    """16-atom superlattice 4x4x2, 1000-point axes, tetrahedron spectra."""
    from oracle import cumulative as C
    T = 300.0
    e, d, Fn, lw, kod, ki = _engine_state(inputs_dir, "sl16", (4, 4, 2), T)
    path = str(tmp_path / "c.hdf5")
    e.write_cumulative_hdf5(path, 0)
    _, h, _ = _read(path)
    f = h["frequencies"]
    assert len(f) == 1000 and len(h["mean_free_path_axis"]) == 1000
    assert abs(C.trapezoid(f, h["dos"]) - d.nb) < 1e-9 * d.nb
    cum = h["cumulative_kappa_vs_mean_free_path"]                       # [3][3][p]
    spec = h["spectral_kappa_vs_frequency"]
    from tdep_amd.engine import KAPPA_AU_TO_SI
    # two q-planes along z, both mirror planes: every v_z vanishes and kappa_zz = 0
    assert ki[2, 2] == 0 and np.all(cum[2, 2] == 0) and np.all(spec[2, 2] == 0)
    for i in range(2):
        assert np.all(np.diff(cum[i, i]) >= 0)                          # monotone in the mfp
        total = cum[i, i, -1]
        assert abs(total - ki[i, i] * KAPPA_AU_TO_SI) < 1e-9 * total    # the cumulative ends at kappa_iter
        assert abs(C.trapezoid(f, spec[i, i]) - total) < 1e-9 * total   # same total, other route
        # the parts are rescaled to the total where it exceeds 1e-5 W/mK and zero below
        assert np.allclose(h["cumulative_kappa_vs_mean_free_path_per_mode"][i, i].sum(axis=0), cum[i, i],
                           rtol=1e-9, atol=1e-5)
        assert np.allclose(h["cumulative_kappa_vs_mean_free_path_per_atom"][i, i].sum(axis=0), cum[i, i],
                           rtol=1e-9, atol=1e-5)
        assert np.allclose(h["spectral_kappa_vs_frequency_per_mode"][i, i].sum(axis=0), spec[i, i],
                           rtol=1e-9, atol=1e-12 * spec[i, i].max())
        # kappa with boundary scattering rises with the length; the shortest length (a tenth
        # of the shortest mean free path) leaves under a tenth of the diagonal part, the
        # longest leaves kappa_iter + kappa_offdiag within 1e-5
        bk = h["boundary_scattering_kappa"][i, i]
        od = kod[i, i] * KAPPA_AU_TO_SI
        assert np.all(np.diff(bk) >= -1e-12 * bk[-1])
        assert bk[0] - od < 0.1 * (bk[-1] - od)
        assert abs(bk[-1] - (ki[i, i] * KAPPA_AU_TO_SI + od)) < 1e-5 * bk[-1]
