"""libtc3.so on the CPU: exported ABI, input readers and the host-side preprocessing
(space group, q-mesh, irreducible triplets) against the oracle.  No device calls."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import lattice as L
from oracle import scattering as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _lib():
    from tdep_amd.lib import load
    return load()


def test_library_exports_every_declared_symbol():
    lib = _lib()
    hdr = open(os.path.join(ROOT, "include", "tc3.h")).read()
    names = re.findall(r"^\s*(?:int|void|const char \*)\s*(tc3_\w+)\(", hdr, re.M)
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    from tdep_amd.lib import SIGNATURES
    assert set(names) == set(SIGNATURES)


def test_device_calls_fail_cleanly_on_host_only_context(inputs_dir):
    from tdep_amd.engine import Engine
    from tdep_amd.lib import Tc3Error
    e = Engine.from_directory(inputs_dir("fcc_al"), (4, 4, 4), device=-1)
    with pytest.raises(Tc3Error) as ei:
        e.dispersions()
    assert ei.value.code == 7


@pytest.mark.parametrize("name,grid", [("fcc_al", (12, 12, 12)), ("rocksalt", (8, 8, 8)),
                                       ("sl16", (10, 10, 4))])
def test_host_mesh_and_triplets_equal_oracle(inputs_dir, name, grid):
    from tdep_amd.engine import Engine
    d = inputs_dir(name)
    e = Engine.from_directory(d, grid, device=-1)
    p = L.read_ucposcar(os.path.join(d, "infile.ucposcar"))
    L.read_isotopes(p, os.path.join(d, "infile.isotopes"))
    ops = L.spacegroup(p)
    m = L.generate_mesh(p, ops, grid)
    mm = e.mesh()
    sy = e.symmetry()
    assert e.nsym == len(ops) and e.nirr == m.nirr and e.nfull == m.nfull
    assert np.array_equal(mm["irr_full_index"], m.irr_full_index)
    assert np.array_equal(mm["irr_of_full"], m.irr_of_full)
    assert np.array_equal(mm["op_of_full"], m.op_of_full)
    assert np.allclose(mm["weight_irr"], m.weight_irr, rtol=0, atol=1e-15)
    assert np.abs(mm["qcart"] - m.qcart).max() < 1e-14
    for o, r, t, f in zip(ops, sy["rot"], sy["tr"], sy["fmap"]):
        assert np.array_equal(o.m, r) and np.allclose(o.tr, t, atol=1e-14) and np.array_equal(o.fmap, f)
    rot = S.rotation_table(p, m, ops)
    nt = sum(len(S.irreducible_triplets(m, rot, q)) for q in range(m.nirr))
    assert e.ntriplets == nt


def test_readers_on_reference_example_fixture():
    """examples/example_1_fcc_al infiles of the reference (a real TDEP fit: 43 pairs,
    85 triplets including three-body clusters)."""
    from tdep_amd import io as tio
    d = os.path.join(GOLDEN, "fcc_al_example")
    p = tio.read_ucposcar(os.path.join(d, "infile.ucposcar"))
    p.mass = np.array([26.98 * tio.AMU_TO_EMU])
    fc2 = tio.read_fc2(p, os.path.join(d, "infile.forceconstant"))
    fc3 = tio.read_fc3(p, os.path.join(d, "infile.forceconstant_thirdorder"))
    assert len(fc2["a1"]) == 43 and len(fc3["a1"]) == 85
    # acoustic sum rules of a fitted model hold to the fit's precision
    assert np.abs(fc2["m"].sum(axis=0)).max() < 1e-6 * np.abs(fc2["m"]).max()
    # the oracle's reader agrees with the product reader
    po = L.read_ucposcar(os.path.join(d, "infile.ucposcar"))
    po.invsqrtmass = np.array([1 / np.sqrt(p.mass[0])])
    f3o = L.read_fc3(po, os.path.join(d, "infile.forceconstant_thirdorder"))
    assert np.allclose(f3o["lv2"], fc3["lv2"]) and np.allclose(f3o["m"], fc3["m"], rtol=1e-15)
    f2o = L.read_fc2(po, os.path.join(d, "infile.forceconstant"))
    assert np.allclose(f2o["m"], fc2["m"], rtol=1e-15) and np.allclose(f2o["lv2"], fc2["lv2"])


def test_outfile_format_matches_reference_fixture():
    """The reference's own outfile.thermal_conductivity (tests/thermal_conductivity/reference)
    fixes the layout: header lines, four titled blocks of six E24.12 numbers."""
    from tdep_amd.engine import _fe
    ref = open(os.path.join(GOLDEN, "outfile.thermal_conductivity.reference")).read().splitlines()
    assert ref[0].startswith("# Unit:") and ref[1].startswith("# Temperature:")
    vals = [ln for ln in ref if not ln.startswith("#")]
    assert len(vals) == 4
    for ln in vals:
        nums = [float(x) for x in ln.split()]
        assert len(nums) == 6
        assert ln == " " + "".join(" " + _fe(x) for x in nums)
    t = float(ref[1].split()[-1])
    assert ref[1] == "# Temperature:" + " " * 8 + _fe(t, 20, 12)


@pytest.mark.parametrize("name,grid", [("rocksalt", (8, 8, 8)), ("sl16", (10, 10, 4))])
def test_fortran_host_reads_inputs_like_the_python_readers(inputs_dir, name, grid):
    """The Fortran host's readers (tdep_amd/fortran/tc_host.f90) give the same mesh and
    triplet counts as the Python readers through the same library (host-only context), and
    the same input checksums."""
    import subprocess
    from tdep_amd.engine import Engine
    from tdep_amd import io as tio
    exe = os.path.join(ROOT, "tdep_amd", "tc_host")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (no Fortran compiler)")
    d = inputs_dir(name)
    r = subprocess.run([exe, "-qg"] + [str(g) for g in grid] + ["--device", "-1", "--inputs-only", "--readiso"],
                       cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    e = Engine.from_directory(d, grid, device=-1)
    assert [int(x) for x in out["mesh"]] == [e.nsym, e.nfull, e.nirr, e.ntriplets, e.nrows]
    p, fc2, fc3 = tio.read_directory(d)
    f = lambda k: np.array([float(x) for x in out[k][1:]])
    assert int(out["structure"][0]) == p.na
    assert np.allclose(f("structure"), [np.abs(p.latvec).sum(), p.rfrac.sum(), p.mass.sum()], rtol=1e-13)
    assert int(out["fc2"][0]) == len(fc2["a1"]) and int(out["fc3"][0]) == len(fc3["a1"])
    assert np.allclose(f("fc2"), [np.abs(fc2["m"]).sum(), fc2["m"][:, 0, 1].sum(), np.abs(fc2["lv2"]).sum()],
                       rtol=1e-12, atol=1e-16)
    assert np.allclose(f("fc3"), [np.abs(fc3["m"]).sum(), fc3["m"][:, 0, 1, 2].sum(), np.abs(fc3["lv2"]).sum(),
                                  np.abs(fc3["lv3"]).sum()], rtol=1e-12, atol=1e-16)


# standard atomic weights (IUPAC) and Tamura's mass-variance parameters, independent of the table
_ATOMIC_WEIGHT = {"H": 1.00794, "C": 12.0107, "N": 14.0067, "O": 15.9994, "Na": 22.98976928, "Mg": 24.3050,
                  "Al": 26.9815386, "Si": 28.0855, "P": 30.973762, "S": 32.065, "Cl": 35.453, "K": 39.0983,
                  "Ca": 40.078, "Ti": 47.867, "Fe": 55.845, "Cu": 63.546, "Ga": 69.723, "Ge": 72.64,
                  "As": 74.92160, "Ag": 107.8682, "Sn": 118.710, "Au": 196.966569, "Pb": 207.2, "Bi": 208.98040}


def test_natural_isotope_table_against_standard_atomic_weights():
    """tc3_natural_isotopes (the reference's naturaldistribution table): abundances sum to one,
    the mean masses are the standard atomic weights, and the mass-disorder parameters of Si and
    Ge are Tamura's (Phys. Rev. B 27, 858 (1983)): 2.01e-4 and 5.87e-4; single-isotope elements
    have none.  Unknown symbols fail with the reference's parameter error."""
    from tdep_amd import io as tio
    for sym, w in _ATOMIC_WEIGHT.items():
        c, m = tio.natural_isotopes(sym)
        assert abs(c.sum() - 1.0) < 2e-4, sym
        assert np.all(m > 0) and np.all(np.diff(m) > 0), sym
        mean = (c / c.sum() * m).sum()
        assert abs(mean - w) < 2e-4 * w, (sym, mean, w)
    emu = tio.AMU_TO_EMU
    for sym, g in (("Si", 2.01e-4), ("Ge", 5.87e-4)):
        c, m = tio.natural_isotopes(sym)
        mean, dis = tio._distribution(c, m)
        assert abs(mean / emu - _ATOMIC_WEIGHT[sym]) < 1e-3
        assert abs(dis - g) < 0.02 * g, (sym, dis)
    c, m = tio.natural_isotopes("Al")
    assert len(c) == 1 and tio._distribution(c, m)[1] == 0.0
    with pytest.raises(ValueError):
        tio.natural_isotopes("Xx")
    lib = _lib()
    n = ctypes.c_int(-1)
    assert lib.tc3_natural_isotopes(b"Sn", 0, None, None, ctypes.byref(n)) == 0 and n.value == 10
    buf = np.zeros(4)
    from tdep_amd.lib import dptr
    assert lib.tc3_natural_isotopes(b"Sn", 4, dptr(buf), dptr(buf), ctypes.byref(n)) == 1


@pytest.mark.parametrize("name,grid", [("rocksalt", (4, 4, 4)), ("sl16", (4, 4, 2))])
def test_natural_masses_without_readiso_in_both_hosts(inputs_dir, name, grid):
    """Without --readiso both hosts take the masses and mass-disorder parameters from the natural
    isotope distributions (type_crystalstructure.f90:650-652): the Fortran host's checksums equal
    the Python readers', and differ from the infile.isotopes ones."""
    import subprocess
    from tdep_amd import io as tio
    exe = os.path.join(ROOT, "tdep_amd", "tc_host")
    assert os.path.exists(exe), "Fortran host not built"
    d = inputs_dir(name)
    r = subprocess.run([exe, "-qg"] + [str(g) for g in grid] + ["--device", "-1", "--inputs-only"], cwd=d,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    p, _, _ = tio.read_directory(d, readiso=False)
    q, _, _ = tio.read_directory(d, readiso=True)
    assert float(out["structure"][3]) == pytest.approx(p.mass.sum(), rel=1e-14)
    assert float(out["disorder"][0]) == pytest.approx(p.mass_disorder.sum(), rel=1e-13, abs=1e-300)
    assert abs(p.mass.sum() - q.mass.sum()) > 1e-6 * q.mass.sum() or name == "rocksalt"
    for a in range(p.na):
        c, m = tio.natural_isotopes(p.symbols[a])
        assert p.mass[a] == pytest.approx((c / c.sum() * m).sum() * tio.AMU_TO_EMU, rel=1e-14)


def test_nosym_mesh_is_reduced_by_time_reversal_only(inputs_dir):
    """--nosym (type_qpointmesh_gridgeneration.f90:54-58): the mesh keeps only the identity and
    time reversal, so the irreducible points are the orbits {q, -q}; the crystal keeps its space
    group (48 operations for rocksalt) for the kappa averages."""
    from tdep_amd.engine import Engine
    d = inputs_dir("rocksalt")
    e = Engine.from_directory(d, (4, 4, 4), device=-1, qpsymmetry=False)
    assert e.nsym == 48 and e.nfull == 64 and e.nirr == (64 + 8) // 2
    m = e.mesh()
    assert np.all(np.abs(m["op_of_full"]) == 1)
    e2 = Engine.from_directory(d, (4, 4, 4), device=-1, qpsymmetry=False, timereversal=False)
    assert e2.nirr == 64
