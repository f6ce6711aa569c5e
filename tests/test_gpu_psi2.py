"""Three-phonon matrix elements |Psi|^2 of the contraction kernels (tc3_get_psi2) against the
oracle's, and the error of the mixed-precision modes against FP64.

The oracle forms Psi[b1, b2, b3] = sum E1 E2 E3 Phi(q2, q3) from its own dispersions and its
restatement of pretransform_phi3 (scattering_threephonon.f90:96-112, 275-317), pinned there
against the reference's zgeru/dot_product form element by element.  Only elements whose three
frequencies are above lo_freqtol enter (the reference never forms the others).

Tolerances: FP64 1e-8 of the triplet's largest element (the device eigenvectors agree with the
oracle's to 1e-9; measured 1e-13 .. 7e-12).  Against the FP64 kernels, normwise per triplet
and as the median relative error of the elements above 1e-3 of the triplet's largest: FP32
2e-6 (measured 3.7e-7 and 2.4e-7), split-BF16x3 3e-5 (measured 8.9e-6 and 8.1e-6: the split
products drop the lo x lo term, 2^-16 relative).
"""
import numpy as np
import pytest

from oracle import scattering as S
from oracle import units as U

pytestmark = pytest.mark.gpu


def _engine(inputs_dir, name, grid):
    from tdep_amd.engine import Engine
    print(f"[case] psi2 {name} {grid}", flush=True)
    e = Engine.from_directory(inputs_dir(name), grid, device=0)
    e.dispersions()
    return e


def _oracle_psi2(fc3, mesh, disp, q1, q2, q3):
    nb = disp.nb

    def E(egv, om):
        return egv / np.sqrt(np.where(om > 0, om, 1.0))[None, :]
    ptf = S.pretransform_phi3(fc3, mesh.qcart[q2], mesh.qcart[q3], nb)
    psi = S.psi_chain(E(disp.egv_irr[q1], disp.om_irr[q1]), E(disp.egv_full[q2], disp.om_full[q2]),
                      E(disp.egv_full[q3], disp.om_full[q3]), ptf)
    return S.THREEPHONON_PREFACTOR * np.abs(psi) ** 2 / mesh.nfull


def _mask(disp, q1, q2, q3):
    w1, w2, w3 = disp.om_irr[q1], disp.om_full[q2], disp.om_full[q3]
    return ((w1 >= U.FREQTOL)[:, None, None] & (w2 >= U.FREQTOL)[None, :, None]
            & (w3 >= U.FREQTOL)[None, None, :])


@pytest.mark.parametrize("name,grid,nsample", [("sl16", (4, 4, 2), 6), ("sl32", (4, 4, 1), 3)])
def test_fp64_psi2_matches_oracle(inputs_dir, oracle_setup, name, grid, nsample):
    p, fc2, fc3, ops, mesh, disp = oracle_setup(name, grid)
    e = _engine(inputs_dir, name, grid)
    trips = np.linspace(0, e.ntriplets - 1, nsample).astype(np.int64)
    got, q = e.psi2(trips, "fp64")
    for i in range(nsample):
        want = _oracle_psi2(fc3, mesh, disp, *q[i])
        m = _mask(disp, *q[i])
        scale = np.abs(want[m]).max()
        err = np.abs(got[i][m] - want[m]).max() / scale
        print(f"[psi2] fp64 {name} triplet {trips[i]} q {tuple(q[i])}: {err:.2e} of the largest", flush=True)
        assert err <= 1e-8, (i, q[i])


def psi2_errors(mixed, ref, masks):
    """Per-triplet normwise relative error, and the relative error of every element above
    1e-3 of its triplet's largest."""
    norm, elem = [], []
    for a, b, m in zip(mixed, ref, masks):
        a, b = a[m], b[m]
        norm.append(np.linalg.norm(a - b) / np.linalg.norm(b))
        big = b > 1e-3 * b.max()
        elem.append(np.abs(a[big] / b[big] - 1))
    return np.array(norm), np.concatenate(elem)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_mixed_precision_psi2_error_16_atom_10x10x4(inputs_dir, oracle_setup, precision):
    """The 16-atom superlattice at its full 10x10x4 mesh: 16 triplets spread over the set."""
    e = _engine(inputs_dir, "sl16", (10, 10, 4))
    trips = np.linspace(0, e.ntriplets - 1, 16).astype(np.int64)
    ref, q = e.psi2(trips, "fp64")
    got, q2 = e.psi2(trips, precision)
    assert np.array_equal(q, q2)
    om, _, _ = e.harmonic()
    omf, _, _ = e.harmonic(full=True)
    masks = [((om[a] >= U.FREQTOL)[:, None, None] & (omf[b] >= U.FREQTOL)[None, :, None]
              & (omf[c] >= U.FREQTOL)[None, None, :]) for a, b, c in q]
    norm, elem = psi2_errors(got, ref, masks)
    print(f"[psi2] {precision}: norm max {norm.max():.2e} median {np.median(norm):.2e}; "
          f"element max {elem.max():.2e} median {np.median(elem):.2e}", flush=True)
    bound = {"fp32": 2e-6, "bf16x3": 3e-5}[precision]
    assert norm.max() < bound
    assert np.median(elem) < bound
