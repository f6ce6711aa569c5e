"""tc3_kappa_noalloc (working buffers in a fixed device array, no device allocation) against
tc3_kappa through the C ABI: the same single-mode and iterative kappa (FP64 arithmetic, same
kernels in the same order: bit for bit), iteration counts and final F, for two calls in a row
on one GPU, the second with a smaller workload, and for a single-precision scattering matrix."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(e, nT):
    from tdep_amd.lib import dptr, iptr
    out = []
    for fn in ("tc3_kappa", "tc3_kappa_noalloc"):
        ks, ki, nit = np.zeros(9 * nT), np.zeros(9 * nT), np.zeros(nT, np.int32)
        if fn == "tc3_kappa":
            r = e.lib.tc3_kappa(e.ctx, 200, 1e-5, 0, dptr(ks), dptr(ki), None, iptr(nit))
        else:
            r = e.lib.tc3_kappa_noalloc(e.ctx, 200, 1e-5, 0, dptr(ks), dptr(ki), iptr(nit))
        assert r == 0, e.lib.tc3_last_error(e.ctx).decode()
        Fn = [e.modes(i)[0] for i in range(nT)]
        out.append((ks, ki, nit, Fn))
    return out


@pytest.mark.parametrize("name,grid,temps,prec", [("sl16", (4, 4, 2), [100.0, 300.0, 700.0], "fp64"),
                                                   ("rocksalt", (4, 4, 4), [300.0], "fp64"),
                                                   ("sl16", (4, 4, 2), [300.0], "bf16x3")])
def test_noalloc_kappa_equals_kappa(inputs_dir, name, grid, temps, prec):
    from tdep_amd.engine import Engine
    print(f"[case] noalloc {name} {grid} {temps} {prec}", flush=True)
    e = Engine.from_directory(inputs_dir(name), grid, device=0)
    e.dispersions()
    e.scattering_rates(temps, precision=prec)
    (ks, ki, nit, Fn), (ks2, ki2, nit2, Fn2) = _both(e, len(temps))
    assert np.array_equal(nit, nit2) and nit.min() > 0
    assert np.array_equal(ks, ks2)
    assert np.abs(ki2 - ki).max() <= 1e-13 * np.abs(ki).max()
    for a, b in zip(Fn, Fn2):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
    # a second call on the same context reuses the array: same results again
    ks3, ki3, nit3 = np.zeros(9 * len(temps)), np.zeros(9 * len(temps)), np.zeros(len(temps), np.int32)
    from tdep_amd.lib import dptr, iptr
    assert e.lib.tc3_kappa_noalloc(e.ctx, 200, 1e-5, 0, dptr(ks3), dptr(ki3), iptr(nit3)) == 0
    assert np.array_equal(ki3, ki2) and np.array_equal(nit3, nit2)
