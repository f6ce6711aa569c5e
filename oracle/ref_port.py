"""Build and drive oracle/scatter_ref.c (the per-element restatement of the reference's
row loop).  TEST INFRASTRUCTURE ONLY: used by tests and by bench.py's cpu_baseline leg."""
import ctypes
import os
import subprocess
import time

import numpy as np

from . import scattering as S
from . import units as U

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "scatter_ref.c")
LIB = os.path.join(HERE, "libscatter_ref.so")


def build():
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-shared", "-fPIC", SRC, "-o", LIB + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def _load():
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        build()
    lib = ctypes.CDLL(LIB)
    lib.ref_rows.restype = ctypes.c_long
    lib.ref_threads.restype = ctypes.c_int
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def run_rows(pre, temperature, rows, max_trip=None):
    """Linewidth of the given (q_irr, band) rows over their first max_trip visited triplets.
    Returns (gamma, seconds, threads, n_elements) with n_elements = sum rows*triplets*nb^2."""
    p, fc2, fc3, ops, mesh, disp = pre
    lib = _load()
    nb = disp.nb
    be, sigsq, _ = S.occupations_and_smearing(p, mesh, disp, temperature)
    rot = S.rotation_table(p, mesh, ops)
    cache = {}
    off, tq2, tq3, perm, neq = [0], [], [], [], []
    for (q1, b1) in rows:
        if q1 not in cache:
            cache[q1] = S.irreducible_triplets(mesh, rot, q1)
        tl = cache[q1][:max_trip] if max_trip else cache[q1]
        for (q2, q3, pm, red) in tl:
            tq2.append(q2); tq3.append(q3); perm.append(pm); neq.append(len(red))
        off.append(len(tq2))
    i32 = lambda x: np.ascontiguousarray(x, dtype=np.int32)
    f64 = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    c128 = lambda x: np.ascontiguousarray(x, dtype=np.complex128)
    gam = np.zeros(len(rows))
    args = [i32(fc3["a1"]), i32(fc3["a2"]), i32(fc3["a3"]), f64(fc3["lv2"]), f64(fc3["lv3"]),
            f64(fc3["mwm"]), f64(mesh.qcart), f64(disp.om_full), c128(disp.egv_full),
            i32(mesh.irr_of_full), f64(be), f64(sigsq)]
    rq = i32([r[0] for r in rows]); rb = i32([r[1] for r in rows])
    rom = f64([disp.om_irr[q, b] for q, b in rows])
    regv = c128(np.array([disp.egv_irr[q][:, b] for q, b in rows]))
    trip = [i32(off), i32(tq2), i32(tq3), f64(perm), i32(neq)]
    t0 = time.perf_counter()
    lib.ref_rows(ctypes.c_int(nb), ctypes.c_int(mesh.nfull), ctypes.c_int(len(fc3["a1"])),
                 *[_p(a) for a in args], ctypes.c_int(len(rows)), _p(rq), _p(rb), _p(rom), _p(regv),
                 *[_p(a) for a in trip], ctypes.c_double(1.0 / mesh.nfull), _p(gam))
    dt = time.perf_counter() - t0
    return gam, dt, lib.ref_threads(), int((len(tq2)) * nb * nb)
