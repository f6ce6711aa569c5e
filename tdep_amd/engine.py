"""Python mirror of the reference driver for the three-phonon path.

`Engine` owns one libtc3 context and exposes the reference's steps under their
own names:

  Engine.from_directory(...)      uc%readfromfile / fc%readfromfile / fct%readfromfile
                                  (main.f90:94-133) and lo_generate_qmesh (main.f90:139)
  engine.dispersions()            dr%generate (main.f90:146)
  engine.scattering_rates(...)    sr%generate (main.f90:191) for several temperatures
  engine.kappa(...)               get_kappa / get_kappa_offdiag / iterative_solution /
                                  symmetrize_kappa (main.f90:219-244)
  engine.outfile_text(iT)         outfile.thermal_conductivity (main.f90:248-294)

Errors from the library raise Tc3Error with the reference's exit code.
"""
import ctypes
import os

import numpy as np

from . import io as tio
from .lib import Tc3Error, dptr, iptr, load

PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}
KAPPA_AU_TO_SI = (27.21138602 * 1.6021766208e-19) / (
    (6.582119514e-16 * 1.6021766208e-19) / (27.21138602 * 1.6021766208e-19)) / 0.52917721067e-10
M_TO_BOHR = 1.0 / 0.52917721067e-10   # lo_m_to_Bohr (konstanter.f90:45-46)


def _fe(x, w=24, d=12):
    """Fortran Ew.d edit descriptor (0.ddddE+xx)."""
    if x == 0.0 or not np.isfinite(x):
        s = "0." + "0" * d + "E+00"
    else:
        e = int(np.floor(np.log10(abs(x)))) + 1
        ms = f"{abs(x) / 10.0 ** e:.{d}f}"
        if ms.startswith("1."):
            e += 1
            ms = f"{abs(x) / 10.0 ** e:.{d}f}"
        s = ("-" if x < 0 else "") + ms + ("E+" if e >= 0 else "E-") + f"{abs(e):02d}"
    return s.rjust(w)


class Engine:
    def __init__(self, device=0):
        self.lib = load()
        self.ctx = ctypes.c_void_p()
        r = self.lib.tc3_create(ctypes.byref(self.ctx), int(device))
        if r:
            raise Tc3Error(r, "could not create a context on device %d" % device)

    def __del__(self):
        try:
            if self.ctx:
                self.lib.tc3_destroy(self.ctx)
                self.ctx = ctypes.c_void_p()
        except Exception:
            pass

    def _check(self, r):
        if r:
            raise Tc3Error(r, self.lib.tc3_last_error(self.ctx).decode())

    # ---------------------------------------------------------------- inputs
    @classmethod
    def from_directory(cls, indir, qgrid, device=0, timereversal=True, readiso=True, qpsymmetry=True):
        """readiso: masses from infile.isotopes (--readiso), else from the natural isotope
        distributions (the reference's default); qpsymmetry=False is the reference's --nosym."""
        eng = cls(device)
        p, fc2, fc3 = tio.read_directory(indir, readiso=readiso)
        eng.set_structure(p)
        eng.set_fc2(fc2)
        eng.set_fc3(fc3)
        eng.set_qmesh(qgrid, timereversal, qpsymmetry)
        return eng

    def set_thirdorder(self, enable=True):
        """Three-phonon term on or off (enable=False is the reference's --nothirdorder)."""
        self._check(self.lib.tc3_set_thirdorder(self.ctx, int(bool(enable))))

    def set_montecarlo(self, qg3ph, seed):
        """Monte-Carlo q2 grid of the three-phonon term (-qg3ph, --seed): a stratified draw of
        prod(qg3ph) points per row; None (or qg3ph = qgrid) for the full mesh."""
        if qg3ph is None:
            self._check(self.lib.tc3_set_montecarlo(self.ctx, None, int(seed or 0)))
        else:
            q = np.ascontiguousarray(qg3ph, dtype=np.int32)
            self._check(self.lib.tc3_set_montecarlo(self.ctx, iptr(q), int(seed)))

    def set_boundary(self, max_mfp):
        """Boundary scattering |v| / L with L = max_mfp in metres (the reference's --max_mfp,
        scattering.f90:245-250); None or a value <= 0 switches it off."""
        L = -1.0 if max_mfp is None or max_mfp <= 0 else float(max_mfp) * M_TO_BOHR
        self._check(self.lib.tc3_set_boundary(self.ctx, L))

    def set_isotopes(self, disorder, enable=True):
        """Isotope scattering with per-atom mass-disorder parameters (scattering_isotope.f90);
        enable=False is the reference's --noisotope."""
        g = np.ascontiguousarray(disorder, dtype=np.float64)
        self._check(self.lib.tc3_set_isotopes(self.ctx, dptr(g), int(bool(enable))))

    def set_structure(self, p):
        self.structure = p
        lat = np.ascontiguousarray(p.latvec.T.reshape(-1))  # columns -> consecutive triples
        rf = np.ascontiguousarray(p.rfrac.reshape(-1))
        sp = np.ascontiguousarray(p.species, dtype=np.int32)
        ms = np.ascontiguousarray(p.mass, dtype=np.float64)
        self._check(self.lib.tc3_set_structure(self.ctx, p.na, dptr(lat), dptr(rf), iptr(sp), dptr(ms)))
        self.na = p.na
        self.nb = 3 * p.na

    def set_fc2(self, fc2):
        m = np.ascontiguousarray(np.transpose(fc2["m"], (0, 2, 1)).reshape(-1))  # m[9n+3j+i]
        self._check(self.lib.tc3_set_fc2(self.ctx, len(fc2["a1"]), iptr(fc2["a1"]), iptr(fc2["a2"]),
                                         dptr(np.ascontiguousarray(fc2["lv2"].reshape(-1))), dptr(m)))

    def set_fc3(self, fc3):
        self.fc3 = fc3
        self.nfc3 = len(fc3["a1"])
        m = np.ascontiguousarray(np.transpose(fc3["m"], (0, 3, 2, 1)).reshape(-1))  # m[27t+9k+3j+i]
        self._check(self.lib.tc3_set_fc3(self.ctx, len(fc3["a1"]), iptr(fc3["a1"]), iptr(fc3["a2"]),
                                         iptr(fc3["a3"]), dptr(np.ascontiguousarray(fc3["lv2"].reshape(-1))),
                                         dptr(np.ascontiguousarray(fc3["lv3"].reshape(-1))), dptr(m)))

    def set_fc4(self, fc4):
        """Fourth-order IFCs (io.read_fc4): four-phonon scattering in the following
        scattering_rates calls (--fourthorder, full q-grid for the quartets); None switches
        it off."""
        if fc4 is None:
            z = np.zeros(1, np.int32)
            self._check(self.lib.tc3_set_fc4(self.ctx, 0, iptr(z), iptr(z), iptr(z), iptr(z), None, None, None, None))
            return
        m = np.ascontiguousarray(np.transpose(fc4["m"], (0, 4, 3, 2, 1)).reshape(-1))   # m[81t+27l+9k+3j+i]
        flat = lambda k: dptr(np.ascontiguousarray(fc4[k].reshape(-1)))
        self._check(self.lib.tc3_set_fc4(self.ctx, len(fc4["a1"]), iptr(fc4["a1"]), iptr(fc4["a2"]), iptr(fc4["a3"]),
                                         iptr(fc4["a4"]), flat("lv2"), flat("lv3"), flat("lv4"), dptr(m)))

    def set_montecarlo4(self, qg4ph):
        """Monte-Carlo grids of the four-phonon term (-qg4ph; the seed is set_montecarlo's):
        every row draws its three-phonon grid and two four-phonon grids in turn; None for the
        full mesh."""
        if qg4ph is None:
            self._check(self.lib.tc3_set_montecarlo4(self.ctx, None))
        else:
            q = np.ascontiguousarray(qg4ph, dtype=np.int32)
            self._check(self.lib.tc3_set_montecarlo4(self.ctx, iptr(q)))

    def set_fourthorder(self, indir):
        """Read indir/infile.forceconstant_fourthorder and switch four-phonon scattering on."""
        from tdep_amd import io as tio
        self.set_fc4(tio.read_fc4(self.structure, os.path.join(indir, "infile.forceconstant_fourthorder")))

    def set_qmesh(self, qgrid, timereversal=True, qpsymmetry=True):
        d = np.ascontiguousarray(qgrid, dtype=np.int32)
        self._check(self.lib.tc3_set_qmesh_sym(self.ctx, iptr(d), int(bool(timereversal)), int(bool(qpsymmetry))))
        self.qgrid = tuple(int(x) for x in qgrid)
        self.mesh_info()

    def mesh_info(self):
        nsym, nfull, nirr, nrows = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ntrip = ctypes.c_int64()
        self._check(self.lib.tc3_mesh_info(self.ctx, ctypes.byref(nsym), ctypes.byref(nfull), ctypes.byref(nirr),
                                           ctypes.byref(ntrip), ctypes.byref(nrows)))
        self.nsym, self.nfull, self.nirr, self.ntriplets = nsym.value, nfull.value, nirr.value, ntrip.value
        self.nrows = nrows.value
        return dict(nsym=self.nsym, nfull=self.nfull, nirr=self.nirr, ntriplets=self.ntriplets,
                    nrows=self.nrows)

    def mesh(self):
        irr = np.zeros(self.nirr, np.int32)
        iof = np.zeros(self.nfull, np.int32)
        oof = np.zeros(self.nfull, np.int32)
        w = np.zeros(self.nirr)
        qc = np.zeros(3 * self.nfull)
        self._check(self.lib.tc3_get_mesh(self.ctx, iptr(irr), iptr(iof), iptr(oof), dptr(w), dptr(qc)))
        return dict(irr_full_index=irr, irr_of_full=iof, op_of_full=oof, weight_irr=w,
                    qcart=qc.reshape(self.nfull, 3))

    def symmetry(self):
        rot = np.zeros(9 * self.nsym)
        tr = np.zeros(3 * self.nsym)
        fm = np.zeros(self.na * self.nsym, np.int32)
        self._check(self.lib.tc3_get_symmetry(self.ctx, dptr(rot), dptr(tr), iptr(fm)))
        return dict(rot=rot.reshape(self.nsym, 3, 3).transpose(0, 2, 1), tr=tr.reshape(self.nsym, 3),
                    fmap=fm.reshape(self.nsym, self.na))

    # ------------------------------------------------------------- harmonic
    def dispersions(self):
        om = ctypes.c_double()
        self._check(self.lib.tc3_harmonic(self.ctx, ctypes.byref(om)))
        self.omega_min = om.value
        self.mesh_info()
        return self.omega_min

    def harmonic(self, full=False):
        nq = self.nfull if full else self.nirr
        nb = self.nb
        om = np.zeros(nq * nb)
        vel = np.zeros(nq * nb * 3)
        eg = np.zeros(nq * nb * nb * 2)
        self._check(self.lib.tc3_get_harmonic(self.ctx, int(full), dptr(om), dptr(vel), dptr(eg)))
        eg = eg.reshape(nq, nb, nb, 2)
        egv = (eg[..., 0] + 1j * eg[..., 1]).transpose(0, 2, 1)  # [q, component, mode]
        return om.reshape(nq, nb), vel.reshape(nq, nb, 3), egv

    # ----------------------------------------------------------- scattering
    def scattering_rates(self, temperatures, precision="fp64", integrationtype=2, sigma=1.0,
                         classical=False):
        """Three-phonon (plus isotope and boundary) scattering for all temperatures at once.
        integrationtype: 1 fixed Gaussian of width sigma (THz), 2 adaptive, 6 adaptive from
        group-velocity differences (scattering_threephonon.f90:242-262)."""
        T = np.ascontiguousarray(np.atleast_1d(temperatures), dtype=np.float64)
        self.temperatures = T
        self.classical = bool(classical)
        self._check(self.lib.tc3_scatter(self.ctx, len(T), dptr(T), PREC[precision], int(integrationtype),
                                         float(sigma), int(bool(classical))))

    def linewidths(self, iT=0):
        qb = self.nirr * self.nb
        lw = np.zeros(qb)
        F0 = np.zeros(qb * 3)
        sm = np.zeros(qb)
        self._check(self.lib.tc3_get_scattering(self.ctx, iT, dptr(lw), dptr(F0), dptr(sm)))
        return lw.reshape(self.nirr, self.nb), F0.reshape(self.nirr, self.nb, 3), sm.reshape(self.nirr, self.nb)

    def scattering_matrix(self, iT=0, row0=0, nrow=None):
        nrow = self.nrows - row0 if nrow is None else nrow
        xi = np.zeros(nrow * self.nfull * self.nb)
        rq = np.zeros(2 * self.nrows, np.int32)
        self._check(self.lib.tc3_get_xi(self.ctx, iT, row0, nrow, dptr(xi), iptr(rq)))
        return xi.reshape(nrow, self.nfull * self.nb), rq.reshape(self.nrows, 2)

    # ---------------------------------------------------------------- kappa
    def kappa(self, maxiter=200, tol=1e-5, offdiag=True):
        nT = len(self.temperatures)
        ks, ki, ko = np.zeros(9 * nT), np.zeros(9 * nT), np.zeros(9 * nT)
        nit = np.zeros(nT, np.int32)
        self._check(self.lib.tc3_kappa(self.ctx, int(maxiter), float(tol), int(self.classical), dptr(ks), dptr(ki),
                                       dptr(ko) if offdiag else None, iptr(nit)))
        f = lambda a: a.reshape(nT, 3, 3).transpose(0, 2, 1).copy()
        self.kappa_sma, self.kappa_iter, self.kappa_offdiag = f(ks), f(ki), f(ko)
        self.niter = nit
        return self.kappa_sma, self.kappa_iter, self.kappa_offdiag

    def kappa_log(self, iT=0):
        buf = np.zeros(7 * 1024)
        n = ctypes.c_int()
        self._check(self.lib.tc3_get_kappa_log(self.ctx, iT, 1024, dptr(buf), ctypes.byref(n)))
        return buf[:7 * min(n.value, 1024)].reshape(-1, 7)

    def modes(self, iT=0):
        qb = self.nirr * self.nb
        Fn = np.zeros(qb * 3)
        km = np.zeros(qb * 9)
        self._check(self.lib.tc3_get_modes(self.ctx, iT, dptr(Fn), dptr(km)))
        return Fn.reshape(self.nirr, self.nb, 3), km.reshape(self.nirr, self.nb, 3, 3).transpose(0, 1, 3, 2)

    # -------------------------------------------------------------- adjoint
    def adjoint_gamma(self, dL_dlinewidth, precision="fp64"):
        """dL/dm [ntrip, 3, 3, 3] (m as in fc3["m"], Ha/Bohr^3) for a cotangent of the
        linewidths [nT, nirr, nb] of the last scattering_rates call; |Psi|^2 recomputed."""
        nT = len(self.temperatures)
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(dL_dlinewidth, np.float64),
                                                 (nT, self.nirr, self.nb)).reshape(-1))
        out = np.zeros(27 * self.nfc3)
        self._check(self.lib.tc3_adjoint_gamma(self.ctx, dptr(g), PREC[precision], dptr(out)))
        return out.reshape(self.nfc3, 3, 3, 3).transpose(0, 3, 2, 1).copy()

    def adjoint_kappa_rta(self, dL_dkappa, precision="fp64"):
        """dL/dm [ntrip, 3, 3, 3] for a cotangent of kappa_sma [nT, 3, 3] (as kappa() returns
        it, atomic units); needs scattering_rates, not kappa."""
        nT = len(self.temperatures)
        d = np.broadcast_to(np.asarray(dL_dkappa, np.float64), (nT, 3, 3))
        d = np.ascontiguousarray(d.transpose(0, 2, 1).reshape(-1))   # column-major per T
        out = np.zeros(27 * self.nfc3)
        self._check(self.lib.tc3_adjoint_kappa_rta(self.ctx, dptr(d), PREC[precision], dptr(out)))
        return out.reshape(self.nfc3, 3, 3, 3).transpose(0, 3, 2, 1).copy()

    def jvp(self, dfc2=None, dfc3=None):
        """Changes along dfc2 [npair, 3, 3] and dfc3 [ntrip, 3, 3, 3] (the layouts of set_fc2 and
        set_fc3's m, Hartree/Bohr^n; None for no change) of the linewidths [nT, nirr, nb], the
        frequencies [nirr, nb], the group velocities [nirr, nb, 3] and kappa_sma [nT, 3, 3] (as
        kappa() returns it), by forward-mode differentiation on the device (tc3_jvp); needs
        scattering_rates."""
        nT, qb = len(self.temperatures), self.nirr * self.nb
        d2 = None if dfc2 is None else np.ascontiguousarray(np.transpose(dfc2, (0, 2, 1)).reshape(-1))
        d3 = None if dfc3 is None else np.ascontiguousarray(np.transpose(dfc3, (0, 3, 2, 1)).reshape(-1))
        dlw, dom, dv, dk = np.zeros(nT * qb), np.zeros(qb), np.zeros(3 * qb), np.zeros(9 * nT)
        self._check(self.lib.tc3_jvp(self.ctx, None if d2 is None else dptr(d2), None if d3 is None else dptr(d3),
                                     dptr(dlw), dptr(dom), dptr(dv), dptr(dk)))
        return dict(linewidth=dlw.reshape(nT, self.nirr, self.nb), omega=dom.reshape(self.nirr, self.nb),
                    velocity=dv.reshape(self.nirr, self.nb, 3), kappa_sma=dk.reshape(nT, 3, 3).transpose(0, 2, 1).copy())

    def psi2(self, triplets, precision="fp64"):
        """threephonon_prefactor |Psi|^2 / N_q [n, nb, nb, nb] (b1, b2, b3) of the irreducible
        triplets with the given indices (0 .. ntriplets - 1 of the full grid), formed by the
        contraction kernels in the given precision (tc3_get_psi2; 48 and 96 bands), and their
        (q1 irreducible, q2, q3) indices [n, 3]."""
        t = np.ascontiguousarray(np.asarray(triplets, dtype=np.int64).reshape(-1))
        n, nb = len(t), self.nb
        out = np.zeros(n * nb ** 3)
        q = np.zeros(3 * n, dtype=np.int32)
        self._check(self.lib.tc3_get_psi2(self.ctx, PREC[precision], n,
                                          t.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), dptr(out),
                                          q.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out.reshape(n, nb, nb, nb), q.reshape(n, 3)

    def timings(self):
        t = np.zeros(8)
        self._check(self.lib.tc3_get_timings(self.ctx, dptr(t), 8))
        return dict(harmonic=t[0], scatter=t[1], kappa=t[2], bte=t[3], offdiag=t[4], scatter_kernels=t[5],
                    adjoint=t[6], adjoint_kernels=t[7])

    def tuning(self):
        """Batch autotuner probes so far: list of (triplets per launch, best time ms)."""
        n = ctypes.c_int()
        self._check(self.lib.tc3_get_tuning(self.ctx, 0, None, None, ctypes.byref(n)))
        b = np.zeros(max(1, n.value))
        ms = np.zeros(max(1, n.value))
        self._check(self.lib.tc3_get_tuning(self.ctx, n.value, dptr(b), dptr(ms), ctypes.byref(n)))
        return [(int(b[i]), float(ms[i])) for i in range(n.value)]

    def xi_storage(self):
        """Device storage of the scattering matrix of the last scattering_rates call: (elements
        per temperature, bytes per element, rows stored per orbit)."""
        n, es, po = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.tc3_get_xi_storage(self.ctx, ctypes.byref(n), ctypes.byref(es), ctypes.byref(po)))
        return int(n.value), int(es.value), bool(po.value)

    def write_grid_hdf5(self, path, iT=0):
        self._check(self.lib.tc3_write_grid_hdf5(self.ctx, iT, path.encode()))

    def write_cumulative_hdf5(self, path, iT=0, mfppts=1000, freqpts=1000, dosintegrationtype=3, sigma=1.0,
                              dossigma=1.0, unit="thz"):
        """outfile.cumulative_thermal_conductivity.hdf5 (main.f90:298-340): DOS, cumulative
        kappa against mean free path, kappa with boundary scattering and spectral kappa of
        temperature iT, with thermal_conductivity's --mfppts, --freqpts,
        --dosintegrationtype, --sigma, --dossigma and --unit."""
        species = []
        for s in getattr(self.structure, "symbols", []):
            if s not in species:
                species.append(s)
        self._check(self.lib.tc3_write_cumulative_hdf5(self.ctx, iT, path.encode(), int(mfppts), int(freqpts),
                                                        int(dosintegrationtype), float(sigma), float(dossigma),
                                                        unit.encode(), " ".join(species).encode()))

    def outfile_text(self, iT=0):
        """outfile.thermal_conductivity (main.f90:248-294)."""
        def six(m):
            m = m * KAPPA_AU_TO_SI
            v = [m[0, 0], m[1, 1], m[2, 2], m[0, 1], m[0, 2], m[1, 2]]
            return " " + "".join(" " + _fe(x) for x in v) + "\n"
        hdr = "#" + "".join(" " + s.rjust(24) for s in ("kxx", "kyy", "kzz", "kxy", "kxz", "kyz")) + "\n"
        s, it, od = self.kappa_sma[iT], self.kappa_iter[iT], self.kappa_offdiag[iT]
        t = "# Unit:" + " " * 15 + "W/m/K\n"
        t += "# Temperature:" + " " * 8 + _fe(float(self.temperatures[iT]), 20, 12) + "\n"
        t += "# Single mode approximation\n" + hdr + six(s)
        t += "# Collective contribution\n" + hdr + six(it - s)
        t += "# Off diagonal (coherence) contribution\n" + hdr + six(od)
        t += "# Total thermal conductivity\n" + hdr + six(it + od)
        return t
