// outfile.thermal_conductivity_grid.hdf5: the layout of lo_phonon_dispersions%write_to_hdf5
// (type_phonon_dispersions.f90:316-641) as thermal_conductivity calls it (main.f90:323-328,
// 337): file attributes, then per full-grid q-point frequencies, velocities, eigenvectors,
// lifetimes, linewidths, mean free paths, per-mode thermal conductivity and heat capacity.
// Fortran arrays a(i, j, ..) appear in the file in C order (.., j, i), as the reference's
// Fortran HDF5 calls write them; 2-D and larger datasets are chunked (chunk = whole array)
// and deflate-6 compressed like lo_hdf5_helper's (hdf5_wrappers.f90:704-737); units and
// dimension names are string attributes 'unit' and 'dimensions'.
//
// The HDF5 C library (1.10 ABI) is loaded at run time, so libtc3 has no link dependency on
// it; TC3_HDF5_LIB names the library when it is not on the loader's path.
#include <dlfcn.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tc3_ctx.hpp"
#include "tc3_cumulative.hpp"

namespace tc3 {

namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

constexpr unsigned kAccTrunc = 0x0002u;   // H5F_ACC_TRUNC
constexpr hid_t kDefault = 0;             // H5P_DEFAULT, H5S_ALL
constexpr int kScalar = 0;                // H5S_SCALAR
constexpr int kSpacePad = 2;              // H5T_STR_SPACEPAD

struct H5 {
    void *lib = nullptr;
    herr_t (*open)();
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*screate)(int);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    hid_t (*pcreate)(hid_t);
    herr_t (*pset_chunk)(hid_t, int, const hsize_t *);
    herr_t (*pset_deflate)(hid_t, unsigned);
    herr_t (*pclose)(hid_t);
    hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dclose)(hid_t);
    hid_t (*acreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
    herr_t (*awrite)(hid_t, hid_t, const void *);
    herr_t (*aclose)(hid_t);
    hid_t (*tcopy)(hid_t);
    herr_t (*tset_size)(hid_t, size_t);
    herr_t (*tset_strpad)(hid_t, int);
    herr_t (*tclose)(hid_t);
    hid_t *native_double, *native_int, *fortran_s1, *dataset_create;

    template <typename F>
    bool sym(F &f, const char *name) {
        f = (F)dlsym(lib, name);
        return f != nullptr;
    }
    bool load(std::string &err) {
        const char *env = std::getenv("TC3_HDF5_LIB");
        const char *names[] = {env, "libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103"};
        for (const char *n : names)
            if (n && (lib = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
        if (!lib) {
            err = "the HDF5 library (1.10) could not be loaded; set TC3_HDF5_LIB";
            return false;
        }
        bool ok = sym(open, "H5open") && sym(fcreate, "H5Fcreate") && sym(fclose, "H5Fclose") &&
                  sym(screate, "H5Screate") && sym(screate_simple, "H5Screate_simple") && sym(sclose, "H5Sclose") &&
                  sym(pcreate, "H5Pcreate") && sym(pset_chunk, "H5Pset_chunk") &&
                  sym(pset_deflate, "H5Pset_deflate") && sym(pclose, "H5Pclose") && sym(dcreate, "H5Dcreate2") &&
                  sym(dwrite, "H5Dwrite") && sym(dclose, "H5Dclose") && sym(acreate, "H5Acreate2") &&
                  sym(awrite, "H5Awrite") && sym(aclose, "H5Aclose") && sym(tcopy, "H5Tcopy") &&
                  sym(tset_size, "H5Tset_size") && sym(tset_strpad, "H5Tset_strpad") && sym(tclose, "H5Tclose") &&
                  sym(native_double, "H5T_NATIVE_DOUBLE_g") && sym(native_int, "H5T_NATIVE_INT_g") &&
                  sym(fortran_s1, "H5T_FORTRAN_S1_g") && sym(dataset_create, "H5P_CLS_DATASET_CREATE_ID_g");
        if (!ok || open() < 0) {
            err = "the loaded HDF5 library lacks the 1.10 C API";
            return false;
        }
        return true;
    }
    ~H5() {
        if (lib) dlclose(lib);
    }

    // string attribute: fixed-length, space padded, scalar (store_char_as_attribute)
    bool str_attr(hid_t obj, const char *name, const std::string &v) {
        const hid_t t = tcopy(*fortran_s1);
        if (t < 0) return false;
        bool ok = tset_size(t, std::max<size_t>(v.size(), 1)) >= 0 && tset_strpad(t, kSpacePad) >= 0;
        const hid_t s = screate(kScalar);
        const hid_t a = acreate(obj, name, t, s, kDefault, kDefault);
        std::string buf = v.empty() ? std::string(" ") : v;
        ok = ok && a >= 0 && awrite(a, t, buf.data()) >= 0;
        if (a >= 0) aclose(a);
        sclose(s);
        tclose(t);
        return ok;
    }
    // numeric attribute of one element (store_int_as_attribute / store_float_as_attribute)
    bool num_attr(hid_t obj, const char *name, hid_t type, const void *v) {
        const hsize_t one = 1;
        const hid_t s = screate_simple(1, &one, nullptr);
        const hid_t a = acreate(obj, name, type, s, kDefault, kDefault);
        const bool ok = a >= 0 && awrite(a, type, v) >= 0;
        if (a >= 0) aclose(a);
        sclose(s);
        return ok;
    }
    // FP64 dataset with C-order dims; rank >= 2 chunked (whole array) and deflated
    bool data(hid_t file, const char *name, const std::vector<hsize_t> &dims, const std::vector<double> &v,
              const char *unit, const char *dimnames) {
        const hid_t s = screate_simple((int)dims.size(), dims.data(), nullptr);
        hid_t p = kDefault;
        if (dims.size() >= 2) {
            p = pcreate(*dataset_create);
            pset_chunk(p, (int)dims.size(), dims.data());
            pset_deflate(p, 6);
        }
        const hid_t d = dcreate(file, name, *native_double, s, kDefault, p, kDefault);
        bool ok = d >= 0 && dwrite(d, *native_double, kDefault, kDefault, kDefault, v.data()) >= 0;
        if (d >= 0) {
            ok = ok && str_attr(d, "unit", unit) && str_attr(d, "dimensions", dimnames);
            dclose(d);
        }
        if (p != kDefault) pclose(p);
        sclose(s);
        return ok;
    }
};

double cv_of(double T, double om) {
    // lo_harmonic_oscillator_cv (gottochblandat_physics.f90:331-367)
    if (om < kFreqTol || T < kTemperatureTol) return 0.0;
    const double x = om / T / kKbHartree;
    if (x > 200.0) return 0.0;
    if (x > 21.0) return kKbHartree * x * x * std::exp(-x);
    if (x < 2e-2) return kKbHartree * (1.0 - x * x / 12.0 + x * x * x * x / 240.0);
    const double ex = std::exp(x);
    return kKbHartree * x * x * ex / ((ex - 1) * (ex - 1));
}

double planck_of(double T, double om) {
    // lo_planck (gottochblandat_physics.f90:160-188)
    if (T < kTemperatureTol || om < kFreqTol) return 0.0;
    const double x = om / (kKbHartree * T);
    if (x > 1e2) return 0.0;
    return 1.0 / (std::exp(x) - 1.0);
}

}  // namespace

int write_grid_hdf5(const tc3_ctx *c, int iT, const char *path, std::string &err) {
    H5 h;
    if (!h.load(err)) return 6;
    const int nb = c->nb, nq = c->mesh.nfull, nirr = c->mesh.nirr;
    const double T = c->temps[iT];
    const double hz = 1.0 / kHbarHartree;   // lo_frequency_Hartree_to_Hz
    const hsize_t Q = (hsize_t)nq, B = (hsize_t)nb;
    const hid_t f = h.fcreate(path, kAccTrunc, kDefault, kDefault);
    if (f < 0) {
        err = std::string("cannot create ") + path;
        return 6;
    }
    bool ok = true;
    const int na = nb / 3;
    ok = ok && h.num_attr(f, "number_of_atoms", *h.native_int, &na);
    ok = ok && h.num_attr(f, "number_of_bands", *h.native_int, &nb);
    ok = ok && h.num_attr(f, "number_of_qpoints", *h.native_int, &nq);
    ok = ok && h.str_attr(f, "meshtype", "FFT");
    // q-points: stored as held internally (Cartesian 1/Bohr without 2 pi) under the
    // reference's label
    ok = ok && h.data(f, "qpoints", {Q, 3}, c->mesh.qcart, "1/A", "q-vector,xyz");
    std::vector<double> d((size_t)nq * nb);
    for (size_t i = 0; i < d.size(); ++i) d[i] = c->om_full[i] * hz;
    ok = ok && h.data(f, "frequencies", {Q, B}, d, "Hz (angular)", "q-vector,mode");
    std::vector<double> v3((size_t)nq * nb * 3);
    for (size_t i = 0; i < v3.size(); ++i) v3[i] = c->vel_full[i] * kGroupvelAuToMs;
    ok = ok && h.data(f, "group_velocities", {Q, B, 3}, v3, "m/s", "q-vector,mode,xyz");
    std::vector<double> er((size_t)nq * nb * nb), ei(er.size());
    for (int q = 0; q < nq; ++q)
        for (int m = 0; m < nb; ++m)
            for (int k = 0; k < nb; ++k) {
                const cd x = c->egv_full[((size_t)q * nb + k) * nb + m];
                er[((size_t)q * nb + m) * nb + k] = x.x;
                ei[((size_t)q * nb + m) * nb + k] = x.y;
            }
    ok = ok && h.data(f, "eigenvectors_re", {Q, B, B}, er, "dimensionless", "q-vector,mode,atom-xyz");
    ok = ok && h.data(f, "eigenvectors_im", {Q, B, B}, ei, "dimensionless", "q-vector,mode,atom-xyz");
    const double *lw = &c->linewidth[(size_t)iT * nirr * nb];
    for (int q = 0; q < nq; ++q)
        for (int j = 0; j < nb; ++j) {
            const double g = lw[(size_t)c->mesh.irr_of_full[q] * nb + j];
            d[(size_t)q * nb + j] = (g > kFreqTol / 10) ? 0.5 / (g * hz) : 0.0;
        }
    ok = ok && h.data(f, "lifetimes", {Q, B}, d, "s", "q-vector,mode");
    std::vector<double> lat(9);
    for (int i = 0; i < 9; ++i) lat[i] = kBohrToM * 1e10 * c->cr.lat[i];
    ok = ok && h.data(f, "lattice_vectors", {3, 3}, lat, "A", "xyz,xyz");
    for (int q = 0; q < nq; ++q)
        for (int j = 0; j < nb; ++j) d[(size_t)q * nb + j] = lw[(size_t)c->mesh.irr_of_full[q] * nb + j] * hz;
    ok = ok && h.data(f, "linewidths", {Q, B}, d, "Hz", "q-vector,mode");
    for (int q = 0; q < nq; ++q)
        for (int j = 0; j < nb; ++j) {
            const double g = lw[(size_t)c->mesh.irr_of_full[q] * nb + j];
            double x = 0.0;
            if (g > kFreqTol) {
                const double *v = &c->vel_full[((size_t)q * nb + j) * 3];
                x = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * 0.5 / g;
                x = (x > kTol) ? x * kBohrToM : 0.0;
            }
            d[(size_t)q * nb + j] = x;
        }
    ok = ok && h.data(f, "mean_free_paths", {Q, B}, d, "m", "q-vector,mode");
    // per-mode kappa from Fn (with the omega/T prefactor of main.f90:323-328) rotated to the
    // full grid
    if (c->have_kappa && !c->Fn.empty()) {
        std::vector<double> k4((size_t)nq * nb * 9, 0.0);
        const double *Fn = &c->Fn[(size_t)iT * nirr * nb * 3];
        for (int q = 0; q < nq; ++q) {
            const int l = c->mesh.irr_of_full[q], k = c->mesh.op_of_full[q];
            const double *m = c->ops[std::abs(k) - 1].m;
            const double sg = (k > 0) ? 1.0 : -1.0;
            for (int j = 0; j < nb; ++j) {
                if (c->om_full[(size_t)q * nb + j] < c->omega_min * 0.5) continue;
                const double om = c->om_irr[(size_t)l * nb + j];
                const double pre = (om < kFreqTol) ? 1.0 : om / T;
                const double *F = &Fn[((size_t)l * nb + j) * 3];
                const double *v = &c->vel_irr[((size_t)l * nb + j) * 3];
                double v0[3], v1[3];
                for (int i = 0; i < 3; ++i) {
                    v0[i] = sg * pre * (m[3 * i] * F[0] + m[3 * i + 1] * F[1] + m[3 * i + 2] * F[2]);
                    v1[i] = sg * (m[3 * i] * v[0] + m[3 * i + 1] * v[1] + m[3 * i + 2] * v[2]);
                }
                const double n = planck_of(T, om);
                const double f0 = om * (n + 1) * n / (c->cr.volume * kKbHartree * T) * kKappaAuToSi;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) k4[(((size_t)q * nb + j) * 3 + a) * 3 + b] = f0 * v0[b] * v1[a];
            }
        }
        ok = ok && h.data(f, "thermal_conductivity", {Q, B, 3, 3}, k4, "W/mK", "q-vector,mode,xyz,xyz");
    }
    ok = ok && h.num_attr(f, "temperature", *h.native_double, &T);
    for (size_t i = 0; i < d.size(); ++i) d[i] = cv_of(T, c->om_full[i]) * kHartreeToJoule;
    ok = ok && h.data(f, "harmonic_heat_capacity", {Q, B}, d, "J/K", "q-vector,mode");
    ok = (h.fclose(f) >= 0) && ok;
    if (!ok) {
        err = std::string("HDF5 write failed for ") + path;
        return 6;
    }
    return 0;
}

// outfile.cumulative_thermal_conductivity.hdf5 (lo_cumulative_kappa%write_to_hdf5,
// cumulative_kappa.f90:580-777, as main.f90:298-340 fills it) for temperature iT.  Fortran
// arrays d(p, .., i, j) appear in C order (j, i, .., p); the mean-free-path and boundary
// tensors are expanded from Voigt order (xx yy zz yz xz xy), the spectral ones from the
// nine components in row order; the angular-momentum spectrum from its nine components in
// column order (the reshape of cumulative_kappa.f90:762) and the tensor as stored, in
// atomic units (:770-771).
int write_cumulative_hdf5(const tc3_ctx *c, int iT, const char *path, const CumulativeOpts &o,
                          const std::string &species, std::string &err) {
    double uf;
    const char *omstr, *dosstr;
    if (o.unit == "thz") {
        uf = kFreqHartreeToThz; omstr = "THz"; dosstr = "States/THz";
    } else if (o.unit == "mev") {
        uf = 1000.0 * kHartreeToEv; omstr = "meV"; dosstr = "States/meV";
    } else if (o.unit == "icm") {
        uf = 1.0 / kHbarHartree / kTwoPi / 299792458.0 / 100.0; omstr = "cm^-1"; dosstr = "States/cm^-1";
    } else {
        err = "Unknown unit";
        return 5;
    }
    std::vector<int> tc;
    std::vector<double> tw;
    if (o.dosint == 3) irreducible_tetrahedra(c, tc, tw);
    PhononDos d;
    phonon_dos(c, o, tc, tw, d);
    CumulativeResult r;
    cumulative_kappa(c, iT, o, r);
    boundary_kappa(c, iT, o, r);
    spectral_kappa(c, iT, o, d, tc, tw, r);
    angular_momentum_matrix(c, iT, r.alpha);
    spectral_angular_momentum(c, iT, o, d, tc, tw, r);

    H5 h;
    if (!h.load(err)) return 6;
    const hid_t f = h.fcreate(path, kAccTrunc, kDefault, kDefault);
    if (f < 0) {
        err = std::string("cannot create ") + path;
        return 6;
    }
    const int na = c->cr.na, nb = c->nb, n = o.freqpts, np = o.mfppts;
    const hsize_t N = (hsize_t)n, P = (hsize_t)np, B = (hsize_t)nb, A = (hsize_t)na;
    bool ok = h.str_attr(f, "atomic_species", species);
    // the reference stores the word itself (cumulative_kappa.f90:624)
    ok = ok && h.str_attr(f, "temperature", "temperature");
    std::vector<double> v(n);
    for (int i = 0; i < n; ++i) v[i] = d.omega[i] * uf;
    ok = ok && h.data(f, "frequencies", {N}, v, omstr, "frequency");
    for (int i = 0; i < n; ++i) v[i] = d.dos[i] / uf;
    ok = ok && h.data(f, "dos", {N}, v, dosstr, "frequency");
    v.resize(np);
    for (int i = 0; i < np; ++i) v[i] = r.mfaxis[i] * kBohrToM;
    ok = ok && h.data(f, "mean_free_path_axis", {P}, v, "m", "mfp");
    static const int voigt[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
    // [j][i][m][p] of a [p][m][6] Voigt array
    auto expand6 = [&](const std::vector<double> &src, int np_, int nm) {
        std::vector<double> out((size_t)9 * nm * np_);
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i)
                for (int m = 0; m < nm; ++m)
                    for (int p = 0; p < np_; ++p)
                        out[(((size_t)j * 3 + i) * nm + m) * np_ + p] =
                            src[((size_t)p * nm + m) * 6 + voigt[i][j]] * kKappaAuToSi;
        return out;
    };
    auto expand9 = [&](const std::vector<double> &src, int nx, int nm) {
        std::vector<double> out((size_t)9 * nm * nx);
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i)
                for (int m = 0; m < nm; ++m)
                    for (int x = 0; x < nx; ++x)
                        out[(((size_t)j * 3 + i) * nm + m) * nx + x] = src[((size_t)x * nm + m) * 9 + 3 * i + j] * kKappaAuToSi / uf;
        return out;
    };
    ok = ok && h.data(f, "cumulative_kappa_vs_mean_free_path", {3, 3, P}, expand6(r.mf, np, 1), "W/m/K", "xyz,xyz,mfp");
    ok = ok && h.data(f, "cumulative_kappa_vs_mean_free_path_per_mode", {3, 3, B, P}, expand6(r.mf_band, np, nb), "W/m/K",
                      "xyz,xyz,mode,mfp");
    ok = ok && h.data(f, "cumulative_kappa_vs_mean_free_path_per_atom", {3, 3, A, P}, expand6(r.mf_atom, np, na), "W/m/K",
                      "xyz,xyz,atom,mfp");
    ok = ok && h.data(f, "spectral_kappa_vs_frequency", {3, 3, N}, expand9(r.spec, n, 1), "W/m/K", "xyz,xyz,frequency");
    ok = ok && h.data(f, "spectral_kappa_vs_frequency_per_mode", {3, 3, B, N}, expand9(r.spec_band, n, nb), "W/m/K",
                      "xyz,xyz,mode,frequency");
    ok = ok && h.data(f, "spectral_kappa_vs_frequency_per_atom", {3, 3, A, N}, expand9(r.spec_atom, n, na), "W/m/K",
                      "xyz,xyz,atom,frequency");
    for (int i = 0; i < np; ++i) v[i] = r.baxis[i] * kBohrToM;
    ok = ok && h.data(f, "boundary_scattering_lengths", {P}, v, "m", "domainsize");
    std::vector<double> bt((size_t)np * 6);
    for (int p = 0; p < np; ++p)
        for (int k = 0; k < 6; ++k) bt[(size_t)p * 6 + k] = r.bkappa[(size_t)k * np + p];
    ok = ok && h.data(f, "boundary_scattering_kappa", {3, 3, P}, expand6(bt, np, 1), "W/m/K", "xyz,xyz,domainsize");
    std::vector<double> am((size_t)9 * n);
    for (int cc = 0; cc < 3; ++cc)
        for (int rr = 0; rr < 3; ++rr)
            for (int x = 0; x < n; ++x) am[((size_t)cc * 3 + rr) * n + x] = r.am_spec[(size_t)x * 9 + 3 * cc + rr] / uf / kBohrToM;
    ok = ok && h.data(f, "generating_angular_momentum_tensor_vs_frequency", {3, 3, N}, am, "hbar/m/K",
                      "xyz,xyz,frequency");
    std::vector<double> al(9);
    for (int cc = 0; cc < 3; ++cc)
        for (int rr = 0; rr < 3; ++rr) al[3 * cc + rr] = r.alpha[3 * rr + cc];
    ok = ok && h.data(f, "generating_angular_momentum_tensor", {3, 3}, al, "hbar/m/K", "xyz,xyz");
    ok = (h.fclose(f) >= 0) && ok;
    if (!ok) {
        err = std::string("HDF5 write failed for ") + path;
        return 6;
    }
    return 0;
}

}  // namespace tc3
