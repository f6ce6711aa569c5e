// Cumulative, boundary and spectral thermal conductivity and the phonon density of states
// (tc3_cumulative.cpp) and their HDF5 file (tc3_hdf5.cpp).
#pragma once

#include <string>
#include <vector>

struct tc3_ctx;

namespace tc3 {

// thermal_conductivity's --mfppts, --freqpts, --dosintegrationtype (1 Gaussian, 2 adaptive
// Gaussian, 3 tetrahedron), --sigma, --dossigma, --unit (options.f90:86-147)
struct CumulativeOpts {
    int mfppts = 1000, freqpts = 1000, dosint = 3;
    double sigma = 1.0, dossigma = 1.0;
    std::string unit = "thz";
};

struct PhononDos {
    std::vector<double> omega, dos, pdos_mode, pdos_site;   // [n], [n], [n][nb], [n][na]
    std::vector<double> ds;                                 // default smearing per mode
    double omax = 0, omin = 0;
};

struct CumulativeResult {
    double kappa_total[6];                        // Voigt: xx yy zz yz xz xy
    std::vector<double> mfp;                      // [nirr][nb] scalar mean free paths
    std::vector<double> mfaxis, mf, mf_band, mf_atom;   // [np], [np][6], [np][nb][6], [np][na][6]
    std::vector<double> baxis, bkappa;            // [np], [6][np]
    double spec_total[9];
    std::vector<double> spec, spec_band, spec_atom;     // [n][9], [n][nb][9], [n][na][9]
    double alpha[9];                              // angular-momentum tensor, row-major
    double am_total[9];
    std::vector<double> am_spec;                  // [n][9] spectral angular momentum
};

void lv_tetrahedron_weights(const double *ein, double z, double tol, double sigma, double *w);
void irreducible_tetrahedra(const tc3_ctx *c, std::vector<int> &corners, std::vector<double> &weights);
void phonon_dos(const tc3_ctx *c, const CumulativeOpts &o, const std::vector<int> &tc, const std::vector<double> &tw,
                PhononDos &d);
void cumulative_kappa(const tc3_ctx *c, int iT, const CumulativeOpts &o, CumulativeResult &r);
void boundary_kappa(const tc3_ctx *c, int iT, const CumulativeOpts &o, CumulativeResult &r);
void spectral_kappa(const tc3_ctx *c, int iT, const CumulativeOpts &o, const PhononDos &d, const std::vector<int> &tc,
                    const std::vector<double> &tw, CumulativeResult &r);
void angular_momentum_matrix(const tc3_ctx *c, int iT, double *alpha);
void spectral_angular_momentum(const tc3_ctx *c, int iT, const CumulativeOpts &o, const PhononDos &d,
                               const std::vector<int> &tc, const std::vector<double> &tw, CumulativeResult &r);
int write_cumulative_hdf5(const tc3_ctx *c, int iT, const char *path, const CumulativeOpts &o,
                          const std::string &species, std::string &err);

}  // namespace tc3
