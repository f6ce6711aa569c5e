// Four-phonon scattering terms (tc3_fourphonon.hip).
#pragma once

#include <vector>

struct tc3_ctx;

namespace tc3 {

struct F4Inputs {
    int nT = 0, integrationtype = 2;
    int precision = 0;   // contraction of cells above two atoms: 0 FP64, else FP32 (the mixed modes)
    double sigma_fixed = 0, thresh = 0;
    const double *d_be = nullptr, *d_sigsq = nullptr, *d_uirr = nullptr, *d_ufull = nullptr;
    const int *d_rowof = nullptr;
};

// Linewidth terms g4[T][row] and off-diagonal scattering-matrix terms od4[T][row][q][b],
// already averaged over the little group of each row's q1.
int fourphonon_terms(tc3_ctx *c, const F4Inputs &in, std::vector<double> &g4, double *d_od4);

}  // namespace tc3
