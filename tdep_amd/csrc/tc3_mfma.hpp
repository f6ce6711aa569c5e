// Interface between the scattering driver and the MFMA contraction kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "tc3_device.hpp"

namespace tc3 {

struct MfmaEpiArgs {
    int nb, nT, t0, nt, nirr, integrationtype, cb;
    double sigma_fixed, psi_scale;
    const int *tq1, *tq2, *tq3, *eq_off, *eq_q2, *eq_q3, *irr_of_full, *row_of;
    const double *perm, *om_irr, *om_full, *be, *sigsq;
    double *xi, *gpart;
    size_t xi_row, xi_T;
};

struct MfmaPlan {
    int na, nb, nirr, nfull, maxa1;
    long ntrip, batch;
    const cd *egv_irr, *egv_full;
    const double *om_irr, *om_full;
    void *E1re, *E1im, *Efre, *Efim, *work;
    const int *tq2, *tq3, *a1off, *blk3;
    const double *qcart, *lv2, *lv3, *mwm;
    size_t psi_shmem;
    MfmaEpiArgs epi;
    hipEvent_t *ev;   // 3 events per batch (before k_uw, between, after k_psi), or null
};

size_t mfma_elem_size(int precision);
bool mfma_supported(int nb, int precision);
int mfma_choose_chunk(int nb, int nT, int precision, size_t *shmem);
int run_mfma_chain(const MfmaPlan &p, int precision, hipStream_t st);

}  // namespace tc3
