// Interface between the scattering driver and the MFMA contraction kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "tc3_device.hpp"

namespace tc3 {

struct MfmaEpiArgs {
    int nb, nT, t0, nt, nirr, integrationtype;
    int gT, T0;              // temperature stride and offset of gpart (temperature chunks)
    double sigma_fixed, psi_scale;
    const int *tq1, *tq2, *tq3, *eq_off, *eq_q2, *eq_q3, *irr_of_full, *row_of;
    const double *perm, *om_irr, *om_full, *be, *sigsq;   // be: first temperature of the chunk
    void *xi;                // first temperature of the chunk; FP64 entries in FP64 mode, else FP32
    double *gpart;
    size_t xi_row, xi_T;
    // Compact per-triplet sums instead of scattering-matrix entries, when not null:
    // comp[T][t][b1][0][b2] = sum_b3 f_T (the q2' entries before their frequency factor) and
    // comp[T][t][b1][1][b3] = sum_b2 f_T (the q3' entries), in the epilogue type; first
    // temperature of the chunk, comp_T elements per temperature.
    void *comp;
    size_t comp_T;
    const double *u_irr, *u_full;   // integration type 6: reduced group velocities [q][mode][3]
    double thresh6;                 // and the baseline width
    const float *tw;                // Monte-Carlo draw counts [t][b1], or null (all 1)
    void *psi_out;                  // run_mfma_psi2 only: |Psi|^2 [t][b1][b2][b3], epilogue type
};

struct MfmaPlan {
    int na, nb, nirr, nfull, maxa1, maxblk, nty;
    long ntrip, batch;
    const cd *egv_irr, *egv_full;
    const double *om_irr, *om_full;
    void *E1re, *E1im, *Efre, *Efim, *work;   // planar e/sqrt(w), [q][component][mode]
    void *Eti, *Etf;                          // split bf16 planes, [q][plane][mode][k] (BF16X3)
    const int *tq2, *tq3, *a1off;
    const int4 *blist;                        // nonzero (a2, a3) blocks of every a1
    const int *a1_boff, *a12_boff;
    const double *qcart, *lv2, *lv3, *mwm;
    MfmaEpiArgs epi;
    hipEvent_t *ev;   // 3 events per launch pair (before k_uw, between, after k_psi), or null
    int psi_bb = 1;   // blocks of 16 b1 per k_psi workgroup: 1, or nb/16 (split-BF16x3, 48 bands)
};

// Whether the MFMA kernels handle this band count and precision, and the temperature
// block (4, 8 or 16) a pass uses for nT temperatures (0 if none fits the LDS).
bool mfma_supported(int nb, int precision);
int mfma_temperature_block(int nb, int nT, int precision);
size_t mfma_elem_size(int precision);
size_t mfma_split_elems(int nb);   // bf16 elements of one q-point's split planes
int run_mfma_chain(const MfmaPlan &p, int precision, hipStream_t st);
// |Psi|^2 with its prefactor (psi_scale, perm) of the plan's triplets, every element, into
// p.epi.psi_out ([t][b1][b2][b3], FP64 in FP64 mode, FP32 otherwise), by the production
// k_uw and the k_psi contraction; p.batch triplets per launch pair.
int run_mfma_psi2(const MfmaPlan &p, int precision, hipStream_t st);

}  // namespace tc3
