// Interface between the scattering driver and the fused 48-band contraction (tc3_fused.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "tc3_device.hpp"

namespace tc3 {

struct FusedArgs {
    int ntrip, nirr, nT, gT, T0, integrationtype, nfc3, nblk, npair, npblk;
    double sigma_fixed, psi_scale;
    const int *tq1, *tq2, *tq3, *eq_off, *irr_of_full, *row_of;
    const double *perm, *om_irr, *om_full, *be, *sigsq;   // be: first temperature of the chunk
    const double *qcart, *lv2, *lv3, *mwm;                // IFC entries sorted by (a1, a2, a3)
    const int4 *blocks;      // nonzero (a1, a2, a3) blocks in entry order: (first entry, end entry,
                             // slot position of its first value, unused); the slot holds per pair,
                             // per block, the 27 values (i, j, k), so that V entry (j, k) of a pair
                             // reads its blocks' values at stride 9
    const int4 *pairs;       // (a2, a3) pairs with blocks: (a2, a3, slot base, number of blocks)
    const int *pa1;          // a1 of every block, pair by pair
    const __bf16 *ftab;      // E fragments of every full-grid point
    const float *e1re, *e1im;   // E1 planar [q_irr][component][mode]
    float2 *phi_slot;        // per workgroup: Phi blocks of its current triplet [nblk][27]
    int *queue;              // triplet counter
    float *comp;             // compact sums, first temperature of the chunk
    size_t comp_T;
    double *gpart;
    const float *tw;         // Monte-Carlo draw counts [t][b1], or null
};

struct FusedPlan {
    int nirr, nfull, nT, nty, nwg;
    const cd *egv_irr, *egv_full;
    const double *om_irr, *om_full, *be;
    __bf16 *ftab;
    float *e1re, *e1im;
    void *comp;
    FusedArgs args;
    hipEvent_t *ev;   // 3 events: before the tables, before the contraction, after it (or null)
};

// Whether the fused kernel handles this case (48 bands, split-BF16x3, integration type 1 or
// 2; IFC entries, blocks, pairs, entries per block maxe and blocks per pair maxb within its
// LDS and register budget), the bytes of its E table for nq points and the temperature block
// (4 or 8) for nT temperatures.
bool fused_supported(int nb, int precision, int integrationtype, int nfc3, int nblk, int npair, int maxe, int maxb);
size_t fused_table_bytes(int nq);
int fused_temperature_block(int nT);
int run_fused(const FusedPlan &p, hipStream_t st);

}  // namespace tc3
