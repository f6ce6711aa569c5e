// |Psi|^2 contraction on MFMA for band counts nb = 48 and 96 (16 and 32 atoms).
//
// Per batch of visited triplets t = (q1, q2, q3) two kernels run:
//
//   k_uw   workgroup per (t, ia).  Phi(q2,q3) is block sparse in the atoms (a1, a2, a3):
//          only the (a2, a3) blocks of a1 = ia/3 that hold IFCs are summed
//          (pretransform_phi3, scattering_threephonon.f90:275-317).  U[ib][b3] =
//          sum_ic Phi[ia][ib][ic] E3[ic][b3] is formed block by block on the vector ALUs,
//          W[b2][b3] = sum_ib E2[ib][b2] U[ib][b3] on MFMA, and W[t][ia][b2][b3] is stored.
//   k_psi  workgroup per (t, 16 values of b1) with one wave per 16 values of b3.  For each
//          b2 a wave forms the tile Psi^T[(b2, b3)][b1] = sum_ia W[ia][b2][b3] E1[ia][b1]
//          on MFMA (:88-113) and weights |Psi|^2 with the smeared energy conservation
//          (get_dirac :219-271).  With P = psisq (d0 + d1) and M = psisq (d1 - d0) the
//          per-temperature rate f0 + f1 of :116-123 is
//              f_T = P (n2_T + 1/2) + M (n3_T + 1/2),
//          so the occupations enter only through the two sums taken afterwards:
//            * over b3 (linewidth, q2' entries of the scattering matrix, :126-133): one
//              more MFMA per tile whose output columns are the temperatures;
//            * over b2 (q3' entries, :134): per-lane registers, since a wave keeps its b3
//              range for the whole triplet.
//          Cross-wave sums go through LDS in a fixed order, so results do not depend on
//          scheduling.
//
// Complex products are four real products (re = ar br - ai bi, im = ar bi + ai br).
// Precision modes:
//   FP64    v_mfma_f64_16x16x4_f64, FP64 epilogue
//   FP32    v_mfma_f32_16x16x4_f32, FP32 epilogue
//   BF16X3  v_mfma_f32_16x16x32_bf16 with every FP32 operand split as hi + lo (two
//           bf16), a.b ~ ah.bh + ah.bl + al.bh, FP32 accumulation and epilogue.
// Lane layouts (16x16 output tile, 64 lanes, g = lane/16, c = lane%16): A[m = c][k],
// B[k][n = c] with k = k0 + g (K-step 4) or k = k0 + 8g + j, j < 8 (K-step 32); D[m][n = c]
// in register r < 4 with m = 4g + r (FP32 accumulation) or m = 4r + g (FP64).  Checked on
// the device by an in-bounds probe.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "tc3_ctx.hpp"
#include "tc3_mfma.hpp"

namespace tc3 {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kPlanes = 6;   // split planes: re hi, re lo, im hi, im lo, -im hi, -im lo
constexpr int kChunk = 8;    // b2 values between cross-wave reductions in k_psi
constexpr double kTol = 1e-9;

constexpr int kpad(int nb) { return (nb + 31) / 32 * 32; }

template <int PREC> struct Traits;
template <> struct Traits<0> { typedef double T; typedef v4d Acc; };
template <> struct Traits<1> { typedef float T; typedef v4f Acc; };
template <> struct Traits<2> { typedef float T; typedef v4f Acc; };

// K-steps of a contraction over nb: 4 per step (FP64, FP32) or 32, zero padded (BF16).
template <int PREC, int NB> struct KSteps { static constexpr int n = NB / 4; };
template <int NB> struct KSteps<2, NB> { static constexpr int n = (NB + 31) / 32; };

// Epilogue type: FP64 in FP64 mode, FP32 otherwise.
template <int PREC> struct Stage { typedef float T; };
template <> struct Stage<0> { typedef double T; };

struct Split { v8bf h, l; };

typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

// x ~ hi + lo with hi = bf16(x) (round to nearest even) and lo = bf16(x - hi), two values
// per conversion instruction.
__device__ inline void split8(const float *x, Split &s) {
    v4u hu, lu;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const v2f v = {x[2 * p], x[2 * p + 1]};
        const unsigned u = __builtin_bit_cast(unsigned, __builtin_convertvector(v, v2bf));
        const v2f hf = {__builtin_bit_cast(float, u << 16), __builtin_bit_cast(float, u & 0xffff0000u)};
        hu[p] = u;
        lu[p] = __builtin_bit_cast(unsigned, __builtin_convertvector(v - hf, v2bf));
    }
    s.h = __builtin_bit_cast(v8bf, hu);
    s.l = __builtin_bit_cast(v8bf, lu);
}

__device__ inline v8bf ld8(const __bf16 *p) { return *(const v8bf *)p; }

__device__ inline void mma3(v4f &c, const Split &a, const Split &b) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.h, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.l, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.l, b.h, c, 0, 0, 0);
}
__device__ inline void mma1(v4f &c, float a, float b) { c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
__device__ inline void mma1(v4d &c, double a, double b) { c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

template <int PREC>
__device__ inline int out_row(int lane, int r) {
    if constexpr (PREC == 0) return 4 * r + (lane >> 4);
    else return 4 * (lane >> 4) + r;
}

__device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// e[q][component][mode] / sqrt(omega), transposed to [q][plane][mode][k] and split into
// bf16 hi/lo planes of re, im and -im; k padded with zeros to a multiple of 32.
__global__ void k_split_t(long nq, int nb, int kp, const cd *__restrict__ egv, const double *__restrict__ om,
                          __bf16 *__restrict__ et) {
    const long n = nq * nb * kp;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int k = (int)(i % kp);
        const long qm = i / kp;
        const int mode = (int)(qm % nb);
        const long q = qm / nb;
        float re = 0.f, im = 0.f;
        if (k < nb) {
            const double w = om[q * nb + mode];
            const double s = (w < kTol) ? 0.0 : 1.0 / sqrt(w);
            const cd e = egv[(q * nb + k) * nb + mode];
            re = (float)(e.x * s);
            im = (float)(e.y * s);
        }
        const __bf16 rh = (__bf16)re, ih = (__bf16)im;
        const __bf16 rl = (__bf16)(re - (float)rh), il = (__bf16)(im - (float)ih);
        const size_t pstride = (size_t)nb * kp;
        __bf16 *o = et + (size_t)q * kPlanes * pstride + (size_t)mode * kp + k;
        o[0] = rh;
        o[pstride] = rl;
        o[2 * pstride] = ih;
        o[3 * pstride] = il;
        o[4 * pstride] = -ih;
        o[5 * pstride] = -il;
    }
}

// e[q][component][mode] / sqrt(omega) as planar re / im in the contraction precision.
template <typename T>
__global__ void k_scale_planar(long n, int nb, const cd *__restrict__ egv, const double *__restrict__ om,
                               T *__restrict__ re, T *__restrict__ im) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long q = i / ((long)nb * nb);
        const int b = (int)(i % nb);
        const double w = om[q * nb + b];
        const double s = (w < kTol) ? 0.0 : 1.0 / sqrt(w);
        re[i] = (T)(egv[i].x * s);
        im[i] = (T)(egv[i].y * s);
    }
}

struct UwArgs {
    int t0, maxblk;
    const int *tq2, *tq3;
    const double *qcart, *lv2, *lv3, *mwm;
    const int *a1_off;      // IFC entries of atom a1: [a1_off[a1], a1_off[a1 + 1])
    const int4 *blist;      // (a2, a3, first entry, end entry) of every nonzero block, by (a1, a2, a3)
    const int *a1_boff;     // blocks of a1: [a1_boff[a1], a1_boff[a1 + 1])
    const int *a12_boff;    // blocks of (a1, a2): [a12_boff[a1 na + a2], a12_boff[a1 na + a2 + 1])
    const void *Ef_re, *Ef_im;
    const __bf16 *Et;
    void *W;                // [tl][re/im][ia][b2 nb + b3]
};

template <typename T> struct C2 { T x, y; };

// U on MFMA (dense Phi_i . E3 in split BF16) for nb = 48 in the BF16X3 mode; on the
// vector ALUs over the nonzero IFC blocks otherwise.
template <int PREC, int NB> struct UwMfmaU { static constexpr bool on = (PREC == 2 && NB == 48); };

// LDS of k_uw: the U region (BF16X3: split planes [4][NB][RS]; else planar [2][NB][NB]),
// which also holds the staged IFC entries of a1 and, with U on MFMA, Phi_i; the blocks of
// a1 (int4); the Phi blocks of a1 [block][i][j][k]; the phases of the entries of a1.
template <int PREC, int NB>
__host__ __device__ constexpr size_t uw_region() {
    typedef typename Traits<PREC>::T T;
    constexpr int RS = kpad(NB) + 8;
    return (PREC == 2) ? sizeof(__bf16) * 4 * NB * RS : sizeof(T) * 2 * NB * NB;
}
template <int PREC, int NB>
size_t uw_lds(int maxblk, int maxa1) {
    typedef typename Traits<PREC>::T T;
    size_t s = uw_region<PREC, NB>() + sizeof(int4) * (size_t)maxblk + sizeof(C2<T>) * 27 * (size_t)maxblk;
    s = (s + 15) & ~(size_t)15;
    return s + sizeof(cd) * (size_t)std::max(maxa1, 1);
}

// Raw buffer descriptor over [p, p + bytes) (gfx9 word 3).
__device__ inline __amdgpu_buffer_rsrc_t rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ inline float bload(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ inline double bload2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <typename T>
__device__ inline T bload_t(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    if constexpr (sizeof(T) == 8) return bload2(r, voff, soff);
    else return bload(r, voff, soff);
}

// -x for a split operand: both halves change sign.
__device__ inline v8bf negbf(v8bf x) {
    v4u u = __builtin_bit_cast(v4u, x);
    u ^= 0x80008000u;
    return __builtin_bit_cast(v8bf, u);
}

// Four FP32 values split into bf16 hi and lo (as split8), packed two per word.
__device__ inline void split4(const v4f x, uint2 &h, uint2 &l) {
    unsigned hu[2], lu[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const v2f v = {x[2 * p], x[2 * p + 1]};
        const unsigned u = __builtin_bit_cast(unsigned, __builtin_convertvector(v, v2bf));
        const v2f hf = {__builtin_bit_cast(float, u << 16), __builtin_bit_cast(float, u & 0xffff0000u)};
        hu[p] = u;
        lu[p] = __builtin_bit_cast(unsigned, __builtin_convertvector(v - hf, v2bf));
    }
    h = make_uint2(hu[0], hu[1]);
    l = make_uint2(lu[0], lu[1]);
}

// Workgroup per (triplet, atom a1), one wave per 16 values of b2.
//   Phi blocks: the IFC entries of a1 and their phases exp(-2 pi i (q2.R2 + q3.R3)) are
//   staged in LDS, and every nonzero block (a2, a3) of a1 is summed once for all 27 (i, j,
//   k) (pretransform_phi3, scattering_threephonon.f90:275-317).
//   For i = 0, 1, 2 (ia = 3 a1 + i):
//     U[ib][b3] = sum_ic Phi[ia][ib][ic] E3[ic][b3]: BF16X3 with nb = 48 on MFMA, Phi_i
//     written densely into LDS (zero outside the blocks), each wave one tile of 16 b3 with
//     its E3 fragments in registers; otherwise on the vector ALUs over the nonzero blocks
//     (four b3 per thread);
//     W[b2][b3] = sum_ib E2[ib][b2] U[ib][b3] on MFMA, each wave its 16 b2 with the E2
//     fragments in registers for the whole workgroup; W[t][ia][b2][b3] is stored.
template <int PREC, int NB>
__global__ void __launch_bounds__(4 * NB) k_uw(UwArgs u) {
    typedef typename Traits<PREC>::T T;
    typedef typename Traits<PREC>::Acc Acc;
    constexpr int NA = NB / 3, NB2 = NB * NB, NT = NB / 16;
    constexpr int KP = kpad(NB), RS = KP + 8, KS = KSteps<PREC, NB>::n;
    constexpr bool MU = UwMfmaU<PREC, NB>::on;
    extern __shared__ __align__(16) unsigned char smem[];
    const int tl = blockIdx.x / NA, a1 = blockIdx.x % NA;
    const int t = u.t0 + tl;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, c16 = lane & 15;
    constexpr size_t ub = uw_region<PREC, NB>();
    __bf16 *ut = (__bf16 *)smem;
    T *us = (T *)smem;
    T *mst = (T *)smem;                    // staged IFC entries of a1 (before the i loop)
    int4 *pbl = (int4 *)(smem + ub);
    C2<T> *pb = (C2<T> *)(pbl + u.maxblk);
    cd *ph = (cd *)(smem + align16((size_t)((unsigned char *)(pb + 27 * u.maxblk) - smem)));
    const int q2 = u.tq2[t], q3 = u.tq3[t];
    const double *Q2 = u.qcart + 3 * q2, *Q3 = u.qcart + 3 * q3;
    const int f0 = u.a1_off[a1], f1 = u.a1_off[a1 + 1], nf = f1 - f0;
    const int bo0 = u.a1_boff[a1], nblk = u.a1_boff[a1 + 1] - bo0;
    const bool staged = (size_t)27 * nf * sizeof(T) <= ub;
    if (staged)
        for (int x = threadIdx.x; x < 27 * nf; x += blockDim.x) mst[x] = (T)u.mwm[27 * (size_t)f0 + x];
    for (int f = f0 + threadIdx.x; f < f1; f += blockDim.x) {
        const double arg = -kTwoPi * (Q2[0] * u.lv2[3 * f] + Q2[1] * u.lv2[3 * f + 1] + Q2[2] * u.lv2[3 * f + 2] +
                                      Q3[0] * u.lv3[3 * f] + Q3[1] * u.lv3[3 * f + 1] + Q3[2] * u.lv3[3 * f + 2]);
        double s, c;
        sincos(arg, &s, &c);
        ph[f - f0] = {c, s};
    }
    for (int bl = threadIdx.x; bl < nblk; bl += blockDim.x) pbl[bl] = u.blist[bo0 + bl];
    // E2 operand (A[m = b2][k = ib] = E2[ib][b2]) of this wave's b2 tile, for all i
    const int m0 = 16 * w;
    Split ea_r[PREC == 2 ? KS : 1], ea_i[PREC == 2 ? KS : 1], ea_n[PREC == 2 ? KS : 1];
    Split eb_r[MU ? KS : 1], eb_i[MU ? KS : 1];
    T fa_r[PREC == 2 ? 1 : KS], fa_i[PREC == 2 ? 1 : KS];
    if constexpr (PREC == 2) {
        const size_t ps = (size_t)NB * KP;
        const __bf16 *pa = u.Et + (size_t)q2 * kPlanes * ps + (size_t)(m0 + c16) * KP + 8 * g;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            ea_r[ks].h = ld8(pa + 32 * ks);
            ea_r[ks].l = ld8(pa + ps + 32 * ks);
            ea_i[ks].h = ld8(pa + 2 * ps + 32 * ks);
            ea_i[ks].l = ld8(pa + 3 * ps + 32 * ks);
            ea_n[ks].h = ld8(pa + 4 * ps + 32 * ks);
            ea_n[ks].l = ld8(pa + 5 * ps + 32 * ks);
        }
        if constexpr (MU) {
            // E3 operand (B[k = ic][n = b3] = E3[ic][b3]) of this wave's b3 tile 16 w
            const __bf16 *pe = u.Et + (size_t)q3 * kPlanes * ps + (size_t)(m0 + c16) * KP + 8 * g;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                eb_r[ks].h = ld8(pe + 32 * ks);
                eb_r[ks].l = ld8(pe + ps + 32 * ks);
                eb_i[ks].h = ld8(pe + 2 * ps + 32 * ks);
                eb_i[ks].l = ld8(pe + 3 * ps + 32 * ks);
            }
        }
    } else {
        const T *E2r = (const T *)u.Ef_re + (size_t)q2 * NB2 + m0 + c16;
        const T *E2i = (const T *)u.Ef_im + (size_t)q2 * NB2 + m0 + c16;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            fa_r[ks] = E2r[(4 * ks + g) * NB];
            fa_i[ks] = E2i[(4 * ks + g) * NB];
        }
    }
    __syncthreads();
    // Phi blocks of a1, all (i, j, k): pb[block][9 i + 3 j + k]
    for (int it = threadIdx.x; it < 27 * nblk; it += blockDim.x) {
        const int bl = it / 27, ijk = it % 27;
        const int4 b = pbl[bl];
        double ax = 0, ay = 0;
        for (int f = b.z; f < b.w; ++f) {
            const double m = staged ? (double)mst[27 * (f - f0) + ijk] : u.mwm[27 * (size_t)f + ijk];
            const cd p = ph[f - f0];
            ax += m * p.x;
            ay += m * p.y;
        }
        pb[it] = {(T)ax, (T)ay};
    }
    __syncthreads();
    if constexpr (PREC == 2) {
        // zero the region including the K padding read by the MFMA fragments (the staged
        // entries are no longer needed)
        uint4 *z = (uint4 *)ut;
        for (int x = threadIdx.x; x < 4 * NB * RS / 8; x += blockDim.x) z[x] = make_uint4(0, 0, 0, 0);
        __syncthreads();
    }
    const T *E3r = (const T *)u.Ef_re + (size_t)q3 * NB2, *E3i = (const T *)u.Ef_im + (size_t)q3 * NB2;
    T *Wt = (T *)u.W + (size_t)tl * 2 * NB * NB2;
    for (int i = 0; i < 3; ++i) {
        if constexpr (MU) {
            // Phi_i[ib][ic] in split planes (re hi, re lo, im hi, im lo), zero outside the blocks
            if (i > 0) {
                for (int x = threadIdx.x; x < 4 * NB * (NB / 8); x += blockDim.x) {
                    const int row = x / (NB / 8), c8 = x % (NB / 8);
                    *(uint4 *)(ut + (size_t)row * RS + 8 * c8) = make_uint4(0, 0, 0, 0);
                }
                __syncthreads();
            }
            for (int it = threadIdx.x; it < 9 * nblk; it += blockDim.x) {
                const int bl = it / 9, jk = it % 9;
                const int4 b = pbl[bl];
                const C2<T> v = pb[27 * bl + 9 * i + jk];
                const int ib = 3 * b.x + jk / 3, ic = 3 * b.y + jk % 3;
                const __bf16 rh = (__bf16)v.x, ih = (__bf16)v.y;
                __bf16 *o = ut + (size_t)ib * RS + ic;
                o[0] = rh;
                o[(size_t)NB * RS] = (__bf16)(v.x - (float)rh);
                o[(size_t)2 * NB * RS] = ih;
                o[(size_t)3 * NB * RS] = (__bf16)(v.y - (float)ih);
            }
            __syncthreads();
            // U[ib][b3] for this wave's b3 tile, every ib tile
            v4f ur[NT], ui[NT];
#pragma unroll
            for (int mt = 0; mt < NT; ++mt) {
                ur[mt] = {0, 0, 0, 0};
                ui[mt] = {0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const __bf16 *pa = ut + (size_t)(16 * mt + c16) * RS + 32 * ks + 8 * g;
                    Split ar, ai, an;
                    ar.h = ld8(pa);
                    ar.l = ld8(pa + (size_t)NB * RS);
                    ai.h = ld8(pa + (size_t)2 * NB * RS);
                    ai.l = ld8(pa + (size_t)3 * NB * RS);
                    an.h = negbf(ai.h);
                    an.l = negbf(ai.l);
                    mma3(ur[mt], ar, eb_r[ks]);
                    mma3(ur[mt], an, eb_i[ks]);
                    mma3(ui[mt], ar, eb_i[ks]);
                    mma3(ui[mt], ai, eb_r[ks]);
                }
            }
            __syncthreads();
            // U into the same region as the W operand: [plane][b3][ib]; this lane holds
            // ib = 16 mt + 4 g + r, b3 = 16 w + c16
#pragma unroll
            for (int mt = 0; mt < NT; ++mt) {
                uint2 rh, rl, ih, il;
                split4(ur[mt], rh, rl);
                split4(ui[mt], ih, il);
                __bf16 *o = ut + (size_t)(m0 + c16) * RS + 16 * mt + 4 * g;
                *(uint2 *)o = rh;
                *(uint2 *)(o + (size_t)NB * RS) = rl;
                *(uint2 *)(o + (size_t)2 * NB * RS) = ih;
                *(uint2 *)(o + (size_t)3 * NB * RS) = il;
            }
        } else {
            for (int e = threadIdx.x; e < NB2 / 4; e += blockDim.x) {
                const int ib = e / (NB / 4), b3 = 4 * (e % (NB / 4)), a2 = ib / 3, j = ib % 3;
                const int s0 = u.a12_boff[a1 * NA + a2] - bo0, s1 = u.a12_boff[a1 * NA + a2 + 1] - bo0;
                T ur[4] = {0, 0, 0, 0}, ui[4] = {0, 0, 0, 0};
                for (int bl = s0; bl < s1; ++bl) {
                    const int a3 = pbl[bl].y;
                    const C2<T> *pp = pb + 27 * bl + 9 * i + 3 * j;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const C2<T> p = pp[k];
                        const T *er = E3r + (3 * a3 + k) * NB + b3, *ei = E3i + (3 * a3 + k) * NB + b3;
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            ur[v] += p.x * er[v] - p.y * ei[v];
                            ui[v] += p.x * ei[v] + p.y * er[v];
                        }
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if constexpr (PREC == 2) {
                        const __bf16 rh = (__bf16)ur[v], ih = (__bf16)ui[v];
                        __bf16 *o = ut + (size_t)(b3 + v) * RS + ib;
                        o[0] = rh;
                        o[(size_t)NB * RS] = (__bf16)(ur[v] - (float)rh);
                        o[(size_t)2 * NB * RS] = ih;
                        o[(size_t)3 * NB * RS] = (__bf16)(ui[v] - (float)ih);
                    } else {
                        us[ib * NB + b3 + v] = ur[v];
                        us[NB2 + ib * NB + b3 + v] = ui[v];
                    }
                }
            }
        }
        __syncthreads();
        const int ia = 3 * a1 + i;
        T *Wr = Wt + (size_t)ia * NB2, *Wi = Wr + (size_t)NB * NB2;
#pragma unroll 1
        for (int nt = 0; nt < NT; ++nt) {
            const int n0 = 16 * nt;
            Acc cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
            if constexpr (PREC == 2) {
                const __bf16 *pu = ut + (size_t)(n0 + c16) * RS + 8 * g;
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    Split br, bi;
                    br.h = ld8(pu + 32 * ks);
                    br.l = ld8(pu + (size_t)NB * RS + 32 * ks);
                    bi.h = ld8(pu + (size_t)2 * NB * RS + 32 * ks);
                    bi.l = ld8(pu + (size_t)3 * NB * RS + 32 * ks);
                    mma3(cr, ea_r[ks], br);
                    mma3(cr, ea_n[ks], bi);
                    mma3(ci, ea_r[ks], bi);
                    mma3(ci, ea_i[ks], br);
                }
            } else {
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const int k = 4 * ks + g;
                    const T br = us[k * NB + n0 + c16], bi = us[NB2 + k * NB + n0 + c16];
                    mma1(cr, fa_r[ks], br);
                    mma1(cr, -fa_i[ks], bi);
                    mma1(ci, fa_r[ks], bi);
                    mma1(ci, fa_i[ks], br);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int b2 = m0 + out_row<PREC>(lane, r);
                Wr[b2 * NB + n0 + c16] = cr[r];
                Wi[b2 * NB + n0 + c16] = ci[r];
            }
        }
        __syncthreads();
    }
}

__device__ inline double gaussd(double x, double mu, double sigma) {
    // lo_gauss (gottochblandat_calculus.f90:268-288)
    const double d = x - mu;
    if (fabs(d) > 8.0 * sigma) return 0.0;
    const double inv = 1.0 / sigma;
    return exp(-d * d * 0.5 * inv * inv) * 0.398942280401433 * inv;
}

// d0 = g(w1; w3 - w2) - g(w1; w2 - w3), d1 = g(w1; w2 + w3) - g(w1; -w2 - w3) (get_dirac,
// scattering_threephonon.f90:247-270) for sigma^2 = sig2, with g zero beyond 8 sigma, for the
// four elements (one b1, one b2, four b3) of a lane.  The four Gaussians are named by the
// distance they test: A = w1 + w2 - w3, B = w1 - w2 + w3, C = w1 - w2 - w3, D = w1 + w2 + w3.
// Bit k of the wave-uniform mask tm clear means term k lies beyond the 8 sigma cut for every
// element of the wave's tile (interval bound in k_psi), where lo_gauss returns exactly 0: the
// term is not evaluated, with bit-identical results.
__device__ inline void delta4(double w1, double w2, const double *w3, const double *sig2, unsigned tm, double *d0,
                              double *d1) {
    double sig[4], gA[4] = {0, 0, 0, 0}, gB[4] = {0, 0, 0, 0}, gC[4] = {0, 0, 0, 0}, gD[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < 4; ++r) sig[r] = sqrt(sig2[r]);
    if (tm & 1u) {
#pragma unroll
        for (int r = 0; r < 4; ++r) gA[r] = gaussd(w1, -w2 + w3[r], sig[r]);
    }
    if (tm & 2u) {
#pragma unroll
        for (int r = 0; r < 4; ++r) gB[r] = gaussd(w1, w2 - w3[r], sig[r]);
    }
    if (tm & 4u) {
#pragma unroll
        for (int r = 0; r < 4; ++r) gC[r] = gaussd(w1, w2 + w3[r], sig[r]);
    }
    if (tm & 8u) {
#pragma unroll
        for (int r = 0; r < 4; ++r) gD[r] = gaussd(w1, -w2 - w3[r], sig[r]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        d0[r] = gA[r] - gB[r];
        d1[r] = gC[r] - gD[r];
    }
}
__device__ inline void delta4(float x1, float x2, const float *x3, const float *sig2, unsigned tm, float *d0,
                              float *d1) {
    float inv[4], ce[4], lim[4], dd[4], ss[4];
    float gA[4] = {0, 0, 0, 0}, gB[4] = {0, 0, 0, 0}, gC[4] = {0, 0, 0, 0}, gD[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        inv[r] = __builtin_amdgcn_rsqf(sig2[r]);
        ce[r] = -0.72134752044448170f * inv[r] * inv[r];   // -log2(e) / (2 sigma^2)
        lim[r] = 64.0f * sig2[r];                           // (8 sigma)^2
        dd[r] = x2 - x3[r];
        ss[r] = x2 + x3[r];
    }
    auto g = [&](float d, int r) {
        const float d2 = d * d;
        return d2 > lim[r] ? 0.0f : __builtin_amdgcn_exp2f(ce[r] * d2);
    };
    if (tm & 1u) {
#pragma unroll
        for (int r = 0; r < 4; ++r) gA[r] = g(x1 + dd[r], r);
    }
    if (tm & 2u) {
#pragma unroll
        for (int r = 0; r < 4; ++r) gB[r] = g(x1 - dd[r], r);
    }
    if (tm & 4u) {
#pragma unroll
        for (int r = 0; r < 4; ++r) gC[r] = g(x1 - ss[r], r);
    }
    if (tm & 8u) {
#pragma unroll
        for (int r = 0; r < 4; ++r) gD[r] = g(x1 + ss[r], r);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float nrm = 0.398942280401433f * inv[r];
        d0[r] = nrm * (gA[r] - gB[r]);
        d1[r] = nrm * (gC[r] - gD[r]);
    }
}

template <int NB> struct MaskWords { static constexpr int n = (NB + 31) / 32; };

// The value v holds in lane l of the wave (every lane active).
__device__ inline float lane_value(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ inline double lane_value(double v, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <int PREC, int NB, int NTY, int BB = 1>
size_t psi_lds() {
    typedef typename Stage<PREC>::T S;
    constexpr int NW = NB / 16, NMW = MaskWords<NB>::n;
    size_t s = sizeof(S) * (2 * NTY * NB + 2 * NB + BB * NW * kChunk * 16 * NTY) + sizeof(float) * NB;
    s = (s + 15) & ~(size_t)15;
    s += (PREC == 2) ? (size_t)BB * (NB / 16) * 4 * 64 * 16 : 0;   // E1 fragments (stacked K)
    s += sizeof(double) * (NB + BB * 16 + BB * 16 * NTY) + sizeof(float) * NB + sizeof(int) * BB * 16 +
         sizeof(unsigned) * (size_t)(NB + BB * 16) * NMW;
    s = (s + 15) & ~(size_t)15;
    return s + sizeof(S) * 3 * (BB * 16 + 2 * NB);    // reduced velocities (integration type 6)
}

// DUMP: every tile is formed (no dead-tile skip) and |Psi|^2 with its prefactor is stored to
// a.psi_out[t][b1][b2][b3] (diagnostic export; the production instantiations have DUMP off).
// BB: blocks of 16 b1 per workgroup (NW waves each, one per 16 b3): with BB = NB/16 one
// workgroup takes a whole triplet and W is fetched once for all its b1.
template <int PREC, int NB, int NTY, bool DUMP = false, int BB = 1>
__global__ void __launch_bounds__(4 * NB * BB) __attribute__((amdgpu_waves_per_eu(PREC == 2 ? 3 : 1))) k_psi(MfmaEpiArgs a, const void *E1re_, const void *E1im_,
                                                const __bf16 *__restrict__ Et1, const void *W_) {
    typedef typename Traits<PREC>::T T;
    typedef typename Traits<PREC>::Acc Acc;
    typedef typename Stage<PREC>::T S;
    typedef S XT;                          // scattering-matrix entries: FP64 in FP64 mode, else FP32
    XT *const xi = (XT *)a.xi;
    constexpr int NW = NB / 16, NB2 = NB * NB, KS = KSteps<PREC, NB>::n, KP = kpad(NB), NMW = MaskWords<NB>::n;
    extern __shared__ __align__(16) unsigned char smem[];
    const int nT = a.nT;
    // XCD-aware order: the workgroups (b1 block groups) of one triplet share blockIdx % 8.
    // The batch's triplets are taken last first: the W that k_uw wrote last is the part still
    // in the last-level cache when this launch starts.
    constexpr int NG = NW / BB;
    static_assert(NW % BB == 0, "b1 blocks per workgroup divide the b1 blocks");
    const int gid = blockIdx.x;
    const int tl = a.nt - 1 - ((gid / (8 * NG)) * 8 + (gid % 8));
    const int lane = threadIdx.x & 63, wg = threadIdx.x >> 6, g = lane >> 4, c16 = lane & 15;
    // this wave: b1 block bl of the workgroup (block bb of the triplet), b3 block w
    const int bl = wg / NW, w = wg % NW;
    const int bb0 = ((gid / 8) % NG) * BB, bb = bb0 + bl;
    if (tl < 0) return;
    const int t = a.t0 + tl;
    const int q1 = a.tq1[t], q2 = a.tq2[t], q3 = a.tq3[t];
    const int i2 = a.irr_of_full[q2], i3 = a.irr_of_full[q3];
    const int e0 = a.eq_off[t], e1 = a.eq_off[t + 1], neq = e1 - e0;
    // LDS
    S *n2h = (S *)smem;                    // [NTY][NB]     n2 + 1/2
    S *n3h = n2h + NTY * NB;               // [NTY][NB]     n3 + 1/2
    S *sg2s = n3h + NTY * NB;              // [NB]          sigma^2 of q2 modes
    S *sg3s = sg2s + NB;                   // [NB]          sigma^2 of q3 modes
    S *Xp = sg3s + NB;                     // [BB NW][kChunk][16][NTY] per-wave partial sums over b3
    float *w3f = (float *)(Xp + BB * NW * kChunk * 16 * NTY);   // [NB] omega of q3 modes
    unsigned char *p8 = smem + align16(sizeof(S) * (size_t)(2 * NTY * NB + 2 * NB + BB * NW * kChunk * 16 * NTY) +
                                       sizeof(float) * NB);
    v8bf *e1sa = (v8bf *)p8;               // [BB][NB/16 K-steps][4 operands][64 lanes] (BF16X3)
    if constexpr (PREC == 2) p8 += (size_t)BB * (NB / 16) * 4 * 64 * 16;
    v8bf *e1s = e1sa + (size_t)bl * (NB / 16) * 4 * 64;   // this wave's b1 block
    double *w2s = (double *)p8;            // [NB]
    double *w1s = w2s + NB;                // [BB 16]
    double *gam = w1s + BB * 16;           // [BB][NTY][16] linewidth partial sums
    float *w2f = (float *)(gam + BB * 16 * NTY);   // [NB]
    int *rows = (int *)(w2f + NB);         // [BB 16]
    unsigned *v23 = (unsigned *)(rows + BB * 16);   // [NB][NMW]  |w2 - w3| >= tol, by b3
    unsigned *v12 = v23 + NB * NMW;                 // [BB 16][NMW]  b2 allowed for b1, by b2
    // [BB 16 b1, NB b2, NB b3][3] reduced group velocities (integration type 6)
    S *u6 = (S *)(smem + align16((size_t)((unsigned char *)(v12 + BB * 16 * NMW) - smem)));
    constexpr int M1 = BB * 16;            // b1 of the workgroup
    const bool itype6 = a.integrationtype == 6;
    if (itype6)
        for (int x = threadIdx.x; x < 3 * (M1 + 2 * NB); x += blockDim.x) {
            const int m = x / 3, cc = x % 3;
            const double v = (m < M1)        ? a.u_irr[((size_t)q1 * NB + 16 * bb0 + m) * 3 + cc]
                             : (m < M1 + NB) ? a.u_full[((size_t)q2 * NB + m - M1) * 3 + cc]
                                             : a.u_full[((size_t)q3 * NB + m - M1 - NB) * 3 + cc];
            u6[x] = (S)v;
        }
    const double *om2 = a.om_full + (size_t)q2 * NB, *om3 = a.om_full + (size_t)q3 * NB;
    for (int x = threadIdx.x; x < nT * NB; x += blockDim.x) {
        const int T_ = x / NB, b = x % NB;
        n2h[T_ * NB + b] = (S)(a.be[((size_t)T_ * a.nirr + i2) * NB + b] + 0.5);
        n3h[T_ * NB + b] = (S)(a.be[((size_t)T_ * a.nirr + i3) * NB + b] + 0.5);
    }
    for (int b = threadIdx.x; b < NB; b += blockDim.x) {
        w2s[b] = om2[b];
        w2f[b] = (float)om2[b];
        sg2s[b] = (S)a.sigsq[(size_t)i2 * NB + b];
        sg3s[b] = (S)a.sigsq[(size_t)i3 * NB + b];
        w3f[b] = (float)om3[b];
    }
    if (threadIdx.x < M1) {
        const int b1 = 16 * bb0 + threadIdx.x;
        w1s[threadIdx.x] = a.om_irr[(size_t)q1 * NB + b1];
        rows[threadIdx.x] = a.row_of[(size_t)q1 * NB + b1];
    }
    for (int x = threadIdx.x; x < M1 * NTY; x += blockDim.x) gam[x] = 0.0;
    // pair masks (scattering_threephonon.f90:90-91, 100-101)
    for (int x = threadIdx.x; x < NB * NMW; x += blockDim.x) {
        const int b2 = x / NMW, wd = x % NMW;
        const double w2 = om2[b2];
        unsigned m = 0;
        for (int k = 0; k < 32 && 32 * wd + k < NB; ++k)
            if (!(fabs(w2 - om3[32 * wd + k]) < kTol)) m |= 1u << k;
        v23[x] = m;
    }
    for (int x = threadIdx.x; x < M1 * NMW; x += blockDim.x) {
        const int m1 = x / NMW, wd = x % NMW;
        const int b1 = 16 * bb0 + m1;
        const double w1 = a.om_irr[(size_t)q1 * NB + b1];
        const bool r_ok = a.row_of[(size_t)q1 * NB + b1] >= 0;
        unsigned m = 0;
        for (int k = 0; k < 32 && 32 * wd + k < NB; ++k) {
            const double w2 = om2[32 * wd + k];
            if (r_ok && !(w2 < kTol) && !(fabs(w1 - w2) < kTol)) m |= 1u << k;
        }
        v12[x] = m;
    }
    if constexpr (PREC == 2) {
        // E1 operands with real and imaginary parts stacked in K: K-step s covers the 16
        // components ia = 16 s .. 16 s + 15, lane (g, c) supplies k = 8 g + j with ia = 16 s + 4 g
        // + (j & 3), the real part of W for j < 4 and the imaginary part for j >= 4 (A), so that
        // B_re = [E1r; -E1i] gives Re(Psi) and B_im = [E1i; E1r] gives Im(Psi) with no padding.
        // Slots per K-step: B_re hi, B_re lo, B_im hi, B_im lo (planes of the split table: re hi,
        // re lo, im hi, im lo, -im hi, -im lo).
        const size_t ps = (size_t)NB * KP;
        const __bf16 *pe = Et1 + (size_t)q1 * kPlanes * ps + (size_t)(16 * bb + c16) * KP + 4 * g;
        if (w == 0) {   // the first wave of each b1 block
            auto cat = [&](int pa, int pb, int s) {
                const uint2 lo = *(const uint2 *)(pe + pa * ps + 16 * s), hi = *(const uint2 *)(pe + pb * ps + 16 * s);
                return __builtin_bit_cast(v8bf, make_uint4(lo.x, lo.y, hi.x, hi.y));
            };
#pragma unroll
            for (int s = 0; s < NB / 16; ++s) {
                e1s[(4 * s + 0) * 64 + lane] = cat(0, 4, s);
                e1s[(4 * s + 1) * 64 + lane] = cat(1, 5, s);
                e1s[(4 * s + 2) * 64 + lane] = cat(2, 0, s);
                e1s[(4 * s + 3) * 64 + lane] = cat(3, 1, s);
            }
        }
    }
    __syncthreads();
    // per-lane constants: b1 = 16 bb + c16 and four values of b3
    const int m1l = 16 * bl + c16;         // the lane's b1 in the workgroup
    const double w1 = w1s[m1l];
    const float x1 = (float)w1;
    const int row1 = rows[m1l];
    const S sg1 = (S)a.sigsq[(size_t)q1 * NB + 16 * bb + c16];
    unsigned v12w[NMW];
#pragma unroll
    for (int k = 0; k < NMW; ++k) v12w[k] = v12[m1l * NMW + k];
    const int b3w = 16 * w;                         // first b3 of this wave
    const int sh3 = (b3w & 31);                     // bit of b3w in its mask word
    unsigned ok13 = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int b3 = b3w + out_row<PREC>(lane, r);
        const double w3 = om3[b3];
        if (row1 >= 0 && !(w3 < kTol) && !(fabs(w1 - w3) < kTol)) ok13 |= 1u << (sh3 + out_row<PREC>(lane, r));
    }
    // loop-invariant per-lane values of b3 are read from LDS in the loop (register budget)
    const int tcol = (c16 < nT) ? c16 : 0;
    // W operand (A[m = (b2, b3)][k = ia]), prefetched one tile ahead by buffer loads
    const T *Wt = (const T *)W_ + (size_t)tl * 2 * NB * NB2;
    const __amdgpu_buffer_rsrc_t wr = rsrc(Wt, sizeof(T) * NB * NB2);
    const __amdgpu_buffer_rsrc_t wi = rsrc(Wt + (size_t)NB * NB2, sizeof(T) * NB * NB2);
    // BF16X3: per K-step of 16 components the lane's 4 (re and im) values of W; else one value
    constexpr int KW = (PREC == 2) ? NB / 16 : KS, RW = (PREC == 2) ? 4 : 1;
    T rr[KW][RW], ri[KW][RW];
    const T *E1r = (const T *)E1re_ + (size_t)q1 * NB2 + 16 * bb + c16;
    const T *E1i = (const T *)E1im_ + (size_t)q1 * NB2 + 16 * bb + c16;
    auto load_w = [&](int b2) {
        const int n = b2 * NB + b3w + c16;
#pragma unroll
        for (int ks = 0; ks < KW; ++ks) {
            if constexpr (PREC == 2) {
                const int vo = (int)sizeof(T) * ((16 * ks + 4 * g) * NB2 + n);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    rr[ks][j] = bload(wr, vo, (int)sizeof(T) * j * NB2);
                    ri[ks][j] = bload(wi, vo, (int)sizeof(T) * j * NB2);
                }
            } else {
                const int vo = (int)sizeof(T) * (g * NB2 + n);
                rr[ks][0] = bload_t<T>(wr, vo, (int)sizeof(T) * 4 * ks * NB2);
                ri[ks][0] = bload_t<T>(wi, vo, (int)sizeof(T) * 4 * ks * NB2);
            }
        }
    };
    S y[NTY][4], yM[4];
#pragma unroll
    for (int T_ = 0; T_ < NTY; ++T_)
#pragma unroll
        for (int r = 0; r < 4; ++r) y[T_][r] = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) yM[r] = 0;
    // this lane's b1: prefactor, permutation factor and Monte-Carlo draw count
    const S sc = (S)(a.psi_scale * a.perm[t] * (a.tw ? (double)a.tw[(size_t)t * NB + 16 * bb + c16] : 1.0));
    const S sfix2 = (S)(a.sigma_fixed * a.sigma_fixed);
    const bool itype2 = a.integrationtype == 2;
    // Tiles (16 b1 x 16 b3 at one b2) whose every element lies beyond the Gaussians' 8 sigma
    // cut in all four terms contribute exactly zero (lo_gauss returns 0 there): a wave skips
    // them.  Interval bounds over the tile's w1, w3 and largest sigma, with a margin far above
    // the rounding of the element-wise test, decide per b2; the decision is wave-uniform.
    constexpr int NLW = (NB + 63) / 64;
    unsigned long long live[NLW];
    // per Gaussian term (A, B, C, D of delta4) the b2 whose tile it reaches; live = their union
    unsigned long long lterm[4][NLW];
    if constexpr (DUMP) {
#pragma unroll
        for (int wd = 0; wd < NLW; ++wd) {
            live[wd] = (NB - 64 * wd >= 64) ? ~0ull : ((1ull << (NB - 64 * wd)) - 1);
#pragma unroll
            for (int k = 0; k < 4; ++k) lterm[k][wd] = live[wd];
        }
    } else {
        double w1lo = 1e300, w1hi = -1e300, s1m = 0, w3lo = 1e300, w3hi = -1e300, s3m = 0, u1m = 0, u3m = 0;
        for (int m = 0; m < 16; ++m) {
            w1lo = fmin(w1lo, w1s[16 * bl + m]);
            w1hi = fmax(w1hi, w1s[16 * bl + m]);
            s1m = fmax(s1m, a.sigsq[(size_t)q1 * NB + 16 * bb + m]);
            w3lo = fmin(w3lo, om3[b3w + m]);
            w3hi = fmax(w3hi, om3[b3w + m]);
            s3m = fmax(s3m, a.sigsq[(size_t)i3 * NB + b3w + m]);
            if (itype6)
                for (int cc = 0; cc < 3; ++cc) {
                    u1m = fmax(u1m, fabs(a.u_irr[((size_t)q1 * NB + 16 * bb + m) * 3 + cc]));
                    u3m = fmax(u3m, fabs(a.u_full[((size_t)q3 * NB + b3w + m) * 3 + cc]));
                }
        }
#pragma unroll
        for (int wd = 0; wd < NLW; ++wd) {
            const int b2 = 64 * wd + lane;
            bool lt[4] = {false, false, false, false};
            if (b2 < NB) {
                const double w2 = w2s[b2];
                // type 6: each pair width is at most sqrt(1/2) (|u_a|_inf + |u_b|_inf)
                double u2m = 0;
                if (itype6)
                    for (int cc = 0; cc < 3; ++cc) u2m = fmax(u2m, fabs(a.u_full[((size_t)q2 * NB + b2) * 3 + cc]));
                const double smax = (a.integrationtype == 2) ? sqrt(s1m + a.sigsq[(size_t)i2 * NB + b2] + s3m)
                                    : itype6 ? 0.7071067811865476 * (2.0 / 3.0) * (u1m + u2m + u3m) + a.thresh6
                                             : a.sigma_fixed;
                const double L = 8.0 * smax * (1.0 + 1e-4) + 1e-6 * (fabs(w1hi) + fabs(w2) + fabs(w3hi));
                auto hit = [&](double lo, double hi) { return lo <= L && hi >= -L; };
                lt[0] = hit(w1lo + w2 - w3hi, w1hi + w2 - w3lo);   // A = w1 + w2 - w3
                lt[1] = hit(w1lo - w2 + w3lo, w1hi - w2 + w3hi);   // B = w1 - w2 + w3
                lt[2] = hit(w1lo - w2 - w3hi, w1hi - w2 - w3lo);   // C = w1 - w2 - w3
                lt[3] = hit(w1lo + w2 + w3lo, w1hi + w2 + w3hi);   // D = w1 + w2 + w3
            }
            live[wd] = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const unsigned long long bl = __ballot(lt[k]);
                lterm[k][wd] = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(bl >> 32)) << 32) |
                               (unsigned)__builtin_amdgcn_readfirstlane((unsigned)bl);
                live[wd] |= lterm[k][wd];
            }
        }
    }
    auto next_live = [&](int from) -> int {
#pragma unroll
        for (int wd = 0; wd < NLW; ++wd) {
            if (from >= 64 * (wd + 1)) continue;
            const int sh = from > 64 * wd ? from - 64 * wd : 0;
            const unsigned long long m = live[wd] >> sh;
            if (m) return 64 * wd + sh + __builtin_ctzll(m);
        }
        return NB;
    };
    // per-lane values of this lane's four b3, the same for every b2: frequencies and sigma^2
    // (4 consecutive values outside FP64) and the occupations n3 + 1/2 of the lane's
    // temperature column
    S w3v[4], s3v[4], bmv[4];
    if constexpr (PREC == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int b3 = b3w + out_row<PREC>(lane, r);
            w3v[r] = om3[b3];
            s3v[r] = sg3s[b3];
        }
    } else {
        const int b30 = b3w + out_row<PREC>(lane, 0);
        const float4 wv = *(const float4 *)(w3f + b30), sv = *(const float4 *)(sg3s + b30);
        w3v[0] = wv.x; w3v[1] = wv.y; w3v[2] = wv.z; w3v[3] = wv.w;
        s3v[0] = sv.x; s3v[1] = sv.y; s3v[2] = sv.z; s3v[3] = sv.w;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) bmv[r] = (c16 < nT) ? n3h[tcol * NB + b3w + out_row<PREC>(lane, r)] : (S)0;
    int nxt = next_live(0);
    if (nxt < NB) load_w(nxt);
    for (int cb = 0; cb < NB; cb += kChunk) {
    // this wave's partial sums of the chunk start at zero; live b2 overwrite their slots
    for (int x = lane; x < kChunk * 16 * NTY; x += 64) Xp[wg * kChunk * 16 * NTY + x] = (S)0;
    unsigned cm = (unsigned)((live[cb >> 6] >> (cb & 63)) & ((1ull << kChunk) - 1));
    while (cm) {
        const int cs = __builtin_ctz(cm);
        cm &= cm - 1;
        const int b2 = cb + cs;
        Acc cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
        if constexpr (PREC == 2) {
            Split a8[KW];
#pragma unroll
            for (int ks = 0; ks < KW; ++ks) {
                const float x[8] = {rr[ks][0], rr[ks][1], rr[ks][2], rr[ks][3],
                                    ri[ks][0], ri[ks][1], ri[ks][2], ri[ks][3]};
                split8(x, a8[ks]);
            }
#pragma unroll
            for (int ks = 0; ks < KW; ++ks) {
                Split bre, bim;
                bre.h = e1s[(4 * ks + 0) * 64 + lane];
                bre.l = e1s[(4 * ks + 1) * 64 + lane];
                bim.h = e1s[(4 * ks + 2) * 64 + lane];
                bim.l = e1s[(4 * ks + 3) * 64 + lane];
                mma3(cr, a8[ks], bre);
                mma3(ci, a8[ks], bim);
            }
            // next live tile's W is in flight during the epilogue
            nxt = next_live(b2 + 1);
            if (nxt < NB) load_w(nxt);
        } else {
            T ar[KS], ai[KS];
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                ar[ks] = rr[ks][0];
                ai[ks] = ri[ks][0];
            }
            nxt = next_live(b2 + 1);
            if (nxt < NB) load_w(nxt);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int k = 4 * ks + g;
                const T br = E1r[k * NB], bi = E1i[k * NB];
                mma1(cr, ar[ks], br);
                mma1(cr, ai[ks], -bi);
                mma1(ci, ar[ks], bi);
                mma1(ci, ai[ks], br);
            }
        }
        // smeared energy conservation for the 4 elements (b1, b2, b3) of this lane; the pair
        // masks select the results (a masked element's value is computed and discarded)
        const unsigned ok2 = (v12w[b2 >> 5] >> (b2 & 31)) & 1u;
        const unsigned okm = ok2 ? (ok13 & v23[b2 * NMW + (b3w >> 5)]) : 0u;
        const S sg12 = sg1 + sg2s[b2];
        const S x2 = (PREC == 0) ? (S)w2s[b2] : (S)w2f[b2];
        // Gaussian terms that reach this tile (wave-uniform)
        const int wb = b2 >> 6, sb = b2 & 63;
        const unsigned tm = (unsigned)((lterm[0][wb] >> sb) & 1ull) | ((unsigned)((lterm[1][wb] >> sb) & 1ull) << 1) |
                            ((unsigned)((lterm[2][wb] >> sb) & 1ull) << 2) | ((unsigned)((lterm[3][wb] >> sb) & 1ull) << 3);
        S P[4], M[4], psq[4], sgv[4], d0[4], d1[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const S re = (S)cr[r], im = (S)ci[r];
            psq[r] = sc * (re * re + im * im);
            if constexpr (DUMP)
                ((S *)a.psi_out)[(((size_t)t * NB + 16 * bb + c16) * NB + b2) * NB + b3w + out_row<PREC>(lane, r)] = psq[r];
            sgv[r] = itype2 ? sg12 + s3v[r] : sfix2;
            if (itype6) {
                const S s6 = sigma_type6<S>(u6 + 3 * m1l, u6 + 3 * (M1 + b2), u6 + 3 * (M1 + NB + b3w + out_row<PREC>(lane, r)),
                                            (S)a.thresh6);
                sgv[r] = s6 * s6;
            }
        }
        if constexpr (PREC == 0)
            delta4(w1, x2, w3v, sgv, tm, d0, d1);
        else
            delta4(x1, x2, w3v, sgv, tm, d0, d1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const S A = psq[r] * d0[r], B = psq[r] * d1[r];
            const bool on = (okm >> (sh3 + out_row<PREC>(lane, r))) & 1u;
            P[r] = on ? A + B : (S)0;
            M[r] = on ? B - A : (S)0;
        }
        // sum over this wave's b3 of f_T, temperature T in output column c16
        const S bp = (c16 < nT) ? n2h[c16 * NB + b2] : (S)0;
        Acc xa = {0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            mma1(xa, M[r], bmv[r]);
            mma1(xa, P[r], bp);
        }
        if (c16 < nT) {
#pragma unroll
            for (int r = 0; r < 4; ++r) Xp[((wg * kChunk + cs) * 16 + out_row<PREC>(lane, r)) * NTY + c16] = (S)xa[r];
        }
        // sums over b2, kept per lane; n2 + 1/2 of temperature T at this b2 is bp of lane T
#pragma unroll
        for (int T_ = 0; T_ < NTY; ++T_) {
            if (T_ < nT) {
                const S nh = lane_value(bp, T_);
#pragma unroll
                for (int r = 0; r < 4; ++r) y[T_][r] += P[r] * nh;
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) yM[r] += M[r];
    }
        {
            // sums over b3 of the chunk's b2: linewidth partials and the q2' entries of the
            // scattering matrix (:126-133).  Every target (row, q', band) belongs to one
            // triplet and, inside it, to one equivalent pair, so values are stored.
            __syncthreads();
            for (int x = threadIdx.x; x < M1 * nT; x += blockDim.x) {
                const int mb = x % M1, T_ = x / M1, m = mb % 16, xb = mb / 16;
                S v[kChunk];
                double gs = 0;
#pragma unroll
                for (int c = 0; c < kChunk; ++c) {
                    S s = 0;
#pragma unroll
                    for (int ww = 0; ww < NW; ++ww) s += Xp[(((xb * NW + ww) * kChunk + c) * 16 + m) * NTY + T_];
                    v[c] = s;
                    gs += (double)s;
                }
                gam[(xb * NTY + T_) * 16 + m] += gs;
                const int r = rows[mb];
                if (r < 0) continue;
                if (a.comp) {
                    S *cp = (S *)a.comp + (size_t)T_ * a.comp_T + ((size_t)t * NB + 16 * bb0 + mb) * 2 * NB + cb;
#pragma unroll
                    for (int c = 0; c < kChunk; ++c) cp[c] = v[c];
                    continue;
                }
                const double iw1 = 2.0 / w1s[mb];
                for (int e = e0; e < e1; ++e) {
                    XT *xr = xi + (size_t)T_ * a.xi_T + (size_t)r * a.xi_row + (size_t)a.eq_q2[e] * NB + cb;
#pragma unroll
                    for (int c = 0; c < kChunk; ++c) xr[c] = (XT)((double)v[c] * w2s[cb + c] * iw1);
                }
            }
            __syncthreads();
        }
    }
    // linewidth partials (g0 += f0 + f1 for every equivalent triplet, :126-128)
    for (int x = threadIdx.x; x < nT * M1; x += blockDim.x) {
        const int T_ = x / M1, mb = x % M1, m = mb % 16, xb = mb / 16;
        a.gpart[((size_t)t * a.gT + a.T0 + T_) * NB + 16 * bb0 + mb] =
            (rows[mb] >= 0) ? gam[(xb * NTY + T_) * 16 + m] * (double)neq : 0.0;
    }
    // q3' entries (:134) from this lane's registers; where q3' = q2' the entry already
    // holds the q2' term written above by this workgroup, and the q3' term is added.
    if (row1 >= 0 && a.comp) {
        S *cp = (S *)a.comp + ((size_t)t * NB + 16 * bb + c16) * 2 * NB + NB + b3w;
#pragma unroll
        for (int T_ = 0; T_ < NTY; ++T_) {
            if (T_ >= nT) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int b3 = b3w + out_row<PREC>(lane, r);
                cp[(size_t)T_ * a.comp_T + out_row<PREC>(lane, r)] = y[T_][r] + n3h[T_ * NB + b3] * yM[r];
            }
        }
    } else if (row1 >= 0) {
        const double iw1 = 2.0 / w1;
        for (int e = e0; e < e1; ++e) {
            const int x2 = a.eq_q2[e], x3q = a.eq_q3[e];
#pragma unroll
            for (int T_ = 0; T_ < NTY; ++T_) {
                if (T_ >= nT) continue;
                XT *xr = xi + (size_t)T_ * a.xi_T + (size_t)row1 * a.xi_row + (size_t)x3q * NB;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int b3 = b3w + out_row<PREC>(lane, r);
                    const S s3 = y[T_][r] + n3h[T_ * NB + b3] * yM[r];
                    const double v = (double)s3 * om3[b3] * iw1;
                    if (x2 == x3q) xr[b3] = (XT)((double)xr[b3] + v);
                    else xr[b3] = (XT)v;
                }
            }
        }
    }
}

template <int PREC, int NB, int NTY, bool DUMP = false, int BB = 1>
int run_chain(const MfmaPlan &p, hipStream_t st) {
    constexpr int NW = NB / 16;
    UwArgs u;
    u.maxblk = std::max(p.maxblk, 1);
    u.tq2 = p.tq2; u.tq3 = p.tq3; u.qcart = p.qcart; u.lv2 = p.lv2; u.lv3 = p.lv3; u.mwm = p.mwm;
    u.a1_off = p.a1off; u.blist = p.blist; u.a1_boff = p.a1_boff; u.a12_boff = p.a12_boff;
    u.Ef_re = p.Efre; u.Ef_im = p.Efim; u.Et = (const __bf16 *)p.Etf; u.W = p.work;
    const size_t sh_uw = uw_lds<PREC, NB>(u.maxblk, p.maxa1);
    const size_t sh_psi = psi_lds<PREC, NB, NTY, BB>();
    hipError_t e = hipFuncSetAttribute((const void *)k_uw<PREC, NB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)sh_uw);
    if (e != hipSuccess) return device_error(e, "k_uw LDS size", __FILE__, __LINE__);
    e = hipFuncSetAttribute((const void *)k_psi<PREC, NB, NTY, DUMP, BB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sh_psi);
    if (e != hipSuccess) return device_error(e, "k_psi LDS size", __FILE__, __LINE__);
    MfmaEpiArgs ea = p.epi;
    for (long t0 = 0; t0 < p.ntrip; t0 += p.batch) {
        const int nt = (int)std::min<long>(p.batch, p.ntrip - t0);
        hipEvent_t *ev = p.ev ? p.ev + 3 * (t0 / p.batch) : nullptr;
        if (ev) (void)hipEventRecord(ev[0], st);
        u.t0 = (int)t0;
        hipLaunchKernelGGL((k_uw<PREC, NB>), dim3(nt * (NB / 3)), dim3(4 * NB), sh_uw, st, u);
        if (ev) (void)hipEventRecord(ev[1], st);
        ea.t0 = (int)t0;
        ea.nt = nt;
        const int ngrp = (nt + 7) / 8;
        hipLaunchKernelGGL((k_psi<PREC, NB, NTY, DUMP, BB>), dim3(ngrp * 8 * (NW / BB)), dim3(64 * NW * BB), sh_psi, st,
                           ea, p.E1re, p.E1im,
                           (const __bf16 *)p.Eti, (const void *)p.work);
        if (ev) (void)hipEventRecord(ev[2], st);
        e = hipGetLastError();
        if (e != hipSuccess) return device_error(e, "MFMA contraction launch", __FILE__, __LINE__);
    }
    return 0;
}

template <int PREC>
int prepare(const MfmaPlan &p, hipStream_t st) {
    typedef typename Traits<PREC>::T T;
    const int nb = p.nb;
    const size_t nb2 = (size_t)nb * nb;
    hipLaunchKernelGGL(k_scale_planar<T>, dim3(1024), dim3(256), 0, st, (long)p.nirr * nb2, nb, p.egv_irr, p.om_irr,
                       (T *)p.E1re, (T *)p.E1im);
    hipLaunchKernelGGL(k_scale_planar<T>, dim3(4096), dim3(256), 0, st, (long)p.nfull * nb2, nb, p.egv_full,
                       p.om_full, (T *)p.Efre, (T *)p.Efim);
    if (PREC == 2) {
        hipLaunchKernelGGL(k_split_t, dim3(1024), dim3(256), 0, st, (long)p.nirr, nb, kpad(nb), p.egv_irr, p.om_irr,
                           (__bf16 *)p.Eti);
        hipLaunchKernelGGL(k_split_t, dim3(4096), dim3(256), 0, st, (long)p.nfull, nb, kpad(nb), p.egv_full,
                           p.om_full, (__bf16 *)p.Etf);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : device_error(e, "operand preparation", __FILE__, __LINE__);
}

template <int PREC, int NB>
int dispatch_nty(const MfmaPlan &p, hipStream_t st) {
    // one workgroup per triplet in k_psi: split-BF16x3 at 48 bands
    if constexpr (PREC == 2 && NB == 48) {
        if (p.psi_bb == NB / 16) {
            switch (p.nty) {
            case 4: return run_chain<PREC, NB, 4, false, NB / 16>(p, st);
            case 8: return run_chain<PREC, NB, 8, false, NB / 16>(p, st);
            default: return run_chain<PREC, NB, 16, false, NB / 16>(p, st);
            }
        }
    }
    switch (p.nty) {
    case 4: return run_chain<PREC, NB, 4>(p, st);
    case 8: return run_chain<PREC, NB, 8>(p, st);
    default: return run_chain<PREC, NB, 16>(p, st);
    }
}

template <int PREC>
int dispatch_nb(const MfmaPlan &p, hipStream_t st) {
    int r = prepare<PREC>(p, st);
    if (r) return r;
    switch (p.nb) {
    case 48: return dispatch_nty<PREC, 48>(p, st);
    case 96: return dispatch_nty<PREC, 96>(p, st);
    default: return 5;
    }
}

template <int PREC>
int dispatch_psi2(const MfmaPlan &p, hipStream_t st) {
    int r = prepare<PREC>(p, st);
    if (r) return r;
    switch (p.nb) {
    case 48: return run_chain<PREC, 48, 4, true>(p, st);
    case 96: return run_chain<PREC, 96, 4, true>(p, st);
    default: return 5;
    }
}

// LDS bytes of the two kernels for band count nb and temperature block nty (k_psi with
// bb blocks of 16 b1 per workgroup).
size_t lds_need(int nb, int nty, int precision, int maxblk, int maxa1, int bb = 1) {
    const size_t es = (precision == TC3_PREC_FP64) ? 8 : 4;
    const size_t kp = kpad(nb), nw = nb / 16, nmw = (nb + 31) / 32;
    size_t uw = (precision == TC3_PREC_BF16X3) ? 2 * 4 * nb * (kp + 8) : es * 2 * nb * nb;
    uw += 2 * es * 27 * (size_t)maxblk + 16 * (size_t)maxblk;
    uw = ((uw + 15) & ~(size_t)15) + 16 * (size_t)std::max(maxa1, 1);
    size_t ps = es * (2 * nty * nb + 2 * nb + bb * nw * kChunk * 16 * nty) + 4 * nb;
    ps = ((ps + 15) & ~(size_t)15);
    if (precision == TC3_PREC_BF16X3) ps += (size_t)bb * (nb / 16) * 4 * 64 * 16;
    ps += 8 * (nb + 16 * bb + 16 * bb * nty) + 4 * nb + 4 * 16 * bb + 4 * (nb + 16 * bb) * nmw;
    ps = ((ps + 15) & ~(size_t)15) + es * 3 * (16 * bb + 2 * (size_t)nb);
    return std::max(uw, ps);
}

}  // namespace

size_t mfma_elem_size(int precision) { return precision == TC3_PREC_FP64 ? sizeof(double) : sizeof(float); }

size_t mfma_split_elems(int nb) { return (size_t)kPlanes * nb * kpad(nb); }

bool mfma_supported(int nb, int precision) {
    if (nb != 48 && nb != 96) return false;   // nb = 3 na and a multiple of 16
    // IFC blocks and entries per atom are bounded separately by the caller
    return lds_need(nb, 4, precision, 0, 0) + 8 * 1024 <= 160 * 1024;
}

int mfma_temperature_block(int nb, int nT, int precision) {
    for (int nty : {4, 8, 16})
        if (nty >= nT && lds_need(nb, nty, precision, 0, 0) + 8 * 1024 <= 160 * 1024) return nty;
    for (int nty : {16, 8, 4})
        if (lds_need(nb, nty, precision, 0, 0) + 8 * 1024 <= 160 * 1024) return nty;
    return 0;
}

int run_mfma_psi2(const MfmaPlan &p, int precision, hipStream_t st) {
    if (p.nty != 4 || p.epi.psi_out == nullptr) return 5;
    if (lds_need(p.nb, 4, precision, p.maxblk, p.maxa1) > 160 * 1024) return 5;
    switch (precision) {
    case TC3_PREC_FP64: return dispatch_psi2<0>(p, st);
    case TC3_PREC_FP32: return dispatch_psi2<1>(p, st);
    default: return dispatch_psi2<2>(p, st);
    }
}

int run_mfma_chain(const MfmaPlan &p, int precision, hipStream_t st) {
    if (lds_need(p.nb, p.nty, precision, p.maxblk, p.maxa1, p.psi_bb) > 160 * 1024) return 5;
    switch (precision) {
    case TC3_PREC_FP64: return dispatch_nb<0>(p, st);
    case TC3_PREC_FP32: return dispatch_nb<1>(p, st);
    default: return dispatch_nb<2>(p, st);
    }
}

}  // namespace tc3
