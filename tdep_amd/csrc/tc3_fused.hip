// Fused three-phonon contraction for 48 bands (16-atom cells) in split-BF16x3: one persistent
// kernel, fed by a device work queue of triplets, that forms Phi(q2,q3), |Psi|^2, the smeared
// energy conservation and the compact per-(triplet, b1) sums without any HBM intermediate.
//
// Psi[b1][b2][b3] = sum_{ia,ib,ic} E1[ia][b1] E2[ib][b2] E3[ic][b3] Phi[ia][ib][ic]
// (scattering_threephonon.f90:88-113, E = e / sqrt(omega), no conjugation) is contracted in the
// order
//   V_b1[ib][ic] = sum_ia E1[ia][b1] Phi[ia][ib][ic]       vector ALUs, over the nonzero IFC
//                                                           blocks only (pretransform_phi3 :275-317)
//   X_b1[ib][b3] = sum_ic V_b1[ib][ic] E3[ic][b3]          MFMA
//   Psi_b1[b2][b3] = sum_ib E2[ib][b2] X_b1[ib][b3]        MFMA
// with one wave per b1, so that both sums the scattering matrix needs (over b3 for the q2'
// entries, over b2 for the q3' entries, :126-135) are complete inside the wave.
//
// Complex products on real MFMAs: the K dimension carries real and imaginary parts together.
// For a K-step of 16 complex indices k0..k0+15, lane (g, c) supplies k = 8g + j (j < 8) of
// v_mfma_f32_16x16x32_bf16, and k = 8g + j stands for index k0 + 4g + (j & 3), real part for
// j < 4 and imaginary part for j >= 4.  With A = [Vr | Vi] and B_re = [E3r | -E3i],
// B_im = [E3i | E3r] one K-step gives Re and Im of a complex product with no padding (48 = 3
// steps of 16), and the accumulator of X (lane (g, c): rows 4g + r, column c) is, value for
// value, the B operand of the next product in the same order, so X never leaves the registers.
// Every E operand (E3 as B, E2 as A: A[m = c][k] and B[k][n = c] share the lane layout) comes
// from one table per q-point (k_ftab).  Split-BF16x3: every FP32 operand is hi + lo bf16 and
// a.b ~ ah.bh + ah.bl + al.bh, FP32 accumulation (as tc3_mfma.hip).
//
// Epilogue per element (get_dirac :219-271, rates :116-123): with P = psisq (d0 + d1) and
// M = psisq (d1 - d0), f_T = P (n2_T + 1/2) + M (n3_T + 1/2).  The q3' sums (over b2, the
// accumulator rows) go through v_mfma_f32_16x16x4f32 with the occupations as the second
// operand (temperature = output column); the q2' sums (over b3, the accumulator columns) are
// summed over the 16 lanes of a row with DPP.  Outputs are the compact sums of the two-kernel
// path (comp[T][t][b1][0: b2 | 1: b3], the linewidth partials gpart), assembled into rows by
// k_xi_asm.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "tc3_ctx.hpp"
#include "tc3_fused.hpp"

namespace tc3 {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

constexpr int NB = 48;          // bands
constexpr int NTL = NB / 16;    // 16-wide tiles of a band index
constexpr int NWV = 4;          // waves per workgroup (one b1 each)
constexpr int FRAG = 64 * 8;    // bf16 of one operand fragment (64 lanes x 8)
constexpr int FQ = 2 * NTL * NTL * 2;   // fragments per q-point: [part][K-step][tile][hi/lo]
constexpr int VF = NTL * NTL * 2;       // fragments of one V_b1: [row tile][K-step][hi/lo]
constexpr double kTol = 1e-9;
constexpr double kTwoPiD = 6.283185307179586;

struct Split { v8bf h, l; };

__device__ inline void mma3(v4f &c, const Split &a, const Split &b) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.h, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.l, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.l, b.h, c, 0, 0, 0);
}
__device__ inline void mma4(v4f &c, float a, float b) { c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// x ~ hi + lo with hi = bf16(x) (round to nearest even) and lo = bf16(x - hi)
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

__device__ inline Split frag(const __bf16 *base, int fi, int lane) {
    Split s;
    s.h = *(const v8bf *)(base + (size_t)fi * FRAG + lane * 8);
    s.l = *(const v8bf *)(base + (size_t)(fi + 1) * FRAG + lane * 8);
    return s;
}

// fragment index of an E table: part (0 re, 1 im), K-step s, tile, hi
__device__ inline int ffi(int part, int s, int tile) { return ((part * NTL + s) * NTL + tile) * 2; }

// sum over the 16 lanes of a DPP row (lanes 16 k .. 16 k + 15)
__device__ inline float rowsum16(float x) {
    x += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0xB1, 0xF, 0xF, false));
    x += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x4E, 0xF, 0xF, false));
    x += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x141, 0xF, 0xF, false));
    x += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x140, 0xF, 0xF, false));
    return x;
}

// d0 = g(w1; w3 - w2) - g(w1; w2 - w3), d1 = g(w1; w2 + w3) - g(w1; -w2 - w3), g zero beyond
// 8 sigma (lo_gauss, gottochblandat_calculus.f90:268-288)
__device__ inline void delta_pair(float x1, float x2, float x3, float sig2, float &d0, float &d1) {
    const float inv = __builtin_amdgcn_rsqf(sig2);
    const float ce = -0.72134752044448170f * inv * inv;  // -log2(e) / (2 sigma^2)
    const float lim = 64.0f * sig2;
    const float dd = x2 - x3, ss = x2 + x3;
    auto g = [&](float d) {
        const float d2 = d * d;
        return d2 > lim ? 0.0f : __builtin_amdgcn_exp2f(ce * d2);
    };
    const float nrm = 0.398942280401433f * inv;
    d0 = nrm * (g(x1 + dd) - g(x1 - dd));
    d1 = nrm * (g(x1 - ss) - g(x1 + ss));
}

// E table of every q-point: for part re, K-step s, tile, lane (g, c) and j < 8, the value
// j < 4 ? Er[16 s + 4 g + j][16 tile + c] : -Ei[16 s + 4 g + j - 4][16 tile + c]  (part re)
// j < 4 ? Ei[16 s + 4 g + j][16 tile + c] :  Er[16 s + 4 g + j - 4][16 tile + c]  (part im)
// with E[component][mode] = e / sqrt(omega) (0 below the frequency tolerance), split into bf16
// hi and lo.
__global__ void k_ftab(int nq, const cd *__restrict__ egv, const double *__restrict__ om, __bf16 *__restrict__ ft) {
    const long n = (long)nq * 2 * NTL * NTL * 64 * 8;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) {
        const int j = (int)(x & 7), lane = (int)((x >> 3) & 63);
        long y = x >> 9;
        const int tile = (int)(y % NTL);
        y /= NTL;
        const int s = (int)(y % NTL);
        y /= NTL;
        const int part = (int)(y & 1);
        const long q = y >> 1;
        const int g = lane >> 4, c = lane & 15;
        const int comp = 16 * s + 4 * g + (j & 3), mode = 16 * tile + c;
        const double w = om[q * NB + mode];
        const double sc = (w < kTol) ? 0.0 : 1.0 / sqrt(w);
        const cd e = egv[(q * NB + comp) * NB + mode];
        const float er = (float)(e.x * sc), ei = (float)(e.y * sc);
        const float v = part == 0 ? (j < 4 ? er : -ei) : (j < 4 ? ei : er);
        const __bf16 h = (__bf16)v;
        const __bf16 l = (__bf16)(v - (float)h);
        const size_t fi = (size_t)q * FQ + (size_t)((part * NTL + s) * NTL + tile) * 2;
        ft[fi * FRAG + lane * 8 + j] = h;
        ft[(fi + 1) * FRAG + lane * 8 + j] = l;
    }
}

// e / sqrt(omega) of the irreducible points, planar FP32 [q][component][mode]
__global__ void k_e1(long n, const cd *__restrict__ egv, const double *__restrict__ om, float *__restrict__ re,
                     float *__restrict__ im) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long q = i / (NB * NB);
        const int b = (int)(i % NB);
        const double w = om[q * NB + b];
        const double s = (w < kTol) ? 0.0 : 1.0 / sqrt(w);
        re[i] = (float)(egv[i].x * s);
        im[i] = (float)(egv[i].y * s);
    }
}

__global__ void k_zero_queue(int *q) { *q = 0; }

constexpr int MAXE = 16;   // IFC entries per block the Phi sums prefetch at once

template <int NTY>
constexpr size_t fused_lds_fixed() {
    return sizeof(__bf16) * ((size_t)NWV * VF * FRAG + 2 * (size_t)FQ * FRAG) + sizeof(float) * 2 * NWV * NB +
           sizeof(float) * (2 * NTY * NB + 4 * NB) + sizeof(unsigned) * NB * 2 +
           (sizeof(double) + 3 * sizeof(float) + 4 * sizeof(unsigned)) * NB + 16;
}
size_t fused_lds(int nty, int npair, int npblk) {
    const size_t f = nty <= 4 ? fused_lds_fixed<4>() : fused_lds_fixed<8>();
    return f + sizeof(int4) * (size_t)npair + sizeof(int) * (size_t)npblk;
}

// One persistent workgroup per CU, four waves.  Per triplet taken from the queue: the phases
// of the IFC entries and the Phi blocks (summed once for all 27 (i, j, k) into the
// workgroup's global slot, ordered so that the values one V entry needs are contiguous), the
// E tables of q2 and q3 and the per-mode data into LDS, then twelve groups of four b1: V of
// the four b1 (all lanes, over the nonzero blocks of every (a2, a3) pair) into the waves' LDS
// regions, and per wave the two MFMA contractions and the epilogue of its b1.  With one wave
// per SIMD every global read is issued in batches before its values are used.
template <int NTY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) k_fused(FusedArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    __bf16 *Vreg = (__bf16 *)smem;                       // [wave][VF fragments]
    __bf16 *F2 = Vreg + (size_t)NWV * VF * FRAG;          // E2 fragments (A operand)
    __bf16 *F3 = F2 + (size_t)FQ * FRAG;                  // E3 fragments (B operand)
    float2 *E1c = (float2 *)(F3 + (size_t)FQ * FRAG);     // [ia][wave] E1 column of each wave's b1
    float *n2h = (float *)(E1c + NWV * NB);               // [T][b2]  n2 + 1/2
    float *n3h = n2h + NTY * NB;                          // [T][b3]  n3 + 1/2
    float *sg2 = n3h + NTY * NB, *sg3 = sg2 + NB, *w2f = sg3 + NB, *w3f = w2f + NB;
    unsigned *v23 = (unsigned *)(w3f + NB);               // [b2][2]  |w2 - w3| >= tol, by b3
    double *w1d = (double *)(v23 + 2 * NB);               // [b1] omega_1
    float *sg1L = (float *)(w1d + NB), *scL = sg1L + NB;  // [b1] sigma^2, psi prefactor
    int *rowL = (int *)(scL + NB);                        // [b1] row or -1
    unsigned *m12L = (unsigned *)(rowL + NB);             // [b1][2] b2 allowed for b1
    unsigned *m13L = m12L + 2 * NB;                       // [b1][2] b3 allowed for b1
    int *qslot = (int *)(m13L + 2 * NB);
    int4 *pairsL = (int4 *)(qslot + 4);                   // (a2, a3, slot base, blocks)
    int *pa1L = (int *)(pairsL + a.npair);                // a1 of every block, pair by pair
    double2 *ph = (double2 *)smem;                        // phases, in the V regions before V
    int4 *blocksL = (int4 *)(ph + a.nfc3);                // (first entry, end entry, slot position, -)
    float2 *slot = a.phi_slot + (size_t)blockIdx.x * a.nblk * 27;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, c = lane & 15;
    const int nT = a.nT;
    for (int x = tid; x < a.npair; x += 256) pairsL[x] = a.pairs[x];
    for (int x = tid; x < a.npblk; x += 256) pa1L[x] = a.pa1[x];
    for (;;) {
        if (tid == 0) *qslot = atomicAdd(a.queue, 1);
        __syncthreads();
        const int t = *qslot;
        if (t >= a.ntrip) break;
        const int q1 = a.tq1[t], q2 = a.tq2[t], q3 = a.tq3[t];
        const int i2 = a.irr_of_full[q2], i3 = a.irr_of_full[q3];
        // phases exp(-2 pi i (q2.R2 + q3.R3)) of every IFC entry (FP64) and the block table
        {
            const double *Q2 = a.qcart + 3 * (size_t)q2, *Q3 = a.qcart + 3 * (size_t)q3;
            for (int f = tid; f < a.nfc3; f += 256) {
                const double arg =
                    -kTwoPiD * (Q2[0] * a.lv2[3 * f] + Q2[1] * a.lv2[3 * f + 1] + Q2[2] * a.lv2[3 * f + 2] +
                                Q3[0] * a.lv3[3 * f] + Q3[1] * a.lv3[3 * f + 1] + Q3[2] * a.lv3[3 * f + 2]);
                double sn, cs;
                sincos(arg, &sn, &cs);
                ph[f] = make_double2(cs, sn);
            }
            for (int x = tid; x < a.nblk; x += 256) blocksL[x] = a.blocks[x];
        }
        __syncthreads();
        // Phi blocks, every (i, j, k): four blocks' entries in flight per thread (indices clamped
        // into the block, values masked: no branches around the loads)
        for (int x0 = tid; x0 < a.nblk * 27; x0 += 4 * 256) {
            double m[4][MAXE];
            int4 bl[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int x = min(x0 + 256 * u, a.nblk * 27 - 1);
                bl[u] = blocksL[x / 27];
                const int ijk = x % 27;
#pragma unroll
                for (int e = 0; e < MAXE; ++e)
                    m[u][e] = a.mwm[27 * (size_t)min(bl[u].x + e, bl[u].y - 1) + ijk];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int x = x0 + 256 * u;
                double sx = 0, sy = 0;
#pragma unroll
                for (int e = 0; e < MAXE; ++e) {
                    const double2 p = ph[min(bl[u].x + e, bl[u].y - 1)];
                    const double mm = (bl[u].x + e < bl[u].y) ? m[u][e] : 0.0;
                    sx += mm * p.x;
                    sy += mm * p.y;
                }
                if (x < a.nblk * 27) slot[bl[u].z + x % 27] = make_float2((float)sx, (float)sy);
            }
        }
        // the slot's new values reach L2 before the barrier and no lane reads a line of the
        // previous triplet's values from its L1 after it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        // E tables of q2 and q3, per-mode data of the triplet
        {
            const uint4 *s2 = (const uint4 *)(a.ftab + (size_t)q2 * FQ * FRAG);
            const uint4 *s3 = (const uint4 *)(a.ftab + (size_t)q3 * FQ * FRAG);
            uint4 *d2 = (uint4 *)F2, *d3 = (uint4 *)F3;
            for (int x = tid; x < FQ * FRAG / 8; x += 256) {
                d2[x] = s2[x];
                d3[x] = s3[x];
            }
        }
        for (int x = tid; x < NTY * NB; x += 256) {
            const int T_ = x / NB, b = x % NB;
            n2h[T_ * NB + b] = T_ < nT ? (float)(a.be[((size_t)T_ * a.nirr + i2) * NB + b] + 0.5) : 0.0f;
            n3h[T_ * NB + b] = T_ < nT ? (float)(a.be[((size_t)T_ * a.nirr + i3) * NB + b] + 0.5) : 0.0f;
        }
        for (int b = tid; b < NB; b += 256) {
            w2f[b] = (float)a.om_full[(size_t)q2 * NB + b];
            w3f[b] = (float)a.om_full[(size_t)q3 * NB + b];
            sg2[b] = (float)a.sigsq[(size_t)i2 * NB + b];
            sg3[b] = (float)a.sigsq[(size_t)i3 * NB + b];
            const int r = a.row_of[(size_t)q1 * NB + b];
            rowL[b] = r;
            w1d[b] = a.om_irr[(size_t)q1 * NB + b];
            sg1L[b] = (float)a.sigsq[(size_t)q1 * NB + b];
            scL[b] = (float)(a.psi_scale * a.perm[t] * (a.tw ? (double)a.tw[(size_t)t * NB + b] : 1.0));
        }
        // pair masks (scattering_threephonon.f90:90-91, 100-101): |w2 - w3| >= tol by (b2, b3),
        // and for every b1 the b2 and b3 it allows
        for (int x = tid; x < 2 * NB; x += 256) {
            const int b = x >> 1, wd = x & 1;
            const double w1 = a.om_irr[(size_t)q1 * NB + b], w2 = a.om_full[(size_t)q2 * NB + b];
            unsigned m = 0, m2 = 0, m3 = 0;
            for (int k = 0; k < 32 && 32 * wd + k < NB; ++k) {
                const double x2 = a.om_full[(size_t)q2 * NB + 32 * wd + k], x3 = a.om_full[(size_t)q3 * NB + 32 * wd + k];
                if (!(fabs(w2 - x3) < kTol)) m |= 1u << k;
                if (!(x2 < kTol) && !(fabs(w1 - x2) < kTol)) m2 |= 1u << k;
                if (!(x3 < kTol) && !(fabs(w1 - x3) < kTol)) m3 |= 1u << k;
            }
            v23[x] = m;
            m12L[x] = m2;
            m13L[x] = m3;
        }
        __syncthreads();
        // V regions start at zero: the positions of (ib, ic) without IFC blocks are never written
        for (int x = tid; x < NWV * VF * FRAG / 8; x += 256) ((uint4 *)Vreg)[x] = make_uint4(0, 0, 0, 0);
        const int neq = a.eq_off[t + 1] - a.eq_off[t];
        for (int grp = 0; grp < NB / NWV; ++grp) {
            // E1 columns of the group's four b1, [ia][b1 of the group]
            for (int x = tid; x < NWV * NB; x += 256) {
                const int ia = x / NWV, ww = x % NWV;
                const size_t o = ((size_t)q1 * NB + ia) * NB + NWV * grp + ww;
                E1c[x] = make_float2(a.e1re[o], a.e1im[o]);
            }
            __syncthreads();
            // V_b1[ib][ic] = sum_{a1, i} E1[3 a1 + i][b1] Phi[blk (a1, a2, a3)][i][j][k] for the
            // four b1, written in the A-fragment order of the X product.  The pairs come sorted
            // by their number of blocks, so the lanes of a wave loop over about as many blocks
            // as they have, four blocks' values in flight at a time.
            for (int it = tid; it < a.npair * 9; it += 256) {
                const int4 pr = pairsL[it / 9];
                const int jk = it % 9;
                const int nbw = __builtin_amdgcn_readfirstlane(pr.w);   // largest count of the wave
                const float2 *sp = slot + pr.z + jk;                     // value v = 3 bb + i at sp[9 v]
                const int pb0 = pr.z / 27;
                float vr[NWV], vi[NWV];
#pragma unroll
                for (int ww = 0; ww < NWV; ++ww) vr[ww] = vi[ww] = 0.0f;
                for (int bb0 = 0; bb0 < nbw; bb0 += 4) {
                    float2 f[12];
                    int a1v[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int cb = min(bb0 + k, pr.w - 1);
                        a1v[k] = pa1L[pb0 + cb];
#pragma unroll
                        for (int i = 0; i < 3; ++i) f[3 * k + i] = sp[9 * (3 * cb + i)];
                    }
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const float on = bb0 + k < pr.w ? 1.0f : 0.0f;
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            const float fx = on * f[3 * k + i].x, fy = on * f[3 * k + i].y;
                            const float4 *e4 = (const float4 *)(E1c + (3 * a1v[k] + i) * NWV);
                            const float4 e01 = e4[0], e23 = e4[1];
                            const float er[NWV] = {e01.x, e01.z, e23.x, e23.z}, ei[NWV] = {e01.y, e01.w, e23.y, e23.w};
#pragma unroll
                            for (int ww = 0; ww < NWV; ++ww) {
                                vr[ww] += er[ww] * fx - ei[ww] * fy;
                                vi[ww] += er[ww] * fy + ei[ww] * fx;
                            }
                        }
                    }
                }
                const int ib = 3 * pr.x + jk / 3, ic = 3 * pr.y + jk % 3;
                const int tt = ib >> 4, s = ic >> 4;
                const int pos = (16 * ((ic & 15) >> 2) + (ib & 15)) * 8 + (ic & 3);
                const int f0 = (tt * NTL + s) * 2;
#pragma unroll
                for (int ww = 0; ww < NWV; ++ww) {
                    __bf16 *o = Vreg + (size_t)ww * VF * FRAG + (size_t)f0 * FRAG + pos;
                    const __bf16 rh = (__bf16)vr[ww], ih = (__bf16)vi[ww];
                    o[0] = rh;
                    o[FRAG] = (__bf16)(vr[ww] - (float)rh);
                    o[4] = ih;
                    o[FRAG + 4] = (__bf16)(vi[ww] - (float)ih);
                }
            }
            __syncthreads();
            // ---- this wave's b1
            const int b1 = NWV * grp + w;
            if (rowL[b1] < 0) {
                if (lane < nT) a.gpart[((size_t)t * a.gT + a.T0 + lane) * NB + b1] = 0.0;
                continue;
            }
            const double w1 = w1d[b1];
            const float x1 = (float)w1;
            const float sg1 = sg1L[b1];
            const float sc = scL[b1];
            const float sfix2 = (float)(a.sigma_fixed * a.sigma_fixed);
            const bool itype2 = a.integrationtype == 2;
            const unsigned long long m12 = (unsigned long long)m12L[2 * b1] | ((unsigned long long)m12L[2 * b1 + 1] << 32);
            const unsigned long long m13 = (unsigned long long)m13L[2 * b1] | ((unsigned long long)m13L[2 * b1 + 1] << 32);
            // (b2 tile, b3 tile) pairs whose every element lies beyond the Gaussians' 8 sigma cut
            // in all four terms contribute exactly zero: interval bounds with a margin far above
            // rounding decide per tile, wave-uniformly
            unsigned live = 0;
            {
                bool lv = false;
                if (lane < NTL * NTL) {
                    const int bt = lane / NTL, ct = lane % NTL;
                    float l2 = 1e30f, h2 = -1e30f, l3 = 1e30f, h3 = -1e30f, s2m = 0, s3m = 0;
                    for (int m = 0; m < 16; ++m) {
                        l2 = fminf(l2, w2f[16 * bt + m]);
                        h2 = fmaxf(h2, w2f[16 * bt + m]);
                        l3 = fminf(l3, w3f[16 * ct + m]);
                        h3 = fmaxf(h3, w3f[16 * ct + m]);
                        s2m = fmaxf(s2m, sg2[16 * bt + m]);
                        s3m = fmaxf(s3m, sg3[16 * ct + m]);
                    }
                    const double smax = itype2 ? sqrt((double)sg1 + s2m + s3m) : a.sigma_fixed;
                    const double L = 8.0 * smax * (1.0 + 1e-4) +
                                     1e-6 * (fabs(w1) + fabs((double)h2) + fabs((double)h3) + fabs((double)l2) +
                                             fabs((double)l3));
                    auto hit = [&](double lo, double hi) { return lo <= L && hi >= -L; };
                    lv = hit(w1 + l2 - h3, w1 + h2 - l3) || hit(w1 - h2 + l3, w1 - l2 + h3) ||
                         hit(w1 - h2 - h3, w1 - l2 - l3) || hit(w1 + l2 + l3, w1 + h2 + h3);
                }
                live = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)__ballot(lv));
            }
            // X = V E3 and Psi = E2 X, ib tile by ib tile
            v4f Pr[NTL][NTL], Pi[NTL][NTL];
#pragma unroll
            for (int bt = 0; bt < NTL; ++bt)
#pragma unroll
                for (int ct = 0; ct < NTL; ++ct) Pr[bt][ct] = Pi[bt][ct] = (v4f){0, 0, 0, 0};
            const __bf16 *Vw = Vreg + (size_t)w * VF * FRAG;
            if (live) {
#pragma unroll
                for (int tt = 0; tt < NTL; ++tt) {
                    v4f Xr[NTL], Xi[NTL];
#pragma unroll
                    for (int ct = 0; ct < NTL; ++ct) Xr[ct] = Xi[ct] = (v4f){0, 0, 0, 0};
#pragma unroll
                    for (int s = 0; s < NTL; ++s) {
                        const Split A = frag(Vw, (tt * NTL + s) * 2, lane);
#pragma unroll
                        for (int ct = 0; ct < NTL; ++ct) {
                            const Split Br = frag(F3, ffi(0, s, ct), lane), Bi = frag(F3, ffi(1, s, ct), lane);
                            mma3(Xr[ct], A, Br);
                            mma3(Xi[ct], A, Bi);
                        }
                    }
                    Split Bx[NTL];
#pragma unroll
                    for (int ct = 0; ct < NTL; ++ct) {
                        const float x8[8] = {Xr[ct][0], Xr[ct][1], Xr[ct][2], Xr[ct][3],
                                             Xi[ct][0], Xi[ct][1], Xi[ct][2], Xi[ct][3]};
                        split8(x8, Bx[ct]);
                    }
#pragma unroll
                    for (int bt = 0; bt < NTL; ++bt) {
                        const Split Ar = frag(F2, ffi(0, tt, bt), lane), Ai = frag(F2, ffi(1, tt, bt), lane);
#pragma unroll
                        for (int ct = 0; ct < NTL; ++ct) {
                            if (!((live >> (bt * NTL + ct)) & 1u)) continue;
                            mma3(Pr[bt][ct], Ar, Bx[ct]);
                            mma3(Pi[bt][ct], Ai, Bx[ct]);
                        }
                    }
                }
            }
            // epilogue: lane (g, c) holds Psi[b2 = 16 bt + 4 g + r][b3 = 16 ct + c]; the lane's
            // three b3 columns' data is held in registers
            v4f DP[NTL], DM[NTL];
            float x3v[NTL], s3v[NTL], n3v[NTL][NTY];
            unsigned ok3v = 0;
#pragma unroll
            for (int ct = 0; ct < NTL; ++ct) {
                DP[ct] = DM[ct] = (v4f){0, 0, 0, 0};
                const int b3 = 16 * ct + c;
                x3v[ct] = w3f[b3];
                s3v[ct] = sg3[b3];
                ok3v |= (unsigned)((m13 >> b3) & 1ull) << ct;
#pragma unroll
                for (int T_ = 0; T_ < NTY; ++T_) n3v[ct][T_] = n3h[T_ * NB + b3];
            }
            const float one = (c < nT) ? 1.0f : 0.0f;
            const int cT = c < NTY ? c : 0;
            const size_t cbase = ((size_t)t * NB + b1) * 2 * NB;
#pragma unroll
            for (int bt = 0; bt < NTL; ++bt) {
                float SP[4], SM[NTY][4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    SP[r] = 0;
#pragma unroll
                    for (int T_ = 0; T_ < NTY; ++T_) SM[T_][r] = 0;
                    const int b2 = 16 * bt + 4 * g + r;
                    const float x2 = w2f[b2], s2 = sg1 + sg2[b2];
                    const float wn2 = one * n2h[cT * NB + b2];
                    const bool ok2 = (m12 >> b2) & 1ull;
                    const unsigned long long m23 = (unsigned long long)v23[2 * b2] | ((unsigned long long)v23[2 * b2 + 1] << 32);
#pragma unroll
                    for (int ct = 0; ct < NTL; ++ct) {
                        if (!((live >> (bt * NTL + ct)) & 1u)) continue;
                        const int b3 = 16 * ct + c;
                        const bool on = ok2 && ((ok3v >> ct) & 1u) && ((m23 >> b3) & 1ull);
                        const float re = Pr[bt][ct][r], im = Pi[bt][ct][r];
                        const float psisq = sc * (re * re + im * im);
                        const float sig2 = itype2 ? s2 + s3v[ct] : sfix2;
                        float d0, d1;
                        delta_pair(x1, x2, x3v[ct], sig2, d0, d1);
                        const float A_ = psisq * d0, B_ = psisq * d1;
                        const float P = on ? A_ + B_ : 0.0f, M = on ? B_ - A_ : 0.0f;
                        // q3' sums over b2 (rows): temperatures as output columns
                        mma4(DP[ct], P, wn2);
                        mma4(DM[ct], M, one);
                        SP[r] += P;
#pragma unroll
                        for (int T_ = 0; T_ < NTY; ++T_) SM[T_][r] += M * n3v[ct][T_];
                    }
                }
                // q2' sums over b3 (columns): the 16 lanes of the row, then one lane per
                // temperature writes (scattering_threephonon.f90:126-133)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    SP[r] = rowsum16(SP[r]);
#pragma unroll
                    for (int T_ = 0; T_ < NTY; ++T_) SM[T_][r] = rowsum16(SM[T_][r]);
                }
                if (c < nT) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        float sm = 0;
#pragma unroll
                        for (int T_ = 0; T_ < NTY; ++T_) sm = (T_ == c) ? SM[T_][r] : sm;
                        const int b2 = 16 * bt + 4 * g + r;
                        a.comp[(size_t)c * a.comp_T + cbase + b2] = n2h[c * NB + b2] * SP[r] + sm;
                    }
                }
            }
            // q3' entries (:134) and the linewidth partial of every temperature (:126-128)
            float gam = 0;
            if (c < nT) {
#pragma unroll
                for (int ct = 0; ct < NTL; ++ct)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int b3 = 16 * ct + 4 * g + r;
                        const float v = DP[ct][r] + n3h[c * NB + b3] * DM[ct][r];
                        a.comp[(size_t)c * a.comp_T + cbase + NB + b3] = v;
                        gam += v;
                    }
            }
            gam += __shfl_xor(gam, 16);
            gam += __shfl_xor(gam, 32);
            if (lane < nT) a.gpart[((size_t)t * a.gT + a.T0 + lane) * NB + b1] = (double)gam * (double)neq;
        }
        __syncthreads();
    }
}

}  // namespace

bool fused_supported(int nb, int precision, int integrationtype, int nfc3, int nblk, int npair, int maxe, int maxb) {
    // the phases of all IFC entries and the block table share the V regions; the tables of the
    // (a2, a3) pairs and the rest stay in LDS for the whole triplet
    return nb == NB && precision == TC3_PREC_BF16X3 && (integrationtype == 1 || integrationtype == 2) &&
           (size_t)nfc3 * 16 + (size_t)nblk * 16 <= sizeof(__bf16) * (size_t)NWV * VF * FRAG && maxe <= MAXE &&
           maxb > 0 && fused_lds(8, npair, nblk) <= 160 * 1024;
}

size_t fused_table_bytes(int nq) { return sizeof(__bf16) * (size_t)nq * FQ * FRAG; }

int fused_temperature_block(int nT) { return nT <= 4 ? 4 : 8; }

int run_fused(const FusedPlan &p, hipStream_t st) {
    // operand tables: E of every full-grid point as split fragments, E1 planar
    hipLaunchKernelGGL(k_ftab, dim3(4096), dim3(256), 0, st, p.nfull, p.egv_full, p.om_full, p.ftab);
    hipLaunchKernelGGL(k_e1, dim3(1024), dim3(256), 0, st, (long)p.nirr * NB * NB, p.egv_irr, p.om_irr, p.e1re,
                       p.e1im);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return device_error(e, "fused contraction tables", __FILE__, __LINE__);
    if (p.ev) (void)hipEventRecord(p.ev[1], st);
    FusedArgs a = p.args;
    const size_t sh = fused_lds(p.nty, a.npair, a.npblk);
    const void *kf = p.nty == 4 ? (const void *)k_fused<4> : (const void *)k_fused<8>;
    e = hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return device_error(e, "fused kernel LDS size", __FILE__, __LINE__);
    for (int T0 = 0; T0 < p.nT; T0 += p.nty) {
        a.T0 = T0;
        a.nT = std::min(p.nty, p.nT - T0);
        a.be = p.be + (size_t)T0 * p.nirr * NB;
        a.comp = (float *)p.comp + (size_t)T0 * a.comp_T;
        hipLaunchKernelGGL(k_zero_queue, dim3(1), dim3(1), 0, st, a.queue);
        if (p.nty == 4) hipLaunchKernelGGL(k_fused<4>, dim3(p.nwg), dim3(256), sh, st, a);
        else hipLaunchKernelGGL(k_fused<8>, dim3(p.nwg), dim3(256), sh, st, a);
        e = hipGetLastError();
        if (e != hipSuccess) return device_error(e, "fused contraction launch", __FILE__, __LINE__);
    }
    if (p.ev) (void)hipEventRecord(p.ev[2], st);
    return 0;
}

}  // namespace tc3
