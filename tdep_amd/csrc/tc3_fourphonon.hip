// Four-phonon scattering: compute_fourphonon_scattering (scattering_fourphonon.f90:1-328)
// with qg4ph = qgrid (the Monte-Carlo grids are the identity, weight 1/N_q).  For every
// irreducible q1 the quartets q2 <= q3 <= q4 = -q1-q2-q3 that quartet_is_irreducible
// (:330-456) keeps, with their images under the little group of q1; per quartet one
// workgroup forms pretransform_phi4 (:459-506) and Psi = sum phi4 e1 e2 e3 e4 (e / sqrt(w),
// unconjugated) in four contraction stages in LDS, FP64, and the four occupation and
// smeared-delta channels with the permutation multiplicities (:96-175) for every
// temperature, summed over the other three bands per leg.  The host adds the terms of every
// image to the row linewidths and to the off-diagonal scattering-matrix terms and averages
// those over the little group (:143-205).  Cells of one or two atoms (nb <= 6: phi4 of one
// quartet, 6^4 complex values, sits in LDS twice) use k_fourphonon; larger cells the batched
// kernels below (k_f4_phase, k_f4_phi, four k_f4_stage steps on FP64 MFMA, k_f4_epi), up to
// 96 bands.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "tc3_chain.hpp"
#include "tc3_ctx.hpp"
#include "tc3_fourphonon.hpp"

namespace tc3 {

namespace {

constexpr int kF4MaxNb = 6, kF4MaxN4 = 1296, kF4Block = 256, kF4Chunk = 512;

struct F4Args {
    int nb, na, nT, nirr;
    const int *fq;                      // [t][5] q1 (irreducible), q2, q3, q4 (full), multiplicity class
    const int *blk4;                    // [na^4 + 1]
    const double *lv, *mwm;             // [p][9], [p][81]
    const double *qcart;                // [full q][3]
    const cd *egv_irr, *egv_full;       // [q][component][mode]
    const double *om_irr, *om_full;
    const double *be, *sigsq;           // [T][nirr][nb], [nirr][nb]
    const int *irr_of_full, *row_of;
    const double *u_irr, *u_full;       // reduced group velocities (integration type 6)
    int itype;
    double sigma_fixed, thresh, pref;   // pref = pi/96 / N^2 (N points of one grid)
    const float *tw;                    // [t][nb] draw counts of Monte-Carlo rows, or null
    double *g, *s;                      // [t][T][nb], [t][T][3][nb][nb]
};

__device__ inline double pair_w6(const double *ua, const double *ub) {
    double m = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double d = ua[k] - ub[k];
        m = fmax(m, d * d);
    }
    return sqrt(m * 0.5) / 6.0;
}

__global__ void __launch_bounds__(kF4Block) k_fourphonon(F4Args a) {
    __shared__ cd P[kF4MaxN4], Q[kF4MaxN4], E[4][kF4MaxNb * kF4MaxNb];
    __shared__ double ps[kF4MaxN4];
    __shared__ cd ph[kF4Chunk];
    double *const fb = (double *)Q;   // per-temperature terms, once Psi is formed
    const int t = blockIdx.x, nb = a.nb, na = a.na, nb2 = nb * nb, nb3 = nb2 * nb, n4 = nb3 * nb;
    const int q[4] = {a.fq[5 * t], a.fq[5 * t + 1], a.fq[5 * t + 2], a.fq[5 * t + 3]};
    const int mc = a.fq[5 * t + 4];
    for (int x = threadIdx.x; x < 4 * nb2; x += blockDim.x) {
        const int k = x / nb2, r = x % nb2, comp = r / nb, mode = r % nb;
        const cd e = (k == 0) ? a.egv_irr[((size_t)q[0] * nb + comp) * nb + mode] : a.egv_full[((size_t)q[k] * nb + comp) * nb + mode];
        const double w = (k == 0) ? a.om_irr[(size_t)q[0] * nb + mode] : a.om_full[(size_t)q[k] * nb + mode];
        const double sc = (w >= kFreqTol) ? 1.0 / sqrt(w) : 0.0;
        E[k][r] = {e.x * sc, e.y * sc};
    }
    // pretransform_phi4: phases exp(-2 pi i (q2.R2 + q3.R3 + q4.R4)), q Cartesian without 2 pi,
    // once per IFC entry into an LDS table, in chunks of kF4Chunk entries
    const double *v2 = a.qcart + 3 * (size_t)q[1], *v3 = a.qcart + 3 * (size_t)q[2], *v4 = a.qcart + 3 * (size_t)q[3];
    const int nent = a.blk4[na * na * na * na];
    for (int x = threadIdx.x; x < n4; x += blockDim.x) P[x] = {0.0, 0.0};
    for (int c0 = 0; c0 < nent; c0 += kF4Chunk) {
        const int c1 = min(nent, c0 + kF4Chunk);
        __syncthreads();
        for (int p = c0 + threadIdx.x; p < c1; p += blockDim.x) {
            const double *l = a.lv + 9 * (size_t)p;
            const double iqr = -(v2[0] * l[0] + v2[1] * l[1] + v2[2] * l[2] + v3[0] * l[3] + v3[1] * l[4] +
                                 v3[2] * l[5] + v4[0] * l[6] + v4[1] * l[7] + v4[2] * l[8]) * kTwoPi;
            ph[p - c0] = {cos(iqr), sin(iqr)};
        }
        __syncthreads();
        for (int x = threadIdx.x; x < n4; x += blockDim.x) {
            const int ia = x / nb3, ib = (x / nb2) % nb, ic = (x / nb) % nb, id = x % nb;
            const int blk = (((ia / 3) * na + ib / 3) * na + ic / 3) * na + id / 3;
            const int pa = max(a.blk4[blk], c0), pb = min(a.blk4[blk + 1], c1);
            if (pa >= pb) continue;
            const int cmp = 27 * (ia % 3) + 9 * (ib % 3) + 3 * (ic % 3) + id % 3;
            double re = P[x].x, im = P[x].y;
            for (int p = pa; p < pb; ++p) {
                const double m = a.mwm[81 * (size_t)p + cmp];
                re += m * ph[p - c0].x;
                im += m * ph[p - c0].y;
            }
            P[x] = {re, im};
        }
    }
    __syncthreads();
    // Psi[b1][b2][b3][b4]: contract id, ic, ib, ia in turn (P -> Q -> P -> Q -> P)
    auto stage = [&](const cd *src, cd *dst, const cd *e, int lvl) {
        // lvl 3: dst[ia,ib,ic,b4] = sum_id src[ia,ib,ic,id] e[id][b4]
        // lvl 2: dst[ia,ib,b3,b4] = sum_ic src[ia,ib,ic,b4] e[ic][b3]
        // lvl 1: dst[ia,b2,b3,b4] = sum_ib src[ia,ib,b3,b4] e[ib][b2]
        // lvl 0: dst[b1,b2,b3,b4] = sum_ia src[ia,b2,b3,b4] e[ia][b1]
        int st = 1;
        for (int k = lvl; k < 3; ++k) st *= nb;
        for (int x = threadIdx.x; x < n4; x += blockDim.x) {
            const int bo = (x / st) % nb, base = x - bo * st;
            double re = 0.0, im = 0.0;
            for (int k = 0; k < nb; ++k) {
                const cd u = src[base + k * st], v = e[k * nb + bo];
                re += u.x * v.x - u.y * v.y;
                im += u.x * v.y + u.y * v.x;
            }
            dst[x] = {re, im};
        }
        __syncthreads();
    };
    stage(P, Q, E[3], 3);
    stage(Q, P, E[2], 2);
    stage(P, Q, E[1], 1);
    stage(Q, P, E[0], 0);
    for (int x = threadIdx.x; x < n4; x += blockDim.x) ps[x] = a.pref * (P[x].x * P[x].x + P[x].y * P[x].y);
    const double m0 = (mc == 0) ? 1.0 : (mc == 1 ? 3.0 : 6.0), m1 = (mc == 2) ? 2.0 : 1.0;
    const int i2 = a.irr_of_full[q[1]], i3 = a.irr_of_full[q[2]], i4 = a.irr_of_full[q[3]];
    __syncthreads();
    for (int T = 0; T < a.nT; ++T) {
        for (int x = threadIdx.x; x < n4; x += blockDim.x) {
            const int b1 = x / nb3, b2 = (x / nb2) % nb, b3 = (x / nb) % nb, b4 = x % nb;
            const double w1 = a.om_irr[(size_t)q[0] * nb + b1], w2 = a.om_full[(size_t)q[1] * nb + b2],
                         w3 = a.om_full[(size_t)q[2] * nb + b3], w4 = a.om_full[(size_t)q[3] * nb + b4];
            double f = 0.0;
            if (a.row_of[(size_t)q[0] * nb + b1] >= 0 && w2 >= kFreqTol && w3 >= kFreqTol && w4 >= kFreqTol) {
                const double n2 = a.be[((size_t)T * a.nirr + i2) * nb + b2], n3 = a.be[((size_t)T * a.nirr + i3) * nb + b3],
                             n4 = a.be[((size_t)T * a.nirr + i4) * nb + b4];
                const double n2p = n2 + 1.0, n3p = n3 + 1.0, n4p = n4 + 1.0;
                const double plf0 = n2p * n3p * n4p - n2 * n3 * n4;
                const double plf1 = 3.0 * n2 * n3p * n4p - n2p * n3 * n4;
                const double plf2 = 3.0 * n3 * n2p * n4p - n3p * n2 * n4;
                const double plf3 = 3.0 * n4 * n3p * n2p - n4p * n3 * n2;
                double sg;
                if (a.itype == 1) {
                    sg = a.sigma_fixed;
                } else if (a.itype == 2) {
                    sg = sqrt(a.sigsq[(size_t)q[0] * nb + b1] + a.sigsq[(size_t)i2 * nb + b2] +
                              a.sigsq[(size_t)i3 * nb + b3] + a.sigsq[(size_t)i4 * nb + b4]);
                } else {
                    const double *u1 = a.u_irr + ((size_t)q[0] * nb + b1) * 3, *u2 = a.u_full + ((size_t)q[1] * nb + b2) * 3,
                                 *u3 = a.u_full + ((size_t)q[2] * nb + b3) * 3, *u4 = a.u_full + ((size_t)q[3] * nb + b4) * 3;
                    sg = 0.0;
                    sg += pair_w6(u1, u2);
                    sg += pair_w6(u1, u3);
                    sg += pair_w6(u1, u4);
                    sg += pair_w6(u2, u3);
                    sg += pair_w6(u2, u4);
                    sg += pair_w6(u3, u4);
                    sg += a.thresh;
                }
                const double d0 = gauss(w1, w2 + w3 + w4, sg) - gauss(w1, -w2 - w3 - w4, sg);
                const double d1 = gauss(w1, -w2 + w3 + w4, sg) - gauss(w1, w2 - w3 - w4, sg);
                const double d2 = gauss(w1, w2 - w3 + w4, sg) - gauss(w1, -w2 + w3 - w4, sg);
                const double d3 = gauss(w1, w2 + w3 - w4, sg) - gauss(w1, -w2 - w3 + w4, sg);
                const double p = ps[x] * (a.tw ? (double)a.tw[(size_t)t * nb + b1] : 1.0);
                f = m0 * p * plf0 * d0 + m1 * p * plf1 * d1 + m1 * p * plf2 * d2 + m1 * p * plf3 * d3;
            }
            fb[x] = f;
        }
        __syncthreads();
        double *g = a.g + ((size_t)t * a.nT + T) * nb;
        double *s = a.s + ((size_t)t * a.nT + T) * 3 * nb2;
        for (int y = threadIdx.x; y < nb; y += blockDim.x) {
            double tot = 0.0;
            for (int z = 0; z < nb3; ++z) tot += fb[y * nb3 + z];
            g[y] = tot;
        }
        for (int y = threadIdx.x; y < nb2; y += blockDim.x) {
            const int b1 = y / nb, b = y % nb;
            double s2 = 0.0, s3 = 0.0, s4 = 0.0;
            for (int u = 0; u < nb; ++u)
                for (int v = 0; v < nb; ++v) {
                    s2 += fb[((b1 * nb + b) * nb + u) * nb + v];
                    s3 += fb[((b1 * nb + u) * nb + b) * nb + v];
                    s4 += fb[((b1 * nb + u) * nb + v) * nb + b];
                }
            s[y] = s2;
            s[nb2 + y] = s3;
            s[2 * nb2 + y] = s4;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// Cells of more than two atoms (nb > 6): quartets in batches, each step over the whole batch.
//   k_f4_phase   phases exp(-2 pi i (q2.R2 + q3.R3 + q4.R4)) of every IFC entry, per quartet
//   k_f4_phi     pretransform_phi4 (:459-506): phi4[ia][ib][ic][id], FP64 complex, dense (zero
//                outside the atom blocks that hold IFCs)
//   k_f4_stage   four contraction steps on FP64 MFMA, Y[r][b] = sum_k X[k][r] E[k][b] with X
//                read as [nb][nb^3]: each step contracts the leading index with its leg's
//                e / sqrt(w) and appends the band index last, so after ia, ib, ic, id the
//                batch holds Psi[b1][b2][b3][b4] (unconjugated, :88-113 / :140-175)
//   k_f4_epi     per (quartet, b1): the four occupation / smeared-delta channels with the
//                permutation multiplicities for a block of temperatures and the sums per leg,
//                in fixed order (the same outputs as k_fourphonon)
constexpr int kF4Tb = 4;          // temperatures per epilogue pass
constexpr int kF4EpiBlock = 256;

__global__ void k_f4_phase(int nent, const int *__restrict__ fq, int t0, const double *__restrict__ qcart,
                           const double *__restrict__ lv, cd *__restrict__ ph) {
    const int tl = blockIdx.y, t = t0 + tl;
    const double *v2 = qcart + 3 * (size_t)fq[5 * t + 1], *v3 = qcart + 3 * (size_t)fq[5 * t + 2],
                 *v4 = qcart + 3 * (size_t)fq[5 * t + 3];
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nent; p += gridDim.x * blockDim.x) {
        const double *l = lv + 9 * (size_t)p;
        const double iqr = -(v2[0] * l[0] + v2[1] * l[1] + v2[2] * l[2] + v3[0] * l[3] + v3[1] * l[4] + v3[2] * l[5] +
                             v4[0] * l[6] + v4[1] * l[7] + v4[2] * l[8]) *
                           kTwoPi;
        ph[(size_t)tl * nent + p] = {cos(iqr), sin(iqr)};
    }
}

// phi4 summed in FP64, stored in the contraction's type T (double, or float in the mixed modes)
template <typename T>
__global__ void k_f4_phi(int nb, int na, int nent, const int *__restrict__ blk4, const double *__restrict__ mwm,
                         const cd *__restrict__ ph, cplx<T> *__restrict__ X) {
    const int tl = blockIdx.y;
    const long nb2 = (long)nb * nb, nb3 = nb2 * nb, n4 = nb3 * nb;
    const cd *pt = ph + (size_t)tl * nent;
    cplx<T> *out = X + (size_t)tl * n4;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < n4; x += (long)gridDim.x * blockDim.x) {
        const int ia = (int)(x / nb3), ib = (int)((x / nb2) % nb), ic = (int)((x / nb) % nb), id = (int)(x % nb);
        const int blk = (((ia / 3) * na + ib / 3) * na + ic / 3) * na + id / 3;
        const int cmp = 27 * (ia % 3) + 9 * (ib % 3) + 3 * (ic % 3) + id % 3;
        double re = 0.0, im = 0.0;
        for (int p = blk4[blk]; p < blk4[blk + 1]; ++p) {
            const double m = mwm[81 * (size_t)p + cmp];
            re += m * pt[p].x;
            im += m * pt[p].y;
        }
        out[x] = {(T)re, (T)im};
    }
}

typedef double v4d_f4 __attribute__((ext_vector_type(4)));
typedef float v4f_f4 __attribute__((ext_vector_type(4)));
template <typename T> struct F4Acc { typedef v4d_f4 V; };
template <> struct F4Acc<float> { typedef v4f_f4 V; };
__device__ inline void f4_mma(v4d_f4 &c, double a, double b) { c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline void f4_mma(v4f_f4 &c, float a, float b) { c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
// output row of accumulator register i in a 16 x 16 tile: m = 4 i + g (FP64), 4 g + i (FP32)
template <typename T> __device__ inline int f4_row(int i, int g) { return sizeof(T) == 8 ? 4 * i + g : 4 * g + i; }

// Y[tl][r][b] = sum_k X[tl][k][r] E_q[k][b], r < nb^3, k, b < nb, complex, on
// v_mfma_f64_16x16x4f64 (T = double) or v_mfma_f32_16x16x4f32 (T = float): A[m = lane % 16]
// [k = k0 + lane / 16], B[k][n = lane % 16], D rows f4_row.  Workgroup: 4 waves of 16 rows r
// each, every column tile of b.  E_q: the leg's table, E + q nb^2 with q = fq[5 t + leg].
template <typename T, int NT>
__global__ void __launch_bounds__(256) k_f4_stage(int nb, const int *__restrict__ fq, int t0, int leg,
                                                  const cplx<T> *__restrict__ E, const cplx<T> *__restrict__ X,
                                                  cplx<T> *__restrict__ Y) {
    typedef typename F4Acc<T>::V V;
    const int tl = blockIdx.y;
    const long nb2 = (long)nb * nb, R = nb2 * nb;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, c16 = lane & 15;
    const long r0 = (long)blockIdx.x * 64 + 16 * w;
    if (r0 >= R) return;
    const cplx<T> *Xt = X + (size_t)tl * R * nb;
    const cplx<T> *Eq = E + (size_t)fq[5 * (t0 + tl) + leg] * nb2;
    V ar[NT], ai[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) ar[n] = ai[n] = (V){0, 0, 0, 0};
    const long rr = r0 + c16;
    for (int k0 = 0; k0 < nb; k0 += 4) {
        const int k = k0 + g;
        cplx<T> a = {0, 0};
        if (k < nb && rr < R) a = Xt[(size_t)k * R + rr];
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int col = 16 * n + c16;
            cplx<T> b = {0, 0};
            if (k < nb && col < nb) b = Eq[(size_t)k * nb + col];
            f4_mma(ar[n], a.x, b.x);
            f4_mma(ar[n], -a.y, b.y);
            f4_mma(ai[n], a.x, b.y);
            f4_mma(ai[n], a.y, b.x);
        }
    }
    cplx<T> *Yt = Y + (size_t)tl * R * nb;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int col = 16 * n + c16;
        if (col >= nb) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long row = r0 + f4_row<T>(i, g);
            if (row < R) Yt[(size_t)row * nb + col] = {ar[n][i], ai[n][i]};
        }
    }
}

struct F4EpiArgs {
    int nb, nT, nirr, t0, T0, nTb, tb, itype;   // nTb temperatures in this pass, tb <= kF4Tb LDS slots
    const int *fq;
    const double *om_irr, *om_full, *be, *sigsq;
    const int *irr_of_full, *row_of;
    const double *u_irr, *u_full;
    double sigma_fixed, thresh, pref;
    const float *tw;
    const void *psi;      // [tl][b1][b2][b3][b4], cplx<T> of the contraction
    double *g, *s;        // [t][T][nb], [t][T][3][nb][nb]
};

// fixed-order sum over a workgroup of 256 threads (wave sums, then the four waves in order)
__device__ inline double f4_block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    return ((red[0] + red[1]) + red[2]) + red[3];
}

// Workgroup per (quartet of the batch, b1), temperatures T0 .. T0 + nTb - 1 (nTb <= kF4Tb).
// Every thread keeps the (b3, b4) pairs p = tid + 256 j; their sums over b2 go to LDS
// ([T][b3][b4], each slot written by its own thread), the sums over (b3, b4) at every b2
// are workgroup sums.
template <typename CT>
__global__ void __launch_bounds__(kF4EpiBlock) k_f4_epi(F4EpiArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nb = a.nb, nb2 = nb * nb, tl = blockIdx.y, t = a.t0 + tl, b1 = blockIdx.x;
    const int tb = a.tb;
    double *acc = (double *)smem;                          // [tb][nb][nb]
    double *w234 = acc + (size_t)tb * nb2;                 // [3][nb] frequencies of legs 2, 3, 4
    double *nn = w234 + 3 * nb;                            // [tb][3][nb] occupations of legs 2, 3, 4
    double *sg = nn + tb * 3 * nb;                         // [3][nb] sigma^2 of legs 2, 3, 4
    double *uu = sg + 3 * nb;                              // [3][nb][3] reduced velocities (type 6)
    double *red = uu + 9 * nb;                             // [4]
    double *s2 = red + 4;                                  // [tb][nb]
    const int q1 = a.fq[5 * t], mc = a.fq[5 * t + 4];
    const int qf[3] = {a.fq[5 * t + 1], a.fq[5 * t + 2], a.fq[5 * t + 3]};
    const int qi[3] = {a.irr_of_full[qf[0]], a.irr_of_full[qf[1]], a.irr_of_full[qf[2]]};
    for (int x = threadIdx.x; x < 3 * nb; x += blockDim.x) {
        const int leg = x / nb, b = x % nb;
        w234[x] = a.om_full[(size_t)qf[leg] * nb + b];
        sg[x] = a.sigsq[(size_t)qi[leg] * nb + b];
        if (a.itype == 6)
            for (int k = 0; k < 3; ++k) uu[3 * x + k] = a.u_full[((size_t)qf[leg] * nb + b) * 3 + k];
        for (int T = 0; T < a.nTb; ++T)
            nn[(T * 3 + leg) * nb + b] = a.be[((size_t)(a.T0 + T) * a.nirr + qi[leg]) * nb + b];
    }
    for (int x = threadIdx.x; x < tb * nb2; x += blockDim.x) acc[x] = 0.0;
    __syncthreads();
    const double w1 = a.om_irr[(size_t)q1 * nb + b1];
    const bool row_ok = a.row_of[(size_t)q1 * nb + b1] >= 0;
    const double m0 = (mc == 0) ? 1.0 : (mc == 1 ? 3.0 : 6.0), m1 = (mc == 2) ? 2.0 : 1.0;
    const double tw = a.tw ? (double)a.tw[(size_t)t * nb + b1] : 1.0;
    const double sg1 = a.sigsq[(size_t)q1 * nb + b1];
    const double *u1 = (a.itype == 6) ? a.u_irr + ((size_t)q1 * nb + b1) * 3 : nullptr;
    const cplx<CT> *ps = (const cplx<CT> *)a.psi + ((size_t)tl * nb + b1) * (size_t)nb2 * nb;
    for (int b2 = 0; b2 < nb; ++b2) {
        double part[kF4Tb] = {0.0, 0.0, 0.0, 0.0};
        const double w2 = w234[b2];
        for (int p = threadIdx.x; p < nb2; p += blockDim.x) {
            const int b3 = p / nb, b4 = p % nb;
            const double w3 = w234[nb + b3], w4 = w234[2 * nb + b4];
            if (!(row_ok && w2 >= kFreqTol && w3 >= kFreqTol && w4 >= kFreqTol)) continue;
            const cplx<CT> v = ps[(size_t)b2 * nb2 + p];
            double sgm;
            if (a.itype == 1) {
                sgm = a.sigma_fixed;
            } else if (a.itype == 2) {
                sgm = sqrt(sg1 + sg[b2] + sg[nb + b3] + sg[2 * nb + b4]);
            } else {
                const double *u2 = uu + 3 * b2, *u3 = uu + 3 * (nb + b3), *u4 = uu + 3 * (2 * nb + b4);
                sgm = 0.0;
                sgm += pair_w6(u1, u2);
                sgm += pair_w6(u1, u3);
                sgm += pair_w6(u1, u4);
                sgm += pair_w6(u2, u3);
                sgm += pair_w6(u2, u4);
                sgm += pair_w6(u3, u4);
                sgm += a.thresh;
            }
            const double d0 = gauss(w1, w2 + w3 + w4, sgm) - gauss(w1, -w2 - w3 - w4, sgm);
            const double d1 = gauss(w1, -w2 + w3 + w4, sgm) - gauss(w1, w2 - w3 - w4, sgm);
            const double d2 = gauss(w1, w2 - w3 + w4, sgm) - gauss(w1, -w2 + w3 - w4, sgm);
            const double d3 = gauss(w1, w2 + w3 - w4, sgm) - gauss(w1, -w2 - w3 + w4, sgm);
            const double vx = (double)v.x, vy = (double)v.y;
            const double pq = a.pref * (vx * vx + vy * vy) * tw;
#pragma unroll
            for (int T = 0; T < kF4Tb; ++T) {
                if (T >= a.nTb) break;
                const double n2 = nn[(T * 3 + 0) * nb + b2], n3 = nn[(T * 3 + 1) * nb + b3], n4 = nn[(T * 3 + 2) * nb + b4];
                const double n2p = n2 + 1.0, n3p = n3 + 1.0, n4p = n4 + 1.0;
                const double plf0 = n2p * n3p * n4p - n2 * n3 * n4;
                const double plf1 = 3.0 * n2 * n3p * n4p - n2p * n3 * n4;
                const double plf2 = 3.0 * n3 * n2p * n4p - n3p * n2 * n4;
                const double plf3 = 3.0 * n4 * n3p * n2p - n4p * n3 * n2;
                const double f = m0 * pq * plf0 * d0 + m1 * pq * plf1 * d1 + m1 * pq * plf2 * d2 + m1 * pq * plf3 * d3;
                part[T] += f;
                acc[(size_t)T * nb2 + p] += f;
            }
        }
#pragma unroll
        for (int T = 0; T < kF4Tb; ++T) {
            if (T >= a.nTb) break;
            const double s = f4_block_sum(part[T], red);
            if (threadIdx.x == 0) s2[T * nb + b2] = s;
        }
    }
    __syncthreads();
    for (int T = 0; T < a.nTb; ++T) {
        double *g = a.g + ((size_t)t * a.nT + a.T0 + T) * nb;
        double *s = a.s + ((size_t)t * a.nT + a.T0 + T) * 3 * (size_t)nb2;
        if (threadIdx.x == 0) {
            double tot = 0.0;
            for (int b2 = 0; b2 < nb; ++b2) tot += s2[T * nb + b2];
            g[b1] = tot;
        }
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            double s3 = 0.0, s4 = 0.0;
            for (int u = 0; u < nb; ++u) {
                s3 += acc[(size_t)T * nb2 + b * nb + u];
                s4 += acc[(size_t)T * nb2 + u * nb + b];
            }
            s[(size_t)b1 * nb + b] = s2[T * nb + b];
            s[nb2 + (size_t)b1 * nb + b] = s3;
            s[2 * (size_t)nb2 + (size_t)b1 * nb + b] = s4;
        }
    }
}

size_t f4_epi_lds(int nb, int tb) {
    return sizeof(double) * ((size_t)tb * nb * nb + 3 * nb + (size_t)tb * 3 * nb + 3 * nb + 9 * nb + 4 + (size_t)tb * nb);
}

int fourth_grid_index(int i1, int i2, int i3, const int *dims) {
    const int n2 = dims[1], n3 = dims[2];
    int g[3] = {0, 0, 0};
    for (int i : {i1, i2, i3}) {
        g[0] += i / (n2 * n3);
        g[1] += (i / n3) % n2;
        g[2] += i % n3;
    }
    for (int l = 0; l < 3; ++l) g[l] = ((-g[l]) % dims[l] + dims[l]) % dims[l];
    return (g[0] * n2 + g[1]) * n3 + g[2];
}

// quartet_is_irreducible (scattering_fourphonon.f90:330-456): false for a reducible quartet,
// else its images (permutations removed) appended to img
bool quartet_images(const Mesh &m, int q1, int q2, int q3, int q4, std::vector<int> &img) {
    const int np = m.nfull;
    const std::vector<int> &lg = m.littlegroup[q1];
    std::vector<std::array<int, 3>> nq(lg.size(), {q2, q3, q4}), ns(lg.size(), {q2, q3, q4});
    for (size_t j = 0; j < lg.size(); ++j) {
        const int o = lg[j];
        const int a = m.rot[(size_t)o * np + q2], b = m.rot[(size_t)o * np + q3], cc = m.rot[(size_t)o * np + q4];
        if (a < 0 || b < 0 || cc < 0) continue;
        std::array<int, 3> s = {a, b, cc};
        std::sort(s.begin(), s.end());
        nq[j] = {a, b, cc};
        ns[j] = s;
        if (s[0] > q2 || s[1] > q3) return false;
    }
    for (size_t j = 0; j < lg.size(); ++j) {
        bool later = false;
        for (size_t k = j + 1; k < lg.size(); ++k) later = later || (ns[j] == ns[k]);
        if (!later) img.insert(img.end(), {nq[j][0], nq[j][1], nq[j][2]});
    }
    return true;
}

int multiplicity_class(int q2, int q3, int q4) {
    if (q2 == q3 && q3 == q4) return 0;
    if (q3 == q4 || q2 == q4 || q2 == q3) return 1;
    return 2;
}

// the irreducible quartets of every irreducible q1 on the full mesh
void build_quartets(tc3_ctx *c) {
    const Mesh &m = c->mesh;
    const int np = m.nfull;
    c->quart.clear();
    c->quart_img_off.assign(1, 0);
    c->quart_img.clear();
    for (int q1 = 0; q1 < m.nirr; ++q1) {
        const int q1f = m.irr_full_index[q1];
        for (int q2 = 0; q2 < np; ++q2)
            for (int q3 = q2; q3 < np; ++q3) {
                const int q4 = fourth_grid_index(q1f, q2, q3, m.dims);
                if (q4 < q3) continue;
                if (!quartet_images(m, q1, q2, q3, q4, c->quart_img)) continue;
                c->quart.insert(c->quart.end(), {q1, q2, q3, q4, multiplicity_class(q2, q3, q4)});
                c->quart_img_off.push_back((int)c->quart_img.size() / 3);
            }
    }
    c->quart_gen = c->gen_model;
}

// the quartets visited by the Monte-Carlo rows (the (q2, q3) pairs of each row's two grids,
// compute_fourphonon_scattering :87-95), with the draw count of every row in tw [t][nb]
void build_quartets_mc(const tc3_ctx *c, std::vector<int> &quart, std::vector<int> &img_off, std::vector<int> &img,
                       std::vector<float> &tw) {
    const Mesh &m = c->mesh;
    const int np = m.nfull, nb = c->nb;
    const long n4 = c->mc4_npoints;
    quart.clear();
    img_off.assign(1, 0);
    img.clear();
    tw.clear();
    // (q2, q3) pairs seen for this q1: -1 reducible, else the quartet (at most 2 nb n4^2 draws
    // per q1, so a hash map instead of an np^2 table)
    std::unordered_map<long long, int> slot;
    for (int q1 = 0; q1 < m.nirr; ++q1) {
        slot.clear();
        const int q1f = m.irr_full_index[q1];
        for (int b1 = 0; b1 < nb; ++b1) {
            if (c->row_of[(size_t)q1 * nb + b1] < 0) continue;
            const long *ga = &c->g4[((size_t)q1 * nb + b1) * 2 * n4], *gb = ga + n4;
            for (long qi = 0; qi < n4; ++qi)
                for (long qj = 0; qj < n4; ++qj) {
                    const int q2 = (int)ga[qi], q3 = (int)gb[qj];
                    if (q3 < q2) continue;
                    const int q4 = fourth_grid_index(q1f, q2, q3, m.dims);
                    if (q4 < q3) continue;
                    const auto ins = slot.emplace((long long)q2 * np + q3, -2);
                    int &s = ins.first->second;
                    if (ins.second) {
                        if (quartet_images(m, q1, q2, q3, q4, img)) {
                            s = (int)quart.size() / 5;
                            quart.insert(quart.end(), {q1, q2, q3, q4, multiplicity_class(q2, q3, q4)});
                            img_off.push_back((int)img.size() / 3);
                            tw.resize(tw.size() + nb, 0.0f);
                        } else {
                            s = -1;
                        }
                    }
                    if (s >= 0) tw[(size_t)s * nb + b1] += 1.0f;
                }
        }
    }
}

}  // namespace

int fourphonon_terms(tc3_ctx *c, const F4Inputs &in, std::vector<double> &g4, double *d_od4) {
    const int nb = c->nb, na = c->cr.na, nT = in.nT, nirr = c->mesh.nirr, np = c->mesh.nfull, nrows = (int)c->row_q.size();
    const bool mc = c->mc4_on;
    if (mc && c->mquart_gen != c->gen_model) {
        build_quartets_mc(c, c->mquart, c->mquart_img_off, c->mquart_img, c->mquart_tw);
        c->mquart_gen = c->gen_model;
    } else if (!mc && c->quart_gen != c->gen_model) {
        build_quartets(c);
    }
    const std::vector<int> &quart = mc ? c->mquart : c->quart, &qoff = mc ? c->mquart_img_off : c->quart_img_off,
                           &qimg = mc ? c->mquart_img : c->quart_img;
    const std::vector<float> &mtw = c->mquart_tw;
    const int nqt = (int)quart.size() / 5;
    g4.assign((size_t)nT * nrows, 0.0);
    hipStream_t st = c->stream;
    hipError_t e;
#define TRY(x) do { e = (x); if (e != hipSuccess) return device_error(e, #x, __FILE__, __LINE__); } while (0)
    const size_t ncol = (size_t)np * nb;
    if (nqt == 0) {
        TRY(hipMemsetAsync(d_od4, 0, sizeof(double) * nT * nrows * ncol, st));
        return 0;
    }
    auto up = [&](const char *name, const void *src, size_t bytes, void **dst) -> hipError_t {
        hipError_t r = c->pool_get(name, std::max<size_t>(bytes, 16), dst);
        if (r == hipSuccess && bytes) r = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, st);
        return r;
    };
    F4Args a{};
    a.nb = nb; a.na = na; a.nT = nT; a.nirr = nirr;
    void *p;
    TRY(up("f4_q", quart.data(), sizeof(int) * quart.size(), &p)); a.fq = (const int *)p;
    a.tw = nullptr;
    if (mc) {
        TRY(up("f4_tw", mtw.data(), sizeof(float) * mtw.size(), &p));
        a.tw = (const float *)p;
    }
    TRY(up("f4_blk", c->blk4_off.data(), sizeof(int) * c->blk4_off.size(), &p)); a.blk4 = (const int *)p;
    TRY(up("f4_lv", c->f4_lv.data(), sizeof(double) * c->f4_lv.size(), &p)); a.lv = (const double *)p;
    TRY(up("f4_mwm", c->f4_mwm.data(), sizeof(double) * c->f4_mwm.size(), &p)); a.mwm = (const double *)p;
    TRY(up("f4_qc", c->mesh.qcart.data(), sizeof(double) * c->mesh.qcart.size(), &p)); a.qcart = (const double *)p;
    a.egv_irr = c->d_egv_irr; a.egv_full = c->d_egv_full; a.om_irr = c->d_om_irr; a.om_full = c->d_om_full;
    a.be = in.d_be; a.sigsq = in.d_sigsq; a.irr_of_full = c->d_irr_of_full; a.row_of = in.d_rowof;
    a.u_irr = in.d_uirr; a.u_full = in.d_ufull;
    a.itype = in.integrationtype; a.sigma_fixed = in.sigma_fixed; a.thresh = in.thresh;
    const double npg = mc ? (double)c->mc4_npoints : (double)np;   // points of one grid (mcg%weight)
    a.pref = (kPi / 96.0) / (npg * npg);
    TRY(c->pool_get("f4_g", sizeof(double) * (size_t)nqt * nT * nb, &p)); a.g = (double *)p;
    TRY(c->pool_get("f4_s", sizeof(double) * (size_t)nqt * nT * 3 * nb * nb, &p)); a.s = (double *)p;
    // TC3_F4_BATCHED: the batched kernels also for cells of up to two atoms; TC3_F4_NQB: at
    // most that many quartets per batch (both for tests)
    if (nb <= kF4MaxNb && std::getenv("TC3_F4_BATCHED") == nullptr) {
        hipLaunchKernelGGL(k_fourphonon, dim3(nqt), dim3(kF4Block), 0, st, a);
        TRY(hipGetLastError());
    } else {
        // batched: phi4 and the contraction steps of up to nqb quartets at once, two buffers of
        // nb^4 complex values per quartet (at most 1 GiB each); FP64, or FP32 in the mixed modes
        auto batched = [&](auto tag) -> int {
            typedef decltype(tag) T;
            const size_t n4 = (size_t)nb * nb * nb * nb, per = n4 * sizeof(cplx<T>);
            int nqb = (int)std::max<size_t>(1, std::min<size_t>((size_t)nqt, ((size_t)1 << 30) / per));
            if (const char *env = std::getenv("TC3_F4_NQB")) nqb = std::max(1, std::min(nqb, std::atoi(env)));
            const int nent = std::max(c->nfc4, 1);
            void *pE1, *pEf, *pph, *pX, *pY;
            const char *tg = sizeof(T) == 8 ? "d" : "f";
            TRY(c->pool_get((std::string("f4_E1") + tg).c_str(), sizeof(cplx<T>) * (size_t)nirr * nb * nb, &pE1));
            TRY(c->pool_get((std::string("f4_Ef") + tg).c_str(), sizeof(cplx<T>) * (size_t)np * nb * nb, &pEf));
            TRY(c->pool_get("f4_ph", sizeof(cd) * (size_t)nqb * nent, &pph));
            TRY(c->pool_get("f4_X", per * nqb, &pX));
            TRY(c->pool_get("f4_Y", per * nqb, &pY));
            hipLaunchKernelGGL(k_scale<T>, dim3(256), dim3(256), 0, st, (long)nirr * nb * nb, nb, c->d_egv_irr,
                               c->d_om_irr, (cplx<T> *)pE1);
            hipLaunchKernelGGL(k_scale<T>, dim3(1024), dim3(256), 0, st, (long)np * nb * nb, nb, c->d_egv_full,
                               c->d_om_full, (cplx<T> *)pEf);
            TRY(hipGetLastError());
            const cplx<T> *E1 = (const cplx<T> *)pE1, *Ef = (const cplx<T> *)pEf;
            cplx<T> *X = (cplx<T> *)pX, *Y = (cplx<T> *)pY;
            // epilogue: temperature slots that fit 100 KiB of LDS (at least one)
            int tb = kF4Tb;
            while (tb > 1 && f4_epi_lds(nb, tb) > 100 * 1024) --tb;
            const size_t shm = f4_epi_lds(nb, tb);
            if (shm > 160 * 1024) return c->fail(5, "band count too large for the four-phonon epilogue's LDS");
            TRY(hipFuncSetAttribute((const void *)k_f4_epi<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
            const long R = (long)nb * nb * nb;
            const unsigned sgx = (unsigned)((R + 63) / 64);
            const int NT = (nb + 15) / 16;
            if (NT > 6) return c->fail(5, "four-phonon contraction handles at most 96 bands");
            auto stage = [&](int leg, const cplx<T> *src, cplx<T> *dst, int t0, int nq) -> hipError_t {
                const dim3 gs(sgx, (unsigned)nq);
                const cplx<T> *E = leg == 0 ? E1 : Ef;
                switch (NT) {
                case 1: hipLaunchKernelGGL((k_f4_stage<T, 1>), gs, dim3(256), 0, st, nb, a.fq, t0, leg, E, src, dst); break;
                case 2: hipLaunchKernelGGL((k_f4_stage<T, 2>), gs, dim3(256), 0, st, nb, a.fq, t0, leg, E, src, dst); break;
                case 3: hipLaunchKernelGGL((k_f4_stage<T, 3>), gs, dim3(256), 0, st, nb, a.fq, t0, leg, E, src, dst); break;
                case 4: hipLaunchKernelGGL((k_f4_stage<T, 4>), gs, dim3(256), 0, st, nb, a.fq, t0, leg, E, src, dst); break;
                case 5: hipLaunchKernelGGL((k_f4_stage<T, 5>), gs, dim3(256), 0, st, nb, a.fq, t0, leg, E, src, dst); break;
                default: hipLaunchKernelGGL((k_f4_stage<T, 6>), gs, dim3(256), 0, st, nb, a.fq, t0, leg, E, src, dst); break;
                }
                return hipGetLastError();
            };
            F4EpiArgs ep{};
            ep.nb = nb; ep.nT = nT; ep.nirr = nirr; ep.tb = tb; ep.itype = a.itype;
            ep.fq = a.fq; ep.om_irr = a.om_irr; ep.om_full = a.om_full; ep.be = a.be; ep.sigsq = a.sigsq;
            ep.irr_of_full = a.irr_of_full; ep.row_of = a.row_of; ep.u_irr = a.u_irr; ep.u_full = a.u_full;
            ep.sigma_fixed = a.sigma_fixed; ep.thresh = a.thresh; ep.pref = a.pref; ep.tw = a.tw;
            ep.psi = X; ep.g = a.g; ep.s = a.s;
            for (int t0 = 0; t0 < nqt; t0 += nqb) {
                const int nq = std::min(nqb, nqt - t0);
                hipLaunchKernelGGL(k_f4_phase, dim3((unsigned)((nent + 255) / 256), (unsigned)nq), dim3(256), 0, st,
                                   nent, a.fq, t0, a.qcart, a.lv, (cd *)pph);
                hipLaunchKernelGGL(k_f4_phi<T>, dim3((unsigned)std::min<size_t>((n4 + 255) / 256, 65535), (unsigned)nq),
                                   dim3(256), 0, st, nb, na, nent, a.blk4, a.mwm, (const cd *)pph, X);
                TRY(hipGetLastError());
                // ia with e1 (q1), ib with e2, ic with e3, id with e4: X -> Y -> X -> Y -> X
                TRY(stage(0, X, Y, t0, nq));
                TRY(stage(1, Y, X, t0, nq));
                TRY(stage(2, X, Y, t0, nq));
                TRY(stage(3, Y, X, t0, nq));
                ep.t0 = t0;
                for (int T0 = 0; T0 < nT; T0 += tb) {
                    ep.T0 = T0;
                    ep.nTb = std::min(tb, nT - T0);
                    hipLaunchKernelGGL(k_f4_epi<T>, dim3((unsigned)nb, (unsigned)nq), dim3(kF4EpiBlock), shm, st, ep);
                    TRY(hipGetLastError());
                }
            }
            return 0;
        };
        const int rb = in.precision == 0 ? batched(double{}) : batched(float{});
        if (rb) return rb;
    }
    std::vector<double> hg((size_t)nqt * nT * nb), hs((size_t)nqt * nT * 3 * nb * nb);
    TRY(hipMemcpyAsync(hg.data(), a.g, sizeof(double) * hg.size(), hipMemcpyDeviceToHost, st));
    TRY(hipMemcpyAsync(hs.data(), a.s, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    // one temperature at a time on the host (rows x nfull nb), then to its slice of d_od4
    std::vector<int> row_of((size_t)nirr * nb, -1);
    for (int r = 0; r < nrows; ++r) row_of[(size_t)c->row_q[r] * nb + c->row_b[r]] = r;
    std::vector<double> odT((size_t)nrows * ncol);
    std::vector<double> buf(nb);
    std::vector<int> nn(nb), imgs;
    const long nlist = mc ? c->mc4_npoints : np;
    for (int T = 0; T < nT; ++T) {
        std::fill(odT.begin(), odT.end(), 0.0);
        // the terms of every image (scattering_fourphonon.f90:176-187)
        for (int t = 0; t < nqt; ++t) {
            const int q1 = quart[5 * t], qq[3] = {quart[5 * t + 1], quart[5 * t + 2], quart[5 * t + 3]};
            const double *gt = &hg[((size_t)t * nT + T) * nb];
            const double *stt = &hs[((size_t)t * nT + T) * 3 * nb * nb];
            for (int b1 = 0; b1 < nb; ++b1) {
                const int r = row_of[(size_t)q1 * nb + b1];
                if (r < 0) continue;
                const double w1 = c->om_irr[(size_t)q1 * nb + b1];
                double *od = &odT[(size_t)r * ncol];
                for (int i = qoff[t]; i < qoff[t + 1]; ++i) {
                    g4[(size_t)T * nrows + r] += gt[b1];
                    for (int leg = 0; leg < 3; ++leg) {
                        const int qp = qimg[3 * i + leg];
                        // frequency of the quartet's own point (the host holds those of the
                        // irreducible points; a full-grid point has its irreducible point's)
                        const double *wq = &c->om_irr[(size_t)c->mesh.irr_of_full[qq[leg]] * nb];
                        for (int b = 0; b < nb; ++b)
                            od[(size_t)qp * nb + b] += 2.0 * stt[(size_t)leg * nb * nb + b1 * nb + b] * wq[b] / w1;
                    }
                }
            }
        }
        // symmetrize_and_distribute (scattering_fourphonon.f90:143-205), over the row's first
        // grid in drawn order (every point on the full mesh)
        for (int r = 0; r < nrows; ++r) {
            const int q1 = c->row_q[r];
            const std::vector<int> &lg = c->mesh.littlegroup[q1];
            double *od = &odT[(size_t)r * ncol];
            const long *ga = mc ? &c->g4[((size_t)q1 * nb + c->row_b[r]) * 2 * c->mc4_npoints] : nullptr;
            for (long qi = 0; qi < nlist; ++qi) {
                const int q2 = mc ? (int)ga[qi] : (int)qi;
                imgs.clear();
                bool skip = false;
                for (int o : lg) {
                    const int x = c->mesh.rot[(size_t)o * np + q2];
                    if (x < 0) continue;
                    if (x < q2) { skip = true; break; }
                    imgs.push_back(x);
                }
                if (skip) continue;
                std::fill(buf.begin(), buf.end(), 0.0);
                std::fill(nn.begin(), nn.end(), 0);
                for (int x : imgs)
                    for (int b = 0; b < nb; ++b)
                        if (od[(size_t)x * nb + b] > 0.0) {
                            ++nn[b];
                            buf[b] += od[(size_t)x * nb + b];
                        }
                for (int x : imgs)
                    for (int b = 0; b < nb; ++b)
                        if (nn[b]) od[(size_t)x * nb + b] = buf[b] / (double)nn[b];
            }
        }
        TRY(hipMemcpy(d_od4 + (size_t)T * nrows * ncol, odT.data(), sizeof(double) * odT.size(), hipMemcpyHostToDevice));
    }
#undef TRY
    return 0;
}

}  // namespace tc3
