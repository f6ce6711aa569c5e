// Kernels shared by the forward three-phonon pass and its adjoint: the Phi3 pretransform,
// a batched complex GEMM (optionally with split-BF16x3 products) and the eigenvector scaling,
// and the three-GEMM chain that turns Phi3 into Psi for a batch of triplets.
#pragma once

#include <hip/hip_runtime.h>

#include "tc3_ctx.hpp"

namespace tc3 {
namespace {

constexpr int kBlock = 256;

template <typename T>
struct cplx { T x, y; };

// Phi[tl][ia][ib][ic] for the triplets of one batch; workgroup per (triplet, atom a1).
template <typename T>
__global__ void k_phi(int na, int t0, const int *__restrict__ tq2, const int *__restrict__ tq3,
                      const double *__restrict__ qcart, const int *__restrict__ a1_off,
                      const int *__restrict__ blk_off, const double *__restrict__ lv2,
                      const double *__restrict__ lv3, const double *__restrict__ mwm,
                      cplx<T> *__restrict__ phi) {
    extern __shared__ cd ph[];
    const int tl = blockIdx.x / na, a1 = blockIdx.x % na;
    const int t = t0 + tl;
    const int nb = 3 * na;
    const double *q2 = qcart + 3 * tq2[t];
    const double *q3 = qcart + 3 * tq3[t];
    const int f0 = a1_off[a1], f1 = a1_off[a1 + 1];
    for (int f = f0 + threadIdx.x; f < f1; f += blockDim.x) {
        const double arg = -kTwoPi * (q2[0] * lv2[3 * f] + q2[1] * lv2[3 * f + 1] + q2[2] * lv2[3 * f + 2] +
                                      q3[0] * lv3[3 * f] + q3[1] * lv3[3 * f + 1] + q3[2] * lv3[3 * f + 2]);
        double s, c;
        sincos(arg, &s, &c);
        ph[f - f0] = {c, s};
    }
    __syncthreads();
    const size_t nb2 = (size_t)nb * nb;
    cplx<T> *out = phi + (size_t)tl * nb * nb2 + (size_t)(3 * a1) * nb2;
    for (size_t e = threadIdx.x; e < 3 * nb2; e += blockDim.x) {
        const int i = (int)(e / nb2);
        const int rem = (int)(e % nb2);
        const int ib = rem / nb, ic = rem % nb;
        const int a2 = ib / 3, j = ib % 3, a3 = ic / 3, k = ic % 3;
        const int blk = (a1 * na + a2) * na + a3;
        double ax = 0, ay = 0;
        for (int f = blk_off[blk]; f < blk_off[blk + 1]; ++f) {
            const double m = mwm[27 * f + 9 * i + 3 * j + k];
            const cd p = ph[f - f0];
            ax += m * p.x;
            ay += m * p.y;
        }
        out[e] = {(T)ax, (T)ay};
    }
}

// x = hi + lo with hi = bf16(x), lo = bf16(x - hi), returned as FP32 values.
__device__ inline void split_bf16(float x, float &h, float &l) {
    h = (float)(__bf16)x;
    l = (float)(__bf16)(x - h);
}
// a.b ~ ah.bh + ah.bl + al.bh: the split-BF16x3 product (each term exact in FP32).
__device__ inline float mul3(float ah, float al, float bh, float bl) { return fmaf(ah, bh, fmaf(ah, bl, al * bh)); }

// Generic batched complex GEMM, one 16x16 output tile per workgroup.
// C_z(m,n) = sum_k A_z(m,k) B_z(k,n); matrices addressed by element strides, the batch
// index z selecting A/B through optional index tables (table[z / div]).  SPLIT: every
// FP32 product taken as the split-BF16x3 product, as the matrix cores form it.
template <typename T, bool SPLIT = false>
__global__ void k_cgemm(int M, int N, int K, int ntm, int ntn,
                        const cplx<T> *__restrict__ A, const int *__restrict__ aidx, int adiv, long sA, int am, int ak,
                        const cplx<T> *__restrict__ B, const int *__restrict__ bidx, int bdiv, long sB, int bk, int bn,
                        cplx<T> *__restrict__ C, long sC, int cm, int cn) {
    __shared__ cplx<T> As[16][17], Bs[16][17];
    const int tiles = ntm * ntn;
    const long z = blockIdx.x / tiles;
    const int tile = blockIdx.x % tiles;
    const int m0 = (tile / ntn) * 16, n0 = (tile % ntn) * 16;
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const cplx<T> *Az = A + (aidx ? (long)aidx[z / adiv] : z) * sA;
    const cplx<T> *Bz = B + (bidx ? (long)bidx[z / bdiv] : z) * sB;
    T rx = 0, ry = 0;
    for (int k0 = 0; k0 < K; k0 += 16) {
        {
            const int m = m0 + ty, k = k0 + tx;
            As[ty][tx] = (m < M && k < K) ? Az[(long)m * am + (long)k * ak] : cplx<T>{0, 0};
            const int kk = k0 + ty, n = n0 + tx;
            Bs[ty][tx] = (kk < K && n < N) ? Bz[(long)kk * bk + (long)n * bn] : cplx<T>{0, 0};
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            const cplx<T> a = As[ty][kk], b = Bs[kk][tx];
            if constexpr (SPLIT) {
                float arh, arl, aih, ail, brh, brl, bih, bil;
                split_bf16(a.x, arh, arl);
                split_bf16(a.y, aih, ail);
                split_bf16(b.x, brh, brl);
                split_bf16(b.y, bih, bil);
                rx += mul3(arh, arl, brh, brl) - mul3(aih, ail, bih, bil);
                ry += mul3(arh, arl, bih, bil) + mul3(aih, ail, brh, brl);
            } else {
                rx += a.x * b.x - a.y * b.y;
                ry += a.x * b.y + a.y * b.x;
            }
        }
        __syncthreads();
    }
    const int m = m0 + ty, n = n0 + tx;
    if (m < M && n < N) C[z * sC + (long)m * cm + (long)n * cn] = {rx, ry};
}

// E[q][ia][b] = egv[q][ia][b] / sqrt(omega_b), zero for omega_b < lo_freqtol.
template <typename T>
__global__ void k_scale(long n, int nb, const cd *__restrict__ egv, const double *__restrict__ om,
                        cplx<T> *__restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long q = i / ((long)nb * nb);
        const int b = (int)(i % nb);
        const double w = om[q * nb + b];
        const double s = (w < 1e-9) ? 0.0 : 1.0 / sqrt(w);
        out[i] = {(T)(egv[i].x * s), (T)(egv[i].y * s)};
    }
}

__device__ inline double gauss(double x, double mu, double sigma) {
    // lo_gauss (gottochblandat_calculus.f90:268-288)
    const double d = x - mu;
    if (fabs(d) > 8.0 * sigma) return 0.0;
    const double inv = 1.0 / sigma;
    return exp(-d * d * 0.5 * inv * inv) * 0.398942280401433 * inv;
}

template <typename T, bool SPLIT = false>
int launch_chain(tc3_ctx *c, hipStream_t st, int t0, int nt, const cplx<T> *E1, const cplx<T> *Ef,
                 cplx<T> *phi, cplx<T> *t1, cplx<T> *t2, cplx<T> *psi, const int *d_tq1, const int *d_tq2,
                 const int *d_tq3, const double *d_qcart, const int *d_a1off, const int *d_blk3,
                 const double *d_lv2, const double *d_lv3, const double *d_mwm, size_t phi_shmem) {
    const int na = c->cr.na, nb = 3 * na;
    const long nb2 = (long)nb * nb, nb3 = nb2 * nb;
    hipLaunchKernelGGL(k_phi<T>, dim3(nt * na), dim3(kBlock), phi_shmem, st, na, t0, d_tq2, d_tq3, d_qcart,
                       d_a1off, d_blk3, d_lv2, d_lv3, d_mwm, phi);
    // stage A: t1[tl](b1, n) = sum_ia E1[q1](ia, b1) phi[tl](ia, n)
    {
        const int ntm = (nb + 15) / 16, ntn = (int)((nb2 + 15) / 16);
        hipLaunchKernelGGL((k_cgemm<T, SPLIT>), dim3((unsigned)((long)nt * ntm * ntn)), dim3(256), 0, st, nb, (int)nb2, nb,
                           ntm, ntn, E1, d_tq1 + t0, 1, nb2, 1, nb, phi, (const int *)nullptr, 1, nb3, (int)nb2, 1,
                           t1, nb3, (int)nb2, 1);
    }
    const int ntm = (nb + 15) / 16, ntn = (nb + 15) / 16;
    // stage B: t2[tl,b1](b2, ic) = sum_ib E2[q2](ib, b2) t1[tl,b1](ib, ic)
    hipLaunchKernelGGL((k_cgemm<T, SPLIT>), dim3((unsigned)((long)nt * nb * ntm * ntn)), dim3(256), 0, st, nb, nb, nb, ntm,
                       ntn, Ef, d_tq2 + t0, nb, nb2, 1, nb, t1, (const int *)nullptr, 1, nb2, nb, 1, t2, nb2, nb, 1);
    // stage C: psi[tl,b1](b2, b3) = sum_ic t2[tl,b1](b2, ic) E3[q3](ic, b3)
    hipLaunchKernelGGL((k_cgemm<T, SPLIT>), dim3((unsigned)((long)nt * nb * ntm * ntn)), dim3(256), 0, st, nb, nb, nb, ntm,
                       ntn, t2, (const int *)nullptr, 1, nb2, nb, 1, Ef, d_tq3 + t0, nb, nb2, nb, 1, psi, nb2, nb, 1);
    return 0;
}

}  // namespace
}  // namespace tc3
