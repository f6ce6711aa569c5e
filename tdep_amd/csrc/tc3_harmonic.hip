// Harmonic properties on the q-mesh, FP64 on the GPU.
//
//   k_dynmat     dynamical matrix and its q-gradient per irreducible point
//                (type_forceconstant_secondorder_dynamicalmatrix.f90:484-788, non-polar)
//   k_jacobi     Hermitian eigen-solve: parallel cyclic Jacobi, one workgroup per matrix
//                (replaces zheev, …_dynamicalmatrix.f90:113-124)
//   k_post       Gamma zeroing, sorting, degenerate-subspace rotation to a fixed gauge,
//                group velocities with degenerate averaging (…:145-215, :287-374, :472)
//   k_unfold     rotation of eigenvectors/velocities to the full grid
//                (type_phonon_dispersions_generation.f90:117-173 with
//                 lo_eigenvector_transformation_matrix, type_symmetryoperation.f90:243-293)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "tc3_ctx.hpp"

namespace tc3 {

__device__ const double kWellDefinedDev[69] = {
    0.0, 0.1, -0.1, 0.1111111111111111, -0.1111111111111111, 0.125, -0.125,
    0.1428571428571428, -0.1428571428571428, 0.1666666666666667, -0.1666666666666667,
    0.2, -0.2, 0.2222222222222222, -0.2222222222222222, 0.25, -0.25,
    0.2857142857142857, -0.2857142857142857, 0.3, -0.3, 0.3333333333333333,
    -0.3333333333333333, 0.375, -0.375, 0.4, -0.4, 0.4285714285714285,
    -0.4285714285714285, 0.4444444444444444, -0.4444444444444444, 0.5, -0.5,
    0.5555555555555556, -0.5555555555555556, 0.5714285714285714, -0.5714285714285714,
    0.6, -0.6, 0.625, -0.625, 0.6666666666666667, -0.6666666666666667, 0.7, -0.7,
    0.7071067811865475, -0.7071067811865475, 0.7142857142857143, -0.7142857142857143,
    0.75, -0.75, 0.7777777777777778, -0.7777777777777778, 0.8, -0.8,
    0.8333333333333334, -0.8333333333333334, 0.8571428571428571, -0.8571428571428571,
    0.8660254037844386, -0.8660254037844386, 0.875, -0.875, 0.8888888888888888,
    -0.8888888888888888, 0.9, -0.9, 1.0, -1.0};

__device__ double chop_dev(double x, double tol) {
    for (int i = 0; i < 69; ++i)
        if (fabs(x - kWellDefinedDev[i]) < tol) return kWellDefinedDev[i];
    return x;
}

namespace {

constexpr int kBlock = 256;
constexpr double kGold = 0.6180339887498949;
constexpr double kPlastic = 0.7548776662466927;

// D[q][(a2,j)][(a1,i)] = sum_pairs m_ij e^{i qv.R} / sqrt(m1 m2); gD likewise with i R_alpha.
__global__ void k_dynmat(int na, const double *__restrict__ qv, const int *__restrict__ blk_off,
                         const double *__restrict__ lv, const double *__restrict__ pm,
                         const double *__restrict__ ism, cd *__restrict__ D, cd *__restrict__ gD) {
    const int q = blockIdx.x;
    const int nb = 3 * na;
    const double q0 = qv[3 * q], q1 = qv[3 * q + 1], q2 = qv[3 * q + 2];
    cd *Dq = D + (size_t)q * nb * nb;
    cd *gq = gD + (size_t)q * 3 * nb * nb;
    for (int blk = threadIdx.x; blk < na * na; blk += blockDim.x) {
        const int a1 = blk / na, a2 = blk % na;
        cd acc[9], gacc[3][9];
        for (int k = 0; k < 9; ++k) {
            acc[k] = {0, 0};
            for (int al = 0; al < 3; ++al) gacc[al][k] = {0, 0};
        }
        for (int p = blk_off[blk]; p < blk_off[blk + 1]; ++p) {
            const double r0 = lv[3 * p], r1 = lv[3 * p + 1], r2 = lv[3 * p + 2];
            const double qr = r0 * q0 + r1 * q1 + r2 * q2;
            double s, c;
            sincos(qr, &s, &c);
            const double R[3] = {r0, r1, r2};
            for (int k = 0; k < 9; ++k) {
                const double m = pm[9 * p + k];
                acc[k].x += c * m;
                acc[k].y += s * m;
                for (int al = 0; al < 3; ++al) {
                    // i R e^{iqr} m
                    gacc[al][k].x += -R[al] * s * m;
                    gacc[al][k].y += R[al] * c * m;
                }
            }
        }
        const double w = ism[a1] * ism[a2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const size_t idx = (size_t)(3 * a2 + j) * nb + (3 * a1 + i);
                Dq[idx] = cscale(acc[3 * i + j], w);
                for (int al = 0; al < 3; ++al) gq[(size_t)al * nb * nb + idx] = cscale(gacc[al][3 * i + j], w);
            }
    }
}

// Parallel cyclic Jacobi for one Hermitian matrix per workgroup.  Round-robin pairing
// (circle method) gives nb/2 disjoint (p,q) rotations per round; all rotations of a
// round are applied to the columns, then to the rows, then to V.
#ifndef TC3_JAC_BLOCK
#define TC3_JAC_BLOCK 1024
#endif
constexpr int kJacBlock = TC3_JAC_BLOCK;   // threads per matrix
template <bool USE_LDS>
__global__ void __launch_bounds__(kJacBlock) k_jacobi(int nb, cd *__restrict__ Aall, cd *__restrict__ Vall, double *__restrict__ lam,
                         int maxsweep, int *__restrict__ status) {
    constexpr bool use_lds = USE_LDS;
    const int q = blockIdx.x;
    // A and V live in LDS when they fit (use_lds), else in global memory
    extern __shared__ cd jsm[];
    cd *A = use_lds ? jsm : Aall + (size_t)q * nb * nb;
    cd *V = use_lds ? jsm + (size_t)nb * nb : Vall + (size_t)q * nb * nb;
    if (use_lds)
        for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) A[i] = Aall[(size_t)q * nb * nb + i];
    const int mpad = nb + (nb & 1);
    const int half = mpad / 2;
    __shared__ int sp[128], sq[128];
    __shared__ double sc[128], ss[128], sphx[128], sphy[128];
    __shared__ double red[2 * (kJacBlock / 64)];
    __shared__ int done;
    for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) V[i] = {(i / nb) == (i % nb) ? 1.0 : 0.0, 0.0};
    __syncthreads();
    int sweep = 0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    for (; sweep < maxsweep; ++sweep) {
        // convergence test: off-diagonal vs diagonal Frobenius norm (wave sums, then LDS)
        double off = 0, dia = 0;
        for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) {
            const int r = i / nb, c = i % nb;
            if (r == c) dia += cabs2(A[i]);
            else off += cabs2(A[i]);
        }
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o, 64);
            dia += __shfl_xor(dia, o, 64);
        }
        if (lane == 0) {
            red[wv] = off;
            red[nwv + wv] = dia;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double offt = 0, diat = 0;
            for (int w = 0; w < nwv; ++w) {
                offt += red[w];
                diat += red[nwv + w];
            }
            done = (offt <= 1e-28 * diat) || offt == 0.0;
        }
        __syncthreads();
        if (done) break;
        for (int r = 0; r < mpad - 1; ++r) {
            // pairs of this round and their rotations
            for (int k = threadIdx.x; k < half; k += blockDim.x) {
                int p, qq;
                if (k == 0) { p = r; qq = mpad - 1; }
                else { p = (r + k) % (mpad - 1); qq = (r - k + mpad - 1) % (mpad - 1); }
                if (p > qq) { int t = p; p = qq; qq = t; }
                double c = 1, s = 0, ex = 1, ey = 0;
                if (qq < nb) {
                    const cd apq = A[(size_t)p * nb + qq];
                    const double b = sqrt(cabs2(apq));
                    if (b > 1e-300) {
                        const double app = A[(size_t)p * nb + p].x, aqq = A[(size_t)qq * nb + qq].x;
                        const double tau = (aqq - app) / (2.0 * b);
                        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        ex = apq.x / b;
                        ey = apq.y / b;
                    }
                } else {
                    qq = -1;
                }
                sp[k] = p; sq[k] = qq; sc[k] = c; ss[k] = s; sphx[k] = ex; sphy[k] = ey;
            }
            __syncthreads();
            // A <- U^H A U with U = diag(1, e^{-i phi}) [[c, s], [-s, c]] per pair: the rows
            // of pair P and the columns of pair Q form a 2x2 block that only the rotations of
            // P and Q touch; each block gets its column rotation, then its row rotation.
            for (int it = threadIdx.x; it < half * half; it += blockDim.x) {
                const int kP = it / half, kQ = it % half;
                const int p = sp[kP], qq = sq[kP], p2 = sp[kQ], q2 = sq[kQ];
                const bool rowP = qq >= 0 && ss[kP] != 0.0, colQ = q2 >= 0 && ss[kQ] != 0.0;
                if (!rowP && !colQ) continue;
                const bool hq = qq >= 0, hq2 = q2 >= 0;
                const cd zero = {0.0, 0.0};
                cd x00 = A[(size_t)p * nb + p2], x01 = zero, x10 = zero, x11 = zero;
                if (hq2) x01 = A[(size_t)p * nb + q2];
                if (hq) x10 = A[(size_t)qq * nb + p2];
                if (hq && hq2) x11 = A[(size_t)qq * nb + q2];
                if (colQ) {
                    const double c = sc[kQ], s = ss[kQ];
                    const cd em = {sphx[kQ], -sphy[kQ]};  // e^{-i phi}
                    cd ap = x00, emaq = cmul(em, x01);
                    x00 = csub(cscale(ap, c), cscale(emaq, s));
                    x01 = cadd(cscale(ap, s), cscale(emaq, c));
                    ap = x10;
                    emaq = cmul(em, x11);
                    x10 = csub(cscale(ap, c), cscale(emaq, s));
                    x11 = cadd(cscale(ap, s), cscale(emaq, c));
                }
                if (rowP) {
                    const double c = sc[kP], s = ss[kP];
                    const cd ep = {sphx[kP], sphy[kP]};  // e^{i phi}
                    cd ap = x00, epaq = cmul(ep, x10);
                    x00 = csub(cscale(ap, c), cscale(epaq, s));
                    x10 = cadd(cscale(ap, s), cscale(epaq, c));
                    ap = x01;
                    epaq = cmul(ep, x11);
                    x01 = csub(cscale(ap, c), cscale(epaq, s));
                    x11 = cadd(cscale(ap, s), cscale(epaq, c));
                }
                A[(size_t)p * nb + p2] = x00;
                if (hq2) A[(size_t)p * nb + q2] = x01;
                if (hq) A[(size_t)qq * nb + p2] = x10;
                if (hq && hq2) A[(size_t)qq * nb + q2] = x11;
            }
            // V <- V U
            for (int it = threadIdx.x; it < nb * half; it += blockDim.x) {
                const int row = it / half, k = it % half;
                const int p = sp[k], qq = sq[k];
                if (qq < 0 || ss[k] == 0.0) continue;
                const double c = sc[k], s = ss[k];
                const cd em = {sphx[k], -sphy[k]};
                const cd vp = V[(size_t)row * nb + p], emvq = cmul(em, V[(size_t)row * nb + qq]);
                V[(size_t)row * nb + p] = csub(cscale(vp, c), cscale(emvq, s));
                V[(size_t)row * nb + qq] = cadd(cscale(vp, s), cscale(emvq, c));
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < nb; i += blockDim.x) lam[(size_t)q * nb + i] = A[(size_t)i * nb + i].x;
    if (use_lds)
        for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) Vall[(size_t)q * nb * nb + i] = V[i];
    if (threadIdx.x == 0) status[q] = (sweep >= maxsweep) ? 1 : 0;
}

// Serial Jacobi for a small Hermitian matrix (degenerate-subspace probe), ascending order.
__device__ void small_eigh(int k, cd *M, cd *W, double *ev, int *ord) {
    for (int i = 0; i < k * k; ++i) W[i] = {(i / k) == (i % k) ? 1.0 : 0.0, 0.0};
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, dia = 0;
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) (i == j ? dia : off) += cabs2(M[i * k + j]);
        if (off <= 1e-32 * dia || off == 0.0) break;
        for (int p = 0; p < k; ++p)
            for (int q = p + 1; q < k; ++q) {
                const cd apq = M[p * k + q];
                const double b = sqrt(cabs2(apq));
                if (b <= 1e-300) continue;
                const double tau = (M[q * k + q].x - M[p * k + p].x) / (2.0 * b);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                const cd em = {apq.x / b, -apq.y / b}, ep = {apq.x / b, apq.y / b};
                for (int r = 0; r < k; ++r) {
                    cd ap = M[r * k + p], aq = cmul(em, M[r * k + q]);
                    M[r * k + p] = csub(cscale(ap, c), cscale(aq, s));
                    M[r * k + q] = cadd(cscale(ap, s), cscale(aq, c));
                    cd wp = W[r * k + p], wq = cmul(em, W[r * k + q]);
                    W[r * k + p] = csub(cscale(wp, c), cscale(wq, s));
                    W[r * k + q] = cadd(cscale(wp, s), cscale(wq, c));
                }
                for (int r = 0; r < k; ++r) {
                    cd ap = M[p * k + r], aq = cmul(ep, M[q * k + r]);
                    M[p * k + r] = csub(cscale(ap, c), cscale(aq, s));
                    M[q * k + r] = cadd(cscale(ap, s), cscale(aq, c));
                }
            }
    }
    for (int i = 0; i < k; ++i) { ev[i] = M[i * k + i].x; ord[i] = i; }
    for (int i = 1; i < k; ++i) {  // stable insertion sort by eigenvalue
        int o = ord[i];
        int j = i - 1;
        while (j >= 0 && ev[ord[j]] > ev[o]) { ord[j + 1] = ord[j]; --j; }
        ord[j + 1] = o;
    }
}

// Per irreducible point: Gamma zeroing, sort by signed frequency, canonical gauge of the
// degenerate subspaces, phase fix, chop, group velocities.
template <bool USE_LDS>
__global__ void k_post(int nb, const double *__restrict__ lam_in, const cd *__restrict__ Vin,
                       const cd *__restrict__ gD, const int *__restrict__ at_gamma,
                       double *__restrict__ omega, cd *__restrict__ egv, double *__restrict__ vel,
                       cd *__restrict__ scratch, double veltol) {
    constexpr bool use_lds = USE_LDS;
    const int q = blockIdx.x;
    __shared__ double lam[128], om[128], hf[3][128];
    __shared__ int ord[128], grp[129];
    __shared__ int ngrp;
    // work arrays of the degenerate-group and subspace passes
    __shared__ int s_so[128], s_first[128];
    __shared__ double s_ev[128];
    // eigenvectors in LDS when they fit (use_lds), else in the output array
    extern __shared__ cd psm[];
    const cd *V = Vin + (size_t)q * nb * nb;
    cd *Eg = egv + (size_t)q * nb * nb;
    cd *E = use_lds ? psm : Eg;
    cd *scr = scratch + (size_t)q * (2 * nb * nb + nb);
    // sorts as ranks (ties by index, as a stable sort): the three smallest |lambda| at Gamma
    // set to zero, then the modes ordered by signed frequency
    __shared__ double lam0[128];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) lam0[i] = lam_in[(size_t)q * nb + i];
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        double l = lam0[i];
        if (at_gamma[q]) {
            int rk = 0;
            for (int j = 0; j < nb; ++j) rk += (fabs(lam0[j]) < fabs(l) || (fabs(lam0[j]) == fabs(l) && j < i)) ? 1 : 0;
            if (rk < 3) l = 0.0;
        }
        lam[i] = l;
        om[i] = sqrt(fabs(l)) * (l >= 0 ? 1.0 : -1.0);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        int rk = 0;
        for (int j = 0; j < nb; ++j) rk += (om[j] < om[i] || (om[j] == om[i] && j < i)) ? 1 : 0;
        ord[rk] = i;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) {
        const int r = i / nb, c = i % nb;
        E[i] = V[(size_t)r * nb + ord[c]];
    }
    for (int i = threadIdx.x; i < nb; i += blockDim.x) lam0[i] = lam[ord[i]];
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        lam[i] = lam0[i];
        om[i] = sqrt(fabs(lam0[i])) * (lam0[i] >= 0 ? 1.0 : -1.0);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // groups of consecutive modes closer than lo_freqtol
        int n = 0;
        grp[0] = 0;
        for (int i = 1; i < nb; ++i)
            if (!(om[i] - om[i - 1] < 1e-9)) grp[++n] = i;
        grp[++n] = nb;
        ngrp = n;
    }
    // fixed weights of the gauge probe and of the phase fix, one per component row
    __shared__ double prw[128];
    __shared__ cd phw[128];
    for (int r = threadIdx.x; r < nb; r += blockDim.x) {
        const double kk = (double)(r + 1);
        prw[r] = fmod(kk * kGold, 1.0);
        const double amp = 1.0 + fmod(kk * kPlastic, 1.0);
        double sn, cs;
        sincos(0.9 * kk, &sn, &cs);
        phw[r] = {amp * cs, amp * sn};
    }
    __syncthreads();
    // canonical gauge inside degenerate groups: one thread per group (the groups own disjoint
    // columns of E and disjoint parts of the scratch)
    for (int g = threadIdx.x; g < ngrp; g += blockDim.x) {
        const int i0 = grp[g], k = grp[g + 1] - grp[g];
        if (k < 2) continue;
        cd *M = scr + (size_t)i0 * nb;
        cd *W = scr + (size_t)nb * nb + (size_t)i0 * nb;
        double *ev = s_ev + i0;
        int *so = s_so + i0;
        for (int a = 0; a < k; ++a)
            for (int b = 0; b < k; ++b) {
                cd s = {0, 0};
                for (int r = 0; r < nb; ++r)
                    s = cadd(s, cscale(cmulc(E[(size_t)r * nb + i0 + a], E[(size_t)r * nb + i0 + b]), prw[r]));
                M[a * k + b] = s;
            }
        for (int a = 0; a < k; ++a)
            for (int b = a + 1; b < k; ++b) {
                cd x = M[a * k + b], y = cconj(M[b * k + a]);
                M[a * k + b] = cscale(cadd(x, y), 0.5);
                M[b * k + a] = cconj(M[a * k + b]);
            }
        for (int a = 0; a < k; ++a) M[a * k + a].y = 0.0;
        small_eigh(k, M, W, ev, so);
        cd *tmp = scr + 2 * (size_t)nb * nb + i0;
        for (int r = 0; r < nb; ++r) {
            for (int b = 0; b < k; ++b) {
                cd s = {0, 0};
                for (int a = 0; a < k; ++a) s = cadd(s, cmul(E[(size_t)r * nb + i0 + a], W[a * k + so[b]]));
                tmp[b] = s;
            }
            for (int b = 0; b < k; ++b) E[(size_t)r * nb + i0 + b] = tmp[b];
        }
    }
    __syncthreads();
    // phase fix and chop
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        cd s = {0, 0};
        for (int r = 0; r < nb; ++r) s = cadd(s, cmul(phw[r], E[(size_t)r * nb + b]));
        const double a = sqrt(cabs2(s));
        const cd ph = {s.x / a, -s.y / a};
        for (int r = 0; r < nb; ++r) {
            cd e = cmul(E[(size_t)r * nb + b], ph);
            if (fabs(e.x) < 1e-13) e.x = 0.0;
            if (fabs(e.y) < 1e-13) e.y = 0.0;
            E[(size_t)r * nb + b] = e;
        }
    }
    __syncthreads();
    // Hellmann-Feynman: hf[al][b] = Re e_b^H gD_al e_b, partial sums over rows i per thread
    // (work items (i-block, b)) then a fixed-order sum; gD_al staged in LDS when E is
    const cd *g = gD + (size_t)q * 3 * nb * nb;
    constexpr int kIb = 8;
    __shared__ double hpart[kIb][128];
    const int nib = (nb + kIb - 1) / kIb;
    cd *gs = use_lds ? psm + (size_t)nb * nb : nullptr;
    for (int al = 0; al < 3; ++al) {
        const cd *ga = g + (size_t)al * nb * nb;
        if (use_lds) {
            for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) gs[i] = ga[i];
            ga = gs;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < kIb * nb; t += blockDim.x) {
            const int b = t % nb, ib = t / nb;
            double acc = 0;
            for (int i = ib * nib; i < min(nb, (ib + 1) * nib); ++i) {
                cd s = {0, 0};
                const cd *gr = ga + (size_t)i * nb;
                for (int j = 0; j < nb; ++j) s = cadd(s, cmul(gr[j], E[(size_t)j * nb + b]));
                const cd ei = E[(size_t)i * nb + b];
                acc += ei.x * s.x + ei.y * s.y;
            }
            hpart[ib][b] = acc;
        }
        __syncthreads();
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            double acc = 0;
            for (int ib = 0; ib < kIb; ++ib) acc += hpart[ib][b];
            hf[al][b] = acc;
        }
        __syncthreads();
    }
    if (use_lds)
        for (int i = threadIdx.x; i < nb * nb; i += blockDim.x) Eg[i] = E[i];
    __syncthreads();
    // return_degenerate_subspaces grouping (…_dynamicalmatrix.f90:911-980): mode i joins the
    // subspace of the first earlier mode within lo_freqtol, else starts one (s_first: that
    // earlier mode, then the subspace's first mode)
    int *first = s_first;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        int j = 0;
        while (j < i && !(fabs(om[j] - om[i]) < 1e-9)) ++j;
        first[i] = j < i ? j : -1;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < nb; ++i) first[i] = first[i] < 0 ? i : first[first[i]];
    __syncthreads();
    // velocities averaged over each subspace, its modes summed in order
    for (int s0 = threadIdx.x; s0 < nb; s0 += blockDim.x) {
        if (first[s0] != s0) continue;
        double m0 = 0, m1 = 0, m2 = 0;
        int cnt = 0;
        for (int i = s0; i < nb; ++i)
            if (first[i] == s0) { m0 += hf[0][i]; m1 += hf[1][i]; m2 += hf[2][i]; ++cnt; }
        for (int i = s0; i < nb; ++i)
            if (first[i] == s0) {
                double v[3] = {m0 / cnt, m1 / cnt, m2 / cnt};
                for (int al = 0; al < 3; ++al) {
                    double x = (om[i] > 1e-9) ? v[al] / (2.0 * om[i]) : 0.0;
                    vel[((size_t)q * nb + i) * 3 + al] = chop_dev(x, veltol);
                }
            }
    }
    for (int i = threadIdx.x; i < nb; i += blockDim.x) omega[(size_t)q * nb + i] = om[i];
}

// Full-grid eigenvectors e_full = Gamma(R) e_irr (conjugated under time reversal),
// velocities +-R v_irr.
__global__ void k_unfold(int nb, int na, const int *__restrict__ irr_of_full,
                         const int *__restrict__ op_of_full, const double *__restrict__ opm,
                         const int *__restrict__ src, const double *__restrict__ v0tab,
                         const double *__restrict__ qirr, const cd *__restrict__ egv_irr,
                         const double *__restrict__ vel_irr, const double *__restrict__ om_irr,
                         cd *__restrict__ egv_full, double *__restrict__ vel_full,
                         double *__restrict__ om_full) {
    const int q = blockIdx.x;
    const int k = irr_of_full[q];
    const int iop = op_of_full[q];
    const int o = (iop > 0 ? iop : -iop) - 1;
    const bool tr = iop < 0;
    const double *m = opm + 9 * o;
    const double *qi = qirr + 3 * k;
    const cd *Ei = egv_irr + (size_t)k * nb * nb;
    cd *Ef = egv_full + (size_t)q * nb * nb;
    for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
        const int row = idx / nb, b = idx % nb;
        const int a1 = row / 3, al = row % 3;
        const int a2 = src[o * na + a1];
        const double *v0 = v0tab + 3 * (o * na + a2);
        const double ph = -(v0[0] * qi[0] + v0[1] * qi[1] + v0[2] * qi[2]) * kTwoPi;
        double sn, cs;
        sincos(ph, &sn, &cs);
        cd s = {0, 0};
        for (int be = 0; be < 3; ++be) s = cadd(s, cscale(Ei[(size_t)(3 * a2 + be) * nb + b], m[3 * al + be]));
        s = cmul(s, {cs, sn});
        if (tr) s.y = -s.y;
        Ef[idx] = s;
    }
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        om_full[(size_t)q * nb + b] = om_irr[(size_t)k * nb + b];
        const double *v = vel_irr + ((size_t)k * nb + b) * 3;
        for (int al = 0; al < 3; ++al) {
            double x = m[3 * al] * v[0] + m[3 * al + 1] * v[1] + m[3 * al + 2] * v[2];
            vel_full[((size_t)q * nb + b) * 3 + al] = tr ? -x : x;
        }
    }
}

// First-order change of the harmonic properties of one irreducible point along a change dD of
// its dynamical matrix (and dG of dD/dq), workgroup per point, working arrays in global
// scratch (Y = dD E, then X = E^H dD E):
//   d lambda_n = mean over the degenerate group of n of Re X_mm (the group stays degenerate
//     under a symmetric change: E_S^H dD E_S is then a multiple of 1),
//   de_n = sum over m outside the group of n of e_m X_mn / (lambda_n - lambda_m),
//   d omega = d lambda / (2 omega), lambda = omega |omega|,
//   dv = mean over the group of [Re e^H dG e + 2 Re e^H G de] / (2 omega) - v d omega / omega,
// the derivative of the group-averaged Hellmann-Feynman velocity of k_post.  Rotations inside
// a degenerate group and phases of single modes are left out: the linewidths averaged over
// degenerate modes do not depend on them.  Modes below lo_freqtol get zero.
__global__ void k_eig_jvp(int nb, const cd *__restrict__ egv, const double *__restrict__ om,
                          const double *__restrict__ vel, const cd *__restrict__ gD, const cd *__restrict__ dD,
                          const cd *__restrict__ dgD, cd *__restrict__ scr, cd *__restrict__ de,
                          double *__restrict__ dom, double *__restrict__ dvel) {
    const int q = blockIdx.x;
    const size_t nn = (size_t)nb * nb;
    const cd *E = egv + (size_t)q * nn;
    const cd *A = dD + (size_t)q * nn;
    cd *Y = scr + (size_t)q * 2 * nn, *X = Y + nn;
    cd *dE = de + (size_t)q * nn;
    const double *w = om + (size_t)q * nb;
    __shared__ int gfirst[128], glast[128];
    __shared__ double dl[128], dw[128];
    for (int i = threadIdx.x; i < (int)nn; i += blockDim.x) {
        const int r = i / nb, b = i % nb;
        cd s = {0, 0};
        for (int c = 0; c < nb; ++c) s = cadd(s, cmul(A[(size_t)r * nb + c], E[(size_t)c * nb + b]));
        Y[i] = s;
    }
    if (threadIdx.x == 0) {
        // consecutive modes closer than lo_freqtol, as the gauge groups of k_post
        int g0 = 0;
        for (int i = 1; i <= nb; ++i)
            if (i == nb || !(w[i] - w[i - 1] < 1e-9)) {
                for (int k = g0; k < i; ++k) { gfirst[k] = g0; glast[k] = i; }
                g0 = i;
            }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < (int)nn; i += blockDim.x) {
        const int m = i / nb, n = i % nb;
        cd s = {0, 0};
        for (int r = 0; r < nb; ++r) s = cadd(s, cmulc(E[(size_t)r * nb + m], Y[(size_t)r * nb + n]));
        X[i] = s;
    }
    __syncthreads();
    for (int n = threadIdx.x; n < nb; n += blockDim.x) {
        double s = 0;
        for (int m = gfirst[n]; m < glast[n]; ++m) s += X[(size_t)m * nb + m].x;
        dl[n] = s / (glast[n] - gfirst[n]);
        dw[n] = (w[n] > 1e-9) ? dl[n] / (2.0 * w[n]) : 0.0;
        dom[(size_t)q * nb + n] = dw[n];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < (int)nn; i += blockDim.x) {
        const int c = i / nb, n = i % nb;
        cd s = {0, 0};
        if (w[n] > 1e-9) {
            const double ln = w[n] * fabs(w[n]);
            for (int m = 0; m < nb; ++m) {
                if (m >= gfirst[n] && m < glast[n]) continue;
                const double lm = w[m] * fabs(w[m]);
                s = cadd(s, cscale(cmul(E[(size_t)c * nb + m], X[(size_t)m * nb + n]), 1.0 / (ln - lm)));
            }
        }
        dE[i] = s;
    }
    __syncthreads();
    // group-averaged velocity derivative: one thread per (mode, direction) for the
    // Hellmann-Feynman terms, then the mean over the group
    __shared__ double dhf[3][128];
    for (int t = threadIdx.x; t < 3 * nb; t += blockDim.x) {
        const int al = t / nb, n = t % nb;
        const cd *G = gD + ((size_t)q * 3 + al) * nn, *dG = dgD + ((size_t)q * 3 + al) * nn;
        double a = 0, b = 0;
        for (int r = 0; r < nb; ++r) {
            cd g = {0, 0}, h = {0, 0};
            for (int c = 0; c < nb; ++c) {
                g = cadd(g, cmul(dG[(size_t)r * nb + c], E[(size_t)c * nb + n]));
                h = cadd(h, cmul(G[(size_t)r * nb + c], dE[(size_t)c * nb + n]));
            }
            const cd er = E[(size_t)r * nb + n];
            a += er.x * g.x + er.y * g.y;
            b += er.x * h.x + er.y * h.y;
        }
        dhf[al][n] = a + 2.0 * b;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * nb; t += blockDim.x) {
        const int al = t / nb, n = t % nb;
        double x = 0;
        if (w[n] > 1e-9) {
            double s = 0;
            for (int m = gfirst[n]; m < glast[n]; ++m) s += dhf[al][m];
            s /= (glast[n] - gfirst[n]);
            x = s / (2.0 * w[n]) - vel[((size_t)q * nb + n) * 3 + al] * dw[n] / w[n];
        }
        dvel[((size_t)q * nb + n) * 3 + al] = x;
    }
}

}  // namespace

int run_harmonic(tc3_ctx *c) {
    const int na = c->cr.na, nb = 3 * na;
    const int nirr = c->mesh.nirr, nfull = c->mesh.nfull;
    const int nsym = (int)c->ops.size();
    if (nb > 128) return c->fail(1, "at most 42 atoms per cell are supported by the harmonic kernels");
    hipStream_t st = c->stream;
    // host tables
    std::vector<double> qv(3 * nirr), qirr(3 * nirr);
    std::vector<int> atg(nirr);
    for (int k = 0; k < nirr; ++k) {
        const double *q = &c->mesh.qcart[3 * c->mesh.irr_full_index[k]];
        double n2 = 0;
        for (int i = 0; i < 3; ++i) {
            qv[3 * k + i] = chop_scalar(q[i] * kTwoPi, 1e-13);
            qirr[3 * k + i] = q[i];
            n2 += q[i] * q[i];
        }
        atg[k] = n2 < kSqTol ? 1 : 0;
    }
    std::vector<double> opm(9 * nsym), v0tab(3 * nsym * na);
    std::vector<int> src(nsym * na);
    for (int o = 0; o < nsym; ++o) {
        const SymOp &op = c->ops[o];
        std::memcpy(&opm[9 * o], op.m, 9 * sizeof(double));
        for (int a2 = 0; a2 < na; ++a2) {
            const int a1 = op.fmap[a2];
            src[o * na + a1] = a2;
            double d[3];
            for (int i = 0; i < 3; ++i) d[i] = c->cr.rcart[3 * a1 + i] - op.tr[i];
            // im = transpose(m)
            for (int i = 0; i < 3; ++i)
                v0tab[3 * (o * na + a2) + i] =
                    op.m[0 * 3 + i] * d[0] + op.m[1 * 3 + i] * d[1] + op.m[2 * 3 + i] * d[2] - c->cr.rcart[3 * a2 + i];
        }
    }
    // device buffers
    double *d_qv = nullptr, *d_lv = nullptr, *d_pm = nullptr, *d_ism = nullptr, *d_lam = nullptr;
    double *d_opm = nullptr, *d_v0 = nullptr, *d_qirr = nullptr;
    int *d_blk = nullptr, *d_stat = nullptr, *d_atg = nullptr, *d_src = nullptr, *d_opf = nullptr;
    cd *d_D = nullptr, *d_gD = nullptr, *d_V = nullptr, *d_scr = nullptr;
    const size_t nn = (size_t)nb * nb;
    const int npair = (int)c->p_a1.size();
    auto cleanup = [&]() {};
    hipError_t e = hipSuccess;
#define TRY(x) do { e = (x); if (e != hipSuccess) { cleanup(); return device_error(e, #x, __FILE__, __LINE__); } } while (0)
    // work buffers from the context's pool (kept across calls); the model's tables are uploaded
    // when the model changed since the last upload or a buffer is new
    bool fresh_any = false;
    auto work = [&](const char *name, void **p, size_t bytes) {
        bool fresh = false;
        const hipError_t r = c->pool_get(name, std::max<size_t>(bytes, 16), p, &fresh);
        fresh_any = fresh_any || fresh;
        return r;
    };
    TRY(work("h_qv", (void **)&d_qv, sizeof(double) * 3 * nirr));
    TRY(work("h_qirr", (void **)&d_qirr, sizeof(double) * 3 * nirr));
    TRY(work("h_lv", (void **)&d_lv, sizeof(double) * 3 * std::max(npair, 1)));
    TRY(work("h_pm", (void **)&d_pm, sizeof(double) * 9 * std::max(npair, 1)));
    TRY(work("h_ism", (void **)&d_ism, sizeof(double) * na));
    TRY(work("h_blk", (void **)&d_blk, sizeof(int) * (na * na + 1)));
    TRY(work("h_D", (void **)&d_D, sizeof(cd) * nn * nirr));
    TRY(work("h_gD", (void **)&d_gD, sizeof(cd) * 3 * nn * nirr));
    TRY(work("h_V", (void **)&d_V, sizeof(cd) * nn * nirr));
    TRY(work("h_lam", (void **)&d_lam, sizeof(double) * nb * nirr));
    TRY(work("h_stat", (void **)&d_stat, sizeof(int) * nirr));
    TRY(work("h_atg", (void **)&d_atg, sizeof(int) * nirr));
    TRY(work("h_scr", (void **)&d_scr, sizeof(cd) * (2 * nn + nb) * nirr));
    TRY(work("h_opm", (void **)&d_opm, sizeof(double) * 9 * nsym));
    TRY(work("h_v0", (void **)&d_v0, sizeof(double) * 3 * nsym * na));
    TRY(work("h_src", (void **)&d_src, sizeof(int) * nsym * na));
    TRY(work("h_opf", (void **)&d_opf, sizeof(int) * nfull));
    // persistent outputs, reallocated only when their sizes change
    if (c->harm_dims[0] != nb || c->harm_dims[1] != nirr || c->harm_dims[2] != nfull) {
        hipFree(c->d_om_irr); hipFree(c->d_om_full); hipFree(c->d_vel_irr); hipFree(c->d_vel_full);
        hipFree(c->d_egv_irr); hipFree(c->d_egv_full); hipFree(c->d_irr_of_full);
        c->d_om_irr = c->d_om_full = c->d_vel_irr = c->d_vel_full = nullptr;
        c->d_egv_irr = c->d_egv_full = nullptr;
        c->d_irr_of_full = nullptr;
        c->harm_dims[0] = c->harm_dims[1] = c->harm_dims[2] = -1;
        TRY(hipMalloc(&c->d_om_irr, sizeof(double) * nb * nirr));
        TRY(hipMalloc(&c->d_vel_irr, sizeof(double) * 3 * nb * nirr));
        TRY(hipMalloc(&c->d_egv_irr, sizeof(cd) * nn * nirr));
        TRY(hipMalloc(&c->d_om_full, sizeof(double) * nb * nfull));
        TRY(hipMalloc(&c->d_vel_full, sizeof(double) * 3 * nb * nfull));
        TRY(hipMalloc(&c->d_egv_full, sizeof(cd) * nn * nfull));
        TRY(hipMalloc(&c->d_irr_of_full, sizeof(int) * nfull));
        c->harm_dims[0] = nb; c->harm_dims[1] = nirr; c->harm_dims[2] = nfull;
        fresh_any = true;
    }
    if (fresh_any || c->harm_tables_gen != c->gen_model) {
        c->harm_tables_gen = -1;
        TRY(hipMemcpyAsync(d_qv, qv.data(), sizeof(double) * 3 * nirr, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(d_qirr, qirr.data(), sizeof(double) * 3 * nirr, hipMemcpyHostToDevice, st));
        if (npair) {
            TRY(hipMemcpyAsync(d_lv, c->p_lv.data(), sizeof(double) * 3 * npair, hipMemcpyHostToDevice, st));
            TRY(hipMemcpyAsync(d_pm, c->p_m.data(), sizeof(double) * 9 * npair, hipMemcpyHostToDevice, st));
        }
        TRY(hipMemcpyAsync(d_ism, c->cr.invsqrtmass.data(), sizeof(double) * na, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(d_blk, c->blk2_off.data(), sizeof(int) * (na * na + 1), hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(d_atg, atg.data(), sizeof(int) * nirr, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(d_opm, opm.data(), sizeof(double) * 9 * nsym, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(d_v0, v0tab.data(), sizeof(double) * 3 * nsym * na, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(d_src, src.data(), sizeof(int) * nsym * na, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(d_opf, c->mesh.op_of_full.data(), sizeof(int) * nfull, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(c->d_irr_of_full, c->mesh.irr_of_full.data(), sizeof(int) * nfull, hipMemcpyHostToDevice, st));
        c->harm_tables_gen = c->gen_model;
    }

    hipLaunchKernelGGL(k_dynmat, dim3(nirr), dim3(kBlock), 0, st, na, d_qv, d_blk, d_lv, d_pm, d_ism, d_D, d_gD);
    TRY(hipGetLastError());
    // matrices in LDS when they fit next to the kernels' static arrays
    const size_t jac_lds = 2 * sizeof(cd) * (size_t)nb * nb, post_lds = 2 * sizeof(cd) * (size_t)nb * nb;
    const int jac_use = jac_lds <= 144 * 1024, post_use = post_lds <= 96 * 1024;
    if (jac_use) {
        TRY(hipFuncSetAttribute((const void *)k_jacobi<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)jac_lds));
        hipLaunchKernelGGL(k_jacobi<true>, dim3(nirr), dim3(kJacBlock), jac_lds, st, nb, d_D, d_V, d_lam, 100, d_stat);
    } else {
        hipLaunchKernelGGL(k_jacobi<false>, dim3(nirr), dim3(kJacBlock), 0, st, nb, d_D, d_V, d_lam, 100, d_stat);
    }
    TRY(hipGetLastError());
    const double veltol = 1e-9 / kGroupvelAuToMs;
    if (post_use) {
        TRY(hipFuncSetAttribute((const void *)k_post<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)post_lds));
        hipLaunchKernelGGL(k_post<true>, dim3(nirr), dim3(kBlock), post_lds, st, nb, d_lam, d_V, d_gD, d_atg,
                           c->d_om_irr, c->d_egv_irr, c->d_vel_irr, d_scr, veltol);
    } else {
        hipLaunchKernelGGL(k_post<false>, dim3(nirr), dim3(kBlock), 0, st, nb, d_lam, d_V, d_gD, d_atg, c->d_om_irr,
                           c->d_egv_irr, c->d_vel_irr, d_scr, veltol);
    }
    TRY(hipGetLastError());
    hipLaunchKernelGGL(k_unfold, dim3(nfull), dim3(kBlock), 0, st, nb, na, c->d_irr_of_full, d_opf, d_opm,
                       d_src, d_v0, d_qirr, c->d_egv_irr, c->d_vel_irr, c->d_om_irr, c->d_egv_full,
                       c->d_vel_full, c->d_om_full);
    TRY(hipGetLastError());
    // solver status, frequencies and velocities down through the pinned buffer
    const size_t b_om = sizeof(double) * nb * nirr, b_vel = 3 * b_om;
    TRY(c->pinned(b_om + b_vel + sizeof(int) * nirr));
    double *h_om = (double *)c->pin, *h_vel = h_om + (size_t)nb * nirr;
    int *h_stat = (int *)(h_vel + (size_t)3 * nb * nirr);
    TRY(hipMemcpyAsync(h_stat, d_stat, sizeof(int) * nirr, hipMemcpyDeviceToHost, st));
    TRY(hipMemcpyAsync(h_om, c->d_om_irr, b_om, hipMemcpyDeviceToHost, st));
    TRY(hipMemcpyAsync(h_vel, c->d_vel_irr, b_vel, hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
#undef TRY
    std::vector<int> stat(h_stat, h_stat + nirr);
    c->om_irr.assign(h_om, h_om + (size_t)nb * nirr);
    c->vel_irr.assign(h_vel, h_vel + (size_t)3 * nb * nirr);
    // eigenvectors and full-grid results stay on the device; host copies are made on first use
    c->full_host_valid = false;
    for (int k = 0; k < nirr; ++k)
        if (stat[k]) return c->fail(2, "Jacobi eigen-solver did not converge");
    // degeneracy table (type_phonon_dispersions.f90:1190-1214)
    c->degen.assign(nirr, std::vector<std::vector<int>>(nb));
    c->omega_min = 1e300;
    for (int k = 0; k < nirr; ++k)
        for (int b = 0; b < nb; ++b) {
            const double w = c->om_irr[(size_t)k * nb + b];
            for (int b2 = 0; b2 < nb; ++b2)
                if (std::fabs(c->om_irr[(size_t)k * nb + b2] - w) < kFreqTol) c->degen[k][b].push_back(b2);
            if (std::fabs(w) > kFreqTol) c->omega_min = std::min(c->omega_min, w);
        }
    // rows
    c->row_q.clear(); c->row_b.clear();
    c->row_of.assign((size_t)nirr * nb, -1);
    for (int k = 0; k < nirr; ++k)
        for (int b = 0; b < nb; ++b)
            if (!(c->om_irr[(size_t)k * nb + b] < kFreqTol)) {
                c->row_of[(size_t)k * nb + b] = (int)c->row_q.size();
                c->row_q.push_back(k);
                c->row_b.push_back(b);
            }
    c->nrows = (int)c->row_q.size();
    return 0;
}

// Host copies of the full-grid frequencies, velocities and eigenvectors (on demand).
// Derivatives of the frequencies, velocities and eigenvectors of the irreducible and the full
// grid along a change dpm of the second-order force constants (9 per pair in the context's
// sorted order, rows as p_m), into pool buffers: the dynamical matrices of the model and of
// dpm (k_dynmat is linear in the force constants), the perturbation per irreducible point
// (k_eig_jvp), and the symmetry unfolding of k_unfold applied to the derivatives (it is
// linear in them).  The tables of the last harmonic step (q-points, pairs, operations) are
// reused from the pool.
int run_harmonic_jvp(tc3_ctx *c, const std::vector<double> &dpm, HarmonicJvp &out) {
    const int na = c->cr.na, nb = 3 * na, nirr = c->mesh.nirr, nfull = c->mesh.nfull;
    const int nsym = (int)c->ops.size(), npair = (int)c->p_a1.size();
    const size_t nn = (size_t)nb * nb;
    hipStream_t st = c->stream;
    hipError_t e = hipSuccess;
#define TRY(x) do { e = (x); if (e != hipSuccess) return device_error(e, #x, __FILE__, __LINE__); } while (0)
    auto work = [&](const char *name, void **p, size_t bytes) { return c->pool_get(name, std::max<size_t>(bytes, 16), p); };
    double *d_qv, *d_lv, *d_pm, *d_dpm, *d_ism, *d_opm, *d_v0, *d_qirr;
    int *d_blk, *d_src, *d_opf;
    cd *d_D, *d_gD, *d_dD, *d_dgD, *d_scr;
    TRY(work("h_qv", (void **)&d_qv, sizeof(double) * 3 * nirr));
    TRY(work("h_qirr", (void **)&d_qirr, sizeof(double) * 3 * nirr));
    TRY(work("h_lv", (void **)&d_lv, sizeof(double) * 3 * std::max(npair, 1)));
    TRY(work("h_pm", (void **)&d_pm, sizeof(double) * 9 * std::max(npair, 1)));
    TRY(work("h_ism", (void **)&d_ism, sizeof(double) * na));
    TRY(work("h_blk", (void **)&d_blk, sizeof(int) * (na * na + 1)));
    TRY(work("h_opm", (void **)&d_opm, sizeof(double) * 9 * nsym));
    TRY(work("h_v0", (void **)&d_v0, sizeof(double) * 3 * nsym * na));
    TRY(work("h_src", (void **)&d_src, sizeof(int) * nsym * na));
    TRY(work("h_opf", (void **)&d_opf, sizeof(int) * nfull));
    TRY(work("j_dpm", (void **)&d_dpm, sizeof(double) * 9 * std::max(npair, 1)));
    TRY(work("j_D", (void **)&d_D, sizeof(cd) * nn * nirr));
    TRY(work("j_gD", (void **)&d_gD, sizeof(cd) * 3 * nn * nirr));
    TRY(work("j_dD", (void **)&d_dD, sizeof(cd) * nn * nirr));
    TRY(work("j_dgD", (void **)&d_dgD, sizeof(cd) * 3 * nn * nirr));
    TRY(work("j_scr", (void **)&d_scr, sizeof(cd) * 2 * nn * nirr));
    TRY(work("j_de_irr", (void **)&out.de_irr, sizeof(cd) * nn * nirr));
    TRY(work("j_dom_irr", (void **)&out.dom_irr, sizeof(double) * nb * nirr));
    TRY(work("j_dv_irr", (void **)&out.dv_irr, sizeof(double) * 3 * nb * nirr));
    TRY(work("j_de_full", (void **)&out.de_full, sizeof(cd) * nn * nfull));
    TRY(work("j_dom_full", (void **)&out.dom_full, sizeof(double) * nb * nfull));
    TRY(work("j_dv_full", (void **)&out.dv_full, sizeof(double) * 3 * nb * nfull));
    if (npair) {
        TRY(hipMemcpyAsync(d_pm, c->p_m.data(), sizeof(double) * 9 * npair, hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(d_dpm, dpm.data(), sizeof(double) * 9 * npair, hipMemcpyHostToDevice, st));
    }
    hipLaunchKernelGGL(k_dynmat, dim3(nirr), dim3(kBlock), 0, st, na, d_qv, d_blk, d_lv, d_pm, d_ism, d_D, d_gD);
    hipLaunchKernelGGL(k_dynmat, dim3(nirr), dim3(kBlock), 0, st, na, d_qv, d_blk, d_lv, d_dpm, d_ism, d_dD, d_dgD);
    hipLaunchKernelGGL(k_eig_jvp, dim3(nirr), dim3(kBlock), 0, st, nb, c->d_egv_irr, c->d_om_irr, c->d_vel_irr, d_gD,
                       d_dD, d_dgD, d_scr, out.de_irr, out.dom_irr, out.dv_irr);
    hipLaunchKernelGGL(k_unfold, dim3(nfull), dim3(kBlock), 0, st, nb, na, c->d_irr_of_full, d_opf, d_opm, d_src, d_v0,
                       d_qirr, out.de_irr, out.dv_irr, out.dom_irr, out.de_full, out.dv_full, out.dom_full);
    TRY(hipGetLastError());
#undef TRY
    return 0;
}

int ensure_full_host(tc3_ctx *c) {
    if (c->full_host_valid) return 0;
    const int nb = c->nb, nfull = c->mesh.nfull, nirr = c->mesh.nirr;
    const size_t nn = (size_t)nb * nb;
    c->om_full.resize((size_t)nb * nfull);
    c->vel_full.resize((size_t)3 * nb * nfull);
    c->egv_full.resize(nn * nfull);
    c->egv_irr.resize(nn * nirr);
    hipError_t e = hipMemcpyAsync(c->egv_irr.data(), c->d_egv_irr, sizeof(cd) * nn * nirr, hipMemcpyDeviceToHost,
                                  c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->om_full.data(), c->d_om_full, sizeof(double) * nb * nfull, hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->vel_full.data(), c->d_vel_full, sizeof(double) * 3 * nb * nfull, hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->egv_full.data(), c->d_egv_full, sizeof(cd) * nn * nfull, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return device_error(e, "full-grid harmonic results to the host", __FILE__, __LINE__);
    c->full_host_valid = true;
    return 0;
}

}  // namespace tc3
