// |Psi|^2 contraction on MFMA for band counts that are multiples of 16.
//
// Two kernels per batch of triplets (t = q1, q2, q3):
//
//   k_uw   workgroup per (t, ia):  builds the row Phi(q2,q3)[ia][ib][ic] in LDS from the
//          third-order IFC list (pretransform_phi3, scattering_threephonon.f90:275-317),
//          then U = Phi_ia E3 and W_ia = E2^T U on MFMA; writes W[t][ia][b2][b3].
//   k_psi  workgroup per (t, 16 values of b1): Psi[b1][b2][b3] = sum_ia E1[ia][b1] W[ia][b2][b3]
//          on MFMA, chunk by chunk of b2, fused with the smeared energy conservation and
//          occupations of every temperature (get_dirac :219-271, :113-123) and the
//          linewidth / scattering-matrix accumulation (:126-135).  Reductions are done in
//          a fixed order in LDS, so results do not depend on scheduling.
//
// Complex products are formed from four real MFMA products (re = ar br - ai bi,
// im = ar bi + ai br).  Precision modes:
//   FP64    v_mfma_f64_16x16x4 on FP64 operands
//   FP32    v_mfma_f32_16x16x4 on FP32 operands, FP32 accumulation
//   BF16X3  v_mfma_f32_16x16x32_bf16 with every FP32 operand split as hi + lo (two bf16);
//           a.b ~ ah.bh + ah.bl + al.bh, FP32 accumulation
// Lane layouts (16x16 output tile, 64 lanes): A[m = l%16][k], B[k][n = l%16] with
// k = k0 + l/16 (K-step 4) or k = k0 + 8*(l/16) + j (K-step 32, j < 8), and
// D[m][n = l%16] in register r < 4 with m = 4*(l/16) + r (FP32 accumulation) or
// m = 4*r + l/16 (FP64); both checked on the device by an in-bounds probe.
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

template <int PREC> struct Traits;
template <> struct Traits<0> { typedef double T; typedef v4d Acc; static constexpr int KS = 4; };
template <> struct Traits<1> { typedef float T; typedef v4f Acc; static constexpr int KS = 4; };
template <> struct Traits<2> { typedef float T; typedef v4f Acc; static constexpr int KS = 32; };

// Fragment of one real operand for one K-step.
template <int PREC> struct Frag;
template <> struct Frag<0> { double v; };
template <> struct Frag<1> { float v; };
template <> struct Frag<2> { v8bf h, l; };

template <int PREC>
__device__ inline void mma(typename Traits<PREC>::Acc &c, const Frag<PREC> &a, const Frag<PREC> &b);
template <>
__device__ inline void mma<0>(v4d &c, const Frag<0> &a, const Frag<0> &b) {
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a.v, b.v, c, 0, 0, 0);
}
template <>
__device__ inline void mma<1>(v4f &c, const Frag<1> &a, const Frag<1> &b) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a.v, b.v, c, 0, 0, 0);
}
template <>
__device__ inline void mma<2>(v4f &c, const Frag<2> &a, const Frag<2> &b) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.h, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.h, b.l, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.l, b.h, c, 0, 0, 0);
}

// Output row of register r in lane l: 4*(l/16) + r for the FP32-accumulating
// instructions, 4*r + l/16 for v_mfma_f64_16x16x4_f64.
template <int PREC>
__device__ inline int out_row(int lane, int r) {
    if constexpr (PREC == 0) return 4 * r + (lane >> 4);
    else return 4 * (lane >> 4) + r;
}

template <int PREC>
__device__ inline Frag<PREC> neg(const Frag<PREC> &a) {
    Frag<PREC> r;
    if constexpr (PREC == 2) {
        r.h = -a.h;
        r.l = -a.l;
    } else {
        r.v = -a.v;
    }
    return r;
}

// Load the A fragment of a K-step: element (m, k) from accessor fa(m, k) -> (re, im).
template <int PREC, typename F>
__device__ inline void load_a(int lane, int m0, int k0, int K, F fa, Frag<PREC> &re, Frag<PREC> &im) {
    const int m = m0 + (lane & 15);
    if constexpr (PREC == 2) {
        const int kb = k0 + 8 * (lane >> 4);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float x = 0, y = 0;
            if (kb + j < K) fa(m, kb + j, x, y);
            const __bf16 xh = (__bf16)x, yh = (__bf16)y;
            re.h[j] = xh;
            re.l[j] = (__bf16)(x - (float)xh);
            im.h[j] = yh;
            im.l[j] = (__bf16)(y - (float)yh);
        }
    } else {
        const int k = k0 + (lane >> 4);
        typename Traits<PREC>::T x = 0, y = 0;
        if (k < K) fa(m, k, x, y);
        re.v = x;
        im.v = y;
    }
}

template <int PREC, typename F>
__device__ inline void load_b(int lane, int k0, int n0, int K, F fb, Frag<PREC> &re, Frag<PREC> &im) {
    const int n = n0 + (lane & 15);
    if constexpr (PREC == 2) {
        const int kb = k0 + 8 * (lane >> 4);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float x = 0, y = 0;
            if (kb + j < K) fb(kb + j, n, x, y);
            const __bf16 xh = (__bf16)x, yh = (__bf16)y;
            re.h[j] = xh;
            re.l[j] = (__bf16)(x - (float)xh);
            im.h[j] = yh;
            im.l[j] = (__bf16)(y - (float)yh);
        }
    } else {
        const int k = k0 + (lane >> 4);
        typename Traits<PREC>::T x = 0, y = 0;
        if (k < K) fb(k, n, x, y);
        re.v = x;
        im.v = y;
    }
}

// acc_re += ar br - ai bi ; acc_im += ar bi + ai br
template <int PREC>
__device__ inline void cmma(typename Traits<PREC>::Acc &cr, typename Traits<PREC>::Acc &ci, const Frag<PREC> &ar,
                            const Frag<PREC> &ai, const Frag<PREC> &br, const Frag<PREC> &bi) {
    mma<PREC>(cr, ar, br);
    mma<PREC>(cr, neg<PREC>(ai), bi);
    mma<PREC>(ci, ar, bi);
    mma<PREC>(ci, ai, br);
}

// C(nb x nb) = A(nb x nb) B(nb x nb), complex, 16x16 tiles spread over the 4 waves.
// fa/fb are element accessors, fc(m, n, re, im) stores one result.
template <int PREC, typename FA, typename FB, typename FC>
__device__ inline void cgemm_nb(int nb, FA fa, FB fb, FC fc) {
    typedef typename Traits<PREC>::Acc Acc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nt = nb / 16;
    for (int tile = wave; tile < nt * nt; tile += 4) {
        const int m0 = (tile / nt) * 16, n0 = (tile % nt) * 16;
        Acc cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
        for (int k0 = 0; k0 < nb; k0 += Traits<PREC>::KS) {
            Frag<PREC> ar, ai, br, bi;
            load_a<PREC>(lane, m0, k0, nb, fa, ar, ai);
            load_b<PREC>(lane, k0, n0, nb, fb, br, bi);
            cmma<PREC>(cr, ci, ar, ai, br, bi);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) fc(m0 + out_row<PREC>(lane, r), n0 + (lane & 15), cr[r], ci[r]);
    }
}

template <int PREC>
__global__ void __launch_bounds__(256) k_uw(int na, int t0, const int *__restrict__ tq2, const int *__restrict__ tq3,
                                            const double *__restrict__ qcart, const int *__restrict__ a1_off,
                                            const int *__restrict__ blk_off, const double *__restrict__ lv2,
                                            const double *__restrict__ lv3, const double *__restrict__ mwm,
                                            const typename Traits<PREC>::T *__restrict__ Ef_re,
                                            const typename Traits<PREC>::T *__restrict__ Ef_im,
                                            typename Traits<PREC>::T *__restrict__ W) {
    typedef typename Traits<PREC>::T T;
    extern __shared__ __align__(16) unsigned char smem[];
    const int nb = 3 * na;
    const int tl = blockIdx.x / nb, ia = blockIdx.x % nb;
    const int t = t0 + tl;
    const int a1 = ia / 3, i = ia % 3;
    const size_t nb2 = (size_t)nb * nb;
    T *phr = (T *)smem, *phi = phr + nb2, *ur = phi + nb2, *ui = ur + nb2;
    cd *ph = (cd *)(ui + nb2);
    const int q2 = tq2[t], q3 = tq3[t];
    const double *Q2 = qcart + 3 * q2, *Q3 = qcart + 3 * q3;
    const int f0 = a1_off[a1], f1 = a1_off[a1 + 1];
    for (int f = f0 + threadIdx.x; f < f1; f += blockDim.x) {
        const double arg = -kTwoPi * (Q2[0] * lv2[3 * f] + Q2[1] * lv2[3 * f + 1] + Q2[2] * lv2[3 * f + 2] +
                                      Q3[0] * lv3[3 * f] + Q3[1] * lv3[3 * f + 1] + Q3[2] * lv3[3 * f + 2]);
        double s, c;
        sincos(arg, &s, &c);
        ph[f - f0] = {c, s};
    }
    __syncthreads();
    for (int e = threadIdx.x; e < (int)nb2; e += blockDim.x) {
        const int ib = e / nb, ic = e % nb;
        const int blk = (a1 * na + ib / 3) * na + ic / 3;
        const int j = ib % 3, k = ic % 3;
        double ax = 0, ay = 0;
        for (int f = blk_off[blk]; f < blk_off[blk + 1]; ++f) {
            const double m = mwm[27 * f + 9 * i + 3 * j + k];
            ax += m * ph[f - f0].x;
            ay += m * ph[f - f0].y;
        }
        phr[e] = (T)ax;
        phi[e] = (T)ay;
    }
    __syncthreads();
    const T *E3r = Ef_re + (size_t)q3 * nb2, *E3i = Ef_im + (size_t)q3 * nb2;
    const T *E2r = Ef_re + (size_t)q2 * nb2, *E2i = Ef_im + (size_t)q2 * nb2;
    // U[ib][b3] = sum_ic Phi[ib][ic] E3[ic][b3]
    cgemm_nb<PREC>(
        nb, [&](int m, int k, T &x, T &y) { x = phr[m * nb + k]; y = phi[m * nb + k]; },
        [&](int k, int n, T &x, T &y) { x = E3r[k * nb + n]; y = E3i[k * nb + n]; },
        [&](int m, int n, T x, T y) { ur[m * nb + n] = x; ui[m * nb + n] = y; });
    __syncthreads();
    // W[b2][b3] = sum_ib E2[ib][b2] U[ib][b3]
    T *Wr = W + (size_t)tl * 2 * nb * nb2 + (size_t)ia * nb2;
    T *Wi = Wr + nb * nb2;
    cgemm_nb<PREC>(
        nb, [&](int m, int k, T &x, T &y) { x = E2r[k * nb + m]; y = E2i[k * nb + m]; },
        [&](int k, int n, T &x, T &y) { x = ur[k * nb + n]; y = ui[k * nb + n]; },
        [&](int m, int n, T x, T y) { Wr[m * nb + n] = x; Wi[m * nb + n] = y; });
}

__device__ inline double gaussd(double x, double mu, double sigma) {
    const double d = x - mu;
    if (fabs(d) > 8.0 * sigma) return 0.0;
    const double inv = 1.0 / sigma;
    return exp(-d * d * 0.5 * inv * inv) * 0.398942280401433 * inv;
}

// Staging/accumulation type of the epilogue: FP64 in FP64 mode, FP32 otherwise.
template <int PREC> struct Stage { typedef float T; };
template <> struct Stage<0> { typedef double T; };

template <int PREC>
__global__ void __launch_bounds__(256) k_psi(MfmaEpiArgs a, const typename Traits<PREC>::T *__restrict__ E1_re,
                                             const typename Traits<PREC>::T *__restrict__ E1_im,
                                             const typename Traits<PREC>::T *__restrict__ W) {
    typedef typename Traits<PREC>::T T;
    typedef typename Traits<PREC>::Acc Acc;
    typedef typename Stage<PREC>::T S;
    extern __shared__ __align__(16) unsigned char smem[];
    const int nb = a.nb, nT = a.nT, CB = a.cb;
    const int nbp = nb + 1;                         // padded row stride of the staging arrays
    const int nbb = nb / 16;
    // XCD-aware order: the b1 blocks of one triplet share blockIdx % 8
    const int g = blockIdx.x;
    const int tl = (g / (8 * nbb)) * 8 + (g % 8);
    const int bb = (g / 8) % nbb;
    if (tl >= a.nt) return;
    const int t = a.t0 + tl;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q1 = a.tq1[t], q2 = a.tq2[t], q3 = a.tq3[t];
    const int i2 = a.irr_of_full[q2], i3 = a.irr_of_full[q3];
    const double perm = a.perm[t];
    const int e0 = a.eq_off[t], e1 = a.eq_off[t + 1];
    const size_t nb2 = (size_t)nb * nb;
    // LDS layout
    S *A = (S *)smem;                               // [16][CB][nbp]  psisq*d0 (masked)
    S *B = A + 16 * CB * nbp;                       // [16][CB][nbp]  psisq*d1
    S *S2 = B + 16 * CB * nbp;                      // [nT][16][nb]   sum over b3
    S *S3 = S2 + nT * 16 * nb;                      // [nT][16][nb]   sum over b2
    double *n2t = (double *)(((size_t)(S3 + nT * 16 * nb) + 15) & ~(size_t)15);  // [nT][nb]
    double *n3t = n2t + nT * nb;                    // [nT][nb]
    double *w1 = n3t + nT * nb;                     // 16
    double *sg1 = w1 + 16;                          // 16
    int *row = (int *)(sg1 + 16);                   // 16
    for (int x = threadIdx.x; x < nT * nb; x += blockDim.x) {
        const int T_ = x / nb, b = x % nb;
        n2t[x] = a.be[((size_t)T_ * a.nirr + i2) * nb + b];
        n3t[x] = a.be[((size_t)T_ * a.nirr + i3) * nb + b];
    }
    for (int x = threadIdx.x; x < nT * 16 * nb; x += blockDim.x) S3[x] = 0;
    if (threadIdx.x < 16) {
        const int b1 = 16 * bb + threadIdx.x;
        w1[threadIdx.x] = a.om_irr[(size_t)q1 * nb + b1];
        sg1[threadIdx.x] = a.sigsq[(size_t)q1 * nb + b1];
        row[threadIdx.x] = a.row_of[(size_t)q1 * nb + b1];
    }
    __syncthreads();
    const T *E1r = E1_re + (size_t)q1 * nb2, *E1i = E1_im + (size_t)q1 * nb2;
    const T *Wr = W + (size_t)tl * 2 * nb * nb2;
    const T *Wi = Wr + nb * nb2;
    const double *om2 = a.om_full + (size_t)q2 * nb, *om3 = a.om_full + (size_t)q3 * nb;
    const double *sgq2 = a.sigsq + (size_t)i2 * nb, *sgq3 = a.sigsq + (size_t)i3 * nb;
    const int ntile = CB * nb / 16;
    for (int c0 = 0; c0 < nb; c0 += CB) {
        // Psi[b1][(b2,b3)] for b2 in [c0, c0+CB): tiles of 16 consecutive b3
        for (int tile = wave; tile < ntile; tile += 4) {
            const int n0 = c0 * nb + tile * 16;
            Acc cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
            for (int k0 = 0; k0 < nb; k0 += Traits<PREC>::KS) {
                Frag<PREC> ar, ai, br, bi;
                load_a<PREC>(lane, 16 * bb, k0, nb,
                             [&](int m, int k, T &x, T &y) { x = E1r[k * nb + m]; y = E1i[k * nb + m]; }, ar, ai);
                load_b<PREC>(lane, k0, n0, nb,
                             [&](int k, int n, T &x, T &y) { x = Wr[(size_t)k * nb2 + n]; y = Wi[(size_t)k * nb2 + n]; },
                             br, bi);
                cmma<PREC>(cr, ci, ar, ai, br, bi);
            }
            const int n = n0 + (lane & 15);
            const int b2 = n / nb, b3 = n % nb;
            const double w2 = om2[b2], w3 = om3[b3];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = out_row<PREC>(lane, r);       // local b1
                const double w1v = w1[m];
                double va = 0.0, vb = 0.0;
                if (row[m] >= 0 && !(w2 < 1e-9) && !(fabs(w1v - w2) < 1e-9) && !(w3 < 1e-9) &&
                    !(fabs(w1v - w3) < 1e-9) && !(fabs(w2 - w3) < 1e-9)) {
                    const double re = (double)cr[r], im = (double)ci[r];
                    const double psisq = a.psi_scale * (re * re + im * im) * perm;
                    const double sig = (a.integrationtype == 2) ? sqrt(sg1[m] + sgq2[b2] + sgq3[b3]) : a.sigma_fixed;
                    const double d0 = gaussd(w1v, -w2 + w3, sig) - gaussd(w1v, w2 - w3, sig);
                    const double d1 = gaussd(w1v, w2 + w3, sig) - gaussd(w1v, -w2 - w3, sig);
                    va = psisq * d0;
                    vb = psisq * d1;
                }
                const int loc = (m * CB + (b2 - c0)) * nbp + b3;
                A[loc] = (S)va;
                B[loc] = (S)vb;
            }
        }
        __syncthreads();
        // S2[T][b1][b2] = sum_b3 f  (work item per (T, b1, b2 of the chunk))
        for (int x = threadIdx.x; x < nT * 16 * CB; x += blockDim.x) {
            const int T_ = x / (16 * CB), mj = x % (16 * CB);
            const int m = mj / CB, j = mj % CB, b2 = c0 + j;
            const double *n3 = n3t + T_ * nb;
            const double n2 = n2t[T_ * nb + b2];
            const S *pa = A + (size_t)mj * nbp, *pb = B + (size_t)mj * nbp;
            S s = 0;
            for (int b3 = 0; b3 < nb; ++b3)
                s += (S)(n2 - n3[b3]) * pa[b3] + (S)(n2 + n3[b3] + 1.0) * pb[b3];
            S2[(T_ * 16 + m) * nb + b2] = s;
        }
        // S3[T][b1][b3] += sum_{b2 in chunk} f  (work item per (T, b1, b3))
        for (int x = threadIdx.x; x < nT * 16 * nb; x += blockDim.x) {
            const int T_ = x / (16 * nb), mb = x % (16 * nb);
            const int m = mb / nb, b3 = mb % nb;
            const double *n2 = n2t + T_ * nb;
            const double n3 = n3t[T_ * nb + b3];
            S s = 0;
            for (int j = 0; j < CB; ++j) {
                const int loc = (m * CB + j) * nbp + b3;
                s += (S)(n2[c0 + j] - n3) * A[loc] + (S)(n2[c0 + j] + n3 + 1.0) * B[loc];
            }
            S3[x] += s;
        }
        __syncthreads();
    }
    // linewidth partials (g0 += f0 + f1 per equivalent triplet, :126-128)
    for (int x = threadIdx.x; x < nT * 16; x += blockDim.x) {
        const int T_ = x / 16, m = x % 16;
        double s = 0;
        for (int b2 = 0; b2 < nb; ++b2) s += (double)S2[(T_ * 16 + m) * nb + b2];
        a.gpart[((size_t)t * nT + T_) * nb + 16 * bb + m] = (row[m] >= 0) ? s * (double)(e1 - e0) : 0.0;
    }
    // scattering-matrix entries of this (triplet, b1 block) (:133-134).  Every target
    // (row, q', band) belongs to exactly one triplet, and inside a triplet q' appears in
    // one equivalent pair only, so the values are stored, not accumulated.
    const int neq = e1 - e0;
    for (long x = threadIdx.x; x < (long)nT * 16 * neq * nb; x += blockDim.x) {
        const int b = (int)(x % nb);
        long y = x / nb;
        const int e = e0 + (int)(y % neq);
        y /= neq;
        const int m = (int)(y % 16), T_ = (int)(y / 16);
        const int r = row[m];
        if (r < 0) continue;
        double *xr = a.xi + (size_t)T_ * a.xi_T + (size_t)r * a.xi_row;
        const int x2 = a.eq_q2[e], x3 = a.eq_q3[e];
        const double v2 = 2.0 * (double)S2[(T_ * 16 + m) * nb + b] * om2[b] / w1[m];
        const double v3 = 2.0 * (double)S3[(T_ * 16 + m) * nb + b] * om3[b] / w1[m];
        if (x2 == x3) {
            xr[(size_t)x2 * nb + b] = v2 + v3;
        } else {
            xr[(size_t)x2 * nb + b] = v2;
            xr[(size_t)x3 * nb + b] = v3;
        }
    }
}

template <int PREC>
__global__ void k_scale_planar(long n, int nb, const cd *__restrict__ egv, const double *__restrict__ om,
                               typename Traits<PREC>::T *__restrict__ re, typename Traits<PREC>::T *__restrict__ im) {
    typedef typename Traits<PREC>::T T;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long q = i / ((long)nb * nb);
        const int b = (int)(i % nb);
        const double w = om[q * nb + b];
        const double s = (w < 1e-9) ? 0.0 : 1.0 / sqrt(w);
        re[i] = (T)(egv[i].x * s);
        im[i] = (T)(egv[i].y * s);
    }
}

template <int PREC>
int run_mfma(const MfmaPlan &p, hipStream_t st) {
    typedef typename Traits<PREC>::T T;
    const int nb = p.nb;
    const size_t nb2 = (size_t)nb * nb;
    hipLaunchKernelGGL(k_scale_planar<PREC>, dim3(1024), dim3(256), 0, st, (long)p.nirr * nb2, nb, p.egv_irr, p.om_irr,
                       (T *)p.E1re, (T *)p.E1im);
    hipLaunchKernelGGL(k_scale_planar<PREC>, dim3(4096), dim3(256), 0, st, (long)p.nfull * nb2, nb, p.egv_full, p.om_full,
                       (T *)p.Efre, (T *)p.Efim);
    const size_t sh_uw = sizeof(T) * 4 * nb2 + sizeof(cd) * p.maxa1;
    const size_t sh_psi = p.psi_shmem;
    hipFuncSetAttribute((const void *)k_uw<PREC>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh_uw);
    hipFuncSetAttribute((const void *)k_psi<PREC>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh_psi);
    MfmaEpiArgs ea = p.epi;
    const int nbb = nb / 16;
    for (long t0 = 0; t0 < p.ntrip; t0 += p.batch) {
        const int nt = (int)std::min<long>(p.batch, p.ntrip - t0);
        hipEvent_t *ev = p.ev ? p.ev + 3 * (t0 / p.batch) : nullptr;
        if (ev) hipEventRecord(ev[0], st);
        hipLaunchKernelGGL(k_uw<PREC>, dim3(nt * nb), dim3(256), sh_uw, st, p.na, (int)t0, p.tq2, p.tq3, p.qcart,
                           p.a1off, p.blk3, p.lv2, p.lv3, p.mwm, (const T *)p.Efre, (const T *)p.Efim, (T *)p.work);
        if (ev) hipEventRecord(ev[1], st);
        ea.t0 = (int)t0;
        ea.nt = nt;
        const int ngrp = (nt + 7) / 8;
        hipLaunchKernelGGL(k_psi<PREC>, dim3(ngrp * 8 * nbb), dim3(256), sh_psi, st, ea, (const T *)p.E1re,
                           (const T *)p.E1im, (const T *)p.work);
        if (ev) hipEventRecord(ev[2], st);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return device_error(e, "MFMA contraction launch", __FILE__, __LINE__);
    }
    return 0;
}

}  // namespace

size_t mfma_elem_size(int precision) { return precision == TC3_PREC_FP64 ? sizeof(double) : sizeof(float); }

bool mfma_supported(int nb, int precision) {
    if (nb % 16 != 0) return false;
    const size_t es = mfma_elem_size(precision);
    return 4 * (size_t)nb * nb * es + 16 * 4096 <= 160 * 1024;
}

int mfma_choose_chunk(int nb, int nT, int precision, size_t *shmem) {
    const size_t es = (precision == TC3_PREC_FP64) ? sizeof(double) : sizeof(float);
    const size_t fixed = es * (2 * (size_t)nT * 16 * nb) + 16 +
                         sizeof(double) * (2 * (size_t)nT * nb + 32) + 16 * sizeof(int);
    // largest chunk of b2 that keeps the workgroup within 80 KB (two per CU), else 150 KB
    for (size_t cap : {(size_t)80 * 1024, (size_t)150 * 1024})
        for (int cb : {16, 8, 4, 2, 1}) {
            if (nb % cb) continue;
            const size_t s = fixed + es * (2 * 16 * (size_t)cb * (nb + 1));
            if (s <= cap) {
                *shmem = s;
                return cb;
            }
        }
    return 0;
}

int run_mfma_chain(const MfmaPlan &p, int precision, hipStream_t st) {
    switch (precision) {
    case TC3_PREC_FP64: return run_mfma<0>(p, st);
    case TC3_PREC_FP32: return run_mfma<1>(p, st);
    default: return run_mfma<2>(p, st);
    }
}

}  // namespace tc3
