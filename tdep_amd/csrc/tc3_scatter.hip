// Three-phonon scattering on the GPU.
//
// Per visited triplet (q1 irreducible, q2, q3 = -q1-q2) the engine forms
//   Phi(q2,q3)[ia,ib,ic] = sum_t mwm_t e^{-2 pi i (q2.lv2_t + q3.lv3_t)}     (pretransform_phi3,
//                                                   scattering_threephonon.f90:275-317)
//   Psi[b1,b2,b3] = sum E1[ia,b1] E2[ib,b2] E3[ic,b3] Phi[ia,ib,ic]             (:88-113)
// for all b1 at once as three batched complex GEMMs, then weights |Psi|^2 with the
// Gaussian-smeared energy conservation and the occupations of every temperature
// (get_dirac :219-271, :113-123) and accumulates the linewidth and the off-diagonal
// scattering-matrix rows (:126-135).  The off-diagonal rows are symmetrised over the
// little group of q1 (:143-205, scattering.f90:296-357).
//
// Each (row, q') target of the scattering matrix receives contributions from exactly
// one triplet (checked on the host), so rows are accumulated without atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "tc3_chain.hpp"
#include "tc3_ctx.hpp"
#include "tc3_fourphonon.hpp"
#include "tc3_fused.hpp"
#include "tc3_mfma.hpp"

namespace tc3 {

namespace {

struct EpiArgs {
    int nb, nT, t0, nfull, integrationtype;
    double sigma_fixed, psi_scale;  // pi/16 / N_q
    const int *tq1, *tq2, *tq3, *eq_off, *eq_q2, *eq_q3, *irr_of_full, *row_of;
    const double *perm, *om_irr, *om_full, *be, *sigsq;
    void *xi;        // [T][row][nfull*nb], entries of the epilogue's type P
    double *gpart;   // [t][T][nb]
    size_t xi_row, xi_T;
    int nirr;
    const double *u_irr, *u_full;   // integration type 6: reduced group velocities [q][mode][3]
    double thresh6;                 // and the baseline width
    const float *tw;                // Monte-Carlo draw counts [t][b1], or null (all 1)
};

// Reduced group velocities u_j = sum_i v_i reclat_ij / dims_j of every mode (the operand of
// the integration type 6 widths, scattering.f90:138-140).
struct Mat3 { double m[9]; };
__global__ void k_reduced_vel(long nmode, const double *__restrict__ vel, Mat3 r, double *__restrict__ u) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nmode; i += (long)gridDim.x * blockDim.x) {
        const double *v = vel + 3 * i;
        for (int j = 0; j < 3; ++j) u[3 * i + j] = v[0] * r.m[j] + v[1] * r.m[3 + j] + v[2] * r.m[6 + j];
    }
}

// |Psi|^2 -> smeared-delta weighted rates for every temperature; linewidth partials and
// scattering-matrix rows.  Workgroup per (triplet, b1).  psi_ld: stride of b1 rows.
template <typename P>
__global__ void k_epilogue(EpiArgs a, const P *__restrict__ psi, int psi_complex_planar) {
    extern __shared__ double sm[];
    const int nb = a.nb;
    double *fs = sm;               // nb*nb
    double *s2 = fs + nb * nb;     // nb
    double *s3 = s2 + nb;          // nb
    const int tl = blockIdx.x / nb, b1 = blockIdx.x % nb;
    const int t = a.t0 + tl;
    const int q1 = a.tq1[t], q2 = a.tq2[t], q3 = a.tq3[t];
    const int row = a.row_of[q1 * nb + b1];
    if (row < 0) {
        for (int T = 0; T < a.nT && threadIdx.x == 0; ++T) a.gpart[((size_t)t * a.nT + T) * nb + b1] = 0.0;
        return;
    }
    const int i2 = a.irr_of_full[q2], i3 = a.irr_of_full[q3];
    const double w1 = a.om_irr[q1 * nb + b1];
    const double perm = a.perm[t] * (a.tw ? (double)a.tw[(size_t)t * nb + b1] : 1.0);
    const double sg1 = a.sigsq[q1 * nb + b1];
    const int e0 = a.eq_off[t], e1 = a.eq_off[t + 1];
    const size_t base = ((size_t)tl * nb + b1) * nb * nb;
    for (int T = 0; T < a.nT; ++T) {
        const double *be = a.be + (size_t)T * a.nirr * nb;
        for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
            const int b2 = e / nb, b3 = e % nb;
            const double w2 = a.om_full[(size_t)q2 * nb + b2], w3 = a.om_full[(size_t)q3 * nb + b3];
            double f = 0.0;
            if (!(w2 < 1e-9) && !(fabs(w1 - w2) < 1e-9) && !(w3 < 1e-9) && !(fabs(w1 - w3) < 1e-9) &&
                !(fabs(w2 - w3) < 1e-9)) {
                double re, im;
                if (psi_complex_planar) {
                    re = (double)psi[base * 2 + e];
                    im = (double)psi[base * 2 + nb * nb + e];
                } else {
                    re = (double)psi[(base + e) * 2];
                    im = (double)psi[(base + e) * 2 + 1];
                }
                const double psisq = a.psi_scale * (re * re + im * im);
                double sig;
                if (a.integrationtype == 2)
                    sig = sqrt(sg1 + a.sigsq[i2 * nb + b2] + a.sigsq[i3 * nb + b3]);
                else if (a.integrationtype == 6)
                    sig = sigma_type6(a.u_irr + ((size_t)q1 * nb + b1) * 3, a.u_full + ((size_t)q2 * nb + b2) * 3,
                                      a.u_full + ((size_t)q3 * nb + b3) * 3, a.thresh6);
                else
                    sig = a.sigma_fixed;
                const double d0 = gauss(w1, -w2 + w3, sig) - gauss(w1, w2 - w3, sig);
                const double d1 = gauss(w1, w2 + w3, sig) - gauss(w1, -w2 - w3, sig);
                const double n2 = be[i2 * nb + b2], n3 = be[i3 * nb + b3];
                f = perm * psisq * ((n2 - n3) * d0 + (n2 + n3 + 1.0) * d1);
            }
            fs[e] = f;
        }
        __syncthreads();
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            double x = 0, y = 0;
            for (int k = 0; k < nb; ++k) {
                x += fs[b * nb + k];
                y += fs[k * nb + b];
            }
            s2[b] = x;
            s3[b] = y;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot = 0;
            for (int b = 0; b < nb; ++b) tot += s2[b];
            a.gpart[((size_t)t * a.nT + T) * nb + b1] = tot * (double)(e1 - e0);
        }
        // every (row, q', band) target belongs to one triplet and, inside it, to one
        // equivalent pair: store the values (q2' == q3' receives both terms)
        P *xr = (P *)a.xi + (size_t)T * a.xi_T + (size_t)row * a.xi_row;
        for (int e = e0; e < e1; ++e) {
            const int x2 = a.eq_q2[e], x3 = a.eq_q3[e];
            for (int b = threadIdx.x; b < nb; b += blockDim.x) {
                const double v2 = 2.0 * s2[b] * a.om_full[(size_t)q2 * nb + b] / w1;
                const double v3 = 2.0 * s3[b] * a.om_full[(size_t)q3 * nb + b] / w1;
                if (x2 == x3) {
                    xr[(size_t)x2 * nb + b] = (P)(v2 + v3);
                } else {
                    xr[(size_t)x2 * nb + b] = (P)v2;
                    xr[(size_t)x3 * nb + b] = (P)v3;
                }
            }
        }
        __syncthreads();
    }
}

// Small cells (nb <= 6): one persistent kernel fed by a work queue.  Every wave takes the
// next triplet index from a device counter and does the whole chain for it in LDS: phases,
// Phi (pretransform_phi3, scattering_threephonon.f90:275-317), the three contractions into
// Psi (:88-113), the smeared-delta weights of every temperature (get_dirac :219-271) and
// the linewidth partial and scattering-matrix entries the triplet owns (:115-135) -- the
// same arithmetic as k_phi + k_cgemm + k_epilogue, without batch buffers in HBM.  Triplets
// finish in any order; each writes only its own outputs, so results do not depend on it.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, bool SPLIT>
__device__ inline void cmac(T &rx, T &ry, cplx<T> a, cplx<T> b) {
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

// LDS bytes per wave of k_small.
template <int NB, typename T>
size_t small_wave_lds(int nfc3) {
    constexpr int NB2 = NB * NB, NB3 = NB2 * NB;
    size_t b = sizeof(cd) * (size_t)std::max(nfc3, 1);
    b += sizeof(cplx<T>) * (3 * (size_t)NB3 + 3 * (size_t)NB2);
    b += sizeof(double) * ((size_t)NB3 + 2 * (size_t)NB2);
    return (b + 15) & ~(size_t)15;
}

__global__ void k_zero_int(int *p) { *p = 0; }

template <int NB, typename T, bool SPLIT>
__global__ void __launch_bounds__(256) k_small(EpiArgs a, int ntrip, int *__restrict__ queue, size_t wave_lds,
                                               const cplx<T> *__restrict__ E1all, const cplx<T> *__restrict__ Efall,
                                               const double *__restrict__ qcart, const int *__restrict__ blk_off,
                                               const double *__restrict__ lv2, const double *__restrict__ lv3,
                                               const double *__restrict__ mwm, int nfc3) {
    constexpr int NA = NB / 3, NB2 = NB * NB, NB3 = NB2 * NB;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned char *base = smem + (size_t)w * wave_lds;
    cd *ph = (cd *)base;
    cplx<T> *phi = (cplx<T> *)(base + sizeof(cd) * (size_t)std::max(nfc3, 1));
    cplx<T> *t1 = phi + NB3, *t2 = t1 + NB3;
    cplx<T> *e1 = t2 + NB3, *e2 = e1 + NB2, *e3 = e2 + NB2;
    double *fs = (double *)(e3 + NB2);
    double *s2 = fs + NB3, *s3 = s2 + NB2;
    for (;;) {
        int t = 0;
        if (lane == 0) t = atomicAdd(queue, 1);
        t = __shfl(t, 0, 64);
        if (t >= ntrip) break;
        const int q1 = a.tq1[t], q2 = a.tq2[t], q3 = a.tq3[t];
        for (int e = lane; e < NB2; e += 64) {
            e1[e] = E1all[(size_t)q1 * NB2 + e];
            e2[e] = Efall[(size_t)q2 * NB2 + e];
            e3[e] = Efall[(size_t)q3 * NB2 + e];
        }
        const double *Q2 = qcart + 3 * q2, *Q3 = qcart + 3 * q3;
        for (int f = lane; f < nfc3; f += 64) {
            const double arg = -kTwoPi * (Q2[0] * lv2[3 * f] + Q2[1] * lv2[3 * f + 1] + Q2[2] * lv2[3 * f + 2] +
                                          Q3[0] * lv3[3 * f] + Q3[1] * lv3[3 * f + 1] + Q3[2] * lv3[3 * f + 2]);
            double sn, cs;
            sincos(arg, &sn, &cs);
            ph[f] = {cs, sn};
        }
        wave_sync();
        // Phi[ia][ib][ic]
        for (int e = lane; e < NB3; e += 64) {
            const int ia = e / NB2, ib = (e / NB) % NB, ic = e % NB;
            const int blk = ((ia / 3) * NA + ib / 3) * NA + ic / 3;
            const int comp = 9 * (ia % 3) + 3 * (ib % 3) + ic % 3;
            double ax = 0, ay = 0;
            for (int f = blk_off[blk]; f < blk_off[blk + 1]; ++f) {
                const double m = mwm[27 * f + comp];
                const cd p = ph[f];
                ax += m * p.x;
                ay += m * p.y;
            }
            phi[e] = {(T)ax, (T)ay};
        }
        wave_sync();
        // t1[b1][ib][ic] = sum_ia E1[ia][b1] Phi[ia][ib][ic]
        for (int e = lane; e < NB3; e += 64) {
            const int b1 = e / NB2, r = e % NB2;
            T rx = 0, ry = 0;
#pragma unroll
            for (int k = 0; k < NB; ++k) cmac<T, SPLIT>(rx, ry, e1[k * NB + b1], phi[k * NB2 + r]);
            t1[e] = {rx, ry};
        }
        wave_sync();
        // t2[b1][b2][ic] = sum_ib E2[ib][b2] t1[b1][ib][ic]
        for (int e = lane; e < NB3; e += 64) {
            const int b1 = e / NB2, b2 = (e / NB) % NB, ic = e % NB;
            T rx = 0, ry = 0;
#pragma unroll
            for (int k = 0; k < NB; ++k) cmac<T, SPLIT>(rx, ry, e2[k * NB + b2], t1[b1 * NB2 + k * NB + ic]);
            t2[e] = {rx, ry};
        }
        wave_sync();
        // Psi[b1][b2][b3] = sum_ic t2[b1][b2][ic] E3[ic][b3], into the Phi buffer
        for (int e = lane; e < NB3; e += 64) {
            const int r = e / NB, b3 = e % NB;
            T rx = 0, ry = 0;
#pragma unroll
            for (int k = 0; k < NB; ++k) cmac<T, SPLIT>(rx, ry, t2[r * NB + k], e3[k * NB + b3]);
            phi[e] = {rx, ry};
        }
        wave_sync();
        const int i2 = a.irr_of_full[q2], i3 = a.irr_of_full[q3];
        const double perm = a.perm[t];
        const int e0 = a.eq_off[t], e1q = a.eq_off[t + 1];
        for (int T_ = 0; T_ < a.nT; ++T_) {
            const double *be = a.be + (size_t)T_ * a.nirr * NB;
            for (int e = lane; e < NB3; e += 64) {
                const int b1 = e / NB2, b2 = (e / NB) % NB, b3 = e % NB;
                const double w1 = a.om_irr[q1 * NB + b1];
                const double w2 = a.om_full[(size_t)q2 * NB + b2], w3 = a.om_full[(size_t)q3 * NB + b3];
                double f = 0.0;
                if (a.row_of[q1 * NB + b1] >= 0 && !(w2 < 1e-9) && !(fabs(w1 - w2) < 1e-9) && !(w3 < 1e-9) &&
                    !(fabs(w1 - w3) < 1e-9) && !(fabs(w2 - w3) < 1e-9)) {
                    const double re = (double)phi[e].x, im = (double)phi[e].y;
                    const double psisq = a.psi_scale * (re * re + im * im);
                    const double sig =
                        (a.integrationtype == 2)
                            ? sqrt(a.sigsq[q1 * NB + b1] + a.sigsq[i2 * NB + b2] + a.sigsq[i3 * NB + b3])
                        : (a.integrationtype == 6)
                            ? sigma_type6(a.u_irr + ((size_t)q1 * NB + b1) * 3, a.u_full + ((size_t)q2 * NB + b2) * 3,
                                          a.u_full + ((size_t)q3 * NB + b3) * 3, a.thresh6)
                            : a.sigma_fixed;
                    const double d0 = gauss(w1, -w2 + w3, sig) - gauss(w1, w2 - w3, sig);
                    const double d1 = gauss(w1, w2 + w3, sig) - gauss(w1, -w2 - w3, sig);
                    const double n2 = be[i2 * NB + b2], n3 = be[i3 * NB + b3];
                    f = perm * (a.tw ? (double)a.tw[(size_t)t * NB + b1] : 1.0) * psisq * ((n2 - n3) * d0 + (n2 + n3 + 1.0) * d1);
                }
                fs[e] = f;
            }
            wave_sync();
            for (int x = lane; x < NB2; x += 64) {
                const int b1 = x / NB, b = x % NB;
                double u = 0, v = 0;
                for (int k = 0; k < NB; ++k) {
                    u += fs[b1 * NB2 + b * NB + k];
                    v += fs[b1 * NB2 + k * NB + b];
                }
                s2[x] = u;
                s3[x] = v;
            }
            wave_sync();
            for (int b1 = lane; b1 < NB; b1 += 64) {
                double tot = 0;
                for (int b = 0; b < NB; ++b) tot += s2[b1 * NB + b];
                a.gpart[((size_t)t * a.nT + T_) * NB + b1] =
                    (a.row_of[q1 * NB + b1] >= 0) ? tot * (double)(e1q - e0) : 0.0;
            }
            // every (row, q', band) target belongs to one triplet and one equivalent pair:
            // store the values (q2' == q3' receives both terms)
            T *xiT = (T *)a.xi + (size_t)T_ * a.xi_T;
            for (int e = e0; e < e1q; ++e) {
                const int x2 = a.eq_q2[e], x3 = a.eq_q3[e];
                for (int x = lane; x < NB2; x += 64) {
                    const int b1 = x / NB, b = x % NB;
                    const int row = a.row_of[q1 * NB + b1];
                    if (row < 0) continue;
                    const double w1 = a.om_irr[q1 * NB + b1];
                    T *xr = xiT + (size_t)row * a.xi_row;
                    const double v2 = 2.0 * s2[x] * a.om_full[(size_t)q2 * NB + b] / w1;
                    const double v3 = 2.0 * s3[x] * a.om_full[(size_t)q3 * NB + b] / w1;
                    if (x2 == x3) {
                        xr[(size_t)x2 * NB + b] = (T)(v2 + v3);
                    } else {
                        xr[(size_t)x2 * NB + b] = (T)v2;
                        xr[(size_t)x3 * NB + b] = (T)v3;
                    }
                }
            }
            wave_sync();
        }
    }
}

// Bose-Einstein occupations be[T][q][b] of the irreducible modes (scattering.f90:148-152:
// lo_planck, gottochblandat_physics.f90:160-188; classical: k_B T / omega), zero below
// lo_freqtol.
__global__ void k_occupations(int nT, int nqb, const double *__restrict__ temps, const double *__restrict__ om,
                              int classical, double *__restrict__ be) {
    const long n = (long)nT * nqb;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double T = temps[i / nqb], w = om[i % nqb];
        double v = 0.0;
        if (w >= kFreqTol) {
            if (classical) {
                v = kKbHartree * T / w;
            } else if (T >= kTemperatureTol) {
                const double x = w / (kKbHartree * T);
                if (!(x > 1e2)) v = 1.0 / (exp(x) - 1.0);
            }
        }
        be[i] = v;
    }
}

// Linewidth of every row: deterministic sum over the triplets of its q1.  Workgroup per
// (irreducible q1, temperature), one thread per band; the triplets of q1 are summed in
// order, eight loads in flight.
__global__ void k_gamma(int nrows, int nT, int nb, const int *__restrict__ row_of, const int *__restrict__ q1_off,
                        const double *__restrict__ gpart, double *__restrict__ lw) {
    const int q = blockIdx.x, T = blockIdx.y, b = threadIdx.x;
    if (b >= nb) return;
    const int r = row_of[(size_t)q * nb + b];
    if (r < 0) return;
    const double *gp = gpart + (size_t)T * nb + b;
    const size_t st = (size_t)nT * nb;
    const int t0 = q1_off[q], t1 = q1_off[q + 1];
    double s = 0;
    int t = t0;
    for (; t + 8 <= t1; t += 8) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = gp[(size_t)(t + k) * st];
#pragma unroll
        for (int k = 0; k < 8; ++k) s += v[k];
    }
    for (; t < t1; ++t) s += gp[(size_t)t * st];
    lw[(size_t)T * nrows + r] = s;
}

// Isotope scattering inputs (compute_isotope_scattering, scattering_isotope.f90:2-70): the
// mass-disorder parameter of every atom; null g means no isotope term.
struct IsoArgs {
    const double *g;               // na
    const cd *egv_irr, *egv_full;  // [q][component][mode]
    const double *om_irr, *om_full, *sigsq;
    const int *row_b, *irr_of_full;
    double sigma_fixed, pref;      // pi/4 / N_q
    int na, integrationtype;
    const double *u_full;          // integration type 6: reduced group velocities, and baseline
    double thresh6;
    double *gam;                   // [row] linewidth term (written by the T = 0 workgroup)
};

__device__ inline double gauss_iso(double x, double mu, double sigma) {
    // lo_gauss (gottochblandat_calculus.f90:268-288)
    const double d = x - mu;
    if (fabs(d) > 8.0 * sigma) return 0.0;
    const double inv = 1.0 / sigma;
    return exp(-d * d * 0.5 * inv * inv) * 0.398942280401433 * inv;
}

// Orbit averaging of one scattering-matrix row: symmetrize_and_distribute (mean of the
// positive three-phonon entries broadcast to the orbit, scattering_threephonon.f90:156-196),
// then the isotope term added (it enters the row directly, scattering_isotope.f90:62-66),
// then symmetrize_scatmat (plain mean, scattering.f90:312-354).  Without isotopes the first
// step leaves an orbit with a positive entry holding one value, so the mean is evaluated
// on that value in registers; each entry is read once and written once.  Workgroup per
// (row, T).
template <typename XT, bool ISO>
__global__ void k_xi_sym(int nrows, int nb, const int *__restrict__ row_q, const int *__restrict__ q1_off,
                         const int *__restrict__ orb_off, const int *__restrict__ members,
                         XT *__restrict__ xi, size_t xi_row, size_t xi_T, IsoArgs iso,
                         const double *__restrict__ add4 = nullptr, size_t add_row = 0) {
    // add4: four-phonon terms [T][row][q * nb + b], already averaged over the little group of
    // q1 and added, like the isotope term, after the positive mean of the three-phonon terms
    // and before the plain mean (compute_fourphonon_scattering adds them to Xi after
    // compute_threephonon_scattering has averaged its own, scattering.f90:233-242)
    const int r = blockIdx.x % nrows, T = blockIdx.x / nrows;
    const double *a4 = add4 ? add4 + ((size_t)T * nrows + r) * add_row : nullptr;
    const int q1 = row_q[r];
    XT *x = xi + (size_t)T * xi_T + (size_t)r * xi_row;
    const int o0 = q1_off[q1], o1 = q1_off[q1 + 1];
    const long work = (long)(o1 - o0) * nb;
    __shared__ cd e1[128];
    __shared__ double gsh[64], red[kBlock];
    int b1 = 0;
    double om1 = 0, sg1 = 0;
    if constexpr (ISO) {
        b1 = iso.row_b[r];
        om1 = iso.om_irr[(size_t)q1 * nb + b1];
        sg1 = iso.sigsq[(size_t)q1 * nb + b1];
        for (int k = threadIdx.x; k < nb; k += blockDim.x) e1[k] = iso.egv_irr[((size_t)q1 * nb + k) * nb + b1];
        for (int a = threadIdx.x; a < iso.na; a += blockDim.x) gsh[a] = iso.g[a];
        __syncthreads();
    }
    double gsum = 0;
    for (long w = threadIdx.x; w < work; w += blockDim.x) {
        const int o = o0 + (int)(w / nb), b = (int)(w % nb);
        const int m0 = orb_off[o], m1 = orb_off[o + 1];
        double sp = 0, sa = 0;
        int n = 0;
        for (int m = m0; m < m1; ++m) {
            const double v = (double)x[(size_t)members[m] * nb + b];
            sa += v;
            if (v > 0.0) { sp += v; ++n; }
        }
        double s = sa;
        if constexpr (!ISO) {
            if (a4) {
                const double avg = (n > 0) ? sp / n : 0.0;
                s = 0;
                for (int m = m0; m < m1; ++m) {
                    const int q2 = members[m];
                    s += ((n > 0) ? avg : (double)x[(size_t)q2 * nb + b]) + a4[(size_t)q2 * nb + b];
                }
            } else if (n > 0) {
                const double avg = sp / n;
                s = 0;
                for (int m = m0; m < m1; ++m) s += avg;
            }
        } else {
            const double avg = (n > 0) ? sp / n : 0.0;
            s = 0;
            for (int m = m0; m < m1; ++m) {
                const int q2 = members[m];
                double v = (n > 0) ? avg : (double)x[(size_t)q2 * nb + b];
                if (a4) v += a4[(size_t)q2 * nb + b];
                const double om2 = iso.om_full[(size_t)q2 * nb + b];
                if (!(om2 < kFreqTol)) {
                    double st = 0;
                    for (int a = 0; a < iso.na; ++a) {
                        double f0 = 0;
                        for (int j = 0; j < 3; ++j) {
                            const cd p = e1[3 * a + j], e2 = iso.egv_full[((size_t)q2 * nb + 3 * a + j) * nb + b];
                            const double re = p.x * e2.x + p.y * e2.y, im = p.x * e2.y - p.y * e2.x;
                            f0 += sqrt(re * re + im * im);
                        }
                        st += f0 * f0 * gsh[a];
                    }
                    const double sig =
                        (iso.integrationtype == 2) ? sqrt(sg1 + iso.sigsq[(size_t)iso.irr_of_full[q2] * nb + b])
                        : (iso.integrationtype == 6) ? sigma_type6_iso(iso.u_full + ((size_t)q2 * nb + b) * 3, iso.thresh6)
                                                     : iso.sigma_fixed;
                    const double f = st * iso.pref * om1 * om2 * gauss_iso(om1, om2, sig);
                    gsum += f;
                    v += f * om2 / om1;
                }
                s += v;
            }
        }
        const XT out = (XT)(s / (m1 - m0));
        for (int m = m0; m < m1; ++m) x[(size_t)members[m] * nb + b] = out;
    }
    if (ISO && T == 0) {
        red[threadIdx.x] = gsum;
        __syncthreads();
        for (int st = blockDim.x / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) iso.gam[r] = red[0];
    }
}

// The same orbit averaging without isotopes, for nb <= 256, over chunks of whole orbits:
// the members of a chunk (orbits [chk[c], chk[c + 1]) of q1) and their entries are staged
// in LDS with independent loads, every (orbit, band) is averaged there in the member order
// of k_xi_sym, and the chunk is written back.  Each thread keeps one band; the workgroup
// has nb * (256 / nb) threads.  Workgroup per (row, T).
#ifndef TC3_XI_DEPTH
#define TC3_XI_DEPTH 16
#endif
constexpr int kXiDepth = TC3_XI_DEPTH;
#ifndef TC3_XI_CHUNK
#define TC3_XI_CHUNK 4096
#endif
constexpr int kXiChunk = TC3_XI_CHUNK;   // entries per orbit chunk staged in LDS

template <typename XT>
__global__ void __launch_bounds__(kBlock) k_xi_sym_lds(int nrows, int nb, const int *__restrict__ row_q,
                                                       const int *__restrict__ chk_off, const int *__restrict__ chk,
                                                       const int *__restrict__ orb_off,
                                                       const int *__restrict__ members, XT *__restrict__ xi,
                                                       size_t xi_row, size_t xi_T, int capm) {
    extern __shared__ __align__(16) unsigned char smem[];
    XT *val = (XT *)smem;
    int *mem = (int *)(val + (size_t)capm * nb);
    int *ooff = mem + capm;                  // member offsets of the chunk's orbits
    const int r = blockIdx.x % nrows, T = blockIdx.x / nrows;
    const int q1 = row_q[r];
    XT *x = xi + (size_t)T * xi_T + (size_t)r * xi_row;
    const int b = threadIdx.x % nb, p0 = threadIdx.x / nb, ps = blockDim.x / nb;
    for (int c = chk_off[q1]; c + 1 < chk_off[q1 + 1]; ++c) {
        const int oa = chk[c], ob = chk[c + 1];
        const int pa = orb_off[oa], n = orb_off[ob] - pa;
        for (int p = threadIdx.x; p < n; p += blockDim.x) mem[p] = members[pa + p];
        for (int o = threadIdx.x; o <= ob - oa; o += blockDim.x) ooff[o] = orb_off[oa + o] - pa;
        __syncthreads();
        // kXiDepth independent loads in flight per thread before their LDS stores
        for (int p = p0; p < n; p += kXiDepth * ps) {
            XT v[kXiDepth];
#pragma unroll
            for (int k = 0; k < kXiDepth; ++k) {
                const int pk = p + k * ps;
                v[k] = (pk < n) ? x[(size_t)mem[pk] * nb + b] : (XT)0;
            }
#pragma unroll
            for (int k = 0; k < kXiDepth; ++k) {
                const int pk = p + k * ps;
                if (pk < n) val[(size_t)pk * nb + b] = v[k];
            }
        }
        __syncthreads();
        for (int o = oa + p0; o < ob; o += ps) {
            const int m0 = ooff[o - oa], m1 = ooff[o - oa + 1];
            double sp = 0, sa = 0;
            int cnt = 0;
            for (int m = m0; m < m1; ++m) {
                const double v = (double)val[(size_t)m * nb + b];
                sa += v;
                if (v > 0.0) { sp += v; ++cnt; }
            }
            double sum = sa;
            if (cnt > 0) {
                const double avg = sp / cnt;
                sum = 0;
                for (int m = m0; m < m1; ++m) sum += avg;
            }
            const XT out = (XT)(sum / (m1 - m0));
            for (int m = m0; m < m1; ++m) val[(size_t)m * nb + b] = out;
        }
        __syncthreads();
        for (int p = p0; p < n; p += kXiDepth * ps) {
#pragma unroll
            for (int k = 0; k < kXiDepth; ++k) {
                const int pk = p + k * ps;
                if (pk < n) x[(size_t)mem[pk] * nb + b] = val[(size_t)pk * nb + b];
            }
        }
        __syncthreads();
    }
}

// Scattering-matrix rows from the compact per-triplet sums of the MFMA kernels, then (AVG)
// the orbit averaging of k_xi_sym_lds.  msrc, in the member order of the orbits, holds for
// every q' the owner 4 t + mode (the triplet t of q1 whose equivalent pairs own the entries
// (q', b) of the rows of q1: mode 1 the q2' sums, mode 2 the q3' sums, mode 3 both, q2' =
// q3' within one pair; 0 none, entry zero) and the frequency offsets nb q of the owning
// role and of q3': (q', owner, offset, offset q3').  The entries are formed with the
// arithmetic of the direct stores of k_psi (frequency of the mode times 2 / omega_1,
// rounding to the entry type, the q3' term added to the rounded q2' term), so the rows are
// the same bits.
// Raw buffer access with 32-bit byte offsets from a uniform base (gfx9 descriptor word 3).
typedef unsigned xv2u __attribute__((ext_vector_type(2)));
typedef unsigned v4u32 __attribute__((ext_vector_type(4)));
__device__ inline __amdgpu_buffer_rsrc_t xrsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ inline T xload(__amdgpu_buffer_rsrc_t r, unsigned off) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <typename T>
__device__ inline void xstore(T v, __amdgpu_buffer_rsrc_t r, unsigned off) {
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(xv2u, v), r, (int)off, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 0);
}

// No LDS staging: three bands per thread (nb = 3 na, so na lanes cover a member's nb
// entries), the orbit sums in registers.  Without AVG a lane group takes one member.  With
// AVG the orbits of one member (ordered first among the orbits of q1; their average is
// 0 + x) are taken the same way, and each orbit of several members by nb lanes, one band
// each, which form the entries of every member in member order, accumulate the sums of the
// averaging in that order and write the orbit's value to every member.
template <typename T>
__device__ inline void xload3(T (&v)[3], __amdgpu_buffer_rsrc_t r, unsigned off) {
#pragma unroll
    for (int j = 0; j < 3; ++j) v[j] = xload<T>(r, off + (unsigned)sizeof(T) * j);
}
template <typename T>
__device__ inline void xstore3(const T (&v)[3], __amdgpu_buffer_rsrc_t r, unsigned off) {
#pragma unroll
    for (int j = 0; j < 3; ++j) xstore<T>(v[j], r, off + (unsigned)sizeof(T) * j);
}

#ifndef TC3_ORB_DEPTH
#define TC3_ORB_DEPTH 2
#endif
constexpr int kOrbDepth = TC3_ORB_DEPTH;
#ifndef TC3_ORB_TEMPS
#define TC3_ORB_TEMPS 4
#endif
constexpr int kOrbTemps = TC3_ORB_TEMPS;   // temperatures per workgroup

// Workgroup per (row, block of kOrbTemps temperatures): the owner table and the frequencies
// are read once for the block's temperatures, whose sums are independent loads in flight.
// xo_row non-null (AVG only): the rows are stored per orbit (o nb + b for the o-th orbit of
// q1, row r at xo_row[r]), each orbit's value once.
template <typename XT, typename S, bool AVG>
__global__ void __launch_bounds__(kBlock) k_xi_asm(int nrows, int nb, int nfull, int nT, const int *__restrict__ row_q,
                                                   const int *__restrict__ row_b, const double *__restrict__ om_irr,
                                                   const double *__restrict__ om_full, const int4 *__restrict__ msrc,
                                                   const S *__restrict__ comp, size_t comp_T,
                                                   const int *__restrict__ orb_q1, const int *__restrict__ orb_off,
                                                   const int *__restrict__ orb_ns, XT *__restrict__ xi, size_t xi_row,
                                                   size_t xi_T, const long long *__restrict__ xo_row) {
    constexpr int TB = kOrbTemps;
    const int L = nb / 3;
    const int gs = blockDim.x / L, g0 = threadIdx.x / L, b = 3 * (threadIdx.x % L);
    if (g0 >= gs) return;
    const int r = blockIdx.x % nrows, T0 = (blockIdx.x / nrows) * TB, nt = min(TB, nT - T0);
    const int q1 = row_q[r], b1 = row_b[r];
    const double iw1 = 2.0 / om_irr[(size_t)q1 * nb + b1];
    const bool orbout = AVG && xo_row != nullptr;
    const size_t rbase = orbout ? (size_t)xo_row[r] : (size_t)r * xi_row;
    const size_t rlen = orbout ? (size_t)(orb_q1[q1 + 1] - orb_q1[q1]) * nb : xi_row;
    // byte offsets below 2^32 (checked by the caller)
    __amdgpu_buffer_rsrc_t rx[TB], rc[TB];
#pragma unroll
    for (int t = 0; t < TB; ++t) {
        const int Tt = min(T0 + t, nT - 1);
        rx[t] = xrsrc(xi + (size_t)Tt * xi_T + rbase, (unsigned)(rlen * sizeof(XT)));
        rc[t] = xrsrc(comp + (size_t)Tt * comp_T, (unsigned)(comp_T * sizeof(S)));
    }
    const __amdgpu_buffer_rsrc_t ro = xrsrc(om_full, (unsigned)((size_t)nfull * nb * sizeof(double)));
    const unsigned tstride = (unsigned)(2 * nb * nb * sizeof(S));
    const unsigned cbase = (unsigned)((2 * nb * b1 + b) * sizeof(S)), c3 = (unsigned)(nb * sizeof(S));
    // the three entries of member position m at the block's temperatures (zero without owner)
    auto entries = [&](const int4 ms, XT (&e)[TB][3]) {
        const int mode = ms.y & 3;
        const unsigned tb = (unsigned)(ms.y >> 2) * tstride + cbase;
        S v[TB][3];
        double f[3];
#pragma unroll
        for (int t = 0; t < TB; ++t)
            if (t < nt) xload3<S>(v[t], rc[t], tb + (mode == 2 ? c3 : 0u));
        xload3<double>(f, ro, (unsigned)(ms.z + b) * 8u);
#pragma unroll
        for (int t = 0; t < TB; ++t)
#pragma unroll
            for (int j = 0; j < 3; ++j) e[t][j] = mode ? (XT)((double)v[t][j] * f[j] * iw1) : (XT)0;
        if (mode == 3) {
            // q2' = q3' within one pair: the q3' term added to the rounded q2' term
            S v3[TB][3];
            double f3[3];
#pragma unroll
            for (int t = 0; t < TB; ++t)
                if (t < nt) xload3<S>(v3[t], rc[t], tb + c3);
            xload3<double>(f3, ro, (unsigned)(ms.w + b) * 8u);
#pragma unroll
            for (int t = 0; t < TB; ++t)
#pragma unroll
                for (int j = 0; j < 3; ++j) e[t][j] = (XT)((double)e[t][j] + (double)v3[t][j] * f3[j] * iw1);
        }
    };
    auto store = [&](const XT (&e)[TB][3], int mem) {
#pragma unroll
        for (int t = 0; t < TB; ++t)
            if (t < nt) xstore3<XT>(e[t], rx[t], (unsigned)(mem * nb + b) * (unsigned)sizeof(XT));
    };
    // members one per lane group, kOrbDepth in flight; with AVG the orbits of one member (the
    // first orbits of q1), whose average reduces to 0 + x
    const int pa = orb_off[orb_q1[q1]], pb = AVG ? pa + orb_ns[q1] : orb_off[orb_q1[q1 + 1]];
    for (int m = pa + g0; m < pb; m += kOrbDepth * gs) {
        int4 ms[kOrbDepth];
#pragma unroll
        for (int k = 0; k < kOrbDepth; ++k)
            if (m + k * gs < pb) ms[k] = msrc[m + k * gs];
#pragma unroll
        for (int k = 0; k < kOrbDepth; ++k) {
            if (m + k * gs >= pb) break;
            XT e[TB][3];
            entries(ms[k], e);
            if constexpr (AVG) {
#pragma unroll
                for (int t = 0; t < TB; ++t)
#pragma unroll
                    for (int j = 0; j < 3; ++j) e[t][j] = (XT)(0.0 + (double)e[t][j]);
            }
            // per orbit: these orbits of one member are the first orbits of q1, in member order
            store(e, orbout ? m + k * gs - pa : ms[k].x);
        }
    }
    if constexpr (AVG) {
        // orbits of several members: nb lanes per orbit, one band each, the sums of the block's
        // temperatures in member order
        const int g1 = threadIdx.x / nb, gs1 = blockDim.x / nb, b0 = threadIdx.x % nb;
        const unsigned cbase1 = (unsigned)((2 * nb * b1 + b0) * sizeof(S));
        for (int o = orb_q1[q1] + orb_ns[q1] + g1; g1 < gs1 && o < orb_q1[q1 + 1]; o += gs1) {
            const int m0 = orb_off[o], m1 = orb_off[o + 1];
            double sa[TB] = {}, sp[TB] = {};
            int cnt[TB] = {};
            int mfirst[kOrbDepth];   // members of the first batch, for the stores
            for (int m = m0; m < m1; m += kOrbDepth) {
                int4 ms[kOrbDepth];
#pragma unroll
                for (int k = 0; k < kOrbDepth; ++k)
                    if (m + k < m1) ms[k] = msrc[m + k];
                S v[kOrbDepth][TB], v3[kOrbDepth][TB];
                double f[kOrbDepth], f3[kOrbDepth];
#pragma unroll
                for (int k = 0; k < kOrbDepth; ++k) {
                    if (m + k >= m1) break;
                    const int mode = ms[k].y & 3;
                    const unsigned tb = (unsigned)(ms[k].y >> 2) * tstride + cbase1;
#pragma unroll
                    for (int t = 0; t < TB; ++t)
                        if (t < nt) v[k][t] = xload<S>(rc[t], tb + (mode == 2 ? c3 : 0u));
                    f[k] = xload<double>(ro, (unsigned)(ms[k].z + b0) * 8u);
                    if (mode == 3) {
#pragma unroll
                        for (int t = 0; t < TB; ++t)
                            if (t < nt) v3[k][t] = xload<S>(rc[t], tb + c3);
                        f3[k] = xload<double>(ro, (unsigned)(ms[k].w + b0) * 8u);
                    }
                }
#pragma unroll
                for (int k = 0; k < kOrbDepth; ++k) {
                    if (m + k >= m1) break;
                    if (m == m0) mfirst[k] = ms[k].x;
                    const int mode = ms[k].y & 3;
#pragma unroll
                    for (int t = 0; t < TB; ++t) {
                        XT e = mode ? (XT)((double)v[k][t] * f[k] * iw1) : (XT)0;
                        // q2' = q3' within one pair: the q3' term added to the rounded q2' term
                        if (mode == 3) e = (XT)((double)e + (double)v3[k][t] * f3[k] * iw1);
                        const double x = (double)e;
                        sa[t] += x;
                        if (x > 0.0) { sp[t] += x; ++cnt[t]; }
                    }
                }
            }
            XT out[TB];
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                double sum = sa[t];
                if (cnt[t] > 0) {
                    const double avg = sp[t] / cnt[t];
                    sum = 0;
                    for (int m = m0; m < m1; ++m) sum += avg;
                }
                out[t] = (XT)(sum / (m1 - m0));
            }
            if (orbout) {
#pragma unroll
                for (int t = 0; t < TB; ++t)
                    if (t < nt) xstore<XT>(out[t], rx[t], (unsigned)((o - orb_q1[q1]) * nb + b0) * (unsigned)sizeof(XT));
                continue;
            }
            for (int m = m0; m < m1; m += kOrbDepth) {
                int mm[kOrbDepth];
#pragma unroll
                for (int k = 0; k < kOrbDepth; ++k)
                    if (m + k < m1) mm[k] = (m == m0) ? mfirst[k] : msrc[m + k].x;
#pragma unroll
                for (int k = 0; k < kOrbDepth; ++k) {
                    if (m + k >= m1) break;
#pragma unroll
                    for (int t = 0; t < TB; ++t)
                        if (t < nt) xstore<XT>(out[t], rx[t], (unsigned)(mm[k] * nb + b0) * (unsigned)sizeof(XT));
                }
            }
        }
    }
}

// Owner table of k_xi_asm in the member order of the orbits of every irreducible q1, from
// the triplets' equivalent pairs in the order of the direct stores of k_psi (all q2'
// entries of a triplet, then its q3' entries: a store, or an addition where q2' = q3' in the
// pair): four ints per member position (see k_xi_asm).  False if an entry would be owned by
// two triplets or take a form the table cannot express.
bool xi_source_map(const tc3::TripletSet &tr, const tc3::OrbitSet &orb, int nirr, int nfull, int nb,
                   std::vector<int> &msrc) {
    std::vector<int> s(nfull);
    msrc.assign(4 * orb.members.size(), 0);
    for (int q1 = 0; q1 < nirr; ++q1) {
        std::fill(s.begin(), s.end(), 0);
        for (int t = tr.q1_off[q1]; t < tr.q1_off[q1 + 1]; ++t) {
            for (int e = tr.eq_off[t]; e < tr.eq_off[t + 1]; ++e) {
                int &v = s[tr.eq_q2[e]];
                if ((v & 3) && (v >> 2) != t) return false;
                v = 4 * t + 1;
            }
            for (int e = tr.eq_off[t]; e < tr.eq_off[t + 1]; ++e) {
                const int q3 = tr.eq_q3[e];
                int &v = s[q3];
                if ((v & 3) && (v >> 2) != t) return false;
                if (tr.eq_q2[e] == q3) {
                    if ((v & 3) != 1) return false;
                    v = 4 * t + 3;
                } else {
                    v = 4 * t + 2;
                }
            }
        }
        const int pa = orb.orb_off[orb.q1_off[q1]], pb = orb.orb_off[orb.q1_off[q1 + 1]];
        for (int p = pa; p < pb; ++p) {
            const int q = orb.members[p], sm = s[q], t = sm >> 2;
            int *m = msrc.data() + 4 * (size_t)p;
            m[0] = q;
            m[1] = sm;
            m[2] = ((sm & 3) == 2 ? tr.q3[t] : tr.q2[t]) * nb;
            m[3] = tr.q3[t] * nb;
        }
    }
    return true;
}

}  // namespace

// Nonzero (a2, a3) IFC blocks of every atom a1, in (a1, a2, a3) order, with the offsets of
// each a1 and each (a1, a2) and the most blocks of one a1 (the MFMA kernels' Phi blocks).
// Tables of the scattering matrix stored per orbit (tc3_ctx::xi_orb), from the orbits of the
// little group of every irreducible point.  False if an offset would not fit the tables' ints.
static bool build_orbit_layout(tc3_ctx *c) {
    if (c->xo_gen == c->gen_model) return true;
    const int nirr = c->mesh.nirr, nfull = c->mesh.nfull, nb = c->nb, nrows = c->nrows;
    const OrbitSet &o = c->orb;
    std::vector<int> grp(nirr, -1);
    c->xo_goff.clear();
    c->xo_ent.clear();
    c->xo_mem.clear();
    std::map<std::vector<int>, int> groups;
    long long foff = 0;
    for (int q = 0; q < nirr; ++q) {
        const int oa = o.q1_off[q], ob = o.q1_off[q + 1], norb = ob - oa;
        // the orbit list as (size, members...) per orbit
        std::vector<int> key;
        bool ident = norb == nfull;
        for (int x = oa; x < ob; ++x) {
            key.push_back(o.orb_off[x + 1] - o.orb_off[x]);
            for (int m = o.orb_off[x]; m < o.orb_off[x + 1]; ++m) key.push_back(o.members[m]);
            ident = ident && o.orb_off[x + 1] - o.orb_off[x] == 1 && o.members[o.orb_off[x]] == x - oa;
        }
        if (ident) continue;
        auto it = groups.find(key);
        if (it == groups.end()) {
            it = groups.emplace(key, (int)c->xo_goff.size()).first;
            c->xo_goff.push_back(foff);
            for (int x = oa; x < ob; ++x) {
                const int m0 = (int)c->xo_mem.size();
                for (int m = o.orb_off[x]; m < o.orb_off[x + 1]; ++m) c->xo_mem.push_back(o.members[m]);
                c->xo_ent.push_back(make_int4(m0, (int)c->xo_mem.size(), (int)(foff + (long long)(x - oa) * nb),
                                              norb * nb));
            }
            foff += 3ll * norb * nb;
            if (foff >= (1ll << 31)) return false;
        }
        grp[q] = it->second;
    }
    c->xo_F = (size_t)foff;
    c->xo_row.assign(nrows, 0);
    c->xo_rg.clear();
    long long off = 0;
    for (int r = 0; r < nrows; ++r) {
        c->xo_row[r] = off;
        const int q = c->row_q[r];
        off += (long long)(o.q1_off[q + 1] - o.q1_off[q]) * nb;
    }
    c->xo_T = (size_t)off;
    for (int r = 0; r < nrows;) {
        const int q = c->row_q[r];
        int n = 0;
        while (r + n < nrows && c->row_q[r + n] == q && n < 16) ++n;
        c->xo_rg.push_back(make_int4(r, n, grp[q], (o.q1_off[q + 1] - o.q1_off[q]) * nb));
        r += n;
    }
    c->xo_gen = c->gen_model;
    return true;
}

// Every argument the contraction's launches take (the plan, its epilogue arguments and the
// events of each launch pair): the key of the pass's HIP graph.
static void chain_key(std::vector<long> &k, const MfmaPlan &p, int precision, int nev) {
    auto P = [&](const void *x) { k.push_back((long)(uintptr_t)x); };
    auto I = [&](long x) { k.push_back(x); };
    auto D = [&](double x) {
        long v;
        std::memcpy(&v, &x, sizeof(v));
        k.push_back(v);
    };
    I(precision); I(p.na); I(p.nb); I(p.nirr); I(p.nfull); I(p.maxa1); I(p.maxblk); I(p.nty); I(p.ntrip);
    I(p.batch); I(p.psi_bb);
    P(p.egv_irr); P(p.egv_full); P(p.om_irr); P(p.om_full); P(p.E1re); P(p.E1im); P(p.Efre); P(p.Efim);
    P(p.work); P(p.Eti); P(p.Etf); P(p.tq2); P(p.tq3); P(p.a1off); P(p.blist); P(p.a1_boff); P(p.a12_boff);
    P(p.qcart); P(p.lv2); P(p.lv3); P(p.mwm);
    const MfmaEpiArgs &e = p.epi;
    I(e.nb); I(e.nT); I(e.t0); I(e.nt); I(e.nirr); I(e.integrationtype); I(e.gT); I(e.T0);
    D(e.sigma_fixed); D(e.psi_scale);
    P(e.tq1); P(e.tq2); P(e.tq3); P(e.eq_off); P(e.eq_q2); P(e.eq_q3); P(e.irr_of_full); P(e.row_of);
    P(e.perm); P(e.om_irr); P(e.om_full); P(e.be); P(e.sigsq); P(e.xi); P(e.gpart); I((long)e.xi_row);
    I((long)e.xi_T); P(e.comp); I((long)e.comp_T); P(e.u_irr); P(e.u_full); D(e.thresh6); P(e.tw);
    P(e.psi_out);
    for (int i = 0; i < nev; ++i) P(p.ev ? (const void *)p.ev[i] : nullptr);
}

static void mfma_block_lists(const tc3_ctx *c, std::vector<int4> &blist, std::vector<int> &a1_boff,
                             std::vector<int> &a12_boff, int &maxblk) {
    const int na = c->cr.na;
    blist.clear();
    a1_boff.assign(na + 1, 0);
    a12_boff.assign((size_t)na * na + 1, 0);
    maxblk = 0;
    for (int a1 = 0; a1 < na; ++a1) {
        a1_boff[a1] = (int)blist.size();
        for (int a2 = 0; a2 < na; ++a2) {
            a12_boff[(size_t)a1 * na + a2] = (int)blist.size();
            for (int a3 = 0; a3 < na; ++a3) {
                const size_t bk = ((size_t)a1 * na + a2) * na + a3;
                if (c->blk3_off[bk + 1] > c->blk3_off[bk])
                    blist.push_back(make_int4(a2, a3, c->blk3_off[bk], c->blk3_off[bk + 1]));
            }
        }
        maxblk = std::max(maxblk, (int)blist.size() - a1_boff[a1]);
    }
    a1_boff[na] = (int)blist.size();
    a12_boff[(size_t)na * na] = (int)blist.size();
}

int run_psi2(tc3_ctx *c, int precision, int n, const int64_t *trips, double *out, int *qpts) {
    const int na = c->cr.na, nb = 3 * na;
    const int nirr = c->mesh.nirr, nfull = c->mesh.nfull;
    const TripletSet &tr = c->trip_full;
    const long ntf = (long)tr.q1.size();
    if (!mfma_supported(nb, precision))
        return c->fail(5, "the |Psi|^2 export covers the MFMA contraction (48 and 96 bands)");
    if (n <= 0 || !trips || !out) return c->fail(5, "need at least one triplet and an output buffer");
    for (int i = 0; i < n; ++i)
        if (trips[i] < 0 || trips[i] >= ntf) return c->fail(5, "triplet index out of range");
    hipStream_t st = c->stream;
    hipError_t e = hipSuccess;
#define TRY(x) do { e = (x); if (e != hipSuccess) return device_error(e, #x, __FILE__, __LINE__); } while (0)
    auto up = [&](const char *name, void **p, const void *src, size_t bytes) -> hipError_t {
        hipError_t r = c->pool_get(name, bytes, p);
        if (r == hipSuccess) r = hipMemcpyAsync(*p, src, bytes, hipMemcpyHostToDevice, st);
        return r;
    };
    // the listed triplets, each with permutation factor 1 and no equivalent pairs (nothing
    // of the scattering matrix is written: every row index is -1)
    std::vector<int> tq1(n), tq2(n), tq3(n), eqo(n + 1, 0), rowof((size_t)nirr * nb, -1);
    for (int i = 0; i < n; ++i) {
        const long t = (long)trips[i];
        tq1[i] = tr.q1[t];
        tq2[i] = tr.q2[t];
        tq3[i] = tr.q3[t];
        if (qpts) {
            qpts[3 * i] = tq1[i];
            qpts[3 * i + 1] = tq2[i];
            qpts[3 * i + 2] = tq3[i];
        }
    }
    const std::vector<double> perm(n, 1.0), be((size_t)nirr * nb, 0.0), sg((size_t)nirr * nb, 1.0);
    std::vector<int4> blist;
    std::vector<int> a1_boff, a12_boff;
    int maxblk = 0, maxa1 = 0;
    mfma_block_lists(c, blist, a1_boff, a12_boff, maxblk);
    for (int a = 0; a < na; ++a) maxa1 = std::max(maxa1, c->a1_off[a + 1] - c->a1_off[a]);
    int *d_tq1, *d_tq2, *d_tq3, *d_eqo, *d_rowof, *d_a1off, *d_a1boff, *d_a12boff;
    int4 *d_blist;
    double *d_perm, *d_be, *d_sg, *d_qcart, *d_lv2, *d_lv3, *d_mwm, *d_gpart;
    TRY(up("p2_tq1", (void **)&d_tq1, tq1.data(), sizeof(int) * n));
    TRY(up("p2_tq2", (void **)&d_tq2, tq2.data(), sizeof(int) * n));
    TRY(up("p2_tq3", (void **)&d_tq3, tq3.data(), sizeof(int) * n));
    TRY(up("p2_eqo", (void **)&d_eqo, eqo.data(), sizeof(int) * (n + 1)));
    TRY(up("p2_rowof", (void **)&d_rowof, rowof.data(), sizeof(int) * rowof.size()));
    TRY(up("p2_perm", (void **)&d_perm, perm.data(), sizeof(double) * n));
    TRY(up("p2_be", (void **)&d_be, be.data(), sizeof(double) * be.size()));
    TRY(up("p2_sg", (void **)&d_sg, sg.data(), sizeof(double) * sg.size()));
    TRY(up("p2_blist", (void **)&d_blist, blist.data(), sizeof(int4) * std::max<size_t>(blist.size(), 1)));
    TRY(up("p2_a1boff", (void **)&d_a1boff, a1_boff.data(), sizeof(int) * a1_boff.size()));
    TRY(up("p2_a12boff", (void **)&d_a12boff, a12_boff.data(), sizeof(int) * a12_boff.size()));
    TRY(up("p2_a1off", (void **)&d_a1off, c->a1_off.data(), sizeof(int) * (na + 1)));
    TRY(up("p2_qcart", (void **)&d_qcart, c->mesh.qcart.data(), sizeof(double) * 3 * nfull));
    TRY(up("p2_lv2", (void **)&d_lv2, c->t_lv2.data(), sizeof(double) * 3 * c->nfc3));
    TRY(up("p2_lv3", (void **)&d_lv3, c->t_lv3.data(), sizeof(double) * 3 * c->nfc3));
    TRY(up("p2_mwm", (void **)&d_mwm, c->t_mwm.data(), sizeof(double) * 27 * c->nfc3));
    TRY(c->pool_get("p2_gpart", sizeof(double) * (size_t)n * nb, (void **)&d_gpart));
    TRY(hipStreamSynchronize(st));   // the host vectors above are sources of the uploads
    const size_t es = mfma_elem_size(precision);
    const long batch = std::min(n, 16);
    MfmaPlan mp{};
    mp.na = na; mp.nb = nb; mp.nirr = nirr; mp.nfull = nfull; mp.maxa1 = std::max(maxa1, 1);
    mp.maxblk = std::max(maxblk, 1); mp.nty = 4; mp.ntrip = n; mp.batch = batch;
    mp.egv_irr = c->d_egv_irr; mp.egv_full = c->d_egv_full; mp.om_irr = c->d_om_irr; mp.om_full = c->d_om_full;
    TRY(c->pool_get("p2_E1re", es * (size_t)nirr * nb * nb, &mp.E1re));
    TRY(c->pool_get("p2_E1im", es * (size_t)nirr * nb * nb, &mp.E1im));
    TRY(c->pool_get("p2_Efre", es * (size_t)nfull * nb * nb, &mp.Efre));
    TRY(c->pool_get("p2_Efim", es * (size_t)nfull * nb * nb, &mp.Efim));
    if (precision == TC3_PREC_BF16X3) {
        TRY(c->pool_get("p2_Eti", 2 * mfma_split_elems(nb) * nirr, &mp.Eti));
        TRY(c->pool_get("p2_Etf", 2 * mfma_split_elems(nb) * nfull, &mp.Etf));
    }
    const size_t nb3 = (size_t)nb * nb * nb;
    TRY(c->pool_get("p2_work", 2 * nb3 * es * batch, &mp.work));
    mp.tq2 = d_tq2; mp.tq3 = d_tq3; mp.a1off = d_a1off;
    mp.blist = d_blist; mp.a1_boff = d_a1boff; mp.a12_boff = d_a12boff;
    mp.qcart = d_qcart; mp.lv2 = d_lv2; mp.lv3 = d_lv3; mp.mwm = d_mwm;
    mp.ev = nullptr;
    MfmaEpiArgs &me = mp.epi;
    me.nb = nb; me.nT = 1; me.nirr = nirr; me.integrationtype = 1; me.gT = 1; me.T0 = 0;
    me.sigma_fixed = 1.0;
    // threephonon_prefactor * mcg%weight of the full grid (scattering_threephonon.f90:112)
    me.psi_scale = (kPi / 16.0) / (double)nfull;
    me.tq1 = d_tq1; me.tq2 = d_tq2; me.tq3 = d_tq3; me.eq_off = d_eqo; me.eq_q2 = d_eqo; me.eq_q3 = d_eqo;
    me.irr_of_full = c->d_irr_of_full; me.row_of = d_rowof;
    me.perm = d_perm; me.om_irr = c->d_om_irr; me.om_full = c->d_om_full; me.be = d_be; me.sigsq = d_sg;
    me.xi = nullptr; me.gpart = d_gpart; me.xi_row = 0; me.xi_T = 0;
    me.comp = nullptr; me.comp_T = 0; me.u_irr = nullptr; me.u_full = nullptr; me.thresh6 = 0.0; me.tw = nullptr;
    TRY(c->pool_get("p2_out", es * nb3 * n, &me.psi_out));
    const int r = run_mfma_psi2(mp, precision, st);
    if (r) return r == 5 ? c->fail(5, "IFC blocks per atom exceed the MFMA kernels' LDS") : r;
    if (precision == TC3_PREC_FP64) {
        TRY(hipMemcpyAsync(out, me.psi_out, sizeof(double) * nb3 * n, hipMemcpyDeviceToHost, st));
        TRY(hipStreamSynchronize(st));
    } else {
        std::vector<float> h(nb3 * n);
        TRY(hipMemcpyAsync(h.data(), me.psi_out, sizeof(float) * h.size(), hipMemcpyDeviceToHost, st));
        TRY(hipStreamSynchronize(st));
        for (size_t i = 0; i < h.size(); ++i) out[i] = (double)h[i];
    }
#undef TRY
    return 0;
}

int run_scatter(tc3_ctx *c, int precision, int integrationtype, double sigma) {
    const int na = c->cr.na, nb = 3 * na;
    const int nirr = c->mesh.nirr, nfull = c->mesh.nfull, nT = c->nT;
    const long ntrip = (long)c->trip.q1.size();
    const int nrows = c->nrows;
    hipStream_t st = c->stream;
    const auto wall0 = std::chrono::steady_clock::now();
    // device tables: kept in the context's pool, uploaded again only when the model changed
    std::vector<void *> tmp;
    auto cleanup = [&]() {
        for (void *p : tmp) hipFree(p);
        tmp.clear();
    };
    hipError_t e = hipSuccess;
#define TRY(x) do { e = (x); if (e != hipSuccess) { cleanup(); return device_error(e, #x, __FILE__, __LINE__); } } while (0)
    const bool upload = c->scatter_tables_gen != c->gen_model;
    auto table = [&](const char *name, void **p, const void *src, size_t bytes) -> hipError_t {
        bool fresh = false;
        hipError_t r = c->pool_get(name, bytes, p, &fresh);
        if (r == hipSuccess && (upload || fresh) && bytes > 0)
            r = hipMemcpyAsync(*p, src, bytes, hipMemcpyHostToDevice, st);
        return r;
    };
    int *d_tq1, *d_tq2, *d_tq3, *d_eqoff, *d_eq2, *d_eq3, *d_rowof, *d_rowq, *d_rowb, *d_q1off;
    int *d_orboff, *d_orbq1, *d_members, *d_a1off, *d_blk3;
    double *d_perm, *d_qcart, *d_be, *d_sigsq, *d_lv2, *d_lv3, *d_mwm, *d_gpart, *d_lw;
    TRY(table("tq1", (void **)&d_tq1, c->trip.q1.data(), sizeof(int) * ntrip));
    TRY(table("tq2", (void **)&d_tq2, c->trip.q2.data(), sizeof(int) * ntrip));
    TRY(table("tq3", (void **)&d_tq3, c->trip.q3.data(), sizeof(int) * ntrip));
    TRY(table("perm", (void **)&d_perm, c->trip.perm.data(), sizeof(double) * ntrip));
    TRY(table("eqoff", (void **)&d_eqoff, c->trip.eq_off.data(), sizeof(int) * (ntrip + 1)));
    TRY(table("eq2", (void **)&d_eq2, c->trip.eq_q2.data(), sizeof(int) * c->trip.eq_q2.size()));
    TRY(table("eq3", (void **)&d_eq3, c->trip.eq_q3.data(), sizeof(int) * c->trip.eq_q3.size()));
    TRY(table("rowof", (void **)&d_rowof, c->row_of.data(), sizeof(int) * nirr * nb));
    TRY(table("rowq", (void **)&d_rowq, c->row_q.data(), sizeof(int) * nrows));
    TRY(table("rowb", (void **)&d_rowb, c->row_b.data(), sizeof(int) * nrows));
    TRY(table("q1off", (void **)&d_q1off, c->trip.q1_off.data(), sizeof(int) * (nirr + 1)));
    TRY(table("orboff", (void **)&d_orboff, c->orb.orb_off.data(), sizeof(int) * c->orb.orb_off.size()));
    TRY(table("orbq1", (void **)&d_orbq1, c->orb.q1_off.data(), sizeof(int) * c->orb.q1_off.size()));
    TRY(table("members", (void **)&d_members, c->orb.members.data(), sizeof(int) * c->orb.members.size()));
    int *d_orbns;
    TRY(table("orbns", (void **)&d_orbns, c->orb.nsingle.data(), sizeof(int) * nirr));
    // chunks of whole orbits of every q1 for k_xi_sym_lds: at most capm
    // members (kXiChunk entries, or the largest orbit)
    if (c->xchk_gen != c->gen_model) {
        std::vector<int> &chk_off = c->xchk_off, &chk = c->xchk;
        int &capm = c->xcapm;
        chk_off.assign(nirr + 1, 0);
        chk.clear();
        int maxorb = 1;
        for (size_t o = 0; o + 1 < c->orb.orb_off.size(); ++o)
            maxorb = std::max(maxorb, c->orb.orb_off[o + 1] - c->orb.orb_off[o]);
        capm = std::max(maxorb, kXiChunk / std::max(nb, 1));
        for (int q = 0; q < nirr; ++q) {
            chk_off[q] = (int)chk.size();
            const int oa = c->orb.q1_off[q], ob = c->orb.q1_off[q + 1];
            chk.push_back(oa);
            for (int o = oa; o < ob; ++o)
                if (c->orb.orb_off[o + 1] - c->orb.orb_off[chk.back()] > capm) chk.push_back(o);
            if (chk.back() != ob) chk.push_back(ob);
        }
        chk_off[nirr] = (int)chk.size();
        c->xchk_gen = c->gen_model;
    }
    const int capm = c->xcapm;
    const std::vector<int> &chk_off = c->xchk_off, &chk = c->xchk;
    int *d_chkoff, *d_chk;
    TRY(table("xchkoff", (void **)&d_chkoff, chk_off.data(), sizeof(int) * chk_off.size()));
    TRY(table("xchk", (void **)&d_chk, chk.data(), sizeof(int) * chk.size()));
    TRY(table("a1off", (void **)&d_a1off, c->a1_off.data(), sizeof(int) * (na + 1)));
    TRY(table("blk3", (void **)&d_blk3, c->blk3_off.data(), sizeof(int) * ((size_t)na * na * na + 1)));
    TRY(table("qcart", (void **)&d_qcart, c->mesh.qcart.data(), sizeof(double) * 3 * nfull));
    TRY(table("lv2", (void **)&d_lv2, c->t_lv2.data(), sizeof(double) * 3 * c->nfc3));
    TRY(table("lv3", (void **)&d_lv3, c->t_lv3.data(), sizeof(double) * 3 * c->nfc3));
    TRY(table("mwm", (void **)&d_mwm, c->t_mwm.data(), sizeof(double) * 27 * c->nfc3));
    // per call: occupations, smearing, partial sums
    TRY(c->pool_get("be", sizeof(double) * nT * nirr * nb, (void **)&d_be));
    TRY(c->pool_get("sigsq", sizeof(double) * nirr * nb, (void **)&d_sigsq));
    TRY(c->pool_get("gpart", sizeof(double) * std::max<long>(ntrip, 1) * nT * nb, (void **)&d_gpart));
    TRY(c->pool_get("lw", sizeof(double) * nrows * nT, (void **)&d_lw));
    // per-call transfers go through the context's pinned staging buffer: occupations and
    // widths up, linewidths (and the isotope term) down
    const size_t n_be = (size_t)nT * nirr * nb, n_sg = (size_t)nirr * nb, n_lw = (size_t)nrows * nT;
    TRY(c->pinned(sizeof(double) * (nT + n_sg + n_lw + nrows)));
    double *h_T = (double *)c->pin, *h_sg = h_T + nT, *h_lw = h_sg + n_sg, *h_iso = h_lw + n_lw;
    std::memcpy(h_T, c->temps.data(), sizeof(double) * nT);
    std::memcpy(h_sg, c->sigsq.data(), sizeof(double) * n_sg);
    double *d_temps;
    TRY(c->pool_get("temps", sizeof(double) * nT, (void **)&d_temps));
    TRY(hipMemcpyAsync(d_temps, h_T, sizeof(double) * nT, hipMemcpyHostToDevice, st));
    TRY(hipMemcpyAsync(d_sigsq, h_sg, sizeof(double) * n_sg, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_occupations, dim3((unsigned)std::min<size_t>(1024, (n_be + 255) / 256)), dim3(256), 0, st, nT,
                       nirr * nb, d_temps, c->d_om_irr, (int)c->classical, d_be);
    TRY(hipGetLastError());
    const bool use_mfma = mfma_supported(nb, precision) && std::getenv("TC3_NO_MFMA") == nullptr;
    c->xi_row = (size_t)nfull * nb;
    c->xi_es = (precision == TC3_PREC_FP64) ? 8 : 4;
    // compact per-triplet sums, assembled into the rows by k_xi_asm, when every entry has one
    // owner (TC3_XI_DIRECT: direct stores); k_xi_asm addresses one temperature's sums and one
    // row with 32-bit byte offsets
    const size_t comp_T = (size_t)ntrip * nb * 2 * nb;
    bool comp_ok = false;
    if (use_mfma && c->thirdorder) {
        if (c->xi_src_gen != c->gen_model) {
            c->xi_src_ok = xi_source_map(c->trip, c->orb, nirr, nfull, nb, c->xi_src);
            c->xi_src_gen = c->gen_model;
        }
        comp_ok = c->xi_src_ok && (double)c->xi_es * comp_T < 4.0e9 && (double)c->xi_es * c->xi_row < 4.0e9 &&
                  std::getenv("TC3_XI_DIRECT") == nullptr;
    }
    // isotope scattering (scattering.f90:229-232): on when requested and some atom has mass
    // disorder
    bool use_iso = false;
    for (double g : c->iso_g) use_iso = use_iso || (c->iso_on && g > 0.0);
    // rows stored per orbit when k_xi_asm writes the averaged rows and nothing is added to
    // them afterwards (no isotope or four-phonon terms); TC3_XI_FULL keeps the full rows
    c->xi_orb = comp_ok && !use_iso && c->nfc4 == 0 && std::getenv("TC3_XI_FULL") == nullptr &&
                build_orbit_layout(c);
    // scattering matrix: every entry that any triplet owns is rewritten by every pass, so
    // the full rows are zeroed only when they are new or the model changed; k_xi_asm writes
    // every element of the rows stored per orbit
    const size_t xi_T = c->xi_orb ? c->xo_T : c->xi_row * nrows;
    if (c->xi_bytes < (size_t)c->xi_es * xi_T * nT) {
        hipFree(c->d_xi);
        c->d_xi = nullptr;
        c->xi_bytes = 0;
        TRY(hipMalloc(&c->d_xi, (size_t)c->xi_es * xi_T * nT));
        c->xi_bytes = (size_t)c->xi_es * xi_T * nT;
        c->xi_gen = -1;
    }
    if (c->xi_orb) {
        c->xi_gen = -1;
    } else if (c->xi_gen != c->gen_model || c->xi_es_zeroed != c->xi_es) {
        TRY(hipMemsetAsync(c->d_xi, 0, c->xi_bytes, st));
        c->xi_gen = c->gen_model;
        c->xi_es_zeroed = c->xi_es;
    }
    c->scatter_tables_gen = c->gen_model;

    // generic path: batch of intermediates of 4 * nb^3 complex per triplet
    const long nb3 = (long)nb * nb * nb;
    const size_t elem = (precision == TC3_PREC_FP64) ? sizeof(cd) : sizeof(cf);
    const size_t per_trip = 4 * (size_t)nb3 * elem;
    long batch = std::max<long>(1, std::min<long>(ntrip, (long)((size_t)2 << 30) / (long)per_trip));
    batch = std::min<long>(batch, 1 << 20);
    // small cells go to the persistent kernel when its per-wave LDS fits (4 waves per
    // workgroup while at least two workgroups fit a CU, fewer otherwise)
    size_t small_lds = 0;
    int small_waves = 0;
    if (!use_mfma && (nb == 3 || nb == 6) && std::getenv("TC3_NO_SMALL") == nullptr) {
        small_lds = (nb == 3) ? ((precision == TC3_PREC_FP64) ? small_wave_lds<3, double>(c->nfc3)
                                                              : small_wave_lds<3, float>(c->nfc3))
                              : ((precision == TC3_PREC_FP64) ? small_wave_lds<6, double>(c->nfc3)
                                                              : small_wave_lds<6, float>(c->nfc3));
        for (int nw = 4; nw >= 1 && small_waves == 0; --nw)
            if (small_lds * nw <= 80 * 1024 || (nw == 1 && small_lds <= 160 * 1024)) small_waves = nw;
    }
    const bool use_small = small_waves > 0;
    void *d_work = nullptr;
    void *d_E1 = nullptr, *d_Ef = nullptr;
    if (!use_mfma) {
        if (!use_small) TRY(c->pool_get("g_work", per_trip * batch, &d_work));
        TRY(c->pool_get("g_E1", elem * (size_t)nirr * nb * nb, &d_E1));
        TRY(c->pool_get("g_Ef", elem * (size_t)nfull * nb * nb, &d_Ef));
        if (precision == TC3_PREC_FP64) {
            hipLaunchKernelGGL(k_scale<double>, dim3(1024), dim3(256), 0, st, (long)nirr * nb * nb, nb, c->d_egv_irr,
                               c->d_om_irr, (cplx<double> *)d_E1);
            hipLaunchKernelGGL(k_scale<double>, dim3(4096), dim3(256), 0, st, (long)nfull * nb * nb, nb, c->d_egv_full,
                               c->d_om_full, (cplx<double> *)d_Ef);
        } else {
            hipLaunchKernelGGL(k_scale<float>, dim3(1024), dim3(256), 0, st, (long)nirr * nb * nb, nb, c->d_egv_irr,
                               c->d_om_irr, (cplx<float> *)d_E1);
            hipLaunchKernelGGL(k_scale<float>, dim3(4096), dim3(256), 0, st, (long)nfull * nb * nb, nb, c->d_egv_full,
                               c->d_om_full, (cplx<float> *)d_Ef);
        }
        TRY(hipGetLastError());
    }
    int maxa1 = 0;
    for (int a = 0; a < na; ++a) maxa1 = std::max(maxa1, c->a1_off[a + 1] - c->a1_off[a]);
    const size_t phi_shmem = sizeof(cd) * std::max(maxa1, 1);

    // integration type 6: reduced group velocities of the irreducible and full meshes
    double *d_uirr = nullptr, *d_ufull = nullptr;
    if (integrationtype == 6) {
        Mat3 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[3 * i + j] = c->cr.rec[3 * i + j] / c->mesh.dims[j];
        TRY(c->pool_get("u_irr", sizeof(double) * 3 * nirr * nb, (void **)&d_uirr));
        TRY(c->pool_get("u_full", sizeof(double) * 3 * nfull * nb, (void **)&d_ufull));
        hipLaunchKernelGGL(k_reduced_vel, dim3(256), dim3(256), 0, st, (long)nirr * nb, c->d_vel_irr, r, d_uirr);
        hipLaunchKernelGGL(k_reduced_vel, dim3(1024), dim3(256), 0, st, (long)nfull * nb, c->d_vel_full, r, d_ufull);
        TRY(hipGetLastError());
    }
    EpiArgs ea;
    ea.nb = nb; ea.nT = nT; ea.nfull = nfull; ea.integrationtype = integrationtype;
    ea.u_irr = d_uirr; ea.u_full = d_ufull; ea.thresh6 = c->thresh6;
    ea.sigma_fixed = sigma * kFreqThzToHartree;
    // weight of a draw: 1 / N_q, or 1 / npoints of the Monte-Carlo grid (mcg%weight)
    ea.psi_scale = (kPi / 16.0) / (double)(c->trip_w.empty() ? (long)nfull : c->mc_npoints);
    ea.tw = nullptr;
    if (!c->trip_w.empty())
        TRY(table("tw", (void **)&ea.tw, c->trip_w.data(), sizeof(float) * c->trip_w.size()));
    ea.tq1 = d_tq1; ea.tq2 = d_tq2; ea.tq3 = d_tq3; ea.eq_off = d_eqoff; ea.eq_q2 = d_eq2; ea.eq_q3 = d_eq3;
    ea.irr_of_full = c->d_irr_of_full; ea.row_of = d_rowof;
    ea.perm = d_perm; ea.om_irr = c->d_om_irr; ea.om_full = c->d_om_full; ea.be = d_be; ea.sigsq = d_sigsq;
    ea.xi = c->d_xi; ea.gpart = d_gpart; ea.xi_row = c->xi_row; ea.xi_T = xi_T; ea.nirr = nirr;
    const size_t epi_shmem = sizeof(double) * ((size_t)nb * nb + 2 * nb);

    int *d_xisrc = nullptr;
    void *d_comp = nullptr;
    // events come from the context's pool (created once, kept until the context ends)
    auto pooled_events = [&](size_t n) -> hipError_t {
        while (c->ev_pool.size() < n) {
            hipEvent_t e;
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            c->ev_pool.push_back(e);
        }
        return hipSuccess;
    };
    auto take_events = [&](std::vector<hipEvent_t> &v, size_t n) -> hipError_t {
        const hipError_t r = pooled_events(4 + n);
        if (r == hipSuccess) v.assign(c->ev_pool.begin() + 4, c->ev_pool.begin() + 4 + n);
        return r;
    };
    TRY(pooled_events(4));
    hipEvent_t ev0 = c->ev_pool[0], ev1 = c->ev_pool[1];
    std::vector<hipEvent_t> kev;
    c->ks_pending = false;
    for (double &x : c->kstat_ms) x = 0;
    for (int &x : c->kstat_n) x = 0;
    if (!c->thirdorder) {
        // no three-phonon term (--nothirdorder, scattering.f90:234-238): zero partial sums and
        // scattering matrix, to which the isotope, four-phonon and boundary terms are added
        TRY(hipEventRecord(ev0, st));
        TRY(hipMemsetAsync(d_gpart, 0, sizeof(double) * std::max<long>(ntrip, 1) * nT * nb, st));
        TRY(hipMemsetAsync(c->d_xi, 0, (size_t)c->xi_es * xi_T * nT, st));
        c->batch = 0;
    } else if (use_mfma) {
        std::vector<int4> blist;
        std::vector<int> a1_boff, a12_boff;
        int maxblk = 0;
        mfma_block_lists(c, blist, a1_boff, a12_boff, maxblk);
        int4 *d_blist;
        int *d_a1boff, *d_a12boff;
        TRY(table("blist", (void **)&d_blist, blist.data(), sizeof(int4) * blist.size()));
        TRY(table("a1boff", (void **)&d_a1boff, a1_boff.data(), sizeof(int) * a1_boff.size()));
        TRY(table("a12boff", (void **)&d_a12boff, a12_boff.data(), sizeof(int) * a12_boff.size()));
        // W of one triplet: 2 * nb^3 reals.  Triplets per launch pair: at most 8 GB of W;
        // TC3_BATCH_MB forces a size, otherwise the autotuner below picks it.
        const size_t es = mfma_elem_size(precision);
        const size_t wtrip = 2 * (size_t)nb3 * es;
        const long bmax = std::max<long>(1, std::min<long>(ntrip, (long)((size_t)8 << 30) / (long)wtrip));
        long mbatch = bmax;
        const char *env_mb = std::getenv("TC3_BATCH_MB");
        if (env_mb) mbatch = std::max<long>(1, std::min<long>(bmax, (std::max(1l, std::atol(env_mb)) << 20) / (long)wtrip));
        MfmaPlan mp;
        mp.na = na; mp.nb = nb; mp.nirr = nirr; mp.nfull = nfull; mp.maxa1 = std::max(maxa1, 1);
        mp.maxblk = std::max(maxblk, 1);
        mp.nty = mfma_temperature_block(nb, nT, precision);
        if (mp.nty == 0) { cleanup(); return c->fail(5, "band count too large for the MFMA kernels' LDS"); }
        mp.ntrip = ntrip; mp.batch = mbatch;
        mp.egv_irr = c->d_egv_irr; mp.egv_full = c->d_egv_full; mp.om_irr = c->d_om_irr; mp.om_full = c->d_om_full;
        if (const char *eb = std::getenv("TC3_PSI_BB"))
            mp.psi_bb = (precision == TC3_PREC_BF16X3 && nb == 48 && std::atoi(eb) == nb / 16) ? nb / 16 : 1;
        // the fused kernel (48 bands, split-BF16x3, compact sums) needs neither W nor the
        // chain's operand planes
        // IFC blocks (a1, a2, a3) in entry order and the (a2, a3) pairs with their blocks, for the
        // fused kernel: its Phi slot holds, per pair, per (j, k), per block, the three i values
        if (c->fused_gen != c->gen_model) {
            std::vector<std::vector<int>> per((size_t)na * na);   // block indices of every pair
            std::vector<int4> blocks;
            std::vector<int> a1of;
            int maxe = 0;
            for (int a1 = 0; a1 < na; ++a1)
                for (int a2 = 0; a2 < na; ++a2)
                    for (int a3 = 0; a3 < na; ++a3) {
                        const size_t bk = ((size_t)a1 * na + a2) * na + a3;
                        if (c->blk3_off[bk + 1] == c->blk3_off[bk]) continue;
                        per[(size_t)a2 * na + a3].push_back((int)blocks.size());
                        blocks.push_back(make_int4(c->blk3_off[bk], c->blk3_off[bk + 1], 0, 0));
                        a1of.push_back(a1);
                        maxe = std::max(maxe, c->blk3_off[bk + 1] - c->blk3_off[bk]);
                    }
            std::vector<int4> pairs;
            std::vector<int> pa1;
            int base = 0, maxb = 0;
            // pairs by decreasing number of blocks (the kernel's lanes then loop over similar counts)
            std::vector<int> order;
            for (int x = 0; x < na * na; ++x)
                if (!per[x].empty()) order.push_back(x);
            std::stable_sort(order.begin(), order.end(),
                             [&](int x, int y) { return per[x].size() > per[y].size(); });
            for (int x : order) {
                {
                    const int a2 = x / na, a3 = x % na;
                    const auto &v = per[x];
                    const int nbp = (int)v.size();
                    pairs.push_back(make_int4(a2, a3, base, nbp));
                    for (int bb = 0; bb < nbp; ++bb) {
                        // value (bb, i, j, k) of the pair at base + 27 bb + 9 i + 3 j + k
                        blocks[v[bb]].z = base + 27 * bb;
                        blocks[v[bb]].w = 0;
                        pa1.push_back(a1of[v[bb]]);
                    }
                    base += 27 * nbp;
                    maxb = std::max(maxb, nbp);
                }
            }
            c->fused_blocks = blocks;
            c->fused_pairs = pairs;
            c->fused_pa1 = pa1;
            c->fused_maxe = maxe;
            c->fused_maxb = maxb;
            c->fused_gen = c->gen_model;
        }
        // opt-in (TC3_FUSED=1): no W round trip, but with one wave per SIMD (its LDS) it is
        // latency-bound and slower than the two kernels below
        const bool fused = comp_ok &&
                           fused_supported(nb, precision, integrationtype, c->nfc3, (int)c->fused_blocks.size(),
                                           (int)c->fused_pairs.size(), c->fused_maxe, c->fused_maxb) &&
                           std::getenv("TC3_FUSED") != nullptr;
        mp.Eti = mp.Etf = nullptr;
        if (!fused) {
            TRY(c->pool_get("m_E1re", es * (size_t)nirr * nb * nb, &mp.E1re));
            TRY(c->pool_get("m_E1im", es * (size_t)nirr * nb * nb, &mp.E1im));
            TRY(c->pool_get("m_Efre", es * (size_t)nfull * nb * nb, &mp.Efre));
            TRY(c->pool_get("m_Efim", es * (size_t)nfull * nb * nb, &mp.Efim));
            if (precision == TC3_PREC_BF16X3) {
                TRY(c->pool_get("m_Eti", 2 * mfma_split_elems(nb) * nirr, &mp.Eti));
                TRY(c->pool_get("m_Etf", 2 * mfma_split_elems(nb) * nfull, &mp.Etf));
            }
            TRY(c->pool_get("m_work", wtrip * bmax, &mp.work));
        }
        mp.tq2 = d_tq2; mp.tq3 = d_tq3; mp.a1off = d_a1off;
        mp.blist = d_blist; mp.a1_boff = d_a1boff; mp.a12_boff = d_a12boff;
        mp.qcart = d_qcart; mp.lv2 = d_lv2; mp.lv3 = d_lv3; mp.mwm = d_mwm;
        MfmaEpiArgs &me = mp.epi;
        me.nb = nb; me.nirr = nirr; me.integrationtype = integrationtype; me.gT = nT;
        me.sigma_fixed = ea.sigma_fixed; me.psi_scale = ea.psi_scale;
        me.tq1 = d_tq1; me.tq2 = d_tq2; me.tq3 = d_tq3; me.eq_off = d_eqoff; me.eq_q2 = d_eq2; me.eq_q3 = d_eq3;
        me.irr_of_full = c->d_irr_of_full; me.row_of = d_rowof;
        me.u_irr = d_uirr; me.u_full = d_ufull; me.thresh6 = c->thresh6;
        me.tw = ea.tw;
        me.perm = d_perm; me.om_irr = c->d_om_irr; me.om_full = c->d_om_full; me.sigsq = d_sigsq;
        me.gpart = d_gpart; me.xi_row = c->xi_row; me.xi_T = xi_T;
        me.comp = nullptr;
        me.comp_T = comp_T;
        if (comp_ok) {
            TRY(table("xisrc", (void **)&d_xisrc, c->xi_src.data(), sizeof(int4) * (c->xi_src.size() / 4)));
            TRY(c->pool_get("xi_comp", (size_t)c->xi_es * nT * me.comp_T, &me.comp));
            d_comp = me.comp;
        }
        if (fused) {
            int4 *d_fblocks, *d_fpairs;
            int *d_fpa1;
            TRY(table("f_blocks", (void **)&d_fblocks, c->fused_blocks.data(), sizeof(int4) * c->fused_blocks.size()));
            TRY(table("f_pairs", (void **)&d_fpairs, c->fused_pairs.data(), sizeof(int4) * c->fused_pairs.size()));
            TRY(table("f_pa1", (void **)&d_fpa1, c->fused_pa1.data(), sizeof(int) * c->fused_pa1.size()));
            int ncu = 0;
            TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
            FusedPlan fp;
            fp.nirr = nirr; fp.nfull = nfull; fp.nT = nT; fp.nty = fused_temperature_block(nT);
            fp.nwg = std::max(ncu, 1);
            fp.egv_irr = c->d_egv_irr; fp.egv_full = c->d_egv_full; fp.om_irr = c->d_om_irr; fp.om_full = c->d_om_full;
            fp.be = d_be;
            TRY(c->pool_get("f_tab", fused_table_bytes(nfull), (void **)&fp.ftab));
            TRY(c->pool_get("f_e1re", sizeof(float) * (size_t)nirr * nb * nb, (void **)&fp.e1re));
            TRY(c->pool_get("f_e1im", sizeof(float) * (size_t)nirr * nb * nb, (void **)&fp.e1im));
            fp.comp = d_comp;
            FusedArgs &fa = fp.args;
            fa.ntrip = (int)ntrip; fa.nirr = nirr; fa.gT = nT; fa.integrationtype = integrationtype;
            fa.nfc3 = c->nfc3; fa.nblk = (int)c->fused_blocks.size(); fa.npair = (int)c->fused_pairs.size();
            fa.npblk = (int)c->fused_pa1.size();
            fa.sigma_fixed = ea.sigma_fixed; fa.psi_scale = ea.psi_scale;
            fa.tq1 = d_tq1; fa.tq2 = d_tq2; fa.tq3 = d_tq3; fa.eq_off = d_eqoff; fa.irr_of_full = c->d_irr_of_full;
            fa.row_of = d_rowof; fa.perm = d_perm; fa.om_irr = c->d_om_irr; fa.om_full = c->d_om_full;
            fa.sigsq = d_sigsq; fa.qcart = d_qcart; fa.lv2 = d_lv2; fa.lv3 = d_lv3; fa.mwm = d_mwm;
            fa.blocks = d_fblocks; fa.pairs = d_fpairs; fa.pa1 = d_fpa1; fa.ftab = fp.ftab;
            fa.e1re = fp.e1re; fa.e1im = fp.e1im;
            TRY(c->pool_get("f_slots", sizeof(float2) * (size_t)fp.nwg * std::max(fa.nblk, 1) * 27, (void **)&fa.phi_slot));
            TRY(c->pool_get("f_queue", sizeof(int), (void **)&fa.queue));
            fa.comp_T = comp_T; fa.gpart = d_gpart; fa.tw = ea.tw;
            TRY(take_events(kev, 3));
            fp.ev = kev.data();
            TRY(hipEventRecord(ev0, st));
            TRY(hipEventRecord(kev[0], st));
            const int r = run_fused(fp, st);
            if (r) { cleanup(); return r; }
            c->batch = ntrip;
        } else {
        // Autotuner: on the first pass of a shape (band count, precision, temperature block,
        // triplet count), time each candidate batch (the largest, then halving, at least 64)
        // on the same first triplets (as many as the largest batch) and keep the fastest.
        // The probes compute real results for those triplets; the pass below rewrites them.
        if (!env_mb) {
            const std::string key = std::to_string(nb) + "/" + std::to_string(precision) + "/" +
                                    std::to_string(mp.nty) + "/" + std::to_string(ntrip);
            // the workgroup shape of k_psi: one b1 block (the default) or the whole triplet
            // (split-BF16x3, 48 bands), timed at the best batch unless TC3_PSI_BB forces one
            const bool try_bb = precision == TC3_PREC_BF16X3 && nb == 48 && std::getenv("TC3_PSI_BB") == nullptr;
            auto it = c->tuned_batch.find(key);
            if (it != c->tuned_batch.end()) {
                mbatch = std::min<long>(it->second, bmax);
                if (try_bb) mp.psi_bb = c->tuned_bb[key];
            } else {
                std::vector<long> cand;
                for (long b = bmax; cand.size() < 4 && (b >= 64 || cand.empty()); b /= 2) cand.push_back(b);
                hipEvent_t p0, p1;
                TRY(hipEventCreate(&p0));
                TRY(hipEventCreate(&p1));
                double best = 1e300;
                MfmaPlan pp = mp;
                pp.ev = nullptr;
                pp.epi.T0 = 0;
                pp.epi.nT = std::min(mp.nty, nT);
                pp.epi.be = d_be;
                pp.epi.xi = c->d_xi;
                // one untimed launch pair first: code loading and cold caches are not the
                // candidates' cost
                pp.ntrip = std::min<long>(ntrip, 64);
                pp.batch = pp.ntrip;
                {
                    int r = run_mfma_chain(pp, precision, st);
                    if (r) { hipEventDestroy(p0); hipEventDestroy(p1); cleanup(); return r; }
                }
                for (long b : cand) {
                    // every candidate runs the same triplets (work per triplet varies with q1);
                    // best of two timings
                    pp.ntrip = bmax;
                    pp.batch = b;
                    float msb = 1e30f;
                    for (int rep = 0; rep < 2; ++rep) {
                        TRY(hipEventRecord(p0, st));
                        int r = run_mfma_chain(pp, precision, st);
                        if (r) { hipEventDestroy(p0); hipEventDestroy(p1); cleanup(); return r; }
                        TRY(hipEventRecord(p1, st));
                        TRY(hipEventSynchronize(p1));
                        float ms = 0;
                        hipEventElapsedTime(&ms, p0, p1);
                        msb = std::min(msb, ms);
                    }
                    c->tune_log.push_back({key, (double)b, (double)msb});
                    if (msb < best) { best = msb; mbatch = b; }
                }
                if (try_bb) {
                    MfmaPlan pb = pp;
                    pb.psi_bb = nb / 16;
                    // untimed first launch of the other kernel, then best of two at the best batch
                    pb.ntrip = std::min<long>(ntrip, 64);
                    pb.batch = pb.ntrip;
                    int r = run_mfma_chain(pb, precision, st);
                    pb.ntrip = bmax;
                    pb.batch = mbatch;
                    float msb = 1e30f;
                    for (int rep = 0; r == 0 && rep < 2; ++rep) {
                        TRY(hipEventRecord(p0, st));
                        r = run_mfma_chain(pb, precision, st);
                        TRY(hipEventRecord(p1, st));
                        TRY(hipEventSynchronize(p1));
                        float ms = 0;
                        hipEventElapsedTime(&ms, p0, p1);
                        msb = std::min(msb, ms);
                    }
                    if (r) { hipEventDestroy(p0); hipEventDestroy(p1); cleanup(); return r; }
                    c->tune_log.push_back({key + "/whole-triplet k_psi", (double)mbatch, (double)msb});
                    c->tuned_bb[key] = (msb < best) ? nb / 16 : 1;
                    mp.psi_bb = c->tuned_bb[key];
                }
                hipEventDestroy(p0);
                hipEventDestroy(p1);
                c->tuned_batch[key] = mbatch;
            }
        }
        mp.batch = mbatch;
        const long nbatch = (ntrip + mbatch - 1) / mbatch;
        const int ntchunk = (nT + mp.nty - 1) / mp.nty;
        TRY(take_events(kev, 3 * nbatch * ntchunk));
        TRY(hipEventRecord(ev0, st));
        // temperatures in blocks of mp.nty; each block recomputes the contraction
        auto chunk = [&](int tc) {
            const int T0 = tc * mp.nty;
            me.T0 = T0;
            me.nT = std::min(mp.nty, nT - T0);
            me.be = d_be + (size_t)T0 * nirr * nb;
            me.xi = (char *)c->d_xi + (size_t)c->xi_es * T0 * xi_T;
            if (d_comp) me.comp = (char *)d_comp + (size_t)c->xi_es * T0 * me.comp_T;
            mp.ev = kev.data() + 3 * nbatch * tc;
        };
        auto chain = [&]() -> int {
            for (int tc = 0; tc < ntchunk; ++tc) {
                chunk(tc);
                const int r = run_mfma_chain(mp, precision, st);
                if (r) return r;
            }
            return 0;
        };
        auto chain_noev = [&]() -> int {
            for (int tc = 0; tc < ntchunk; ++tc) {
                chunk(tc);
                mp.ev = nullptr;
                const int r = run_mfma_chain(mp, precision, st);
                if (r) return r;
            }
            return 0;
        };
        // the launches of the pass (operand preparation, every launch pair with its events) are
        // captured once as a HIP graph and replayed while every argument they take stays the
        // same; direct launches where capture fails (TC3_NO_GRAPH: always)
        std::vector<long> gkey;
        for (int tc = 0; tc < ntchunk; ++tc) {
            chunk(tc);
            mp.ev = nullptr;
            chain_key(gkey, mp, precision, 0);
        }
        const bool use_graph = std::getenv("TC3_NO_GRAPH") == nullptr;
        if (c->sc_exec && c->sc_key != gkey) {
            (void)hipGraphExecDestroy(c->sc_exec);
            c->sc_exec = nullptr;
        }
        if (use_graph && !c->sc_exec && gkey != c->sc_key_failed &&
            hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            const int rc = chain_noev();
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &graph);
            if (rc == 0 && ec == hipSuccess && graph &&
                hipGraphInstantiate(&c->sc_exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                c->sc_key = gkey;
            } else {
                c->sc_exec = nullptr;
                c->sc_key_failed = gkey;
            }
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
        }
        if (use_graph && c->sc_exec) {
            // the graph records no per-launch events: no per-kernel statistics for this pass
            TRY(hipGraphLaunch(c->sc_exec, st));
            kev.clear();
        } else {
            const int r = chain();
            if (r) { cleanup(); return r == 5 ? c->fail(5, "IFC blocks per atom exceed the MFMA kernels' LDS") : r; }
        }
        c->batch = mbatch;
        }
    } else if (use_small) {
        // persistent work-queue kernel for small cells
        TRY(hipEventRecord(ev0, st));
        TRY(take_events(kev, 3));
        int *d_queue;
        TRY(c->pool_get("small_queue", sizeof(int), (void **)&d_queue));
        int ncu = 0;
        TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
        const int nwg = std::max(1, ncu) * 8;
        TRY(hipEventRecord(kev[0], st));
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, d_queue);
        TRY(hipEventRecord(kev[1], st));
        const dim3 g(nwg), b(64 * small_waves);
        const size_t shm = small_lds * small_waves;
        auto launch = [&](auto kern, auto E1p, auto Efp) -> hipError_t {
            hipError_t r = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            if (r != hipSuccess) return r;
            hipLaunchKernelGGL(kern, g, b, shm, st, ea, (int)ntrip, d_queue, small_lds, E1p, Efp, d_qcart, d_blk3,
                               d_lv2, d_lv3, d_mwm, c->nfc3);
            return hipGetLastError();
        };
        using C64 = const cplx<double> *;
        using C32 = const cplx<float> *;
        if (nb == 3) {
            if (precision == TC3_PREC_FP64) TRY(launch(k_small<3, double, false>, (C64)d_E1, (C64)d_Ef));
            else if (precision == TC3_PREC_FP32) TRY(launch(k_small<3, float, false>, (C32)d_E1, (C32)d_Ef));
            else TRY(launch(k_small<3, float, true>, (C32)d_E1, (C32)d_Ef));
        } else {
            if (precision == TC3_PREC_FP64) TRY(launch(k_small<6, double, false>, (C64)d_E1, (C64)d_Ef));
            else if (precision == TC3_PREC_FP32) TRY(launch(k_small<6, float, false>, (C32)d_E1, (C32)d_Ef));
            else TRY(launch(k_small<6, float, true>, (C32)d_E1, (C32)d_Ef));
        }
        TRY(hipEventRecord(kev[2], st));
        c->batch = ntrip;
    } else {
    TRY(hipEventRecord(ev0, st));
    TRY(take_events(kev, 3 * ((ntrip + batch - 1) / batch)));
    for (long t0 = 0; t0 < ntrip; t0 += batch) {
        const int nt = (int)std::min<long>(batch, ntrip - t0);
        hipEvent_t *ev = &kev[3 * (t0 / batch)];
        TRY(hipEventRecord(ev[0], st));
        ea.t0 = (int)t0;
        if (precision == TC3_PREC_FP64) {
            auto *w = (cplx<double> *)d_work;
            cplx<double> *phi = w, *t1 = w + (size_t)batch * nb3, *t2 = t1 + (size_t)batch * nb3,
                         *psi = t2 + (size_t)batch * nb3;
            launch_chain<double>(c, st, (int)t0, nt, (const cplx<double> *)d_E1, (const cplx<double> *)d_Ef, phi, t1,
                                 t2, psi, d_tq1, d_tq2, d_tq3, d_qcart, d_a1off, d_blk3, d_lv2, d_lv3, d_mwm,
                                 phi_shmem);
            TRY(hipEventRecord(ev[1], st));
            hipLaunchKernelGGL(k_epilogue<double>, dim3(nt * nb), dim3(kBlock), epi_shmem, st, ea,
                               (const double *)psi, 0);
        } else if (precision == TC3_PREC_FP32 || precision == TC3_PREC_BF16X3) {
            auto *w = (cplx<float> *)d_work;
            cplx<float> *phi = w, *t1 = w + (size_t)batch * nb3, *t2 = t1 + (size_t)batch * nb3,
                        *psi = t2 + (size_t)batch * nb3;
            if (precision == TC3_PREC_FP32)
                launch_chain<float>(c, st, (int)t0, nt, (const cplx<float> *)d_E1, (const cplx<float> *)d_Ef, phi,
                                    t1, t2, psi, d_tq1, d_tq2, d_tq3, d_qcart, d_a1off, d_blk3, d_lv2, d_lv3, d_mwm,
                                    phi_shmem);
            else
                launch_chain<float, true>(c, st, (int)t0, nt, (const cplx<float> *)d_E1, (const cplx<float> *)d_Ef,
                                          phi, t1, t2, psi, d_tq1, d_tq2, d_tq3, d_qcart, d_a1off, d_blk3, d_lv2,
                                          d_lv3, d_mwm, phi_shmem);
            TRY(hipEventRecord(ev[1], st));
            hipLaunchKernelGGL(k_epilogue<float>, dim3(nt * nb), dim3(kBlock), epi_shmem, st, ea,
                               (const float *)psi, 0);
        } else {
            cleanup();
            return c->fail(5, "precision mode not available in this build");
        }
        TRY(hipEventRecord(ev[2], st));
        TRY(hipGetLastError());
    }
    c->batch = batch;
    }
    // four-phonon terms (scattering.f90:239-242): off-diagonal terms into the rows before the
    // symmetrisation below, linewidth terms after the isotope term
    std::vector<double> g4;
    double *d_od4 = nullptr;   // four-phonon terms for the symmetrisation below
    if (c->nfc4 > 0) {
        F4Inputs fi;
        fi.nT = nT; fi.integrationtype = integrationtype; fi.sigma_fixed = ea.sigma_fixed; fi.thresh = c->thresh6;
        fi.precision = precision;
        fi.d_be = d_be; fi.d_sigsq = d_sigsq; fi.d_uirr = d_uirr; fi.d_ufull = d_ufull; fi.d_rowof = d_rowof;
        // the off-diagonal terms [T][row][q nb + b], formed a temperature at a time on the host
        TRY(c->pool_get("f4_od", sizeof(double) * std::max<size_t>((size_t)nT * nrows * nfull * nb, 1),
                        (void **)&d_od4));
        const int r4 = fourphonon_terms(c, fi, g4, d_od4);
        if (r4) { cleanup(); return r4; }
    }
    hipEvent_t evs0 = c->ev_pool[2], evs1 = c->ev_pool[3];
    hipLaunchKernelGGL(k_gamma, dim3((unsigned)nirr, (unsigned)nT), dim3((unsigned)((nb + 63) / 64 * 64)), 0, st, nrows,
                       nT, nb, d_rowof, d_q1off, d_gpart, d_lw);
    TRY(hipGetLastError());
    TRY(hipEventRecord(evs0, st));
    IsoArgs iso{};
    if (use_iso) {
        double *d_g;
        TRY(c->pool_get("iso_g", sizeof(double) * na, (void **)&d_g));
        TRY(c->pool_get("iso_gam", sizeof(double) * std::max(nrows, 1), (void **)&iso.gam));
        TRY(hipMemcpyAsync(d_g, c->iso_g.data(), sizeof(double) * na, hipMemcpyHostToDevice, st));
        iso.g = d_g;
        iso.egv_irr = c->d_egv_irr; iso.egv_full = c->d_egv_full;
        iso.om_irr = c->d_om_irr; iso.om_full = c->d_om_full; iso.sigsq = d_sigsq;
        iso.row_b = d_rowb; iso.irr_of_full = c->d_irr_of_full;
        iso.sigma_fixed = ea.sigma_fixed; iso.pref = (kPi / 4.0) / (double)nfull;
        iso.na = na; iso.integrationtype = integrationtype;
        iso.u_full = d_ufull; iso.thresh6 = c->thresh6;
    }
    {
        const dim3 g(nrows * nT), b(kBlock), bl(nb <= kBlock ? nb * (kBlock / nb) : kBlock);
        const size_t shl = (size_t)capm * nb * c->xi_es + sizeof(int) * (2 * (size_t)capm + 1);
        if (d_comp && nb > kBlock) { cleanup(); return c->fail(5, "scattering-matrix assembly needs nb <= 256"); }
        if (d_comp) {
            // rows from the compact sums; with isotope or four-phonon terms the averaging below
            // follows
            const size_t cT = (size_t)ntrip * nb * 2 * nb;
            const int L = nb / 3;
            const dim3 bo(L * (kBlock / L)), go(nrows * ((nT + kOrbTemps - 1) / kOrbTemps));
            const bool avg = !(use_iso || d_od4);
            long long *d_xorow = nullptr;
            if (c->xi_orb) TRY(table("xo_row", (void **)&d_xorow, c->xo_row.data(), sizeof(long long) * nrows));
            auto launch = [&](auto kf, auto *cp, auto *xp) {
                hipLaunchKernelGGL(kf, go, bo, 0, st, nrows, nb, nfull, nT, d_rowq, d_rowb, c->d_om_irr, c->d_om_full,
                                   (const int4 *)d_xisrc, cp, cT, d_orbq1, d_orboff, d_orbns, xp, c->xi_row, xi_T,
                                   d_xorow);
            };
            if (c->xi_es == 8)
                launch(avg ? k_xi_asm<double, double, true> : k_xi_asm<double, double, false>, (const double *)d_comp,
                       (double *)c->d_xi);
            else
                launch(avg ? k_xi_asm<float, float, true> : k_xi_asm<float, float, false>, (const float *)d_comp,
                       (float *)c->d_xi);
            TRY(hipGetLastError());
        }
        const bool use_lds = nb <= kBlock && shl <= 80 * 1024 && std::getenv("TC3_XI_GLOBAL") == nullptr && !d_od4;
        if (use_lds) {
            const void *kf = (c->xi_es == 8) ? (const void *)k_xi_sym_lds<double> : (const void *)k_xi_sym_lds<float>;
            TRY(hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shl));
        }
        if (d_comp && !use_iso && !d_od4) {
            // averaged by k_xi_asm
        } else if (c->xi_es == 8) {
            if (use_iso)
                hipLaunchKernelGGL((k_xi_sym<double, true>), g, b, 0, st, nrows, nb, d_rowq, d_orbq1, d_orboff,
                                   d_members, (double *)c->d_xi, c->xi_row, xi_T, iso, (const double *)d_od4,
                                   (size_t)nfull * nb);
            else if (use_lds)
                hipLaunchKernelGGL((k_xi_sym_lds<double>), g, bl, shl, st, nrows, nb, d_rowq, d_chkoff, d_chk,
                                   d_orboff, d_members, (double *)c->d_xi, c->xi_row, xi_T, capm);
            else
                hipLaunchKernelGGL((k_xi_sym<double, false>), g, b, 0, st, nrows, nb, d_rowq, d_orbq1, d_orboff,
                                   d_members, (double *)c->d_xi, c->xi_row, xi_T, iso, (const double *)d_od4,
                                   (size_t)nfull * nb);
        } else {
            if (use_iso)
                hipLaunchKernelGGL((k_xi_sym<float, true>), g, b, 0, st, nrows, nb, d_rowq, d_orbq1, d_orboff,
                                   d_members, (float *)c->d_xi, c->xi_row, xi_T, iso, (const double *)d_od4,
                                   (size_t)nfull * nb);
            else if (use_lds)
                hipLaunchKernelGGL((k_xi_sym_lds<float>), g, bl, shl, st, nrows, nb, d_rowq, d_chkoff, d_chk,
                                   d_orboff, d_members, (float *)c->d_xi, c->xi_row, xi_T, capm);
            else
                hipLaunchKernelGGL((k_xi_sym<float, false>), g, b, 0, st, nrows, nb, d_rowq, d_orbq1, d_orboff,
                                   d_members, (float *)c->d_xi, c->xi_row, xi_T, iso, (const double *)d_od4,
                                   (size_t)nfull * nb);
        }
    }
    TRY(hipGetLastError());
    TRY(hipEventRecord(evs1, st));
    TRY(hipEventRecord(ev1, st));
    TRY(hipMemcpyAsync(h_lw, d_lw, sizeof(double) * n_lw, hipMemcpyDeviceToHost, st));
    if (use_iso) TRY(hipMemcpyAsync(h_iso, iso.gam, sizeof(double) * nrows, hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    std::vector<double> lw(h_lw, h_lw + n_lw), lw_iso(use_iso ? nrows : 0);
    if (use_iso) std::copy(h_iso, h_iso + nrows, lw_iso.begin());
    // isotope linewidth first, three-phonon added (scattering.f90:226-236), then boundary
    // scattering |v| / L for group velocities above lo_phonongroupveltol (:245-250)
    if (use_iso)
        for (int T = 0; T < nT; ++T)
            for (int r = 0; r < nrows; ++r) lw[(size_t)T * nrows + r] = lw_iso[r] + lw[(size_t)T * nrows + r];
    if (!g4.empty())
        for (size_t i = 0; i < g4.size(); ++i) lw[i] += g4[i];
    if (c->mfp_max > 0.0)
        for (int r = 0; r < nrows; ++r) {
            const double *v = &c->vel_irr[((size_t)c->row_q[r] * nb + c->row_b[r]) * 3];
            const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (vn > kGroupVelTol)
                for (int T = 0; T < nT; ++T) lw[(size_t)T * nrows + r] += vn / c->mfp_max;
        }
    // kernel statistics and the pass's device time: read from these events on request
    c->ks_kev = kev;
    c->ks_ev0 = ev0; c->ks_ev1 = ev1; c->ks_evs0 = evs0; c->ks_evs1 = evs1;
    c->ks_comp = d_comp != nullptr;
    c->ks_pending = true;

#undef TRY
    cleanup();
    // reduce + degeneracy average + F0 (scattering.f90:260-292)
    const size_t qb = (size_t)nirr * nb;
    c->linewidth.assign(qb * nT, 0.0);
    c->qs.assign(qb * nT, 0.0);
    c->F0.assign(qb * 3 * nT, 0.0);
    c->mfp.assign(qb * 3 * nT, 0.0);
    c->smfp.assign(qb * nT, 0.0);
    for (int T = 0; T < nT; ++T) {
        std::vector<double> buf(qb, 0.0);
        for (int r = 0; r < nrows; ++r) buf[(size_t)c->row_q[r] * nb + c->row_b[r]] = lw[(size_t)T * nrows + r];
        for (int q = 0; q < nirr; ++q)
            for (int b = 0; b < nb; ++b) {
                const double w = c->om_irr[(size_t)q * nb + b];
                if (w < kFreqTol) continue;
                const auto &dm = c->degen[q][b];
                double f0 = 0;
                for (int b2 : dm) f0 += buf[(size_t)q * nb + b2];
                f0 /= (double)dm.size();
                for (int b2 : dm) buf[(size_t)q * nb + b2] = f0;
                const size_t i = (size_t)T * qb + (size_t)q * nb + b;
                c->linewidth[i] = buf[(size_t)q * nb + b];
                c->qs[i] = 2.0 * c->linewidth[i];
                if (c->linewidth[i] > kFreqTol) {
                    const double *v = &c->vel_irr[((size_t)q * nb + b) * 3];
                    const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                    for (int k = 0; k < 3; ++k) {
                        c->mfp[i * 3 + k] = v[k] / c->qs[i];
                        c->F0[i * 3 + k] = c->mfp[i * 3 + k];
                    }
                    c->smfp[i] = vn / c->qs[i];
                }
            }
    }
    c->timings[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    return 0;
}

}  // namespace tc3
