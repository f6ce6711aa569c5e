// Thermal conductivity: RTA (host), iterative BTE (FP64 on the GPU) and the
// off-diagonal coherence term (host).
//
//   get_kappa          kappa.f90:28-75
//   iterative_solution kappa.f90:326-495 (fold-out, Xi.F, update on the GPU)
//   get_kappa_offdiag  kappa.f90:77-306
//   symmetrize_kappa   kappa.f90:308-324
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <cstdlib>

#include "tc3_ctx.hpp"

namespace tc3 {

namespace {

using zc = std::complex<double>;

double cv_host(double T, double om) {
    // lo_harmonic_oscillator_cv (gottochblandat_physics.f90:331-367)
    if (om < kFreqTol) return 0.0;
    if (T < kTemperatureTol) return 0.0;
    const double x = om / T / kKbHartree;
    if (x > 200.0) return 0.0;
    if (x > 21.0) return kKbHartree * x * x * std::exp(-x);
    if (x < 2e-2) return kKbHartree * (1.0 - x * x / 12.0 + x * x * x * x / 240.0);
    const double ex = std::exp(x);
    return kKbHartree * x * x * ex / ((ex - 1) * (ex - 1));
}

void chop_kappa(double *k) {
    double s = 0;
    for (int i = 0; i < 9; ++i) s += std::fabs(k[i]);
    for (int i = 0; i < 9; ++i) k[i] = chop_scalar(k[i], s * 1e-6);
}

// The operation average of get_kappa, (1/N) sum_R (R F)_i (R v)_j = (1/N) sum_R (R X R^T)_ij
// with X = F v^T, is a fixed linear map on 3x3 matrices: S[3 j + i][3 b + a] =
// (1/N) sum_R R_ia R_jb (column-major indices, like the kappa arrays).
void op_average_map(const tc3_ctx *c, double *S) {
    const int nsym = (int)c->ops.size();
    std::fill(S, S + 81, 0.0);
    for (int o = 0; o < nsym; ++o) {
        const double *m = c->ops[o].m;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i)
                for (int b = 0; b < 3; ++b)
                    for (int a = 0; a < 3; ++a) S[9 * (3 * j + i) + 3 * b + a] += m[3 * i + a] * m[3 * j + b];
    }
    for (int x = 0; x < 81; ++x) S[x] /= nsym;
}

// get_kappa (kappa.f90:28-75): kappa (column-major 3x3) and optional per-mode tensors.
// Without per-mode output the operation average is applied once to the weighted sum of
// F v^T (it is linear); with it, per mode.
void get_kappa(const tc3_ctx *c, const double *Fn, double T, bool classical, double *kappa, double *kmode) {
    const int nb = c->nb, nirr = c->mesh.nirr;
    double S[81];
    op_average_map(c, S);
    double X[9] = {0}, k[9] = {0};
    for (int q = 0; q < nirr; ++q)
        for (int b = 0; b < nb; ++b) {
            const double om = c->om_irr[(size_t)q * nb + b];
            if (kmode) std::fill(kmode + ((size_t)q * nb + b) * 9, kmode + ((size_t)q * nb + b) * 9 + 9, 0.0);
            if (om < kFreqTol) continue;
            const double cv = classical ? kKbHartree : cv_host(T, om);
            const double *F = Fn + ((size_t)q * nb + b) * 3;
            const double *v = &c->vel_irr[((size_t)q * nb + b) * 3];
            const double f = cv / c->cr.volume;
            if (kmode) {
                double *km = kmode + ((size_t)q * nb + b) * 9;
                for (int x = 0; x < 9; ++x) {
                    double sx = 0;
                    for (int bb = 0; bb < 3; ++bb)
                        for (int a = 0; a < 3; ++a) sx += S[9 * x + 3 * bb + a] * F[a] * v[bb];
                    km[x] = f * sx;
                    k[x] += km[x] * c->mesh.weight_irr[q];
                }
            } else {
                const double w = f * c->mesh.weight_irr[q];
                for (int bb = 0; bb < 3; ++bb)
                    for (int a = 0; a < 3; ++a) X[3 * bb + a] += w * F[a] * v[bb];
            }
        }
    if (!kmode)
        for (int x = 0; x < 9; ++x) {
            double sx = 0;
            for (int y = 0; y < 9; ++y) sx += S[9 * x + y] * X[y];
            k[x] = sx;
        }
    chop_kappa(k);
    std::memcpy(kappa, k, sizeof(k));
}

void symmetrize_kappa(const tc3_ctx *c, double *k) {
    const int nsym = (int)c->ops.size();
    double t[9] = {0};
    for (int o = 0; o < nsym; ++o) {
        const double *m = c->ops[o].m;
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                double s = 0;
                for (int jj = 0; jj < 3; ++jj)
                    for (int ii = 0; ii < 3; ++ii) s += k[3 * jj + ii] * m[3 * i + ii] * m[3 * j + jj];
                t[3 * j + i] += s;
            }
    }
    for (int i = 0; i < 9; ++i) k[i] = t[i] / nsym;
    chop_kappa(k);
}

// get_kappa_offdiag (kappa.f90:77-306) with the coherence velocities formed as
// avg_op[egw^H G egw] (entry for entry the product the reference forms through its
// Kronecker matrix).  The velocities are temperature independent: V_z(q) = sum over the
// little group of q of egw^H G_z egw, with G_z the z-derivative of the dynamical matrix
// without masses, G[(a2, j)][(a1, i)] = sum_pairs i R_z exp(i q.R) phi_ij, and egw the
// rotated eigenvectors scaled by 1 / sqrt(m 2 omega).  od_velocity_host forms one q on the
// host; k_od_velocity forms them on the device.
void od_velocity_host(const tc3_ctx *c, int iq, zc *V) {
    const int na = c->cr.na, nb = c->nb;
    std::vector<zc> G(3 * (size_t)nb * nb), egw((size_t)nb * nb), tmp((size_t)nb * nb);
    std::vector<double> mw(nb);
    for (int a = 0; a < na; ++a)
        for (int i = 0; i < 3; ++i) mw[3 * a + i] = c->cr.invsqrtmass[a];
    const double *q = &c->mesh.qcart[3 * c->mesh.irr_full_index[iq]];
    double qv[3];
    for (int i = 0; i < 3; ++i) qv[i] = chop_scalar(q[i] * kTwoPi, 1e-13);
    for (size_t p = 0; p < c->p_a1.size(); ++p) {
        const int a1 = c->p_a1[p], a2 = c->p_a2[p];
        const double *R = &c->p_lv[3 * p];
        const double qr = R[0] * qv[0] + R[1] * qv[1] + R[2] * qv[2];
        const zc e(std::cos(qr), std::sin(qr));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double m = c->p_m[9 * p + 3 * i + j];
                for (int al = 0; al < 3; ++al)
                    G[(size_t)al * nb * nb + (size_t)(3 * a2 + j) * nb + (3 * a1 + i)] += zc(0, R[al]) * e * m;
            }
    }
    const double *om = &c->om_irr[(size_t)iq * nb];
    std::fill(V, V + 3 * (size_t)nb * nb, zc(0, 0));
    for (int o : c->mesh.littlegroup[iq]) {
        const SymOp &op = c->ops[o];
        // egw = Gamma(R) egv, scaled by invsqrtmass / sqrt(2 omega)
        for (int a1 = 0; a1 < na; ++a1) {
            int a2 = 0;
            for (int x = 0; x < na; ++x)
                if (op.fmap[x] == a1) a2 = x;
            double d[3], v0[3];
            for (int i = 0; i < 3; ++i) d[i] = c->cr.rcart[3 * a1 + i] - op.tr[i];
            for (int i = 0; i < 3; ++i)
                v0[i] = op.m[i] * d[0] + op.m[3 + i] * d[1] + op.m[6 + i] * d[2] - c->cr.rcart[3 * a2 + i];
            const double ph = -(v0[0] * q[0] + v0[1] * q[1] + v0[2] * q[2]) * kTwoPi;
            const zc e(std::cos(ph), std::sin(ph));
            for (int al = 0; al < 3; ++al)
                for (int b = 0; b < nb; ++b) {
                    zc s(0, 0);
                    for (int be = 0; be < 3; ++be) {
                        const cd x = c->egv_irr[((size_t)iq * nb + 3 * a2 + be) * nb + b];
                        s += op.m[3 * al + be] * zc(x.x, x.y);
                    }
                    s *= e;
                    const double sc = (om[b] > kFreqTol) ? mw[3 * a1 + al] / std::sqrt(om[b] * 2.0) : 0.0;
                    egw[(size_t)(3 * a1 + al) * nb + b] = s * sc;
                }
        }
        for (int z = 0; z < 3; ++z) {
            // tmp = G egw ; V += egw^H tmp
            const zc *Gz = &G[(size_t)z * nb * nb];
            for (int r = 0; r < nb; ++r)
                for (int b = 0; b < nb; ++b) {
                    zc s(0, 0);
                    for (int x = 0; x < nb; ++x) s += Gz[(size_t)r * nb + x] * egw[(size_t)x * nb + b];
                    tmp[(size_t)r * nb + b] = s;
                }
            for (int i = 0; i < nb; ++i)
                for (int j = 0; j < nb; ++j) {
                    zc s(0, 0);
                    for (int r = 0; r < nb; ++r) s += std::conj(egw[(size_t)r * nb + i]) * tmp[(size_t)r * nb + j];
                    V[(size_t)z * nb * nb + (size_t)i * nb + j] += s;
                }
        }
    }
}

// lo_harmonic_oscillator_cv (gottochblandat_physics.f90:331-367), as cv_host
__device__ double cv_dev(double T, double om) {
    const double x = om / T / kKbHartree;
    if (x > 200.0) return 0.0;
    if (x > 21.0) return kKbHartree * x * x * exp(-x);
    if (x < 2e-2) return kKbHartree * (1.0 - x * x / 12.0 + x * x * x * x / 240.0);
    const double ex = exp(x);
    return kKbHartree * x * x * ex / ((ex - 1) * (ex - 1));
}

struct OdArgs {
    int nb, na;
    const double *qv, *qirr;        // [nirr][3]: 2 pi q (chopped) and q of the irreducible points
    const cd *egv;                  // [q][component][mode]
    const double *om;               // [q][mode]
    const int *blk2;                // pairs of (a1, a2): [blk2[a1 na + a2], blk2[a1 na + a2 + 1])
    const double *plv, *pm;         // 3 / 9 per pair
    const double *opm, *optr;       // 9 / 3 per operation
    const int *invf;                // [op][a1]: the atom a2 with fmap[a2] = a1
    const int *lg_off, *lg;         // little group of every irreducible point
    const double *rcart, *mw;       // [na][3], [nb]
    cd *V;                          // [q][3][nb][nb]
};

__device__ inline cd zmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

constexpr int kOdSlots = 9;   // nb^2 / 256 entries of V per thread in LDS mode (nb <= 48)

// Workgroup per (irreducible q, direction z).  SH: G_z, egw and G_z egw in LDS (3 nb^2
// complex) and the sums over the little group in registers; otherwise (any nb) the three
// matrices in the workgroup's slice of a global scratch and the sums accumulated in V, each
// entry by the thread that owns it.  The operations and their order are the same in both
// forms, in the order of od_velocity_host.
template <bool SH>
__global__ void __launch_bounds__(256) k_od_velocity(OdArgs a, cd *__restrict__ scr) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nb = a.nb, na = a.na, nb2 = nb * nb;
    const int iq = blockIdx.x / 3, z = blockIdx.x % 3;
    cd *G = SH ? (cd *)smem : scr + (size_t)blockIdx.x * 3 * nb2;
    cd *E = G + nb2, *Tm = E + nb2;
    cd *Vz = a.V + ((size_t)iq * 3 + z) * nb2;
    const double *qv = a.qv + 3 * iq, *q = a.qirr + 3 * iq;
    for (int x = threadIdx.x; x < nb2; x += blockDim.x) {
        const int r = x / nb, col = x % nb;
        const int a2 = r / 3, j = r % 3, a1 = col / 3, i = col % 3;
        cd g = {0, 0};
        for (int p = a.blk2[a1 * na + a2]; p < a.blk2[a1 * na + a2 + 1]; ++p) {
            const double *R = a.plv + 3 * p;
            const double qr = R[0] * qv[0] + R[1] * qv[1] + R[2] * qv[2];
            const cd e = {cos(qr), sin(qr)};
            const cd t = zmul({0.0, R[z]}, e);
            const double m = a.pm[9 * p + 3 * i + j];
            g.x += t.x * m;
            g.y += t.y * m;
        }
        G[x] = g;
        if constexpr (!SH) Vz[x] = {0, 0};
    }
    cd acc[SH ? kOdSlots : 1];
    for (int s = 0; s < (SH ? kOdSlots : 1); ++s) acc[s] = {0, 0};
    for (int l = a.lg_off[iq]; l < a.lg_off[iq + 1]; ++l) {
        const int o = a.lg[l];
        const double *m = a.opm + 9 * o, *tr = a.optr + 3 * o;
        for (int x = threadIdx.x; x < nb2; x += blockDim.x) {
            const int row = x / nb, b = x % nb, a1 = row / 3, al = row % 3;
            const int a2 = a.invf[o * na + a1];
            double d[3], v0[3];
            for (int i = 0; i < 3; ++i) d[i] = a.rcart[3 * a1 + i] - tr[i];
            for (int i = 0; i < 3; ++i) v0[i] = m[i] * d[0] + m[3 + i] * d[1] + m[6 + i] * d[2] - a.rcart[3 * a2 + i];
            const double ph = -(v0[0] * q[0] + v0[1] * q[1] + v0[2] * q[2]) * kTwoPi;
            cd s = {0, 0};
            for (int be = 0; be < 3; ++be) {
                const cd e = a.egv[((size_t)iq * nb + 3 * a2 + be) * nb + b];
                s.x += m[3 * al + be] * e.x;
                s.y += m[3 * al + be] * e.y;
            }
            s = zmul(s, {cos(ph), sin(ph)});
            const double w = a.om[(size_t)iq * nb + b];
            const double sc = (w > kFreqTol) ? a.mw[row] / sqrt(w * 2.0) : 0.0;
            E[x] = {s.x * sc, s.y * sc};
        }
        __syncthreads();
        for (int x = threadIdx.x; x < nb2; x += blockDim.x) {
            const int r = x / nb, b = x % nb;
            cd s = {0, 0};
            for (int k = 0; k < nb; ++k) {
                const cd t = zmul(G[r * nb + k], E[k * nb + b]);
                s.x += t.x;
                s.y += t.y;
            }
            Tm[x] = s;
        }
        __syncthreads();
        for (int sl = 0;; ++sl) {
            const int x = threadIdx.x + sl * blockDim.x;
            if (x >= nb2 || (SH && sl >= kOdSlots)) break;
            const int i = x / nb, j = x % nb;
            cd s = {0, 0};
            for (int r = 0; r < nb; ++r) {
                const cd e = E[r * nb + i];
                const cd t = zmul({e.x, -e.y}, Tm[r * nb + j]);
                s.x += t.x;
                s.y += t.y;
            }
            if constexpr (SH) {
                acc[sl].x += s.x;
                acc[sl].y += s.y;
            } else {
                cd v = Vz[x];
                v.x += s.x;
                v.y += s.y;
                Vz[x] = v;
            }
        }
        __syncthreads();
    }
    if constexpr (SH)
        for (int sl = 0; sl < kOdSlots; ++sl) {
            const int x = threadIdx.x + sl * blockDim.x;
            if (x < nb2) Vz[x] = acc[sl];
        }
}

// The coherence velocities of every mode pair on the device, vel[q][j][k][z] = Re V_z(q)[j][k]
// / |lg(q)|, below vtol set to zero (kappa.f90:203-213), as od_real_velocities.
__global__ void k_od_real(int nirr, int nb, const int *__restrict__ lg_off, const cd *__restrict__ V, double vtol,
                          double *__restrict__ vel) {
    const size_t nb2 = (size_t)nb * nb;
    const size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (x >= (size_t)nirr * nb2 * 3) return;
    const int z = (int)(x % 3);
    const size_t jk = (x / 3) % nb2, iq = x / 3 / nb2;
    const double ng = (double)(lg_off[iq + 1] - lg_off[iq]);
    const double re = V[(iq * 3 + z) * nb2 + jk].x / ng;
    vel[x] = fabs(re) < vtol ? 0.0 : re;
}

// M[T][q][9]: the sum over the mode pairs j != k of q, both above lo_freqtol in frequency and
// linewidth, of w v_jk v_jk^T with w = f0 (G_j + G_k) / ((G_j + G_k)^2 + (om_j - om_k)^2),
// f0 the mean heat capacity (kB classically) (kappa.f90:220-287, as kappa_offdiag); workgroup
// per (q, T), per-thread sums over a fixed pair set added by a fixed tree.
__global__ void __launch_bounds__(256) k_od_pairsum(int nirr, int nb, int classical, const double *__restrict__ temps,
                                                    const double *__restrict__ om_, const double *__restrict__ lw_,
                                                    const double *__restrict__ vel_, double *__restrict__ M) {
    __shared__ double red[9][256];
    const int iq = blockIdx.x, T = blockIdx.y;
    const double tk = temps[T];
    const double *om = om_ + (size_t)iq * nb, *lw = lw_ + ((size_t)T * nirr + iq) * nb;
    const double *vel = vel_ + (size_t)iq * nb * nb * 3;
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int p = threadIdx.x; p < nb * nb; p += blockDim.x) {
        const int j = p / nb, k = p % nb;
        if (j == k) continue;
        const double o1 = om[j], o2 = om[k], t1 = lw[j], t2 = lw[k];
        if (o1 < kFreqTol || t1 < kFreqTol || o2 < kFreqTol || t2 < kFreqTol) continue;
        const double c1 = classical ? 0.0 : cv_dev(tk, o1), c2 = classical ? 0.0 : cv_dev(tk, o2);
        const double f0 = classical ? kKbHartree : 0.5 * (c1 + c2);
        const double tau = (t1 + t2) / ((t1 + t2) * (t1 + t2) + (o1 - o2) * (o1 - o2));
        const double w = tau * f0;
        const double *v = vel + (size_t)p * 3;
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) m[3 * x + y] += w * v[x] * v[y];
    }
    for (int i = 0; i < 9; ++i) red[i][threadIdx.x] = m[i];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int i = 0; i < 9; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 9) M[((size_t)T * nirr + iq) * 9 + threadIdx.x] = red[threadIdx.x][0];
}

// The coherence velocities of every mode pair, vel[q][j][k][z] = Re V_z(q)[j][k] / |lg(q)|,
// below 1e-10 km/s set to zero (kappa.f90:203-213).
void od_real_velocities(const tc3_ctx *c, const std::vector<zc> &V, std::vector<double> &vel) {
    const int nb = c->nb, nirr = c->mesh.nirr;
    const size_t nb2 = (size_t)nb * nb;
    const double vtol = 1e-10 / (kGroupvelAuToMs / 1000.0);
    vel.assign((size_t)nirr * nb2 * 3, 0.0);
    for (int iq = 0; iq < nirr; ++iq) {
        const double ng = (double)c->mesh.littlegroup[iq].size();
        for (int z = 0; z < 3; ++z)
            for (size_t jk = 0; jk < nb2; ++jk) {
                const double re = (V[((size_t)iq * 3 + z) * nb2 + jk] / ng).real();
                vel[((size_t)iq * nb2 + jk) * 3 + z] = std::fabs(re) < vtol ? 0.0 : re;
            }
    }
}

// Host form of the coherence velocities (TC3_OD_HOST; the check of the device form).
void od_velocities_host(tc3_ctx *c, std::vector<double> &vel) {
    if (ensure_full_host(c)) return;   // the irreducible eigenvectors on the host
    const int nb = c->nb, nirr = c->mesh.nirr;
    const size_t nb2 = (size_t)nb * nb;
    std::vector<zc> V((size_t)nirr * 3 * nb2, zc(0, 0));
    for (int iq = 0; iq < nirr; ++iq) od_velocity_host(c, iq, V.data() + (size_t)iq * 3 * nb2);
    od_real_velocities(c, V, vel);
}

// The off-diagonal term on the device for every temperature: the coherence velocities V of
// every irreducible point (k_od_velocity: LDS form when 3 nb^2 complex values fit and nb^2 <=
// 256 kOdSlots, else the global-scratch form), their real parts (k_od_real) and the pair sums
// M[T][q][9] (k_od_pairsum), downloaded for the star rotation.
int od_device(tc3_ctx *c, std::vector<double> &M) {
    const int na = c->cr.na, nb = c->nb, nirr = c->mesh.nirr, nsym = (int)c->ops.size(), nT = c->nT;
    const size_t nb2 = (size_t)nb * nb;
    const size_t lds = 3 * nb2 * sizeof(cd);
    const bool sh = nb2 <= 256 * (size_t)kOdSlots && lds <= 150 * 1024;
    hipStream_t st = c->stream;
    hipError_t e;
#define TRY(x) do { e = (x); if (e != hipSuccess) return device_error(e, #x, __FILE__, __LINE__); } while (0)
    std::vector<double> qv(3 * (size_t)nirr), qirr(3 * (size_t)nirr), opm(9 * (size_t)nsym), optr(3 * (size_t)nsym), mw(nb);
    std::vector<int> invf((size_t)nsym * na), lg_off(nirr + 1, 0), lg;
    for (int iq = 0; iq < nirr; ++iq) {
        const double *q = &c->mesh.qcart[3 * c->mesh.irr_full_index[iq]];
        for (int i = 0; i < 3; ++i) {
            qirr[3 * iq + i] = q[i];
            qv[3 * iq + i] = chop_scalar(q[i] * kTwoPi, 1e-13);
        }
        lg_off[iq] = (int)lg.size();
        lg.insert(lg.end(), c->mesh.littlegroup[iq].begin(), c->mesh.littlegroup[iq].end());
    }
    lg_off[nirr] = (int)lg.size();
    for (int o = 0; o < nsym; ++o) {
        std::memcpy(&opm[9 * o], c->ops[o].m, sizeof(double) * 9);
        std::memcpy(&optr[3 * o], c->ops[o].tr, sizeof(double) * 3);
        for (int a1 = 0; a1 < na; ++a1) {
            int a2 = 0;
            for (int x = 0; x < na; ++x)
                if (c->ops[o].fmap[x] == a1) a2 = x;
            invf[(size_t)o * na + a1] = a2;
        }
    }
    for (int a = 0; a < na; ++a)
        for (int i = 0; i < 3; ++i) mw[3 * a + i] = c->cr.invsqrtmass[a];
    auto up = [&](const char *name, const void *src, size_t bytes, void **dst) -> hipError_t {
        hipError_t r = c->pool_get(name, std::max<size_t>(bytes, 16), dst);
        if (r == hipSuccess && bytes) r = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, st);
        return r;
    };
    OdArgs a;
    a.nb = nb;
    a.na = na;
    void *p;
    TRY(up("od_qv", qv.data(), sizeof(double) * qv.size(), &p)); a.qv = (const double *)p;
    TRY(up("od_qirr", qirr.data(), sizeof(double) * qirr.size(), &p)); a.qirr = (const double *)p;
    TRY(up("od_blk2", c->blk2_off.data(), sizeof(int) * c->blk2_off.size(), &p)); a.blk2 = (const int *)p;
    TRY(up("od_plv", c->p_lv.data(), sizeof(double) * c->p_lv.size(), &p)); a.plv = (const double *)p;
    TRY(up("od_pm", c->p_m.data(), sizeof(double) * c->p_m.size(), &p)); a.pm = (const double *)p;
    TRY(up("od_opm", opm.data(), sizeof(double) * opm.size(), &p)); a.opm = (const double *)p;
    TRY(up("od_optr", optr.data(), sizeof(double) * optr.size(), &p)); a.optr = (const double *)p;
    TRY(up("od_invf", invf.data(), sizeof(int) * invf.size(), &p)); a.invf = (const int *)p;
    TRY(up("od_lgoff", lg_off.data(), sizeof(int) * lg_off.size(), &p)); a.lg_off = (const int *)p;
    TRY(up("od_lg", lg.data(), sizeof(int) * lg.size(), &p)); a.lg = (const int *)p;
    TRY(up("od_rcart", c->cr.rcart.data(), sizeof(double) * 3 * na, &p)); a.rcart = (const double *)p;
    TRY(up("od_mw", mw.data(), sizeof(double) * nb, &p)); a.mw = (const double *)p;
    a.egv = c->d_egv_irr;
    a.om = c->d_om_irr;
    TRY(c->pool_get("od_V", sizeof(cd) * (size_t)nirr * 3 * nb2, &p));
    a.V = (cd *)p;
    if (sh) {
        TRY(hipFuncSetAttribute((const void *)k_od_velocity<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_od_velocity<true>, dim3(3 * nirr), dim3(256), lds, st, a, (cd *)nullptr);
    } else {
        cd *scr;
        TRY(c->pool_get("od_scr", sizeof(cd) * 3 * nb2 * 3 * (size_t)nirr, (void **)&scr));
        hipLaunchKernelGGL(k_od_velocity<false>, dim3(3 * nirr), dim3(256), 0, st, a, scr);
    }
    TRY(hipGetLastError());
    double *d_vel, *d_lw, *d_T, *d_M;
    const size_t nvel = (size_t)nirr * nb2 * 3;
    TRY(c->pool_get("od_vel", sizeof(double) * nvel, (void **)&d_vel));
    const double vtol = 1e-10 / (kGroupvelAuToMs / 1000.0);
    hipLaunchKernelGGL(k_od_real, dim3((unsigned)((nvel + 255) / 256)), dim3(256), 0, st, nirr, nb, a.lg_off, a.V,
                       vtol, d_vel);
    TRY(up("od_lw", c->linewidth.data(), sizeof(double) * (size_t)nT * nirr * nb, (void **)&d_lw));
    TRY(up("od_T", c->temps.data(), sizeof(double) * nT, (void **)&d_T));
    TRY(c->pool_get("od_M", sizeof(double) * 9 * (size_t)nT * nirr, (void **)&d_M));
    hipLaunchKernelGGL(k_od_pairsum, dim3(nirr, nT), dim3(256), 0, st, nirr, nb, (int)c->classical, d_T, c->d_om_irr,
                       d_lw, d_vel, d_M);
    TRY(hipGetLastError());
    M.assign(9 * (size_t)nT * nirr, 0.0);
    TRY(hipMemcpyAsync(M.data(), d_M, sizeof(double) * M.size(), hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
#undef TRY
    return 0;
}

// The star rotation of get_kappa_offdiag (kappa.f90:220-287): the sum over the star of q of
// (m v)(m v)^T is linear in v v^T, so the weighted v v^T of all mode pairs of a q (M[q][9]) is
// rotated through the star once.
void od_rotate(const tc3_ctx *c, const double *Mq, double *kod) {
    const int nirr = c->mesh.nirr;
    double k[9] = {0};
    for (int iq = 0; iq < nirr; ++iq) {
        const double *M = Mq + 9 * (size_t)iq;
        const auto &so = c->mesh.star_ops[iq];
        const double pref = c->mesh.weight_irr[iq] / c->cr.volume / (double)so.size();
        for (int s : so) {
            const double *m = c->ops[std::abs(s) - 1].m;
            double mM[9];   // m M
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y)
                    mM[3 * x + y] = m[3 * x] * M[y] + m[3 * x + 1] * M[3 + y] + m[3 * x + 2] * M[6 + y];
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y)
                    k[3 * x + y] += pref * (mM[3 * x] * m[3 * y] + mM[3 * x + 1] * m[3 * y + 1] + mM[3 * x + 2] * m[3 * y + 2]);
        }
    }
    chop_kappa(k);
    std::memcpy(kod, k, sizeof(k));
}

// The temperature-dependent part of get_kappa_offdiag on the host (TC3_OD_HOST) for the
// velocities of od_velocities_host: the pair sums of every q (as k_od_pairsum), then od_rotate.
void kappa_offdiag(const tc3_ctx *c, int iT, double T, bool classical, const double *vel, double *kod) {
    const int nb = c->nb, nirr = c->mesh.nirr;
    std::vector<double> Mq(9 * (size_t)nirr, 0.0), cv(nb);
    for (int iq = 0; iq < nirr; ++iq) {
        const double *om = &c->om_irr[(size_t)iq * nb];
        const double *lw = &c->linewidth[(size_t)iT * nirr * nb + (size_t)iq * nb];
        for (int j = 0; j < nb; ++j) cv[j] = (classical || om[j] < kFreqTol) ? 0.0 : cv_host(T, om[j]);
        double *M = &Mq[9 * (size_t)iq];
        for (int j = 0; j < nb; ++j) {
            const double om1 = om[j];
            if (om1 < kFreqTol) continue;
            const double t1 = lw[j];
            if (t1 < kFreqTol) continue;
            for (int kk = 0; kk < nb; ++kk) {
                if (kk == j) continue;
                const double om2 = om[kk];
                if (om2 < kFreqTol) continue;
                const double t2 = lw[kk];
                if (t2 < kFreqTol) continue;
                const double *v = &vel[(((size_t)iq * nb + j) * nb + kk) * 3];
                const double f0 = classical ? kKbHartree : 0.5 * (cv[j] + cv[kk]);
                const double tau = (t1 + t2) / ((t1 + t2) * (t1 + t2) + (om1 - om2) * (om1 - om2));
                const double w = tau * f0;
                for (int x = 0; x < 3; ++x)
                    for (int y = 0; y < 3; ++y) M[3 * x + y] += w * v[x] * v[y];
            }
        }
    }
    od_rotate(c, Mq.data(), kod);
}

constexpr int kBlock = 256;

// Iteration state of the BTE solve, kept on the device so that chunks of iterations run
// without a host round trip: iterations past convergence leave everything unchanged.
struct BteState {
    int it;        // iterations done
    int done;      // converged or maxiter reached
    int maxiter;
    int pad;
    double mix;    // mixing factor (kappa.f90:364, 487-489)
    double tol;
};

// All temperatures iterate together: blockIdx.z is the temperature, every per-temperature
// array is strided by it, and a temperature whose state says done is skipped.
template <typename FT>
__global__ void k_foldout(int nfull, int nb, const int *__restrict__ irr_of_full, const int *__restrict__ op_of_full,
                          const double *__restrict__ opm, const double *__restrict__ Fn_, FT *__restrict__ Fbb_,
                          const BteState *__restrict__ st_, long sF, long sFbb) {
    const int T = blockIdx.z;
    if (st_[T].done) return;
    const double *Fn = Fn_ + T * sF;
    FT *Fbb = Fbb_ + T * sFbb;
    const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (j >= (long)nfull * nb) return;
    const int q = (int)(j / nb), b = (int)(j % nb);
    const int k = irr_of_full[q], iop = op_of_full[q];
    const double *m = opm + 9 * ((iop > 0 ? iop : -iop) - 1);
    const double *F = Fn + ((size_t)k * nb + b) * 3;
    const double s = iop > 0 ? 1.0 : -1.0;
    const long ncol = (long)nfull * nb;
    for (int i = 0; i < 3; ++i) Fbb[i * ncol + j] = (FT)(s * (m[3 * i] * F[0] + m[3 * i + 1] * F[1] + m[3 * i + 2] * F[2]));
}

// part[slice][row][0..2] = sum over the slice's columns j of Xi[row][j] Fbb[0..2][j]
// (lo_gemv x3, kappa.f90:411-413); the folded-out F is planar, [component][column], in FP32
// next to FP32 Xi (each product of an FP32 Xi and an FP32 F is exact in FP64; the sums are
// FP64) and in FP64 next to FP64 Xi.  The slices are added in slice order by k_bte_update, so
// the result does not depend on scheduling.
#ifndef TC3_BTE_ROWS
#define TC3_BTE_ROWS 4
#endif
#ifndef TC3_BTE_SLICES
#define TC3_BTE_SLICES 8
#endif
constexpr int kRows = TC3_BTE_ROWS, kSlices = TC3_BTE_SLICES;

// Workgroup per (4 kRows rows, column slice): each wave takes kRows rows and all waves sweep
// the same columns together, so every F value fetched from L2 serves the four waves (from
// L1) and kRows rows each; VW columns per load (16 bytes of FP32 Xi), and the per-row sums
// need only a wave reduction.
template <typename XT, int VW, typename FT>
__global__ void __launch_bounds__(kBlock) k_xi_gemv_w(int nrows, long ncol, const XT *__restrict__ xi_,
                                                      const FT *__restrict__ Fbb_, double *__restrict__ part_,
                                                      const BteState *__restrict__ st_, size_t sXi, long sFbb,
                                                      long sPart) {
    typedef XT XV __attribute__((ext_vector_type(VW)));
    typedef FT DV __attribute__((ext_vector_type(VW)));
    const int T = blockIdx.z;
    if (st_[T].done) return;
    const XT *xi = xi_ + T * sXi;
    const FT *Fbb = Fbb_ + T * sFbb;
    double *part = part_ + T * sPart;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * (kBlock / 64) + wave) * kRows, sl = blockIdx.y;
    const int nr = min(kRows, nrows - r0);
    if (nr <= 0) return;
    const long nv = ncol / VW;
    const long v0 = nv * sl / kSlices, v1 = nv * (sl + 1) / kSlices;
    const XT *x = xi + (size_t)r0 * ncol;
    double s[kRows][3];
#pragma unroll
    for (int r = 0; r < kRows; ++r) s[r][0] = s[r][1] = s[r][2] = 0;
    for (long v = v0 + lane; v < v1; v += 64) {
        const long j = v * VW;
        const DV f0 = *(const DV *)(Fbb + j), f1 = *(const DV *)(Fbb + ncol + j), f2 = *(const DV *)(Fbb + 2 * ncol + j);
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            if (r < nr) {
                const XV xv = *(const XV *)(x + (size_t)r * ncol + j);
#pragma unroll
                for (int k = 0; k < VW; ++k) {
                    const double e = (double)xv[k];
                    s[r][0] += e * (double)f0[k];
                    s[r][1] += e * (double)f1[k];
                    s[r][2] += e * (double)f2[k];
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kRows; ++r)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double v = s[r][i];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            s[r][i] = v;
        }
    if (lane < 3 * nr) {
        double v = 0;
#pragma unroll
        for (int r = 0; r < kRows; ++r)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                if (lane == 3 * r + i) v = s[r][i];
        part[((size_t)sl * nrows + r0) * 3 + lane] = v;
    }
}

// Rows stored per orbit (one value per orbit of the little group of the row's q1 and band):
// the product needs, per orbit, the sum of the folded-out F over its members.  Entry e is
// one orbit of one group: Fo[i][ent.z + b] (plane stride ent.w = norb nb of the group) =
// sum over the members p of Fbb[i][p nb + b], in member order, in FP64 (rounded to FT once).
template <typename FT>
__global__ void k_orbsum(int nent, int nfull, int nb, const int4 *__restrict__ ent, const int *__restrict__ mem,
                         const FT *__restrict__ Fbb_, FT *__restrict__ Fo_, const BteState *__restrict__ st_, long sFbb,
                         long sFo) {
    const int T = blockIdx.z;
    if (st_[T].done) return;
    const FT *Fbb = Fbb_ + T * sFbb;
    FT *Fo = Fo_ + T * sFo;
    const long ncol = (long)nfull * nb;
    for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < (long)nent * nb; x += (long)gridDim.x * blockDim.x) {
        const int e = (int)(x / nb), b = (int)(x % nb);
        const int4 en = ent[e];
        double s0 = 0, s1 = 0, s2 = 0;
        for (int m = en.x; m < en.y; ++m) {
            const long j = (long)mem[m] * nb + b;
            s0 += (double)Fbb[j];
            s1 += (double)Fbb[ncol + j];
            s2 += (double)Fbb[2 * ncol + j];
        }
        FT *o = Fo + en.z + b;
        o[0] = (FT)s0;
        o[en.w] = (FT)s1;
        o[2 * (long)en.w] = (FT)s2;
    }
}

// k_xi_gemv_w over rows stored per orbit: workgroup per (work unit, temperature), a unit being
// one column slice of a row group (up to 16 rows of one q1, consecutive, norb nb elements
// each) against the orbit sums of the group's orbit list (the folded-out F itself where every
// orbit is one point).  A row of ncol elements is cut into ns = ceil(S ncol / ncol_max)
// slices (S = 4 for FP32 rows, kSlices for FP64 rows), so that every slice holds about as many
// columns as one of the longest rows'; the last slice's workgroup writes zero partial sums into
// the slots after it, up to kSlices.  Units come in
// decreasing size with the temperature fastest, so the short ones fill the end of the launch.
static_assert(kRows * (kBlock / 64) >= 16, "a workgroup takes a whole row group");
template <typename XT, int VW, typename FT>
__global__ void __launch_bounds__(kBlock) k_xi_gemv_o(int nrows, int nT, const int4 *__restrict__ units,
                                                      const int4 *__restrict__ rg_,
                                                      const long long *__restrict__ xo_row,
                                                      const long long *__restrict__ goff, const XT *__restrict__ xi_,
                                                      const FT *__restrict__ Fbb_, const FT *__restrict__ Fo_,
                                                      double *__restrict__ part_, const BteState *__restrict__ st_,
                                                      size_t sXi, long sFbb, long sFo, long sPart) {
    typedef XT XV __attribute__((ext_vector_type(VW)));
    typedef FT DV __attribute__((ext_vector_type(VW)));
    const int T = blockIdx.x % nT;
    if (st_[T].done) return;
    const int4 un = units[blockIdx.x / nT];   // row group, slice, slices
    const int4 rg = rg_[un.x];                // first row, rows, group, norb nb
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = rg.x + wave * kRows, sl = un.y, ns = un.z;
    const int nr = min(kRows, rg.x + rg.y - r0);
    if (nr <= 0) return;
    const long ncol = rg.w;
    double *part = part_ + T * sPart;
    if (sl == ns - 1 && lane < 3 * nr)
        for (int z = ns; z < kSlices; ++z) part[((size_t)z * nrows + r0) * 3 + lane] = 0.0;
    const XT *x = xi_ + T * sXi + xo_row[r0];
    const FT *Fbb = rg.z < 0 ? Fbb_ + T * sFbb : Fo_ + T * sFo + goff[rg.z];
    const long nv = ncol / VW;
    const long v0 = nv * sl / ns, v1 = nv * (sl + 1) / ns;
    double s[kRows][3];
#pragma unroll
    for (int r = 0; r < kRows; ++r) s[r][0] = s[r][1] = s[r][2] = 0;
    for (long v = v0 + lane; v < v1; v += 64) {
        const long j = v * VW;
        const DV f0 = *(const DV *)(Fbb + j), f1 = *(const DV *)(Fbb + ncol + j), f2 = *(const DV *)(Fbb + 2 * ncol + j);
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            if (r < nr) {
                const XV xv = *(const XV *)(x + (size_t)r * ncol + j);
#pragma unroll
                for (int k = 0; k < VW; ++k) {
                    const double e = (double)xv[k];
                    s[r][0] += e * (double)f0[k];
                    s[r][1] += e * (double)f1[k];
                    s[r][2] += e * (double)f2[k];
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kRows; ++r)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double v = s[r][i];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            s[r][i] = v;
        }
    if (lane < 3 * nr) {
        double v = 0;
#pragma unroll
        for (int r = 0; r < kRows; ++r)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                if (lane == 3 * r + i) v = s[r][i];
        part[((size_t)sl * nrows + r0) * 3 + lane] = v;
    }
}

// Fnb = -buf/qs; degeneracy average then +F0, band by band in order (kappa.f90:415-448);
// convergence partial and mixing (kappa.f90:456-468).  Workgroup per irreducible point;
// the ordered degeneracy pass runs over LDS copies: where the degenerate sets of q partition
// its bands (dgpart), one thread per set takes its bands in order (the other sets' bands are
// untouched by them, so this is the band-by-band pass), else one thread takes every band.
// The products of the slices are added here in slice order (Xi.F of the rows of q), and the
// new F of q is folded out to the full-grid points of its star for the next product:
// Fbb[i][p nb + b] = s (R F(q, b))_i for p = R q (s = -1 for the operations with time reversal).
template <typename FT>
__global__ void k_bte_update(int nb, int nrows, const int *__restrict__ row_of, const double *__restrict__ part_,
                             const double *__restrict__ qs_, const double *__restrict__ F0_,
                             const int *__restrict__ dg_off, const int *__restrict__ dg,
                             const int *__restrict__ dgpart, const int *__restrict__ fzero,
                             const int *__restrict__ star_off, const int *__restrict__ star,
                             const int *__restrict__ op_of_full, const double *__restrict__ opm, long ncol,
                             double *__restrict__ Fn_, FT *__restrict__ Fbb_, double *__restrict__ g0part_,
                             const BteState *__restrict__ st_,
                             long sPart, long sQs, long sF, long sFbb, long sG) {
    const int T = blockIdx.z;
    if (st_[T].done) return;
    const double mix = st_[T].mix;
    const double *part = part_ + T * sPart, *qs = qs_ + T * sQs, *F0 = F0_ + T * sF;
    double *Fn = Fn_ + T * sF, *g0part = g0part_ + T * sG;
    FT *Fbb = Fbb_ + T * sFbb;
    const int q = blockIdx.x;
    __shared__ double Fnb[128 * 3], F0s[128 * 3], g1s[128];
    __shared__ int dgo[129], dgl[128 * 8];
    const int o_base = dg_off[q * nb];
    const int ndg = dg_off[q * nb + nb] - o_base;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const int r = row_of[q * nb + b];
        const double s = qs[q * nb + b];
        for (int i = 0; i < 3; ++i) {
            double v = 0;
            if (r >= 0)
                for (int sl = 0; sl < kSlices; ++sl) v += part[((size_t)sl * nrows + r) * 3 + i];
            Fnb[3 * b + i] = (r < 0 || s < 1e-9) ? 0.0 : -v / s;
            F0s[3 * b + i] = F0[((size_t)q * nb + b) * 3 + i];
        }
    }
    for (int b = threadIdx.x; b <= nb; b += blockDim.x) dgo[b] = dg_off[q * nb + b] - o_base;
    for (int x = threadIdx.x; x < ndg; x += blockDim.x) dgl[x] = dg[o_base + x];
    __syncthreads();
    // the pass for band b1 with the set dgl[o0..o1)
    auto step = [&](int b1, int o0, int o1) {
        double v0[3] = {0, 0, 0};
        for (int o = o0; o < o1; ++o)
            for (int i = 0; i < 3; ++i) v0[i] += Fnb[3 * dgl[o] + i];
        for (int i = 0; i < 3; ++i) v0[i] /= (double)(o1 - o0);
        for (int o = o0; o < o1; ++o)
            for (int i = 0; i < 3; ++i) Fnb[3 * dgl[o] + i] = v0[i];
        for (int i = 0; i < 3; ++i) Fnb[3 * b1 + i] = F0s[3 * b1 + i] + Fnb[3 * b1 + i];
    };
    if (dgpart[q]) {
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            const int o0 = dgo[b], o1 = dgo[b + 1];
            if (dgl[o0] != b) continue;   // the set's first band takes it
            for (int o = o0; o < o1; ++o) step(dgl[o], o0, o1);
        }
    } else if (threadIdx.x == 0) {
        for (int b1 = 0; b1 < nb; ++b1) step(b1, dgo[b1], dgo[b1 + 1]);
    }
    __syncthreads();
    // fzero[q]: the little group of q fixes no vector (its rotations sum to zero), so F(q) = R F(q)
    // for every R in it forces F(q) = 0; what the sum over columns leaves there is rounding,
    // which single-precision scattering matrices raise above the |F|^2 > 1e-10 test below
    const bool fz = fzero[q] != 0;
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        double *F = Fn + ((size_t)q * nb + b) * 3;
        if (fz) {
            F[0] = F[1] = F[2] = 0.0;
            g1s[b] = 0.0;
            F0s[3 * b] = F0s[3 * b + 1] = F0s[3 * b + 2] = 0.0;
            continue;
        }
        double g1 = 0, g2 = 0, f[3];
        for (int i = 0; i < 3; ++i) {
            f[i] = F[i];
            g1 += (f[i] - Fnb[3 * b + i]) * (f[i] - Fnb[3 * b + i]);
            g2 += f[i] * f[i];
        }
        g1s[b] = (g2 > 1e-10) ? g1 / g2 : 0.0;
        for (int i = 0; i < 3; ++i) {
            F[i] = f[i] * (1.0 - mix) + mix * Fnb[3 * b + i];
            F0s[3 * b + i] = F[i];   // the new F, for the fold-out below
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double g0 = 0;
        for (int b = 0; b < nb; ++b) g0 += g1s[b];
        g0part[q] = g0;
    }
    const int p0 = star_off[q], np = star_off[q + 1] - p0;
    for (int x = threadIdx.x; x < np * nb; x += blockDim.x) {
        const int p = star[p0 + x / nb], b = x % nb, iop = op_of_full[p];
        const double *m = opm + 9 * ((iop > 0 ? iop : -iop) - 1);
        const double sg = iop > 0 ? 1.0 : -1.0;
        const double *F = F0s + 3 * b;
        const long j = (long)p * nb + b;
        for (int i = 0; i < 3; ++i)
            Fbb[i * ncol + j] = (FT)(sg * (m[3 * i] * F[0] + m[3 * i + 1] * F[1] + m[3 * i + 2] * F[2]));
    }
}

// kappa of the current Fn for the iteration log (get_kappa, kappa.f90:28-75, before the
// chop): workgroup per slice of modes, each writing its 9 partial sums; the host adds
// the slices in order.  cvw[q][b] = cv(omega) * weight_q / volume (0 below lo_freqtol).
constexpr int kKapSlices = 32;
__global__ void __launch_bounds__(256) k_kappa_sum(int nmode, int nsym, const double *__restrict__ cvw_,
                                                   const double *__restrict__ Fn_, const double *__restrict__ vel,
                                                   const double *__restrict__ opm, double *__restrict__ klog_,
                                                   const BteState *__restrict__ st_, long sCvw, long sF, long sLog) {
    const int T = blockIdx.z;
    if (st_[T].done) return;
    const double *cvw = cvw_ + T * sCvw, *Fn = Fn_ + T * sF;
    double *out = klog_ + T * sLog + (size_t)9 * gridDim.x * st_[T].it;
    __shared__ double red[9][256];
    double k[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int m0 = (int)((long)nmode * blockIdx.x / gridDim.x), m1 = (int)((long)nmode * (blockIdx.x + 1) / gridDim.x);
    for (int x = m0 + threadIdx.x; x < m1; x += blockDim.x) {
        const double w = cvw[x];
        if (w == 0.0) continue;
        const double *F = Fn + 3 * (size_t)x, *v = vel + 3 * (size_t)x;
        double v2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int o = 0; o < nsym; ++o) {
            const double *m = opm + 9 * o;
            double v0[3], v1[3];
            for (int i = 0; i < 3; ++i) {
                v0[i] = m[3 * i] * F[0] + m[3 * i + 1] * F[1] + m[3 * i + 2] * F[2];
                v1[i] = m[3 * i] * v[0] + m[3 * i + 1] * v[1] + m[3 * i + 2] * v[2];
            }
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) v2[3 * j + i] += v0[i] * v1[j] / nsym;
        }
        for (int i = 0; i < 9; ++i) k[i] += w * v2[i];
    }
    for (int i = 0; i < 9; ++i) red[i][threadIdx.x] = k[i];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int i = 0; i < 9; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 9) out[(size_t)blockIdx.x * 9 + threadIdx.x] = red[threadIdx.x][0];
}

// Convergence test and mixing schedule after one iteration (kappa.f90:456-489):
// scfcheck = mean over modes of |Fn - Fnb|^2 / |Fn|^2 (partials per point added in order),
// stop when the last three add up to less than tol; after iteration 15 the mixing factor
// shrinks by 2% per iteration while above 0.5.

// Heat capacity of every mode and temperature, and cv w_q / V, the weight of the kappa log
// (zero below lo_freqtol or at T = 0).
__global__ void k_cv(int qb, int nb, int classical, const double *__restrict__ temps, const double *__restrict__ om,
                     const double *__restrict__ wq, double volume, double *__restrict__ cvm, double *__restrict__ cvw,
                     long sQ) {
    const int T = blockIdx.z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= qb) return;
    const double tk = temps[T];
    double cv = 0, w = 0;
    if (!(om[i] < kFreqTol) && !(tk < kTemperatureTol)) {
        cv = classical ? kKbHartree : cv_dev(tk, om[i]);
        w = cv * wq[i / nb] / volume;
    }
    cvm[T * sQ + i] = cv;
    cvw[T * sQ + i] = w;
}

// get_kappa (kappa.f90:52-70) on the device: per mode km[x] = (cv / V) sum_ba S[9 x + 3 b + a]
// F_a v_b (S the operation average of op_average_map), products and sums unfused in the order
// of the host restatement get_kappa (so the per-mode tensors are its bits), then per
// temperature kappa[x] = sum over the modes of km[x] w_q in a fixed order.
__global__ void k_kappa_mode(int qb, const double *__restrict__ S, const double *__restrict__ om,
                             const double *__restrict__ cv_, double volume, const double *__restrict__ Fn_,
                             const double *__restrict__ vel, double *__restrict__ km_, long sQ, long sF) {
    const int T = blockIdx.z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= qb) return;
    double *km = km_ + (size_t)T * qb * 9 + (size_t)i * 9;
    if (om[i] < kFreqTol) {
        for (int x = 0; x < 9; ++x) km[x] = 0.0;
        return;
    }
    const double f = __ddiv_rn(cv_[T * sQ + i], volume);
    const double *F = Fn_ + T * sF + (size_t)i * 3, *v = vel + (size_t)i * 3;
    for (int x = 0; x < 9; ++x) {
        double sx = 0;
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) sx = __dadd_rn(sx, __dmul_rn(__dmul_rn(S[9 * x + 3 * b + a], F[a]), v[b]));
        km[x] = __dmul_rn(f, sx);
    }
}

// The sum over the modes: each thread sums a contiguous range of modes in order, then the
// ranges are added by a fixed tree (wave shuffles, then the waves in order), so the result does
// not depend on scheduling.
__global__ void __launch_bounds__(256) k_kappa_total(int nb, int qb, const double *__restrict__ om,
                                                     const double *__restrict__ km_, const double *__restrict__ wq,
                                                     double *__restrict__ kt) {
    const int T = blockIdx.z;
    const double *km = km_ + (size_t)T * qb * 9;
    const int per = (qb + blockDim.x - 1) / blockDim.x;
    const int i0 = threadIdx.x * per, i1 = min(qb, i0 + per);
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = i0; i < i1; ++i) {
        if (om[i] < kFreqTol) continue;
        const double w = wq[i / nb];
        for (int x = 0; x < 9; ++x) acc[x] = __dadd_rn(acc[x], __dmul_rn(km[(size_t)i * 9 + x], w));
    }
    __shared__ double red[9][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int x = 0; x < 9; ++x) {
        double v = acc[x];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[x][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        double v = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[threadIdx.x][w];
        kt[9 * T + threadIdx.x] = v;
    }
}

__global__ void k_bte_check(int nirr, int nb, const double *__restrict__ g0part_, double *__restrict__ scf_,
                            BteState *__restrict__ st_, long sG, long sScf) {
    BteState *st = st_ + blockIdx.z;
    if (threadIdx.x != 0 || st->done) return;
    const double *g0part = g0part_ + blockIdx.z * sG;
    double *scf = scf_ + blockIdx.z * sScf;
    double s = 0;
    for (int q = 0; q < nirr; ++q) s += g0part[q];
    const int it = st->it + 1;   // 1-based number of the iteration just done
    scf[it - 1] = s / nirr / nb;
    st->it = it;
    if (it >= 3 && scf[it - 1] + scf[it - 2] + scf[it - 3] < st->tol) st->done = 1;
    else if (it >= st->maxiter) st->done = 1;
    else if (it > 15 && st->mix > 0.50) st->mix *= 0.98;
}

}  // namespace

// Cotangent of the linewidths for L = sum_ab G_ab kappa_RTA_ab at temperature iT, kappa_RTA
// the symmetrised single-mode kappa from F0 = v / (2 linewidth) (get_kappa kappa.f90:28-75,
// symmetrize_kappa :308-324, F0 from scattering.f90:278-292).  G and kappa column-major
// (G[3 j + i] pairs with kappa_ij); lo_chop is taken as the identity.  The symmetrisation
// averages R k R^T, so its adjoint averages R^T G R; each mode's kappa is linear in
// 1 / linewidth, so d kappa / d linewidth = -w_q kappa_mode / linewidth.
void kappa_rta_cotangent(const tc3_ctx *c, int iT, const double *G, double *dlw) {
    const int nb = c->nb, nirr = c->mesh.nirr, nsym = (int)c->ops.size();
    const size_t qb = (size_t)nirr * nb;
    std::fill(dlw, dlw + qb, 0.0);
    const double T = c->temps[iT];
    if (T < kTemperatureTol) return;
    double Gs[9] = {0};
    for (int o = 0; o < nsym; ++o) {
        const double *m = c->ops[o].m;  // row-major R
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
                double s = 0;
                for (int j = 0; j < 3; ++j)
                    for (int i = 0; i < 3; ++i) s += m[3 * i + a] * G[3 * j + i] * m[3 * j + b];
                Gs[3 * b + a] += s / nsym;
            }
    }
    for (int q = 0; q < nirr; ++q)
        for (int b = 0; b < nb; ++b) {
            const size_t i = (size_t)q * nb + b;
            const double om = c->om_irr[i];
            const double lw = c->linewidth[(size_t)iT * qb + i];
            if (om < kFreqTol || !(lw > kFreqTol)) continue;
            const double cv = c->classical ? kKbHartree : cv_host(T, om);
            const double *F = &c->F0[((size_t)iT * qb + i) * 3];
            const double *v = &c->vel_irr[i * 3];
            double s = 0;
            for (int o = 0; o < nsym; ++o) {
                const double *m = c->ops[o].m;
                double v0[3], v1[3];
                for (int x = 0; x < 3; ++x) {
                    v0[x] = m[3 * x] * F[0] + m[3 * x + 1] * F[1] + m[3 * x + 2] * F[2];
                    v1[x] = m[3 * x] * v[0] + m[3 * x + 1] * v[1] + m[3 * x + 2] * v[2];
                }
                for (int y = 0; y < 3; ++y)
                    for (int x = 0; x < 3; ++x) s += Gs[3 * y + x] * v0[x] * v1[y];
            }
            const double kmode = cv * s / nsym / c->cr.volume;
            dlw[i] = -c->mesh.weight_irr[q] * kmode / lw;
        }
}

void symmetrize_kappa_tensor(const tc3_ctx *c, double *k) { symmetrize_kappa(c, k); }

// Working memory of tc3_kappa_noalloc: every buffer of the iterative solve is carved from this
// array (one copy per GPU, held until the process ends).  Host code never uses the name as an
// address: bte_scratch_ptr() resolves the device address of the current GPU's copy once.
constexpr size_t kBteScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char bte_scratch[kBteScratchBytes];

size_t bte_scratch_bytes() { return kBteScratchBytes; }

int bte_scratch_ptr(int device, void **ptr) {
    static std::mutex mu;
    static std::map<int, void *> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
        // the device is current (the caller selected it)
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(bte_scratch));
        if (e != hipSuccess || p == nullptr) return device_error(e, "hipGetSymbolAddress(bte_scratch)", __FILE__, __LINE__);
        it = addr.emplace(device, p).first;
    }
    *ptr = it->second;
    return 0;
}

int run_kappa(tc3_ctx *c, int maxiter, double tol, bool offdiag, void *scratch, size_t scratch_bytes) {
    const int nb = c->nb, nirr = c->mesh.nirr, nfull = c->mesh.nfull, nT = c->nT, nsym = (int)c->ops.size();
    const int nrows = c->nrows;
    const size_t qb = (size_t)nirr * nb;
    hipStream_t st = c->stream;
    const auto w0 = std::chrono::steady_clock::now();
    c->kappa_sma.assign(9 * nT, 0.0);
    c->kappa_iter.assign(9 * nT, 0.0);
    c->kappa_od.assign(9 * nT, 0.0);
    c->Fn.assign(qb * 3 * nT, 0.0);
    c->kappa_mode.assign(qb * 9 * nT, 0.0);
    c->kappa_log.assign(nT, {});
    c->niter.assign(nT, 0);
    std::vector<void *> tmp;
    auto cleanup = [&]() {
        for (void *p : tmp) hipFree(p);
        tmp.clear();
    };
    hipError_t e = hipSuccess;
#define TRY(x) do { e = (x); if (e != hipSuccess) { cleanup(); return device_error(e, #x, __FILE__, __LINE__); } } while (0)
    // device buffers of the solve live in the context's pool, so that repeated calls reuse
    // them (and the captured iteration graph below); with a scratch array they are carved
    // from it instead (256-byte aligned), nothing is allocated and no graph is captured
    int nbuf = 0;
    size_t used = 0;
    bool overflow = false;
    if (scratch && offdiag) return c->fail(5, "the off-diagonal term needs device allocations");
    auto dalloc = [&](void **p, size_t n) -> hipError_t {
        if (!scratch) return c->pool_get("bte" + std::to_string(nbuf++), std::max<size_t>(n, 16), p);
        const size_t need = (std::max<size_t>(n, 16) + 255) & ~(size_t)255;
        if (used + need > scratch_bytes) {
            overflow = true;
            *p = scratch;   // never used: the solve stops below
            return hipSuccess;
        }
        *p = (char *)scratch + used;
        used += need;
        return hipSuccess;
    };
    double *d_Fn, *d_F0, *d_qs, *d_Fbb, *d_opm, *d_g0, *d_cvw, *d_vel, *d_klog;
    int *d_opf, *d_rowof, *d_dgoff, *d_dg, *d_dgp, *d_star_off, *d_star;
    std::vector<double> opm(9 * nsym);
    for (int o = 0; o < nsym; ++o) std::memcpy(&opm[9 * o], c->ops[o].m, sizeof(double) * 9);
    std::vector<int> dgoff(qb + 1, 0), dg, dgpart(std::max(nirr, 1), 1);
    for (int q = 0; q < nirr; ++q) {
        for (int b = 0; b < nb; ++b) {
            for (int x : c->degen[q][b]) dg.push_back(x);
            dgoff[(size_t)q * nb + b + 1] = (int)dg.size();
            // the sets partition the bands of q when every member lists the same set (the lists
            // are ascending); then k_bte_update takes the sets in parallel
            const auto &set = c->degen[q][b];
            if (set.empty() || std::find(set.begin(), set.end(), b) == set.end()) dgpart[q] = 0;
            for (int x : set)
                if (c->degen[q][x] != set) dgpart[q] = 0;
        }
        // k_bte_update holds one point's degenerate-set lists in LDS (1024 entries)
        if (dgoff[(size_t)q * nb + nb] - dgoff[(size_t)q * nb] > 1024)
            return c->fail(1, "degenerate sets of one q-point exceed the BTE kernel's capacity");
    }
    const int mi = std::max(maxiter, 1);
    const long sF = (long)qb * 3, sQ = (long)qb, sFbb = (long)nfull * nb * 3,
               sPart = 3l * std::max(nrows, 1) * kSlices, sG = nirr, sScf = mi, sLog = 9l * kKapSlices * mi;
    // the full-grid points of every irreducible point's star, for the fold-out in k_bte_update
    std::vector<int> star_off(nirr + 1, 0), star(std::max(nfull, 1));
    for (int p = 0; p < nfull; ++p) ++star_off[c->mesh.irr_of_full[p] + 1];
    for (int q = 0; q < nirr; ++q) star_off[q + 1] += star_off[q];
    {
        std::vector<int> fill(star_off.begin(), star_off.end() - 1);
        for (int p = 0; p < nfull; ++p) star[fill[c->mesh.irr_of_full[p]]++] = p;
    }
    TRY(dalloc((void **)&d_Fn, sizeof(double) * sF * nT));
    TRY(dalloc((void **)&d_F0, sizeof(double) * sF * nT));
    TRY(dalloc((void **)&d_qs, sizeof(double) * sQ * nT));
    TRY(dalloc((void **)&d_Fbb, sizeof(double) * sFbb * nT));
    double *d_S, *d_cvm, *d_wq, *d_kmode, *d_kt, *d_ks, *d_temps;
    TRY(dalloc((void **)&d_ks, sizeof(double) * 9 * nT));
    TRY(dalloc((void **)&d_temps, sizeof(double) * nT));
    TRY(dalloc((void **)&d_S, sizeof(double) * 81));
    TRY(dalloc((void **)&d_cvm, sizeof(double) * sQ * nT));
    TRY(dalloc((void **)&d_wq, sizeof(double) * std::max(nirr, 1)));
    TRY(dalloc((void **)&d_kmode, sizeof(double) * qb * 9 * nT));
    TRY(dalloc((void **)&d_kt, sizeof(double) * 9 * nT));
    TRY(dalloc((void **)&d_star_off, sizeof(int) * (nirr + 1)));
    TRY(dalloc((void **)&d_star, sizeof(int) * star.size()));
    TRY(dalloc((void **)&d_opm, sizeof(double) * 9 * nsym));
    TRY(dalloc((void **)&d_g0, sizeof(double) * sG * nT));
    TRY(dalloc((void **)&d_opf, sizeof(int) * nfull));
    TRY(dalloc((void **)&d_rowof, sizeof(int) * qb));
    TRY(dalloc((void **)&d_dgoff, sizeof(int) * (qb + 1)));
    TRY(dalloc((void **)&d_dg, sizeof(int) * dg.size()));
    TRY(dalloc((void **)&d_dgp, sizeof(int) * dgpart.size()));
    int *d_fz;
    std::vector<int> fzero(nirr, 0);
    {
        // points where symmetry fixes no vector, so that F vanishes: F(q) = R F(q) for R q = q
        // (the scattering-matrix rows are averaged over these R, so the reference's iteration
        // keeps F(q) in their fixed space up to rounding) and, with time reversal (F(-q) =
        // -F(q)), F(q) = -R F(q) for R q = -q; F = 0 where the signed sum of these R is zero.
        // Time reversal holds for the scattering matrix only when every q2 is visited: a
        // Monte-Carlo grid draws q2 and -q2 independently, the reference keeps the F of
        // Monte-Carlo-noise size there, and so only the proper operations count then.
        const bool tr_fix = c->mesh.timereversal && !c->mc_on && !(c->nfc4 > 0 && c->mc4_on);
        auto matvec3 = [](const double *m, const double *x, double *y) {
            for (int i = 0; i < 3; ++i) y[i] = m[3 * i] * x[0] + m[3 * i + 1] * x[1] + m[3 * i + 2] * x[2];
        };
        for (int q = 0; q < nirr; ++q) {
            const double *qc = &c->mesh.qcart[3 * (size_t)c->mesh.irr_full_index[q]];
            double P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, mx = 0;
            int cnt = 0;
            // only the operations the mesh (and so the row symmetrisation) was built with
            for (int o = 0; o < c->mesh.nops; ++o) {
                double rq[3];
                matvec3(c->ops[o].m, qc, rq);
                for (int sgn = 1; sgn >= (tr_fix ? -1 : 1); sgn -= 2) {
                    double d[3], f[3], s = 0;
                    for (int i = 0; i < 3; ++i) d[i] = qc[i] - sgn * rq[i];
                    matvec3(c->cr.invrec, d, f);
                    for (int i = 0; i < 3; ++i) { f[i] -= std::nearbyint(f[i]); s += f[i] * f[i]; }
                    if (s < kSqTol) {
                        for (int k = 0; k < 9; ++k) P[k] += sgn * c->ops[o].m[k];
                        ++cnt;
                    }
                }
            }
            for (int k = 0; k < 9; ++k) mx = std::max(mx, std::fabs(P[k]));
            fzero[q] = (cnt > 0 && mx < 1e-8 * (double)cnt) ? 1 : 0;
        }
        TRY(dalloc((void **)&d_fz, sizeof(int) * std::max(nirr, 1)));
    }
    TRY(dalloc((void **)&d_cvw, sizeof(double) * sQ * nT));
    TRY(dalloc((void **)&d_vel, sizeof(double) * qb * 3));
    TRY(dalloc((void **)&d_klog, sizeof(double) * sLog * nT));
    // rows stored per orbit: row groups, row offsets, the groups' orbit entries and the F
    // orbit sums
    const bool orb = c->xi_orb;
    const long sFo = orb ? (long)std::max<size_t>(c->xo_F, 1) : 0;
    int4 *d_xorg = nullptr, *d_xoent = nullptr;
    long long *d_xorow = nullptr, *d_xogoff = nullptr;
    int *d_xomem = nullptr;
    double *d_Fo = nullptr;
    // work units of the product over rows stored per orbit: (row group, slice, slices), in
    // decreasing size
    std::vector<int4> units;
    int4 *d_units = nullptr;
    if (orb) {
        const long ncol_max = (long)nfull * nb;
        std::vector<std::pair<long, int4>> wu;
        for (size_t g = 0; g < c->xo_rg.size(); ++g) {
            const int4 rg = c->xo_rg[g];
            const long smax = (c->xi_es == 4) ? std::min(4, kSlices) : kSlices;
            const int ns = (int)std::min<long>(smax, (smax * (long)rg.w + ncol_max - 1) / ncol_max);
            for (int sl = 0; sl < ns; ++sl) wu.push_back({(long)rg.y * ((long)rg.w / ns), make_int4((int)g, sl, ns, 0)});
        }
        std::stable_sort(wu.begin(), wu.end(), [](const std::pair<long, int4> &a, const std::pair<long, int4> &b) {
            return a.first > b.first;
        });
        for (const auto &w : wu) units.push_back(w.second);
        TRY(dalloc((void **)&d_units, sizeof(int4) * units.size()));
    }
    if (orb) {
        TRY(dalloc((void **)&d_xorg, sizeof(int4) * c->xo_rg.size()));
        TRY(dalloc((void **)&d_xorow, sizeof(long long) * c->xo_row.size()));
        TRY(dalloc((void **)&d_xogoff, sizeof(long long) * c->xo_goff.size()));
        TRY(dalloc((void **)&d_xoent, sizeof(int4) * c->xo_ent.size()));
        TRY(dalloc((void **)&d_xomem, sizeof(int) * c->xo_mem.size()));
        TRY(dalloc((void **)&d_Fo, (size_t)c->xi_es * sFo * nT));
    }
    double *d_part, *d_scf;
    BteState *d_st;
    TRY(dalloc((void **)&d_part, sizeof(double) * sPart * nT));
    TRY(dalloc((void **)&d_scf, sizeof(double) * sScf * nT));
    TRY(dalloc((void **)&d_st, sizeof(BteState) * nT));
    if (overflow)
        return c->fail(1, "the iterative solution needs " + std::to_string(used >> 20) + "+ MiB, more than the " +
                              std::to_string(scratch_bytes >> 20) + " MiB scratch array holds");
    // transfers staged through the context's pinned buffer (a host vector without device
    // allocation in scratch mode), 256-byte slots: the model's tables first, then the per-call
    // inputs below; the downloads reuse it after the uploads, which the stream has finished by
    // then
    struct Upload {
        void *dst;
        const void *src;
        size_t n;
    };
    std::vector<Upload> tables = {{d_vel, c->vel_irr.data(), sizeof(double) * qb * 3},
                                  {d_opm, opm.data(), sizeof(double) * 9 * nsym},
                                  {d_opf, c->mesh.op_of_full.data(), sizeof(int) * nfull},
                                  {d_rowof, c->row_of.data(), sizeof(int) * qb},
                                  {d_dgoff, dgoff.data(), sizeof(int) * (qb + 1)},
                                  {d_dg, dg.data(), sizeof(int) * dg.size()},
                                  {d_dgp, dgpart.data(), sizeof(int) * dgpart.size()},
                                  {d_star_off, star_off.data(), sizeof(int) * (nirr + 1)},
                                  {d_star, star.data(), sizeof(int) * star.size()},
                                  {d_fz, fzero.data(), sizeof(int) * nirr}};
    if (orb) {
        tables.push_back({d_units, units.data(), sizeof(int4) * units.size()});
        tables.push_back({d_xorg, c->xo_rg.data(), sizeof(int4) * c->xo_rg.size()});
        tables.push_back({d_xorow, c->xo_row.data(), sizeof(long long) * c->xo_row.size()});
        tables.push_back({d_xogoff, c->xo_goff.data(), sizeof(long long) * c->xo_goff.size()});
        tables.push_back({d_xoent, c->xo_ent.data(), sizeof(int4) * c->xo_ent.size()});
        tables.push_back({d_xomem, c->xo_mem.data(), sizeof(int) * c->xo_mem.size()});
    }
    auto slot = [](size_t n) { return (n + 255) & ~(size_t)255; };
    size_t b_tab = 0;
    for (const Upload &u : tables) b_tab += slot(u.n);
    const size_t b_qT = sizeof(double) * qb * nT, b_F = sizeof(double) * qb * 3 * nT;
    const size_t b_up = slot(b_qT) + slot(b_F) + slot(sizeof(BteState) * nT) + slot(sizeof(double) * 81) +
                        slot(sizeof(double) * std::max(nirr, 1)) + slot(sizeof(double) * nT);
    const size_t b_down = slot(sizeof(double) * sScf * nT) + slot(sizeof(double) * sLog * nT) + slot(b_F) +
                          2 * slot(sizeof(double) * 9 * nT) + slot(sizeof(double) * qb * 9 * nT);
    std::vector<char> stage_host;
    char *stage = nullptr;
    if (scratch) {
        stage_host.resize(std::max(b_tab + b_up, b_down));
        stage = stage_host.data();
    } else {
        TRY(c->pinned(std::max(b_tab + b_up, b_down)));
        stage = (char *)c->pin;
    }
    char *h = stage;
    auto up = [&](void *dst, const void *src, size_t n) -> hipError_t {
        if (n == 0) return hipSuccess;
        std::memcpy(h, src, n);
        const hipError_t r = hipMemcpyAsync(dst, h, n, hipMemcpyHostToDevice, st);
        h += slot(n);
        return r;
    };
    for (const Upload &u : tables) TRY(up(u.dst, u.src, u.n));
    double t_bte = 0, t_od = 0;
    // off-diagonal term: on the device for all temperatures at once (the star rotation of the
    // per-point sums on the host), or all on the host with TC3_OD_HOST
    const bool od_host = std::getenv("TC3_OD_HOST") != nullptr;
    std::vector<double> od_V, od_M;
    if (offdiag && !od_host) {
        const auto a = std::chrono::steady_clock::now();
        const int r = od_device(c, od_M);
        if (r) return r;
        t_od += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    }
    // per temperature on the host: the off-diagonal term and the initial iteration state; the
    // heat capacities, the SMA kappa (get_kappa of F0) and the kappa of the final F come from
    // the device
    std::vector<BteState> hs(nT);
    for (int iT = 0; iT < nT; ++iT) {
        const double T = c->temps[iT];
        hs[iT] = {0, 1, maxiter, 0, 0.95, tol};
        if (T < kTemperatureTol) continue;
        if (offdiag) {
            const auto a = std::chrono::steady_clock::now();
            if (od_host) {
                if (od_V.empty()) od_velocities_host(c, od_V);
                kappa_offdiag(c, iT, T, c->classical, od_V.data(), &c->kappa_od[9 * iT]);
            } else {
                od_rotate(c, &od_M[9 * (size_t)iT * nirr], &c->kappa_od[9 * iT]);
            }
            t_od += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
        }
        if (maxiter <= 0) continue;
        hs[iT].done = 0;
    }
    const auto a = std::chrono::steady_clock::now();
    bool any = false;
    for (int iT = 0; iT < nT; ++iT) any = any || !hs[iT].done;
    {
        double S[81];
        op_average_map(c, S);
        TRY(up(d_F0, c->F0.data(), b_F));
        TRY(hipMemcpyAsync(d_Fn, d_F0, b_F, hipMemcpyDeviceToDevice, st));
        if (any) {
            TRY(up(d_qs, c->qs.data(), b_qT));
            TRY(up(d_st, hs.data(), sizeof(BteState) * nT));
        }
        TRY(up(d_S, S, sizeof(S)));
        TRY(up(d_wq, c->mesh.weight_irr.data(), sizeof(double) * nirr));
        TRY(up(d_temps, c->temps.data(), sizeof(double) * nT));
    }
    const dim3 gm((unsigned)((qb + 255) / 256), 1, nT);
    hipLaunchKernelGGL(k_cv, gm, dim3(256), 0, st, (int)qb, nb, (int)c->classical, d_temps, c->d_om_irr, d_wq,
                       c->cr.volume, d_cvm, d_cvw, sQ);
    // SMA kappa: get_kappa of F0 (kappa.f90:52-70); the per-mode values are overwritten below
    hipLaunchKernelGGL(k_kappa_mode, gm, dim3(256), 0, st, (int)qb, d_S, c->d_om_irr, d_cvm, c->cr.volume, d_F0, d_vel,
                       d_kmode, sQ, sF);
    hipLaunchKernelGGL(k_kappa_total, dim3(1, 1, nT), dim3(256), 0, st, nb, (int)qb, c->d_om_irr, d_kmode, d_wq, d_ks);
    TRY(hipGetLastError());
    if (any) {
        const size_t sXi = c->xi_row * (size_t)nrows;
        const bool f32 = c->xi_es == 4;   // FP32 folded-out F next to FP32 Xi
        const long nj = (long)nfull * nb;
        // F0 folded out for the first product; each update folds out the F it makes
        if (f32)
            hipLaunchKernelGGL(k_foldout<float>, dim3((unsigned)((nj + 255) / 256), 1, nT), dim3(256), 0, st, nfull, nb,
                               c->d_irr_of_full, d_opf, d_opm, d_Fn, (float *)d_Fbb, d_st, sF, sFbb);
        else
            hipLaunchKernelGGL(k_foldout<double>, dim3((unsigned)((nj + 255) / 256), 1, nT), dim3(256), 0, st, nfull,
                               nb, c->d_irr_of_full, d_opf, d_opm, d_Fn, d_Fbb, d_st, sF, sFbb);
        TRY(hipGetLastError());
        // one iteration for every temperature still running: Xi.F, update and fold-out, kappa
        // log, convergence test (kappa.f90:395-489)
        auto iteration = [&]() {
            if (nrows > 0 && orb) {
                const int nent = (int)c->xo_ent.size();
                const long ne = (long)nent * nb;
                const dim3 gs((unsigned)std::max<long>(1, std::min<long>(1024, (ne + 255) / 256)), 1, nT);
                const dim3 gw((unsigned)(units.size() * nT));
                if (f32) {
                    if (nent > 0)
                        hipLaunchKernelGGL(k_orbsum<float>, gs, dim3(256), 0, st, nent, nfull, nb, d_xoent, d_xomem,
                                           (const float *)d_Fbb, (float *)d_Fo, d_st, sFbb, sFo);
                    if (nb % 4 == 0)
                        hipLaunchKernelGGL((k_xi_gemv_o<float, 4, float>), gw, dim3(kBlock), 0, st, nrows, nT, d_units, d_xorg,
                                           d_xorow, d_xogoff, (const float *)c->d_xi, (const float *)d_Fbb,
                                           (const float *)d_Fo, d_part, d_st, c->xo_T, sFbb, sFo, sPart);
                    else
                        hipLaunchKernelGGL((k_xi_gemv_o<float, 1, float>), gw, dim3(kBlock), 0, st, nrows, nT, d_units, d_xorg,
                                           d_xorow, d_xogoff, (const float *)c->d_xi, (const float *)d_Fbb,
                                           (const float *)d_Fo, d_part, d_st, c->xo_T, sFbb, sFo, sPart);
                } else {
                    if (nent > 0)
                        hipLaunchKernelGGL(k_orbsum<double>, gs, dim3(256), 0, st, nent, nfull, nb, d_xoent, d_xomem,
                                           (const double *)d_Fbb, d_Fo, d_st, sFbb, sFo);
                    if (nb % 2 == 0)
                        hipLaunchKernelGGL((k_xi_gemv_o<double, 2, double>), gw, dim3(kBlock), 0, st, nrows, nT, d_units, d_xorg,
                                           d_xorow, d_xogoff, (const double *)c->d_xi, (const double *)d_Fbb,
                                           (const double *)d_Fo, d_part, d_st, c->xo_T, sFbb, sFo, sPart);
                    else
                        hipLaunchKernelGGL((k_xi_gemv_o<double, 1, double>), gw, dim3(kBlock), 0, st, nrows, nT, d_units, d_xorg,
                                           d_xorow, d_xogoff, (const double *)c->d_xi, (const double *)d_Fbb,
                                           (const double *)d_Fo, d_part, d_st, c->xo_T, sFbb, sFo, sPart);
                }
            } else if (nrows > 0) {
                const dim3 gw((nrows + 4 * kRows - 1) / (4 * kRows), kSlices, nT);
                const long ncol = (long)c->xi_row;
                if (!f32 && ncol % 2 == 0)
                    hipLaunchKernelGGL((k_xi_gemv_w<double, 2, double>), gw, dim3(kBlock), 0, st, nrows, ncol,
                                       (const double *)c->d_xi, d_Fbb, d_part, d_st, sXi, sFbb, sPart);
                else if (!f32)
                    hipLaunchKernelGGL((k_xi_gemv_w<double, 1, double>), gw, dim3(kBlock), 0, st, nrows, ncol,
                                       (const double *)c->d_xi, d_Fbb, d_part, d_st, sXi, sFbb, sPart);
                else if (ncol % 4 == 0)
                    hipLaunchKernelGGL((k_xi_gemv_w<float, 4, float>), gw, dim3(kBlock), 0, st, nrows, ncol,
                                       (const float *)c->d_xi, (const float *)d_Fbb, d_part, d_st, sXi, sFbb, sPart);
                else
                    hipLaunchKernelGGL((k_xi_gemv_w<float, 1, float>), gw, dim3(kBlock), 0, st, nrows, ncol,
                                       (const float *)c->d_xi, (const float *)d_Fbb, d_part, d_st, sXi, sFbb, sPart);
            }
            if (f32)
                hipLaunchKernelGGL(k_bte_update<float>, dim3(nirr, 1, nT), dim3(64), 0, st, nb, nrows, d_rowof, d_part,
                                   d_qs, d_F0, d_dgoff, d_dg, d_dgp, d_fz, d_star_off, d_star, d_opf, d_opm, nj, d_Fn,
                                   (float *)d_Fbb, d_g0, d_st, sPart, sQ, sF, sFbb, sG);
            else
                hipLaunchKernelGGL(k_bte_update<double>, dim3(nirr, 1, nT), dim3(64), 0, st, nb, nrows, d_rowof, d_part,
                                   d_qs, d_F0, d_dgoff, d_dg, d_dgp, d_fz, d_star_off, d_star, d_opf, d_opm, nj, d_Fn,
                                   d_Fbb, d_g0, d_st, sPart, sQ, sF, sFbb, sG);
            hipLaunchKernelGGL(k_kappa_sum, dim3(kKapSlices, 1, nT), dim3(256), 0, st, (int)qb, nsym, d_cvw, d_Fn,
                               d_vel, d_opm, d_klog, d_st, sQ, sF, sLog);
            hipLaunchKernelGGL(k_bte_check, dim3(1, 1, nT), dim3(64), 0, st, nirr, nb, d_g0, d_scf, d_st, sG, sScf);
        };
        // a chunk of kChunkIt iterations, then chunks of kChunkIt2 until the device reports
        // every temperature done (iterations past convergence are no-ops), each chunk captured
        // once as a graph and replayed; direct launches if capture is not available.  The
        // instantiated graphs are kept in the context while every argument they captured
        // (buffers, sizes, strides) stays the same.
        constexpr int kChunkIt = 8, kChunkIt2 = 4;
        const std::vector<long> key = {(long)(uintptr_t)c->d_xi, (long)c->xi_es, (long)c->xi_row, nrows, nfull, nb,
                                       nirr, nT, nsym, (long)(uintptr_t)c->d_irr_of_full, (long)(uintptr_t)d_Fn,
                                       (long)(uintptr_t)d_F0, (long)(uintptr_t)d_qs, (long)(uintptr_t)d_Fbb,
                                       (long)(uintptr_t)d_star_off, (long)(uintptr_t)d_star,
                                       (long)(uintptr_t)d_opm, (long)(uintptr_t)d_g0,
                                       (long)(uintptr_t)d_opf, (long)(uintptr_t)d_rowof, (long)(uintptr_t)d_dgoff, (long)(uintptr_t)d_dgp,
                                       (long)(uintptr_t)d_dg, (long)(uintptr_t)d_cvw, (long)(uintptr_t)d_vel,
                                       (long)(uintptr_t)d_klog, (long)(uintptr_t)d_part, (long)(uintptr_t)d_scf,
                                       (long)(uintptr_t)d_st, sF, sFbb, sPart, sG, sScf, sLog, (long)orb,
                                       (long)c->xo_T, (long)c->xo_rg.size(), (long)c->xo_ent.size(), sFo,
                                       (long)(uintptr_t)d_xorg, (long)(uintptr_t)d_xorow, (long)(uintptr_t)d_xogoff,
                                       (long)(uintptr_t)d_xoent, (long)(uintptr_t)d_xomem, (long)(uintptr_t)d_Fo,
                                       (long)units.size(), (long)(uintptr_t)d_units};
        if (!scratch && c->bte_key != key) {
            for (hipGraphExec_t *g : {&c->bte_exec, &c->bte_exec2})
                if (*g) {
                    (void)hipGraphExecDestroy(*g);
                    *g = nullptr;
                }
        }
        auto capture = [&](hipGraphExec_t *exec, int nit) {
            if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return;
            hipGraph_t graph = nullptr;
            for (int k = 0; k < nit; ++k) iteration();
            if (hipStreamEndCapture(st, &graph) != hipSuccess || hipGraphInstantiate(exec, graph, nullptr, nullptr, 0) != hipSuccess)
                *exec = nullptr;
            if (graph) (void)hipGraphDestroy(graph);
        };
        if (!scratch && (!c->bte_exec || !c->bte_exec2)) {
            if (!c->bte_exec) capture(&c->bte_exec, kChunkIt);
            if (!c->bte_exec2) capture(&c->bte_exec2, kChunkIt2);
            c->bte_key = key;
        }
        (void)hipGetLastError();
        for (int chunk = 0, done = 0; !done; ++chunk) {
            const hipGraphExec_t gexec = scratch ? nullptr : (chunk == 0 ? c->bte_exec : c->bte_exec2);
            if (gexec) TRY(hipGraphLaunch(gexec, st));
            else
                for (int k = 0; k < (chunk == 0 ? kChunkIt : kChunkIt2); ++k) iteration();
            TRY(hipGetLastError());
            TRY(hipMemcpyAsync(hs.data(), d_st, sizeof(BteState) * nT, hipMemcpyDeviceToHost, st));
            TRY(hipStreamSynchronize(st));
            done = 1;
            for (int iT = 0; iT < nT; ++iT) done = done && hs[iT].done;
        }
    }
    // kappa of the final F, per mode and in total, on the device; then the final F, the kappa
    // log (kappa.f90:473-474) and the kappa tensors to the host in one batch
    hipLaunchKernelGGL(k_kappa_mode, dim3((unsigned)((qb + 255) / 256), 1, nT), dim3(256), 0, st, (int)qb, d_S,
                       c->d_om_irr, d_cvm, c->cr.volume, d_Fn, d_vel, d_kmode, sQ, sF);
    hipLaunchKernelGGL(k_kappa_total, dim3(1, 1, nT), dim3(256), 0, st, nb, (int)qb, c->d_om_irr, d_kmode, d_wq, d_kt);
    TRY(hipGetLastError());
    {
        char *h = stage;
        auto down = [&](const void *src, size_t n, double **dst) -> hipError_t {
            *dst = (double *)h;
            h += slot(n);
            return hipMemcpyAsync(*dst, src, n, hipMemcpyDeviceToHost, st);
        };
        double *scf, *klog, *Fall, *ksd, *kt, *km;
        TRY(down(d_scf, sizeof(double) * sScf * nT, &scf));
        TRY(down(d_klog, sizeof(double) * sLog * nT, &klog));
        TRY(down(d_Fn, b_F, &Fall));
        TRY(down(d_ks, sizeof(double) * 9 * nT, &ksd));
        TRY(down(d_kt, sizeof(double) * 9 * nT, &kt));
        TRY(down(d_kmode, sizeof(double) * qb * 9 * nT, &km));
        TRY(hipStreamSynchronize(st));
        std::memcpy(c->Fn.data(), Fall, b_F);
        for (int iT = 0; iT < nT; ++iT) {
            if (c->temps[iT] < kTemperatureTol) continue;
            std::memcpy(&c->kappa_mode[(size_t)iT * qb * 9], &km[(size_t)iT * qb * 9], sizeof(double) * qb * 9);
            double *ks = &c->kappa_sma[9 * iT], *ki = &c->kappa_iter[9 * iT], *ko = &c->kappa_od[9 * iT];
            std::memcpy(ks, &ksd[9 * iT], sizeof(double) * 9);
            chop_kappa(ks);
            std::memcpy(ki, &kt[9 * iT], sizeof(double) * 9);
            chop_kappa(ki);
            if (any && maxiter > 0) {
                const int it_done = hs[iT].it;
                c->niter[iT] = it_done;
                auto &log = c->kappa_log[iT];
                // the first record is the kappa of F0 (kappa.f90:377), i.e. the SMA value
                log.insert(log.end(), {ks[0], ks[4], ks[8], ks[3], ks[6], ks[7], 0.0});
                for (int it = 0; it < it_done; ++it) {
                    double kk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                    for (int sl = 0; sl < kKapSlices; ++sl)
                        for (int i = 0; i < 9; ++i) kk[i] += klog[iT * sLog + (9 * kKapSlices) * (size_t)it + 9 * sl + i];
                    chop_kappa(kk);
                    log.insert(log.end(), {kk[0], kk[4], kk[8], kk[3], kk[6], kk[7], scf[iT * sScf + it]});
                }
            }
            symmetrize_kappa(c, ki);
            symmetrize_kappa(c, ko);
            symmetrize_kappa(c, ks);
        }
    }
    t_bte += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
#undef TRY
    cleanup();
    c->timings[2] = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
    c->timings[3] = t_bte;
    c->timings[4] = t_od;
    return 0;
}

}  // namespace tc3
