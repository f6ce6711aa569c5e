// Adjoint of the three-phonon linewidths with respect to the third-order force constants.
//
// The linewidth of row (q1, b1) is, before the degeneracy average (scattering.f90:260-277),
//   Gamma_raw = sum_trip sum_{b2,b3} c(b1,b2,b3) |Psi(b1,b2,b3)|^2,
// with c the occupation-weighted smeared energy conservation of every temperature times the
// prefactor, the permutation factor and the number of equivalent pairs
// (scattering_threephonon.f90:88-135), Psi = sum E1 E2 E3 Phi and
// Phi(q2,q3) = sum_t mwm_t e^{-2 pi i (q2.lv2_t + q3.lv3_t)} (pretransform_phi3 :275-317).
// For a cotangent g of the raw linewidths (all temperatures) the gradient is
//   dL/dmwm_t[i,j,k] = sum_trip 2 Re( ph_t Y[3a1+i, 3a2+j, 3a3+k] ),
//   Y[ia,ib,ic] = sum_b E1[ia,b1] E2[ib,b2] E3[ic,b3] Z[b],   Z = H conj(Psi),
//   H[b] = sum_T g_T(q1,b1) c_T(b).
// Psi is recomputed per batch of triplets (never stored), Z is formed in place, Y comes from
// the same GEMM chain run backwards, and the gather over triplets is a deterministic
// two-level reduction: per (IFC-chunk, triplet-chunk) partial sums, then a sum over chunks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "tc3_chain.hpp"
#include "tc3_ctx.hpp"

namespace tc3 {

namespace {

struct AdjArgs {
    int nb, nT, t0, nirr, nrows, integrationtype;
    double sigma_fixed, psi_scale;  // pi/16 / N_q
    const int *tq1, *tq2, *tq3, *eq_off, *irr_of_full, *row_of;
    const double *perm, *om_irr, *om_full, *be, *sigsq;
    const double *g;                // [T][row] cotangent of the raw linewidths
    const double *u_irr, *u_full;   // integration type 6: reduced group velocities [q][mode][3]
    double thresh6;
    const float *tw;                // Monte-Carlo draw counts [t][b1], or null (all 1)
};

// Z[tl,b1](b2,b3) = H conj(Psi) in place; H as in the file comment, zero where the forward
// pass skips the element (scattering_threephonon.f90:91, :101) or b1 is not a row.
// Workgroup per (triplet of the batch, b1).
template <typename T>
__global__ void k_adj_weight(AdjArgs a, cplx<T> *__restrict__ psi) {
    const int nb = a.nb;
    const int tl = blockIdx.x / nb, b1 = blockIdx.x % nb;
    const int t = a.t0 + tl;
    const int q1 = a.tq1[t], q2 = a.tq2[t], q3 = a.tq3[t];
    const int row = a.row_of[q1 * nb + b1];
    cplx<T> *z = psi + ((size_t)tl * nb + b1) * nb * nb;
    if (row < 0) {
        for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) z[e] = {(T)0, (T)0};
        return;
    }
    const int i2 = a.irr_of_full[q2], i3 = a.irr_of_full[q3];
    const double w1 = a.om_irr[q1 * nb + b1];
    const double sg1 = a.sigsq[q1 * nb + b1];
    const double scale = a.perm[t] * a.psi_scale * (double)(a.eq_off[t + 1] - a.eq_off[t]) *
                         (a.tw ? (double)a.tw[(size_t)t * nb + b1] : 1.0);
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        const int b2 = e / nb, b3 = e % nb;
        const double w2 = a.om_full[(size_t)q2 * nb + b2], w3 = a.om_full[(size_t)q3 * nb + b3];
        double h = 0.0;
        if (!(w2 < 1e-9) && !(fabs(w1 - w2) < 1e-9) && !(w3 < 1e-9) && !(fabs(w1 - w3) < 1e-9) &&
            !(fabs(w2 - w3) < 1e-9)) {
            const double sig =
                (a.integrationtype == 2) ? sqrt(sg1 + a.sigsq[i2 * nb + b2] + a.sigsq[i3 * nb + b3])
                : (a.integrationtype == 6)
                    ? sigma_type6(a.u_irr + ((size_t)q1 * nb + b1) * 3, a.u_full + ((size_t)q2 * nb + b2) * 3,
                                  a.u_full + ((size_t)q3 * nb + b3) * 3, a.thresh6)
                    : a.sigma_fixed;
            const double d0 = gauss(w1, -w2 + w3, sig) - gauss(w1, w2 - w3, sig);
            const double d1 = gauss(w1, w2 + w3, sig) - gauss(w1, -w2 - w3, sig);
            for (int T_ = 0; T_ < a.nT; ++T_) {
                const double *be = a.be + (size_t)T_ * a.nirr * nb;
                const double n2 = be[i2 * nb + b2], n3 = be[i3 * nb + b3];
                h += a.g[(size_t)T_ * a.nrows + row] * ((n2 - n3) * d0 + (n2 + n3 + 1.0) * d1);
            }
            h *= scale;
        }
        const cplx<T> p = z[e];
        z[e] = {(T)(h * (double)p.x), (T)(-h * (double)p.y)};
    }
}

constexpr int kAdjF = 9;       // IFC triplets per gather workgroup: 9 x 27 components = 243 lanes
constexpr int kAdjStage = 32;  // triplets whose phases are staged in LDS at a time

// part[chunk][27 f + 9 i + 3 j + k] = sum over the chunk's triplets of 2 Re(ph_f Y[...]).
// blockIdx.x: group of kAdjF IFC triplets; blockIdx.y: chunk of per_chunk batch triplets.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_gather(int nb, int nfc3, int nt, int t0, int per_chunk,
                                                    const int *__restrict__ tq2, const int *__restrict__ tq3,
                                                    const double *__restrict__ qcart, const int *__restrict__ fa,
                                                    const double *__restrict__ lv2, const double *__restrict__ lv3,
                                                    const cplx<T> *__restrict__ Y, double *__restrict__ part) {
    __shared__ double phx[kAdjStage][kAdjF], phy[kAdjStage][kAdjF];
    const int f0 = blockIdx.x * kAdjF;
    const int c0 = blockIdx.y * per_chunk;
    const int c1 = min(nt, c0 + per_chunk);
    const int fl = threadIdx.x / 27, comp = threadIdx.x % 27;
    const int f = f0 + fl;
    const bool act = threadIdx.x < 27 * kAdjF && f < nfc3;
    long off = 0;
    if (act) {
        const int i = comp / 9, j = (comp / 3) % 3, k = comp % 3;
        off = ((long)(3 * fa[3 * f] + i) * nb + (3 * fa[3 * f + 1] + j)) * nb + (3 * fa[3 * f + 2] + k);
    }
    const long nb3 = (long)nb * nb * nb;
    double acc = 0.0;
    for (int s0 = c0; s0 < c1; s0 += kAdjStage) {
        const int ns = min(kAdjStage, c1 - s0);
        for (int p = threadIdx.x; p < ns * kAdjF; p += blockDim.x) {
            const int tt = p / kAdjF, ff = p % kAdjF, fg = f0 + ff;
            if (fg < nfc3) {
                const double *q2 = qcart + 3 * tq2[t0 + s0 + tt];
                const double *q3 = qcart + 3 * tq3[t0 + s0 + tt];
                const double arg = -kTwoPi * (q2[0] * lv2[3 * fg] + q2[1] * lv2[3 * fg + 1] + q2[2] * lv2[3 * fg + 2] +
                                              q3[0] * lv3[3 * fg] + q3[1] * lv3[3 * fg + 1] + q3[2] * lv3[3 * fg + 2]);
                double s, c;
                sincos(arg, &s, &c);
                phx[tt][ff] = c;
                phy[tt][ff] = s;
            }
        }
        __syncthreads();
        if (act)
            for (int tt = 0; tt < ns; ++tt) {
                const cplx<T> y = Y[(long)(s0 + tt) * nb3 + off];
                acc += phx[tt][fl] * (double)y.x - phy[tt][fl] * (double)y.y;
            }
        __syncthreads();
    }
    if (act) part[(size_t)blockIdx.y * 27 * nfc3 + 27 * (size_t)f + comp] = 2.0 * acc;
}

// grad[i] += sum_c part[c][i], chunks in order (deterministic).
__global__ void k_adj_reduce(int n, int nchunk, const double *__restrict__ part, double *__restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * n + i];
    grad[i] += s;
}

// One batch: Psi by the forward chain, Z in place, Y by the chain run backwards (into the
// Phi buffer), then the gather into the gradient.
template <typename T, bool SPLIT>
int adjoint_batches(tc3_ctx *c, hipStream_t st, long ntrip, long batch, void *d_work, const void *d_E1,
                    const void *d_Ef, AdjArgs &aa, const int *d_tq1, const int *d_tq2, const int *d_tq3,
                    const double *d_qcart, const int *d_a1off, const int *d_blk3, const double *d_lv2,
                    const double *d_lv3, const double *d_mwm, const int *d_fa, double *d_part, double *d_grad,
                    int ntchunk, size_t phi_shmem) {
    const int nb = c->nb, nfc3 = c->nfc3;
    const long nb2 = (long)nb * nb, nb3 = nb2 * nb;
    const int nfchunk = (nfc3 + kAdjF - 1) / kAdjF;
    const auto *E1 = (const cplx<T> *)d_E1;
    const auto *Ef = (const cplx<T> *)d_Ef;
    auto *w = (cplx<T> *)d_work;
    cplx<T> *phi = w, *t1 = w + (size_t)batch * nb3, *t2 = t1 + (size_t)batch * nb3, *psi = t2 + (size_t)batch * nb3;
    const int tm = (nb + 15) / 16;
    for (long t0 = 0; t0 < ntrip; t0 += batch) {
        const int nt = (int)std::min<long>(batch, ntrip - t0);
        launch_chain<T, SPLIT>(c, st, (int)t0, nt, E1, Ef, phi, t1, t2, psi, d_tq1, d_tq2, d_tq3, d_qcart, d_a1off,
                               d_blk3, d_lv2, d_lv3, d_mwm, phi_shmem);
        aa.t0 = (int)t0;
        hipLaunchKernelGGL(k_adj_weight<T>, dim3(nt * nb), dim3(kBlock), 0, st, aa, psi);
        // t2[tl,b1](b2, ic) = sum_b3 Z[tl,b1](b2, b3) E3[q3](ic, b3)
        hipLaunchKernelGGL((k_cgemm<T, SPLIT>), dim3((unsigned)((long)nt * nb * tm * tm)), dim3(256), 0, st, nb, nb,
                           nb, tm, tm, (const cplx<T> *)psi, (const int *)nullptr, 1, nb2, nb, 1, Ef, d_tq3 + t0, nb,
                           nb2, 1, nb, t2, nb2, nb, 1);
        // t1[tl,b1](ib, ic) = sum_b2 E2[q2](ib, b2) t2[tl,b1](b2, ic)
        hipLaunchKernelGGL((k_cgemm<T, SPLIT>), dim3((unsigned)((long)nt * nb * tm * tm)), dim3(256), 0, st, nb, nb,
                           nb, tm, tm, Ef, d_tq2 + t0, nb, nb2, nb, 1, (const cplx<T> *)t2, (const int *)nullptr, 1,
                           nb2, nb, 1, t1, nb2, nb, 1);
        // Y[tl](ia, (ib, ic)) = sum_b1 E1[q1](ia, b1) t1[tl](b1, (ib, ic))
        {
            const int ntn = (int)((nb2 + 15) / 16);
            hipLaunchKernelGGL((k_cgemm<T, SPLIT>), dim3((unsigned)((long)nt * tm * ntn)), dim3(256), 0, st, nb,
                               (int)nb2, nb, tm, ntn, E1, d_tq1 + t0, 1, nb2, nb, 1, (const cplx<T> *)t1,
                               (const int *)nullptr, 1, nb3, (int)nb2, 1, phi, nb3, (int)nb2, 1);
        }
        const int per_chunk = (nt + ntchunk - 1) / ntchunk;
        const int nch = (nt + per_chunk - 1) / per_chunk;
        hipLaunchKernelGGL(k_adj_gather<T>, dim3(nfchunk, nch), dim3(256), 0, st, nb, nfc3, nt, (int)t0, per_chunk,
                           d_tq2, d_tq3, d_qcart, d_fa, d_lv2, d_lv3, (const cplx<T> *)phi, d_part);
        const int n = 27 * nfc3;
        hipLaunchKernelGGL(k_adj_reduce, dim3((n + 255) / 256), dim3(256), 0, st, n, nch, d_part, d_grad);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return device_error(e, "adjoint batch", __FILE__, __LINE__);
    }
    return 0;
}

}  // namespace

// Gradient of L = sum_T sum_row g[T][row] Gamma_raw_T(row) with respect to the mass-weighted
// IFC tensors in the context's sorted order (27 per triplet, [i][j][k]).
int run_adjoint_gamma(tc3_ctx *c, int precision, const double *g, std::vector<double> &grad) {
    const int na = c->cr.na, nb = 3 * na;
    const int nirr = c->mesh.nirr, nfull = c->mesh.nfull, nT = c->nT, nrows = c->nrows, nfc3 = c->nfc3;
    const long ntrip = (long)c->trip.q1.size();
    hipStream_t st = c->stream;
    const auto wall0 = std::chrono::steady_clock::now();
    grad.assign(27 * (size_t)nfc3, 0.0);
    if (ntrip == 0 || nrows == 0 || nfc3 == 0) return 0;
    hipError_t e = hipSuccess;
#define TRY(x) do { e = (x); if (e != hipSuccess) return device_error(e, #x, __FILE__, __LINE__); } while (0)
    // the forward pass's tables are current: every model change clears have_scatter
    const bool upload = c->scatter_tables_gen != c->gen_model;
    const bool upload_adj = c->adjoint_tables_gen != c->gen_model;
    auto table = [&](const char *name, void **p, const void *src, size_t bytes, bool up) -> hipError_t {
        bool fresh = false;
        hipError_t r = c->pool_get(name, bytes, p, &fresh);
        if (r == hipSuccess && (up || fresh) && bytes > 0)
            r = hipMemcpyAsync(*p, src, bytes, hipMemcpyHostToDevice, st);
        return r;
    };
    int *d_tq1, *d_tq2, *d_tq3, *d_eqoff, *d_rowof, *d_a1off, *d_blk3, *d_fa;
    double *d_perm, *d_qcart, *d_lv2, *d_lv3, *d_mwm, *d_be, *d_sigsq, *d_g, *d_part, *d_grad;
    TRY(table("tq1", (void **)&d_tq1, c->trip.q1.data(), sizeof(int) * ntrip, upload));
    TRY(table("tq2", (void **)&d_tq2, c->trip.q2.data(), sizeof(int) * ntrip, upload));
    TRY(table("tq3", (void **)&d_tq3, c->trip.q3.data(), sizeof(int) * ntrip, upload));
    TRY(table("perm", (void **)&d_perm, c->trip.perm.data(), sizeof(double) * ntrip, upload));
    TRY(table("eqoff", (void **)&d_eqoff, c->trip.eq_off.data(), sizeof(int) * (ntrip + 1), upload));
    TRY(table("rowof", (void **)&d_rowof, c->row_of.data(), sizeof(int) * nirr * nb, upload));
    TRY(table("a1off", (void **)&d_a1off, c->a1_off.data(), sizeof(int) * (na + 1), upload));
    TRY(table("blk3", (void **)&d_blk3, c->blk3_off.data(), sizeof(int) * ((size_t)na * na * na + 1), upload));
    TRY(table("qcart", (void **)&d_qcart, c->mesh.qcart.data(), sizeof(double) * 3 * nfull, upload));
    TRY(table("lv2", (void **)&d_lv2, c->t_lv2.data(), sizeof(double) * 3 * nfc3, upload));
    TRY(table("lv3", (void **)&d_lv3, c->t_lv3.data(), sizeof(double) * 3 * nfc3, upload));
    TRY(table("mwm", (void **)&d_mwm, c->t_mwm.data(), sizeof(double) * 27 * nfc3, upload));
    c->scatter_tables_gen = c->gen_model;
    std::vector<int> fa(3 * (size_t)nfc3);
    for (int f = 0; f < nfc3; ++f) {
        fa[3 * f] = c->t_a1[f];
        fa[3 * f + 1] = c->t_a2[f];
        fa[3 * f + 2] = c->t_a3[f];
    }
    TRY(table("adj_fa", (void **)&d_fa, fa.data(), sizeof(int) * fa.size(), upload_adj));
    c->adjoint_tables_gen = c->gen_model;
    TRY(c->pool_get("be", sizeof(double) * nT * nirr * nb, (void **)&d_be));
    TRY(c->pool_get("sigsq", sizeof(double) * nirr * nb, (void **)&d_sigsq));
    TRY(c->pool_get("adj_g", sizeof(double) * (size_t)nT * nrows, (void **)&d_g));
    host_occupations(c);
    TRY(hipMemcpyAsync(d_be, c->be.data(), sizeof(double) * nT * nirr * nb, hipMemcpyHostToDevice, st));
    TRY(hipMemcpyAsync(d_sigsq, c->sigsq.data(), sizeof(double) * nirr * nb, hipMemcpyHostToDevice, st));
    TRY(hipMemcpyAsync(d_g, g, sizeof(double) * (size_t)nT * nrows, hipMemcpyHostToDevice, st));

    // batch of four nb^3 complex intermediates per triplet, as in the forward generic path
    const long nb3 = (long)nb * nb * nb;
    const size_t elem = (precision == TC3_PREC_FP64) ? sizeof(cd) : sizeof(cf);
    const size_t per_trip = 4 * (size_t)nb3 * elem;
    long batch = std::max<long>(1, std::min<long>(ntrip, (long)((size_t)2 << 30) / (long)per_trip));
    batch = std::min<long>(batch, 1 << 20);
    void *d_work, *d_E1, *d_Ef;
    TRY(c->pool_get("g_work", per_trip * batch, &d_work));
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
    // gather grid: about 2048 workgroups over (IFC chunks) x (triplet chunks)
    const int nfchunk = (nfc3 + kAdjF - 1) / kAdjF;
    const int ntchunk = (int)std::max<long>(1, std::min<long>(batch, 2048 / nfchunk));
    TRY(c->pool_get("adj_part", sizeof(double) * (size_t)ntchunk * 27 * nfc3, (void **)&d_part));
    TRY(c->pool_get("adj_grad", sizeof(double) * 27 * (size_t)nfc3, (void **)&d_grad));
    TRY(hipMemsetAsync(d_grad, 0, sizeof(double) * 27 * (size_t)nfc3, st));
    int maxa1 = 0;
    for (int a = 0; a < na; ++a) maxa1 = std::max(maxa1, c->a1_off[a + 1] - c->a1_off[a]);
    const size_t phi_shmem = sizeof(cd) * std::max(maxa1, 1);

    AdjArgs aa;
    aa.nb = nb; aa.nT = nT; aa.t0 = 0; aa.nirr = nirr; aa.nrows = nrows; aa.integrationtype = c->integrationtype;
    aa.sigma_fixed = c->sigma * kFreqThzToHartree;
    // reduced group velocities of integration type 6, left in the pool by the scattering pass
    aa.u_irr = aa.u_full = nullptr;
    aa.thresh6 = c->thresh6;
    if (c->integrationtype == 6) {
        double *ui, *uf;
        TRY(c->pool_get("u_irr", sizeof(double) * 3 * nirr * nb, (void **)&ui));
        TRY(c->pool_get("u_full", sizeof(double) * 3 * nfull * nb, (void **)&uf));
        aa.u_irr = ui;
        aa.u_full = uf;
    }
    aa.psi_scale = (kPi / 16.0) / (double)(c->trip_w.empty() ? (long)nfull : c->mc_npoints);
    aa.tw = nullptr;
    if (!c->trip_w.empty()) {
        float *tw;
        TRY(table("tw", (void **)&tw, c->trip_w.data(), sizeof(float) * c->trip_w.size(), upload));
        aa.tw = tw;
    }
    aa.tq1 = d_tq1; aa.tq2 = d_tq2; aa.tq3 = d_tq3; aa.eq_off = d_eqoff; aa.irr_of_full = c->d_irr_of_full;
    aa.row_of = d_rowof; aa.perm = d_perm; aa.om_irr = c->d_om_irr; aa.om_full = c->d_om_full; aa.be = d_be;
    aa.sigsq = d_sigsq; aa.g = d_g;
    hipEvent_t ev0, ev1;
    TRY(hipEventCreate(&ev0));
    TRY(hipEventCreate(&ev1));
    TRY(hipEventRecord(ev0, st));
    int r;
    if (precision == TC3_PREC_FP64)
        r = adjoint_batches<double, false>(c, st, ntrip, batch, d_work, d_E1, d_Ef, aa, d_tq1, d_tq2, d_tq3, d_qcart,
                                           d_a1off, d_blk3, d_lv2, d_lv3, d_mwm, d_fa, d_part, d_grad, ntchunk,
                                           phi_shmem);
    else if (precision == TC3_PREC_FP32)
        r = adjoint_batches<float, false>(c, st, ntrip, batch, d_work, d_E1, d_Ef, aa, d_tq1, d_tq2, d_tq3, d_qcart,
                                          d_a1off, d_blk3, d_lv2, d_lv3, d_mwm, d_fa, d_part, d_grad, ntchunk,
                                          phi_shmem);
    else
        r = adjoint_batches<float, true>(c, st, ntrip, batch, d_work, d_E1, d_Ef, aa, d_tq1, d_tq2, d_tq3, d_qcart,
                                         d_a1off, d_blk3, d_lv2, d_lv3, d_mwm, d_fa, d_part, d_grad, ntchunk,
                                         phi_shmem);
    if (r) {
        hipEventDestroy(ev0);
        hipEventDestroy(ev1);
        return r;
    }
    TRY(hipEventRecord(ev1, st));
    TRY(hipMemcpyAsync(grad.data(), d_grad, sizeof(double) * 27 * (size_t)nfc3, hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    float ms = 0;
    hipEventElapsedTime(&ms, ev0, ev1);
    hipEventDestroy(ev0);
    hipEventDestroy(ev1);
#undef TRY
    c->timings[6] = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    c->timings[7] = ms * 1e-3;
    return 0;
}

}  // namespace tc3
