// Directional derivatives (forward mode) of the three-phonon linewidths and of the RTA kappa
// along a change (dPhi2, dPhi3) of the force constants, for gradients with respect to a few
// parameters of the force constants (the Morse parameters of a pair potential, say), one
// direction per call.
//
// The raw linewidth of row (q1, b1) is (scattering_threephonon.f90:88-135)
//   Gamma_raw_T = sum_trip scale sum_{b2,b3} c_T(b) |Psi(b)|^2,
//   c_T = (n2 - n3) d0 + (n2 + n3 + 1) d1,
//   d0 = g(w1; w3 - w2) - g(w1; w2 - w3),  d1 = g(w1; w2 + w3) - g(w1; -w2 - w3),
// g the Gaussian of width sigma (adaptive: sigma^2 = s1 + s2 + s3 of the per-mode widths of
// scattering.f90:131-167), n the Bose occupations and Psi = sum Phi E1 E2 E3 with E = e /
// sqrt(w).  Its change is
//   d Gamma_raw_T = sum scale [dc_T |Psi|^2 + c_T 2 Re(conj(Psi) dPsi)],
//   dPsi = sum (dPhi E1 E2 E3 + Phi dE1 E2 E3 + Phi E1 dE2 E3 + Phi E1 E2 dE3),
// with dE = d(e / sqrt(w)) from the eigenvector and frequency changes of the harmonic step
// (run_harmonic_jvp), dn = -n (n + 1) dw / (kB T) (classical: -n dw / w), and dc from the
// changes of the frequencies and widths in the Gaussians.  The chain of complex GEMMs of the
// forward generic path is run on Psi's operands and, by linearity, on their changes; the sum
// over (b2, b3) is a workgroup reduction per (triplet, b1) and the sum over triplets runs per
// row in triplet order (deterministic).  All in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "tc3_chain.hpp"
#include "tc3_ctx.hpp"

namespace tc3 {

namespace {

constexpr int kJvpT = 8;   // temperatures per pass of the weight kernel

struct JvpArgs {
    int nb, nT, T0, nt, t0, nirr, integrationtype;
    double sigma_fixed, psi_scale;
    const int *tq1, *tq2, *tq3, *eq_off, *irr_of_full, *row_of;
    const double *perm, *om_irr, *om_full, *be, *sigsq;
    const double *dom_irr, *dom_full, *dbe, *dsigsq;
    double *part;   // [t][b1][T]
};

// dE[q][ia][b] = de / sqrt(w) - e dw / (2 w^(3/2)), zero below lo_freqtol (as k_scale)
__global__ void k_dscale(long n, int nb, const cd *__restrict__ egv, const cd *__restrict__ de,
                         const double *__restrict__ om, const double *__restrict__ dom, cplx<double> *__restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long q = i / ((long)nb * nb);
        const int b = (int)(i % nb);
        const double w = om[q * nb + b];
        if (w < 1e-9) {
            out[i] = {0.0, 0.0};
            continue;
        }
        const double s = 1.0 / sqrt(w), h = 0.5 * dom[q * nb + b] / w * s;
        out[i] = {de[i].x * s - egv[i].x * h, de[i].y * s - egv[i].y * h};
    }
}

__global__ void k_caxpy(long n, cplx<double> *__restrict__ y, const cplx<double> *__restrict__ x) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = {y[i].x + x[i].x, y[i].y + x[i].y};
}

// d n / d w dw per irreducible mode and temperature
__global__ void k_dbe(int n, int nT, int classical, const double *__restrict__ temps, const double *__restrict__ om,
                      const double *__restrict__ be, const double *__restrict__ dom, double *__restrict__ dbe) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int T = 0; T < nT; ++T) {
        const double nn = be[(size_t)T * n + i], w = om[i];
        double d = 0.0;
        if (!(w < kFreqTol) && !(temps[T] < kTemperatureTol))
            d = classical ? -nn / w * dom[i] : -nn * (nn + 1.0) / (kKbHartree * temps[T]) * dom[i];
        dbe[(size_t)T * n + i] = d;
    }
}

// lo_gauss and its change along (dx, dmu, dsigma)
__device__ inline void dgauss(double x, double mu, double sigma, double dx, double dmu, double dsigma, double &g,
                              double &dg) {
    const double d = x - mu;
    if (fabs(d) > 8.0 * sigma) {
        g = dg = 0.0;
        return;
    }
    const double inv = 1.0 / sigma;
    g = exp(-d * d * 0.5 * inv * inv) * 0.398942280401433 * inv;
    dg = g * (-d * inv * inv * (dx - dmu) + (d * d * inv * inv * inv - inv) * dsigma);
}

// part[t][b1][T] = sum over (b2, b3) of the change of scale c_T |Psi|^2, workgroup per
// (triplet of the batch, b1), kJvpT temperatures from T0.
__global__ void __launch_bounds__(256) k_jvp_weight(JvpArgs a, const cplx<double> *__restrict__ psi,
                                                    const cplx<double> *__restrict__ dpsi) {
    const int nb = a.nb;
    const int tl = blockIdx.x / nb, b1 = blockIdx.x % nb;
    const int t = a.t0 + tl;
    const int q1 = a.tq1[t], q2 = a.tq2[t], q3 = a.tq3[t];
    const int row = a.row_of[q1 * nb + b1];
    double acc[kJvpT];
    for (int k = 0; k < kJvpT; ++k) acc[k] = 0.0;
    if (row >= 0) {
        const int i2 = a.irr_of_full[q2], i3 = a.irr_of_full[q3];
        const double w1 = a.om_irr[q1 * nb + b1], dw1 = a.dom_irr[q1 * nb + b1];
        const double sg1 = a.sigsq[q1 * nb + b1], dsg1 = a.dsigsq[q1 * nb + b1];
        const double scale = a.perm[t] * a.psi_scale * (double)(a.eq_off[t + 1] - a.eq_off[t]);
        const cplx<double> *p = psi + ((size_t)tl * nb + b1) * nb * nb;
        const cplx<double> *dp = dpsi + ((size_t)tl * nb + b1) * nb * nb;
        for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
            const int b2 = e / nb, b3 = e % nb;
            const double w2 = a.om_full[(size_t)q2 * nb + b2], w3 = a.om_full[(size_t)q3 * nb + b3];
            if (w2 < 1e-9 || fabs(w1 - w2) < 1e-9 || w3 < 1e-9 || fabs(w1 - w3) < 1e-9 || fabs(w2 - w3) < 1e-9)
                continue;
            const double dw2 = a.dom_full[(size_t)q2 * nb + b2], dw3 = a.dom_full[(size_t)q3 * nb + b3];
            double sig = a.sigma_fixed, dsig = 0.0;
            if (a.integrationtype == 2) {
                sig = sqrt(sg1 + a.sigsq[i2 * nb + b2] + a.sigsq[i3 * nb + b3]);
                dsig = (dsg1 + a.dsigsq[i2 * nb + b2] + a.dsigsq[i3 * nb + b3]) / (2.0 * sig);
            }
            double ga, dga, gb, dgb, gc, dgc, gd, dgd;
            dgauss(w1, -w2 + w3, sig, dw1, -dw2 + dw3, dsig, ga, dga);
            dgauss(w1, w2 - w3, sig, dw1, dw2 - dw3, dsig, gb, dgb);
            dgauss(w1, w2 + w3, sig, dw1, dw2 + dw3, dsig, gc, dgc);
            dgauss(w1, -w2 - w3, sig, dw1, -dw2 - dw3, dsig, gd, dgd);
            const double d0 = ga - gb, dd0 = dga - dgb, d1 = gc - gd, dd1 = dgc - dgd;
            const cplx<double> x = p[e], dx = dp[e];
            const double p2 = x.x * x.x + x.y * x.y, dp2 = 2.0 * (x.x * dx.x + x.y * dx.y);
            for (int k = 0; k < kJvpT; ++k) {
                const int T_ = a.T0 + k;
                if (T_ >= a.nT) break;
                const double *be = a.be + (size_t)T_ * a.nirr * nb, *dbe = a.dbe + (size_t)T_ * a.nirr * nb;
                const double n2 = be[i2 * nb + b2], n3 = be[i3 * nb + b3];
                const double dn2 = dbe[i2 * nb + b2], dn3 = dbe[i3 * nb + b3];
                const double cT = (n2 - n3) * d0 + (n2 + n3 + 1.0) * d1;
                const double dcT = (dn2 - dn3) * d0 + (n2 - n3) * dd0 + (dn2 + dn3) * d1 + (n2 + n3 + 1.0) * dd1;
                acc[k] += scale * (dcT * p2 + cT * dp2);
            }
        }
    }
    __shared__ double red[kJvpT][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = 0; k < kJvpT; ++k) {
        double v = acc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[k][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < kJvpT && a.T0 + (int)threadIdx.x < a.nT) {
        const int k = threadIdx.x;
        double v = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[k][w];
        a.part[((size_t)t * nb + b1) * a.nT + a.T0 + k] = v;
    }
}

// dg[T][row] = sum over the triplets of q1, in order, of part[t][b1][T]
__global__ void k_jvp_rows(int nb, int nT, int nrows, const int *__restrict__ row_q, const int *__restrict__ row_b,
                           const int *__restrict__ q1_off, const double *__restrict__ part, double *__restrict__ dg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows * nT) return;
    const int r = i / nT, T = i % nT;
    const int q1 = row_q[r], b1 = row_b[r];
    double s = 0.0;
    for (int t = q1_off[q1]; t < q1_off[q1 + 1]; ++t) s += part[((size_t)t * nb + b1) * nT + T];
    dg[(size_t)T * nrows + r] = s;
}

double cv_of(double T, double om) {
    // lo_harmonic_oscillator_cv (gottochblandat_physics.f90:331-367)
    if (om < kFreqTol || T < kTemperatureTol) return 0.0;
    const double x = om / T / kKbHartree;
    if (x > 200.0) return 0.0;
    if (x > 21.0) return kKbHartree * x * x * std::exp(-x);
    if (x < 2e-2) return kKbHartree * (1.0 - x * x / 12.0 + x * x * x * x / 240.0);
    const double ex = std::exp(x);
    return kKbHartree * x * x * ex / ((ex - 1) * (ex - 1));
}

// d cv / d omega, branch by branch of cv_of
double dcv_of(double T, double om) {
    if (om < kFreqTol || T < kTemperatureTol) return 0.0;
    const double x = om / T / kKbHartree, dxdw = 1.0 / (T * kKbHartree);
    if (x > 200.0) return 0.0;
    if (x > 21.0) return kKbHartree * std::exp(-x) * (2.0 * x - x * x) * dxdw;
    if (x < 2e-2) return kKbHartree * (-x / 6.0 + x * x * x / 60.0) * dxdw;
    const double ex = std::exp(x), em = ex - 1.0;
    return kKbHartree * x * ex / (em * em) * (2.0 + x - 2.0 * x * ex / em) * dxdw;
}

}  // namespace

int run_jvp(tc3_ctx *c, const double *dfc2, const double *dfc3, std::vector<double> &dlw, std::vector<double> &dom,
            std::vector<double> &dvel, std::vector<double> &dks) {
    const int na = c->cr.na, nb = 3 * na;
    const int nirr = c->mesh.nirr, nfull = c->mesh.nfull, nT = c->nT, nrows = c->nrows, nfc3 = c->nfc3;
    const int npair = (int)c->p_a1.size();
    const long ntrip = (long)c->trip.q1.size();
    const size_t qb = (size_t)nirr * nb, nn = (size_t)nb * nb;
    hipStream_t st = c->stream;
    hipError_t e = hipSuccess;
#define TRY(x) do { e = (x); if (e != hipSuccess) return device_error(e, #x, __FILE__, __LINE__); } while (0)
    // the change of the second-order force constants in the sorted order of the pairs, rows as
    // p_m (tc3_set_fc2), and of the third-order ones mass-weighted as t_mwm (tc3_set_fc3)
    std::vector<double> dpm(9 * (size_t)std::max(npair, 1), 0.0), dmwm(27 * (size_t)std::max(nfc3, 1), 0.0);
    if (dfc2)
        for (int n = 0; n < npair; ++n) {
            const int p = c->p_src[n];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) dpm[9 * n + 3 * i + j] = dfc2[9 * p + 3 * j + i];
        }
    if (dfc3) {
        const auto &ism = c->cr.invsqrtmass;
        for (int n = 0; n < nfc3; ++n) {
            const int t = c->t_src[n];
            const double w = ism[c->t_a1[n]] * ism[c->t_a2[n]] * ism[c->t_a3[n]];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) dmwm[27 * n + 9 * i + 3 * j + k] = dfc3[27 * t + 9 * k + 3 * j + i] * w;
        }
    }
    HarmonicJvp hj;
    int r = run_harmonic_jvp(c, dpm, hj);
    if (r) return r;
    dom.assign(qb, 0.0);
    dvel.assign(qb * 3, 0.0);
    TRY(hipMemcpyAsync(dom.data(), hj.dom_irr, sizeof(double) * qb, hipMemcpyDeviceToHost, st));
    TRY(hipMemcpyAsync(dvel.data(), hj.dv_irr, sizeof(double) * qb * 3, hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    // change of the adaptive widths (scattering.f90:131-167): s = max_j (sum_i |v_i| rec_ij /
    // n_j)^2 / 2 per mode, plus 1e-4 of the largest mode-weighted mean
    std::vector<double> dsigsq(qb, 0.0), dsgavg(nb, 0.0), sgavg(nb, 0.0);
    for (int q = 0; q < nirr; ++q)
        for (int b = 0; b < nb; ++b) {
            const size_t i = (size_t)q * nb + b;
            if (c->om_irr[i] < kFreqTol) continue;
            const double *v = &c->vel_irr[i * 3], *dv = &dvel[i * 3];
            double sg = -1e300, dsg = 0.0;
            for (int j = 0; j < 3; ++j) {
                double s = 0, ds = 0;
                for (int k = 0; k < 3; ++k) {
                    const double f = c->cr.rec[3 * k + j] / c->mesh.dims[j];
                    s += std::fabs(v[k]) * f;
                    ds += (v[k] > 0 ? 1.0 : (v[k] < 0 ? -1.0 : 0.0)) * dv[k] * f;
                }
                if (s * s * 0.5 > sg) {
                    sg = s * s * 0.5;
                    dsg = s * ds;
                }
            }
            dsigsq[i] = dsg;
            sgavg[b] += sg * c->mesh.weight_irr[q];
            dsgavg[b] += dsg * c->mesh.weight_irr[q];
        }
    {
        const int bmax = (int)(std::max_element(sgavg.begin(), sgavg.end()) - sgavg.begin());
        for (double &x : dsigsq) x += dsgavg[bmax] * 1e-4;
    }
    // device tables of the forward pass (pool), occupations and widths
    const bool upload = c->scatter_tables_gen != c->gen_model;
    auto table = [&](const char *name, void **p, const void *src, size_t bytes) -> hipError_t {
        bool fresh = false;
        hipError_t rr = c->pool_get(name, bytes, p, &fresh);
        if (rr == hipSuccess && (upload || fresh) && bytes > 0)
            rr = hipMemcpyAsync(*p, src, bytes, hipMemcpyHostToDevice, st);
        return rr;
    };
    int *d_tq1, *d_tq2, *d_tq3, *d_eqoff, *d_rowof, *d_rowq, *d_rowb, *d_q1off, *d_a1off, *d_blk3;
    double *d_perm, *d_qcart, *d_lv2, *d_lv3, *d_mwm;
    TRY(table("tq1", (void **)&d_tq1, c->trip.q1.data(), sizeof(int) * ntrip));
    TRY(table("tq2", (void **)&d_tq2, c->trip.q2.data(), sizeof(int) * ntrip));
    TRY(table("tq3", (void **)&d_tq3, c->trip.q3.data(), sizeof(int) * ntrip));
    TRY(table("perm", (void **)&d_perm, c->trip.perm.data(), sizeof(double) * ntrip));
    TRY(table("eqoff", (void **)&d_eqoff, c->trip.eq_off.data(), sizeof(int) * (ntrip + 1)));
    TRY(table("rowof", (void **)&d_rowof, c->row_of.data(), sizeof(int) * nirr * nb));
    TRY(table("rowq", (void **)&d_rowq, c->row_q.data(), sizeof(int) * nrows));
    TRY(table("rowb", (void **)&d_rowb, c->row_b.data(), sizeof(int) * nrows));
    TRY(table("q1off", (void **)&d_q1off, c->trip.q1_off.data(), sizeof(int) * (nirr + 1)));
    TRY(table("a1off", (void **)&d_a1off, c->a1_off.data(), sizeof(int) * (na + 1)));
    TRY(table("blk3", (void **)&d_blk3, c->blk3_off.data(), sizeof(int) * ((size_t)na * na * na + 1)));
    TRY(table("qcart", (void **)&d_qcart, c->mesh.qcart.data(), sizeof(double) * 3 * nfull));
    TRY(table("lv2", (void **)&d_lv2, c->t_lv2.data(), sizeof(double) * 3 * nfc3));
    TRY(table("lv3", (void **)&d_lv3, c->t_lv3.data(), sizeof(double) * 3 * nfc3));
    TRY(table("mwm", (void **)&d_mwm, c->t_mwm.data(), sizeof(double) * 27 * nfc3));
    c->scatter_tables_gen = c->gen_model;
    double *d_dmwm, *d_be, *d_dbe, *d_sigsq, *d_dsigsq, *d_temps, *d_part, *d_dg;
    TRY(c->pool_get("j_dmwm", sizeof(double) * dmwm.size(), (void **)&d_dmwm));
    TRY(c->pool_get("be", sizeof(double) * nT * qb, (void **)&d_be));
    TRY(c->pool_get("j_dbe", sizeof(double) * nT * qb, (void **)&d_dbe));
    TRY(c->pool_get("sigsq", sizeof(double) * qb, (void **)&d_sigsq));
    TRY(c->pool_get("j_dsigsq", sizeof(double) * qb, (void **)&d_dsigsq));
    TRY(c->pool_get("j_temps", sizeof(double) * nT, (void **)&d_temps));
    TRY(c->pool_get("j_part", sizeof(double) * std::max<long>(ntrip, 1) * nb * nT, (void **)&d_part));
    TRY(c->pool_get("j_dg", sizeof(double) * std::max(nrows, 1) * nT, (void **)&d_dg));
    TRY(hipMemcpyAsync(d_dmwm, dmwm.data(), sizeof(double) * dmwm.size(), hipMemcpyHostToDevice, st));
    host_occupations(c);
    TRY(hipMemcpyAsync(d_be, c->be.data(), sizeof(double) * nT * qb, hipMemcpyHostToDevice, st));
    TRY(hipMemcpyAsync(d_sigsq, c->sigsq.data(), sizeof(double) * qb, hipMemcpyHostToDevice, st));
    TRY(hipMemcpyAsync(d_dsigsq, dsigsq.data(), sizeof(double) * qb, hipMemcpyHostToDevice, st));
    TRY(hipMemcpyAsync(d_temps, c->temps.data(), sizeof(double) * nT, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_dbe, dim3((unsigned)((qb + 255) / 256)), dim3(256), 0, st, (int)qb, nT, (int)c->classical,
                       d_temps, c->d_om_irr, d_be, hj.dom_irr, d_dbe);
    // eigenvectors scaled by 1 / sqrt(w) and their changes, FP64
    cplx<double> *d_E1, *d_Ef, *d_dE1, *d_dEf;
    TRY(c->pool_get("g_E1", sizeof(cd) * nn * nirr, (void **)&d_E1));
    TRY(c->pool_get("g_Ef", sizeof(cd) * nn * nfull, (void **)&d_Ef));
    TRY(c->pool_get("j_dE1", sizeof(cd) * nn * nirr, (void **)&d_dE1));
    TRY(c->pool_get("j_dEf", sizeof(cd) * nn * nfull, (void **)&d_dEf));
    hipLaunchKernelGGL(k_scale<double>, dim3(1024), dim3(256), 0, st, (long)(nn * nirr), nb, c->d_egv_irr, c->d_om_irr,
                       d_E1);
    hipLaunchKernelGGL(k_scale<double>, dim3(4096), dim3(256), 0, st, (long)(nn * nfull), nb, c->d_egv_full,
                       c->d_om_full, d_Ef);
    hipLaunchKernelGGL(k_dscale, dim3(1024), dim3(256), 0, st, (long)(nn * nirr), nb, c->d_egv_irr, hj.de_irr,
                       c->d_om_irr, hj.dom_irr, d_dE1);
    hipLaunchKernelGGL(k_dscale, dim3(4096), dim3(256), 0, st, (long)(nn * nfull), nb, c->d_egv_full, hj.de_full,
                       c->d_om_full, hj.dom_full, d_dEf);
    TRY(hipGetLastError());
    // nine nb^3 complex intermediates per triplet of a batch
    const long nb2 = (long)nb * nb, nb3 = nb2 * nb;
    const size_t per_trip = 9 * (size_t)nb3 * sizeof(cd);
    long batch = std::max<long>(1, std::min<long>(ntrip, (long)((size_t)1 << 30) / (long)per_trip));
    cplx<double> *w;
    TRY(c->pool_get("j_work", per_trip * std::max<long>(batch, 1), (void **)&w));
    cplx<double> *phi = w, *dphi = phi + batch * nb3, *t1 = dphi + batch * nb3, *u1 = t1 + batch * nb3,
                 *t2 = u1 + batch * nb3, *u2 = t2 + batch * nb3, *psi = u2 + batch * nb3, *dpsi = psi + batch * nb3,
                 *tmp = dpsi + batch * nb3;
    int maxa1 = 0;
    for (int a = 0; a < na; ++a) maxa1 = std::max(maxa1, c->a1_off[a + 1] - c->a1_off[a]);
    const size_t phi_shmem = sizeof(cd) * std::max(maxa1, 1);
    JvpArgs ja;
    ja.nb = nb; ja.nT = nT; ja.nirr = nirr; ja.integrationtype = c->integrationtype;
    ja.sigma_fixed = c->sigma * kFreqThzToHartree;
    ja.psi_scale = (kPi / 16.0) / (double)nfull;
    ja.tq1 = d_tq1; ja.tq2 = d_tq2; ja.tq3 = d_tq3; ja.eq_off = d_eqoff; ja.irr_of_full = c->d_irr_of_full;
    ja.row_of = d_rowof; ja.perm = d_perm; ja.om_irr = c->d_om_irr; ja.om_full = c->d_om_full; ja.be = d_be;
    ja.sigsq = d_sigsq; ja.dom_irr = hj.dom_irr; ja.dom_full = hj.dom_full; ja.dbe = d_dbe; ja.dsigsq = d_dsigsq;
    ja.part = d_part;
    const int tm = (nb + 15) / 16, tn2 = (int)((nb2 + 15) / 16);
    // the stages of launch_chain: A contracts ia with E1 (irreducible, by tq1), B ib with E2
    // (full grid, by tq2), C ic with E3 (full grid, by tq3)
    auto stageA = [&](int t0, int nt, const cplx<double> *in, const cplx<double> *E, cplx<double> *out) {
        hipLaunchKernelGGL((k_cgemm<double, false>), dim3((unsigned)((long)nt * tm * tn2)), dim3(256), 0, st, nb,
                           (int)nb2, nb, tm, tn2, E, d_tq1 + t0, 1, nb2, 1, nb, in, (const int *)nullptr, 1, nb3,
                           (int)nb2, 1, out, nb3, (int)nb2, 1);
    };
    auto stageB = [&](int t0, int nt, const cplx<double> *in, const cplx<double> *E, cplx<double> *out) {
        hipLaunchKernelGGL((k_cgemm<double, false>), dim3((unsigned)((long)nt * nb * tm * tm)), dim3(256), 0, st, nb,
                           nb, nb, tm, tm, E, d_tq2 + t0, nb, nb2, 1, nb, in, (const int *)nullptr, 1, nb2, nb, 1, out,
                           nb2, nb, 1);
    };
    auto stageC = [&](int t0, int nt, const cplx<double> *in, const cplx<double> *E, cplx<double> *out) {
        hipLaunchKernelGGL((k_cgemm<double, false>), dim3((unsigned)((long)nt * nb * tm * tm)), dim3(256), 0, st, nb,
                           nb, nb, tm, tm, in, (const int *)nullptr, 1, nb2, nb, 1, E, d_tq3 + t0, nb, nb2, nb, 1, out,
                           nb2, nb, 1);
    };
    auto add = [&](cplx<double> *y, const cplx<double> *x, long n) {
        hipLaunchKernelGGL(k_caxpy, dim3((unsigned)std::min<long>((n + 255) / 256, 4096)), dim3(256), 0, st, n, y, x);
    };
    for (long t0 = 0; t0 < ntrip; t0 += batch) {
        const int nt = (int)std::min<long>(batch, ntrip - t0);
        const long n3 = (long)nt * nb3;
        hipLaunchKernelGGL(k_phi<double>, dim3(nt * na), dim3(kBlock), phi_shmem, st, na, (int)t0, d_tq2, d_tq3,
                           d_qcart, d_a1off, d_blk3, d_lv2, d_lv3, d_mwm, phi);
        hipLaunchKernelGGL(k_phi<double>, dim3(nt * na), dim3(kBlock), phi_shmem, st, na, (int)t0, d_tq2, d_tq3,
                           d_qcart, d_a1off, d_blk3, d_lv2, d_lv3, d_dmwm, dphi);
        stageA((int)t0, nt, phi, d_E1, t1);
        stageA((int)t0, nt, dphi, d_E1, u1);
        stageA((int)t0, nt, phi, d_dE1, tmp);
        add(u1, tmp, n3);
        stageB((int)t0, nt, t1, d_Ef, t2);
        stageB((int)t0, nt, u1, d_Ef, u2);
        stageB((int)t0, nt, t1, d_dEf, tmp);
        add(u2, tmp, n3);
        stageC((int)t0, nt, t2, d_Ef, psi);
        stageC((int)t0, nt, u2, d_Ef, dpsi);
        stageC((int)t0, nt, t2, d_dEf, tmp);
        add(dpsi, tmp, n3);
        ja.t0 = (int)t0;
        ja.nt = nt;
        for (int T0 = 0; T0 < nT; T0 += kJvpT) {
            ja.T0 = T0;
            hipLaunchKernelGGL(k_jvp_weight, dim3(nt * nb), dim3(kBlock), 0, st, ja, (const cplx<double> *)psi,
                               (const cplx<double> *)dpsi);
        }
        TRY(hipGetLastError());
    }
    if (nrows > 0)
        hipLaunchKernelGGL(k_jvp_rows, dim3((nrows * nT + 255) / 256), dim3(256), 0, st, nb, nT, nrows, d_rowq, d_rowb,
                           d_q1off, d_part, d_dg);
    TRY(hipGetLastError());
    std::vector<double> dg((size_t)nT * nrows);
    if (nrows > 0) TRY(hipMemcpyAsync(dg.data(), d_dg, sizeof(double) * dg.size(), hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
#undef TRY
    // degeneracy average of the linewidths (scattering.f90:260-277), linear in them
    dlw.assign(qb * nT, 0.0);
    for (int T = 0; T < nT; ++T) {
        std::vector<double> buf(qb, 0.0);
        for (int rr = 0; rr < nrows; ++rr) buf[(size_t)c->row_q[rr] * nb + c->row_b[rr]] = dg[(size_t)T * nrows + rr];
        for (int q = 0; q < nirr; ++q)
            for (int b = 0; b < nb; ++b) {
                if (c->om_irr[(size_t)q * nb + b] < kFreqTol) continue;
                const auto &dm = c->degen[q][b];
                double f0 = 0;
                for (int b2 : dm) f0 += buf[(size_t)q * nb + b2];
                f0 /= (double)dm.size();
                for (int b2 : dm) buf[(size_t)q * nb + b2] = f0;
                dlw[(size_t)T * qb + (size_t)q * nb + b] = buf[(size_t)q * nb + b];
            }
    }
    // change of the single-mode kappa (get_kappa of F0 = v / (2 Gamma), kappa.f90:28-75, then
    // symmetrize_kappa): S applied to sum_modes w / V [dcv F0 v^T + cv (dF0 v^T + F0 dv^T)]
    dks.assign(9 * (size_t)nT, 0.0);
    double S[81];
    {
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
    for (int T = 0; T < nT; ++T) {
        const double temp = c->temps[T];
        if (temp < kTemperatureTol) continue;
        double X[9] = {0};
        for (int q = 0; q < nirr; ++q)
            for (int b = 0; b < nb; ++b) {
                const size_t i = (size_t)q * nb + b, it = (size_t)T * qb + i;
                const double om = c->om_irr[i], lw = c->linewidth[it];
                if (om < kFreqTol || !(lw > kFreqTol)) continue;
                const double cv = c->classical ? kKbHartree : cv_of(temp, om);
                const double dcv = c->classical ? 0.0 : dcv_of(temp, om) * dom[i];
                const double qs = 2.0 * lw, dqs = 2.0 * dlw[it];
                const double *v = &c->vel_irr[i * 3], *dv = &dvel[i * 3];
                const double wv = c->mesh.weight_irr[q] / c->cr.volume;
                for (int bb = 0; bb < 3; ++bb)
                    for (int a = 0; a < 3; ++a) {
                        const double F = v[a] / qs, dF = dv[a] / qs - v[a] * dqs / (qs * qs);
                        X[3 * bb + a] += wv * (dcv * F * v[bb] + cv * (dF * v[bb] + F * dv[bb]));
                    }
            }
        double k[9];
        for (int x = 0; x < 9; ++x) {
            double s = 0;
            for (int y = 0; y < 9; ++y) s += S[9 * x + y] * X[y];
            k[x] = s;
        }
        symmetrize_kappa_tensor(c, k);
        std::memcpy(&dks[9 * (size_t)T], k, sizeof(k));
    }
    return 0;
}

}  // namespace tc3
