// C ABI of libtc3.so (declared in include/tc3.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <numeric>
#include <string>

#include "../../include/tc3.h"
#include "tc3_ctx.hpp"
#include "tc3_cumulative.hpp"

using namespace tc3;

namespace tc3 {
thread_local std::string g_device_error;

int device_error(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error '%s' in %s (%s:%d)", hipGetErrorString(e), what, file, line);
    g_device_error = buf;
    return 7;
}
}  // namespace tc3

tc3_ctx::~tc3_ctx() {
    if (bte_exec) hipGraphExecDestroy(bte_exec);
    if (bte_exec2) hipGraphExecDestroy(bte_exec2);
    if (sc_exec) hipGraphExecDestroy(sc_exec);
    for (hipEvent_t e : ev_pool) hipEventDestroy(e);
    if (pin) hipHostFree(pin);
    hipFree(d_om_irr);
    hipFree(d_om_full);
    hipFree(d_vel_irr);
    hipFree(d_vel_full);
    hipFree(d_egv_irr);
    hipFree(d_egv_full);
    hipFree(d_xi);
    hipFree(d_irr_of_full);
    for (auto &kv : pool) hipFree(kv.second.p);
    if (stream) hipStreamDestroy(stream);
}

namespace {

// Run f(), turning exceptions and device errors into status codes.
template <typename F>
int guarded(tc3_ctx *c, F f) {
    if (!c) return 5;
    try {
        if (c->device >= 0 && hipSetDevice(c->device) != hipSuccess) return c->fail(7, "cannot select the HIP device");
        g_device_error.clear();
        int r = f();
        if (r == 7 && !g_device_error.empty()) c->err = g_device_error;
        return r;
    } catch (const std::bad_alloc &) {
        return c->fail(5, "host allocation failed");
    } catch (const std::exception &e) {
        return c->fail(5, e.what());
    } catch (...) {
        return c->fail(5, "unknown failure");
    }
}

double planck(double T, double om) {
    // lo_planck (gottochblandat_physics.f90:160-188)
    if (T < kTemperatureTol) return 0.0;
    if (om < kFreqTol) return 0.0;
    const double x = om / (kKbHartree * T);
    if (x > 1e2) return 0.0;
    return 1.0 / (std::exp(x) - 1.0);
}


}  // namespace

namespace tc3 {
// Bose-Einstein occupations of the irreducible modes at the pass's temperatures on the host
// (scattering.f90:148-152; classical: k_B T / omega), zero below lo_freqtol.
void host_occupations(tc3_ctx *ctx) {
    if (ctx->be_host) return;
    const int nb = ctx->nb, nirr = ctx->mesh.nirr, nT = ctx->nT;
    ctx->be.assign((size_t)nT * nirr * nb, 0.0);
    for (int q = 0; q < nirr; ++q)
        for (int b = 0; b < nb; ++b) {
            const double om = ctx->om_irr[(size_t)q * nb + b];
            if (om < kFreqTol) continue;
            for (int T = 0; T < nT; ++T)
                ctx->be[((size_t)T * nirr + q) * nb + b] =
                    ctx->classical ? kKbHartree * ctx->temps[T] / om : planck(ctx->temps[T], om);
        }
    ctx->be_host = true;
}
}  // namespace tc3

extern "C" {

int tc3_create(tc3_ctx **ctx, int device) {
    if (!ctx) return 5;
    *ctx = nullptr;
    if (device == -1) {
        // host-only context: structure, symmetry, mesh and triplet tables, no device work
        tc3_ctx *c = new (std::nothrow) tc3_ctx();
        if (!c) return 5;
        c->device = -1;
        *ctx = c;
        return 0;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 7;
    if (device < 0 || device >= n) return 5;
    tc3_ctx *c = new (std::nothrow) tc3_ctx();
    if (!c) return 5;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return 7;
    }
    *ctx = c;
    return 0;
}

void tc3_destroy(tc3_ctx *ctx) {
    if (!ctx) return;
    if (ctx->device >= 0) hipSetDevice(ctx->device);
    delete ctx;
}

const char *tc3_last_error(const tc3_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tc3_set_structure(tc3_ctx *ctx, int na, const double *latvec, const double *rfrac, const int *species,
                      const double *mass) {
    return guarded(ctx, [&]() -> int {
        if (!latvec || !rfrac || !species || !mass) return ctx->fail(5, "null array");
        Status s = build_crystal(ctx->cr, na, latvec, rfrac, species, mass);
        if (s.code) return ctx->fail(s.code, s.msg);
        ctx->nb = 3 * na;
        ctx->iso_g.assign(na, 0.0);
        ctx->iso_on = false;
        ctx->have_structure = true;
        ctx->have_fc2 = ctx->have_fc3 = ctx->have_mesh = ctx->have_harmonic = ctx->have_scatter = false;
        ++ctx->gen_model;
        return 0;
    });
}

int tc3_set_fc2(tc3_ctx *ctx, int npair, const int *a1, const int *a2, const double *lv2, const double *m) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_structure) return ctx->fail(5, "set the structure first");
        const int na = ctx->cr.na;
        std::vector<int> idx(npair);
        std::iota(idx.begin(), idx.end(), 0);
        for (int p = 0; p < npair; ++p)
            if (a1[p] < 0 || a1[p] >= na || a2[p] < 0 || a2[p] >= na) return ctx->fail(1, "pair atom index out of range");
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
            return std::make_pair(a1[x], a2[x]) < std::make_pair(a1[y], a2[y]);
        });
        ctx->p_a1.resize(npair); ctx->p_a2.resize(npair);
        ctx->p_src = idx;
        ctx->p_lv.resize(3 * (size_t)npair); ctx->p_m.resize(9 * (size_t)npair);
        ctx->blk2_off.assign(na * na + 1, 0);
        for (int n = 0; n < npair; ++n) {
            const int p = idx[n];
            ctx->p_a1[n] = a1[p];
            ctx->p_a2[n] = a2[p];
            for (int i = 0; i < 3; ++i) ctx->p_lv[3 * n + i] = lv2[3 * p + i];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) ctx->p_m[9 * n + 3 * i + j] = m[9 * p + 3 * j + i];
            ctx->blk2_off[a1[p] * na + a2[p] + 1]++;
        }
        for (int b = 0; b < na * na; ++b) ctx->blk2_off[b + 1] += ctx->blk2_off[b];
        ctx->have_fc2 = true;
        ++ctx->gen_model;
        ctx->have_harmonic = ctx->have_scatter = false;
        return 0;
    });
}

int tc3_set_fc3(tc3_ctx *ctx, int ntrip, const int *a1, const int *a2, const int *a3, const double *lv2,
                const double *lv3, const double *m) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_structure) return ctx->fail(5, "set the structure first");
        const int na = ctx->cr.na;
        for (int t = 0; t < ntrip; ++t)
            if (a1[t] < 0 || a1[t] >= na || a2[t] < 0 || a2[t] >= na || a3[t] < 0 || a3[t] >= na)
                return ctx->fail(1, "triplet atom index out of range");
        std::vector<int> idx(ntrip);
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
            return std::make_tuple(a1[x], a2[x], a3[x]) < std::make_tuple(a1[y], a2[y], a3[y]);
        });
        ctx->nfc3 = ntrip;
        ctx->t_a1.resize(ntrip); ctx->t_a2.resize(ntrip); ctx->t_a3.resize(ntrip);
        ctx->t_lv2.resize(3 * (size_t)ntrip); ctx->t_lv3.resize(3 * (size_t)ntrip);
        ctx->t_mwm.resize(27 * (size_t)ntrip);
        ctx->t_src = idx;
        ctx->blk3_off.assign((size_t)na * na * na + 1, 0);
        ctx->a1_off.assign(na + 1, 0);
        const auto &ism = ctx->cr.invsqrtmass;
        for (int n = 0; n < ntrip; ++n) {
            const int t = idx[n];
            ctx->t_a1[n] = a1[t]; ctx->t_a2[n] = a2[t]; ctx->t_a3[n] = a3[t];
            for (int i = 0; i < 3; ++i) {
                ctx->t_lv2[3 * n + i] = lv2[3 * t + i];
                ctx->t_lv3[3 * n + i] = lv3[3 * t + i];
            }
            // type_forceconstant_thirdorder.f90:277
            const double w = ism[a1[t]] * ism[a2[t]] * ism[a3[t]];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k)
                        ctx->t_mwm[27 * n + 9 * i + 3 * j + k] = m[27 * t + 9 * k + 3 * j + i] * w;
            ctx->blk3_off[((size_t)a1[t] * na + a2[t]) * na + a3[t] + 1]++;
            ctx->a1_off[a1[t] + 1]++;
        }
        for (size_t b = 0; b < (size_t)na * na * na; ++b) ctx->blk3_off[b + 1] += ctx->blk3_off[b];
        for (int a = 0; a < na; ++a) ctx->a1_off[a + 1] += ctx->a1_off[a];
        ctx->have_fc3 = true;
        ++ctx->gen_model;
        ctx->have_scatter = false;
        return 0;
    });
}

int tc3_set_qmesh(tc3_ctx *ctx, const int *dims, int timereversal) {
    return tc3_set_qmesh_sym(ctx, dims, timereversal, 1);
}

int tc3_set_qmesh_sym(tc3_ctx *ctx, const int *dims, int timereversal, int qpsymmetry) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_structure) return ctx->fail(5, "set the structure first");
        if (!dims) return ctx->fail(5, "q-grid dimensions missing");
        Status s = spacegroup(ctx->cr, ctx->ops);
        if (s.code) return ctx->fail(s.code, s.msg);
        // without q-point symmetry the mesh is reduced by the identity (the first operation)
        // and time reversal only (build_grid with nosymmetry, type_qpointmesh_gridgeneration.f90:
        // 54-58); the crystal keeps its space group for the kappa averages
        const std::vector<SymOp> mops(ctx->ops.begin(), ctx->ops.begin() + (qpsymmetry ? ctx->ops.size() : 1));
        s = build_mesh(ctx->cr, mops, dims, timereversal != 0, ctx->mesh);
        if (s.code) return ctx->fail(s.code, s.msg);
        s = build_triplets(ctx->mesh, ctx->trip_full);
        if (s.code) return ctx->fail(s.code, s.msg);
        ctx->trip = ctx->trip_full;
        ctx->trip_w.clear();
        ctx->mc_gen = -1;
        // the Monte-Carlo grids default to the full mesh (options.f90:226-236): a three- or
        // four-phonon grid left unset draws every point
        for (int i = 0; i < 3; ++i) ctx->mc_dims[i] = ctx->mc4_dims[i] = ctx->mesh.dims[i];
        ctx->mc_on = ctx->mc4_on = false;
        s = build_orbits(ctx->mesh, ctx->orb);
        if (s.code) return ctx->fail(s.code, s.msg);
        ctx->have_mesh = true;
        ++ctx->gen_model;
        ctx->have_harmonic = ctx->have_scatter = false;
        return 0;
    });
}

int tc3_set_montecarlo4(tc3_ctx *ctx, const int *qg4ph) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_mesh) return ctx->fail(5, "set the q-mesh first");
        bool on = false;
        for (int i = 0; i < 3; ++i) {
            const int d = (!qg4ph || qg4ph[i] < 1 || qg4ph[i] > ctx->mesh.dims[i]) ? ctx->mesh.dims[i] : qg4ph[i];
            ctx->mc4_dims[i] = d;
            on = on || d < ctx->mesh.dims[i];
        }
        ctx->mc4_on = on;
        ctx->mc_gen = -1;
        ++ctx->gen_model;
        ctx->have_scatter = false;
        return 0;
    });
}

int tc3_set_montecarlo(tc3_ctx *ctx, const int *qg3ph, int seed) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_mesh) return ctx->fail(5, "set the q-mesh first");
        bool on = qg3ph != nullptr;
        if (on) {
            on = false;
            for (int i = 0; i < 3; ++i) {
                // options.f90:226-230: outside 1..qgrid means the full grid
                const int d = (qg3ph[i] < 1 || qg3ph[i] > ctx->mesh.dims[i]) ? ctx->mesh.dims[i] : qg3ph[i];
                ctx->mc_dims[i] = d;
                on = on || d < ctx->mesh.dims[i];
            }
        }
        if (on && seed <= 0)
            return ctx->fail(5, "Monte-Carlo grids need a positive seed (a wall-clock seed is not reproducible)");
        ctx->mc_on = on;
        ctx->mc_seed = seed;
        ctx->mc_gen = -1;
        ++ctx->gen_model;
        ctx->have_scatter = false;
        return 0;
    });
}

int tc3_set_thirdorder(tc3_ctx *ctx, int enable) {
    return guarded(ctx, [&]() -> int {
        ctx->thirdorder = enable != 0;
        ctx->mc_gen = -1;
        ctx->have_scatter = false;
        return 0;
    });
}

int tc3_set_boundary(tc3_ctx *ctx, double mfp_max) {
    return guarded(ctx, [&]() -> int {
        if (!(mfp_max == mfp_max)) return ctx->fail(3, "boundary length is not a number");
        ctx->mfp_max = mfp_max;
        ctx->have_scatter = false;
        return 0;
    });
}

int tc3_set_isotopes(tc3_ctx *ctx, const double *disorder, int enable) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_structure) return ctx->fail(5, "set the structure first");
        if (enable && !disorder) return ctx->fail(5, "isotope scattering needs the mass-disorder parameters");
        ctx->iso_g.assign(ctx->cr.na, 0.0);
        if (disorder)
            for (int a = 0; a < ctx->cr.na; ++a) {
                if (!(disorder[a] >= 0.0)) return ctx->fail(3, "negative mass-disorder parameter");
                ctx->iso_g[a] = disorder[a];
            }
        ctx->iso_on = enable != 0;
        ctx->have_scatter = false;
        return 0;
    });
}

int tc3_mesh_info(const tc3_ctx *ctx, int *nsym, int *nfull, int *nirr, int64_t *ntriplets, int *nrows) {
    if (!ctx) return 5;
    if (!ctx->have_mesh) return 5;
    if (nsym) *nsym = (int)ctx->ops.size();
    if (nfull) *nfull = ctx->mesh.nfull;
    if (nirr) *nirr = ctx->mesh.nirr;
    if (ntriplets) *ntriplets = (int64_t)ctx->trip.q1.size();
    if (nrows) *nrows = ctx->have_harmonic ? ctx->nrows : -1;
    return 0;
}

int tc3_get_mesh(const tc3_ctx *ctx, int *irr_full_index, int *irr_of_full, int *op_of_full, double *weight_irr,
                 double *qcart) {
    if (!ctx || !ctx->have_mesh) return 5;
    const auto &m = ctx->mesh;
    if (irr_full_index) std::copy(m.irr_full_index.begin(), m.irr_full_index.end(), irr_full_index);
    if (irr_of_full) std::copy(m.irr_of_full.begin(), m.irr_of_full.end(), irr_of_full);
    if (op_of_full) std::copy(m.op_of_full.begin(), m.op_of_full.end(), op_of_full);
    if (weight_irr) std::copy(m.weight_irr.begin(), m.weight_irr.end(), weight_irr);
    if (qcart) std::copy(m.qcart.begin(), m.qcart.end(), qcart);
    return 0;
}

int tc3_get_symmetry(const tc3_ctx *ctx, double *rot, double *tr, int *fmap) {
    if (!ctx || !ctx->have_mesh) return 5;
    const int na = ctx->cr.na;
    for (size_t o = 0; o < ctx->ops.size(); ++o) {
        const SymOp &op = ctx->ops[o];
        if (rot)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) rot[9 * o + 3 * j + i] = op.m[3 * i + j];
        if (tr)
            for (int i = 0; i < 3; ++i) tr[3 * o + i] = op.tr[i];
        if (fmap)
            for (int a = 0; a < na; ++a) fmap[na * o + a] = op.fmap[a];
    }
    return 0;
}

int tc3_harmonic(tc3_ctx *ctx, double *omega_min) {
    return guarded(ctx, [&]() -> int {
        if (ctx->device < 0) return ctx->fail(7, "host-only context has no device");
        if (!ctx->have_mesh || !ctx->have_fc2) return ctx->fail(5, "set force constants and q-mesh first");
        const auto t0 = std::chrono::steady_clock::now();
        int r = run_harmonic(ctx);
        if (r) return r;
        ctx->timings[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (omega_min) *omega_min = ctx->omega_min;
        ctx->have_harmonic = true;
        ctx->have_scatter = false;
        return 0;
    });
}

int tc3_get_harmonic(const tc3_ctx *ctx, int full, double *omega, double *vel, double *egv) {
    if (!ctx || !ctx->have_harmonic) return 5;
    if (full || egv) {
        // eigenvectors and full-grid results are copied from the device on first request
        tc3_ctx *mc = const_cast<tc3_ctx *>(ctx);
        const int r = guarded(mc, [&]() -> int { return ensure_full_host(mc); });
        if (r) return r;
    }
    const int nb = ctx->nb;
    const int nq = full ? ctx->mesh.nfull : ctx->mesh.nirr;
    const auto &om = full ? ctx->om_full : ctx->om_irr;
    const auto &ve = full ? ctx->vel_full : ctx->vel_irr;
    const auto &eg = full ? ctx->egv_full : ctx->egv_irr;
    if (omega) std::copy(om.begin(), om.end(), omega);
    if (vel) std::copy(ve.begin(), ve.end(), vel);
    if (egv)
        for (int q = 0; q < nq; ++q)
            for (int b = 0; b < nb; ++b)
                for (int i = 0; i < nb; ++i) {
                    const cd x = eg[((size_t)q * nb + i) * nb + b];
                    egv[2 * (((size_t)q * nb + b) * nb + i)] = x.x;
                    egv[2 * (((size_t)q * nb + b) * nb + i) + 1] = x.y;
                }
    return 0;
}

int tc3_scatter(tc3_ctx *ctx, int nT, const double *temperatures, int precision, int integrationtype, double sigma,
                int classical) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_harmonic || !ctx->have_fc3) return ctx->fail(5, "run tc3_harmonic and set fc3 first");
        if (nT <= 0 || !temperatures) return ctx->fail(5, "need at least one temperature");
        if (integrationtype != 1 && integrationtype != 2 && integrationtype != 6)
            return ctx->fail(5, "integrationtype must be 1, 2 or 6");
        if (precision < 0 || precision > 2) return ctx->fail(5, "unknown precision mode");
        if (ctx->omega_min < -kFreqTol) return ctx->fail(3, "FOUND UNSTABLE MODES");
        const int nb = ctx->nb, nirr = ctx->mesh.nirr;
        ctx->nT = nT;
        ctx->temps.assign(temperatures, temperatures + nT);
        ctx->precision = precision;
        ctx->classical = classical != 0;
        ctx->integrationtype = integrationtype;
        ctx->sigma = sigma;
        // adaptive smearing (scattering.f90:131-167); the occupations (:148-152) are formed on
        // the device by the pass (k_occupations), on the host when a derivative needs them
        ctx->be_host = false;
        ctx->sigsq.assign((size_t)nirr * nb, 0.0);
        std::vector<double> sigavg(nb, 0.0);
        const int *dims = ctx->mesh.dims;
        for (int q = 0; q < nirr; ++q)
            for (int b = 0; b < nb; ++b) {
                const double om = ctx->om_irr[(size_t)q * nb + b];
                if (om < kFreqTol) continue;
                const double *v = &ctx->vel_irr[((size_t)q * nb + b) * 3];
                double sg = -1e300;
                for (int j = 0; j < 3; ++j) {
                    double s = 0;
                    for (int i = 0; i < 3; ++i) s += std::fabs(v[i]) * ctx->cr.rec[3 * i + j] / dims[j];
                    sg = std::max(sg, s * s * 0.5);
                }
                ctx->sigsq[(size_t)q * nb + b] = sg;
                sigavg[b] += sg * ctx->mesh.weight_irr[q];
            }
        const double base = *std::max_element(sigavg.begin(), sigavg.end()) * 1e-4;
        for (double &s : ctx->sigsq) s += base;
        ctx->thresh6 = base;
        // triplets of the pass: the Monte-Carlo subset with its draw counts, or all
        // with four-phonon scattering every row draws its three-phonon grid and then its two
        // four-phonon grids from the one generator (full-mesh grids included)
        const bool f4 = ctx->nfc4 > 0;
        if (!ctx->thirdorder && !f4 && !ctx->iso_on && !(ctx->mfp_max > 0.0))
            return ctx->fail(5, "no scattering term is switched on");
        if ((ctx->mc_on && ctx->thirdorder) || (f4 && ctx->mc4_on)) {
            if (ctx->mc_seed <= 0)
                return ctx->fail(5, "Monte-Carlo grids need a positive seed (a wall-clock seed is not reproducible)");
            if (ctx->mc_gen != ctx->gen_model) {
                int d3[3];
                for (int i = 0; i < 3; ++i) d3[i] = ctx->mc_on ? ctx->mc_dims[i] : ctx->mesh.dims[i];
                Status ms = montecarlo_triplets(ctx->mesh, ctx->trip_full, ctx->om_irr, nb, d3, ctx->mc_seed, ctx->trip,
                                                ctx->trip_w, ctx->mc_npoints, f4 ? ctx->mc4_dims : nullptr, &ctx->g4,
                                                &ctx->mc4_npoints, ctx->thirdorder);
                if (ms.code) return ctx->fail(ms.code, ms.msg);
                ctx->mc_gen = ctx->gen_model;
            }
        } else if (!ctx->trip_w.empty()) {
            ctx->trip = ctx->trip_full;
            ctx->trip_w.clear();
        }
        int r = run_scatter(ctx, precision, integrationtype, sigma);
        if (r) return r;
        ctx->have_scatter = true;
        ctx->have_kappa = false;
        return 0;
    });
}

int tc3_get_scattering(const tc3_ctx *ctx, int iT, double *linewidth, double *F0, double *scalar_mfp) {
    if (!ctx || !ctx->have_scatter || iT < 0 || iT >= ctx->nT) return 5;
    const size_t qb = (size_t)ctx->mesh.nirr * ctx->nb;
    if (linewidth) std::copy(&ctx->linewidth[iT * qb], &ctx->linewidth[iT * qb] + qb, linewidth);
    if (F0) std::copy(&ctx->F0[iT * qb * 3], &ctx->F0[iT * qb * 3] + qb * 3, F0);
    if (scalar_mfp) std::copy(&ctx->smfp[iT * qb], &ctx->smfp[iT * qb] + qb, scalar_mfp);
    return 0;
}

int tc3_get_xi(const tc3_ctx *ctx, int iT, int row0, int nrow, double *xi, int *row_qb) {
    if (!ctx || !ctx->have_scatter || iT < 0 || iT >= ctx->nT) return 5;
    if (row_qb)
        for (int r = 0; r < ctx->nrows; ++r) {
            row_qb[2 * r] = ctx->row_q[r];
            row_qb[2 * r + 1] = ctx->row_b[r];
        }
    if (xi && nrow > 0) {
        if (row0 < 0 || row0 + nrow > ctx->nrows) return 1;
        hipSetDevice(ctx->device);
        if (ctx->xi_orb) {
            // rows stored per orbit: expanded to every member of each orbit of the row's q1
            const size_t e0 = (size_t)ctx->xo_row[row0];
            const size_t e1 = (row0 + nrow < ctx->nrows) ? (size_t)ctx->xo_row[row0 + nrow] : ctx->xo_T;
            std::vector<double> rows(e1 - e0);
            const char *src = (const char *)ctx->d_xi + (size_t)ctx->xi_es * ((size_t)iT * ctx->xo_T + e0);
            if (ctx->xi_es == 8) {
                if (hipMemcpy(rows.data(), src, sizeof(double) * rows.size(), hipMemcpyDeviceToHost) != hipSuccess)
                    return 7;
            } else {
                std::vector<float> tmp(rows.size());
                if (hipMemcpy(tmp.data(), src, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost) != hipSuccess)
                    return 7;
                std::copy(tmp.begin(), tmp.end(), rows.begin());
            }
            const int nb = ctx->nb;
            const tc3::OrbitSet &o = ctx->orb;
            for (int r = 0; r < nrow; ++r) {
                const int q = ctx->row_q[row0 + r];
                const double *x = rows.data() + ((size_t)ctx->xo_row[row0 + r] - e0);
                double *out = xi + (size_t)r * ctx->xi_row;
                for (int k = o.q1_off[q]; k < o.q1_off[q + 1]; ++k)
                    for (int m = o.orb_off[k]; m < o.orb_off[k + 1]; ++m)
                        std::copy(x + (size_t)(k - o.q1_off[q]) * nb, x + (size_t)(k - o.q1_off[q] + 1) * nb,
                                  out + (size_t)o.members[m] * nb);
            }
            return 0;
        }
        const size_t n = ctx->xi_row * (size_t)nrow;
        const char *src = (const char *)ctx->d_xi + (size_t)ctx->xi_es * ((size_t)iT * ctx->xi_row * ctx->nrows +
                                                                          (size_t)row0 * ctx->xi_row);
        if (ctx->xi_es == 8) {
            if (hipMemcpy(xi, src, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) return 7;
        } else {
            std::vector<float> tmp(n);
            if (hipMemcpy(tmp.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess) return 7;
            for (size_t i = 0; i < n; ++i) xi[i] = (double)tmp[i];
        }
    }
    return 0;
}

int tc3_kappa(tc3_ctx *ctx, int maxiter, double tol, int classical, double *kappa_sma, double *kappa_iter,
              double *kappa_offdiag, int *niter) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_scatter) return ctx->fail(5, "run tc3_scatter first");
        if ((classical != 0) != ctx->classical) return ctx->fail(5, "classical flag differs from tc3_scatter");
        int r = run_kappa(ctx, maxiter, tol, kappa_offdiag != nullptr);
        if (r) return r;
        const int nT = ctx->nT;
        if (kappa_sma) std::copy(ctx->kappa_sma.begin(), ctx->kappa_sma.end(), kappa_sma);
        if (kappa_iter) std::copy(ctx->kappa_iter.begin(), ctx->kappa_iter.end(), kappa_iter);
        if (kappa_offdiag) std::copy(ctx->kappa_od.begin(), ctx->kappa_od.end(), kappa_offdiag);
        if (niter) std::copy(ctx->niter.begin(), ctx->niter.begin() + nT, niter);
        ctx->have_kappa = true;
        return 0;
    });
}

int tc3_kappa_noalloc(tc3_ctx *ctx, int maxiter, double tol, int classical, double *kappa_sma, double *kappa_iter,
                      int *niter) {
    return guarded(ctx, [&]() -> int {
        if (ctx->device < 0) return ctx->fail(7, "host-only context has no device");
        if (!ctx->have_scatter) return ctx->fail(5, "run tc3_scatter first");
        if ((classical != 0) != ctx->classical) return ctx->fail(5, "classical flag differs from tc3_scatter");
        void *ptr = nullptr;
        int r = bte_scratch_ptr(ctx->device, &ptr);
        if (r) return r;
        r = run_kappa(ctx, maxiter, tol, false, ptr, bte_scratch_bytes());
        if (r) return r;
        const int nT = ctx->nT;
        if (kappa_sma) std::copy(ctx->kappa_sma.begin(), ctx->kappa_sma.end(), kappa_sma);
        if (kappa_iter) std::copy(ctx->kappa_iter.begin(), ctx->kappa_iter.end(), kappa_iter);
        if (niter) std::copy(ctx->niter.begin(), ctx->niter.begin() + nT, niter);
        ctx->have_kappa = true;
        return 0;
    });
}

int tc3_get_kappa_log(const tc3_ctx *ctx, int iT, int maxlen, double *log, int *nlog) {
    if (!ctx || !ctx->have_kappa || iT < 0 || iT >= ctx->nT) return 5;
    const auto &l = ctx->kappa_log[iT];
    const int n = (int)(l.size() / 7);
    const int m = std::min(n, maxlen);
    if (log) std::copy(l.begin(), l.begin() + 7 * m, log);
    if (nlog) *nlog = n;
    return 0;
}

int tc3_get_modes(const tc3_ctx *ctx, int iT, double *Fn, double *kappa_mode) {
    if (!ctx || !ctx->have_kappa || iT < 0 || iT >= ctx->nT) return 5;
    const size_t qb = (size_t)ctx->mesh.nirr * ctx->nb;
    if (Fn) std::copy(&ctx->Fn[iT * qb * 3], &ctx->Fn[iT * qb * 3] + qb * 3, Fn);
    if (kappa_mode) std::copy(&ctx->kappa_mode[iT * qb * 9], &ctx->kappa_mode[iT * qb * 9] + qb * 9, kappa_mode);
    return 0;
}

// dL/dm for a cotangent of the (degeneracy-averaged) linewidths of every temperature.
static int adjoint_from_linewidths(tc3_ctx *ctx, const double *dlw, int precision, double *dL_dfc3) {
    const int nb = ctx->nb, nirr = ctx->mesh.nirr, nT = ctx->nT, nrows = ctx->nrows;
    const size_t qb = (size_t)nirr * nb;
    // The forward averages the raw linewidths over degenerate bands band by band, in place
    // (scattering.f90:264-277); each step is a symmetric averaging map, so the cotangent
    // goes through the same steps in reverse order.
    std::vector<double> g((size_t)nT * nrows, 0.0), buf(qb);
    for (int T = 0; T < nT; ++T) {
        for (size_t i = 0; i < qb; ++i) buf[i] = (ctx->om_irr[i] < kFreqTol) ? 0.0 : dlw[(size_t)T * qb + i];
        for (int q = 0; q < nirr; ++q)
            for (int b = nb - 1; b >= 0; --b) {
                if (ctx->om_irr[(size_t)q * nb + b] < kFreqTol) continue;
                const auto &dm = ctx->degen[q][b];
                double f0 = 0;
                for (int b2 : dm) f0 += buf[(size_t)q * nb + b2];
                f0 /= (double)dm.size();
                for (int b2 : dm) buf[(size_t)q * nb + b2] = f0;
            }
        for (int r = 0; r < nrows; ++r) g[(size_t)T * nrows + r] = buf[(size_t)ctx->row_q[r] * nb + ctx->row_b[r]];
    }
    std::vector<double> grad;
    int r = run_adjoint_gamma(ctx, precision, g.data(), grad);
    if (r) return r;
    // sorted mass-weighted [i][j][k] -> caller's order and layout m[27 t + 9 k + 3 j + i];
    // mwm = m / sqrt(m1 m2 m3) (type_forceconstant_thirdorder.f90:277)
    const auto &ism = ctx->cr.invsqrtmass;
    std::fill(dL_dfc3, dL_dfc3 + 27 * (size_t)ctx->nfc3, 0.0);
    for (int n = 0; n < ctx->nfc3; ++n) {
        const int t = ctx->t_src[n];
        const double w = ism[ctx->t_a1[n]] * ism[ctx->t_a2[n]] * ism[ctx->t_a3[n]];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k)
                    dL_dfc3[27 * (size_t)t + 9 * k + 3 * j + i] = grad[27 * (size_t)n + 9 * i + 3 * j + k] * w;
    }
    return 0;
}

int tc3_adjoint_gamma(tc3_ctx *ctx, const double *dL_dlinewidth, int precision, double *dL_dfc3) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_scatter) return ctx->fail(5, "run tc3_scatter first");
        if (!ctx->thirdorder) return ctx->fail(5, "the three-phonon term is switched off");
        if (!dL_dlinewidth || !dL_dfc3) return ctx->fail(5, "null argument");
        if (precision < 0 || precision > 2) return ctx->fail(5, "unknown precision mode");
        return adjoint_from_linewidths(ctx, dL_dlinewidth, precision, dL_dfc3);
    });
}

int tc3_adjoint_kappa_rta(tc3_ctx *ctx, const double *dL_dkappa, int precision, double *dL_dfc3) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_scatter) return ctx->fail(5, "run tc3_scatter first");
        if (!ctx->thirdorder) return ctx->fail(5, "the three-phonon term is switched off");
        if (!dL_dkappa || !dL_dfc3) return ctx->fail(5, "null argument");
        if (precision < 0 || precision > 2) return ctx->fail(5, "unknown precision mode");
        const size_t qb = (size_t)ctx->mesh.nirr * ctx->nb;
        std::vector<double> dlw(qb * ctx->nT);
        for (int T = 0; T < ctx->nT; ++T) kappa_rta_cotangent(ctx, T, dL_dkappa + 9 * (size_t)T, &dlw[qb * T]);
        return adjoint_from_linewidths(ctx, dlw.data(), precision, dL_dfc3);
    });
}

int tc3_jvp(tc3_ctx *ctx, const double *dfc2, const double *dfc3, double *dlinewidth, double *domega, double *dvel,
            double *dkappa_sma) {
    return guarded(ctx, [&]() -> int {
        if (ctx->device < 0) return ctx->fail(7, "host-only context has no device");
        if (!ctx->have_scatter) return ctx->fail(5, "run tc3_scatter first");
        if (!ctx->thirdorder) return ctx->fail(5, "the three-phonon term is switched off");
        bool iso = false;
        for (double g : ctx->iso_g) iso = iso || (ctx->iso_on && g > 0.0);
        if (iso || ctx->mfp_max > 0.0 || ctx->nfc4 > 0 || !ctx->trip_w.empty() || ctx->integrationtype == 6)
            return ctx->fail(5, "derivatives cover three-phonon scattering on the full grid with integration type 1 "
                                "or 2 (no isotope disorder, boundary or four-phonon scattering)");
        std::vector<double> dlw, dom, dv, dks;
        const int r = run_jvp(ctx, dfc2, dfc3, dlw, dom, dv, dks);
        if (r) return r;
        if (dlinewidth) std::copy(dlw.begin(), dlw.end(), dlinewidth);
        if (domega) std::copy(dom.begin(), dom.end(), domega);
        if (dvel) std::copy(dv.begin(), dv.end(), dvel);
        if (dkappa_sma) std::copy(dks.begin(), dks.end(), dkappa_sma);
        return 0;
    });
}

int tc3_get_psi2(tc3_ctx *ctx, int precision, int n, const int64_t *triplets, double *psi2, int *qpoints) {
    return guarded(ctx, [&]() -> int {
        if (ctx->device < 0) return ctx->fail(7, "host-only context has no device");
        if (!ctx->have_harmonic || !ctx->have_fc3) return ctx->fail(5, "run tc3_harmonic and set fc3 first");
        if (precision < 0 || precision > 2) return ctx->fail(5, "unknown precision mode");
        return run_psi2(ctx, precision, n, triplets, psi2, qpoints);
    });
}

int tc3_get_timings(const tc3_ctx *ctx, double *t, int n) {
    if (!ctx || !t) return 5;
    const_cast<tc3_ctx *>(ctx)->resolve_kstats();
    for (int i = 0; i < n && i < 8; ++i) t[i] = ctx->timings[i];
    return 0;
}

int tc3_get_kernel_stats(const tc3_ctx *ctx, int which, double *total_ms, int *launches, int64_t *batch) {
    if (!ctx || which < 0 || which > 4) return 5;
    const_cast<tc3_ctx *>(ctx)->resolve_kstats();
    if (total_ms) *total_ms = ctx->kstat_ms[which];
    if (launches) *launches = ctx->kstat_n[which];
    if (batch) *batch = ctx->batch;
    return 0;
}

int tc3_get_xi_storage(const tc3_ctx *ctx, int64_t *elems_per_T, int *entry_bytes, int *per_orbit) {
    if (!ctx || !ctx->have_scatter) return 5;
    if (elems_per_T) *elems_per_T = ctx->xi_orb ? (int64_t)ctx->xo_T : (int64_t)(ctx->xi_row * ctx->nrows);
    if (entry_bytes) *entry_bytes = ctx->xi_es;
    if (per_orbit) *per_orbit = ctx->xi_orb ? 1 : 0;
    return 0;
}

int tc3_get_tuning(const tc3_ctx *ctx, int maxn, double *batch, double *ms, int *n) {
    if (!ctx || maxn < 0) return 5;
    const int m = (int)ctx->tune_log.size();
    if (n) *n = m;
    for (int i = 0; i < m && i < maxn; ++i) {
        if (batch) batch[i] = ctx->tune_log[i].batch;
        if (ms) ms[i] = ctx->tune_log[i].ms;
    }
    return 0;
}

int tc3_write_grid_hdf5(const tc3_ctx *ctx, int iT, const char *path) {
    if (!ctx || !ctx->have_kappa || iT < 0 || iT >= ctx->nT || !path) return 5;
    tc3_ctx *mc = const_cast<tc3_ctx *>(ctx);
    const int rf = guarded(mc, [&]() -> int { return ensure_full_host(mc); });
    if (rf) return rf;
    std::string err;
    int r = write_grid_hdf5(ctx, iT, path, err);
    if (r) const_cast<tc3_ctx *>(ctx)->err = err;
    return r;
}

int tc3_set_fc4(tc3_ctx *ctx, int nquart, const int *a1, const int *a2, const int *a3, const int *a4,
                const double *lv2, const double *lv3, const double *lv4, const double *m) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->have_structure) return ctx->fail(5, "set the structure first");
        if (nquart < 0) return ctx->fail(5, "negative quartet count");
        const int na = ctx->cr.na;
        for (int t = 0; t < nquart; ++t)
            if (a1[t] < 0 || a1[t] >= na || a2[t] < 0 || a2[t] >= na || a3[t] < 0 || a3[t] >= na || a4[t] < 0 ||
                a4[t] >= na)
                return ctx->fail(1, "quartet atom index out of range");
        std::vector<int> idx(nquart);
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) {
            return std::make_tuple(a1[x], a2[x], a3[x], a4[x]) < std::make_tuple(a1[y], a2[y], a3[y], a4[y]);
        });
        const size_t nb4 = (size_t)na * na * na * na;
        ctx->nfc4 = nquart;
        ctx->blk4_off.assign(nb4 + 1, 0);
        ctx->f4_lv.resize(9 * (size_t)nquart);
        ctx->f4_mwm.resize(81 * (size_t)nquart);
        const auto &ism = ctx->cr.invsqrtmass;
        for (int n = 0; n < nquart; ++n) {
            const int t = idx[n];
            for (int i = 0; i < 3; ++i) {
                ctx->f4_lv[9 * (size_t)n + i] = lv2[3 * t + i];
                ctx->f4_lv[9 * (size_t)n + 3 + i] = lv3[3 * t + i];
                ctx->f4_lv[9 * (size_t)n + 6 + i] = lv4[3 * t + i];
            }
            // type_forceconstant_fourthorder.f90:296
            const double w = ism[a1[t]] * ism[a2[t]] * ism[a3[t]] * ism[a4[t]];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k)
                        for (int l = 0; l < 3; ++l)
                            ctx->f4_mwm[81 * (size_t)n + 27 * i + 9 * j + 3 * k + l] =
                                m[81 * (size_t)t + 27 * l + 9 * k + 3 * j + i] * w;
            ctx->blk4_off[(((size_t)a1[t] * na + a2[t]) * na + a3[t]) * na + a4[t] + 1]++;
        }
        for (size_t b = 0; b < nb4; ++b) ctx->blk4_off[b + 1] += ctx->blk4_off[b];
        ++ctx->gen_model;
        return 0;
    });
}

int tc3_write_cumulative_hdf5(const tc3_ctx *ctx, int iT, const char *path, int mfppts, int freqpts,
                              int dosintegrationtype, double sigma, double dossigma, const char *unit,
                              const char *atomic_species) {
    if (!ctx || !ctx->have_kappa || iT < 0 || iT >= ctx->nT || !path || mfppts < 2 || freqpts < 2 ||
        dosintegrationtype < 1 || dosintegrationtype > 3 || !(sigma > 0) || !(dossigma > 0) || ctx->Fn.empty())
        return 5;
    tc3_ctx *mc = const_cast<tc3_ctx *>(ctx);
    const int rf = guarded(mc, [&]() -> int { return ensure_full_host(mc); });
    if (rf) return rf;
    CumulativeOpts o;
    o.mfppts = mfppts;
    o.freqpts = freqpts;
    o.dosint = dosintegrationtype;
    o.sigma = sigma;
    o.dossigma = dossigma;
    o.unit = unit ? unit : "thz";
    std::string err;
    const int r = guarded(mc, [&]() -> int {
        return write_cumulative_hdf5(ctx, iT, path, o, atomic_species ? atomic_species : "", err);
    });
    if (r && !err.empty()) mc->err = err;
    return r;
}

}  // extern "C"
