// The tc3 context: host-side model data, mesh/triplet tables and device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/tc3.h"
#include "tc3_device.hpp"
#include "tc3_internal.hpp"

struct tc3_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    bool have_structure = false, have_fc2 = false, have_fc3 = false, have_mesh = false,
         have_harmonic = false, have_scatter = false, have_kappa = false;

    tc3::Crystal cr;
    std::vector<tc3::SymOp> ops;
    int nb = 0;

    // second order, sorted by (a1, a2); blk2_off[(a1*na + a2)] .. [+1]
    std::vector<int> p_a1, p_a2, blk2_off, p_src;  // p_src: the caller's index of each sorted pair
    std::vector<double> p_lv, p_m;  // 3 / 9 per pair (m row-major [i][j])

    // third order, sorted by (a1, a2, a3); blk3_off over na^3 blocks; a1_off over atoms
    int nfc3 = 0;
    std::vector<int> t_a1, t_a2, t_a3, blk3_off, a1_off;
    std::vector<double> t_lv2, t_lv3, t_mwm;  // 3 / 3 / 27 per triplet (mwm [i][j][k])

    // fourth order (four-phonon scattering when nfc4 > 0), sorted by (a1, a2, a3, a4);
    // blk4_off over na^4 blocks; lattice vectors 9 per quartet (lv2, lv3, lv4), mwm 81 [i][j][k][l]
    int nfc4 = 0;
    std::vector<int> blk4_off;
    std::vector<double> f4_lv, f4_mwm;
    // irreducible quartets of every irreducible q1 (rebuilt when gen_model changes):
    // 5 ints each (q1, q2, q3, q4, multiplicity class), images 3 ints each from img_off
    std::vector<int> quart, quart_img_off, quart_img;
    long quart_gen = -1;
    // the quartets the Monte-Carlo rows visit and their draw counts [quartet][b1] (rebuilt
    // when the model or the draws change)
    std::vector<int> mquart, mquart_img_off, mquart_img;
    std::vector<float> mquart_tw;
    long mquart_gen = -1;
    std::vector<int> t_src;                   // caller's index of each sorted triplet

    tc3::Mesh mesh;
    tc3::TripletSet trip;        // triplets a pass visits (the full list, or the Monte-Carlo subset)
    tc3::TripletSet trip_full;   // every irreducible triplet of the mesh
    // Monte-Carlo q2 grids (-qg3ph, --seed): draw counts per (triplet, b1) of the subset
    int mc_dims[3] = {0, 0, 0};
    int mc_seed = 0;
    bool mc_on = false;
    long mc_gen = -1, mc_npoints = 0;
    std::vector<float> trip_w;
    // Monte-Carlo grids of the four-phonon term (-qg4ph): per row q1 * nb + b1 the two drawn
    // grids, 2 * mc4_npoints full-grid indices
    int mc4_dims[3] = {0, 0, 0};
    bool mc4_on = false;
    long mc4_npoints = 0;
    std::vector<long> g4;
    tc3::OrbitSet orb;

    // harmonic results (host copies); egv stored [q][component][mode]
    std::vector<double> om_irr, vel_irr, om_full, vel_full;
    std::vector<tc3::cd> egv_irr, egv_full;
    bool full_host_valid = false;             // egv_irr and om/vel/egv_full hold the device results
    int harm_dims[3] = {-1, -1, -1};          // nb, nirr, nfull of the device harmonic arrays
    std::vector<std::vector<std::vector<int>>> degen;  // [q_irr][band] -> degenerate bands
    double omega_min = 0;

    // rows (q_irr, band) with omega >= freqtol, q-major
    int nrows = 0;
    std::vector<int> row_q, row_b, row_of;  // row_of[q*nb + b] = row or -1

    // scattering results
    int nT = 0;
    std::vector<double> temps;
    int precision = 0;
    bool classical = false;
    int integrationtype = 2;
    double sigma = 1.0;                       // fixed smearing (THz) of integrationtype 1
    std::vector<double> be, sigsq;            // [T][q][b], [q][b]
    // be is formed on the device by the scattering pass; the host copy (adjoint, forward-mode
    // derivatives) on first use after it (host_occupations)
    bool be_host = false;
    std::vector<double> iso_g;                // mass-disorder parameter per atom
    bool iso_on = false;                      // isotope scattering requested
    double mfp_max = -1.0;                    // boundary scattering length (Bohr), off if <= 0
    bool thirdorder = true;                   // three-phonon term on (--nothirdorder switches it off)
    double thresh6 = 0.0;                     // baseline width of integration type 6 (scattering.f90:165-167)
    std::vector<double> linewidth, qs, F0, mfp, smfp;  // [T][q][b](x3)

    // kappa results
    std::vector<double> kappa_sma, kappa_iter, kappa_od;  // [T][9]
    std::vector<double> Fn, kappa_mode;                   // [T][q][b][3], [T][q][b][9]
    std::vector<std::vector<double>> kappa_log;           // [T][iter*7]
    std::vector<int> niter;

    double timings[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // per-kernel-class device time of the last calls (HIP events): 0 Phi+U+W kernel,
    // 1 Psi + smeared-delta epilogue kernel, 2 scattering-matrix symmetrisation, 3 BTE Xi.F
    double kstat_ms[5] = {0, 0, 0, 0, 0};
    int kstat_n[5] = {0, 0, 0, 0, 0};
    // events of the last scattering pass, read into kstat_ms / timings[5] when the statistics
    // are asked for (querying them inside the pass costs host time after every pass)
    std::vector<hipEvent_t> ks_kev;            // 3 per launch pair: before, between, after
    hipEvent_t ks_ev0 = nullptr, ks_ev1 = nullptr, ks_evs0 = nullptr, ks_evs1 = nullptr;
    bool ks_pending = false, ks_comp = false;
    void resolve_kstats() {
        if (!ks_pending) return;
        ks_pending = false;
        float ms = 0;
        hipEventElapsedTime(&ms, ks_ev0, ks_ev1);
        timings[5] = ms * 1e-3;
        for (size_t b = 0; b + 2 < ks_kev.size(); b += 3) {
            float m1 = 0, m2 = 0;
            hipEventElapsedTime(&m1, ks_kev[b], ks_kev[b + 1]);
            hipEventElapsedTime(&m2, ks_kev[b + 1], ks_kev[b + 2]);
            kstat_ms[0] += m1; kstat_n[0] += 1;
            kstat_ms[1] += m2; kstat_n[1] += 1;
        }
        float m3 = 0;
        hipEventElapsedTime(&m3, ks_evs0, ks_evs1);
        kstat_ms[2] += m3; kstat_n[2] += 1;
        if (ks_comp) { kstat_ms[4] += m3; kstat_n[4] += 1; }
    }
    long batch = 0;

    // device buffers
    double *d_om_irr = nullptr, *d_om_full = nullptr, *d_vel_irr = nullptr, *d_vel_full = nullptr;
    tc3::cd *d_egv_irr = nullptr, *d_egv_full = nullptr;
    void *d_xi = nullptr;       // [T][row][q*nb + b]: FP64 in FP64 mode, FP32 in the mixed modes
    int xi_es = 8;              // bytes per entry of d_xi (8 or 4)
    size_t xi_row = 0;          // nfull*nb
    size_t xi_bytes = 0;        // allocated size of d_xi
    long xi_gen = -1;           // model generation d_xi was zeroed for
    int xi_es_zeroed = 0;       // entry size d_xi was zeroed for
    int *d_irr_of_full = nullptr;

    // Device buffers kept across calls, by name (grown on demand, freed with the context),
    // and the model generation: bumped by every structure / IFC / mesh change so that
    // tables uploaded for an older model are uploaded again.
    struct DevBuf { void *p = nullptr; size_t bytes = 0; };
    std::map<std::string, DevBuf> pool;
    long gen_model = 0;
    long scatter_tables_gen = -1;
    long adjoint_tables_gen = -1;
    // autotuned triplets per launch of the MFMA contraction, by shape, and the probes timed
    std::map<std::string, long> tuned_batch;
    std::map<std::string, int> tuned_bb;   // b1 blocks per k_psi workgroup, by shape
    // owner of every scattering-matrix entry of the rows of each irreducible q1 (k_xi_asm)
    std::vector<int> xi_src;
    bool xi_src_ok = false;
    long xi_src_gen = -1;
    // Scattering matrix stored per orbit (xi_orb): once symmetrised, a row is constant over
    // every orbit of the little group of its q1, so row r holds one value per orbit and band,
    // o nb + b for the o-th orbit of q1, at xo_row[r] (elements per temperature: xo_T).  The
    // irreducible points with the same orbit list share a group whose F orbit sums (3 planes
    // of norb nb, at xo_goff[g] of the xo_F elements per temperature) the BTE product reads;
    // group -1 where every orbit is the single point of its index (the folded-out F itself).
    // Row groups of the product, at most 16 rows of one q1: (first row, rows, group, norb nb).
    // Orbit entries of the sums: (first, end of its members in xo_mem, offset of the orbit's
    // values in the sums, norb nb of its group).
    long harm_tables_gen = -1;  // model generation the harmonic step's tables were uploaded for
    bool xi_orb = false;
    long xo_gen = -1;
    std::vector<long long> xo_row, xo_goff;
    std::vector<int4> xo_rg, xo_ent;
    std::vector<int> xo_mem;
    size_t xo_T = 0, xo_F = 0;
    // IFC block tables of the fused 48-band kernel (per model)
    std::vector<int4> fused_blocks, fused_pairs;
    std::vector<int> fused_pa1;
    int fused_maxe = 0, fused_maxb = 0;
    long fused_gen = -1;
    // orbit chunks of the scattering-matrix kernels (per model) and pooled HIP events
    std::vector<int> xchk_off, xchk;
    int xcapm = 0;
    long xchk_gen = -1;
    std::vector<hipEvent_t> ev_pool;
    // pinned host staging buffer of the per-call transfers (grown as needed)
    void *pin = nullptr;
    size_t pin_bytes = 0;
    hipError_t pinned(size_t bytes) {
        if (bytes <= pin_bytes) return hipSuccess;
        if (pin) hipHostFree(pin);
        pin = nullptr;
        pin_bytes = 0;
        const hipError_t r = hipHostMalloc(&pin, bytes, hipHostMallocDefault);
        if (r == hipSuccess) pin_bytes = bytes;
        return r;
    }
    struct TuneProbe { std::string key; double batch, ms; };
    std::vector<TuneProbe> tune_log;
    // instantiated BTE iteration graphs and the arguments they were captured with
    hipGraphExec_t bte_exec = nullptr, bte_exec2 = nullptr;   // first chunk, later chunks
    std::vector<long> bte_key;
    // instantiated graph of the scattering pass's contraction launches and the arguments it
    // was captured with (sc_key_failed: the arguments a capture failed for, not retried)
    hipGraphExec_t sc_exec = nullptr;
    std::vector<long> sc_key, sc_key_failed;
    hipError_t pool_get(const std::string &name, size_t bytes, void **p, bool *fresh = nullptr) {
        DevBuf &b = pool[name];
        if (fresh) *fresh = false;
        if (b.bytes < bytes || b.p == nullptr) {
            if (b.p) (void)hipFree(b.p);
            b.p = nullptr;
            b.bytes = 0;
            const hipError_t e = hipMalloc(&b.p, bytes < 16 ? 16 : bytes);
            if (e != hipSuccess) { b.p = nullptr; return e; }
            b.bytes = bytes;
            if (fresh) *fresh = true;
        }
        *p = b.p;
        return hipSuccess;
    }

    ~tc3_ctx();
    int fail(int code, const std::string &msg) {
        err = msg;
        return code;
    }
};

namespace tc3 {

int device_error(hipError_t e, const char *what, const char *file, int line);
extern thread_local std::string g_device_error;

// tc3_harmonic.hip
int run_harmonic(tc3_ctx *c);
int ensure_full_host(tc3_ctx *c);
// device derivatives of the harmonic properties along a change of the second-order force
// constants (pool buffers of the context)
struct HarmonicJvp {
    cd *de_irr = nullptr, *de_full = nullptr;       // [q][component][mode]
    double *dom_irr = nullptr, *dom_full = nullptr; // [q][mode]
    double *dv_irr = nullptr, *dv_full = nullptr;   // [q][mode][3]
};
int run_harmonic_jvp(tc3_ctx *c, const std::vector<double> &dpm, HarmonicJvp &out);
// tc3_scatter.hip
int run_scatter(tc3_ctx *c, int precision, int integrationtype, double sigma);
int run_psi2(tc3_ctx *c, int precision, int n, const int64_t *trips, double *out, int *qpts);
// tc3_kappa.hip
int run_kappa(tc3_ctx *c, int maxiter, double tol, bool offdiag, void *scratch = nullptr, size_t scratch_bytes = 0);
size_t bte_scratch_bytes();
int bte_scratch_ptr(int device, void **ptr);
void kappa_rta_cotangent(const tc3_ctx *c, int iT, const double *G, double *dlw);
void symmetrize_kappa_tensor(const tc3_ctx *c, double *k);   // symmetrize_kappa, kappa.f90:308-324
// tc3_jvp.hip
int run_jvp(tc3_ctx *c, const double *dfc2, const double *dfc3, std::vector<double> &dlw, std::vector<double> &dom,
            std::vector<double> &dvel, std::vector<double> &dks);
// tc3_adjoint.hip
int run_adjoint_gamma(tc3_ctx *c, int precision, const double *g, std::vector<double> &grad);
// tc3_hdf5.cpp
int write_grid_hdf5(const tc3_ctx *c, int iT, const char *path, std::string &err);
// tc3_api.cpp: the occupations of the last pass on the host (c->be), formed on first use
void host_occupations(tc3_ctx *c);

}  // namespace tc3
