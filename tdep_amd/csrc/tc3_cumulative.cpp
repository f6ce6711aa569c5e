// Cumulative, boundary and spectral thermal conductivity and the phonon density of states of
// outfile.cumulative_thermal_conductivity.hdf5, for one temperature after tc3_kappa:
//   get_cumulative_kappa   cumulative_kappa.f90:58-414
//   get_boundary_kappa     cumulative_kappa.f90:416-508
//   get_spectral_kappa     cumulative_kappa.f90:510-561, type_phonon_dos_integrations.f90:428-809
//   lo_phonon_dos%generate type_phonon_dos.f90:110-297 (tetrahedron :812-924, Gaussian
//                          :927-1021)
//   irreducible tetrahedra type_qpointmesh_gridgeneration.f90:700-1260, 1342-1470
//   LV weights             type_qpointmesh_integrationweights.f90:148-262
//   angular momentum       type_phonon_dispersions.f90:868-971,
//                          type_phonon_dos_integrations.f90:8-425
// Host FP64: the work is O(tetrahedra x modes x frequency points) once per output file.  Two
// values the reference reads before it sets them are fixed here: the boundary rate of a mode
// without group velocity before any mode had one (cumulative_kappa.f90:464-467) starts at
// zero, and the full-grid per-mode kappa of modes below lo_freqtol (:538-553) is zero.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "tc3_ctx.hpp"
#include "tc3_cumulative.hpp"

namespace tc3 {

namespace {

constexpr double kKappaSiToAu = 1.0 / kKappaAuToSi;

// lo_gauss (gottochblandat_calculus.f90:268-287)
double gauss(double x, double mu, double sigma) {
    if (std::fabs(x - mu) > 8 * sigma) return 0.0;
    const double inv = 1.0 / sigma;
    return std::exp(-(x - mu) * (x - mu) * 0.5 * inv * inv) * 0.398942280401433 * inv;
}

// lo_linspace / lo_logspace (gottochblandat_calculus.f90:188-249)
std::vector<double> linspace(double a, double b, int n) {
    std::vector<double> x(n);
    if (n == 1) {
        x[0] = (a + b) * 0.5;
        return x;
    }
    const double f1 = b - a;
    for (int i = 0; i < n; ++i) x[i] = a + ((i + 1) * 1.0 - 1.0) / (n * 1.0 - 1.0) * f1;
    x[0] = a;
    x[n - 1] = b;
    return x;
}

std::vector<double> logspace(double a, double b, int n) {
    if (n == 1) return std::vector<double>(1, (a + b) * 0.5);
    const double shift = (a < kTol) ? 1.0 - a : 0.0;
    std::vector<double> d = linspace(std::log(a + shift), std::log(b + shift), n);
    for (double &v : d) v = std::exp(v) - shift;
    return d;
}

// lo_trapezoid_integration (gottochblandat_calculus.f90:252-265)
double trapezoid(const std::vector<double> &x, const double *y, size_t stride = 1) {
    double f = 0.0;
    for (size_t i = 0; i + 1 < x.size(); ++i) f += (y[i * stride] + y[(i + 1) * stride]) * 0.5 * (x[i + 1] - x[i]);
    return f;
}

// lo_linear_interpolation (gottochblandat_calculus.f90:75-151) on a sorted x; the first i
// with x(i) - xi > thres is found by bisection (the predicate is monotone on a sorted x)
double linear_interpolation(const double *x, const double *y, int n, double xi, double thres = kSqTol) {
    if (std::fabs(xi - x[0]) < thres) return y[0];
    if (std::fabs(xi - x[n - 1]) < thres) return y[n - 1];
    const double mn = *std::min_element(x, x + n), mx = *std::max_element(x, x + n);
    if (mn - xi > 2 * thres || std::fabs(mx) - xi < -2 * thres) return 0.0;
    int lo = 2, hi = n + 1;   // 1-based
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (x[mid - 1] - xi > thres) hi = mid;
        else lo = mid + 1;
    }
    int ii, jj;
    if (lo <= n) { ii = lo; jj = lo - 1; }
    else { ii = n; jj = 0; }
    if (ii < 2) { ii = 2; jj = 1; }
    if (ii >= n) { ii = n; jj = n - 1; }
    const double A = (x[ii - 1] - xi) / (x[ii - 1] - x[jj - 1]);
    return A * y[jj - 1] + (1.0 - A) * y[ii - 1];
}

// sym%degeneracy / sym%degenerate_atom (type_symmetryoperation.f90:651-676)
void atom_classes(const tc3_ctx *c, std::vector<int> &deg, std::vector<std::vector<int>> &mem) {
    const int na = c->cr.na;
    std::vector<char> di((size_t)na * na, 0);
    for (const SymOp &o : c->ops)
        for (int i = 0; i < na; ++i) {
            const int j = o.fmap[i];
            di[(size_t)i * na + j] = di[(size_t)j * na + i] = 1;
        }
    deg.assign(na, 0);
    mem.assign(na, {});
    for (int i = 0; i < na; ++i)
        for (int j = 0; j < na; ++j)
            if (di[(size_t)j * na + i]) {
                ++deg[i];
                mem[i].push_back(j);
            }
}

// dr%default_smearing, omega_max, omega_min (type_phonon_dispersions_generation.f90:195-233)
void default_smearing(const tc3_ctx *c, std::vector<double> &ds, double &omax, double &omin) {
    const int nb = c->nb, nirr = c->mesh.nirr;
    ds.assign(nb, 0.0);
    std::vector<double> x(nirr);
    for (int b = 0; b < nb; ++b) {
        for (int q = 0; q < nirr; ++q) x[q] = c->om_irr[(size_t)q * nb + b];
        std::sort(x.begin(), x.end());
        double f0 = 0.0;
        for (int q = 0; q + 1 < nirr; ++q) f0 = std::max(f0, std::fabs(x[q + 1] - x[q]));
        ds[b] = f0;
    }
    const double f0 = *std::max_element(ds.begin(), ds.end()) / 5.0;
    for (double &v : ds) v = std::max(v, f0);
    omax = -std::numeric_limits<double>::max();
    omin = std::numeric_limits<double>::infinity();
    for (double w : c->om_irr) {
        omax = std::max(omax, w);
        if (std::fabs(w) > kFreqTol) omin = std::min(omin, w);
    }
}

double site_projection(const cd *egv, int nb, int atom, int mode) {
    double s = 0.0;
    for (int i = 3 * atom; i < 3 * atom + 3; ++i) {
        const cd x = egv[(size_t)i * nb + mode];
        s += x.x * x.x + x.y * x.y;
    }
    return s;
}

double cv_of(double T, double om, bool classical) {
    if (classical) return kKbHartree;
    // lo_harmonic_oscillator_cv (gottochblandat_physics.f90:331-367)
    if (om < kFreqTol || T < kTemperatureTol) return 0.0;
    const double x = om / T / kKbHartree;
    if (x > 200.0) return 0.0;
    if (x > 21.0) return kKbHartree * x * x * std::exp(-x);
    if (x < 2e-2) return kKbHartree * (1.0 - x * x / 12.0 + x * x * x * x / 240.0);
    const double ex = std::exp(x);
    return kKbHartree * x * x * ex / ((ex - 1) * (ex - 1));
}

void rotate(const tc3_ctx *c, int k, const double *v, double *out) {
    const double *m = c->ops[std::abs(k) - 1].m;
    const double s = (k > 0) ? 1.0 : -1.0;
    for (int i = 0; i < 3; ++i) out[i] = s * (m[3 * i] * v[0] + m[3 * i + 1] * v[1] + m[3 * i + 2] * v[2]);
}

// lo_clean_fractional_coordinates (gottochblandat.f90:1261-1283), default tolerance
double clean_frac(double y) {
    for (;;) {
        if (std::fabs(y) < kSqTol) y = 0.0;
        if (std::fabs(y - 1.0) < kSqTol) y = 0.0;
        if (y > 1.0) y = y - 1.0;
        if (y < 0.0) y = y + 1.0;
        if (y > -kSqTol && y < 1.0) return y;
    }
}

}  // namespace

// lo_LV_tetrahedron_weights (type_qpointmesh_integrationweights.f90:148-262): Lambin and
// Vigneron's (PRB 29, 3430 (1984)) corner weights of a linear tetrahedron for the density of
// states at z, degenerate tetrahedra (spread below tol) as four Gaussians of width sigma.
void lv_tetrahedron_weights(const double *ein, double z, double tol, double sigma, double *w) {
    int ind[4] = {0, 1, 2, 3};
    std::stable_sort(ind, ind + 4, [&](int a, int b) { return ein[a] < ein[b]; });
    double ei[4];
    for (int i = 0; i < 4; ++i) ei[i] = ein[ind[i]];
    for (int i = 0; i < 4; ++i) w[i] = 0.0;
    if (z <= ei[0] || z >= ei[3]) return;
    const double EZ1 = z - ei[0], EZ2 = z - ei[1], EZ3 = z - ei[2], EZ4 = z - ei[3];
    double wt[4];
    if (ei[3] - ei[0] < tol) {
        wt[0] = gauss(EZ1, 0.0, sigma) * 0.25;
        wt[1] = gauss(EZ2, 0.0, sigma) * 0.25;
        wt[2] = gauss(EZ3, 0.0, sigma) * 0.25;
        wt[3] = gauss(EZ4, 0.0, sigma) * 0.25;
    } else if (z < ei[1]) {
        const double a = 1.0 / (ei[1] - ei[0]), b = 1.0 / (ei[2] - ei[0]), c = 1.0 / (ei[3] - ei[0]);
        wt[0] = a * b * c * (EZ1 * EZ1) * (-a * EZ2 - b * EZ3 - c * EZ4);
        wt[1] = (a * a) * b * c * (EZ1 * EZ1 * EZ1);
        wt[2] = a * (b * b) * c * (EZ1 * EZ1 * EZ1);
        wt[3] = a * b * (c * c) * (EZ1 * EZ1 * EZ1);
    } else if (z < ei[2]) {
        const double b = 1.0 / (ei[2] - ei[0]), c = 1.0 / (ei[3] - ei[0]);
        const double d = 1.0 / (ei[2] - ei[1]), e = 1.0 / (ei[3] - ei[1]);
        const double ff0 = -(b * b) * EZ3, ff2 = -(c * c) * EZ4;
        const double gg0 = -(d * d) * EZ3, gg2 = -(e * e) * EZ4;
        const double hh0 = (d * d) * EZ2, hh2 = (b * b) * EZ1;
        const double ii0 = (e * e) * EZ2, ii2 = (c * c) * EZ1;
        const double ff1 = -c * d * EZ1 * EZ3 - d * e * EZ2 * EZ3 - c * e * EZ1 * EZ4;
        const double ff3 = -b * d * EZ1 * EZ3 - b * e * EZ1 * EZ4 - d * e * EZ2 * EZ4;
        const double gg1 = -b * c * EZ1 * EZ3 - b * e * EZ2 * EZ3 - c * e * EZ2 * EZ4;
        const double gg3 = -b * d * EZ2 * EZ3 - b * c * EZ1 * EZ4 - c * d * EZ2 * EZ4;
        const double hh1 = -b * c * EZ1 * EZ3 - b * e * EZ2 * EZ3 - c * e * EZ2 * EZ4;
        const double hh3 = -c * d * EZ1 * EZ3 - d * e * EZ2 * EZ3 - c * e * EZ1 * EZ4;
        const double ii1 = -b * d * EZ2 * EZ3 - b * c * EZ1 * EZ4 - c * d * EZ2 * EZ4;
        const double ii3 = -b * d * EZ1 * EZ3 - b * e * EZ1 * EZ4 - d * e * EZ2 * EZ4;
        wt[0] = 0.5 * (ff0 * ff1 + ff2 * ff3);
        wt[1] = 0.5 * (gg0 * gg1 + gg2 * gg3);
        wt[2] = 0.5 * (hh0 * hh1 + hh2 * hh3);
        wt[3] = 0.5 * (ii0 * ii1 + ii2 * ii3);
    } else {
        const double c = 1.0 / (ei[3] - ei[0]), e = 1.0 / (ei[3] - ei[1]), f = 1.0 / (ei[3] - ei[2]);
        const double e3 = EZ4 * EZ4 * EZ4;
        wt[0] = -((c * c) * e * e3 * f);
        wt[1] = -(c * (e * e) * e3 * f);
        wt[2] = -(c * e * e3 * (f * f));
        wt[3] = c * e * (EZ4 * EZ4) * f * (c * EZ1 + e * EZ2 + EZ3 * f);
    }
    for (int i = 0; i < 4; ++i) w[ind[i]] = wt[i];
}

// The six tetrahedra of a box (neighbour_corners_from_center and chop_box_into_tetrahedrons,
// type_qpointmesh_gridgeneration.f90:1342-1470): corners x slowest, z fastest, offsets 0/1
// from the grid point below the centre; the tetrahedra share the shortest main diagonal
// (the first of equal lengths in corner order, with the reference's arithmetic) and take
// consecutive corners around it.
void chop_box(const tc3_ctx *c, const int *dims, const double *center, int off[8][3], int tets[6][4]) {
    double box[8][3];
    for (int i = 0, l = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k, ++l) {
                const int o[3] = {i, j, k};
                for (int d = 0; d < 3; ++d) {
                    const double h = 0.5 / dims[d];
                    box[l][d] = center[d] + (o[d] ? h : -h);
                    off[l][d] = o[d];
                }
            }
    const double *B = c->cr.rec;   // columns are the reciprocal vectors
    double r[8][3], com[3] = {0, 0, 0};
    for (int i = 0; i < 8; ++i) {
        const double v1[3] = {box[i][0] - box[0][0], box[i][1] - box[0][1], box[i][2] - box[0][2]};
        for (int x = 0; x < 3; ++x) r[i][x] = B[3 * x] * v1[0] + B[3 * x + 1] * v1[1] + B[3 * x + 2] * v1[2];
    }
    for (int i = 0; i < 8; ++i)
        for (int x = 0; x < 3; ++x) com[x] = com[x] + r[i][x] / 8.0;
    double f2 = std::numeric_limits<double>::infinity();
    int d0 = 0, d1 = 7;
    for (int i = 0; i < 8; ++i)
        for (int j = i + 1; j < 8; ++j) {
            const double v0[3] = {r[i][0] - com[0], r[i][1] - com[1], r[i][2] - com[2]};
            const double v1[3] = {r[j][0] - com[0], r[j][1] - com[1], r[j][2] - com[2]};
            const double d01 = v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2];
            const double d00 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
            const double f0 = std::fabs(d01 + d00) + std::fabs(d01 + d00);
            if (f0 < 1e-10) {
                const double dv[3] = {r[j][0] - r[i][0], r[j][1] - r[i][1], r[j][2] - r[i][2]};
                const double f1 = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
                if (f1 < f2) {
                    f2 = f1;
                    d0 = i;
                    d1 = j;
                }
            }
        }
    int other[6], no = 0;
    for (int i = 0; i < 8; ++i)
        if (i != d0 && i != d1) other[no++] = i;
    double n[3] = {r[d0][0] - r[d1][0], r[d0][1] - r[d1][1], r[d0][2] - r[d1][2]};
    const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (double &x : n) x /= nn;
    double u[3] = {1, 0, 0};
    if (std::fabs(n[0]) >= 0.9) { u[0] = 0; u[1] = 1; }
    const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    double e1[3] = {u[0] - un * n[0], u[1] - un * n[1], u[2] - un * n[2]};
    const double e1n = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (double &x : e1) x /= e1n;
    const double e2[3] = {n[1] * e1[2] - n[2] * e1[1], n[2] * e1[0] - n[0] * e1[2], n[0] * e1[1] - n[1] * e1[0]};
    double ang[6];
    for (int i = 0; i < 6; ++i) {
        const double v[3] = {r[other[i]][0] - com[0], r[other[i]][1] - com[1], r[other[i]][2] - com[2]};
        ang[i] = std::atan2(v[0] * e2[0] + v[1] * e2[1] + v[2] * e2[2], v[0] * e1[0] + v[1] * e1[1] + v[2] * e1[2]);
    }
    int order[6] = {0, 1, 2, 3, 4, 5};
    std::stable_sort(order, order + 6, [&](int a, int b) { return ang[a] < ang[b]; });
    for (int i = 0; i < 6; ++i) {
        tets[i][0] = d0;
        tets[i][1] = d1;
        tets[i][2] = other[order[i]];
        tets[i][3] = other[order[(i + 1) % 6]];
    }
}

// Irreducible tetrahedra of the FFT mesh (type_qpointmesh_gridgeneration.f90:700-1260): the
// boxes above the grid points reduced by the point group and time reversal, the lowest box
// of each class as prototype, its six tetrahedra weighted by the size of the class.
void irreducible_tetrahedra(const tc3_ctx *c, std::vector<int> &corners, std::vector<double> &weights) {
    const int *dims = c->mesh.dims;
    const int n1 = dims[0], n2 = dims[1], n3 = dims[2], np = n1 * n2 * n3;
    const int nsym = (int)c->ops.size();
    const bool tr = c->mesh.timereversal;
    const int ncol = tr ? 2 * nsym : nsym;
    std::vector<double> ctr(3 * (size_t)np);
    for (int i = 0, l = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j)
            for (int k = 0; k < n3; ++k, ++l) {
                const int g[3] = {i, j, k};
                for (int d = 0; d < 3; ++d) {
                    double v = (g[d] + 1 - 0.5) / (double)dims[d];
                    for (int rep = 0; rep < 2; ++rep) v = clean_frac(chop_scalar(v, 1e-11));
                    ctr[3 * (size_t)l + d] = v;
                }
            }
    std::vector<int> conn((size_t)np * ncol, 0);
    for (int ip = 0; ip < np; ++ip) {
        const double *v0 = &ctr[3 * (size_t)ip];
        for (int t = 0; t < (tr ? 2 : 1); ++t)
            for (int o = 0; o < nsym; ++o) {
                const double *R = c->ops[o].rfm;
                const double s = t ? -1.0 : 1.0;
                double w0[3], err = 0.0;
                int g[3];
                for (int d = 0; d < 3; ++d) {
                    const double v1 = s * (R[3 * d] * v0[0] + R[3 * d + 1] * v0[1] + R[3 * d + 2] * v0[2]);
                    w0[d] = (1.0 + 2.0 * dims[d] * v1) * 0.5;
                    const double rw = std::nearbyint(w0[d]);
                    err += std::fabs(w0[d] - rw);
                    g[d] = (((int)rw - 1) % dims[d] + dims[d]) % dims[d];
                }
                if (err > kSqTol) continue;
                conn[(size_t)ip * ncol + o + (t ? nsym : 0)] = (g[0] * n2 + g[1]) * n3 + g[2] + 1;
            }
    }
    std::vector<char> keep(np, 1);
    for (int ip = 0; ip < np; ++ip)
        for (int x = 0; x < ncol; ++x) {
            const int j = std::abs(conn[(size_t)ip * ncol + x]) - 1;
            if (j > ip) keep[j] = 0;
        }
    corners.clear();
    weights.clear();
    for (int ip = 0; ip < np; ++ip) {
        if (!keep[ip]) continue;
        std::vector<int> members;
        for (int x = 0; x < ncol; ++x) {
            const int v = conn[(size_t)ip * ncol + x];
            if (v != 0 && std::find(members.begin(), members.end(), v) == members.end()) members.push_back(v);
        }
        int off[8][3], tets[6][4];
        chop_box(c, dims, &ctr[3 * (size_t)ip], off, tets);
        const int gi = ip / (n2 * n3), gj = (ip / n3) % n2, gk = ip % n3;
        int cidx[8];
        for (int l = 0; l < 8; ++l)
            cidx[l] = (((gi + off[l][0]) % n1) * n2 + (gj + off[l][1]) % n2) * n3 + (gk + off[l][2]) % n3;
        for (int t = 0; t < 6; ++t) {
            for (int x = 0; x < 4; ++x) corners.push_back(cidx[tets[t][x]]);
            weights.push_back((double)members.size());
        }
    }
}

// lo_phonon_dos%generate (type_phonon_dos.f90:110-297)
void phonon_dos(const tc3_ctx *c, const CumulativeOpts &o, const std::vector<int> &tc, const std::vector<double> &tw,
                PhononDos &d) {
    const int na = c->cr.na, nb = c->nb, n = o.freqpts, nirr = c->mesh.nirr;
    default_smearing(c, d.ds, d.omax, d.omin);
    double dosmin = *std::min_element(c->om_irr.begin(), c->om_irr.end());
    if (dosmin > -kFreqTol) dosmin = 0.0;
    const double dosmax = d.omax + 4.01 * *std::max_element(d.ds.begin(), d.ds.end());
    d.omega = linspace(dosmin, dosmax, n);
    std::vector<double> pm((size_t)n * nb, 0.0), ps((size_t)n * na, 0.0);
    const double tmax = *std::max_element(d.omega.begin(), d.omega.end());
    const double tmin = *std::min_element(d.omega.begin(), d.omega.end());
    const double invf = n / (tmax - tmin);
    const cd *E = c->egv_irr.data();
    std::vector<double> proj(na);
    if (o.dosint == 3) {
        const double de = (d.omega[1] - d.omega[0]) * 0.5;
        for (size_t it = 0; it < tw.size(); ++it) {
            int irr[4];
            for (int x = 0; x < 4; ++x) irr[x] = c->mesh.irr_of_full[tc[4 * it + x]];
            for (int j = 0; j < nb; ++j) {
                const double sg = d.ds[j];
                double tete[4];
                for (int x = 0; x < 4; ++x) tete[x] = c->om_irr[(size_t)irr[x] * nb + j];
                std::fill(proj.begin(), proj.end(), 0.0);
                for (int x = 0; x < 4; ++x)
                    for (int a = 0; a < na; ++a) proj[a] += site_projection(E + (size_t)irr[x] * nb * nb, nb, a, j);
                const double mn = *std::min_element(tete, tete + 4) - tmin, mx = *std::max_element(tete, tete + 4) - tmin;
                const int ii = std::max((int)std::floor(mn * invf), 1), jj = std::min((int)std::ceil(mx * invf), n);
                for (int k = ii; k <= jj; ++k) {
                    const double e = d.omega[k - 1];
                    double w1[4], w2[4];
                    lv_tetrahedron_weights(tete, e - de, kFreqTol, sg, w1);
                    lv_tetrahedron_weights(tete, e + de, kFreqTol, sg, w2);
                    double s = 0.0;
                    for (int x = 0; x < 4; ++x) s += w1[x] + w2[x];
                    const double f0 = s * tw[it] * 0.5;
                    pm[(size_t)(k - 1) * nb + j] += f0;
                    for (int a = 0; a < na; ++a) ps[(size_t)(k - 1) * na + a] += proj[a] * f0;
                }
            }
        }
    } else {
        const double radius = std::cbrt(3.0 / c->cr.volume / (double)c->mesh.nfull / 4.0 / kPi);
        for (int q = 0; q < nirr; ++q) {
            const double wq = c->mesh.weight_irr[q];
            for (int j = 0; j < nb; ++j) {
                const double f1 = c->om_irr[(size_t)q * nb + j];
                if (f1 < d.omin * 0.5) continue;
                double sg;
                if (o.dosint == 1) {
                    sg = d.ds[j] * o.dossigma;
                } else {
                    const double *v = &c->vel_irr[((size_t)q * nb + j) * 3];
                    sg = o.dossigma * (kTwoPi / std::sqrt(2.0)) * radius * std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                    sg = std::max(d.ds[j] * 0.25 * o.dossigma, sg);
                    sg = std::min(d.ds[j] * 4.0 * o.dossigma, sg);
                }
                for (int a = 0; a < na; ++a) proj[a] = site_projection(E + (size_t)q * nb * nb, nb, a, j);
                const int ii = std::max((int)std::floor((f1 - tmin - 4 * sg) * invf), 1);
                const int jj = std::min((int)std::floor((f1 - tmin + 4 * sg) * invf), n);
                for (int k = ii; k <= jj; ++k) {
                    const double dy = gauss(f1, d.omega[k - 1], sg);
                    pm[(size_t)(k - 1) * nb + j] += dy * wq;
                    for (int a = 0; a < na; ++a) ps[(size_t)(k - 1) * na + a] += dy * proj[a];
                }
            }
        }
    }
    // makenice (type_phonon_dos.f90:213-290)
    auto floor_line = [&](std::vector<double> &p, int ncol) {
        for (int j = 0; j < ncol; ++j) {
            const double f1 = p[j];
            for (int i = 1; i <= n; ++i) {
                double f0 = (n - i) * 1.0 / ((n - 1) * 1.0);
                f0 = f0 * f0;
                double &x = p[(size_t)(i - 1) * ncol + j];
                x = x - f0 * f1;
                if (x < kFreqTol) x = 0.0;
            }
            p[j] = 0.0;
        }
    };
    floor_line(pm, nb);
    floor_line(ps, na);
    d.dos.assign(n, 0.0);
    for (int j = 0; j < nb; ++j) {
        const double f0 = 1.0 / trapezoid(d.omega, &pm[j], nb);
        for (int i = 0; i < n; ++i) {
            d.dos[i] = d.dos[i] + pm[(size_t)i * nb + j];
            pm[(size_t)i * nb + j] *= f0;
        }
    }
    d.dos[0] = 0.0;
    const double dostol = std::accumulate(d.dos.begin(), d.dos.end(), 0.0) / n * kSqTol;
    for (double &x : d.dos) x = chop_scalar(x, dostol);
    {
        const double f0 = trapezoid(d.omega, d.dos.data());
        for (double &x : d.dos) x = x * nb / f0;
    }
    for (int i = 0; i < n; ++i) {
        double f0 = 0.0;
        for (int j = 0; j < nb; ++j) f0 += pm[(size_t)i * nb + j];
        for (int j = 0; j < nb; ++j) pm[(size_t)i * nb + j] = (f0 > dostol) ? pm[(size_t)i * nb + j] * d.dos[i] / f0 : 0.0;
    }
    std::vector<int> deg;
    std::vector<std::vector<int>> mem;
    atom_classes(c, deg, mem);
    std::vector<double> sb((size_t)n * na, 0.0);
    for (int i = 0; i < na; ++i) {
        for (int j = 0; j < deg[i]; ++j)
            for (int k = 0; k < n; ++k) sb[(size_t)k * na + i] = sb[(size_t)k * na + i] + ps[(size_t)k * na + mem[i][j]] / (1.0 * deg[j]);
        const double f0 = trapezoid(d.omega, &sb[i], na);
        for (int k = 0; k < n; ++k) sb[(size_t)k * na + i] = sb[(size_t)k * na + i] * 3.0 / f0;
    }
    for (int i = 0; i < n; ++i) {
        double f0 = 0.0;
        for (int a = 0; a < na; ++a) f0 += sb[(size_t)i * na + a];
        for (int a = 0; a < na; ++a) sb[(size_t)i * na + a] = (f0 > dostol) ? sb[(size_t)i * na + a] * d.dos[i] / f0 : 0.0;
    }
    d.pdos_mode = std::move(pm);
    d.pdos_site = std::move(sb);
}

// get_cumulative_kappa (cumulative_kappa.f90:58-414)
void cumulative_kappa(const tc3_ctx *c, int iT, const CumulativeOpts &o, CumulativeResult &r) {
    const int na = c->cr.na, nb = c->nb, nirr = c->mesh.nirr, nf = c->mesh.nfull, npts = o.mfppts;
    const double T = c->temps[iT];
    const double kappatol = 1e-5 * kKappaSiToAu;
    const double *Fn = &c->Fn[(size_t)iT * nirr * nb * 3];
    std::vector<double> kflat((size_t)nirr * nb * 6, 0.0);
    for (int fq = 0; fq < nf; ++fq) {
        const int iq = c->mesh.irr_of_full[fq], k = c->mesh.op_of_full[fq];
        for (int b = 0; b < nb; ++b) {
            const double om = c->om_irr[(size_t)iq * nb + b];
            if (om < kFreqTol) continue;
            const double cv = cv_of(T, om, c->classical);
            double v0[3], v1[3];
            rotate(c, k, &Fn[((size_t)iq * nb + b) * 3], v0);
            rotate(c, k, &c->vel_irr[((size_t)iq * nb + b) * 3], v1);
            double *kf = &kflat[((size_t)iq * nb + b) * 6];
            auto buf = [&](int i, int j) { return cv * (v0[i] * v1[j]) / c->cr.volume; };
            kf[0] += buf(0, 0); kf[1] += buf(1, 1); kf[2] += buf(2, 2);
            kf[3] += buf(1, 2); kf[4] += buf(0, 2); kf[5] += buf(0, 1);
        }
    }
    std::vector<double> kchop((size_t)nirr * nb * na * 6, 0.0);   // [iq][b][atom][6]
    std::vector<double> prj(na);
    for (int iq = 0; iq < nirr; ++iq)
        for (int b = 0; b < nb; ++b) {
            double s = 0.0;
            for (int a = 0; a < na; ++a) s += (prj[a] = site_projection(&c->egv_irr[(size_t)iq * nb * nb], nb, a, b));
            for (int a = 0; a < na; ++a) prj[a] = chop_scalar(prj[a] / s, kSqTol);
            for (int a = 0; a < na; ++a)
                for (int i = 0; i < 6; ++i)
                    kchop[(((size_t)iq * nb + b) * na + a) * 6 + i] = kflat[((size_t)iq * nb + b) * 6 + i] * prj[a] / nf;
        }
    double ktot[6];
    bool live[6];
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (size_t x = i; x < kchop.size(); x += 6) s += kchop[x];
        ktot[i] = chop_scalar(s, kappatol);
        live[i] = ktot[i] >= kappatol;
    }
    // scalar mean free paths |v| / (2 Gamma) (scattering.f90:284-289)
    const double *lw = &c->linewidth[(size_t)iT * nirr * nb];
    std::vector<double> mfp((size_t)nirr * nb, 0.0);
    for (size_t i = 0; i < mfp.size(); ++i) {
        const double *v = &c->vel_irr[3 * i];
        if (lw[i] > kFreqTol) mfp[i] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / (2.0 * lw[i]);
    }
    const double mfptol = 1e-12 / kGroupvelAuToMs;
    std::vector<double> x(mfp.size());
    for (size_t i = 0; i < mfp.size(); ++i) x[i] = std::log(mfp[i] > mfptol ? mfp[i] : mfptol);
    std::sort(x.begin(), x.end());
    const size_t nx = x.size();
    const double a = x[(size_t)std::ceil(nx * 0.5) - 1];
    const double bq = x[std::min(nx, (size_t)std::ceil(nx * 0.98)) - 1];
    double logsig = 0.0;
    for (size_t i = 0; i + 1 < nx; ++i)
        if (x[i] > a && x[i] < bq) logsig = std::max(logsig, x[i + 1] - x[i]);
    logsig *= o.sigma;
    const int nn = npts * 20;
    std::vector<double> xm((size_t)na * nb * nn), ym((size_t)na * nb * nn * 6, 0.0);   // [atom][mode][k]([6])
    std::vector<int> ind(nirr);
    std::vector<double> xq(nirr), y((size_t)nirr * 6);
    for (int at = 0; at < na; ++at)
        for (int md = 0; md < nb; ++md) {
            for (int iq = 0; iq < nirr; ++iq) ind[iq] = iq;
            std::stable_sort(ind.begin(), ind.end(), [&](int p, int q) { return mfp[(size_t)p * nb + md] < mfp[(size_t)q * nb + md]; });
            double minx = std::numeric_limits<double>::infinity(), maxx = -std::numeric_limits<double>::infinity();
            for (int iq = 0; iq < nirr; ++iq) {
                xq[iq] = mfp[(size_t)ind[iq] * nb + md];
                for (int i = 0; i < 6; ++i) y[(size_t)iq * 6 + i] = live[i] ? kchop[(((size_t)ind[iq] * nb + md) * na + at) * 6 + i] : 0.0;
                if (xq[iq] > kTol) minx = std::min(minx, xq[iq]);
                maxx = std::max(maxx, xq[iq]);
            }
            minx *= 0.1;
            maxx *= 1.5;
            double *xmp = &xm[((size_t)at * nb + md) * nn];
            double *ymp = &ym[((size_t)at * nb + md) * nn * 6];
            const std::vector<double> ls = linspace(minx, maxx, nn);
            std::copy(ls.begin(), ls.end(), xmp);
            for (int iq = 0; iq < nirr; ++iq) {
                const double xv = xq[iq];
                if (xv < minx || xv > maxx) continue;
                const double la = std::log(xv);
                const double s1 = std::exp(la - 4 * logsig) - xv, s2 = std::exp(la + 4 * logsig) - xv;
                const double isig = 1.0 / logsig;
                const int ii = std::max(1, (int)std::floor(nn * (xv + s1 - minx) / maxx));
                const int jj = std::min(nn, (int)std::ceil(nn * (xv + s2 - minx) / maxx));
                if (jj <= ii) continue;
                const double *yq = &y[(size_t)iq * 6];
                for (int k = ii; k <= jj; ++k) {
                    const double f1 = (1.0 + std::erf((std::log(xmp[k - 1]) - std::log(xv)) * isig)) * 0.5;
                    for (int i = 0; i < 6; ++i) ymp[(size_t)(k - 1) * 6 + i] += f1 * yq[i];
                }
                for (int k = jj + 1; k <= nn; ++k)
                    for (int i = 0; i < 6; ++i) ymp[(size_t)(k - 1) * 6 + i] += yq[i];
            }
        }
    const double xmin = *std::min_element(xm.begin(), xm.end()), xmax = *std::max_element(xm.begin(), xm.end());
    r.mfaxis = logspace(xmin, xmax * 2.0, npts);
    r.mf.assign((size_t)npts * 6, 0.0);
    r.mf_band.assign((size_t)npts * nb * 6, 0.0);
    r.mf_atom.assign((size_t)npts * na * 6, 0.0);
    std::vector<double> col(nn);
    double f0[6] = {0, 0, 0, 0, 0, 0};
    for (int md = 0; md < nb; ++md)
        for (int at = 0; at < na; ++at) {
            const double *xmp = &xm[((size_t)at * nb + md) * nn];
            const double *ymp = &ym[((size_t)at * nb + md) * nn * 6];
            for (int i = 0; i < npts; ++i) {
                const double ax = r.mfaxis[i];
                if (ax < xmp[0] * 1.01) {
                    for (double &v : f0) v = 0.0;
                } else if (ax > xmp[nn - 1] * 0.999) {
                    for (int d = 0; d < 6; ++d) f0[d] = ymp[(size_t)(nn - 1) * 6 + d];
                } else {
                    for (int d = 0; d < 6; ++d) {
                        if (!live[d]) continue;
                        for (int k = 0; k < nn; ++k) col[k] = ymp[(size_t)k * 6 + d];
                        f0[d] = linear_interpolation(xmp, col.data(), nn, ax);
                    }
                }
                for (int d = 0; d < 6; ++d) {
                    r.mf[(size_t)i * 6 + d] += f0[d];
                    r.mf_band[((size_t)i * nb + md) * 6 + d] += f0[d];
                    r.mf_atom[((size_t)i * na + at) * 6 + d] += f0[d];
                }
            }
        }
    for (int i = 0; i + 1 < npts; ++i)
        for (int d = 0; d < 6; ++d) {
            if (!live[d]) continue;
            double *m0 = &r.mf[(size_t)i * 6 + d], *m1 = m0 + 6;
            if (*m1 < *m0) *m1 = *m0;
            for (int at = 0; at < na; ++at) {
                double &p0 = r.mf_atom[((size_t)i * na + at) * 6 + d], &p1 = r.mf_atom[((size_t)(i + 1) * na + at) * 6 + d];
                if (p1 < p0) p1 = p0;
            }
            for (int md = 0; md < nb; ++md) {
                double &p0 = r.mf_band[((size_t)i * nb + md) * 6 + d], &p1 = r.mf_band[((size_t)(i + 1) * nb + md) * 6 + d];
                if (p1 < p0) p1 = p0;
            }
        }
    std::vector<int> deg;
    std::vector<std::vector<int>> mem;
    atom_classes(c, deg, mem);
    std::vector<double> sb((size_t)npts * na);
    for (int d = 0; d < 6; ++d) {
        if (!live[d]) continue;
        std::fill(sb.begin(), sb.end(), 0.0);
        for (int i = 0; i < na; ++i)
            for (int j = 0; j < deg[i]; ++j)
                for (int k = 0; k < npts; ++k)
                    sb[(size_t)k * na + i] = sb[(size_t)k * na + i] + r.mf_atom[((size_t)k * na + mem[i][j]) * 6 + d] / (double)deg[j];
        for (int k = 0; k < npts; ++k)
            for (int i = 0; i < na; ++i) r.mf_atom[((size_t)k * na + i) * 6 + d] = sb[(size_t)k * na + i];
    }
    for (int d = 0; d < 6; ++d) {
        if (!live[d]) continue;
        const double last = r.mf[(size_t)(npts - 1) * 6 + d];
        for (int i = 0; i < npts; ++i) r.mf[(size_t)i * 6 + d] = r.mf[(size_t)i * 6 + d] * ktot[d] / last;
    }
    for (int i = 0; i < npts; ++i)
        for (int d = 0; d < 6; ++d) {
            if (!live[d]) continue;
            const double f2 = r.mf[(size_t)i * 6 + d];
            double f1 = 0.0;
            for (int md = 0; md < nb; ++md) f1 += r.mf_band[((size_t)i * nb + md) * 6 + d];
            for (int md = 0; md < nb; ++md) {
                double &v = r.mf_band[((size_t)i * nb + md) * 6 + d];
                v = (f1 > kappatol && f2 > kappatol) ? v * f2 / f1 : 0.0;
            }
            f1 = 0.0;
            for (int at = 0; at < na; ++at) f1 += r.mf_atom[((size_t)i * na + at) * 6 + d];
            for (int at = 0; at < na; ++at) {
                double &v = r.mf_atom[((size_t)i * na + at) * 6 + d];
                v = (f1 > kappatol && f2 > kappatol) ? v * f2 / f1 : 0.0;
            }
        }
    std::memcpy(r.kappa_total, ktot, sizeof(ktot));
    r.mfp = std::move(mfp);
}

// get_boundary_kappa (cumulative_kappa.f90:416-508)
void boundary_kappa(const tc3_ctx *c, int iT, const CumulativeOpts &o, CumulativeResult &r) {
    const int nb = c->nb, nirr = c->mesh.nirr, npts = o.mfppts;
    const double T = c->temps[iT];
    const double *Fn = &c->Fn[(size_t)iT * nirr * nb * 3];
    const double *lwv = &c->linewidth[(size_t)iT * nirr * nb];
    double minx = std::numeric_limits<double>::max(), maxx = -std::numeric_limits<double>::max();
    for (double f : r.mfp)
        if (f > kTol) {
            minx = std::min(minx, f);
            maxx = std::max(maxx, f);
        }
    minx = std::max(minx * 1e-1, kGroupVelTol);
    maxx = maxx * 1e6;
    r.baxis = logspace(minx, maxx, npts);
    r.bkappa.assign((size_t)6 * npts, 0.0);
    const int nops = (int)c->ops.size();
    // symmetrised off-diagonal kappa of this temperature, column-major 3x3
    double kod[9] = {0};
    if (c->kappa_od.size() >= (size_t)(iT + 1) * 9) std::memcpy(kod, &c->kappa_od[(size_t)iT * 9], sizeof(kod));
    double f0 = 0.0;
    for (int ll = 0; ll < npts; ++ll) {
        double K[9];   // row-major
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) K[3 * i + j] = kod[3 * j + i];
        for (int iq = 0; iq < nirr; ++iq)
            for (int b = 0; b < nb; ++b) {
                const double om = c->om_irr[(size_t)iq * nb + b];
                if (om < kFreqTol) continue;
                const double cv = cv_of(T, om, c->classical);
                const double *vel = &c->vel_irr[((size_t)iq * nb + b) * 3], *fn = &Fn[((size_t)iq * nb + b) * 3];
                const double lw = lwv[(size_t)iq * nb + b];
                const double vn = std::sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
                if (vn > kGroupVelTol) f0 = vn / r.baxis[ll];
                const double f1 = (lw + f0 > kFreqTol) ? lw / (lw + f0) : 1.0;
                double v2[9] = {0};
                for (int j = 0; j < nops; ++j) {
                    const double *m = c->ops[j].m;
                    double a0[3], a1[3];
                    for (int i = 0; i < 3; ++i) {
                        a0[i] = m[3 * i] * fn[0] + m[3 * i + 1] * fn[1] + m[3 * i + 2] * fn[2];
                        a1[i] = m[3 * i] * vel[0] + m[3 * i + 1] * vel[1] + m[3 * i + 2] * vel[2];
                    }
                    for (int i = 0; i < 3; ++i)
                        for (int k = 0; k < 3; ++k) v2[3 * i + k] = v2[3 * i + k] + a0[i] * a1[k] / nops;
                }
                const double s = cv * c->mesh.weight_irr[iq] * f1 / c->cr.volume;
                for (int i = 0; i < 9; ++i) K[i] = K[i] + v2[i] * s;
            }
        // symmetrize_kappa (kappa.f90:308-324), then chop again (cumulative_kappa.f90:490)
        double S[9] = {0};
        for (int j = 0; j < nops; ++j) {
            const double *m = c->ops[j].m;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    double s = 0.0;
                    for (int p = 0; p < 3; ++p)
                        for (int q = 0; q < 3; ++q) s += m[3 * a + p] * K[3 * p + q] * m[3 * b + q];
                    S[3 * a + b] += s;
                }
        }
        for (int rep = 0; rep < 2; ++rep) {
            double t = 0.0;
            for (int i = 0; i < 9; ++i) {
                if (rep == 0) S[i] /= nops;
                t += std::fabs(S[i]);
            }
            for (int i = 0; i < 9; ++i) S[i] = chop_scalar(S[i], t * 1e-6);
        }
        double *bk = &r.bkappa[0];
        bk[0 * npts + ll] = S[0];
        bk[1 * npts + ll] = S[4];
        bk[2 * npts + ll] = S[8];
        bk[3 * npts + ll] = S[5];
        bk[4 * npts + ll] = S[2];
        bk[5 * npts + ll] = S[1];
    }
}

// get_spectral_kappa and lo_phonon_dos%spectral_kappa (cumulative_kappa.f90:510-561,
// type_phonon_dos_integrations.f90:428-809)
void spectral_kappa(const tc3_ctx *c, int iT, const CumulativeOpts &o, const PhononDos &d, const std::vector<int> &tc,
                    const std::vector<double> &tw, CumulativeResult &r) {
    const int na = c->cr.na, nb = c->nb, nirr = c->mesh.nirr, nf = c->mesh.nfull, n = o.freqpts;
    const double T = c->temps[iT];
    const double kappatol = 1e-6 * kKappaSiToAu;
    const double *Fn = &c->Fn[(size_t)iT * nirr * nb * 3];
    std::vector<double> kf((size_t)nirr * nb * 9, 0.0);
    for (int aq = 0; aq < nf; ++aq) {
        const int iq = c->mesh.irr_of_full[aq], k = c->mesh.op_of_full[aq];
        for (int b = 0; b < nb; ++b) {
            const double om = c->om_irr[(size_t)iq * nb + b];
            if (om < kFreqTol) continue;
            const double cv = cv_of(T, om, c->classical);
            double v0[3], v1[3];
            rotate(c, k, &Fn[((size_t)iq * nb + b) * 3], v0);
            rotate(c, k, &c->vel_irr[((size_t)iq * nb + b) * 3], v1);
            double *p = &kf[((size_t)iq * nb + b) * 9];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) p[3 * i + j] += cv * (v0[i] * v1[j]) / c->cr.volume;
        }
    }
    std::vector<double> kflat((size_t)nf * nb * 9);
    for (int aq = 0; aq < nf; ++aq) {
        const int iq = c->mesh.irr_of_full[aq];
        const double f0 = 1.0 / c->mesh.weight_irr[iq];
        for (int b = 0; b < nb; ++b)
            for (int i = 0; i < 9; ++i)
                kflat[((size_t)aq * nb + b) * 9 + i] = chop_scalar(kf[((size_t)iq * nb + b) * 9 + i] * f0 / nf, kappatol);
    }
    double ktot[9];
    for (int i = 0; i < 9; ++i) {
        double s = 0.0;
        for (size_t x = i; x < kflat.size(); x += 9) s += kflat[x];
        ktot[i] = s / nf;
    }
    std::vector<double> sb((size_t)n * nb * 9, 0.0), sa((size_t)n * na * 9, 0.0);   // [x][mode|atom][9]
    const double invf = n / *std::max_element(d.omega.begin(), d.omega.end());
    if (o.dosint == 1 || o.dosint == 2) {
        double scaled[9];   // 2 pi reciprocal vectors / grid, columns
        for (int x = 0; x < 3; ++x)
            for (int i = 0; i < 3; ++i) scaled[3 * x + i] = c->cr.rec[3 * x + i] * kTwoPi / c->mesh.dims[i];
        for (int dr = 0; dr < 9; ++dr) {
            if (std::fabs(ktot[dr]) < kappatol) continue;
            for (int iq = 0; iq < nirr; ++iq) {
                const double wq = c->mesh.weight_irr[iq];
                const int fi = c->mesh.irr_full_index[iq];
                for (int b = 0; b < nb; ++b) {
                    const double f1 = c->om_irr[(size_t)iq * nb + b];
                    if (f1 < d.omin * 0.5) continue;
                    double sg;
                    if (o.dosint == 1) {
                        sg = d.ds[b] * o.dossigma;
                    } else {
                        const double *v = &c->vel_irr[((size_t)iq * nb + b) * 3];
                        const double g[3] = {std::fabs(v[0]), std::fabs(v[1]), std::fabs(v[2])};
                        double s = 0.0;
                        for (int i = 0; i < 3; ++i) {
                            const double dp = g[0] * scaled[i] + g[1] * scaled[3 + i] + g[2] * scaled[6 + i];
                            s += dp * dp;
                        }
                        sg = std::sqrt(s * 0.083333333333) * o.dossigma;
                        sg = std::min(sg, o.dossigma * d.ds[b] * 4.0);
                        sg = std::max(sg, o.dossigma * d.ds[b] * 0.25);
                    }
                    const double qk = kflat[((size_t)fi * nb + b) * 9 + dr];
                    const int ii = std::max((int)std::floor((f1 - 4 * sg) * invf), 1);
                    const int jj = std::min((int)std::floor((f1 + 4 * sg) * invf), n);
                    for (int x = ii; x <= jj; ++x) {
                        const double f0 = gauss(f1, d.omega[x - 1], sg) * wq;
                        sb[((size_t)(x - 1) * nb + b) * 9 + dr] += qk * f0;
                        for (int a = 0; a < na; ++a)
                            sa[((size_t)(x - 1) * na + a) * 9 + dr] +=
                                qk * site_projection(&c->egv_irr[(size_t)iq * nb * nb], nb, a, b) * f0;
                    }
                }
            }
        }
    } else {
        const size_t ntet = tw.size();
        std::vector<double> tp(ntet * nb * 4 * na);   // [tet][mode][corner][atom]
        for (size_t it = 0; it < ntet; ++it)
            for (int b = 0; b < nb; ++b)
                for (int x = 0; x < 4; ++x)
                    for (int a = 0; a < na; ++a)
                        tp[((it * nb + b) * 4 + x) * na + a] = site_projection(&c->egv_full[(size_t)tc[4 * it + x] * nb * nb], nb, a, b);
        for (int dr = 0; dr < 9; ++dr)
            for (int b = 0; b < nb; ++b) {
                const double sg = d.ds[b];
                for (size_t it = 0; it < ntet; ++it) {
                    double tete[4], tk[4];
                    for (int x = 0; x < 4; ++x) {
                        const int cq = tc[4 * it + x];
                        tete[x] = c->om_full[(size_t)cq * nb + b];
                        tk[x] = kflat[((size_t)cq * nb + b) * 9 + dr];
                    }
                    const int ii = std::max((int)std::floor(*std::min_element(tete, tete + 4) * invf), 1);
                    const int jj = std::min((int)std::ceil(*std::max_element(tete, tete + 4) * invf), n);
                    const double *p = &tp[(it * nb + b) * 4 * na];
                    for (int x = ii; x <= jj; ++x) {
                        double w[4];
                        lv_tetrahedron_weights(tete, d.omega[x - 1], kFreqTol, sg, w);
                        double s = 0.0;
                        for (int k = 0; k < 4; ++k) {
                            w[k] *= tw[it];
                            s += w[k] * tk[k];
                        }
                        sb[((size_t)(x - 1) * nb + b) * 9 + dr] += s;
                        for (int a = 0; a < na; ++a) {
                            double sa1 = 0.0;
                            for (int k = 0; k < 4; ++k) sa1 += w[k] * tk[k] * p[k * na + a];
                            sa[((size_t)(x - 1) * na + a) * 9 + dr] += sa1;
                        }
                    }
                }
            }
    }
    // makenice (type_phonon_dos_integrations.f90:695-807)
    double ker[13], ks = 0.0;
    for (int i = -6; i <= 6; ++i) ks += (ker[i + 6] = gauss(0.0, i * 1.0, 1.5));
    for (double &v : ker) v /= ks;
    for (double &v : sa) v = std::max(v, 0.0);
    for (double &v : sb) v = std::max(v, 0.0);
    r.spec.assign((size_t)n * 9, 0.0);
    std::vector<int> deg;
    std::vector<std::vector<int>> mem;
    atom_classes(c, deg, mem);
    std::vector<double> ysm(n + 12);
    auto smooth = [&](std::vector<double> &v, int ncol, int col, int dr) {
        std::fill(ysm.begin(), ysm.end(), 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = -6; j <= 6; ++j) ysm[i + 6 + j] += ker[j + 6] * v[((size_t)i * ncol + col) * 9 + dr];
        for (int i = 0; i < n; ++i) v[((size_t)i * ncol + col) * 9 + dr] = ysm[i + 6];
    };
    auto zero_at_zero = [&](std::vector<double> &v, int ncol, int col, int dr) {
        const double f0 = v[(size_t)col * 9 + dr];
        for (int i = 1; i <= n; ++i) {
            const double f1 = 1.0 - (i - 1.0) / (double)(n - 1);
            double &x = v[((size_t)(i - 1) * ncol + col) * 9 + dr];
            x = std::max(x - f1 * f0, 0.0);
        }
        v[(size_t)col * 9 + dr] = 0.0;
    };
    std::vector<double> ys(n);
    for (int dr = 0; dr < 9; ++dr) {
        if (ktot[dr] < kappatol) continue;
        for (int a = 0; a < na; ++a) smooth(sa, na, a, dr);
        for (int b = 0; b < nb; ++b) smooth(sb, nb, b, dr);
        // site degeneracies as written (type_phonon_dos_integrations.f90:740-751): the
        // spectrum of atom i, summed deg(i) times over deg(i), given to its whole class
        for (int i = 0; i < na; ++i) {
            std::fill(ys.begin(), ys.end(), 0.0);
            for (int j = 0; j < deg[i]; ++j)
                for (int x = 0; x < n; ++x) ys[x] = ys[x] + sa[((size_t)x * na + i) * 9 + dr] / (1.0 * deg[i]);
            for (int j = 0; j < deg[i]; ++j)
                for (int x = 0; x < n; ++x) sa[((size_t)x * na + mem[i][j]) * 9 + dr] = ys[x];
        }
        for (int a = 0; a < na; ++a) zero_at_zero(sa, na, a, dr);
        for (int b = 0; b < nb; ++b) zero_at_zero(sb, nb, b, dr);
        for (int x = 0; x < n; ++x) {
            double s = 0.0;
            for (int b = 0; b < nb; ++b) s = s + sb[((size_t)x * nb + b) * 9 + dr];
            r.spec[(size_t)x * 9 + dr] = s;
        }
        const double f0 = trapezoid(d.omega, &r.spec[dr], 9);
        for (int x = 0; x < n; ++x)
            r.spec[(size_t)x * 9 + dr] = (f0 > kappatol) ? r.spec[(size_t)x * 9 + dr] * ktot[dr] / f0 : 0.0;
        for (int x = 0; x < n; ++x) {
            double f0a = 0.0, f1b = 0.0;
            for (int a = 0; a < na; ++a) f0a += sa[((size_t)x * na + a) * 9 + dr];
            for (int b = 0; b < nb; ++b) f1b += sb[((size_t)x * nb + b) * 9 + dr];
            const double f2 = r.spec[(size_t)x * 9 + dr];
            if (f2 > kappatol) {
                for (int a = 0; a < na; ++a) {
                    double &v = sa[((size_t)x * na + a) * 9 + dr];
                    v = (f0a > kappatol) ? v * f2 / f0a : 0.0;
                }
                for (int b = 0; b < nb; ++b) {
                    double &v = sb[((size_t)x * nb + b) * 9 + dr];
                    v = (f1b > kappatol) ? v * f2 / f1b : 0.0;
                }
            } else {
                r.spec[(size_t)x * 9 + dr] = 0.0;
                for (int a = 0; a < na; ++a) sa[((size_t)x * na + a) * 9 + dr] = 0.0;
                for (int b = 0; b < nb; ++b) sb[((size_t)x * nb + b) * 9 + dr] = 0.0;
            }
        }
    }
    r.spec_band = std::move(sb);
    r.spec_atom = std::move(sa);
    std::memcpy(r.spec_total, ktot, sizeof(ktot));
}

namespace {

// lo_get_small_group_of_qpoint (type_qpointmesh.f90:660-700), no time reversal
std::vector<int> small_group(const tc3_ctx *c, const double *q) {
    std::vector<int> out;
    for (int o = 0; o < (int)c->ops.size(); ++o) {
        const double *m = c->ops[o].m;
        double d[3], f[3], s = 0.0;
        for (int i = 0; i < 3; ++i) d[i] = q[i] - (m[3 * i] * q[0] + m[3 * i + 1] * q[1] + m[3 * i + 2] * q[2]);
        for (int i = 0; i < 3; ++i) {
            f[i] = c->cr.invrec[3 * i] * d[0] + c->cr.invrec[3 * i + 1] * d[1] + c->cr.invrec[3 * i + 2] * d[2];
            f[i] -= std::nearbyint(f[i]);
            s += f[i] * f[i];
        }
        if (s < kSqTol) out.push_back(o);
    }
    return out;
}

// sum over atoms of e^H M_a e with M_a = -i (spin-1 generators) of
// phonon_angular_momentum_matrix (type_phonon_dispersions.f90:886-897):
// 2 Im(conj(e_y) e_z), 2 Im(conj(e_z) e_x), 2 Im(conj(e_x) e_y)
void mode_angular_momentum(const cd *egv, int nb, int na, int mode, double *l) {
    l[0] = l[1] = l[2] = 0.0;
    for (int k = 0; k < na; ++k) {
        const cd x = egv[(size_t)(3 * k) * nb + mode], y = egv[(size_t)(3 * k + 1) * nb + mode],
                 z = egv[(size_t)(3 * k + 2) * nb + mode];
        auto im_conj_mul = [](cd a, cd b) { return a.x * b.y - a.y * b.x; };   // Im(conj(a) b)
        l[0] += 2.0 * im_conj_mul(y, z);
        l[1] += 2.0 * im_conj_mul(z, x);
        l[2] += 2.0 * im_conj_mul(x, y);
    }
}

}  // namespace

// phonon_angular_momentum_matrix (type_phonon_dispersions.f90:868-971) with the linewidths of
// temperature iT; alpha row-major (alpha[3a + b] = sum l_a v_b ...)
void angular_momentum_matrix(const tc3_ctx *c, int iT, double *alpha) {
    const int na = c->cr.na, nb = c->nb, nf = c->mesh.nfull, nirr = c->mesh.nirr;
    const double T = c->temps[iT];
    const double *lw = &c->linewidth[(size_t)iT * nirr * nb];
    for (int i = 0; i < 9; ++i) alpha[i] = 0.0;
    for (int i = 0; i < nf; ++i) {
        const std::vector<int> lg = small_group(c, &c->mesh.qcart[3 * (size_t)i]);
        const int k = c->mesh.irr_of_full[i];
        for (int j = 0; j < nb; ++j) {
            const double om = c->om_full[(size_t)i * nb + j];
            if (om < kFreqTol) continue;
            double v1[3], v0[3] = {0, 0, 0}, w0[3] = {0, 0, 0};
            mode_angular_momentum(&c->egv_full[(size_t)i * nb * nb], nb, na, j, v1);
            const double *w1 = &c->vel_full[((size_t)i * nb + j) * 3];
            for (int o : lg) {
                const double *m = c->ops[o].m;
                for (int a = 0; a < 3; ++a) {
                    v0[a] = v0[a] + (m[3 * a] * v1[0] + m[3 * a + 1] * v1[1] + m[3 * a + 2] * v1[2]);
                    w0[a] = w0[a] + (m[3 * a] * w1[0] + m[3 * a + 1] * w1[1] + m[3 * a + 2] * w1[2]);
                }
            }
            for (int a = 0; a < 3; ++a) {
                v0[a] /= (double)lg.size();
                w0[a] /= (double)lg.size();
            }
            double f0 = cv_of(T, om, false) / om;
            f0 = f0 * (1.0 / nf);
            f0 = f0 / (2.0 * lw[(size_t)k * nb + j]);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) alpha[3 * a + b] = alpha[3 * a + b] + v0[a] * w0[b] * f0;
        }
    }
    double nrm = 0.0;
    for (int i = 0; i < 9; ++i) {
        alpha[i] /= c->cr.volume;
        nrm += alpha[i] * alpha[i];
    }
    nrm = std::sqrt(nrm);
    for (int i = 0; i < 9; ++i) alpha[i] = chop_scalar(alpha[i], nrm * 1e-10);
}

// spectral_angular_momentum (type_phonon_dos_integrations.f90:8-425) as written: the mode's
// angular momentum (not its little-group average) times c_v / omega times twice the
// linewidth, outer the little-group-averaged velocity
void spectral_angular_momentum(const tc3_ctx *c, int iT, const CumulativeOpts &o, const PhononDos &d,
                               const std::vector<int> &tc, const std::vector<double> &tw, CumulativeResult &r) {
    const int na = c->cr.na, nb = c->nb, nirr = c->mesh.nirr, nf = c->mesh.nfull, n = o.freqpts;
    const double T = c->temps[iT];
    const double tol = 1e-30;
    const double *lw = &c->linewidth[(size_t)iT * nirr * nb];
    std::vector<double> kf((size_t)nirr * nb * 9, 0.0);
    for (int iq = 0; iq < nf; ++iq) {
        const int ii = c->mesh.irr_of_full[iq];
        const std::vector<int> lg = small_group(c, &c->mesh.qcart[3 * (size_t)iq]);
        for (int b = 0; b < nb; ++b) {
            const double w = c->om_full[(size_t)iq * nb + b];
            if (w < kFreqTol) continue;
            double l[3], w0[3] = {0, 0, 0};
            mode_angular_momentum(&c->egv_full[(size_t)iq * nb * nb], nb, na, b, l);
            const double *w1 = &c->vel_full[((size_t)iq * nb + b) * 3];
            for (int op : lg) {
                const double *m = c->ops[op].m;
                for (int a = 0; a < 3; ++a) w0[a] = w0[a] + (m[3 * a] * w1[0] + m[3 * a + 1] * w1[1] + m[3 * a + 2] * w1[2]);
            }
            for (double &x : w0) x /= (double)lg.size();
            double f0 = cv_of(T, w, false);
            f0 = f0 / w;
            f0 = f0 * (1.0 / nf);
            f0 = f0 * (2.0 * lw[(size_t)ii * nb + b]);
            double *p = &kf[((size_t)ii * nb + b) * 9];
            for (int a = 0; a < 3; ++a)
                for (int e = 0; e < 3; ++e) p[3 * a + e] += (l[a] * f0) * w0[e];
        }
    }
    std::vector<double> flat((size_t)nf * nb * 9);
    for (int iq = 0; iq < nf; ++iq) {
        const int ii = c->mesh.irr_of_full[iq];
        const double f0 = 1.0 / c->mesh.weight_irr[ii];
        for (int b = 0; b < nb; ++b)
            for (int i = 0; i < 9; ++i) flat[((size_t)iq * nb + b) * 9 + i] = kf[((size_t)ii * nb + b) * 9 + i] * f0;
    }
    double total[9], nrm = 0.0;
    for (int i = 0; i < 9; ++i) {
        double s = 0.0;
        for (size_t x = i; x < flat.size(); x += 9) s += flat[x];
        total[i] = s / nf;
        nrm += total[i] * total[i];
    }
    nrm = std::sqrt(nrm);
    for (double &v : total) v = chop_scalar(v, nrm * 1e-9);
    std::vector<double> sb((size_t)n * nb * 9, 0.0), sa((size_t)n * na * 9, 0.0);
    const double invf = n / *std::max_element(d.omega.begin(), d.omega.end());
    if (o.dosint == 1 || o.dosint == 2) {
        double scaled[9];
        for (int x = 0; x < 3; ++x)
            for (int i = 0; i < 3; ++i) scaled[3 * x + i] = c->cr.rec[3 * x + i] * kTwoPi / c->mesh.dims[i];
        for (int dr = 0; dr < 9; ++dr) {
            if (std::fabs(total[dr]) < tol) continue;
            for (int iq = 0; iq < nirr; ++iq) {
                const double wq = c->mesh.weight_irr[iq];
                const int fi = c->mesh.irr_full_index[iq];
                for (int b = 0; b < nb; ++b) {
                    const double f1 = c->om_irr[(size_t)iq * nb + b];
                    if (f1 < d.omin * 0.5) continue;
                    double sg;
                    if (o.dosint == 1) {
                        sg = d.ds[b] * o.dossigma;
                    } else {
                        const double *v = &c->vel_irr[((size_t)iq * nb + b) * 3];
                        const double g[3] = {std::fabs(v[0]), std::fabs(v[1]), std::fabs(v[2])};
                        double s = 0.0;
                        for (int i = 0; i < 3; ++i) {
                            const double dp = g[0] * scaled[i] + g[1] * scaled[3 + i] + g[2] * scaled[6 + i];
                            s += dp * dp;
                        }
                        sg = std::sqrt(s * 0.083333333333) * o.dossigma;
                        sg = std::min(sg, o.dossigma * d.ds[b] * 4.0);
                        sg = std::max(sg, o.dossigma * d.ds[b] * 0.25);
                    }
                    const double qa = flat[((size_t)fi * nb + b) * 9 + dr];
                    const int ii = std::max((int)std::floor((f1 - 4 * sg) * invf), 1);
                    const int jj = std::min((int)std::floor((f1 + 4 * sg) * invf), n);
                    for (int x = ii; x <= jj; ++x) {
                        const double f0 = gauss(f1, d.omega[x - 1], sg) * wq;
                        sb[((size_t)(x - 1) * nb + b) * 9 + dr] += qa * f0;
                        for (int a = 0; a < na; ++a)
                            sa[((size_t)(x - 1) * na + a) * 9 + dr] +=
                                qa * site_projection(&c->egv_irr[(size_t)iq * nb * nb], nb, a, b) * f0;
                    }
                }
            }
        }
    } else {
        const size_t ntet = tw.size();
        std::vector<double> tp(ntet * nb * 4 * na);
        for (size_t it = 0; it < ntet; ++it)
            for (int b = 0; b < nb; ++b)
                for (int x = 0; x < 4; ++x)
                    for (int a = 0; a < na; ++a)
                        tp[((it * nb + b) * 4 + x) * na + a] = site_projection(&c->egv_full[(size_t)tc[4 * it + x] * nb * nb], nb, a, b);
        for (int dr = 0; dr < 9; ++dr)
            for (int b = 0; b < nb; ++b) {
                const double sg = d.ds[b];
                for (size_t it = 0; it < ntet; ++it) {
                    double tete[4], ta[4];
                    for (int x = 0; x < 4; ++x) {
                        const int cq = tc[4 * it + x];
                        tete[x] = c->om_full[(size_t)cq * nb + b];
                        ta[x] = flat[((size_t)cq * nb + b) * 9 + dr];
                    }
                    const int ii = std::max((int)std::floor(*std::min_element(tete, tete + 4) * invf), 1);
                    const int jj = std::min((int)std::ceil(*std::max_element(tete, tete + 4) * invf), n);
                    const double *p = &tp[(it * nb + b) * 4 * na];
                    for (int x = ii; x <= jj; ++x) {
                        double w[4];
                        lv_tetrahedron_weights(tete, d.omega[x - 1], kFreqTol, sg, w);
                        double s = 0.0;
                        for (int k = 0; k < 4; ++k) {
                            w[k] *= tw[it];
                            s += w[k] * ta[k];
                        }
                        sb[((size_t)(x - 1) * nb + b) * 9 + dr] += s;
                        for (int a = 0; a < na; ++a) {
                            double s1 = 0.0;
                            for (int k = 0; k < 4; ++k) s1 += w[k] * ta[k] * p[k * na + a];
                            sa[((size_t)(x - 1) * na + a) * 9 + dr] += s1;
                        }
                    }
                }
            }
    }
    double ker[13], ks = 0.0;
    for (int i = -6; i <= 6; ++i) ks += (ker[i + 6] = gauss(0.0, i * 1.0, 1.5));
    for (double &v : ker) v /= ks;
    r.am_spec.assign((size_t)n * 9, 0.0);
    std::vector<int> deg;
    std::vector<std::vector<int>> mem;
    atom_classes(c, deg, mem);
    std::vector<double> ysm(n + 12), ys(n);
    auto smooth = [&](std::vector<double> &v, int ncol, int col, int dr) {
        std::fill(ysm.begin(), ysm.end(), 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = -6; j <= 6; ++j) ysm[i + 6 + j] += ker[j + 6] * v[((size_t)i * ncol + col) * 9 + dr];
        for (int i = 0; i < n; ++i) v[((size_t)i * ncol + col) * 9 + dr] = ysm[i + 6];
    };
    for (int dr = 0; dr < 9; ++dr) {
        if (std::fabs(total[dr]) < tol) continue;
        for (int a = 0; a < na; ++a) smooth(sa, na, a, dr);
        for (int b = 0; b < nb; ++b) smooth(sb, nb, b, dr);
        for (int i = 0; i < na; ++i) {
            std::fill(ys.begin(), ys.end(), 0.0);
            for (int j = 0; j < deg[i]; ++j)
                for (int x = 0; x < n; ++x) ys[x] = ys[x] + sa[((size_t)x * na + i) * 9 + dr] / (1.0 * deg[i]);
            for (int j = 0; j < deg[i]; ++j)
                for (int x = 0; x < n; ++x) sa[((size_t)x * na + mem[i][j]) * 9 + dr] = ys[x];
        }
        for (int x = 0; x < n; ++x) {
            double s = 0.0;
            for (int b = 0; b < nb; ++b) s = s + sb[((size_t)x * nb + b) * 9 + dr];
            r.am_spec[(size_t)x * 9 + dr] = s;
        }
        const double f0 = trapezoid(d.omega, &r.am_spec[dr], 9);
        for (int x = 0; x < n; ++x)
            r.am_spec[(size_t)x * 9 + dr] = (std::fabs(f0) > tol) ? r.am_spec[(size_t)x * 9 + dr] * total[dr] / f0 : 0.0;
        for (int x = 0; x < n; ++x) {
            double f0a = 0.0, f1b = 0.0;
            for (int a = 0; a < na; ++a) f0a += sa[((size_t)x * na + a) * 9 + dr];
            for (int b = 0; b < nb; ++b) f1b += sb[((size_t)x * nb + b) * 9 + dr];
            const double f2 = r.am_spec[(size_t)x * 9 + dr];
            if (std::fabs(f2) > tol) {
                for (int a = 0; a < na; ++a) {
                    double &v = sa[((size_t)x * na + a) * 9 + dr];
                    v = (std::fabs(f0a) > tol) ? v * f2 / f0a : 0.0;
                }
                for (int b = 0; b < nb; ++b) {
                    double &v = sb[((size_t)x * nb + b) * 9 + dr];
                    v = (std::fabs(f1b) > tol) ? v * f2 / f1b : 0.0;
                }
            } else {
                r.am_spec[(size_t)x * 9 + dr] = 0.0;
                for (int a = 0; a < na; ++a) sa[((size_t)x * na + a) * 9 + dr] = 0.0;
                for (int b = 0; b < nb; ++b) sb[((size_t)x * nb + b) * 9 + dr] = 0.0;
            }
        }
    }
    std::memcpy(r.am_total, total, sizeof(total));
}

}  // namespace tc3
