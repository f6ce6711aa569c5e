// Host-side preprocessing of the tc3 engine: crystal, space group, FFT q-mesh,
// little groups, irreducible triplets and little-group orbits.
//
// The rules follow the reference:
//   space group contract      type_symmetryoperation.f90:296-~700 (own enumeration order)
//   q-mesh, irreducible set   type_qpointmesh_gridgeneration.f90:8-545, 612-680
//   little group              type_qpointmesh.f90:660-700
//   q3 = -q1-q2 index         lo_fftgrid_helper.f90:136-163
//   irreducible triplets      scattering_threephonon.f90:72-84, 319-448
//   orbit averaging sets      scattering_threephonon.f90:156-196, scattering.f90:312-354
#include "tc3_internal.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <tuple>

namespace tc3 {

namespace {

// gottochblandat.f90:1286-1374: the reference's "well defined" values, in list order
const double kWellDefined[69] = {
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

const double kDegenVector[3] = {1.0e-8, 1.33e-8, 1.45624623e-8};

void matmul3(const double *a, const double *b, double *c) {
    double t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
            t[3 * i + j] = s;
        }
    std::memcpy(c, t, sizeof(t));
}

void matvec3(const double *a, const double *v, double *w) {
    double t[3];
    for (int i = 0; i < 3; ++i) t[i] = a[3 * i] * v[0] + a[3 * i + 1] * v[1] + a[3 * i + 2] * v[2];
    w[0] = t[0]; w[1] = t[1]; w[2] = t[2];
}

bool inv3(const double *a, double *b) {
    double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                 a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (std::fabs(det) < 1e-300) return false;
    double id = 1.0 / det;
    double t[9];
    t[0] = (a[4] * a[8] - a[5] * a[7]) * id;
    t[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    t[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    t[3] = (a[5] * a[6] - a[3] * a[8]) * id;
    t[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    t[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    t[6] = (a[3] * a[7] - a[4] * a[6]) * id;
    t[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    t[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    std::memcpy(b, t, sizeof(t));
    return true;
}

double det3(const double *a) {
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
           a[2] * (a[3] * a[7] - a[4] * a[6]);
}

double clean_frac(double v) {
    // lo_clean_fractional_coordinates (gottochblandat.f90:1261-1283)
    const double tol = kSqTol;
    for (;;) {
        if (std::fabs(v) < tol) v = 0.0;
        if (std::fabs(v - 1.0) < tol) v = 0.0;
        if (v > 1.0) v -= 1.0;
        if (v < 0.0) v += 1.0;
        if (v > -tol && v < 1.0) break;
    }
    return v;
}

bool on_grid(const double *v, const int *dims) {
    double s = 0;
    for (int i = 0; i < 3; ++i) {
        double w = v[i] * dims[i] + 1.0;
        s += std::fabs(w - std::nearbyint(w));
    }
    return s < kTol;
}

int index_from_coordinate(const double *v, const int *dims) {
    int gi[3];
    for (int i = 0; i < 3; ++i) gi[i] = (int)std::nearbyint(clean_frac(v[i]) * dims[i] + 1.0);
    return (gi[0] - 1) * dims[1] * dims[2] + (gi[1] - 1) * dims[2] + (gi[2] - 1);
}

}  // namespace

double chop_scalar(double x, double tol) {
    for (int i = 0; i < 69; ++i)
        if (std::fabs(x - kWellDefined[i]) < tol) return kWellDefined[i];
    return x;
}

Status build_crystal(Crystal &c, int na, const double *latvec, const double *rfrac,
                     const int *species, const double *mass) {
    Status st;
    if (na <= 0) { st.code = 1; st.msg = "number of atoms must be positive"; return st; }
    c.na = na;
    double lat[9];
    for (int col = 0; col < 3; ++col)
        for (int r = 0; r < 3; ++r) lat[3 * r + col] = latvec[3 * col + r];
    double il[9];
    if (!inv3(lat, il)) { st.code = 5; st.msg = "singular lattice"; return st; }
    double rec[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) rec[3 * i + j] = il[3 * j + i];
    double ir[9];
    inv3(rec, ir);
    for (int i = 0; i < 9; ++i) {
        c.lat[i] = chop_scalar(lat[i], kSqTol);
        c.invlat[i] = chop_scalar(il[i], kSqTol);
        c.rec[i] = chop_scalar(rec[i], kSqTol);
        c.invrec[i] = chop_scalar(ir[i], kSqTol);
    }
    c.volume = std::fabs(det3(c.lat));
    c.rfrac.assign(3 * na, 0.0);
    c.rcart.assign(3 * na, 0.0);
    c.species.assign(species, species + na);
    c.mass.assign(mass, mass + na);
    c.invsqrtmass.resize(na);
    for (int a = 0; a < na; ++a) {
        if (!(mass[a] > 0)) { st.code = 3; st.msg = "masses must be positive"; return st; }
        c.invsqrtmass[a] = 1.0 / std::sqrt(mass[a]);
        for (int i = 0; i < 3; ++i) c.rfrac[3 * a + i] = chop_scalar(clean_frac(rfrac[3 * a + i]), kSqTol);
        double rc[3];
        matvec3(c.lat, &c.rfrac[3 * a], rc);
        for (int i = 0; i < 3; ++i) c.rcart[3 * a + i] = chop_scalar(rc[i], kSqTol);
    }
    return st;
}

Status spacegroup(const Crystal &c, std::vector<SymOp> &ops) {
    Status st;
    ops.clear();
    const double *A = c.lat;
    double G[9], At[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) At[3 * i + j] = A[3 * j + i];
    matmul3(At, A, G);
    double gsum = 0;
    for (double g : G) gsum += std::fabs(g);
    // lattice point group: integer matrices with entries in {-1,0,1} that keep the metric
    std::vector<std::array<double, 9>> Ws;
    for (int n = 0; n < 19683; ++n) {
        std::array<double, 9> W;
        int x = n;
        for (int k = 8; k >= 0; --k) { W[k] = (double)(x % 3) - 1.0; x /= 3; }
        double WT[9], t1[9], GW[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) WT[3 * i + j] = W[3 * j + i];
        matmul3(WT, G, t1);
        matmul3(t1, W.data(), GW);
        double d = 0;
        for (int k = 0; k < 9; ++k) d += std::fabs(GW[k] - G[k]);
        if (d < 1e-6 * gsum) Ws.push_back(W);
    }
    // identity first, the rest in enumeration order
    std::stable_partition(Ws.begin(), Ws.end(), [](const std::array<double, 9> &W) {
        return W[0] == 1 && W[4] == 1 && W[8] == 1 && W[1] == 0 && W[2] == 0 && W[3] == 0 &&
               W[5] == 0 && W[6] == 0 && W[7] == 0;
    });
    const int na = c.na;
    double Ainv[9];
    inv3(A, Ainv);
    for (const auto &W : Ws) {
        std::vector<double> rot(3 * na);
        for (int a = 0; a < na; ++a) matvec3(W.data(), &c.rfrac[3 * a], &rot[3 * a]);
        struct Found { std::array<double, 3> key; double t[3]; std::vector<int> fmap; };
        std::vector<Found> found;
        for (int b = 0; b < na; ++b) {
            if (c.species[b] != c.species[0]) continue;
            double t[3];
            for (int i = 0; i < 3; ++i) t[i] = clean_frac(c.rfrac[3 * b + i] - rot[i]);
            std::vector<int> fmap(na, -1);
            bool good = true;
            for (int a = 0; a < na && good; ++a) {
                double best = 1e300;
                int bj = -1;
                for (int j = 0; j < na; ++j) {
                    double d[3];
                    for (int i = 0; i < 3; ++i) {
                        d[i] = c.rfrac[3 * j + i] - (rot[3 * a + i] + t[i]);
                        d[i] -= std::nearbyint(d[i]);
                    }
                    double dc[3];
                    matvec3(A, d, dc);
                    double nn = std::sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
                    if (nn < best) { best = nn; bj = j; }
                }
                if (best > kTol || c.species[bj] != c.species[a]) good = false;
                else fmap[a] = bj;
            }
            if (!good) continue;
            Found f;
            for (int i = 0; i < 3; ++i) { f.key[i] = std::nearbyint(t[i] * 1e8) / 1e8; f.t[i] = t[i]; }
            f.fmap = fmap;
            found.push_back(f);
        }
        std::stable_sort(found.begin(), found.end(),
                         [](const Found &x, const Found &y) { return x.key < y.key; });
        for (const auto &f : found) {
            SymOp o;
            std::memcpy(o.fm, W.data(), sizeof(o.fm));
            double t1[9], mm[9];
            matmul3(A, W.data(), t1);
            matmul3(t1, Ainv, mm);
            for (int k = 0; k < 9; ++k) o.m[k] = chop_scalar(mm[k], 1e-10);
            double t2[9], rf[9];
            matmul3(c.invrec, o.m, t2);
            matmul3(t2, c.rec, rf);
            for (int k = 0; k < 9; ++k) o.rfm[k] = chop_scalar(rf[k], 1e-10);
            for (int i = 0; i < 3; ++i) o.ftr[i] = f.t[i];
            double tc[3];
            matvec3(A, f.t, tc);
            for (int i = 0; i < 3; ++i) o.tr[i] = chop_scalar(tc[i], 1e-11);
            o.fmap = f.fmap;
            ops.push_back(o);
        }
    }
    if (ops.empty()) { st.code = 4; st.msg = "no symmetry operations found"; }
    return st;
}

int third_grid_index(int i1, int i2, const int *dims) {
    const int n2 = dims[1], n3 = dims[2];
    int g1[3] = {i1 / (n2 * n3), (i1 / n3) % n2, i1 % n3};
    int g2[3] = {i2 / (n2 * n3), (i2 / n3) % n2, i2 % n3};
    int g3[3];
    for (int l = 0; l < 3; ++l) {
        int v = -g1[l] - g2[l];
        v %= dims[l];
        if (v < 0) v += dims[l];
        g3[l] = v;
    }
    return g3[0] * n2 * n3 + g3[1] * n3 + g3[2];
}

namespace {

void bz_fold(const Crystal &c, const double *v0, double *out) {
    // type_qpointmesh_gridgeneration.f90:115-133: subtract the reciprocal lattice vector
    // nearest to q + lo_degenvector.  The reference's inscribed-sphere shortcut is left
    // out: it only decides between two zone-boundary images by rounding.
    const double *B = c.rec;
    double q[3];
    matvec3(B, v0, q);
    double v1[3] = {q[0] + kDegenVector[0], q[1] + kDegenVector[1], q[2] + kDegenVector[2]};
    double best = 1e300;
    int bn[3] = {0, 0, 0};
    for (int i = -2; i <= 2; ++i)
        for (int j = -2; j <= 2; ++j)
            for (int k = -2; k <= 2; ++k) {
                double n[3] = {(double)i, (double)j, (double)k}, g[3];
                matvec3(B, n, g);
                double d = std::sqrt((v1[0] - g[0]) * (v1[0] - g[0]) + (v1[1] - g[1]) * (v1[1] - g[1]) +
                                     (v1[2] - g[2]) * (v1[2] - g[2]));
                if (d < best) { best = d; bn[0] = i; bn[1] = j; bn[2] = k; }
            }
    for (int i = 0; i < 3; ++i) out[i] = v0[i] - bn[i];
}

}  // namespace

Status build_mesh(const Crystal &c, const std::vector<SymOp> &ops, const int *dims,
                  bool timereversal, Mesh &m) {
    Status st;
    for (int i = 0; i < 3; ++i) {
        if (dims[i] <= 0) { st.code = 5; st.msg = "q-grid dimensions must be positive"; return st; }
        m.dims[i] = dims[i];
    }
    const int n1 = dims[0], n2 = dims[1], n3 = dims[2];
    const int np = n1 * n2 * n3;
    const int nsym = (int)ops.size();
    m.nfull = np;
    m.timereversal = timereversal;
    m.nops = nsym;
    std::vector<double> pts(3 * np);
    for (int i = 0, l = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j)
            for (int k = 0; k < n3; ++k, ++l) {
                double v0[3] = {(double)i / n1, (double)j / n2, (double)k / n3};
                bz_fold(c, v0, &pts[3 * l]);
            }
    std::vector<std::array<double, 9>> irfm(nsym);
    for (int o = 0; o < nsym; ++o) inv3(ops[o].rfm, irfm[o].data());
    const int ncol = timereversal ? 2 * nsym : nsym;
    std::vector<int> conn((size_t)np * ncol, 0);
    for (int i = 0; i < np; ++i) {
        const double *v0 = &pts[3 * i];
        for (int j = 0; j < nsym; ++j) {
            for (int tr = 0; tr < (timereversal ? 2 : 1); ++tr) {
                double s = tr ? -1.0 : 1.0;
                double v1[3];
                matvec3(ops[j].rfm, v0, v1);
                for (double &x : v1) x *= s;
                if (!on_grid(v1, dims)) continue;
                int ii = index_from_coordinate(v1, dims);
                double v2[3];
                matvec3(irfm[j].data(), &pts[3 * ii], v2);
                for (double &x : v2) x *= s;
                if (!on_grid(v2, dims)) continue;
                int jj = index_from_coordinate(v2, dims);
                if (jj != i) continue;
                if (tr) conn[(size_t)i * ncol + j + nsym] = -(ii + 1);
                else conn[(size_t)i * ncol + j] = ii + 1;
            }
        }
    }
    std::vector<char> ind(np, 1);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < ncol; ++j) {
            int k = std::abs(conn[(size_t)i * ncol + j]) - 1;
            if (k > i) ind[k] = 0;
        }
    m.irr_full_index.clear();
    for (int i = 0; i < np; ++i)
        if (ind[i]) m.irr_full_index.push_back(i);
    m.nirr = (int)m.irr_full_index.size();
    m.irr_of_full.assign(np, -1);
    m.op_of_full.assign(np, 0);
    for (int k = 0; k < m.nirr; ++k) {
        int qi = m.irr_full_index[k];
        for (int j = 0; j < nsym; ++j) {
            int i = conn[(size_t)qi * ncol + j] - 1;
            if (i < 0 || m.irr_of_full[i] >= 0) continue;
            m.irr_of_full[i] = k;
            m.op_of_full[i] = j + 1;
        }
        if (timereversal)
            for (int j = 0; j < nsym; ++j) {
                int i = -conn[(size_t)qi * ncol + j + nsym] - 1;
                if (i < 0 || m.irr_of_full[i] >= 0) continue;
                m.irr_of_full[i] = k;
                m.op_of_full[i] = -(j + 1);
            }
    }
    for (int i = 0; i < np; ++i)
        if (m.irr_of_full[i] < 0) { st.code = 4; st.msg = "failed mapping the full grid to the irreducible points"; return st; }
    m.weight_irr.assign(m.nirr, 0.0);
    for (int i = 0; i < np; ++i) m.weight_irr[m.irr_of_full[i]] += 1.0;
    for (double &w : m.weight_irr) w /= (double)np;
    m.qcart.resize(3 * np);
    for (int i = 0; i < np; ++i) {
        double q[3];
        matvec3(c.rec, &pts[3 * i], q);
        for (int k = 0; k < 3; ++k) m.qcart[3 * i + k] = chop_scalar(q[k], 1e-14);
    }
    // little group of each irreducible point (type_qpointmesh.f90:660-700)
    m.littlegroup.assign(m.nirr, {});
    for (int k = 0; k < m.nirr; ++k) {
        const double *q = &m.qcart[3 * m.irr_full_index[k]];
        for (int o = 0; o < nsym; ++o) {
            double rq[3], d[3], f[3];
            matvec3(ops[o].m, q, rq);
            for (int i = 0; i < 3; ++i) d[i] = q[i] - rq[i];
            matvec3(c.invrec, d, f);
            double s = 0;
            for (int i = 0; i < 3; ++i) { f[i] -= std::nearbyint(f[i]); s += f[i] * f[i]; }
            if (s < kSqTol) m.littlegroup[k].push_back(o);
        }
    }
    // rotation table on the grid (reciprocal fractional operation, no translation)
    m.rot.assign((size_t)nsym * np, -1);
    std::vector<double> frac(3 * np);
    for (int i = 0; i < np; ++i) matvec3(c.invrec, &m.qcart[3 * i], &frac[3 * i]);
    for (int o = 0; o < nsym; ++o)
        for (int i = 0; i < np; ++i) {
            double v1[3];
            matvec3(ops[o].rfm, &frac[3 * i], v1);
            if (on_grid(v1, dims)) m.rot[(size_t)o * np + i] = index_from_coordinate(v1, dims);
        }
    // star of every irreducible point (index_full_point / operation_full_point)
    m.star_ops.assign(m.nirr, {});
    for (int k = 0; k < m.nirr; ++k) {
        const double *v0 = &frac[3 * m.irr_full_index[k]];
        std::vector<int> seen;
        for (int tr = 0; tr < (timereversal ? 2 : 1); ++tr)
            for (int o = 0; o < nsym; ++o) {
                double v1[3];
                matvec3(ops[o].rfm, v0, v1);
                if (tr) for (double &x : v1) x = -x;
                if (!on_grid(v1, dims)) continue;
                int jj = index_from_coordinate(v1, dims);
                if (std::find(seen.begin(), seen.end(), jj) != seen.end()) continue;
                seen.push_back(jj);
                m.star_ops[k].push_back(tr ? -(o + 1) : (o + 1));
            }
    }
    return st;
}

Status build_triplets(const Mesh &m, TripletSet &t) {
    Status st;
    const int np = m.nfull;
    t = TripletSet();
    t.eq_off.push_back(0);
    t.q1_off.push_back(0);
    std::vector<std::pair<int, int>> newqp;
    for (int q1 = 0; q1 < m.nirr; ++q1) {
        const auto &lg = m.littlegroup[q1];
        const int q1f = m.irr_full_index[q1];
        for (int q2 = 0; q2 < np; ++q2) {
            const int q3 = third_grid_index(q1f, q2, m.dims);
            if (q3 < q2) continue;
            newqp.clear();
            bool isred = false;
            for (int k : lg) {
                int a = m.rot[(size_t)k * np + q2];
                int b = (a < 0) ? -1 : m.rot[(size_t)k * np + q3];
                if (a < 0 || b < 0) { newqp.push_back({-1, -1}); continue; }
                if (std::min(a, b) > q2) { isred = true; break; }
                newqp.push_back({a, b});
            }
            if (isred) continue;
            for (auto &x : newqp)
                if (x.first < 0) x = {q2, q3};
            const int n = (int)newqp.size();
            for (int j = 0; j < n; ++j) {
                auto sj = std::minmax(newqp[j].first, newqp[j].second);
                bool later = false;
                for (int k = j + 1; k < n && !later; ++k) {
                    auto sk = std::minmax(newqp[k].first, newqp[k].second);
                    if (sj == sk) later = true;
                }
                if (!later) { t.eq_q2.push_back(newqp[j].first); t.eq_q3.push_back(newqp[j].second); }
            }
            t.q1.push_back(q1);
            t.q2.push_back(q2);
            t.q3.push_back(q3);
            t.perm.push_back(q2 == q3 ? 1.0 : 2.0);
            t.eq_off.push_back((int)t.eq_q2.size());
        }
        t.q1_off.push_back((int)t.q1.size());
    }
    // The off-diagonal accumulation writes every (q', band) target of a row from exactly
    // one triplet; check that no grid point is shared by two triplets of the same q1.
    std::vector<int> owner(np, -1);
    for (int q1 = 0; q1 < m.nirr; ++q1) {
        for (int it = t.q1_off[q1]; it < t.q1_off[q1 + 1]; ++it)
            for (int e = t.eq_off[it]; e < t.eq_off[it + 1]; ++e)
                for (int x : {t.eq_q2[e], t.eq_q3[e]}) {
                    if (owner[x] >= 0 && owner[x] != it) {
                        st.code = 4;
                        st.msg = "q-grid does not respect the crystal symmetry (overlapping triplet orbits)";
                        return st;
                    }
                    owner[x] = it;
                }
        for (int it = t.q1_off[q1]; it < t.q1_off[q1 + 1]; ++it)
            for (int e = t.eq_off[it]; e < t.eq_off[it + 1]; ++e) {
                owner[t.eq_q2[e]] = -1;
                owner[t.eq_q3[e]] = -1;
            }
    }
    return st;
}

Status build_orbits(const Mesh &m, OrbitSet &o) {
    Status st;
    const int np = m.nfull;
    o = OrbitSet();
    o.q1_off.push_back(0);
    o.orb_off.push_back(0);
    std::vector<int> imgs, multi, multi_len;
    for (int q1 = 0; q1 < m.nirr; ++q1) {
        const auto &lg = m.littlegroup[q1];
        multi.clear();
        multi_len.clear();
        int ns = 0;
        // orbits of one member first, then the others, each in the order of their first member
        for (int q2 = 0; q2 < np; ++q2) {
            imgs.clear();
            bool skip = false;
            for (int k : lg) {
                int x = m.rot[(size_t)k * np + q2];
                if (x < 0) continue;
                if (x < q2) { skip = true; break; }
                imgs.push_back(x);
            }
            if (skip || imgs.empty()) continue;
            std::sort(imgs.begin(), imgs.end());
            imgs.erase(std::unique(imgs.begin(), imgs.end()), imgs.end());
            if (imgs.size() == 1) {
                o.members.push_back(imgs[0]);
                o.orb_off.push_back((int)o.members.size());
                ++ns;
            } else {
                multi.insert(multi.end(), imgs.begin(), imgs.end());
                multi_len.push_back((int)imgs.size());
            }
        }
        size_t p = 0;
        for (int n : multi_len) {
            o.members.insert(o.members.end(), multi.begin() + p, multi.begin() + p + n);
            o.orb_off.push_back((int)o.members.size());
            p += n;
        }
        o.q1_off.push_back((int)o.orb_off.size() - 1);
        o.nsingle.push_back(ns);
    }
    return st;
}

}  // namespace tc3

namespace tc3 {

void MersenneTwister64::init_int(uint64_t seed) {
    mt[0] = seed;
    for (int i = 1; i < 312; ++i) mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
    mti = 312;
}

void MersenneTwister64::init_tdep(int iseed, double rseed) {
    // lo_randomnumbers.f90:216-227
    double f0 = rseed - std::floor(rseed) - 0.5;
    if (iseed > 0) f0 = std::fmod(f0 + std::log((double)iseed), 1.0) - 0.5;
    else if (iseed < 0) f0 = std::fmod(f0 - std::log((double)-iseed), 1.0) - 0.5;
    const int64_t s = (int64_t)std::round(f0 * (double)(1 << 30));
    init_int((uint64_t)s);
}

uint64_t MersenneTwister64::next() {
    constexpr int NN = 312, MM = 156;
    constexpr uint64_t MATRIX_A = 0xB5026F5AA96619E9ULL, UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    if (mti >= NN) {
        auto mag = [](uint64_t x) { return (x & 1ULL) ? MATRIX_A : 0ULL; };
        int i = 0;
        for (; i < NN - MM; ++i) {
            const uint64_t x = (mt[i] & UM) | (mt[i + 1] & LM);
            mt[i] = mt[i + MM] ^ (x >> 1) ^ mag(x);
        }
        for (; i < NN - 1; ++i) {
            const uint64_t x = (mt[i] & UM) | (mt[i + 1] & LM);
            mt[i] = mt[i + (MM - NN)] ^ (x >> 1) ^ mag(x);
        }
        const uint64_t x = (mt[NN - 1] & UM) | (mt[0] & LM);
        mt[NN - 1] = mt[MM - 1] ^ (x >> 1) ^ mag(x);
        mti = 0;
    }
    uint64_t x = mt[mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

double MersenneTwister64::real1() { return (double)(next() >> 11) * (1.0 / 9007199254740991.0); }

Status montecarlo_triplets(const Mesh &m, const TripletSet &full, const std::vector<double> &om_irr, int nb,
                           const int *mc_dims_in, int seed, TripletSet &sub, std::vector<float> &w, long &npoints,
                           const int *mc4_dims, std::vector<long> *g4, long *npoints4, bool draw3) {
    Status st;
    int mcd[3], mcd4[3];
    double ratio[3], ratio4[3];
    for (int i = 0; i < 3; ++i) {
        // a grid dimension outside 1..qgrid means the full mesh (options.f90:226-236)
        auto dim = [&](int d) { return (d < 1 || d > m.dims[i]) ? m.dims[i] : d; };
        mcd[i] = dim(mc_dims_in[i]);
        ratio[i] = (double)m.dims[i] / (double)mcd[i];
        mcd4[i] = mc4_dims ? dim(mc4_dims[i]) : 1;
        ratio4[i] = (double)m.dims[i] / (double)mcd4[i];
    }
    npoints = (long)mcd[0] * mcd[1] * mcd[2];
    const long np4 = (long)mcd4[0] * mcd4[1] * mcd4[2];
    if (mc4_dims) {
        *npoints4 = np4;
        g4->assign((size_t)m.nirr * nb * 2 * np4, 0);
    }
    // scattering.f90:107-115, one MPI rank: integer seed 0, real seed 1 / seed
    MersenneTwister64 rng;
    rng.init_tdep(0, 1.0 / (double)seed);
    // mc_point_to_full (lo_fftgrid_helper.f90:56-73), 0-based point indices
    auto point_to_full_d = [&](long imc, const int *md, const double *rt) {
        const int k = (int)(imc % md[2]), j = (int)((imc / md[2]) % md[1]), i = (int)(imc / ((long)md[1] * md[2]));
        const int gm[3] = {i, j, k};
        int gf[3];
        for (int d = 0; d < 3; ++d) {
            const double r = rng.real1();
            // 1-based ceiling((g - 1) ratio + r ratio); a draw of exactly 0 (probability 2^-53)
            // is kept on its own cell
            const int g = (int)std::ceil((double)gm[d] * rt[d] + r * rt[d]);
            gf[d] = std::max(1, std::min(g, m.dims[d])) - 1;
        }
        return ((long)gf[0] * m.dims[1] + gf[1]) * m.dims[2] + gf[2];
    };
    auto point_to_full = [&](long imc) { return point_to_full_d(imc, mcd, ratio); };
    // generate_grid (lo_fftgrid_helper.f90:75-99) of the four-phonon grids
    auto draw4 = [&](long *out) {
        out[0] = point_to_full_d(0, mcd4, ratio4);
        long prev = out[0], test = out[0];
        for (long qi = 1; qi < np4; ++qi) {
            while (test == prev) test = point_to_full_d(qi, mcd4, ratio4);
            out[qi] = test;
            prev = test;
        }
    };
    const long np = m.nfull;
    std::vector<long> tof(np);
    std::vector<float> wfull;
    wfull.assign(full.q1.size() * (size_t)nb, 0.0f);
    std::vector<long> grid(npoints);
    for (int q1 = 0; q1 < m.nirr; ++q1) {
        std::fill(tof.begin(), tof.end(), -1);
        for (int t = full.q1_off[q1]; t < full.q1_off[q1 + 1]; ++t) tof[full.q2[t]] = t;
        for (int b1 = 0; b1 < nb; ++b1) {
            if (om_irr[(size_t)q1 * nb + b1] < kFreqTol) continue;
            // generate_grid (:75-99): no point equal to the previous one; without the
            // three-phonon term (--nothirdorder) its grid is not drawn (scattering.f90:119-121, 234)
            if (draw3) {
                grid[0] = point_to_full(0);
                long prev = grid[0], test = grid[0];
                for (long qi = 1; qi < npoints; ++qi) {
                    while (test == prev) test = point_to_full(qi);
                    grid[qi] = test;
                    prev = test;
                }
                for (long qi = 0; qi < npoints; ++qi) {
                    const long t = tof[grid[qi]];
                    if (t >= 0) wfull[(size_t)t * nb + b1] += 1.0f;
                }
            }
            // four-phonon grids of the row (scattering_fourphonon.f90:84-85)
            if (mc4_dims) {
                long *o = &(*g4)[((size_t)q1 * nb + b1) * 2 * np4];
                draw4(o);
                draw4(o + np4);
            }
        }
    }
    // triplets drawn for any row, in the order of the full list
    sub = TripletSet();
    sub.eq_off.push_back(0);
    sub.q1_off.push_back(0);
    w.clear();
    for (int q1 = 0; q1 < m.nirr; ++q1) {
        for (int t = full.q1_off[q1]; t < full.q1_off[q1 + 1]; ++t) {
            bool any = false;
            for (int b = 0; b < nb && !any; ++b) any = wfull[(size_t)t * nb + b] > 0.0f;
            if (!any) continue;
            sub.q1.push_back(full.q1[t]);
            sub.q2.push_back(full.q2[t]);
            sub.q3.push_back(full.q3[t]);
            sub.perm.push_back(full.perm[t]);
            for (int e = full.eq_off[t]; e < full.eq_off[t + 1]; ++e) {
                sub.eq_q2.push_back(full.eq_q2[e]);
                sub.eq_q3.push_back(full.eq_q3[e]);
            }
            sub.eq_off.push_back((int)sub.eq_q2.size());
            w.insert(w.end(), wfull.begin() + (size_t)t * nb, wfull.begin() + (size_t)(t + 1) * nb);
        }
        sub.q1_off.push_back((int)sub.q1.size());
    }
    return st;
}

}  // namespace tc3
