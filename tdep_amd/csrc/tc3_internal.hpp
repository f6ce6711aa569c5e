// Internal state of a tc3 context and the host-side (CPU) preprocessing:
// space group, FFT q-mesh, little groups, irreducible triplets and orbits.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace tc3 {

// Tolerances and constants of src/libolle/konstanter.f90:26-135
constexpr double kPi = 3.141592653589793;
constexpr double kTwoPi = 6.283185307179586;
constexpr double kTol = 1e-5;
constexpr double kSqTol = kTol * kTol;
constexpr double kFreqTol = kTol * 1e-4;
constexpr double kGroupVelTol = kTol * 1e-5;   // lo_phonongroupveltol
constexpr double kTemperatureTol = 1e-3;
constexpr double kHartreeToEv = 27.21138602;
constexpr double kEvToHartree = 1.0 / kHartreeToEv;
constexpr double kKbHartree = 8.6173303e-5 * kEvToHartree;
constexpr double kEvToJoule = 1.6021766208e-19;
constexpr double kHartreeToJoule = kHartreeToEv * kEvToJoule;
constexpr double kBohrToM = 0.52917721067e-10;
constexpr double kHbarJoule = 6.582119514e-16 * kEvToJoule;
constexpr double kTimeAuToS = kHbarJoule / kHartreeToJoule;
constexpr double kGroupvelAuToMs = kBohrToM / kTimeAuToS;
constexpr double kHbarHartree = 6.582119514e-16 * kEvToHartree;
constexpr double kFreqHartreeToThz = 1e-12 / kTwoPi / kHbarHartree;
constexpr double kFreqThzToHartree = 1.0 / kFreqHartreeToThz;
constexpr double kKappaAuToSi = kHartreeToJoule / kTimeAuToS / kBohrToM;

struct Status {
    int code = 0;
    std::string msg;
};

struct SymOp {
    double fm[9];   // fractional rotation, row-major
    double m[9];    // Cartesian rotation, row-major
    double rfm[9];  // reciprocal fractional rotation, row-major
    double ftr[3];  // fractional translation
    double tr[3];   // Cartesian translation (Bohr)
    std::vector<int> fmap;  // fmap[a] = image of atom a
};

struct Crystal {
    int na = 0;
    double lat[9];       // columns are lattice vectors, row-major storage lat[3*r + c]
    double invlat[9];
    double rec[9];       // reciprocal vectors (no 2 pi) as columns
    double invrec[9];
    double volume = 0;
    std::vector<double> rfrac;   // 3*na
    std::vector<double> rcart;   // 3*na
    std::vector<int> species;
    std::vector<double> mass, invsqrtmass;
};

struct Mesh {
    int dims[3] = {0, 0, 0};
    int nfull = 0, nirr = 0;
    bool timereversal = true;
    int nops = 0;                      // operations the mesh was reduced with (1 with --nosym)
    std::vector<int> irr_full_index;   // nirr
    std::vector<int> irr_of_full;      // nfull
    std::vector<int> op_of_full;       // nfull, 1-based, negative = time reversal
    std::vector<double> weight_irr;    // nirr
    std::vector<double> qcart;         // 3*nfull
    std::vector<std::vector<int>> littlegroup;  // per irreducible point, 0-based ops
    std::vector<std::vector<int>> star_ops;     // per irreducible point, signed 1-based
    std::vector<int> rot;              // nsym*nfull: grid index of R_o q, -1 off grid
};

// One visited (q1, q2, q3) triplet with its list of equivalent (q2', q3').
struct TripletSet {
    std::vector<int> q1, q2, q3;     // q1 irreducible index, q2/q3 full indices
    std::vector<double> perm;
    std::vector<int> eq_off;         // ntrip+1 offsets into eq_q2/eq_q3
    std::vector<int> eq_q2, eq_q3;
    std::vector<int> q1_off;         // nirr+1 offsets: triplets of q1 are contiguous
};

// Orbits of the little group of each irreducible q1 acting on the full grid.
struct OrbitSet {
    std::vector<int> q1_off;         // nirr+1 offsets into orbit list
    std::vector<int> orb_off;        // norbit+1 offsets into members
    std::vector<int> members;
    std::vector<int> nsingle;        // nirr: orbits of one member, the first of every q1
};

// Host-side preprocessing (tc3_host.cpp)
double chop_scalar(double x, double tol);
Status spacegroup(const Crystal &c, std::vector<SymOp> &ops);
Status build_crystal(Crystal &c, int na, const double *latvec_colmajor, const double *rfrac,
                     const int *species, const double *mass);
Status build_mesh(const Crystal &c, const std::vector<SymOp> &ops, const int *dims,
                  bool timereversal, Mesh &m);
int third_grid_index(int i1, int i2, const int *dims);
Status build_triplets(const Mesh &m, TripletSet &t);

// 64-bit Mersenne twister (MT19937-64, Matsumoto and Nishimura) with the seeding of
// lo_mersennetwister%init (lo_randomnumbers.f90:204-236): a real seed and an integer seed
// (the MPI rank) give the integer that starts init_genrand64.
struct MersenneTwister64 {
    uint64_t mt[312];
    int mti = 313;
    void init_int(uint64_t seed);
    void init_tdep(int iseed, double rseed);
    uint64_t next();
    double real1();   // [0, 1]: (x >> 11) / (2^53 - 1), genrand64_real1
};

// Monte-Carlo q2 grids (lo_montecarlo_grid, lo_fftgrid_helper.f90:30-99): for every row (q1,
// b1) with omega >= freqtol, in (q1, b1) order, one stratified draw of prod(mc_dims) points
// of the full grid; the triplets of `full` whose first point is drawn form `sub`, and
// w[t * nb + b1] counts how often triplet t is drawn for row (q1, b1).  npoints = the
// number of points per draw (the weight of a draw is 1 / npoints).  With draw3 false (no
// three-phonon term) only the four-phonon grids are drawn and `sub` is empty.
Status montecarlo_triplets(const Mesh &m, const TripletSet &full, const std::vector<double> &om_irr, int nb,
                           const int *mc_dims, int seed, TripletSet &sub, std::vector<float> &w, long &npoints,
                           const int *mc4_dims = nullptr, std::vector<long> *g4 = nullptr, long *npoints4 = nullptr,
                           bool draw3 = true);
Status build_orbits(const Mesh &m, OrbitSet &o);

}  // namespace tc3
