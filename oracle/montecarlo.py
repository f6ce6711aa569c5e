"""Monte-Carlo q2 grids of the three-phonon term (test infrastructure, see oracle/__init__.py).

lo_mersennetwister (lo_randomnumbers.f90:80-300): the 64-bit Mersenne twister of Matsumoto
and Nishimura (MT19937-64) with TDEP's seeding (init_genrand64 from a real and an integer
seed, :204-236) and genrand64_real1 on [0, 1]; lo_montecarlo_grid (lo_fftgrid_helper.f90:
30-99): a stratified draw of prod(mc_dims) points of the full grid, one point per cell of
the coarse grid, no point equal to the previous one.
"""
import math

import numpy as np

M64 = (1 << 64) - 1
NN, MM = 312, 156
MATRIX_A = 0xB5026F5AA96619E9
UM, LM = 0xFFFFFFFF80000000, 0x7FFFFFFF


class MersenneTwister64:
    def __init__(self):
        self.mt = [0] * NN
        self.mti = NN + 1

    def init_int(self, seed):
        """init_genrand64: seed as a 64-bit pattern (negative integers two's complement)."""
        mt = self.mt
        mt[0] = seed & M64
        for i in range(1, NN):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & M64
        self.mti = NN

    def init_tdep(self, iseed, rseed):
        """lo_randomnumbers.f90:216-227."""
        f0 = rseed - math.floor(rseed) - 0.5
        if iseed > 0:
            f0 = math.fmod(f0 + math.log(float(iseed)), 1.0) - 0.5
        elif iseed < 0:
            f0 = math.fmod(f0 - math.log(float(-iseed)), 1.0) - 0.5
        x = f0 * float(1 << 30)
        seed = int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)   # anint: half away from zero
        self.init_int(seed)

    def next(self):
        mt = self.mt
        if self.mti >= NN:
            for i in range(NN):
                x = (mt[i] & UM) | (mt[(i + 1) % NN] & LM)
                mt[i] = mt[(i + MM) % NN] ^ (x >> 1) ^ (MATRIX_A if x & 1 else 0)
            self.mti = 0
        x = mt[self.mti]
        self.mti += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000 & M64
        x ^= (x << 37) & 0xFFF7EEE000000000 & M64
        x ^= x >> 43
        return x & M64

    def real1(self):
        """genrand64_real1: [0, 1]."""
        return float(self.next() >> 11) * (1.0 / 9007199254740991.0)


def mc_dims_of(dims, qg3ph):
    """options.f90:226-230 and lo_montecarlo_grid%initialize (:30-52)."""
    return [dims[i] if (qg3ph[i] < 1 or qg3ph[i] > dims[i]) else qg3ph[i] for i in range(3)]


def row_grids(dims, rows, mc_dims, seed):
    """The q2 grid (0-based full-grid indices) of every row (q1, b1), in the order of
    `rows`, from one generator seeded as scattering.f90:107-115 with one MPI rank."""
    rng = MersenneTwister64()
    rng.init_tdep(0, 1.0 / float(seed))
    ratio = [dims[i] / mc_dims[i] for i in range(3)]
    npoints = mc_dims[0] * mc_dims[1] * mc_dims[2]

    def point_to_full(imc):
        gm = (imc // (mc_dims[1] * mc_dims[2]), (imc // mc_dims[2]) % mc_dims[1], imc % mc_dims[2])
        gf = []
        for d in range(3):
            g = math.ceil(gm[d] * ratio[d] + rng.real1() * ratio[d])
            gf.append(min(max(g, 1), dims[d]) - 1)
        return (gf[0] * dims[1] + gf[1]) * dims[2] + gf[2]

    out = {}
    for r in rows:
        grid = [point_to_full(0)]
        prev = test = grid[0]
        for qi in range(1, npoints):
            while test == prev:
                test = point_to_full(qi)
            grid.append(test)
            prev = test
        out[r] = np.array(grid, dtype=np.int64)
    return out, npoints


def row_grids_4ph(dims, rows, mc3, mc4, seed):
    """The grids of every row (q1, b1) when four-phonon scattering is on: per row the
    three-phonon q2 grid (drawn even on the full mesh, compute_threephonon_scattering
    :69-70), then the two four-phonon grids (scattering_fourphonon.f90:84-85), all from one
    generator in row order.  Returns {row: (g3, g4a, g4b)}, npoints3, npoints4."""
    rng = MersenneTwister64()
    rng.init_tdep(0, 1.0 / float(seed))

    def draw(mcd):
        ratio = [dims[i] / mcd[i] for i in range(3)]
        npts = mcd[0] * mcd[1] * mcd[2]

        def point_to_full(imc):
            gm = (imc // (mcd[1] * mcd[2]), (imc // mcd[2]) % mcd[1], imc % mcd[2])
            gf = []
            for d in range(3):
                g = math.ceil(gm[d] * ratio[d] + rng.real1() * ratio[d])
                gf.append(min(max(g, 1), dims[d]) - 1)
            return (gf[0] * dims[1] + gf[1]) * dims[2] + gf[2]

        grid = [point_to_full(0)]
        prev = test = grid[0]
        for qi in range(1, npts):
            while test == prev:
                test = point_to_full(qi)
            grid.append(test)
            prev = test
        return np.array(grid, dtype=np.int64)

    out = {}
    for r in rows:
        g3 = draw(mc3)
        out[r] = (g3, draw(mc4), draw(mc4))
    return out, mc3[0] * mc3[1] * mc3[2], mc4[0] * mc4[1] * mc4[2]
