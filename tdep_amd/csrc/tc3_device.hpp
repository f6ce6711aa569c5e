// Device helpers shared by the tc3 HIP sources.
#pragma once

#include <hip/hip_runtime.h>

namespace tc3 {

struct cd { double x, y; };
struct cf { float x, y; };

__host__ __device__ inline cd cmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ inline cd cmulc(cd a, cd b) { return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x}; }  // conj(a)*b
__host__ __device__ inline cd cadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cd csub(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cd cscale(cd a, double s) { return {a.x * s, a.y * s}; }
__host__ __device__ inline cd cconj(cd a) { return {a.x, -a.y}; }
__host__ __device__ inline double cabs2(cd a) { return a.x * a.x + a.y * a.y; }

// Integration type 6 widths (get_dirac, scattering_threephonon.f90:251-261; the isotope term,
// scattering_isotope.f90:52-54) from the reduced group velocities u = v reclat / dims (three
// per mode): for each pair of modes sqrt(max_c (u_a - u_b)_c^2 / 2), the three pairs
// averaged, plus the small baseline thresh; the isotope term uses the second mode alone.
template <typename S>
__host__ __device__ inline S pair_width(const S *a, const S *b) {
    S m = 0;
    for (int c = 0; c < 3; ++c) {
        const S d = a[c] - b[c];
        m = m > d * d ? m : d * d;
    }
    return sqrt(m * (S)0.5);
}
template <typename S>
__host__ __device__ inline S sigma_type6(const S *u1, const S *u2, const S *u3, S thresh) {
    S s = 0;
    s += pair_width(u1, u2) / (S)3;
    s += pair_width(u1, u3) / (S)3;
    s += pair_width(u2, u3) / (S)3;
    return s + thresh;
}
template <typename S>
__host__ __device__ inline S sigma_type6_iso(const S *u2, S thresh) {
    S m = 0;
    for (int c = 0; c < 3; ++c) m = m > u2[c] * u2[c] ? m : u2[c] * u2[c];
    return sqrt(m * (S)0.5) + thresh;
}

// lo_chop_real for |x| small relative to the listed values: the device copy of the
// reference's well-defined value list (gottochblandat.f90:1286-1374).
__device__ double chop_dev(double x, double tol);

#define TC3_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return device_error(e_, #call, __FILE__, __LINE__);        \
    } while (0)

}  // namespace tc3
