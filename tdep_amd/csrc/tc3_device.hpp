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

// lo_chop_real for |x| small relative to the listed values: the device copy of the
// reference's well-defined value list (gottochblandat.f90:1286-1374).
__device__ double chop_dev(double x, double tol);

#define TC3_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return device_error(e_, #call, __FILE__, __LINE__);        \
    } while (0)

}  // namespace tc3
