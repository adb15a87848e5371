// bbpd_internal.hpp -- shared helpers of libbbpd_hip (error reporting, HIP checks, reductions).
#pragma once
#include <hip/hip_runtime.h>
#include <string>

#include "bbpd.h"

namespace bbpd {
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
} // namespace bbpd

#define BBPD_HIP(call)                                                                                             \
    do {                                                                                                           \
        hipError_t e_ = (call);                                                                                    \
        if (e_ != hipSuccess) return bbpd::hip_fail(e_, #call);                                                    \
    } while (0)

#define BBPD_REQUIRE(cond, msg)                                                                                    \
    do {                                                                                                           \
        if (!(cond)) {                                                                                             \
            bbpd::set_error(msg);                                                                                  \
            return BBPD_ERR_ARG;                                                                                   \
        }                                                                                                          \
    } while (0)

namespace bbpd {

// Fixed-order block sum of one double per thread (blockDim.x a power of two <= 1024).  The tree
// shape depends only on blockDim, so results are reproducible run to run.
template <int BLOCK>
__device__ inline double block_sum_fixed(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// Two block sums in one pass of the same fixed tree (each equals block_sum_fixed of its value);
// sh holds 2 * BLOCK doubles.
template <int BLOCK>
__device__ inline void block_sum_fixed2(double &a, double &b, double *sh)
{
    sh[threadIdx.x] = a;
    sh[BLOCK + threadIdx.x] = b;
    __syncthreads();
#pragma unroll
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
            sh[BLOCK + threadIdx.x] = sh[BLOCK + threadIdx.x] + sh[BLOCK + threadIdx.x + w];
        }
        __syncthreads();
    }
    a = sh[0];
    b = sh[BLOCK];
    __syncthreads();
}

// Three block sums in one pass of the same fixed tree (each equals block_sum_fixed of its value);
// sh holds 3 * BLOCK doubles.
template <int BLOCK>
__device__ inline void block_sum_fixed3(double &a, double &b, double &c, double *sh)
{
    sh[threadIdx.x] = a;
    sh[BLOCK + threadIdx.x] = b;
    sh[2 * BLOCK + threadIdx.x] = c;
    __syncthreads();
#pragma unroll
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
            sh[BLOCK + threadIdx.x] = sh[BLOCK + threadIdx.x] + sh[BLOCK + threadIdx.x + w];
            sh[2 * BLOCK + threadIdx.x] = sh[2 * BLOCK + threadIdx.x] + sh[2 * BLOCK + threadIdx.x + w];
        }
        __syncthreads();
    }
    a = sh[0];
    b = sh[BLOCK];
    c = sh[2 * BLOCK];
    __syncthreads();
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(long n, int block)
{
    long g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

} // namespace bbpd
