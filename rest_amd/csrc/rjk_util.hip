// rjk_util.hip — layout and housekeeping kernels: synthetic B fill, packing
// of full B columns (REST ri3fn) into the packed pair layout, padding of
// compact packed columns to the device leading dimension, FP32 copies,
// symmetrisation of D, a plain copy (counter calibration), a finite check and
// the helpers of the density-only K on a general D.
#include <algorithm>

#include "rjk_internal.hpp"

namespace rjk {
namespace {

__device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ inline void unpair(size_t p, size_t* i, size_t* j) {
    size_t jj = (size_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while (jj * (jj + 1) / 2 > p) --jj;
    while ((jj + 1) * (jj + 2) / 2 <= p) ++jj;
    *j = jj;
    *i = p - jj * (jj + 1) / 2;
}

// b[P][p] = (2u-1)·sqrt(3/naux), u from the hash of the compact flat index;
// padding elements (p >= npair) are zero.
__global__ void k_fill_synth(uint64_t seed, size_t npair, size_t naux, size_t ldb, double* __restrict__ b) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = naux * ldb;
    const double a = sqrt(3.0 / (double)naux);
    for (; t < total; t += (size_t)gridDim.x * blockDim.x) {
        size_t P = t / ldb, p = t % ldb;
        double v = 0.0;
        if (p < npair) {
            uint64_t z = mix64(seed + (P * npair + p + 1) * 0x9E3779B97F4A7C15ULL);
            double u = (double)(z >> 11) * 0x1.0p-53;
            v = (2.0 * u - 1.0) * a;
        }
        b[t] = v;
    }
}

// full column-major B_P (nbf×nbf) -> packed upper (row <= col), padded ld
__global__ void k_pack_ri3fn(const double* __restrict__ b3, size_t nbf, size_t nP, size_t ldb,
                             double* __restrict__ b) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = nP * ldb;
    size_t npair = nbf * (nbf + 1) / 2;
    for (; t < total; t += (size_t)gridDim.x * blockDim.x) {
        size_t P = t / ldb, p = t % ldb;
        double v = 0.0;
        if (p < npair) {
            size_t i, j;
            unpair(p, &i, &j);
            v = b3[P * nbf * nbf + i + j * nbf];
        }
        b[t] = v;
    }
}

__global__ void k_pad_rimatr(const double* __restrict__ src, size_t npair, size_t nP, size_t ldb,
                             double* __restrict__ b) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t total = nP * ldb;
    for (; t < total; t += (size_t)gridDim.x * blockDim.x) {
        size_t P = t / ldb, p = t % ldb;
        b[t] = (p < npair) ? src[P * npair + p] : 0.0;
    }
}

__global__ void k_to_fp32(const double* __restrict__ b, size_t n, float* __restrict__ out) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < n; t += (size_t)gridDim.x * blockDim.x) out[t] = (float)b[t];
}

__global__ void k_symmetrize(const double* __restrict__ dm, size_t nbf, double* __restrict__ out) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbf * nbf) return;
    size_t i = t % nbf, j = t / nbf;
    out[t] = 0.5 * (dm[i + j * nbf] + dm[j + i * nbf]);
}

__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < n16; t += (size_t)gridDim.x * blockDim.x) dst[t] = src[t];
}
__global__ void k_copy1(const unsigned char* __restrict__ src, unsigned char* __restrict__ dst, size_t n) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

__global__ void k_check_finite(const double* __restrict__ x, size_t n, int* __restrict__ flag) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    for (; t < n; t += (size_t)gridDim.x * blockDim.x)
        if (!isfinite(x[t])) bad = 1;
    if (bad) *flag = 1;   // benign race: every writer stores 1
}

// Per-block maxima of |D(i,j) − D(j,i)| and |D(i,j)| over a column-major nbf×nbf D:
// out[2·block] and out[2·block+1]; a non-finite entry makes both +inf.
__global__ __launch_bounds__(256) void k_asym_stats(const double* __restrict__ dm, size_t nbf,
                                                    double* __restrict__ out) {
    const size_t n = nbf * nbf;
    double a = 0.0, m = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t i = t % nbf, j = t / nbf;
        const double x = dm[t], y = dm[i * nbf + j];
        if (!isfinite(x)) { a = INFINITY; m = INFINITY; continue; }
        a = fmax(a, fabs(x - y));
        m = fmax(m, fabs(x));
    }
    __shared__ double ra[256], rm[256];
    ra[threadIdx.x] = a;
    rm[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            ra[threadIdx.x] = fmax(ra[threadIdx.x], ra[threadIdx.x + h]);
            rm[threadIdx.x] = fmax(rm[threadIdx.x], rm[threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = ra[0];
        out[2 * blockIdx.x + 1] = rm[0];
    }
}

// Column sets of the general density-only route, row-major [nrow][nk] like C':
// cx[ν][c] = D(ν, c) (column c of the column-major D), cy[ν][c] = δ_νc, both zero
// outside nbf × nbf.
__global__ void k_prep_general(const double* __restrict__ dm, size_t nbf, size_t nk, size_t nrow,
                               double* __restrict__ cx, double* __restrict__ cy) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrow * nk) return;
    const size_t c = t % nk, nu = t / nk;
    const bool in = nu < nbf && c < nbf;
    cx[t] = in ? dm[nu + c * nbf] : 0.0;
    cy[t] = (in && nu == c) ? 1.0 : 0.0;
}

// vk[p(i,j)] = scale · G(i, j) for i <= j, G column-major with leading dimension ldg.
__global__ void k_pack_upper(const double* __restrict__ g, size_t nbf, size_t ldg, double scale,
                             double* __restrict__ vk) {
    const size_t npair = nbf * (nbf + 1) / 2;
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npair) return;
    size_t i, j;
    unpair(p, &i, &j);
    vk[p] = scale * g[i + j * ldg];
}

unsigned grid_for(size_t n) {
    size_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    if (g == 0) g = 1;
    return (unsigned)g;
}

}  // namespace

hipError_t launch_fill_synth(hipStream_t st, uint64_t seed, size_t npair, size_t naux, size_t ldb, double* b) {
    hipLaunchKernelGGL(k_fill_synth, dim3(grid_for(naux * ldb)), dim3(256), 0, st, seed, npair, naux, ldb, b);
    return hipGetLastError();
}
hipError_t launch_pack_ri3fn(hipStream_t st, const double* b3, size_t nbf, size_t nP, size_t ldb, double* b) {
    hipLaunchKernelGGL(k_pack_ri3fn, dim3(grid_for(nP * ldb)), dim3(256), 0, st, b3, nbf, nP, ldb, b);
    return hipGetLastError();
}
hipError_t launch_pad_rimatr(hipStream_t st, const double* src, size_t npair, size_t nP, size_t ldb, double* b) {
    hipLaunchKernelGGL(k_pad_rimatr, dim3(grid_for(nP * ldb)), dim3(256), 0, st, src, npair, nP, ldb, b);
    return hipGetLastError();
}
hipError_t launch_to_fp32(hipStream_t st, const double* b, size_t n, float* out) {
    hipLaunchKernelGGL(k_to_fp32, dim3(grid_for(n)), dim3(256), 0, st, b, n, out);
    return hipGetLastError();
}
hipError_t launch_symmetrize(hipStream_t st, const double* dm, size_t nbf, double* out) {
    size_t n = nbf * nbf;
    hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dm, nbf, out);
    return hipGetLastError();
}
hipError_t launch_copy(hipStream_t st, void* dst, const void* src, size_t bytes) {
    size_t n16 = bytes / 16;
    if (n16)
        hipLaunchKernelGGL(k_copy16, dim3(grid_for(n16)), dim3(256), 0, st, (const uint4*)src, (uint4*)dst, n16);
    size_t rem = bytes - n16 * 16;
    if (rem)
        hipLaunchKernelGGL(k_copy1, dim3(1), dim3(256), 0, st, (const unsigned char*)src + n16 * 16,
                           (unsigned char*)dst + n16 * 16, rem);
    return hipGetLastError();
}
unsigned asym_stats_blocks(size_t nbf) {
    size_t g = (nbf * nbf + 255) / 256;
    return (unsigned)std::min<size_t>(std::max<size_t>(g, 1), kAsymBlocks);
}
hipError_t launch_asym_stats(hipStream_t st, const double* dm, size_t nbf, double* out) {
    hipLaunchKernelGGL(k_asym_stats, dim3(asym_stats_blocks(nbf)), dim3(256), 0, st, dm, nbf, out);
    return hipGetLastError();
}
hipError_t launch_prep_general(hipStream_t st, const double* dm, size_t nbf, size_t nk, size_t nrow, double* cx,
                               double* cy) {
    size_t n = nrow * nk;
    hipLaunchKernelGGL(k_prep_general, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dm, nbf, nk, nrow, cx,
                       cy);
    return hipGetLastError();
}
hipError_t launch_pack_upper(hipStream_t st, const double* g, size_t nbf, size_t ldg, double scale, double* vk) {
    size_t n = nbf * (nbf + 1) / 2;
    hipLaunchKernelGGL(k_pack_upper, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, nbf, ldg, scale, vk);
    return hipGetLastError();
}
hipError_t launch_check_finite(hipStream_t st, const double* x, size_t n, int* flag) {
    hipLaunchKernelGGL(k_check_finite, dim3(grid_for(n)), dim3(256), 0, st, x, n, flag);
    return hipGetLastError();
}

}  // namespace rjk
