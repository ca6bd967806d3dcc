// serve_kernels.hip -- small device helpers of the serving loop: operand generation,
// zero fill and result checksums.
#include "internal.hpp"

namespace cfgmm {
namespace {

__device__ inline unsigned long long splitmix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// x[i] = uniform[-1, 1) from the top 53 bits of splitmix64(seed * 2^32 + i)
__global__ void fill_uniform_kernel(double *__restrict__ x, long n, unsigned long long seed) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const unsigned long long r = splitmix64((seed << 32) + (unsigned long long)i);
        x[i] = (double)(r >> 11) * 0x1.0p-52 - 1.0;
    }
}

__global__ void fill_zero_kernel(unsigned *__restrict__ x, long nwords) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (long)gridDim.x * blockDim.x) x[i] = 0u;
}

// out += sum_i splitmix64(bits(x_i) ^ splitmix64(i)) (mod 2^64) over the rows x cols matrix
// x[r*ld + c], i = r*cols + c: order-independent, so the value depends only on the bits of
// the result in compact row-major order.
__global__ void checksum_kernel(const double *__restrict__ x, long rows, long cols, long ld,
                                unsigned long long *__restrict__ out) {
    unsigned long long h = 0;
    const long n = rows * cols;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        h += splitmix64((unsigned long long)__double_as_longlong(x[(i / cols) * ld + i % cols]) ^
                        splitmix64((unsigned long long)i));
    for (int off = 32; off > 0; off >>= 1) h += __shfl_xor(h, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, h);
}

unsigned grid_for(long n) {
    const long g = (n + 255) / 256;
    return (unsigned)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}

}  // namespace

int launch_fill_uniform(hipStream_t stream, double *x, size_t n, unsigned long long seed) {
    if (n == 0) return CFGMM_OK;
    hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for((long)n)), dim3(256), 0, stream, x, (long)n, seed);
    return hip_status(hipGetLastError(), "fill_uniform_kernel launch");
}

int launch_fill_zero(hipStream_t stream, void *x, size_t bytes) {
    if (bytes % 4 != 0) return fail(CFGMM_EINVAL, "fill_zero: size must be a multiple of 4 bytes");
    if (bytes == 0) return CFGMM_OK;
    const long nwords = (long)(bytes / 4);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for(nwords)), dim3(256), 0, stream, static_cast<unsigned *>(x),
                       nwords);
    return hip_status(hipGetLastError(), "fill_zero_kernel launch");
}

// *out = checksum of the rows x cols matrix x[r*ld + c] (out is cleared first, on the same stream).
int launch_checksum(hipStream_t stream, const double *x, size_t rows, size_t cols, size_t ld,
                    unsigned long long *out) {
    if (rows > 1 && ld < cols) return fail(CFGMM_EINVAL, "checksum: ld < cols");
    int st = launch_fill_zero(stream, out, sizeof(unsigned long long));
    const size_t n = rows * cols;
    if (st != CFGMM_OK || n == 0) return st;
    hipLaunchKernelGGL(checksum_kernel, dim3(grid_for((long)n)), dim3(256), 0, stream, x, (long)rows, (long)cols,
                       (long)ld, out);
    return hip_status(hipGetLastError(), "checksum_kernel launch");
}

}  // namespace cfgmm
