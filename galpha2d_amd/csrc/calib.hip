// galpha_calib -- streaming kernels of known size for scaling the HBM counters of a profile.
//
// FETCH_SIZE and WRITE_SIZE count L2-to-memory traffic in units that differ from bytes by a
// factor that depends on the access kind.  This program moves a known number of bytes three
// ways, each in its own kernel: a read-only reduction over `bytes`, a write-only fill of `bytes`,
// and a copy (read `bytes`, write `bytes`).  Profiling it with the same counters gives the factors
// that turn the counters of the pipeline's kernels into bytes.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

// one partial per block; the buffer is read once, 16 bytes per lane and step
__global__ void k_calib_read(const int4 *__restrict__ a, size_t n, int *__restrict__ part)
{
    int acc = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        int4 v = a[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    for (int d = 32; d > 0; d >>= 1) acc ^= __shfl_xor(acc, d, 64);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void k_calib_fill(int4 *__restrict__ a, size_t n, int v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = make_int4(v, v + 1, v + 2, v + 3);
}

__global__ void k_calib_copy(const int4 *__restrict__ a, int4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

static void check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        fprintf(stderr, "galpha_calib: %s: %s\n", what, hipGetErrorString(e));
        exit(1);
    }
}

int main(int argc, char **argv)
{
    size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 256;
    size_t bytes = mib << 20, n = bytes / 16;
    const int grid = 4096, block = 256;
    int4 *a = nullptr, *b = nullptr;
    int *part = nullptr;
    check(hipMalloc(&a, bytes), "hipMalloc");
    check(hipMalloc(&b, bytes), "hipMalloc");
    check(hipMalloc(&part, sizeof(int) * grid), "hipMalloc");
    hipLaunchKernelGGL(k_calib_fill, dim3(grid), dim3(block), 0, 0, a, n, 1);   // first touch
    check(hipDeviceSynchronize(), "fill");
    for (int r = 0; r < 3; r++) {
        hipLaunchKernelGGL(k_calib_read, dim3(grid), dim3(block), 0, 0, (const int4 *)a, n, part);
        hipLaunchKernelGGL(k_calib_fill, dim3(grid), dim3(block), 0, 0, b, n, r);
        hipLaunchKernelGGL(k_calib_copy, dim3(grid), dim3(block), 0, 0, (const int4 *)a, b, n);
    }
    check(hipDeviceSynchronize(), "kernels");
    printf("galpha_calib: %zu MiB read, filled and copied 3 times each (k_calib_read, k_calib_fill, k_calib_copy)\n",
           mib);
    hipFree(a);
    hipFree(b);
    hipFree(part);
    return 0;
}
