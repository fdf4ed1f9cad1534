// scan.hip -- exclusive prefix sum of int32 arrays (replaces Cuda/cudaScanLargeArray.cu, which is
// not reused).  Three passes: per-block reduction of 4096 elements (256 threads x 16), a
// recursive scan of the block sums, and a block-local scan that adds the block offset.
// Wave-level prefix sums use 64-lane shuffles.
#include "dev_common.h"
#include "kernels.h"

#define SCAN_BLOCK 256
#define SCAN_PER_THREAD 16
#define SCAN_TILE (SCAN_BLOCK * SCAN_PER_THREAD)

GA_INLINE int wave_incl_scan(int v)
{
    int lane = threadIdx.x & 63;
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// block-wide exclusive scan of one value per thread; returns the block total in *total
GA_INLINE int block_excl_scan(int v, int *total)
{
    __shared__ int wsum[SCAN_BLOCK / 64];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < SCAN_BLOCK / 64; i++) {
        if (i < w) off += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const int *__restrict__ in, size_t n,
                                                            int *__restrict__ sums)
{
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_PER_THREAD;
    int s = 0;
    for (int k = 0; k < SCAN_PER_THREAD; k++)
        if (base + k < n) s += in[base + k];
    int tot;
    (void)block_excl_scan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const int *__restrict__ in, size_t n,
                                                           const int *__restrict__ offs, int *__restrict__ out)
{
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_PER_THREAD;
    int v[SCAN_PER_THREAD];
    int s = 0;
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int tot;
    int ex = block_excl_scan(s, &tot) + (offs ? offs[blockIdx.x] : 0);
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
}

static size_t nblocks(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

size_t scan_temp_bytes(size_t n)
{
    size_t total = 0;
    while (n > SCAN_TILE) {
        size_t b = nblocks(n);
        total += 2 * b * sizeof(int) + 256;
        n = b;
    }
    return total + 256;
}

// exclusive scan; in and out may alias
void launch_exclusive_scan(const int *in, int *out, size_t n, void *temp, hipStream_t st)
{
    if (n == 0) return;
    if (n <= SCAN_TILE) {
        GA_LAUNCH(k_scan_apply, dim3(1), dim3(SCAN_BLOCK), 0, st, in, n, (const int *)nullptr, out);
        return;
    }
    size_t b = nblocks(n);
    int *sums = (int *)temp;
    int *offs = sums + b;
    void *rest = (void *)(((uintptr_t)(offs + b) + 255) & ~(uintptr_t)255);
    GA_LAUNCH(k_scan_reduce, dim3((unsigned)b), dim3(SCAN_BLOCK), 0, st, in, n, sums);
    launch_exclusive_scan(sums, offs, b, rest, st);
    GA_LAUNCH(k_scan_apply, dim3((unsigned)b), dim3(SCAN_BLOCK), 0, st, in, n, (const int *)offs, out);
}
