// util.hip -- bounding boxes of point sets on the device (min x, min y, max x, max y, and a flag
// for non-finite coordinates, which the host rejects).
#include "dev_common.h"
#include "kernels.h"
#include <algorithm>

// reduce one block's box; returns it in thread 0
GA_INLINE void block_box(double &a, double &b, double &c, double &d, double &bad)
{
    for (int s = 32; s > 0; s >>= 1) {
        a = fmin(a, __shfl_down(a, s, 64)); b = fmin(b, __shfl_down(b, s, 64));
        c = fmax(c, __shfl_down(c, s, 64)); d = fmax(d, __shfl_down(d, s, 64));
        bad = fmax(bad, __shfl_down(bad, s, 64));
    }
    __shared__ double sh[16][5];
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[w][0] = a; sh[w][1] = b; sh[w][2] = c; sh[w][3] = d; sh[w][4] = bad; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
            a = fmin(a, sh[k][0]); b = fmin(b, sh[k][1]); c = fmax(c, sh[k][2]); d = fmax(d, sh[k][3]);
            bad = fmax(bad, sh[k][4]);
        }
    }
}

GA_INLINE void box_add(double2 p, double &a, double &b, double &c, double &d, double &bad)
{
    if (!isfinite(p.x) || !isfinite(p.y)) bad = 1.0;
    a = fmin(a, p.x); b = fmin(b, p.y); c = fmax(c, p.x); d = fmax(d, p.y);
}

// one set, 64 blocks of 1,024 threads: partial boxes out5[5 * block]; four loads in flight per
// thread (the box is read before anything else can run, so its latency is all exposed)
__global__ __launch_bounds__(1024) void k_bbox(const double2 *__restrict__ xy, int n, double *out5)
{
    double a = INFINITY, b = INFINITY, c = -INFINITY, d = -INFINITY, bad = 0.0;
    const int stride = gridDim.x * blockDim.x;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const double2 p0 = xy[i], p1 = xy[i + stride], p2 = xy[i + 2 * stride], p3 = xy[i + 3 * stride];
        box_add(p0, a, b, c, d, bad);
        box_add(p1, a, b, c, d, bad);
        box_add(p2, a, b, c, d, bad);
        box_add(p3, a, b, c, d, bad);
    }
    for (; i < n; i += stride) box_add(xy[i], a, b, c, d, bad);
    block_box(a, b, c, d, bad);
    if (threadIdx.x == 0) {
        double *o = out5 + 5 * blockIdx.x;
        o[0] = a; o[1] = b; o[2] = c; o[3] = d; o[4] = bad;
    }
}

// many sets, one block each: set s covers points [offs[s], offs[s + 1])
__global__ __launch_bounds__(256) void k_bbox_sets(const double2 *__restrict__ xy, const int *__restrict__ offs,
                                                   double *out5)
{
    int s = blockIdx.x;
    double a = INFINITY, b = INFINITY, c = -INFINITY, d = -INFINITY, bad = 0.0;
    for (int i = offs[s] + threadIdx.x; i < offs[s + 1]; i += blockDim.x) box_add(xy[i], a, b, c, d, bad);
    block_box(a, b, c, d, bad);
    if (threadIdx.x == 0) {
        double *o = out5 + 5 * s;
        o[0] = a; o[1] = b; o[2] = c; o[3] = d; o[4] = bad;
    }
}

// box of one set (tmp: >= 5 * 64 doubles of device memory); out5 on the host
void bbox_device(const double *d_xy, int n, double *out5, void *tmp, hipStream_t st)
{
    const int blocks = 64;
    double *part = (double *)tmp;
    GA_LAUNCH(k_bbox, dim3(blocks), dim3(1024), 0, st, (const double2 *)d_xy, n, part);
    double h[5 * blocks];
    hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
    double a = INFINITY, b = INFINITY, c = -INFINITY, d = -INFINITY, bad = 0.0;
    for (int k = 0; k < blocks; k++) {
        a = fmin(a, h[5 * k]); b = fmin(b, h[5 * k + 1]); c = fmax(c, h[5 * k + 2]); d = fmax(d, h[5 * k + 3]);
        bad = fmax(bad, h[5 * k + 4]);
    }
    out5[0] = a; out5[1] = b; out5[2] = c; out5[3] = d; out5[4] = bad;
}

// boxes of nsets sets (d_offs: device offsets[nsets + 1]; tmp: >= 5 * nsets doubles); out5 on the host
void bbox_sets_device(const double *d_xy, const int *d_offs, int nsets, double *out5, void *tmp, hipStream_t st)
{
    if (nsets <= 0) return;
    GA_LAUNCH(k_bbox_sets, dim3(nsets), dim3(256), 0, st, (const double2 *)d_xy, d_offs, (double *)tmp);
    hipMemcpyAsync(out5, tmp, sizeof(double) * 5 * (size_t)nsets, hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
}

// segments of device ints gathered into one contiguous array (for a single host read)
// dst may be pinned host memory (the host waits on the stream before it reads)
__global__ void k_gather_segments(IntSegments g, int *dst)
{
    int base = 0;
    for (int k = 0; k < g.n; k++) {
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.len[k]; i += gridDim.x * blockDim.x)
            dst[base + i] = g.src[k][i];
        base += g.len[k];
    }
    __threadfence_system();
}

void launch_gather_segments(const IntSegments &g, int *dst, hipStream_t st)
{
    int tot = 0;
    for (int k = 0; k < g.n; k++) tot += g.len[k];
    int blocks = std::max(1, std::min(256, (tot + 255) / 256));
    GA_LAUNCH(k_gather_segments, dim3(blocks), dim3(256), 0, st, g, dst);
}

// several device int arrays, each set to one value, in one launch (16-byte stores where a
// segment's start is 16-byte aligned)
__global__ void k_fill_multi(IntFills f)
{
    const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, nth = (size_t)gridDim.x * blockDim.x;
    for (int k = 0; k < f.n; k++) {
        int *p = f.p[k];
        const size_t n = f.len[k];
        const int v = f.v[k];
        size_t head = 0;
        if (((uintptr_t)p & 15) == 0) {
            const size_t n4 = n / 4;
            int4 *p4 = (int4 *)p;
            for (size_t i = tid; i < n4; i += nth) p4[i] = make_int4(v, v, v, v);
            head = 4 * n4;
        }
        for (size_t i = head + tid; i < n; i += nth) p[i] = v;
    }
}

void launch_fill_multi(const IntFills &f, hipStream_t st)
{
    size_t tot = 0;
    for (int k = 0; k < f.n; k++) tot += f.len[k];
    if (!tot) return;
    int blocks = (int)std::max((size_t)1, std::min((size_t)2048, (tot / 4 + 255) / 256));
    GA_LAUNCH(k_fill_multi, dim3(blocks), dim3(256), 0, st, f);
}
