// util.hip -- small device utilities: bounding box of the input points.
#include "dev_common.h"
#include "kernels.h"

__global__ __launch_bounds__(256) void k_bbox(const double2 *__restrict__ xy, int n, double *out4)
{
    double a = INFINITY, b = INFINITY, c = -INFINITY, d = -INFINITY;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double2 p = xy[i];
        a = fmin(a, p.x); b = fmin(b, p.y); c = fmax(c, p.x); d = fmax(d, p.y);
    }
    for (int s = 32; s > 0; s >>= 1) {
        a = fmin(a, __shfl_down(a, s, 64)); b = fmin(b, __shfl_down(b, s, 64));
        c = fmax(c, __shfl_down(c, s, 64)); d = fmax(d, __shfl_down(d, s, 64));
    }
    __shared__ double sh[4][4];
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[w][0] = a; sh[w][1] = b; sh[w][2] = c; sh[w][3] = d; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            a = fmin(a, sh[k][0]); b = fmin(b, sh[k][1]); c = fmax(c, sh[k][2]); d = fmax(d, sh[k][3]);
        }
        out4[4 * blockIdx.x + 0] = a; out4[4 * blockIdx.x + 1] = b;
        out4[4 * blockIdx.x + 2] = c; out4[4 * blockIdx.x + 3] = d;
    }
}

void bbox_device(const double *d_xy, int n, double *out4, void *tmp, hipStream_t st)
{
    const int blocks = 64;
    double *part = (double *)tmp;
    GA_LAUNCH(k_bbox, dim3(blocks), dim3(256), 0, st, (const double2 *)d_xy, n, part);
    double h[4 * blocks];
    hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
    double a = INFINITY, b = INFINITY, c = -INFINITY, d = -INFINITY;
    for (int k = 0; k < blocks; k++) {
        a = fmin(a, h[4 * k]); b = fmin(b, h[4 * k + 1]); c = fmax(c, h[4 * k + 2]); d = fmax(d, h[4 * k + 3]);
    }
    out4[0] = a; out4[1] = b; out4[2] = c; out4[3] = d;
}
