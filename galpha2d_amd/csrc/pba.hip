// pba.hip -- discrete Voronoi diagram on a T x T texture (gfx950).
//
// Replaces the reference's PBA path (Cuda/pba2DHost.cu:82-155, Cuda/pba2DKernel.h:48-477,
// Cuda/kernelVoronoi.h:48-191, Cuda/cudaVoronoi.cu:181-307) with a layout and kernels of
// its own:
//   * textures are x-major (pixel (x, y) at x*T + y), so a wave whose lanes are 64
//     consecutive rows reads and writes 256 contiguous bytes per column step;
//   * phase 1 (nearest site in each column) is one wave per column: a 64-bit ballot per
//     64-row chunk marks the sites, and the nearest site above/below each pixel comes from
//     bit scans over those masks held in LDS -- the site map is read once;
//   * phase 2 builds per-(row, band) stacks of proximate sites with lanes = rows, stored as
//     arrays (no pointer chasing), integer-exact dominance tests;
//   * phase 2b merges band stacks per row and compacts the envelope contiguously;
//   * phase 3 colours pixels with lanes = rows, binary-searching the start element of its
//     pixel band in the compacted envelope.
// All distances are exact integers; every pixel receives a site at minimal Euclidean
// distance (ties go to the lower column, then the lower row).
//
// The island certificate replaces the reference's iterate-and-recolour loop
// (kernelIslandDetection, CPU colorIsland): a pixel must have a same-labelled neighbour
// one step toward its site (dx, dy or the diagonal).  Sites failing it are taken out of
// the site map and the diagram is recomputed, so every label stays an exact nearest site.
#include "dev_common.h"
#include "kernels.h"

#define PBA_EMPTY 0x7fffffff

__global__ void k_fill_i32(int *a, int v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) a[i] = v;
}

// Map real points to pixels.  pix = (px << 16) | py; the smallest index wins a pixel.
__global__ void k_splat(const double2 *__restrict__ xy, int n, double shx, double shy,
                        double inv, int margin, int logT, int *__restrict__ smap,
                        int *__restrict__ pix)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 p = xy[i];
    int px = (int)floor((p.x - shx) * inv) + margin;
    int py = (int)floor((p.y - shy) * inv) + margin;
    int T = 1 << logT;
    px = px < margin ? margin : (px > T - 1 - margin ? T - 1 - margin : px);
    py = py < margin ? margin : (py > T - 1 - margin ? T - 1 - margin : py);
    pix[i] = (px << 16) | py;
    atomicMin(&smap[((size_t)px << logT) + py], i);
}

// Frame sites (ids n .. n+nf-1) at their border pixels.
__global__ void k_splat_frame(const int *__restrict__ fpix, int nf, int n, int logT,
                              int *__restrict__ smap, int *__restrict__ pix)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nf) return;
    int pp = fpix[k];
    pix[n + k] = pp;
    smap[((size_t)(pp >> 16) << logT) + (pp & 0xffff)] = n + k;
}

// ---- phase 1: nearest site per column --------------------------------------------------
// block = 256 (4 waves, one column each); g[x*T + y] = row of the nearest site in column x
// (ties to the lower row), -1 for an empty column.
__global__ __launch_bounds__(256) void k_pba_cols(const int *__restrict__ smap,
                                                  short *__restrict__ g, int logT)
{
    __shared__ unsigned long long masks[4][256];
    __shared__ short prv[4][256], nxt[4][256];
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int T = 1 << logT, nch = T >> 6;
    int x = blockIdx.x * 4 + wave;
    const int *col = smap + ((size_t)x << logT);
    for (int c = 0; c < nch; c++) {
        int v = col[(c << 6) + lane];
        unsigned long long m = __ballot(v != PBA_EMPTY);
        if (lane == 0) masks[wave][c] = m;
    }
    __syncthreads();
    // previous / next nonempty chunk, 64 chunks per ballot group
    int carry_prev = -1;
    for (int gbase = 0; gbase < nch; gbase += 64) {
        int c = gbase + lane;
        bool nz = c < nch && masks[wave][c] != 0ull;
        unsigned long long b = __ballot(nz);
        unsigned long long below = lane ? (b & ((1ull << lane) - 1ull)) : 0ull;
        int pv = below ? gbase + 63 - __clzll(below) : carry_prev;
        if (c < nch) prv[wave][c] = (short)pv;
        if (b) carry_prev = gbase + 63 - __clzll(b);
    }
    int carry_next = nch;
    for (int gbase = ((nch - 1) >> 6) << 6; gbase >= 0; gbase -= 64) {
        int c = gbase + lane;
        bool nz = c < nch && masks[wave][c] != 0ull;
        unsigned long long b = __ballot(nz);
        unsigned long long above = (lane == 63) ? 0ull : (b & (~0ull << (lane + 1)));
        int nx = above ? gbase + __ffsll((long long)above) - 1 : carry_next;
        if (c < nch) nxt[wave][c] = (short)nx;
        if (b) carry_next = gbase + __ffsll((long long)b) - 1;
    }
    __syncthreads();
    short *gc = g + ((size_t)x << logT);
    for (int c = 0; c < nch; c++) {
        unsigned long long m = masks[wave][c];
        int y = (c << 6) + lane;
        unsigned long long mb = (lane == 63) ? m : (m & ((2ull << lane) - 1ull));
        int yb = -1, ya = -1;
        if (mb) yb = (c << 6) + 63 - __clzll(mb);
        else {
            int p = prv[wave][c];
            if (p >= 0) yb = (p << 6) + 63 - __clzll(masks[wave][p]);
        }
        unsigned long long ma = m & (~0ull << lane);
        if (ma) ya = (c << 6) + __ffsll((long long)ma) - 1;
        else {
            int q = nxt[wave][c];
            if (q < nch) ya = (q << 6) + __ffsll((long long)masks[wave][q]) - 1;
        }
        int r;
        if (yb < 0) r = ya;
        else if (ya < 0) r = yb;
        else r = (y - yb <= ya - y) ? yb : ya;
        gc[y] = (short)r;
    }
}

// ---- phase 2: proximate-site stacks per (row, band) ------------------------------------
GA_INLINE int pk(int x, int h) { return (x << 16) | (h & 0xffff); }
GA_INLINE int pk_x(int e) { return (int)((unsigned)e >> 16); }
GA_INLINE int pk_h(int e) { return (int)(short)(e & 0xffff); }

// b is dominated by a (left) and c (right): the a|b bisector is not left of the b|c one
GA_INLINE bool dominated(int a, int b, int c)
{
    long long xa = pk_x(a), ha = pk_h(a), xb = pk_x(b), hb = pk_h(b), xc = pk_x(c), hc = pk_h(c);
    long long fa = xa * xa + ha * ha, fb = xb * xb + hb * hb, fc = xc * xc + hc * hc;
    return (fb - fa) * (xc - xb) >= (fc - fb) * (xb - xa);
}

// grid (T/64, m2), block 64: lanes are rows.  Stack entries of band b live in
// S[(b*Wb + k) * T + y], k < cnt[b*T + y].
__global__ __launch_bounds__(64) void k_pba_stack(const short *__restrict__ g, int *__restrict__ S,
                                                  int *__restrict__ cnt, int logT, int Wb)
{
    int T = 1 << logT;
    int y = blockIdx.x * 64 + threadIdx.x;
    int b = blockIdx.y;
    int x0 = b * Wb;
    int top = 0, e0 = 0, e1 = 0;
    for (int x = x0; x < x0 + Wb; x++) {
        int gy = g[((size_t)x << logT) + y];
        if (gy < 0) continue;
        int c = pk(x, gy - y);
        while (top >= 2 && dominated(e0, e1, c)) {
            top--;
            e1 = e0;
            if (top >= 2) e0 = S[((size_t)(x0 + top - 2) << logT) + y];
        }
        S[((size_t)(x0 + top) << logT) + y] = c;
        top++;
        e0 = e1;
        e1 = c;
    }
    cnt[(size_t)b * T + y] = top;
}

// ---- phase 2b: merge band stacks per row, compact the envelope -------------------------
#define PBA_MAXB 64
__global__ __launch_bounds__(64) void k_pba_merge(const int *__restrict__ S, const int *__restrict__ cnt,
                                                  int *__restrict__ E, int *__restrict__ ecnt,
                                                  int logT, int Wb, int m2)
{
    __shared__ short lo_s[PBA_MAXB][64], hi_s[PBA_MAXB][64];
    __shared__ unsigned char bst[PBA_MAXB][64];
    int T = 1 << logT;
    int lane = threadIdx.x;
    int y = blockIdx.x * 64 + lane;
    int nb = 0;
#define ELEM(B, I) S[((size_t)((B) * Wb + (I)) << logT) + y]
    for (int b = 0; b < m2; b++) {
        int lo = 0, hi = cnt[(size_t)b * T + y];
        if (hi == 0) continue;
        while (nb > 0) {
            int L = bst[nb - 1][lane];
            int Lhi = hi_s[L][lane], Llo = lo_s[L][lane];
            int l1 = ELEM(L, Lhi - 1);
            bool has0 = false;
            int l0 = 0;
            if (Lhi - 1 > Llo) { l0 = ELEM(L, Lhi - 2); has0 = true; }
            else if (nb >= 2) { int L2 = bst[nb - 2][lane]; l0 = ELEM(L2, hi_s[L2][lane] - 1); has0 = true; }
            int r1 = ELEM(b, lo);
            if (has0 && dominated(l0, l1, r1)) {
                hi_s[L][lane] = (short)(Lhi - 1);
                if (Lhi - 1 == Llo) nb--;
                continue;
            }
            if (lo + 1 < hi) {
                int r2 = ELEM(b, lo + 1);
                if (dominated(l1, r1, r2)) { lo++; continue; }
            }
            break;
        }
        bst[nb][lane] = (unsigned char)b;
        lo_s[b][lane] = (short)lo;
        hi_s[b][lane] = (short)hi;
        nb++;
    }
    int k = 0;
    for (int s = 0; s < nb; s++) {
        int B = bst[s][lane];
        int lo = lo_s[B][lane], hi = hi_s[B][lane];
        for (int i = lo; i < hi; i++) {
            E[((size_t)k << logT) + y] = ELEM(B, i);
            k++;
        }
    }
    ecnt[y] = k;
#undef ELEM
}

// ---- phase 3: colour ------------------------------------------------------------------
GA_INLINE long long fdist(int e, int i)
{
    long long dx = i - pk_x(e), h = pk_h(e);
    return dx * dx + h * h;
}

// grid (T/64, m3), block 64; vor[x*T + y] = site id
__global__ __launch_bounds__(64) void k_pba_color(const int *__restrict__ E, const int *__restrict__ ecnt,
                                                  const int *__restrict__ smap, int *__restrict__ vor,
                                                  int logT, int Wp)
{
    int y = blockIdx.x * 64 + threadIdx.x;
    int ps = blockIdx.y * Wp, pe = ps + Wp;
    int n = ecnt[y];
    if (n <= 0) {
        for (int i = ps; i < pe; i++) vor[((size_t)i << logT) + y] = -1;
        return;
    }
#define EL(K) E[((size_t)(K) << logT) + y]
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (fdist(EL(mid + 1), ps) < fdist(EL(mid), ps)) lo = mid + 1;
        else hi = mid;
    }
    int k = lo;
    int e = EL(k);
    int en = (k + 1 < n) ? EL(k + 1) : 0;
    int id = smap[((size_t)pk_x(e) << logT) + (y + pk_h(e))];
    for (int i = ps; i < pe; i++) {
        while (k + 1 < n && fdist(en, i) < fdist(e, i)) {
            k++;
            e = en;
            en = (k + 1 < n) ? EL(k + 1) : 0;
            id = smap[((size_t)pk_x(e) << logT) + (y + pk_h(e))];
        }
        vor[((size_t)i << logT) + y] = id;
    }
#undef EL
}

// ---- island certificate ------------------------------------------------------------------
__global__ void k_islands(const int *__restrict__ vor, const int *__restrict__ pix, int logT,
                          int *__restrict__ bad, int *__restrict__ flag)
{
    int T = 1 << logT;
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ((size_t)T << logT)) return;
    int x = (int)(idx >> logT), y = (int)(idx & (T - 1));
    int s = vor[idx];
    if (s < 0) return;
    int sp = pix[s];
    int sx = sp >> 16, sy = sp & 0xffff;
    int dx = (sx > x) - (sx < x), dy = (sy > y) - (sy < y);
    if (dx == 0 && dy == 0) return;
    bool ok = false;
    if (dx != 0 && vor[idx + ((long long)dx << logT)] == s) ok = true;
    if (!ok && dy != 0 && vor[idx + dy] == s) ok = true;
    if (!ok && dx != 0 && dy != 0 && vor[idx + ((long long)dx << logT) + dy] == s) ok = true;
    if (!ok) {
        bad[s] = 1;
        *flag = 1;
    }
}

// Take sites flagged by the certificate out of the site map.
__global__ void k_remove_bad(int *__restrict__ bad, const int *__restrict__ pix, int n, int logT,
                             int *__restrict__ smap)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || !bad[s]) return;
    bad[s] = 0;
    int sp = pix[s];
    smap[((size_t)(sp >> 16) << logT) + (sp & 0xffff)] = PBA_EMPTY;
}

// ---- host-side launchers ---------------------------------------------------------------
void launch_fill_i32(int *a, int v, size_t n, hipStream_t st)
{
    int blocks = (int)((n + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    GA_LAUNCH(k_fill_i32, dim3(blocks), dim3(256), 0, st, a, v, n);
}

void launch_splat(const double2 *xy, int n, double shx, double shy, double inv, int margin,
                  int logT, int *smap, int *pix, const int *fpix, int nf, hipStream_t st)
{
    if (n > 0)
        GA_LAUNCH(k_splat, dim3((n + 255) / 256), dim3(256), 0, st, xy, n, shx, shy, inv,
                           margin, logT, smap, pix);
    if (nf > 0)
        GA_LAUNCH(k_splat_frame, dim3((nf + 255) / 256), dim3(256), 0, st, fpix, nf, n, logT,
                           smap, pix);
}

void launch_pba(const PbaBuffers &b, int logT, int m2, int m3, hipStream_t st)
{
    int T = 1 << logT;
    GA_LAUNCH(k_pba_cols, dim3(T / 4), dim3(256), 0, st, b.smap, b.g, logT);
    GA_LAUNCH(k_pba_stack, dim3(T / 64, m2), dim3(64), 0, st, b.g, b.S, b.cnt, logT, T / m2);
    GA_LAUNCH(k_pba_merge, dim3(T / 64), dim3(64), 0, st, b.S, b.cnt, b.E, b.ecnt, logT, T / m2, m2);
    GA_LAUNCH(k_pba_color, dim3(T / 64, m3), dim3(64), 0, st, b.E, b.ecnt, b.smap, b.vor, logT, T / m3);
}

void launch_islands(const int *vor, const int *pix, int logT, int *bad, int *flag, hipStream_t st)
{
    size_t P = (size_t)1 << (2 * logT);
    GA_LAUNCH(k_islands, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, vor, pix, logT, bad, flag);
}

void launch_remove_bad(int *bad, const int *pix, int n, int logT, int *smap, hipStream_t st)
{
    if (n > 0)
        GA_LAUNCH(k_remove_bad, dim3((n + 255) / 256), dim3(256), 0, st, bad, pix, n, logT, smap);
}
