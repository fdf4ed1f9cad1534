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

// Set of every vertex: real points by their poff range, frame vertices by their fbase range.
__global__ void k_vset(const SetDesc *__restrict__ sets, int nsets, int ntot, int V, int *__restrict__ vset)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    int lo = 0, hi = nsets - 1;
    if (v < ntot) {
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (sets[mid].poff <= v) lo = mid;
            else hi = mid - 1;
        }
    } else {
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (sets[mid].fbase <= v) lo = mid;
            else hi = mid - 1;
        }
    }
    vset[v] = lo;
}

// Map real points to pixels of their set's texture.  pix = (px << 16) | py; the smallest
// index wins a pixel.
__global__ void k_splat(const double2 *__restrict__ xy, BatchDev b, int *__restrict__ smap,
                        int *__restrict__ pix)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.ntot) return;
    const SetDesc &sd = b.sets[b.vset[i]];
    double2 p = xy[i];
    int margin = sd.margin;
    int px = (int)floor((p.x - sd.shx) * sd.inv) + margin;
    int py = (int)floor((p.y - sd.shy) * sd.inv) + margin;
    int T = 1 << sd.logT;
    px = px < margin ? margin : (px > T - 1 - margin ? T - 1 - margin : px);
    py = py < margin ? margin : (py > T - 1 - margin ? T - 1 - margin : py);
    pix[i] = (px << 16) | py;
    atomicMin(&smap[sd.toff + ((size_t)px << sd.logT) + py], i);
}

// Frame sites (global ids ntot .. V-1) at their border pixels (fpix: packed local pixels).
__global__ void k_splat_frame(const int *__restrict__ fpix, BatchDev b, int *__restrict__ smap,
                              int *__restrict__ pix)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int f = b.ntot + k;
    if (f >= b.V) return;
    const SetDesc &sd = b.sets[b.vset[f]];
    int pp = fpix[k];
    pix[f] = pp;
    smap[sd.toff + ((size_t)(pp >> 16) << sd.logT) + (pp & 0xffff)] = f;
}

// ---- phase 1: nearest site per column --------------------------------------------------
// block = 256 (4 waves, one column each); g[x*T + y] = row of the nearest site in column x
// (ties to the lower row), -1 for an empty column.
__global__ __launch_bounds__(256) void k_pba_cols(const int *__restrict__ smap_all,
                                                  short *__restrict__ g_all, BatchDev bd)
{
    __shared__ unsigned long long masks[4][256];
    __shared__ short prv[4][256], nxt[4][256];
    const int *bm = bd.bmap(BM_COLS);
    int set = set_of(bm, bd.nsets, blockIdx.x);
    const SetDesc &sd = bd.sets[set];
    int logT = sd.logT;
    const int *smap = smap_all + sd.toff;
    short *g = g_all + sd.toff;
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int T = 1 << logT, nch = T >> 6;
    int x = (blockIdx.x - bm[set]) * 4 + wave;
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
__global__ __launch_bounds__(64) void k_pba_stack(const short *__restrict__ g_all, int *__restrict__ S_all,
                                                  int *__restrict__ cnt_all, BatchDev bd)
{
    const int *bm = bd.bmap(BM_STACK);
    int set = set_of(bm, bd.nsets, blockIdx.x);
    const SetDesc &sd = bd.sets[set];
    int logT = sd.logT;
    int T = 1 << logT, nbx = T >> 6, Wb = T / sd.m2;
    int lb = blockIdx.x - bm[set];
    const short *g = g_all + sd.toff;
    int *S = S_all + sd.toff;
    int *cnt = cnt_all + (size_t)16 * sd.roff;
    int y = (lb % nbx) * 64 + threadIdx.x;
    int b = lb / nbx;
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
__global__ __launch_bounds__(64) void k_pba_merge(const int *__restrict__ S_all, const int *__restrict__ cnt_all,
                                                  int *__restrict__ E_all, int *__restrict__ ecnt_all,
                                                  BatchDev bd)
{
    __shared__ short lo_s[PBA_MAXB][64], hi_s[PBA_MAXB][64];
    __shared__ unsigned char bst[PBA_MAXB][64];
    const int *bm = bd.bmap(BM_MERGE);
    int set = set_of(bm, bd.nsets, blockIdx.x);
    const SetDesc &sd = bd.sets[set];
    int logT = sd.logT;
    int T = 1 << logT, m2 = sd.m2, Wb = T / m2;
    const int *S = S_all + sd.toff;
    const int *cnt = cnt_all + (size_t)16 * sd.roff;
    int *E = E_all + sd.toff;
    int *ecnt = ecnt_all + sd.roff;
    int lane = threadIdx.x;
    int y = (blockIdx.x - bm[set]) * 64 + lane;
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
__global__ __launch_bounds__(64) void k_pba_color(const int *__restrict__ E_all, const int *__restrict__ ecnt_all,
                                                  const int *__restrict__ smap_all, int *__restrict__ vor_all,
                                                  BatchDev bd)
{
    const int *bm = bd.bmap(BM_COLOR);
    int set = set_of(bm, bd.nsets, blockIdx.x);
    const SetDesc &sd = bd.sets[set];
    int logT = sd.logT;
    int T = 1 << logT, nbx = T >> 6, Wp = T / sd.m3;
    int lb = blockIdx.x - bm[set];
    const int *E = E_all + sd.toff;
    const int *ecnt = ecnt_all + sd.roff;
    const int *smap = smap_all + sd.toff;
    int *vor = vor_all + sd.toff;
    int y = (lb % nbx) * 64 + threadIdx.x;
    int ps = (lb / nbx) * Wp, pe = ps + Wp;
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
// blocks of 256 pixels (a texture has T*T >= 4096 pixels, a multiple of 256)
__global__ __launch_bounds__(256) void k_islands(const int *__restrict__ vor_all, const int *__restrict__ pix,
                                                 BatchDev bd, int *__restrict__ bad, int *__restrict__ flag)
{
    const int *bm = bd.bmap(BM_PIX);
    int set = set_of(bm, bd.nsets, blockIdx.x);
    const SetDesc &sd = bd.sets[set];
    int logT = sd.logT;
    int T = 1 << logT;
    const int *vor = vor_all + sd.toff;
    size_t idx = (size_t)(blockIdx.x - bm[set]) * 256 + threadIdx.x;
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

// Take sites flagged by the certificate out of the site map (counted in *removed).
__global__ void k_remove_bad(int *__restrict__ bad, const int *__restrict__ pix, BatchDev bd,
                             int *__restrict__ smap, int *__restrict__ removed)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    bool r = s < bd.ntot && bad[s];
    wave_count_full(removed, r ? 1 : 0);
    if (!r) return;
    bad[s] = 0;
    const SetDesc &sd = bd.sets[bd.vset[s]];
    int sp = pix[s];
    smap[sd.toff + ((size_t)(sp >> 16) << sd.logT) + (sp & 0xffff)] = PBA_EMPTY;
}

// ---- host-side launchers ---------------------------------------------------------------
void launch_fill_i32(int *a, int v, size_t n, hipStream_t st)
{
    int blocks = (int)((n + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    GA_LAUNCH(k_fill_i32, dim3(blocks), dim3(256), 0, st, a, v, n);
}

void launch_vset(const BatchDev &b, int *vset, hipStream_t st)
{
    if (b.V > 0)
        GA_LAUNCH(k_vset, dim3((b.V + 255) / 256), dim3(256), 0, st, b.sets, b.nsets, b.ntot, b.V, vset);
}

void launch_splat(const double2 *xy, const BatchDev &b, int *smap, int *pix, const int *fpix, hipStream_t st)
{
    if (b.ntot > 0)
        GA_LAUNCH(k_splat, dim3((b.ntot + 255) / 256), dim3(256), 0, st, xy, b, smap, pix);
    int nf = b.V - b.ntot;
    if (nf > 0)
        GA_LAUNCH(k_splat_frame, dim3((nf + 255) / 256), dim3(256), 0, st, fpix, b, smap, pix);
}

void launch_pba(const PbaBuffers &p, const BatchDev &b, hipStream_t st)
{
    GA_LAUNCH(k_pba_cols, dim3(b.blocks[BM_COLS]), dim3(256), 0, st, p.smap, p.g, b);
    GA_LAUNCH(k_pba_stack, dim3(b.blocks[BM_STACK]), dim3(64), 0, st, p.g, p.S, p.cnt, b);
    GA_LAUNCH(k_pba_merge, dim3(b.blocks[BM_MERGE]), dim3(64), 0, st, p.S, p.cnt, p.E, p.ecnt, b);
    GA_LAUNCH(k_pba_color, dim3(b.blocks[BM_COLOR]), dim3(64), 0, st, p.E, p.ecnt, p.smap, p.vor, b);
}

void launch_islands(const int *vor, const int *pix, const BatchDev &b, int *bad, int *flag, hipStream_t st)
{
    GA_LAUNCH(k_islands, dim3(b.blocks[BM_PIX]), dim3(256), 0, st, vor, pix, b, bad, flag);
}

void launch_remove_bad(int *bad, const int *pix, const BatchDev &b, int *smap, int *removed, hipStream_t st)
{
    if (b.ntot > 0)
        GA_LAUNCH(k_remove_bad, dim3((b.ntot + 255) / 256), dim3(256), 0, st, bad, pix, b, smap, removed);
}
