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

// pixel of a real point in its set's texture (clamped inside the margin)
GA_INLINE void point_pixel(const SetDesc &sd, double2 p, int *px_, int *py_)
{
    int margin = sd.margin;
    int px = (int)floor((p.x - sd.shx) * sd.inv) + margin;
    int py = (int)floor((p.y - sd.shy) * sd.inv) + margin;
    int T = 1 << sd.logT;
    *px_ = px < margin ? margin : (px > T - 1 - margin ? T - 1 - margin : px);
    *py_ = py < margin ? margin : (py > T - 1 - margin ? T - 1 - margin : py);
}

// Spatial order of the real points: a counting sort by 16 x 16-pixel cell, cells of a set in
// Morton order after the cells of the previous sets (SetDesc::coff).  The order inside a cell is
// arbitrary; rules that depend on point order (the pixel's site, insertion claims) use the
// input index.
GA_INLINE unsigned spread14(unsigned v)
{
    v &= 0x3fffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__global__ void k_cell_count(const double2 *__restrict__ xy, BatchDev b, int *__restrict__ cell,
                             int *__restrict__ ccount)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.ntot) return;
    const SetDesc &sd = b.sets[b.vset[i]];
    int px, py;
    point_pixel(sd, xy[i], &px, &py);
    int c = sd.coff + (int)((spread14((unsigned)px >> 4) << 1) | spread14((unsigned)py >> 4));
    cell[i] = c;
    atomicAdd(&ccount[c], 1);
}

// cursor = exclusive prefix of the cell counts; perm[new index] = input index
__global__ void k_cell_scatter(const int *__restrict__ cell, int n, int *__restrict__ cursor, int *__restrict__ perm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[atomicAdd(&cursor[cell[i]], 1)] = i;
}

__global__ void k_gather_xy(const double2 *__restrict__ xy, const int *__restrict__ perm, int n,
                            double2 *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = xy[perm[i]];
}

// Map real points to pixels of their set's texture.  pix = (px << 16) | py; the point with the
// smallest input index (perm[i], or i without a renumbering) wins a pixel; k_splat_own then
// stores the winner's own index.
__global__ void k_splat(const double2 *__restrict__ xy, BatchDev b, const int *__restrict__ perm,
                        int *__restrict__ smap, int *__restrict__ pix)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.ntot) return;
    const SetDesc &sd = b.sets[b.vset[i]];
    int px, py;
    point_pixel(sd, xy[i], &px, &py);
    pix[i] = (px << 16) | py;
    atomicMin(&smap[sd.toff + ((size_t)px << sd.logT) + py], perm ? perm[i] : i);
}

__global__ void k_splat_own(BatchDev b, const int *__restrict__ perm, const int *__restrict__ pix,
                            int *__restrict__ smap)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.ntot) return;
    const SetDesc &sd = b.sets[b.vset[i]];
    int pp = pix[i];
    int *cellp = &smap[sd.toff + ((size_t)(pp >> 16) << sd.logT) + (pp & 0xffff)];
    if (*cellp == perm[i]) *cellp = i;    // one winner per pixel: no other thread writes here
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
#define PBA_MAXB 64   // bands per texture row at most (m2, m3 <= 64)
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
// S[(b*Wb + k) * T + y], k < hi[b*T + y]; lo[b*T + y] = 0 (the merge steps move both ends).
// The column reads of a band do not depend on the stack, so they are issued 8 at a time ahead
// of the dominance tests.
__global__ __launch_bounds__(64) void k_pba_stack(const short *__restrict__ g_all, int *__restrict__ S_all,
                                                  int *__restrict__ hi_all, int *__restrict__ lo_all, BatchDev bd)
{
    const int *bm = bd.bmap(BM_STACK);
    int set = set_of(bm, bd.nsets, blockIdx.x);
    const SetDesc &sd = bd.sets[set];
    int logT = sd.logT;
    int T = 1 << logT, nbx = T >> 6, Wb = T / sd.m2;
    int lb = blockIdx.x - bm[set];
    const short *g = g_all + sd.toff;
    int *S = S_all + sd.toff;
    size_t roff = (size_t)PBA_MAXB * sd.roff;
    int y = (lb % nbx) * 64 + threadIdx.x;
    int b = lb / nbx;
    int x0 = b * Wb;
    int top = 0, e0 = 0, e1 = 0;
    for (int xb = x0; xb < x0 + Wb; xb += 8) {
        short gv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) gv[u] = xb + u < x0 + Wb ? g[((size_t)(xb + u) << logT) + y] : (short)-1;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            int gy = gv[u];
            if (gy < 0) continue;
            int c = pk(xb + u, gy - y);
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
    }
    hi_all[roff + (size_t)b * T + y] = top;
    lo_all[roff + (size_t)b * T + y] = 0;
}

// ---- phase 2b: merge band envelopes pairwise, then compact ------------------------------
// Step s merges, for every row, the envelope of bands [j*2^s, j*2^s + 2^(s-1)) with that of
// the next 2^(s-1) bands: the left chain's last element and the right chain's first are
// popped while the triple at the junction shows one of them dominated (the same exact test
// as the stacks).  A chain runs through the nonempty [lo, hi) ranges of its bands in order.
// The lower envelope of a row is unique, so the merge order does not change it.
// grid (T/64) * (m2/2) blocks of 64 (lanes are rows); blocks whose group does not exist at
// this step return.
__global__ __launch_bounds__(64) void k_pba_merge_step(const int *__restrict__ S_all, int *__restrict__ hi_all,
                                                       int *__restrict__ lo_all, BatchDev bd, int s)
{
    const int *bm = bd.bmap(BM_MERGE);
    int set = set_of(bm, bd.nsets, blockIdx.x);
    const SetDesc &sd = bd.sets[set];
    int logT = sd.logT;
    int T = 1 << logT, nbx = T >> 6, m2 = sd.m2, Wb = T / m2;
    int gsz = 1 << s, half = gsz >> 1;
    int lb = blockIdx.x - bm[set];
    int j = lb / nbx;
    if ((j + 1) * gsz > m2) return;
    const int *S = S_all + sd.toff;
    size_t roff = (size_t)PBA_MAXB * sd.roff;
    int y = (lb % nbx) * 64 + threadIdx.x;
    int *HI = hi_all + roff + y, *LO = lo_all + roff + y;   // band b at [b * T]
#define ELEM(B, I) S[((size_t)((B) * Wb + (I)) << logT) + y]
    int b0 = j * gsz, bmid = b0 + half, b1 = b0 + gsz;
    // L: last nonempty band of the left half; R: first nonempty band of the right half
    int L = bmid - 1;
    while (L >= b0 && HI[(size_t)L * T] <= LO[(size_t)L * T]) L--;
    int R = bmid;
    while (R < b1 && HI[(size_t)R * T] <= LO[(size_t)R * T]) R++;
    if (L < b0 || R >= b1) return;
    int Llo = LO[(size_t)L * T], Lhi = HI[(size_t)L * T];
    int Rlo = LO[(size_t)R * T], Rhi = HI[(size_t)R * T];
    for (;;) {
        int l1 = ELEM(L, Lhi - 1), r1 = ELEM(R, Rlo);
        // element before l1 on the left chain
        bool has0 = false;
        int l0 = 0;
        if (Lhi - 1 > Llo) { l0 = ELEM(L, Lhi - 2); has0 = true; }
        else {
            int P = L - 1;
            while (P >= b0 && HI[(size_t)P * T] <= LO[(size_t)P * T]) P--;
            if (P >= b0) { l0 = ELEM(P, HI[(size_t)P * T] - 1); has0 = true; }
        }
        if (has0 && dominated(l0, l1, r1)) {
            Lhi--;
            HI[(size_t)L * T] = Lhi;
            if (Lhi == Llo) {
                L--;
                while (L >= b0 && HI[(size_t)L * T] <= LO[(size_t)L * T]) L--;
                if (L < b0) break;
                Llo = LO[(size_t)L * T]; Lhi = HI[(size_t)L * T];
            }
            continue;
        }
        // element after r1 on the right chain
        bool has2 = false;
        int r2 = 0;
        if (Rlo + 1 < Rhi) { r2 = ELEM(R, Rlo + 1); has2 = true; }
        else {
            int Q = R + 1;
            while (Q < b1 && HI[(size_t)Q * T] <= LO[(size_t)Q * T]) Q++;
            if (Q < b1) { r2 = ELEM(Q, LO[(size_t)Q * T]); has2 = true; }
        }
        if (has2 && dominated(l1, r1, r2)) {
            Rlo++;
            LO[(size_t)R * T] = Rlo;
            if (Rlo == Rhi) {
                R++;
                while (R < b1 && HI[(size_t)R * T] <= LO[(size_t)R * T]) R++;
                if (R >= b1) break;
                Rlo = LO[(size_t)R * T]; Rhi = HI[(size_t)R * T];
            }
            continue;
        }
        break;
    }
#undef ELEM
}

// grid (T/64) * m2 blocks of 64 (lanes are rows): band b's [lo, hi) goes to the row's compacted
// envelope after the ranges of bands 0 .. b-1; the last band writes the envelope size
__global__ __launch_bounds__(64) void k_pba_compact(const int *__restrict__ S_all, const int *__restrict__ hi_all,
                                                    const int *__restrict__ lo_all, int *__restrict__ E_all,
                                                    int *__restrict__ ecnt_all, BatchDev bd)
{
    const int *bm = bd.bmap(BM_STACK);
    int set = set_of(bm, bd.nsets, blockIdx.x);
    const SetDesc &sd = bd.sets[set];
    int logT = sd.logT;
    int T = 1 << logT, nbx = T >> 6, m2 = sd.m2, Wb = T / m2;
    int lb = blockIdx.x - bm[set];
    int b = lb / nbx;
    int y = (lb % nbx) * 64 + threadIdx.x;
    size_t roff = (size_t)PBA_MAXB * sd.roff;
    const int *HI = hi_all + roff + y, *LO = lo_all + roff + y;
    int off = 0;
    for (int q = 0; q < b; q++) off += HI[(size_t)q * T] - LO[(size_t)q * T];
    int lo = LO[(size_t)b * T], hi = HI[(size_t)b * T];
    const int *S = S_all + sd.toff;
    int *E = E_all + sd.toff;
    for (int i = lo; i < hi; i++) E[((size_t)(off + i - lo) << logT) + y] = S[((size_t)(b * Wb + i) << logT) + y];
    if (b == m2 - 1) ecnt_all[sd.roff + y] = off + (hi > lo ? hi - lo : 0);
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

// ccount (ncells + 1 entries) must be zero on entry
void launch_renumber(const double2 *xy, const BatchDev &b, int *cell, int *ccount, int *cursor, int ncells,
                     void *scan_tmp, int *perm, double2 *xys, hipStream_t st)
{
    int n = b.ntot;
    if (n <= 0) return;
    GA_LAUNCH(k_cell_count, dim3((n + 255) / 256), dim3(256), 0, st, xy, b, cell, ccount);
    launch_exclusive_scan(ccount, cursor, (size_t)ncells + 1, scan_tmp, st);
    GA_LAUNCH(k_cell_scatter, dim3((n + 255) / 256), dim3(256), 0, st, (const int *)cell, n, cursor, perm);
    GA_LAUNCH(k_gather_xy, dim3((n + 255) / 256), dim3(256), 0, st, xy, (const int *)perm, n, xys);
}

void launch_splat(const double2 *xy, const BatchDev &b, const int *perm, int *smap, int *pix, const int *fpix,
                  hipStream_t st)
{
    if (b.ntot > 0) {
        GA_LAUNCH(k_splat, dim3((b.ntot + 255) / 256), dim3(256), 0, st, xy, b, perm, smap, pix);
        if (perm)
            GA_LAUNCH(k_splat_own, dim3((b.ntot + 255) / 256), dim3(256), 0, st, b, perm, (const int *)pix, smap);
    }
    int nf = b.V - b.ntot;
    if (nf > 0)
        GA_LAUNCH(k_splat_frame, dim3((nf + 255) / 256), dim3(256), 0, st, fpix, b, smap, pix);
}

void launch_pba(const PbaBuffers &p, const BatchDev &b, hipStream_t st)
{
    GA_LAUNCH(k_pba_cols, dim3(b.blocks[BM_COLS]), dim3(256), 0, st, p.smap, p.g, b);
    GA_LAUNCH(k_pba_stack, dim3(b.blocks[BM_STACK]), dim3(64), 0, st, p.g, p.S, p.cnt, p.lo, b);
    for (int s = 1; (1 << s) <= b.max_m2; s++)
        GA_LAUNCH(k_pba_merge_step, dim3(b.blocks[BM_MERGE]), dim3(64), 0, st, (const int *)p.S, p.cnt, p.lo, b, s);
    GA_LAUNCH(k_pba_compact, dim3(b.blocks[BM_STACK]), dim3(64), 0, st, (const int *)p.S, (const int *)p.cnt,
              (const int *)p.lo, p.E, p.ecnt, b);
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
