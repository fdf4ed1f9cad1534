// mesh.hip -- triangle-mesh stages of the Delaunay pipeline on gfx950.
//
// Reference stages replaced here (paths relative to the reference root):
//   reconstruction      Cuda/cudaReconstruction.cu:134-574 (row counting + CPU prefix + CPU hull
//                       fix + CPU vertex fans).  Here: one pass over 2x2 pixel blocks with a
//                       tile scan, triangles emitted counter-clockwise; the frame sites at the
//                       texture border make the mesh boundary a convex rectangle, so no hull fix.
//   adjacency           cpuVertexArray + kernelFindNextTriangles (:348-436): vertex fans built
//                       on the device as a CSR (degree count, scan, fill), then edge matching.
//   shifting            Cuda/cudaShifting.cu:117-599: bulk move of every vertex to its true
//                       position with exact orientation checks, then greedy independent rounds,
//                       then deletion (ear clipping of the star-shaped hole) of the rest.
//   missing insertion   Cuda/cudaMissing.cu:125-808: stochastic visibility walk, 1->3 / 2->4
//                       splits, one insertion per triangle per round by 64-bit vote words.
//   flipping            Cuda/cudaFlipping.cu:227-685: Lawson flips with an FP64 filter and an
//                       exact integer fallback on the device (no CPU pass), lifted-perturbation
//                       tie-break for co-circular quadruples.
//   boundary removal    Cuda/cudaBoundary.cu:106-608: distance-2 independent frame-vertex
//                       deletions with a ccw-stack fill, then a hull convexification pass.
// Links after a modification are repaired by one generic pass: a modified triangle keeps the
// old external link of each outer edge (stored as -(e+2)); if that neighbour was modified in
// the same operation the new location is read from its remap[] entry, otherwise the link is
// kept and the neighbour's back-link is written.
#include "dev_common.h"
#include "kernels.h"
#include <algorithm>

// Workgroups go to the 8 XCDs round-robin by blockIdx.x, and each XCD has its own L2.  Grid-
// stride loops renumber their blocks so that XCD k runs the blocks [k G / 8, (k + 1) G / 8) of
// every sweep: neighbouring indices (triangle slots and vertices are numbered in spatial order)
// then share one L2 instead of being spread over all eight.
GA_INLINE int xcd_block()
{
    unsigned g = gridDim.x, b = blockIdx.x;
    return (g & 7u) ? (int)b : (int)((b & 7u) * (g >> 3) + (b >> 3));
}
#define GRID_STRIDE(i, n) for (int i = xcd_block() * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)
#define FIXED_GRID 2048

GA_INLINE int ic(const int4 &a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
GA_INLINE void sc(int4 &a, int i, int v)
{
    if (i == 0) a.x = v;
    else if (i == 1) a.y = v;
    else a.z = v;
}
GA_INLINE int slot_of(const int4 &t, int v) { return t.x == v ? 0 : (t.y == v ? 1 : 2); }
GA_INLINE int mark_old(int e) { return e < 0 ? -1 : -(e + 2); }

GA_INLINE double2 snapped(const MeshDev &m, int v)
{
    const SetDesc &sd = m.sets[m.vset[v]];
    int pp = m.pix[v];
    int px = pp >> 16, py = pp & 0xffff;
    return make_double2(sd.shx + (double)(px - sd.margin) * sd.scale, sd.shy + (double)(py - sd.margin) * sd.scale);
}

// a triangle of v's set that is certainly alive while frame vertices exist: one at the set's
// first frame vertex
GA_INLINE int set_anchor(const MeshDev &m, int v) { return m.vt[m.sets[m.vset[v]].fbase]; }

// input index of vertex v: real vertices are numbered in spatial order (perm maps them back),
// frame vertices keep their ids (above every real index)
GA_INLINE int oid(const MeshDev &m, int v) { return v < m.n ? m.perm[v] : v; }

// triangle slots in use; a reservation past the capacity sets GA_ERR_CAPACITY and leaves the
// counter above it, so loops over the slots stop at the capacity until the host sees the error
GA_INLINE int ntri_used(const MeshDev &m)
{
    int n = *m.ntri;
    return n < m.cap ? n : m.cap;
}

__device__ __noinline__ int orient_slow(double2 a, double2 b, double2 c, int *status)
{
    int ovf = 0;
    int s = orient_exact(a, b, c, &ovf);
    if (ovf) atomicOr(status, GA_ERR_EXACT_OVERFLOW);
    return s;
}

GA_INLINE int orient_dev(const MeshDev &m, double2 a, double2 b, double2 c)
{
    int s = orient_filter(a, b, c);
    if (s != 0) return s;
    stat_inc(m.stats, C_EXACT_ORIENT, true);
    return orient_slow(a, b, c, m.status);
}

__device__ __noinline__ int incircle_slow(double2 a, int ia, double2 b, int ib, double2 c, int ic_,
                                          double2 d, int id, int *status)
{
    int ovf = 0;
    int s = incircle_exact_sos(a, ia, b, ib, c, ic_, d, id, &ovf);
    if (ovf) atomicOr(status, GA_ERR_EXACT_OVERFLOW);
    return s;
}

// a, b, c, d: the positions of vertices ia, ib, ic_, id, already loaded by the caller
GA_INLINE int incircle_at(const MeshDev &m, double2 a, double2 b, double2 c, double2 d, int ia, int ib,
                          int ic_, int id)
{
    int s = incircle_filter(a, b, c, d);
    if (s != 0) return s;
    stat_inc(m.stats, C_EXACT_INCIRCLE, true);
    // the symbolic tie-break ranks vertices by input index
    return incircle_slow(a, oid(m, ia), b, oid(m, ib), c, oid(m, ic_), d, oid(m, id), m.status);
}

GA_INLINE int incircle_dev(const MeshDev &m, int ia, int ib, int ic_, int id)
{
    return incircle_at(m, m.pos[ia], m.pos[ib], m.pos[ic_], m.pos[id], ia, ib, ic_, id);
}

// ---- reconstruction --------------------------------------------------------------------
// Block corners counter-clockwise: a=(x,y) b=(x+1,y) c=(x+1,y+1) d=(x,y+1).
GA_INLINE int recon_count(int a, int b, int c, int d)
{
    if (a == c || b == d) return 0;
    int neq = (a == b) + (b == c) + (c == d) + (d == a);
    if (neq >= 2) return 0;
    return neq == 1 ? 1 : 2;
}

GA_INLINE long long iorient(int pa, int pb, int pc)
{
    long long ax = pa >> 16, ay = pa & 0xffff, bx = pb >> 16, by = pb & 0xffff, cx = pc >> 16, cy = pc & 0xffff;
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
}

// bits at the even positions of v, packed (inverse of the cell numbering's bit spread)
GA_INLINE unsigned compact_bits(unsigned v)
{
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0f0f0f0fu;
    v = (v | (v >> 4)) & 0x00ff00ffu;
    v = (v | (v >> 8)) & 0x0000ffffu;
    return v;
}

// Tiles of 16 x 16 2x2-block origins, one per 16 x 16-pixel cell, in the Morton order of the
// cells (the order the real points are numbered in), so triangles are emitted in the spatial
// order of their vertices.  Set s owns tiles [bm[s], bm[s + 1]) of the BM_RECON map; thread k of
// a tile takes the block at (16 cx + k / 16, 16 cy + k % 16).
GA_INLINE void recon_block(int lt, int *x, int *y)
{
    *x = (int)(compact_bits((unsigned)lt >> 1) << 4) + (int)(threadIdx.x >> 4);
    *y = (int)(compact_bits((unsigned)lt) << 4) + (int)(threadIdx.x & 15);
}

__global__ __launch_bounds__(256) void k_recon_count(const int *__restrict__ vor_all, BatchDev bd,
                                                     int *__restrict__ tile_cnt)
{
    const int *bm = bd.bmap(BM_RECON);
    int set = set_of(bm, bd.nsets, blockIdx.x);
    const SetDesc &sd = bd.sets[set];
    int logT = sd.logT;
    int T = 1 << logT;
    const int *vor = vor_all + sd.toff;
    int tile = blockIdx.x;
    int x, y;
    recon_block(tile - bm[set], &x, &y);
    int c = 0;
    if (x < T - 1 && y < T - 1) {
        size_t i0 = ((size_t)x << logT) + y, i1 = i0 + T;
        c = recon_count(vor[i0], vor[i1], vor[i1 + 1], vor[i0 + 1]);
    }
    __shared__ int ws[4];
    int s = c;
    for (int d = 32; d > 0; d >>= 1) s += __shfl_down(s, d, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[tile] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(256) void k_recon_emit(const int *__restrict__ vor_all, const int *__restrict__ pix,
                                                    BatchDev bd, const int *__restrict__ tile_off,
                                                    int4 *__restrict__ tv, int4 *__restrict__ tn,
                                                    int *__restrict__ status, int cap)
{
    const int *bm = bd.bmap(BM_RECON);
    int set = set_of(bm, bd.nsets, blockIdx.x);
    const SetDesc &sd = bd.sets[set];
    int logT = sd.logT;
    int T = 1 << logT;
    const int *vor = vor_all + sd.toff;
    int tile = blockIdx.x;
    int x, y;
    recon_block(tile - bm[set], &x, &y);
    int c = 0, a = 0, b = 0, cc = 0, d = 0;
    if (x < T - 1 && y < T - 1) {
        size_t i0 = ((size_t)x << logT) + y, i1 = i0 + T;
        a = vor[i0]; b = vor[i1]; cc = vor[i1 + 1]; d = vor[i0 + 1];
        c = recon_count(a, b, cc, d);
    }
    // block exclusive scan of c
    __shared__ int ws[4];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = c;
    for (int k = 1; k < 64; k <<= 1) {
        int t = __shfl_up(inc, k, 64);
        if (lane >= k) inc += t;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    int off = tile_off[tile];
    for (int i = 0; i < w; i++) off += ws[i];
    off += inc - c;
    if (c == 0) return;
    if (off + c > cap) { atomicOr(status, GA_ERR_CAPACITY); return; }
    int4 none = make_int4(-1, -1, -1, 0);
    if (c == 1) {
        int t0, t1, t2;
        if (a == b) { t0 = a; t1 = cc; t2 = d; }
        else if (b == cc) { t0 = a; t1 = b; t2 = d; }
        else { t0 = a; t1 = b; t2 = cc; }   // c == d or d == a
        if (iorient(pix[t0], pix[t1], pix[t2]) <= 0) atomicOr(status, GA_ERR_RECON);
        tv[off] = make_int4(t0, t1, t2, 0);
        tn[off] = none;
    } else {
        int pa = pix[a], pb = pix[b], pc = pix[cc], pd = pix[d];
        if (iorient(pa, pb, pc) > 0 && iorient(pa, pc, pd) > 0) {
            tv[off] = make_int4(a, b, cc, 0);
            tv[off + 1] = make_int4(a, cc, d, 0);
        } else if (iorient(pb, pc, pd) > 0 && iorient(pb, pd, pa) > 0) {
            tv[off] = make_int4(b, cc, d, 0);
            tv[off + 1] = make_int4(b, d, a, 0);
        } else {
            atomicOr(status, GA_ERR_RECON);
            tv[off] = make_int4(a, b, cc, 0);
            tv[off + 1] = make_int4(a, cc, d, 0);
        }
        tn[off] = none;
        tn[off + 1] = none;
    }
}

__global__ void k_set_ntri(const int *tile_cnt, const int *tile_off, int ntiles, int *ntri, int cap)
{
    int F = tile_off[ntiles - 1] + tile_cnt[ntiles - 1];
    *ntri = F > cap ? cap : F;
}

void launch_recon(const int *vor, const int *pix, const BatchDev &b, MeshDev &m, int *tile_cnt, int *tile_off,
                  void *scan_tmp, hipStream_t st)
{
    int ntiles = b.blocks[BM_RECON];
    GA_LAUNCH(k_recon_count, dim3(ntiles), dim3(256), 0, st, vor, b, tile_cnt);
    launch_exclusive_scan(tile_cnt, tile_off, ntiles, scan_tmp, st);
    GA_LAUNCH(k_recon_emit, dim3(ntiles), dim3(256), 0, st, vor, pix, b, tile_off, m.tv, m.tn,
                       m.status, m.cap);
    GA_LAUNCH(k_set_ntri, dim3(1), dim3(1), 0, st, tile_cnt, tile_off, ntiles, m.ntri, m.cap);
}

// ---- adjacency from scratch ----------------------------------------------------------------
// Degrees and fans with the per-vertex atomics combined in LDS first.  A block takes 1024
// consecutive triangle slots; slots and vertices are both in spatial (Morton) order, so their
// vertices fall in a short id range.  Counts for ids in [vmin, vmin + ADJ_WIN) go to an LDS
// window and reach HBM as one atomic per distinct vertex and block; other ids take the global
// atomic directly.  (Device-scope atomics are performed beyond the per-XCD L2, about 11 ns each
// on one address.)
#define ADJ_WIN 8192
#define ADJ_CHUNK 1024

// smallest live vertex id of slots [c0, c0 + ADJ_CHUNK) (block-uniform; every thread calls it)
GA_INLINE int adj_chunk_min(const int4 *__restrict__ tv, int c0, int n, int *s_red)
{
    int mn = 0x7fffffff;
    for (int j = 0; j < ADJ_CHUNK / 256; j++) {
        int t = c0 + j * 256 + (int)threadIdx.x;
        if (t >= n) break;
        int4 v = tv[t];
        if (v.x == GA_DEAD) continue;
        mn = min(mn, min(v.x, min(v.y, v.z)));
    }
    for (int d = 32; d > 0; d >>= 1) mn = min(mn, __shfl_xor(mn, d, 64));
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = mn;
    __syncthreads();
    mn = min(min(s_red[0], s_red[1]), min(s_red[2], s_red[3]));
    return mn;
}

__global__ __launch_bounds__(256) void k_deg(const int4 *__restrict__ tv, const int *__restrict__ ntri,
                                             int *__restrict__ deg)
{
    __shared__ int cnt[ADJ_WIN];
    __shared__ int s_red[4];
    const int n = *ntri;
    for (int c0 = xcd_block() * ADJ_CHUNK; c0 < n; c0 += gridDim.x * ADJ_CHUNK) {
        const int base = adj_chunk_min(tv, c0, n, s_red);
        for (int i = threadIdx.x; i < ADJ_WIN; i += 256) cnt[i] = 0;
        __syncthreads();
        for (int j = 0; j < ADJ_CHUNK / 256; j++) {
            int t = c0 + j * 256 + (int)threadIdx.x;
            if (t >= n) break;
            int4 v = tv[t];
            if (v.x == GA_DEAD) continue;
            for (int i = 0; i < 3; i++) {
                int a = ic(v, i);
                unsigned r = (unsigned)(a - base);
                if (r < ADJ_WIN) atomicAdd(&cnt[r], 1);
                else atomicAdd(&deg[a], 1);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < ADJ_WIN; i += 256) {
            int c = cnt[i];
            if (c) atomicAdd(&deg[base + i], c);
        }
        __syncthreads();
    }
}

// fan entry of corner i of triangle t: (enc(t, i), the vertex that follows it counter-clockwise);
// a corner's place in its vertex's fan is its block's base for the vertex (one global atomic per
// distinct vertex) plus its rank among the block's corners of that vertex (an LDS atomic)
__global__ __launch_bounds__(256) void k_fanfill(const int4 *__restrict__ tv, const int *__restrict__ ntri,
                                                 const int *__restrict__ off, int *__restrict__ fill,
                                                 int2 *__restrict__ fan, int *__restrict__ vt)
{
    __shared__ int cnt[ADJ_WIN];
    __shared__ int s_red[4];
    const int n = *ntri;
    for (int c0 = xcd_block() * ADJ_CHUNK; c0 < n; c0 += gridDim.x * ADJ_CHUNK) {
        const int base = adj_chunk_min(tv, c0, n, s_red);
        for (int i = threadIdx.x; i < ADJ_WIN; i += 256) cnt[i] = 0;
        __syncthreads();
        int lr[ADJ_CHUNK / 256][3];
#pragma unroll
        for (int j = 0; j < ADJ_CHUNK / 256; j++) {
            lr[j][0] = lr[j][1] = lr[j][2] = -1;
            int t = c0 + j * 256 + (int)threadIdx.x;
            if (t >= n) continue;
            int4 v = tv[t];
            if (v.x == GA_DEAD) continue;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                int a = ic(v, i);
                unsigned r = (unsigned)(a - base);
                if (r < ADJ_WIN) {
                    lr[j][i] = atomicAdd(&cnt[r], 1);
                } else {
                    int k = atomicAdd(&fill[a], 1);
                    fan[off[a] + k] = make_int2(enc(t, i), ic(v, p1(i)));
                    if (k == 0) vt[a] = t;
                }
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < ADJ_WIN; i += 256) {
            int c = cnt[i];
            if (c) cnt[i] = atomicAdd(&fill[base + i], c);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ADJ_CHUNK / 256; j++) {
            int t = c0 + j * 256 + (int)threadIdx.x;
            if (t >= n) continue;
            if (lr[j][0] < 0 && lr[j][1] < 0 && lr[j][2] < 0) continue;
            int4 v = tv[t];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                if (lr[j][i] < 0) continue;
                int a = ic(v, i);
                int k = cnt[a - base] + lr[j][i];
                fan[off[a] + k] = make_int2(enc(t, i), ic(v, p1(i)));
                if (k == 0) vt[a] = t;
            }
        }
        __syncthreads();
    }
}

// Edge i of t runs a -> b (a = v[i + 1], b = v[i + 2]); its twin runs b -> a in the triangle u
// whose fan entry at b has a as its next vertex, and lies opposite the corner before b in u.
__global__ void k_match(const int4 *__restrict__ tv, const int *__restrict__ ntri, const int *__restrict__ off,
                        const int2 *__restrict__ fan, int4 *__restrict__ tn)
{
    int n = *ntri;
    GRID_STRIDE(t, n) {
        int4 v = tv[t];
        if (v.x == GA_DEAD) continue;
        int4 r = make_int4(-1, -1, -1, 0);
        for (int i = 0; i < 3; i++) {
            int a = ic(v, p1(i)), b = ic(v, m1(i));
            for (int k = off[b], k1 = off[b + 1]; k < k1; k++) {
                int2 f = fan[k];
                if (f.y == a) { sc(r, i, enc(enc_t(f.x), m1(enc_j(f.x)))); break; }
            }
        }
        tn[t] = r;
    }
}

void launch_adjacency(MeshDev &m, int *deg, int *off, int *fill, int2 *fan, void *scan_tmp, hipStream_t st)
{
    IntFills f;
    f.add(deg, (size_t)m.V + 1, 0);
    f.add(fill, (size_t)m.V + 1, 0);
    launch_fill_multi(f, st);
    GA_LAUNCH(k_deg, dim3(FIXED_GRID), dim3(256), 0, st, m.tv, m.ntri, deg);
    launch_exclusive_scan(deg, off, m.V + 1, scan_tmp, st);
    GA_LAUNCH(k_fanfill, dim3(FIXED_GRID), dim3(256), 0, st, m.tv, m.ntri, off, fill, fan, m.vt);
    GA_LAUNCH(k_match, dim3(FIXED_GRID), dim3(256), 0, st, m.tv, m.ntri, off, fan, m.tn);
}

// ---- positions / states after reconstruction ------------------------------------------------
// deg (from the adjacency build, still valid) says which vertices are in the mesh.
__global__ void k_init_positions(MeshDev m, const int *__restrict__ deg)
{
    GRID_STRIDE(v, m.V) {
        int in = deg[v] > 0;
        if (v >= m.n) {
            m.vstate[v] = VS_IN_MESH;
            m.pos[v] = snapped(m, v);
        } else if (in) {
            m.vstate[v] = VS_PENDING;
            m.pos[v] = snapped(m, v);
        } else {
            m.vstate[v] = VS_MISSING;
            m.pos[v] = m.xy[v];
            m.vt[v] = -1;
        }
    }
}

void launch_init_positions(MeshDev &m, hipStream_t st)
{
    GA_LAUNCH(k_init_positions, dim3(FIXED_GRID), dim3(256), 0, st, m, (const int *)m.deg);
}

// ---- shifting ----------------------------------------------------------------------------------
GA_INLINE double2 cand_pos(const MeshDev &m, const unsigned char *moved, int v)
{
    if (v < m.n && moved[v]) return m.xy[v];
    return m.pos[v];
}

// Bulk shifting: every pending vertex tries its true position at once (moved = 1); a triangle
// that is not counter-clockwise with the current choice reverts its moved vertices, until no
// triangle does.  Reverting only ever clears moved[], so the first pass looks at every triangle
// and later passes only at the stars of the vertices reverted by the pass before (a worklist;
// mark[] lists each vertex once).
// also zeroes the revert marks, the count of k_shift_bulk's output list and that of the first
// star pass's
__global__ void k_shift_init(MeshDev m, unsigned char *moved, int *mark, int cnt_out_slot, int cnt_next_slot)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        m.counters[cnt_out_slot] = 0;
        m.counters[cnt_next_slot] = 0;
    }
    GRID_STRIDE(v, m.n) moved[v] = (m.vstate[v] == VS_PENDING) ? 1 : 0;
    GRID_STRIDE(v, m.V) mark[v] = 0;
}

// test triangle v (vertex ids) with the current choice; revert and list its moved vertices when
// it is not counter-clockwise
GA_INLINE void shift_check(const MeshDev &m, unsigned char *moved, int *mark, int4 v, int *wl_out,
                           int cnt_out_slot)
{
    bool any = (v.x < m.n && moved[v.x]) || (v.y < m.n && moved[v.y]) || (v.z < m.n && moved[v.z]);
    unsigned got = 0;   // bit k: vertex k of the triangle reverted here and listed
    if (any) {
        double2 a = cand_pos(m, moved, v.x), b = cand_pos(m, moved, v.y), cc = cand_pos(m, moved, v.z);
        if (orient_dev(m, a, b, cc) <= 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                int u = ic(v, k);
                if (u < m.n && moved[u]) {
                    moved[u] = 0;
                    if (atomicExch(&mark[u], 1) == 0) got |= 1u << k;
                }
            }
        }
    }
    int q = wave_reserve_n<2>(&m.counters[cnt_out_slot], __popc(got));
#pragma unroll
    for (int k = 0; k < 3; k++)
        if ((got >> k) & 1) wl_out[q++] = ic(v, k);
}

__global__ void k_shift_bulk(MeshDev m, unsigned char *moved, int *mark, int *wl_out, int cnt_out_slot)
{
    int n = ntri_used(m);
    GRID_STRIDE(t, n) {
        int4 v = m.tv[t];
        if (v.x == GA_DEAD) continue;
        shift_check(m, moved, mark, v, wl_out, cnt_out_slot);
    }
}

// cnt_next_slot: the output count of the next pass (neither this pass's input nor its output),
// zeroed here so that no launch is needed between passes
__global__ void k_shift_star(MeshDev m, unsigned char *moved, int *mark, const int *wl_in, int cnt_in_slot,
                             int *wl_out, int cnt_out_slot, int cnt_next_slot)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) m.counters[cnt_next_slot] = 0;
    int nw = m.counters[cnt_in_slot];
    GRID_STRIDE(k, nw) {
        int v = wl_in[k];
        int t0 = m.vt[v], t = t0, i = slot_of(m.tv[t0], v), guard = 0;
        do {
            shift_check(m, moved, mark, m.tv[t], wl_out, cnt_out_slot);
            int e = ic(m.tn[t], p1(i));
            if (e < 0) { atomicOr(m.status, GA_ERR_WALK); break; }
            t = enc_t(e);
            i = p1(enc_j(e));
            if (++guard > 100000) { atomicOr(m.status, GA_ERR_WALK); break; }
        } while (t != t0);
    }
}

void launch_shift_bulk(MeshDev &m, unsigned char *moved, int *mark, int *wl_out, int cnt_out_slot,
                       int cnt_next_slot, hipStream_t st)
{
    GA_LAUNCH(k_shift_init, dim3(FIXED_GRID), dim3(256), 0, st, m, moved, mark, cnt_out_slot, cnt_next_slot);
    GA_LAUNCH(k_shift_bulk, dim3(FIXED_GRID), dim3(256), 0, st, m, moved, mark, wl_out, cnt_out_slot);
}

void launch_shift_star(MeshDev &m, unsigned char *moved, int *mark, const int *wl_in, int cnt_in_slot, int *wl_out,
                       int cnt_out_slot, int cnt_next_slot, hipStream_t st)
{
    GA_LAUNCH(k_shift_star, dim3(FIXED_GRID), dim3(256), 0, st, m, moved, mark, wl_in, cnt_in_slot, wl_out,
              cnt_out_slot, cnt_next_slot);
}

__global__ void k_shift_commit(MeshDev m, const unsigned char *moved)
{
    GRID_STRIDE(v, m.n) {
        bool pend = m.vstate[v] == VS_PENDING;
        if (pend && moved[v]) {
            m.pos[v] = m.xy[v];
            m.vstate[v] = VS_IN_MESH;
            pend = false;
        }
        stat_inc(m.stats, C_PEND, pend);
    }
}

void launch_shift_commit(MeshDev &m, const unsigned char *moved, hipStream_t st)
{
    GA_LAUNCH(k_shift_commit, dim3(FIXED_GRID), dim3(256), 0, st, m, moved);
}

// The greedy and deletion rounds only look at vertices that are still pending after bulk
// shifting (a few percent), listed once; a listed vertex that has left the pending state since is
// skipped by its state test.
__global__ __launch_bounds__(256) void k_pending_list(MeshDev m)
{
    for (int b0 = blockIdx.x * 1024; b0 < m.n; b0 += gridDim.x * 1024) {
        unsigned bits = 0;
        for (int j = 0; j < 4; j++) {
            int v = b0 + j * 256 + (int)threadIdx.x;
            if (v < m.n && m.vstate[v] == VS_PENDING) bits |= 1u << j;
        }
        int k = block_reserve(&m.counters[C_PLIST], __popc(bits));
        for (int j = 0; j < 4; j++)
            if ((bits >> j) & 1) m.plist[k++] = b0 + j * 256 + (int)threadIdx.x;
    }
}

void launch_pending_list(MeshDev &m, hipStream_t st)
{
    GA_LAUNCH(k_pending_list, dim3(FIXED_GRID), dim3(256), 0, st, m);
}

GA_INLINE uint32_t vprio(int v, int round) { return hash32((uint32_t)v * 0x9E3779B1u ^ hash32((uint32_t)round + 0x1234567u)); }

// Greedy round: a pending vertex whose pending neighbours all have lower priority moves to
// its true position when all its star triangles stay counter-clockwise.  Decisions only;
// the commit kernel applies them (so every decision sees the same snapshot).
__global__ void k_shift_greedy(MeshDev m, int round, int *decide)
{
    int np = m.counters[C_PLIST];
    GRID_STRIDE(k, np) {
        decide[k] = 0;
        int v = m.plist[k];
        if (m.vstate[v] != VS_PENDING) continue;
        uint32_t p = vprio(v, round);
        int t0 = m.vt[v];
        int t = t0, i = slot_of(m.tv[t0], v);
        bool ok = true, winner = true;
        double2 q = m.xy[v];
        int guard = 0;
        do {
            int4 tt = m.tv[t];
            int a = ic(tt, p1(i)), b = ic(tt, m1(i));
            if (a < m.n && m.vstate[a] == VS_PENDING) {
                uint32_t pa = vprio(a, round);
                if (pa > p || (pa == p && a > v)) { winner = false; break; }
            }
            if (ok && orient_dev(m, q, m.pos[a], m.pos[b]) <= 0) ok = false;
            int e = ic(m.tn[t], p1(i));
            if (e < 0) { atomicOr(m.status, GA_ERR_WALK); winner = false; break; }
            t = enc_t(e);
            i = p1(enc_j(e));
            if (++guard > 100000) { atomicOr(m.status, GA_ERR_WALK); winner = false; break; }
        } while (t != t0);
        if (winner && ok) decide[k] = 1;
    }
}

__global__ void k_shift_greedy_commit(MeshDev m, const int *decide)
{
    int np = m.counters[C_PLIST];
    GRID_STRIDE(k, np) {
        bool d = decide[k] != 0;
        if (d) {
            int v = m.plist[k];
            m.pos[v] = m.xy[v];
            m.vstate[v] = VS_IN_MESH;
        }
        stat_inc(m.stats, C_MOVED, d);
    }
}

void launch_shift_greedy(MeshDev &m, int round, int grid, hipStream_t st)
{
    GA_LAUNCH(k_shift_greedy, dim3(grid), dim3(256), 0, st, m, round, m.decide);
    GA_LAUNCH(k_shift_greedy_commit, dim3(grid), dim3(256), 0, st, m, (const int *)m.decide);
}

// ---- deletion of vertices that cannot move (star-shaped hole, ear clipping) --------------------
// Priorities of one round of vertex removal: the round number (< 2^20, increasing within a run)
// in the top bits, so the per-vertex maxima left by earlier rounds always lose and mx[] is
// cleared once per run; then 12 random bits; the vertex id breaks ties.
GA_INLINE unsigned long long prio64(int v, int round)
{
    return ((unsigned long long)round << 44) | ((unsigned long long)(vprio(v, round) & 0xfffu) << 32) | (unsigned)v;
}

// distance-2 independence: every vertex records the largest priority in its closed neighbourhood
// also zeroes the scratch and modified-list counts that k_del_apply of this round fills
__global__ void k_del_prio(MeshDev m, unsigned long long *mx, int round)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        m.counters[C_SCR] = 0;
        m.counters[C_MODN] = 0;
    }
    int np = m.counters[C_PLIST];
    GRID_STRIDE(k, np) {
        int v = m.plist[k];
        if (m.vstate[v] != VS_PENDING) continue;
        unsigned long long p = prio64(v, round);
        atomicMax(&mx[v], p);
        int t0 = m.vt[v], t = t0, i = slot_of(m.tv[t0], v), guard = 0;
        do {
            int a = ic(m.tv[t], p1(i));
            atomicMax(&mx[a], p);
            int e = ic(m.tn[t], p1(i));
            if (e < 0) { atomicOr(m.status, GA_ERR_WALK); break; }
            t = enc_t(e);
            i = p1(enc_j(e));
            if (++guard > 100000) break;
        } while (t != t0);
    }
}

// link two (triangle, slot) sides of one edge; X < 0 means hull
// single-word stores: other threads may write other slots of the same neighbour record
GA_INLINE void set_link(const MeshDev &m, int T, int k, int v) { ((int *)(m.tn + T))[k] = v; }
GA_INLINE void bond(const MeshDev &m, int T, int k, int X)
{
    set_link(m, T, k, X);
    if (X >= 0) set_link(m, enc_t(X), enc_j(X), enc(T, k));
}

__global__ void k_del_apply(MeshDev m, const unsigned long long *mx, int round, int op, int *list_out)
{
    int np = m.counters[C_PLIST];
    GRID_STRIDE(kk, np) {
        int v = m.plist[kk];
        if (m.vstate[v] != VS_PENDING) continue;
        unsigned long long p = prio64(v, round);
        if (mx[v] != p) continue;
        int t0 = m.vt[v], t = t0, i = slot_of(m.tv[t0], v), d = 0, guard = 0;
        bool win = true;
        do {
            int a = ic(m.tv[t], p1(i));
            if (mx[a] != p) { win = false; break; }
            int e = ic(m.tn[t], p1(i));
            if (e < 0) { atomicOr(m.status, GA_ERR_WALK); win = false; break; }
            t = enc_t(e);
            i = p1(enc_j(e));
            d++;
            if (++guard > 100000) { win = false; break; }
        } while (t != t0);
        if (!win) continue;
        // scratch: poly[d], plink[d], slots[d]; list slots for the d - 2 new triangles
        int base = wave_reserve_n<28>(&m.counters[C_SCR], 3 * d);
        int lbase = wave_reserve_n<28>(&m.counters[C_MODN], d >= 2 ? d - 2 : 0);
        if ((size_t)(base + 3 * d) > m.scratch_cap) {
            atomicOr(m.status, GA_ERR_SCRATCH);
            for (int k = 0; k < d - 2; k++) list_out[lbase + k] = -1;
            continue;
        }
        int *poly = m.scratch + base, *plink = poly + d, *slots = plink + d;
        t = t0; i = slot_of(m.tv[t0], v);
        for (int k = 0; k < d; k++) {
            int4 tt = m.tv[t];
            poly[k] = ic(tt, p1(i));                 // triangle (v, l_k, l_k+1)
            plink[k] = ic(m.tn[t], i);               // outer side of edge (l_k, l_k+1)
            slots[k] = t;
            int e = ic(m.tn[t], p1(i));
            t = enc_t(e);
            i = p1(enc_j(e));
        }
        // The hole is star-shaped from K (v's position): K lies left of every polygon edge.
        // An ear (A, B, C) is cut when it is counter-clockwise and K is left of or on the new
        // edge A -> C; that keeps K left of or on every edge of the remaining polygon, so the
        // ear holds no other polygon vertex (they are in angular order around K) and an ear
        // always exists.  (K strictly inside the ear is excluded by the same test.)
        double2 K = m.pos[v];
        int msz = d, used = 0;
        while (msz > 3) {
            int pick = -1;
            for (int k = 0; k < msz && pick < 0; k++) {
                int kp = k == 0 ? msz - 1 : k - 1, kn = k == msz - 1 ? 0 : k + 1;
                double2 A = m.pos[poly[kp]], B = m.pos[poly[k]], C = m.pos[poly[kn]];
                if (orient_dev(m, A, B, C) <= 0) continue;
                if (orient_dev(m, A, C, K) < 0) continue;
                pick = k;
            }
            if (pick < 0) {
                atomicOr(m.status, GA_ERR_PATCH);
                if (atomicCAS(&m.counters[C_DBG], 0, 1) == 0) {
                    m.counters[C_DBG + 1] = v;
                    m.counters[C_DBG + 2] = msz;
                    for (int k = 0; k < msz && k < 12; k++) m.counters[C_DBG + 3 + k] = poly[k];
                    m.dbg[0] = K.x; m.dbg[1] = K.y;
                    for (int k = 0; k < msz && k < 15; k++) {
                        m.dbg[2 + 2 * k] = m.pos[poly[k]].x;
                        m.dbg[3 + 2 * k] = m.pos[poly[k]].y;
                        m.dbg[32 + k] = (double)m.vstate[poly[k]];
                    }
                }
                break;
            }
            int kp = pick == 0 ? msz - 1 : pick - 1, kn = pick == msz - 1 ? 0 : pick + 1;
            int T = slots[used++];
            int A = poly[kp], B = poly[pick], C = poly[kn];
            m.tv[T] = make_int4(A, B, C, 0);
            m.tn[T] = make_int4(-1, -1, -1, 0);
            // slot 2 (opp C) = edge (A,B) -> plink[kp]; slot 0 (opp A) = edge (B,C) -> plink[pick]
            bond(m, T, 2, plink[kp]);
            bond(m, T, 0, plink[pick]);
            m.vt[A] = T; m.vt[B] = T; m.vt[C] = T;
            m.stamp[T] = op;
            list_out[lbase + used - 1] = T;
            // new polygon edge (A, C): its other side is slot 1 (opp B) of T
            plink[kp] = enc(T, 1);
            for (int k = pick; k < msz - 1; k++) { poly[k] = poly[k + 1]; plink[k] = plink[k + 1]; }
            msz--;
        }
        if (msz == 3) {
            if (orient_dev(m, m.pos[poly[0]], m.pos[poly[1]], m.pos[poly[2]]) <= 0) atomicOr(m.status, GA_ERR_PATCH);
            int T = slots[used++];
            m.tv[T] = make_int4(poly[0], poly[1], poly[2], 0);
            m.tn[T] = make_int4(-1, -1, -1, 0);
            bond(m, T, 2, plink[0]);   // (p0,p1)
            bond(m, T, 0, plink[1]);   // (p1,p2)
            bond(m, T, 1, plink[2]);   // (p2,p0)
            m.vt[poly[0]] = T; m.vt[poly[1]] = T; m.vt[poly[2]] = T;
            m.stamp[T] = op;
            list_out[lbase + used - 1] = T;
        }
        for (int k = used; k < d - 2; k++) list_out[lbase + k] = -1;   // only after a failure
        for (int k = used; k < d; k++) m.tv[slots[k]] = make_int4(GA_DEAD, GA_DEAD, GA_DEAD, 0);
        m.vstate[v] = VS_MISSING;
        m.vt[v] = slots[0];      // anchor for the re-insertion walk
        m.pos[v] = m.xy[v];
        stat_inc(m.stats, C_DELETED, true);
    }
}

void launch_delete_round(MeshDev &m, unsigned int *mxraw, int round, int op, int grid, hipStream_t st)
{
    unsigned long long *mx = (unsigned long long *)mxraw;
    GA_LAUNCH(k_del_prio, dim3(grid), dim3(256), 0, st, m, mx, round);
    GA_LAUNCH(k_del_apply, dim3(grid), dim3(256), 0, st, m, (const unsigned long long *)mx, round, op,
                       m.dellist);
}

// ---- missing-site insertion ------------------------------------------------------------------
// list position k for missing vertex v and its walk start: the anchor a deleted vertex keeps,
// else a triangle of the nearest labelled pixel
GA_INLINE void missing_anchor(const MeshDev &m, const int *__restrict__ vor, int v, int k, int *mlist, int *loc)
{
    mlist[k] = v;
    // a deleted vertex keeps an anchor triangle near its old position in vt[]
    int t = m.vt[v];
    if (t < 0 || m.tv[t].x == GA_DEAD) {
        // the nearest labelled pixel (square rings around v's pixel, radius <= 16) whose
        // site is in the mesh: its triangles are next to v, so the walk stays short
        const SetDesc &sd = m.sets[m.vset[v]];
        int T = 1 << sd.logT;
        int pp = m.pix[v];
        int px = pp >> 16, py = pp & 0xffff;
        t = -1;
        for (int r = 0; r <= 16 && t < 0; r++) {
            for (int q = 0; q < (r == 0 ? 1 : 8 * r) && t < 0; q++) {
                int dx, dy;
                if (r == 0) { dx = 0; dy = 0; }
                else {
                    int side = q / (2 * r), o = q % (2 * r) - r;
                    if (side == 0) { dx = o; dy = -r; }
                    else if (side == 1) { dx = r; dy = o; }
                    else if (side == 2) { dx = -o; dy = r; }
                    else { dx = -r; dy = -o; }
                }
                int x = px + dx, y = py + dy;
                if (x < 0 || y < 0 || x >= T || y >= T) continue;
                int s = vor[sd.toff + ((size_t)x << sd.logT) + y];
                if (s >= 0 && m.vstate[s] == VS_IN_MESH) t = m.vt[s];
            }
        }
    }
    if (t < 0) t = set_anchor(m, v);
    loc[k] = t;
}

__global__ __launch_bounds__(256) void k_missing_init(MeshDev m, const int *__restrict__ vor, int *mlist, int *loc)
{
    // block-uniform loop: 1024 vertices per block and step, one list reservation per step
    for (int b0 = blockIdx.x * 1024; b0 < m.n; b0 += gridDim.x * 1024) {
        unsigned bits = 0;
        for (int j = 0; j < 4; j++) {
            int v = b0 + j * 256 + (int)threadIdx.x;
            if (v < m.n && m.vstate[v] == VS_MISSING) bits |= 1u << j;
        }
        int k0 = block_reserve(&m.counters[C_MISS], __popc(bits));
        for (int j = 0; j < 4; j++) {
            if (!((bits >> j) & 1)) continue;
            int v = b0 + j * 256 + (int)threadIdx.x;
            missing_anchor(m, vor, v, k0 + __popc(bits & ((1u << j) - 1u)), mlist, loc);
        }
    }
}

void launch_missing_init(MeshDev &m, const int *vor, int *mlist, int *loc, hipStream_t st)
{
    GA_LAUNCH(k_missing_init, dim3(FIXED_GRID), dim3(256), 0, st, m, vor, mlist, loc);
}

// located: lt = (type << 2) | edge ; type 1 inside, 2 on edge, 3 duplicate ; -1 not located
// also zeroes the modified-slot count that k_insert of this round fills (the previous round's
// relink, which read it, has finished)
__global__ void k_locate(MeshDev m, const int *mlist, int *loc, int *lt, int op)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) m.counters[C_MODN] = 0;
    int nm = m.counters[C_MISS];
    GRID_STRIDE(k, nm) {
        int v = mlist[k];
        lt[k] = -1;
        if (m.vstate[v] != VS_MISSING) continue;
        double2 p = m.xy[v];
        int t = loc[k];
        if (t < 0 || m.tv[t].x == GA_DEAD) t = set_anchor(m, v);
        uint32_t rng = hash32((uint32_t)v ^ ((uint32_t)op * 0x85ebca6bu));
        int located = 0, o[3];
        for (int step = 0; step < 4096; step++) {
            int4 tt = m.tv[t];
            rng = rng * 1664525u + 1013904223u;
            int r = (rng >> 16) % 3;
            int moved = 0;
            for (int e = 0; e < 3; e++) {
                int i = (r + e) % 3;
                o[i] = orient_dev(m, m.pos[ic(tt, p1(i))], m.pos[ic(tt, m1(i))], p);
                if (o[i] < 0) {
                    int nb = ic(m.tn[t], i);
                    if (nb < 0) { atomicOr(m.status, GA_ERR_WALK); moved = -1; break; }
                    t = enc_t(nb);
                    moved = 1;
                    break;
                }
            }
            if (moved == 0) { located = 1; break; }
            if (moved < 0) break;
        }
        loc[k] = t;
        if (!located) continue;
        int z = (o[0] == 0) + (o[1] == 0) + (o[2] == 0);
        int4 tt = m.tv[t];
        if (z >= 2) {
            int kk = o[0] != 0 ? 0 : (o[1] != 0 ? 1 : 2);
            int w = ic(tt, kk);
            m.vstate[v] = VS_DUP;
            m.vt[v] = -1 - w;      // duplicate of w
            stat_inc(m.stats, C_DUPS, true);
            continue;
        }
        unsigned long long key = ((unsigned long long)op << 32) | (0xffffffffu - (unsigned)oid(m, v));
        if (z == 0) {
            atomicMax(&m.claim[t], key);
            lt[k] = (1 << 2);
        } else {
            int i = o[0] == 0 ? 0 : (o[1] == 0 ? 1 : 2);
            int nb = ic(m.tn[t], i);
            if (nb < 0) { atomicOr(m.status, GA_ERR_WALK); continue; }
            atomicMax(&m.claim[t], key);
            atomicMax(&m.claim[enc_t(nb)], key);
            lt[k] = (2 << 2) | i;
        }
    }
}

__global__ void k_insert(MeshDev m, const int *mlist, const int *loc, const int *lt, int op, int *list_out)
{
    int nm = m.counters[C_MISS];
    GRID_STRIDE(k, nm) {
        int L = lt[k];
        if (L < 0) continue;
        int v = mlist[k];
        int t = loc[k];
        unsigned long long key = ((unsigned long long)op << 32) | (0xffffffffu - (unsigned)oid(m, v));
        int type = L >> 2, i = L & 3;
        int4 tt = m.tv[t], nt = m.tn[t];
        bool win = m.claim[t] == key;
        if (win && type == 2) win = m.claim[enc_t(ic(nt, i))] == key;
        // four list slots and two new triangle slots per winner, one atomic each per wave
        int q0 = wave_reserve_mult(&m.counters[C_MODN], win, 4);
        int q1 = q0 + 1, q2 = q0 + 2, q3 = q0 + 3;
        int s = wave_reserve_mult(m.ntri, win, 2);
        if (!win) continue;
        if (type == 1) {
            if (s + 2 > m.cap) {
                atomicOr(m.status, GA_ERR_CAPACITY);
                list_out[q0] = list_out[q1] = list_out[q2] = list_out[q3] = -1;
                continue;
            }
            int s1 = s, s2 = s + 1;
            int a = tt.x, b = tt.y, c = tt.z;
            m.tv[t] = make_int4(a, b, v, 0);
            m.tv[s1] = make_int4(b, c, v, 0);
            m.tv[s2] = make_int4(c, a, v, 0);
            m.tn[t] = make_int4(enc(s1, 1), enc(s2, 0), mark_old(nt.z), 0);
            m.tn[s1] = make_int4(enc(s2, 1), enc(t, 0), mark_old(nt.x), 0);
            m.tn[s2] = make_int4(enc(t, 1), enc(s1, 0), mark_old(nt.y), 0);
            m.remap[t] = make_int4(enc(s1, 2), enc(s2, 2), enc(t, 2), 0);
            m.stamp[t] = op; m.stamp[s1] = op; m.stamp[s2] = op;
            list_out[q0] = t; list_out[q1] = s1; list_out[q2] = s2; list_out[q3] = -1;
            m.vt[v] = t; m.vt[c] = s1;
        } else {
            int nb = ic(nt, i);
            int u = enc_t(nb), j = enc_j(nb);
            if (s + 2 > m.cap) {
                atomicOr(m.status, GA_ERR_CAPACITY);
                list_out[q0] = list_out[q1] = list_out[q2] = list_out[q3] = -1;
                continue;
            }
            int s1 = s, s2 = s + 1;
            int4 ut = m.tv[u], un = m.tn[u];
            int c = ic(tt, i), a = ic(tt, p1(i)), b = ic(tt, m1(i));
            int d = ic(ut, j);
            int o_bc = ic(nt, p1(i)), o_ca = ic(nt, m1(i));
            int o_ad = ic(un, p1(j)), o_db = ic(un, m1(j));
            m.tv[t] = make_int4(c, a, v, 0);
            m.tv[s1] = make_int4(c, v, b, 0);
            m.tv[u] = make_int4(d, b, v, 0);
            m.tv[s2] = make_int4(d, v, a, 0);
            m.tn[t] = make_int4(enc(s2, 0), enc(s1, 2), mark_old(o_ca), 0);
            m.tn[s1] = make_int4(enc(u, 0), mark_old(o_bc), enc(t, 1), 0);
            m.tn[u] = make_int4(enc(s1, 0), enc(s2, 2), mark_old(o_db), 0);
            m.tn[s2] = make_int4(enc(t, 0), mark_old(o_ad), enc(u, 1), 0);
            int4 rt = make_int4(0, 0, 0, 0), ru = make_int4(0, 0, 0, 0);
            sc(rt, i, -1); sc(rt, p1(i), enc(s1, 1)); sc(rt, m1(i), enc(t, 2));
            sc(ru, j, -1); sc(ru, p1(j), enc(s2, 1)); sc(ru, m1(j), enc(u, 2));
            m.remap[t] = rt; m.remap[u] = ru;
            m.stamp[t] = op; m.stamp[u] = op; m.stamp[s1] = op; m.stamp[s2] = op;
            list_out[q0] = t; list_out[q1] = s1; list_out[q2] = u; list_out[q3] = s2;
            m.vt[v] = t; m.vt[a] = t; m.vt[c] = t; m.vt[b] = u; m.vt[d] = u;
        }
        m.vstate[v] = VS_IN_MESH;
        m.pos[v] = m.xy[v];
        stat_inc(m.stats, C_INSERTED, true);
    }
}

void launch_insert_round(MeshDev &m, int *mlist, int *loc, int op, int grid, hipStream_t st)
{
    GA_LAUNCH(k_locate, dim3(grid), dim3(256), 0, st, m, (const int *)mlist, loc, m.lt, op);
    GA_LAUNCH(k_insert, dim3(grid), dim3(256), 0, st, m, (const int *)mlist, (const int *)loc,
                       (const int *)m.lt, op, m.modA);
}

// ---- generic link repair -------------------------------------------------------------------
GA_INLINE void relink_body(const MeshDev &m, const int *list, int n, int op)
{
    GRID_STRIDE(q, n) {
        int t = list[q];
        if (t < 0) continue;
        int4 r = m.tn[t];
        bool ch = false;
        for (int k = 0; k < 3; k++) {
            int e = ic(r, k);
            if (e > -2) continue;
            int old = -e - 2;
            int B = enc_t(old), jb = enc_j(old);
            if (m.stamp[B] == op) {
                sc(r, k, ic(m.remap[B], jb));
            } else {
                sc(r, k, old);
                set_link(m, B, jb, enc(t, k));
            }
            ch = true;
        }
        if (ch) m.tn[t] = r;
    }
}

__global__ void k_relink(MeshDev m, const int *list, int cnt_slot, int op, int zero_slot)
{
    int n = m.counters[cnt_slot];
    if (zero_slot >= 0 && blockIdx.x == 0 && threadIdx.x == 0) m.counters[zero_slot] = 0;
    relink_body(m, list, n, op);
}

void launch_relink(MeshDev &m, const int *modlist, int cnt_slot, int op, int grid, hipStream_t st)
{
    GA_LAUNCH(k_relink, dim3(grid), dim3(256), 0, st, m, modlist, cnt_slot, op, -1);
}

// ---- constraint edge set ----------------------------------------------------------------------
// key of the undirected edge {a, b}; never 0 (the empty mark)
GA_INLINE unsigned long long cons_key(int a, int b)
{
    unsigned lo = (unsigned)min(a, b), hi = (unsigned)max(a, b);
    return (((unsigned long long)lo << 32) | hi) + 1ull;
}

GA_INLINE unsigned long long cons_mix(unsigned long long k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    return k ^ (k >> 33);
}

// true when {a, b} is a constraint; the table is at least twice the number of constraints, so
// every probe sequence reaches an empty slot
GA_INLINE bool cons_has(const MeshDev &m, int a, int b)
{
    unsigned long long k = cons_key(a, b);
    for (unsigned long long h = cons_mix(k);; h++) {
        unsigned long long s = m.chash[h & (unsigned long long)m.chmask];
        if (s == k) return true;
        if (s == 0) return false;
    }
}

// ---- flipping ---------------------------------------------------------------------------------
// every slot in use (dead ones are skipped by the test); no atomics: the list is the identity
__global__ void k_flip_all(MeshDev m, int op, int *list, int cnt_slot)
{
    int n = ntri_used(m);
    GRID_STRIDE(t, n) {
        m.stamp[t] = op;
        list[t] = t;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) m.counters[cnt_slot] = n;
}

void launch_flip_all(MeshDev &m, int op, int *list, hipStream_t st)
{
    GA_LAUNCH(k_flip_all, dim3(FIXED_GRID), dim3(256), 0, st, m, op, list, C_MODN);
}

// Candidates are the listed triangles' edges that fail the incircle test (each edge once: the
// lower slot of a pair that were both modified in the previous pass skips it).  Block-uniform
// loop over 256 * per list entries per block and step with one reservation of candidate slots
// per step: per = 4 for lists longer than four times the grid (fewer reservations), else 1 (more
// waves in flight for the latency-bound gathers).
// block-uniform: every thread of the block runs the same iterations (block_reserve inside)
GA_INLINE void flip_test_body(const MeshDev &m, const int *list, int n, int prev_op, int op)
{
    int tests = 0;
    const int bd = (int)blockDim.x;
    int per = n > 4 * (int)(gridDim.x * bd) ? 4 : 1;
    for (int b0 = xcd_block() * bd * per; b0 < n; b0 += gridDim.x * bd * per) {
        unsigned bits = 0;   // bit 3j + i: edge i of the j-th triangle of this thread wants a flip
        int tj[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            tj[j] = -1;
            if (j >= per) continue;
            int q = b0 + j * bd + (int)threadIdx.x;
            int t = q < n ? list[q] : -1;
            tj[j] = t;
            if (t < 0) continue;
            int4 tt = m.tv[t];
            if (tt.x == GA_DEAD) continue;
            int4 nn = m.tn[t];
            // The edges to test first, then every gather they need before any predicate: the
            // corner positions once, the three opposite vertices, then their positions, so the
            // three dependent chains are in flight together instead of one after another behind
            // the claim atomics.
            int uu[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                int e = ic(nn, i);
                uu[i] = -1;
                if (e < 0) continue;
                int u = enc_t(e);
                if (m.stamp[u] == prev_op && u < t) continue;
                if (m.chash && cons_has(m, ic(tt, p1(i)), ic(tt, m1(i)))) continue;   // constraint edge
                uu[i] = u;
            }
            if ((uu[0] & uu[1] & uu[2]) < 0) continue;   // no edge to test
            int dd[3];
#pragma unroll
            for (int i = 0; i < 3; i++) dd[i] = uu[i] >= 0 ? ic(m.tv[uu[i]], enc_j(ic(nn, i))) : 0;
            double2 pa = m.pos[tt.x], pb = m.pos[tt.y], pc = m.pos[tt.z];
            double2 pd[3];
#pragma unroll
            for (int i = 0; i < 3; i++) pd[i] = uu[i] >= 0 ? m.pos[dd[i]] : make_double2(0.0, 0.0);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                if (uu[i] < 0) continue;
                tests++;
                if (incircle_at(m, pa, pb, pc, pd[i], tt.x, tt.y, tt.z, dd[i]) > 0) {
                    unsigned long long key = ((unsigned long long)op << 32) | (0xffffffffu - (unsigned)(t * 3 + i));
                    atomicMax(&m.claim[t], key);
                    atomicMax(&m.claim[uu[i]], key);
                    bits |= 1u << (3 * j + i);
                }
            }
        }
        int slot = block_reserve(&m.counters[C_CAND], __popc(bits));
#pragma unroll
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 3; i++)
                if ((bits >> (3 * j + i)) & 1) m.cand[slot++] = tj[j] * 3 + i;
    }
    stat_add_full(m.stats, C_TESTS_INCIRCLE, tests);
}

__global__ __launch_bounds__(256) void k_flip_test(MeshDev m, const int *list, int cnt_slot, int prev_op, int op,
                                                   int cnt_out_slot)
{
    int n = m.counters[cnt_slot];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&m.counters[C_FLIP_LISTED], n);
        // the commit of this pass counts its output list and its flips from zero
        m.counters[cnt_out_slot] = 0;
        m.counters[C_FLIPS] = 0;
    }
    flip_test_body(m, list, n, prev_op, op);
}

// Block-uniform loop (block_reserve inside): up to four candidates per thread and step, and one
// reservation of output-list slots per block and step (a reservation per wave put thousands of
// atomics on the one counter).
GA_INLINE void flip_commit_body(const MeshDev &m, int n, int op, int *list_out, int cnt_out_slot)
{
    int flips = 0;
    const int bd = (int)blockDim.x;
    const int per = n > 4 * (int)(gridDim.x * bd) ? 4 : 1;
    for (int b0 = xcd_block() * bd * per; b0 < n; b0 += gridDim.x * bd * per) {
        int tq[4], uq[4];
        unsigned outs = 0;   // bit 2j: slot t of the j-th candidate joins the next list, 2j + 1: slot u
#pragma unroll
        for (int j = 0; j < 4; j++) {
            tq[j] = 0;
            uq[j] = 0;
            if (j >= per) continue;
            int q = b0 + j * bd + (int)threadIdx.x;
            if (q >= n) continue;
            int c = m.cand[q];
            int t = c / 3, i = c % 3;
            unsigned long long key = ((unsigned long long)op << 32) | (0xffffffffu - (unsigned)c);
            bool win = m.claim[t] == key;
            int4 tt = m.tv[t], nt = m.tn[t];
            int e = ic(nt, i);
            int u = enc_t(e), jj = enc_j(e);
            win = win && m.claim[u] == key;
            // Next pass's list: both triangles of a flip, and the first triangle of a losing
            // candidate (its edge may stay non-Delaunay when the candidate that beat it lost too);
            // lmark keeps every slot at most once per list, so no candidate key repeats.
            if (atomicExch(&m.lmark[t], op) != op) outs |= 1u << (2 * j);
            if (win && atomicExch(&m.lmark[u], op) != op) outs |= 2u << (2 * j);
            tq[j] = t;
            uq[j] = u;
            if (!win) continue;
            int4 ut = m.tv[u], un = m.tn[u];
            int cc = ic(tt, i), x = ic(tt, p1(i)), y = ic(tt, m1(i));
            int d = ic(ut, jj);
            int nt1 = ic(nt, p1(i)), nt2 = ic(nt, m1(i));   // (y,c) , (c,x)
            int nu1 = ic(un, p1(jj)), nu2 = ic(un, m1(jj));   // (x,d) , (d,y)
            m.tv[t] = make_int4(cc, x, d, 0);
            m.tv[u] = make_int4(d, y, cc, 0);
            m.tn[t] = make_int4(mark_old(nu1), enc(u, 1), mark_old(nt2), 0);
            m.tn[u] = make_int4(mark_old(nt1), enc(t, 1), mark_old(nu2), 0);
            int4 rt = make_int4(0, 0, 0, 0), ru = make_int4(0, 0, 0, 0);
            sc(rt, i, -1); sc(rt, p1(i), enc(u, 0)); sc(rt, m1(i), enc(t, 2));
            sc(ru, jj, -1); sc(ru, p1(jj), enc(t, 0)); sc(ru, m1(jj), enc(u, 2));
            m.remap[t] = rt; m.remap[u] = ru;
            m.stamp[t] = op; m.stamp[u] = op;
            m.vt[x] = t; m.vt[cc] = t; m.vt[y] = u; m.vt[d] = u;
            flips++;
        }
        int slot = block_reserve(&m.counters[cnt_out_slot], __popc(outs));
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if ((outs >> (2 * j)) & 1) list_out[slot++] = tq[j];
            if ((outs >> (2 * j + 1)) & 1) list_out[slot++] = uq[j];
        }
    }
    // C_FLIPS only tells the host whether the pass flipped anything: a plain store
    int wf = flips;
    for (int d = 32; d > 0; d >>= 1) wf += __shfl_xor(wf, d, 64);
    if (__lane_id() == 0 && wf) m.counters[C_FLIPS] = 1;
    stat_add_full(m.stats, C_FLIPS_TOTAL, flips);
}

__global__ void k_flip_commit(MeshDev m, int op, int *list_out, int cnt_out_slot)
{
    flip_commit_body(m, m.counters[C_CAND], op, list_out, cnt_out_slot);
}

// Short lists: the remaining passes in one workgroup, phases separated by block barriers instead
// of launches, until a pass flips nothing (counters[C_TAIL_DONE] = 1) or the list outgrows
// TAIL_LIST or max_passes run out (the host goes on with grid passes).  Pass k uses op op0 + k;
// counters[C_TAIL_PASSES] = passes run (odd: the lists have swapped).
#define TAIL_THREADS 1024
__global__ __launch_bounds__(TAIL_THREADS) void k_flip_tail(MeshDev m, int *lA, int sA, int *lB, int sB,
                                                            int prev_op, int op0, int max_passes)
{
    int *lin = lA, *lout = lB, sin = sA, sout = sB;
    int prev = prev_op, pass = 0, done = 0;
    for (; pass < max_passes; pass++) {
        const int n = m.counters[sin];
        if (n > TAIL_LIST) break;
        const int op = op0 + pass;
        __syncthreads();   // every thread has read n before the counters are reset
        if (threadIdx.x == 0) {
            atomicAdd(&m.counters[C_FLIP_LISTED], n);
            atomicAdd(&m.counters[C_TAIL_LISTED], n);
            m.counters[sout] = 0;
            m.counters[C_CAND] = 0;
            m.counters[C_FLIPS] = 0;
        }
        __syncthreads();
        flip_test_body(m, lin, n, prev, op);
        __syncthreads();
        flip_commit_body(m, m.counters[C_CAND], op, lout, sout);
        __syncthreads();
        relink_body(m, lout, m.counters[sout], op);
        __syncthreads();
        const bool any = m.counters[C_FLIPS] != 0;
        prev = op;
        int *tl = lin; lin = lout; lout = tl;
        int ts = sin; sin = sout; sout = ts;
        if (!any) {
            pass++;
            done = 1;
            break;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        m.counters[C_CAND] = 0;
        m.counters[C_TAIL_PASSES] = pass;
        m.counters[C_TAIL_DONE] = done;
    }
}

void launch_flip_tail(MeshDev &m, int *lA, int sA, int *lB, int sB, int prev_op, int op0, int max_passes,
                      hipStream_t st)
{
    GA_LAUNCH(k_flip_tail, dim3(1), dim3(TAIL_THREADS), 0, st, m, lA, sA, lB, sB, prev_op, op0, max_passes);
}

// Three launches per pass and no counter resets of their own: the test clears the output list's
// count and the flip flag, and the relink clears the candidate count for the next pass (it must
// be zero before the first pass).
void launch_flip_pass(MeshDev &m, const int *list_in, int cnt_in_slot, int *list_out, int cnt_out_slot,
                      int prev_op, int op, int grid, hipStream_t st)
{
    GA_LAUNCH(k_flip_test, dim3(grid), dim3(256), 0, st, m, list_in, cnt_in_slot, prev_op, op, cnt_out_slot);
    GA_LAUNCH(k_flip_commit, dim3(grid), dim3(256), 0, st, m, op, list_out, cnt_out_slot);
    GA_LAUNCH(k_relink, dim3(grid), dim3(256), 0, st, m, (const int *)list_out, cnt_out_slot, op, C_CAND);
}

// ---- frame removal ---------------------------------------------------------------------------
// fan of a hull vertex f: first triangle whose cw side (edge (f, l0)) is a hull edge
GA_INLINE int fan_first(const MeshDev &m, int f, int *slot)
{
    int t = m.vt[f], i = slot_of(m.tv[t], f), guard = 0;
    while (true) {
        int e = ic(m.tn[t], m1(i));
        if (e < 0) break;
        t = enc_t(e);
        i = slot_of(m.tv[t], f);
        if (++guard > 1000000) { atomicOr(m.status, GA_ERR_WALK); break; }
    }
    *slot = i;
    return t;
}

// Every remaining frame vertex marks its closed neighbourhood with its priority and records its
// first fan triangle and fan length in ffan[] for k_frame_remove of the same round.
__global__ void k_frame_prio(MeshDev m, unsigned long long *mx, int round)
{
    int nf = m.V - m.n;
    GRID_STRIDE(k, nf) {
        int f = m.n + k;
        if (m.vstate[f] != VS_IN_MESH) continue;
        unsigned long long p = prio64(f, round);
        atomicMax(&mx[f], p);
        int i, t = fan_first(m, f, &i);
        m.ffan[2 * k] = enc(t, i);
        atomicMax(&mx[ic(m.tv[t], p1(i))], p);
        int kf = 0;
        for (int guard = 0; guard < 1000000; guard++) {
            kf++;
            atomicMax(&mx[ic(m.tv[t], m1(i))], p);
            int e = ic(m.tn[t], p1(i));
            if (e < 0) break;
            t = enc_t(e);
            i = slot_of(m.tv[t], f);
        }
        m.ffan[2 * k + 1] = kf;
    }
}

// Read-only: a frame vertex wins when every vertex of its closed neighbourhood holds its
// priority; ffan[2k + 1] keeps the fan length of a winner and -1 otherwise.
__global__ void k_frame_decide(MeshDev m, const unsigned long long *mx, int round)
{
    int nf = m.V - m.n;
    GRID_STRIDE(k, nf) {
        int f = m.n + k;
        if (m.vstate[f] != VS_IN_MESH) continue;
        unsigned long long p = prio64(f, round);
        int kf = m.ffan[2 * k + 1];
        bool win = mx[f] == p;
        int t = enc_t(m.ffan[2 * k]), i = enc_j(m.ffan[2 * k]);
        if (win) win = mx[ic(m.tv[t], p1(i))] == p;
        for (int q = 0; q < kf && win; q++) {
            int4 tt = m.tv[t];
            if (mx[ic(tt, m1(i))] != p) win = false;
            else if (q + 1 < kf) {
                int e = ic(m.tn[t], p1(i));
                if (e < 0) { win = false; atomicOr(m.status, GA_ERR_WALK); break; }
                t = enc_t(e);
                i = slot_of(m.tv[t], f);
            }
        }
        if (!win) m.ffan[2 * k + 1] = -1;
    }
}

__global__ void k_frame_remove(MeshDev m, const unsigned long long *mx, int round, int op, int *list_out,
                               int cnt_out_slot)
{
    int nf = m.V - m.n;
    int seen = 0, chainsum = 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < ((nf + 63) & ~63); k += gridDim.x * blockDim.x) {
        // (full-wave loop so the work counters can be summed per wave at the end)
        if (k >= nf) continue;
        int f = m.n + k;
        if (m.vstate[f] != VS_IN_MESH) continue;
        seen++;
        // winner decided by k_frame_decide this round; its fan is untouched by other winners
        // (distance-2 independence), so the chain is read without races
        int kf = m.ffan[2 * k + 1];
        if (kf < 0) continue;
        int t0 = enc_t(m.ffan[2 * k]), i0 = enc_j(m.ffan[2 * k]);
        // chain l_0..l_kf, outer links o_0..o_{kf-1}, fan slots
        int need = 4 * (kf + 1);
        int base = atomicAdd(&m.counters[C_SCR], need);
        if ((size_t)(base + need) > m.scratch_cap) { atomicOr(m.status, GA_ERR_SCRATCH); continue; }
        int *chain = m.scratch + base, *olink = chain + (kf + 1), *slots = olink + (kf + 1), *st = slots + (kf + 1);
        int t = t0, i = i0;
        chain[0] = ic(m.tv[t0], p1(i0));
        for (int q = 0; q < kf; q++) {
            int4 tt = m.tv[t], nn = m.tn[t];
            chain[q + 1] = ic(tt, m1(i));
            olink[q] = ic(nn, i);
            slots[q] = t;
            int e = ic(nn, p1(i));
            if (q + 1 < kf) {
                if (e < 0) { atomicOr(m.status, GA_ERR_WALK); break; }
                t = enc_t(e);
                i = slot_of(m.tv[t], f);
            }
        }
        // chain vertices keep an incident triangle: the outer one
        for (int q = 0; q <= kf; q++) {
            int o = q < kf ? olink[q] : -1;
            if (o < 0 && q > 0) o = olink[q - 1];
            if (o >= 0) m.vt[chain[q]] = enc_t(o);
        }
        // ccw-stack fill: stack st[], slink = olink reused as stack-edge links
        int *slink = olink;   // slink[s] = outer side of stack edge (st[s], st[s+1]); olink[q] read before overwrite
        int top = 0, used = 0;
        st[0] = chain[0];
        for (int q = 1; q <= kf; q++) {
            int c = chain[q];
            int cur = olink[q - 1];
            while (top >= 1 && orient_dev(m, m.pos[st[top - 1]], m.pos[st[top]], m.pos[c]) > 0) {
                int A = st[top - 1], B = st[top];
                int T = slots[used++];
                m.tv[T] = make_int4(A, B, c, 0);
                m.tn[T] = make_int4(-1, -1, -1, 0);
                bond(m, T, 2, slink[top - 1]);   // edge (A, B)
                bond(m, T, 0, cur);              // edge (B, c)
                m.vt[A] = T; m.vt[B] = T; m.vt[c] = T;
                m.stamp[T] = op;
                if (atomicExch(&m.lmark[T], m.list_op) != m.list_op)
                    list_out[atomicAdd(&m.counters[cnt_out_slot], 1)] = T;
                cur = enc(T, 1);                 // edge (c, A) seen from T
                top--;
            }
            slink[top] = cur;
            st[++top] = c;
        }
        // remaining stack edges become hull edges
        for (int s = 0; s < top; s++) {
            int X = slink[s];
            if (X >= 0) set_link(m, enc_t(X), enc_j(X), -1);
        }
        for (int q = used; q < kf; q++) m.tv[slots[q]] = make_int4(GA_DEAD, GA_DEAD, GA_DEAD, 0);
        m.vstate[f] = VS_REMOVED;
        m.vt[f] = -1;
        stat_inc(m.stats, C_FRAME_DONE, true);
        chainsum += kf;
    }
    stat_add_full(m.stats, C_FRAME_SEEN, seen);
    stat_add_full(m.stats, C_FRAME_CHAIN, chainsum);
}

void launch_frame_round(MeshDev &m, unsigned int *mxraw, int round, int *list_out, int cnt_out_slot, int op,
                        hipStream_t st)
{
    unsigned long long *mx = (unsigned long long *)mxraw;
    GA_LAUNCH(k_frame_prio, dim3(FIXED_GRID), dim3(256), 0, st, m, mx, round);
    GA_LAUNCH(k_frame_decide, dim3(FIXED_GRID), dim3(256), 0, st, m, (const unsigned long long *)mx, round);
    GA_LAUNCH(k_frame_remove, dim3(FIXED_GRID), dim3(256), 0, st, m, (const unsigned long long *)mx, round, op,
                       list_out, cnt_out_slot);
}

// ---- hull convexification (one workgroup per set) --------------------------------------------
// Hull cycle of set `blockIdx.x` in the scratch arena at 4 * hull_off[set]: verts[], owner[] (enc
// of the triangle side of hull edge (verts[k] -> verts[k+1])), nxt[], prv[]; hull_cnt[set] is the
// number of hull edges, so the cycle fits exactly.
__global__ __launch_bounds__(256) void k_convexify(MeshDev m, const int *hnext, const int *hedge,
                                                   const int *hull_cnt, const int *hull_off,
                                                   const int *hull_start, int *list_out, int cnt_out_slot,
                                                   int op)
{
    __shared__ int s_flag, s_start, s_h, s_cur, s_block;
    int set = blockIdx.x;
    int hcap = hull_cnt[set];
    int *verts = m.scratch + 4 * (size_t)hull_off[set], *owner = verts + hcap, *nxt = owner + hcap,
        *prv = nxt + hcap;
    if (threadIdx.x == 0) {
        int start = hull_start[set];
        if (start == 0x7fffffff) start = -1;
        int h = 0, v = start;
        if (start >= 0) {
            do {
                if (h >= hcap) { atomicOr(m.status, GA_ERR_SCRATCH); h = -1; break; }
                verts[h] = v;
                owner[h] = hedge[v];
                h++;
                v = hnext[v];
            } while (v != start && v >= 0);
        }
        if (h > 0) for (int k = 0; k < h; k++) { nxt[k] = (k + 1) % h; prv[k] = (k + h - 1) % h; }
        s_h = h;
        s_start = 0;
    }
    __syncthreads();
    int h = s_h;
    if (h < 3) return;
    for (int pass = 0; pass < h + 2; pass++) {
        if (threadIdx.x == 0) s_flag = 0;
        __syncthreads();
        int cur = s_start;
        // one sweep around the cycle
        for (int it = 0; it < 4 * h + 4; it++) {
            __syncthreads();
            if (threadIdx.x == 0) {
                s_cur = cur;
                s_block = 0;
            }
            __syncthreads();
            int kc = s_cur, kp = prv[kc], kn = nxt[kc];
            double2 Pp = m.pos[verts[kp]], C = m.pos[verts[kc]], N = m.pos[verts[kn]];
            bool reflex = orient_dev(m, Pp, C, N) < 0;
            if (reflex) {
                // blocked if any other live hull vertex lies in the closed notch (P, N, C) minus P, N, C
                for (int q = threadIdx.x; q < h; q += blockDim.x) {
                    if (q == kc || q == kp || q == kn || nxt[q] < 0) continue;
                    double2 Q = m.pos[verts[q]];
                    if (orient_dev(m, Pp, N, Q) >= 0 && orient_dev(m, N, C, Q) >= 0 && orient_dev(m, C, Pp, Q) >= 0)
                        s_block = 1;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                if (reflex && !s_block) {
                    int T = atomicAdd(m.ntri, 1);
                    if (T >= m.cap) { atomicOr(m.status, GA_ERR_CAPACITY); }
                    else {
                        int vp = verts[kp], vc = verts[kc], vn = verts[kn];
                        m.tv[T] = make_int4(vp, vn, vc, 0);
                        m.tn[T] = make_int4(-1, -1, -1, 0);
                        bond(m, T, 0, owner[kc]);   // edge (vn, vc) = reverse of hull edge (vc -> vn)
                        bond(m, T, 1, owner[kp]);   // edge (vc, vp) = reverse of hull edge (vp -> vc)
                        m.stamp[T] = op;
                        if (atomicExch(&m.lmark[T], m.list_op) != m.list_op)
                            list_out[atomicAdd(&m.counters[cnt_out_slot], 1)] = T;
                        owner[kp] = enc(T, 2);       // new hull edge (vp -> vn)
                        nxt[kp] = kn; prv[kn] = kp;
                        nxt[kc] = -1;
                        if (s_start == kc) s_start = kp;
                        s_flag = 1;
                        cur = kp;
                    }
                } else {
                    cur = kn;
                }
                s_cur = cur;
            }
            __syncthreads();
            cur = s_cur;
            if (cur == s_start && it > 0) break;
        }
        __syncthreads();
        if (!s_flag) break;
    }
}

// zero up to four counter slots in one launch (negative slot: unused)
__global__ void k_reset_counters(int *c, int a, int b, int d, int e)
{
    if (threadIdx.x == 0) {
        if (a >= 0) c[a] = 0;
        if (b >= 0) c[b] = 0;
        if (d >= 0) c[d] = 0;
        if (e >= 0) c[e] = 0;
    }
}

void launch_reset_counters(int *counters, int a, int b, int d, int e, hipStream_t st)
{
    GA_LAUNCH(k_reset_counters, dim3(1), dim3(64), 0, st, counters, a, b, d, e);
}

__global__ void k_fold_stats(int *c)
{
    int k = threadIdx.x, sum = 0;
    for (int s = 0; s < GA_STRIPES; s++) sum += c[GA_STRIPE_INTS * (1 + s) + k];
    c[C_FOLD + k] = sum;
}

void launch_fold_stats(int *counters, hipStream_t st)
{
    GA_LAUNCH(k_fold_stats, dim3(1), dim3(64), 0, st, counters);
}

// counters[slot, slot + n) straight into pinned host memory (n <= 256); with fold, the slots
// C_FOLD + k hold the stripe sum of statistic k (also stored back on the device, as k_fold_stats
// does).  The host waits on the stream before it reads.
__global__ void k_counters_to_host(int *c, int slot, int n, int fold, int *host)
{
    const int i = threadIdx.x;
    if (i < n) {
        const int q = slot + i;
        int v;
        if (fold && q >= C_FOLD && q < C_FOLD + 64) {
            v = 0;
            for (int s = 0; s < GA_STRIPES; s++) v += c[GA_STRIPE_INTS * (1 + s) + (q - C_FOLD)];
            c[q] = v;
        } else {
            v = c[q];
        }
        host[i] = v;
    }
    __threadfence_system();
}

void launch_counters_to_host(int *counters, int slot, int n, int fold, int *host, hipStream_t st)
{
    GA_LAUNCH(k_counters_to_host, dim3(1), dim3(256), 0, st, counters, slot, n, fold, host);
}

__global__ void k_hull_edges(MeshDev m, int *hnext, int *hedge, int *hull_cnt, int *hull_start)
{
    int n = ntri_used(m);
    GRID_STRIDE(t, n) {
        int4 v = m.tv[t];
        if (v.x == GA_DEAD) continue;
        int4 nb = m.tn[t];
        for (int i = 0; i < 3; i++)
            if (ic(nb, i) < 0) {
                int a = ic(v, p1(i)), b = ic(v, m1(i));
                hnext[a] = b;
                hedge[a] = enc(t, i);
                int s = m.vset[a];
                atomicMin(&hull_start[s], a);
                atomicAdd(&hull_cnt[s], 1);
            }
    }
}

void launch_convexify(MeshDev &m, int *list_out, int cnt_out_slot, int op, int *hull_cnt, int *hull_off,
                      int *hull_start, void *scan_tmp, hipStream_t st)
{
    IntFills f;
    f.add(m.hnext, (size_t)m.V, -1);
    f.add(hull_cnt, (size_t)m.nsets, 0);
    f.add(hull_start, (size_t)m.nsets, 0x7fffffff);
    launch_fill_multi(f, st);
    GA_LAUNCH(k_hull_edges, dim3(FIXED_GRID), dim3(256), 0, st, m, m.hnext, m.hedge, hull_cnt, hull_start);
    launch_exclusive_scan(hull_cnt, hull_off, m.nsets, scan_tmp, st);
    GA_LAUNCH(k_convexify, dim3(m.nsets), dim3(256), 0, st, m, (const int *)m.hnext, (const int *)m.hedge,
              (const int *)hull_cnt, (const int *)hull_off, (const int *)hull_start, list_out, cnt_out_slot, op);
}

// ---- bulk frame removal ------------------------------------------------------------------------
// Every triangle of the current mesh whose vertices are all real is a triangle of the final
// triangulation (its circumcircle holds no point at all).  When the region R of those triangles
// has, in every set, a boundary that is one simple cycle through every meshed real point, all
// frame triangles are dropped at once and R is made convex by parallel ear rounds on that cycle.
// The checks below run before anything is modified; if one fails, the caller removes the frame
// vertices round by round instead.
GA_INLINE bool is_frame_tri(const MeshDev &m, int4 v) { return v.x >= m.n || v.y >= m.n || v.z >= m.n; }

// boundary edges of R (a -> b with R on the left), their count and smallest tail per set; a vertex
// with two outgoing boundary edges (a pinch) or a meshed real point in no R triangle marks its set
// Each set's boundary-edge count and smallest boundary vertex are accumulated per thread (while
// the set stays the same) and added once per wave when the wave's edges share one set: atomics
// of every boundary edge on the set's two words would serialise.
__global__ void k_bnd_edges(MeshDev m, int *outdeg, int *cover, int *hull_cnt, int *hull_start, int *badset)
{
    int n = ntri_used(m);
    int cs = -1, cc = 0, cmin = 0x7fffffff;   // pending count and minimum of set cs
    GRID_STRIDE(t, n) {
        int4 v = m.tv[t];
        if (v.x == GA_DEAD || is_frame_tri(m, v)) continue;
        cover[v.x] = 1; cover[v.y] = 1; cover[v.z] = 1;
        int4 nb = m.tn[t];
        for (int i = 0; i < 3; i++) {
            int e = ic(nb, i);
            bool bnd = e < 0 || is_frame_tri(m, m.tv[enc_t(e)]);
            if (!bnd) continue;
            int a = ic(v, p1(i)), b = ic(v, m1(i));
            m.hnext[a] = b;
            m.hedge[a] = enc(t, i);
            int s = m.vset[a];
            if (atomicAdd(&outdeg[a], 1) != 0) badset[s] = 1;
            if (s != cs) {
                if (cc) {
                    atomicMin(&hull_start[cs], cmin);
                    atomicAdd(&hull_cnt[cs], cc);
                }
                cs = s;
                cc = 0;
                cmin = 0x7fffffff;
            }
            cc++;
            cmin = min(cmin, a);
        }
    }
    // every lane of the wave is here
    const bool has = cc > 0;
    const unsigned long long bal = __ballot(has);
    if (!bal) return;
    const int lane = threadIdx.x & 63, leader = __ffsll((long long)bal) - 1;
    const int s0 = __shfl(cs, leader, 64);
    if (__all(!has || cs == s0)) {
        int sum = has ? cc : 0, mn = has ? cmin : 0x7fffffff;
        for (int d = 32; d > 0; d >>= 1) {
            sum += __shfl_xor(sum, d, 64);
            mn = min(mn, __shfl_xor(mn, d, 64));
        }
        if (lane == leader) {
            atomicMin(&hull_start[s0], mn);
            atomicAdd(&hull_cnt[s0], sum);
        }
    } else if (has) {
        atomicMin(&hull_start[cs], cmin);
        atomicAdd(&hull_cnt[cs], cc);
    }
}

__global__ void k_bnd_cover(MeshDev m, const int *cover, int *badset, int *blist, int *bidx, int *bcount)
{
    GRID_STRIDE(v, m.n) {
        if (m.vstate[v] == VS_IN_MESH && !cover[v]) badset[m.vset[v]] = 1;
        bool b = m.hnext[v] >= 0;
        int k = wave_reserve(bcount, b);
        if (b) { blist[k] = v; bidx[v] = k; }
    }
}

// list ranking along each cycle, cut before its start vertex: d = steps to the cut
__global__ void k_bnd_rank_init(MeshDev m, int B, const int *blist, const int *bidx, const int *hull_start,
                                int *succ, int *d, int *badset)
{
    GRID_STRIDE(k, B) {
        int a = blist[k];
        int nx = m.hnext[a];
        bool last = nx == hull_start[m.vset[a]];
        int si = last ? -1 : bidx[nx];
        if (!last && (si < 0 || si >= B)) {   // successor without an outgoing boundary edge
            badset[m.vset[a]] = 1;
            si = -1;
            last = true;
        }
        succ[k] = si;
        d[k] = last ? 0 : 1;
    }
}

__global__ void k_bnd_rank_step(int B, const int *succ, const int *d, int *succ2, int *d2)
{
    GRID_STRIDE(k, B) {
        int s = succ[k];
        if (s < 0) { succ2[k] = -1; d2[k] = d[k]; }
        else { succ2[k] = succ[s]; d2[k] = d[k] + d[s]; }
    }
}

// the same rounds in one workgroup, in place (each round reads every entry's successor into
// registers, then writes), for cycles of up to RANK_BLOCK * RANK_PER positions: one launch instead
// of one per round
#define RANK_BLOCK 1024
#define RANK_PER 16
__global__ __launch_bounds__(RANK_BLOCK) void k_bnd_rank_block(int B, int rounds, int *succ, int *d)
{
    int sn[RANK_PER], dn[RANK_PER];
    for (int r = 0; r < rounds; r++) {
#pragma unroll
        for (int j = 0; j < RANK_PER; j++) {
            int k = threadIdx.x + j * RANK_BLOCK;
            if (k >= B) continue;
            int sk = succ[k], dk = d[k];
            sn[j] = sk < 0 ? -1 : succ[sk];
            dn[j] = sk < 0 ? dk : dk + d[sk];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RANK_PER; j++) {
            int k = threadIdx.x + j * RANK_BLOCK;
            if (k >= B) continue;
            succ[k] = sn[j];
            d[k] = dn[j];
        }
        __syncthreads();
    }
}

// a cycle that does not return to its start within the rounds (a second cycle in the set), or two
// vertices ranked to one position, marks the set; verts[] starts at -1, so a position of a marked
// set may stay unwritten (verts -1, nxt/prv unset) and later kernels skip it
__global__ void k_bnd_place(MeshDev m, int B, const int *blist, const int *succ, const int *d, const int *hull_cnt,
                            const int *hull_off, int *badset, int *verts, int *owner, int *nxt, int *prv,
                            int *alive, int *pset, double2 *bpos)
{
    GRID_STRIDE(k, B) {
        int a = blist[k];
        int s = m.vset[a];
        int h = hull_cnt[s];
        if (succ[k] >= 0 || d[k] >= h) { badset[s] = 1; continue; }
        int pos = h - 1 - d[k];
        int p = hull_off[s] + pos;
        if (atomicExch(&verts[p], a) != -1) badset[s] = 1;
        bpos[p] = m.pos[a];
        owner[p] = m.hedge[a];
        nxt[p] = hull_off[s] + (pos + 1 == h ? 0 : pos + 1);
        prv[p] = hull_off[s] + (pos == 0 ? h - 1 : pos - 1);
        alive[p] = 1;
        pset[p] = s;
    }
}

// drop every frame triangle; R's triangles across a dropped edge get a hull link; frame vertices
// leave the mesh; boundary vertices take their outgoing boundary triangle as anchor
__global__ void k_kill_frame(MeshDev m)
{
    int n = ntri_used(m);
    GRID_STRIDE(t, n) {
        int4 v = m.tv[t];
        if (v.x == GA_DEAD || !is_frame_tri(m, v)) continue;
        int4 nb = m.tn[t];
        for (int i = 0; i < 3; i++) {
            int e = ic(nb, i);
            if (e < 0) continue;
            int u = enc_t(e);
            if (!is_frame_tri(m, m.tv[u])) set_link(m, u, enc_j(e), -1);
        }
        m.tv[t] = make_int4(GA_DEAD, GA_DEAD, GA_DEAD, 0);
    }
    GRID_STRIDE(k, m.V - m.n) {
        int f = m.n + k;
        if (m.vstate[f] == VS_IN_MESH) m.vstate[f] = VS_REMOVED;
        m.vt[f] = -1;
    }
    GRID_STRIDE(v, m.n) {
        if (m.hnext[v] >= 0) m.vt[v] = enc_t(m.hedge[v]);
    }
}

// bnd layout: succ[2][B], d[2][B], verts, owner, nxt, prv, alive, pset, cand (win reuses d[0]), then
// the positions of the cycle (double2[B]) from int offset bnd_pos_off(B), 16-byte aligned
GA_INLINE_HOST_DEVICE int bnd_pos_off(int B) { return (11 * B + 3) & ~3; }

// Ear rounds.  A cycle position p (vertex C between P = prev and N = next) is a candidate when C
// is reflex (P, C, N clockwise) and no other live vertex of its cycle lies in the closed triangle
// (P, N, C); winners are candidates above every candidate within cycle distance 2 (so two winners
// share no vertex), and each winner cuts its ear.  Rounds repeat until one has no candidate: the
// cycle is then convex.  The emptiness test looks only at the vertices in the grid cells the ear's
// bounding box covers: a uniform grid over each set's cycle, about two vertices per cell, built by
// a counting sort before the first round.  Few large sets run their rounds as grid-wide kernels
// (two per round, with fan ears over straight runs of the cycle, below); many small sets run all
// their rounds in one launch, one workgroup per set.
GA_INLINE uint32_t ear_prio(int p, int round) { return hash32((uint32_t)p * 0x9E3779B1u ^ (uint32_t)round); }
GA_INLINE bool ear_beats(int p, int q, int round)
{
    uint32_t a = ear_prio(p, round), b = ear_prio(q, round);
    return a > b || (a == b && p > q);
}

GA_INLINE int ear_cell(double v, double lo, double sc, int G)
{
    int c = (int)((v - lo) * sc);
    return c < 0 ? 0 : (c >= G ? G - 1 : c);
}

// The cycle arrays of bulk frame removal (bnd layout: succ[2][B], d[2][B], verts, owner, nxt, prv,
// alive, pset, cand, positions) and the grid of set s: cell ends in succ (2 h ints from 2 off; the
// ranking is done), positions by cell in d[1], and (lx, ly, sx, sy, G) after the positions.
struct EarCycle {
    const int *verts;
    int *owner, *nxt, *prv, *alive, *pset, *cand, *win;
    const double2 *bpos;
    double *param;
    int *cells;
    int B;
};
GA_INLINE EarCycle ear_cycle(int *bnd, int B)
{
    EarCycle e;
    e.verts = bnd + 4 * B; e.owner = bnd + 5 * B; e.nxt = bnd + 6 * B; e.prv = bnd + 7 * B;
    e.alive = bnd + 8 * B; e.pset = bnd + 9 * B; e.cand = bnd + 10 * B; e.win = bnd + 2 * B;
    e.bpos = (const double2 *)(bnd + bnd_pos_off(B));
    e.param = (double *)(bnd + bnd_pos_off(B) + 4 * B);
    e.cells = bnd;
    e.B = B;
    return e;
}
GA_INLINE int *ear_cend(const EarCycle &e, int off) { return e.cells + 2 * (size_t)off; }
GA_INLINE int *ear_citem(const EarCycle &e, int off) { return e.cells + 3 * (size_t)e.B + off; }

#define EAR_THREADS 1024
#define EAR_LDS_CELLS 16384   // grid cells k_ear_grid counts in LDS (64 KB); larger grids count in memory
// The grid of the set of this workgroup (h positions from off): bounding box, counting sort of the
// positions by cell (counts, exclusive starts, then a fill that counts each start up to its cell's
// end: cell c holds [c ? cend[c - 1] : 0, cend[c])), parameters stored for other kernels.
// lcells: LDS cells of the calling kernel (NULL: none); the counting sort runs there when the
// grid fits in lcap cells, and its result is copied out
GA_INLINE void ear_build_grid(const EarCycle &e, int s, int h, int off, int *lcells = nullptr, int lcap = 0)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    __shared__ double red[4][EAR_THREADS / 64];
    __shared__ int s_total[EAR_THREADS / 64];
    int *cend = ear_cend(e, off), *citem = ear_citem(e, off);
    double lx = INFINITY, ly = INFINITY, hx = -INFINITY, hy = -INFINITY;
    for (int k = tid; k < h; k += EAR_THREADS) {
        double2 q = e.bpos[off + k];
        lx = fmin(lx, q.x); ly = fmin(ly, q.y); hx = fmax(hx, q.x); hy = fmax(hy, q.y);
    }
    for (int d = 32; d > 0; d >>= 1) {
        lx = fmin(lx, __shfl_xor(lx, d, 64)); ly = fmin(ly, __shfl_xor(ly, d, 64));
        hx = fmax(hx, __shfl_xor(hx, d, 64)); hy = fmax(hy, __shfl_xor(hy, d, 64));
    }
    if (lane == 0) { red[0][wv] = lx; red[1][wv] = ly; red[2][wv] = hx; red[3][wv] = hy; }
    __syncthreads();
    lx = red[0][0]; ly = red[1][0]; hx = red[2][0]; hy = red[3][0];
    for (int w = 1; w < EAR_THREADS / 64; w++) {
        lx = fmin(lx, red[0][w]); ly = fmin(ly, red[1][w]); hx = fmax(hx, red[2][w]); hy = fmax(hy, red[3][w]);
    }
    int G = 1;
    while (2 * G * G < h) G++;   // about two vertices per cell; G * G <= 2 h for h >= 3
    const int ncell = G * G;
    const double sx = hx > lx ? G / (hx - lx) : 0.0, sy = hy > ly ? G / (hy - ly) : 0.0;
    if (tid == 0) {
        double *pr = e.param + 5 * (size_t)s;
        pr[0] = lx; pr[1] = ly; pr[2] = sx; pr[3] = sy; pr[4] = (double)G;
    }
    int *cnt = (lcells && ncell <= lcap) ? lcells : cend;
    for (int c = tid; c < ncell; c += EAR_THREADS) cnt[c] = 0;
    __syncthreads();
    for (int k = tid; k < h; k += EAR_THREADS) {
        double2 q = e.bpos[off + k];
        atomicAdd(&cnt[ear_cell(q.y, ly, sy, G) * G + ear_cell(q.x, lx, sx, G)], 1);
    }
    __syncthreads();
    {
        int per = (ncell + EAR_THREADS - 1) / EAR_THREADS, c0 = tid * per, c1 = min(ncell, c0 + per);
        int sum = 0;
        for (int c = c0; c < c1; c++) sum += cnt[c];
        int inc = sum;
        for (int d = 1; d < 64; d <<= 1) {
            int t = __shfl_up(inc, d, 64);
            if (lane >= d) inc += t;
        }
        if (lane == 63) s_total[wv] = inc;
        __syncthreads();
        int base = inc - sum;
        for (int w = 0; w < wv; w++) base += s_total[w];
        for (int c = c0; c < c1; c++) {
            int v = cnt[c];
            cnt[c] = base;
            base += v;
        }
    }
    __syncthreads();
    for (int k = tid; k < h; k += EAR_THREADS) {
        double2 q = e.bpos[off + k];
        int c = ear_cell(q.y, ly, sy, G) * G + ear_cell(q.x, lx, sx, G);
        citem[atomicAdd(&cnt[c], 1)] = off + k;
    }
    __syncthreads();
    if (cnt != cend)
        for (int c = tid; c < ncell; c += EAR_THREADS) cend[c] = cnt[c];
}

// 1 when live position p of set s is reflex and its ear is empty
GA_INLINE int ear_candidate(const MeshDev &m, const EarCycle &e, int p, int off)
{
    if (!e.alive[p]) return 0;
    int kp = e.prv[p], kn = e.nxt[p];
    if (kp == kn) return 0;
    double2 P = e.bpos[kp], C = e.bpos[p], N = e.bpos[kn];
    if (orient_dev(m, P, C, N) >= 0) return 0;
    const double *pr = e.param + 5 * (size_t)e.pset[p];
    const double lx = pr[0], ly = pr[1], sx = pr[2], sy = pr[3];
    const int G = (int)pr[4];
    const int *cend = ear_cend(e, off), *citem = ear_citem(e, off);
    double bx0 = fmin(P.x, fmin(C.x, N.x)), bx1 = fmax(P.x, fmax(C.x, N.x));
    double by0 = fmin(P.y, fmin(C.y, N.y)), by1 = fmax(P.y, fmax(C.y, N.y));
    int cx0 = ear_cell(bx0, lx, sx, G), cx1 = ear_cell(bx1, lx, sx, G);
    int cy0 = ear_cell(by0, ly, sy, G), cy1 = ear_cell(by1, ly, sy, G);
    for (int cy = cy0; cy <= cy1; cy++)
        for (int cx = cx0; cx <= cx1; cx++) {
            int cc = cy * G + cx;
            for (int j = cc ? cend[cc - 1] : 0, j1 = cend[cc]; j < j1; j++) {
                int q = citem[j];
                if (q == p || q == kp || q == kn || !e.alive[q]) continue;
                double2 Q = e.bpos[q];
                if (Q.x < bx0 || Q.x > bx1 || Q.y < by0 || Q.y > by1) continue;
                if (orient_dev(m, P, N, Q) >= 0 && orient_dev(m, N, C, Q) >= 0 && orient_dev(m, C, P, Q) >= 0)
                    return 0;
            }
        }
    return 1;
}

// winner: a candidate above every candidate within cycle distance 2
GA_INLINE int ear_winner(const EarCycle &e, int p, int round)
{
    if (!e.cand[p]) return 0;
    int a = e.prv[p], b = e.nxt[p];
    int q[4] = {a, b, e.prv[a], e.nxt[b]};
    for (int j = 0; j < 4; j++)
        if (q[j] != p && e.cand[q[j]] && !ear_beats(p, q[j], round)) return 0;
    return 1;
}

// cut the ear of winner p: triangle (prev, next, p), new boundary edge prev -> next; false when the
// triangle slots are exhausted (status set)
GA_INLINE bool ear_cut(const MeshDev &m, const EarCycle &e, int p, int *list_out, int cnt_out_slot, int op)
{
    int kp = e.prv[p], kn = e.nxt[p];
    int T = atomicAdd(m.ntri, 1);
    if (T >= m.cap) {
        atomicOr(m.status, GA_ERR_CAPACITY);
        return false;
    }
    int vp = e.verts[kp], vc = e.verts[p], vn = e.verts[kn];
    m.tv[T] = make_int4(vp, vn, vc, 0);
    m.tn[T] = make_int4(-1, -1, -1, 0);
    bond(m, T, 0, e.owner[p]);    // edge (vn, vc) = reverse of boundary edge (vc -> vn)
    bond(m, T, 1, e.owner[kp]);   // edge (vc, vp) = reverse of boundary edge (vp -> vc)
    m.stamp[T] = op;
    if (atomicExch(&m.lmark[T], m.list_op) != m.list_op)
        list_out[atomicAdd(&m.counters[cnt_out_slot], 1)] = T;
    e.owner[kp] = enc(T, 2);      // new boundary edge (vp -> vn)
    e.nxt[kp] = kn;
    e.prv[kn] = kp;
    e.alive[p] = 0;
    m.vt[vp] = T; m.vt[vc] = T; m.vt[vn] = T;
    return true;
}

// one workgroup per set: the grid
__global__ __launch_bounds__(EAR_THREADS) void k_ear_grid(MeshDev m, int B, int *bnd, const int *hull_cnt,
                                                          const int *hull_off)
{
    __shared__ int s_cells[EAR_LDS_CELLS];
    int s = blockIdx.x, h = hull_cnt[s];
    if (h < 3) return;
    ear_build_grid(ear_cycle(bnd, B), s, h, hull_off[s], s_cells, EAR_LDS_CELLS);
}

// Grid-wide rounds: round r counts its candidates in counters[C_EAR + ...] by parity; a round
// after one without candidates returns at once (so several run per host read), and the apply of
// round r clears the counter round r + 1 will use.
GA_INLINE int *ear_count(const MeshDev &m, int round) { return m.counters + ((round & 1) ? C_EAR_PREV : C_EAR); }

// Fan ears.  Where the cycle runs straight (C, N, N2, ... collinear) only one of the run's vertices
// is reflex at a time, and plain ears would take the run one vertex per round (a long thin pocket
// between a straight side of lattice points and a hull edge takes hundreds).  A candidate C whose
// neighbour N starts such a run instead cuts the fan (P, N, C), (P, N2, N), ... up to the farthest
// run vertex E for which the closed triangle (P, E, C) holds no live vertex of the cycle outside
// the positions P .. E (the run lies on its edge C E) and no fan triangle has P's other
// neighbour in its circumcircle; the triangles nest, so E is found by binary search.  Mirrored for a run through P, with apex N.  A candidate's span is the positions from its
// first to its last vertex; a winner holds the largest claim key (plain ear before fan, priority,
// position) at every position of its span, so the spans of two winners never meet.  Plain ears
// first: the apex of a pocket's fan is then its last reflex vertex left, not the first one met
// (a fan from a shallow apex is far from Delaunay, and the flips that repair it run one
// triangle per pass along the fan).
#define FAN_CAP 1024    // most run vertices one fan takes (a longer run takes more rounds)
#define EAR_WAVES 4     // waves per workgroup of the wave-per-position round kernels

GA_INLINE unsigned long long ear_key(int p, int round, bool fan)
{
    return ((unsigned long long)!fan << 63) | ((unsigned long long)(ear_prio(p, round) >> 1) << 32) |
           (unsigned)(p + 1);
}
// position k of the set's cyclic range [off, off + h), k within one turn of it
GA_INLINE int ear_wrap(int k, int off, int h) { return k >= off + h ? k - h : (k < off ? k + h : k); }
// positions from lo forward to hi in the cyclic range (both included)
GA_INLINE int ear_span(int lo, int hi, int h) { return (hi >= lo ? hi - lo : hi - lo + h) + 1; }

// wave-uniform: no live vertex of the set at a position outside [lo, hi] lies in the closed
// triangle (A, Bq, C), clockwise order A, C, Bq
GA_INLINE bool ear_empty_wave(const MeshDev &m, const EarCycle &e, int off, int h, const double *pr, double2 A,
                              double2 Bq, double2 C, int lo, int hi)
{
    const int lane = threadIdx.x & 63;
    const double lx = pr[0], ly = pr[1], sx = pr[2], sy = pr[3];
    const int G = (int)pr[4];
    const int *cend = ear_cend(e, off), *citem = ear_citem(e, off);
    const int last = ear_span(lo, hi, h) - 1;
    double bx0 = fmin(A.x, fmin(C.x, Bq.x)), bx1 = fmax(A.x, fmax(C.x, Bq.x));
    double by0 = fmin(A.y, fmin(C.y, Bq.y)), by1 = fmax(A.y, fmax(C.y, Bq.y));
    int cx0 = ear_cell(bx0, lx, sx, G), cx1 = ear_cell(bx1, lx, sx, G);
    int cy0 = ear_cell(by0, ly, sy, G), cy1 = ear_cell(by1, ly, sy, G);
    bool blocked = false;
    for (int cy = cy0; cy <= cy1 && !__any(blocked); cy++) {
        int c0 = cy * G + cx0, c1 = cy * G + cx1;
        int j0 = c0 ? cend[c0 - 1] : 0, j1 = cend[c1];
        for (int j = j0 + lane; j < j1; j += 64) {
            int q = citem[j];
            int dq = q - lo;
            if (dq < 0) dq += h;
            if (dq <= last || !e.alive[q]) continue;
            double2 Q = e.bpos[q];
            if (Q.x < bx0 || Q.x > bx1 || Q.y < by0 || Q.y > by1) continue;
            if (orient_dev(m, A, Bq, Q) >= 0 && orient_dev(m, Bq, C, Q) >= 0 && orient_dev(m, C, A, Q) >= 0)
                blocked = true;
        }
    }
    return !__any(blocked);
}

// The run beyond X on the line (C, X): the live positions after x in direction d (+1 forward)
// while they lie on the line beyond X; written to run[] in order (at most FAN_CAP).  Returns
// their count (wave-uniform).  The scan ends at the latest at the candidate's other neighbour,
// which is off the line.
GA_INLINE int ear_run_wave(const MeshDev &m, const EarCycle &e, int off, int h, int x, int d, double2 C, double2 X,
                           int *run)
{
    const int lane = threadIdx.x & 63;
    const double dx = X.x - C.x, dy = X.y - C.y;
    int R = 0;
    for (int base = 1; base < h && R < FAN_CAP; base += 64) {
        int k = base + lane, q = -1;
        bool live = false, brk = false;
        if (k < h) {
            q = ear_wrap(x + d * k, off, h);
            live = e.alive[q] != 0;
            if (live) {
                double2 Q = e.bpos[q];
                brk = orient_dev(m, C, X, Q) != 0 || (Q.x - X.x) * dx + (Q.y - X.y) * dy <= 0.0;
            }
        }
        unsigned long long bm = __ballot(brk), lm = __ballot(live);
        if (bm) lm &= (1ull << (__ffsll((long long)bm) - 1)) - 1;   // live lanes before the first break
        int idx = R + __popcll(lm & ((1ull << lane) - 1));
        if (((lm >> lane) & 1) && idx < FAN_CAP) run[idx] = q;
        R += __popcll(lm);
        if (bm) break;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return min(R, FAN_CAP);
}

// The usable part of a run: the fan's triangles up to run[k] keep the apex's other neighbour W
// outside their circumcircles (exact, with the mesh's symbolic tie-break).  Going on past that
// point would give the apex run edges that belong to W or to vertices beyond it in the final
// mesh, and the flips that hand them over run one triangle per pass along the fan; stopped
// there, the next rounds cut the apex and continue the run with W as apex.  d = +1: triangle
// (apex, run[k], run[k - 1]); d = -1: (run[k], apex, run[k - 1]); run[-1] = x.
GA_INLINE int ear_fan_delaunay(const MeshDev &m, const EarCycle &e, const int *run, int R, int d, int apex, int w,
                               int x)
{
    const int lane = threadIdx.x & 63;
    const double2 A = e.bpos[apex], W = e.bpos[w];
    const int ia = e.verts[apex], iw = e.verts[w];
    for (int b = 0; b < R; b += 64) {
        int k = b + lane;
        bool bad = false;
        if (k < R) {
            int qa = run[k], qb = k ? run[k - 1] : x;
            double2 Qa = e.bpos[qa], Qb = e.bpos[qb];
            int ja = e.verts[qa], jb = e.verts[qb];
            bad = d > 0 ? incircle_at(m, A, Qa, Qb, W, ia, ja, jb, iw) > 0
                        : incircle_at(m, Qa, A, Qb, W, ja, ia, jb, iw) > 0;
        }
        unsigned long long bm = __ballot(bad);
        if (bm) return b + __ffsll((long long)bm) - 1;
    }
    return R;
}

// Largest k < R whose fan (end run[k]) is empty; -1 when even run[0]'s is not.  d = +1: apex P,
// triangle (P, E, C), span kp .. E; d = -1: apex N, triangle (E, N, C), span E .. kn.
GA_INLINE int ear_fan_end(const MeshDev &m, const EarCycle &e, int off, int h, const double *pr, const int *run,
                          int R, int d, double2 P, double2 C, double2 N, int kp, int kn)
{
    auto empty = [&](int k) {
        int q = run[k];
        double2 E = e.bpos[q];
        return d > 0 ? ear_empty_wave(m, e, off, h, pr, P, E, C, kp, q)
                     : ear_empty_wave(m, e, off, h, pr, E, N, C, q, kn);
    };
    if (empty(R - 1)) return R - 1;
    int lo = -1, hi = R - 1;   // run[lo] empty (-1: N or P itself), run[hi] blocked
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (empty(mid)) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One wave per cycle position: cand[p] = 0 (none), 1 (plain ear), 2 (fan through N) or 3 (fan
// through P), fan[p] = the fan's far end; candidates claim their spans in claims[round & 1].
__global__ __launch_bounds__(64 * EAR_WAVES) void k_ear_cand(MeshDev m, int B, int *bnd, const int *hull_off,
                                                            const int *hull_cnt, unsigned long long *claims,
                                                            int round)
{
    if (round > 0 && *ear_count(m, round - 1) == 0) return;
    __shared__ int s_run[EAR_WAVES][FAN_CAP];
    EarCycle e = ear_cycle(bnd, B);
    const int lane = threadIdx.x & 63;
    int *run = s_run[threadIdx.x >> 6];
    unsigned long long *claim = claims + (size_t)(round & 1) * B;
    const int wave = (xcd_block() * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (int p = wave; p < B; p += nw) {
        int c = 0, lo = 0, hi = 0, end = -1, off = 0, h = 0;
        if (e.alive[p]) {
            int kp = e.prv[p], kn = e.nxt[p];
            if (kp != kn) {
                double2 P = e.bpos[kp], C = e.bpos[p], N = e.bpos[kn];
                if (orient_dev(m, P, C, N) < 0) {
                    const int s = e.pset[p];
                    off = hull_off[s];
                    h = hull_cnt[s];
                    const double *pr = e.param + 5 * (size_t)s;
                    if (ear_empty_wave(m, e, off, h, pr, P, N, C, kp, kn)) {
                        c = 1; lo = kp; hi = kn;
                        int k2 = e.nxt[kn];
                        if (k2 != kp && orient_dev(m, C, N, e.bpos[k2]) == 0) {
                            int R = ear_run_wave(m, e, off, h, kn, 1, C, N, run);
                            if (R > 0) R = ear_fan_delaunay(m, e, run, R, 1, kp, e.prv[kp], kn);
                            int f = R > 0 ? ear_fan_end(m, e, off, h, pr, run, R, 1, P, C, N, kp, kn) : -1;
                            if (f >= 0) { c = 2; end = run[f]; hi = end; }
                        }
                        int k0 = e.prv[kp];
                        if (c == 1 && k0 != kn && orient_dev(m, C, P, e.bpos[k0]) == 0) {
                            int R = ear_run_wave(m, e, off, h, kp, -1, C, P, run);
                            if (R > 0) R = ear_fan_delaunay(m, e, run, R, -1, kn, e.nxt[kn], kp);
                            int f = R > 0 ? ear_fan_end(m, e, off, h, pr, run, R, -1, P, C, N, kp, kn) : -1;
                            if (f >= 0) { c = 3; end = run[f]; lo = end; }
                        }
                    }
                }
            }
        }
        if (c) {
            const unsigned long long key = ear_key(p, round, c > 1);
            const int len = ear_span(lo, hi, h);
            for (int k = lane; k < len; k += 64) atomicMax(&claim[ear_wrap(lo + k, off, h)], key);
        }
        if (lane == 0) {
            e.cand[p] = c;
            e.win[p] = end;
            if (c) atomicAdd(ear_count(m, round), 1);
        }
    }
}

// The fan of winner p in direction d (+1: run through N, apex P; -1: run through P, apex N) up to
// position `end`, by the whole wave: the run's live positions q_1 .. q_R after p, triangle i
// between q_i and q_{i+1} (q_0 = p), consecutive triangles linked across their shared apex edges.
GA_INLINE void ear_fan_cut(const MeshDev &m, const EarCycle &e, int p, int d, int end, int off, int h, int kp, int kn,
                           int *list_out, int cnt_out_slot, int op)
{
    const int lane = threadIdx.x & 63;
    const int len = ear_span(d > 0 ? p : end, d > 0 ? end : p, h) - 1;   // positions p + d .. end
    int R = 0;
    for (int b = 0; b < len; b += 64) {
        int k = b + lane + 1;
        bool live = k <= len && e.alive[ear_wrap(p + d * k, off, h)];
        R += __popcll(__ballot(live));
    }
    int T0 = 0;
    if (lane == 0) T0 = atomicAdd(m.ntri, R);
    T0 = __shfl(T0, 0, 64);
    if (T0 + R > m.cap) {
        if (lane == 0) atomicOr(m.status, GA_ERR_CAPACITY);
        return;
    }
    const int apex = d > 0 ? kp : kn, va = e.verts[apex];
    int carry = p, i0 = 0;
    for (int b = 0; b < len; b += 64) {
        int k = b + lane + 1;
        int q = k <= len ? ear_wrap(p + d * k, off, h) : -1;
        bool live = q >= 0 && e.alive[q];
        unsigned long long lm = __ballot(live), below = lm & ((1ull << lane) - 1);
        int src = below ? 63 - __clzll((long long)below) : lane;
        int qprev = __shfl(q, src, 64);
        if (live) {
            const int i = i0 + __popcll(below), T = T0 + i;
            const int qi = below ? qprev : carry;
            const int vi = e.verts[qi], vn = e.verts[q];
            int4 tn;
            if (d > 0) {
                // (P, q_{i+1}, q_i): edge 0 faces the boundary edge q_i -> q_{i+1}, edge 1 the apex edge
                const int X0 = e.owner[qi], X1 = i == 0 ? e.owner[kp] : enc(T - 1, 2);
                m.tv[T] = make_int4(va, vn, vi, 0);
                tn = make_int4(X0, X1, i == R - 1 ? -1 : enc(T + 1, 1), 0);
                if (X0 >= 0) set_link(m, enc_t(X0), enc_j(X0), enc(T, 0));
                if (i == 0 && X1 >= 0) set_link(m, enc_t(X1), enc_j(X1), enc(T, 1));
            } else {
                // (q_{i+1}, N, q_i): edge 1 faces the boundary edge q_{i+1} -> q_i, edge 0 the apex edge
                const int Y0 = i == 0 ? e.owner[p] : enc(T - 1, 2), Y1 = e.owner[q];
                m.tv[T] = make_int4(vn, va, vi, 0);
                tn = make_int4(Y0, Y1, i == R - 1 ? -1 : enc(T + 1, 0), 0);
                if (Y1 >= 0) set_link(m, enc_t(Y1), enc_j(Y1), enc(T, 1));
                if (i == 0 && Y0 >= 0) set_link(m, enc_t(Y0), enc_j(Y0), enc(T, 0));
            }
            m.tn[T] = tn;
            m.stamp[T] = op;
            if (atomicExch(&m.lmark[T], m.list_op) != m.list_op)
                list_out[atomicAdd(&m.counters[cnt_out_slot], 1)] = T;
            e.alive[qi] = 0;
            m.vt[vi] = T;
            m.vt[vn] = T;
            if (i == 0) m.vt[va] = T;
        }
        if (lm) carry = __shfl(q, 63 - __clzll((long long)lm), 64);
        i0 += __popcll(lm);
    }
    if (lane == 0) {
        const int Tl = T0 + R - 1;
        if (d > 0) {
            e.owner[kp] = enc(Tl, 2);
            e.nxt[kp] = end;
            e.prv[end] = kp;
        } else {
            e.owner[end] = enc(Tl, 2);
            e.nxt[end] = kn;
            e.prv[kn] = end;
        }
    }
}

// One wave per candidate: a winner (its key at every position of its span) cuts its ear or fan.
// Clears the claims of the next round (written two rounds ago, read by the previous apply).
__global__ __launch_bounds__(64 * EAR_WAVES) void k_ear_apply(MeshDev m, int B, int *bnd, const int *hull_off,
                                                             const int *hull_cnt, unsigned long long *claims,
                                                             int round, int *list_out, int cnt_out_slot, int op)
{
    bool any = *ear_count(m, round) != 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *ear_count(m, round + 1) = 0;
    if (!any) return;
    EarCycle e = ear_cycle(bnd, B);
    const unsigned long long *claim = claims + (size_t)(round & 1) * B;
    unsigned long long *other = claims + (size_t)((round + 1) & 1) * B;
    GRID_STRIDE(k, B) other[k] = 0;
    const int lane = threadIdx.x & 63;
    const int wave = (xcd_block() * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (int p = wave; p < B; p += nw) {
        const int c = e.cand[p];
        if (!c) continue;
        const unsigned long long key = ear_key(p, round, c > 1);
        // a winner holds p, so no other winner changes p's links or its span
        if (claim[p] != key) continue;
        const int s = e.pset[p], off = hull_off[s], h = hull_cnt[s];
        const int kp = e.prv[p], kn = e.nxt[p], end = e.win[p];
        const int lo = c == 3 ? end : kp, hi = c == 2 ? end : kn, len = ear_span(lo, hi, h);
        bool lost = false;
        for (int k = lane; k < len; k += 64)
            if (claim[ear_wrap(lo + k, off, h)] != key) lost = true;
        if (__any(lost)) continue;
        if (c == 1) {
            if (lane == 0) ear_cut(m, e, p, list_out, cnt_out_slot, op);
        } else {
            ear_fan_cut(m, e, p, c == 2 ? 1 : -1, end, off, h, kp, kn, list_out, cnt_out_slot, op);
        }
    }
}

// All rounds of every set in one launch, one workgroup per set (batches of many small sets);
// counters[C_EAR] = the most rounds a set ran.
__global__ __launch_bounds__(EAR_THREADS) void k_ear_sets(MeshDev m, int B, int *bnd, const int *hull_cnt,
                                                          const int *hull_off, int *list_out, int cnt_out_slot,
                                                          int op)
{
    const int s = blockIdx.x, h = hull_cnt[s], off = hull_off[s];
    if (h < 3) return;
    EarCycle e = ear_cycle(bnd, B);
    ear_build_grid(e, s, h, off);
    const int tid = threadIdx.x;
    __shared__ int s_cut, s_stop;
    for (int round = 0;; round++) {
        if (round > h + 8) {   // every round with a candidate cuts at least one ear
            if (tid == 0) atomicOr(m.status, GA_ERR_WALK);
            break;
        }
        for (int k = tid; k < h; k += EAR_THREADS) e.cand[off + k] = ear_candidate(m, e, off + k, off);
        __syncthreads();
        if (tid == 0) { s_cut = 0; s_stop = 0; }
        for (int k = tid; k < h; k += EAR_THREADS) e.win[off + k] = ear_winner(e, off + k, round);
        __syncthreads();
        for (int k = tid; k < h; k += EAR_THREADS) {
            if (!e.win[off + k]) continue;
            if (ear_cut(m, e, off + k, list_out, cnt_out_slot, op)) s_cut = 1;
            else s_stop = 1;
        }
        __syncthreads();
        bool done = !s_cut || s_stop;
        if (done) {
            if (tid == 0) atomicMax(&m.counters[C_EAR], round + 1);
            break;
        }
        __syncthreads();
    }
}

void launch_bnd_detect(MeshDev &m, int *outdeg, int *cover, int *hull_cnt, int *hull_start, int *badset,
                       int *blist, int *bidx, hipStream_t st)
{
    GA_LAUNCH(k_bnd_edges, dim3(FIXED_GRID), dim3(256), 0, st, m, outdeg, cover, hull_cnt, hull_start, badset);
    GA_LAUNCH(k_bnd_cover, dim3(FIXED_GRID), dim3(256), 0, st, m, (const int *)cover, badset, blist, bidx,
              m.counters + C_BND);
}

void launch_bnd_rank(MeshDev &m, int B, int rounds, const int *blist, const int *bidx, const int *hull_start,
                     const int *hull_cnt, const int *hull_off, int *badset, int *bnd, hipStream_t st)
{
    int *succ[2] = {bnd, bnd + B}, *d[2] = {bnd + 2 * B, bnd + 3 * B};
    int *verts = bnd + 4 * B, *owner = bnd + 5 * B, *nxt = bnd + 6 * B, *prv = bnd + 7 * B, *alive = bnd + 8 * B,
        *pset = bnd + 9 * B;
    int g = std::min(4096, (B + 255) / 256);
    GA_LAUNCH(k_bnd_rank_init, dim3(g), dim3(256), 0, st, m, B, blist, bidx, hull_start, succ[0], d[0], badset);
    int cur = 0;
    if (B <= RANK_BLOCK * RANK_PER) {
        GA_LAUNCH(k_bnd_rank_block, dim3(1), dim3(RANK_BLOCK), 0, st, B, rounds, succ[0], d[0]);
    } else {
        for (int r = 0; r < rounds; r++) {
            GA_LAUNCH(k_bnd_rank_step, dim3(g), dim3(256), 0, st, B, (const int *)succ[cur], (const int *)d[cur],
                      succ[1 - cur], d[1 - cur]);
            cur = 1 - cur;
        }
    }
    launch_fill_i32(verts, -1, (size_t)B, st);
    GA_LAUNCH(k_bnd_place, dim3(g), dim3(256), 0, st, m, B, blist, (const int *)succ[cur], (const int *)d[cur],
              hull_cnt, hull_off, badset, verts, owner, nxt, prv, alive, pset, (double2 *)(bnd + bnd_pos_off(B)));
}

void launch_kill_frame(MeshDev &m, hipStream_t st)
{
    GA_LAUNCH(k_kill_frame, dim3(FIXED_GRID), dim3(256), 0, st, m);
}

// claim keys of the grid-wide rounds: two arrays of B, after the grid parameters of the sets
static unsigned long long *ear_claims(int *bnd, int B, int nsets)
{
    return (unsigned long long *)(bnd + ((bnd_pos_off(B) + 4 * B + 10 * nsets + 1) & ~1));
}

// the grid of every set, and cleared claims for the grid-wide rounds
void launch_ear_grid(MeshDev &m, int B, int *bnd, const int *hull_cnt, const int *hull_off, int nsets, hipStream_t st)
{
    GA_LAUNCH(k_ear_grid, dim3(nsets), dim3(EAR_THREADS), 0, st, m, B, bnd, hull_cnt, hull_off);
    // claims zero, and the first round's candidate count
    IntFills f;
    f.add(ear_claims(bnd, B, nsets), 4 * (size_t)B, 0);
    f.add(m.counters + C_EAR, 1, 0);
    launch_fill_multi(f, st);
}

// rounds r0 .. r0 + k - 1 as grid-wide kernels, one wave per cycle position
void launch_ear_rounds(MeshDev &m, int B, int *bnd, const int *hull_cnt, const int *hull_off, int nsets, int r0,
                       int k, int *list_out, int cnt_out_slot, int op, hipStream_t st)
{
    unsigned long long *claims = ear_claims(bnd, B, nsets);
    int gw = std::min(4096, (B + EAR_WAVES - 1) / EAR_WAVES);
    gw = (gw + 7) & ~7;
    for (int r = r0; r < r0 + k; r++) {
        GA_LAUNCH(k_ear_cand, dim3(gw), dim3(64 * EAR_WAVES), 0, st, m, B, bnd, hull_off, hull_cnt, claims, r);
        GA_LAUNCH(k_ear_apply, dim3(gw), dim3(64 * EAR_WAVES), 0, st, m, B, bnd, hull_off, hull_cnt, claims, r,
                  list_out, cnt_out_slot, op);
    }
}

// every set's ear rounds in one workgroup; counters[C_EAR] = the most rounds any set ran
void launch_ear_sets(MeshDev &m, int B, int *bnd, const int *hull_cnt, const int *hull_off, int nsets,
                     int *list_out, int cnt_out_slot, int op, hipStream_t st)
{
    GA_LAUNCH(k_ear_sets, dim3(nsets), dim3(EAR_THREADS), 0, st, m, B, bnd, hull_cnt, hull_off, list_out,
              cnt_out_slot, op);
}

// ---- canonical output ----------------------------------------------------------------------------
// Output order: (v0, v1, v2) of input indices with v0 the smallest vertex; input indices of the sets
// of a batch are consecutive ranges in set order, so this is also set order.  Every directed edge
// belongs to one triangle, so (v0, v1) alone is unique: the key is v0 << B | v1 (indices below 2^B),
// sorted on its low 2B bits.  Dead slots get all ones there, above every live key (v0 <= n - 3).
__global__ void k_canon_keys(MeshDev m, unsigned long long *keys, int *vals, int *rot, int B)
{
    int n = ntri_used(m);
    int alive = 0;
    GRID_STRIDE(t, n) {
        int4 v = m.tv[t];
        vals[t] = t;
        if (v.x == GA_DEAD) { keys[t] = ~0ull; rot[t] = 0; continue; }
        alive++;
        int4 o = make_int4(oid(m, v.x), oid(m, v.y), oid(m, v.z), 0);   // input indices
        int r = 0;
        if (o.y < ic(o, r)) r = 1;
        if (o.z < ic(o, r)) r = 2;
        rot[t] = r;
        unsigned long long a = (unsigned)ic(o, r), b = (unsigned)ic(o, (r + 1) % 3);
        keys[t] = (a << B) | b;
    }
    stat_add_full(m.stats, C_RECON_TRIS, alive);
}

void launch_canon_keys(MeshDev &m, unsigned long long *keys, int *vals, int *rot, int B, hipStream_t st)
{
    GA_LAUNCH(k_canon_keys, dim3(FIXED_GRID), dim3(256), 0, st, m, keys, vals, rot, B);
}

// 32-bit keys (indices below 2^B, B <= 20): v0 in the top B bits, the top 32 - B bits of v1
// (of its B bits) below, so one sort of 4 radix passes (B <= 16: the whole key, 2B bits) orders
// the triangles up to runs of equal keys: the same v0 and v1 within 2^(2B - 32) of each other, at
// most 256 triangles (every directed edge is in one triangle).  v1[t] keeps the whole v1 for
// k_canon_newid, which orders those runs.  Dead slots get all ones, above every live key.
__global__ void k_canon_keys32(MeshDev m, unsigned *keys, unsigned *v1s, int *vals, int *rot, int B)
{
    int n = ntri_used(m);
    int alive = 0;
    const int lo = 2 * B > 32 ? 2 * B - 32 : 0;   // v1 bits dropped from the key
    GRID_STRIDE(t, n) {
        int4 v = m.tv[t];
        vals[t] = t;
        if (v.x == GA_DEAD) { keys[t] = 0xffffffffu; rot[t] = 0; continue; }
        alive++;
        int4 o = make_int4(oid(m, v.x), oid(m, v.y), oid(m, v.z), 0);   // input indices
        int r = 0;
        if (o.y < ic(o, r)) r = 1;
        if (o.z < ic(o, r)) r = 2;
        rot[t] = r;
        const unsigned a = (unsigned)ic(o, r), b = (unsigned)ic(o, (r + 1) % 3);
        keys[t] = (a << (B - lo)) | (b >> lo);
        v1s[t] = b;
    }
    stat_add_full(m.stats, C_RECON_TRIS, alive);
}

void launch_canon_keys32(MeshDev &m, unsigned *keys, unsigned *v1s, int *vals, int *rot, int B, hipStream_t st)
{
    GA_LAUNCH(k_canon_keys32, dim3(FIXED_GRID), dim3(256), 0, st, m, keys, v1s, vals, rot, B);
}

// newid[slot] = output position << 2 | rotation of the slot's triangle: one gather per neighbour
// in k_canon_emit.  With 32-bit keys (skey, v1s non-null) a run of equal sorted keys is put in
// v1 order by the thread of its first position (insertion sort in place, at most 256 entries).
__global__ void k_canon_newid(int *order, const unsigned *skey, const unsigned *v1s, const int *rot, const int *dF,
                              int *newid)
{
    int F = *dF;
    GRID_STRIDE(r, F) {
        if (!skey) {
            int s = order[r];
            newid[s] = (r << 2) | rot[s];
            continue;
        }
        const unsigned key = skey[r];
        if (r > 0 && skey[r - 1] == key) continue;   // inside a run: its first position orders it
        int e = r + 1;
        while (e < F && skey[e] == key) e++;
        for (int i = r + 1; i < e; i++) {
            const int x = order[i];
            const unsigned vx = v1s[x];
            int j = i - 1;
            while (j >= r && v1s[order[j]] > vx) { order[j + 1] = order[j]; j--; }
            order[j + 1] = x;
        }
        for (int i = r; i < e; i++) {
            const int s = order[i];
            newid[s] = (i << 2) | rot[s];
        }
    }
}

// out9 holds global vertex ids and global triangle positions.  One thread per output record r,
// from the slot order[r] that k_canon_newid left there (the inverse of newid): neighbouring lanes
// write whole records (75 MB written at 1M points, 319 MB when each slot scattered its record),
// and the slot reads become gathers, since the canonical order follows input ids.
// Batches: rset[r] = set of output triangle r, for k_canon_set_first.
__global__ void k_canon_emit(MeshDev m, const int *order, const int *newid, const int *dF, int *out9, int *rset)
{
    const int F = *dF;
    GRID_STRIDE(r, F) {
        const int s = order[r];
        const int4 v = m.tv[s];
        const int ro = newid[s] & 3;
        const int4 nb = m.tn[s];
        int rec[9];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int si = (i + ro) % 3;
            rec[3 + i] = oid(m, ic(v, si));
            const int e = ic(nb, si);
            if (e < 0) rec[i] = -1;
            else {
                const int nu = newid[enc_t(e)];
                rec[i] = enc(nu >> 2, (enc_j(e) - (nu & 3) + 3) % 3);
            }
            rec[6 + i] = 0;
        }
        int *o = out9 + (size_t)r * 9;
#pragma unroll
        for (int i = 0; i < 9; i++) o[i] = rec[i];
        if (m.nsets > 1) rset[r] = m.vset[v.x];
    }
}

// set_first[s] = first output triangle of set s (left at -1 for a set without triangles)
__global__ void k_canon_set_first(MeshDev m, const int *rset, const int *dF, int *set_first)
{
    int F = *dF;
    if (m.nsets == 1) {
        if (blockIdx.x == 0 && threadIdx.x == 0 && F > 0) set_first[0] = 0;
        return;
    }
    GRID_STRIDE(r, F) {
        int set = rset[r];
        if (r == 0 || rset[r - 1] != set) set_first[set] = r;
    }
}

void launch_canon_emit(MeshDev &m, int *order, const unsigned *skey, const unsigned *v1s, const int *rot, int *newid,
                       const int *dF, int *out9, int *set_first, int *rset, hipStream_t st)
{
    launch_fill_i32(set_first, -1, (size_t)m.nsets, st);
    GA_LAUNCH(k_canon_newid, dim3(FIXED_GRID), dim3(256), 0, st, order, skey, v1s, rot, dF, newid);
    GA_LAUNCH(k_canon_emit, dim3(FIXED_GRID), dim3(256), 0, st, m, (const int *)order, (const int *)newid, dF, out9,
              rset);
    GA_LAUNCH(k_canon_set_first, dim3(FIXED_GRID), dim3(256), 0, st, m, (const int *)rset, dF, set_first);
}

// ---- constrained DT ---------------------------------------------------------------------------
// Constraint edges go into the finished DT by cavity retriangulation.  A constraint (a, b) walks
// from a to b and collects the triangles it crosses (its cavity) with the chains of vertices
// left and right of it.  Cavities claim their slots; every constraint that holds all of its slots
// replaces its cavity by triangulations of the two pseudo-polygons on either side of the new edge.
// Each is built by taking the chain vertex whose circle through the base edge holds no other chain
// vertex (tie-broken incircle, the order the flip pass uses) and recursing on both sides.  A cavity
// of K triangles has K + 2 vertices, so the new triangles reuse its K slots.  A constrained flip
// pass afterwards (k_flip_test skips constraint edges) makes every other edge locally Delaunay.
// The reference inserts by parallel flips instead (Cuda/cudaConstraint.cu:143-939) and marks the
// constraint edges for its flip pass (Cuda/cudaFlipping.cu:111-220); the result, the constrained
// DT, is unique under the tie-break.

__global__ void k_cons_inv(MeshDev m, int *inv)
{
    GRID_STRIDE(v, m.n) inv[m.perm[v]] = v;
}

__global__ void k_cons_map(MeshDev m, const int2 *cin, int nc, const int *inv, int2 *cv, int *done,
                           unsigned long long *chash, int chmask)
{
    GRID_STRIDE(c, nc) {
        int a = inv[cin[c].x], b = inv[cin[c].y];
        cv[c] = make_int2(a, b);
        done[c] = 0;
        if (m.vstate[a] == VS_DUP || m.vstate[b] == VS_DUP) {
            atomicOr(m.status, GA_ERR_CONS_DUP);
            done[c] = 1;
            continue;
        }
        unsigned long long k = cons_key(a, b);
        for (unsigned long long h = cons_mix(k);; h++) {
            unsigned long long prev = atomicCAS(&chash[h & (unsigned long long)chmask], 0ull, k);
            if (prev == 0 || prev == k) break;
        }
    }
}

// every vertex's anchor triangle from the live slots (frame removal leaves some anchors stale)
__global__ void k_cons_vt(MeshDev m)
{
    int n = ntri_used(m);
    GRID_STRIDE(t, n) {
        int4 tt = m.tv[t];
        if (tt.x == GA_DEAD) continue;
        m.vt[tt.x] = t;
        m.vt[tt.y] = t;
        m.vt[tt.z] = t;
    }
}

// ints of a cavity region of K triangles: K, ml, mr, pad; slots[K]; Lx[K + 2]; Rx[K + 2];
// lb[K + 1]; rb[K + 1]; ol[K + 1]; orr[K + 1]; stack[3 (K + 2)]
GA_INLINE int cons_region(int K) { return 10 * K + 18; }

struct ConsRegion {
    int *hdr, *slots, *Lx, *Rx, *lb, *rb, *ol, *orr, *stk;
    __device__ ConsRegion(int *r, int K)
    {
        hdr = r;
        slots = r + 4;
        Lx = slots + K;
        Rx = Lx + K + 2;
        lb = Rx + K + 2;
        rb = lb + K + 1;
        ol = rb + K + 1;
        orr = ol + K + 1;
        stk = orr + K + 1;
    }
};

// Walk of constraint (a, b).  Returns the cavity size K >= 1, 0 when the edge exists already, -1
// on an error (status set).  With REC the cavity is written to the region sized for Kc triangles:
// slots in walk order; Lx = a, left chain, b and Rx = a, right chain, b; lb[i] / rb[i] = the
// cavity side (t << 2 | j) on the boundary edge between Lx[i], Lx[i + 1] / Rx[i], Rx[i + 1].
template <bool REC>
__device__ int cons_walk(const MeshDev &m, int a, int b, ConsRegion *R, int Kc)
{
    double2 A = m.pos[a], B = m.pos[b];
    int t0 = m.vt[a], t = -1, k = 0;
    if (t0 < 0 || t0 >= m.cap) {
        atomicOr(m.status, GA_ERR_CONS_WALK);
        return -1;
    }
    // the triangle (a, p, q) of a's star with b strictly inside its angle at a
    for (int dir = 0; dir < 2 && t < 0; dir++) {
        int s = t0;
        for (int guard = 0; guard < 1000000; guard++) {
            int4 tt = m.tv[s];
            int ka = tt.x == a ? 0 : (tt.y == a ? 1 : 2);
            int p = ic(tt, p1(ka)), q = ic(tt, m1(ka));
            if (p == b || q == b) return 0;
            double2 P = m.pos[p], Q = m.pos[q];
            int op = orient_dev(m, A, P, B), oq = orient_dev(m, A, Q, B);
            // a vertex on the ray from a towards b lies inside the segment (b itself is not on
            // an edge of a's star)
            if ((op == 0 && (P.x - A.x) * (B.x - A.x) + (P.y - A.y) * (B.y - A.y) > 0) ||
                (oq == 0 && (Q.x - A.x) * (B.x - A.x) + (Q.y - A.y) * (B.y - A.y) > 0)) {
                atomicOr(m.status, GA_ERR_CONS_POINT);
                return -1;
            }
            if (op > 0 && oq < 0) {
                t = s;
                k = ka;
                break;
            }
            // counter-clockwise: across (a, q), the side opposite p; clockwise: across (a, p)
            int e = ic(m.tn[s], dir == 0 ? p1(ka) : m1(ka));
            if (e < 0) break;
            s = enc_t(e);
            if (s == t0) break;
        }
    }
    if (t < 0) {
        atomicOr(m.status, GA_ERR_CONS_WALK);
        return -1;
    }
    int4 tt = m.tv[t];
    int p = ic(tt, p1(k)), q = ic(tt, m1(k));   // p right of a -> b, q left
    int K = 1, ml = 1, mr = 1;
    if (REC) {
        R->slots[0] = t;
        R->Lx[0] = a; R->Lx[1] = q;
        R->Rx[0] = a; R->Rx[1] = p;
        R->lb[0] = enc(t, p1(k));   // edge (q, a)
        R->rb[0] = enc(t, m1(k));   // edge (a, p)
    }
    int cur = t, ex = k;
    for (int guard = 0; guard < m.cap; guard++) {
        int e = ic(m.tn[cur], ex);
        if (e < 0) break;
        int u = enc_t(e), j = enc_j(e);
        int4 ut = m.tv[u];
        // entered across (r1, r2): r1 left of a -> b, r2 right; w is the far vertex
        int w = ic(ut, j), r1 = ic(ut, p1(j)), r2 = ic(ut, m1(j));
        if (cons_has(m, r1, r2)) {
            atomicOr(m.status, GA_ERR_CONS_CROSS);
            return -1;
        }
        if (REC && K >= Kc) break;   // the mesh changed since the count: report below
        if (REC) R->slots[K] = u;
        K++;
        if (w == b) {
            if (REC) {
                R->lb[ml] = enc(u, m1(j));   // edge (r1, b)
                R->rb[mr] = enc(u, p1(j));   // edge (b, r2)
                R->Lx[ml + 1] = b;
                R->Rx[mr + 1] = b;
                R->hdr[0] = K; R->hdr[1] = ml; R->hdr[2] = mr;
            }
            return K;
        }
        int o = orient_dev(m, A, B, m.pos[w]);
        if (o == 0) {   // w is on the segment (the walk has not reached b)
            atomicOr(m.status, GA_ERR_CONS_POINT);
            return -1;
        }
        if (o > 0) {   // w joins the left chain; leave across (w, r2)
            if (REC) { R->lb[ml] = enc(u, m1(j)); R->Lx[ml + 1] = w; }
            ml++;
            ex = p1(j);
        } else {       // w joins the right chain; leave across (r1, w)
            if (REC) { R->rb[mr] = enc(u, p1(j)); R->Rx[mr + 1] = w; }
            mr++;
            ex = m1(j);
        }
        cur = u;
    }
    atomicOr(m.status, GA_ERR_CONS_WALK);
    return -1;
}

__global__ void k_cons_count(MeshDev m, const int2 *cv, int nc, int *done, int *cnt, int *rsize)
{
    GRID_STRIDE(c, nc) {
        int K = 0;
        if (!done[c]) {
            K = cons_walk<false>(m, cv[c].x, cv[c].y, nullptr, 0);
            if (K <= 0) {
                done[c] = 1;   // present already, or an error (status set)
                K = 0;
            }
        }
        cnt[c] = K;
        rsize[c] = K ? cons_region(K) : 0;
        if (K) atomicAdd(&m.counters[C_CONS_PEND], 1);
    }
}

// walks again, recording the cavity, and claims its slots (larger key = smaller constraint id)
__global__ void k_cons_collect(MeshDev m, const int2 *cv, int nc, const int *cnt, const int *roff, int *work,
                               int op)
{
    GRID_STRIDE(c, nc) {
        int K = cnt[c];
        if (!K) continue;
        ConsRegion R(work + roff[c], K);
        R.hdr[0] = 0;
        if (cons_walk<true>(m, cv[c].x, cv[c].y, &R, K) != K) {
            atomicOr(m.status, GA_ERR_CONS_WALK);
            continue;
        }
        unsigned long long key = ((unsigned long long)op << 32) | (0xffffffffu - (unsigned)c);
        for (int i = 0; i < K; i++) atomicMax(&m.claim[R.slots[i]], key);
    }
}

// boundary edge of the cavity on side k of new triangle T: the old cavity side `side` had the
// outside neighbour `out` (-1 on the hull); k_relink repairs the outside end
GA_INLINE void cons_boundary(const MeshDev &m, int T, int k, int side, int out)
{
    set_link(m, T, k, mark_old(out));
    ((int *)(m.remap + enc_t(side)))[enc_j(side)] = enc(T, k);
}

// Triangulates the pseudo-polygon E(0), E(1..M), E(M + 1) lying left of E(0) -> E(M + 1), with
// E(i) = X[i] (or X[M + 1 - i] when rev); bs / bo: cavity side and outside neighbour of the
// boundary edge between E(i) and E(i + 1) at index i (M - i when rev).  New triangles take slots
// from slots[*s]; returns the triangle on the base edge E(0) E(M + 1) (its side 2).
// Run by one thread (width 1) or one wave (width 64).  In a wave every lane runs the same
// sequential control flow on the same values (the stack in stk is written with identical values
// by all lanes, and each lane reads its own writes); only the choice of the chain vertex is split
// across lanes, and only lane 0 writes the mesh.
__device__ int cons_fill(const MeshDev &m, const int *X, int M, bool rev, const int *bs, const int *bo,
                         const int *slots, int *s, int *stk, int lane, int width)
{
#define CE(i) (rev ? X[M + 1 - (i)] : X[(i)])
#define CB(i) (rev ? M - (i) : (i))
    int top = -1, sp = 0;
    stk[0] = 1; stk[1] = M; stk[2] = -1;
    sp = 1;
    while (sp > 0) {
        sp--;
        int lo = stk[3 * sp], hi = stk[3 * sp + 1], pe = stk[3 * sp + 2];
        int x = CE(lo - 1), y = CE(hi + 1);
        // the chain vertex whose circle through x and y holds no other chain vertex: the minimum
        // of the order "inside the circle of" (the circles through x and y on one side are
        // nested, and the tie-break makes the order strict).  Each lane scans a strided part,
        // a butterfly combines them, and every lane ends with the index a sequential scan finds.
        int ci = -1;
        for (int i = lo + lane; i <= hi; i += width)
            if (ci < 0 || incircle_dev(m, x, y, CE(ci), CE(i)) > 0) ci = i;
        if (width == 64)
            for (int d = 32; d > 0; d >>= 1) {
                int o = __shfl_xor(ci, d, 64);
                if (o >= 0 && (ci < 0 || incircle_dev(m, x, y, CE(ci), CE(o)) > 0)) ci = o;
            }
        int cv = CE(ci);
        int T = slots[(*s)++];
        if (lane == 0) {
            m.tv[T] = make_int4(x, y, cv, 0);
            if (orient_dev(m, m.pos[x], m.pos[y], m.pos[cv]) <= 0) atomicOr(m.status, GA_ERR_CONS_TRI);
            if (pe >= 0) {
                set_link(m, T, 2, pe);
                set_link(m, enc_t(pe), enc_j(pe), enc(T, 2));
            }
            // side 1: edge (cv, x); side 0: edge (y, cv)
            if (ci == lo) cons_boundary(m, T, 1, bs[CB(lo - 1)], bo[CB(lo - 1)]);
            if (ci == hi) cons_boundary(m, T, 0, bs[CB(hi)], bo[CB(hi)]);
            m.vt[x] = T;
            m.vt[y] = T;
            m.vt[cv] = T;
        }
        if (pe < 0) top = T;
        if (ci > lo) {
            stk[3 * sp] = lo; stk[3 * sp + 1] = ci - 1; stk[3 * sp + 2] = enc(T, 1);
            sp++;
        }
        if (ci < hi) {
            stk[3 * sp] = ci + 1; stk[3 * sp + 1] = hi; stk[3 * sp + 2] = enc(T, 0);
            sp++;
        }
    }
#undef CE
#undef CB
    return top;
}

// Cavities below CONS_WAVE_K triangles: one thread per constraint; larger ones: one wave per
// constraint (wave-uniform control flow throughout).  A slot belongs to at most one winning
// cavity, so the two kernels never touch the same slots.
#define CONS_WAVE_K 32

__device__ void cons_commit_one(const MeshDev &m, int c, int K, int *done, const int *roff, int *work,
                                int *modlist, int op, int lane, int width)
{
    ConsRegion R(work + roff[c], K);
    if (R.hdr[0] != K) return;   // collect failed (status set)
    unsigned long long key = ((unsigned long long)op << 32) | (0xffffffffu - (unsigned)c);
    bool lost = false;
    for (int i = lane; i < K; i += width) lost |= m.claim[R.slots[i]] != key;
    if (width == 64 ? __ballot(lost) != 0 : lost) return;
    int ml = R.hdr[1], mr = R.hdr[2];
    // outside neighbours of the boundary, read before any slot is rewritten (every lane writes
    // the whole list, so it reads back only its own writes)
    for (int i = 0; i <= ml; i++) R.ol[i] = ic(m.tn[enc_t(R.lb[i])], enc_j(R.lb[i]));
    for (int i = 0; i <= mr; i++) R.orr[i] = ic(m.tn[enc_t(R.rb[i])], enc_j(R.rb[i]));
    for (int i = lane; i < K; i += width) m.stamp[R.slots[i]] = op;
    int s = 0;
    int tl = cons_fill(m, R.Lx, ml, false, R.lb, R.ol, R.slots, &s, R.stk, lane, width);
    int tr = cons_fill(m, R.Rx, mr, true, R.rb, R.orr, R.slots, &s, R.stk, lane, width);
    int q = 0;
    if (lane == 0) {
        set_link(m, tl, 2, enc(tr, 2));   // the constraint edge (a, b)
        set_link(m, tr, 2, enc(tl, 2));
        q = atomicAdd(&m.counters[C_CONS_MOD], K);
        done[c] = 1;
    }
    if (width == 64) q = __shfl(q, 0, 64);
    for (int i = lane; i < K; i += width) modlist[q + i] = R.slots[i];
}

__global__ void k_cons_commit_small(MeshDev m, int nc, int *done, const int *cnt, const int *roff, int *work,
                                    int *modlist, int op)
{
    GRID_STRIDE(c, nc) {
        int K = cnt[c];
        if (K > 0 && K < CONS_WAVE_K) cons_commit_one(m, c, K, done, roff, work, modlist, op, 0, 1);
    }
}

__global__ __launch_bounds__(256) void k_cons_commit_wave(MeshDev m, int nc, int *done, const int *cnt,
                                                          const int *roff, int *work, int *modlist, int op)
{
    int lane = threadIdx.x & 63;
    int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (int c = wave; c < nc; c += nw) {
        int K = cnt[c];
        if (K >= CONS_WAVE_K) cons_commit_one(m, c, K, done, roff, work, modlist, op, lane, 64);
    }
}

void launch_cons_setup(MeshDev &m, const int2 *cin, int nc, int *inv, int2 *cv, int *done,
                       unsigned long long *chash, int chmask, hipStream_t st)
{
    GA_LAUNCH(k_cons_inv, dim3(FIXED_GRID), dim3(256), 0, st, m, inv);
    GA_LAUNCH(k_cons_map, dim3(FIXED_GRID), dim3(256), 0, st, m, cin, nc, (const int *)inv, cv, done, chash, chmask);
    GA_LAUNCH(k_cons_vt, dim3(FIXED_GRID), dim3(256), 0, st, m);
}

void launch_cons_count(MeshDev &m, const int2 *cv, int nc, int *done, int *cnt, int *rsize, hipStream_t st)
{
    GA_LAUNCH(k_cons_count, dim3(FIXED_GRID), dim3(256), 0, st, m, cv, nc, done, cnt, rsize);
}

void launch_cons_insert(MeshDev &m, const int2 *cv, int nc, int *done, const int *cnt, const int *roff,
                        int *work, int *modlist, int op, hipStream_t st)
{
    GA_LAUNCH(k_cons_collect, dim3(FIXED_GRID), dim3(256), 0, st, m, cv, nc, cnt, roff, work, op);
    GA_LAUNCH(k_cons_commit_small, dim3(FIXED_GRID), dim3(256), 0, st, m, nc, done, cnt, roff, work, modlist, op);
    GA_LAUNCH(k_cons_commit_wave, dim3(FIXED_GRID), dim3(256), 0, st, m, nc, done, cnt, roff, work, modlist, op);
    GA_LAUNCH(k_relink, dim3(FIXED_GRID), dim3(256), 0, st, m, (const int *)modlist, C_CONS_MOD, op, -1);
}
