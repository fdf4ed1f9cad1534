// alpha.hip -- alpha-shape spectrum, intervals and master list on gfx950.
//
// Reference semantics (paths relative to the reference root):
//   edge ownership / ids      AlphaShape/kernel.cu:160-217, :252-308  -- triangle t owns the edge
//                             opposite vtx[i] iff vtx[i+1] < vtx[i+2] or tri[i] == -1; edges are
//                             numbered by (t, i) through an exclusive scan of per-triangle counts.
//   rho                       AlphaShape/predicates.h:1297-1357 (birth_time): same FP64 expression
//                             order (no FMA), so spectrum values match the reference bit for bit.
//   attachment                AlphaShape/predicates.h:1359-1444, kernel.cu:310-358: attached iff
//                             the exact |i-j|^2 - |i+j-2k|^2 is >= 0 for the owning triangle's apex,
//                             else for the opposite apex.
//   exact order + merging     Alpha_Shapes_2D.cu:308-522, kernel.cu:360-401: the reference sorts by
//                             FP64 value, runs odd-even exact fix-up rounds on the device with a
//                             200-double expansion budget and sends overflow to a CPU insertion
//                             sort.  Here every entry carries a certified interval; a radix sort by
//                             FP64 value, a prefix-max / suffix-min scan of the interval ends and a
//                             per-cluster exact insertion sort (integer arithmetic on the device)
//                             give the exact order with no host pass.  Equal exact values share a
//                             rank; rank = 1 + number of smaller distinct values.
//   intervals                 kernel.cu:427-462 (edges), :518-624 (vertices; here by per-edge
//                             atomic min/max instead of a fan walk -- same result).
//   master list               kernel.cu:626-690 + Alpha_Shapes_2D.cu:641-727: stable radix sort of
//                             rank keys; entries with rank < 1 dropped; spec_pointers = [-1] + the
//                             first index of each rank.
#include <cstring>
#include <rocprim/rocprim.hpp>
#include "dev_common.h"
#include "kernels.h"
#include "alpha_host.h"
#include <algorithm>
#include <initializer_list>
#include <vector>

#define GS(i, n) for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)
#define AGRID 2048

// Radix sort of 32-bit keys with 32-bit values: 1024-thread blocks of 8 items and 8-bit digits.
// rocPRIM carries no tuned onesweep configuration for gfx950; this one sorts 13.5M rank keys in
// 332 us against 391 us for its default on MI355X.
using SortCfg32 = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 8>, rocprim::kernel_config<1024, 8>, 8,
                                        rocprim::block_radix_rank_algorithm::match>>;
#define U53 1.1102230246251565e-16

GA_INLINE double2 P(const double2 *xy, int v) { return xy[v]; }

// ---- rho in the reference's expression order -------------------------------------------------
GA_INLINE double rho_tri(double2 pa, double2 pb, double2 pc)
{
    double x1 = pa.x, y1 = pa.y, x2 = pb.x, y2 = pb.y, x3 = pc.x, y3 = pc.y;
    double ax = x1 - x2, ay = y1 - y2, bx = x2 - x3, by = y2 - y3, cx = x3 - x1, cy = y3 - y1;
    ax = ax * ax; ay = ay * ay; bx = bx * bx; by = by * by; cx = cx * cx; cy = cy * cy;
    double a = ax + ay, b = bx + by, c = cx + cy;
    double ar1 = x2 * y3 - x3 * y2 - x1 * y3 + x3 * y1 + x1 * y2 - x2 * y1;
    double den = 4 * ar1 * ar1;
    return (a * b * c) / den;
}

GA_INLINE double rho_edge(double2 pa, double2 pb)
{
    double ax = pa.x - pb.x, ay = pa.y - pb.y;
    ax = ax * ax; ay = ay * ay;
    return (ax + ay) / 4;
}

// ---- exact helpers (integer arithmetic over a common unit) ------------------------------------
#define ANL 40
typedef BigInt<ANL> BI;

GA_INLINE void bi_sq_dist(BI &r, const BI &ax, const BI &ay, const BI &bx, const BI &by, int *ovf)
{
    BI dx, dy, t;
    bi_sub(dx, ax, bx, ovf);
    bi_sub(dy, ay, by, ovf);
    bi_mul(t, dx, dx, ovf);
    bi_mul(dy, dy, dy, ovf);
    bi_add(r, t, dy, ovf);
}

// entries: st 0 = edge (p, q), 1 = triangle (a, b, c).  Returns sign(rho_A - rho_B).
__device__ __noinline__ int rho_cmp_exact(int sta, const double2 *pa, int stb, const double2 *pb, int *ovf)
{
    // Equal values are certified by residues (the common case of key-equal neighbours; a
    // different pair fails at the first prime), the sign of a different pair in fixed-width
    // registers when the differences fit 58 bits; the big integers take the rest.  Comparing in
    // fixed width first made the 2M lattice set's key-equal neighbours 6x slower.
    if (rho_equal_res(sta, pa, stb, pb) == 1) return 0;
    const int f = rho_cmp_fixed(sta, pa, stb, pb);
    if (f != 2) return f;
    double v[12];
    int na = sta ? 3 : 2, nb = stb ? 3 : 2, k = 0;
    for (int i = 0; i < na; i++) { v[k++] = pa[i].x; v[k++] = pa[i].y; }
    for (int i = 0; i < nb; i++) { v[k++] = pb[i].x; v[k++] = pb[i].y; }
    int s = min_ulp_exp(v, k);
    BI X[6], Y[6];
    for (int i = 0; i < na; i++) { bi_from_double(X[i], pa[i].x, s, ovf); bi_from_double(Y[i], pa[i].y, s, ovf); }
    for (int i = 0; i < nb; i++) { bi_from_double(X[3 + i], pb[i].x, s, ovf); bi_from_double(Y[3 + i], pb[i].y, s, ovf); }
    // value = num / den (den > 0): edge num = |pq|^2, den = 1 ; triangle num = l1 l2 l3, den = ar^2
    BI num[2], den[2], t1, t2;
    int stv[2] = {sta, stb};
    for (int e = 0; e < 2; e++) {
        int o = e * 3;
        if (!stv[e]) {
            bi_sq_dist(num[e], X[o], Y[o], X[o + 1], Y[o + 1], ovf);
            bi_zero(den[e]);
            den[e].l[0] = 1; den[e].n = 1;
        } else {
            BI l1, l2, l3;
            bi_sq_dist(l1, X[o], Y[o], X[o + 1], Y[o + 1], ovf);
            bi_sq_dist(l2, X[o + 1], Y[o + 1], X[o + 2], Y[o + 2], ovf);
            bi_sq_dist(l3, X[o + 2], Y[o + 2], X[o], Y[o], ovf);
            bi_mul(t1, l1, l2, ovf);
            bi_mul(num[e], t1, l3, ovf);
            // ar = (bx-ax)(cy-ay) - (by-ay)(cx-ax)
            BI ux, uy, vx, vy;
            bi_sub(ux, X[o + 1], X[o], ovf); bi_sub(uy, Y[o + 1], Y[o], ovf);
            bi_sub(vx, X[o + 2], X[o], ovf); bi_sub(vy, Y[o + 2], Y[o], ovf);
            bi_mul(t1, ux, vy, ovf);
            bi_mul(t2, uy, vx, ovf);
            bi_sub(t1, t1, t2, ovf);
            bi_mul(den[e], t1, t1, ovf);
        }
    }
    bi_mul(t1, num[0], den[1], ovf);
    bi_mul(t2, num[1], den[0], ovf);
    bi_sub(t1, t1, t2, ovf);
    return bi_sign(t1);
}

// attachment: exact sign of |i-j|^2 - |i+j-2k|^2 (>= 0 attached)
__device__ __noinline__ int attached_exact(double2 i, double2 j, double2 k, int *ovf)
{
    const int q = attached_i128(i, j, k);
    if (q >= 0) return q;
    double v[6] = {i.x, i.y, j.x, j.y, k.x, k.y};
    int s = min_ulp_exp(v, 6) - 1;   // -1: 2k stays an integer multiple of the unit
    BigInt<16> ix, iy, jx, jy, kx, ky, d1, sx, sy, t;
    bi_from_double(ix, i.x, s, ovf); bi_from_double(iy, i.y, s, ovf);
    bi_from_double(jx, j.x, s, ovf); bi_from_double(jy, j.y, s, ovf);
    bi_from_double(kx, k.x, s, ovf); bi_from_double(ky, k.y, s, ovf);
    BigInt<16> dx, dy;
    bi_sub(dx, ix, jx, ovf); bi_sub(dy, iy, jy, ovf);
    bi_mul(dx, dx, dx, ovf); bi_mul(dy, dy, dy, ovf);
    bi_add(d1, dx, dy, ovf);
    bi_add(sx, ix, jx, ovf); bi_sub(sx, sx, kx, ovf); bi_sub(sx, sx, kx, ovf);
    bi_add(sy, iy, jy, ovf); bi_sub(sy, sy, ky, ovf); bi_sub(sy, sy, ky, ovf);
    bi_mul(sx, sx, sx, ovf); bi_mul(sy, sy, sy, ovf);
    bi_add(t, sx, sy, ovf);
    bi_sub(t, d1, t, ovf);
    return bi_sign(t) >= 0;
}

// per spectrum entry id: certified interval of its value and the value (32 bytes, one line)
struct alignas(32) EntryIv {
    double lo, hi, v, pad;
};

struct AlphaK {
    const double2 *xy;
    double2 *cxy;        // [3 F] corner points of every triangle in vertex order (k_tri_rho)
    const int *tris;     // F x 9
    int n, F;
    int *cnt, *eoff;     // [F]
    int *edge_t, *edge_v, *edge_att, *edge_rank, *edge_numin, *edge_numax;  // [E]
    int *tri_rank;       // [F]
    unsigned long long *key;   // [S0] single set: FP64 value bits of the triangles; batches: sort keys
    int *id, *id2;       // [S0] entry ids (t or F + e)
    unsigned long long *key2;
    EntryIv *iv;         // [F + E] per entry id: certified interval (lo, hi) and FP64 value
    double *slo, *shi;   // [S0] sorted
    int *flag;           // [S0] cluster starts
    int *gflag;          // [S0] group starts (exact)
    int *rank;           // [S0]
    unsigned long long *specbits;  // [S0 + 1]
    int *vnumin, *vnumax;  // [n]
    int2 *vnum;            // [n] (vnumin, vnumax) side by side while k_edge_intervals builds them
    int *status;
    int *counters;       // [ALPHA_COUNTER_INTS]: 64 counters, then the stripes
    int *bigstart;             // large clusters (start position)
    int *xclu;                 // small clusters that need exact comparisons (start position)
    unsigned long long *xpair; // their pairwise exact order, 2 bits per pair (k_clusters_pairs)
    int *xatt;                 // edges whose attachment needs the exact predicate
    int bigcap;
    int nS0;
    int skip_if_big;     // first pass of a single set: kernels after k_clusters return when it found large clusters
    // batches: set of every vertex, first spectrum slot and rank base of every set
    const int *vset;     // [n] (NULL for a single set)
    int nsets;
    const int *specstart;  // [nsets] slot of rank 0 of set s in the concatenated spectra
};

#define A_NUNATT 0
#define A_EXACT_CMP 1
#define A_EXACT_ATT 2
#define A_CLUSTERS 3
#define A_MAXCLUSTER 4
#define A_BIGCLUSTER 5
#define A_BADMESH 6      // caller triangles that fail k_mesh_check
#define A_EXCLU 7        // small clusters left to the exact kernel (list in xclu)
#define A_ATT_LIST 8     // edges whose attachment the filter leaves open (list in xatt)
#define A_MK 9           // master-list entries with a rank (the sorted list's length), set by k_master_decode
// Counters that every wave of a large grid updates live in stripes 128 bytes apart (atomics on
// one address or line from many waves serialise at about 11 ns each), chosen by the wave's
// global index.  After the 64 counters (ints): the single set's largest FP64 bit pattern of a
// value and the largest complement of one (u64 stripes), then the statistics ST_* (int stripes),
// all zero before the run; the final read folds the statistics into their counters.
#define URANGE_STRIPES 64
#define STRIPE_U64 16    // u64 between stripes
#define STRIPE_INT 32    // ints between stripes
#define UR_MAX_OFF 64                                            // ints from counters
#define UR_MINC_OFF (UR_MAX_OFF + 2 * STRIPE_U64 * URANGE_STRIPES)
#define ST_OFF (UR_MINC_OFF + 2 * STRIPE_U64 * URANGE_STRIPES)
#define ALPHA_COUNTER_INTS (ST_OFF + STRIPE_INT * URANGE_STRIPES)
#define ST_CLUSTERS 0    // sum -> A_CLUSTERS
#define ST_MAXCLUSTER 1  // max -> A_MAXCLUSTER
#define ST_EXACT_CMP 2   // sum -> A_EXACT_CMP
#define ST_EXACT_ATT 3   // sum -> A_EXACT_ATT
GA_INLINE int wave_stripe()
{
    return (int)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % URANGE_STRIPES);
}
// wave sum of v into statistic k (call with every lane of the wave active)
GA_INLINE void stat_add(int *counters, int k, int v)
{
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    if (__lane_id() == 0 && v) atomicAdd(&counters[ST_OFF + wave_stripe() * STRIPE_INT + k], v);
}

GA_INLINE int tv(const AlphaK &a, int t, int i) { return a.tris[(size_t)t * 9 + 3 + i]; }
GA_INLINE int tn(const AlphaK &a, int t, int i) { return a.tris[(size_t)t * 9 + i]; }
GA_INLINE int pl1(int i) { return i == 2 ? 0 : i + 1; }
// first pass of a single set that found large clusters: the rest is redone after their order
GA_INLINE bool big_pending(const AlphaK &a) { return a.skip_if_big && a.counters[A_BIGCLUSTER] > 0; }
// corner i of triangle t (its point, from the per-triangle copy: one read near the triangle's
// other data instead of a gather at the vertex id)
GA_INLINE double2 cpt(const AlphaK &a, int t, int i) { return a.cxy[(size_t)t * 3 + i]; }
GA_INLINE int mi1(int i) { return i == 0 ? 2 : i - 1; }

// Every edge is listed once, by its owner: the triangle in which it runs from the smaller vertex
// id to the larger, or the only triangle of a hull edge.  Edges of one owner are consecutive,
// in slot order, from eoff[owner].
GA_INLINE bool edge_owned(const AlphaK &a, int t, int i)
{
    return tv(a, t, pl1(i)) < tv(a, t, mi1(i)) || tn(a, t, i) == -1;
}

// Caller triangles, before any kernel follows a link: three distinct vertex ids in [0, n), and
// every neighbour either -1 or a slot of a triangle in [0, F) that links back over the same two
// vertices in the other direction.  Counts the triangles that fail in A_BADMESH.
__global__ void k_mesh_check(AlphaK a)
{
    GS(t, a.F) {
        bool bad = false;
        int v[3];
        for (int i = 0; i < 3; i++) {
            v[i] = tv(a, t, i);
            bad |= v[i] < 0 || v[i] >= a.n;
        }
        bad |= v[0] == v[1] || v[1] == v[2] || v[2] == v[0];
        for (int i = 0; i < 3 && !bad; i++) {
            int nb = tn(a, t, i);
            if (nb == -1) continue;
            int u = nb >> 2, j = nb & 3;
            if (nb < 0 || u >= a.F || j > 2) { bad = true; break; }
            bad |= tn(a, u, j) != t * 4 + i || tv(a, u, pl1(j)) != v[mi1(i)] || tv(a, u, mi1(j)) != v[pl1(i)];
        }
        if (bad) atomicAdd(&a.counters[A_BADMESH], 1);
    }
}

__global__ void k_edge_count(AlphaK a)
{
    GS(t, a.F) {
        int c = 0;
        for (int i = 0; i < 3; i++) c += edge_owned(a, t, i);
        a.cnt[t] = c;
    }
}

__global__ void k_edges(AlphaK a)
{
    GS(t, a.F) {
        int e = a.eoff[t];
        for (int i = 0; i < 3; i++)
            if (edge_owned(a, t, i)) {
                a.edge_t[e] = t;
                a.edge_v[e] = i;
                e++;
            }
    }
}

// Single set: the radix sort orders 32-bit keys (key32).  Nonnegative doubles order like their
// bit patterns u, so key32 = (u >> sh) - (umin >> sh), with the smallest shift sh that fits the
// value range [umin, umax] of the run in 32 bits, is monotone in the value; values within 2^sh
// of each other's bits may tie.  The cluster boundaries below hold for any order, so ties only
// make clusters, which are put in exact order there.  (Uniform points keep about 26 mantissa
// bits.)  The range comes from a wave maximum of u and of ~u, one atomic per wave.
// wave maxima of the thread maxima (call with every lane of the wave active), one atomic each
GA_INLINE void range_note(const AlphaK &a, unsigned long long mx, unsigned long long mc)
{
    for (int d = 32; d > 0; d >>= 1) {
        const unsigned long long ox = __shfl_xor(mx, d, 64), oc = __shfl_xor(mc, d, 64);
        mx = ox > mx ? ox : mx;
        mc = oc > mc ? oc : mc;
    }
    if (__lane_id() == 0) {
        const int sl = wave_stripe();
        if (mx) atomicMax((unsigned long long *)(a.counters + UR_MAX_OFF) + sl * STRIPE_U64, mx);
        if (mc) atomicMax((unsigned long long *)(a.counters + UR_MINC_OFF) + sl * STRIPE_U64, mc);
    }
}
struct Key32Map {
    unsigned long long base;
    int sh;
};
// the map of the run from the stripes; every thread of the block calls it (one barrier)
__device__ Key32Map key32_map(const AlphaK &a)
{
    __shared__ unsigned long long s_r[2];
    if (threadIdx.x < 64) {
        unsigned long long mx = 0, mc = 0;
        for (int i = threadIdx.x; i < URANGE_STRIPES; i += 64) {
            mx = max(mx, ((const unsigned long long *)(a.counters + UR_MAX_OFF))[i * STRIPE_U64]);
            mc = max(mc, ((const unsigned long long *)(a.counters + UR_MINC_OFF))[i * STRIPE_U64]);
        }
        for (int d = 32; d > 0; d >>= 1) {
            mx = max(mx, (unsigned long long)__shfl_xor(mx, d, 64));
            mc = max(mc, (unsigned long long)__shfl_xor(mc, d, 64));
        }
        if (threadIdx.x == 0) { s_r[0] = mx; s_r[1] = mc; }
    }
    __syncthreads();
    const unsigned long long umax = s_r[0], umin = ~s_r[1];
    int sh = 0;
    while (sh < 32 && (umax >> sh) - (umin >> sh) > 0xffffffffull) sh++;
    return Key32Map{umin >> sh, sh};
}
GA_INLINE unsigned key32(const Key32Map &m, unsigned long long u) { return (unsigned)((u >> m.sh) - m.base); }
GA_INLINE unsigned long long value_bits(double v) { return (unsigned long long)__double_as_longlong(v); }
// Batches order entries by (set, value): the set above the value's top 40 bits, so certified
// clusters never span sets (equal values in different sets, e.g. repeated requests, are never
// compared).  The low value bits are not kept; they are recomputed where needed.
GA_INLINE unsigned long long batch_key(int set, double v)
{
    return ((unsigned long long)set << 40) | ((unsigned long long)__double_as_longlong(v) >> 24);
}

// relative error bound of the certified rho estimate used for ordering
__global__ void k_tri_rho(AlphaK a)
{
    unsigned long long mx = 0, mc = 0;   // value-bit range of this thread's triangles
    GS(t, a.F) {
        {
            double2 A = a.xy[tv(a, t, 0)], B = a.xy[tv(a, t, 1)], Cc = a.xy[tv(a, t, 2)];
            a.cxy[(size_t)t * 3] = A;
            a.cxy[(size_t)t * 3 + 1] = B;
            a.cxy[(size_t)t * 3 + 2] = Cc;
            double v = rho_tri(A, B, Cc);
            const unsigned long long u = value_bits(v);
            a.key[t] = a.vset ? batch_key(a.vset[tv(a, t, 0)], v) : u;
            a.id[t] = t;
            // ar with Shewchuk's orient2d bound: |ar - det| <= eb
            double detleft = (A.x - Cc.x) * (B.y - Cc.y), detright = (A.y - Cc.y) * (B.x - Cc.x);
            double det = detleft - detright;
            double eb = GA_CCWERRA * (fabs(detleft) + fabs(detright));
            double la = (A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y);
            double lb = (B.x - Cc.x) * (B.x - Cc.x) + (B.y - Cc.y) * (B.y - Cc.y);
            double lc = (Cc.x - A.x) * (Cc.x - A.x) + (Cc.y - A.y) * (Cc.y - A.y);
            double lo, hi;
            if (eb < 0.25 * fabs(det)) {
                double r = eb / fabs(det);
                double est = (la * lb * lc) / (4.0 * det * det);
                double rel = 32.0 * U53 + 3.0 * r;
                lo = est * (1.0 - rel);
                hi = est * (1.0 + rel);
            } else {
                lo = 0.0;
                hi = INFINITY;
            }
            a.iv[t] = EntryIv{lo, hi, v, 0.0};
            mx = max(mx, u);
            mc = max(mc, ~u);
        }
    }
    if (!a.vset) range_note(a, mx, mc);
}

// attachment of edge (p, q) to the apex k by the FP64 filter: 1 attached, 0 not, -1 open
GA_INLINE int attached_filter(double2 p, double2 q, double2 k)
{
    double dx = p.x - q.x, dy = p.y - q.y;
    double d1 = dx * dx + dy * dy;
    double sx = p.x + q.x - 2 * k.x, sy = p.y + q.y - 2 * k.y;
    double d2 = sx * sx + sy * sy;
    double diff = d1 - d2;
    double px = fabs(p.x + q.x) + 2 * fabs(k.x), py = fabs(p.y + q.y) + 2 * fabs(k.y);
    double eb = 32.0 * U53 * (d1 + px * px + py * py);
    if (diff > eb) return 1;
    if (-diff > eb) return 0;
    return -1;
}

// Edges decided by the filter; an edge it leaves open goes to xatt for k_attached_exact (which
// carries the expansion arithmetic and its scratch, so this kernel runs at full occupancy).
__global__ void k_attached(AlphaK a, int Emax)
{
    int E = a.eoff[a.F - 1] + a.cnt[a.F - 1];
    unsigned long long mx = 0, mc = 0;   // value-bit range of every edge (a superset of the entries)
    GS(e, Emax) {
        if (e >= E) a.flag[e] = 0;   // the scan of unattached edges runs to Emax
        else {
            int t = a.edge_t[e], i = a.edge_v[e];
            double2 p = cpt(a, t, pl1(i)), q = cpt(a, t, mi1(i));
            int nb = tn(a, t, i);
            int att = attached_filter(p, q, cpt(a, t, i));
            if (att == 0 && nb >= 0) att = attached_filter(p, q, cpt(a, nb >> 2, nb & 3));
            const bool open = att < 0;
            int slot = wave_reserve(&a.counters[A_ATT_LIST], open);
            if (open) {
                a.xatt[slot] = e;
                att = 0;   // set by k_attached_exact before the flags are scanned again
            }
            a.edge_att[e] = att;
            a.flag[e] = !att;
            double v = rho_edge(p, q);
            const unsigned long long u = value_bits(v);
            a.iv[a.F + e] = EntryIv{v * (1.0 - 8.0 * U53), v * (1.0 + 8.0 * U53), v, 0.0};
            mx = max(mx, u);
            mc = max(mc, ~u);
        }
    }
    if (!a.vset) range_note(a, mx, mc);
}

// the edges left open by k_attached: both sides with the exact predicate where the filter fails
__global__ void k_attached_exact(AlphaK a)
{
    const int nx = a.counters[A_ATT_LIST];
    int natt = 0;
    GS(q, nx) {
        int e = a.xatt[q];
        int t = a.edge_t[e], i = a.edge_v[e];
        double2 p = cpt(a, t, pl1(i)), pq = cpt(a, t, mi1(i));
        int nb = tn(a, t, i);
        int att = 0;
        for (int side = 0; side < 2 && !att; side++) {
            double2 k;
            if (side == 0) k = cpt(a, t, i);
            else {
                if (nb < 0) break;
                k = cpt(a, nb >> 2, nb & 3);
            }
            att = attached_filter(p, pq, k);
            if (att < 0) {
                natt++;
                int ovf = 0;
                att = attached_exact(p, pq, k, &ovf);
                if (ovf) atomicOr(a.status, GA_ERR_EXACT_OVERFLOW);
            }
        }
        a.edge_att[e] = att;
        a.flag[e] = !att;
    }
    stat_add(a.counters, ST_EXACT_ATT, natt);
}

// unattached edges appended after the triangles (flag scanned into rank[] positions).  Single
// set: the 32-bit sort keys of every entry go to k32 (the triangles' from their value bits in
// key[]); batches: 64-bit (set, value) keys in key[].
__global__ void k_edge_entries(AlphaK a, int E, unsigned *k32)
{
    Key32Map m{0, 0};
    if (!a.vset) {
        m = key32_map(a);
        GS(t, a.F) k32[t] = key32(m, a.key[t]);
    }
    GS(e, E) {
        if (a.edge_att[e]) continue;
        int k = a.F + a.rank[e];
        int t = a.edge_t[e], i = a.edge_v[e];
        double v = rho_edge(cpt(a, t, pl1(i)), cpt(a, t, mi1(i)));
        if (a.vset) a.key[k] = batch_key(a.vset[tv(a, t, 0)], v);
        else k32[k] = key32(m, value_bits(v));
        a.id[k] = a.F + e;
    }
}

// batches: sorted intervals, the set of every position (and reversed) and the negated lo values
// reversed, for the segmented scans of the cluster flags
__global__ void k_gather_iv(AlphaK a, int S0, double *neg_lo_rev, int *sset, int *sset_rev)
{
    GS(k, S0) {
        if (sset) {
            int s = (int)(a.key2[k] >> 40);
            sset[k] = s;
            sset_rev[S0 - 1 - k] = s;
        }
        int id = a.id2[k];
        const EntryIv w = a.iv[id];
        a.shi[k] = w.hi;
        a.slo[k] = w.lo;
        neg_lo_rev[S0 - 1 - k] = -w.lo;
    }
}

// cluster starts: boundary before k when max(hi[0..k-1]) < min(lo[k..]); with sets (scans
// segmented by set), also where the set changes
// ---- single set: cluster flags from tile bounds instead of full-length scans ----------------
// Tiles of CF_TILE sorted entries; a block of 256 threads holds CF_ITEMS consecutive entries per
// thread.  k_gather_iv_tiles writes the sorted intervals and each tile's max hi and min lo;
// k_cluster_flags_tiles sets flag[k] = max(hi[0..k-1]) < min(lo[k..S0-1]) from the bounds of the
// tiles before and after, a block scan of the thread bounds and the thread's own entries.
#define CF_ITEMS 8
#define CF_TILE (256 * CF_ITEMS)

__global__ __launch_bounds__(256) void k_gather_iv_tiles(AlphaK a, int S0, double *tmax, double *tmin)
{
    __shared__ double s_mx[4], s_mn[4];
    const int base = blockIdx.x * CF_TILE;
    double mx = -INFINITY, mn = INFINITY;
    for (int i = 0; i < CF_ITEMS; i++) {
        const int k = base + i * 256 + threadIdx.x;
        if (k >= S0) break;
        const EntryIv w = a.iv[a.id2[k]];
        a.shi[k] = w.hi;
        a.slo[k] = w.lo;
        a.key2[k] = value_bits(w.v);   // the value, for spec_value
        mx = fmax(mx, w.hi);
        mn = fmin(mn, w.lo);
    }
    for (int d = 32; d > 0; d >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, d, 64));
        mn = fmin(mn, __shfl_xor(mn, d, 64));
    }
    if ((threadIdx.x & 63) == 0) { s_mx[threadIdx.x >> 6] = mx; s_mn[threadIdx.x >> 6] = mn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) { mx = fmax(mx, s_mx[w]); mn = fmin(mn, s_mn[w]); }
        tmax[blockIdx.x] = mx;
        tmin[blockIdx.x] = mn;
    }
}

// tile bounds -> exclusive bounds: tmax[t] <- max hi of the tiles before t, tmin[t] <- min lo of
// the tiles after t (one block, chunks of 1024 tiles with a carry)
__global__ __launch_bounds__(1024) void k_tile_bounds_scan(int ntiles, double *tmax, double *tmin)
{
    __shared__ double s_w[16];
    __shared__ double s_carry;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // prefix max, forward
    if (threadIdx.x == 0) s_carry = -INFINITY;
    __syncthreads();
    for (int c0 = 0; c0 < ntiles; c0 += 1024) {
        const int j = c0 + threadIdx.x;
        const double x = j < ntiles ? tmax[j] : -INFINITY;
        double inc = x;
        for (int d = 1; d < 64; d <<= 1) {
            const double y = __shfl_up(inc, d, 64);
            if (lane >= d) inc = fmax(inc, y);
        }
        if (lane == 63) s_w[w] = inc;
        __syncthreads();
        double ex = __shfl_up(inc, 1, 64);
        if (lane == 0) ex = -INFINITY;
        for (int v = 0; v < w; v++) ex = fmax(ex, s_w[v]);
        ex = fmax(ex, s_carry);
        double tot = s_carry;
        for (int v = 0; v < 16; v++) tot = fmax(tot, s_w[v]);
        __syncthreads();
        if (j < ntiles) tmax[j] = ex;
        if (threadIdx.x == 0) s_carry = tot;
        __syncthreads();
    }
    // suffix min, backward
    if (threadIdx.x == 0) s_carry = INFINITY;
    __syncthreads();
    for (int c1 = ntiles; c1 > 0; c1 -= 1024) {
        const int j = c1 - 1 - threadIdx.x;   // thread 0 takes the chunk's last tile
        const double x = j >= 0 ? tmin[j] : INFINITY;
        double inc = x;
        for (int d = 1; d < 64; d <<= 1) {
            const double y = __shfl_up(inc, d, 64);
            if (lane >= d) inc = fmin(inc, y);
        }
        if (lane == 63) s_w[w] = inc;
        __syncthreads();
        double ex = __shfl_up(inc, 1, 64);
        if (lane == 0) ex = INFINITY;
        for (int v = 0; v < w; v++) ex = fmin(ex, s_w[v]);
        ex = fmin(ex, s_carry);
        double tot = s_carry;
        for (int v = 0; v < 16; v++) tot = fmin(tot, s_w[v]);
        __syncthreads();
        if (j >= 0) tmin[j] = ex;
        if (threadIdx.x == 0) s_carry = tot;
        __syncthreads();
    }
}

// The tile's hi and lo go through LDS (coalesced loads, then each thread's CF_ITEMS consecutive
// entries; one pad slot per CF_ITEMS keeps the per-thread reads off one bank), and the flags go
// back the same way.  P / Q: max hi before and min lo after the tile (k_tile_bounds_scan).
#define CF_PAD(i) ((i) + (i) / CF_ITEMS)
__global__ __launch_bounds__(256) void k_cluster_flags_tiles(AlphaK a, int S0, const double *texmax,
                                                             const double *texmin)
{
    __shared__ double s_hi[CF_PAD(CF_TILE)], s_lo[CF_PAD(CF_TILE)];
    __shared__ double s_wpre[4], s_wsuf[4];
    const int tile = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double P = texmax[tile], Q = texmin[tile];
    const int base = tile * CF_TILE;
#pragma unroll
    for (int j = 0; j < CF_ITEMS; j++) {
        const int idx = j * 256 + threadIdx.x, k = base + idx;
        const bool in = k < S0;
        s_hi[CF_PAD(idx)] = in ? a.shi[k] : -INFINITY;
        s_lo[CF_PAD(idx)] = in ? a.slo[k] : INFINITY;
    }
    __syncthreads();
    double hi[CF_ITEMS], lo[CF_ITEMS];
#pragma unroll
    for (int i = 0; i < CF_ITEMS; i++) {
        hi[i] = s_hi[CF_PAD(threadIdx.x * CF_ITEMS + i)];
        lo[i] = s_lo[CF_PAD(threadIdx.x * CF_ITEMS + i)];
    }
    // lo[i] <- min(lo[i..]) within the thread; thread bounds H (max hi), L (min lo)
    double H = -INFINITY;
#pragma unroll
    for (int i = 0; i < CF_ITEMS; i++) H = fmax(H, hi[i]);
#pragma unroll
    for (int i = CF_ITEMS - 2; i >= 0; i--) lo[i] = fmin(lo[i], lo[i + 1]);
    const double L = lo[0];
    // exclusive prefix max of H and exclusive suffix min of L over the threads of the block
    double pre = H, suf = L;
    for (int d = 1; d < 64; d <<= 1) {
        const double u = __shfl_up(pre, d, 64), v = __shfl_down(suf, d, 64);
        if (lane >= d) pre = fmax(pre, u);
        if (lane + d < 64) suf = fmin(suf, v);
    }
    if (lane == 63) s_wpre[w] = pre;   // wave totals
    if (lane == 0) s_wsuf[w] = suf;
    double ex_pre = __shfl_up(pre, 1, 64), ex_suf = __shfl_down(suf, 1, 64);
    if (lane == 0) ex_pre = -INFINITY;
    if (lane == 63) ex_suf = INFINITY;
    __syncthreads();   // also: every thread holds its entries, s_hi is free for the flags
    for (int v = 0; v < w; v++) ex_pre = fmax(ex_pre, s_wpre[v]);
    for (int v = w + 1; v < 4; v++) ex_suf = fmin(ex_suf, s_wsuf[v]);
    double before = fmax(P, ex_pre);         // max hi of every entry before this thread's first
    const double after_t = fmin(Q, ex_suf);  // min lo of every entry after this thread's last
    int *s_f = (int *)s_hi;
#pragma unroll
    for (int i = 0; i < CF_ITEMS; i++) {
        const int k = base + threadIdx.x * CF_ITEMS + i;
        s_f[CF_PAD(threadIdx.x * CF_ITEMS + i)] = k == 0 ? 1 : (before < fmin(after_t, lo[i]));
        before = fmax(before, hi[i]);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CF_ITEMS; j++) {
        const int idx = j * 256 + threadIdx.x, k = base + idx;
        if (k < S0) {
            const int f = s_f[CF_PAD(idx)];
            a.flag[k] = f;
            a.gflag[k] = f;
        }
    }
}

GA_INLINE int cluster_flag(int k, int S0, const double *pmax, const double *neg_smin_rev, const int *sset)
{
    if (k == 0 || (sset && sset[k] != sset[k - 1])) return 1;
    return pmax[k - 1] < -neg_smin_rev[S0 - 1 - k];
}

// cluster flags, into flag and gflag (the group flags start from them)
__global__ void k_cluster_flags(AlphaK a, int S0, const double *pmax, const double *neg_smin_rev, const int *sset)
{
    GS(k, S0) {
        const int f = cluster_flag(k, S0, pmax, neg_smin_rev, sset);
        a.flag[k] = f;
        a.gflag[k] = f;
    }
}

GA_INLINE void entry_pts(const AlphaK &a, int id, int *st, double2 *p)
{
    if (id < a.F) {
        *st = 1;
        p[0] = cpt(a, id, 0); p[1] = cpt(a, id, 1); p[2] = cpt(a, id, 2);
    } else {
        int e = id - a.F;
        int t = a.edge_t[e], i = a.edge_v[e];
        *st = 0;
        p[0] = cpt(a, t, pl1(i)); p[1] = cpt(a, t, mi1(i));
    }
}

// exact comparisons are counted in *ncmp (the caller adds its count with stat_add)
GA_INLINE int cmp_ids(const AlphaK &a, int ia, int ib, int *ncmp)
{
    const double la = a.iv[ia].lo, ha = a.iv[ia].hi, lb = a.iv[ib].lo, hb = a.iv[ib].hi;
    if (ha < lb) return -1;
    if (hb < la) return 1;
    (*ncmp)++;
    int sa, sb;
    double2 pa[3], pb[3];
    entry_pts(a, ia, &sa, pa);
    entry_pts(a, ib, &sb, pb);
    int ovf = 0;
    int r = rho_cmp_exact(sa, pa, sb, pb, &ovf);
    if (ovf) atomicOr(a.status, GA_ERR_EXACT_OVERFLOW);
    return r;
}

// equality of two entries whose 256-bit keys are equal (1 equal, 0 different, -1 unknown)
GA_INLINE int key_equal_values(const AlphaK &a, int ia, int ib, int *ncmp)
{
    (*ncmp)++;
    int sa, sb;
    double2 pa[3], pb[3];
    entry_pts(a, ia, &sa, pa);
    entry_pts(a, ib, &sb, pb);
    return rho_equal_res(sa, pa, sb, pb, 1);
}

// Clusters of up to MAX_SMALL_CLUSTER entries are put in order by one thread (exact insertion
// sort: up to m (m - 1) / 2 exact comparisons in sequence); larger ones by the exact-key path,
// which works on all their entries in parallel.  On the 2M lattice + cluster set a threshold of
// 64 left one thread with thousands of exact comparisons (36 ms); at 8 the worst is 28.
#define MAX_SMALL_CLUSTER 8

// ---- large clusters: exact 256-bit keys --------------------------------------------------------
// For every entry of a large cluster, rho = N / (4 D) * 2^(2s) exactly (big integers).  The key is
// the normalised truncated quotient q = floor(N 2^k / D) in [2^255, 2^256) with the exponent
// E = 2s - 2 - k: equal values give equal keys, and truncation is monotone, so different keys are
// ordered correctly.  Only key-equal neighbours need the exact comparator.
template <int NL>
GA_INLINE int bi_bitlen(const BigInt<NL> &b)
{
    if (b.n == 0) return 0;
    return 32 * (b.n - 1) + (32 - __clz(b.l[b.n - 1]));
}
// q = floor(u / v) for u of m limbs and v of n limbs (1 <= n <= m <= DIV_MAXU, v[n - 1] != 0); q gets
// m - n + 1 limbs.  Knuth's algorithm D with 32-bit digits: normalise v's top bit, estimate each
// quotient digit from the top two remainder digits, correct it at most twice, multiply-subtract,
// and add back on the rare overdraft.
#define DIV_MAXU 52
GA_INLINE void bi_divq(const uint32_t *u, int m, const uint32_t *v, int n, uint32_t *q)
{
    if (n == 1) {
        uint64_t r = 0;
        for (int j = m - 1; j >= 0; j--) {
            uint64_t cur = (r << 32) | u[j];
            q[j] = (uint32_t)(cur / v[0]);
            r = cur - (uint64_t)q[j] * v[0];
        }
        return;
    }
    const uint64_t b = 1ull << 32;
    uint32_t un[DIV_MAXU + 1], vn[ANL];
    const int s = __clz(v[n - 1]);
    for (int i = n - 1; i > 0; i--) vn[i] = (uint32_t)(((uint64_t)v[i] << s) | ((uint64_t)v[i - 1] >> (32 - s)));
    vn[0] = v[0] << s;
    un[m] = (uint32_t)((uint64_t)u[m - 1] >> (32 - s));
    for (int i = m - 1; i > 0; i--) un[i] = (uint32_t)(((uint64_t)u[i] << s) | ((uint64_t)u[i - 1] >> (32 - s)));
    un[0] = u[0] << s;
    for (int j = m - n; j >= 0; j--) {
        const uint64_t top = ((uint64_t)un[j + n] << 32) | un[j + n - 1];
        uint64_t qhat = top / vn[n - 1], rhat = top - qhat * vn[n - 1];
        while (qhat >= b || qhat * vn[n - 2] > ((rhat << 32) | un[j + n - 2])) {
            qhat--;
            rhat += vn[n - 1];
            if (rhat >= b) break;
        }
        int64_t t, kk = 0;
        for (int i = 0; i < n; i++) {
            const uint64_t pr = qhat * vn[i];
            t = (int64_t)un[i + j] - kk - (int64_t)(pr & 0xffffffffull);
            un[i + j] = (uint32_t)t;
            kk = (int64_t)(pr >> 32) - (t >> 32);
        }
        t = (int64_t)un[j + n] - kk;
        un[j + n] = (uint32_t)t;
        q[j] = (uint32_t)qhat;
        if (t < 0) {
            q[j]--;
            uint64_t c = 0;
            for (int i = 0; i < n; i++) {
                const uint64_t s2 = (uint64_t)un[i + j] + vn[i] + c;
                un[i + j] = (uint32_t)s2;
                c = s2 >> 32;
            }
            un[j + n] = (uint32_t)((uint64_t)un[j + n] + c);
        }
    }
}

__device__ __noinline__ void rho_key_exact(int st, const double2 *p, unsigned long long *q4, int *E, int *ovf)
{
    if (rho_key_fixed(st, p, q4, E)) return;   // differences within 58 bits: registers
    double v[6];
    int np = st ? 3 : 2;
    for (int i = 0; i < np; i++) { v[2 * i] = p[i].x; v[2 * i + 1] = p[i].y; }
    int s = min_ulp_exp(v, 2 * np);
    BI X[3], Y[3], num, den, t1, t2;
    for (int i = 0; i < np; i++) { bi_from_double(X[i], p[i].x, s, ovf); bi_from_double(Y[i], p[i].y, s, ovf); }
    if (!st) {
        bi_sq_dist(num, X[0], Y[0], X[1], Y[1], ovf);
        bi_zero(den);
        den.l[0] = 1; den.n = 1;
    } else {
        BI l1, l2, l3, ux, uy, vx, vy;
        bi_sq_dist(l1, X[0], Y[0], X[1], Y[1], ovf);
        bi_sq_dist(l2, X[1], Y[1], X[2], Y[2], ovf);
        bi_sq_dist(l3, X[2], Y[2], X[0], Y[0], ovf);
        bi_mul(t1, l1, l2, ovf);
        bi_mul(num, t1, l3, ovf);
        bi_sub(ux, X[1], X[0], ovf); bi_sub(uy, Y[1], Y[0], ovf);
        bi_sub(vx, X[2], X[0], ovf); bi_sub(vy, Y[2], Y[0], ovf);
        bi_mul(t1, ux, vy, ovf);
        bi_mul(t2, uy, vx, ovf);
        bi_sub(t1, t1, t2, ovf);
        bi_mul(den, t1, t1, ovf);
    }
    // long division of N * 2^k by D in 32-bit digits (Knuth's algorithm D); the quotient has 256
    // or 257 bits
    int Ln = bi_bitlen(num), Ld = bi_bitlen(den);
    int k = 256 - Ln + Ld;
    unsigned long long q[5] = {0, 0, 0, 0, 0};
    const int ls = k >> 5, bs = k & 31;
    int m = num.n + ls + 1;
    if (k < 0 || den.n == 0 || m > DIV_MAXU) {
        *ovf = 1;
        for (int w = 0; w < 4; w++) q4[w] = 0;
        *E = 0;
        return;
    }
    uint32_t u[DIV_MAXU], qd[DIV_MAXU];
    for (int i = 0; i < ls; i++) u[i] = 0;
    uint32_t carry = 0;
    for (int i = 0; i < num.n; i++) {
        uint64_t w = ((uint64_t)num.l[i] << bs) | carry;
        u[ls + i] = (uint32_t)w;
        carry = (uint32_t)(w >> 32);
    }
    u[ls + num.n] = carry;
    while (m > 1 && u[m - 1] == 0) m--;
    bi_divq(u, m, den.l, den.n, qd);
    const int qn = m - den.n + 1;
    for (int i = 0; i < qn; i++) {
        if (i < 10) q[i >> 1] |= (unsigned long long)qd[i] << (32 * (i & 1));
        else if (qd[i]) *ovf = 1;
    }
    // normalise to exactly 256 bits (top bit 255)
    if (q[4] & 1ull) {
        for (int w = 0; w < 4; w++) q[w] = (q[w] >> 1) | (q[w + 1] << 63);
        q[4] >>= 1;
        k -= 1;
    }
    for (int w = 0; w < 4; w++) q4[w] = q[w];
    *E = 2 * s - 2 - k;
}

struct BigK {
    int *cstart, *csize, *coff;     // [nbig]
    int *eid, *epos, *ej;           // [Mbig]
    unsigned long long *key[5];     // [Mbig] each: mantissa words 0..3 (low..high), top = cluster | exponent
    unsigned long long *kt, *kt2;
    int *vj, *vj2;
};

// one thread per entry of all large clusters (grid-stride); the cluster is found by binary
// search over the entry offsets coff[0..nbig) (coff[0] = 0, increasing)
__global__ void k_big_elems(AlphaK a, BigK b, int nbig, int Mb, int maxsize)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&a.counters[A_MAXCLUSTER], maxsize);
    GS(k, Mb) {
        int lo = 0, hi = nbig - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (b.coff[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        int c = lo;
        int j = k - b.coff[c];
        int gpos = b.cstart[c] + j;
        int id = a.id2[gpos];
        int st;
        double2 p[3];
        entry_pts(a, id, &st, p);
        int ovf = 0, E = 0;
        unsigned long long q[4];
        rho_key_exact(st, p, q, &E, &ovf);
        if (ovf) atomicOr(a.status, GA_ERR_EXACT_OVERFLOW);
        b.eid[k] = id;
        b.epos[k] = gpos;
        b.ej[k] = k;
        for (int w = 0; w < 4; w++) b.key[w][k] = q[w];
        b.key[4][k] = ((unsigned long long)c << 32) | (unsigned)(E + 0x40000000);
    }
}

__global__ void k_gather_keys(int M, const int *perm, const unsigned long long *src, unsigned long long *dst)
{
    GS(k, M) dst[k] = src[perm[k]];
}

__global__ void k_big_scatter(AlphaK a, BigK b, int Mbig, const int *sorted_j)
{
    GS(k, Mbig) a.id2[b.epos[k]] = b.eid[sorted_j[k]];
}

// neighbours inside a large cluster after the key sort: different keys are ordered and
// distinct; equal keys go to the exact comparator (equal values share a rank).  An exact
// inversion (values closer than 2^-255 relative) marks the pair for the fix-up.
__global__ void k_big_pairs(AlphaK a, BigK b, int Mbig, const int *sorted_j, int *closef, int *inv)
{
    int ncmp = 0;
    GS(k, Mbig) {
        int gpos = b.epos[k];
        int jc = sorted_j[k];
        int c = (int)(b.key[4][jc] >> 32);
        closef[gpos] = 0;
        inv[gpos] = 0;
        if (gpos == b.cstart[c]) continue;   // cluster start keeps its flag
        int jp = sorted_j[k - 1];
        bool same = true;
        for (int w = 0; w < 5; w++) same = same && b.key[w][jp] == b.key[w][jc];
        int f = 1;
        if (same) {
            closef[gpos] = 1;
            // equal keys: equality needs fewer primes (the values differ by less than 2^-255
            // relative); a different pair goes on to the full exact comparison
            int r = 0;
            if (key_equal_values(a, b.eid[jp], b.eid[jc], &ncmp) != 1) r = cmp_ids(a, b.eid[jp], b.eid[jc], &ncmp);
            if (r > 0) inv[gpos] = 1;
            f = r != 0;
        }
        a.gflag[gpos] = f;
    }
    stat_add(a.counters, ST_EXACT_CMP, ncmp);
}

// runs of key-equal neighbours containing an exact inversion: exact insertion sort (one thread)
__global__ void k_big_fix(AlphaK a, BigK b, int Mbig, const int *closef, const int *inv, int *claimed)
{
    int ncmp = 0;
    GS(k, Mbig) {
        int gpos = b.epos[k];
        if (!inv[gpos]) continue;
        int s = gpos - 1;
        while (closef[s]) s--;            // closef is 0 at every cluster start
        if (atomicCAS(&claimed[s], 0, 1) != 0) continue;
        int e = gpos;
        while (e + 1 < a.nS0 && closef[e + 1]) e++;
        int *ids = a.id2 + s;
        int m = e - s + 1;
        for (int i = 1; i < m; i++) {
            int x = ids[i], j = i - 1;
            while (j >= 0 && cmp_ids(a, ids[j], x, &ncmp) > 0) { ids[j + 1] = ids[j]; j--; }
            ids[j + 1] = x;
        }
        for (int i = 1; i < m; i++) a.gflag[s + i] = cmp_ids(a, ids[i - 1], ids[i], &ncmp) != 0;
    }
    stat_add(a.counters, ST_EXACT_CMP, ncmp);
}

// One thread per cluster start with size >= 2.  A small cluster whose certified intervals are
// pairwise disjoint is ordered by them (its values are distinct: every entry starts a group);
// one with overlapping intervals goes to xclu for k_clusters_pairs, which holds the expansion
// arithmetic (so this kernel needs no scratch).  The cluster count and largest size go to the
// counters once per wave.
__global__ void k_clusters(AlphaK a, int S0)
{
    int nclu = 0, mmax = 0;
    GS(k, S0) {
        if (!a.flag[k]) continue;
        if (k + 1 >= S0 || a.flag[k + 1]) { continue; }
        // the walk stops after MAX_SMALL_CLUSTER entries: larger clusters get their ends from
        // k_big_ends (one cluster can hold most of the spectrum, e.g. the equal circumradii of a
        // lattice's triangles)
        int end = k + 1;
        const int lim = min(S0, k + MAX_SMALL_CLUSTER);
        while (end < lim && !a.flag[end]) end++;
        nclu++;
        if (end < S0 && !a.flag[end]) {
            int q = atomicAdd(&a.counters[A_BIGCLUSTER], 1);
            if (q < a.bigcap) a.bigstart[q] = k;
            else atomicOr(a.status, GA_ERR_CAPACITY);
            continue;
        }
        const int m = end - k;
        mmax = max(mmax, m);
        int *ids = a.id2 + k;
        int idv[MAX_SMALL_CLUSTER];
        double lov[MAX_SMALL_CLUSTER], hiv[MAX_SMALL_CLUSTER];
#pragma unroll
        for (int i = 0; i < MAX_SMALL_CLUSTER; i++) {
            idv[i] = 0; lov[i] = 0.0; hiv[i] = 0.0;
            if (i < m) {
                idv[i] = ids[i];
                lov[i] = a.slo[k + i];
                hiv[i] = a.shi[k + i];
            }
        }
        bool open = false;
#pragma unroll
        for (int i = 0; i < MAX_SMALL_CLUSTER; i++)
#pragma unroll
            for (int j = i + 1; j < MAX_SMALL_CLUSTER; j++)
                if (j < m && !(hiv[i] < lov[j] || hiv[j] < lov[i])) open = true;
        if (open) {
            const int q = atomicAdd(&a.counters[A_EXCLU], 1);
            a.xclu[q] = k;
            a.xpair[q] = 0;   // k_clusters_pairs ORs its results in
            continue;
        }
#pragma unroll
        for (int i = 0; i < MAX_SMALL_CLUSTER; i++) {
            if (i >= m) continue;
            int r = 0;
#pragma unroll
            for (int j = 0; j < MAX_SMALL_CLUSTER; j++)
                if (j < m && lov[j] < lov[i]) r++;
            ids[r] = idv[i];
            if (i > 0) a.gflag[k + i] = 1;
        }
    }
    for (int d = 32; d > 0; d >>= 1) {
        mmax = max(mmax, __shfl_xor(mmax, d, 64));
        nclu += __shfl_xor(nclu, d, 64);
    }
    if (__lane_id() == 0) {
        int *row = a.counters + ST_OFF + wave_stripe() * STRIPE_INT;
        if (mmax) atomicMax(row + ST_MAXCLUSTER, mmax);
        if (nclu) atomicAdd(row + ST_CLUSTERS, nclu);
    }
}

// the statistics from their stripes into their counters (one wave, at the end of the run); every
// lane returns the folded (clusters, largest cluster, exact compares, exact attachments)
__device__ int4 fold_stats(const AlphaK &a)
{
    int c = 0, m = 0, x = 0, t = 0;
    for (int i = threadIdx.x; i < URANGE_STRIPES; i += 64) {
        const int *row = a.counters + ST_OFF + i * STRIPE_INT;
        c += row[ST_CLUSTERS];
        m = max(m, row[ST_MAXCLUSTER]);
        x += row[ST_EXACT_CMP];
        t += row[ST_EXACT_ATT];
    }
    for (int d = 32; d > 0; d >>= 1) {
        c += __shfl_xor(c, d, 64);
        m = max(m, __shfl_xor(m, d, 64));
        x += __shfl_xor(x, d, 64);
        t += __shfl_xor(t, d, 64);
    }
    m = max(m, a.counters[A_MAXCLUSTER]);   // also k_big_elems' atomicMax
    if (threadIdx.x == 0) {   // these counters change only through their stripes: folding twice is safe
        a.counters[A_CLUSTERS] = c;
        a.counters[A_MAXCLUSTER] = m;
        a.counters[A_EXACT_CMP] = x;
        a.counters[A_EXACT_ATT] = t;
    }
    return make_int4(c, m, x, t);
}

// The small clusters with overlapping intervals, in two launches.  k_clusters_pairs runs one
// exact comparison per thread: work item w is pair p = w % 28 (i < j, row by row: (0,1) .. (0,7),
// (1,2) ..) of cluster q = w / 28, with code 1 (equal), 2 (entry i smaller) or 3 (entry j smaller)
// in bits 2p of xpair[q].  k_clusters_rank then puts each cluster in order: entry i goes to
// position #{j : v_j < v_i} + #{j < i : v_j = v_i}, the order of a stable sort by value, and starts
// a group unless an earlier entry has its value.  One thread ran a whole cluster's insertion sort
// before (up to 35 exact comparisons in sequence): 2.9 ms on the 2M lattice + cluster set.
static_assert(MAX_SMALL_CLUSTER == 8, "28 pairs of 2 bits in one 64-bit word");
__device__ __forceinline__ int pair_index(int i, int j) { return i * (15 - i) / 2 + j - i - 1; }

// the entries of the cluster that starts at k (2 .. MAX_SMALL_CLUSTER)
__device__ __forceinline__ int small_cluster_size(const AlphaK &a, int k, int S0)
{
    int end = k + 1;
    while (end < S0 && end < k + MAX_SMALL_CLUSTER && !a.flag[end]) end++;
    return end - k;
}

__global__ void k_clusters_pairs(AlphaK a, int S0)
{
    const int nx = a.counters[A_EXCLU];
    int ncmp = 0;
    const long long nw = 28ll * nx;   // (above 2^31 for 77M clusters)
    for (long long w = blockIdx.x * (long long)blockDim.x + threadIdx.x; w < nw;
         w += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(w / 28);
        int r = (int)(w - 28ll * q), i = 0;
        while (r >= 7 - i) { r -= 7 - i; i++; }
        const int j = i + 1 + r;
        const int k = a.xclu[q];
        if (j >= small_cluster_size(a, k, S0)) continue;
        const int c = cmp_ids(a, a.id2[k + i], a.id2[k + j], &ncmp);
        atomicOr(&a.xpair[q], (unsigned long long)(c == 0 ? 1 : (c < 0 ? 2 : 3)) << (2 * pair_index(i, j)));
    }
    stat_add(a.counters, ST_EXACT_CMP, ncmp);
}

__global__ void k_clusters_rank(AlphaK a, int S0)
{
    const int nx = a.counters[A_EXCLU];
    GS(q, nx) {
        const int k = a.xclu[q];
        const int m = small_cluster_size(a, k, S0);
        const unsigned long long code = a.xpair[q];
        int *ids = a.id2 + k;
        int idv[MAX_SMALL_CLUSTER];
#pragma unroll
        for (int i = 0; i < MAX_SMALL_CLUSTER; i++) idv[i] = i < m ? ids[i] : 0;
#pragma unroll
        for (int i = 0; i < MAX_SMALL_CLUSTER; i++) {
            if (i >= m) continue;
            int pos = 0;
            bool eq_before = false;
#pragma unroll
            for (int j = 0; j < MAX_SMALL_CLUSTER; j++) {
                if (j >= m || j == i) continue;
                const int lo = i < j ? i : j, hi = i < j ? j : i;
                const int c = (int)((code >> (2 * pair_index(lo, hi))) & 3ull);
                // c: 1 equal, 2 entry lo smaller, 3 entry hi smaller
                const bool j_smaller = (c == 2 && lo == j) || (c == 3 && hi == j);
                if (j_smaller || (c == 1 && j < i)) pos++;
                if (c == 1 && j < i) eq_before = true;
            }
            ids[pos] = idv[i];
            if (pos > 0) a.gflag[k + pos] = eq_before ? 0 : 1;
        }
    }
}

// End (one past the last entry) of every large cluster, cstart[0..nbig) in increasing order: a
// cluster start j (or j = S0) ends a large cluster when no start lies in the MAX_SMALL_CLUSTER
// entries before it; that cluster is the last large one starting before j.  The backward walks
// stop at the first start, so they read each small cluster once.
__global__ void k_big_ends(const int *flag, int S0, const int *cstart, int nbig, int *cend)
{
    GS(j, S0 + 1) {
        if (j == 0 || (j < S0 && !flag[j])) continue;
        int i = j - 1;
        const int lim = j - MAX_SMALL_CLUSTER;
        while (i >= lim && i >= 0 && !flag[i]) i--;
        if (i >= lim) continue;   // the cluster before j has at most MAX_SMALL_CLUSTER entries
        int lo = 0, hi = nbig - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (cstart[mid] < j) lo = mid;
            else hi = mid - 1;
        }
        cend[lo] = j;
    }
}

// inclusive = exclusive + flag (in place over the flag array)
// incl[i] = excl[i] + flag[i] (incl may be flag)
__global__ void k_incl_from_excl(const int *excl, const int *flag, int *incl, int n)
{
    GS(i, n) incl[i] = excl[i] + flag[i];
}

GA_INLINE int entry_tri(const AlphaK &a, int id) { return id < a.F ? id : a.edge_t[id - a.F]; }
GA_INLINE int tri_set(const AlphaK &a, int t) { return a.vset ? a.vset[tv(a, t, 0)] : 0; }

GA_INLINE double entry_value(const AlphaK &a, int id)
{
    if (id < a.F) return rho_tri(cpt(a, id, 0), cpt(a, id, 1), cpt(a, id, 2));
    int e = id - a.F;
    int t = a.edge_t[e], i = a.edge_v[e];
    return rho_edge(cpt(a, t, pl1(i)), cpt(a, t, mi1(i)));
}

// single set: rank = incl[k] (1-based); the spectrum value of a rank is the smallest FP64 value
// of its members
// An entry alone in its certified cluster is never moved by the exact sorts, so key2[k] (single
// set) still holds its FP64 value bits, and it is alone in its rank (plain store).  Entries of larger clusters
// recompute their value and take the minimum: first over the lanes of the wave that hold the same
// slot (slots do not decrease from lane to lane), then one atomicMin per slot and wave (a
// lattice's triangles can put a million entries in one rank).  Called by every lane of the wave;
// `act` is false past the end.
GA_INLINE void spec_value(const AlphaK &a, int S0, int k, int id, int slot, bool act)
{
    const bool alone = act && a.flag[k] && (k + 1 == S0 || a.flag[k + 1]);
    if (alone && !a.vset) a.specbits[slot] = a.key2[k];
    else if (alone) a.specbits[slot] = (unsigned long long)__double_as_longlong(entry_value(a, id));
    const bool need = act && !alone;
    if (!__any(need)) return;
    const int lane = threadIdx.x & 63;
    unsigned long long v = need ? (unsigned long long)__double_as_longlong(entry_value(a, id)) : ~0ull;
    const int sl = need ? slot : -1;
    for (int d = 1; d < 64; d <<= 1) {
        unsigned long long ov = __shfl_up(v, d, 64);
        int os = __shfl_up(sl, d, 64);
        if (lane >= d && os == sl && ov < v) v = ov;
    }
    const int next = __shfl_down(sl, 1, 64);
    if (need && (lane == 63 || next != sl)) atomicMin(&a.specbits[slot], v);
}

__global__ void k_ranks(AlphaK a, int S0, const int *incl)
{
    if (big_pending(a)) return;
    for (int base = blockIdx.x * blockDim.x; base < S0; base += gridDim.x * blockDim.x) {
        const int k = base + threadIdx.x;
        const bool act = k < S0;
        int r = 0, id = 0;
        if (act) {
            r = incl[k];
            id = a.id2[k];
            if (id < a.F) a.tri_rank[id] = r;
            else a.edge_rank[id - a.F] = r;
        }
        spec_value(a, S0, k, id, r, act);
    }
}

// batches: entries are in (set, exact value) order, so the global group ranks count up through
// the sets; local ranks count the groups inside the set.
__global__ void k_iota(int n, int *a)
{
    GS(k, n) a[k] = k;
}

__global__ void k_batch_flags(int S0, const unsigned int *sset2, const int *order2, const int *incl, int *flags2)
{
    GS(j, S0) {
        int f = 1;
        if (j > 0) f = sset2[j] != sset2[j - 1] || incl[order2[j]] != incl[order2[j - 1]];
        flags2[j] = f;
    }
}

// first / last inclusive local count of every set (sets without entries keep -1)
__global__ void k_set_bounds(int S0, const unsigned int *sset2, const int *incl2, int *first, int *last)
{
    GS(j, S0) {
        int s = (int)sset2[j];
        if (j == 0 || sset2[j - 1] != sset2[j]) first[s] = incl2[j];
        if (j == S0 - 1 || sset2[j + 1] != sset2[j]) last[s] = incl2[j];
    }
}

__global__ void k_ranks_batch(AlphaK a, int S0, const unsigned int *sset2, const int *order2, const int *incl2,
                              const int *first)
{
    for (int base = blockIdx.x * blockDim.x; base < S0; base += gridDim.x * blockDim.x) {
        const int j = base + threadIdx.x;
        const bool act = j < S0;
        int k = 0, id = 0, slot = 0;
        if (act) {
            int s = (int)sset2[j];
            int r = incl2[j] - first[s] + 1;
            k = order2[j];
            id = a.id2[k];
            if (id < a.F) a.tri_rank[id] = r;
            else a.edge_rank[id - a.F] = r;
            slot = a.specstart[s] + r;
        }
        spec_value(a, S0, k, id, slot, act);
    }
}

__global__ void k_edge_intervals(AlphaK a, int E)
{
    if (big_pending(a)) return;
    GS(e, E) {
        int t = a.edge_t[e], i = a.edge_v[e];
        if (a.edge_att[e]) a.edge_rank[e] = 0;
        int r1 = a.tri_rank[t];
        int nb = tn(a, t, i);
        int mn, mx;
        if (nb < 0) { mn = r1; mx = 0; }
        else {
            int r2 = a.tri_rank[nb >> 2];
            mn = r1 <= r2 ? r1 : r2;
            mx = r1 <= r2 ? r2 : r1;
        }
        a.edge_numin[e] = mn;
        a.edge_numax[e] = mx;
        int key = a.edge_rank[e] != 0 ? a.edge_rank[e] : mn;
        int u = tv(a, t, pl1(i)), w = tv(a, t, mi1(i));
        // vnumin only decreases and vnumax only increases, so a plain read that already shows
        // a value at or beyond this edge's (stale reads are larger / smaller) makes the atomic
        // a no-op and it is skipped; both extremes of a vertex are one 8-byte read
        const int2 vu = a.vnum[u], vw = a.vnum[w];
        if (vu.x > key) atomicMin(&a.vnum[u].x, key);
        if (vw.x > key) atomicMin(&a.vnum[w].x, key);
        if (vu.y < mx) atomicMax(&a.vnum[u].y, mx);
        if (vw.y < mx) atomicMax(&a.vnum[w].y, mx);
    }
}

// master keys: q < 2n vertices (2v: vnumin, 2v + 1: vnumax), < 2n + 3E edges (2n + 3e + 0/1/2: rank,
// numin, numax), then triangles; entries without a rank get the key `none`, above every rank, so
// they sort to the end.  One loop per simplex kind, so no index is divided (the host keeps
// 2n + 3E + F below 2^31 - 2^20).
GA_INLINE unsigned master_key(const AlphaK &a, int r, int set, unsigned none)
{
    if (r < 1) return none;
    return (unsigned)((a.vset ? a.specstart[set] : 0) + r);
}

__global__ void k_master_keys(AlphaK a, int E, unsigned int *mkey, int *mval, unsigned int none)
{
    if (big_pending(a)) return;
    const int n2 = 2 * a.n, e0 = n2 + 3 * E;
    GS(v, a.n) {
        // the vertex's interval from its working pair (a vertex on no edge: -1, -1)
        int2 wp = a.vnum[v];
        if (wp.x == 0x7fffffff) wp = make_int2(-1, -1);
        a.vnumin[v] = wp.x;
        a.vnumax[v] = wp.y;
        const int set = a.vset ? a.vset[v] : 0;
        const int r0 = wp.x, r1 = wp.y;
        *(uint2 *)&mkey[2 * v] = make_uint2(master_key(a, r0, set, none), master_key(a, r1, set, none));
        *(int2 *)&mval[2 * v] = make_int2(2 * v, 2 * v + 1);
    }
    GS(e, E) {
        const int set = a.vset ? tri_set(a, a.edge_t[e]) : 0;
        const int r0 = a.edge_rank[e], r1 = a.edge_numin[e], r2 = a.edge_numax[e];
        const int q = n2 + 3 * e;
        mkey[q] = master_key(a, r0, set, none);
        mkey[q + 1] = master_key(a, r1, set, none);
        mkey[q + 2] = master_key(a, r2, set, none);
        mval[q] = q;
        mval[q + 1] = q + 1;
        mval[q + 2] = q + 2;
    }
    GS(t, a.F) {
        const int r = a.tri_rank[t];
        mkey[e0 + t] = master_key(a, r, tri_set(a, t), none);
        mval[e0 + t] = e0 + t;
    }
}

// number of sorted keys below `none` (the ranked entries), by binary search; every block
// computes it, block 0 stores it in *pMk
__device__ int ranked_count(int M, const unsigned int *skey, unsigned int none, int *pMk)
{
    __shared__ int s_mk;
    if (threadIdx.x == 0) {
        int lo = 0, hi = M;
        while (lo < hi) {
            const int mid = (int)(((unsigned)lo + (unsigned)hi) >> 1);
            if (skey[mid] < none) lo = mid + 1;
            else hi = mid;
        }
        s_mk = lo;
        if (blockIdx.x == 0) *pMk = lo;
    }
    __syncthreads();
    return s_mk;
}

// the ranked sorted entries (keys below `none`; their count goes to *pMk): simplex id, type
// (2 vertex, 0 edge, 1 triangle) and attribute of each
__global__ void k_master_decode(int n, int E, int M, unsigned int none, const unsigned int *skey, int *pMk,
                                const int *sval, int *simplex, int *st, int *at)
{
    const int Mk = ranked_count(M, skey, none, pMk);
    const unsigned n2 = 2u * (unsigned)n, e3 = 3u * (unsigned)E;
    GS(k, Mk) {
        const unsigned q = (unsigned)sval[k];
        int sp, ty, att;
        if (q < n2) { sp = (int)(q >> 1); ty = 2; att = (int)(q & 1) + 1; }
        else if (q < n2 + e3) { const unsigned w = q - n2, e = w / 3u; sp = (int)e; ty = 0; att = (int)(w - 3u * e); }
        else { sp = (int)(q - n2 - e3); ty = 1; att = 0; }
        simplex[k] = sp;
        st[k] = ty;
        at[k] = att;
    }
}

// sp[1 + i] = first master entry of the i-th distinct key.  Every rank r >= 1 of a set holds at
// least the entry of its own value, so the keys of set s are exactly specstart[s] + 1 ..
// specstart[s + 1] - 1 (single set: 1 .. R), and key K of set s is distinct key number
// K - s - 1: no scan over the entries is needed.  specstart[0] = 0.
__global__ void k_spec_pointers(AlphaK a, int M, unsigned int none, const unsigned int *skey, const int *specstart,
                                int nsets, int *sp, int cap)
{
    if (big_pending(a)) return;
    GS(k, M) {
        const unsigned K = skey[k];
        if (K >= none || (k > 0 && skey[k - 1] == K)) continue;
        int s = 0;
        if (nsets > 1) {
            // largest s with specstart[s] < K
            int lo = 0, hi = nsets - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if ((unsigned)specstart[mid] < K) lo = mid;
                else hi = mid - 1;
            }
            s = lo;
        }
        if ((int)K - s < cap) sp[(int)K - s] = k;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) sp[0] = -1;
}

// slots without a member (rank 0 of every set) read as 0; ~0 is a NaN pattern, never a value.
// nslots < 0: a single set whose rank count R is incl[0] (the inclusive rank of the last entry),
// nslots = R + 1
// also sets every vertex's working interval pair to (INT_MAX, -1) for k_edge_intervals
__global__ void k_spec_values(const unsigned long long *bits, int nslots, const int *incl_last, double *spec,
                              int2 *vnum, int n)
{
    if (nslots < 0) nslots = *incl_last + 1;
    GS(r, nslots) spec[r] = bits[r] == ~0ull ? 0.0 : __longlong_as_double((long long)bits[r]);
    GS(v, n) vnum[v] = make_int2(0x7fffffff, -1);
}

// first master entry of every set: keys are sorted and the keys of set s lie in
// (specstart[s], specstart[s + 1]), so it is the first position with key > specstart[s]
// (Mk when the set has no master entry); one binary search per set
__global__ void k_master_set_first(const int *pMk, const unsigned int *skey, const int *specstart, int nsets,
                                   int *mfirst)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsets) return;
    const int Mk = *pMk;
    unsigned int key = (unsigned)specstart[s];
    int lo = 0, hi = Mk;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (skey[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    mfirst[s] = lo;
}

// first edge of every set = edge offset of its first triangle (E when the set has no triangle)
__global__ void k_edge_set_first(const int *tfirst, int nsets, int F, const int *eoff, int E, int *efirst)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsets) return;
    int t = tfirst[s];
    efirst[s] = t < F ? eoff[t] : E;
}

// ---- host side ------------------------------------------------------------------------------------
enum {
    AB_CNT, AB_EOFF, AB_ET, AB_EV, AB_EATT, AB_ERANK, AB_ENMIN, AB_ENMAX, AB_TRANK, AB_KEY, AB_KEY2, AB_ID,
    AB_ID2, AB_LO, AB_HI, AB_SLO, AB_SHI, AB_FLAG, AB_RANK, AB_SPECBITS, AB_VMIN, AB_VMAX, AB_STATUS, AB_CNTRS,
    AB_TMP, AB_SCAN, AB_NEG, AB_PMAX, AB_NSMIN, AB_MKEY, AB_MVAL, AB_MFLAG, AB_MPOS, AB_OKEY, AB_OVAL,
    AB_SKEY, AB_SVAL, AB_SPEC, AB_SP, AB_MISC, AB_BIG0, AB_BIG1, AB_BIG2, AB_BIG3, AB_BIGL,
    AB_BS0, AB_BS1, AB_BS2, AB_BS3, AB_BS4, AB_BSET, AB_MSET, AB_VHULL, AB_RANKQ, AB_RD, AB_XATT, AB_XCLU, AB_XPAIR, AB_VNUM, AB_CXY, AB_COUNT
};
static_assert(AB_COUNT <= 64, "AlphaDev holds 64 buffers");

static void ab(AlphaDev &d, int i, size_t bytes)
{
    AlphaBuf &b = d.b[i];
    if (bytes <= b.bytes) return;
    if (galpha_arena_active()) throw GalphaError("alpha: no device allocation is allowed in this call");
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    size_t nb = bytes + bytes / 8 + 256;
    if (hipMalloc(&b.p, nb) != hipSuccess) throw GalphaError("alpha: device allocation failed");
    b.bytes = nb;
}
template <class T> static T *AP(AlphaDev &d, int i) { return (T *)d.b[i].p; }

void alpha_release(AlphaDev &d)
{
    for (auto &b : d.b)
        if (b.p) { hipFree(b.p); b.p = nullptr; b.bytes = 0; }
}

void alpha_destroy_events(AlphaDev &d)
{
    for (auto &e : d.ev)
        if (e) { hipEventDestroy(e); e = nullptr; }
    if (d.hpin) { hipHostFree(d.hpin); d.hpin = nullptr; }
}

static void chk(hipError_t e, const char *w)
{
    if (e != hipSuccess) throw GalphaError(std::string("alpha: ") + w + ": " + hipGetErrorString(e));
}

// A few device ints in one transfer: one wave gathers them into a device buffer and one copy
// brings them to pinned host memory (a copy into pageable memory waits for a round trip of its
// own, about 25 us each).
struct IntReads {
    const int *p[16];
    int n;
};
// into pinned host memory (the host waits on the stream before it reads); with fold, the
// statistics are folded from their stripes first and read as folded (one wave)
__device__ int4 fold_stats(const AlphaK &a);
__global__ void k_gather_ints(IntReads r, int *host, AlphaK a, int fold)
{
    const int4 f = fold ? fold_stats(a) : make_int4(0, 0, 0, 0);
    int i = threadIdx.x;
    if (i < r.n) {
        const int *p = r.p[i];
        int v = *p;
        if (fold && p == a.counters + A_CLUSTERS) v = f.x;
        if (fold && p == a.counters + A_MAXCLUSTER) v = f.y;
        if (fold && p == a.counters + A_EXACT_CMP) v = f.z;
        if (fold && p == a.counters + A_EXACT_ATT) v = f.w;
        host[i] = v;
    }
    __threadfence_system();
}

static void read_ints(AlphaDev &d, hipStream_t st, std::initializer_list<const int *> ptrs, int *out,
                      const AlphaK *fold = nullptr)
{
    IntReads r{};
    for (const int *q : ptrs) {
        if (r.n == 16) throw GalphaError("alpha: too many values in one read");
        r.p[r.n++] = q;
    }
    if (!d.hpin) chk(hipHostMalloc((void **)&d.hpin, 16 * sizeof(int), hipHostMallocDefault), "pinned alloc");
    GA_LAUNCH(k_gather_ints, dim3(1), dim3(64), 0, st, r, d.hpin, fold ? *fold : AlphaK{}, fold ? 1 : 0);
    chk(hipStreamSynchronize(st), "sync");
    memcpy(out, d.hpin, 4 * (size_t)r.n);
}

// the clusters above MAX_SMALL_CLUSTER entries (nbig of them, listed by k_clusters in bigstart)
static void big_clusters(AlphaDev &d, AlphaK &a, int S0, int nbig, hipStream_t st)
{
    if (nbig == 0) return;
    if (nbig > a.bigcap) throw GalphaError("alpha: too many large clusters");
    std::vector<int> cs2(nbig), ce(nbig), sz2(nbig), off(nbig + 1, 0);
    chk(hipMemcpyAsync(cs2.data(), a.bigstart, 4 * (size_t)nbig, hipMemcpyDeviceToHost, st), "copy");
    chk(hipStreamSynchronize(st), "sync");
    // order clusters by position so that cluster ids follow spectrum order; then their ends
    std::sort(cs2.begin(), cs2.end());
    ab(d, AB_BIG0, 4 * (size_t)nbig * 3 + 64);
    int *cstart = AP<int>(d, AB_BIG0), *csize = cstart + nbig, *coff = csize + nbig;
    chk(hipMemcpyAsync(cstart, cs2.data(), 4 * (size_t)nbig, hipMemcpyHostToDevice, st), "copy");
    GA_LAUNCH(k_big_ends, dim3(AGRID), dim3(256), 0, st, (const int *)a.flag, S0, (const int *)cstart, nbig, csize);
    chk(hipMemcpyAsync(ce.data(), csize, 4 * (size_t)nbig, hipMemcpyDeviceToHost, st), "copy");
    chk(hipStreamSynchronize(st), "sync");
    int maxsize = 0;
    for (int i = 0; i < nbig; i++) {
        sz2[i] = ce[i] - cs2[i];
        if (sz2[i] <= MAX_SMALL_CLUSTER || ce[i] > S0) throw GalphaError("alpha: large cluster bounds");
        off[i + 1] = off[i] + sz2[i];
        maxsize = std::max(maxsize, sz2[i]);
    }
    int Mb = off[nbig];
    chk(hipMemcpyAsync(csize, sz2.data(), 4 * (size_t)nbig, hipMemcpyHostToDevice, st), "copy");
    chk(hipMemcpyAsync(coff, off.data(), 4 * (size_t)nbig, hipMemcpyHostToDevice, st), "copy");
    ab(d, AB_BIG1, (size_t)Mb * (4 * 5 + 8 * 7) + 1024);
    char *p = (char *)d.b[AB_BIG1].p;
    BigK b{};
    b.cstart = cstart; b.csize = csize; b.coff = coff;
    auto take = [&](size_t bytes) { char *r = p; p += (bytes + 255) & ~(size_t)255; return (void *)r; };
    b.eid = (int *)take(4 * (size_t)Mb);
    b.epos = (int *)take(4 * (size_t)Mb);
    b.ej = (int *)take(4 * (size_t)Mb);
    b.vj = (int *)take(4 * (size_t)Mb);
    b.vj2 = (int *)take(4 * (size_t)Mb);
    for (int w = 0; w < 5; w++) b.key[w] = (unsigned long long *)take(8 * (size_t)Mb);
    b.kt = (unsigned long long *)take(8 * (size_t)Mb);
    b.kt2 = (unsigned long long *)take(8 * (size_t)Mb);
    GA_LAUNCH(k_big_elems, dim3((Mb + 255) / 256), dim3(256), 0, st, a, b, nbig, Mb, maxsize);
    // LSD radix passes: mantissa word 0 (lowest) .. 3, then cluster | exponent
    size_t tb = 0;
    chk(rocprim::radix_sort_pairs(nullptr, tb, b.kt, b.kt2, b.ej, b.vj, (unsigned)Mb, 0u, 64u, st), "sort");
    ab(d, AB_TMP, tb + 256);
    int *perm_in = b.ej, *perm_out = b.vj;
    // the top word (cluster << 32 | exponent + 2^30) has no bits above 32 + bitlen(nbig - 1)
    int cbits = 0;
    while ((1 << cbits) < nbig) cbits++;
    for (int w = 0; w < 5; w++) {
        GA_LAUNCH(k_gather_keys, dim3(AGRID), dim3(256), 0, st, Mb, (const int *)perm_in,
                  (const unsigned long long *)b.key[w], b.kt);
        size_t t2 = d.b[AB_TMP].bytes;
        const unsigned end_bit = w == 4 ? 32u + (unsigned)cbits : 64u;
        chk(rocprim::radix_sort_pairs(d.b[AB_TMP].p, t2, b.kt, b.kt2, perm_in, perm_out, (unsigned)Mb, 0u, end_bit,
                                      st),
            "sort");
        std::swap(perm_in, perm_out);
        if (w == 0) perm_out = b.vj2;
    }
    // perm_in: element indices in (cluster, value) order
    GA_LAUNCH(k_big_scatter, dim3(AGRID), dim3(256), 0, st, a, b, Mb, (const int *)perm_in);
    ab(d, AB_BIG3, 12 * (size_t)S0 + 64);
    int *closef = AP<int>(d, AB_BIG3), *inv = closef + S0, *claimed = inv + S0;
    chk(hipMemsetAsync(closef, 0, 12 * (size_t)S0, st), "memset");
    a.nS0 = S0;
    GA_LAUNCH(k_big_pairs, dim3(AGRID), dim3(256), 0, st, a, b, Mb, (const int *)perm_in, closef, inv);
    GA_LAUNCH(k_big_fix, dim3(AGRID), dim3(256), 0, st, a, b, Mb, (const int *)closef, (const int *)inv, claimed);
}

void alpha_run(AlphaDev &d, const double *d_xy, int n, const int *d_tris9, int F, hipStream_t st, galpha_stats *stats,
               bool check_mesh)
{
    AlphaBatchInfo one;
    one.nsets = 1;
    one.poff = {0, n};
    one.tfirst = {0, F};
    one.check_mesh = check_mesh;
    alpha_run_batch(d, d_xy, n, d_tris9, F, one, st, stats);
}

void alpha_run_batch(AlphaDev &d, const double *d_xy, int n, const int *d_tris9, int F, const AlphaBatchInfo &info,
                     hipStream_t st, galpha_stats *stats)
{
    d.n = n; d.F = F; d.E = 0; d.S = 0; d.specSize = 1; d.M = 0; d.nsp = 1;
    d.tris9 = d_tris9;
    const int nsets = info.nsets;
    d.nsets = nsets;
    d.poff = info.poff;
    d.tfirst = info.tfirst;
    d.efirst.assign(nsets + 1, 0);
    d.specstart.assign(nsets + 1, 0);
    d.mfirst.assign(nsets + 1, 0);
    for (int s = 0; s <= nsets; s++) d.specstart[s] = s;   // empty spectra until ranks exist
    ab(d, AB_STATUS, 16);
    ab(d, AB_CNTRS, 4 * (size_t)ALPHA_COUNTER_INTS);
    {
        IntFills f;
        f.add(d.b[AB_STATUS].p, 4, 0);
        f.add(d.b[AB_CNTRS].p, ALPHA_COUNTER_INTS, 0);
        launch_fill_multi(f, st);
    }
    ab(d, AB_VMIN, 4 * (size_t)(n + 1));
    ab(d, AB_VMAX, 4 * (size_t)(n + 1));
    ab(d, AB_SPEC, 8 * 2);
    if (F <= 0) {
        ab(d, AB_SPEC, 8 * (size_t)(nsets + 1));
        d.specSize = nsets;
        chk(hipMemsetAsync(d.b[AB_SPEC].p, 0, 8 * (size_t)nsets, st), "memset");
        chk(hipMemsetAsync(d.b[AB_VMIN].p, 0xff, 4 * (size_t)n + 4, st), "memset");
        chk(hipMemsetAsync(d.b[AB_VMAX].p, 0xff, 4 * (size_t)n + 4, st), "memset");
        ab(d, AB_SP, 8);
        int m1 = -1;
        chk(hipMemcpyAsync(d.b[AB_SP].p, &m1, 4, hipMemcpyHostToDevice, st), "copy");
        chk(hipStreamSynchronize(st), "sync");
        return;
    }
    for (int k = 0; k < 4; k++)
        if (!d.ev[k]) chk(hipEventCreate(&d.ev[k]), "event");
    chk(hipEventRecord(d.ev[0], st), "event");
    ab(d, AB_CNT, 4 * (size_t)F);
    ab(d, AB_EOFF, 4 * (size_t)F);
    ab(d, AB_SCAN, scan_temp_bytes((size_t)3 * F + 2 * (size_t)n + 16) + 4096);
    AlphaK a{};
    a.xy = (const double2 *)d_xy; a.tris = d_tris9; a.n = n; a.F = F;
    a.cnt = AP<int>(d, AB_CNT); a.eoff = AP<int>(d, AB_EOFF);
    a.status = AP<int>(d, AB_STATUS); a.counters = AP<int>(d, AB_CNTRS);
    if (info.check_mesh) GA_LAUNCH(k_mesh_check, dim3(AGRID), dim3(256), 0, st, a);
    GA_LAUNCH(k_edge_count, dim3(AGRID), dim3(256), 0, st, a);
    launch_exclusive_scan(a.cnt, a.eoff, F, d.b[AB_SCAN].p, st);
    if (info.check_mesh) {   // caller triangles: no kernel follows their links before the check
        int bad = 0;
        chk(hipMemcpyAsync(&bad, a.counters + A_BADMESH, 4, hipMemcpyDeviceToHost, st), "copy");
        chk(hipStreamSynchronize(st), "sync");
        if (bad)
            throw GalphaError("alpha: the triangles are not a mesh of the points (" + std::to_string(bad) +
                              " triangles with a vertex or neighbour out of range, or a neighbour link that does "
                              "not link back over the same edge)");
    }
    // Edge arrays are sized by a bound and the edge count E is read with the unattached count
    // below: a triangulation of n points with F triangles has E = n' + F - 1 <= n + F edges (n'
    // points in it); a caller mesh has at most 3F.
    const int Emax = (int)std::min<long long>(3LL * F, info.check_mesh ? 3LL * F : (long long)n + F);
    int E = Emax;
    size_t NE = (size_t)F + Emax;
    ab(d, AB_ET, 4 * (size_t)E); ab(d, AB_EV, 4 * (size_t)E); ab(d, AB_EATT, 4 * (size_t)E);
    ab(d, AB_ERANK, 4 * (size_t)E); ab(d, AB_ENMIN, 4 * (size_t)E); ab(d, AB_ENMAX, 4 * (size_t)E);
    ab(d, AB_TRANK, 4 * (size_t)F);
    ab(d, AB_KEY, 8 * NE); ab(d, AB_KEY2, 8 * NE); ab(d, AB_ID, 4 * NE); ab(d, AB_ID2, 4 * NE);
    ab(d, AB_LO, 32 * NE); ab(d, AB_SLO, 8 * NE); ab(d, AB_SHI, 8 * NE);
    ab(d, AB_FLAG, 4 * NE); ab(d, AB_RANK, 4 * NE); ab(d, AB_SPECBITS, 8 * (NE + (size_t)nsets + 2));
    ab(d, AB_NEG, 8 * NE); ab(d, AB_PMAX, 8 * NE); ab(d, AB_NSMIN, 8 * NE); ab(d, AB_MISC, 4 * NE);
    a.edge_t = AP<int>(d, AB_ET); a.edge_v = AP<int>(d, AB_EV); a.edge_att = AP<int>(d, AB_EATT);
    a.edge_rank = AP<int>(d, AB_ERANK); a.edge_numin = AP<int>(d, AB_ENMIN); a.edge_numax = AP<int>(d, AB_ENMAX);
    a.tri_rank = AP<int>(d, AB_TRANK);
    a.key = AP<unsigned long long>(d, AB_KEY); a.key2 = AP<unsigned long long>(d, AB_KEY2);
    a.id = AP<int>(d, AB_ID); a.id2 = AP<int>(d, AB_ID2);
    a.iv = AP<EntryIv>(d, AB_LO);
    a.slo = AP<double>(d, AB_SLO); a.shi = AP<double>(d, AB_SHI);
    a.flag = AP<int>(d, AB_FLAG); a.rank = AP<int>(d, AB_RANK); a.gflag = AP<int>(d, AB_MISC);

    a.specbits = AP<unsigned long long>(d, AB_SPECBITS);
    a.vset = nsets > 1 ? info.vset : nullptr;
    a.nsets = nsets;
    a.specstart = nullptr;
    a.vnumin = AP<int>(d, AB_VMIN); a.vnumax = AP<int>(d, AB_VMAX);
    ab(d, AB_VNUM, 8 * (size_t)n + 64);
    a.vnum = AP<int2>(d, AB_VNUM);
    GA_LAUNCH(k_edges, dim3(AGRID), dim3(256), 0, st, a);
    ab(d, AB_CXY, 48 * (size_t)F + 64);
    a.cxy = AP<double2>(d, AB_CXY);
    GA_LAUNCH(k_tri_rho, dim3(AGRID), dim3(256), 0, st, a);
    ab(d, AB_XATT, 4 * (size_t)Emax + 64);
    a.xatt = AP<int>(d, AB_XATT);
    GA_LAUNCH(k_attached, dim3(AGRID), dim3(256), 0, st, a, Emax);
    // positions of unattached edges (flags past E are zero)
    launch_exclusive_scan(a.flag, a.rank, Emax, d.b[AB_SCAN].p, st);
    int h3[5];
    read_ints(d, st, {a.rank + Emax - 1, a.flag + Emax - 1, a.eoff + F - 1, a.cnt + F - 1, a.counters + A_ATT_LIST},
              h3);
    if (h3[4] > 0) {
        // edges the filter left open: the exact predicate, then the positions again
        GA_LAUNCH(k_attached_exact, dim3(std::min(AGRID, (h3[4] + 255) / 256)), dim3(256), 0, st, a);
        launch_exclusive_scan(a.flag, a.rank, Emax, d.b[AB_SCAN].p, st);
        read_ints(d, st, {a.rank + Emax - 1, a.flag + Emax - 1, a.eoff + F - 1, a.cnt + F - 1}, h3);
    }
    int nun = h3[0] + h3[1];
    E = h3[2] + h3[3];
    if (E > Emax) throw GalphaError("alpha: more edges than the Euler bound (" + std::to_string(E) + ")");
    d.E = E;
    int S0 = F + nun;
    d.S = S0;
    // single set: 32-bit keys in and out side by side in key2 (free until the gather below)
    unsigned *k32 = (unsigned *)a.key2;
    GA_LAUNCH(k_edge_entries, dim3(AGRID), dim3(256), 0, st, a, E, k32);
    chk(hipEventRecord(d.ev[1], st), "event");
    // radix sort by value: single set, 32-bit monotone keys (4 passes; see key32); batches, the
    // set above the top 40 bits of the FP64 value.  The cluster boundaries below only need
    // "everything before is certainly smaller than everything after", which holds for any order,
    // so entries that tie on their key end in one cluster and are put in exact order there.
    size_t tb = 0;
    if (nsets == 1) {
        chk(rocprim::radix_sort_pairs<SortCfg32>(nullptr, tb, k32, k32 + S0, a.id, a.id2, (unsigned)S0, 0u, 32u, st),
            "sort");
        ab(d, AB_TMP, tb + 256);
        tb = d.b[AB_TMP].bytes;
        chk(rocprim::radix_sort_pairs<SortCfg32>(d.b[AB_TMP].p, tb, k32, k32 + S0, a.id, a.id2, (unsigned)S0, 0u,
                                                 32u, st),
            "sort");
    } else {
        unsigned sbits = 0;   // set bits above the 40 value bits
        while ((1LL << sbits) < (long long)nsets) sbits++;
        const unsigned kbits = 40u + sbits;
        chk(rocprim::radix_sort_pairs(nullptr, tb, a.key, a.key2, a.id, a.id2, (unsigned)S0, 0u, kbits, st), "sort");
        ab(d, AB_TMP, tb + 256);
        tb = d.b[AB_TMP].bytes;
        chk(rocprim::radix_sort_pairs(d.b[AB_TMP].p, tb, a.key, a.key2, a.id, a.id2, (unsigned)S0, 0u, kbits, st),
            "sort");
    }
    // certified intervals in sorted order, prefix max of hi, suffix min of lo (per set in batches)
    double *neg = AP<double>(d, AB_NEG), *pmax = AP<double>(d, AB_PMAX), *nsmin = AP<double>(d, AB_NSMIN);
    int *sset = nullptr, *sset_rev = nullptr;
    if (nsets > 1) {
        ab(d, AB_BS0, 4 * (size_t)S0 + 64);
        ab(d, AB_BS1, 4 * (size_t)S0 + 64);
        sset = AP<int>(d, AB_BS0);
        sset_rev = AP<int>(d, AB_BS1);
    }
    if (nsets == 1) {
        const int ntiles = (S0 + CF_TILE - 1) / CF_TILE;
        double *tmax = pmax, *tmin = nsmin;   // tile bounds (ntiles each)
        GA_LAUNCH(k_gather_iv_tiles, dim3(ntiles), dim3(256), 0, st, a, S0, tmax, tmin);
        GA_LAUNCH(k_tile_bounds_scan, dim3(1), dim3(1024), 0, st, ntiles, tmax, tmin);
        GA_LAUNCH(k_cluster_flags_tiles, dim3(ntiles), dim3(256), 0, st, a, S0, (const double *)tmax,
                  (const double *)tmin);
    } else {
        GA_LAUNCH(k_gather_iv, dim3(AGRID), dim3(256), 0, st, a, S0, neg, sset, sset_rev);
        tb = 0;
        chk(rocprim::inclusive_scan_by_key(nullptr, tb, sset, a.shi, pmax, (size_t)S0, rocprim::maximum<double>(),
                                           rocprim::equal_to<int>(), st), "scan");
        ab(d, AB_TMP, tb + 256);
        tb = d.b[AB_TMP].bytes;
        chk(rocprim::inclusive_scan_by_key(d.b[AB_TMP].p, tb, sset, a.shi, pmax, (size_t)S0,
                                           rocprim::maximum<double>(), rocprim::equal_to<int>(), st), "scan");
        tb = d.b[AB_TMP].bytes;
        chk(rocprim::inclusive_scan_by_key(d.b[AB_TMP].p, tb, sset_rev, neg, nsmin, (size_t)S0,
                                           rocprim::maximum<double>(), rocprim::equal_to<int>(), st), "scan");
        GA_LAUNCH(k_cluster_flags, dim3(AGRID), dim3(256), 0, st, a, S0, (const double *)pmax,
                  (const double *)nsmin, (const int *)sset);
    }
    // every large cluster has more than MAX_SMALL_CLUSTER entries, every cluster at least two
    a.bigcap = S0 / (MAX_SMALL_CLUSTER + 1) + 1;
    ab(d, AB_BIGL, 4 * (size_t)a.bigcap + 64);
    a.bigstart = AP<int>(d, AB_BIGL);
    ab(d, AB_XCLU, 4 * ((size_t)S0 / 2 + 1) + 64);
    a.xclu = AP<int>(d, AB_XCLU);
    ab(d, AB_XPAIR, 8 * ((size_t)S0 / 2 + 1) + 64);
    a.xpair = AP<unsigned long long>(d, AB_XPAIR);
    GA_LAUNCH(k_clusters, dim3(AGRID), dim3(256), 0, st, a, S0);
    // small clusters with overlapping intervals (the kernels read their count)
    GA_LAUNCH(k_clusters_pairs, dim3(AGRID), dim3(256), 0, st, a, S0);
    GA_LAUNCH(k_clusters_rank, dim3(256), dim3(256), 0, st, a, S0);
    // Large clusters: a batch reads their count here.  A single set goes on as if there were none
    // and reads the count with its final counters; when there are some, it runs their exact
    // order and everything after it again (the first pass's kernels below k_clusters return at
    // once, but its library sorts and scans run).
    if (nsets > 1) {
        int nb = 0;
        read_ints(d, st, {a.counters + A_BIGCLUSTER}, &nb);
        big_clusters(d, a, S0, nb, st);
    }
    bool big_done = nsets > 1;
    for (;;) {
    a.skip_if_big = big_done ? 0 : 1;
    // ranks: inclusive scan of group starts.  A single set reads its rank count R with the final
    // counters (buffers are sized by R <= S0); a batch reads it here.
    launch_exclusive_scan(a.gflag, a.rank, S0, d.b[AB_SCAN].p, st);
    int R = -1, nslots = -1;
    if (nsets == 1) {
        // spectrum slots (minimum by atomics)
        IntFills f;
        f.add(a.specbits, 2 * (size_t)(S0 + 2), -1);
        launch_fill_multi(f, st);
    } else {
        int h4[2];
        read_ints(d, st, {a.rank + S0 - 1, a.gflag + S0 - 1}, h4);
        R = h4[0] + h4[1];
        nslots = R + 1;
    }
    // incl = excl + flag: single set into the free NEG buffer (gflag stays for a second pass),
    // batches into the group-flag buffer
    int *incl = nsets == 1 ? AP<int>(d, AB_NEG) : a.gflag;
    GA_LAUNCH(k_incl_from_excl, dim3(AGRID), dim3(256), 0, st, (const int *)a.rank, (const int *)a.gflag, incl, S0);
    if (nsets == 1) {
        GA_LAUNCH(k_ranks, dim3(AGRID), dim3(256), 0, st, a, S0, (const int *)incl);
    } else {
        // entries are already in (set, exact value) order; sset[j] is the set of position j
        ab(d, AB_BS3, 4 * (size_t)S0 + 64);
        ab(d, AB_BS4, 4 * (size_t)S0 + 64);
        ab(d, AB_BSET, 4 * 3 * (size_t)(nsets + 1));
        const unsigned *sset2 = (const unsigned *)sset;
        int *pos = sset_rev;   // free: the reversed sets were only needed by the scan
        int *order2 = AP<int>(d, AB_BS3), *flags2 = AP<int>(d, AB_BS4);
        int *first = AP<int>(d, AB_BSET), *last = first + (nsets + 1), *sstart = last + (nsets + 1);
        GA_LAUNCH(k_iota, dim3(AGRID), dim3(256), 0, st, S0, order2);
        GA_LAUNCH(k_batch_flags, dim3(AGRID), dim3(256), 0, st, S0, (const unsigned *)sset2, (const int *)order2,
                  (const int *)incl, flags2);
        int *incl2 = pos;   // pos is free after the sort
        launch_exclusive_scan(flags2, incl2, S0, d.b[AB_SCAN].p, st);
        GA_LAUNCH(k_incl_from_excl, dim3(AGRID), dim3(256), 0, st, (const int *)incl2, (const int *)flags2, flags2,
                  S0);
        incl2 = flags2;
        launch_fill_i32(first, -1, 2 * (size_t)(nsets + 1), st);
        GA_LAUNCH(k_set_bounds, dim3(AGRID), dim3(256), 0, st, S0, (const unsigned *)sset2, (const int *)incl2,
                  first, last);
        std::vector<int> hf(nsets), hl(nsets);
        chk(hipMemcpyAsync(hf.data(), first, 4 * (size_t)nsets, hipMemcpyDeviceToHost, st), "copy");
        chk(hipMemcpyAsync(hl.data(), last, 4 * (size_t)nsets, hipMemcpyDeviceToHost, st), "copy");
        chk(hipStreamSynchronize(st), "sync");
        long long acc = 0;
        for (int q = 0; q < nsets; q++) {
            d.specstart[q] = (int)acc;
            int Rq = hf[q] >= 0 ? hl[q] - hf[q] + 1 : 0;
            acc += Rq + 1;
        }
        d.specstart[nsets] = (int)acc;
        if (acc > 0x7fffffffLL) throw GalphaError("alpha: batch spectrum too large");
        nslots = (int)acc;
        d.specSize = nslots;
        chk(hipMemcpyAsync(sstart, d.specstart.data(), 4 * (size_t)nsets, hipMemcpyHostToDevice, st), "copy");
        a.specstart = sstart;
        IntFills f;
        f.add(a.specbits, 2 * (size_t)nslots, -1);
        launch_fill_multi(f, st);
        GA_LAUNCH(k_ranks_batch, dim3(AGRID), dim3(256), 0, st, a, S0, (const unsigned *)sset2, (const int *)order2,
                  (const int *)incl2, (const int *)first);
    }
    const int slot_bound = nsets == 1 ? S0 + 1 : nslots;   // spectrum slots, or a bound of them
    ab(d, AB_SPEC, 8 * (size_t)(slot_bound + 1));
    GA_LAUNCH(k_spec_values, dim3(AGRID), dim3(256), 0, st, (const unsigned long long *)a.specbits, nslots,
              (const int *)(incl + S0 - 1), AP<double>(d, AB_SPEC), a.vnum, n);
    chk(hipEventRecord(d.ev[2], st), "event");
    // intervals (k_master_keys splits the vertex pairs into the output arrays)
    GA_LAUNCH(k_edge_intervals, dim3(AGRID), dim3(256), 0, st, a, E);
    // master list
    long long M = 2LL * n + 3LL * E + F;
    // (grid-stride loops over int indices: keep the last index plus the grid below 2^31)
    if (M > 0x7fffffffLL - (1LL << 20)) throw GalphaError("alpha: master list too large");
    ab(d, AB_MKEY, 4 * (size_t)M); ab(d, AB_MVAL, 4 * (size_t)M);
    unsigned *mkey = AP<unsigned>(d, AB_MKEY);
    int *mval = AP<int>(d, AB_MVAL);
    // every entry is sorted, the unranked ones with a key above every rank (no compaction and no
    // host read of the ranked count before the sort)
    int bits = 1;
    while ((1LL << bits) <= (long long)slot_bound) bits++;
    const unsigned none = (unsigned)((1LL << bits) - 1);
    GA_LAUNCH(k_master_keys, dim3(AGRID), dim3(256), 0, st, a, E, mkey, mval, none);
    ab(d, AB_OKEY, 4 * (size_t)M + 16); ab(d, AB_OVAL, 4 * (size_t)M + 16);
    ab(d, AB_SKEY, 4 * (size_t)M + 16); ab(d, AB_SVAL, 4 * (size_t)M + 16);
    unsigned *skey = AP<unsigned>(d, AB_SKEY);
    int *sval = AP<int>(d, AB_SVAL);
    tb = 0;
    chk(rocprim::radix_sort_pairs<SortCfg32>(nullptr, tb, mkey, skey, mval, sval, (unsigned)M, 0u, (unsigned)bits, st),
        "sort");
    ab(d, AB_TMP, tb + 256);
    tb = d.b[AB_TMP].bytes;
    chk(rocprim::radix_sort_pairs<SortCfg32>(d.b[AB_TMP].p, tb, mkey, skey, mval, sval, (unsigned)M, 0u,
                                             (unsigned)bits, st),
        "sort");
    // decode: simplex -> AB_OKEY, st -> AB_OVAL, at -> AB_MKEY (free after the sort)
    int *msimp = AP<int>(d, AB_OKEY), *mst = AP<int>(d, AB_OVAL), *mat = (int *)mkey;
    int *pMk = a.counters + A_MK;
    GA_LAUNCH(k_master_decode, dim3(AGRID), dim3(256), 0, st, n, E, (int)M, none, (const unsigned *)skey, pMk,
              (const int *)sval, msimp, mst, mat);
    // spec_pointers: one entry per distinct key, and every rank >= 1 of a set holds at least the
    // entry of its own value, so there are slot_bound - nsets of them (read with the final
    // counters for a single set)
    ab(d, AB_SP, 4 * (size_t)(slot_bound + 2));
    GA_LAUNCH(k_spec_pointers, dim3(AGRID), dim3(256), 0, st, a, (int)M, none, (const unsigned *)skey,
              (const int *)a.specstart, nsets, AP<int>(d, AB_SP), slot_bound + 2);
    int Mk = -1;   // read below
    if (nsets > 1) {
        // first master entry and first edge of every set
        ab(d, AB_MSET, 4 * 3 * (size_t)(nsets + 1));
        int *mf = AP<int>(d, AB_MSET), *tf = mf + (nsets + 1), *ef = tf + (nsets + 1);
        chk(hipMemcpyAsync(tf, d.tfirst.data(), 4 * (size_t)nsets, hipMemcpyHostToDevice, st), "copy");
        GA_LAUNCH(k_master_set_first, dim3((nsets + 255) / 256), dim3(256), 0, st, pMk, (const unsigned *)skey,
                  a.specstart, nsets, mf);
        GA_LAUNCH(k_edge_set_first, dim3((nsets + 255) / 256), dim3(256), 0, st, (const int *)tf, nsets, F,
                  (const int *)a.eoff, E, ef);
        std::vector<int> hm(nsets), he(nsets);
        chk(hipMemcpyAsync(hm.data(), mf, 4 * (size_t)nsets, hipMemcpyDeviceToHost, st), "copy");
        chk(hipMemcpyAsync(he.data(), ef, 4 * (size_t)nsets, hipMemcpyDeviceToHost, st), "copy");
        chk(hipMemcpyAsync(&Mk, pMk, 4, hipMemcpyDeviceToHost, st), "copy");
        chk(hipStreamSynchronize(st), "sync");
        d.mfirst[nsets] = Mk;
        for (int q = 0; q < nsets; q++) d.mfirst[q] = hm[q];
        for (int q = 0; q < nsets; q++) d.efirst[q] = he[q];
        d.efirst[nsets] = E;
    } else {
        d.efirst = {0, E};
    }
    chk(hipEventRecord(d.ev[3], st), "event");
    // counters, status, the single set's rank count and the ranked master entries in one read
    int fin[11];
    const int *zero = a.counters + A_BADMESH;   // 0 here: only caller meshes with errors set it
    read_ints(d, st,
              {a.counters + 0, a.counters + 1, a.counters + 2, a.counters + 3, a.counters + 4, a.counters + 5,
               a.counters + 6, a.counters + 7, a.status, nsets == 1 ? incl + S0 - 1 : zero, pMk},
              fin, &a);   // the statistics folded on the way
    if (!big_done && fin[A_BIGCLUSTER] > 0) {
        big_clusters(d, a, S0, fin[A_BIGCLUSTER], st);
        big_done = true;
        continue;
    }
    Mk = fin[10];
    d.M = Mk;
    if (nsets == 1) d.mfirst = {0, Mk};
    int *hc = fin, hs = fin[8];
    if (nsets == 1) {
        R = fin[9];
        nslots = R + 1;
        d.specSize = R + 1;
        d.specstart = {0, R + 1};
    }
    // one spec pointer per distinct key (every rank >= 1 of every set), after the leading -1
    d.nsp = (nsets == 1 ? R : nslots - nsets) + 1;
    if (stats) {
        stats->nEdges = E;
        stats->specSize = R + 1;
        stats->alphaExactCompares = hc[A_EXACT_CMP];
        stats->alphaExactAttach = hc[A_EXACT_ATT];
        stats->alphaClusters = hc[A_CLUSTERS];
        stats->alphaMaxCluster = hc[A_MAXCLUSTER];
        float ms[3] = {0, 0, 0};
        for (int k = 0; k < 3; k++) chk(hipEventElapsedTime(&ms[k], d.ev[k], d.ev[k + 1]), "event");
        stats->ms_alpha_rho = ms[0];
        stats->ms_alpha_order = ms[1];
        stats->ms_alpha_lists = ms[2];
    }
    if (hs) throw GalphaError("alpha: device status " + std::to_string(hs) +
                              (hc[A_BIGCLUSTER] ? " (exact cluster larger than supported)" : ""));
    return;
    }
}

PGPUDTALPHA alpha_copy_to_host(AlphaDev &d, HostStage &stage, hipStream_t st)
{
    PGPUDTALPHA A = (PGPUDTALPHA)calloc(1, sizeof(GPUDTALPHA));
    A->nPoints = d.n; A->nEdges = d.E; A->nTris = d.F; A->specSize = d.specSize;
    A->masterSize = d.M; A->nSpecPointers = d.nsp;
    auto cp = [&](void **dst, int idx, size_t bytes) {
        *dst = galpha_host_alloc(bytes);
        if (!*dst) throw GalphaError("host allocation failed");
        if (bytes) stage_d2h(stage, *dst, d.b[idx].p, bytes, st);
    };
    cp((void **)&A->spectrum, AB_SPEC, 8 * (size_t)d.specSize);
    cp((void **)&A->tri_ro_rank, AB_TRANK, 4 * (size_t)d.F);
    cp((void **)&A->edge_t, AB_ET, 4 * (size_t)d.E);
    cp((void **)&A->edge_vopp, AB_EV, 4 * (size_t)d.E);
    cp((void **)&A->edge_ro_rank, AB_ERANK, 4 * (size_t)d.E);
    cp((void **)&A->edge_nu_min_rank, AB_ENMIN, 4 * (size_t)d.E);
    cp((void **)&A->edge_nu_max_rank, AB_ENMAX, 4 * (size_t)d.E);
    cp((void **)&A->vert_nu_min_rank, AB_VMIN, 4 * (size_t)d.n);
    cp((void **)&A->vert_nu_max_rank, AB_VMAX, 4 * (size_t)d.n);
    cp((void **)&A->master_simplex, AB_OKEY, 4 * (size_t)d.M);
    cp((void **)&A->master_st, AB_OVAL, 4 * (size_t)d.M);
    cp((void **)&A->master_at, AB_MKEY, 4 * (size_t)d.M);
    cp((void **)&A->spec_pointers, AB_SP, 4 * (size_t)d.nsp);
    return A;
}

PGPUDTALPHA alpha_copy_set_to_host(AlphaDev &d, int set, HostStage &stage, hipStream_t st)
{
    if (set < 0 || set >= d.nsets) throw GalphaError("alpha: set index out of range");
    if (d.nsets == 1) return alpha_copy_to_host(d, stage, st);
    int p0 = d.poff[set], np = d.poff[set + 1] - p0;
    int t0 = d.tfirst[set], nt = d.tfirst[set + 1] - t0;
    int e0 = d.efirst[set], ne = d.efirst[set + 1] - e0;
    int s0 = d.specstart[set], ns = d.specstart[set + 1] - s0;
    int m0 = d.mfirst[set], nm = d.mfirst[set + 1] - m0;
    PGPUDTALPHA A = (PGPUDTALPHA)calloc(1, sizeof(GPUDTALPHA));
    if (!A) throw GalphaError("host allocation failed");
    A->nPoints = np; A->nEdges = ne; A->nTris = nt; A->specSize = ns;
    A->masterSize = nm; A->nSpecPointers = ns;
    auto cp = [&](void **dst, int idx, size_t elem, size_t first, size_t count) {
        *dst = galpha_host_alloc(count * elem);
        if (!*dst) throw GalphaError("host allocation failed");
        if (count) stage_d2h(stage, *dst, (char *)d.b[idx].p + first * elem, count * elem, st);
    };
    cp((void **)&A->spectrum, AB_SPEC, 8, s0, ns);
    cp((void **)&A->tri_ro_rank, AB_TRANK, 4, t0, nt);
    cp((void **)&A->edge_t, AB_ET, 4, e0, ne);
    cp((void **)&A->edge_vopp, AB_EV, 4, e0, ne);
    cp((void **)&A->edge_ro_rank, AB_ERANK, 4, e0, ne);
    cp((void **)&A->edge_nu_min_rank, AB_ENMIN, 4, e0, ne);
    cp((void **)&A->edge_nu_max_rank, AB_ENMAX, 4, e0, ne);
    cp((void **)&A->vert_nu_min_rank, AB_VMIN, 4, p0, np);
    cp((void **)&A->vert_nu_max_rank, AB_VMAX, 4, p0, np);
    cp((void **)&A->master_simplex, AB_OKEY, 4, m0, nm);
    cp((void **)&A->master_st, AB_OVAL, 4, m0, nm);
    cp((void **)&A->master_at, AB_MKEY, 4, m0, nm);
    // spec pointers of rank r >= 1 of this set: distinct key number specstart - set + r - 1
    int R = ns - 1;
    cp((void **)&A->spec_pointers, AB_SP, 4, R > 0 ? (size_t)(1 + s0 - set) - 1 : 0, R > 0 ? (size_t)R + 1 : 1);
    A->spec_pointers[0] = -1;
    for (int r = 1; r <= R; r++) A->spec_pointers[r] -= m0;
    for (int k = 0; k < ne; k++) A->edge_t[k] -= t0;
    for (int k = 0; k < nm; k++) {
        int stp = A->master_st[k];
        A->master_simplex[k] -= stp == 2 ? p0 : (stp == 0 ? e0 : t0);
    }
    return A;
}

PGPUDTALPHA alpha_empty_host(int np)
{
    PGPUDTALPHA A = (PGPUDTALPHA)calloc(1, sizeof(GPUDTALPHA));
    if (!A) throw GalphaError("host allocation failed");
    A->nPoints = np; A->specSize = 1; A->nSpecPointers = 1;
    A->spectrum = (double *)calloc(1, sizeof(double));
    A->spec_pointers = (int *)malloc(sizeof(int));
    A->spec_pointers[0] = -1;
    A->vert_nu_min_rank = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
    A->vert_nu_max_rank = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
    for (int i = 0; i < np; i++) { A->vert_nu_min_rank[i] = -1; A->vert_nu_max_rank[i] = -1; }
    A->tri_ro_rank = (int *)malloc(1); A->edge_t = (int *)malloc(1); A->edge_vopp = (int *)malloc(1);
    A->edge_ro_rank = (int *)malloc(1); A->edge_nu_min_rank = (int *)malloc(1); A->edge_nu_max_rank = (int *)malloc(1);
    A->master_simplex = (int *)malloc(1); A->master_st = (int *)malloc(1); A->master_at = (int *)malloc(1);
    return A;
}

void alpha_free_host(PGPUDTALPHA A)
{
    if (!A) return;
    free(A->spectrum); free(A->tri_ro_rank); free(A->edge_t); free(A->edge_vopp); free(A->edge_ro_rank);
    free(A->edge_nu_min_rank); free(A->edge_nu_max_rank); free(A->vert_nu_min_rank); free(A->vert_nu_max_rank);
    free(A->master_simplex); free(A->master_st); free(A->master_at); free(A->spec_pointers);
    free(A);
}

// ---- alpha-complex classification at a rank ----------------------------------------------------
// (no counterpart in the reference beyond setAlpha/getAlpha, Alpha_Shapes_2D.cu:739-747.)  Class
// codes: 0 exterior, 1 singular, 2 regular, 3 interior, from the interval ranks:
//   triangle  interior once r >= rho;
//   edge      enters at rho (unattached) or nu_min (attached); singular until nu_min; regular
//             until nu_max, for ever on the hull (nu_max = 0); then interior;
//   vertex    not in the triangulation (nu_min = -1): exterior; singular while r < nu_min;
//             regular until nu_max, for ever on the hull; then interior.
__global__ void k_classify_edges(const int *erank, const int *enmin, const int *enmax, const int *et, const int *ev,
                                 const int *tris9, int E, int r, unsigned char *ecls, unsigned char *vhull)
{
    GS(e, E) {
        int ro = erank[e], mn = enmin[e], mx = enmax[e];
        int enter = ro ? ro : mn;
        unsigned char c = r < enter ? 0 : ((ro && r < mn) ? 1 : ((mx == 0 || r < mx) ? 2 : 3));
        if (ecls) ecls[e] = c;
        if (mx == 0) {   // hull edge: both vertices lie on the hull
            int t = et[e], i = ev[e];
            vhull[tris9[(size_t)t * 9 + 3 + pl1(i)]] = 1;
            vhull[tris9[(size_t)t * 9 + 3 + mi1(i)]] = 1;
        }
    }
}

__global__ void k_classify_tris(const int *trank, int F, int r, unsigned char *tcls)
{
    GS(t, F) tcls[t] = r >= trank[t] ? 3 : 0;
}

__global__ void k_classify_verts(const int *vmin, const int *vmax, const unsigned char *vhull, int n, int r,
                                 unsigned char *vcls)
{
    GS(v, n) {
        int mn = vmin[v], mx = vmax[v];
        vcls[v] = mn < 0 ? 0 : (r < mn ? 1 : ((vhull[v] || r < mx) ? 2 : 3));
    }
}

// largest r with spectrum[r] <= alpha (-1 when there is none).  The FP64 spectrum need not be
// monotone (a rank's value is the smallest FP64 value of its exact group, and neighbouring exact
// values can round in the other order), so every rank is tested: one max per wave.
__global__ void k_alpha_rank(const double *spec, int nspec, double alpha, int *out)
{
    int best = -1;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < ((nspec + 63) & ~63); r += gridDim.x * blockDim.x)
        if (r < nspec && spec[r] <= alpha) best = r;
    for (int d = 32; d > 0; d >>= 1) best = max(best, __shfl_xor(best, d, 64));
    if ((threadIdx.x & 63) == 0 && best >= 0) atomicMax(out, best);
}

int alpha_rank_of(AlphaDev &d, double alpha, HostStage &stage, hipStream_t st)
{
    if (d.nsets != 1) throw GalphaError("alpha: rank lookup needs a single-set result");
    ab(d, AB_RANKQ, 16);
    launch_fill_i32(AP<int>(d, AB_RANKQ), -1, 1, st);
    int g = std::max(1, std::min(1024, (d.specSize + 255) / 256));
    GA_LAUNCH(k_alpha_rank, dim3(g), dim3(256), 0, st, (const double *)d.b[AB_SPEC].p, d.specSize, alpha,
              AP<int>(d, AB_RANKQ));
    int r = -1;
    stage_d2h(stage, &r, d.b[AB_RANKQ].p, 4, st);
    return r;
}

void alpha_classify(AlphaDev &d, int r, unsigned char *vcls, unsigned char *ecls, unsigned char *tcls,
                    hipStream_t st)
{
    if (d.nsets != 1) throw GalphaError("alpha: classification needs a single-set result");
    if (r < 0 || r >= d.specSize) throw GalphaError("alpha: rank out of range");
    if (!d.tris9 && d.E > 0) throw GalphaError("alpha: no triangulation");
    // cleared by a fill kernel, not a memset node, so a caller may capture this on a stream
    size_t hull_words = ((size_t)d.n + 16 + 3) / 4;
    ab(d, AB_VHULL, hull_words * 4);
    unsigned char *vhull = AP<unsigned char>(d, AB_VHULL);
    launch_fill_i32((int *)vhull, 0, hull_words, st);
    if (d.E > 0)
        GA_LAUNCH(k_classify_edges, dim3(AGRID), dim3(256), 0, st, AP<const int>(d, AB_ERANK), AP<const int>(d, AB_ENMIN),
                  AP<const int>(d, AB_ENMAX), AP<const int>(d, AB_ET), AP<const int>(d, AB_EV), d.tris9, d.E, r, ecls,
                  vhull);
    if (d.F > 0 && tcls) GA_LAUNCH(k_classify_tris, dim3(AGRID), dim3(256), 0, st, AP<const int>(d, AB_TRANK), d.F, r, tcls);
    if (d.n > 0 && vcls)
        GA_LAUNCH(k_classify_verts, dim3(AGRID), dim3(256), 0, st, AP<const int>(d, AB_VMIN), AP<const int>(d, AB_VMAX),
                  (const unsigned char *)vhull, d.n, r, vcls);
}
