// kernels.h -- launchers shared between the HIP translation units and the host orchestrator.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define GA_INLINE_HOST_DEVICE __host__ __device__ __forceinline__

void ga_probe_begin(const char *kernel, hipStream_t st);
void ga_probe_end(const char *kernel, hipStream_t st);
// kernel launch with the optional timing probe around it
#define GA_LAUNCH(K, G, B, SH, ST, ...)                      \
    do {                                                     \
        ga_probe_begin(#K, ST);                              \
        hipLaunchKernelGGL(K, G, B, SH, ST, __VA_ARGS__);    \
        ga_probe_end(#K, ST);                                \
    } while (0)

// One point set of a batch.  Global vertex ids: the real points of all sets first (set s at
// [poff, poff + n)), then the frame vertices of all sets (set s at [fbase, fbase + nf)), so that
// inside a set the index order of the single-set layout is kept (real points before frame
// vertices, each in their own order).  Textures of all sets are concatenated in one pixel array.
struct SetDesc {
    size_t toff;         // first pixel of the set's texture
    int logT;            // texture side T = 1 << logT
    int n, poff;         // real points and the global id of the first one
    int nf, fbase;       // frame vertices and the global id of the first one
    int roff;            // sum of the texture sides of the previous sets (row / column arrays)
    int m2, m3;          // PBA bands for phases 2 and 3
    int coff;            // first 16 x 16-pixel cell of the set (spatial numbering of the points)
    int margin;
    double shx, shy, scale, inv;   // snap grid: snapped = sh + (pixel - margin) * scale
};

// block ranges of one launch over all sets: set s owns blocks [start[s], start[s + 1])
GA_INLINE_HOST_DEVICE int set_of(const int *start, int nsets, int b)
{
    int lo = 0, hi = nsets - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// per-launch block maps (device arrays of nsets + 1 entries each)
enum { BM_COLS, BM_STACK, BM_MERGE, BM_COLOR, BM_PIX, BM_RECON, BM_COUNT };
struct BatchDev {
    const SetDesc *sets;   // [nsets]
    int nsets;
    const int *bm;         // [BM_COUNT][nsets + 1]
    int blocks[BM_COUNT];  // total blocks per map (host copy of bm[k][nsets])
    const int *vset;       // [V] set of every vertex
    int ntot, V;           // real points, all vertices
    int max_m2;            // largest phase-2 band count of the sets
    size_t P;              // pixels of all textures
    GA_INLINE_HOST_DEVICE const int *bmap(int k) const { return bm + (size_t)k * (nsets + 1); }
};

struct PbaBuffers {
    int *smap;     // T*T site id per pixel (0x7fffffff = empty)
    short *g;      // T*T nearest-site row per column
    int *S;        // T*T band stacks
    int *cnt;      // 64*T per set: end (hi) of each band's envelope range
    int *lo;       // 64*T per set: start (lo) of each band's envelope range
    int *E;        // T*T compacted envelopes
    int *ecnt;     // T envelope sizes
    int *vor;      // T*T nearest site id
};

// Mesh state of one point set on the device.
struct MeshDev {
    int4 *tv;            // triangle vertices (x == -2: dead slot)
    int4 *tn;            // neighbour encodings (t << 2 | j), -1 hull
    int4 *remap;         // per slot: new encoding of each old edge after a modification
    int *stamp;          // per slot: id of the last operation that modified it
    unsigned long long *claim;  // per slot: vote word
    int cap;             // slot capacity
    int *ntri;           // device: slots in use
    double2 *pos;        // current vertex positions [V]
    const double2 *xy;   // true positions of the real points [n]
    int *vt;             // vertex -> an incident triangle [V]
    int *vstate;         // vertex state [V]
    int *pix;            // vertex pixel (px << 16 | py) [V]
    int n, V;            // real points of all sets, all vertices (real + frame)
    const SetDesc *sets; // per set (snap grid, frame range, texture)
    const int *vset;     // [V] set of every vertex
    int nsets;
    int *status;         // device error bits
    int *counters;       // device scratch counters [64]; [64, 128) folded statistics
    int *stats;          // striped statistics: counters + GA_STRIPE_INTS, GA_STRIPES stripes
    int *plist;          // vertices still pending after bulk shifting (counters[C_PLIST] entries)
    const int *perm;     // input index of real vertex v (real vertices are numbered in spatial order)
    int *modA, *modB;    // modified-triangle lists (capacity cap)
    int *cand;           // candidate list (capacity 3*cap)
    int *scratch;        // per-winner scratch arena
    size_t scratch_cap;
    double *dbg;         // failure dump: positions (K + up to 15 polygon vertices) and states
    // per-context work buffers (never process globals: contexts run concurrently)
    int *deg;            // [V + 1] vertex degrees of the reconstruction
    int *decide;         // [V] greedy shifting decisions
    int *dellist;        // [cap] triangles created by a deletion round
    int *lt;             // [n] located type of every missing site
    int *hnext, *hedge;  // [V] hull successor and hull edge of every hull vertex
    int *lmark;          // [cap] operation that last put the slot in a flip list (one entry per pass)
    int *ffan;           // [2 * (V - n)] per frame vertex: first fan triangle (enc) and fan length
    int list_op;         // lmark tag of the list that frame removal and convexification fill
    // constraint edges (set of (min << 32 | max) + 1 keys, open addressing, 0 = empty); NULL when
    // the run has no constraints.  k_flip_test never flips an edge in this set.
    const unsigned long long *chash;
    int chmask;
};

// vertex states
#define VS_IN_MESH 1       // in the mesh, position final
#define VS_PENDING 2       // in the mesh at its snapped position
#define VS_MISSING 3       // to be inserted
#define VS_DUP 4           // exact duplicate of another point, dropped
#define VS_REMOVED 5       // frame vertex removed

// counters[] slots
#define C_FLAG 0
#define C_CAND 1
#define C_MODN 2
#define C_MODN2 3
#define C_WIN 4
#define C_SCR 5
#define C_MISS 6
#define C_PEND 7
#define C_EXACT_ORIENT 8
#define C_EXACT_INCIRCLE 9
#define C_FLIPS 10
#define C_INSERTED 11
#define C_DUPS 12
#define C_FRAME 13
#define C_TESTS_INCIRCLE 14
#define C_TESTS_ORIENT 15
#define C_RECON_TRIS 16
#define C_SHIFT_BAD 17
#define C_DELETED 18
#define C_ISLAND_SITES 19
#define C_HSTART 20
#define C_FLIPS_TOTAL 21
#define C_FRAME_SEEN 22       // frame vertices examined by k_frame_remove (all launches)
#define C_FRAME_CHAIN 23      // sum of the fan lengths of removed frame vertices
#define C_FLIP_LISTED 24      // triangles listed for k_flip_test (all launches)
#define C_FRAME_LIST 25       // triangles created by frame removal and convexification
#define C_BND 26              // boundary vertices of the real region (bulk frame removal)
#define C_EAR 27              // ear candidates of the current round
#define C_MOVED 28            // vertices moved by greedy shifting rounds (statistic)
#define C_FRAME_DONE 29       // frame vertices removed round by round (statistic)
#define C_PLIST 30            // entries of the pending-vertex list (MeshDev::plist)
#define C_EAR_PREV 31         // ear candidates of the previous round
#define C_STATUS 32           // device status word (GA_ERR_* bits); MeshDev::status points here
#define C_CONS_PEND 33        // constraints still to insert (the current insertion round's walk)
#define C_CONS_MOD 34         // slots rewritten by the current insertion round
#define C_TAIL_PASSES 35      // passes run by the last single-workgroup flip tail
#define C_TAIL_DONE 36        // 1 when that tail ended with a pass without flips
#define C_TAIL_LISTED 37      // triangles listed by flip tails (also counted in C_FLIP_LISTED)
#define C_MODN3 38            // third list count of the bulk shifting passes (with C_MODN, C_MODN2)
#define TAIL_LIST 2048        // longest flip list the single-workgroup tail takes
// status bits of constraint insertion (the other GA_ERR_* bits are in dev_common.h)
#define GA_ERR_CONS_POINT 128     // a point lies inside a constraint segment
#define GA_ERR_CONS_CROSS 256     // two constraints cross
#define GA_ERR_CONS_WALK 512      // the walk along a constraint left the mesh
#define GA_ERR_CONS_DUP 1024      // a constraint endpoint is a dropped duplicate point
#define GA_ERR_CONS_TRI 2048      // a retriangulated cavity produced a triangle that is not ccw
#ifndef GA_STRIPES
#define GA_STRIPES 16         // statistic stripes, GA_STRIPE_INTS ints apart (also in dev_common.h)
#define GA_STRIPE_INTS 1024
#endif
#define C_FOLD 64             // counters[C_FOLD + k] = stripe sum of statistic k after a fold
// Counters C_PEND, C_EXACT_*, C_INSERTED, C_DUPS, C_TESTS_INCIRCLE, C_RECON_TRIS, C_DELETED,
// C_FLIPS_TOTAL, C_FRAME_SEEN, C_FRAME_CHAIN, C_MOVED and C_FRAME_DONE are striped statistics
// (MeshDev::stats); the others are single words.
#define C_DBG 48          // 48..63: first failure dump
#define C_NCOUNTERS 64

// pba.hip
void launch_fill_i32(int *a, int v, size_t n, hipStream_t st);
void launch_vset(const BatchDev &b, int *vset, hipStream_t st);
// perm: input index of each real point (NULL: points in input order)
void launch_splat(const double2 *xy, const BatchDev &b, const int *perm, int *smap, int *pix, const int *fpix,
                  hipStream_t st);
// spatial numbering of the real points: perm[new] = input index, xys[new] = xy[perm[new]];
// ccount (ncells + 1 ints) zero on entry
void launch_renumber(const double2 *xy, const BatchDev &b, int *cell, int *ccount, int *cursor, int ncells,
                     void *scan_tmp, int *perm, double2 *xys, hipStream_t st);
void launch_pba(const PbaBuffers &p, const BatchDev &b, hipStream_t st);
void launch_islands(const int *vor, const int *pix, const BatchDev &b, int *bad, int *flag, hipStream_t st);
void launch_remove_bad(int *bad, const int *pix, const BatchDev &b, int *smap, int *removed, hipStream_t st);

// util.hip: up to four segments of device ints copied into one contiguous device array
struct IntSegments {
    const int *src[4];
    int len[4];
    int n;
};
void launch_gather_segments(const IntSegments &g, int *dst, hipStream_t st);
// up to eight device int arrays, each set to one value, in one launch
struct IntFills {
    int *p[8];
    size_t len[8];
    int v[8];
    int n = 0;
    void add(void *ptr, size_t ints, int value)
    {
        p[n] = (int *)ptr;
        len[n] = ints;
        v[n] = value;
        n++;
    }
};
void launch_fill_multi(const IntFills &f, hipStream_t st);

// scan.hip
size_t scan_temp_bytes(size_t n);
void launch_exclusive_scan(const int *in, int *out, size_t n, void *temp, hipStream_t st);

// mesh.hip
void launch_recon(const int *vor, const int *pix, const BatchDev &b, MeshDev &m, int *tile_cnt, int *tile_off,
                  void *scan_tmp, hipStream_t st);
void launch_adjacency(MeshDev &m, int *deg, int *off, int *fill, int2 *fan, void *scan_tmp, hipStream_t st);
void launch_init_positions(MeshDev &m, hipStream_t st);
// bulk shifting: one pass over all triangles (reverted vertices listed in wl_out, mark[] zero on
// entry), then passes over the stars of the vertices the previous pass listed
void launch_shift_bulk(MeshDev &m, unsigned char *moved, int *mark, int *wl_out, int cnt_out_slot,
                       int cnt_next_slot, hipStream_t st);
void launch_shift_star(MeshDev &m, unsigned char *moved, int *mark, const int *wl_in, int cnt_in_slot, int *wl_out,
                       int cnt_out_slot, int cnt_next_slot, hipStream_t st);
// lists the vertices still pending into m.plist (count in counters[C_PLIST])
void launch_pending_list(MeshDev &m, hipStream_t st);
void launch_shift_commit(MeshDev &m, const unsigned char *moved, hipStream_t st);
void launch_shift_greedy(MeshDev &m, int round, int grid, hipStream_t st);
void launch_delete_round(MeshDev &m, unsigned int *mx, int round, int op, int grid, hipStream_t st);
void launch_missing_init(MeshDev &m, const int *vor, int *mlist, int *loc, hipStream_t st);
void launch_insert_round(MeshDev &m, int *mlist, int *loc, int op, int grid, hipStream_t st);
void launch_relink(MeshDev &m, const int *modlist, int cnt_slot, int op, int grid, hipStream_t st);
void launch_reset_counters(int *counters, int a, int b, int d, int e, hipStream_t st);
// counters[C_FOLD + k] = sum over the stripes of statistic k, for k < 64
void launch_fold_stats(int *counters, hipStream_t st);
// counters[slot, slot + n) (n <= 256) into pinned host memory, statistics folded when fold is set
void launch_counters_to_host(int *counters, int slot, int n, int fold, int *host, hipStream_t st);
void launch_flip_all(MeshDev &m, int op, int *list, hipStream_t st);
// flip passes on short lists in one workgroup (see k_flip_tail)
void launch_flip_tail(MeshDev &m, int *lA, int sA, int *lB, int sB, int prev_op, int op0, int max_passes,
                      hipStream_t st);
void launch_flip_pass(MeshDev &m, const int *list_in, int cnt_in_slot, int *list_out, int cnt_out_slot,
                      int prev_op, int op, int grid, hipStream_t st);
// Blocks of 256 threads for a round over about `items` list entries: enough for one entry per
// thread, a multiple of 8 (one share per XCD), at most 2048 (grid-stride loops cover any rest).
// Lists that shrink from round to round then do not launch thousands of idle blocks.
inline int round_grid(long long items)
{
    long long b = (items + 255) / 256;
    b = (b + 7) & ~7LL;
    return (int)(b < 8 ? 8 : (b > 2048 ? 2048 : b));
}
// constrained DT (see mesh.hip).  Set-up: cv = constraints as vertex ids, done flags, the key set
// (chash, size chmask + 1, cleared by the caller), vt rebuilt from the live slots.
void launch_cons_setup(MeshDev &m, const int2 *cin, int nc, int *inv, int2 *cv, int *done,
                       unsigned long long *chash, int chmask, hipStream_t st);
// one insertion round: walk (counts, region sizes) -> scan -> collect + claim -> commit -> relink
void launch_cons_count(MeshDev &m, const int2 *cv, int nc, int *done, int *cnt, int *rsize, hipStream_t st);
void launch_cons_insert(MeshDev &m, const int2 *cv, int nc, int *done, const int *cnt, const int *roff,
                        int *work, int *modlist, int op, hipStream_t st);
void launch_frame_round(MeshDev &m, unsigned int *mx, int round, int *list_out, int cnt_out_slot, int op,
                        hipStream_t st);
// bulk frame removal (see mesh.hip); bnd = 11 * B ints (rounded up to 4) + B double2 cycle positions of
// work space once B is known
void launch_bnd_detect(MeshDev &m, int *outdeg, int *cover, int *hull_cnt, int *hull_start, int *badset,
                       int *blist, int *bidx, hipStream_t st);
void launch_bnd_rank(MeshDev &m, int B, int rounds, const int *blist, const int *bidx, const int *hull_start,
                     const int *hull_cnt, const int *hull_off, int *badset, int *bnd, hipStream_t st);
void launch_kill_frame(MeshDev &m, hipStream_t st);
// bulk frame removal's ear rounds: the per-set grids, then rounds [r0, r0 + k) as grid-wide kernels
// (counters[C_EAR] / [C_EAR_PREV] count the candidates of even / odd rounds)
void launch_ear_grid(MeshDev &m, int B, int *bnd, const int *hull_cnt, const int *hull_off, int nsets, hipStream_t st);
void launch_ear_rounds(MeshDev &m, int B, int *bnd, const int *hull_cnt, const int *hull_off, int nsets, int r0,
                       int k, int *list_out, int cnt_out_slot, int op, hipStream_t st);
// every set's ear rounds in one workgroup per set (bulk frame removal); counters[C_EAR] = most rounds
void launch_ear_sets(MeshDev &m, int B, int *bnd, const int *hull_cnt, const int *hull_off, int nsets,
                     int *list_out, int cnt_out_slot, int op, hipStream_t st);
void launch_convexify(MeshDev &m, int *list_out, int cnt_out_slot, int op, int *hull_cnt, int *hull_off,
                      int *hull_start, void *scan_tmp, hipStream_t st);
// canonical order: key (v0 << B) | v1 of input indices below 2^B, sorted on its low 2B bits
void launch_canon_keys(MeshDev &m, unsigned long long *keys, int *vals, int *rot, int B, hipStream_t st);
// canonical keys of 32 bits up to input indices below 2^CANON_KEY32_BITS (ties within 256 triangles)
#define CANON_KEY32_BITS 20
void launch_canon_keys32(MeshDev &m, unsigned *keys, unsigned *v1s, int *vals, int *rot, int B, hipStream_t st);
void launch_canon_emit(MeshDev &m, int *order, const unsigned *skey, const unsigned *v1s, const int *rot, int *newid,
                       const int *dF, int *out9,
                       int *set_first, int *rset, hipStream_t st);
int sort_pairs_u64(unsigned long long *keys_in, int *vals_in, unsigned long long *keys_out, int *vals_out,
                   int n, int begin_bit, int end_bit, void *temp, size_t *temp_bytes, hipStream_t st);
int sort_pairs_u32(unsigned int *keys_in, int *vals_in, unsigned int *keys_out, int *vals_out,
                   int n, int begin_bit, int end_bit, void *temp, size_t *temp_bytes, hipStream_t st);
