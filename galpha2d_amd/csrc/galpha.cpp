// galpha.cpp -- host orchestrator and C ABI of libgalpha.so.
//
// Replaces the reference's host driver gpudt/gpudt.cpp:771-924 (gpudtComputeDT) and its
// global-state stage calls (Cuda/cuda.h:64-75).  State lives in a context (no globals except
// the lazily created default context behind the reference-compatible entry points), every
// stage runs on one HIP stream, and the only host reads inside the pipeline are 4-byte
// counters that decide loop termination.  Snap grid: gpudt.cpp:722-769 (scale with few
// mantissa bits so snapped coordinates are exact doubles) -- restated with a rounded-up
// scale of (53 - log2 T - 2) bits and a shift on the scale's unit.
#include "kernels.h"
#include "../../include/gpudt.h"
#include "alpha_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

void set_deg_for_init(int *deg);
void set_decide_buffer(int *d);
void set_del_list(int *l);
void set_lt_buffer(int *b);
void set_hull_buffers(int *hn, int *he);

static thread_local std::string g_last_error;

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t _e = (x);                                                                   \
        if (_e != hipSuccess) {                                                                \
            throw GalphaError(std::string(#x) + ": " + hipGetErrorString(_e));                 \
        }                                                                                      \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void ensure(size_t b)
    {
        if (b <= bytes) return;
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t nb = b + b / 8 + 256;
        HIPCHK(hipMalloc(&p, nb));
        bytes = nb;
    }
    void release()
    {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() { return (T *)p; }
};

// Frame sites sit on the texture border every frame_step(T) pixels; real points are mapped at
// least margin(T) pixels inside, with frame_step/2 < margin, so every border pixel is nearer
// to a frame site than to any real site.
static int frame_step(int T) { return T >= 2048 ? T / 256 : 8; }
static int margin_for(int T) { return frame_step(T) / 2 + 1; }

struct galpha_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // PBA
    DevBuf smap, g, S, cnt, E, ecnt, vor, bad, fpix;
    // mesh
    DevBuf tv, tn, remap, stamp, claim, modA, modB, cand, pos, vt, vstate, pix, deg, off, fill, fan,
        moved, decide, mx, mlist, loc, lt, scratch, hnext, hedge, keys, keys2, vals, vals2, rot, newid,
        out9, counters, status, xy, scanTmp, sortTmp, tileCnt, tileOff, dellist, dbg;
    AlphaDev alpha;
    int last_n = 0, last_F = 0, last_T = 0;
    bool last_alpha = false;
    galpha_stats stats{};
    hipEvent_t ev[16];
    int op = 1;
    std::vector<int> host_counters;
};

static void ctx_counters_read(galpha_ctx *c, int *dst, int slot, int n)
{
    HIPCHK(hipMemcpyAsync(dst, c->counters.as<int>() + slot, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}
static int ctx_counter(galpha_ctx *c, int slot)
{
    int v = 0;
    ctx_counters_read(c, &v, slot, 1);
    return v;
}
// stream-ordered read of device memory (the caller's stream may be non-blocking)
static void ctx_read(galpha_ctx *c, void *dst, const void *src, size_t bytes)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
}
static int ctx_status(galpha_ctx *c)
{
    int v = 0;
    ctx_read(c, &v, c->status.p, sizeof(int));
    return v;
}
static void ctx_counter_set(galpha_ctx *c, int slot, int v)
{
    HIPCHK(hipMemsetD32Async((hipDeviceptr_t)(c->counters.as<int>() + slot), v, 1, c->stream));
}

struct SnapGrid {
    double shx, shy, scale, inv;
};

// scale = range / (T - 2M - 3) rounded up to (53 - log2T - 2) significant bits; shift on a
// multiple of the scale's unit so that shift + k * scale is exact for |k| <= T.
static SnapGrid snap_grid(double minx, double miny, double maxx, double maxy, int T)
{
    int logT = 0;
    while ((1 << logT) < T) logT++;
    double range = std::max(maxx - minx, maxy - miny);
    if (!(range > 0)) range = std::max(1.0, std::fabs(minx) + std::fabs(miny));
    double raw = range / (double)(T - 2 * margin_for(T) - 3);
    for (int B = 53 - logT - 2; B >= 4; B--) {
        int e;
        std::frexp(raw, &e);                 // raw = f * 2^e, 0.5 <= f < 1
        double unit = std::ldexp(1.0, e - B);
        double scale = std::ceil(raw / unit) * unit;
        double sx = std::floor(minx / unit) * unit, sy = std::floor(miny / unit) * unit;
        double lim = std::ldexp(1.0, 52);
        if (std::fabs(sx / unit) + (double)T * (scale / unit) < lim &&
            std::fabs(sy / unit) + (double)T * (scale / unit) < lim) {
            return SnapGrid{sx, sy, scale, 1.0 / scale};
        }
    }
    throw GalphaError("coordinate range cannot be snapped exactly (data too far from the origin)");
}

static std::vector<int> frame_pixels(int T)
{
    std::vector<int> f;
    auto add = [&](int x, int y) { f.push_back((x << 16) | y); };
    // bottom (y = 0) left->right, right (x = T-1) bottom->top, top right->left, left top->bottom
    for (int x = 0; x < T - 1; x += frame_step(T)) add(x, 0);
    for (int y = 0; y < T - 1; y += frame_step(T)) add(T - 1, y);
    for (int x = T - 1; x > 0; x -= frame_step(T)) add(x, T - 1);
    for (int y = T - 1; y > 0; y -= frame_step(T)) add(0, y);
    return f;
}

static int log2i(int T)
{
    int l = 0;
    while ((1 << l) < T) l++;
    return l;
}

static void check_texture(int T)
{
    if (T < 64 || T > 16384 || (T & (T - 1))) throw GalphaError("fboSize must be a power of two in [64, 16384]");
}

int galphaAutoTexture(int n)
{
    // about 16 pixels per point, at least 256
    int T = 256;
    while (T < 16384 && (double)T * T < 16.0 * (double)n) T <<= 1;
    return T;
}

static void ensure_pba(galpha_ctx *c, int T, int n, int nf)
{
    size_t P = (size_t)T * T;
    c->smap.ensure(4 * P);
    c->g.ensure(2 * P);
    c->S.ensure(4 * P);
    c->cnt.ensure(4 * (size_t)64 * T);
    c->E.ensure(4 * P);
    c->ecnt.ensure(4 * (size_t)T);
    c->vor.ensure(4 * P);
    c->bad.ensure(4 * (size_t)(n + nf + 1));
    c->fpix.ensure(4 * (size_t)(nf + 1));
    c->pix.ensure(4 * (size_t)(n + nf + 1));
    c->counters.ensure(sizeof(int) * C_NCOUNTERS);
    c->status.ensure(sizeof(int) * 4);
    size_t ntiles = (size_t)(T - 1) * ((T - 1 + 255) / 256);
    c->tileCnt.ensure(4 * ntiles);
    c->tileOff.ensure(4 * ntiles);
    c->scanTmp.ensure(scan_temp_bytes(std::max(ntiles, (size_t)(n + nf + 2))) + 4096);
}

static void run_pba(galpha_ctx *c, const double *d_xy, int n, int T, const SnapGrid &sg, int nf, int *iters,
                    int *island_sites)
{
    hipStream_t st = c->stream;
    int logT = log2i(T);
    size_t P = (size_t)T * T;
    launch_fill_i32(c->smap.as<int>(), 0x7fffffff, P, st);
    launch_splat((const double2 *)d_xy, n, sg.shx, sg.shy, sg.inv, margin_for(T), logT, c->smap.as<int>(),
                 c->pix.as<int>(), c->fpix.as<int>(), nf, st);
    HIPCHK(hipMemsetAsync(c->bad.p, 0, 4 * (size_t)(n + nf + 1), st));
    PbaBuffers b{c->smap.as<int>(), c->g.as<short>(), c->S.as<int>(), c->cnt.as<int>(),
                 c->E.as<int>(), c->ecnt.as<int>(), c->vor.as<int>()};
    int m2 = std::min(16, T / 64), m3 = std::min(16, T / 64);
    int it = 0, removed = 0;
    for (; it < 16; it++) {
        launch_pba(b, logT, m2, m3, st);
        ctx_counter_set(c, C_FLAG, 0);
        launch_islands(c->vor.as<int>(), c->pix.as<int>(), logT, c->bad.as<int>(), c->counters.as<int>() + C_FLAG, st);
        if (ctx_counter(c, C_FLAG) == 0) break;
        std::vector<int> hb(n);
        HIPCHK(hipMemcpyAsync(hb.data(), c->bad.p, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int i = 0; i < n; i++) removed += hb[i] != 0;
        launch_remove_bad(c->bad.as<int>(), c->pix.as<int>(), n, logT, c->smap.as<int>(), st);
    }
    if (it == 16) throw GalphaError("discrete Voronoi diagram did not settle (island certificate)");
    *iters = it + 1;
    *island_sites = removed;
}

static float ev_ms(galpha_ctx *c, int a, int b)
{
    float ms = 0;
    hipEventElapsedTime(&ms, c->ev[a], c->ev[b]);
    return ms;
}

static void bbox_host(const double *xy, int n, double *mnx, double *mny, double *mxx, double *mxy)
{
    double a = INFINITY, b = INFINITY, cc = -INFINITY, d = -INFINITY;
    for (int i = 0; i < n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        if (!std::isfinite(x) || !std::isfinite(y)) throw GalphaError("non-finite input coordinate");
        a = std::min(a, x); b = std::min(b, y); cc = std::max(cc, x); d = std::max(d, y);
    }
    *mnx = a; *mny = b; *mxx = cc; *mxy = d;
}

void bbox_device(const double *d_xy, int n, double *out4, void *tmp, hipStream_t st);

// Delaunay triangulation of d_xy (device).  Leaves canonical triangles in c->out9 (9 ints each).
static int run_dt(galpha_ctx *c, const double *d_xy, int n, int T, const double *bbox)
{
    check_texture(T);
    hipStream_t st = c->stream;
    galpha_stats &S = c->stats;
    memset(&S, 0, sizeof(S));
    c->last_n = n;
    c->last_T = T;
    c->last_F = 0;
    if (n < 3) return 0;
    int logT = log2i(T);
    SnapGrid sg = snap_grid(bbox[0], bbox[1], bbox[2], bbox[3], T);
    std::vector<int> fp = frame_pixels(T);
    int nf = (int)fp.size();
    int V = n + nf;
    ensure_pba(c, T, n, nf);
    HIPCHK(hipMemcpyAsync(c->fpix.p, fp.data(), 4 * (size_t)nf, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(c->counters.p, 0, sizeof(int) * C_NCOUNTERS, st));
    HIPCHK(hipMemsetAsync(c->status.p, 0, sizeof(int) * 4, st));
    HIPCHK(hipEventRecord(c->ev[0], st));
    run_pba(c, d_xy, n, T, sg, nf, &S.pbaIterations, &S.islandSites);
    HIPCHK(hipEventRecord(c->ev[1], st));

    // ---- mesh buffers
    // dual triangles <= 2V; each re-inserted point adds 2 slots, each deleted vertex leaves 2 dead
    int cap = 3 * V + 8192;
    c->tv.ensure(16 * (size_t)cap);
    c->tn.ensure(16 * (size_t)cap);
    c->remap.ensure(16 * (size_t)cap);
    c->stamp.ensure(4 * (size_t)cap);
    c->claim.ensure(8 * (size_t)cap);
    c->modA.ensure(4 * (size_t)cap);
    c->modB.ensure(4 * (size_t)cap);
    c->cand.ensure(4 * 3 * (size_t)cap);
    c->pos.ensure(16 * (size_t)V);
    c->vt.ensure(4 * (size_t)V);
    c->vstate.ensure(4 * (size_t)V);
    c->deg.ensure(4 * (size_t)(V + 1));
    c->off.ensure(4 * (size_t)(V + 1));
    c->fill.ensure(4 * (size_t)(V + 1));
    c->fan.ensure(4 * 3 * (size_t)cap);
    c->moved.ensure((size_t)V);
    c->decide.ensure(4 * (size_t)V);
    c->mx.ensure(8 * (size_t)V);
    c->mlist.ensure(4 * (size_t)n);
    c->loc.ensure(4 * (size_t)n);
    c->lt.ensure(4 * (size_t)n);
    size_t scap = std::max((size_t)4 * 8192 + 64, (size_t)(8 * V) + 4096);
    c->scratch.ensure(4 * scap);
    c->hnext.ensure(4 * (size_t)V);
    c->hedge.ensure(4 * (size_t)V);
    c->dellist.ensure(4 * (size_t)cap);
    HIPCHK(hipMemsetAsync(c->stamp.p, 0, 4 * (size_t)cap, st));
    HIPCHK(hipMemsetAsync(c->claim.p, 0, 8 * (size_t)cap, st));
    HIPCHK(hipMemsetAsync(c->vt.p, 0xff, 4 * (size_t)V, st));

    MeshDev m{};
    m.tv = c->tv.as<int4>(); m.tn = c->tn.as<int4>(); m.remap = c->remap.as<int4>();
    m.stamp = c->stamp.as<int>(); m.claim = c->claim.as<unsigned long long>();
    m.cap = cap; m.ntri = c->counters.as<int>() + 40;   // slot 40 holds the slot count
    m.pos = c->pos.as<double2>(); m.xy = (const double2 *)d_xy; m.vt = c->vt.as<int>();
    m.vstate = c->vstate.as<int>(); m.pix = c->pix.as<int>();
    m.n = n; m.V = V; m.shx = sg.shx; m.shy = sg.shy; m.scale = sg.scale; m.margin = margin_for(T);
    m.status = c->status.as<int>(); m.counters = c->counters.as<int>();
    m.modA = c->modA.as<int>(); m.modB = c->modB.as<int>(); m.cand = c->cand.as<int>();
    m.scratch = c->scratch.as<int>(); m.scratch_cap = scap;
    c->dbg.ensure(64 * sizeof(double));
    m.dbg = c->dbg.as<double>();
    set_decide_buffer(c->decide.as<int>());
    set_del_list(c->dellist.as<int>());
    set_lt_buffer(c->lt.as<int>());
    set_hull_buffers(c->hnext.as<int>(), c->hedge.as<int>());

    // ---- reconstruction + adjacency
    launch_recon(c->vor.as<int>(), c->pix.as<int>(), logT, m, c->tileCnt.as<int>(), c->tileOff.as<int>(),
                 c->scanTmp.p, st);
    launch_adjacency(m, c->deg.as<int>(), c->off.as<int>(), c->fill.as<int>(), c->fan.as<int>(), c->scanTmp.p, st);
    set_deg_for_init(c->deg.as<int>());
    launch_init_positions(m, st);
    S.reconTris = ctx_counter(c, 40);
    if (ctx_counter(c, 40) + 0 > cap) throw GalphaError("triangle capacity exceeded in reconstruction");
    {
        int stt = 0;
        stt = ctx_status(c);
        if (stt) throw GalphaError("reconstruction produced an invalid triangle (status " + std::to_string(stt) + ")");
    }
    HIPCHK(hipEventRecord(c->ev[2], st));

    // ---- shifting
    // revert moved vertices of non-ccw triangles until every triangle is ccw (monotone)
    for (int it = 0;; it++) {
        if (it > 100000) throw GalphaError("bulk shifting did not converge");
        ctx_counter_set(c, C_FLAG, 0);
        launch_shift_bulk(m, c->moved.as<unsigned char>(), it == 0, st);
        if (ctx_counter(c, C_FLAG) == 0) break;
    }
    ctx_counter_set(c, C_PEND, 0);
    launch_shift_commit(m, c->moved.as<unsigned char>(), st);
    int pend = ctx_counter(c, C_PEND);
    S.shiftBad = pend;
    for (int r = 0; r < 48 && pend > 0; r++) {
        ctx_counter_set(c, C_FLAG, 0);
        launch_shift_greedy(m, r, st);
        int movedn = ctx_counter(c, C_FLAG);
        pend -= movedn;
        if (movedn == 0 && r >= 8) break;
    }
    for (int r = 0; pend > 0; r++) {
        if (r > 10000) throw GalphaError("vertex deletion did not finish");
        ctx_counter_set(c, C_SCR, 0);
        ctx_counter_set(c, C_MODN, 0);
        int before = ctx_counter(c, C_DELETED);
        launch_delete_round(m, c->mx.as<unsigned int>(), 1000 + r, c->op++, st);
        int after = ctx_counter(c, C_DELETED);
        pend -= after - before;
        int stt = 0;
        stt = ctx_status(c);
        if (stt) {
            int dbg[16];
            ctx_counters_read(c, dbg, C_DBG, 16);
            std::string d = " [v=" + std::to_string(dbg[1]) + " poly(" + std::to_string(dbg[2]) + "):";
            for (int k = 0; k < 12 && k < dbg[2]; k++) d += " " + std::to_string(dbg[3 + k]);
            double dp[48];
            ctx_read(c, dp, c->dbg.p, sizeof(dp));
            char buf[64];
            d += " pos:";
            for (int k = 0; k < 2 + 2 * std::min(dbg[2], 15); k++) {
                snprintf(buf, sizeof(buf), " %.17g", dp[k]);
                d += buf;
            }
            d += " states:";
            for (int k = 0; k < std::min(dbg[2], 15); k++) d += " " + std::to_string((int)dp[32 + k]);
            throw GalphaError("vertex deletion failed (status " + std::to_string(stt) + ")" + d + "]");
        }
    }
    S.deleted = ctx_counter(c, C_DELETED);
    HIPCHK(hipEventRecord(c->ev[3], st));

    // ---- flips to a Delaunay mesh of the current vertices
    auto flip_until_done = [&](int *listA, int cntA, int *listB, int cntB, int prev) {
        int *lin = listA, *lout = listB, sin = cntA, sout = cntB;
        for (int pass = 0; pass < 100000; pass++) {
            int opn = c->op++;
            ctx_counter_set(c, C_CAND, 0);
            ctx_counter_set(c, sout, 0);
            ctx_counter_set(c, C_FLIPS, 0);
            launch_flip_pass(m, lin, sin, lout, sout, prev, opn, st);
            S.flipPasses++;
            int fl = ctx_counter(c, C_FLIPS);
            S.flips += fl;
            prev = opn;
            std::swap(lin, lout);
            std::swap(sin, sout);
            if (fl == 0) return;
        }
        throw GalphaError("flipping did not converge");
    };
    auto flip_all = [&]() {
        int opn = c->op++;
        ctx_counter_set(c, C_MODN, 0);
        launch_flip_all(m, opn, m.modA, st);
        flip_until_done(m.modA, C_MODN, m.modB, C_MODN2, opn);
    };
    flip_all();
    HIPCHK(hipEventRecord(c->ev[4], st));

    // ---- missing sites
    ctx_counter_set(c, C_MISS, 0);
    launch_missing_init(m, c->vor.as<int>(), logT, c->mlist.as<int>(), c->loc.as<int>(), st);
    int nmiss = ctx_counter(c, C_MISS);
    S.missing = nmiss;
    for (int r = 0; r < 100000; r++) {
        int cnts[C_NCOUNTERS];
        ctx_counters_read(c, cnts, 0, C_NCOUNTERS);
        int left = nmiss - cnts[C_INSERTED] - cnts[C_DUPS];
        if (left <= 0) break;
        {
            int stt = 0;
            stt = ctx_status(c);
            if (stt) throw GalphaError("missing-site insertion failed (status " + std::to_string(stt) + ")");
        }
        if (r > 5000) throw GalphaError("missing-site insertion did not finish");
        int opn = c->op++;
        ctx_counter_set(c, C_MODN, 0);
        launch_insert_round(m, c->mlist.as<int>(), n, c->loc.as<int>(), opn, st);
        launch_relink(m, m.modA, C_MODN, opn, st);
        S.insertRounds++;
        flip_until_done(m.modA, C_MODN, m.modB, C_MODN2, opn);
    }
    S.inserted = ctx_counter(c, C_INSERTED);
    S.duplicates = ctx_counter(c, C_DUPS);
    HIPCHK(hipEventRecord(c->ev[5], st));

    // ---- frame removal + hull convexification
    ctx_counter_set(c, C_FRAME, nf);
    for (int r = 0;; r++) {
        if (r > 100000) throw GalphaError("frame removal did not finish");
        ctx_counter_set(c, C_SCR, 0);
        ctx_counter_set(c, C_MODN, 0);
        launch_frame_round(m, c->mx.as<unsigned int>(), 5000 + r, m.modA, C_MODN, c->op++, st);
        int left = ctx_counter(c, C_FRAME);
        if (left <= 0) break;
    }
    ctx_counter_set(c, C_MODN, 0);
    launch_convexify(m, m.modA, C_MODN, c->op++, st);
    HIPCHK(hipEventRecord(c->ev[6], st));
    flip_all();
    HIPCHK(hipEventRecord(c->ev[7], st));
    {
        int stt = 0;
        stt = ctx_status(c);
        if (stt) throw GalphaError("device pipeline error (status " + std::to_string(stt) + ")");
    }

    // ---- canonical output
    int nslots = ctx_counter(c, 40);
    c->keys.ensure(8 * (size_t)nslots + 64);
    c->keys2.ensure(8 * (size_t)nslots + 64);
    c->vals.ensure(4 * (size_t)nslots + 64);
    c->vals2.ensure(4 * (size_t)nslots + 64);
    c->rot.ensure(4 * (size_t)nslots + 64);
    c->newid.ensure(4 * (size_t)nslots + 64);
    ctx_counter_set(c, C_RECON_TRIS, 0);
    launch_canon_keys(m, c->keys.as<unsigned long long>(), c->vals.as<int>(), c->rot.as<int>(), st);
    size_t tb = 0;
    sort_pairs_u64(c->keys.as<unsigned long long>(), c->vals.as<int>(), c->keys2.as<unsigned long long>(),
                   c->vals2.as<int>(), nslots, 0, 64, nullptr, &tb, st);
    c->sortTmp.ensure(tb + 256);
    tb = c->sortTmp.bytes;
    int se = sort_pairs_u64(c->keys.as<unsigned long long>(), c->vals.as<int>(), c->keys2.as<unsigned long long>(),
                            c->vals2.as<int>(), nslots, 0, 64, c->sortTmp.p, &tb, st);
    if (se) throw GalphaError("radix sort failed");
    int F = ctx_counter(c, C_RECON_TRIS);
    c->out9.ensure(36 * (size_t)F + 64);
    launch_canon_emit(m, c->vals2.as<int>(), c->rot.as<int>(), c->newid.as<int>(), F, c->out9.as<int>(), st);
    HIPCHK(hipEventRecord(c->ev[8], st));
    HIPCHK(hipStreamSynchronize(st));
    c->last_F = F;
    S.nTris = F;
    int cnts[C_NCOUNTERS];
    ctx_counters_read(c, cnts, 0, C_NCOUNTERS);
    S.incircleTests = cnts[C_TESTS_INCIRCLE];
    S.incircleExact = cnts[C_EXACT_INCIRCLE];
    S.orientExact = cnts[C_EXACT_ORIENT];
    S.ms_pba = ev_ms(c, 0, 1);
    S.ms_recon = ev_ms(c, 1, 2);
    S.ms_shift = ev_ms(c, 2, 3) + ev_ms(c, 3, 4);
    S.ms_missing = ev_ms(c, 4, 5);
    S.ms_boundary = ev_ms(c, 5, 6);
    S.ms_flip = ev_ms(c, 6, 7);
    S.ms_canon = ev_ms(c, 7, 8);
    S.ms_total = ev_ms(c, 0, 8);
    return F;
}

// ---------------------------------------------------------------------------------------------
// C ABI
extern "C" {

galpha_ctx *galphaCreate(int device)
{
    try {
        galpha_ctx *c = new galpha_ctx();
        c->device = device;
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (int i = 0; i < 16; i++) HIPCHK(hipEventCreate(&c->ev[i]));
        return c;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

static void release_all(galpha_ctx *c)
{
    DevBuf *bufs[] = {&c->smap, &c->g, &c->S, &c->cnt, &c->E, &c->ecnt, &c->vor, &c->bad, &c->fpix, &c->tv,
                      &c->tn, &c->remap, &c->stamp, &c->claim, &c->modA, &c->modB, &c->cand, &c->pos, &c->vt,
                      &c->vstate, &c->pix, &c->deg, &c->off, &c->fill, &c->fan, &c->moved, &c->decide, &c->mx,
                      &c->mlist, &c->loc, &c->lt, &c->scratch, &c->hnext, &c->hedge, &c->keys, &c->keys2,
                      &c->vals, &c->vals2, &c->rot, &c->newid, &c->out9, &c->counters, &c->status, &c->xy,
                      &c->scanTmp, &c->sortTmp, &c->tileCnt, &c->tileOff, &c->dellist, &c->dbg};
    for (DevBuf *b : bufs) b->release();
    alpha_release(c->alpha);
}

void galphaDestroy(galpha_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    release_all(c);
    for (int i = 0; i < 16; i++) hipEventDestroy(c->ev[i]);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char *galphaLastError(galpha_ctx *c) { return c ? c->err.c_str() : g_last_error.c_str(); }
const char *gpudtGetLastError(void) { return g_last_error.c_str(); }

static int fail(galpha_ctx *c, const std::exception &e)
{
    g_last_error = e.what();
    if (c) c->err = e.what();
    return -1;
}

int galphaComputeDevice(galpha_ctx *c, const double *d_xy, int n, int fboSize, int with_alpha, void *stream)
{
    if (!c) return -1;
    try {
        HIPCHK(hipSetDevice(c->device));
        hipStream_t own = c->stream;
        if (stream) c->stream = (hipStream_t)stream;
        double bb[4];
        if (n > 0) {
            c->scanTmp.ensure(4096);
            bbox_device(d_xy, n, bb, c->scanTmp.p, c->stream);
        }
        int F = run_dt(c, d_xy, n, fboSize, bb);
        c->last_alpha = false;
        if (with_alpha) {
            auto t0 = std::chrono::steady_clock::now();
            alpha_run(c->alpha, d_xy, n, c->out9.as<int>(), F, c->stream, &c->stats);
            c->last_alpha = true;
            HIPCHK(hipStreamSynchronize(c->stream));
            auto t1 = std::chrono::steady_clock::now();
            c->stats.ms_alpha = std::chrono::duration<double, std::milli>(t1 - t0).count();
        }
        c->stream = own;
        return F;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int galphaDeviceTriangles(galpha_ctx *c, const int **d_tris9, int *ntris)
{
    if (!c) return -1;
    *d_tris9 = c->out9.as<int>();
    *ntris = c->last_F;
    return 0;
}

int galphaCopyTriangles(galpha_ctx *c, int *out, int capacity)
{
    if (!c) return -1;
    try {
        int F = c->last_F;
        if (capacity < F) throw GalphaError("output capacity too small");
        if (F > 0) ctx_read(c, out, c->out9.p, 36 * (size_t)F);
        return F;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int galphaGetStats(galpha_ctx *c, galpha_stats *s)
{
    if (!c || !s) return -1;
    *s = c->stats;
    return 0;
}

int galphaComputeHost(galpha_ctx *c, const double *xy, int n, int fboSize, int *tris_out9, int *ntris_out)
{
    if (!c) return -1;
    try {
        HIPCHK(hipSetDevice(c->device));
        double bb[4] = {0, 0, 1, 1};
        if (n > 0) bbox_host(xy, n, &bb[0], &bb[1], &bb[2], &bb[3]);
        c->xy.ensure(16 * (size_t)std::max(n, 1));
        if (n > 0) HIPCHK(hipMemcpyAsync(c->xy.p, xy, 16 * (size_t)n, hipMemcpyHostToDevice, c->stream));
        int F = run_dt(c, c->xy.as<double>(), n, fboSize, bb);
        if (tris_out9) {
            if (*ntris_out < F) throw GalphaError("output capacity too small");
            if (F > 0) HIPCHK(hipMemcpyAsync(tris_out9, c->out9.p, 36 * (size_t)F, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        *ntris_out = F;
        return F;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int galphaVoronoi(galpha_ctx *c, const double *xy, int n, int T, int *labels, int *pix, int pix_capacity)
{
    if (!c) return -1;
    try {
        HIPCHK(hipSetDevice(c->device));
        check_texture(T);
        double bb[4] = {0, 0, 1, 1};
        if (n > 0) bbox_host(xy, n, &bb[0], &bb[1], &bb[2], &bb[3]);
        SnapGrid sg = snap_grid(bb[0], bb[1], bb[2], bb[3], T);
        std::vector<int> fp = frame_pixels(T);
        int nf = (int)fp.size();
        if (pix_capacity < n + nf) throw GalphaError("pix capacity too small");
        ensure_pba(c, T, n, nf);
        c->xy.ensure(16 * (size_t)std::max(n, 1));
        HIPCHK(hipMemcpyAsync(c->xy.p, xy, 16 * (size_t)n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->fpix.p, fp.data(), 4 * (size_t)nf, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemsetAsync(c->counters.p, 0, sizeof(int) * C_NCOUNTERS, c->stream));
        int iters = 0, isl = 0;
        run_pba(c, c->xy.as<double>(), n, T, sg, nf, &iters, &isl);
        c->stats.pbaIterations = iters;
        c->stats.islandSites = isl;
        HIPCHK(hipMemcpyAsync(labels, c->vor.p, 4 * (size_t)T * T, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(pix, c->pix.p, 4 * (size_t)(n + nf), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return nf;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int galphaCopyAlpha(galpha_ctx *c, PGPUDTALPHA *pA)
{
    if (!c || !pA) return -1;
    try {
        if (!c->last_alpha) throw GalphaError("no alpha result on this context");
        *pA = alpha_copy_to_host(c->alpha, c->stream);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

// ---- reference-compatible entry points on a lazily created default context --------------------
static std::mutex g_default_mu;
static galpha_ctx *g_default = nullptr;
static galpha_ctx *default_ctx()
{
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default) g_default = galphaCreate(0);
    return g_default;
}

PGPUDTOUTPUT gpudtComputeDT(PGPUDTPARAMS p)
{
    try {
        g_last_error.clear();
        if (!p) throw GalphaError("null parameters");
        if (p->nConstraints > 0) throw GalphaError("constrained Delaunay triangulation is not supported");
        if (p->nPoints < 0 || (p->nPoints > 0 && !p->points)) throw GalphaError("invalid point list");
        galpha_ctx *c = default_ctx();
        if (!c) throw GalphaError(g_last_error);
        std::lock_guard<std::mutex> lk(g_default_mu);
        int n = p->nPoints;
        int cap = 0;
        int F = galphaComputeHost(c, (const double *)p->points, n, p->fboSize, nullptr, &cap);
        if (F < 0) throw GalphaError(c->err);
        PGPUDTOUTPUT out = new GPUDTOUTPUT;
        out->nTris = F;
        out->triangles = new gpudtTriangle[F > 0 ? F : 1];
        if (F > 0) ctx_read(c, out->triangles, c->out9.p, 36 * (size_t)F);
        return out;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void gpudtReleaseDTOutput(PGPUDTOUTPUT o)
{
    if (!o) return;
    delete[] o->triangles;
    delete o;
}

int gpudtComputeAlpha(PGPUDTPARAMS p, PGPUDTOUTPUT dt, PGPUDTALPHA *pA)
{
    try {
        g_last_error.clear();
        if (!p || !dt || !pA) throw GalphaError("null argument");
        galpha_ctx *c = default_ctx();
        if (!c) throw GalphaError(g_last_error);
        std::lock_guard<std::mutex> lk(g_default_mu);
        int n = p->nPoints;
        c->xy.ensure(16 * (size_t)std::max(n, 1));
        if (n > 0) HIPCHK(hipMemcpyAsync(c->xy.p, p->points, 16 * (size_t)n, hipMemcpyHostToDevice, c->stream));
        c->out9.ensure(36 * (size_t)std::max(dt->nTris, 1));
        if (dt->nTris > 0)
            HIPCHK(hipMemcpyAsync(c->out9.p, dt->triangles, 36 * (size_t)dt->nTris, hipMemcpyHostToDevice, c->stream));
        alpha_run(c->alpha, c->xy.as<double>(), n, c->out9.as<int>(), dt->nTris, c->stream, &c->stats);
        *pA = alpha_copy_to_host(c->alpha, c->stream);
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

void gpudtReleaseAlpha(PGPUDTALPHA a) { alpha_free_host(a); }

void galphaFree(void *p) { free(p); }

// Batched entry point: every set is triangulated (and optionally gets its alpha spectrum) on
// this context; results are concatenated.  tris9_out / tri_offsets_out are malloc'ed (release
// with galphaFree).
int galphaBatch(galpha_ctx *c, const double *xy, const int *offsets, int nsets, const int *fboSizes,
                int with_alpha, int **tris9_out, int **tri_offsets_out)
{
    if (!c || !xy || !offsets || nsets < 0 || !tris9_out || !tri_offsets_out) return -1;
    try {
        HIPCHK(hipSetDevice(c->device));
        std::vector<int> toff(nsets + 1, 0);
        std::vector<int> all;
        for (int s = 0; s < nsets; s++) {
            int b = offsets[s], e = offsets[s + 1];
            if (e < b) throw GalphaError("offsets must be non-decreasing");
            int n = e - b;
            int T = (fboSizes && fboSizes[s] > 0) ? fboSizes[s] : galphaAutoTexture(n);
            double bb[4] = {0, 0, 1, 1};
            if (n > 0) bbox_host(xy + 2 * (size_t)b, n, &bb[0], &bb[1], &bb[2], &bb[3]);
            c->xy.ensure(16 * (size_t)std::max(n, 1));
            if (n > 0)
                HIPCHK(hipMemcpyAsync(c->xy.p, xy + 2 * (size_t)b, 16 * (size_t)n, hipMemcpyHostToDevice, c->stream));
            int F = run_dt(c, c->xy.as<double>(), n, T, bb);
            if (with_alpha) alpha_run(c->alpha, c->xy.as<double>(), n, c->out9.as<int>(), F, c->stream, &c->stats);
            size_t old = all.size();
            all.resize(old + 9 * (size_t)F);
            if (F > 0)
                HIPCHK(hipMemcpyAsync(all.data() + old, c->out9.p, 36 * (size_t)F, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            toff[s + 1] = toff[s] + F;
        }
        *tris9_out = (int *)malloc(sizeof(int) * std::max<size_t>(all.size(), 1));
        *tri_offsets_out = (int *)malloc(sizeof(int) * (nsets + 1));
        memcpy(*tris9_out, all.data(), sizeof(int) * all.size());
        memcpy(*tri_offsets_out, toff.data(), sizeof(int) * (nsets + 1));
        return nsets;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

}  // extern "C"
