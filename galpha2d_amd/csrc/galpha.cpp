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
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <string>
#include <vector>


static thread_local std::string g_last_error;

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t _e = (x);                                                                   \
        if (_e != hipSuccess) {                                                                \
            throw GalphaError(std::string(#x) + ": " + hipGetErrorString(_e));                 \
        }                                                                                      \
    } while (0)

// Fixed device region that buffers are carved from instead of hipMalloc while a no-allocation
// run is active on this thread (bump allocation, reset at the start of each run).
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0, high = 0;
};
static thread_local Arena *t_arena = nullptr;

struct ArenaScope {
    Arena *prev;
    explicit ArenaScope(Arena *a) : prev(t_arena) { t_arena = a; }
    ~ArenaScope() { t_arena = prev; }
};

__attribute__((visibility("hidden"))) bool galpha_arena_active() { return t_arena != nullptr; }

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool carved = false;   // p lies in an arena: never freed
    void ensure(size_t b)
    {
        if (b <= bytes) return;
        if (p && !carved) hipFree(p);
        p = nullptr;
        bytes = 0;
        carved = false;
        if (t_arena) {
            size_t nb = (b + 255) & ~(size_t)255;
            if (t_arena->used + nb > t_arena->cap)
                throw GalphaError("scratch exhausted: the run needs more than the " +
                                  std::to_string(t_arena->cap >> 20) + " MiB of the no-allocation scratch");
            p = t_arena->base + t_arena->used;
            t_arena->used += nb;
            t_arena->high = std::max(t_arena->high, t_arena->used);
            bytes = nb;
            carved = true;
            return;
        }
        size_t nb = b + b / 8 + 256;
        HIPCHK(hipMalloc(&p, nb));
        bytes = nb;
    }
    void release()
    {
        if (p && !carved) hipFree(p);
        p = nullptr;
        bytes = 0;
        carved = false;
    }
    template <class T> T *as() { return (T *)p; }
};

// Frame sites sit on the texture border every frame_step(T) pixels; real points are mapped at
// least margin(T) pixels inside, with frame_step/2 < margin, so every border pixel is nearer
// to a frame site than to any real site.
#define GALPHA_TUNE_CLASSES 4
static int frame_step(int T) { return T >= 2048 ? T / 256 : 8; }
static int margin_for(int T) { return frame_step(T) / 2 + 1; }

struct galpha_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // PBA
    DevBuf smap, g, S, cnt, E, ecnt, vor, bad, fpix;
    // mesh
    DevBuf plo;   // phase-2 band range starts
    DevBuf perm, xys, pcell, ccount, ccursor;   // spatial numbering of the real points
    DevBuf cls;                                  // alpha-complex classes (host queries)
    DevBuf tv, tn, remap, stamp, claim, modA, modB, cand, pos, vt, vstate, pix, deg, off, fill, fan,
        moved, decide, mx, mlist, loc, lt, scratch, hnext, hedge, keys, keys2, vals, vals2, rot, newid,
        out9, counters, status, xy, scanTmp, sortTmp, tileCnt, tileOff, dellist, dbg;
    // batch plan on the device
    DevBuf setsD, bmD, vset, hullCnt, hullOff, hullStart, setFirst, offsD, bboxTmp, lmark, ffan, modF, badset, bnd;
    // constrained DT: input pairs, vertex-id pairs, inverse numbering, flags, walk counts, region
    // sizes and offsets, cavity regions, constraint key set
    DevBuf kin, kv, kinv, kdone, kcnt, krs, kroff, kwork, khash;
    DevBuf rdbuf;   // segments gathered for one host read
    std::vector<int> cons;            // constraint pairs (input indices) for single-set runs
    int cons_rounds = 0;              // insertion rounds of the last constrained run
    std::vector<int> set_first;       // [nsets + 1] first output triangle of each planned set
    std::vector<SetDesc> plan_sets;   // host copy of the last plan
    int plan_ntot = 0;
    std::vector<int> batch_sets;      // input set of each planned set of the last batch
    std::vector<int> batch_npts;      // points of every input set of the last batch
    const double *batch_xy = nullptr; // device points of the planned sets of the last batch
    AlphaDev alpha;
    double density[GALPHA_TUNE_CLASSES] = {16.0, 16.0, 16.0, 16.0};   // pixels per point by size class
    int last_n = 0, last_F = 0, last_T = 0;
    bool last_alpha = false;
    galpha_stats stats{};
    hipEvent_t ev[16] = {};
    int op = 1;
    // the slot-tag arrays (stamp, claim, lmark) and the per-vertex removal maxima (mx) as last
    // cleared, and the removal round number, which grows across runs like op
    void *tag_p[3] = {nullptr, nullptr, nullptr};
    void *mx_p = nullptr;
    int prio_round = 0;
    int *hsnap = nullptr;   // staging for counter reads (SNAP_INTS ints)
    bool hsnap_pinned = false;   // pinned (kernels write it directly), else pageable
    std::vector<int> hsnap_pageable;   // counter reads of a no-allocation context
    // the last gpudtComputeDT on this (default) context: the caller's point array and the returned
    // triangle array with their sizes and content hashes, so that gpudtComputeAlpha of that output
    // runs on the points and triangles still resident on the device
    struct {
        const void *points = nullptr, *tris = nullptr;
        int n = -1, F = -1;
        unsigned long long hp = 0, ht = 0;
    } last_dt;
    HostStage stage;        // pinned staging for large host <-> device copies
};

// ---- pinned staging ------------------------------------------------------------------------
static const size_t STAGE_HALF = (size_t)32 << 20;

// Position-dependent hash of bytes [base, base + n) of an array whose bytes from `base` are at p:
// a sum of mixed 8-byte words (and the tail bytes), each mixed with its offset, so that parts of an
// array hashed separately add up to the hash of the whole.  Detects changed, moved or swapped
// data; not meant against adversaries.
static unsigned long long content_hash(const char *p, size_t n, size_t base)
{
    auto mix = [](unsigned long long x) {
        x ^= x >> 31;
        x *= 0x7fb5d329728ea185ull;
        x ^= x >> 27;
        x *= 0x81dadef4bc2dd44dull;
        return x ^ (x >> 33);
    };
    unsigned long long h = 0;
    size_t i = 0;
    // words at offsets that are multiples of 8 in the whole array (base is a multiple of 4096)
    for (; i + 8 <= n; i += 8) {
        unsigned long long w;
        memcpy(&w, p + i, 8);
        h += mix(w ^ ((base + i) * 0x9E3779B97F4A7C15ull));
    }
    for (; i < n; i++) h += mix((unsigned long long)(unsigned char)p[i] ^ ((base + i) * 0xC2B2AE3D27D4EB4Full));
    return h;
}

// Host copies between the pinned staging halves and caller arrays run on several threads: one
// thread copies about 10 GB/s, far below what the host's memory and PCIe give, and the results of
// a 1M-point run (triangles and alpha arrays) are about 330 MB.  The pool belongs to one context
// (serving slots copy concurrently, each with its own).
struct CopyPool {
    static const int kThreads = 7;      // workers; the calling thread takes a share too
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable go, done;
    char *dst = nullptr;
    const char *src = nullptr;
    size_t len = 0, part = 0;
    int nparts = 0, next = 0, pending = 0;
    unsigned gen = 0;
    bool stop = false;
    size_t hbase = 0;                        // offset of src in the hashed array
    std::atomic<unsigned long long> *hash = nullptr;

    CopyPool()
    {
        for (int k = 0; k < kThreads; k++) th.emplace_back([this] { work(); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        go.notify_all();
        for (auto &t : th) t.join();
    }
    // takes parts of the current copy until none is left
    void drain()
    {
        for (;;) {
            int k;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (next >= nparts) return;
                k = next++;
            }
            size_t off = (size_t)k * part, n = std::min(part, len - off);
            if (dst) memcpy(dst + off, src + off, n);
            if (hash) hash->fetch_add(content_hash(src + off, n, hbase + off));
            std::lock_guard<std::mutex> lk(mu);
            if (--pending == 0) done.notify_all();
        }
    }
    void work()
    {
        unsigned seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                go.wait(lk, [&] { return stop || gen != seen; });
                if (stop) return;
                seen = gen;
            }
            drain();
        }
    }
    // dst NULL: hash only; h non-NULL: add the content hash of src (at offset base of its array)
    void copy(void *d, const void *s_, size_t n, std::atomic<unsigned long long> *h = nullptr, size_t base = 0)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            dst = (char *)d;
            src = (const char *)s_;
            hash = h;
            hbase = base;
            len = n;
            nparts = kThreads + 1;
            part = (n + nparts - 1) / nparts;
            part = (part + 4095) & ~(size_t)4095;
            nparts = (int)((n + part - 1) / part);
            next = 0;
            pending = nparts;
            gen++;
        }
        go.notify_all();
        drain();
        std::unique_lock<std::mutex> lk(mu);
        done.wait(lk, [&] { return pending == 0; });
    }
};

static void host_copy(HostStage &s, void *dst, const void *src, size_t n,
                      std::atomic<unsigned long long> *hash = nullptr, size_t base = 0)
{
    if (n < ((size_t)4 << 20)) {
        if (dst) memcpy(dst, src, n);
        if (hash) hash->fetch_add(content_hash((const char *)src, n, base));
        return;
    }
    if (!s.pool) s.pool = new CopyPool();
    s.pool->copy(dst, src, n, hash, base);
}

// content hash of a host array (sum of per-part hashes, so any split gives the same value)
static unsigned long long host_hash(HostStage &s, const void *p, size_t n)
{
    std::atomic<unsigned long long> h{0};
    host_copy(s, nullptr, p, n, &h, 0);
    return h.load();
}

static void stage_ensure(HostStage &s)
{
    if (s.p) return;
    if (t_arena) throw GalphaError("pinned staging is not available in a no-allocation call");
    HIPCHK(hipHostMalloc((void **)&s.p, 2 * STAGE_HALF, hipHostMallocDefault));
    s.half = STAGE_HALF;
    for (int k = 0; k < 2; k++) HIPCHK(hipEventCreateWithFlags(&s.ev[k], hipEventDisableTiming));
}

void stage_free(HostStage &s)
{
    delete s.pool;
    if (s.p) hipHostFree(s.p);
    for (int k = 0; k < 2; k++)
        if (s.ev[k]) hipEventDestroy(s.ev[k]);
    s = HostStage();
}

void stage_d2h(HostStage &s, void *dst, const void *src, size_t bytes, hipStream_t st)
{
    stage_d2h_hash(s, dst, src, bytes, st, nullptr);
}

void stage_d2h_hash(HostStage &s, void *dst, const void *src, size_t bytes, hipStream_t st,
                    std::atomic<unsigned long long> *hash)
{
    if (bytes == 0) return;
    stage_ensure(s);
    size_t nch = (bytes + s.half - 1) / s.half;
    for (size_t i = 0; i <= nch; i++) {
        if (i < nch) {   // DMA of chunk i into half i % 2 (its previous chunk was copied out below)
            size_t off = i * s.half, len = std::min(s.half, bytes - off);
            HIPCHK(hipMemcpyAsync(s.p + (i & 1) * s.half, (const char *)src + off, len, hipMemcpyDeviceToHost, st));
            HIPCHK(hipEventRecord(s.ev[i & 1], st));
        }
        if (i > 0) {     // host copy of chunk i - 1 while chunk i is in flight
            size_t j = i - 1, off = j * s.half, len = std::min(s.half, bytes - off);
            HIPCHK(hipEventSynchronize(s.ev[j & 1]));
            host_copy(s, (char *)dst + off, s.p + (j & 1) * s.half, len, hash, off);
        }
    }
}

void stage_h2d(HostStage &s, void *dst, const void *src, size_t bytes, hipStream_t st)
{
    stage_h2d_hash(s, dst, src, bytes, st, nullptr);
}

void stage_h2d_hash(HostStage &s, void *dst, const void *src, size_t bytes, hipStream_t st,
                    std::atomic<unsigned long long> *hash)
{
    if (bytes == 0) return;
    stage_ensure(s);
    size_t nch = (bytes + s.half - 1) / s.half;
    for (size_t i = 0; i < nch; i++) {
        size_t off = i * s.half, len = std::min(s.half, bytes - off);
        if (i >= 2) HIPCHK(hipEventSynchronize(s.ev[i & 1]));   // the DMA of chunk i - 2 left this half
        host_copy(s, s.p + (i & 1) * s.half, (const char *)src + off, len, hash, off);
        HIPCHK(hipMemcpyAsync((char *)dst + off, s.p + (i & 1) * s.half, len, hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(s.ev[i & 1], st));
    }
    // the staging halves are free again once their last DMAs ended
    HIPCHK(hipEventSynchronize(s.ev[(nch - 1) & 1]));
    if (nch >= 2) HIPCHK(hipEventSynchronize(s.ev[nch & 1]));
}

static const int SNAP_INTS = 256;

// Counter reads: a kernel writes the counters into pinned host memory, then one wait (a no-
// allocation context has pageable staging: one device-to-host copy instead).  fold: the slots
// C_FOLD + k come from the statistics stripes.
static void ctx_counters_read(galpha_ctx *c, int *dst, int slot, int n, bool fold = false)
{
    if (n > SNAP_INTS) throw GalphaError("counter read too large");
    if (c->hsnap_pinned) {
        launch_counters_to_host(c->counters.as<int>(), slot, n, fold ? 1 : 0, c->hsnap, c->stream);
    } else {
        if (fold) launch_fold_stats(c->counters.as<int>(), c->stream);
        HIPCHK(hipMemcpyAsync(c->hsnap, c->counters.as<int>() + slot, sizeof(int) * n, hipMemcpyDeviceToHost,
                              c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    memcpy(dst, c->hsnap, sizeof(int) * n);
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
    stage_d2h(c->stage, dst, src, bytes, c->stream);
}
static int ctx_status(galpha_ctx *c) { return ctx_counter(c, C_STATUS); }

// Device int segments (up to four) in one host read: gathered into one device array, then one copy
// through pinned memory and one wait (each copy into pageable memory waits a round trip of its own).
static void ctx_read_segments(galpha_ctx *c, std::initializer_list<std::pair<const int *, int>> segs, int *dst)
{
    IntSegments g{};
    size_t tot = 0;
    for (auto &sg : segs) {
        if (g.n == 4) throw GalphaError("too many segments in one read");
        g.src[g.n] = sg.first;
        g.len[g.n++] = sg.second;
        tot += (size_t)sg.second;
    }
    if (tot == 0) return;
    if (c->hsnap_pinned && tot <= (size_t)SNAP_INTS) {   // gathered straight into pinned memory
        launch_gather_segments(g, c->hsnap, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
        memcpy(dst, c->hsnap, 4 * tot);
        return;
    }
    c->rdbuf.ensure(4 * tot + 64);
    launch_gather_segments(g, c->rdbuf.as<int>(), c->stream);
    if (tot <= (size_t)SNAP_INTS) {
        HIPCHK(hipMemcpyAsync(c->hsnap, c->rdbuf.p, 4 * tot, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        memcpy(dst, c->hsnap, 4 * tot);
    } else if (t_arena) {   // no pinned staging in a no-allocation run
        HIPCHK(hipMemcpyAsync(dst, c->rdbuf.p, 4 * tot, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } else {
        stage_d2h(c->stage, dst, c->rdbuf.p, 4 * tot, c->stream);
    }
}
static void ctx_counter_set(galpha_ctx *c, int slot, int v)
{
    HIPCHK(hipMemsetD32Async((hipDeviceptr_t)(c->counters.as<int>() + slot), v, 1, c->stream));
}
// striped statistics (see kernels.h): fold the stripes on the device, then read
static int ctx_stat(galpha_ctx *c, int k)
{
    int v = 0;
    ctx_counters_read(c, &v, C_FOLD + k, 1, true);
    return v;
}
// all single-word counters, the status word and the folded statistics in one read:
// snap[k] = counter k, snap[C_FOLD + k] = statistic k
struct Snap {
    int v[C_FOLD + 64];
    int stat(int k) const { return v[C_FOLD + k]; }
    int status() const { return v[C_STATUS]; }
};
static Snap ctx_snapshot(galpha_ctx *c)
{
    Snap s;
    ctx_counters_read(c, s.v, 0, C_FOLD + 64, true);
    return s;
}
static const size_t COUNTER_BYTES = sizeof(int) * GA_STRIPE_INTS * (1 + GA_STRIPES);

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

// Triangle slots are named by (t << 2) | j in int32 links, so slot ids stay below 2^29.  A pass
// reserves 3 V + 8192 slots (V = real points + frame sites of all its sets): at most 178,954,240
// vertices per pass.  Checked from the sizes alone, before any device work touches the points.
static const long long kMaxSlots = 1LL << 29;
static long long frame_count(int T) { return (long long)frame_pixels(T).size(); }
static void check_slot_capacity(long long V)
{
    if (3 * V + 8192 > kMaxSlots)
        throw GalphaError("too many points in one pass: " + std::to_string(V) + " vertices need " +
                          std::to_string(3 * V + 8192) + " triangle slots, and slot ids (t << 2 | j) must stay "
                          "below 2^29 (at most 178954240 vertices, frame sites included)");
}

// smallest power-of-two side in [256, 16384] with at least `density` pixels per point
static int texture_for_density(int n, double density)
{
    int T = 256;
    while (T < 16384 && (double)T * T < density * (double)n) T <<= 1;
    return T;
}

int galphaAutoTexture(int n) { return texture_for_density(n, 16.0); }

// size classes of the texture autotuner (upper bounds in points, exclusive)
static const int kTuneClassMax[GALPHA_TUNE_CLASSES] = {4096, 32768, 262144, 0x7fffffff};
static int tune_class(int n)
{
    int k = 0;
    while (k + 1 < GALPHA_TUNE_CLASSES && n >= kTuneClassMax[k]) k++;
    return k;
}

// ---- batch plan: per-set descriptors, block maps and frame pixels ----------------------------
struct BatchPlan {
    std::vector<SetDesc> sets;
    std::vector<int> bm;              // BM_COUNT x (nsets + 1)
    int blocks[BM_COUNT] = {0};
    int max_m2 = 1;                   // largest phase-2 band count
    int ncells = 0;                   // 16 x 16-pixel cells of all textures
    std::vector<int> fpix;            // frame pixels of all sets, in set order
    int ntot = 0, V = 0, sumT = 0;
    size_t P = 0;
    int nsets() const { return (int)sets.size(); }
};

// n[s] points (>= 3), texture side T[s], box bb5[5 s .. 5 s + 4] (min x, min y, max x, max y, bad)
static BatchPlan plan_batch(const int *n, const int *T, const double *bb5, int nsets)
{
    BatchPlan b;
    b.sets.resize(nsets);
    long long poff = 0, nftot = 0, roff = 0;
    size_t toff = 0;
    std::vector<int> nfs(nsets);
    for (int s = 0; s < nsets; s++) {
        check_texture(T[s]);
        if (bb5[5 * s + 4] != 0.0) throw GalphaError("non-finite input coordinate");
        SnapGrid sg = snap_grid(bb5[5 * s], bb5[5 * s + 1], bb5[5 * s + 2], bb5[5 * s + 3], T[s]);
        std::vector<int> fp = frame_pixels(T[s]);
        SetDesc &d = b.sets[s];
        d.toff = toff;
        d.logT = log2i(T[s]);
        d.n = n[s];
        d.poff = (int)poff;
        d.nf = (int)fp.size();
        d.roff = (int)roff;
        // bands per row: 64 columns each, at most 64 (T / 64 is a power of two)
        d.m2 = d.m3 = std::max(1, std::min(64, T[s] / 64));
        d.coff = b.ncells;
        b.ncells += (T[s] / 16) * (T[s] / 16);
        b.max_m2 = std::max(b.max_m2, d.m2);
        d.margin = margin_for(T[s]);
        d.shx = sg.shx; d.shy = sg.shy; d.scale = sg.scale; d.inv = sg.inv;
        nfs[s] = d.nf;
        b.fpix.insert(b.fpix.end(), fp.begin(), fp.end());
        poff += n[s];
        nftot += d.nf;
        roff += T[s];
        toff += (size_t)T[s] * T[s];
    }
    check_slot_capacity(poff + nftot);
    b.ntot = (int)poff;
    b.V = (int)(poff + nftot);
    b.sumT = (int)roff;
    b.P = toff;
    long long fb = poff;
    for (int s = 0; s < nsets; s++) { b.sets[s].fbase = (int)fb; fb += nfs[s]; }
    b.bm.assign((size_t)BM_COUNT * (nsets + 1), 0);
    long long acc[BM_COUNT] = {0};
    for (int s = 0; s <= nsets; s++) {
        for (int k = 0; k < BM_COUNT; k++) {
            if (acc[k] > 0x7fffffffLL) throw GalphaError("batch too large (launch blocks)");
            b.bm[(size_t)k * (nsets + 1) + s] = (int)acc[k];
        }
        if (s == nsets) break;
        long long Ts = T[s], m = std::max(1, std::min(64, T[s] / 64));
        acc[BM_COLS] += Ts / 4;
        acc[BM_STACK] += (Ts / 64) * m;
        acc[BM_MERGE] += (Ts / 64) * std::max(1LL, m / 2);
        acc[BM_COLOR] += (Ts / 64) * m;
        acc[BM_PIX] += Ts * Ts / 256;
        acc[BM_RECON] += (Ts / 16) * (Ts / 16);   // one tile per 16 x 16-pixel cell
    }
    for (int k = 0; k < BM_COUNT; k++) b.blocks[k] = (int)acc[k];
    return b;
}

static void ensure_pba(galpha_ctx *c, const BatchPlan &b)
{
    size_t P = b.P;
    c->smap.ensure(4 * P);
    c->g.ensure(2 * P);
    c->S.ensure(4 * P);
    c->cnt.ensure(4 * (size_t)64 * b.sumT);
    c->plo.ensure(4 * (size_t)64 * b.sumT);
    c->E.ensure(4 * P);
    c->ecnt.ensure(4 * (size_t)b.sumT);
    c->vor.ensure(4 * P);
    c->bad.ensure(4 * (size_t)(b.V + 1));
    c->fpix.ensure(4 * (size_t)(b.V - b.ntot + 1));
    c->pix.ensure(4 * (size_t)(b.V + 1));
    c->vset.ensure(4 * (size_t)(b.V + 1));
    c->counters.ensure(COUNTER_BYTES);
    c->status.ensure(sizeof(int) * 4);
    size_t ntiles = (size_t)std::max(b.blocks[BM_RECON], 1);
    c->tileCnt.ensure(4 * ntiles);
    c->tileOff.ensure(4 * ntiles);
    c->scanTmp.ensure(scan_temp_bytes(std::max(ntiles, (size_t)(b.V + 2))) + 4096);
    c->setsD.ensure(sizeof(SetDesc) * b.sets.size());
    c->bmD.ensure(sizeof(int) * b.bm.size());
    int ns = b.nsets();
    c->hullCnt.ensure(4 * (size_t)ns + 16);
    c->hullOff.ensure(4 * (size_t)ns + 16);
    c->hullStart.ensure(4 * (size_t)ns + 16);
    c->setFirst.ensure(4 * (size_t)ns + 16);
}

// upload the plan; returns the device view
static BatchDev upload_plan(galpha_ctx *c, const BatchPlan &b)
{
    hipStream_t st = c->stream;
    HIPCHK(hipMemcpyAsync(c->setsD.p, b.sets.data(), sizeof(SetDesc) * b.sets.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c->bmD.p, b.bm.data(), sizeof(int) * b.bm.size(), hipMemcpyHostToDevice, st));
    if (!b.fpix.empty())
        HIPCHK(hipMemcpyAsync(c->fpix.p, b.fpix.data(), 4 * b.fpix.size(), hipMemcpyHostToDevice, st));
    BatchDev d{};
    d.sets = c->setsD.as<SetDesc>();
    d.nsets = b.nsets();
    d.bm = c->bmD.as<int>();
    for (int k = 0; k < BM_COUNT; k++) d.blocks[k] = b.blocks[k];
    d.vset = c->vset.as<int>();
    d.ntot = b.ntot;
    d.V = b.V;
    d.max_m2 = b.max_m2;
    d.P = b.P;
    return d;
}

// perm: input index of each real point when they were renumbered (NULL: input order)
static void run_pba(galpha_ctx *c, const double *d_xy, const BatchDev &bd, int *iters, int *island_sites,
                    const int *perm)
{
    hipStream_t st = c->stream;
    launch_fill_i32(c->smap.as<int>(), 0x7fffffff, bd.P, st);
    launch_vset(bd, c->vset.as<int>(), st);
    launch_splat((const double2 *)d_xy, bd, perm, c->smap.as<int>(), c->pix.as<int>(), c->fpix.as<int>(), st);
    {
        IntFills f;   // island marks, and the island counters of the first pass
        f.add(c->bad.p, (size_t)bd.V + 1, 0);
        f.add(c->counters.as<int>() + C_ISLAND_SITES, 1, 0);
        f.add(c->counters.as<int>() + C_FLAG, 1, 0);
        launch_fill_multi(f, st);
    }
    PbaBuffers b{c->smap.as<int>(), c->g.as<short>(), c->S.as<int>(), c->cnt.as<int>(), c->plo.as<int>(),
                 c->E.as<int>(), c->ecnt.as<int>(), c->vor.as<int>()};
    int it = 0, isl = 0;
    for (; it < 16; it++) {
        launch_pba(b, bd, st);
        if (it > 0) ctx_counter_set(c, C_FLAG, 0);
        launch_islands(c->vor.as<int>(), c->pix.as<int>(), bd, c->bad.as<int>(), c->counters.as<int>() + C_FLAG, st);
        int cv[C_ISLAND_SITES + 1];   // the flag and the island count in one read
        ctx_counters_read(c, cv, 0, C_ISLAND_SITES + 1);
        isl = cv[C_ISLAND_SITES];
        if (cv[C_FLAG] == 0) break;
        launch_remove_bad(c->bad.as<int>(), c->pix.as<int>(), bd, c->smap.as<int>(),
                          c->counters.as<int>() + C_ISLAND_SITES, st);
    }
    if (it == 16) throw GalphaError("discrete Voronoi diagram did not settle (island certificate)");
    *iters = it + 1;
    *island_sites = isl;
}

static float ev_ms(galpha_ctx *c, int a, int b)
{
    float ms = 0;
    if (c->ev[a] && c->ev[b]) hipEventElapsedTime(&ms, c->ev[a], c->ev[b]);
    return ms;
}

// stage boundary k on stream st (contexts created without events skip the timing)
static void ev_mark(galpha_ctx *c, int k, hipStream_t st)
{
    if (c->ev[k]) HIPCHK(hipEventRecord(c->ev[k], st));
}

// box of host points into bb5 (min x, min y, max x, max y, non-finite flag)
static void bbox_host(const double *xy, int n, double *bb5)
{
    double a = INFINITY, b = INFINITY, cc = -INFINITY, d = -INFINITY, bad = 0.0;
    for (int i = 0; i < n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        if (!std::isfinite(x) || !std::isfinite(y)) bad = 1.0;
        a = std::min(a, x); b = std::min(b, y); cc = std::max(cc, x); d = std::max(d, y);
    }
    bb5[0] = a; bb5[1] = b; bb5[2] = cc; bb5[3] = d; bb5[4] = bad;
}

void bbox_device(const double *d_xy, int n, double *out5, void *tmp, hipStream_t st);
void bbox_sets_device(const double *d_xy, const int *d_offs, int nsets, double *out5, void *tmp, hipStream_t st);

// Bulk frame removal: when the real region of every set is bounded by one simple cycle through all
// of its meshed points, drop all frame triangles and convexify each cycle by parallel ear rounds.
// Returns false (mesh untouched) when some set does not qualify.
// On success *ear_rounds = the rounds run (-1: read later) and *frame_list = the length of the
// list of triangles it created (-1: not read).
static bool bulk_frame_removal(galpha_ctx *c, MeshDev &m, int nsets, hipStream_t st, int *ear_rounds,
                               long long *frame_list)
{
    *frame_list = -1;
    int V = m.V, n = m.n;
    int *outdeg = m.decide, *cover = m.deg, *bidx = c->fill.as<int>(), *blist = c->mlist.as<int>();
    c->badset.ensure(4 * (size_t)nsets + 16);
    int *badset = c->badset.as<int>(), *hcnt = c->hullCnt.as<int>(), *hstart = c->hullStart.as<int>();
    {
        IntFills f;
        f.add(m.hnext, (size_t)V, -1);
        f.add(outdeg, (size_t)V, 0);
        f.add(cover, (size_t)V, 0);
        f.add(bidx, (size_t)V, -1);
        f.add(hcnt, (size_t)nsets, 0);
        f.add(badset, (size_t)nsets, 0);
        f.add(hstart, (size_t)nsets, 0x7fffffff);
        f.add(c->counters.as<int>() + C_BND, 1, 0);
        launch_fill_multi(f, st);
    }
    launch_bnd_detect(m, outdeg, cover, hcnt, hstart, badset, blist, bidx, st);
    std::vector<int> rd(2 * (size_t)nsets + 1);
    ctx_read_segments(c, {{badset, nsets}, {hcnt, nsets}, {c->counters.as<int>() + C_BND, 1}}, rd.data());
    std::vector<int> bad(rd.begin(), rd.begin() + nsets), cnt(rd.begin() + nsets, rd.begin() + 2 * nsets);
    int B = rd[2 * (size_t)nsets];
    long long tot = 0;
    int maxh = 0;
    std::vector<int> off(nsets);
    for (int s = 0; s < nsets; s++) {
        // a set needs a cycle of at least 3 vertices; very long cycles make the ear test's
        // per-candidate scan too long, so those sets keep the round-by-round removal
        if (bad[s] || cnt[s] < 3 || cnt[s] > 65536) return false;
        off[s] = (int)tot;
        tot += cnt[s];
        maxh = std::max(maxh, cnt[s]);
    }
    if (tot != B) return false;
    (void)n;
    // cycle arrays, the cycle's positions, 5 doubles of grid parameters per set, then two arrays of
    // claim keys
    c->bnd.ensure(4 * (size_t)((11 * (size_t)B + 3) & ~(size_t)3) + 32 * (size_t)B + 40 * (size_t)nsets + 64);
    HIPCHK(hipMemcpyAsync(c->hullOff.p, off.data(), 4 * (size_t)nsets, hipMemcpyHostToDevice, st));
    int rounds = 1;
    while ((1 << rounds) < maxh) rounds++;
    rounds += 1;
    launch_bnd_rank(m, B, rounds, blist, bidx, hstart, hcnt, c->hullOff.as<int>(), badset, c->bnd.as<int>(), st);
    ctx_read_segments(c, {{badset, nsets}}, rd.data());
    for (int s = 0; s < nsets; s++)
        if (rd[s]) return false;
    // from here on the mesh changes
    launch_kill_frame(m, st);
    int *bnd = c->bnd.as<int>(), *hoff = c->hullOff.as<int>();
    if (nsets >= 32) {
        // many sets: every set's rounds in one workgroup, one launch (the most rounds a set ran is
        // read with the next snapshot)
        launch_reset_counters(c->counters.as<int>(), C_EAR, -1, -1, -1, st);
        launch_ear_sets(m, B, bnd, hcnt, hoff, nsets, c->modF.as<int>(), C_FRAME_LIST, c->op++, st);
        *ear_rounds = -1;
        return true;
    }
    // few sets: grid-wide rounds, eight per host read (a round after one without candidates
    // returns at once).  Sets of 256K points and more take 20-24 rounds (1M uniform points,
    // the 2M lattice + cluster set), so their first read comes after sixteen.
    launch_ear_grid(m, B, bnd, hcnt, hoff, nsets, st);   // also zeroes counters[C_EAR]
    int round = 0, per_read = m.n >= (1 << 18) ? 16 : 8;
    for (;; round += per_read, per_read = 8) {
        if (round > 200000) throw GalphaError("hull convexification did not finish");
        launch_ear_rounds(m, B, bnd, hcnt, hoff, nsets, round, per_read, c->modF.as<int>(), C_FRAME_LIST, c->op++,
                          st);
        int cv[C_EAR_PREV - C_FRAME_LIST + 1];
        ctx_counters_read(c, cv, C_FRAME_LIST, C_EAR_PREV - C_FRAME_LIST + 1);
        if (cv[(((round + per_read - 1) & 1) ? C_EAR_PREV : C_EAR) - C_FRAME_LIST] == 0) {
            *frame_list = cv[0];
            break;
        }
    }
    *ear_rounds = round + per_read;
    return true;
}

// Real points are renumbered in spatial order (set, then 16 x 16-pixel cells in Morton order;
// any order inside a cell), so that vertex-indexed reads of neighbouring vertices share cache
// lines.  Returns the coordinates in that order; c->perm maps the new index to the input index,
// which the rules that depend on point order use (a pixel's site, insertion claims, the
// symbolic tie-break) and the output reports.
static const double *renumber_points(galpha_ctx *c, const double *d_xy, const BatchDev &bd, const BatchPlan &plan)
{
    hipStream_t st = c->stream;
    int n = bd.ntot;
    size_t nc = (size_t)plan.ncells + 1;
    c->perm.ensure(4 * (size_t)n + 64);
    c->xys.ensure(16 * (size_t)n + 64);
    c->pcell.ensure(4 * (size_t)n + 64);
    c->ccount.ensure(4 * nc + 64);
    c->ccursor.ensure(4 * nc + 64);
    c->scanTmp.ensure(scan_temp_bytes(nc) + 4096);
    launch_vset(bd, c->vset.as<int>(), st);
    HIPCHK(hipMemsetAsync(c->ccount.p, 0, 4 * nc, st));
    launch_renumber((const double2 *)d_xy, bd, c->pcell.as<int>(), c->ccount.as<int>(), c->ccursor.as<int>(),
                    plan.ncells, c->scanTmp.p, c->perm.as<int>(), c->xys.as<double2>(), st);
    return c->xys.as<double>();
}

// Delaunay triangulation of every set of the plan (points in d_xy, concatenated in set order).
// Leaves the canonical triangles of all sets in c->out9 (9 ints each, global vertex ids and
// global triangle positions, set-major) and the first triangle of each set in c->set_first.
static int run_batch(galpha_ctx *c, const double *d_xy, const BatchPlan &plan)
{
    hipStream_t st = c->stream;
    galpha_stats &S = c->stats;
    memset(&S, 0, sizeof(S));
    int nsets = plan.nsets();
    c->last_F = 0;
    c->set_first.assign(nsets + 1, 0);
    c->plan_sets = plan.sets;
    c->plan_ntot = plan.ntot;
    if (nsets == 0) return 0;
    int V = plan.V, n = plan.ntot;
    ensure_pba(c, plan);
    {
        IntFills f;
        f.add(c->counters.p, COUNTER_BYTES / sizeof(int), 0);
        f.add(c->status.p, 4, 0);
        launch_fill_multi(f, st);
    }
    BatchDev bd = upload_plan(c, plan);
    ev_mark(c, 0, st);
    d_xy = renumber_points(c, d_xy, bd, plan);
    run_pba(c, d_xy, bd, &S.pbaIterations, &S.islandSites, c->perm.as<int>());
    ev_mark(c, 1, st);

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
    c->fan.ensure(8 * 3 * (size_t)cap);
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
    c->lmark.ensure(4 * (size_t)cap);
    c->ffan.ensure(8 * (size_t)(V - n) + 16);
    c->modF.ensure(4 * (size_t)cap);
    // Slot tags (stamp, claim keys, list marks) carry the context's op number, and the removal
    // rounds (deletion, frame) tag their per-vertex priority maxima with a round number; both
    // only grow, across runs too, so values left by earlier runs never match or beat later ones.
    // The arrays are cleared only when their memory is new (or carved from a shared scratch
    // region) and when the numbers would run out.
    {
        IntFills f;
        const bool arena = galpha_arena_active();
        if (arena || c->op > (1 << 30) || c->tag_p[0] != c->stamp.p || c->tag_p[1] != c->claim.p ||
            c->tag_p[2] != c->lmark.p) {
            f.add(c->stamp.p, c->stamp.bytes / 4, 0);
            f.add(c->claim.p, c->claim.bytes / 4, 0);
            f.add(c->lmark.p, c->lmark.bytes / 4, 0);
            c->tag_p[0] = c->stamp.p; c->tag_p[1] = c->claim.p; c->tag_p[2] = c->lmark.p;
            c->op = 1;
        }
        if (arena || c->prio_round > (1 << 19) || c->mx_p != c->mx.p) {
            f.add(c->mx.p, c->mx.bytes / 4, 0);
            c->mx_p = c->mx.p;
            c->prio_round = 0;
        }
        f.add(c->vt.p, (size_t)V, -1);
        launch_fill_multi(f, st);
    }
    auto next_prio_round = [&]() {
        if (++c->prio_round >= (1 << 20)) throw GalphaError("too many removal rounds");
        return c->prio_round;
    };

    MeshDev m{};
    m.tv = c->tv.as<int4>(); m.tn = c->tn.as<int4>(); m.remap = c->remap.as<int4>();
    m.stamp = c->stamp.as<int>(); m.claim = c->claim.as<unsigned long long>();
    m.cap = cap; m.ntri = c->counters.as<int>() + 40;   // slot 40 holds the slot count
    m.pos = c->pos.as<double2>(); m.xy = (const double2 *)d_xy; m.vt = c->vt.as<int>();
    m.perm = c->perm.as<int>();
    m.vstate = c->vstate.as<int>(); m.pix = c->pix.as<int>();
    m.n = n; m.V = V;
    m.sets = bd.sets; m.vset = bd.vset; m.nsets = nsets;
    m.counters = c->counters.as<int>();
    m.status = m.counters + C_STATUS;
    m.stats = m.counters + GA_STRIPE_INTS;
    m.modA = c->modA.as<int>(); m.modB = c->modB.as<int>(); m.cand = c->cand.as<int>();
    m.scratch = c->scratch.as<int>(); m.scratch_cap = scap;
    c->dbg.ensure(64 * sizeof(double));
    m.dbg = c->dbg.as<double>();
    m.deg = c->deg.as<int>();
    m.decide = c->decide.as<int>();
    m.dellist = c->dellist.as<int>();
    m.lt = c->lt.as<int>();
    m.hnext = c->hnext.as<int>();
    m.hedge = c->hedge.as<int>();
    m.lmark = c->lmark.as<int>();
    m.ffan = c->ffan.as<int>();

    // ---- reconstruction + adjacency
    launch_recon(c->vor.as<int>(), c->pix.as<int>(), bd, m, c->tileCnt.as<int>(), c->tileOff.as<int>(),
                 c->scanTmp.p, st);
    launch_adjacency(m, c->deg.as<int>(), c->off.as<int>(), c->fill.as<int>(), c->fan.as<int2>(), c->scanTmp.p, st);
    launch_init_positions(m, st);
    {
        int cv[C_NCOUNTERS];   // slot count (40) and status word
        ctx_counters_read(c, cv, 0, C_NCOUNTERS);
        S.reconTris = cv[40];
        if (cv[40] > cap) throw GalphaError("triangle capacity exceeded in reconstruction");
        if (cv[C_STATUS])
            throw GalphaError("reconstruction produced an invalid triangle (status " + std::to_string(cv[C_STATUS]) + ")");
    }
    ev_mark(c, 2, st);
    // ---- shifting
    // revert moved vertices of non-ccw triangles until every triangle is ccw (monotone)
    // Round loops launch several rounds per host read of their counters: a round after
    // convergence finds nothing to do, so over-running is harmless.
    {
        // worklists in the flip lists (free until the flips), marks in decide[] (free until the
        // greedy rounds)
        // Lists alternate between two buffers; their counts rotate through three slots, so each
        // pass zeroes the next pass's output count (the slot the pass before it read) itself.
        unsigned char *moved = c->moved.as<unsigned char>();
        int *wl[2] = {m.modA, m.modB};
        const int slot[3] = {C_MODN, C_MODN2, C_MODN3};
        launch_shift_bulk(m, moved, m.decide, wl[0], slot[0], slot[1], st);   // (zeroes decide[])
        for (int it = 0;; it += 4) {
            if (it > 100000) throw GalphaError("bulk shifting did not converge");
            for (int k = it; k < it + 4; k++)
                launch_shift_star(m, moved, m.decide, wl[k & 1], slot[k % 3], wl[(k + 1) & 1], slot[(k + 1) % 3],
                                  slot[(k + 2) % 3], st);
            if (ctx_counter(c, slot[(it + 4) % 3]) == 0) break;
        }
    }
    launch_shift_commit(m, c->moved.as<unsigned char>(), st);
    // the missing-site list is free until the missing stage: it holds the pending vertices
    m.plist = c->mlist.as<int>();
    launch_pending_list(m, st);
    int pend = ctx_stat(c, C_PEND);
    S.shiftBad = pend;
    int moved_before = 0;
    // Eight greedy rounds (two batches of four per host read), then deletion takes the rest.
    // Later rounds move few vertices each and cost more than deleting and reinserting them:
    // on 1M uniform points the shifting stage took 1.68 ms with up to 48 rounds, 1.40 ms with 8.
    const int pgrid = round_grid(pend);   // the pending list keeps its length (done entries are skipped)
    // (all eight before one read: a round after convergence moves nothing)
    // Sets of 256K points and more always leave vertices to delete after the greedy rounds, so
    // their moved count comes with the deletion loop's first snapshot instead of a read of its own.
    const bool large = m.n >= (1 << 18);
    if (pend > 0) {
        for (int k = 0; k < 8; k++) launch_shift_greedy(m, k, pgrid, st);
        if (!large) {
            int moved_total = ctx_stat(c, C_MOVED);
            pend -= moved_total - moved_before;
            moved_before = moved_total;
        }
    }
    int deleted_before = 0;
    // three rounds per read; sets of 256K points and more start with six (1M uniform points take
    // six), saving a read
    for (int r = 0, per = large ? 6 : 3; pend > 0; r += per, per = 3) {
        if (r > 10000) throw GalphaError("vertex deletion did not finish");
        for (int k = 0; k < per; k++) {
            // (k_del_prio zeroes C_SCR and C_MODN; rounds with nothing pending return at once)
            launch_delete_round(m, c->mx.as<unsigned int>(), next_prio_round(), c->op++, pgrid, st);
        }
        Snap sn = ctx_snapshot(c);
        int after = sn.stat(C_DELETED);
        const int moved_total = sn.stat(C_MOVED);
        pend -= (after - deleted_before) + (moved_total - moved_before);
        moved_before = moved_total;
        deleted_before = after;
        int stt = sn.status();
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
    S.deleted = deleted_before;
    ev_mark(c, 3, st);

    // Flip helpers: Lawson passes over the list of triangles to test; a pass with no flip has no candidate
    // left (the highest candidate always wins), so three passes run per host read.
    // `last` keeps the snapshot of the final check (counters, statistics and status after the
    // passes), which the callers use instead of another read
    Snap last;
    // `hint`: the length of the first list when the caller knows it (else a full grid); later
    // passes size their grids by the list length of the last read (lists shrink from pass to pass).
    // Lists of at most TAIL_LIST triangles go to the single-workgroup tail, which runs passes
    // back to back without launches or host reads until one flips nothing.
    auto flip_until_done = [&](int *listA, int cntA, int *listB, int cntB, int prev, long long hint) {
        int *lin = listA, *lout = listB, sin = cntA, sout = cntB;
        int pass_count = 0;
        int grid = hint >= 0 ? round_grid(2 * hint) : 2048;
        long long next = hint;   // length of the next pass's list, when known
        launch_reset_counters(c->counters.as<int>(), C_CAND, -1, -1, -1, st);
        const int maxp = 512;
        // the tail's passes after a read: returns true when the flips are done
        auto tail_result = [&](int op0) {
            const int p = last.v[C_TAIL_PASSES];
            c->op = op0 + p;   // tags of passes that did not run are not used
            S.flipPasses += p;
            S.flipTailPasses += p;
            pass_count += p;
            if (p & 1) {
                std::swap(lin, lout);
                std::swap(sin, sout);
            }
            if (p > 0) prev = op0 + p - 1;
            return last.v[C_TAIL_DONE] != 0;
        };
        for (pass_count = 0; pass_count < 100000;) {
            if (next >= 0 && next <= TAIL_LIST) {
                const int op0 = c->op;
                c->op += maxp;
                launch_flip_tail(m, lin, sin, lout, sout, prev, op0, maxp, st);
                last = ctx_snapshot(c);
                const int p = last.v[C_TAIL_PASSES];
                if (tail_result(op0)) return;
                next = last.v[sin];
                grid = round_grid(2LL * next);
                if (p == maxp && next <= TAIL_LIST) continue;
                next = TAIL_LIST + 1;   // the list outgrew the tail: grid passes first
            }
            for (int k = 0; k < 3; k++) {
                int opn = c->op++;
                launch_flip_pass(m, lin, sin, lout, sout, prev, opn, grid, st);
                prev = opn;
                std::swap(lin, lout);
                std::swap(sin, sout);
            }
            S.flipPasses += 3;
            pass_count += 3;
            // the tail right behind the grid passes, before the read: it returns at once while the
            // list is longer than TAIL_LIST (no passes), else it usually finishes the flips
            const int op0 = c->op;
            c->op += maxp;
            launch_flip_tail(m, lin, sin, lout, sout, prev, op0, maxp, st);
            last = ctx_snapshot(c);
            const int p = last.v[C_TAIL_PASSES];
            if (p == 0) {
                c->op = op0;
                if (last.v[C_FLIPS] == 0) return;
                next = last.v[sin];
            } else {
                if (tail_result(op0)) return;
                next = last.v[sin];
                if (p == maxp && next <= TAIL_LIST) continue;
                next = TAIL_LIST + 1;
            }
            grid = round_grid(2LL * next);
        }
        throw GalphaError("flipping did not converge");
    };
    auto flip_all = [&]() {
        int opn = c->op++;
        ctx_counter_set(c, C_MODN, 0);
        launch_flip_all(m, opn, m.modA, st);
        flip_until_done(m.modA, C_MODN, m.modB, C_MODN2, opn, -1);
    };
    // ---- missing sites
    // Inserted into the mesh as it is after shifting (every triangle counter-clockwise, not yet
    // Delaunay): the walks and splits only need a valid triangulation, and one flip phase over
    // all triangles afterwards makes the mesh Delaunay, instead of flip passes after every
    // insertion round.
    // (C_MISS is zero since the run's start: only the missing stage counts it)
    launch_missing_init(m, c->vor.as<int>(), c->mlist.as<int>(), c->loc.as<int>(), st);
    last = ctx_snapshot(c);
    int nmiss = last.v[C_MISS];
    S.missing = nmiss;
    for (int r = 0; r < 100000; r++) {
        // `last`: the snapshot after the previous round (or after the list was built)
        int left = nmiss - last.stat(C_INSERTED) - last.stat(C_DUPS);
        if (left <= 0) break;
        if (last.status())
            throw GalphaError("missing-site insertion failed (status " + std::to_string(last.status()) + ")");
        if (r > 5000) throw GalphaError("missing-site insertion did not finish");
        // two rounds per read (a round after the last insertion finds every entry done); sets of
        // 256K points and more start with four (1M uniform points take four), saving a read
        const int per = r == 0 && m.n >= (1 << 18) ? 4 : 2;
        for (int k = 0; k < per; k++) {
            int opn = c->op++;
            launch_insert_round(m, c->mlist.as<int>(), c->loc.as<int>(), opn, round_grid(nmiss), st);   // (zeroes C_MODN)
            launch_relink(m, m.modA, C_MODN, opn, round_grid(4LL * left), st);
            S.insertRounds++;
        }
        last = ctx_snapshot(c);
    }
    S.inserted = last.stat(C_INSERTED);
    S.duplicates = last.stat(C_DUPS);
    ev_mark(c, 4, st);

    // ---- flips to a Delaunay mesh of the current vertices
    flip_all();
    ev_mark(c, 5, st);
    const Snap after_main = last;   // per-stage statistics: shifting, insertion and these flips

    // ---- frame removal + hull convexification
    // Triangles created by frame removal and convexification go to one list (each slot once);
    // the mesh was Delaunay before, so only their edges can need flips afterwards.
    m.list_op = c->op++;
    launch_reset_counters(c->counters.as<int>(), C_FRAME_LIST, -1, -1, -1, st);
    long long frame_list = -1;   // triangles listed by frame removal, when known
    if (bulk_frame_removal(c, m, nsets, st, &S.earRounds, &frame_list)) {
        S.frameBulk = 1;
    } else {
    for (int r = 0;; r += 2) {
        if (r > 100000) throw GalphaError("frame removal did not finish");
        for (int k = 0; k < 2; k++) {
            launch_reset_counters(c->counters.as<int>(), C_SCR, -1, -1, -1, st);
            launch_frame_round(m, c->mx.as<unsigned int>(), next_prio_round(), c->modF.as<int>(), C_FRAME_LIST,
                               c->op++, st);
        }
        int left = (V - n) - ctx_stat(c, C_FRAME_DONE);
        if (left <= 0) break;
    }
    launch_convexify(m, c->modF.as<int>(), C_FRAME_LIST, c->op++, c->hullCnt.as<int>(), c->hullOff.as<int>(),
                     c->hullStart.as<int>(), c->scanTmp.p, st);
    }
    ev_mark(c, 6, st);
    flip_until_done(c->modF.as<int>(), C_FRAME_LIST, m.modB, C_MODN2, m.list_op, frame_list);
    ev_mark(c, 7, st);
    const Snap after_frame = last;   // per-stage statistics: frame removal and the final flips
    if (S.frameBulk && S.earRounds < 0) S.earRounds = after_frame.v[C_EAR];
    if (last.status()) throw GalphaError("device pipeline error (status " + std::to_string(last.status()) + ")");

    // ---- constraint edges (single-set runs): cavity insertion rounds, then constrained flips
    c->cons_rounds = 0;
    if (!c->cons.empty()) {
        if (nsets != 1) throw GalphaError("constraints apply to single-set runs only");
        int nc = (int)c->cons.size() / 2;
        int hs = 64;
        while (hs < 2 * nc + 64) hs *= 2;
        c->kin.ensure(8 * (size_t)nc + 64);
        c->kv.ensure(8 * (size_t)nc + 64);
        c->kinv.ensure(4 * (size_t)n + 64);
        c->kdone.ensure(4 * (size_t)nc + 64);
        c->kcnt.ensure(4 * (size_t)nc + 64);
        c->krs.ensure(4 * (size_t)nc + 64);
        c->kroff.ensure(4 * (size_t)nc + 64);
        c->khash.ensure(8 * (size_t)hs);
        c->scanTmp.ensure(scan_temp_bytes((size_t)nc) + 256);
        stage_h2d(c->stage, c->kin.p, c->cons.data(), 8 * (size_t)nc, st);
        launch_fill_i32(c->khash.as<int>(), 0, 2 * (size_t)hs, st);
        m.chash = c->khash.as<unsigned long long>();
        m.chmask = hs - 1;
        launch_cons_setup(m, c->kin.as<int2>(), nc, c->kinv.as<int>(), c->kv.as<int2>(), c->kdone.as<int>(),
                          c->khash.as<unsigned long long>(), hs - 1, st);
        for (int r = 0;; r++) {
            if (r > nc + 16) throw GalphaError("constraint insertion did not finish");
            launch_reset_counters(c->counters.as<int>(), C_CONS_PEND, C_CONS_MOD, -1, -1, st);
            launch_cons_count(m, c->kv.as<int2>(), nc, c->kdone.as<int>(), c->kcnt.as<int>(), c->krs.as<int>(), st);
            launch_exclusive_scan(c->krs.as<int>(), c->kroff.as<int>(), (size_t)nc, c->scanTmp.p, st);
            int tail[2];
            stage_d2h(c->stage, &tail[0], c->kroff.as<int>() + nc - 1, 4, st);
            stage_d2h(c->stage, &tail[1], c->krs.as<int>() + nc - 1, 4, st);
            last = ctx_snapshot(c);
            if (int s = last.status()) {
                const char *why = (s & GA_ERR_CONS_POINT) ? "a point lies inside a constraint segment"
                                  : (s & GA_ERR_CONS_CROSS) ? "two constraints cross"
                                  : (s & GA_ERR_CONS_DUP)   ? "a constraint endpoint is a duplicate point"
                                                            : "constraint insertion failed";
                throw GalphaError(std::string(why) + " (status " + std::to_string(s) + ")");
            }
            if (last.v[C_CONS_PEND] == 0) break;
            long long words = (long long)tail[0] + tail[1];
            if (words <= 0 || words > (1LL << 31) - 64) throw GalphaError("constraint work space out of range");
            c->kwork.ensure(4 * (size_t)words + 64);
            launch_cons_insert(m, c->kv.as<int2>(), nc, c->kdone.as<int>(), c->kcnt.as<int>(), c->kroff.as<int>(),
                               c->kwork.as<int>(), m.modA, c->op++, st);
            c->cons_rounds++;
        }
        flip_all();
        if (last.status()) throw GalphaError("constrained flips failed (status " + std::to_string(last.status()) + ")");
    }
    ev_mark(c, 9, st);

    // ---- canonical output
    int nslots = std::min(last.v[40], cap);
    c->keys.ensure(8 * (size_t)nslots + 64);
    c->keys2.ensure(8 * (size_t)nslots + 64);
    c->vals.ensure(4 * (size_t)nslots + 64);
    c->vals2.ensure(4 * (size_t)nslots + 64);
    c->rot.ensure(4 * (size_t)nslots + 64);
    c->newid.ensure(4 * (size_t)nslots + 64);
    // one key (v0, v1) of 2B bits, input indices below 2^B (a batch's sets are consecutive ranges);
    // up to B = 20 a 32-bit key (v0, top bits of v1) whose rare ties k_canon_newid orders
    int kb = 2;
    while ((1LL << kb) < (long long)n) kb++;
    if (2 * kb > 64) throw GalphaError("too many points for the canonical key");
    int *vals = c->vals.as<int>(), *vals2 = c->vals2.as<int>();
    const unsigned *skey32 = nullptr, *v1s = nullptr;
    size_t tb = 0;
    if (kb <= CANON_KEY32_BITS) {
        unsigned *K = c->keys.as<unsigned>(), *Ko = K + nslots, *V1 = c->keys2.as<unsigned>();
        launch_canon_keys32(m, K, V1, vals, c->rot.as<int>(), kb, st);
        const int bits = std::min(32, 2 * kb);
        sort_pairs_u32(K, vals, Ko, vals2, nslots, 0, bits, nullptr, &tb, st);
        c->sortTmp.ensure(tb + 256);
        tb = c->sortTmp.bytes;
        if (sort_pairs_u32(K, vals, Ko, vals2, nslots, 0, bits, c->sortTmp.p, &tb, st))
            throw GalphaError("radix sort failed");
        skey32 = Ko;
        v1s = V1;
    } else {
        unsigned long long *K = c->keys.as<unsigned long long>(), *K2 = c->keys2.as<unsigned long long>();
        launch_canon_keys(m, K, vals, c->rot.as<int>(), kb, st);
        sort_pairs_u64(K, vals, K2, vals2, nslots, 0, 2 * kb, nullptr, &tb, st);
        c->sortTmp.ensure(tb + 256);
        tb = c->sortTmp.bytes;
        if (sort_pairs_u64(K, vals, K2, vals2, nslots, 0, 2 * kb, c->sortTmp.p, &tb, st))
            throw GalphaError("radix sort failed");
    }
    int *order = vals2;
    // the live-triangle count stays on the device (folded statistic) until the final read
    c->out9.ensure(36 * (size_t)nslots + 64);
    launch_fold_stats(c->counters.as<int>(), st);
    launch_canon_emit(m, order, skey32, v1s, c->rot.as<int>(), c->newid.as<int>(),
                      c->counters.as<int>() + C_FOLD + C_RECON_TRIS, c->out9.as<int>(), c->setFirst.as<int>(),
                      c->keys2.as<int>(), st);   // keys2 (v1s) is free after k_canon_newid: output sets
    ev_mark(c, 8, st);
    Snap fin;
    std::vector<int> sfv((size_t)nsets + sizeof(fin.v) / 4);   // first triangle of every set, then the counters
    ctx_read_segments(c, {{c->setFirst.as<int>(), nsets}, {c->counters.as<int>(), (int)(sizeof(fin.v) / 4)}},
                      sfv.data());
    memcpy(fin.v, sfv.data() + nsets, sizeof(fin.v));
    const int *sf = sfv.data();
    if (fin.status()) throw GalphaError("canonical output failed (status " + std::to_string(fin.status()) + ")");
    int F = fin.stat(C_RECON_TRIS);
    // sets without triangles start where the next set starts
    c->set_first[nsets] = F;
    for (int s = nsets - 1; s >= 0; s--) c->set_first[s] = sf[s] >= 0 ? sf[s] : c->set_first[s + 1];
    c->last_F = F;
    S.nTris = F;
    S.flips = fin.stat(C_FLIPS_TOTAL);
    S.frameSeen = fin.stat(C_FRAME_SEEN);
    S.frameChain = fin.stat(C_FRAME_CHAIN);
    S.flipListed = fin.v[C_FLIP_LISTED];
    S.flipTailListed = fin.v[C_TAIL_LISTED];
    S.incircleTests = fin.stat(C_TESTS_INCIRCLE);
    S.incircleExact = fin.stat(C_EXACT_INCIRCLE);
    S.orientExact = fin.stat(C_EXACT_ORIENT);
    S.ms_pba = ev_ms(c, 0, 1);
    S.ms_recon = ev_ms(c, 1, 2);
    S.ms_shift = ev_ms(c, 2, 3);
    S.ms_missing = ev_ms(c, 3, 4);
    S.ms_boundary = ev_ms(c, 5, 6);
    S.ms_flip = ev_ms(c, 4, 5) + ev_ms(c, 6, 7);
    {
        const Snap *cut[4] = {nullptr, &after_main, &after_frame, &fin};
        for (int k = 0; k < 3; k++) {
            auto d = [&](bool striped, int slot) -> long long {
                long long b = striped ? cut[k + 1]->stat(slot) : cut[k + 1]->v[slot];
                long long a = cut[k] ? (striped ? cut[k]->stat(slot) : cut[k]->v[slot]) : 0;
                return b - a;
            };
            S.stageListed[k] = d(false, C_FLIP_LISTED);
            S.stageIncircleTests[k] = d(true, C_TESTS_INCIRCLE);
            S.stageIncircleExact[k] = d(true, C_EXACT_INCIRCLE);
            S.stageOrientExact[k] = d(true, C_EXACT_ORIENT);
        }
    }
    S.ms_constraints = ev_ms(c, 7, 9);
    S.ms_canon = ev_ms(c, 9, 8);
    S.constraintRounds = c->cons_rounds;
    S.ms_total = ev_ms(c, 0, 8);
    return F;
}

// one set as a batch of one (fewer than 3 points: no triangles)
static int run_dt(galpha_ctx *c, const double *d_xy, int n, int T, const double *bb5)
{
    check_texture(T);
    for (size_t k = 0; k < c->cons.size(); k += 2) {
        int a = c->cons[k], b = c->cons[k + 1];
        if (a < 0 || a >= n || b < 0 || b >= n)
            throw GalphaError("constraint " + std::to_string(k / 2) + ": endpoint outside [0, " + std::to_string(n) + ")");
        if (a == b) throw GalphaError("constraint " + std::to_string(k / 2) + ": both endpoints are point " + std::to_string(a));
    }
    if (!c->cons.empty() && n < 3) throw GalphaError("constraints need at least 3 points");
    c->last_n = n;
    c->last_T = T;
    if (n < 3) {
        memset(&c->stats, 0, sizeof(c->stats));
        c->last_F = 0;
        c->set_first.assign(1, 0);
        c->plan_sets.clear();
        c->plan_ntot = 0;
        return 0;
    }
    BatchPlan plan = plan_batch(&n, &T, bb5, 1);
    return run_batch(c, d_xy, plan);
}

// the last plan as the alpha stage sees it
static AlphaBatchInfo batch_info(galpha_ctx *c)
{
    AlphaBatchInfo info;
    int ns = (int)c->plan_sets.size();
    info.nsets = std::max(ns, 1);
    info.vset = ns > 1 ? c->vset.as<int>() : nullptr;
    info.poff.resize(info.nsets + 1, 0);
    info.tfirst.resize(info.nsets + 1, 0);
    for (int s = 0; s < ns; s++) info.poff[s] = c->plan_sets[s].poff;
    info.poff[info.nsets] = c->plan_ntot;
    for (int s = 0; s <= ns; s++) info.tfirst[s] = c->set_first[s];
    if (ns == 0) info.tfirst[1] = 0;
    return info;
}

// Runs the pipeline over nsets host point sets (xy concatenated, offsets[nsets + 1] in points,
// fboSizes[s] <= 0 or NULL: autotuned side).  Sets with fewer than 3 points are left out of the
// plan (no triangles); c->batch_sets maps planned sets to input sets.  Returns the total number
// of triangles.
static int batch_run(galpha_ctx *c, const double *xy, const int *offsets, int nsets, const int *fboSizes,
                     int with_alpha, bool device_input = false)
{
    hipStream_t st = c->stream;
    if (!c->cons.empty()) throw GalphaError("constraints apply to single-set runs (clear them with galphaSetConstraints)");
    c->batch_sets.clear();
    c->batch_npts.assign(nsets, 0);
    std::vector<int> ns, Ts;
    long long total = 0;
    for (int s = 0; s < nsets; s++) {
        int n = offsets[s + 1] - offsets[s];
        if (n < 0) throw GalphaError("offsets must be non-decreasing");
        c->batch_npts[s] = n;
        if (n < 3) continue;
        if (!xy) throw GalphaError("null point array");
        int T = (fboSizes && fboSizes[s] > 0) ? fboSizes[s] : texture_for_density(n, c->density[tune_class(n)]);
        c->batch_sets.push_back(s);
        ns.push_back(n);
        Ts.push_back(T);
        total += n;
    }
    if (total > 0x7fffffffLL / 8) throw GalphaError("batch too large");
    {
        long long V = total;
        for (int T : Ts) V += frame_count(T);
        check_slot_capacity(V);
    }
    int na = (int)ns.size();
    c->xy.ensure(16 * (size_t)std::max<long long>(total, 1));
    std::vector<int> poffs(na + 1, 0);
    for (int k = 0; k < na; k++) poffs[k + 1] = poffs[k] + ns[k];
    hipMemcpyKind kind = device_input ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const double *dx = nullptr;   // device points of the planned sets, contiguous
    if (device_input && na == nsets && nsets > 0) {
        dx = xy + 2 * (size_t)offsets[0];
    } else if (na == nsets && nsets > 0) {
        if (device_input)
            HIPCHK(hipMemcpyAsync(c->xy.p, xy + 2 * (size_t)offsets[0], 16 * (size_t)total, kind, st));
        else
            stage_h2d(c->stage, c->xy.p, xy + 2 * (size_t)offsets[0], 16 * (size_t)total, st);
        dx = c->xy.as<double>();
    } else {
        for (int k = 0; k < na; k++) {
            double *dk = c->xy.as<double>() + 2 * (size_t)poffs[k];
            const double *sk = xy + 2 * (size_t)offsets[c->batch_sets[k]];
            if (device_input) HIPCHK(hipMemcpyAsync(dk, sk, 16 * (size_t)ns[k], kind, st));
            else stage_h2d(c->stage, dk, sk, 16 * (size_t)ns[k], st);
        }
        dx = c->xy.as<double>();
    }
    std::vector<double> bb5((size_t)5 * std::max(na, 1));
    if (na > 0) {
        c->offsD.ensure(4 * (size_t)(na + 1));
        c->bboxTmp.ensure(8 * 5 * (size_t)na);
        HIPCHK(hipMemcpyAsync(c->offsD.p, poffs.data(), 4 * (size_t)(na + 1), hipMemcpyHostToDevice, st));
        bbox_sets_device(dx, c->offsD.as<int>(), na, bb5.data(), c->bboxTmp.p, st);
    }
    BatchPlan plan = plan_batch(ns.data(), Ts.data(), bb5.data(), na);
    int F = run_batch(c, dx, plan);
    c->batch_xy = dx;
    c->last_n = (int)total;
    c->last_alpha = false;
    if (with_alpha) {
        auto t0 = std::chrono::steady_clock::now();
        alpha_run_batch(c->alpha, dx, plan.ntot, c->out9.as<int>(), F, batch_info(c), st, &c->stats);
        c->last_alpha = true;
        HIPCHK(hipStreamSynchronize(st));
        auto t1 = std::chrono::steady_clock::now();
        c->stats.ms_alpha = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
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
        HIPCHK(hipHostMalloc((void **)&c->hsnap, sizeof(int) * SNAP_INTS, hipHostMallocDefault));
        c->hsnap_pinned = true;
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
                      &c->scanTmp, &c->sortTmp, &c->tileCnt, &c->tileOff, &c->dellist, &c->dbg,
                      &c->setsD, &c->bmD, &c->vset, &c->hullCnt, &c->hullOff, &c->hullStart, &c->setFirst,
                      &c->offsD, &c->bboxTmp, &c->lmark, &c->ffan, &c->modF, &c->badset, &c->bnd, &c->plo,
                      &c->perm, &c->xys, &c->pcell, &c->ccount, &c->ccursor, &c->cls,
                      &c->kin, &c->kv, &c->kinv, &c->kdone, &c->kcnt, &c->krs, &c->kroff, &c->kwork, &c->khash, &c->rdbuf};
    for (DevBuf *b : bufs) b->release();
    alpha_release(c->alpha);
}

}  // extern "C"

// device bytes held by a context (all grown buffers)
size_t galpha_ctx_device_bytes(galpha_ctx *c)
{
    DevBuf *bufs[] = {&c->smap, &c->g, &c->S, &c->cnt, &c->E, &c->ecnt, &c->vor, &c->bad, &c->fpix, &c->tv,
                      &c->tn, &c->remap, &c->stamp, &c->claim, &c->modA, &c->modB, &c->cand, &c->pos, &c->vt,
                      &c->vstate, &c->pix, &c->deg, &c->off, &c->fill, &c->fan, &c->moved, &c->decide, &c->mx,
                      &c->mlist, &c->loc, &c->lt, &c->scratch, &c->hnext, &c->hedge, &c->keys, &c->keys2,
                      &c->vals, &c->vals2, &c->rot, &c->newid, &c->out9, &c->counters, &c->status, &c->xy,
                      &c->scanTmp, &c->sortTmp, &c->tileCnt, &c->tileOff, &c->dellist, &c->dbg,
                      &c->setsD, &c->bmD, &c->vset, &c->hullCnt, &c->hullOff, &c->hullStart, &c->setFirst,
                      &c->offsD, &c->bboxTmp, &c->lmark, &c->ffan, &c->modF, &c->badset, &c->bnd, &c->plo,
                      &c->perm, &c->xys, &c->pcell, &c->ccount, &c->ccursor, &c->cls,
                      &c->kin, &c->kv, &c->kinv, &c->kdone, &c->kcnt, &c->krs, &c->kroff, &c->kwork, &c->khash, &c->rdbuf};
    size_t t = 0;
    for (DevBuf *b : bufs) t += b->bytes;
    for (auto &b : c->alpha.b) t += b.bytes;
    return t;
}

// gives every device buffer of an idle context back (they grow again on the next run)
void galpha_ctx_trim(galpha_ctx *c)
{
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    release_all(c);
    c->last_F = 0;
    c->last_alpha = false;
}

// ---- runs that allocate nothing --------------------------------------------------------------
// A context for runs whose device buffers are carved from a fixed region: no stream, no events
// and no pinned memory of its own (counter reads go to pageable host memory).
__attribute__((visibility("hidden"))) galpha_ctx *galpha_noalloc_ctx_new(int device)
{
    galpha_ctx *c = new galpha_ctx();
    c->device = device;
    c->hsnap_pageable.assign(SNAP_INTS, 0);
    c->hsnap = c->hsnap_pageable.data();
    return c;
}

// Upper bound of the bytes a single-set DT run of n points on a T x T texture carves: every
// buffer the run ensures, at its largest size (slots at the capacity 3 V + 8192, boundary cycle
// at V vertices), the scan and sort temporaries as rocPRIM reports them, and 256-byte rounding.
__attribute__((visibility("hidden"))) size_t galpha_noalloc_need(int n, int T)
{
    if (n < 3) return 4096;
    long long P = (long long)T * T, nf = frame_count(T), V = n + nf, cap = 3 * V + 8192;
    long long ntiles = (long long)(T / 16) * (T / 16), nc = P / 256 + 1;
    long long kb = 2;
    while ((1LL << kb) < (long long)n) kb++;
    size_t sort_tmp = 0, sort_tmp32 = 0;
    sort_pairs_u64(nullptr, nullptr, nullptr, nullptr, (int)cap, 0, (int)(2 * kb), nullptr, &sort_tmp, 0);
    sort_pairs_u32(nullptr, nullptr, nullptr, nullptr, (int)cap, 0, (int)std::min(32LL, 2 * kb), nullptr, &sort_tmp32, 0);
    sort_tmp = std::max(sort_tmp, sort_tmp32);
    double b = 0;
    b += 18.0 * P + 516.0 * T + 8.0 * ntiles + 8.0 * nc;          // textures, band ranges, tiles, cells
    b += (116.0 + 68.0) * cap;                                     // mesh slots, canonical keys + output
    b += (89.0 + 76.0 + 12.0) * V + 36.0 * n + 4.0 * nf;           // vertex arrays, boundary cycle, plan
    b += (double)COUNTER_BYTES + 4096.0 + 8.0 * 5 * 64;           // counters, status, box partials
    b += (double)scan_temp_bytes((size_t)std::max(ntiles, V + 2)) + (double)scan_temp_bytes((size_t)nc) + 8192.0;
    b += (double)sort_tmp + 256.0 + 96.0 * 256.0 + 65536.0;        // sort temporary, rounding, slack
    return (size_t)b;
}

// The DT of galphaComputeDevice (no alpha) with every device buffer carved from
// [scratch, scratch + scratch_bytes) and the canonical triangles copied to d_tris9 (9 ints each,
// capacity records).  Calls no hipMalloc or hipFree; returns the triangle count.
__attribute__((visibility("hidden"))) int galpha_noalloc_run(galpha_ctx *c, void *scratch, size_t scratch_bytes,
                                                             const double *d_xy, int n, int fboSize, int *d_tris9,
                                                             int capacity, void *stream, size_t *high_water)
{
    check_texture(fboSize);
    if (n < 0) throw GalphaError("negative point count");
    if (n > 0 && !d_xy) throw GalphaError("null point array");
    if (capacity < 0 || (capacity > 0 && !d_tris9)) throw GalphaError("invalid output buffer");
    if (n >= 3) check_slot_capacity((long long)n + frame_count(fboSize));
    size_t need = galpha_noalloc_need(n, fboSize);
    if (need > scratch_bytes)
        throw GalphaError("no-allocation run of " + std::to_string(n) + " points on a " + std::to_string(fboSize) +
                          "^2 texture needs up to " + std::to_string(need >> 20) + " MiB of scratch; the limit is " +
                          std::to_string(scratch_bytes >> 20) + " MiB");
    Arena ar;
    ar.base = (char *)scratch;
    ar.cap = scratch_bytes;
    ArenaScope scope(&ar);
    release_all(c);   // forget the previous run's carving: the buffers are carved again in order
    c->stream = (hipStream_t)stream;
    c->last_alpha = false;
    double bb[5] = {0, 0, 1, 1, 0};
    if (n > 0) {
        c->bboxTmp.ensure(8 * 5 * 64);
        bbox_device(d_xy, n, bb, c->bboxTmp.p, c->stream);
    }
    int F = run_dt(c, d_xy, n, fboSize, bb);
    if (F > capacity)
        throw GalphaError("output holds " + std::to_string(capacity) + " triangles; the run made " + std::to_string(F));
    if (F > 0) HIPCHK(hipMemcpyAsync(d_tris9, c->out9.p, 36 * (size_t)F, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (high_water) *high_water = ar.high;
    return F;
}

extern "C" {

void galphaDestroy(galpha_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    release_all(c);
    for (int i = 0; i < 16; i++)
        if (c->ev[i]) hipEventDestroy(c->ev[i]);
    alpha_destroy_events(c->alpha);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->hsnap && c->hsnap_pinned) hipHostFree(c->hsnap);
    stage_free(c->stage);
    delete c;
}

const char *galphaLastError(galpha_ctx *c) { return c ? c->err.c_str() : g_last_error.c_str(); }
const char *gpudtGetLastError(void) { return g_last_error.c_str(); }

// sets the thread-local message read by gpudtGetLastError (for the C-ABI calls in other files)
__attribute__((visibility("hidden"))) void galpha_set_last_error(const char *msg) { g_last_error = msg ? msg : ""; }

static int fail(galpha_ctx *c, const std::exception &e)
{
    g_last_error = e.what();
    if (c) c->err = e.what();
    return -1;
}

int galphaSetConstraints(galpha_ctx *c, const int *pairs, int n)
{
    if (!c) return -1;
    try {
        if (n < 0 || (n > 0 && !pairs)) throw GalphaError("invalid constraint list");
        c->cons.assign(pairs, pairs + 2 * (size_t)n);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int galphaComputeDevice(galpha_ctx *c, const double *d_xy, int n, int fboSize, int with_alpha, void *stream)
{
    if (!c) return -1;
    try {
        HIPCHK(hipSetDevice(c->device));
        hipStream_t own = c->stream;
        if (stream) c->stream = (hipStream_t)stream;
        int F;
        try {   // the context's own stream comes back on every exit
            double bb[5] = {0, 0, 1, 1, 0};
            check_texture(fboSize);
            if (n >= 3) check_slot_capacity((long long)n + frame_count(fboSize));
            if (n > 0) {
                c->bboxTmp.ensure(8 * 5 * 64);
                bbox_device(d_xy, n, bb, c->bboxTmp.p, c->stream);
            }
            F = run_dt(c, d_xy, n, fboSize, bb);
            c->last_alpha = false;
            if (with_alpha) {
                auto t0 = std::chrono::steady_clock::now();
                alpha_run(c->alpha, d_xy, n, c->out9.as<int>(), F, c->stream, &c->stats);
                c->last_alpha = true;
                HIPCHK(hipStreamSynchronize(c->stream));
                auto t1 = std::chrono::steady_clock::now();
                c->stats.ms_alpha = std::chrono::duration<double, std::milli>(t1 - t0).count();
            }
        } catch (...) {
            c->stream = own;
            throw;
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
        double bb[5] = {0, 0, 1, 1, 0};
        check_texture(fboSize);
        if (n >= 3) check_slot_capacity((long long)n + frame_count(fboSize));
        if (n > 0) bbox_host(xy, n, bb);
        c->xy.ensure(16 * (size_t)std::max(n, 1));
        if (n > 0) stage_h2d(c->stage, c->xy.p, xy, 16 * (size_t)n, c->stream);
        int F = run_dt(c, c->xy.as<double>(), n, fboSize, bb);
        if (tris_out9) {
            if (*ntris_out < F) throw GalphaError("output capacity too small");
            if (F > 0) stage_d2h(c->stage, tris_out9, c->out9.p, 36 * (size_t)F, c->stream);
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
        if (n < 1) throw GalphaError("no points");
        double bb[5];
        bbox_host(xy, n, bb);
        BatchPlan plan = plan_batch(&n, &T, bb, 1);
        int nf = plan.V - plan.ntot;
        if (pix_capacity < n + nf) throw GalphaError("pix capacity too small");
        ensure_pba(c, plan);
        c->xy.ensure(16 * (size_t)n);
        stage_h2d(c->stage, c->xy.p, xy, 16 * (size_t)n, c->stream);
        HIPCHK(hipMemsetAsync(c->counters.p, 0, COUNTER_BYTES, c->stream));
        BatchDev bd = upload_plan(c, plan);
        int iters = 0, isl = 0;
        run_pba(c, c->xy.as<double>(), bd, &iters, &isl, nullptr);
        c->stats.pbaIterations = iters;
        c->stats.islandSites = isl;
        stage_d2h(c->stage, labels, c->vor.p, 4 * (size_t)T * T, c->stream);
        stage_d2h(c->stage, pix, c->pix.p, 4 * (size_t)(n + nf), c->stream);
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
        *pA = alpha_copy_to_host(c->alpha, c->stage, c->stream);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

// ---- alpha-complex queries on the last single-set alpha result ---------------------------------
int galphaAlphaRank(galpha_ctx *c, double alpha, int *rank)
{
    if (!c || !rank) return -1;
    try {
        if (!c->last_alpha) throw GalphaError("no alpha result on this context");
        HIPCHK(hipSetDevice(c->device));
        *rank = alpha_rank_of(c->alpha, alpha, c->stage, c->stream);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int galphaClassifyDevice(galpha_ctx *c, int rank, unsigned char *d_vert, unsigned char *d_edge, unsigned char *d_tri,
                         void *stream)
{
    if (!c) return -1;
    try {
        if (!c->last_alpha) throw GalphaError("no alpha result on this context");
        HIPCHK(hipSetDevice(c->device));
        alpha_classify(c->alpha, rank, d_vert, d_edge, d_tri, stream ? (hipStream_t)stream : c->stream);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int galphaClassify(galpha_ctx *c, int rank, unsigned char *vert, unsigned char *edge, unsigned char *tri)
{
    if (!c) return -1;
    try {
        if (!c->last_alpha) throw GalphaError("no alpha result on this context");
        HIPCHK(hipSetDevice(c->device));
        size_t n = (size_t)c->alpha.n, E = (size_t)c->alpha.E, F = (size_t)c->alpha.F;
        c->cls.ensure(n + E + F + 64);
        unsigned char *dv = c->cls.as<unsigned char>(), *de = dv + n, *dt = de + E;
        alpha_classify(c->alpha, rank, vert ? dv : nullptr, edge ? de : nullptr, tri ? dt : nullptr, c->stream);
        if (vert) stage_d2h(c->stage, vert, dv, n, c->stream);
        if (edge) stage_d2h(c->stage, edge, de, E, c->stream);
        if (tri) stage_d2h(c->stage, tri, dt, F, c->stream);
        HIPCHK(hipStreamSynchronize(c->stream));
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
        if (p->nPoints < 0 || (p->nPoints > 0 && !p->points)) throw GalphaError("invalid point list");
        if (p->nConstraints < 0 || (p->nConstraints > 0 && !p->constraints)) throw GalphaError("invalid constraint list");
        galpha_ctx *c = default_ctx();
        if (!c) throw GalphaError(g_last_error);
        std::lock_guard<std::mutex> lk(g_default_mu);
        c->last_dt = {};
        int n = p->nPoints;
        c->cons.assign(p->constraints, p->constraints + 2 * (size_t)std::max(p->nConstraints, 0));
        int F;
        std::atomic<unsigned long long> hp{0}, ht{0};
        try {
            HIPCHK(hipSetDevice(c->device));
            double bb[5] = {0, 0, 1, 1, 0};
            check_texture(p->fboSize);
            if (n >= 3) check_slot_capacity((long long)n + frame_count(p->fboSize));
            if (n > 0) bbox_host((const double *)p->points, n, bb);
            c->xy.ensure(16 * (size_t)std::max(n, 1));
            if (n > 0) stage_h2d_hash(c->stage, c->xy.p, p->points, 16 * (size_t)n, c->stream, &hp);
            F = run_dt(c, c->xy.as<double>(), n, p->fboSize, bb);
        } catch (...) {
            c->cons.clear();
            throw;
        }
        c->cons.clear();
        PGPUDTOUTPUT out = new GPUDTOUTPUT;
        out->nTris = F;
        out->triangles = (gpudtTriangle *)galpha_host_alloc(sizeof(gpudtTriangle) * (size_t)(F > 0 ? F : 1));
        if (!out->triangles) {
            delete out;
            throw GalphaError("host allocation failed");
        }
        if (F > 0) stage_d2h_hash(c->stage, out->triangles, c->out9.p, 36 * (size_t)F, c->stream, &ht);
        c->last_dt.points = p->points;
        c->last_dt.tris = out->triangles;
        c->last_dt.n = n;
        c->last_dt.F = F;
        c->last_dt.hp = hp.load();
        c->last_dt.ht = ht.load();
        return out;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void gpudtReleaseDTOutput(PGPUDTOUTPUT o)
{
    if (!o) return;
    free(o->triangles);
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
        if (n < 0 || dt->nTris < 0 || dt->nTris >= (1 << 29))
            throw GalphaError("alpha: point or triangle count out of range");
        if ((n > 0 && !p->points) || (dt->nTris > 0 && !dt->triangles))
            throw GalphaError("alpha: null point or triangle array");
        HIPCHK(hipSetDevice(c->device));
        // The output of the last gpudtComputeDT, unchanged (same arrays, sizes and content): its
        // points and triangles are still on the device, and the mesh is the pipeline's own.
        bool resident = c->last_dt.tris == (const void *)dt->triangles && c->last_dt.F == dt->nTris &&
                        c->last_dt.points == (const void *)p->points && c->last_dt.n == n &&
                        host_hash(c->stage, p->points, 16 * (size_t)std::max(n, 0)) == c->last_dt.hp &&
                        host_hash(c->stage, dt->triangles, 36 * (size_t)std::max(dt->nTris, 0)) == c->last_dt.ht;
        if (!resident) {
            c->last_dt = {};
            c->xy.ensure(16 * (size_t)std::max(n, 1));
            if (n > 0) stage_h2d(c->stage, c->xy.p, p->points, 16 * (size_t)n, c->stream);
            c->out9.ensure(36 * (size_t)std::max(dt->nTris, 1));
            if (dt->nTris > 0) stage_h2d(c->stage, c->out9.p, dt->triangles, 36 * (size_t)dt->nTris, c->stream);
        }
        alpha_run(c->alpha, c->xy.as<double>(), n, c->out9.as<int>(), dt->nTris, c->stream, &c->stats, !resident);
        *pA = alpha_copy_to_host(c->alpha, c->stage, c->stream);
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

void gpudtReleaseAlpha(PGPUDTALPHA a) { alpha_free_host(a); }

void galphaFree(void *p) { free(p); }

// ---- device memory for callers without a HIP runtime of their own ---------------------------
int galphaDeviceCount(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void *galphaDeviceMalloc(int device, size_t bytes)
{
    try {
        g_last_error.clear();
        HIPCHK(hipSetDevice(device));
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, bytes ? bytes : 1));
        return p;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

void galphaDeviceFree(void *d)
{
    if (d) hipFree(d);
}

// kind: 0 host -> device, 1 device -> host, 2 device -> device; synchronous
int galphaDeviceCopy(void *dst, const void *src, size_t bytes, int kind)
{
    try {
        g_last_error.clear();
        if (kind < 0 || kind > 2) throw GalphaError("copy kind must be 0, 1 or 2");
        hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
        if (bytes) HIPCHK(hipMemcpy(dst, src, bytes, k));
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

void *galphaStreamCreate(int device)
{
    try {
        g_last_error.clear();
        HIPCHK(hipSetDevice(device));
        hipStream_t s = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));   // default priority
        return (void *)s;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}

int galphaStreamDestroy(void *stream)
{
    if (!stream) return 0;
    return hipStreamDestroy((hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int galphaStreamSynchronize(void *stream)
{
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) {
        g_last_error = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

int galphaDeviceSynchronize(void)
{
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        g_last_error = std::string("hipDeviceSynchronize: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

// Batched entry point: all sets go through one pass of the pipeline on this context (every launch
// covers every set).  Results are concatenated per input set with set-local vertex and triangle
// indices.  tris9_out / tri_offsets_out are malloc'ed (release with galphaFree).
// triangles of planned set k with set-local indices into out (9 ints each)
static void copy_planned_set(galpha_ctx *c, int k, int *out)
{
    int t0 = c->set_first[k], t1 = c->set_first[k + 1];
    if (t1 <= t0) return;
    ctx_read(c, out, c->out9.as<int>() + (size_t)9 * t0, 36 * (size_t)(t1 - t0));
    int poff = c->plan_sets[k].poff;
    for (int t = 0; t < t1 - t0; t++) {
        int *o = out + (size_t)9 * t;
        for (int i = 0; i < 3; i++) {
            o[3 + i] -= poff;
            if (o[i] >= 0) o[i] = ((o[i] >> 2) - t0) << 2 | (o[i] & 3);
        }
    }
}

int galphaBatchDevice(galpha_ctx *c, const double *d_xy, const int *offsets, int nsets, const int *fboSizes,
                      int with_alpha, void *stream)
{
    if (!c || !offsets || nsets < 0) return -1;
    try {
        HIPCHK(hipSetDevice(c->device));
        hipStream_t own = c->stream;
        if (stream) c->stream = (hipStream_t)stream;
        int F;
        try {
            F = batch_run(c, d_xy, offsets, nsets, fboSizes, with_alpha, true);
        } catch (...) {
            c->stream = own;
            throw;
        }
        c->stream = own;
        return F;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

}  // extern "C"

// One batched pass over host points that leaves every result on the device, for callers that
// then copy set by set (galphaBatchTriangles, galphaBatchCopyAlpha): the serving loop.
int galpha_ctx_batch_host(galpha_ctx *c, const double *xy, const int *offsets, int nsets, const int *fboSizes,
                          int with_alpha)
{
    if (!c || !offsets || nsets < 0) return -1;
    try {
        HIPCHK(hipSetDevice(c->device));
        return batch_run(c, xy, offsets, nsets, fboSizes, with_alpha);
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

extern "C" {

int galphaBatchTriangles(galpha_ctx *c, int set, int *tris9, int capacity)
{
    if (!c) return -1;
    try {
        if (set < 0 || set >= (int)c->batch_npts.size()) throw GalphaError("set index out of range");
        int k = -1;
        for (int q = 0; q < (int)c->batch_sets.size(); q++)
            if (c->batch_sets[q] == set) { k = q; break; }
        if (k < 0) return 0;
        int F = c->set_first[k + 1] - c->set_first[k];
        if (!tris9) return F;
        if (capacity < F) throw GalphaError("output capacity too small");
        copy_planned_set(c, k, tris9);
        return F;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int galphaBatch(galpha_ctx *c, const double *xy, const int *offsets, int nsets, const int *fboSizes,
                int with_alpha, int **tris9_out, int **tri_offsets_out)
{
    if (!c || !offsets || nsets < 0 || !tris9_out || !tri_offsets_out) return -1;
    try {
        HIPCHK(hipSetDevice(c->device));
        int F = batch_run(c, xy, offsets, nsets, fboSizes, with_alpha);
        std::vector<int> all((size_t)9 * F);
        if (F > 0) ctx_read(c, all.data(), c->out9.p, 36 * (size_t)F);
        std::vector<int> toff(nsets + 1, 0);
        int na = (int)c->batch_sets.size();
        std::vector<int> ntri(nsets, 0);
        for (int k = 0; k < na; k++) ntri[c->batch_sets[k]] = c->set_first[k + 1] - c->set_first[k];
        for (int s = 0; s < nsets; s++) toff[s + 1] = toff[s] + ntri[s];
        // planned sets are in input order, so the concatenation order already matches
        for (int k = 0; k < na; k++) {
            int t0 = c->set_first[k], t1 = c->set_first[k + 1];
            int poff = c->plan_sets[k].poff;
            for (int t = t0; t < t1; t++) {
                int *o = all.data() + (size_t)9 * t;
                for (int i = 0; i < 3; i++) {
                    o[3 + i] -= poff;
                    if (o[i] >= 0) o[i] = ((o[i] >> 2) - t0) << 2 | (o[i] & 3);
                }
            }
        }
        *tris9_out = (int *)galpha_host_alloc(sizeof(int) * std::max<size_t>(all.size(), 1));
        *tri_offsets_out = (int *)malloc(sizeof(int) * (nsets + 1));
        if (!*tris9_out || !*tri_offsets_out) throw GalphaError("host allocation failed");
        memcpy(*tris9_out, all.data(), sizeof(int) * all.size());
        memcpy(*tri_offsets_out, toff.data(), sizeof(int) * (nsets + 1));
        return nsets;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

// texture side this context picks for n points (autotuned density of n's size class)
int galphaTextureFor(galpha_ctx *c, int n)
{
    if (!c) return galphaAutoTexture(n);
    return texture_for_density(n, c->density[tune_class(n)]);
}

// Texture autotuner.  For every size class present in the sample (sets of xy / offsets), the
// sample's sets of that class run as one batch (DT + alpha) at each candidate density; the fastest
// density (median of `reps` timed runs after one warm-up) becomes the context's choice for that
// class.  densities_out (optional) receives the 4 chosen densities.  Results never depend on the
// texture size, only the time does.
int galphaAutotune(galpha_ctx *c, const double *xy, const int *offsets, int nsets, int reps, double *densities_out)
{
    if (!c || !xy || !offsets || nsets <= 0) return -1;
    try {
        HIPCHK(hipSetDevice(c->device));
        static const double cand[] = {6.0, 9.0, 12.0, 16.0, 24.0, 32.0};
        if (reps < 1) reps = 1;
        for (int k = 0; k < GALPHA_TUNE_CLASSES; k++) {
            std::vector<double> pts;
            std::vector<int> offs(1, 0);
            for (int s = 0; s < nsets; s++) {
                int n = offsets[s + 1] - offsets[s];
                if (n < 3 || tune_class(n) != k) continue;
                pts.insert(pts.end(), xy + 2 * (size_t)offsets[s], xy + 2 * (size_t)offsets[s + 1]);
                offs.push_back(offs.back() + n);
            }
            int ns = (int)offs.size() - 1;
            if (ns == 0) continue;
            double best = -1, best_ms = 0;
            for (double d : cand) {
                std::vector<int> Ts(ns);
                for (int q = 0; q < ns; q++) Ts[q] = texture_for_density(offs[q + 1] - offs[q], d);
                std::vector<double> ms;
                for (int r = 0; r <= reps; r++) {
                    auto t0 = std::chrono::steady_clock::now();
                    batch_run(c, pts.data(), offs.data(), ns, Ts.data(), 1);
                    HIPCHK(hipStreamSynchronize(c->stream));
                    auto t1 = std::chrono::steady_clock::now();
                    if (r > 0) ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
                }
                std::sort(ms.begin(), ms.end());
                double med = ms[ms.size() / 2];
                if (best < 0 || med < best_ms) { best = d; best_ms = med; }
            }
            c->density[k] = best;
        }
        if (densities_out)
            for (int k = 0; k < GALPHA_TUNE_CLASSES; k++) densities_out[k] = c->density[k];
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

// alpha result of input set `set` of the last galphaBatch(with_alpha = 1) call
int galphaBatchCopyAlpha(galpha_ctx *c, int set, PGPUDTALPHA *pA)
{
    if (!c || !pA) return -1;
    try {
        if (!c->last_alpha) throw GalphaError("no alpha result on this context");
        int k = -1;
        for (int q = 0; q < (int)c->batch_sets.size(); q++)
            if (c->batch_sets[q] == set) { k = q; break; }
        if (k < 0) {
            // a set with fewer than 3 points: no simplices beyond its vertices
            *pA = alpha_empty_host(c->batch_npts.size() > (size_t)set ? c->batch_npts[set] : 0);
            return 0;
        }
        *pA = alpha_copy_set_to_host(c->alpha, k, c->stage, c->stream);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

}  // extern "C"
