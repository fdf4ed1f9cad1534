// serve.cpp -- request loop for alpha-shape serving (no counterpart in the reference, whose API
// is a single blocking call over global state, gpudt/gpudt.cpp:64-77).
//
// A server owns `slots` worker threads, each with its own context (HIP stream, device buffers
// kept resident and reused across requests).  Requests are queued in arrival order.  A free
// worker admits the longest prefix of the queue that fits the HBM budget -- the bytes that all
// contexts hold, with its own holdings grown to what the admitted requests need -- up to
// `max_group` requests (and, when set, up to a number of points), and runs them as one batched
// pass (continuous batching: requests that
// arrive while others are on the GPU join the next pass of a free slot).  When the head request
// does not fit, an idle worker gives back its own buffers and waits for other slots to finish;
// a request larger than the whole budget is refused at submission.  Results (canonical triangles, alpha arrays) are copied to host
// memory owned by the ticket, so a context can take the next group at once.
#include "../../include/gpudt.h"
#include "alpha_host.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

size_t galpha_ctx_device_bytes(galpha_ctx *c);
void galpha_ctx_trim(galpha_ctx *c);
int galpha_ctx_batch_host(galpha_ctx *c, const double *xy, const int *offsets, int nsets, const int *fboSizes,
                          int with_alpha);

namespace {

using clk = std::chrono::steady_clock;

struct Request {
    long long id = 0;
    std::vector<double> xy;
    int n = 0, T = 0, with_alpha = 0;
    size_t need = 0;                   // estimated device bytes of a context that runs it alone
    clk::time_point submitted, started, finished;
    bool done = false;
    int status = 0;                    // >= 0 triangles, < 0 failure
    std::string err;
    std::unique_ptr<int, void (*)(void *)> tris9{nullptr, free};   // 9 ints per triangle (galpha_host_alloc)
    int ntris = 0;
    PGPUDTALPHA alpha = nullptr;
    int group = 0;                     // requests in the pass that served it
};

}  // namespace

struct galpha_server {
    int device = 0, slots = 0, max_group = 0;
    long long max_points = 0;          // points admitted into one pass (0: no limit); a larger
                                       // request still runs, alone
    size_t budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<long long, std::shared_ptr<Request>> tickets;
    std::vector<size_t> held;          // device bytes held by each worker's context
    std::vector<char> busy;            // worker is running a pass (its holdings cannot be trimmed)
    std::vector<galpha_ctx *> ctxs;    // each worker's context, created before the workers start
    std::vector<std::thread> workers;
    long long next_id = 1;
    bool stop = false;
    long long passes = 0, served = 0;
};

// Device bytes a context grows to for `npoints` points on textures of `pixels` pixels, DT + alpha,
// from the buffers each stage ensures (all with the 1/8 growth headroom): textures 20 B per pixel
// (site map, column field, stacks, envelopes, labels); mesh about 650 B per vertex (3V triangle
// slots with links, remaps, votes, lists and fans; vertex arrays; canonical keys and output);
// alpha about 1400 B per point (5n spectrum entries with keys, intervals and scans; 13n master
// entries; sort temporaries); plus a fixed 16 MiB.
static size_t group_footprint(long long npoints, long long pixels)
{
    double b = 20.5 * (double)pixels + 2150.0 * (double)npoints + (16 << 20);
    return (size_t)b;
}
static size_t footprint(long long n, long long T) { return group_footprint(n, T * T); }

static void worker_main(galpha_server *S, int w)
{
    hipSetDevice(S->device);
    galpha_ctx *ctx = S->ctxs[w];
    for (;;) {
        std::vector<std::shared_ptr<Request>> grp;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop) {
                    lk.unlock();
                    galphaDestroy(ctx);
                    return;
                }
                if (!S->queue.empty()) {
                    size_t others = 0;
                    for (int q = 0; q < S->slots; q++)
                        if (q != w) others += S->held[q];
                    size_t need = 0;
                    long long npts = 0, pix = 0;
                    for (auto &r : S->queue) {
                        if ((int)grp.size() >= S->max_group) break;
                        if (!grp.empty() && S->max_points > 0 && npts + r->n > S->max_points) break;
                        size_t nd = group_footprint(npts + r->n, pix + (long long)r->T * r->T);
                        size_t mine = std::max(S->held[w], nd);
                        if (others + mine > S->budget) break;
                        grp.push_back(r);
                        need = nd;
                        npts += r->n;
                        pix += (long long)r->T * r->T;
                    }
                    if (!grp.empty()) {
                        S->queue.erase(S->queue.begin(), S->queue.begin() + grp.size());
                        S->held[w] = std::max(S->held[w], need);
                        S->busy[w] = 1;
                        break;
                    }
                    if (S->held[w] > 0) {
                        // the head request does not fit: an idle slot gives its buffers back
                        // so that some slot can admit it
                        lk.unlock();
                        galpha_ctx_trim(ctx);
                        lk.lock();
                        S->held[w] = 0;
                        S->cv_work.notify_all();
                        continue;
                    }
                    // this slot holds nothing, yet the head does not fit: wake the idle slots
                    // that hold buffers so that they give them back (busy slots wake everyone
                    // when their pass ends).  Once they have trimmed, none is left to wake.
                    bool idle_holder = false;
                    for (int q = 0; q < S->slots; q++)
                        if (q != w && !S->busy[q] && S->held[q] > 0) idle_holder = true;
                    if (idle_holder) S->cv_work.notify_all();
                }
                S->cv_work.wait(lk);
            }
        }
        // one batched pass over the group
        auto t0 = clk::now();
        std::vector<int> offs(grp.size() + 1, 0), Ts(grp.size());
        for (size_t k = 0; k < grp.size(); k++) {
            offs[k + 1] = offs[k] + grp[k]->n;
            Ts[k] = grp[k]->T;
        }
        std::vector<double> xy(2 * (size_t)offs.back());
        for (size_t k = 0; k < grp.size(); k++)
            std::copy(grp[k]->xy.begin(), grp[k]->xy.end(), xy.begin() + 2 * (size_t)offs[k]);
        int with_alpha = 0;
        for (auto &r : grp) with_alpha |= r->with_alpha;
        // results stay on the device; each request's triangles and alpha arrays are then copied
        // straight into its ticket through the context's pinned staging
        int rc = galpha_ctx_batch_host(ctx, xy.data(), offs.data(), (int)grp.size(), Ts.data(), with_alpha);
        std::string err = rc < 0 ? std::string(galphaLastError(ctx)) : std::string();
        for (size_t k = 0; k < grp.size(); k++) {
            auto &r = grp[k];
            r->started = t0;
            r->group = (int)grp.size();
            if (rc < 0) { r->status = -1; r->err = err; continue; }
            int nt = galphaBatchTriangles(ctx, (int)k, nullptr, 0);
            if (nt >= 0) {
                r->tris9.reset((int *)galpha_host_alloc(36 * (size_t)nt));
                nt = r->tris9 ? galphaBatchTriangles(ctx, (int)k, r->tris9.get(), nt) : -1;
                r->ntris = nt;
            }
            if (nt < 0) { r->status = -1; r->err = galphaLastError(ctx); continue; }
            r->status = nt;
            if (r->with_alpha && galphaBatchCopyAlpha(ctx, (int)k, &r->alpha) < 0) {
                r->status = -1;
                r->err = galphaLastError(ctx);
            }
        }
        size_t held_now = galpha_ctx_device_bytes(ctx);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            auto t1 = clk::now();
            for (auto &r : grp) { r->finished = t1; r->done = true; r->xy.clear(); r->xy.shrink_to_fit(); }
            S->held[w] = held_now;
            S->busy[w] = 0;
            S->passes++;
            S->served += (long long)grp.size();
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

extern "C" {

extern "C" void galpha_set_last_error(const char *msg);   // galpha.cpp (library-internal)

galpha_server *galphaServerCreate(int device, int slots, size_t hbm_budget_bytes, int max_group)
{
    if (slots < 1 || slots > 64 || max_group < 1) {
        galpha_set_last_error("server: slots must be in [1, 64] and max_group at least 1");
        return nullptr;
    }
    galpha_server *S = new galpha_server();
    S->device = device;
    S->slots = slots;
    S->max_group = max_group;
    S->budget = hbm_budget_bytes;
    S->held.assign(slots, 0);
    S->busy.assign(slots, 0);
    // every slot's context exists before any worker runs: a device error fails the creation
    for (int w = 0; w < slots; w++) {
        galpha_ctx *c = galphaCreate(device);
        if (!c) {
            std::string why = std::string("server: ") + gpudtGetLastError();
            for (galpha_ctx *x : S->ctxs) galphaDestroy(x);
            delete S;
            galpha_set_last_error(why.c_str());
            return nullptr;
        }
        S->ctxs.push_back(c);
    }
    for (int w = 0; w < slots; w++) S->workers.emplace_back(worker_main, S, w);
    return S;
}

void galphaServerDestroy(galpha_server *S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &t : S->workers) t.join();
    for (auto &kv : S->tickets)
        if (kv.second->alpha) gpudtReleaseAlpha(kv.second->alpha);
    delete S;
}

long long galphaServerSubmit(galpha_server *S, const double *xy, int n, int fboSize, int with_alpha)
{
    if (!S || n < 0 || (n > 0 && !xy)) return -1;
    auto r = std::make_shared<Request>();
    r->xy.assign(xy, xy + 2 * (size_t)n);
    r->n = n;
    r->T = fboSize > 0 ? fboSize : galphaAutoTexture(n);
    r->with_alpha = with_alpha ? 1 : 0;
    r->need = footprint(n, r->T);
    r->submitted = clk::now();
    long long id;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (r->need > S->budget) return -2;   // can never be admitted
        id = S->next_id++;
        r->id = id;
        S->tickets[id] = r;
        S->queue.push_back(r);
    }
    // every waiting slot looks at the queue: the first that can admit the head takes it, and one
    // that holds nothing cannot be the only one woken while idle slots hold the memory it needs
    S->cv_work.notify_all();
    return id;
}

// Waits for a ticket.  Returns its triangle count (< 0 on failure); latency = submit -> done and
// queue = submit -> start of its pass, in ms; group = requests in that pass.
int galphaServerWait(galpha_server *S, long long ticket, double *latency_ms, double *queue_ms, int *group)
{
    if (!S) return -1;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end()) return -1;
    auto r = it->second;
    S->cv_done.wait(lk, [&] { return r->done; });
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->finished - r->submitted).count();
    if (queue_ms) *queue_ms = std::chrono::duration<double, std::milli>(r->started - r->submitted).count();
    if (group) *group = r->group;
    return r->status;
}

// Copies the triangles of a finished ticket (9 ints each, gpudtTriangle layout).
int galphaServerTriangles(galpha_server *S, long long ticket, int *tris9, int capacity)
{
    if (!S) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end() || !it->second->done || it->second->status < 0) return -1;
    auto &r = it->second;
    int F = r->ntris;
    if (!tris9) return F;
    if (capacity < F) return -1;
    if (F > 0) memcpy(tris9, r->tris9.get(), 36 * (size_t)F);
    return F;
}

// Hands the alpha result of a finished ticket to the caller (release with gpudtReleaseAlpha).
int galphaServerAlpha(galpha_server *S, long long ticket, PGPUDTALPHA *pA)
{
    if (!S || !pA) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end() || !it->second->done || !it->second->alpha) return -1;
    *pA = it->second->alpha;
    it->second->alpha = nullptr;
    return 0;
}

// Forgets a finished ticket (its host results are freed).
int galphaServerRelease(galpha_server *S, long long ticket)
{
    if (!S) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end() || !it->second->done) return -1;
    if (it->second->alpha) gpudtReleaseAlpha(it->second->alpha);
    S->tickets.erase(it);
    return 0;
}

// Largest number of points one pass admits (0: no limit).  Short passes free slots sooner, so
// small requests wait less behind large ones; a request above the limit still runs, alone.
int galphaServerSetGroupPoints(galpha_server *S, long long max_points)
{
    if (!S || max_points < 0) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    S->max_points = max_points;
    return 0;
}

// Device bytes the admission rule counts for one request of n points on a fboSize texture (0: the
// default texture for n).
size_t galphaServerFootprint(int n, int fboSize)
{
    if (n < 0) return 0;
    long long T = fboSize > 0 ? fboSize : galphaAutoTexture(n);
    return footprint(n, T);
}

// Passes run, requests served, and device bytes held by all slots.
int galphaServerStats(galpha_server *S, long long *passes, long long *served, size_t *held_bytes)
{
    if (!S) return -1;
    std::lock_guard<std::mutex> lk(S->mu);
    if (passes) *passes = S->passes;
    if (served) *served = S->served;
    if (held_bytes) {
        size_t h = 0;
        for (size_t x : S->held) h += x;
        *held_bytes = h;
    }
    return 0;
}

}  // extern "C"
