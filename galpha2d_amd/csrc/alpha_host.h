// alpha_host.h -- alpha-shape stage (device buffers + host entry points) shared with galpha.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>
#include "../../include/gpudt.h"
#include <cstdlib>
#include <sys/mman.h>

// Host arrays handed to callers (released with free()).  Large ones are 2 MiB aligned and marked
// for transparent huge pages, so the first touch by the copy from the device faults once per
// 2 MiB instead of once per 4 KiB page (the faults, not PCIe, dominate host-to-host calls).
inline void *galpha_host_alloc(size_t bytes)
{
    const size_t huge = (size_t)2 << 20;
    if (bytes < 2 * huge) return malloc(bytes ? bytes : 1);
    void *p = nullptr;
    if (posix_memalign(&p, huge, bytes) != 0) return nullptr;
    madvise(p, (bytes + huge - 1) & ~(huge - 1), MADV_HUGEPAGE);
    return p;
}

struct GalphaError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// true while a no-allocation run (buffers carved from a fixed device region) is active on this
// thread: nothing may call hipMalloc then
bool galpha_arena_active();

struct AlphaBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Pinned host staging for host <-> device copies, owned by a context.  Copies run in chunks
// through two halves used alternately, so the DMA of one chunk overlaps the host copy of the
// other.  (Copies from pageable memory go through the runtime's shared staging buffer and
// serialise the copies of all threads, e.g. of concurrent serving slots.)
struct CopyPool;                 // host threads that split large memcpys (galpha.cpp)
struct HostStage {
    char *p = nullptr;           // 2 * half bytes of pinned memory, allocated on first use
    size_t half = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    CopyPool *pool = nullptr;    // started on the first large copy, joined by stage_free
};
// dst (host) <- src (device); returns when dst holds the data
void stage_d2h(HostStage &s, void *dst, const void *src, size_t bytes, hipStream_t st);
// dst (device) <- src (host); returns when src may be reused (the copies are ordered on st)
void stage_h2d(HostStage &s, void *dst, const void *src, size_t bytes, hipStream_t st);
void stage_free(HostStage &s);
// the same, adding the content hash of the host array to *hash (galpha.cpp)
#include <atomic>
void stage_d2h_hash(HostStage &s, void *dst, const void *src, size_t bytes, hipStream_t st,
                    std::atomic<unsigned long long> *hash);
void stage_h2d_hash(HostStage &s, void *dst, const void *src, size_t bytes, hipStream_t st,
                    std::atomic<unsigned long long> *hash);

#include <vector>

// sets of a batch as the alpha stage sees them (a single set: nsets = 1, vset = NULL)
struct AlphaBatchInfo {
    int nsets = 1;
    const int *vset = nullptr;     // device [n]: set of every real point
    std::vector<int> poff;         // [nsets + 1] first real point of every set
    std::vector<int> tfirst;       // [nsets + 1] first triangle of every set
    bool check_mesh = false;       // caller triangles: reject ids and links that are not a mesh
};

struct AlphaDev {
    AlphaBuf b[64];
    int n = 0, F = 0, E = 0, S = 0, specSize = 0, M = 0, nsp = 0;
    // per set (host): first point, triangle, edge, spectrum slot and master entry; [nsets + 1]
    int nsets = 1;
    std::vector<int> poff, tfirst, efirst, specstart, mfirst;
    const int *tris9 = nullptr;   // device triangles of the last run (the context's output)
    hipEvent_t ev[4] = {};        // sub-stage boundaries (created on the first run)
    int *hpin = nullptr;          // pinned host ints for counter reads (allocated on the first run)
};

void alpha_run(AlphaDev &a, const double *d_xy, int n, const int *d_tris9, int F, hipStream_t st,
               galpha_stats *stats, bool check_mesh = false);
void alpha_run_batch(AlphaDev &a, const double *d_xy, int n, const int *d_tris9, int F, const AlphaBatchInfo &info,
                     hipStream_t st, galpha_stats *stats);
PGPUDTALPHA alpha_copy_to_host(AlphaDev &a, HostStage &stage, hipStream_t st);
// one set of the last batch, with set-local point, triangle and edge indices
PGPUDTALPHA alpha_copy_set_to_host(AlphaDev &a, int set, HostStage &stage, hipStream_t st);
// alpha of a set without triangles: spectrum [0], every vertex interval -1
PGPUDTALPHA alpha_empty_host(int npoints);
void alpha_free_host(PGPUDTALPHA p);
// alpha-complex query of a single-set result: rank of an alpha value, and the class
// (0 exterior, 1 singular, 2 regular, 3 interior) of every vertex, edge and triangle at a rank;
// the class arrays are device pointers (NULL: skipped), written on st
int alpha_rank_of(AlphaDev &a, double alpha, HostStage &stage, hipStream_t st);
void alpha_classify(AlphaDev &a, int rank, unsigned char *vcls, unsigned char *ecls, unsigned char *tcls,
                    hipStream_t st);
void alpha_release(AlphaDev &a);
// events and pinned host memory of the alpha stage
void alpha_destroy_events(AlphaDev &a);
