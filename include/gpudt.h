/*
 * gpudt.h -- C ABI of the MI355X-native 2D Delaunay + alpha-shape library (libgalpha.so).
 *
 * Drop-in for the reference host API gpudt/gpudt.h (ashwin/gAlpha2D): the structs below
 * are byte-compatible with gpudt/gpudt.h:77-128 (REAL = double, gpudtTriangle = 36 bytes,
 * GPUDTPARAMS field order unchanged) and gpudtComputeDT / gpudtReleaseDTOutput keep the
 * reference's names, argument meaning and ownership (gpudt/gpudt.h:131-132,
 * gpudt/gpudt.cpp:771-929).  Differences, all on error paths:
 *   - failures return NULL (or a negative code) and set a thread-local message readable
 *     with gpudtGetLastError(); the reference calls exit(-1) (Cuda/common.h:60-93);
 *   - constraints that the reference leaves undefined are errors: an endpoint outside
 *     [0, nPoints) or equal to the other, an endpoint that is a dropped duplicate point, a
 *     point inside a constraint segment, and two constraints that cross;
 *   - minX/minY/maxX/maxY are recomputed from the points;
 *   - gpudtReleaseDTOutput frees the struct too (the reference leaks it, gpudt.cpp:926-929).
 * Output triangles are counter-clockwise, each rotated so vtx[0] is its smallest vertex and
 * listed in increasing (vtx[0], vtx[1], vtx[2]) order; tri[i] = (t << 2) | j names the
 * neighbour across the edge opposite vtx[i] (t.vtx[j] is its far vertex), -1 on the hull;
 * tmp[] is zero.  Exact duplicate points keep their smallest index; the others are left out
 * (the reference drops duplicates too, Cuda/cudaMissing.cu:285-289).
 * With nConstraints > 0 (constraints: pairs of point indices, gpudt/gpudt.h:113-114) the output
 * is the constrained DT: every constraint is an edge and every other edge is locally Delaunay
 * (Cuda/cudaConstraint.cu:828-939, Cuda/cudaFlipping.cu:111-220, gpudt.cpp:88-155); it is unique
 * under the same tie-break, so the output does not depend on the constraint order.
 *
 * The alpha-shape entry points replace the C++ class Alpha_Shapes_2D
 * (AlphaShape/Alpha_Shapes_2D.h:43-121, computeAlphaShapes Alpha_Shapes_2D.cu:206-233) with
 * a plain struct of arrays (layouts of AlphaShape/headers.h:48-90).
 *
 * The galpha* entry points (contexts, device-resident runs, batches, serving) have no
 * counterpart in the reference; they are for callers that keep data on the GPU.
 */
#ifndef GALPHA_GPUDT_H
#define GALPHA_GPUDT_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REAL double
#define MARKER -32768

typedef struct {
    REAL x, y;
} gpudtVertex;

typedef struct gpudtTriangle {
    int tri[3]; /* neighbour across the edge opposite vtx[i], (t << 2) | j, or -1 */
    int vtx[3]; /* counter-clockwise vertices (input point indices) */
    int tmp[3]; /* zero */
} gpudtTriangle;

typedef struct {
    REAL minX, minY;
    REAL maxX, maxY;
    int nPoints;
    gpudtVertex *points;
    int nConstraints;
    int *constraints;
    int fboSize; /* texture side: power of two, 64..16384 */
} GPUDTPARAMS, *PGPUDTPARAMS;

typedef struct {
    gpudtTriangle *triangles;
    int nTris;
} GPUDTOUTPUT, *PGPUDTOUTPUT;

/* Reference API (gpudt/gpudt.h:131-132). */
PGPUDTOUTPUT gpudtComputeDT(PGPUDTPARAMS pParam);
void gpudtReleaseDTOutput(PGPUDTOUTPUT pGpudtOutput);

/* Thread-local message of the last failure ("" if none). */
const char *gpudtGetLastError(void);

#ifdef __cplusplus
}
#endif

/* ---- mesh navigation (gpudt/gpudt.h:137-244) ------------------------------------------------
 * Same names, arguments and results as the reference, so a caller written against the reference
 * header (its validator checkResult, main.cu:163-343) compiles unchanged.  An oriented triangle
 * (tri, orient) names one edge of triangle `tri`: the edge opposite vtx[orient], directed from
 * org = vtx[(orient + 1) % 3] to dest = vtx[(orient + 2) % 3]; apex = vtx[orient].  It packs into
 * one int as (tri << 2) | orient, the encoding of gpudtTriangle::tri[], so -1 (a hull edge)
 * decodes to tri = -1.
 *   lnext / lprev   the next / previous edge of the same triangle (counter-clockwise / clockwise)
 *   sym             the same edge seen from the neighbouring triangle (tri < 0 on the hull)
 *   onext / oprev   the next edge around org counter-clockwise / clockwise
 *   setorg / setdest / setapex, bond   write a vertex, or link two oriented triangles both ways
 * As in the reference, sym, symself, onext, onextself, oprev and oprevself assign the packed
 * neighbour to a variable named `ptr` (an int) that the caller declares.  Each macro here is one
 * expression, so it also works as the body of an unbraced if or loop. */
typedef struct {
    int tri;
    int orient;
} gpudtOrientedTriangle;

static const int gpudt_plus1mod3[3] = {1, 2, 0};
static const int gpudt_minus1mod3[3] = {2, 0, 1};

#define Encode(otri) (((otri).tri << 2) | (otri).orient)
#define decode(code, otri) ((otri).orient = (int)((unsigned int)(code) & 3u), (otri).tri = (code) >> 2)
#define otricopy(from, to) ((to).tri = (from).tri, (to).orient = (from).orient)
#define otriequal(a, b) ((a).tri == (b).tri && (a).orient == (b).orient)

#define org(triList, otri, v) ((v) = (triList)[(otri).tri].vtx[gpudt_plus1mod3[(otri).orient]])
#define dest(triList, otri, v) ((v) = (triList)[(otri).tri].vtx[gpudt_minus1mod3[(otri).orient]])
#define apex(triList, otri, v) ((v) = (triList)[(otri).tri].vtx[(otri).orient])
#define setorg(triList, otri, v) ((triList)[(otri).tri].vtx[gpudt_plus1mod3[(otri).orient]] = (v))
#define setdest(triList, otri, v) ((triList)[(otri).tri].vtx[gpudt_minus1mod3[(otri).orient]] = (v))
#define setapex(triList, otri, v) ((triList)[(otri).tri].vtx[(otri).orient] = (v))

#define lnext(from, to) ((to).tri = (from).tri, (to).orient = gpudt_plus1mod3[(from).orient])
#define lnextself(otri) ((otri).orient = gpudt_plus1mod3[(otri).orient])
#define lprev(from, to) ((to).tri = (from).tri, (to).orient = gpudt_minus1mod3[(from).orient])
#define lprevself(otri) ((otri).orient = gpudt_minus1mod3[(otri).orient])

#define sym(triList, from, to) (ptr = (triList)[(from).tri].tri[(from).orient], decode(ptr, to))
#define symself(triList, otri) (ptr = (triList)[(otri).tri].tri[(otri).orient], decode(ptr, otri))
#define onext(triList, from, to) (lprev(from, to), symself(triList, to))
#define onextself(triList, otri) (lprevself(otri), symself(triList, otri))
#define oprev(triList, from, to) (sym(triList, from, to), lnextself(to))
#define oprevself(triList, otri) (symself(triList, otri), lnextself(otri))

/* links a and b across their shared edge; a side with tri < 0 (the hull) is not written */
#define bond(triList, a, b)                                                                   \
    (((a).tri >= 0 ? ((triList)[(a).tri].tri[(a).orient] = Encode(b)) : 0),                  \
     ((b).tri >= 0 ? ((triList)[(b).tri].tri[(b).orient] = Encode(a)) : 0))

#ifdef __cplusplus
extern "C" {
#endif

/* Alpha shapes of a DT returned by gpudtComputeDT (AlphaShape/headers.h:48-90 layouts).
 * Ranks index spectrum[]; rank 0 is the attached/void slot (spectrum[0] = 0). */
typedef struct {
    int nPoints, nEdges, nTris;
    int specSize;
    double *spectrum;        /* [specSize] unique exact values, ascending */
    int *tri_ro_rank;        /* [nTris] */
    int *edge_t;             /* [nEdges] owning triangle (triangle t owns the edge opposite vtx[i]
                                iff vtx[i+1] < vtx[i+2] or tri[i] == -1; edges ordered by (t, i)) */
    int *edge_vopp;          /* [nEdges] i */
    int *edge_ro_rank;       /* [nEdges] 0 when attached */
    int *edge_nu_min_rank;   /* [nEdges] */
    int *edge_nu_max_rank;   /* [nEdges] 0 on the hull */
    int *vert_nu_min_rank;   /* [nPoints] -1 for dropped duplicates */
    int *vert_nu_max_rank;   /* [nPoints] */
    int masterSize;
    int *master_simplex;     /* [masterSize] */
    int *master_st;          /* 0 edge, 1 triangle, 2 vertex (simplex_type) */
    int *master_at;          /* 0 ro, 1 nu_min, 2 nu_max (alpha_type) */
    int nSpecPointers;
    int *spec_pointers;      /* [nSpecPointers] [0] = -1, then first master index of each rank */
} GPUDTALPHA, *PGPUDTALPHA;

/* Replaces Alpha_Shapes_2D(pInput, pDT) + computeAlphaShapes() (AlphaShape/Alpha_Shapes_2D.cu:58-74,
 * 206-233).  pDT may be any triangle array of the points in the gpudtComputeDT layout: it is checked
 * on the device first (distinct vertex ids in [0, nPoints), neighbours -1 or a triangle slot that
 * links back over the same edge), and one that fails returns -1 with a message, before any kernel
 * follows its links.  Returns 0 and sets *pAlpha (free with gpudtReleaseAlpha). */
int gpudtComputeAlpha(PGPUDTPARAMS pParam, PGPUDTOUTPUT pDT, PGPUDTALPHA *pAlpha);
void gpudtReleaseAlpha(PGPUDTALPHA pAlpha);

/* ---- context API (no reference counterpart) ------------------------------------------- */
typedef struct galpha_ctx galpha_ctx;

/* stats of the last run on a context */
typedef struct {
    int nTris, nEdges, specSize;
    int islandSites, pbaIterations;
    int reconTris, shiftBad, deleted, missing, inserted, duplicates, insertRounds;
    int flipPasses;
    long long flips;
    long long incircleTests, incircleExact, orientExact;
    long long alphaExactCompares, alphaExactAttach, alphaClusters, alphaMaxCluster;
    double ms_pba, ms_recon, ms_shift, ms_missing, ms_boundary, ms_flip, ms_canon, ms_alpha, ms_total;
    /* work counts of the last run: frame vertices examined by the frame-removal kernel, sum of the
       fan lengths of removed frame vertices, triangles listed for the flip tests */
    long long frameSeen, frameChain, flipListed;
    /* 1 when the frame was removed in bulk (one boundary cycle per set), 0 round by round */
    int frameBulk, earRounds;
    /* constrained runs: cavity insertion rounds, and the time of the constraint stage (insertion
       and the constrained flips) */
    int constraintRounds;
    double ms_constraints;
    /* alpha sub-stages: edges, rho and attachment; spectrum order, ranks and values; intervals
       and the master list */
    double ms_alpha_rho, ms_alpha_order, ms_alpha_lists;
    /* per mesh stage (0: shifting, missing-site insertion and the flips after them; 1: frame
       removal and the final flips; 2: constraints and canonical output): triangles listed for
       flip tests, incircle tests, exact incircle and exact orientation evaluations */
    long long stageListed[3], stageIncircleTests[3], stageIncircleExact[3], stageOrientExact[3];
    /* flip passes run back to back by the single-workgroup tail on short lists, and the triangles
       they listed (flipPasses and flipListed include both) */
    int flipTailPasses;
    long long flipTailListed;
} galpha_stats;

galpha_ctx *galphaCreate(int device);
void galphaDestroy(galpha_ctx *ctx);
const char *galphaLastError(galpha_ctx *ctx);

/* Host-to-host: DT (+ alpha when alpha_out != NULL).  tris_out9: nTris * 9 ints in the
 * gpudtTriangle layout; *ntris_out on input is its capacity (in triangles). */
int galphaComputeHost(galpha_ctx *ctx, const double *xy, int n, int fboSize,
                      int *tris_out9, int *ntris_out);

/* Device-resident: d_xy is a device pointer to n (x, y) doubles.  The triangulation (and the
 * alpha spectrum when with_alpha) stays on the device; query with the getters below.
 * stream is a hipStream_t (NULL = the context's own stream). */
int galphaComputeDevice(galpha_ctx *ctx, const double *d_xy, int n, int fboSize, int with_alpha,
                        void *stream);
/* Constraint edges (n pairs of point indices, the layout of GPUDTPARAMS::constraints) for the
 * following single-set runs of this context (galphaComputeHost / galphaComputeDevice), until set
 * again; n = 0 clears them.  Batches and serving reject a context that holds constraints. */
int galphaSetConstraints(galpha_ctx *ctx, const int *pairs, int n);
/* device pointers to the last result: triangles in the 9-int layout, nTris */
int galphaDeviceTriangles(galpha_ctx *ctx, const int **d_tris9, int *ntris);
int galphaCopyTriangles(galpha_ctx *ctx, int *tris_out9, int capacity);
int galphaGetStats(galpha_ctx *ctx, galpha_stats *stats);

/* Discrete Voronoi texture only (splat + PBA + island certificate), for checking the
 * diagram: labels[T*T] in x-major order (pixel (x, y) at x*T + y), pix[n + frame] packed
 * (px << 16 | py); returns the number of frame sites or a negative error. */
int galphaVoronoi(galpha_ctx *ctx, const double *xy, int n, int fboSize, int *labels, int *pix,
                  int pix_capacity);

/* Copy the alpha result of the last galphaComputeDevice(with_alpha=1) run to host arrays
 * allocated by the library (release with gpudtReleaseAlpha). */
int galphaCopyAlpha(galpha_ctx *ctx, PGPUDTALPHA *pAlpha);

/* Alpha-complex queries on the last single-set alpha result (the reference keeps an alpha value
 * with setAlpha/getAlpha, AlphaShape/Alpha_Shapes_2D.cu:739-747, and leaves the query to callers
 * walking its master list).  galphaAlphaRank: *rank = the largest r with spectrum[r] <= alpha
 * (-1 when none; the FP64 spectrum values need not be monotone).  galphaClassify: the class of every vertex (nPoints), edge (nEdges, in
 * the edge order of the alpha result) and triangle (nTris) at rank r: 0 exterior, 1 singular,
 * 2 regular, 3 interior; a NULL array is skipped.  galphaClassifyDevice writes device arrays on
 * stream (NULL = the context's stream). */
int galphaAlphaRank(galpha_ctx *ctx, double alpha, int *rank);
int galphaClassify(galpha_ctx *ctx, int rank, unsigned char *vert_class, unsigned char *edge_class,
                   unsigned char *tri_class);
int galphaClassifyDevice(galpha_ctx *ctx, int rank, unsigned char *d_vert_class, unsigned char *d_edge_class,
                         unsigned char *d_tri_class, void *stream);

/* ---- Delaunay without device allocation (no reference counterpart) ------------------------
 * For callers that cannot allocate device memory: code between HIP graph captures, tight serving
 * loops, harnesses that forbid hipMalloc.  Same triangulation as galphaComputeDevice with
 * with_alpha = 0, but every device buffer of the run is carved from a static device array of
 * GALPHA_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the process ends;
 * the call makes no hipMalloc, hipFree or other device or pinned allocation.  d_xy: n (x, y)
 * doubles on `device`; d_tris9: device array of `capacity` records in the gpudtTriangle layout
 * (9 ints each; 2 n records always suffice).  stream: a hipStream_t of `device` (NULL: the null
 * stream); the call returns when its device work has finished.  Returns the triangle count, or
 * -1 with gpudtGetLastError() set, e.g. when the run could need more than the scratch holds
 * (textures up to 1024^2 with up to 65,536 points always fit).  Calls on one GPU share its
 * scratch and run one at a time (a per-device lock). */
#define GALPHA_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int galphaComputeDeviceNoAlloc(int device, const double *d_xy, int n, int fboSize, int *d_tris9, int capacity,
                               void *stream);

/* ---- batches and serving (no reference counterpart) ----------------------------------- */
/* Batched: nsets independent point sets, concatenated in xy (offsets[nsets+1] in points),
 * texture side per set (NULL or 0 = autotuned).  All sets go through one pass of the pipeline
 * (every kernel launch covers every set).  Results: triangles of all sets concatenated with
 * set-local vertex and triangle indices, tri_offsets[nsets+1]; both malloc'ed (galphaFree).
 * Each set's result is identical to a single-set run on the same points. */
int galphaBatch(galpha_ctx *ctx, const double *xy, const int *offsets, int nsets, const int *fboSizes,
                int with_alpha, int **tris9_out, int **tri_offsets_out);
/* Device-resident batch: d_xy is a device pointer; results stay on the device.  Returns the
 * total number of triangles.  Query per set with galphaBatchTriangles / galphaBatchCopyAlpha. */
int galphaBatchDevice(galpha_ctx *ctx, const double *d_xy, const int *offsets, int nsets, const int *fboSizes,
                      int with_alpha, void *stream);
/* triangles of input set `set` of the last batch with set-local indices (NULL tris9: count only) */
int galphaBatchTriangles(galpha_ctx *ctx, int set, int *tris9, int capacity);
/* alpha result of input set `set` of the last batch run with alpha (release with gpudtReleaseAlpha) */
int galphaBatchCopyAlpha(galpha_ctx *ctx, int set, PGPUDTALPHA *pAlpha);
void galphaFree(void *p);

/* Serving: `slots` worker threads, each with its own context and HIP stream (buffers stay
 * resident).  Requests queue in arrival order; a free slot admits the longest prefix of the
 * queue that fits hbm_budget_bytes (bytes held by all slots, its own grown to the admitted
 * requests' footprint), up to max_group requests, and runs them as one batched pass.  Results
 * are copied to host memory owned by the ticket.  Every result equals a single-set run. */
typedef struct galpha_server galpha_server;
galpha_server *galphaServerCreate(int device, int slots, size_t hbm_budget_bytes, int max_group);
void galphaServerDestroy(galpha_server *server);
/* returns a ticket (> 0), -2 when the request alone exceeds the budget, -1 on bad arguments */
long long galphaServerSubmit(galpha_server *server, const double *xy, int n, int fboSize, int with_alpha);
/* blocks until the ticket is done; returns its triangle count (< 0 on failure); latency_ms =
 * submit -> done, queue_ms = submit -> start of its pass, group = requests in that pass */
int galphaServerWait(galpha_server *server, long long ticket, double *latency_ms, double *queue_ms, int *group);
int galphaServerTriangles(galpha_server *server, long long ticket, int *tris9, int capacity);
int galphaServerAlpha(galpha_server *server, long long ticket, PGPUDTALPHA *pAlpha);
int galphaServerRelease(galpha_server *server, long long ticket);
int galphaServerStats(galpha_server *server, long long *passes, long long *served, size_t *held_bytes);
/* device bytes the admission rule counts for one request of n points (fboSize 0: the default
 * texture for n): textures, mesh and alpha buffers of a context that runs it alone */
size_t galphaServerFootprint(int n, int fboSize);
/* At most max_points points per pass (0: no limit, the default); a larger request still runs,
 * alone.  Shorter passes free slots sooner, so small requests wait less behind large ones. */
int galphaServerSetGroupPoints(galpha_server *server, long long max_points);

/* Kernel timing probe: while enabled, HIP events are recorded on the launch stream around every
 * launch of the named kernel; NULL disables.  Result: number of launches and total ms. */
int galphaKernelProbe(const char *kernel_name);
int galphaKernelProbeResult(int *launches, double *total_ms);

/* default texture side for n points (at least 16 pixels per point, power of two >= 256) */
int galphaAutoTexture(int n);
/* texture side a context picks for n points: its autotuned pixels-per-point density for n's size
 * class (< 4096, < 32768, < 262144, larger points), 16 until tuned; used for every set whose
 * texture is not given */
int galphaTextureFor(galpha_ctx *ctx, int n);
/* Texture autotuner: times the sample sets of each size class as one batch at densities
 * {6, 9, 12, 16, 24, 32} pixels per point (median of reps runs) and keeps the fastest per class.
 * densities_out (optional): the 4 chosen densities.  Results do not depend on the texture. */
int galphaAutotune(galpha_ctx *ctx, const double *xy, const int *offsets, int nsets, int reps,
                   double *densities_out);

/* ---- device memory (no reference counterpart) -------------------------------------------
 * For callers of the device-resident entry points that have no HIP runtime of their own (ctypes,
 * plain C): the library's runtime allocates, copies and synchronizes.  galphaDeviceCopy kind:
 * 0 host to device, 1 device to host, 2 device to device; synchronous.  Errors: NULL / -1 and
 * gpudtGetLastError(). */
int galphaDeviceCount(void);
void *galphaDeviceMalloc(int device, size_t bytes);
void galphaDeviceFree(void *d_ptr);
int galphaDeviceCopy(void *dst, const void *src, size_t bytes, int kind);
int galphaDeviceSynchronize(void);
/* a HIP stream of the library's runtime (non-blocking, default priority) to pass as `stream` */
void *galphaStreamCreate(int device);
int galphaStreamDestroy(void *stream);
int galphaStreamSynchronize(void *stream);

/* ---- formats ---------------------------------------------------------------------------- */
/* Point file of the reference (main.cu:117-161, ReadDataFromFileBinary): "N", then N lines
 * "idx x y", point idx stored at points[idx]; min/max computed as there (main.cu:139-159).
 * *pParams is library-allocated (release with galphaReleaseParams), nConstraints = 0 and
 * fboSize = galphaAutoTexture(N).  Unlike the reference, a missing file, a short file, a bad
 * number, an index outside [0, N) or a point with no line is an error (-1, gpudtGetLastError). */
int galphaReadPoints(const char *path, PGPUDTPARAMS *pParams);
void galphaReleaseParams(PGPUDTPARAMS pParams);
/* Writes points in the same format with 17 significant digits, so reading gives the same doubles. */
int galphaWritePoints(const char *path, const gpudtVertex *points, int nPoints);
/* Alpha_Shapes_2D::dumpValues (AlphaShape/Alpha_Shapes_2D.cu:749-793) to path: the spectrum size
 * line.  full != 0 adds the sections that function keeps commented out (VERTICES: x y nu_min
 * nu_max; EDGES: x1 y1 x2 y2 ro nu_min nu_max; TRIANGLES: x0 y0 x1 y1 x2 y2 ro; values are
 * spectrum[rank], "nan" for rank -1) before the size line and the spectrum values after it, all
 * with 17 significant digits (the reference's stream default is 6). */
int galphaDumpValues(const char *path, PGPUDTPARAMS pParams, PGPUDTOUTPUT pDT, PGPUDTALPHA pAlpha, int full);
/* Binary result payload (no reference counterpart): a 40-byte header (magic "G2DR", version 1,
 * flags bit 0 = alpha present, nPoints, nTris, nEdges, specSize, masterSize, nSpecPointers, 0),
 * the arrays of the triangles and of the alpha result, each padded to 8 bytes, then a 64-bit
 * FNV-1a hash of all bytes before it; little-endian.  galphaPackResult returns the payload size
 * and writes it when capacity suffices (alpha may be NULL); -1 on error.  galphaUnpackResult
 * checks the header, size and hash and allocates *tris9 (galphaFree) and *pAlpha
 * (gpudtReleaseAlpha; NULL when the payload has no alpha section). */
long long galphaPackResult(int nPoints, const int *tris9, int nTris, PGPUDTALPHA alpha, void *buf,
                           long long capacity);
int galphaUnpackResult(const void *buf, long long size, int *nPoints, int **tris9, int *nTris,
                       PGPUDTALPHA *pAlpha);

#ifdef __cplusplus
}
#endif
#endif
