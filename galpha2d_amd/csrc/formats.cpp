// On-disk and wire formats: the reference's point file, its dumpValues text output, and a
// binary result payload (triangles + alpha arrays) for files and serving responses.
//
//   point file   main.cu:117-161 (ReadDataFromFileBinary): "N", then N lines "idx x y";
//                point idx is stored at points[idx]; min/max are computed from the points.
//   dumpValues   AlphaShape/Alpha_Shapes_2D.cu:749-793: the active output is one line with the
//                spectrum size; the sections VERTICES / EDGES / TRIANGLES it keeps commented out
//                are written on request, with round-trip precision.
//   payload      no counterpart in the reference (its results stay in process memory).
#include "../../include/gpudt.h"
#include "alpha_host.h"

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct File {
    FILE *f = nullptr;
    explicit File(FILE *p) : f(p) {}
    ~File()
    {
        if (f) fclose(f);
    }
};

// next whitespace-separated token of a text file; false at end of file
bool next_token(FILE *f, std::string &tok)
{
    tok.clear();
    int ch;
    while ((ch = fgetc(f)) != EOF && (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f')) {}
    if (ch == EOF) return false;
    do {
        tok.push_back((char)ch);
    } while ((ch = fgetc(f)) != EOF && !(ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\v' || ch == '\f'));
    return true;
}

long long parse_int(const std::string &t, const char *what)
{
    char *end = nullptr;
    errno = 0;
    long long v = strtoll(t.c_str(), &end, 10);
    if (errno || end == t.c_str() || *end) throw GalphaError(std::string("point file: bad ") + what + " '" + t + "'");
    return v;
}

double parse_real(const std::string &t)
{
    char *end = nullptr;
    double v = strtod(t.c_str(), &end);
    if (end == t.c_str() || *end) throw GalphaError("point file: bad coordinate '" + t + "'");
    if (!std::isfinite(v)) throw GalphaError("point file: coordinate is not finite '" + t + "'");
    return v;
}

// ---- payload layout ---------------------------------------------------------------------------
// little-endian; header of ten 32-bit words, then the arrays in this order, each padded to 8 bytes:
//   spectrum f64[specSize], triangles i32[9 nTris], tri_ro_rank i32[nTris],
//   edge_t, edge_vopp, edge_ro_rank, edge_nu_min_rank, edge_nu_max_rank i32[nEdges] each,
//   vert_nu_min_rank, vert_nu_max_rank i32[nPoints] each,
//   master_simplex, master_st, master_at i32[masterSize] each, spec_pointers i32[nSpecPointers];
// then a 64-bit FNV-1a hash of every byte before it.  Without alpha (flags bit 0 clear) only the
// triangles follow the header, and the alpha counts are zero.
const uint32_t PAYLOAD_MAGIC = 0x52443247u;   // "G2DR"
const uint32_t PAYLOAD_VERSION = 1;
const size_t HEADER_BYTES = 40;

size_t pad8(size_t b) { return (b + 7) & ~(size_t)7; }

uint64_t fnv1a(const unsigned char *p, size_t n)
{
    uint64_t h = 0xcbf29ce484222325ull;   // FNV-1a 64-bit offset basis
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 0x100000001b3ull;   // FNV 64-bit prime
    }
    return h;
}

struct Counts {
    int32_t nPoints, nTris, nEdges, specSize, masterSize, nSpecPointers;
    bool alpha;
};

// byte size of each array section, in payload order
std::vector<size_t> section_bytes(const Counts &c)
{
    std::vector<size_t> s;
    if (c.alpha) s.push_back(8 * (size_t)c.specSize);
    s.push_back(36 * (size_t)c.nTris);
    if (c.alpha) {
        s.push_back(4 * (size_t)c.nTris);
        for (int k = 0; k < 5; k++) s.push_back(4 * (size_t)c.nEdges);
        for (int k = 0; k < 2; k++) s.push_back(4 * (size_t)c.nPoints);
        for (int k = 0; k < 3; k++) s.push_back(4 * (size_t)c.masterSize);
        s.push_back(4 * (size_t)c.nSpecPointers);
    }
    return s;
}

size_t payload_size(const Counts &c)
{
    size_t b = HEADER_BYTES;
    for (size_t s : section_bytes(c)) b += pad8(s);
    return b + 8;
}

void put32(unsigned char *p, uint32_t v)
{
    for (int k = 0; k < 4; k++) p[k] = (unsigned char)(v >> (8 * k));
}

uint32_t get32(const unsigned char *p)
{
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

void put64(unsigned char *p, uint64_t v)
{
    for (int k = 0; k < 8; k++) p[k] = (unsigned char)(v >> (8 * k));
}

uint64_t get64(const unsigned char *p)
{
    uint64_t v = 0;
    for (int k = 7; k >= 0; k--) v = v << 8 | p[k];
    return v;
}

// copies n 32-bit (or 64-bit) values in little-endian order (a plain copy on this host)
void put_words(unsigned char *dst, const void *src, size_t bytes) { if (bytes) memcpy(dst, src, bytes); }

bool host_little_endian()
{
    uint32_t one = 1;
    unsigned char b;
    memcpy(&b, &one, 1);
    return b == 1;
}

}  // namespace

extern "C" void galpha_set_last_error(const char *msg);   // galpha.cpp (library-internal)

extern "C" {

int galphaReadPoints(const char *path, PGPUDTPARAMS *pParams)
{
    try {
        if (!path || !pParams) throw GalphaError("null argument");
        *pParams = nullptr;
        File f(fopen(path, "rb"));
        if (!f.f) throw GalphaError(std::string("cannot open point file '") + path + "': " + strerror(errno));
        std::string tok;
        if (!next_token(f.f, tok)) throw GalphaError("point file: empty");
        long long n = parse_int(tok, "point count");
        if (n < 0 || n > 0x7fffffffLL) throw GalphaError("point file: point count out of range");
        std::vector<gpudtVertex> pts((size_t)n);
        std::vector<unsigned char> seen((size_t)n, 0);
        for (long long i = 0; i < n; i++) {
            std::string ti, tx, ty;
            if (!next_token(f.f, ti) || !next_token(f.f, tx) || !next_token(f.f, ty))
                throw GalphaError("point file: " + std::to_string(n) + " points announced, " + std::to_string(i) +
                                  " read");
            long long idx = parse_int(ti, "point index");
            if (idx < 0 || idx >= n)
                throw GalphaError("point file: index " + std::to_string(idx) + " outside [0, " + std::to_string(n) + ")");
            pts[(size_t)idx].x = parse_real(tx);
            pts[(size_t)idx].y = parse_real(ty);
            seen[(size_t)idx] = 1;
        }
        for (long long i = 0; i < n; i++)
            if (!seen[(size_t)i]) throw GalphaError("point file: no line for point " + std::to_string(i));
        PGPUDTPARAMS p = (PGPUDTPARAMS)calloc(1, sizeof(GPUDTPARAMS));
        if (!p) throw GalphaError("host allocation failed");
        p->points = (gpudtVertex *)malloc(sizeof(gpudtVertex) * (size_t)(n > 0 ? n : 1));
        if (!p->points) {
            free(p);
            throw GalphaError("host allocation failed");
        }
        if (n) memcpy(p->points, pts.data(), sizeof(gpudtVertex) * (size_t)n);
        p->nPoints = (int)n;
        // same bounds as main.cu:139-159 (start at +-1E20, strict comparisons)
        double minx = 1e20, miny = 1e20, maxx = -1e20, maxy = -1e20;
        for (long long i = 0; i < n; i++) {
            if (pts[(size_t)i].x > maxx) maxx = pts[(size_t)i].x;
            if (pts[(size_t)i].x < minx) minx = pts[(size_t)i].x;
            if (pts[(size_t)i].y > maxy) maxy = pts[(size_t)i].y;
            if (pts[(size_t)i].y < miny) miny = pts[(size_t)i].y;
        }
        p->minX = minx; p->minY = miny; p->maxX = maxx; p->maxY = maxy;
        p->nConstraints = 0;
        p->constraints = nullptr;
        p->fboSize = galphaAutoTexture((int)n);
        *pParams = p;
        return 0;
    } catch (const std::exception &e) {
        galpha_set_last_error(e.what());
        return -1;
    }
}

void galphaReleaseParams(PGPUDTPARAMS p)
{
    if (!p) return;
    free(p->points);
    free(p->constraints);
    free(p);
}

int galphaWritePoints(const char *path, const gpudtVertex *points, int nPoints)
{
    try {
        if (!path || nPoints < 0 || (nPoints > 0 && !points)) throw GalphaError("invalid argument");
        File f(fopen(path, "wb"));
        if (!f.f) throw GalphaError(std::string("cannot create '") + path + "': " + strerror(errno));
        fprintf(f.f, "%d\n", nPoints);
        for (int i = 0; i < nPoints; i++) {
            if (!std::isfinite(points[i].x) || !std::isfinite(points[i].y))
                throw GalphaError("point " + std::to_string(i) + " is not finite");
            fprintf(f.f, "%d %.17g %.17g\n", i, points[i].x, points[i].y);
        }
        if (ferror(f.f)) throw GalphaError("write failed");
        return 0;
    } catch (const std::exception &e) {
        galpha_set_last_error(e.what());
        return -1;
    }
}

int galphaDumpValues(const char *path, PGPUDTPARAMS pParams, PGPUDTOUTPUT pDT, PGPUDTALPHA A, int full)
{
    try {
        if (!path || !A) throw GalphaError("null argument");
        if (full && (!pParams || !pDT || !pParams->points)) throw GalphaError("full dump needs the points and the DT");
        if (full && (pParams->nPoints != A->nPoints || pDT->nTris != A->nTris))
            throw GalphaError("alpha result does not belong to these points / this DT");
        File f(fopen(path, "wb"));
        if (!f.f) throw GalphaError(std::string("cannot create '") + path + "': " + strerror(errno));
        auto val = [&](int r) {
            // rank -1 (a point outside the triangulation) has no value
            if (r < 0 || r >= A->specSize) {
                fprintf(f.f, " nan");
                return;
            }
            fprintf(f.f, " %.17g", A->spectrum[r]);
        };
        if (full) {
            const gpudtVertex *Pts = pParams->points;
            const gpudtTriangle *T = pDT->triangles;
            const int np = pParams->nPoints;
            auto P = [&](int v) -> const gpudtVertex & {
                if (v < 0 || v >= np) throw GalphaError("triangle vertex " + std::to_string(v) + " outside the points");
                return Pts[v];
            };
            fprintf(f.f, "VERTICES\n");
            for (int v = 0; v < A->nPoints; v++) {
                fprintf(f.f, "%.17g %.17g", P(v).x, P(v).y);
                val(A->vert_nu_min_rank[v]);
                val(A->vert_nu_max_rank[v]);
                fputc('\n', f.f);
            }
            fprintf(f.f, "EDGES\n");
            for (int e = 0; e < A->nEdges; e++) {
                int t = A->edge_t[e], i = A->edge_vopp[e];
                if (t < 0 || t >= pDT->nTris || i < 0 || i > 2) throw GalphaError("edge " + std::to_string(e) + " is invalid");
                int p1 = T[t].vtx[(i + 1) % 3], p2 = T[t].vtx[(i + 2) % 3];
                fprintf(f.f, "%.17g %.17g %.17g %.17g", P(p1).x, P(p1).y, P(p2).x, P(p2).y);
                val(A->edge_ro_rank[e]);
                val(A->edge_nu_min_rank[e]);
                val(A->edge_nu_max_rank[e]);
                fputc('\n', f.f);
            }
            fprintf(f.f, "TRIANGLES\n");
            for (int t = 0; t < A->nTris; t++) {
                for (int k = 0; k < 3; k++) fprintf(f.f, k ? " %.17g %.17g" : "%.17g %.17g", P(T[t].vtx[k]).x, P(T[t].vtx[k]).y);
                val(A->tri_ro_rank[t]);
                fputc('\n', f.f);
            }
        }
        fprintf(f.f, "%d\n", A->specSize);
        if (full)
            for (int r = 0; r < A->specSize; r++) fprintf(f.f, "%.17g\n", A->spectrum[r]);
        if (ferror(f.f)) throw GalphaError("write failed");
        return 0;
    } catch (const std::exception &e) {
        galpha_set_last_error(e.what());
        return -1;
    }
}

long long galphaPackResult(int nPoints, const int *tris9, int nTris, PGPUDTALPHA A, void *buf, long long capacity)
{
    try {
        if (!host_little_endian()) throw GalphaError("payloads need a little-endian host");
        if (nPoints < 0 || nTris < 0 || (nTris > 0 && !tris9)) throw GalphaError("invalid triangles");
        Counts c{nPoints, nTris, 0, 0, 0, 0, A != nullptr};
        if (A) {
            if (A->nPoints != nPoints || A->nTris != nTris) throw GalphaError("alpha result does not match the triangles");
            c.nEdges = A->nEdges; c.specSize = A->specSize; c.masterSize = A->masterSize;
            c.nSpecPointers = A->nSpecPointers;
            if (c.nEdges < 0 || c.specSize < 0 || c.masterSize < 0 || c.nSpecPointers < 0)
                throw GalphaError("invalid alpha counts");
        }
        size_t need = payload_size(c);
        if (!buf || capacity < (long long)need) return (long long)need;
        unsigned char *o = (unsigned char *)buf;
        memset(o, 0, need);
        uint32_t hdr[10] = {PAYLOAD_MAGIC, PAYLOAD_VERSION, c.alpha ? 1u : 0u, (uint32_t)c.nPoints, (uint32_t)c.nTris,
                            (uint32_t)c.nEdges, (uint32_t)c.specSize, (uint32_t)c.masterSize, (uint32_t)c.nSpecPointers, 0};
        for (int k = 0; k < 10; k++) put32(o + 4 * k, hdr[k]);
        std::vector<const void *> src;
        if (c.alpha) src.push_back(A->spectrum);
        src.push_back(tris9);
        if (c.alpha) {
            const void *rest[] = {A->tri_ro_rank, A->edge_t, A->edge_vopp, A->edge_ro_rank, A->edge_nu_min_rank,
                                  A->edge_nu_max_rank, A->vert_nu_min_rank, A->vert_nu_max_rank, A->master_simplex,
                                  A->master_st, A->master_at, A->spec_pointers};
            for (const void *r : rest) src.push_back(r);
        }
        std::vector<size_t> sb = section_bytes(c);
        size_t off = HEADER_BYTES;
        for (size_t k = 0; k < sb.size(); k++) {
            if (sb[k] && !src[k]) throw GalphaError("alpha result has a missing array");
            put_words(o + off, src[k], sb[k]);
            off += pad8(sb[k]);
        }
        put64(o + off, fnv1a(o, off));
        return (long long)need;
    } catch (const std::exception &e) {
        galpha_set_last_error(e.what());
        return -1;
    }
}

int galphaUnpackResult(const void *buf, long long size, int *nPoints, int **tris9, int *nTris, PGPUDTALPHA *pAlpha)
{
    int *tris = nullptr;
    PGPUDTALPHA A = nullptr;
    try {
        if (!buf || !nPoints || !tris9 || !nTris || !pAlpha) throw GalphaError("null argument");
        *tris9 = nullptr;
        *pAlpha = nullptr;
        if (!host_little_endian()) throw GalphaError("payloads need a little-endian host");
        const unsigned char *b = (const unsigned char *)buf;
        if (size < (long long)HEADER_BYTES + 8) throw GalphaError("payload: truncated header");
        if (get32(b) != PAYLOAD_MAGIC) throw GalphaError("payload: bad magic");
        if (get32(b + 4) != PAYLOAD_VERSION) throw GalphaError("payload: unsupported version " + std::to_string(get32(b + 4)));
        uint32_t flags = get32(b + 8);
        if (flags > 1) throw GalphaError("payload: unknown flags");
        Counts c{(int32_t)get32(b + 12), (int32_t)get32(b + 16), (int32_t)get32(b + 20), (int32_t)get32(b + 24),
                 (int32_t)get32(b + 28), (int32_t)get32(b + 32), (flags & 1) != 0};
        if (c.nPoints < 0 || c.nTris < 0 || c.nEdges < 0 || c.specSize < 0 || c.masterSize < 0 || c.nSpecPointers < 0)
            throw GalphaError("payload: negative count");
        if (!c.alpha && (c.nEdges || c.specSize || c.masterSize || c.nSpecPointers))
            throw GalphaError("payload: alpha counts without an alpha section");
        size_t need = payload_size(c);
        if ((long long)need != size) throw GalphaError("payload: size " + std::to_string(size) + ", counts need " + std::to_string(need));
        if (get64(b + need - 8) != fnv1a(b, need - 8)) throw GalphaError("payload: checksum mismatch");
        std::vector<size_t> sb = section_bytes(c);
        std::vector<void **> dst;
        A = c.alpha ? (PGPUDTALPHA)calloc(1, sizeof(GPUDTALPHA)) : nullptr;
        if (c.alpha && !A) throw GalphaError("host allocation failed");
        if (A) dst.push_back((void **)&A->spectrum);
        dst.push_back((void **)&tris);
        if (A) {
            void **rest[] = {(void **)&A->tri_ro_rank, (void **)&A->edge_t, (void **)&A->edge_vopp, (void **)&A->edge_ro_rank,
                             (void **)&A->edge_nu_min_rank, (void **)&A->edge_nu_max_rank, (void **)&A->vert_nu_min_rank,
                             (void **)&A->vert_nu_max_rank, (void **)&A->master_simplex, (void **)&A->master_st,
                             (void **)&A->master_at, (void **)&A->spec_pointers};
            for (void **r : rest) dst.push_back(r);
        }
        size_t off = HEADER_BYTES;
        for (size_t k = 0; k < sb.size(); k++) {
            *dst[k] = malloc(sb[k] ? sb[k] : 1);
            if (!*dst[k]) throw GalphaError("host allocation failed");
            if (sb[k]) memcpy(*dst[k], b + off, sb[k]);
            off += pad8(sb[k]);
        }
        if (A) {
            A->nPoints = c.nPoints; A->nEdges = c.nEdges; A->nTris = c.nTris; A->specSize = c.specSize;
            A->masterSize = c.masterSize; A->nSpecPointers = c.nSpecPointers;
        }
        *nPoints = c.nPoints;
        *nTris = c.nTris;
        *tris9 = tris;
        *pAlpha = A;
        return 0;
    } catch (const std::exception &e) {
        free(tris);
        alpha_free_host(A);
        galpha_set_last_error(e.what());
        return -1;
    }
}

}  // extern "C"
