// galpha_cli: command-line driver over the C ABI, the counterpart of the reference's main.cu
// (main.cu:364-440: read a point file, gpudtComputeDT, Alpha_Shapes_2D::computeAlphaShapes,
// print the time, checkResult).  The file name is an argument instead of a prompt.
//
//   galpha_cli POINTS.txt [--fbo N] [--no-alpha] [--check] [--repeat K]
//              [--dump FILE] [--dump-full FILE] [--pack FILE]
//
// --check runs the structural part of main.cu's checkResult (main.cu:163-343) on the host: every
// neighbour link is symmetric, every point that is not an exact duplicate is a vertex, and the
// orientation and the empty-circle property hold wherever the FP64 filter is certain (Shewchuk's
// static error bounds).  Cases the filter leaves open are counted, not decided: the library's own
// predicates are exact, so this check is a sanity check of the output, not a proof.
#include "../../include/gpudt.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

const double EPS = 1.1102230246251565e-16;   // 2^-53
const double CCW_BOUND = (3.0 + 16.0 * EPS) * EPS;
const double ICC_BOUND = (10.0 + 96.0 * EPS) * EPS;

// +1 / -1 when the FP64 sign is certain, 0 when the filter cannot decide
int orient_filtered(const gpudtVertex &a, const gpudtVertex &b, const gpudtVertex &c)
{
    double l = (a.x - c.x) * (b.y - c.y), r = (a.y - c.y) * (b.x - c.x), det = l - r;
    double bound = CCW_BOUND * (std::fabs(l) + std::fabs(r));
    return det > bound ? 1 : (det < -bound ? -1 : 0);
}

int incircle_filtered(const gpudtVertex &a, const gpudtVertex &b, const gpudtVertex &c, const gpudtVertex &d)
{
    double adx = a.x - d.x, ady = a.y - d.y, bdx = b.x - d.x, bdy = b.y - d.y, cdx = c.x - d.x, cdy = c.y - d.y;
    double bc = bdx * cdy - cdx * bdy, ca = cdx * ady - adx * cdy, ab = adx * bdy - bdx * ady;
    double al = adx * adx + ady * ady, bl = bdx * bdx + bdy * bdy, cl = cdx * cdx + cdy * cdy;
    double det = al * bc + bl * ca + cl * ab;
    double perm = (std::fabs(bdx * cdy) + std::fabs(cdx * bdy)) * al + (std::fabs(cdx * ady) + std::fabs(adx * cdy)) * bl +
                  (std::fabs(adx * bdy) + std::fabs(bdx * ady)) * cl;
    double bound = ICC_BOUND * perm;
    return det > bound ? 1 : (det < -bound ? -1 : 0);
}

int check(const PGPUDTPARAMS in, const PGPUDTOUTPUT out)
{
    int F = out->nTris, n = in->nPoints, errors = 0;
    long long uncertain = 0;
    const gpudtTriangle *T = out->triangles;
    const gpudtVertex *P = in->points;
    std::vector<char> used(n, 0);
    for (int t = 0; t < F; t++)
        for (int i = 0; i < 3; i++) {
            int v = T[t].vtx[i];
            if (v < 0 || v >= n) {
                printf("*** ERROR *** triangle %d has vertex %d outside [0, %d)\n", t, v, n);
                return 1;
            }
            used[v] = 1;
            int nb = T[t].tri[i];
            if (nb < 0) continue;
            int u = nb >> 2, j = nb & 3;
            if (u >= F || j > 2 || T[u].tri[j] != (t << 2 | i)) {
                if (errors++ < 10) printf("*** ERROR *** neighbour link %d.%d is not symmetric\n", t, i);
            }
        }
    // a point may only be missing when an earlier point has the same coordinates
    {
        std::vector<int> idx(n);
        for (int i = 0; i < n; i++) idx[i] = i;
        std::sort(idx.begin(), idx.end(), [&](int a, int b) {
            return P[a].x != P[b].x ? P[a].x < P[b].x : (P[a].y != P[b].y ? P[a].y < P[b].y : a < b);
        });
        for (int k = 0; k < n; k++) {
            int v = idx[k];
            bool dup = k > 0 && P[idx[k - 1]].x == P[v].x && P[idx[k - 1]].y == P[v].y;
            if (!used[v] && !dup && n >= 3)
                if (errors++ < 10) printf("*** ERROR *** vertex %d is not in the triangulation\n", v);
        }
    }
    if (errors) {   // the geometric checks below follow the links: only on a sound structure
        printf("check: %d triangles, %d structural errors\n", F, errors);
        return 1;
    }
    for (int t = 0; t < F; t++) {
        int o = orient_filtered(P[T[t].vtx[0]], P[T[t].vtx[1]], P[T[t].vtx[2]]);
        if (o < 0) {
            if (errors++ < 10) printf("*** ERROR *** wrong orientation at triangle %d\n", t);
        } else if (o == 0) {
            uncertain++;
        }
        for (int i = 0; i < 3; i++) {
            int nb = T[t].tri[i];
            if (nb < 0) continue;
            int far = T[nb >> 2].vtx[nb & 3];
            int s = incircle_filtered(P[T[t].vtx[0]], P[T[t].vtx[1]], P[T[t].vtx[2]], P[far]);
            if (s > 0) {
                if (errors++ < 10) printf("*** ERROR *** edge %d.%d is not locally Delaunay\n", t, i);
            } else if (s == 0) {
                uncertain++;
            }
        }
    }
    printf("check: %d triangles, %d errors, %lld predicate evaluations left open by the FP64 filter\n", F, errors,
           uncertain);
    return errors ? 1 : 0;
}

bool write_file(const char *path, const std::vector<char> &b)
{
    FILE *f = fopen(path, "wb");
    if (!f) return false;
    bool ok = fwrite(b.data(), 1, b.size(), f) == b.size();
    return fclose(f) == 0 && ok;
}

int usage()
{
    fprintf(stderr,
            "usage: galpha_cli POINTS.txt [--fbo N] [--no-alpha] [--check] [--repeat K]\n"
            "                  [--dump FILE] [--dump-full FILE] [--pack FILE]\n");
    return 2;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) return usage();
    const char *path = argv[1], *dump = nullptr, *pack = nullptr;
    int fbo = 0, with_alpha = 1, do_check = 0, repeat = 1, full = 0;
    for (int k = 2; k < argc; k++) {
        std::string a = argv[k];
        auto val = [&]() -> const char * { return k + 1 < argc ? argv[++k] : nullptr; };
        if (a == "--fbo") { const char *v = val(); if (!v) return usage(); fbo = atoi(v); }
        else if (a == "--no-alpha") with_alpha = 0;
        else if (a == "--check") do_check = 1;
        else if (a == "--repeat") { const char *v = val(); if (!v) return usage(); repeat = std::max(1, atoi(v)); }
        else if (a == "--dump") { dump = val(); full = 0; if (!dump) return usage(); }
        else if (a == "--dump-full") { dump = val(); full = 1; if (!dump) return usage(); }
        else if (a == "--pack") { pack = val(); if (!pack) return usage(); }
        else return usage();
    }
    printf("GAlpha (MI355X) - 2D Delaunay triangulation and alpha shapes\n");
    PGPUDTPARAMS in = nullptr;
    if (galphaReadPoints(path, &in) < 0) {
        fprintf(stderr, "error: %s\n", gpudtGetLastError());
        return 1;
    }
    if (fbo > 0) in->fboSize = fbo;
    printf("Read %d points from %s; texture %d\n", in->nPoints, path, in->fboSize);
    PGPUDTOUTPUT out = nullptr;
    PGPUDTALPHA alpha = nullptr;
    int rc = 0;
    for (int r = 0; r < repeat && rc == 0; r++) {
        if (out) gpudtReleaseDTOutput(out);
        if (alpha) gpudtReleaseAlpha(alpha);
        out = nullptr;
        alpha = nullptr;
        auto t0 = std::chrono::steady_clock::now();
        out = gpudtComputeDT(in);
        if (out && with_alpha && gpudtComputeAlpha(in, out, &alpha) < 0) rc = 1;
        auto t1 = std::chrono::steady_clock::now();
        if (!out) rc = 1;
        if (rc) {
            fprintf(stderr, "error: %s\n", gpudtGetLastError());
            break;
        }
        printf("GAlpha time: %.4fs (%d triangles%s)\n", std::chrono::duration<double>(t1 - t0).count(), out->nTris,
               with_alpha ? (", spectrum " + std::to_string(alpha->specSize)).c_str() : "");
    }
    if (!rc && do_check) rc = check(in, out);
    if (!rc && dump) {
        if (!alpha) {
            fprintf(stderr, "error: --dump needs the alpha result\n");
            rc = 1;
        } else if (galphaDumpValues(dump, in, out, alpha, full) < 0) {
            fprintf(stderr, "error: %s\n", gpudtGetLastError());
            rc = 1;
        }
    }
    if (!rc && pack) {
        long long size = galphaPackResult(in->nPoints, (const int *)out->triangles, out->nTris, alpha, nullptr, 0);
        std::vector<char> buf(size > 0 ? (size_t)size : 0);
        if (size < 0 || galphaPackResult(in->nPoints, (const int *)out->triangles, out->nTris, alpha, buf.data(), size) != size ||
            !write_file(pack, buf)) {
            fprintf(stderr, "error: cannot write payload %s\n", pack);
            rc = 1;
        }
    }
    if (alpha) gpudtReleaseAlpha(alpha);
    if (out) gpudtReleaseDTOutput(out);
    galphaReleaseParams(in);
    return rc;
}
