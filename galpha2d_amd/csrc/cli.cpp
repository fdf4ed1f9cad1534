// galpha_cli: command-line driver over the C ABI, the counterpart of the reference's main.cu
// (main.cu:364-440: read a point file, gpudtComputeDT, Alpha_Shapes_2D::computeAlphaShapes,
// print the time, checkResult).  The file name is an argument instead of a prompt.
//
//   galpha_cli POINTS.txt [--fbo N] [--no-alpha] [--check] [--repeat K]
//              [--dump FILE] [--dump-full FILE] [--pack FILE]
//
// --check runs main.cu's checkResult (main.cu:163-343) on the host, walking the result with the
// mesh-navigation macros of include/gpudt.h (navcheck.h): vertex indices, every point that is not
// an exact duplicate is a vertex, sym across every edge, the vertex fans (onext / oprev), Euler's
// formula, and the orientation and empty-circle properties wherever the FP64 filter is certain
// (Shewchuk's static error bounds).  Cases the filter leaves open are counted, not decided: the
// library's own predicates are exact, so this check is a sanity check of the output, not a proof.
#include "../../include/gpudt.h"
#include "navcheck.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

int check(const PGPUDTPARAMS in, const PGPUDTOUTPUT out)
{
    NavReport r = galpha_nav_check(in->points, in->nPoints, out->triangles, out->nTris);
    printf("check: %d triangles, %d errors, %lld predicate evaluations left open by the FP64 filter\n",
           out->nTris, r.errors, r.open);
    printf("check: V = %d, E = %d (%d on the hull), F = %d + 1, Euler characteristic %d, fan corners %lld\n",
           r.vertices, r.edges, r.hullEdges, out->nTris, r.euler, r.fanCorners);
    return r.errors ? 1 : 0;
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
