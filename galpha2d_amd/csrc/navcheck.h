// navcheck.h -- structural and geometric check of a triangulation in the gpudtTriangle layout,
// written against the mesh-navigation macros of include/gpudt.h the way a caller of the
// reference walks its output (the reference's validator checkResult, main.cu:163-343: a corner
// per vertex, vertex fans, sym across every edge, orientation, empty circles, Euler's formula).
// Header-only host code: the command-line driver's --check uses it, and the tests compile it
// against the oracle's meshes.
//
// The geometric tests use Shewchuk's FP64 static filters: a sign the filter certifies is
// checked, and a case it leaves open is counted, not decided (the library's own predicates are
// exact, so this is a sanity check of the output, not a proof).
#pragma once
#include "../../include/gpudt.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

struct NavReport {
    int errors = 0;          // every failed check (the first ten are printed)
    long long open = 0;      // predicate evaluations the FP64 filter left open
    int vertices = 0;        // points that are vertices of the mesh
    int edges = 0;           // edges (hull edges once)
    int hullEdges = 0;
    int euler = 0;           // vertices - edges + triangles + 1 (the outer face): 2 when sound
    long long fanCorners = 0;   // corners met by the vertex fan walks: 3 * nTris when sound
};

namespace galpha_nav {

const double EPS = 1.1102230246251565e-16;   // 2^-53
const double CCW_BOUND = (3.0 + 16.0 * EPS) * EPS;
const double ICC_BOUND = (10.0 + 96.0 * EPS) * EPS;

// +1 / -1 when the FP64 sign is certain, 0 when the filter cannot decide
inline int orient_filtered(const gpudtVertex &a, const gpudtVertex &b, const gpudtVertex &c)
{
    double l = (a.x - c.x) * (b.y - c.y), r = (a.y - c.y) * (b.x - c.x), det = l - r;
    double bound = CCW_BOUND * (std::fabs(l) + std::fabs(r));
    return det > bound ? 1 : (det < -bound ? -1 : 0);
}

inline int incircle_filtered(const gpudtVertex &a, const gpudtVertex &b, const gpudtVertex &c, const gpudtVertex &d)
{
    double adx = a.x - d.x, ady = a.y - d.y, bdx = b.x - d.x, bdy = b.y - d.y, cdx = c.x - d.x, cdy = c.y - d.y;
    double bc = bdx * cdy - cdx * bdy, ca = cdx * ady - adx * cdy, ab = adx * bdy - bdx * ady;
    double al = adx * adx + ady * ady, bl = bdx * bdx + bdy * bdy, cl = cdx * cdx + cdy * cdy;
    double det = al * bc + bl * ca + cl * ab;
    double perm = (std::fabs(bdx * cdy) + std::fabs(cdx * bdy)) * al +
                  (std::fabs(cdx * ady) + std::fabs(adx * cdy)) * bl +
                  (std::fabs(adx * bdy) + std::fabs(bdx * ady)) * cl;
    double bound = ICC_BOUND * perm;
    return det > bound ? 1 : (det < -bound ? -1 : 0);
}

}  // namespace galpha_nav

// Checks the nTris triangles T of the n points P.  Prints the first ten errors to stdout.
inline NavReport galpha_nav_check(const gpudtVertex *P, int n, const gpudtTriangle *T, int nTris)
{
    using namespace galpha_nav;
    NavReport r;
    auto err = [&](const char *what, int a, int b) {
        if (r.errors++ < 10) printf("*** ERROR *** %s (%d, %d)\n", what, a, b);
    };
    gpudtOrientedTriangle o, s, back;
    int ptr;   // the packed neighbour that sym / symself / onext / oprev assign

    // 1. vertex ids, and one corner per vertex: an oriented triangle whose org is the vertex
    std::vector<int> corner(n > 0 ? n : 1, -1);
    for (o.tri = 0; o.tri < nTris; o.tri++)
        for (o.orient = 0; o.orient < 3; o.orient++) {
            int v;
            org(T, o, v);
            if (v < 0 || v >= n) {
                err("vertex index outside the point range", o.tri, v);
                return r;
            }
            corner[v] = Encode(o);
        }

    // 2. every point is a vertex unless an earlier point has the same coordinates
    {
        std::vector<int> idx(n);
        for (int i = 0; i < n; i++) idx[i] = i;
        std::sort(idx.begin(), idx.end(), [&](int a, int b) {
            return P[a].x != P[b].x ? P[a].x < P[b].x : (P[a].y != P[b].y ? P[a].y < P[b].y : a < b);
        });
        for (int k = 0; k < n; k++) {
            int v = idx[k];
            bool dup = k > 0 && P[idx[k - 1]].x == P[v].x && P[idx[k - 1]].y == P[v].y;
            if (corner[v] >= 0) r.vertices++;
            else if (!dup && n >= 3) err("point is not in the triangulation", v, -1);
        }
    }

    // 3. sym is an involution that reverses the edge
    int twice = 0;   // interior edges count twice, hull edges once here
    for (o.tri = 0; o.tri < nTris; o.tri++)
        for (o.orient = 0; o.orient < 3; o.orient++) {
            sym(T, o, s);
            if (s.tri < 0) {
                r.hullEdges++;
                continue;
            }
            twice++;
            if (s.tri >= nTris || s.orient > 2) {
                err("neighbour link outside the triangle list", o.tri, o.orient);
                continue;
            }
            sym(T, s, back);
            int a, b, c, d;
            org(T, o, a); dest(T, o, b); org(T, s, c); dest(T, s, d);
            if (!otriequal(back, o) || a != d || b != c) err("neighbour link is not symmetric", o.tri, o.orient);
        }
    if (r.errors) return r;   // the walks below follow the links: only on a sound structure
    r.edges = twice / 2 + r.hullEdges;
    r.euler = r.vertices - r.edges + nTris + 1;
    if (nTris > 0 && r.euler != 2) err("Euler characteristic is not 2", r.vertices, r.edges);

    // 4. the fan of every vertex: onext from its corner until it closes or meets the hull, then
    //    oprev from the corner the other way; every triangle met has the vertex as org
    for (int v = 0; v < n; v++) {
        if (corner[v] < 0) continue;
        gpudtOrientedTriangle start;
        decode(corner[v], start);
        otricopy(start, o);
        bool closed = false;
        for (int step = 0; step <= nTris; step++) {
            int w;
            org(T, o, w);
            if (w != v) { err("onext left the vertex fan", v, o.tri); break; }
            r.fanCorners++;
            onextself(T, o);
            if (o.tri < 0) break;
            if (otriequal(o, start)) { closed = true; break; }
        }
        if (closed) continue;
        otricopy(start, o);
        for (int step = 0; step <= nTris; step++) {
            oprevself(T, o);
            if (o.tri < 0) break;
            int w;
            org(T, o, w);
            if (w != v) { err("oprev left the vertex fan", v, o.tri); break; }
            r.fanCorners++;
        }
    }
    if (r.fanCorners != 3LL * nTris) err("vertex fans do not cover every corner once", (int)r.fanCorners, nTris);

    // 5. orientation of every triangle, empty circle across every interior edge
    for (o.tri = 0; o.tri < nTris; o.tri++) {
        int a, b, c;
        o.orient = 0;
        org(T, o, a); dest(T, o, b); apex(T, o, c);
        int so = orient_filtered(P[a], P[b], P[c]);
        if (so < 0) err("triangle is not counter-clockwise", o.tri, -1);
        else if (so == 0) r.open++;
        for (o.orient = 0; o.orient < 3; o.orient++) {
            sym(T, o, s);
            if (s.tri < 0) continue;
            int far;
            org(T, o, a); dest(T, o, b); apex(T, o, c); apex(T, s, far);
            int si = incircle_filtered(P[a], P[b], P[c], P[far]);
            if (si > 0) err("edge is not locally Delaunay", o.tri, o.orient);
            else if (si == 0) r.open++;
        }
    }
    return r;
}
