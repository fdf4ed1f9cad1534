/*
 * canon.c -- canonical triangle order and the mesh validator (test infrastructure only).
 *
 * Canonical form: each triangle rotated so that vtx[0] is its smallest vertex index
 * (rotation keeps it counter-clockwise), triangles sorted by (vtx0, vtx1, vtx2), and
 * every neighbour link re-encoded as (t << 2) | j in the new numbering
 * (encoding of gpudt/gpudt.h:137-244).
 *
 * The validator restates checkResult (main.cu:163-343): indices in range (:175-193),
 * every vertex used (:195-201), counter-clockwise by exact orientation (:205-216),
 * every interior edge passes the incircle test (:288-326, here with the tie-break of
 * ora_incircle_sos), Euler's formula (:328-337).  It adds what the reference lacks:
 * reciprocal adjacency, a single convex hull cycle, and a winding number of one around
 * every interior vertex, which together make the mesh a valid triangulation.
 */
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef struct { int v[3]; int old; } sortrec;

static int cmp_rec(const void *a, const void *b)
{
    const sortrec *p = (const sortrec *)a, *q = (const sortrec *)b;
    for (int i = 0; i < 3; i++)
        if (p->v[i] != q->v[i]) return p->v[i] < q->v[i] ? -1 : 1;
    return 0;
}

void ora_canonicalize(ora_tri *t, int nt)
{
    if (nt <= 0) return;
    int *rot = (int *)malloc(sizeof(int) * nt);
    sortrec *r = (sortrec *)malloc(sizeof(sortrec) * nt);
    for (int k = 0; k < nt; k++) {
        int m = 0;
        if (t[k].vtx[1] < t[k].vtx[m]) m = 1;
        if (t[k].vtx[2] < t[k].vtx[m]) m = 2;
        rot[k] = m;
        for (int i = 0; i < 3; i++) r[k].v[i] = t[k].vtx[(i + m) % 3];
        r[k].old = k;
    }
    qsort(r, nt, sizeof(sortrec), cmp_rec);
    int *newid = (int *)malloc(sizeof(int) * nt);
    for (int k = 0; k < nt; k++) newid[r[k].old] = k;
    ora_tri *o = (ora_tri *)malloc(sizeof(ora_tri) * nt);
    for (int k = 0; k < nt; k++) {
        int old = r[k].old, m = rot[old];
        for (int i = 0; i < 3; i++) {
            int s = (i + m) % 3;
            o[k].vtx[i] = t[old].vtx[s];
            int enc = t[old].tri[s];
            if (enc < 0) o[k].tri[i] = -1;
            else {
                int u = enc >> 2, j = enc & 3;
                o[k].tri[i] = (newid[u] << 2) | ((j - rot[u] + 3) % 3);
            }
            o[k].tmp[i] = 0;
        }
    }
    memcpy(t, o, sizeof(ora_tri) * nt);
    free(o); free(newid); free(r); free(rot);
}

#define FAIL(code, ...) do { if (msg) snprintf(msg, msglen, __VA_ARGS__); rc = code; goto done; } while (0)

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static uint64_t edge_key(int a, int b)
{
    int lo = a < b ? a : b, hi = a < b ? b : a;
    return (uint64_t)(uint32_t)lo << 32 | (uint32_t)hi;
}

/* ckeys: sorted constraint pairs (NULL: none); constraint edges are exempt from the incircle
 * test and must all be present */
static int validate_impl(const double *xy, int n, const ora_tri *t, int nt, const int *dup_of,
                         const uint64_t *ckeys, int nck, char *msg, int msglen);

int ora_validate(const double *xy, int n, const ora_tri *t, int nt,
                 const int *dup_of, char *msg, int msglen)
{
    return validate_impl(xy, n, t, nt, dup_of, NULL, 0, msg, msglen);
}

int ora_validate_constrained(const double *xy, int n, const ora_tri *t, int nt, const int *dup_of,
                             const int *cons, int ncons, char *msg, int msglen)
{
    uint64_t *k = (uint64_t *)malloc(sizeof(uint64_t) * (ncons > 0 ? ncons : 1));
    for (int c = 0; c < ncons; c++) k[c] = edge_key(cons[2 * c], cons[2 * c + 1]);
    qsort(k, ncons, sizeof(uint64_t), cmp_u64);
    int rc = validate_impl(xy, n, t, nt, dup_of, k, ncons, msg, msglen);
    if (rc == 0) {
        /* every constraint is an edge of the mesh (main.cu:218-286) */
        int ne = 3 * nt, m = 0;
        uint64_t *e = (uint64_t *)malloc(sizeof(uint64_t) * (ne > 0 ? ne : 1));
        for (int q = 0; q < nt; q++)
            for (int i = 0; i < 3; i++) e[m++] = edge_key(t[q].vtx[(i + 1) % 3], t[q].vtx[(i + 2) % 3]);
        qsort(e, m, sizeof(uint64_t), cmp_u64);
        for (int c = 0; c < ncons && rc == 0; c++)
            if (!bsearch(&k[c], e, m, sizeof(uint64_t), cmp_u64)) {
                if (msg) snprintf(msg, msglen, "constraint %d (%d, %d) is not an edge", c, cons[2 * c], cons[2 * c + 1]);
                rc = 9;
            }
        free(e);
    }
    free(k);
    return rc;
}

static int validate_impl(const double *xy, int n, const ora_tri *t, int nt, const int *dup_of,
                         const uint64_t *ckeys, int nck, char *msg, int msglen)
{
    int rc = 0;
    ora_exactinit();
    int *used = (int *)calloc(n > 0 ? n : 1, sizeof(int));
    int *hull_next = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int *wind = (int *)calloc(n > 0 ? n : 1, sizeof(int));
    for (int i = 0; i < n; i++) hull_next[i] = -1;
    if (msg && msglen) msg[0] = 0;
    int kept = 0;
    for (int i = 0; i < n; i++) if (!dup_of || dup_of[i] < 0) kept++;
    long hull_edges = 0;
    for (int k = 0; k < nt; k++) {
        for (int i = 0; i < 3; i++) {
            int v = t[k].vtx[i];
            if (v < 0 || v >= n) FAIL(1, "triangle %d: vertex %d out of range", k, v);
            if (dup_of && dup_of[v] >= 0) FAIL(1, "triangle %d uses dropped duplicate %d", k, v);
            used[v] = 1;
        }
        const double *a = xy + 2 * (size_t)t[k].vtx[0], *b = xy + 2 * (size_t)t[k].vtx[1],
                     *c = xy + 2 * (size_t)t[k].vtx[2];
        if (ora_orient_sign(a, b, c) <= 0) FAIL(2, "triangle %d (%d %d %d) not counter-clockwise", k,
                                                t[k].vtx[0], t[k].vtx[1], t[k].vtx[2]);
        for (int i = 0; i < 3; i++) {
            int e0 = t[k].vtx[(i + 1) % 3], e1 = t[k].vtx[(i + 2) % 3];
            int enc = t[k].tri[i];
            if (enc < 0) {
                hull_edges++;
                if (hull_next[e0] >= 0) FAIL(6, "vertex %d starts two hull edges", e0);
                hull_next[e0] = e1;
                continue;
            }
            int u = enc >> 2, j = enc & 3;
            if (u < 0 || u >= nt || j > 2) FAIL(3, "triangle %d: bad link %d", k, enc);
            if (t[u].tri[j] != ((k << 2) | i)) FAIL(3, "triangle %d slot %d: link not reciprocal", k, i);
            if (t[u].vtx[(j + 1) % 3] != e1 || t[u].vtx[(j + 2) % 3] != e0)
                FAIL(3, "triangle %d slot %d: shared edge mismatch", k, i);
            uint64_t ek = edge_key(e0, e1);
            if (u > k && !(ckeys && bsearch(&ek, ckeys, nck, sizeof(uint64_t), cmp_u64))) {
                const double *d = xy + 2 * (size_t)t[u].vtx[j];
                if (ora_incircle_sos(a, t[k].vtx[0], b, t[k].vtx[1], c, t[k].vtx[2], d, t[u].vtx[j]) > 0)
                    FAIL(7, "edge (%d,%d) between triangles %d and %d is not locally Delaunay", e0, e1, k, u);
            }
        }
        /* winding: ray +x from each vertex p through the sector (p, a, b) */
        for (int i = 0; i < 3; i++) {
            int p = t[k].vtx[i], va = t[k].vtx[(i + 1) % 3], vb = t[k].vtx[(i + 2) % 3];
            double py = xy[2 * (size_t)p + 1];
            if (xy[2 * (size_t)va + 1] <= py && xy[2 * (size_t)vb + 1] > py) wind[p]++;
        }
    }
    for (int i = 0; i < n; i++)
        if ((!dup_of || dup_of[i] < 0) && !used[i] && nt > 0) FAIL(4, "point %d not in the triangulation", i);
    if (nt > 0) {
        long V = kept, E = (3L * nt + hull_edges) / 2;
        if ((3L * nt + hull_edges) % 2 != 0 || V - E + nt != 1)
            FAIL(5, "Euler check failed: V=%ld E=%ld F=%d", V, E, nt);
        /* the hull is one cycle of left (or straight) turns, winding once */
        int start = -1;
        for (int i = 0; i < n; i++) if (hull_next[i] >= 0) { start = i; break; }
        if (start < 0) FAIL(6, "no hull");
        long len = 0;
        double turn = 0;
        int x = start;
        do {
            int y = hull_next[x];
            if (y < 0) FAIL(6, "hull chain broken at %d", x);
            int z = hull_next[y];
            if (z < 0) FAIL(6, "hull chain broken at %d", y);
            const double *X = xy + 2 * (size_t)x, *Y = xy + 2 * (size_t)y, *Z = xy + 2 * (size_t)z;
            if (ora_orient_sign(X, Y, Z) < 0) FAIL(6, "hull not convex at %d", y);
            double a1 = atan2(Y[1] - X[1], Y[0] - X[0]), a2 = atan2(Z[1] - Y[1], Z[0] - Y[0]);
            double d = a2 - a1;
            while (d <= -M_PI) d += 2 * M_PI;
            while (d > M_PI) d -= 2 * M_PI;
            turn += d;
            x = y;
            if (++len > hull_edges) FAIL(6, "hull is not a single cycle");
        } while (x != start);
        if (len != hull_edges) FAIL(6, "hull has %ld edges in the cycle of %ld", len, hull_edges);
        if (fabs(turn - 2 * M_PI) > 1e-6) FAIL(6, "hull winds %.3f turns", turn / (2 * M_PI));
        for (int i = 0; i < n; i++) {
            if (dup_of && dup_of[i] >= 0) continue;
            if (wind[i] > 1) FAIL(8, "star of vertex %d winds %d times", i, wind[i]);
            if (hull_next[i] < 0 && wind[i] != 1) FAIL(8, "interior vertex %d has winding %d", i, wind[i]);
        }
    }
done:
    free(used); free(hull_next); free(wind);
    return rc;
}
