/*
 * cdt.c -- constrained Delaunay triangulation (test infrastructure only).
 *
 * The reference inserts constraint edges on the GPU by flipping the triangles a constraint
 * crosses (Cuda/cudaConstraint.cu:828-939), marks them (Cuda/cudaFlipping.cu:111-220) and flips
 * every unmarked edge to local Delaunay, its CPU pass keeping marked edges fixed
 * (gpudt.cpp:88-155).  This file restates the result, not the parallel schedule:
 *   1. the DT of the points (dt.c);
 *   2. each constraint, in input order, inserted by Sloan's edge-flip method (S. W. Sloan, "A
 *      fast algorithm for generating constrained Delaunay triangulations", Computers &
 *      Structures 47(3), 1993): the edges that cross it go to a queue; an edge whose
 *      quadrilateral is strictly convex is flipped and its new diagonal re-queued while it
 *      still crosses the constraint; any other edge goes to the back of the queue;
 *   3. Lawson flips of every edge that is not a constraint, with the tie-broken incircle test,
 *      until every such edge is locally Delaunay.
 * A triangulation that holds every constraint and whose other edges are all locally Delaunay
 * is the constrained DT, which is unique under the tie-break; ora_validate_constrained checks
 * exactly that, independently of how the mesh was built.
 * Inputs the reference leaves undefined are errors here: an endpoint out of range, equal to the
 * other or a dropped duplicate, a point inside the open segment, and constraints that cross.
 */
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    ora_tri *t;
    int nt;
    const double *xy;
    int *vt;              /* vertex -> a triangle containing it */
    uint64_t *hkeys;      /* constraint pairs (min << 32 | max), open addressing, 0 = empty */
    int hmask;
} cmesh;

static const double *PT(const cmesh *m, int v) { return m->xy + 2 * (size_t)v; }
static int orient(const cmesh *m, int a, int b, int c) { return ora_orient_sign(PT(m, a), PT(m, b), PT(m, c)); }

static uint64_t pair_key(int a, int b)
{
    int lo = a < b ? a : b, hi = a < b ? b : a;
    return ((uint64_t)(uint32_t)lo << 32 | (uint32_t)hi) + 1;   /* never 0 */
}

static uint64_t mix(uint64_t k)
{
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33;
    return k;
}

static void hset_add(cmesh *m, int a, int b)
{
    uint64_t k = pair_key(a, b);
    for (uint64_t h = mix(k);; h++) {
        uint64_t *s = &m->hkeys[h & (uint64_t)m->hmask];
        if (*s == k) return;
        if (*s == 0) { *s = k; return; }
    }
}

static int is_constraint(const cmesh *m, int a, int b)
{
    if (!m->hkeys) return 0;
    uint64_t k = pair_key(a, b);
    for (uint64_t h = mix(k);; h++) {
        uint64_t s = m->hkeys[h & (uint64_t)m->hmask];
        if (s == k) return 1;
        if (s == 0) return 0;
    }
}

/* triangle and side of the edge {x, y}: T[t].vtx[(s+1)%3], T[t].vtx[(s+2)%3] are x, y in
 * either order; found by rotating around x from vt[x].  Returns t or -1. */
static int find_edge(const cmesh *m, int x, int y, int *side)
{
    for (int dir = 0; dir < 2; dir++) {
        int t = m->vt[x];
        for (int guard = 0; guard < m->nt + 1 && t >= 0; guard++) {
            const ora_tri *T = &m->t[t];
            int k = T->vtx[0] == x ? 0 : (T->vtx[1] == x ? 1 : 2);
            int p = T->vtx[(k + 1) % 3], q = T->vtx[(k + 2) % 3];
            if (p == y) { *side = (k + 2) % 3; return t; }
            if (q == y) { *side = (k + 1) % 3; return t; }
            /* counter-clockwise: across (x, q), the side opposite p; clockwise: across (x, p) */
            int e = dir == 0 ? T->tri[(k + 1) % 3] : T->tri[(k + 2) % 3];
            t = e < 0 ? -1 : e >> 2;
            if (t == m->vt[x]) break;
        }
    }
    return -1;
}

static void set_back(cmesh *m, int e, int t, int s)
{
    if (e >= 0) m->t[e >> 2].tri[e & 3] = (t << 2) | s;
}

/* flip the edge of triangle t opposite vtx[i]; returns the new diagonal's endpoints */
static void flip(cmesh *m, int t, int i, int *np, int *nq)
{
    ora_tri *A = &m->t[t];
    int e = A->tri[i], u = e >> 2, j = e & 3;
    ora_tri *B = &m->t[u];
    int c = A->vtx[i], x = A->vtx[(i + 1) % 3], y = A->vtx[(i + 2) % 3], d = B->vtx[j];
    int a_yc = A->tri[(i + 1) % 3], a_cx = A->tri[(i + 2) % 3];
    int b_xd = B->tri[(j + 1) % 3], b_dy = B->tri[(j + 2) % 3];
    /* t = (c, x, d), u = (d, y, c); the diagonal (d, c) is side 1 of both */
    A->vtx[0] = c; A->vtx[1] = x; A->vtx[2] = d;
    B->vtx[0] = d; B->vtx[1] = y; B->vtx[2] = c;
    A->tri[0] = b_xd; A->tri[1] = (u << 2) | 1; A->tri[2] = a_cx;
    B->tri[0] = a_yc; B->tri[1] = (t << 2) | 1; B->tri[2] = b_dy;
    set_back(m, b_xd, t, 0); set_back(m, a_cx, t, 2);
    set_back(m, a_yc, u, 0); set_back(m, b_dy, u, 2);
    m->vt[c] = t; m->vt[x] = t; m->vt[d] = u; m->vt[y] = u;
    *np = c; *nq = d;
}

static int crosses(const cmesh *m, int a, int b, int x, int y)
{
    if (x == a || x == b || y == a || y == b) return 0;
    return orient(m, a, b, x) * orient(m, a, b, y) < 0 && orient(m, x, y, a) * orient(m, x, y, b) < 0;
}

typedef struct { int *v; long head, tail, cap; } queue;

static int q_push(queue *q, int x, int y)
{
    if (q->tail - q->head + 1 > q->cap) {
        long n = q->tail - q->head, nc = q->cap * 2 + 16;
        int *nv = (int *)malloc(sizeof(int) * 2 * (size_t)nc);
        if (!nv) return -1;
        for (long k = 0; k < n; k++) {
            long s = (q->head + k) % (q->cap ? q->cap : 1);
            nv[2 * k] = q->v[2 * s]; nv[2 * k + 1] = q->v[2 * s + 1];
        }
        free(q->v);
        q->v = nv; q->cap = nc; q->head = 0; q->tail = n;
    }
    long s = q->tail % q->cap;
    q->v[2 * s] = x; q->v[2 * s + 1] = y;
    q->tail++;
    return 0;
}

static void q_pop(queue *q, int *x, int *y)
{
    long s = q->head % q->cap;
    *x = q->v[2 * s]; *y = q->v[2 * s + 1];
    q->head++;
}

#define CFAIL(...) do { if (msg) snprintf(msg, msglen, __VA_ARGS__); rc = -1; goto done; } while (0)

int ora_cdt(const double *xy, int n, const int *cons, int ncons, ora_tri **out, int *dup_of, char *msg, int msglen)
{
    int rc = 0;
    cmesh m;
    memset(&m, 0, sizeof(m));
    queue q;
    memset(&q, 0, sizeof(q));
    if (msg && msglen) msg[0] = 0;
    int *dup = dup_of ? dup_of : (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    *out = NULL;
    ora_exactinit();
    int nt = ora_delaunay(xy, n, &m.t, dup);
    if (nt < 0) CFAIL("delaunay failed");
    m.nt = nt;
    m.xy = xy;
    m.vt = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) m.vt[i] = -1;
    for (int t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) m.vt[m.t[t].vtx[k]] = t;
    int hs = 16;
    while (hs < 2 * ncons + 16) hs *= 2;
    m.hkeys = (uint64_t *)calloc((size_t)hs, sizeof(uint64_t));
    m.hmask = hs - 1;
    for (int c = 0; c < ncons; c++) {
        int a = cons[2 * c], b = cons[2 * c + 1];
        if (a < 0 || a >= n || b < 0 || b >= n) CFAIL("constraint %d: endpoint out of range", c);
        if (a == b) CFAIL("constraint %d: both endpoints are point %d", c, a);
        if (dup[a] >= 0 || dup[b] >= 0) CFAIL("constraint %d: endpoint is a duplicate point", c);
        if (nt == 0) CFAIL("constraint %d: the points have no triangulation", c);
        hset_add(&m, a, b);
    }
    for (int c = 0; c < ncons; c++) {
        int a = cons[2 * c], b = cons[2 * c + 1];
        /* a point inside the open segment */
        const double *A = PT(&m, a), *B = PT(&m, b);
        for (int w = 0; w < n; w++) {
            if (w == a || w == b || dup[w] >= 0) continue;
            if (orient(&m, a, b, w) != 0) continue;
            const double *W = PT(&m, w);
            int ax = A[0] != B[0] ? 0 : 1;
            double lo = A[ax] < B[ax] ? A[ax] : B[ax], hi = A[ax] < B[ax] ? B[ax] : A[ax];
            if (W[ax] > lo && W[ax] < hi) CFAIL("constraint %d passes through point %d", c, w);
        }
        /* the edges that cross it */
        q.head = q.tail = 0;
        for (int t = 0; t < m.nt; t++)
            for (int s = 0; s < 3; s++) {
                int e = m.t[t].tri[s];
                if (e < 0 || (e >> 2) < t) continue;
                int x = m.t[t].vtx[(s + 1) % 3], y = m.t[t].vtx[(s + 2) % 3];
                if (!crosses(&m, a, b, x, y)) continue;
                if (is_constraint(&m, x, y)) CFAIL("constraint %d crosses constraint (%d, %d)", c, x, y);
                if (q_push(&q, x, y)) CFAIL("allocation failed");
            }
        long budget = 1000000 + 64 * (q.tail - q.head) * (q.tail - q.head);
        while (q.tail > q.head) {
            if (--budget < 0) CFAIL("constraint %d: edge-flip insertion did not finish", c);
            int x, y, s;
            q_pop(&q, &x, &y);
            int t = find_edge(&m, x, y, &s);
            if (t < 0) CFAIL("constraint %d: queued edge (%d, %d) not found", c, x, y);
            int e = m.t[t].tri[s];
            int p = m.t[t].vtx[s], far = m.t[e >> 2].vtx[e & 3];
            if (orient(&m, p, far, x) * orient(&m, p, far, y) < 0) {
                int np, nq;
                flip(&m, t, s, &np, &nq);
                if (crosses(&m, a, b, np, nq) && q_push(&q, np, nq)) CFAIL("allocation failed");
            } else if (q_push(&q, x, y)) {
                CFAIL("allocation failed");
            }
        }
        int s;
        if (find_edge(&m, a, b, &s) < 0) CFAIL("constraint %d: not present after insertion", c);
    }
    /* Lawson flips of the non-constraint edges */
    q.head = q.tail = 0;
    for (int t = 0; t < m.nt; t++)
        for (int s = 0; s < 3; s++) {
            int e = m.t[t].tri[s];
            if (e >= 0 && (e >> 2) > t && q_push(&q, m.t[t].vtx[(s + 1) % 3], m.t[t].vtx[(s + 2) % 3]))
                CFAIL("allocation failed");
        }
    long budget = 100000000;
    while (q.tail > q.head) {
        if (--budget < 0) CFAIL("Lawson flips did not finish");
        int x, y, s;
        q_pop(&q, &x, &y);
        if (is_constraint(&m, x, y)) continue;
        int t = find_edge(&m, x, y, &s);
        if (t < 0) continue;   /* flipped away since it was queued */
        int e = m.t[t].tri[s];
        if (e < 0) continue;
        const ora_tri *T = &m.t[t];
        int d = m.t[e >> 2].vtx[e & 3];
        if (ora_incircle_sos(PT(&m, T->vtx[0]), T->vtx[0], PT(&m, T->vtx[1]), T->vtx[1], PT(&m, T->vtx[2]),
                             T->vtx[2], PT(&m, d), d) <= 0)
            continue;
        int np, nq;
        flip(&m, t, s, &np, &nq);
        /* the four outer edges of the quadrilateral: x-np, np-y, y-nq, nq-x */
        if (q_push(&q, x, np) || q_push(&q, np, y) || q_push(&q, y, nq) || q_push(&q, nq, x))
            CFAIL("allocation failed");
    }
    ora_canonicalize(m.t, m.nt);
    *out = m.t;
    m.t = NULL;
    rc = nt;
done:
    free(m.t);
    free(m.vt);
    free(m.hkeys);
    free(q.v);
    if (!dup_of) free(dup);
    return rc;
}
