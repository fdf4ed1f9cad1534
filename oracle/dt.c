/*
 * dt.c -- sequential Delaunay triangulation for the oracle (test infrastructure only).
 *
 * The reference has no CPU Delaunay path: it triangulates on the GPU (gpudt.cpp:771-924)
 * and leaves "unsure" incircle edges to recursive exact flips on the CPU
 * (gpudt.cpp:191-304, 875-914).  This file is an independent sequential algorithm:
 * Bowyer-Watson insertion with ghost triangles in Hilbert order, on the exact
 * predicates of predicates.c, with the lifted-perturbation tie-break
 * (ora_incircle_sos) so that the result is unique even on co-circular input.
 *
 * Output convention (gpudt/gpudt.h:97-102, :137-244): vtx[] counter-clockwise,
 * tri[i] = (t << 2) | j for the neighbour t across the edge opposite vtx[i], where
 * t.vtx[j] is the vertex of t opposite that edge; -1 across a hull edge.
 * Exact duplicates keep their smallest index; the others are left out
 * (the reference drops duplicates too: Cuda/cudaMissing.cu:285-289).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define GHOST (-1)

typedef struct { int v[3]; int n[3]; } mtri;

typedef struct {
    const double *xy;
    mtri *t;
    int nt, cap;
    int *freelist; int nfree, freecap;
    int *stamp;       /* per triangle: insertion id when visited */
    int *conflict;    /* per triangle: insertion id when in conflict */
    int last;         /* a live triangle to start walks from */
    unsigned rng;
} mesh_t;

static const double *P(const mesh_t *m, int i) { return m->xy + 2 * (size_t)i; }

static int new_tri(mesh_t *m)
{
    if (m->nfree > 0) return m->freelist[--m->nfree];
    if (m->nt == m->cap) {
        m->cap = m->cap ? m->cap * 2 : 1024;
        m->t = (mtri *)realloc(m->t, sizeof(mtri) * m->cap);
        m->stamp = (int *)realloc(m->stamp, sizeof(int) * m->cap);
        m->conflict = (int *)realloc(m->conflict, sizeof(int) * m->cap);
    }
    m->stamp[m->nt] = -1;
    m->conflict[m->nt] = -1;
    return m->nt++;
}
static void free_tri(mesh_t *m, int t)
{
    if (m->nfree == m->freecap) {
        m->freecap = m->freecap ? m->freecap * 2 : 1024;
        m->freelist = (int *)realloc(m->freelist, sizeof(int) * m->freecap);
    }
    m->t[t].v[0] = -2;  /* dead marker */
    m->freelist[m->nfree++] = t;
}

static int is_ghost(const mtri *t) { return t->v[0] == GHOST || t->v[1] == GHOST || t->v[2] == GHOST; }

/* conflict test of point p (index ip) with triangle t */
static int in_conflict(const mesh_t *m, const mtri *t, int ip)
{
    const double *p = P(m, ip);
    int g = -1;
    for (int i = 0; i < 3; i++) if (t->v[i] == GHOST) g = i;
    if (g < 0)
        return ora_incircle_sos(P(m, t->v[0]), t->v[0], P(m, t->v[1]), t->v[1],
                                P(m, t->v[2]), t->v[2], p, ip) > 0;
    /* ghost (x, y, G): the outside of hull edge x->y lies to its left */
    int x = t->v[(g + 1) % 3], y = t->v[(g + 2) % 3];
    const double *a = P(m, x), *b = P(m, y);
    int o = ora_orient_sign(a, b, p);
    if (o > 0) return 1;
    if (o < 0) return 0;
    /* collinear: conflict iff p strictly inside segment xy */
    if (a[0] != b[0]) return (a[0] < p[0] && p[0] < b[0]) || (b[0] < p[0] && p[0] < a[0]);
    return (a[1] < p[1] && p[1] < b[1]) || (b[1] < p[1] && p[1] < a[1]);
}

static unsigned rnd(mesh_t *m) { m->rng = m->rng * 1103515245u + 12345u; return m->rng >> 16; }

/* visibility walk to a triangle in conflict with p: a real triangle containing p
 * (closed) or a ghost whose hull edge p lies strictly beyond. */
static int locate(mesh_t *m, int ip)
{
    const double *p = P(m, ip);
    int t = m->last;
    if (m->t[t].v[0] == -2) t = 0;
    while (m->t[t].v[0] == -2) t++;
    if (is_ghost(&m->t[t])) {
        int g = 0;
        for (int i = 0; i < 3; i++) if (m->t[t].v[i] == GHOST) g = i;
        t = m->t[t].n[g] >> 2;
    }
    for (long guard = 0;; guard++) {
        mtri *T = &m->t[t];
        if (is_ghost(T)) return t;
        int r = rnd(m) % 3, moved = 0;
        for (int k = 0; k < 3; k++) {
            int i = (r + k) % 3;
            if (ora_orient_sign(P(m, T->v[(i + 1) % 3]), P(m, T->v[(i + 2) % 3]), p) < 0) {
                t = T->n[i] >> 2;
                moved = 1;
                break;
            }
        }
        if (!moved) return t;
        if (guard > 100000000L) return -1;
    }
}

typedef struct { int u, v, nb; } bedge;

static int insert_point(mesh_t *m, int ip, int iter, int *first_of, int nvert)
{
    int t0 = locate(m, ip);
    if (t0 < 0) return -1;
    {   /* duplicate of an existing vertex? (only possible if dedup missed it) */
        const double *p = P(m, ip);
        for (int i = 0; i < 3; i++) {
            int v = m->t[t0].v[i];
            if (v >= 0 && P(m, v)[0] == p[0] && P(m, v)[1] == p[1]) return 1;
        }
    }
    if (!in_conflict(m, &m->t[t0], ip)) return -2;  /* cannot happen with exact predicates */

    /* flood the conflict region */
    static __thread int *stack = NULL; static __thread int scap = 0;   /* per thread, kept across calls */
    static __thread bedge *bnd = NULL; static __thread int bcap = 0;
    static __thread int *dead = NULL; static __thread int dcap = 0;
    int sp = 0, nb = 0, nd = 0;
    if (!scap) { scap = 1024; stack = (int *)malloc(sizeof(int) * scap); }
    if (!bcap) { bcap = 1024; bnd = (bedge *)malloc(sizeof(bedge) * bcap); }
    if (!dcap) { dcap = 1024; dead = (int *)malloc(sizeof(int) * dcap); }
    stack[sp++] = t0;
    m->conflict[t0] = iter;
    m->stamp[t0] = iter;
    while (sp > 0) {
        int t = stack[--sp];
        if (nd == dcap) { dcap *= 2; dead = (int *)realloc(dead, sizeof(int) * dcap); }
        dead[nd++] = t;
        for (int i = 0; i < 3; i++) {
            int enc = m->t[t].n[i], u = enc >> 2;
            int isc;
            if (m->stamp[u] == iter) isc = (m->conflict[u] == iter);
            else {
                m->stamp[u] = iter;
                isc = in_conflict(m, &m->t[u], ip);
                if (isc) {
                    m->conflict[u] = iter;
                    if (sp == scap) { scap *= 2; stack = (int *)realloc(stack, sizeof(int) * scap); }
                    stack[sp++] = u;
                }
            }
            if (!isc) {
                if (nb == bcap) { bcap *= 2; bnd = (bedge *)realloc(bnd, sizeof(bedge) * bcap); }
                bnd[nb].u = m->t[t].v[(i + 1) % 3];
                bnd[nb].v = m->t[t].v[(i + 2) % 3];
                bnd[nb].nb = enc;
                nb++;
            }
        }
    }
    for (int k = 0; k < nd; k++) free_tri(m, dead[k]);
    /* new triangles (u, v, p): slot 2 faces the outside neighbour */
    int *nt = (int *)malloc(sizeof(int) * nb);
    for (int k = 0; k < nb; k++) {
        int t = new_tri(m);
        nt[k] = t;
        mtri *T = &m->t[t];
        T->v[0] = bnd[k].u; T->v[1] = bnd[k].v; T->v[2] = ip;
        T->n[2] = bnd[k].nb;
        int o = bnd[k].nb >> 2, oi = bnd[k].nb & 3;
        m->t[o].n[oi] = (t << 2) | 2;
        int key = bnd[k].u == GHOST ? nvert : bnd[k].u;
        first_of[key] = t;
    }
    for (int k = 0; k < nb; k++) {
        int t = nt[k];
        mtri *T = &m->t[t];
        int key = T->v[1] == GHOST ? nvert : T->v[1];
        int t2 = first_of[key];          /* (v, w, p): its edge (p, v) is opposite w (slot 1) */
        T->n[0] = (t2 << 2) | 1;
        m->t[t2].n[1] = (t << 2) | 0;
    }
    m->last = nt[nb - 1];
    free(nt);
    return 0;
}

/* Hilbert index of (x, y) in a 2^16 grid */
static uint64_t hilbert_d(uint32_t x, uint32_t y)
{
    uint64_t d = 0;
    for (uint32_t s = 1u << 15; s > 0; s >>= 1) {
        uint32_t rx = (x & s) > 0, ry = (y & s) > 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) { x = (s - 1) - x; y = (s - 1) - y; }
            uint32_t t = x; x = y; y = t;
        }
    }
    return d;
}

typedef struct { uint64_t key; int idx; } keyidx;
static int cmp_keyidx(const void *a, const void *b)
{
    const keyidx *p = (const keyidx *)a, *q = (const keyidx *)b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return p->idx - q->idx;
}
static __thread const double *g_xy;
static int cmp_xy(const void *a, const void *b)
{
    int i = *(const int *)a, j = *(const int *)b;
    const double *p = g_xy + 2 * (size_t)i, *q = g_xy + 2 * (size_t)j;
    if (p[0] != q[0]) return p[0] < q[0] ? -1 : 1;
    if (p[1] != q[1]) return p[1] < q[1] ? -1 : 1;
    return i - j;
}

int ora_delaunay(const double *xy, int n, ora_tri **out, int *dup_of)
{
    ora_exactinit();
    *out = NULL;
    int *dup = dup_of ? dup_of : (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) dup[i] = -1;
    if (n < 3) { if (!dup_of) free(dup); return 0; }
    {   /* exact duplicates: keep the smallest index */
        int *ord = (int *)malloc(sizeof(int) * n);
        for (int i = 0; i < n; i++) ord[i] = i;
        g_xy = xy;
        qsort(ord, n, sizeof(int), cmp_xy);
        for (int k = 1; k < n; k++) {
            const double *p = xy + 2 * (size_t)ord[k], *q = xy + 2 * (size_t)ord[k - 1];
            if (p[0] == q[0] && p[1] == q[1])
                dup[ord[k]] = dup[ord[k - 1]] >= 0 ? dup[ord[k - 1]] : ord[k - 1];
        }
        free(ord);
    }
    /* Hilbert order of the kept points */
    double minx = INFINITY, miny = INFINITY, maxx = -INFINITY, maxy = -INFINITY;
    for (int i = 0; i < n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        if (x < minx) minx = x; if (x > maxx) maxx = x;
        if (y < miny) miny = y; if (y > maxy) maxy = y;
    }
    double range = fmax(maxx - minx, maxy - miny);
    if (!(range > 0)) range = 1;
    keyidx *ki = (keyidx *)malloc(sizeof(keyidx) * n);
    int nk = 0;
    for (int i = 0; i < n; i++) {
        if (dup[i] >= 0) continue;
        double fx = (xy[2 * i] - minx) / range * 65535.0, fy = (xy[2 * i + 1] - miny) / range * 65535.0;
        uint32_t qx = (uint32_t)fmin(fmax(fx, 0), 65535), qy = (uint32_t)fmin(fmax(fy, 0), 65535);
        ki[nk].key = hilbert_d(qx, qy);
        ki[nk].idx = i;
        nk++;
    }
    qsort(ki, nk, sizeof(keyidx), cmp_keyidx);
    if (nk < 3) { free(ki); if (!dup_of) free(dup); return 0; }
    /* first triangle: two distinct points and the first non-collinear third */
    int a = ki[0].idx, b = ki[1].idx, c = -1, ck = -1;
    for (int k = 2; k < nk; k++) {
        if (ora_orient_sign(xy + 2 * (size_t)a, xy + 2 * (size_t)b, xy + 2 * (size_t)ki[k].idx) != 0) {
            c = ki[k].idx; ck = k; break;
        }
    }
    if (c < 0) { free(ki); if (!dup_of) free(dup); return 0; }  /* all collinear: no triangles */
    if (ora_orient_sign(xy + 2 * (size_t)a, xy + 2 * (size_t)b, xy + 2 * (size_t)c) < 0) { int t = b; b = c; c = t; }
    mesh_t m;
    memset(&m, 0, sizeof(m));
    m.xy = xy;
    m.rng = 12345u;
    int T0 = new_tri(&m), G0 = new_tri(&m), G1 = new_tri(&m), G2 = new_tri(&m);
    /* real (a,b,c); ghosts (b,a,G) (c,b,G) (a,c,G) */
    int rv[3] = {a, b, c};
    m.t[T0].v[0] = a; m.t[T0].v[1] = b; m.t[T0].v[2] = c;
    int G[3] = {G0, G1, G2};
    /* T0 edge opposite c (slot 2) is a->b: ghost (b,a,G) = G0 ; opposite a (slot 0) is b->c:
       ghost (c,b,G) = G1 ; opposite b (slot 1) is c->a: ghost (a,c,G) = G2 */
    m.t[G0].v[0] = b; m.t[G0].v[1] = a; m.t[G0].v[2] = GHOST;
    m.t[G1].v[0] = c; m.t[G1].v[1] = b; m.t[G1].v[2] = GHOST;
    m.t[G2].v[0] = a; m.t[G2].v[1] = c; m.t[G2].v[2] = GHOST;
    m.t[T0].n[2] = (G0 << 2) | 2; m.t[G0].n[2] = (T0 << 2) | 2;
    m.t[T0].n[0] = (G1 << 2) | 2; m.t[G1].n[2] = (T0 << 2) | 0;
    m.t[T0].n[1] = (G2 << 2) | 2; m.t[G2].n[2] = (T0 << 2) | 1;
    /* ghost-ghost links: ghost (x,y,G) slot 0 is edge (y,G), slot 1 is edge (G,x).
       Edge (y,G) of ghost (x,y,G) is shared with the ghost whose x' == y, as its (G,x') (slot 1). */
    for (int i = 0; i < 3; i++) {
        int gi = G[i];
        int y = m.t[gi].v[1];
        for (int j = 0; j < 3; j++)
            if (m.t[G[j]].v[0] == y) { m.t[gi].n[0] = (G[j] << 2) | 1; m.t[G[j]].n[1] = (gi << 2) | 0; }
    }
    (void)rv;
    m.last = T0;
    int *first_of = (int *)malloc(sizeof(int) * (n + 1));
    int iter = 0, bad = 0;
    for (int k = 2; k < nk; k++) {
        if (k == ck) continue;
        int r = insert_point(&m, ki[k].idx, ++iter, first_of, n);
        if (r < 0) { bad = 1; break; }
    }
    free(first_of);
    free(ki);
    if (bad) { free(m.t); free(m.stamp); free(m.conflict); free(m.freelist); if (!dup_of) free(dup); return -1; }
    /* collect real triangles */
    int *newid = (int *)malloc(sizeof(int) * m.nt);
    int nr = 0;
    for (int t = 0; t < m.nt; t++) {
        newid[t] = -1;
        if (m.t[t].v[0] == -2 || is_ghost(&m.t[t])) continue;
        newid[t] = nr++;
    }
    ora_tri *o = (ora_tri *)malloc(sizeof(ora_tri) * (nr > 0 ? nr : 1));
    for (int t = 0; t < m.nt; t++) {
        if (newid[t] < 0) continue;
        ora_tri *T = &o[newid[t]];
        for (int i = 0; i < 3; i++) {
            T->vtx[i] = m.t[t].v[i];
            int u = m.t[t].n[i] >> 2, ui = m.t[t].n[i] & 3;
            T->tri[i] = newid[u] >= 0 ? ((newid[u] << 2) | ui) : -1;
            T->tmp[i] = 0;
        }
    }
    free(newid);
    free(m.t); free(m.stamp); free(m.conflict); free(m.freelist);
    if (!dup_of) free(dup);
    ora_canonicalize(o, nr);
    *out = o;
    return nr;
}

/* seeded uniform generator: splitmix64 -> (u >> 11) * 2^-53 */
static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
void ora_gen_uniform(uint64_t seed, int n, double *xy)
{
    uint64_t s = seed;
    for (int i = 0; i < 2 * n; i++) xy[i] = (double)(splitmix64(&s) >> 11) * 0x1.0p-53;
}
