/*
 * alpha.c -- alpha-shape spectrum and intervals for the oracle (test infrastructure only).
 *
 * Restates GAlpha's semantics on a canonical Delaunay triangulation:
 *   edge ownership and edge numbering   AlphaShape/kernel.cu:160-217 (computeTriEdgeCount)
 *                                       and :252-308 (computeEdgeRo): triangle t owns the edge
 *                                       opposite vtx[i] iff vtx[i+1] < vtx[i+2] or tri[i] == -1;
 *                                       edges are numbered by (owning triangle, i).
 *   rho of a triangle / an edge         AlphaShape/predicates.h:1297-1357 (birth_time), same
 *                                       FP64 expression order, no fused multiply-add.
 *   attachment                          AlphaShape/predicates.h:1359-1444 and kernel.cu:310-358:
 *                                       edge (i,j) is attached by apex k iff the exact value of
 *                                       |i-j|^2 - |i+j-2k|^2 is >= 0; the owning triangle's apex is
 *                                       tested first, then the opposite triangle's.
 *   exact spectrum order and merging    Alpha_Shapes_2D.cu:308-522 (sort by FP64, exact fix-up of
 *                                       uncertain neighbours, equal values merged) and
 *                                       kernel.cu:360-401 (processSpectrum/UpdateBackRefs):
 *                                       rank = 1 + number of distinct smaller exact values;
 *                                       spectrum[0] = 0 is the attached/void slot.
 *   edge intervals                      kernel.cu:427-462 (edgeIntervals)
 *   vertex intervals                    kernel.cu:518-624 (computeVertexIntervals): over incident
 *                                       edges, nu_min = min(ro_rank if unattached else nu_min_rank),
 *                                       nu_max = max(nu_max_rank); points not in the DT get -1.
 *   master list                         kernel.cu:626-690 + Alpha_Shapes_2D.cu:641-727: entries
 *                                       (vertex nu_min, nu_max), (edge ro, nu_min, nu_max),
 *                                       (triangle ro) stably sorted by rank; entries with rank < 1
 *                                       dropped; spec_pointers = [-1] + first index of each rank.
 * The spectrum representative of a group of exactly equal values is the smallest FP64 value
 * in the group (the reference leaves this to thrust's sort order).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>

int ora_attached_exact(const double *i, const double *j, const double *k);
double ora_orient_exact_estimate(const double *a, const double *b, const double *c);

#define U53 (0x1.0p-53)

double ora_rho_tri(const double *pa, const double *pb, const double *pc)
{
    double x1 = pa[0], y1 = pa[1], x2 = pb[0], y2 = pb[1], x3 = pc[0], y3 = pc[1];
    double ax = x1 - x2, ay = y1 - y2, bx = x2 - x3, by = y2 - y3, cx = x3 - x1, cy = y3 - y1;
    ax = ax * ax; ay = ay * ay; bx = bx * bx; by = by * by; cx = cx * cx; cy = cy * cy;
    double a = ax + ay, b = bx + by, c = cx + cy;
    double ar1 = x2 * y3 - x3 * y2 - x1 * y3 + x3 * y1 + x1 * y2 - x2 * y1;
    double den = 4 * ar1 * ar1;
    return (a * b * c) / den;
}

double ora_rho_edge(const double *pa, const double *pb)
{
    double ax = pa[0] - pb[0], ay = pa[1] - pb[1];
    ax = ax * ax; ay = ay * ay;
    return (ax + ay) / 4;
}

static __thread long long g_exact_attach;   /* per thread: the oracle may run on several threads */

/* filtered attachment: the filter decides when |d1 - d2| exceeds a bound on its rounding
 * error; otherwise the exact expansion decides (>= 0 means attached). */
int ora_attached(const double *i, const double *j, const double *k)
{
    double dx = i[0] - j[0], dy = i[1] - j[1];
    double d1 = dx * dx + dy * dy;
    double sx = i[0] + j[0] - 2 * k[0], sy = i[1] + j[1] - 2 * k[1];
    double d2 = sx * sx + sy * sy;
    double diff = d1 - d2;
    double px = fabs(i[0] + j[0]) + 2 * fabs(k[0]), py = fabs(i[1] + j[1]) + 2 * fabs(k[1]);
    double perm = d1 + px * px + py * py;
    double eb = 32.0 * U53 * perm;
    if (diff > eb) return 1;
    if (-diff > eb) return 0;
    g_exact_attach++;
    return ora_attached_exact(i, j, k);
}

typedef struct {
    double val;      /* reference-expression rho (the stored spectrum value) */
    double lo, hi;   /* interval certainly containing the exact value */
    int st, idx;     /* st 0 = edge, 1 = triangle */
} sent;

static __thread const double *g_xy;
static __thread const ora_tri *g_t;
static __thread const int *g_et, *g_ev;
static __thread long long g_exact_cmp;

static void ent_pts(const sent *e, const double **p0, const double **p1, const double **p2)
{
    if (e->st == 1) {
        const ora_tri *T = &g_t[e->idx];
        *p0 = g_xy + 2 * (size_t)T->vtx[0]; *p1 = g_xy + 2 * (size_t)T->vtx[1]; *p2 = g_xy + 2 * (size_t)T->vtx[2];
    } else {
        const ora_tri *T = &g_t[g_et[e->idx]];
        int i = g_ev[e->idx];
        *p0 = g_xy + 2 * (size_t)T->vtx[(i + 1) % 3]; *p1 = g_xy + 2 * (size_t)T->vtx[(i + 2) % 3]; *p2 = NULL;
    }
}

static int exact_cmp(const sent *a, const sent *b)
{
    if (a->hi < b->lo) return -1;
    if (b->hi < a->lo) return 1;
    g_exact_cmp++;
    const double *a0, *a1, *a2, *b0, *b1, *b2;
    ent_pts(a, &a0, &a1, &a2);
    ent_pts(b, &b0, &b1, &b2);
    return ora_rho_compare(a->st, a0, a1, a2, b->st, b0, b1, b2);
}

static int cmp_val(const void *x, const void *y)
{
    const sent *a = (const sent *)x, *b = (const sent *)y;
    if (a->val != b->val) return a->val < b->val ? -1 : 1;
    if (a->st != b->st) return a->st - b->st;
    return a->idx - b->idx;
}

void ora_rho_dd(int st, const double *a0, const double *a1, const double *a2, double *hi, double *lo);

typedef struct { double hi, lo; sent e; } ddent;
static int cmp_dd(const void *x, const void *y)
{
    const ddent *a = (const ddent *)x, *b = (const ddent *)y;
    if (a->hi != b->hi) return a->hi < b->hi ? -1 : 1;
    if (a->lo != b->lo) return a->lo < b->lo ? -1 : 1;
    return 0;
}

static void msort(sent *v, sent *tmp, int n);

/* clusters of more than 4 entries: pre-sort by a double-double value computed from the exact
 * expansions (relative error below 2^-96), then split into groups whose consecutive values lie
 * within 2^-90 of each other; only inside a group can the order differ from the double-double
 * order, so each group is merge-sorted with the exact comparator.  grp[i] receives the group
 * id of entry i (entries of different groups have different exact values). */
static void sort_large_cluster(sent *v, int n, sent *tmp, int *grp, int *ngrp)
{
    ddent *d = (ddent *)malloc(sizeof(ddent) * n);
    for (int i = 0; i < n; i++) {
        const double *p0, *p1, *p2;
        ent_pts(&v[i], &p0, &p1, &p2);
        ora_rho_dd(v[i].st, p0, p1, p2, &d[i].hi, &d[i].lo);
        d[i].e = v[i];
    }
    qsort(d, n, sizeof(ddent), cmp_dd);
    for (int i = 0; i < n; i++) v[i] = d[i].e;
    const double tol = 0x1p-90;
    int gs = 0;
    for (int i = 0; i < n; i++) {
        int end = (i + 1 == n) ||
                  ((d[i + 1].hi - d[i].hi) + (d[i + 1].lo - d[i].lo) > tol * d[i + 1].hi);
        if (end) {
            if (i - gs + 1 >= 2) msort(v + gs, tmp, i - gs + 1);
            for (int k = gs; k <= i; k++) grp[k] = *ngrp;
            (*ngrp)++;
            gs = i + 1;
        }
    }
    free(d);
}

/* stable merge sort with the exact comparator */
static void msort(sent *v, sent *tmp, int n)
{
    if (n < 2) return;
    int h = n / 2;
    msort(v, tmp, h);
    msort(v + h, tmp, n - h);
    if (exact_cmp(&v[h - 1], &v[h]) <= 0) return;   /* halves already in order */
    int i = 0, j = h, k = 0;
    while (i < h && j < n) tmp[k++] = (exact_cmp(&v[j], &v[i]) < 0) ? v[j++] : v[i++];
    while (i < h) tmp[k++] = v[i++];
    while (j < n) tmp[k++] = v[j++];
    memcpy(v, tmp, sizeof(sent) * n);
}

int ora_alpha_compute(const double *xy, int n, const ora_tri *t, int nt, ora_alpha *A)
{
    memset(A, 0, sizeof(*A));
    ora_exactinit();
    g_exact_cmp = 0; g_exact_attach = 0;
    A->nTris = nt;
    A->nPoints = n;
    /* 1. edge ownership / numbering */
    int ne = 0;
    for (int k = 0; k < nt; k++)
        for (int i = 0; i < 3; i++)
            if (t[k].vtx[(i + 1) % 3] < t[k].vtx[(i + 2) % 3] || t[k].tri[i] == -1) ne++;
    A->nEdges = ne;
    A->edge_t = (int *)malloc(sizeof(int) * (ne + 1));
    A->edge_vopp = (int *)malloc(sizeof(int) * (ne + 1));
    A->edge_ro_rank = (int *)malloc(sizeof(int) * (ne + 1));
    A->edge_nu_min = (int *)malloc(sizeof(int) * (ne + 1));
    A->edge_nu_max = (int *)malloc(sizeof(int) * (ne + 1));
    A->edge_attached = (int *)malloc(sizeof(int) * (ne + 1));
    A->tri_ro_rank = (int *)malloc(sizeof(int) * (nt + 1));
    int e = 0;
    for (int k = 0; k < nt; k++)
        for (int i = 0; i < 3; i++)
            if (t[k].vtx[(i + 1) % 3] < t[k].vtx[(i + 2) % 3] || t[k].tri[i] == -1) {
                A->edge_t[e] = k; A->edge_vopp[e] = i; e++;
            }
    /* 2-3. rho, attachment */
    sent *S = (sent *)malloc(sizeof(sent) * (nt + ne + 1));
    int ns = 0;
    for (int k = 0; k < nt; k++) {
        const double *a = xy + 2 * (size_t)t[k].vtx[0], *b = xy + 2 * (size_t)t[k].vtx[1],
                     *c = xy + 2 * (size_t)t[k].vtx[2];
        double v = ora_rho_tri(a, b, c);
        double ar = ora_orient_exact_estimate(a, b, c);
        double la = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]);
        double lb = (b[0] - c[0]) * (b[0] - c[0]) + (b[1] - c[1]) * (b[1] - c[1]);
        double lc = (c[0] - a[0]) * (c[0] - a[0]) + (c[1] - a[1]) * (c[1] - a[1]);
        double acc = (la * lb * lc) / (4 * ar * ar);
        S[ns].val = v;
        S[ns].lo = acc - acc * (64 * U53);
        S[ns].hi = acc + acc * (64 * U53);
        S[ns].st = 1; S[ns].idx = k;
        ns++;
    }
    for (int j = 0; j < ne; j++) {
        const ora_tri *T = &t[A->edge_t[j]];
        int i = A->edge_vopp[j];
        const double *p = xy + 2 * (size_t)T->vtx[(i + 1) % 3], *q = xy + 2 * (size_t)T->vtx[(i + 2) % 3];
        int att = ora_attached(p, q, xy + 2 * (size_t)T->vtx[i]);
        if (!att && T->tri[i] >= 0) {
            const ora_tri *O = &t[T->tri[i] >> 2];
            att = ora_attached(p, q, xy + 2 * (size_t)O->vtx[T->tri[i] & 3]);
        }
        A->edge_attached[j] = att;
        if (att) continue;
        double v = ora_rho_edge(p, q);
        S[ns].val = v;
        S[ns].lo = v - v * (16 * U53);
        S[ns].hi = v + v * (16 * U53);
        S[ns].st = 0; S[ns].idx = j;
        ns++;
    }
    /* 4. exact order: FP64 sort, then exact merge sort inside each cluster of entries whose
       certified intervals overlap (cluster bounds from prefix-max(hi) / suffix-min(lo)). */
    g_xy = xy; g_t = t; g_et = A->edge_t; g_ev = A->edge_vopp;
    qsort(S, ns, sizeof(sent), cmp_val);
    double *sufmin = (double *)malloc(sizeof(double) * (ns + 1));
    sufmin[ns] = INFINITY;
    for (int k = ns - 1; k >= 0; k--) sufmin[k] = fmin(sufmin[k + 1], S[k].lo);
    sent *tmp = (sent *)malloc(sizeof(sent) * (ns + 1));
    /* grp[k] >= 0: id of a double-double group (exact compare needed only inside one) */
    int *grp = (int *)malloc(sizeof(int) * (ns + 1));
    int ngrp = 0;
    for (int k = 0; k < ns; k++) grp[k] = -1;
    double pmax = -INFINITY;
    int cs = 0;
    for (int k = 0; k < ns; k++) {
        pmax = fmax(pmax, S[k].hi);
        if (pmax < sufmin[k + 1]) {       /* boundary after k */
            int m = k - cs + 1;
            if (m > 4) sort_large_cluster(S + cs, m, tmp, grp + cs, &ngrp);
            else if (m >= 2) msort(S + cs, tmp, m);
            cs = k + 1;
        }
    }
    free(tmp); free(sufmin);
    /* 5. ranks */
    A->spectrum = (double *)malloc(sizeof(double) * (ns + 1));
    A->spectrum[0] = 0.0;
    int rank = 0;
    for (int k = 0; k < ns; k++) {
        int same = (k > 0) && (grp[k] < 0 || grp[k] == grp[k - 1]) && exact_cmp(&S[k - 1], &S[k]) == 0;
        if (!same) { rank++; A->spectrum[rank] = S[k].val; }
        else if (S[k].val < A->spectrum[rank]) A->spectrum[rank] = S[k].val;
        if (S[k].st == 1) A->tri_ro_rank[S[k].idx] = rank;
        else A->edge_ro_rank[S[k].idx] = rank;
    }
    A->specSize = rank + 1;
    free(grp);
    for (int j = 0; j < ne; j++) if (A->edge_attached[j]) A->edge_ro_rank[j] = 0;
    free(S);
    /* 6. edge intervals */
    for (int j = 0; j < ne; j++) {
        const ora_tri *T = &t[A->edge_t[j]];
        int r1 = A->tri_ro_rank[A->edge_t[j]];
        int o = T->tri[A->edge_vopp[j]];
        if (o < 0) { A->edge_nu_min[j] = r1; A->edge_nu_max[j] = 0; }
        else {
            int r2 = A->tri_ro_rank[o >> 2];
            A->edge_nu_min[j] = r1 <= r2 ? r1 : r2;
            A->edge_nu_max[j] = r1 <= r2 ? r2 : r1;
        }
    }
    /* 7. vertex intervals over incident edges */
    A->vert_nu_min = (int *)malloc(sizeof(int) * (n + 1));
    A->vert_nu_max = (int *)malloc(sizeof(int) * (n + 1));
    for (int i = 0; i < n; i++) { A->vert_nu_min[i] = -1; A->vert_nu_max[i] = -1; }
    for (int j = 0; j < ne; j++) {
        const ora_tri *T = &t[A->edge_t[j]];
        int i = A->edge_vopp[j];
        int key = A->edge_ro_rank[j] != 0 ? A->edge_ro_rank[j] : A->edge_nu_min[j];
        int ends[2] = {T->vtx[(i + 1) % 3], T->vtx[(i + 2) % 3]};
        for (int s = 0; s < 2; s++) {
            int v = ends[s];
            if (A->vert_nu_min[v] < 0 || key < A->vert_nu_min[v]) A->vert_nu_min[v] = key;
            if (A->vert_nu_max[v] < 0 || A->edge_nu_max[j] > A->vert_nu_max[v]) A->vert_nu_max[v] = A->edge_nu_max[j];
        }
    }
    /* 8. master list: counting sort by rank (stable) */
    long M = 2L * n + 3L * ne + nt;
    int *key = (int *)malloc(sizeof(int) * (M + 1));
    long m = 0;
    for (int i = 0; i < n; i++) { key[m++] = A->vert_nu_min[i]; key[m++] = A->vert_nu_max[i]; }
    for (int j = 0; j < ne; j++) { key[m++] = A->edge_ro_rank[j]; key[m++] = A->edge_nu_min[j]; key[m++] = A->edge_nu_max[j]; }
    for (int k = 0; k < nt; k++) key[m++] = A->tri_ro_rank[k];
    int R = A->specSize;
    long *cnt = (long *)calloc(R + 1, sizeof(long));
    long kept = 0;
    for (long q = 0; q < M; q++) if (key[q] >= 1) { cnt[key[q]]++; kept++; }
    long *start = (long *)malloc(sizeof(long) * (R + 1));
    long acc = 0;
    for (int r = 0; r <= R; r++) { start[r] = acc; acc += cnt[r]; }
    A->masterSize = (int)kept;
    A->master_simplex = (int *)malloc(sizeof(int) * (kept + 1));
    A->master_st = (int *)malloc(sizeof(int) * (kept + 1));
    A->master_at = (int *)malloc(sizeof(int) * (kept + 1));
    long *pos = (long *)malloc(sizeof(long) * (R + 1));
    memcpy(pos, start, sizeof(long) * (R + 1));
    for (long q = 0; q < M; q++) {
        if (key[q] < 1) continue;
        long d = pos[key[q]]++;
        int simplex, st, at;
        if (q < 2L * n) { simplex = (int)(q / 2); st = 2; at = (int)(q % 2) + 1; }
        else if (q < 2L * n + 3L * ne) { long r = q - 2L * n; simplex = (int)(r / 3); st = 0; at = (int)(r % 3); }
        else { simplex = (int)(q - 2L * n - 3L * ne); st = 1; at = 0; }
        A->master_simplex[d] = simplex; A->master_st[d] = st; A->master_at[d] = at;
    }
    int np = 1;
    for (int r = 1; r <= R; r++) if (cnt[r] > 0) np++;
    A->nSpecPointers = np;
    A->spec_pointers = (int *)malloc(sizeof(int) * np);
    A->spec_pointers[0] = -1;
    np = 1;
    for (int r = 1; r <= R; r++) if (cnt[r] > 0) A->spec_pointers[np++] = (int)start[r];
    free(cnt); free(start); free(pos); free(key);
    A->exact_compares = g_exact_cmp;
    A->exact_attach = g_exact_attach;
    return 0;
}

void ora_alpha_free(ora_alpha *A)
{
    free(A->spectrum); free(A->tri_ro_rank); free(A->edge_t); free(A->edge_vopp);
    free(A->edge_ro_rank); free(A->edge_nu_min); free(A->edge_nu_max); free(A->edge_attached);
    free(A->vert_nu_min); free(A->vert_nu_max);
    free(A->master_simplex); free(A->master_st); free(A->master_at); free(A->spec_pointers);
    memset(A, 0, sizeof(*A));
}
