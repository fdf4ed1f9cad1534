/*
 * predicates.c -- exact geometric predicates for the oracle (test infrastructure only).
 *
 * Restates J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
 * Robust Geometric Predicates" (1997): error-free transforms, nonoverlapping
 * expansions with zero elimination, a semi-static first stage, and an exact stage.
 * The reference ships the same published method as gpudt/predicates.cpp
 * (gpudt_exactinit :28, counterclockwiseadapt :242, incircleadapt :373).
 * The error-bound constants are the published ones (they must be alike: they are
 * the proof's constants).
 *
 * Build with -ffp-contract=off: a fused multiply-add breaks the error-free
 * transforms.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static double splitter, epsilon;
static double ccwerrboundA, iccerrboundA;
static int inited = 0;
static __thread long long n_orient, n_orient_exact, n_incircle, n_incircle_exact, n_sos;   /* per thread */

void ora_exactinit(void)
{
    double half = 0.5, check = 1.0, lastcheck;
    int every_other = 1;
    epsilon = 1.0;
    splitter = 1.0;
    do {
        lastcheck = check;
        epsilon *= half;
        if (every_other) splitter *= 2.0;
        every_other = !every_other;
        check = 1.0 + epsilon;
    } while (check != 1.0 && check != lastcheck);
    splitter += 1.0;
    ccwerrboundA = (3.0 + 16.0 * epsilon) * epsilon;
    iccerrboundA = (10.0 + 96.0 * epsilon) * epsilon;
    inited = 1;
}

void ora_pred_counters(long long *oc, long long *oe, long long *ic, long long *ie, long long *st)
{
    if (oc) *oc = n_orient;
    if (oe) *oe = n_orient_exact;
    if (ic) *ic = n_incircle;
    if (ie) *ie = n_incircle_exact;
    if (st) *st = n_sos;
}
void ora_pred_reset_counters(void) { n_orient = n_orient_exact = n_incircle = n_incircle_exact = n_sos = 0; }

/* ---- error-free transforms ------------------------------------------------ */
static inline void two_sum(double a, double b, double *x, double *y)
{
    double s = a + b, bv = s - a, av = s - bv;
    *x = s;
    *y = (a - av) + (b - bv);
}
static inline void split(double a, double *hi, double *lo)
{
    double c = splitter * a, big = c - a;
    *hi = c - big;
    *lo = a - *hi;
}
static inline void two_product(double a, double b, double *x, double *y)
{
    double p = a * b, ahi, alo, bhi, blo;
    split(a, &ahi, &alo);
    split(b, &bhi, &blo);
    double e1 = p - ahi * bhi;
    double e2 = e1 - alo * bhi;
    double e3 = e2 - ahi * blo;
    *x = p;
    *y = alo * blo - e3;
}

/* h = e + f (both nonoverlapping, increasing magnitude).  Components are merged by
 * increasing magnitude and accumulated with error-free two_sum, dropping zeros
 * (the fast_expansion_sum_zeroelim scheme of the paper, with two_sum throughout). */
static int exp_sum(int elen, const double *e, int flen, const double *f, double *h)
{
    int n = elen + flen, ei = 0, fi = 0, hi = 0;
    if (n == 0) { h[0] = 0.0; return 1; }
    double *g = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n; k++) {
        if (fi >= flen || (ei < elen && fabs(e[ei]) <= fabs(f[fi]))) g[k] = e[ei++];
        else g[k] = f[fi++];
    }
    double Q = g[0], Qnew, hh;
    for (int k = 1; k < n; k++) {
        two_sum(Q, g[k], &Qnew, &hh);
        if (hh != 0.0) h[hi++] = hh;
        Q = Qnew;
    }
    if (Q != 0.0 || hi == 0) h[hi++] = Q;
    free(g);
    return hi;
}

/* h = b * e, zero-eliminating. */
static int exp_scale(int elen, const double *e, double b, double *h)
{
    double Q, sum, hh, p1, p0;
    int hi = 0;
    two_product(e[0], b, &Q, &hh);
    if (hh != 0.0) h[hi++] = hh;
    for (int i = 1; i < elen; i++) {
        two_product(e[i], b, &p1, &p0);
        two_sum(Q, p0, &sum, &hh);
        if (hh != 0.0) h[hi++] = hh;
        two_sum(p1, sum, &Q, &hh);
        if (hh != 0.0) h[hi++] = hh;
    }
    if (Q != 0.0 || hi == 0) h[hi++] = Q;
    return hi;
}

/* Generic exact expansion value with heap storage (oracle only). */
typedef struct { int n, cap; double *v; } xexp;

static void xe_init(xexp *x) { x->n = 0; x->cap = 0; x->v = NULL; }
static void xe_free(xexp *x) { free(x->v); x->v = NULL; x->n = x->cap = 0; }
static void xe_reserve(xexp *x, int cap)
{
    if (x->cap < cap) {
        x->cap = cap < 16 ? 16 : cap;
        x->v = (double *)realloc(x->v, sizeof(double) * x->cap);
    }
}
static void xe_set2(xexp *x, double hi, double lo)
{
    xe_reserve(x, 2);
    x->n = 0;
    if (lo != 0.0) x->v[x->n++] = lo;
    x->v[x->n++] = hi;
}
static void xe_add(xexp *r, const xexp *a, const xexp *b)
{
    xexp t; xe_init(&t);
    xe_reserve(&t, a->n + b->n + 1);
    t.n = exp_sum(a->n, a->v, b->n, b->v, t.v);
    xe_free(r);
    *r = t;
}
static void xe_neg(xexp *a) { for (int i = 0; i < a->n; i++) a->v[i] = -a->v[i]; }
static void xe_mul(xexp *r, const xexp *a, const xexp *b)
{
    xexp acc, tmp, part;
    xe_init(&acc); xe_init(&tmp); xe_init(&part);
    xe_reserve(&acc, 1);
    acc.v[0] = 0.0; acc.n = 1;
    for (int i = 0; i < b->n; i++) {
        xe_reserve(&part, 2 * a->n + 1);
        part.n = exp_scale(a->n, a->v, b->v[i], part.v);
        xe_reserve(&tmp, acc.n + part.n + 1);
        tmp.n = exp_sum(acc.n, acc.v, part.n, part.v, tmp.v);
        xexp sw = acc; acc = tmp; tmp = sw;
    }
    xe_free(&tmp); xe_free(&part);
    xe_free(r);
    *r = acc;
}
static int xe_sign(const xexp *a)
{
    double top = a->n ? a->v[a->n - 1] : 0.0;
    return (top > 0) - (top < 0);
}
/* exact a - b as an expansion */
static void xe_diff(xexp *r, double a, double b)
{
    double x, y;
    two_sum(a, -b, &x, &y);
    xe_set2(r, x, y);
}

/* exact orient2d value as expansion: (ax*by - ay*bx) + (bx*cy - by*cx) + (cx*ay - cy*ax) */
static void orient_expansion(const double *a, const double *b, const double *c, xexp *out)
{
    double p[6][2];
    two_product(a[0], b[1], &p[0][0], &p[0][1]);
    two_product(-a[1], b[0], &p[1][0], &p[1][1]);
    two_product(b[0], c[1], &p[2][0], &p[2][1]);
    two_product(-b[1], c[0], &p[3][0], &p[3][1]);
    two_product(c[0], a[1], &p[4][0], &p[4][1]);
    two_product(-c[1], a[0], &p[5][0], &p[5][1]);
    xexp acc, term, tmp;
    xe_init(&acc); xe_init(&term); xe_init(&tmp);
    xe_set2(&acc, p[0][0], p[0][1]);
    for (int i = 1; i < 6; i++) {
        xe_set2(&term, p[i][0], p[i][1]);
        xe_add(&tmp, &acc, &term);
        xexp sw = acc; acc = tmp; tmp = sw;
    }
    xe_free(&term); xe_free(&tmp);
    xe_free(out);
    *out = acc;
}

static int orient_exact_sign(const double *a, const double *b, const double *c)
{
    xexp e; xe_init(&e);
    orient_expansion(a, b, c, &e);
    int s = xe_sign(&e);
    xe_free(&e);
    return s;
}

double ora_orient2d(const double *a, const double *b, const double *c)
{
    if (!inited) ora_exactinit();
    n_orient++;
    double detleft = (a[0] - c[0]) * (b[1] - c[1]);
    double detright = (a[1] - c[1]) * (b[0] - c[0]);
    double det = detleft - detright, detsum;
    if (detleft > 0.0) {
        if (detright <= 0.0) return det;
        detsum = detleft + detright;
    } else if (detleft < 0.0) {
        if (detright >= 0.0) return det;
        detsum = -detleft - detright;
    } else {
        return det;
    }
    double errbound = ccwerrboundA * detsum;
    if (det >= errbound || -det >= errbound) return det;
    n_orient_exact++;
    return (double)orient_exact_sign(a, b, c);
}

int ora_orient_sign(const double *a, const double *b, const double *c)
{
    double d = ora_orient2d(a, b, c);
    return (d > 0) - (d < 0);
}

/* exact lifted 4x4 determinant: za*O(b,c,d) - zb*O(a,c,d) + zc*O(a,b,d) - zd*O(a,b,c) */
static int incircle_exact_sign(const double *a, const double *b, const double *c, const double *d)
{
    const double *P[4] = {a, b, c, d};
    xexp acc, lift, sq, o, prod, tmp;
    xe_init(&acc); xe_init(&lift); xe_init(&sq); xe_init(&o); xe_init(&prod); xe_init(&tmp);
    xe_reserve(&acc, 1); acc.v[0] = 0.0; acc.n = 1;
    for (int k = 0; k < 4; k++) {
        const double *q = P[k];
        double x1, x0, y1, y0;
        two_product(q[0], q[0], &x1, &x0);
        two_product(q[1], q[1], &y1, &y0);
        xe_set2(&lift, x1, x0);
        xe_set2(&sq, y1, y0);
        xe_add(&tmp, &lift, &sq);
        { xexp sw = lift; lift = tmp; tmp = sw; }
        const double *r[3];
        int m = 0;
        for (int j = 0; j < 4; j++) if (j != k) r[m++] = P[j];
        orient_expansion(r[0], r[1], r[2], &o);
        xe_mul(&prod, &lift, &o);
        if (k & 1) xe_neg(&prod);
        xe_add(&tmp, &acc, &prod);
        { xexp sw = acc; acc = tmp; tmp = sw; }
    }
    int s = xe_sign(&acc);
    xe_free(&acc); xe_free(&lift); xe_free(&sq); xe_free(&o); xe_free(&prod); xe_free(&tmp);
    return s;
}

double ora_incircle(const double *a, const double *b, const double *c, const double *d)
{
    if (!inited) ora_exactinit();
    n_incircle++;
    double adx = a[0] - d[0], bdx = b[0] - d[0], cdx = c[0] - d[0];
    double ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double alift = adx * adx + ady * ady;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double blift = bdx * bdx + bdy * bdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double clift = cdx * cdx + cdy * cdy;
    double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
    double permanent = (fabs(bdxcdy) + fabs(cdxbdy)) * alift + (fabs(cdxady) + fabs(adxcdy)) * blift
                     + (fabs(adxbdy) + fabs(bdxady)) * clift;
    double errbound = iccerrboundA * permanent;
    if (det > errbound || -det > errbound) return det;
    n_incircle_exact++;
    return (double)incircle_exact_sign(a, b, c, d);
}

int ora_incircle_sign(const double *a, const double *b, const double *c, const double *d)
{
    double v = ora_incircle(a, b, c, d);
    return (v > 0) - (v < 0);
}

/*
 * Tie-break for exactly co-circular quadruples.  Each lifted height z_k = x_k^2 + y_k^2
 * is perturbed by eps^(k+1) (k = global point index, eps infinitesimal), so a smaller
 * index carries the more significant perturbation.  The lifted determinant is linear in
 * the heights: dD/dz_a = O(b,c,d), dD/dz_b = -O(a,c,d), dD/dz_c = O(a,b,d),
 * dD/dz_d = -O(a,b,c).  On an exact zero, the sign is that of the coefficient of the
 * smallest-index point whose coefficient is nonzero.  This is a regular triangulation
 * with infinitesimal weights, so the result is unique and every point stays a vertex.
 */
int ora_incircle_sos(const double *a, int ia, const double *b, int ib,
                     const double *c, int ic, const double *d, int id)
{
    int s = ora_incircle_sign(a, b, c, d);
    if (s != 0) return s;
    n_sos++;
    int idx[4] = {ia, ib, ic, id};
    int order[4] = {0, 1, 2, 3};
    for (int i = 1; i < 4; i++)
        for (int j = i; j > 0 && idx[order[j]] < idx[order[j - 1]]; j--) {
            int t = order[j]; order[j] = order[j - 1]; order[j - 1] = t;
        }
    for (int r = 0; r < 4; r++) {
        int k = order[r], o;
        switch (k) {
        case 0: o = ora_orient_sign(b, c, d); break;
        case 1: o = -ora_orient_sign(a, c, d); break;
        case 2: o = ora_orient_sign(a, b, d); break;
        default: o = -ora_orient_sign(a, b, c); break;
        }
        if (o != 0) return o;
    }
    return 0;
}

/* ---- exact alpha quantities (used by alpha.c) ----------------------------- */

/* |p - q|^2 exactly */
static void sqdist_expansion(const double *p, const double *q, xexp *out)
{
    xexp dx, dy, sx, sy;
    xe_init(&dx); xe_init(&dy); xe_init(&sx); xe_init(&sy);
    xe_diff(&dx, p[0], q[0]);
    xe_diff(&dy, p[1], q[1]);
    xe_mul(&sx, &dx, &dx);
    xe_mul(&sy, &dy, &dy);
    xe_add(out, &sx, &sy);
    xe_free(&dx); xe_free(&dy); xe_free(&sx); xe_free(&sy);
}

/* numerator N = |ab|^2 |bc|^2 |ca|^2 and D = ar^2 (rho = N / (4 D)) */
static void tri_num_den(const double *a, const double *b, const double *c, xexp *num, xexp *den)
{
    xexp l1, l2, l3, t, ar;
    xe_init(&l1); xe_init(&l2); xe_init(&l3); xe_init(&t); xe_init(&ar);
    sqdist_expansion(a, b, &l1);
    sqdist_expansion(b, c, &l2);
    sqdist_expansion(c, a, &l3);
    xe_mul(&t, &l1, &l2);
    xe_mul(num, &t, &l3);
    orient_expansion(a, b, c, &ar);
    xe_mul(den, &ar, &ar);
    xe_free(&l1); xe_free(&l2); xe_free(&l3); xe_free(&t); xe_free(&ar);
}

int ora_rho_compare(int sta, const double *a0, const double *a1, const double *a2,
                    int stb, const double *b0, const double *b1, const double *b2)
{
    if (!inited) ora_exactinit();
    /* value = num/den with den > 0 (edge: den = 1, num = |pq|^2 ; tri: num = N, den = D)
       both rho carry the same factor 1/4, which cancels. */
    xexp na, da, nb, db, l, r, tmp;
    xe_init(&na); xe_init(&da); xe_init(&nb); xe_init(&db);
    xe_init(&l); xe_init(&r); xe_init(&tmp);
    if (sta == 0) { sqdist_expansion(a0, a1, &na); xe_reserve(&da, 1); da.v[0] = 1.0; da.n = 1; }
    else tri_num_den(a0, a1, a2, &na, &da);
    if (stb == 0) { sqdist_expansion(b0, b1, &nb); xe_reserve(&db, 1); db.v[0] = 1.0; db.n = 1; }
    else tri_num_den(b0, b1, b2, &nb, &db);
    xe_mul(&l, &na, &db);
    xe_mul(&r, &nb, &da);
    xe_neg(&r);
    xe_add(&tmp, &l, &r);
    int s = xe_sign(&tmp);
    xe_free(&na); xe_free(&da); xe_free(&nb); xe_free(&db);
    xe_free(&l); xe_free(&r); xe_free(&tmp);
    return s;
}

/* attachment (AlphaShape/predicates.h:1359-1444): the exact sign of
 * |i-j|^2 - |i+j-2k|^2 ; the edge ij is attached by k iff it is >= 0. */
int ora_attached_exact(const double *i, const double *j, const double *k)
{
    xexp d1, sx, sy, tx, ty, a, b, tmp;
    xe_init(&d1); xe_init(&sx); xe_init(&sy); xe_init(&tx); xe_init(&ty);
    xe_init(&a); xe_init(&b); xe_init(&tmp);
    sqdist_expansion(i, j, &d1);
    /* s = i + j - 2k per coordinate, exactly */
    for (int c = 0; c < 2; c++) {
        double h, l;
        xexp e1, e2, s;
        xe_init(&e1); xe_init(&e2); xe_init(&s);
        two_sum(i[c], j[c], &h, &l);
        xe_set2(&e1, h, l);
        xe_reserve(&e2, 1); e2.v[0] = -2.0 * k[c]; e2.n = 1;   /* exact (power-of-two scale) */
        if (e2.v[0] == 0.0) { e2.n = 0; }
        xe_add(&s, &e1, &e2);
        xe_mul(c == 0 ? &tx : &ty, &s, &s);
        xe_free(&e1); xe_free(&e2); xe_free(&s);
    }
    xe_add(&a, &tx, &ty);
    xe_neg(&a);
    xe_add(&tmp, &d1, &a);
    int s = xe_sign(&tmp);
    xe_free(&d1); xe_free(&sx); xe_free(&sy); xe_free(&tx); xe_free(&ty);
    xe_free(&a); xe_free(&b); xe_free(&tmp);
    return s >= 0;
}

/* value of the exact orientation expansion rounded to double (relative error < 2^-52) */
double ora_orient_exact_estimate(const double *a, const double *b, const double *c)
{
    if (!inited) ora_exactinit();
    xexp e; xe_init(&e);
    orient_expansion(a, b, c, &e);
    double s = 0;
    for (int i = 0; i < e.n; i++) s += e.v[i];
    xe_free(&e);
    return s;
}

/* double-double approximation (hi + lo) of a nonoverlapping expansion stored in increasing
 * magnitude: each component is added with an error-free two_sum and the pair renormalised,
 * so the relative error stays near n * 2^-106. */
static void xe_to_dd(const xexp *e, double *hi, double *lo)
{
    double h = 0.0, l = 0.0;
    for (int i = 0; i < e->n; i++) {
        double s, t;
        two_sum(h, e->v[i], &s, &t);
        t += l;
        h = s + t;
        l = t - (h - s);
    }
    *hi = h;
    *lo = l;
}

/* rho of an entry (st 0 edge, 1 triangle) as a double-double computed from the exact
 * numerator and denominator expansions (relative error about 2^-100). */
void ora_rho_dd(int st, const double *a0, const double *a1, const double *a2, double *hi, double *lo)
{
    if (!inited) ora_exactinit();
    xexp num, den;
    xe_init(&num); xe_init(&den);
    double nh, nl, dh = 1.0, dl = 0.0;
    if (st == 0) sqdist_expansion(a0, a1, &num);
    else tri_num_den(a0, a1, a2, &num, &den);
    xe_to_dd(&num, &nh, &nl);
    if (st != 0) xe_to_dd(&den, &dh, &dl);
    /* (nh + nl) / (4 (dh + dl)) in double-double */
    double q1 = nh / dh;
    double p, pe;
    two_product(q1, dh, &p, &pe);
    double r = ((nh - p) - pe + nl - q1 * dl) / dh;
    double h, l;
    two_sum(q1, r, &h, &l);
    *hi = h * 0.25;
    *lo = l * 0.25;
    xe_free(&num); xe_free(&den);
}
