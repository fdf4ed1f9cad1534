/*
 * oracle.h -- CPU oracle for the 2D Delaunay + alpha-shape path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (galpha2d_amd/)
 * includes, links or calls this code.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * What each part restates (reference paths relative to the reference root):
 *   predicates.c  Shewchuk adaptive orient2d / incircle
 *                 (gpudt/predicates.cpp:28-983), plus a symbolic-perturbation
 *                 tie-break for exactly co-circular quadruples (the reference
 *                 has none: gpudt.cpp:160-189 and Cuda/cudaCCW.h:565-598 use a
 *                 strict > 0 test, so its output is order dependent there).
 *   dt.c          sequential Delaunay triangulation (Bowyer-Watson with ghost
 *                 triangles).  The reference has no CPU DT; the output format
 *                 follows gpudt/gpudt.h:97-102 (tri[i] = encoded neighbour
 *                 across the edge opposite vtx[i], -1 on the hull).
 *   alpha.c       per-simplex rho, attachment, edge ownership, exact spectrum
 *                 ranks, edge/vertex intervals and the master list
 *                 (AlphaShape/predicates.h:1297-1444, AlphaShape/kernel.cu:
 *                 160-690, AlphaShape/Alpha_Shapes_2D.cu:127-727).
 *   canon.c       canonical triangle order and the validator that restates
 *                 checkResult (main.cu:163-343) with extra checks.
 *   cdt.c         constrained DT: DT, then Sloan's edge-flip insertion of each
 *                 constraint, then Lawson flips of the other edges (the result of
 *                 Cuda/cudaConstraint.cu:828-939 + Cuda/cudaFlipping.cu:111-220 +
 *                 gpudt.cpp:88-155).
 */
#ifndef GALPHA_ORACLE_H
#define GALPHA_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y; } ora_pt;
typedef struct { int tri[3]; int vtx[3]; int tmp[3]; } ora_tri; /* gpudtTriangle layout */

/* predicates */
void   ora_exactinit(void);
double ora_orient2d(const double *a, const double *b, const double *c);     /* Shewchuk sign value */
double ora_incircle(const double *a, const double *b, const double *c, const double *d);
int    ora_orient_sign(const double *a, const double *b, const double *c);
int    ora_incircle_sign(const double *a, const double *b, const double *c, const double *d);
/* incircle with the lifted-perturbation tie-break (never 0 unless a,b,c,d collinear) */
int    ora_incircle_sos(const double *a, int ia, const double *b, int ib,
                        const double *c, int ic, const double *d, int id);
/* counters of the exact stage (for fallback-rate reporting) */
void   ora_pred_counters(long long *orient_calls, long long *orient_exact,
                         long long *incircle_calls, long long *incircle_exact,
                         long long *sos_ties);
void   ora_pred_reset_counters(void);

/* sequential DT.  points: n (x,y) pairs.  Returns number of triangles written
 * into *out (malloc'ed, caller frees with free()).  Triangles are canonical:
 * each rotated so vtx[0] is the smallest index, ccw, sorted by (vtx0,vtx1,vtx2).
 * dup_of[i] (optional, n ints) receives -1 for kept points, else the index of
 * the kept point with identical coordinates.  Returns -1 on allocation failure. */
int ora_delaunay(const double *xy, int n, ora_tri **out, int *dup_of);

/* canonicalise an arbitrary valid triangle list in place (rotation + order +
 * neighbour re-encoding). */
void ora_canonicalize(ora_tri *t, int nt);

/* validator.  Returns 0 when every check passes, else a positive code and a
 * message in msg (size msglen).  Checks: index range, ccw (exact), reciprocal
 * adjacency, every non-duplicate point used, Euler count for a disk, hull is a
 * convex cycle, every interior edge locally Delaunay under the tie-break. */
int ora_validate(const double *xy, int n, const ora_tri *t, int nt,
                 const int *dup_of, char *msg, int msglen);

/* validator for a constrained DT: the checks above, with constraint edges exempt from the
 * incircle test, and every constraint (cons: ncons pairs of point indices) an edge of the mesh
 * (restates main.cu:218-286).  A mesh that passes is the constrained DT. */
int ora_validate_constrained(const double *xy, int n, const ora_tri *t, int nt, const int *dup_of,
                             const int *cons, int ncons, char *msg, int msglen);

/* constrained DT (cdt.c): returns the number of canonical triangles written to *out (malloc'ed)
 * or -1 with a message (invalid constraints: endpoint out of range, equal endpoints, duplicate
 * endpoint, a point inside a constraint, crossing constraints). */
int ora_cdt(const double *xy, int n, const int *cons, int ncons, ora_tri **out, int *dup_of,
            char *msg, int msglen);

/* alpha shapes on a canonical DT (layout of AlphaShape/headers.h:48-90) */
typedef struct {
    int nEdges, nTris, nPoints;
    int specSize;              /* unique spectrum values incl. index 0 (= 0.0) */
    double *spectrum;          /* [specSize] */
    int *tri_ro_rank;          /* [nTris] */
    int *edge_t, *edge_vopp;   /* [nEdges] owning triangle and local index of the opposite vertex */
    int *edge_ro_rank, *edge_nu_min, *edge_nu_max, *edge_attached;  /* [nEdges] */
    int *vert_nu_min, *vert_nu_max;     /* [nPoints] */
    int masterSize;
    int *master_simplex, *master_st, *master_at;   /* [masterSize] */
    int nSpecPointers;
    int *spec_pointers;        /* [nSpecPointers], [0] = -1 */
    long long exact_compares;  /* comparisons that needed the exact path */
    long long exact_attach;    /* attachment tests that needed the exact path */
} ora_alpha;

int  ora_alpha_compute(const double *xy, int n, const ora_tri *t, int nt, ora_alpha *out);
void ora_alpha_free(ora_alpha *a);

/* exact comparison of two spectrum entries: st 0 = edge (p,q), 1 = triangle (p,q,r).
 * Returns -1, 0, +1 for rho_a <,=,> rho_b (exact real values). */
int ora_rho_compare(int sta, const double *a0, const double *a1, const double *a2,
                    int stb, const double *b0, const double *b1, const double *b2);
double ora_rho_tri(const double *a, const double *b, const double *c);
double ora_rho_edge(const double *a, const double *b);
int    ora_attached(const double *i, const double *j, const double *k);

/* seeded generators shared with bench/tests (splitmix64 -> (u>>11)*2^-53) */
void ora_gen_uniform(uint64_t seed, int n, double *xy);

#ifdef __cplusplus
}
#endif
#endif
