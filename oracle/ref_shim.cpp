// ref_shim.cpp -- extern "C" wrappers around the reference's own CPU predicates
// (gpudt/predicates.cpp: gpudt_exactinit :28, counterclockwise :322, incircle :942), compiled
// from the reference sources in place into oracle/_ref/.  Test infrastructure only: it lets the
// tests compare the oracle's predicates with the reference's on the same inputs.
#include "predicates.h"

static int g_init = 0;
static void init()
{
    if (!g_init) {
        gpudt_exactinit();
        g_init = 1;
    }
}

extern "C" double ref_orient2d(const double *a, const double *b, const double *c)
{
    init();
    gpudtVertex A = {a[0], a[1]}, B = {b[0], b[1]}, C = {c[0], c[1]};
    return counterclockwise(&A, &B, &C);
}

extern "C" double ref_incircle(const double *a, const double *b, const double *c, const double *d)
{
    init();
    gpudtVertex A = {a[0], a[1]}, B = {b[0], b[1]}, C = {c[0], c[1]}, D = {d[0], d[1]};
    return incircle(&A, &B, &C, &D);
}
