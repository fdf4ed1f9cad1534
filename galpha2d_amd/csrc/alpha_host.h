// alpha_host.h -- alpha-shape stage (device buffers + host entry points) shared with galpha.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>
#include "../../include/gpudt.h"

struct GalphaError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct AlphaBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct AlphaDev {
    AlphaBuf b[48];
    int n = 0, F = 0, E = 0, S = 0, specSize = 0, M = 0, nsp = 0;
};

void alpha_run(AlphaDev &a, const double *d_xy, int n, const int *d_tris9, int F, hipStream_t st,
               galpha_stats *stats);
PGPUDTALPHA alpha_copy_to_host(AlphaDev &a, hipStream_t st);
void alpha_free_host(PGPUDTALPHA p);
void alpha_release(AlphaDev &a);
