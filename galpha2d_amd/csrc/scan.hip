// scan.hip -- exclusive prefix sum of int32 arrays (replaces Cuda/cudaScanLargeArray.cu, which is
// not reused): rocPRIM's single-pass scan with decoupled look-back.
#include <cstring>
#include <stdexcept>
#include <string>
#include <rocprim/rocprim.hpp>
#include "dev_common.h"
#include "kernels.h"

// temp bytes of one exclusive scan of n ints (rocPRIM's single-pass decoupled look-back scan:
// each element is read once and written once)
size_t scan_temp_bytes(size_t n)
{
    size_t bytes = 0;
    (void)rocprim::exclusive_scan(nullptr, bytes, (const int *)nullptr, (int *)nullptr, 0, n, rocprim::plus<int>(),
                                  (hipStream_t)0);
    return bytes + 256;
}

// out[i] = in[0] + ... + in[i - 1]; temp holds at least scan_temp_bytes(n) bytes; in != out
void launch_exclusive_scan(const int *in, int *out, size_t n, void *temp, hipStream_t st)
{
    if (n == 0) return;
    size_t bytes = 0;
    (void)rocprim::exclusive_scan(nullptr, bytes, in, out, 0, n, rocprim::plus<int>(), st);
    hipError_t e = rocprim::exclusive_scan(temp, bytes, in, out, 0, n, rocprim::plus<int>(), st);
    if (e != hipSuccess) throw std::runtime_error(std::string("exclusive scan: ") + hipGetErrorString(e));
}
