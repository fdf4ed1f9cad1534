// sort.hip -- radix sorts of (key, value) pairs via rocPRIM's device radix sort (a library
// primitive, like the thrust sort_by_key calls of AlphaShape/Alpha_Shapes_2D.cu:317,400,658).
#include <cstring>
#include <rocprim/rocprim.hpp>
#include "kernels.h"

int sort_pairs_u64(unsigned long long *keys_in, int *vals_in, unsigned long long *keys_out, int *vals_out,
                   int n, int begin_bit, int end_bit, void *temp, size_t *temp_bytes, hipStream_t st)
{
    hipError_t e = rocprim::radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out,
                                             (unsigned int)n, (unsigned)begin_bit, (unsigned)end_bit, st);
    return (int)e;
}

int sort_pairs_u32(unsigned int *keys_in, int *vals_in, unsigned int *keys_out, int *vals_out,
                   int n, int begin_bit, int end_bit, void *temp, size_t *temp_bytes, hipStream_t st)
{
    hipError_t e = rocprim::radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out,
                                             (unsigned int)n, (unsigned)begin_bit, (unsigned)end_bit, st);
    return (int)e;
}
