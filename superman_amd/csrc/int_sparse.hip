// int_sparse.hip -- orders 1..16 of the sparse integer kernel
// (int_sparse.h) and the lookup of every order's kernel.
#include "int_sparse.h"

namespace spm {

const void* int_sparse_kernel_ptr(int n)
{
    if (n < 1 || n > kMaxN) return nullptr;
    if (n <= 16) return sparse::sparse_range_ptr<1, 16>(n);
    if (n <= 25) return int_sparse_ptr_17(n);
    if (n <= 31) return int_sparse_ptr_26(n);
    if (n <= 36) return int_sparse_ptr_32(n);
    if (n <= 41) return int_sparse_ptr_37(n);
    if (n <= 45) return int_sparse_ptr_42(n);
    if (n <= 48) return int_sparse_ptr_46(n);
    if (n <= 56) return int_sparse_ptr_49(n);
    return int_sparse_ptr_57(n);
}

}  // namespace spm
