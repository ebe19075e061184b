// int_sparse_26.hip -- orders 26..31 of the sparse integer kernel (int_sparse.h).
#include "int_sparse.h"

namespace spm {

const void* int_sparse_ptr_26(int n) { return sparse::sparse_range_ptr<26, 31>(n); }

}  // namespace spm
