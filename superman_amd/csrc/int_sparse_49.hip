// int_sparse_49.hip -- orders 49..56 of the sparse integer kernel (int_sparse.h).
#include "int_sparse.h"

namespace spm {

const void* int_sparse_ptr_49(int n) { return sparse::sparse_range_ptr<49, 56>(n); }

}  // namespace spm
