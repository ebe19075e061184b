// int_sparse_42.hip -- orders 42..45 of the sparse integer kernel (int_sparse.h).
#include "int_sparse.h"

namespace spm {

const void* int_sparse_ptr_42(int n) { return sparse::sparse_range_ptr<42, 45>(n); }

}  // namespace spm
