// int_sparse_57.hip -- orders 57..63 of the sparse integer kernel (int_sparse.h).
#include "int_sparse.h"

namespace spm {

const void* int_sparse_ptr_57(int n) { return sparse::sparse_range_ptr<57, 63>(n); }

}  // namespace spm
