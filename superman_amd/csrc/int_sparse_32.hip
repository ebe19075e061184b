// int_sparse_32.hip -- orders 32..36 of the sparse integer kernel (int_sparse.h).
#include "int_sparse.h"

namespace spm {

const void* int_sparse_ptr_32(int n) { return sparse::sparse_range_ptr<32, 36>(n); }

}  // namespace spm
