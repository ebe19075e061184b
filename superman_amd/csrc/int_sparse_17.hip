// int_sparse_17.hip -- orders 17..25 of the sparse integer kernel (int_sparse.h).
#include "int_sparse.h"

namespace spm {

const void* int_sparse_ptr_17(int n) { return sparse::sparse_range_ptr<17, 25>(n); }

}  // namespace spm
