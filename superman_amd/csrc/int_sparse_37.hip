// int_sparse_37.hip -- orders 37..41 of the sparse integer kernel (int_sparse.h).
#include "int_sparse.h"

namespace spm {

const void* int_sparse_ptr_37(int n) { return sparse::sparse_range_ptr<37, 41>(n); }

}  // namespace spm
