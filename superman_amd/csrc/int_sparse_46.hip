// int_sparse_46.hip -- orders 46..48 of the sparse integer kernel (int_sparse.h).
#include "int_sparse.h"

namespace spm {

const void* int_sparse_ptr_46(int n) { return sparse::sparse_range_ptr<46, 48>(n); }

}  // namespace spm
