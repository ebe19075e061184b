// SparseMatrix.h -- Structures/SparseMatrix.h:31-102: a Matrix that also
// holds CSC (cptrs, rows, cvals) and CSR (rptrs, cols, rvals) copies.  As in
// the reference, the constructor takes over the dense buffer of `matrix`
// and deletes the husk.
#ifndef SUPERMAN_AMD_SPARSEMATRIX_H
#define SUPERMAN_AMD_SPARSEMATRIX_H

#include "Matrix.h"

#include <utility>

template <class S>
class SparseMatrix : public Matrix<S>
{
public:
    SparseMatrix(Matrix<S>* matrix, int nnz) : Matrix<S>(std::move(*matrix)), nnz(nnz)
    {
        const int n = this->nov;
        cptrs = new int[n + 1];
        rptrs = new int[n + 1];
        rows = new int[nnz];
        cols = new int[nnz];
        cvals = new S[nnz];
        rvals = new S[nnz];
        int rn = 0, cn = 0;
        for (int i = 0; i < n; ++i) {
            rptrs[i] = rn;
            cptrs[i] = cn;
            for (int j = 0; j < n; ++j) {
                if (this->mat[i * n + j] != S(0)) {
                    cols[rn] = j;
                    rvals[rn++] = this->mat[i * n + j];
                }
                if (this->mat[j * n + i] != S(0)) {
                    rows[cn] = j;
                    cvals[cn++] = this->mat[j * n + i];
                }
            }
        }
        rptrs[n] = rn;
        cptrs[n] = cn;
        delete matrix;
    }

    SparseMatrix(const SparseMatrix&) = delete;
    SparseMatrix& operator=(const SparseMatrix&) = delete;

    ~SparseMatrix() override
    {
        delete[] cptrs;
        delete[] rptrs;
        delete[] rows;
        delete[] cols;
        delete[] cvals;
        delete[] rvals;
    }

public:
    int* cptrs;
    int* rows;
    S* cvals;
    int* rptrs;
    int* cols;
    S* rvals;
    int nnz;
};

#endif
