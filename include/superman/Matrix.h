// Matrix.h -- dense row-major matrix of Structures/Matrix.h:32-84: `mat`
// owns nov*nov entries, `sparsity` is the percentage of nonzeros.
#ifndef SUPERMAN_AMD_MATRIX_H
#define SUPERMAN_AMD_MATRIX_H

#include <cstring>

template <class S>
class Matrix
{
public:
    Matrix() = default;

    explicit Matrix(int nov) : nov(nov)
    {
        mat = new S[(size_t)nov * nov];
        std::memset(mat, 0, sizeof(S) * (size_t)nov * nov);
    }

    Matrix(const Matrix& other) : nov(other.nov), sparsity(other.sparsity)
    {
        mat = new S[(size_t)nov * nov];
        std::memcpy(mat, other.mat, sizeof(S) * (size_t)nov * nov);
    }

    Matrix(Matrix&& other) : mat(other.mat), nov(other.nov), sparsity(other.sparsity) { other.mat = nullptr; }

    Matrix& operator=(const Matrix&) = delete;
    Matrix& operator=(Matrix&&) = delete;

    virtual ~Matrix() { delete[] mat; }

public:
    S* mat = nullptr;
    int nov = 0;
    double sparsity = 100;
};

// Number of nonzeros (CommonFiles/Helpers.h:178-194).
template <class S>
inline int getNNZ(const Matrix<S>* m)
{
    int nnz = 0;
    for (int e = 0; e < m->nov * m->nov; ++e) nnz += m->mat[e] != S(0);
    return nnz;
}

#endif
