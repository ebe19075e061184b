// Permanent.h -- the job object of CommonFiles/Permanent.h:36-91: owns a
// Matrix, computePermanent() runs the driver's permanentFunction() and
// returns the final (scaled) permanent as a double.  The reference reorders
// the matrix here for its sparse kernels (UTOrder / colSort); in this build
// the orderings happen inside the library calls that need them.
#ifndef SUPERMAN_AMD_PERMANENT_H
#define SUPERMAN_AMD_PERMANENT_H

#include "Matrix.h"
#include "Settings.h"

template <class C, class S>
class Permanent
{
public:
    Permanent(Matrix<S>* matrix, Settings settings) : m_Matrix(matrix), m_Settings(settings) {}

    Permanent(const Permanent&) = delete;
    Permanent& operator=(const Permanent&) = delete;
    Permanent(Permanent&&) = delete;
    Permanent& operator=(Permanent&&) = delete;

    virtual ~Permanent() { delete m_Matrix; }

    double computePermanent() { return this->permanentFunction(); }
    virtual double permanentFunction() = 0;

public:
    Matrix<S>* m_Matrix;
    Settings m_Settings;
};

#endif
