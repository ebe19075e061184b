// GPUKernelWrappers.h -- drop-in replacements for the reference's single-GPU
// entry points, calling libsuperman_amd.so through its C ABI:
//
//   gpuDPSingleGPU<C,S>   Algorithms/GPU/GPUKernelWrappers.cu:208-268
//   gpuSPSingleGPU<C,S>   Algorithms/GPU/GPUKernelWrappers.cu:65-113
//   gpuComputeComplex     Algorithms/Complex/gpuComputeComplex.cu:199-418
//
// and the drivers behind them as Permanent<C,S> subclasses (dpSingleGPU.cuh,
// spSingleGPU.cuh, KernelGenSingleGPU.cuh).  Same signatures and Result
// semantics (time = whole-call wall seconds, permanent = double); a library
// status other than SPM_OK becomes std::runtime_error, as the reference's
// configuration errors are (its CUDA errors exit() instead).  As in the
// reference, every algorithm runs behind DecomposePerman: blocks larger
// than 30 are compressed (degree-1/2) and split (degree-3/4) first
// (spm_permanent_decomposed / spm_decompose); the DM pruning and the
// component split stay with the caller (IO::readMatrix, main.cpp), as
// there.
#ifndef SUPERMAN_AMD_GPUKERNELWRAPPERS_H
#define SUPERMAN_AMD_GPUKERNELWRAPPERS_H

#include "../superman_amd.h"
#include "Matrix.h"
#include "Permanent.h"
#include "Result.h"
#include "Settings.h"

#include <chrono>
#include <cmath>
#include <complex>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace superman_amd {

inline void check(int status, const char* what)
{
    if (status != SPM_OK) throw std::runtime_error(std::string(what) + ": " + spm_strerror(status) + "\n");
}

inline double wall_seconds()
{
    using namespace std::chrono;
    return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

// Integer-valued within the exact path's range?
template <class S>
inline bool integral(const Matrix<S>* m, std::vector<int64_t>& out)
{
    const int nn = m->nov * m->nov;
    out.resize(nn);
    for (int e = 0; e < nn; ++e) {
        const double v = (double)m->mat[e];
        if (!(std::fabs(v) <= 1073741824.0) || v != std::floor(v)) return false;
        out[e] = (int64_t)v;
    }
    return true;
}

template <class S>
inline std::vector<double> as_double(const Matrix<S>* m)
{
    std::vector<double> a((size_t)m->nov * m->nov);
    for (size_t e = 0; e < a.size(); ++e) a[e] = (double)m->mat[e];
    return a;
}

// The MI355X `auto` choice (AlgorithmSelector): REGEFFICIENTCODEGENERATION
// when the whole matrix fits one kernel (order <= 63) and spm_choose_path
// estimates a generated kernel cheapest; else XREGISTERMSHARED, whose
// decomposed driver chooses per leaf.  biggestComponentSize is kept for the
// reference's signature; the per-leaf choice makes it unnecessary.
template <class S>
inline Algorithm auto_algorithm(const Matrix<S>* m, unsigned biggestComponentSize)
{
    (void)biggestComponentSize;
    if (m->nov >= 1 && m->nov <= 63) {
        const std::vector<double> a = as_double(m);
        int path = 0;
        if (spm_choose_path(a.data(), m->nov, &path) == SPM_OK &&
            (path == SPM_PATH_GEN_INT || path == SPM_PATH_GEN_REAL))
            return REGEFFICIENTCODEGENERATION;
    }
    return XREGISTERMSHARED;
}

inline int precision_of(const Settings& s)
{
    if (s.calculationPrecision == DD) return SPM_PREC_PLAIN;
    if (s.calculationPrecision == KAHAN) return SPM_PREC_COMPENSATED;
    throw std::runtime_error("Arbitrary precision is only available on the CPU.\n");
}

}  // namespace superman_amd

// DecomposePerman<..., dpSingleGPU / spSingleGPU> (DecomposePerman.h:76-117
// around dpSingleGPU.cuh:44-219 / spSingleGPU.cuh:44-234): compressions
// above order 30, then every leaf on the GPU by the automatic path choice
// (spm_permanent_auto): exact for integer and row-scaled integer leaves,
// else FP64 with the requested accumulation in a generated kernel or the
// dense walk, whichever is estimated faster.
template <class C, class S>
class DecomposedGPU : public Permanent<C, S>
{
public:
    DecomposedGPU(Matrix<S>* matrix, Settings settings) : Permanent<C, S>(matrix, settings) {}
    double permanentFunction() override
    {
        const std::vector<double> a = superman_amd::as_double(this->m_Matrix);
        const Settings& s = this->m_Settings;
        if (s.scaling)  // every leaf balanced first (DecomposePerman.h:97-107, 189-193)
            superman_amd::check(spm_permanent_decomposed_scaled(a.data(), this->m_Matrix->nov, SPM_DECOMP_COMPRESS,
                                                                superman_amd::precision_of(s), (int)s.scalingIterationNo,
                                                                (double)s.scaleInto, s.deviceID, nullptr, &last),
                                "spm_permanent_decomposed_scaled");
        else
            superman_amd::check(spm_permanent_decomposed(a.data(), this->m_Matrix->nov, SPM_DECOMP_COMPRESS,
                                                         superman_amd::precision_of(s), s.deviceID, nullptr, &last),
                                "spm_permanent_decomposed");
        return last.re;
    }
    spm_result last{};
};

template <class C, class S>
using dpSingleGPU = DecomposedGPU<C, S>;
template <class C, class S>
using spSingleGPU = DecomposedGPU<C, S>;

// Generated per-pattern kernels (register-efficient = UT-ordered, or naive).
template <class C, class S>
class KernelGenSingleGPU : public Permanent<C, S>
{
public:
    KernelGenSingleGPU(Matrix<S>* matrix, Settings settings) : Permanent<C, S>(matrix, settings) {}
    double permanentFunction() override
    {
        // The reference compresses degree-1/2 rows and columns above order 30
        // and generates one kernel for what remains (DecomposePerman.h:124-140).
        const std::vector<double> a = superman_amd::as_double(this->m_Matrix);
        spm_decomp* d = nullptr;
        superman_amd::check(spm_decompose(a.data(), this->m_Matrix->nov, SPM_DECOMP_CODEGEN | SPM_DECOMP_COMPRESS, &d),
                            "spm_decompose");
        int leaves = 0, zero = 0, n = 0;
        spm_decomp_stats(d, nullptr, &leaves, nullptr, nullptr, &zero);
        if (zero || leaves == 0) {
            spm_decomp_destroy(d);
            last = spm_result{};
            return 0.0;
        }
        spm_decomp_leaf(d, 0, nullptr, &n, nullptr);
        Matrix<S>* leaf = new Matrix<S>(n);
        std::vector<double> b((size_t)n * n);
        spm_decomp_leaf(d, 0, nullptr, nullptr, b.data());
        spm_decomp_destroy(d);
        std::vector<double> rs, cs;
        if (this->m_Settings.scaling) {  // the leaf balanced first, as DecomposePerman::addQueue does
            rs.resize(n);
            cs.resize(n);
            superman_amd::check(spm_sinkhorn_scale(b.data(), n, (int)this->m_Settings.scalingIterationNo,
                                                   (double)this->m_Settings.scaleInto, rs.data(), cs.data()),
                                "spm_sinkhorn_scale");
        }
        for (size_t e = 0; e < b.size(); ++e) leaf->mat[e] = (S)b[e];
        const double v = leaf_permanent(leaf, b, n);
        for (int i = 0; i < (int)rs.size(); ++i) {  // DecomposePerman.h:97-107
            last.re /= rs[i];
            last.re /= cs[i];
        }
        return rs.empty() ? v : last.re;
    }
    spm_result last{};

private:
    double leaf_permanent(Matrix<S>* leaf, std::vector<double>& b, int n)
    {
        std::vector<int64_t> ai;
        const int opt = this->m_Settings.algorithm == REGEFFICIENTCODEGENERATION ? SPM_CG_REGISTER_EFFICIENT : 0;
        if (n >= 5 && superman_amd::integral(leaf, ai)) {
            const int st = spm_sparse_codegen(ai.data(), n, opt, this->m_Settings.deviceID, nullptr, &last);
            if (st != SPM_ERANGE) {
                delete leaf;
                superman_amd::check(st, "spm_sparse_codegen");
                return last.re;
            }
        } else if (n >= 5) {  // real values: the generated FP64 kernel (exact hex-float literals)
            const int st = spm_sparse_codegen_real(b.data(), n, opt, superman_amd::precision_of(this->m_Settings),
                                                   this->m_Settings.deviceID, nullptr, &last);
            delete leaf;
            superman_amd::check(st, "spm_sparse_codegen_real");
            return last.re;
        }
        Settings plain = this->m_Settings;
        plain.scaling = false;  // a scaled leaf is not scaled again
        DecomposedGPU<C, S> generic(leaf, plain);
        const double v = generic.computePermanent();
        last = generic.last;
        return v;
    }
};

template <class Job, class C, class S>
inline Result run_job(Matrix<S>* matrix, Settings* settings)
{
    const double t0 = superman_amd::wall_seconds();
    Job job(new Matrix<S>(*matrix), *settings);
    const double v = job.computePermanent();
    return Result(superman_amd::wall_seconds() - t0, v);
}

template <class C, class S>
Result gpuDPSingleGPU(Matrix<S>* matrix, Settings* settings)
{
    switch (settings->algorithm) {
    case XREGISTERMSHARED:
    case XREGISTERMGLOBAL:
    case XSHAREDMGLOBAL:
    case XSHAREDMSHARED:
        return run_job<dpSingleGPU<C, S>, C, S>(matrix, settings);
    default:
        throw std::runtime_error("Algorithm you have selected is not included in the available GPU algorithms list.\n");
    }
}

template <class C, class S>
Result gpuSPSingleGPU(Matrix<S>* matrix, Settings* settings)
{
    switch (settings->algorithm) {
    case XSHAREDMGLOBAL:
    case XSHAREDMSHARED:
    case XGLOBALMGLOBAL:
    case XGLOBALMSHARED:
        return run_job<spSingleGPU<C, S>, C, S>(matrix, settings);
    case NAIVECODEGENERATION:
    case REGEFFICIENTCODEGENERATION:
        return run_job<KernelGenSingleGPU<C, S>, C, S>(matrix, settings);
    case APPROXIMATION:
        throw std::runtime_error("The approximation algorithm is not part of the exact-permanent library.\n");
    default:
        throw std::runtime_error("Algorithm you have selected is not included in the available GPU algorithms list.\n");
    }
}

template <class C, class S>
Result gpuDPMultiGPU(Matrix<S>*, Settings*)
{
    throw std::runtime_error("Multi-GPU modes are not part of this single-GPU build.\n");
}

template <class C, class S>
Result gpuSPMultiGPU(Matrix<S>*, Settings*)
{
    throw std::runtime_error("Multi-GPU modes are not part of this single-GPU build.\n");
}

// Complex: the reference prints the result and returns nothing
// (gpuComputeComplex.cu:406-417); gpuComputeComplexResult returns it.
inline spm_result gpuComputeComplexResult(Matrix<std::complex<double>>* matrix, Settings* settings)
{
    spm_result r{};
    const int prec = settings->calculationPrecision == DD ? SPM_PREC_PLAIN : SPM_PREC_COMPENSATED;
    superman_amd::check(spm_dense_c128(reinterpret_cast<const double*>(matrix->mat), matrix->nov, prec,
                                       settings->deviceID, nullptr, &r),
                        "spm_dense_c128");
    return r;
}

inline void gpuComputeComplex(Matrix<std::complex<double>>* matrix, Settings* settings)
{
    const spm_result r = gpuComputeComplexResult(matrix, settings);
    std::stringstream stream;
    if (std::isnan(r.re) || std::isnan(r.im)) {
        stream << "Permanent computation has overflowed." << " - Computed in: " << r.seconds << " seconds." << std::endl;
    } else {
        stream << "Permanent: (" << std::setprecision(std::numeric_limits<double>::max_digits10) << r.re << ','
               << r.im << ") - Computed in: " << r.seconds << " seconds." << std::endl;
    }
    if (settings->rank == 0) std::cout << stream.str() << std::flush;
}

#endif
