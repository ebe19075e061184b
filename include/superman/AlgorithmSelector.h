// AlgorithmSelector.h -- CommonFiles/AlgorithmSelector.h:38-141 over the
// drop-in GPU entry points: same function-pointer type and the same
// sparse-vs-dense routing (:97-139).  The `auto` rule (:52-67: dense
// xregister_mshared when nnz/row <= 6 or the biggest component is small,
// else register-efficient code generation) weighs the reference's own
// GPUs; here it is retuned for MI355X from the library's cost model
// (spm_choose_path: per-subset rates measured on MI355X for every walk and
// generated kernel, plus the hiprtc compile of a kernel not yet cached):
// one generated kernel for the whole matrix when that is estimated
// cheapest, otherwise the decomposed driver, which makes the same estimate
// per leaf (exact integer walks, row-scaled integers, generated kernels,
// the dense FP64 walk).  The CPU mode and the recompile mailbox of the
// reference are not part of this library.
#ifndef SUPERMAN_AMD_ALGORITHMSELECTOR_H
#define SUPERMAN_AMD_ALGORITHMSELECTOR_H

#include "GPUKernelWrappers.h"

template <class C, class S>
class AlgorithmSelector
{
public:
    typedef Result (*Algorithm)(Matrix<S>* matrix, Settings* settings);

    static Algorithm selectAlgorithm(Matrix<S>* matrix, Settings* settings, unsigned biggestComponentSize)
    {
        if (settings->algorithm == AlgorithmEnds && settings->mode != Mode::CPU)
            settings->algorithm = superman_amd::auto_algorithm(matrix, biggestComponentSize);
        if (settings->mode == Mode::CPU)
            throw std::runtime_error("The CPU mode is not part of the GPU library.\n");
        const bool sparse =
            (settings->algorithm != XREGISTERMSHARED && settings->algorithm != XREGISTERMGLOBAL &&
             matrix->sparsity < 30) ||
            settings->algorithm == NAIVECODEGENERATION || settings->algorithm == REGEFFICIENTCODEGENERATION ||
            settings->algorithm == APPROXIMATION;
        if (settings->mode == Mode::SingleGPU) return sparse ? gpuSPSingleGPU<C, S> : gpuDPSingleGPU<C, S>;
        if (settings->mode == Mode::MultiGPU) return sparse ? gpuSPMultiGPU<C, S> : gpuDPMultiGPU<C, S>;
        throw std::runtime_error(
            "The mode you want to calculate the permanent in is unavailable for launch. Terminating...\n");
    }
};

#endif
