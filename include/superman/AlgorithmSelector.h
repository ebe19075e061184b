// AlgorithmSelector.h -- CommonFiles/AlgorithmSelector.h:38-141 over the
// drop-in GPU entry points: same function-pointer type, same `auto` rule
// (:52-67: dense xregister_mshared when nnz/row <= 6 or the biggest
// component is small, else register-efficient code generation) and the same
// sparse-vs-dense routing (:97-139).  The CPU mode and the recompile mailbox
// of the reference are not part of this library.
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
        if (settings->algorithm == AlgorithmEnds && settings->mode != Mode::CPU) {
            if ((matrix->nov * matrix->sparsity / 100 <= 6) || (biggestComponentSize <= 41) ||
                (biggestComponentSize <= (unsigned)(matrix->nov - 3)))
                settings->algorithm = XREGISTERMSHARED;
            else
                settings->algorithm = REGEFFICIENTCODEGENERATION;
        }
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
