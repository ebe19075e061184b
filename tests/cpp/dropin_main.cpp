// Drives the drop-in C++ headers the way the reference's main.cpp drives its
// own (AlgorithmSelector -> Result (*)(Matrix*, Settings*)).  Input on stdin:
//   <mode: sel|run|cplx> <algorithm|-1 for auto> <precision> <n> <entries...>
// (cplx: 2*n*n entries, interleaved re/im).  Output: one line per result.
#include "superman/AlgorithmSelector.h"

#include <cstdio>
#include <iostream>
#include <string>

int main()
{
    std::string mode;
    int alg, prec, n;
    std::cin >> mode >> alg >> prec >> n;
    Settings s;
    s.algorithm = alg < 0 ? AlgorithmEnds : (Algorithm)alg;
    s.calculationPrecision = (Precision)prec;
    try {
        if (mode == "cplx") {
            auto* m = new Matrix<std::complex<double>>(n);
            for (int e = 0; e < n * n; ++e) {
                double re, im;
                std::cin >> re >> im;
                m->mat[e] = {re, im};
            }
            spm_result r = gpuComputeComplexResult(m, &s);
            std::printf("%.17g %.17g\n", r.re, r.im);
            delete m;
            return 0;
        }
        auto* m = new Matrix<double>(n);
        for (int e = 0; e < n * n; ++e) std::cin >> m->mat[e];
        m->sparsity = 100.0 * getNNZ(m) / (double(n) * n);
        auto fn = AlgorithmSelector<double, double>::selectAlgorithm(m, &s, (unsigned)n);
        std::printf("algorithm %d sparse %d\n", (int)s.algorithm,
                    fn == (AlgorithmSelector<double, double>::Algorithm)gpuSPSingleGPU<double, double> ? 1 : 0);
        if (mode == "run") {
            Result r = fn(m, &s);
            std::printf("%.17g %.9g\n", r.permanent, r.time);
        }
        delete m;
    } catch (const std::exception& e) {
        std::printf("error %s", e.what());
        return 3;
    }
    return 0;
}
