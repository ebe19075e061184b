// cli_main.cpp -- `superman_gpu key=value ...`: the reference's command line
// (main.cpp:42-151 with IO::readSettings, IO/IO.h:163-401, and
// IO::readMatrix / readComplex, IO/IO.h:403-596) over the drop-in headers
// and the C ABI, without its fork/recompile wrapper: generated kernels are
// compiled in-process.
//
// Flow as in main.cpp: read the matrix, prune it (DM), split it into
// connected components, pick the algorithm (AlgorithmSelector, `auto` by
// default), run it per component and multiply -- except for the code
// generation algorithms, which take the whole matrix -- then print
//   Permanent: <value> - Computed in: <seconds> seconds.
// Complex matrices (complex=true) go to gpuComputeComplex, which prints its
// own line.  Keys as the reference's: filename, algorithm, mode (single_gpu
// runs; cpu and the multi-GPU modes end with an error), calculation_precision
// (dd | kahan; a number selects the reference's CPU-only arbitrary
// precision, an error here), complex, binary, undirected, device_id,
// scaling, scaling_iteration_no and scale_into (every leaf balanced before
// its permanent, as DecomposePerman's scaling option); pid, repo_dir,
// thread_count and gpu_num are read and have no effect.
#include "superman/AlgorithmSelector.h"

#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

int int_arg(const std::string& key, const std::string& v)
{
    try {
        return std::stoi(v);
    } catch (const std::exception&) {
        throw std::runtime_error("An integer value should be provided to the " + key + " argument!\n");
    }
}

unsigned unsigned_arg(const std::string& key, const std::string& v)
{
    try {
        return (unsigned)std::stoul(v);
    } catch (const std::exception&) {
        throw std::runtime_error("An integer value should be provided to the " + key + " argument!\n");
    }
}

// key=value settings as IO::readSettings (IO/IO.h:163-401) reads them:
// unknown keys are reported and skipped, bad values end the run.  pid and
// repo_dir, which the reference needs for its recompile wrapper, are
// accepted and not required.
Settings parse(int argc, char** argv)
{
    Settings s;
    s.mode = SingleGPU;
    bool have_file = false;
    for (int i = 1; i < argc; ++i) {
        const std::string kv = argv[i];
        const size_t eq = kv.find('=');
        if (eq == std::string::npos) throw std::runtime_error("Arguments are key=value pairs: " + kv + "\n");
        const std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
        if (k == "filename") {
            s.filename = v;
            have_file = true;
        } else if (k == "algorithm") {
            if (v == "auto") s.algorithm = AlgorithmEnds;
            else if (v == "xregister_mshared") s.algorithm = XREGISTERMSHARED;
            else if (v == "xregister_mglobal") s.algorithm = XREGISTERMGLOBAL;
            else if (v == "xshared_mglobal") s.algorithm = XSHAREDMGLOBAL;
            else if (v == "xshared_mshared") s.algorithm = XSHAREDMSHARED;
            else if (v == "xglobal_mglobal") s.algorithm = XGLOBALMGLOBAL;
            else if (v == "xglobal_mshared") s.algorithm = XGLOBALMSHARED;
            else if (v == "naive_code_generation") s.algorithm = NAIVECODEGENERATION;
            else if (v == "register_efficient_code_generation") s.algorithm = REGEFFICIENTCODEGENERATION;
            else if (v == "approximation") s.algorithm = APPROXIMATION;
            else throw std::runtime_error("UNKNOWN ALGORITHM: " + v + '\n');
        } else if (k == "mode") {
            if (v == "cpu") {
                s.mode = CPU;
                s.algorithm = AlgorithmEnds;
            } else if (v == "single_gpu") s.mode = SingleGPU;
            else if (v == "multi_gpu") s.mode = MultiGPU;
            else if (v == "multi_gpu_mpi") s.mode = MultiGPUMPI;
            else throw std::runtime_error("UNKNOWN MODE: " + v + '\n');
        } else if (k == "calculation_precision") {
            if (v == "kahan") s.calculationPrecision = KAHAN;
            else if (v == "dd") s.calculationPrecision = DD;
            else {
                try {
                    s.arbitraryPrecision = (unsigned)std::stoul(v);
                    s.calculationPrecision = ARBITRARY;
                } catch (const std::exception&) {
                    throw std::runtime_error(
                        "Calculation precision can either be kahan or dd or an arbitrary precision given as an "
                        "unsigned integer that will launch SUperman with arbitrarily selected mantissa (check IEEE "
                        "754)!\n");
                }
            }
        } else if (k == "complex") s.complex = v == "true";
        else if (k == "binary") s.binary = v == "true";
        else if (k == "undirected") s.undirected = v == "true";
        else if (k == "scaling") s.scaling = v == "true";
        else if (k == "scaling_iteration_no") s.scalingIterationNo = unsigned_arg(k, v);
        else if (k == "scale_into") s.scaleInto = unsigned_arg(k, v);
        else if (k == "device_id") s.deviceID = int_arg(k, v);
        else if (k == "thread_count") s.threadC = int_arg(k, v);
        else if (k == "gpu_num") s.gpuNum = int_arg(k, v);
        else if (k == "pid") s.PID = int_arg(k, v);
        else if (k == "repo_dir") s.REPO_DIR = v;
        else std::cout << "UNKNOWN ARGUMENT: " << k << " - skipping it." << std::endl;
    }
    if (!have_file)
        throw std::runtime_error(
            "You are required to provide the absolute path of the matrix file that you want to calculate the "
            "matrix permanent for to the executable of this program!\n");
    return s;
}

// Coordinate MatrixMarket (1-based) or plain coordinate file (0-based), as
// IO::readMatrix / IO::readComplex (IO/IO.h:403-596) read it, with their
// three summary lines.  Real: binary=true (and, here, a `pattern` banner)
// reads every entry as 1, zeros are dropped, undirected=true mirrors
// off-diagonal entries; complex: every entry is (re, im), undirected=true
// mirrors.
template <class S>
int read_matrix(const Settings& s, std::vector<S>& a)
{
    constexpr bool cplx = std::is_same<S, std::complex<double>>::value;
    std::ifstream f(s.filename);
    if (!f.is_open()) throw std::runtime_error("File could not be opened.\n");
    const bool mtx = s.filename.size() > 4 && s.filename.substr(s.filename.size() - 4) == ".mtx";
    bool pattern = !cplx && s.binary;
    std::string line;
    while (f.peek() == '%') {
        std::getline(f, line);
        if (!cplx && line.find("pattern") != std::string::npos) pattern = true;
    }
    int n = 0, m = 0, lines = 0;
    f >> n >> m >> lines;
    if (n != m) throw std::runtime_error("Your matrix is not square, and therefore the permanent of it is trivially equal to 0.");
    a.assign((size_t)n * n, S(0));
    long long nnz = 0;
    for (int e = 0; e < lines; ++e) {
        int i, j;
        f >> i >> j;
        S v = S(1);
        if constexpr (cplx) {
            double re, im;
            f >> re >> im;
            v = S(re, im);
        } else if (!pattern) {
            f >> v;
        }
        if (!cplx && v == S(0)) continue;
        if (mtx) {
            --i;
            --j;
        }
        if (i < 0 || j < 0 || i >= n || j >= n) throw std::runtime_error("Entry outside the matrix.\n");
        a[(size_t)i * n + j] = v;
        ++nnz;
        if (s.undirected && (cplx || i != j)) {
            a[(size_t)j * n + i] = v;
            ++nnz;
        }
    }
    std::cout << "Number of rows/columns of matrix is: " << n << std::endl;
    std::cout << "Total number of nonzeros is: " << nnz << std::endl;
    std::cout << "Sparsity of the matrix is determined to be: " << double(nnz) / (double(n) * n) * 100 << std::endl;
    return n;
}

std::string fmt(double v)
{
    std::ostringstream o;
    o << std::setprecision(std::numeric_limits<double>::max_digits10) << v;
    return o.str();
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        Settings s = parse(argc, argv);
        std::cout << "MATRIX NAME: " << s.filename << std::endl;
        if (s.complex) {
            std::vector<std::complex<double>> a;
            const int n = read_matrix(s, a);
            Matrix<std::complex<double>> m(n);
            std::copy(a.begin(), a.end(), m.mat);
            gpuComputeComplex(&m, &s);
            return 0;
        }
        std::vector<double> a;
        const int n = read_matrix(s, a);
        const auto t0 = std::chrono::steady_clock::now();
        // DM pruning and connected components (IO.h:477-508, dm.h:43-107)
        spm_decomp* d = nullptr;
        superman_amd::check(spm_decompose(a.data(), n, SPM_DECOMP_DM | SPM_DECOMP_COMPONENTS, &d), "spm_decompose");
        int comps = 0, zero = 0;
        spm_decomp_stats(d, &comps, nullptr, nullptr, nullptr, &zero);
        unsigned biggest = 0;
        std::vector<Matrix<double>*> blocks;
        for (int k = 0; k < comps && !zero; ++k) {
            int bn = 0;
            spm_decomp_leaf(d, k, nullptr, &bn, nullptr);
            Matrix<double>* b = new Matrix<double>(bn);
            spm_decomp_leaf(d, k, nullptr, nullptr, b->mat);
            b->sparsity = 100.0 * getNNZ(b) / (double(bn) * bn);
            biggest = std::max(biggest, (unsigned)bn);
            blocks.push_back(b);
        }
        spm_decomp_destroy(d);
        Matrix<double> whole(n);
        std::copy(a.begin(), a.end(), whole.mat);
        whole.sparsity = 100.0 * getNNZ(&whole) / (double(n) * n);
        auto algorithm = AlgorithmSelector<double, double>::selectAlgorithm(&whole, &s, biggest);
        const bool per_component = s.algorithm != REGEFFICIENTCODEGENERATION && s.algorithm != NAIVECODEGENERATION;
        double perm = zero ? 0.0 : 1.0;
        if (!zero && per_component) {
            for (Matrix<double>* b : blocks) perm *= algorithm(b, &s).permanent;
        } else if (!zero) {
            spm_decomp* p = nullptr;
            superman_amd::check(spm_decompose(a.data(), n, SPM_DECOMP_DM, &p), "spm_decompose");
            Matrix<double> pruned(n);
            spm_decomp_leaf(p, 0, nullptr, nullptr, pruned.mat);
            spm_decomp_destroy(p);
            pruned.sparsity = 100.0 * getNNZ(&pruned) / (double(n) * n);
            perm = algorithm(&pruned, &s).permanent;
        }
        for (Matrix<double>* b : blocks) delete b;
        const double secs =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        // The leaf statistics line of main.cpp:122-126 (its counters are
        // filled by DecomposePerman; the same decomposition gives them here).
        {
            spm_decomp* st = nullptr;
            const int opts = per_component ? SPM_DECOMP_DEFAULT
                                           : (SPM_DECOMP_DM | SPM_DECOMP_CODEGEN | SPM_DECOMP_COMPRESS);
            if (spm_decompose(a.data(), n, opts, &st) == SPM_OK) {
                int leaves = 0;
                double avg_n = 0, avg_nnz = 0;
                spm_decomp_stats(st, nullptr, &leaves, &avg_n, &avg_nnz, nullptr);
                std::cout << "Submatrix count: " << leaves << " - Average N: " << avg_n
                          << " - Average NNZ: " << avg_nnz << std::endl;
                spm_decomp_destroy(st);
            }
        }
        if (std::isnan(perm))
            std::cout << "Permanent computation has overflowed. - Computed in: " << secs << " seconds." << std::endl;
        else
            std::cout << "Permanent: " << fmt(perm) << " - Computed in: " << secs << " seconds." << std::endl;
        return 0;
    } catch (const std::exception& e) {
        std::cerr << e.what();
        return 1;
    }
}
