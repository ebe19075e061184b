// Settings.h -- the enums and settings struct of CommonFiles/Settings.h:36-102,
// same names, order and fields, so that code filling a Settings for the
// reference fills this one.  The GPU drop-in reads `algorithm`, `mode`,
// `calculationPrecision` and `deviceID`; the rest is carried for
// compatibility (process-wide flags of the reference -- avgN, sparse, ... --
// are not used: every call here is reentrant).
#ifndef SUPERMAN_AMD_SETTINGS_H
#define SUPERMAN_AMD_SETTINGS_H

#include <string>

enum Algorithm
{
    XREGISTERMSHARED,
    XREGISTERMGLOBAL,
    XGLOBALMGLOBAL,
    XGLOBALMSHARED,
    XSHAREDMGLOBAL,
    XSHAREDMSHARED,
    NAIVECODEGENERATION,
    REGEFFICIENTCODEGENERATION,
    APPROXIMATION,
    AlgorithmEnds
};

enum Mode
{
    CPU,
    SingleGPU,
    MultiGPU,
    MultiGPUMPI,
    ModeEnds
};

enum Precision
{
    DD,      // plain FP64 ("dd" in the reference's CLI)
    KAHAN,   // compensated (the reference's default)
    ARBITRARY,
    PrecisionEnds
};

struct Settings
{
    // General
    std::string filename;
    int PID = 1;
    std::string REPO_DIR;
    std::string pipe;
    Algorithm algorithm = AlgorithmEnds;
    Mode mode = SingleGPU;

    // Matrix related
    bool complex = false;
    bool binary = false;
    bool undirected = false;

    // Scaling related
    bool scaling = false;
    unsigned scalingIterationNo = 100;
    unsigned scaleInto = 2;

    // Precision related
    unsigned arbitraryPrecision = 0;
    Precision calculationPrecision = KAHAN;

    // CPU specific
    int threadC = 1;

    // GPU specific
    int deviceID = 0;
    int gpuNum = 1;
    unsigned partition = 1;

    // MPI specific
    int rank = 0;
    int processorNum = 1;
};

#endif
