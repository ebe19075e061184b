// Result.h -- job result, layout of CommonFiles/Result.h:29-37 kept so that
// callers of the reference's drivers compile unchanged: whole-job wall time
// in seconds and the permanent as a double.  Extended information (exact
// integers, double-double value, error estimate) is in spm_result
// (superman_amd.h), never in this struct.
#ifndef SUPERMAN_AMD_RESULT_H
#define SUPERMAN_AMD_RESULT_H

struct Result
{
    Result(double time = 0, double permanent = 0) : time(time), permanent(permanent) {}

    double time;
    double permanent;
};

#endif
