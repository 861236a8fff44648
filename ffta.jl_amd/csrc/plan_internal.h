// plan_internal.h -- library-internal entry points shared by plan.cpp and server.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include "../../include/fftahip.h"

namespace fftahip {
// Scratch bytes an execution of `plan` on an array of shape dims needs.
fftahip_status plan_scratch_bytes(fftahip_plan plan, int ndims, const int64_t* dims, size_t* bytes);
// mul!(y, p, x) on device arrays with caller-provided scratch (no allocation
// once the shape's schedule exists).
fftahip_status plan_execute_scratch(fftahip_plan plan, int ndims, const int64_t* dims, const void* d_in, void* d_out,
                                    void* scratch, hipStream_t s);
// mul! of a 1-D plan on `count` vectors of its length at arbitrary device
// addresses, in one set of work-list launches (the vectors as the lines of
// one batch).  plan_many_bytes gives the scratch and the descriptor memory
// (pass descriptors and offset tables, uploaded per call) the caller
// provides; the addresses of each side must be element aligned.
fftahip_status plan_many_bytes(fftahip_plan plan, int count, size_t* scratch, size_t* desc, hipStream_t s);
fftahip_status plan_execute_many(fftahip_plan plan, int count, const void* const* ins, void* const* outs,
                                 void* scratch, void* desc, hipStream_t s);
// Record a failure message for fftahip_last_error() and return s.
fftahip_status set_error(fftahip_status s, const std::string& msg);
}  // namespace fftahip
