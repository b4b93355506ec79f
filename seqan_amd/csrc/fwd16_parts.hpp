// fwd16_parts.hpp -- the packed forward kernels are compiled in four units
// (fwd16.hip built with SQMI_FWD_DNA = 0 / 1 and SQMI_FWD_G = 16 / 64); each
// exports its kernels through one of these tables.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace sqmi {

struct Fwd16Part {
    // kFwd16 of the shape (G, R) and tracking of v, or nullptr when the
    // unit does not compile it.
    const void *(*kernel)(const Variant16 &v);
    hipError_t (*launch)(const Variant16 &v, const FwdArgs &args, int gridBlocks, hipStream_t stream,
                         int threadsPerBlock);
    // The trace kernel (G = 16, R = 16; nullptr in the G = 64 units).
    hipError_t (*launchTrace)(uint32_t track, const FwdArgs &args, int gridBlocks, hipStream_t stream);
    const void *(*traceKernel)();
    int (*traceSteps)(int maxN);
};

extern const Fwd16Part kFwdPartDna16, kFwdPartTab16, kFwdPartDna64, kFwdPartTab64;

}  // namespace sqmi
