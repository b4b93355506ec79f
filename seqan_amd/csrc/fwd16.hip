// fwd16.hip -- one instantiation unit of the packed forward kernels
// (fwd16.hpp): the shapes of group size SQMI_FWD_G with the DNA profile
// (SQMI_FWD_DNA = 1) or the substitution table (0); the G = 16 units also
// hold the trace kernels.  Built four times, exporting SQMI_FWD_PART.
#include "fwd16.hpp"
#include "fwd16_parts.hpp"

#if !defined(SQMI_FWD_DNA) || !defined(SQMI_FWD_G) || !defined(SQMI_FWD_PART)
#error "fwd16.hip is built with SQMI_FWD_DNA, SQMI_FWD_G and SQMI_FWD_PART defined"
#endif

namespace sqmi {
namespace {

constexpr bool kDna = SQMI_FWD_DNA != 0;
constexpr int kG = SQMI_FWD_G;

const void *kernelOf(const Variant16 &v)
{
    if (v.dna != kDna)
        return nullptr;
#define X(g, r)                                              \
    if constexpr (g == kG) {                                 \
        if (v.G == g && v.R == r)                            \
            return kernelTrk<g, r, kDna>(v.track);           \
    }
    SQMI_FOR_EACH_GR(X)
#undef X
    return nullptr;
}

hipError_t launchOf(const Variant16 &v, const FwdArgs &args, int gridBlocks, hipStream_t stream,
                    int threadsPerBlock)
{
    if (v.dna != kDna)
        return hipErrorInvalidValue;
#define X(g, r)                                                                          \
    if constexpr (g == kG) {                                                             \
        if (v.G == g && v.R == r)                                                        \
            return launchTrk<g, r, kDna>(v, args, gridBlocks, stream, threadsPerBlock);  \
    }
    SQMI_FOR_EACH_GR(X)
#undef X
    return hipErrorInvalidValue;
}

#if SQMI_FWD_G == 16
hipError_t launchTraceOf(uint32_t track, const FwdArgs &args, int gridBlocks, hipStream_t stream)
{
    const dim3 grid(gridBlocks), block(256);
    switch (track) {
        case kTrackEnd: hipLaunchKernelGGL((kFwdTrace16<kDna, kTrackEnd>), grid, block, 0, stream, args); break;
        case kTrackBottom: hipLaunchKernelGGL((kFwdTrace16<kDna, kTrackBottom>), grid, block, 0, stream, args); break;
        case kTrackRight: hipLaunchKernelGGL((kFwdTrace16<kDna, kTrackRight>), grid, block, 0, stream, args); break;
        case kTrackBottom | kTrackRight:
            hipLaunchKernelGGL((kFwdTrace16<kDna, kTrackBottom | kTrackRight>), grid, block, 0, stream, args);
            break;
        case kTrackLocal: hipLaunchKernelGGL((kFwdTrace16<kDna, kTrackLocal>), grid, block, 0, stream, args); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

const void *traceKernelOf() { return (const void *)&kFwdTrace16<kDna, kTrackEnd>; }

int traceStepsOf(int maxN) { return Fwd16<16, 16, kDna, kTrackEnd, true>::stepsFor(maxN); }
#endif

}  // namespace

// The table holds host function pointers: host pass only (the device pass
// instantiates the kernels through the functions above).
#ifndef __HIP_DEVICE_COMPILE__
#if SQMI_FWD_G == 16
extern const Fwd16Part SQMI_FWD_PART = {kernelOf, launchOf, launchTraceOf, traceKernelOf, traceStepsOf};
#else
extern const Fwd16Part SQMI_FWD_PART = {kernelOf, launchOf, nullptr, nullptr, nullptr};
#endif
#endif

}  // namespace sqmi
