// kernels.hpp -- interface between the host planner (engine.cpp) and the
// gfx950 kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sqmi {

// Sentinel for "minus infinity" inside the packed int16 kernels.  Every real
// value of a pair routed to them is proven (on the host) to stay within
// [kNeg16 + kGuard16, 32767 - kGuard16].
constexpr int kNeg16 = -28672;
constexpr int kGuard16 = 2048;

// Tracking kinds (which cells define the result), from the mode bits.
enum Track : uint32_t {
    kTrackEnd = 1,      // global: cell (m, n)
    kTrackBottom = 2,   // last row free: max over row m, j <= n
    kTrackRight = 4,    // last column free: max over column n, i <= m
    kTrackLocal = 8     // local: max over all cells
};

// A work item: up to `pairs per wave` pairs that share the same number of
// row blocks; processed together by one wavefront.
struct WorkItem {
    int32_t first;      // index of the first pair in `order`
    int32_t count;      // number of pairs in the item
    int32_t maxN;       // max |H| over the item's pairs
    int32_t nblocks;    // row blocks of G*R rows (same for all pairs)
};

struct FwdArgs {
    const uint8_t *hRes;
    const uint8_t *vRes;
    const int64_t *hOff;
    const int64_t *vOff;
    const int32_t *order;       // pair ids grouped by work item
    const WorkItem *items;
    int32_t nItems;
    uint32_t *queue;            // work-queue counter (zeroed before launch)
    uint32_t *rowbuf;           // per wave slot: GROUPS * rowCols * 2 dwords
    int32_t rowCols;            // columns per group row buffer
    int32_t *scores;            // indexed by pair id
    int32_t gapOpen;
    int32_t gapExtend;
    uint32_t mode;              // SQMI_* mode bits
    int32_t alpha;              // alphabet size; residue code `alpha` is padding
    const int16_t *tTable;      // (alpha+1)^2 entries: s(h,v) - gapOpen, pad row/col
    const uint32_t *dnaProfile; // 5 dwords: bytes v=0..3 of s(h,v)-gapOpen, h=0..3, pad
};

// Simple int32 kernel: one thread per pair, H/E column in global scratch.
struct Simple32Args {
    const uint8_t *hRes;
    const uint8_t *vRes;
    const int64_t *hOff;
    const int64_t *vOff;
    const int32_t *ids;         // pair ids
    const int64_t *scratchOff;  // per id slot: offset (in int32 pairs) of its column
    int32_t nIds;
    int32_t *scratch;           // 2 * (m+1) int32 per pair
    int32_t *scores;
    int32_t gapOpen;
    int32_t gapExtend;
    uint32_t mode;
    int32_t alpha;
    const int32_t *table;       // alpha^2 substitution scores
};

// Kernel variant of the packed int16 family.
struct Variant16 {
    int G;      // lanes per pair group
    int R;      // rows per lane
    bool dna;   // v_perm profile path (alphabet <= 4, not local)
    uint32_t track;
};

// Returns the number of work-item slots (waves) the launch should use and the
// device function; launches onto `stream`.
hipError_t launchFwd16(const Variant16 &v, const FwdArgs &args, int gridBlocks, hipStream_t stream,
                       int threadsPerBlock = 256);
int wavesPerBlock16();
// Maximum resident blocks per CU for the variant (occupancy query).
int occupancy16(const Variant16 &v);
bool variantSupported(const Variant16 &v);
hipError_t launchSimple32(const Simple32Args &args, hipStream_t stream);
// scores[ids[k]] = value for k < n.
hipError_t launchFillScores(int32_t *scores, const int32_t *ids, int n, int32_t value, hipStream_t stream);

}  // namespace sqmi
