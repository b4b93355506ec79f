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
// Substitution tables of the packed forward: at most this many codes
// (alphabet + padding code) per side, held in LDS replicated per bank.
constexpr int kTabMaxCodes = 28;
constexpr int kGuard16 = 2048;

// Tracking kinds (which cells define the result), from the mode bits.
enum Track : uint32_t {
    kTrackEnd = 1,      // global: cell (m, n)
    kTrackBottom = 2,   // last row free: max over row m, j <= n
    kTrackRight = 4,    // last column free: max over column n, i <= m
    kTrackLocal = 8,    // local: max over all cells
    // Kernel variant flag beside kTrackLocal (table kernels): every table
    // value t = s - gapOpen is >= 0, so the two lookups and the diagonal
    // add fuse into one 32-bit add (FwdArgs::tNonNeg).
    kLocalAdd3 = 16
};

// A work item: up to `pairs per wave` pairs that share the same number of
// row blocks; processed together by one wavefront.
struct WorkItem {
    int32_t first;      // index of the first pair in `order`
    int32_t count;      // bits 0..15: number of pairs in the item; bits 16..17
                        // (packed kernels with G = 16 / 64): quarters of a block
                        // held by block 0 (0: a whole block); banded kernels:
                        // bits 16..31 the pairs' common |V|
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
    uint32_t *queue;            // work-queue counter of the per-wave items (zeroed before launch)
    int32_t nLong;              // items [0, nLong) are shared by the 4 waves of a workgroup
    uint32_t *queueLong;        // work-queue counter of those (zeroed before launch)
    uint32_t *rowbuf;           // per wave slot: GROUPS * rowCols * 2 dwords
    int32_t rowCols;            // columns per group row buffer
    int32_t *scores;            // indexed by pair id
    int32_t gapOpen;
    int32_t gapExtend;
    uint32_t mode;              // SQMI_* mode bits
    int32_t alpha;              // alphabet size; residue code `alpha` is padding
    const int16_t *tTable;      // (alpha+1)^2 entries: s(h,v) - gapOpen, pad row/col
    const uint32_t *dnaProfile; // dwords 0..4: bytes v=0..3 of s(h,v)-gapOpen, h=0..3, pad;
                                // dwords 8..12: bytes h=0..3 of s(h,v)-gapOpen, v=0..3, pad
    int32_t tNonNeg;            // local table kernels: every s(h,v) - gapOpen >= 0
    int32_t bandLo, bandUp;     // banded kernels: cells with bandLo <= j - i <= bandUp
    int32_t bandCols;           // banded kernels: row-buffer ring columns (a power of two <= rowCols)
    // Trace variant only:
    uint4 *trace;               // 16-byte trace records (64 per step)
    const int64_t *traceOff;    // per item: first record (units of 64 records)
    int32_t *endEF;             // per pair id: E and F of the end cell
    int32_t *endIJ;             // free ends: per pair id, the end cell (i, j)
};

// Where the trace records of one pair live: rows are bottom-aligned in the
// item's blocks of 256 rows (row i is item row p + i), lane group lane0..+15,
// pair A or B (`half`) of the record's byte layout.
struct TraceLoc {
    int64_t rec0;               // first record of the item (units of 64 records)
    int32_t steps;              // column steps per block
    int32_t p;                  // item row of the pair's row 0
    int32_t lane0;              // first lane of the pair's group
    int32_t half;               // 0: low int16 half, 1: high
};

// Traceback walk: one thread per pair.
struct WalkArgs {
    const uint8_t *hRes;
    const uint8_t *vRes;
    const int64_t *hOff;
    const int64_t *vOff;
    const int32_t *ids;         // pair ids, one per thread
    const TraceLoc *locs;       // per thread
    int32_t nPairs;
    const uint4 *trace;
    const int32_t *scores;      // per pair id: H(m, n)
    const int32_t *endEF;       // per pair id: E(m, n), F(m, n)
    int32_t *counts;            // per pair id: alpha^2 + 2 counts
    int32_t alpha;
    int32_t linear;             // gap open == gap extend: the linear-gap walk
    const int32_t *endIJ;       // free ends, local: per pair id, the end cell (i, j); null: (m, n)
    int32_t local;              // local walk: stops where H falls to 0, no leading / trailing gaps
    const int32_t *table;       // local: alpha^2 substitution scores (H along the path)
    int32_t gapOpen;
    int32_t gapExtend;
    int64_t *begins;            // local: per pair id, the window's first H and V position (optional)
    // Rows (optional): pair id p's ops end at ops[opOff[p + 1]] (written
    // backwards), their number goes to opLen[p].
    uint8_t *ops;
    const int64_t *opOff;
    int64_t *opLen;
};

// int32 traceback, one thread per pair, trace bytes in global scratch.
struct Counts32Args {
    const uint8_t *hRes;
    const uint8_t *vRes;
    const int64_t *hOff;
    const int64_t *vOff;
    const int32_t *ids;
    const int64_t *scratchOff;  // per thread: byte offset of its scratch
    int32_t nIds;
    uint8_t *scratch;           // 8 (m+1) bytes of H/E columns + (n+1)(m+1) trace bytes per pair
    int32_t *scores;
    int32_t *counts;
    int32_t gapOpen;
    int32_t gapExtend;
    uint32_t mode;              // free end gaps (SQMI_FREE_*) or SQMI_LOCAL
    int32_t alpha;
    const int32_t *table;
    uint8_t *ops;               // rows, as in WalkArgs (optional)
    const int64_t *opOff;
    int64_t *opLen;
    int64_t *begins;            // local: per pair the window's first H and V position (optional)
    int32_t banded;             // cells outside lower <= j - i <= upper are -inf
    int64_t lower;
    int64_t upper;
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

// One pair couple per lane (lane.hip): Variant16 with G = 1 and R in {16,
// 32}.  Work items hold up to 128 pairs with the same number of R-row
// blocks; the row buffer holds, per wave slot, laneRowCols(maxN, R) columns
// x 64 lanes of (G, F) pairs (8 bytes each).
bool laneSupported(const Variant16 &v);
int laneRowCols(int maxN, int R);
// Packed banded global alignment (band16.hip): one pair couple per lane, R in
// {8, 12, 16} rows per block, bandUp - bandLo >= R; work items hold up to 128
// pairs of one length m (WorkItem.count = pairs | m << 16, nblocks =
// (m + R) / R).  Row buffer per wave slot: a ring of bandRingCols(...)
// columns x 64 lanes x 8 bytes.
int bandRowCols(int maxN);
int bandRingCols(int maxN, int64_t lower, int64_t upper, int R);
// LDS bytes of the rings when they fit in LDS (the kernel's LDS form), else 0.
int bandLdsRingBytes(int R, bool dna, int bandCols);
int occupancyBand16(int R, bool dna, int ringLdsBytes);
hipError_t launchBand16(int R, bool dna, const FwdArgs &args, int gridBlocks, int ringLdsBytes, hipStream_t stream);
int occupancyLane16(const Variant16 &v);
hipError_t launchLane16(const Variant16 &v, const FwdArgs &args, int gridBlocks, hipStream_t stream);

// Banded global scores (banded.hip).  Pairs in `ids` passed the host's band
// checks; their band, clipped to the matrix, is lower..upper.
struct BandedArgs {
    const uint8_t *hRes;
    const uint8_t *vRes;
    const int64_t *hOff;
    const int64_t *vOff;
    const int32_t *ids;
    int32_t nIds;
    int32_t *scores;            // indexed by pair id
    int32_t gapOpen;
    int32_t gapExtend;
    uint32_t mode;              // SQMI_FREE_* bits
    int32_t alpha;
    const int32_t *table;       // alpha^2 substitution scores
    int64_t lower, upper;       // the band: lower <= j - i <= upper
    int32_t *scratch;           // wide kernel: 4 * wideStride int32 per workgroup
    int32_t wideStride;
    // Banded traceback (kBandedTrace): per ids index k, rows 1..m of S trace
    // bytes at trace + traceOff[k] (byte of cell (i, j): row i - 1, lane
    // j - i - clipped lower); per pair id the end cell and its E and F.
    uint8_t *trace;
    const int64_t *traceOff;
    int32_t *endIJ;
    int32_t *endEF;
};
// segLanes: 1..64 (power of two >= every band width: kBandedNarrow), or 0
// (kBandedWide on wideGrid workgroups).
hipError_t launchBanded(const BandedArgs &args, int segLanes, int wideGrid, hipStream_t stream);
// Forward with trace bytes for bands of at most 64 diagonals (segLanes as in
// launchBanded), then one thread per pair walks (kBandedWalk).
hipError_t launchBandedTrace(const BandedArgs &args, int segLanes, hipStream_t stream);
struct BandedWalkArgs {
    const uint8_t *hRes;
    const uint8_t *vRes;
    const int64_t *hOff;
    const int64_t *vOff;
    const int32_t *ids;
    int32_t nIds;
    int32_t segLanes;
    int64_t lower;
    int64_t upper;
    const uint8_t *trace;
    const int64_t *traceOff;
    const int32_t *scores;
    const int32_t *endIJ;
    const int32_t *endEF;
    int32_t *counts;
    int32_t alpha;
    int32_t local;
    int32_t linear;
    uint8_t *ops;
    const int64_t *opOff;
    int64_t *opLen;
    int64_t *begins;
};
hipError_t launchBandedWalk(const BandedWalkArgs &args, hipStream_t stream);
// scores[ids[k]] = value for k < n.
hipError_t launchFillScores(int32_t *scores, const int32_t *ids, int n, int32_t value, hipStream_t stream);
// Rows written at the ends of their segments [off[i], off[i + 1]) of src,
// len[i] bytes each, copied back to back to dst + off[0]; start receives the
// n + 1 exclusive prefix sums of len (device memory throughout).
hipError_t launchPackRows(const uint8_t *src, const int64_t *off, const int64_t *len, int64_t *start, uint8_t *dst,
                          int64_t n, hipStream_t stream);
// p[0 .. n) = 0 by a kernel (no memset node, so the call can be captured in a
// HIP graph).
hipError_t launchZero32(uint32_t *p, size_t n, hipStream_t stream);

// The no-allocation entry point's device array (scratch.hip): its size and
// its device address on `device`, which must be the current device; the
// address is looked up once per device.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
hipError_t scratchAddress(int device, void **out);

// Long pairs (int32, tiled).  A pair with m rows and n columns is cut into
// row blocks of kTileRows (bottom-aligned like the packed kernels: item row
// p + i holds pair row i, p = nblocks * kTileRows - 1 - m) and column tiles of
// tileC columns (tile t holds columns t * tileC + 1 ... (t + 1) * tileC).
// Tile (b, t) needs the bottom row of tile (b - 1, t) and the right column of
// tile (b, t - 1): the forward runs one launch per tile anti-diagonal b + t,
// and keeps every tile border as a checkpoint for the traceback.
constexpr int kTileRows = 256;
struct LongPair {
    int32_t pid;
    int32_t n, m;
    int32_t nblocks, ntiles;
    int32_t p;
    int64_t rowck;      // int32 offset of the row checkpoints: nblocks x (n + 1) x (G, F)
    int64_t colck;      // int32 offset of the column checkpoints: nblocks x ntiles x kTileRows x (G, E)
    int64_t bestck;     // local scores: offset of the pair's nblocks maxima (one per row block)
};
struct TileArgs {
    const uint8_t *hRes;
    const uint8_t *vRes;
    const int64_t *hOff;
    const int64_t *vOff;
    const LongPair *pairs;
    int32_t nPairs;
    int32_t wavesPerPair;       // forward: waves per pair in the grid (4 blocks per wave)
    int32_t diag;               // forward: tile anti-diagonal b + t
    int32_t tileC;
    int32_t *ck;                // checkpoint storage
    int32_t gapOpen;
    int32_t gapExtend;
    const uint32_t *dnaProfile;
    int32_t *scores;            // per pair id
    int32_t *endEF;             // per pair id: E, F of cell (m, n)
    uint32_t mode;              // SQMI_FREE_* / SQMI_LOCAL (int32 forward; the packed one is global only)
    int32_t *best;              // local: per row block of each pair (LongPair::bestck), the maximum G so far
    // traceback
    uint2 *trace;               // per long pair: kRoundSlots tiles x kTileRecSteps steps x 16 lanes records
    int32_t *walk;              // per long pair: walk state (kWalkStateInts ints)
    int32_t *counts;            // per pair id: alpha^2 + 2 counts
    int32_t alpha;
    uint8_t *ops;               // rows, as in WalkArgs (optional)
    const int64_t *opOff;
    int64_t *opLen;
    // Streaming packed forward (kTileFwd16S): items = row-block groups of a
    // couple on an anti-diagonal, claimed in anti-diagonal order.  ranges[r]
    // = (first item, couple, diagonal, first group) of the consecutive
    // groups of one couple on one diagonal.
    const int4 *ranges;
    int32_t nRanges;
    int32_t nItems;
    uint32_t *queue;            // item counter (zeroed before the launch)
    uint32_t *progress;         // per couple x row block: tiles done (zeroed)
    int32_t progStride;         // row blocks per couple in `progress`
    uint32_t *error;            // set when a wait runs past its bound
};
constexpr int kTileSlack = 24;        // steps per tile beyond its width (skew + prefetch)
constexpr int kWalkStateInts = 8;
// Traceback window of one round: kRoundBlocks row blocks x kRoundTiles column
// tiles recomputed with trace records (tileC = kLongTileC).  Tiles of 128
// columns: the forward of 64 pairs of 100 kbp takes 157 ms against 171 ms
// with 256 columns (a lone tile's sweep, the per-anti-diagonal latency, is
// halved) and 291 ms with 64 (per-tile overheads take over).
constexpr int kLongTileC = 128;
constexpr int kTileRecSteps = kLongTileC + kTileSlack;
constexpr int kRoundBlocks = 8, kRoundTiles = 16, kRoundSlots = kRoundBlocks * kRoundTiles;
hipError_t launchTileFwd32(const TileArgs &args, hipStream_t stream);
// One persistent launch of the packed forward over every tile of the
// chunk's pairs (kTileFwd16S); gridBlocks of 256 threads.
hipError_t launchTileFwd16Stream(const TileArgs &args, int gridBlocks, hipStream_t stream);
int occupancyTileFwd16Stream();
// Packed variant: pair couples (2 c, 2 c + 1), values relative to a per-tile
// base; valid when kLong16Span x (max|s| + |o| + |e|) fits the int16 window.
constexpr int kLong16Span = 800;
hipError_t launchTileFwd16(const TileArgs &args, hipStream_t stream);
hipError_t launchTileTrace32(const TileArgs &args, hipStream_t stream);
// Scores of long pairs under free end gaps or local alignment, from the
// forward's checkpoints (free last row / column) or its running maximum.
hipError_t launchLongResult(const TileArgs &args, hipStream_t stream);
hipError_t launchTileWalkInit(const TileArgs &args, hipStream_t stream);
hipError_t launchTileWalk32(const TileArgs &args, hipStream_t stream);

// Trace family.
constexpr int kTraceRows = 256;       // rows per block of the trace kernel (G = 16, R = 16)
constexpr int kTraceLanes = 16;
int traceSteps(int maxN);             // column steps per block for an item of width maxN
// track: kTrackEnd (global) or the bottom / right bits of free end gaps.
hipError_t launchFwdTrace16(bool dna, uint32_t track, const FwdArgs &args, int gridBlocks, hipStream_t stream);
int occupancyTrace16(bool dna);
hipError_t launchWalk16(const WalkArgs &args, hipStream_t stream);
hipError_t launchCounts32(const Counts32Args &args, hipStream_t stream);

}  // namespace sqmi
