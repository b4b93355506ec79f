// engine.cpp -- host side of the C ABI: validation, exact int16/int32
// routing, length bucketing into wavefront work items, launches.
//
// Behaviour mirrors the reference entry points:
//   globalAlignmentScore / localAlignmentScore (ExecutionPolicy<Parallel,Vectorial>)
//     include/seqan/align_parallel/parallel_align_interface.h:100-145, 332-365
// The reference splits the set into chunks of <= 256 pairs for OpenMP threads
// and each chunk into SIMD batches of 8/16 pairs padded to the batch maximum
// (include/seqan/align/dp_align_simd_helper.h:147-221, 277-347).  Here pairs
// are instead grouped by length across the whole set, so that the pairs that
// share a wavefront have nearly the same shape.
#include "kernels.hpp"
#include "seqan_amd.h"

#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <thread>
#include <vector>

using namespace sqmi;

namespace {

// A bump allocator over caller-free device memory (the no-allocation entry
// point's scratch array): buffers carved from it are never freed.
struct Arena
{
    char *base = nullptr;
    size_t cap = 0, used = 0;
};

struct DevBuf
{
    void *p = nullptr;
    size_t cap = 0;
    Arena *arena = nullptr;   // set: carve from this arena instead of hipMalloc
};

// Every device buffer a context keeps between calls.
struct Buffers
{
    DevBuf hRes, vRes, hOff, vOff, order, items, queue, rowbuf, scores, tTable, dnaProf, ids32, scratchOff32,
        scratch32, table32, emptyIds;
    // Traceback (sqmi_align_counts).
    DevBuf trace, traceOff, locs, endEF, endIJ, counts;
    // Long pairs (tiled int32 kernels).
    DevBuf longCk, longPairs, longWalk, longTrace, longBest, longSched;
    // Alignment rows (sqmi_align_rows, sqmi_align_local_rows).
    DevBuf ops, opOff, opLen, begins, opsPacked, opStart;
    // Banded scores (sqmi_align_score_banded).
    DevBuf bandIds, bandScratch;

    template <class F>
    void forEach(F f)
    {
        for (DevBuf *b : {&hRes, &vRes, &hOff, &vOff, &order, &items, &queue, &rowbuf, &scores, &tTable, &dnaProf,
                          &ids32, &scratchOff32, &scratch32, &table32, &emptyIds, &trace, &traceOff, &locs, &endEF,
                          &endIJ, &counts, &longCk, &longPairs, &longWalk, &longTrace, &longBest, &longSched, &ops, &opOff, &opLen,
                          &begins, &opsPacked, &opStart, &bandIds, &bandScratch})
            f(*b);
    }
};

}  // namespace

struct sqmi_context
{
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t ownStream = nullptr;
    int numCUs = 0;
    int64_t l2PerCU = 0;    // bytes of L2 per CU (one L2 per XCD of 32 CUs on gfx950)
    int64_t totalMem = 0;   // device memory, bytes
    std::string lastError;
    sqmi_stats stats{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Long-pair forward: side streams (default priority) that sweep groups
    // of pair couples concurrently, with fork / join events.
    static constexpr int kMaxLongStreams = 4;
    hipStream_t longStreams[kMaxLongStreams] = {};
    hipEvent_t longFork = nullptr, longJoin[kMaxLongStreams] = {};
    // Streaming long-pair forward: its item ranges (kept until the next
    // call, as a pending upload may read them) and whether the call used it
    // (then the call checks the kernel's wait-bound flag after its sync).
    std::vector<int4> longRanges;
    bool longStreamed = false;
    Buffers b;
    // The no-allocation entry point: this device's scratch array (its device
    // address, from hipGetSymbolAddress at creation) and, during such a
    // call, the arena carved from it.
    void *scratchPtr = nullptr;
    Arena arena;
    bool noAlloc = false;   // inside sqmi_align_score_noalloc
    // The last sqmi_align_score call, kept for an identical next call: the
    // key of its inputs, its statistics and its launches as an instantiated
    // HIP graph.  Its plan lives on in the context's device buffers, so any
    // other call on the context drops it.
    struct Replay {
        bool valid = false;
        uint64_t key = 0;
        hipGraphExec_t exec = nullptr;
        sqmi_stats stats{};
        bool longStreamed = false;   // the graph holds the streaming long-pair forward
    } replay;
};

namespace {

sqmi_status fail(sqmi_context *ctx, sqmi_status st, const std::string &msg)
{
    if (ctx)
        ctx->lastError = msg;
    return st;
}

sqmi_status hipFail(sqmi_context *ctx, hipError_t err, const char *what)
{
    std::string msg = std::string(what) + ": " + hipGetErrorString(err);
    return fail(ctx, err == hipErrorOutOfMemory ? SQMI_ERR_OUT_OF_MEMORY : SQMI_ERR_DEVICE, msg);
}

hipError_t ensure(DevBuf &b, size_t bytes)
{
    if (bytes == 0)
        bytes = 16;
    if (b.cap >= bytes)
        return hipSuccess;
    if (b.arena) {
        // 256-byte aligned pieces of the arena; an outgrown piece stays unused.
        const size_t need = (bytes + 255) & ~(size_t)255;
        if (b.arena->used + need > b.arena->cap)
            return hipErrorOutOfMemory;
        b.p = b.arena->base + b.arena->used;
        b.arena->used += need;
        b.cap = need;
        return hipSuccess;
    }
    if (b.p)
        (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e == hipSuccess)
        b.cap = bytes;
    return e;
}

// Launches collected during a call's preparation and run (or captured into
// a HIP graph) once every upload is enqueued.
using Deferred = std::vector<std::function<hipError_t()>>;

// SQMI_GRAPH=0 turns off the replay of identical sqmi_align_score calls.
bool graphsEnabled()
{
    const char *env = std::getenv("SQMI_GRAPH");
    return !(env && *env == '0');
}

void dropReplay(sqmi_context *ctx)
{
    if (ctx->replay.exec)
        (void)hipGraphExecDestroy(ctx->replay.exec);
    ctx->replay = sqmi_context::Replay{};
}

uint64_t mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

// Order-sensitive 64-bit digest of n int64 values (a sum of position-tagged
// mixes, so threads can take contiguous ranges).
uint64_t digest64(const int64_t *a, int64_t n)
{
    uint64_t h = 0;
#pragma omp parallel for schedule(static) reduction(+ : h) if (n > 65536)
    for (int64_t i = 0; i < n; ++i)
        h += mix64((uint64_t)a[i] ^ ((uint64_t)i * 0x9e3779b97f4a7c15ULL));
    return mix64(h ^ (uint64_t)n);
}

// Bytes an arena piece of `bytes` takes (ensure() rounds to 256).
size_t arenaBytes(size_t bytes) { return ((bytes ? bytes : 16) + 255) & ~(size_t)255; }

void release(DevBuf &b)
{
    if (b.p && !b.arena)
        (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

struct ScoringInfo
{
    int alpha = 0;
    int o = 0, e = 0;
    int sMax = 0, sMin = 0;
    bool dna = false;   // v_perm profile path possible
};

// Local table kernels fuse the two lookups and the diagonal add into one
// 32-bit add when every table value s - gapOpen is >= 0.
int32_t tNonNeg(const ScoringInfo &si, uint32_t mode)
{
    return (mode & SQMI_LOCAL) != 0 && !si.dna && si.sMin >= si.o ? 1 : 0;
}

// Conservative range of every value a pair can produce inside the DP (H, E, F
// and H + gapOpen), from the sizes alone.  H <= max(0, sMax) * min(n, m) since
// gap costs are <= 0.  H(i,j) >= the score of the all-gap path along row 0
// and column j, 2*o + (i+j-2)*max(o,e) (a gap chain grows by max(o,e) per
// character); E, F and H + o are at least H + o; the diagonal term at least
// H + sMin.
bool fitsInt16(const ScoringInfo &si, int64_t n, int64_t m, bool local)
{
    const int64_t up = (int64_t)std::max(0, si.sMax) * std::min(n, m);
    const int64_t lowH = local ? 0 : 2LL * si.o + (n + m) * (int64_t)std::min(std::max(si.o, si.e), 0);
    const int64_t low = lowH + std::min({si.o, si.sMin, 0}) + std::min(si.e, 0);
    return up <= 32767 - kGuard16 && low >= (int64_t)kNeg16 + kGuard16 && n <= 30000 && m <= 30000;
}

// The same bound for the tiled int32 kernels of long pairs, whose minus
// infinity is -2^29: every real value stays within +-2^28.
bool fitsLong32(const ScoringInfo &si, int64_t n, int64_t m)
{
    const int64_t up = (int64_t)std::max(0, si.sMax) * std::min(n, m);
    const int64_t lowH = 2LL * si.o + (n + m) * (int64_t)std::min(std::max(si.o, si.e), 0);
    const int64_t low = lowH + std::min({si.o, si.sMin, 0}) + std::min(si.e, 0);
    return up < (1LL << 28) && low > -(1LL << 28) && n < (1LL << 30) && m < (1LL << 30);
}

struct Plan16
{
    Variant16 var;
    std::vector<int32_t> order;
    std::vector<WorkItem> items;
    int maxN = 0;
};

int pairsPerWave(const Variant16 &v) { return 2 * (64 / v.G); }

// Shapes compiled for the packed forward: groups of 16 lanes (kFwd16) or one
// pair couple per lane (kLane16).
bool shapeSupported(const Variant16 &v) { return v.G == 1 ? laneSupported(v) : variantSupported(v); }

// The planner's shape table (seqan_amd/tuning/shapes.json, turned into C++
// by the build): per category (DNA byte profile or substitution table,
// global or local), length bin b (2^(b-1) < |V| + 1 <= 2^b) and count class
// (the bin's pairs in this call, at least kTabCountClasses[c]), an index
// into kTabShapes.
#include "shape_table.inc"

int lengthBin(int64_t m)
{
    // The smallest b <= 15 with 2^b >= m + 1.
    if (m < 1)
        return 0;
    const int b = 64 - __builtin_clzll((unsigned long long)m);
    return b < 15 ? b : 15;
}

int shapeCategory(bool dna, bool local) { return (dna ? 0 : 2) + (local ? 1 : 0); }

// SQMI_SHAPE_RULE=fixed: the rule the table replaced (G = 16; R = 8 when
// |V| + 1 <= 128, else 16), for comparing tuned and untuned plans.
bool fixedShapeRule()
{
    const char *env = std::getenv("SQMI_SHAPE_RULE");
    return env && std::strcmp(env, "fixed") == 0;
}

// The shape of length bin `bin` for a call holding pairs16 pairs on the
// packed route (all bins: the shapes' launches run side by side, so the
// whole call is what fills the GPU).
Variant16 tableShape(int category, int bin, int64_t pairs16, bool dna, uint32_t track)
{
    if (fixedShapeRule())
        return Variant16{16, bin <= 7 ? 8 : 16, dna, track};
    int cls = 0;
    while (cls + 1 < kTabNClasses && pairs16 >= kTabCountClasses[cls + 1])
        ++cls;
    const int idx = kTabChoice[category][bin][cls];
    return Variant16{kTabShapes[idx][0], kTabShapes[idx][1], dna, track};
}

// Row blocks of a pair with m rows under shape v (at most 255 fit the keys).
// Row blocks of G * R rows (a power of two for every compiled shape).
int64_t blocksFor(const Variant16 &v, int64_t m)
{
    const int64_t br = (int64_t)v.G * v.R;
    return (br & (br - 1)) == 0 ? (m + br) >> __builtin_ctzll((unsigned long long)br) : (m + br) / br;
}

// One pair per launch: one wavefront, whose first 16-lane group sweeps the
// pair (the other pair slots of the wavefront stay empty).
Variant16 plainVariant(bool dna, uint32_t track, int64_t m)
{
    Variant16 v{16, 16, dna, track};
    if (m + 1 <= 128)
        v.R = 8;
    return v;
}

sqmi_status checkSet(sqmi_context *ctx, const sqmi_sequences *s, const char *name)
{
    if (!s)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, std::string(name) + " is null");
    if (s->count < 0)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, std::string(name) + ": negative count");
    if (s->count > 0 && (!s->offsets || !s->residues))
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, std::string(name) + ": null offsets or residues");
    if (s->count > INT32_MAX - 1)
        return fail(ctx, SQMI_ERR_UNSUPPORTED, std::string(name) + ": more than 2^31-2 sequences");
    if (s->count > 0 && s->offsets[0] < 0)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, std::string(name) + ": negative offset");
    for (int64_t i = 0; i < s->count; ++i)
        if (s->offsets[i + 1] < s->offsets[i])
            return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, std::string(name) + ": offsets not monotone");
    return SQMI_OK;
}

double msSince(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct Key
{
    uint32_t key;
    int32_t id;
    // Left uninitialised: the planner's key arrays (8 MB for a million
    // pairs) are written in full, in parallel, right after allocation.
    Key() {}
    Key(uint32_t k, int32_t i) : key(k), id(i) {}
};

struct Planned
{
    ScoringInfo si;
    uint32_t flags = 0;
    bool local = false;
    uint32_t mode = 0;
    uint32_t track = 0;
    std::vector<int32_t> emptyIds, ids32;
    std::vector<int32_t> idsLong;   // int32 pairs for the tiled long-pair kernels
    std::vector<int64_t> scratchOff32;
    int64_t scratch32 = 0;
    std::vector<Key> keys16;        // int16 pairs, sorted by bucket key
    std::vector<Variant16> shapes;  // int16 shape slots (key classes 0..)
    std::vector<Plan16> plans;      // one per kernel variant in use
    int64_t cells = 0;
};

// int32 pairs whose sides are both at least this long (and that are DNA,
// global) run the tiled long-pair kernels instead of one thread per pair
// (2048, or SQMI_LONG_MIN_SIDE).
// The packed long-pair forward keeps tile values relative to a base: it
// needs kLong16Span steps of at most D = max|s| + |o| + |e| inside the int16
// window.
bool longPacked(const ScoringInfo &si, uint32_t flags)
{
    const int64_t D = std::max(std::abs(si.sMax), std::abs(si.sMin)) + std::abs(si.o) + std::abs(si.e);
    return !(flags & SQMI_FORCE_INT32) && (int64_t)kLong16Span * D <= std::min(-kNeg16 - kGuard16, 32767 - kGuard16);
}

// Test and tuning switch: SQMI_LONG_ALL=1 sends every pair the long route
// can take there, also those the packed kernels could score.
bool longAll()
{
    const char *env = std::getenv("SQMI_LONG_ALL");
    return env && *env == '1';
}

bool envFlag(const char *name)
{
    const char *env = std::getenv(name);
    return env && *env == '1';
}

int64_t longMinSide()
{
    const char *env = std::getenv("SQMI_LONG_MIN_SIDE");
    if (env && *env && std::atoll(env) > 0)
        return std::atoll(env);
    return 2048;
}

// The packed long-pair forward as one persistent launch over every tile
// (SQMI_LONG_STREAM=1), or one launch per tile anti-diagonal (default).  On
// 64 pairs of 100 kbp the persistent form measured 382 ms against 161 ms:
// 323 ms even without its agent-scope fences, so the dataflow schedule
// itself, not only the cross-XCD coherence, costs more than the launches.
bool longStreaming()
{
    const char *env = std::getenv("SQMI_LONG_STREAM");
    return env && *env == '1';
}

// After a call's final synchronisation: the streaming forward's waits are
// bounded, and one that ran past its bound leaves a flag (never expected:
// every wait is for tiles that running waves hold).
sqmi_status longStreamCheck(sqmi_context *ctx)
{
    if (!ctx->longStreamed)
        return SQMI_OK;
    ctx->longStreamed = false;
    uint32_t flag = 0;
    if (hipMemcpy(&flag, static_cast<const uint32_t *>(ctx->b.longSched.p) + 1, 4, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(ctx, SQMI_ERR_DEVICE, "reading the long-pair forward's wait flag");
    if (flag)
        return fail(ctx, SQMI_ERR_DEVICE, "long-pair forward: a tile wait ran past its bound; scores are not valid");
    return SQMI_OK;
}

// Streams the long-pair forward spreads its pair couples over (1..4).  On
// 64 pairs of 100 kbp with 128-column tiles: 1 stream 263 ms, 2 streams
// 164 ms, 3 streams 157 ms (with 256-column tiles: 220, 170, 170 ms; 4
// streams 260 ms, the 4th sharing a hardware queue); starting the groups
// some anti-diagonals apart only made it slower.
int longStreams()
{
    const char *env = std::getenv("SQMI_LONG_STREAMS");
    const int v = (env && *env) ? std::atoi(env) : 3;
    return std::max(1, std::min(v, sqmi_context::kMaxLongStreams));
}


// Calls with at most this many pairs plan on one thread: waking the thread
// team costs more than it saves.
constexpr int64_t kParallelMin = 32768;

// Stable LSD radix sort of 28-bit keys in two passes of 14 bits, each pass
// split into contiguous per-thread ranges so that equal digits keep their
// input order.
void radixSort(std::vector<Key> &keys)
{
    const size_t n = keys.size();
    if (n < 2)
        return;
    std::vector<Key> tmp(n);
    int T = 1;
#ifdef _OPENMP
    T = std::max(1, std::min(omp_get_max_threads(), (int)(n / 16384) + 1));
#endif
    constexpr int B = 14, NB = 1 << B;
    std::vector<uint32_t> hist((size_t)T * NB);
    for (int shift = 0; shift < 2 * B; shift += B) {
        const uint32_t mask = (uint32_t)NB - 1;
        std::fill(hist.begin(), hist.end(), 0u);
#pragma omp parallel num_threads(T)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            const size_t lo = n * t / T, hi = n * (t + 1) / T;
            uint32_t *h = hist.data() + (size_t)t * NB;
            for (size_t k = lo; k < hi; ++k)
                ++h[(keys[k].key >> shift) & mask];
#pragma omp barrier
#pragma omp single
            {
                uint32_t run = 0;
                for (int d = 0; d < NB; ++d)
                    for (int u = 0; u < T; ++u) {
                        uint32_t &c = hist[(size_t)u * NB + d];
                        const uint32_t v = c;
                        c = run;
                        run += v;
                    }
            }
            for (size_t k = lo; k < hi; ++k)
                tmp[h[(keys[k].key >> shift) & mask]++] = keys[k];
        }
        keys.swap(tmp);
    }
}

// Bucket keys (28 bits): class (3 bits: 0..5 int16 shape slot, 6 int32,
// 7 empty) | 1023 - row units (10 bits) | 32767 - |H| (15 bits).  Ascending
// order puts int16 pairs first, grouped by shape and row units (most first),
// widest first, then the int32 and the empty pairs in input order.  A row
// unit is a quarter of a G*R-row block for the packed kernels with groups
// of 16 or 64 lanes (their block 0 may be 1..4 quarters tall), a whole
// block for the one-pair-per-lane kernel and the trace kernel.
constexpr int kKeyNBits = 15, kKeyUnitBits = 10, kKeyClassShift = kKeyNBits + kKeyUnitBits;
constexpr uint32_t kClassInt32 = 6, kClassEmpty = 7, kMaxShapeSlots = 6;
constexpr int64_t kMaxBlocks16 = 255;   // row blocks per pair of a packed work item
inline uint32_t bucketOf(uint32_t key) { return key >> kKeyNBits; }
inline int32_t widthOf(uint32_t key) { return (int32_t)((1u << kKeyNBits) - 1 - (key & ((1u << kKeyNBits) - 1))); }
inline int32_t unitsOf(uint32_t key) { return (int32_t)(1023u - ((key >> kKeyNBits) & 1023u)); }
// Row units of a shape: quarter blocks where block 0 may be shortened.
inline bool quarterUnits(const Variant16 &v, bool traceShape) { return v.G >= 16 && v.R % 4 == 0 && !traceShape; }
// Quarter-block shapes: a pair takes just enough quarters (fine), or whole
// blocks counted as 4 quarters each (coarse).  Fine units cut the rows of
// block 0 but split the pairs of a length bin over 4 times as many buckets,
// which widens the spread of |H| inside a work item when the bin holds few
// pairs; so a bin takes fine units from kFineUnitPairs pairs on
// (SQMI_ROW_UNITS=block / quarter: always coarse / fine).
constexpr int64_t kFineUnitPairs = 8192;
int rowUnitsMode()   // 0 by bin size, 1 coarse, 2 fine
{
    const char *env = std::getenv("SQMI_ROW_UNITS");
    if (env && std::strcmp(env, "block") == 0)
        return 1;
    if (env && std::strcmp(env, "quarter") == 0)
        return 2;
    return 0;
}
inline int64_t unitsFor(const Variant16 &v, int64_t m, bool quarter, bool fine)
{
    if (!quarter)
        return blocksFor(v, m);
    if (!fine)
        return 4 * blocksFor(v, m);
    const int64_t rows = (int64_t)v.G * v.R / 4;
    return (rows & (rows - 1)) == 0 ? (m + rows) >> __builtin_ctzll((unsigned long long)rows) : (m + rows) / rows;
}

// Routing (empty / int16 / int32) and length bucketing into work items.
// Pure host code: needs no device.
void plan(const sqmi_sequences &H, const sqmi_sequences &V, const sqmi_scoring &scoring, uint32_t mode,
          uint32_t flags, Planned &P, bool traceShape = false, bool noLong = false)
{
    P.flags = flags;
    const int64_t N = H.count;
    ScoringInfo &si = P.si;
    si.alpha = scoring.alphabet_size;
    si.o = scoring.gap_open;
    si.e = scoring.gap_extend;
    si.sMax = INT_MIN;
    si.sMin = INT_MAX;
    for (int k = 0; k < si.alpha * si.alpha; ++k) {
        si.sMax = std::max(si.sMax, scoring.table[k]);
        si.sMin = std::min(si.sMin, scoring.table[k]);
    }
    const bool local = (mode & SQMI_LOCAL) != 0;
    P.local = local;
    const bool costsOk16 = si.o <= 0 && si.e <= 0 && si.o >= -1024 && si.e >= -1024 && si.sMax <= 1024 &&
                           si.sMin >= -1024 && !(flags & SQMI_FORCE_INT32) &&
                           (si.alpha <= 4 || si.alpha + 1 <= kTabMaxCodes);
    // DNA byte profile: t(h, v) - gapOpen in 0..255.  Its padding column
    // (past n) is 0, i.e. t = gapOpen <= 0, so under local alignment the
    // cells past n stay at or below the maximum over columns 1..n, and an
    // equal one follows its source in column order.
    si.dna = si.alpha <= 4 && si.sMin - si.o >= 0 && si.sMax - si.o <= 255;
    uint32_t track;
    if (local)
        track = kTrackLocal;
    else {
        track = 0;
        if (mode & SQMI_FREE_BOTTOM)
            track |= kTrackBottom;
        if (mode & SQMI_FREE_RIGHT)
            track |= kTrackRight;
        if (!track)
            track = kTrackEnd;
    }
    P.track = track;
    P.mode = mode;
    // Route every pair and give it a bucket key (see bucketOf).  The int16
    // pairs' shape slots: the trace shape, a forced shape, or per length bin
    // the shape table's choice for the number of the call's pairs in that
    // bin (a pair that needs more than 255 row blocks of its bin's shape
    // takes G = 16, R = 16).
    std::vector<Key> keys((size_t)N);
    int64_t cells = 0;
    // The long route: every AlignConfig and local for scores, global only
    // for the traceback.
    const bool allLong = longAll() && si.dna && (mode == 0 || !traceShape) && !noLong;
    const int64_t longMin0 = longMinSide();
    auto route16 = [&](int64_t n, int64_t m) {
        return !(allLong && n >= longMin0 && m >= longMin0) && costsOk16 && fitsInt16(si, n, m, local) &&
               n < (1 << kKeyNBits);
    };
    std::vector<Variant16> &shapes = P.shapes;
    shapes.clear();
    uint8_t binSlot[16];
    bool fineBin[16] = {};
    int fallbackSlot = -1;
    const int forcedG = (int)((flags >> 8) & 0xFF), forcedR = (int)((flags >> 16) & 0xFF);
    if (traceShape || (forcedG && forcedR)) {
        shapes.push_back(traceShape ? Variant16{kTraceLanes, kTraceRows / kTraceLanes, si.dna, track}
                                    : Variant16{forcedG, forcedR, si.dna, track});
        std::fill(binSlot, binSlot + 16, (uint8_t)0);
        std::fill(fineBin, fineBin + 16, rowUnitsMode() == 2);
        if (shapes[0].G == 1) {
            // Pairs taller than 255 blocks of the forced lane shape.
            shapes.push_back(Variant16{16, 16, si.dna, track});
            fallbackSlot = 1;
        }
    } else {
        // Pairs per length bin (int16 route only), then one slot per shape.
        int64_t binCount[16] = {};
#pragma omp parallel if (N > kParallelMin)
        {
            int64_t local16[16] = {};
#pragma omp for schedule(static) nowait
            for (int64_t i = 0; i < N; ++i) {
                const int64_t n = H.offsets[i + 1] - H.offsets[i];
                const int64_t m = V.offsets[i + 1] - V.offsets[i];
                if (n != 0 && m != 0 && route16(n, m))
                    ++local16[lengthBin(m)];
            }
#pragma omp critical
            for (int b = 0; b < 16; ++b)
                binCount[b] += local16[b];
        }
        const int cat = shapeCategory(si.dna, local);
        auto slotOf = [&](const Variant16 &v) {
            for (size_t k = 0; k < shapes.size(); ++k)
                if (shapes[k].G == v.G && shapes[k].R == v.R)
                    return (int)k;
            shapes.push_back(v);
            return (int)shapes.size() - 1;
        };
        int64_t pairs16 = 0;
        for (int b = 0; b < 16; ++b)
            pairs16 += binCount[b];
        // SQMI_BIN_SHAPES="b:G:R,...": the shape of bin b, in place of the
        // table's (the shape tuner times whole calls with it).
        int overG[16] = {}, overR[16] = {};
        if (const char *env = std::getenv("SQMI_BIN_SHAPES")) {
            for (const char *c = env; *c;) {
                int b = -1, g = 0, r = 0, used = 0;
                if (std::sscanf(c, "%d:%d:%d%n", &b, &g, &r, &used) != 3)
                    break;
                if (b >= 0 && b < 16) {
                    overG[b] = g;
                    overR[b] = r;
                }
                c += used;
                if (*c == ',')
                    ++c;
            }
        }
        for (int b = 0; b < 16; ++b) {
            Variant16 v = overG[b] ? Variant16{overG[b], overR[b], si.dna, track}
                                   : tableShape(cat, b, pairs16, si.dna, track);
            if (!shapeSupported(v))
                v = Variant16{16, 16, si.dna, track};
            binSlot[b] = (uint8_t)(binCount[b] ? slotOf(v) : 0);
        }
        fallbackSlot = slotOf(Variant16{16, 16, si.dna, track});
        const int unitsMode = rowUnitsMode();
        for (int b = 0; b < 16; ++b)
            fineBin[b] = unitsMode == 2 || (unitsMode == 0 && binCount[b] >= kFineUnitPairs);
    }
#pragma omp parallel for schedule(static) reduction(+ : cells) if (N > kParallelMin)
    for (int64_t i = 0; i < N; ++i) {
        const int64_t n = H.offsets[i + 1] - H.offsets[i];
        const int64_t m = V.offsets[i + 1] - V.offsets[i];
        uint32_t key = kClassEmpty << kKeyClassShift;
        if (n != 0 && m != 0) {
            cells += n * m;
            key = kClassInt32 << kKeyClassShift;
            if (route16(n, m)) {
                uint32_t slot = binSlot[lengthBin(m)];
                int64_t nblocks = blocksFor(shapes[slot], m);
                if (nblocks > kMaxBlocks16 && fallbackSlot >= 0) {
                    slot = (uint32_t)fallbackSlot;
                    nblocks = blocksFor(shapes[slot], m);
                }
                if (nblocks <= kMaxBlocks16) {
                    const int64_t units =
                        unitsFor(shapes[slot], m, quarterUnits(shapes[slot], traceShape), fineBin[lengthBin(m)]);
                    key = (slot << kKeyClassShift) | ((1023u - (uint32_t)units) << kKeyNBits) |
                          ((1u << kKeyNBits) - 1 - (uint32_t)n);
                }
            }
        }
        keys[i] = Key{key, (int32_t)i};
    }
    P.cells = cells;
    radixSort(keys);
    // Segments of the sorted order: int16 (by shape), int32, empty.
    size_t n16 = 0;
    while (n16 < keys.size() && (keys[n16].key >> kKeyClassShift) < kClassInt32) {
        // Jump by bucket boundaries found with a binary search.
        const uint32_t cls = keys[n16].key >> kKeyClassShift;
        n16 = (size_t)(std::upper_bound(keys.begin() + n16, keys.end(), ((cls + 1) << kKeyClassShift) - 1,
                                        [](uint32_t v, const Key &k) { return v < k.key; }) -
                       keys.begin());
    }
    size_t e32 = n16;
    while (e32 < keys.size() && (keys[e32].key >> kKeyClassShift) == kClassInt32)
        ++e32;
    const int64_t longMin = longMinSide();
    const bool longOk = si.dna && (mode == 0 || !traceShape) && !noLong;
    for (size_t k = n16; k < e32; ++k) {
        const int32_t i = keys[k].id;
        const int64_t n = H.offsets[i + 1] - H.offsets[i];
        const int64_t m = V.offsets[i + 1] - V.offsets[i];
        if (longOk && n >= longMin && m >= longMin && fitsLong32(si, n, m)) {
            P.idsLong.push_back(i);
            continue;
        }
        P.ids32.push_back(i);
        P.scratchOff32.push_back(P.scratch32);
        // score: H and E columns (int32 pairs); traceback: the columns plus
        // one trace byte per cell, 16-byte aligned.
        P.scratch32 += traceShape ? ((8 * (m + 1) + (n + 1) * (m + 1) + 15) & ~(int64_t)15) : m + 1;
    }
    for (size_t k = e32; k < keys.size(); ++k)
        P.emptyIds.push_back(keys[k].id);
    keys.resize(n16);
    std::vector<Key> &keys16 = P.keys16;
    keys16.swap(keys);

    // Work items: each bucket (shape, row blocks) is cut into runs of up to
    // pairs-per-wave pairs, in sorted order.  Bucket starts are found in
    // parallel, then every bucket writes its own items.
    std::vector<size_t> starts;
    {
        int T = 1;
#ifdef _OPENMP
        T = std::max(1, std::min(omp_get_max_threads(), (int)(n16 / 16384) + 1));
#endif
        std::vector<std::vector<size_t>> part((size_t)T);
#pragma omp parallel num_threads(T)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            const size_t lo = n16 * t / T, hi = n16 * (t + 1) / T;
            for (size_t k = lo; k < hi; ++k)
                if (k == 0 || bucketOf(keys16[k].key) != bucketOf(keys16[k - 1].key))
                    part[t].push_back(k);
        }
        for (auto &v : part)
            starts.insert(starts.end(), v.begin(), v.end());
        starts.push_back(n16);
    }
    const size_t nb = starts.size() - 1;
    std::vector<size_t> itemOff(nb + 1, 0);
    for (size_t b = 0; b < nb; ++b) {
        const int PP = pairsPerWave(shapes[keys16[starts[b]].key >> kKeyClassShift]);
        itemOff[b + 1] = itemOff[b] + (starts[b + 1] - starts[b] + PP - 1) / PP;
    }
    std::vector<Plan16> &plans = P.plans;
    for (uint32_t v = 0; v < (uint32_t)shapes.size(); ++v) {
        // The shape's contiguous range of buckets and of sorted pairs.
        size_t b0 = 0;
        while (b0 < nb && (keys16[starts[b0]].key >> kKeyClassShift) < v)
            ++b0;
        size_t b1 = b0;
        while (b1 < nb && (keys16[starts[b1]].key >> kKeyClassShift) == v)
            ++b1;
        if (b0 == b1)
            continue;
        plans.push_back(Plan16{});
        Plan16 &pl = plans.back();
        pl.var = shapes[v];
        const size_t p0 = starts[b0], p1 = starts[b1];
        const size_t i0 = itemOff[b0];
        const int PP = pairsPerWave(pl.var);
        const bool quarter = quarterUnits(pl.var, traceShape);
        pl.order.resize(p1 - p0);
        pl.items.resize(itemOff[b1] - i0);
        int32_t maxN = 0;
#pragma omp parallel for schedule(static) reduction(max : maxN) if (p1 - p0 > (size_t)kParallelMin)
        for (int64_t k = (int64_t)p0; k < (int64_t)p1; ++k) {
            pl.order[k - p0] = keys16[k].id;
            maxN = std::max(maxN, widthOf(keys16[k].key));
        }
        pl.maxN = maxN;
#pragma omp parallel for schedule(dynamic, 8) if (p1 - p0 > (size_t)kParallelMin)
        for (int64_t b = (int64_t)b0; b < (int64_t)b1; ++b) {
            size_t it = itemOff[b] - i0;
            const int32_t units = unitsOf(keys16[starts[b]].key);
            const int32_t nblocks = quarter ? (units + 3) / 4 : units;
            const int32_t headQ = quarter ? units - 4 * (nblocks - 1) : 4;   // quarters in block 0
            for (size_t k = starts[b]; k < starts[b + 1]; k += PP, ++it) {
                const size_t e = std::min(starts[b + 1], k + PP);
                // Widest first inside a bucket: the first pair has the item's width.
                pl.items[it] = WorkItem{(int32_t)(k - p0), (int32_t)(e - k) | ((headQ & 3) << 16),
                                        widthOf(keys16[k].key), nblocks};
            }
        }
    }
}

void fillStats(const Planned &P, sqmi_stats &st)
{
    st.pairs = (int64_t)(P.emptyIds.size() + P.ids32.size() + P.idsLong.size() + P.keys16.size());
    st.cells = P.cells;
    st.pairs_int32 = (int64_t)(P.ids32.size() + P.idsLong.size());
    st.pairs_long = (int64_t)P.idsLong.size();
    st.pairs_long_packed = (longPacked(P.si, P.flags) && P.mode == 0) ? st.pairs_long : 0;
    st.pairs_int16 = (int64_t)P.keys16.size();
    int64_t items = 0, padded = 0;
    st.pairs_lane = 0;
    for (const Plan16 &pl : P.plans) {
        items += (int64_t)pl.items.size();
        if (pl.var.G == 1)
            st.pairs_lane += (int64_t)pl.order.size();
        for (const WorkItem &w : pl.items) {
            const int64_t BR = (int64_t)pl.var.G * pl.var.R, headQ = ((w.count >> 16) & 3) ? ((w.count >> 16) & 3) : 4;
            padded += 2LL * (64 / pl.var.G) * (int64_t)w.maxN * ((w.nblocks - 1) * BR + headQ * BR / 4);
        }
    }
    st.work_items = items;
    st.padded_cells = padded;
}

sqmi_status validate(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                     const sqmi_scoring *scoring, uint32_t &mode)
{
    sqmi_status st;
    if ((st = checkSet(ctx, setH, "set_h")) != SQMI_OK || (st = checkSet(ctx, setV, "set_v")) != SQMI_OK)
        return st;
    if (setH->count != setV->count)
        return fail(ctx, SQMI_ERR_LENGTH_MISMATCH, "set_h and set_v differ in length");
    if (!scoring || !scoring->table || scoring->alphabet_size < 1 || scoring->alphabet_size > 32)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "scoring: alphabet_size must be 1..32 with a table");
    if (mode & ~31u)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "mode: unknown bits");
    if (mode & SQMI_LOCAL)
        mode = SQMI_LOCAL;   // localAlignmentScore has no AlignConfig
    return SQMI_OK;
}

sqmi_status checkForced(sqmi_context *ctx, uint32_t flags)
{
    const int g = (int)((flags >> 8) & 0xFF), r = (int)((flags >> 16) & 0xFF);
    if (!g && !r)
        return SQMI_OK;
    for (bool dna : {true, false})
        if (!shapeSupported(Variant16{g, r, dna, kTrackEnd}))
            return fail(ctx, SQMI_ERR_UNSUPPORTED, "forced kernel variant is not compiled");
    return SQMI_OK;
}


// Bytes the traceback may hold in trace records or int32 trace scratch at
// once: a quarter of the device's memory, at least 8 GiB and at most 64 GiB
// (72 GB -> 64 GiB on a 288 GB MI355X), or SQMI_TRACE_BUDGET_MB from the
// environment.
int64_t traceBudget(const sqmi_context *ctx)
{
    const char *env = std::getenv("SQMI_TRACE_BUDGET_MB");
    if (env && *env) {
        const long long mb = std::atoll(env);
        if (mb > 0)
            return (int64_t)mb << 20;
    }
    const int64_t quarter = ctx ? ctx->totalMem / 4 : 0;
    return std::min<int64_t>(std::max<int64_t>(quarter, (int64_t)8 << 30), (int64_t)64 << 30);
}

// Work items with at least this many row blocks are swept by the four waves
// of a workgroup together (10, i.e. more than 2304 rows, or
// SQMI_COOP_MIN_BLOCKS; 0 disables).  Shorter items gain nothing from it.
int coopMinBlocks()
{
    const char *env = std::getenv("SQMI_COOP_MIN_BLOCKS");
    if (env && *env)
        return std::max(0, std::atoi(env));
    return kTabCoopMinBlocks;
}

// Device copies of what every entry point needs: residues (unless already on
// the device), offsets, the output scores buffer (unless the caller's is on
// the device) and the substitution tables: t(h, v) = s(h, v) - gapOpen with a
// padding code (int16 kernels), its byte profile (DNA path) and the plain
// table (int32 kernels).  The host staging data of the asynchronous copies
// lives in the struct, which the caller keeps until its stream is synchronised.
struct Inputs
{
    const uint8_t *dH = nullptr, *dV = nullptr;
    int32_t *dScores = nullptr;
    std::vector<int16_t> tTab;
    // DNA profiles: dwords 0..7 by column residue h (byte v = t(h, v)),
    // dwords 8..15 by row residue v (byte h = t(h, v)); the padding code
    // (alpha) maps to zero bytes.
    uint32_t prof[16] = {};
};

// A host thread that is joined when it goes out of scope (early returns
// included).
struct JoinedThread {
    std::thread t;
    ~JoinedThread()
    {
        if (t.joinable())
            t.join();
    }
    void join()
    {
        if (t.joinable())
            t.join();
    }
};

// Host residues into the context's buffers (in.dH / in.dV).
hipError_t uploadResidues(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV, Inputs &in)
{
    const int64_t N = setH->count;
    const int64_t hTot = setH->offsets[N], vTot = setV->offsets[N];
    hipError_t err;
    if ((err = ensure(ctx->b.hRes, hTot + 64)) != hipSuccess || (err = ensure(ctx->b.vRes, vTot + 64)) != hipSuccess ||
        (err = hipMemcpyAsync(ctx->b.hRes.p, setH->residues, hTot, hipMemcpyHostToDevice, ctx->stream)) !=
            hipSuccess ||
        (err = hipMemcpyAsync(ctx->b.vRes.p, setV->residues, vTot, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
        return err;
    in.dH = static_cast<const uint8_t *>(ctx->b.hRes.p);
    in.dV = static_cast<const uint8_t *>(ctx->b.vRes.p);
    return hipSuccess;
}

// Residues (unless uploadResidues already ran for `in`), offsets, tables.
sqmi_status uploadInputs(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                         const sqmi_scoring *scoring, const ScoringInfo &si, int32_t *scores, uint32_t flags,
                         Inputs &in)
{
    hipStream_t stream = ctx->stream;
    const int64_t N = setH->count;
    hipError_t err;
    const uint8_t *dH = setH->residues, *dV = setV->residues;
    if (!(flags & SQMI_RESIDUES_ON_DEVICE)) {
        if (!in.dH && (err = uploadResidues(ctx, setH, setV, in)) != hipSuccess)
            return hipFail(ctx, err, "copying residues");
        dH = in.dH;
        dV = in.dV;
    }
    if ((err = ensure(ctx->b.hOff, 8 * (N + 1))) != hipSuccess || (err = ensure(ctx->b.vOff, 8 * (N + 1))) != hipSuccess)
        return hipFail(ctx, err, "allocating offsets");
    if ((err = hipMemcpyAsync(ctx->b.hOff.p, setH->offsets, 8 * (N + 1), hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (err = hipMemcpyAsync(ctx->b.vOff.p, setV->offsets, 8 * (N + 1), hipMemcpyHostToDevice, stream)) != hipSuccess)
        return hipFail(ctx, err, "copying offsets");
    int32_t *dScores = scores;
    if (!(flags & SQMI_SCORES_ON_DEVICE)) {
        if ((err = ensure(ctx->b.scores, 4 * N)) != hipSuccess)
            return hipFail(ctx, err, "allocating scores");
        dScores = static_cast<int32_t *>(ctx->b.scores.p);
    }

    // Substitution tables: t(h, v) = s(h, v) - gapOpen with a padding code.
    const int A = si.alpha, TW = A + 1;
    std::vector<int16_t> &tTab = in.tTab;
    tTab.assign((size_t)TW * TW, -1024);
    for (int h = 0; h < A; ++h)
        for (int v = 0; v < A; ++v)
            tTab[h * TW + v] = (int16_t)(scoring->table[h * A + v] - si.o);
    uint32_t *prof = in.prof;
    if (si.dna)
        for (int h = 0; h < A; ++h)
            for (int v = 0; v < A; ++v) {
                prof[h] |= (uint32_t)(scoring->table[h * A + v] - si.o) << (8 * v);
                prof[8 + v] |= (uint32_t)(scoring->table[h * A + v] - si.o) << (8 * h);
            }
    if ((err = ensure(ctx->b.tTable, tTab.size() * 2)) != hipSuccess || (err = ensure(ctx->b.dnaProf, 64)) != hipSuccess ||
        (err = ensure(ctx->b.table32, 4 * A * A)) != hipSuccess)
        return hipFail(ctx, err, "allocating tables");
    if ((err = hipMemcpyAsync(ctx->b.tTable.p, tTab.data(), tTab.size() * 2, hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (err = hipMemcpyAsync(ctx->b.dnaProf.p, prof, 64, hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (err = hipMemcpyAsync(ctx->b.table32.p, scoring->table, 4 * A * A, hipMemcpyHostToDevice, stream)) != hipSuccess)
        return hipFail(ctx, err, "copying tables");
    in.dH = dH;
    in.dV = dV;
    in.dScores = dScores;
    return SQMI_OK;
}

// Long pairs: the tiled int32 forward (one launch per tile anti-diagonal),
// then, with `dCounts`, the checkpointed traceback: per round every
// unfinished pair's current tile is recomputed with trace records and
// walked.  Pairs run in chunks whose checkpoints fit traceBudget().  `pairs`
// holds the host copy of the descriptors until the caller synchronises.
// Device addresses of the alignment rows output (all null: no rows).
struct RowsDev {
    uint8_t *ops = nullptr;
    const int64_t *opOff = nullptr;
    int64_t *opLen = nullptr;
    int64_t *begins = nullptr;   // local windows only
};

sqmi_status runLong(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV, const Planned &P,
                    const Inputs &in, uint32_t flags, int32_t *dCounts, const RowsDev &rows,
                    std::vector<LongPair> &pairs, Deferred *deferred = nullptr)
{
    // Stream operations run at once, or are appended to `deferred` (all
    // uploads and allocations still happen here).
    hipError_t opErr = hipSuccess;
    auto op = [&](std::function<hipError_t()> f) {
        if (deferred)
            deferred->push_back(std::move(f));
        else if (opErr == hipSuccess)
            opErr = f();
        return opErr;
    };
    if (P.idsLong.empty())
        return SQMI_OK;
    hipStream_t stream = ctx->stream;
    const int C = kLongTileC;
    const int64_t budget = traceBudget(ctx);
    const int64_t N = setH->count;
    // The packed forward is global only; free end gaps and local take the
    // int32 forward and a result pass over its checkpoints.
    const bool packed = longPacked(P.si, flags) && P.mode == 0;
    // Longest first, so that the couples of the packed forward pair similar shapes.
    std::vector<int32_t> ids(P.idsLong);
    std::stable_sort(ids.begin(), ids.end(), [&](int32_t x, int32_t y) {
        const int64_t mx = setV->offsets[x + 1] - setV->offsets[x], my = setV->offsets[y + 1] - setV->offsets[y];
        const int64_t nx = setH->offsets[x + 1] - setH->offsets[x], ny = setH->offsets[y + 1] - setH->offsets[y];
        return mx != my ? mx > my : nx > ny;
    });
    // Descriptors with checkpoint offsets relative to the pair's chunk.
    std::vector<size_t> chunkEnd;
    int64_t run = 0, maxRun = 0, bestRun = 0;
    pairs.clear();
    for (int32_t id : ids) {
        LongPair lp{};
        lp.pid = id;
        lp.n = (int32_t)(setH->offsets[id + 1] - setH->offsets[id]);
        lp.m = (int32_t)(setV->offsets[id + 1] - setV->offsets[id]);
        lp.nblocks = (int32_t)(((int64_t)lp.m + 1 + kTileRows - 1) / kTileRows);
        lp.ntiles = (lp.n + C - 1) / C;
        lp.p = lp.nblocks * kTileRows - 1 - lp.m;
        const int64_t need = (int64_t)lp.nblocks * (lp.n + 1) * 2 + (int64_t)lp.nblocks * lp.ntiles * kTileRows * 2;
        if (run > 0 && (run + need) * 4 > budget) {
            chunkEnd.push_back(pairs.size());
            run = 0;
        }
        lp.rowck = run;
        lp.colck = run + (int64_t)lp.nblocks * (lp.n + 1) * 2;
        lp.bestck = bestRun;
        bestRun += lp.nblocks;
        run += need;
        maxRun = std::max(maxRun, run);
        pairs.push_back(lp);
    }
    chunkEnd.push_back(pairs.size());
    const size_t np = pairs.size();
    hipError_t err;
    if ((err = ensure(ctx->b.longCk, (size_t)maxRun * 4)) != hipSuccess ||
        (err = ensure(ctx->b.longPairs, sizeof(LongPair) * np)) != hipSuccess ||
        (err = ensure(ctx->b.endEF, (size_t)8 * N)) != hipSuccess ||
        (P.local && (err = ensure(ctx->b.longBest, (size_t)4 * std::max<int64_t>(bestRun, 1))) != hipSuccess))
        return hipFail(ctx, err, "allocating long-pair buffers");
    if (dCounts && ((err = ensure(ctx->b.longWalk, (size_t)4 * kWalkStateInts * np)) != hipSuccess ||
                    (err = ensure(ctx->b.longTrace, sizeof(uint2) * (size_t)kTileRecSteps * 16 * kRoundSlots * np)) !=
                        hipSuccess))
        return hipFail(ctx, err, "allocating long-pair traceback buffers");
    if ((err = hipMemcpyAsync(ctx->b.longPairs.p, pairs.data(), sizeof(LongPair) * np, hipMemcpyHostToDevice,
                              stream)) != hipSuccess)
        return hipFail(ctx, err, "uploading long pairs");
    size_t b = 0;
    for (size_t e : chunkEnd) {
        TileArgs a{};
        a.hRes = in.dH;
        a.vRes = in.dV;
        a.hOff = static_cast<const int64_t *>(ctx->b.hOff.p);
        a.vOff = static_cast<const int64_t *>(ctx->b.vOff.p);
        a.pairs = static_cast<const LongPair *>(ctx->b.longPairs.p) + b;
        a.nPairs = (int32_t)(e - b);
        a.tileC = C;
        a.ck = static_cast<int32_t *>(ctx->b.longCk.p);
        a.gapOpen = P.si.o;
        a.gapExtend = P.si.e;
        a.dnaProfile = static_cast<const uint32_t *>(ctx->b.dnaProf.p);
        a.scores = in.dScores;
        a.endEF = static_cast<int32_t *>(ctx->b.endEF.p);
        a.mode = P.mode;
        a.best = static_cast<int32_t *>(ctx->b.longBest.p);
        a.alpha = P.si.alpha;
        int maxB = 0, maxT = 0;
        for (size_t k = b; k < e; ++k) {
            maxB = std::max(maxB, pairs[k].nblocks);
            maxT = std::max(maxT, pairs[k].ntiles);
        }
        a.wavesPerPair = (maxB + 3) / 4;
        // (One chunk only: the schedule's upload happens now, and with
        // deferred launches a second chunk's would overwrite the first's.)
        if (packed && longStreaming() && chunkEnd.size() == 1 && maxB + maxT < 32768) {
            // One persistent launch: items = groups of four row blocks of a
            // couple on one anti-diagonal, in anti-diagonal order (ranges of
            // consecutive groups of a couple).
            const int couples = (int)((e - b + 1) / 2);
            std::vector<int> nbC(couples), ntC(couples);
            int stride = 1, maxD = 0;
            for (int c = 0; c < couples; ++c) {
                const LongPair &p0 = pairs[b + 2 * c];
                const LongPair &p1 = b + 2 * c + 1 < e ? pairs[b + 2 * c + 1] : p0;
                nbC[c] = std::max(p0.nblocks, p1.nblocks);
                ntC[c] = std::max(p0.ntiles, p1.ntiles);
                stride = std::max(stride, nbC[c]);
                maxD = std::max(maxD, nbC[c] + ntC[c] - 1);
            }
            std::vector<int4> &ranges = ctx->longRanges;
            ranges.clear();
            int64_t items = 0;
            for (int d = 0; d < maxD; ++d)
                for (int c = 0; c < couples; ++c) {
                    const int blo = std::max(0, d - ntC[c] + 1), bhi = std::min(d, nbC[c] - 1);
                    if (blo > bhi)
                        continue;
                    ranges.push_back(make_int4((int)items, c, d, blo / 4));
                    items += bhi / 4 - blo / 4 + 1;
                }
            if (items >= INT32_MAX)
                return fail(ctx, SQMI_ERR_UNSUPPORTED, "long pairs: too many tiles for one launch");
            // longSched: [0] item counter, [1] wait flag, [16 ..) progress, then the ranges.
            const size_t progWords = (size_t)couples * stride;
            const size_t rangeOff = ((16 + progWords) * 4 + 15) / 16 * 16;
            if ((err = ensure(ctx->b.longSched, rangeOff + sizeof(int4) * ranges.size())) != hipSuccess)
                return hipFail(ctx, err, "allocating the long-pair schedule");
            char *sched = static_cast<char *>(ctx->b.longSched.p);
            if ((err = hipMemcpyAsync(sched + rangeOff, ranges.data(), sizeof(int4) * ranges.size(),
                                      hipMemcpyHostToDevice, stream)) != hipSuccess)
                return hipFail(ctx, err, "uploading the long-pair schedule");
            TileArgs t = a;
            t.ranges = reinterpret_cast<const int4 *>(sched + rangeOff);
            t.nRanges = (int32_t)ranges.size();
            t.nItems = (int32_t)items;
            t.queue = reinterpret_cast<uint32_t *>(sched);
            t.error = reinterpret_cast<uint32_t *>(sched) + 1;
            t.progress = reinterpret_cast<uint32_t *>(sched) + 16;
            t.progStride = stride;
            const int grid = std::max(1, ctx->numCUs * occupancyTileFwd16Stream());
            uint32_t *words = reinterpret_cast<uint32_t *>(sched);
            const size_t nWords = 16 + progWords;
            if ((err = op([=] {
                     hipError_t e2 = launchZero32(words, nWords, stream);
                     return e2 != hipSuccess ? e2 : launchTileFwd16Stream(t, grid, stream);
                 })) != hipSuccess)
                return hipFail(ctx, err, "launching long-pair forward");
            ctx->longStreamed = true;
        } else {
        // Every launch sweeps one tile anti-diagonal and ends with a barrier
        // for its whole grid.  Groups of pair couples on separate streams do
        // not wait for each other's slowest tiles, so their launches fill
        // the SIMDs the others leave idle (the ramps of the wavefront and
        // the last waves of the widest anti-diagonals).  Group 0 stays on the
        // call's stream, so that ns groups use ns hardware queues.
        const int couples = (int)((e - b + 1) / 2);
        const int ns = std::min(longStreams(), couples);
        std::vector<TileArgs> sub(ns, a);
        for (int s = 0; s < ns; ++s) {
            const int c0 = couples * s / ns, c1 = couples * (s + 1) / ns;
            sub[s].pairs = a.pairs + 2 * c0;
            sub[s].nPairs = std::min((int32_t)(2 * c1), a.nPairs) - 2 * c0;
        }
        if (ns > 1) {
            if (!ctx->longFork && (err = hipEventCreateWithFlags(&ctx->longFork, hipEventDisableTiming)) != hipSuccess)
                return hipFail(ctx, err, "creating long-pair events");
            for (int s = 1; s < ns; ++s) {
                if (!ctx->longStreams[s] && (err = hipStreamCreate(&ctx->longStreams[s])) != hipSuccess)
                    return hipFail(ctx, err, "creating long-pair streams");
                if (!ctx->longJoin[s] &&
                    (err = hipEventCreateWithFlags(&ctx->longJoin[s], hipEventDisableTiming)) != hipSuccess)
                    return hipFail(ctx, err, "creating long-pair events");
            }
            hipEvent_t fork = ctx->longFork;
            if ((err = op([=] { return hipEventRecord(fork, stream); })) != hipSuccess)
                return hipFail(ctx, err, "long-pair fork");
            for (int s = 1; s < ns; ++s) {
                hipStream_t side = ctx->longStreams[s];
                if ((err = op([=] { return hipStreamWaitEvent(side, fork, 0); })) != hipSuccess)
                    return hipFail(ctx, err, "long-pair fork");
            }
        }
        for (int d = 0; d < maxB + maxT - 1; ++d) {
            for (int s = 0; s < ns; ++s) {
                TileArgs t = sub[s];
                t.diag = d;
                hipStream_t st = s > 0 ? ctx->longStreams[s] : stream;
                if ((err = op([=] { return packed ? launchTileFwd16(t, st) : launchTileFwd32(t, st); })) !=
                    hipSuccess)
                    return hipFail(ctx, err, "launching long-pair forward");
            }
        }
        if (ns > 1) {
            for (int s = 1; s < ns; ++s) {
                hipEvent_t join = ctx->longJoin[s];
                hipStream_t side = ctx->longStreams[s];
                if ((err = op([=] {
                         hipError_t e = hipEventRecord(join, side);
                         return e != hipSuccess ? e : hipStreamWaitEvent(stream, join, 0);
                     })) != hipSuccess)
                    return hipFail(ctx, err, "long-pair join");
            }
        }
        }
        if (P.mode != 0) {
            const TileArgs ta = a;
            if ((err = op([=] { return launchLongResult(ta, stream); })) != hipSuccess)
                return hipFail(ctx, err, "launching long-pair results");
        }
        if (dCounts) {
            a.trace = static_cast<uint2 *>(ctx->b.longTrace.p);
            a.walk = static_cast<int32_t *>(ctx->b.longWalk.p);
            a.counts = dCounts;
            a.ops = rows.ops;
            a.opOff = rows.opOff;
            a.opLen = rows.opLen;
            const TileArgs ta = a;
            if ((err = op([=] { return launchTileWalkInit(ta, stream); })) != hipSuccess)
                return hipFail(ctx, err, "launching long-pair walk init");
            // Each round moves the walk up kRoundBlocks row blocks or left
            // kRoundTiles tiles (or ends it); extra rounds do nothing.
            const int rounds = (maxB + kRoundBlocks - 1) / kRoundBlocks + (maxT + kRoundTiles - 1) / kRoundTiles + 1;
            for (int r = 0; r < rounds; ++r) {
                if ((err = op([=] {
                         hipError_t e = launchTileTrace32(ta, stream);
                         return e != hipSuccess ? e : launchTileWalk32(ta, stream);
                     })) != hipSuccess)
                    return hipFail(ctx, err, "launching long-pair traceback");
            }
        }
        b = e;
    }
    return SQMI_OK;
}

}  // namespace

extern "C" {

int sqmi_abi_version(void) { return SQMI_ABI_VERSION; }

const char *sqmi_status_string(sqmi_status s)
{
    switch (s) {
        case SQMI_OK: return "ok";
        case SQMI_ERR_INVALID_ARGUMENT: return "invalid argument";
        case SQMI_ERR_LENGTH_MISMATCH: return "sequence sets differ in length";
        case SQMI_ERR_UNSUPPORTED: return "unsupported request";
        case SQMI_ERR_NO_DEVICE: return "no usable gfx950 device";
        case SQMI_ERR_DEVICE: return "device error";
        case SQMI_ERR_OUT_OF_MEMORY: return "out of device memory";
    }
    return "unknown status";
}

const char *sqmi_last_error(const sqmi_context *ctx) { return ctx ? ctx->lastError.c_str() : ""; }

sqmi_status sqmi_create(int device, sqmi_context **out)
{
    if (!out)
        return SQMI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || device < 0 || device >= n)
        return SQMI_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return SQMI_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return SQMI_ERR_NO_DEVICE;
    sqmi_context *ctx = new (std::nothrow) sqmi_context();
    if (!ctx)
        return SQMI_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    ctx->numCUs = prop.multiProcessorCount;
    ctx->l2PerCU = (int64_t)prop.l2CacheSize / 32;
    ctx->totalMem = (int64_t)prop.totalGlobalMem;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->ownStream) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        scratchAddress(device, &ctx->scratchPtr) != hipSuccess) {
        if (ctx->ev0)
            (void)hipEventDestroy(ctx->ev0);
        if (ctx->ev1)
            (void)hipEventDestroy(ctx->ev1);
        if (ctx->ownStream)
            (void)hipStreamDestroy(ctx->ownStream);
        delete ctx;
        return SQMI_ERR_DEVICE;
    }
    ctx->stream = ctx->ownStream;
    *out = ctx;
    return SQMI_OK;
}

sqmi_status sqmi_destroy(sqmi_context *ctx)
{
    if (!ctx)
        return SQMI_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    dropReplay(ctx);
    ctx->b.forEach([](DevBuf &b) { release(b); });
    if (ctx->ev0)
        (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1)
        (void)hipEventDestroy(ctx->ev1);
    for (int s = 0; s < sqmi_context::kMaxLongStreams; ++s) {
        if (ctx->longJoin[s])
            (void)hipEventDestroy(ctx->longJoin[s]);
        if (ctx->longStreams[s]) {
            (void)hipStreamSynchronize(ctx->longStreams[s]);
            (void)hipStreamDestroy(ctx->longStreams[s]);
        }
    }
    if (ctx->longFork)
        (void)hipEventDestroy(ctx->longFork);
    if (ctx->ownStream)
        (void)hipStreamDestroy(ctx->ownStream);
    delete ctx;
    return SQMI_OK;
}

sqmi_status sqmi_set_stream(sqmi_context *ctx, void *stream)
{
    if (!ctx)
        return SQMI_ERR_INVALID_ARGUMENT;
    ctx->stream = stream ? reinterpret_cast<hipStream_t>(stream) : ctx->ownStream;
    return SQMI_OK;
}

sqmi_status sqmi_get_stats(const sqmi_context *ctx, sqmi_stats *out)
{
    if (!ctx || !out)
        return SQMI_ERR_INVALID_ARGUMENT;
    *out = ctx->stats;
    return SQMI_OK;
}

}  // extern "C"

// sqmi_align_score and, with ctx->noAlloc, sqmi_align_score_noalloc (every
// ensure() then carves from ctx->arena; the size of the whole request is
// checked before anything is copied or launched, and the row buffers'
// grid is cut to what the arena holds).
static sqmi_status alignScore(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                              const sqmi_scoring *scoring, uint32_t mode, int32_t *scores, uint32_t flags)
{
    const auto tStart = std::chrono::steady_clock::now();
    if (!ctx)
        return SQMI_ERR_INVALID_ARGUMENT;
    ctx->lastError.clear();
    ctx->stats = sqmi_stats{};
    sqmi_status st;
    if ((st = validate(ctx, setH, setV, scoring, mode)) != SQMI_OK || (st = checkForced(ctx, flags)) != SQMI_OK)
        return st;
    const int64_t N = setH->count;
    if (N == 0)
        return SQMI_OK;
    if (!scores)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "scores is null");
    if (hipSetDevice(ctx->device) != hipSuccess)
        return fail(ctx, SQMI_ERR_DEVICE, "hipSetDevice failed");
    hipStream_t stream = ctx->stream;

    // Replay: a call whose plan depends on nothing that differs from the
    // last call's (lengths, scoring, mode, flags, the device addresses the
    // kernels take, the planner's switches) reuses its plan, which is still
    // in the context's device buffers, and relaunches its graph; only host
    // residues and host scores are copied again.
    const bool replayable = graphsEnabled() && !ctx->noAlloc && !(flags & SQMI_PLAIN_PER_PAIR);
    uint64_t key = 0;
    hipError_t err;
    if (replayable) {
        const auto tKey = std::chrono::steady_clock::now();
        key = digest64(setH->offsets, N + 1) ^ mix64(digest64(setV->offsets, N + 1) + 1);
        std::vector<int64_t> scal = {N, mode, flags, scoring->alphabet_size, scoring->gap_open, scoring->gap_extend,
                                     ctx->device};
        for (int k = 0; k < scoring->alphabet_size * scoring->alphabet_size; ++k)
            scal.push_back(scoring->table[k]);
        if (flags & SQMI_RESIDUES_ON_DEVICE) {
            scal.push_back((int64_t)(uintptr_t)setH->residues);
            scal.push_back((int64_t)(uintptr_t)setV->residues);
        }
        if (flags & SQMI_SCORES_ON_DEVICE)
            scal.push_back((int64_t)(uintptr_t)scores);
        for (const char *name : {"SQMI_COOP_MIN_BLOCKS", "SQMI_LONG_ALL", "SQMI_LONG_MIN_SIDE", "SQMI_LONG_STREAMS",
                                 "SQMI_SHAPE_RULE", "SQMI_TRACE_BUDGET_MB", "SQMI_ROW_UNITS", "SQMI_BIN_SHAPES"}) {
            const char *v = std::getenv(name);
            for (const char *c = v ? v : ""; *c; ++c)
                scal.push_back(*c);
            scal.push_back(-1);
        }
        key ^= mix64(digest64(scal.data(), (int64_t)scal.size()) + 2);
        if (ctx->replay.valid && ctx->replay.key == key) {
            const double keyMs = msSince(tKey);
            const uint8_t *rH = setH->residues, *rV = setV->residues;
            const int64_t hTot = setH->offsets[N], vTot = setV->offsets[N];
            if (!(flags & SQMI_RESIDUES_ON_DEVICE) &&
                ((err = hipMemcpyAsync(ctx->b.hRes.p, rH, hTot, hipMemcpyHostToDevice, stream)) != hipSuccess ||
                 (err = hipMemcpyAsync(ctx->b.vRes.p, rV, vTot, hipMemcpyHostToDevice, stream)) != hipSuccess))
                return hipFail(ctx, err, "copying residues");
            if ((err = hipEventRecord(ctx->ev0, stream)) != hipSuccess ||
                (err = hipGraphLaunch(ctx->replay.exec, stream)) != hipSuccess ||
                (err = hipEventRecord(ctx->ev1, stream)) != hipSuccess)
                return hipFail(ctx, err, "launching the call's graph");
            if (!(flags & SQMI_SCORES_ON_DEVICE) &&
                (err = hipMemcpyAsync(scores, ctx->b.scores.p, 4 * N, hipMemcpyDeviceToHost, stream)) != hipSuccess)
                return hipFail(ctx, err, "copying scores");
            if ((err = hipStreamSynchronize(stream)) != hipSuccess)
                return hipFail(ctx, err, "alignment kernels");
            ctx->longStreamed = ctx->replay.longStreamed;
            if ((st = longStreamCheck(ctx)) != SQMI_OK)
                return st;
            ctx->stats = ctx->replay.stats;
            ctx->stats.host_plan_ms = keyMs;
            ctx->stats.replayed = 1;
            float ms = 0;
            ctx->stats.device_ms = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess ? ms : 0.0;
            ctx->stats.total_ms = msSince(tStart);
            return SQMI_OK;
        }
    }
    dropReplay(ctx);

    Planned P;
    const auto tPlan = std::chrono::steady_clock::now();
    plan(*setH, *setV, *scoring, mode, flags, P);
    ctx->stats.host_plan_ms = msSince(tPlan);
    fillStats(P, ctx->stats);
    const ScoringInfo &si = P.si;
    std::vector<Plan16> &plans = P.plans;
    std::vector<int32_t> &emptyIds = P.emptyIds, &ids32 = P.ids32;
    std::vector<int64_t> &scratchOff32 = P.scratchOff32;
    const int64_t scratch32 = P.scratch32;
    const auto &keys16 = P.keys16;
    const uint32_t track = P.track;

    // The no-allocation call: everything but the row buffers has a fixed
    // size; the row buffers need at least one workgroup's worth per shape.
    size_t rowbufBudget = 0;
    if (ctx->noAlloc) {
        if (!P.idsLong.empty())
            return fail(ctx, SQMI_ERR_UNSUPPORTED,
                        "sqmi_align_score_noalloc: pairs on the long-pair route need sqmi_align_score");
        const int64_t A2 = (int64_t)si.alpha * si.alpha, hTot = setH->offsets[N], vTot = setV->offsets[N];
        size_t need = 0;
        if (!(flags & SQMI_RESIDUES_ON_DEVICE))
            need += arenaBytes(hTot + 64) + arenaBytes(vTot + 64);
        need += 2 * arenaBytes(8 * (N + 1));
        if (!(flags & SQMI_SCORES_ON_DEVICE))
            need += arenaBytes(4 * N);
        need += arenaBytes(2 * (si.alpha + 1) * (si.alpha + 1)) + arenaBytes(64) + arenaBytes(4 * A2);
        size_t nOrder = 0, nItems = 0, minRb = 0;
        for (const Plan16 &pl : plans) {
            nOrder += pl.order.size();
            nItems += pl.items.size();
            const Variant16 &v = pl.var;
            const int rc = v.G == 1 ? laneRowCols(pl.maxN, v.R) : pl.maxN + v.G + 2;
            minRb = std::max(minRb, (size_t)wavesPerBlock16() * 64 / (v.G == 1 ? 1 : v.G) * rc * 8);
        }
        if (!plans.empty())
            need += arenaBytes(4 * nOrder) + arenaBytes(sizeof(WorkItem) * nItems) + arenaBytes(64 * plans.size());
        if (!ids32.empty())
            need += arenaBytes(4 * ids32.size()) + arenaBytes(8 * ids32.size()) + arenaBytes((size_t)8 * scratch32);
        if (!emptyIds.empty())
            need += arenaBytes(4 * emptyIds.size());
        const size_t cap = ctx->arena.cap - ctx->arena.used;
        if (need + arenaBytes(minRb) > cap)
            return fail(ctx, SQMI_ERR_OUT_OF_MEMORY,
                        "sqmi_align_score_noalloc: the request needs " + std::to_string(need + arenaBytes(minRb)) +
                            " bytes of device scratch, more than the " + std::to_string(cap) +
                            " of SQMI_NOALLOC_SCRATCH_BYTES; use sqmi_align_score");
        rowbufBudget = (cap - need) & ~(size_t)255;   // holds arenaBytes(minRb), see above
    }

    // ---- device buffers --------------------------------------------------
    Inputs in;
    if ((st = uploadInputs(ctx, setH, setV, scoring, si, scores, flags, in)) != SQMI_OK)
        return st;
    const uint8_t *dH = in.dH, *dV = in.dV;
    int32_t *dScores = in.dScores;
    const int A = si.alpha;
    // Kernel launches wait in `launches` until every upload is enqueued.
    Deferred launches;

    // ---- plain form: one launch per pair (baseline for the batched form) --
    // Host data staged by hipMemcpyAsync must outlive the copies: the plain
    // form's lists live at function scope, like the planner's.
    std::vector<int32_t> plainIds;
    std::vector<WorkItem> plainItems;
    if ((flags & SQMI_PLAIN_PER_PAIR) && (err = hipEventRecord(ctx->ev0, stream)) != hipSuccess)
        return hipFail(ctx, err, "event");
    if (flags & SQMI_PLAIN_PER_PAIR) {
        std::vector<int32_t> &ids16 = plainIds;
        std::vector<WorkItem> &items = plainItems;
        int maxN = 0;
        for (const Key &k : keys16) {
            const int32_t id = k.id;
            const int64_t n = setH->offsets[id + 1] - setH->offsets[id];
            const int64_t m = setV->offsets[id + 1] - setV->offsets[id];
            const Variant16 v = plainVariant(si.dna, track, m);
            const int64_t BR = (int64_t)v.G * v.R;
            items.push_back(WorkItem{(int32_t)ids16.size(), 1, (int32_t)n, (int32_t)((m + 1 + BR - 1) / BR)});
            ids16.push_back(id);
            maxN = std::max<int>(maxN, (int)n);
        }
        const size_t np = ids16.size();
        if ((err = ensure(ctx->b.order, 4 * np)) != hipSuccess ||
            (err = ensure(ctx->b.items, sizeof(WorkItem) * np)) != hipSuccess ||
            (err = ensure(ctx->b.queue, 4 * np + 64)) != hipSuccess ||
            (err = ensure(ctx->b.rowbuf, (size_t)4 * (maxN + 18) * 2 * sizeof(uint32_t))) != hipSuccess)
            return hipFail(ctx, err, "allocating plain work lists");
        if (np && ((err = hipMemcpyAsync(ctx->b.order.p, ids16.data(), 4 * np, hipMemcpyHostToDevice, stream)) !=
                       hipSuccess ||
                   (err = hipMemcpyAsync(ctx->b.items.p, items.data(), sizeof(WorkItem) * np, hipMemcpyHostToDevice,
                                         stream)) != hipSuccess ||
                   (err = hipMemsetAsync(ctx->b.queue.p, 0, 4 * np, stream)) != hipSuccess))
            return hipFail(ctx, err, "uploading plain work lists");
        for (size_t k = 0; k < np; ++k) {
            const int32_t id = ids16[k];
            const Variant16 v = plainVariant(si.dna, track, setV->offsets[id + 1] - setV->offsets[id]);
            FwdArgs a{};
            a.hRes = dH;
            a.vRes = dV;
            a.hOff = static_cast<const int64_t *>(ctx->b.hOff.p);
            a.vOff = static_cast<const int64_t *>(ctx->b.vOff.p);
            a.order = static_cast<const int32_t *>(ctx->b.order.p);
            a.items = static_cast<const WorkItem *>(ctx->b.items.p) + k;
            a.nItems = 1;
            a.queue = static_cast<uint32_t *>(ctx->b.queue.p) + k;
            a.rowbuf = static_cast<uint32_t *>(ctx->b.rowbuf.p);
            a.rowCols = maxN + 18;
            a.scores = dScores;
            a.gapOpen = si.o;
            a.gapExtend = si.e;
            a.mode = mode;
            a.alpha = A;
            a.tTable = static_cast<const int16_t *>(ctx->b.tTable.p);
            a.dnaProfile = static_cast<const uint32_t *>(ctx->b.dnaProf.p);
            a.tNonNeg = tNonNeg(si, mode);
            if ((err = launchFwd16(v, a, 1, stream, 64)) != hipSuccess)
                return hipFail(ctx, err, "launching plain kernel");
        }
        plans.clear();
    }

    // ---- packed int16 kernels -------------------------------------------
    // One launch per shape in use, in order on the stream.  Each plan has its
    // own slices of the work lists and its own queue counters (cleared by one
    // fill kernel), so the uploads of a later plan never touch what an
    // earlier launch still reads; the launches share the row buffers, which a
    // launch only uses while it runs.
    if (!plans.empty()) {
        const int wpb = wavesPerBlock16();
        std::vector<size_t> orderOff(plans.size() + 1, 0), itemOff(plans.size() + 1, 0);
        std::vector<int> grids(plans.size()), rowColsOf(plans.size());
        size_t rbMax = 0;
        for (size_t k = 0; k < plans.size(); ++k) {
            const Plan16 &pl = plans[k];
            const Variant16 &v = pl.var;
            if (!shapeSupported(v))
                return fail(ctx, SQMI_ERR_UNSUPPORTED, "kernel variant not compiled");
            const int occ = v.G == 1 ? occupancyLane16(v) : occupancy16(v);
            const int64_t nIt = (int64_t)pl.items.size();
            grids[k] = std::max(1, (int)std::min<int64_t>((int64_t)ctx->numCUs * occ, (nIt + wpb - 1) / wpb));
            // Row buffers per wave slot: G = 16, one row of (G, F) per group;
            // G = 1, one per lane, column-major across the 64 lanes.
            rowColsOf[k] = v.G == 1 ? laneRowCols(pl.maxN, v.R) : pl.maxN + v.G + 2;
            const size_t perBlock = (size_t)wpb * 64 / (v.G == 1 ? 1 : v.G) * rowColsOf[k] * 8;
            if (ctx->noAlloc)   // the persistent grid shrinks to the scratch that is left
                grids[k] = (int)std::max<size_t>(1, std::min<size_t>((size_t)grids[k], rowbufBudget / perBlock));
            rbMax = std::max(rbMax, (size_t)grids[k] * perBlock);
            orderOff[k + 1] = orderOff[k] + pl.order.size();
            itemOff[k + 1] = itemOff[k] + pl.items.size();
        }
        const size_t nq = 16 * plans.size();   // queue counters: 16 dwords per plan
        if ((err = ensure(ctx->b.order, 4 * orderOff.back())) != hipSuccess ||
            (err = ensure(ctx->b.items, sizeof(WorkItem) * itemOff.back())) != hipSuccess ||
            (err = ensure(ctx->b.queue, 4 * nq)) != hipSuccess || (err = ensure(ctx->b.rowbuf, rbMax)) != hipSuccess)
            return hipFail(ctx, err, "allocating work lists");
        uint32_t *dQueues = static_cast<uint32_t *>(ctx->b.queue.p);
        launches.push_back([=] { return launchZero32(dQueues, nq, stream); });
        for (size_t k = 0; k < plans.size(); ++k) {
            const Plan16 &pl = plans[k];
            const Variant16 &v = pl.var;
            int32_t *dOrder = static_cast<int32_t *>(ctx->b.order.p) + orderOff[k];
            WorkItem *dItems = static_cast<WorkItem *>(ctx->b.items.p) + itemOff[k];
            uint32_t *dQueue = static_cast<uint32_t *>(ctx->b.queue.p) + 16 * k;
            if ((err = hipMemcpyAsync(dOrder, pl.order.data(), 4 * pl.order.size(), hipMemcpyHostToDevice,
                                      stream)) != hipSuccess ||
                (err = hipMemcpyAsync(dItems, pl.items.data(), sizeof(WorkItem) * pl.items.size(),
                                      hipMemcpyHostToDevice, stream)) != hipSuccess)
                return hipFail(ctx, err, "uploading work lists");
            const int64_t nIt = (int64_t)pl.items.size();
            FwdArgs a{};
            a.hRes = dH;
            a.vRes = dV;
            a.hOff = static_cast<const int64_t *>(ctx->b.hOff.p);
            a.vOff = static_cast<const int64_t *>(ctx->b.vOff.p);
            a.order = dOrder;
            a.items = dItems;
            a.nItems = (int32_t)nIt;
            a.queue = dQueue;
            // Items are ordered by row blocks, most first: the leading ones with
            // at least coopMinBlocks() blocks are shared by a workgroup's waves
            // (groups of 16 lanes only).
            {
                const int minBlocks = v.G == 1 ? 0 : coopMinBlocks();
                int32_t nLong = 0;
                while (minBlocks > 0 && nLong < (int32_t)nIt && pl.items[nLong].nblocks >= minBlocks)
                    ++nLong;
                a.nLong = nLong;
                a.queueLong = dQueue + 1;
            }
            a.rowbuf = static_cast<uint32_t *>(ctx->b.rowbuf.p);
            a.rowCols = rowColsOf[k];
            a.scores = dScores;
            a.gapOpen = si.o;
            a.gapExtend = si.e;
            a.mode = mode;
            a.alpha = A;
            a.tTable = static_cast<const int16_t *>(ctx->b.tTable.p);
            a.dnaProfile = static_cast<const uint32_t *>(ctx->b.dnaProf.p);
            a.tNonNeg = tNonNeg(si, mode);
            const int grid = grids[k];
            launches.push_back(
                [=] { return v.G == 1 ? launchLane16(v, a, grid, stream) : launchFwd16(v, a, grid, stream); });
        }
    }

    // ---- int32 pairs -------------------------------------------------------
    if (!ids32.empty()) {
        const size_t n32 = ids32.size();
        if ((err = ensure(ctx->b.ids32, 4 * n32)) != hipSuccess ||
            (err = ensure(ctx->b.scratchOff32, 8 * n32)) != hipSuccess ||
            (err = ensure(ctx->b.scratch32, (size_t)8 * scratch32)) != hipSuccess)
            return hipFail(ctx, err, "allocating int32 lists");
        if ((err = hipMemcpyAsync(ctx->b.ids32.p, ids32.data(), 4 * n32, hipMemcpyHostToDevice, stream)) != hipSuccess ||
            (err = hipMemcpyAsync(ctx->b.scratchOff32.p, scratchOff32.data(), 8 * n32, hipMemcpyHostToDevice, stream)) !=
                hipSuccess)
            return hipFail(ctx, err, "uploading int32 lists");
        Simple32Args a{};
        a.hRes = dH;
        a.vRes = dV;
        a.hOff = static_cast<const int64_t *>(ctx->b.hOff.p);
        a.vOff = static_cast<const int64_t *>(ctx->b.vOff.p);
        a.ids = static_cast<const int32_t *>(ctx->b.ids32.p);
        a.scratchOff = static_cast<const int64_t *>(ctx->b.scratchOff32.p);
        a.nIds = (int32_t)n32;
        a.scratch = static_cast<int32_t *>(ctx->b.scratch32.p);
        a.scores = dScores;
        a.gapOpen = si.o;
        a.gapExtend = si.e;
        a.mode = mode;
        a.alpha = A;
        a.table = static_cast<const int32_t *>(ctx->b.table32.p);
        launches.push_back([=] { return launchSimple32(a, stream); });
    }
    std::vector<LongPair> longPairs;
    if ((st = runLong(ctx, setH, setV, P, in, flags, nullptr, RowsDev{}, longPairs, &launches)) != SQMI_OK)
        return st;
    if (!emptyIds.empty()) {
        if ((err = ensure(ctx->b.emptyIds, 4 * emptyIds.size())) != hipSuccess)
            return hipFail(ctx, err, "allocating empty list");
        if ((err = hipMemcpyAsync(ctx->b.emptyIds.p, emptyIds.data(), 4 * emptyIds.size(), hipMemcpyHostToDevice,
                                  stream)) != hipSuccess)
            return hipFail(ctx, err, "uploading empty list");
        const int ne = (int)emptyIds.size();
        const int32_t *dEmpty = static_cast<const int32_t *>(ctx->b.emptyIds.p);
        launches.push_back([=] { return launchFillScores(dScores, dEmpty, ne, INT32_MIN, stream); });
    }

    // ---- the launches: captured into a graph that an identical next call
    // relaunches, or run directly (plain form, no-allocation calls, or
    // SQMI_GRAPH=0); a capture the runtime refuses falls back to running them.
    if (!(flags & SQMI_PLAIN_PER_PAIR) && (err = hipEventRecord(ctx->ev0, stream)) != hipSuccess)
        return hipFail(ctx, err, "event");
    bool ran = false;
    if (replayable && !launches.empty()) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        err = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (err == hipSuccess) {
            hipError_t capErr = hipSuccess;
            for (auto &f : launches)
                if (capErr == hipSuccess)
                    capErr = f();
            err = hipStreamEndCapture(stream, &graph);
            if (capErr == hipSuccess && err == hipSuccess && graph &&
                hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess &&
                (err = hipGraphLaunch(exec, stream)) == hipSuccess) {
                ran = true;
                ctx->replay.valid = true;
                ctx->replay.key = key;
                ctx->replay.exec = exec;
                exec = nullptr;
            }
            if (graph)
                (void)hipGraphDestroy(graph);
            if (exec)
                (void)hipGraphExecDestroy(exec);
        }
        (void)hipGetLastError();   // a refused capture leaves its error behind
    }
    if (!ran)
        for (auto &f : launches)
            if ((err = f()) != hipSuccess)
                return hipFail(ctx, err, "launching alignment kernels");
    if ((err = hipEventRecord(ctx->ev1, stream)) != hipSuccess)
        return hipFail(ctx, err, "event");
    if (!(flags & SQMI_SCORES_ON_DEVICE)) {
        if ((err = hipMemcpyAsync(scores, dScores, 4 * N, hipMemcpyDeviceToHost, stream)) != hipSuccess)
            return hipFail(ctx, err, "copying scores");
    }
    if ((err = hipStreamSynchronize(stream)) != hipSuccess)
        return hipFail(ctx, err, "alignment kernels");
    if (ctx->replay.valid)
        ctx->replay.longStreamed = ctx->longStreamed;
    if ((st = longStreamCheck(ctx)) != SQMI_OK)
        return st;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess)
        ctx->stats.device_ms = ms;
    ctx->stats.total_ms = msSince(tStart);
    if (ctx->replay.valid)
        ctx->replay.stats = ctx->stats;
    return SQMI_OK;
}

extern "C" {

sqmi_status sqmi_align_score(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                             const sqmi_scoring *scoring, uint32_t mode, int32_t *scores, uint32_t flags)
{
    return alignScore(ctx, setH, setV, scoring, mode, scores, flags);
}

// The context's buffers are set aside for the call and replaced by empty ones
// that carve from the device's scratch array (ctx->scratchPtr, looked up at
// sqmi_create); afterwards the arena pieces are dropped (nothing to free)
// and the context's own buffers come back untouched.
sqmi_status sqmi_align_score_noalloc(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                                     const sqmi_scoring *scoring, uint32_t mode, int32_t *scores, uint32_t flags)
{
    if (!ctx)
        return SQMI_ERR_INVALID_ARGUMENT;
    if (flags & SQMI_PLAIN_PER_PAIR)
        return fail(ctx, SQMI_ERR_UNSUPPORTED, "sqmi_align_score_noalloc: no plain form");
    if (!ctx->scratchPtr)
        return fail(ctx, SQMI_ERR_DEVICE, "sqmi_align_score_noalloc: no device scratch on this context");
    dropReplay(ctx);
    Buffers own;
    std::swap(own, ctx->b);
    ctx->arena = Arena{static_cast<char *>(ctx->scratchPtr), kNoAllocScratchBytes, 0};
    ctx->b.forEach([&](DevBuf &d) { d.arena = &ctx->arena; });
    ctx->noAlloc = true;
    const sqmi_status st = alignScore(ctx, setH, setV, scoring, mode, scores, flags);
    ctx->noAlloc = false;
    std::swap(own, ctx->b);
    ctx->arena = Arena{};
    return st;
}

sqmi_status sqmi_plan_stats(const sqmi_sequences *setH, const sqmi_sequences *setV, const sqmi_scoring *scoring,
                            uint32_t mode, uint32_t flags, sqmi_stats *out)
{
    if (!out)
        return SQMI_ERR_INVALID_ARGUMENT;
    *out = sqmi_stats{};
    sqmi_status st;
    if ((st = validate(nullptr, setH, setV, scoring, mode)) != SQMI_OK || (st = checkForced(nullptr, flags)) != SQMI_OK)
        return st;
    Planned P;
    const auto t0 = std::chrono::steady_clock::now();
    plan(*setH, *setV, *scoring, mode, flags, P);
    out->host_plan_ms = msSince(t0);
    fillStats(P, *out);
    return SQMI_OK;
}

// Traceback: the packed forward pass writes 4 trace bits per cell into a
// buffer of at most traceBudget() bytes, so the work items run in chunks;
// after each chunk one thread per pair walks its path.  Pairs outside the
// int16 window, every pair under SQMI_FORCE_INT32 and every banded pair
// (band given) run the int32 kernel with one trace byte per cell, chunked by
// the same budget.
static sqmi_status alignBackward(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                                 const sqmi_scoring *scoring, uint32_t mode, int32_t *scores, int32_t *counts,
                                 uint8_t *ops, const int64_t *opOffsets, int64_t *opLengths, int64_t *begins,
                                 uint32_t flags, const int64_t *band = nullptr)
{
    const auto tStart = std::chrono::steady_clock::now();
    const bool wantRows = ops != nullptr;
    if (!ctx)
        return SQMI_ERR_INVALID_ARGUMENT;
    dropReplay(ctx);   // this call rewrites the buffers a replay would reuse
    ctx->lastError.clear();
    ctx->stats = sqmi_stats{};
    sqmi_status st;
    if ((st = validate(ctx, setH, setV, scoring, mode)) != SQMI_OK)
        return st;
    if (flags & (SQMI_PLAIN_PER_PAIR | 0x00FFFF00u))
        return fail(ctx, SQMI_ERR_UNSUPPORTED, "sqmi_align_counts: no plain form or forced shapes");
    const int64_t N = setH->count;
    if (N == 0)
        return SQMI_OK;
    if (!scores || !counts)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "scores or counts is null");
    if (wantRows) {
        if (!opOffsets || !opLengths)
            return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "op_offsets or op_lengths is null");
        if (opOffsets[0] < 0)
            return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "op_offsets[0] is negative");
        for (int64_t i = 0; i < N; ++i) {
            const int64_t need = (setH->offsets[i + 1] - setH->offsets[i]) + (setV->offsets[i + 1] - setV->offsets[i]);
            if (opOffsets[i + 1] - opOffsets[i] < need)
                return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "op segment " + std::to_string(i) + " holds fewer than |H| + |V| bytes");
        }
    }
    if (hipSetDevice(ctx->device) != hipSuccess)
        return fail(ctx, SQMI_ERR_DEVICE, "hipSetDevice failed");
    hipStream_t stream = ctx->stream;

    Planned P;
    Inputs in;
    // Host residue sets of 4 MB and more are copied in on a second host
    // thread while this one plans (the planner reads only the offsets; a
    // pageable copy holds its thread until the data is staged).  On 65,536
    // banded pairs of 150 bp: 3.9-4.2 ms per rows call against 4.4-4.7.
    hipError_t upErr = hipSuccess;
    JoinedThread up;
    if (!(flags & SQMI_RESIDUES_ON_DEVICE) && setH->offsets[N] + setV->offsets[N] >= ((int64_t)4 << 20))
        up.t = std::thread([&] {
            upErr = hipSetDevice(ctx->device);
            if (upErr == hipSuccess)
                upErr = uploadResidues(ctx, setH, setV, in);
        });
    const auto tPlan = std::chrono::steady_clock::now();
    // Banded pairs run on the int32 kernel, band-masked.
    plan(*setH, *setV, *scoring, mode, band ? (flags | SQMI_FORCE_INT32) : flags, P, true, band != nullptr);
    ctx->stats.host_plan_ms = msSince(tPlan);
    up.join();
    if (upErr != hipSuccess)
        return hipFail(ctx, upErr, "copying residues");
    fillStats(P, ctx->stats);
    const ScoringInfo &si = P.si;
    const int A = si.alpha, S = A * A + 2;
    const int64_t budget = traceBudget(ctx);

    // Host-side chunking of the int16 items and of the int32 pairs.
    const Plan16 *pl = P.plans.empty() ? nullptr : &P.plans[0];
    std::vector<int64_t> itemOff;           // per item: first record within its chunk (units of 64 records)
    std::vector<size_t> chunkEnd;           // item index after each chunk
    std::vector<TraceLoc> locs;             // per pair in plan order
    int64_t traceRecs = 0;                  // largest chunk, in units of 64 records
    if (pl) {
        const int64_t recBytes = 64 * 16;
        itemOff.resize(pl->items.size());
        locs.resize(pl->order.size());
        int64_t run = 0;
        for (size_t k = 0; k < pl->items.size(); ++k) {
            const WorkItem &w = pl->items[k];
            const int steps = traceSteps(w.maxN);
            const int64_t recs = (int64_t)w.nblocks * steps;
            if (run > 0 && (run + recs) * recBytes > budget) {
                chunkEnd.push_back(k);
                run = 0;
            }
            itemOff[k] = run;
            for (int x = 0; x < w.count; ++x) {
                const int32_t pid = pl->order[w.first + x];
                const int64_t m = setV->offsets[pid + 1] - setV->offsets[pid];
                locs[w.first + x] = TraceLoc{run, steps, (int32_t)((int64_t)w.nblocks * kTraceRows - 1 - m),
                                             (x / 2) * kTraceLanes, x % 2};
            }
            run += recs;
            traceRecs = std::max(traceRecs, run);
        }
        chunkEnd.push_back(pl->items.size());
    }
    std::vector<size_t> chunk32End;         // int32 pairs per chunk
    std::vector<int64_t> off32(P.ids32.size());
    int64_t scratch32 = 0;
    {
        int64_t run = 0;
        for (size_t k = 0; k < P.ids32.size(); ++k) {
            const int64_t need = (k + 1 < P.ids32.size() ? P.scratchOff32[k + 1] : P.scratch32) - P.scratchOff32[k];
            if (need > std::max<int64_t>(budget, (int64_t)4 << 30))
                return fail(ctx, SQMI_ERR_UNSUPPORTED, "sqmi_align_counts: pair too large for the int32 traceback");
            if (run > 0 && run + need > budget) {
                chunk32End.push_back(k);
                run = 0;
            }
            off32[k] = run;
            run += need;
            scratch32 = std::max(scratch32, run);
        }
        if (!P.ids32.empty())
            chunk32End.push_back(P.ids32.size());
    }

    // ---- device buffers --------------------------------------------------
    hipError_t err;
    if ((st = uploadInputs(ctx, setH, setV, scoring, si, scores, flags, in)) != SQMI_OK)
        return st;
    int32_t *dCounts = counts;
    if (!(flags & SQMI_SCORES_ON_DEVICE)) {
        if ((err = ensure(ctx->b.counts, (size_t)4 * S * N)) != hipSuccess)
            return hipFail(ctx, err, "allocating counts");
        dCounts = static_cast<int32_t *>(ctx->b.counts.p);
    }
    if ((err = hipMemsetAsync(dCounts, 0, (size_t)4 * S * N, stream)) != hipSuccess)
        return hipFail(ctx, err, "clearing counts");
    RowsDev rows;
    if (wantRows) {
        // Offsets always come from the host; ops and lengths live where the
        // scores do.
        const int64_t opBytes = opOffsets[N];
        if ((err = ensure(ctx->b.opOff, 8 * (N + 1))) != hipSuccess)
            return hipFail(ctx, err, "allocating op offsets");
        if ((err = hipMemcpyAsync(ctx->b.opOff.p, opOffsets, 8 * (N + 1), hipMemcpyHostToDevice, stream)) != hipSuccess)
            return hipFail(ctx, err, "uploading op offsets");
        rows.opOff = static_cast<const int64_t *>(ctx->b.opOff.p);
        if (flags & SQMI_SCORES_ON_DEVICE) {
            rows.ops = ops;
            rows.opLen = opLengths;
        } else {
            if ((err = ensure(ctx->b.ops, (size_t)std::max<int64_t>(opBytes, 1))) != hipSuccess ||
                (err = ensure(ctx->b.opLen, 8 * N)) != hipSuccess)
                return hipFail(ctx, err, "allocating rows");
            rows.ops = static_cast<uint8_t *>(ctx->b.ops.p);
            rows.opLen = static_cast<int64_t *>(ctx->b.opLen.p);
        }
        if ((err = hipMemsetAsync(rows.opLen, 0, 8 * N, stream)) != hipSuccess)
            return hipFail(ctx, err, "clearing op lengths");
        if (begins) {
            if (flags & SQMI_SCORES_ON_DEVICE) {
                rows.begins = begins;
            } else {
                if ((err = ensure(ctx->b.begins, 16 * N)) != hipSuccess)
                    return hipFail(ctx, err, "allocating begins");
                rows.begins = static_cast<int64_t *>(ctx->b.begins.p);
            }
            if ((err = hipMemsetAsync(rows.begins, 0, 16 * N, stream)) != hipSuccess)
                return hipFail(ctx, err, "clearing begins");
        }
    }
    if ((err = hipEventRecord(ctx->ev0, stream)) != hipSuccess)
        return hipFail(ctx, err, "event");

    if (pl) {
        const size_t nIt = pl->items.size(), np = pl->order.size();
        const int wpb = wavesPerBlock16();
        const int occ = occupancyTrace16(si.dna);
        const int rowCols = pl->maxN + kTraceLanes + 2;
        int64_t maxChunkItems = 0;
        for (size_t c = 0, b = 0; c < chunkEnd.size(); b = chunkEnd[c], ++c)
            maxChunkItems = std::max<int64_t>(maxChunkItems, (int64_t)(chunkEnd[c] - b));
        const int grid = std::max(1, (int)std::min<int64_t>((int64_t)ctx->numCUs * occ, (maxChunkItems + wpb - 1) / wpb));
        const size_t rbBytes = (size_t)grid * wpb * (64 / kTraceLanes) * rowCols * 2 * sizeof(uint32_t);
        if ((err = ensure(ctx->b.order, 4 * np)) != hipSuccess ||
            (err = ensure(ctx->b.items, sizeof(WorkItem) * nIt)) != hipSuccess ||
            (err = ensure(ctx->b.traceOff, 8 * nIt)) != hipSuccess ||
            (err = ensure(ctx->b.locs, sizeof(TraceLoc) * np)) != hipSuccess ||
            (err = ensure(ctx->b.queue, 4 * chunkEnd.size() + 64)) != hipSuccess ||
            (err = ensure(ctx->b.rowbuf, rbBytes)) != hipSuccess ||
            (err = ensure(ctx->b.endEF, (size_t)8 * N)) != hipSuccess ||
            (P.track != kTrackEnd && (err = ensure(ctx->b.endIJ, (size_t)8 * N)) != hipSuccess) ||
            (err = ensure(ctx->b.trace, (size_t)traceRecs * 64 * 16)) != hipSuccess)
            return hipFail(ctx, err, "allocating traceback buffers");
        int32_t *dEndIJ = P.track != kTrackEnd ? static_cast<int32_t *>(ctx->b.endIJ.p) : nullptr;
        if ((err = hipMemcpyAsync(ctx->b.order.p, pl->order.data(), 4 * np, hipMemcpyHostToDevice, stream)) !=
                hipSuccess ||
            (err = hipMemcpyAsync(ctx->b.items.p, pl->items.data(), sizeof(WorkItem) * nIt, hipMemcpyHostToDevice,
                                  stream)) != hipSuccess ||
            (err = hipMemcpyAsync(ctx->b.traceOff.p, itemOff.data(), 8 * nIt, hipMemcpyHostToDevice, stream)) !=
                hipSuccess ||
            (err = hipMemcpyAsync(ctx->b.locs.p, locs.data(), sizeof(TraceLoc) * np, hipMemcpyHostToDevice, stream)) !=
                hipSuccess ||
            (err = hipMemsetAsync(ctx->b.queue.p, 0, 4 * chunkEnd.size() + 64, stream)) != hipSuccess)
            return hipFail(ctx, err, "uploading traceback work lists");
        const WorkItem *dItems = static_cast<const WorkItem *>(ctx->b.items.p);
        const int32_t *dOrder = static_cast<const int32_t *>(ctx->b.order.p);
        size_t b = 0;
        for (size_t c = 0; c < chunkEnd.size(); ++c) {
            const size_t e = chunkEnd[c];
            FwdArgs a{};
            a.hRes = in.dH;
            a.vRes = in.dV;
            a.hOff = static_cast<const int64_t *>(ctx->b.hOff.p);
            a.vOff = static_cast<const int64_t *>(ctx->b.vOff.p);
            a.order = dOrder;
            a.items = dItems + b;
            a.nItems = (int32_t)(e - b);
            a.queue = static_cast<uint32_t *>(ctx->b.queue.p) + c;
            a.rowbuf = static_cast<uint32_t *>(ctx->b.rowbuf.p);
            a.rowCols = rowCols;
            a.scores = in.dScores;
            a.gapOpen = si.o;
            a.gapExtend = si.e;
            a.mode = mode;
            a.alpha = A;
            a.tTable = static_cast<const int16_t *>(ctx->b.tTable.p);
            a.dnaProfile = static_cast<const uint32_t *>(ctx->b.dnaProf.p);
            a.trace = static_cast<uint4 *>(ctx->b.trace.p);
            a.traceOff = static_cast<const int64_t *>(ctx->b.traceOff.p) + b;
            a.endEF = static_cast<int32_t *>(ctx->b.endEF.p);
            a.endIJ = dEndIJ;
            const int cg = std::max(1, std::min(grid, (int)((e - b + wpb - 1) / wpb)));
            if ((err = launchFwdTrace16(si.dna, P.track, a, cg, stream)) != hipSuccess)
                return hipFail(ctx, err, "launching trace kernel");
            const int32_t p0 = pl->items[b].first;
            const int32_t p1 = pl->items[e - 1].first + pl->items[e - 1].count;
            WalkArgs wa{};
            wa.hRes = in.dH;
            wa.vRes = in.dV;
            wa.hOff = a.hOff;
            wa.vOff = a.vOff;
            wa.ids = dOrder + p0;
            wa.locs = static_cast<const TraceLoc *>(ctx->b.locs.p) + p0;
            wa.nPairs = p1 - p0;
            wa.trace = a.trace;
            wa.scores = in.dScores;
            wa.endEF = a.endEF;
            wa.counts = dCounts;
            wa.alpha = A;
            wa.linear = si.o == si.e;
            wa.endIJ = dEndIJ;
            wa.local = P.local ? 1 : 0;
            wa.table = static_cast<const int32_t *>(ctx->b.table32.p);
            wa.gapOpen = si.o;
            wa.gapExtend = si.e;
            wa.begins = rows.begins;
            wa.ops = rows.ops;
            wa.opOff = rows.opOff;
            wa.opLen = rows.opLen;
            if ((err = launchWalk16(wa, stream)) != hipSuccess)
                return hipFail(ctx, err, "launching walk kernel");
            b = e;
        }
    }
    // Banded pairs whose band, clipped to the matrix, has at most 64
    // diagonals: the lane-per-diagonal trace kernel (S trace bytes per row,
    // S = the widest band rounded up to a power of two) and a walk, in
    // chunks whose trace bytes fit the trace budget (offsets restart at 0 in
    // every chunk; the chunks run one after another on the call's stream).
    // SQMI_BANDED_TRACE_INT32=1 keeps them on the int32 kernel.
    int bandLanes = 0;
    std::vector<int64_t> bandOff;
    std::vector<size_t> bandChunkEnd;
    int64_t bandBytes = 0;
    if (band && !P.ids32.empty() && !envFlag("SQMI_BANDED_TRACE_INT32")) {
        int64_t wmax = 1;
        for (int32_t id : P.ids32) {
            const int64_t n = setH->offsets[id + 1] - setH->offsets[id], m = setV->offsets[id + 1] - setV->offsets[id];
            wmax = std::max(wmax, std::min(band[1], n) - std::max(band[0], -m) + 1);
        }
        if (wmax <= 64) {
            bandLanes = 1;
            while (bandLanes < wmax)
                bandLanes <<= 1;
            bandOff.resize(P.ids32.size());
            int64_t run = 0;
            for (size_t k = 0; k < P.ids32.size(); ++k) {
                const int32_t id = P.ids32[k];
                const int64_t need = (setV->offsets[id + 1] - setV->offsets[id]) * bandLanes;
                if (run > 0 && run + need > budget) {
                    bandChunkEnd.push_back(k);
                    run = 0;
                }
                bandOff[k] = run;
                run += need;
                bandBytes = std::max(bandBytes, run);
            }
            bandChunkEnd.push_back(P.ids32.size());
        }
    }
    if (bandLanes > 0) {
        const size_t n32 = P.ids32.size();
        if ((err = ensure(ctx->b.ids32, 4 * n32)) != hipSuccess || (err = ensure(ctx->b.scratchOff32, 8 * n32)) != hipSuccess ||
            (err = ensure(ctx->b.scratch32, (size_t)std::max<int64_t>(bandBytes, 1))) != hipSuccess ||
            (err = ensure(ctx->b.endEF, (size_t)8 * N)) != hipSuccess || (err = ensure(ctx->b.endIJ, (size_t)8 * N)) != hipSuccess)
            return hipFail(ctx, err, "allocating banded traceback buffers");
        if ((err = hipMemcpyAsync(ctx->b.ids32.p, P.ids32.data(), 4 * n32, hipMemcpyHostToDevice, stream)) != hipSuccess ||
            (err = hipMemcpyAsync(ctx->b.scratchOff32.p, bandOff.data(), 8 * n32, hipMemcpyHostToDevice, stream)) !=
                hipSuccess)
            return hipFail(ctx, err, "uploading banded traceback lists");
        size_t cb = 0;
        for (size_t ce : bandChunkEnd) {
        BandedArgs ba{};
        ba.hRes = in.dH;
        ba.vRes = in.dV;
        ba.hOff = static_cast<const int64_t *>(ctx->b.hOff.p);
        ba.vOff = static_cast<const int64_t *>(ctx->b.vOff.p);
        ba.ids = static_cast<const int32_t *>(ctx->b.ids32.p) + cb;
        ba.nIds = (int32_t)(ce - cb);
        ba.scores = in.dScores;
        ba.gapOpen = si.o;
        ba.gapExtend = si.e;
        ba.mode = mode;
        ba.alpha = A;
        ba.table = static_cast<const int32_t *>(ctx->b.table32.p);
        ba.lower = band[0];
        ba.upper = band[1];
        ba.trace = static_cast<uint8_t *>(ctx->b.scratch32.p);
        ba.traceOff = static_cast<const int64_t *>(ctx->b.scratchOff32.p) + cb;
        ba.endIJ = static_cast<int32_t *>(ctx->b.endIJ.p);
        ba.endEF = static_cast<int32_t *>(ctx->b.endEF.p);
        if ((err = launchBandedTrace(ba, bandLanes, stream)) != hipSuccess)
            return hipFail(ctx, err, "launching banded trace kernel");
        BandedWalkArgs wa{};
        wa.hRes = in.dH;
        wa.vRes = in.dV;
        wa.hOff = ba.hOff;
        wa.vOff = ba.vOff;
        wa.ids = ba.ids;
        wa.nIds = ba.nIds;
        wa.segLanes = bandLanes;
        wa.lower = band[0];
        wa.upper = band[1];
        wa.trace = ba.trace;
        wa.traceOff = ba.traceOff;
        wa.scores = in.dScores;
        wa.endIJ = ba.endIJ;
        wa.endEF = ba.endEF;
        wa.counts = dCounts;
        wa.alpha = A;
        wa.local = (mode & SQMI_LOCAL) ? 1 : 0;
        wa.linear = si.o == si.e;
        wa.ops = rows.ops;
        wa.opOff = rows.opOff;
        wa.opLen = rows.opLen;
        wa.begins = rows.begins;
        if ((err = launchBandedWalk(wa, stream)) != hipSuccess)
            return hipFail(ctx, err, "launching banded walk kernel");
        cb = ce;
        }
    } else if (!P.ids32.empty()) {
        const size_t n32 = P.ids32.size();
        if ((err = ensure(ctx->b.ids32, 4 * n32)) != hipSuccess || (err = ensure(ctx->b.scratchOff32, 8 * n32)) != hipSuccess ||
            (err = ensure(ctx->b.scratch32, (size_t)scratch32)) != hipSuccess)
            return hipFail(ctx, err, "allocating int32 traceback buffers");
        if ((err = hipMemcpyAsync(ctx->b.ids32.p, P.ids32.data(), 4 * n32, hipMemcpyHostToDevice, stream)) != hipSuccess ||
            (err = hipMemcpyAsync(ctx->b.scratchOff32.p, off32.data(), 8 * n32, hipMemcpyHostToDevice, stream)) !=
                hipSuccess)
            return hipFail(ctx, err, "uploading int32 traceback lists");
        size_t b = 0;
        for (size_t e : chunk32End) {
            Counts32Args a{};
            a.hRes = in.dH;
            a.vRes = in.dV;
            a.hOff = static_cast<const int64_t *>(ctx->b.hOff.p);
            a.vOff = static_cast<const int64_t *>(ctx->b.vOff.p);
            a.ids = static_cast<const int32_t *>(ctx->b.ids32.p) + b;
            a.scratchOff = static_cast<const int64_t *>(ctx->b.scratchOff32.p) + b;
            a.nIds = (int32_t)(e - b);
            a.scratch = static_cast<uint8_t *>(ctx->b.scratch32.p);
            a.scores = in.dScores;
            a.counts = dCounts;
            a.gapOpen = si.o;
            a.gapExtend = si.e;
            a.mode = mode;
            a.alpha = A;
            a.table = static_cast<const int32_t *>(ctx->b.table32.p);
            a.ops = rows.ops;
            a.opOff = rows.opOff;
            a.opLen = rows.opLen;
            a.begins = rows.begins;
            if (band) {
                a.banded = 1;
                a.lower = band[0];
                a.upper = band[1];
            }
            if ((err = launchCounts32(a, stream)) != hipSuccess)
                return hipFail(ctx, err, "launching int32 traceback kernel");
            b = e;
        }
    }
    std::vector<LongPair> longPairs;
    if ((st = runLong(ctx, setH, setV, P, in, flags, dCounts, rows, longPairs)) != SQMI_OK)
        return st;
    if (!P.emptyIds.empty()) {
        if ((err = ensure(ctx->b.emptyIds, 4 * P.emptyIds.size())) != hipSuccess)
            return hipFail(ctx, err, "allocating empty list");
        if ((err = hipMemcpyAsync(ctx->b.emptyIds.p, P.emptyIds.data(), 4 * P.emptyIds.size(), hipMemcpyHostToDevice,
                                  stream)) != hipSuccess)
            return hipFail(ctx, err, "uploading empty list");
        if ((err = launchFillScores(in.dScores, static_cast<const int32_t *>(ctx->b.emptyIds.p), (int)P.emptyIds.size(),
                                    INT32_MIN, stream)) != hipSuccess)
            return hipFail(ctx, err, "launching fill kernel");
    }
    const uint8_t *rowsOut = rows.ops;
    if (wantRows && (flags & SQMI_ROWS_PACKED)) {
        // Rows back to back in pair order from ops + op_offsets[0].
        const int64_t opBytes = opOffsets[N];
        if ((err = ensure(ctx->b.opsPacked, (size_t)std::max<int64_t>(opBytes, 1))) != hipSuccess ||
            (err = ensure(ctx->b.opStart, 8 * (N + 1))) != hipSuccess)
            return hipFail(ctx, err, "allocating packed rows");
        uint8_t *packed = static_cast<uint8_t *>(ctx->b.opsPacked.p);
        if ((err = launchPackRows(rows.ops, rows.opOff, rows.opLen, static_cast<int64_t *>(ctx->b.opStart.p), packed,
                                  N, stream)) != hipSuccess)
            return hipFail(ctx, err, "launching row packing");
        if ((flags & SQMI_SCORES_ON_DEVICE) &&
            (err = hipMemcpyAsync(ops + opOffsets[0], packed + opOffsets[0], (size_t)(opBytes - opOffsets[0]),
                                  hipMemcpyDeviceToDevice, stream)) != hipSuccess)
            return hipFail(ctx, err, "copying packed rows");
        rowsOut = packed;
    }
    if ((err = hipEventRecord(ctx->ev1, stream)) != hipSuccess)
        return hipFail(ctx, err, "event");
    if (!(flags & SQMI_SCORES_ON_DEVICE)) {
        if ((err = hipMemcpyAsync(scores, in.dScores, 4 * N, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
            (err = hipMemcpyAsync(counts, dCounts, (size_t)4 * S * N, hipMemcpyDeviceToHost, stream)) != hipSuccess)
            return hipFail(ctx, err, "copying results");
        if (wantRows &&
            ((err = hipMemcpyAsync(opLengths, rows.opLen, 8 * N, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
             (begins && (err = hipMemcpyAsync(begins, rows.begins, 16 * N, hipMemcpyDeviceToHost, stream)) !=
                            hipSuccess)))
            return hipFail(ctx, err, "copying rows");
        if (wantRows) {
            // Packed rows fill only the first sum(op_lengths) bytes: read
            // that total (the scan's last entry) and copy no more.
            int64_t opsBytes = opOffsets[N] - opOffsets[0];
            if (flags & SQMI_ROWS_PACKED) {
                int64_t total = 0;
                if ((err = hipMemcpyAsync(&total, static_cast<const int64_t *>(ctx->b.opStart.p) + N, 8,
                                          hipMemcpyDeviceToHost, stream)) != hipSuccess ||
                    (err = hipStreamSynchronize(stream)) != hipSuccess)
                    return hipFail(ctx, err, "reading the packed rows' length");
                opsBytes = std::min(opsBytes, std::max<int64_t>(total, 0));
            }
            if (opsBytes > 0 &&
                (err = hipMemcpyAsync(ops + opOffsets[0], rowsOut + opOffsets[0], (size_t)opsBytes,
                                      hipMemcpyDeviceToHost, stream)) != hipSuccess)
                return hipFail(ctx, err, "copying rows");
        }
    }
    if ((err = hipStreamSynchronize(stream)) != hipSuccess)
        return hipFail(ctx, err, "traceback kernels");
    if ((st = longStreamCheck(ctx)) != SQMI_OK)
        return st;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess)
        ctx->stats.device_ms = ms;
    ctx->stats.total_ms = msSince(tStart);
    return SQMI_OK;
}

sqmi_status sqmi_align_counts(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                              const sqmi_scoring *scoring, uint32_t mode, int32_t *scores, int32_t *counts,
                              uint32_t flags)
{
    return alignBackward(ctx, setH, setV, scoring, mode, scores, counts, nullptr, nullptr, nullptr, nullptr, flags);
}

sqmi_status sqmi_align_rows(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                            const sqmi_scoring *scoring, uint32_t mode, int32_t *scores, int32_t *counts,
                            uint8_t *ops, const int64_t *op_offsets, int64_t *op_lengths, uint32_t flags)
{
    if (!ops)
        return ctx ? fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "ops is null") : SQMI_ERR_INVALID_ARGUMENT;
    if (mode & SQMI_LOCAL)
        return ctx ? fail(ctx, SQMI_ERR_UNSUPPORTED,
                          "sqmi_align_rows: global modes only; local rows come from sqmi_align_local_rows")
                   : SQMI_ERR_UNSUPPORTED;
    return alignBackward(ctx, setH, setV, scoring, mode, scores, counts, ops, op_offsets, op_lengths, nullptr, flags);
}

// Bands the library computes: the band holds the diagonals of (0, 0) and of
// (m, n) with at least one more diagonal beyond each, and H has at least two
// residues.  There the reference's banded code computes the banded
// recurrence.  Elsewhere it takes special paths of its own
// (dp_algorithm_impl.h:595-700: one-diagonal bands, single-column matrices,
// bands whose edge lies on a corner diagonal) whose results differ from that
// recurrence under free end gaps or for |H| = 1, so such requests are
// refused.  Without free ends the reference itself accepts only bands that
// hold both corners (_checkBandProperties, dp_algorithm_impl.h:192-232).
//
// Local alignment (every end free): |H| >= 2, at least two diagonals, and
// the band reaches past the corner cells (0, n - 1) and (m - 1, 0): bands
// that only touch a corner, or have one diagonal, again take the special
// paths.
static bool bandSupported(int64_t n, int64_t m, int64_t lower, int64_t upper, bool local)
{
    if (local)
        return n >= 2 && lower < upper && lower <= n - 2 && upper >= -(m - 2) &&
               std::min<int64_t>(upper, n) - std::max<int64_t>(lower, -m) >= 1;
    return n >= 2 && lower < std::min<int64_t>(0, n - m) && upper > std::max<int64_t>(0, n - m);
}

sqmi_status sqmi_align_score_banded(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                                    const sqmi_scoring *scoring, uint32_t mode, int64_t lower_diag,
                                    int64_t upper_diag, int32_t *scores, uint32_t flags)
{
    const auto tStart = std::chrono::steady_clock::now();
    if (!ctx)
        return SQMI_ERR_INVALID_ARGUMENT;
    dropReplay(ctx);   // this call rewrites the buffers a replay would reuse
    ctx->lastError.clear();
    ctx->stats = sqmi_stats{};
    sqmi_status st;
    if ((st = validate(ctx, setH, setV, scoring, mode)) != SQMI_OK)
        return st;
    const bool local = (mode & SQMI_LOCAL) != 0;
    if (flags & ~(uint32_t)(SQMI_RESIDUES_ON_DEVICE | SQMI_SCORES_ON_DEVICE | SQMI_FORCE_INT32))
        return fail(ctx, SQMI_ERR_UNSUPPORTED,
                    "sqmi_align_score_banded: flags other than the *_ON_DEVICE ones and SQMI_FORCE_INT32");
    if (lower_diag > upper_diag)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "lower_diag > upper_diag");
    const int64_t N = setH->count;
    if (N == 0)
        return SQMI_OK;
    if (!scores)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "scores is null");
    if (hipSetDevice(ctx->device) != hipSuccess)
        return fail(ctx, SQMI_ERR_DEVICE, "hipSetDevice failed");
    hipStream_t stream = ctx->stream;

    // Large host residue sets are copied in on a second host thread while
    // this one routes the pairs (as in the rows calls).
    Inputs in;
    hipError_t upErr = hipSuccess;
    JoinedThread up;
    if (!(flags & SQMI_RESIDUES_ON_DEVICE) && setH->offsets[N] + setV->offsets[N] >= ((int64_t)4 << 20))
        up.t = std::thread([&] {
            upErr = hipSetDevice(ctx->device);
            if (upErr == hipSuccess)
                upErr = uploadResidues(ctx, setH, setV, in);
        });

    // Route: empty pairs (INT32_MIN), the packed int16 kernel (global
    // alignment, values inside the int16 window, at least 8 diagonals past
    // the first), narrow bands (<= 64 diagonals), wide bands.
    const auto tPlan = std::chrono::steady_clock::now();
    ScoringInfo si{};
    si.alpha = scoring->alphabet_size;
    si.o = scoring->gap_open;
    si.e = scoring->gap_extend;
    si.sMax = INT_MIN;
    si.sMin = INT_MAX;
    for (int k = 0; k < si.alpha * si.alpha; ++k) {
        si.sMax = std::max(si.sMax, scoring->table[k]);
        si.sMin = std::min(si.sMin, scoring->table[k]);
    }
    si.dna = si.alpha <= 4 && si.sMin - si.o >= 0 && si.sMax - si.o <= 255;
    const int64_t W1 = upper_diag - lower_diag;
    // Rows per block: more rows share a block's set-up and boundary columns,
    // fewer compute fewer cells outside the band (W - 1 + R columns per
    // block for W band cells per row).  Band [-16, 16] on 1 M pairs of 150
    // bp: R = 12 1.64 ms, 16 1.74 ms; [-100, 100] on 65,536 pairs of 1 kbp:
    // R = 16 4.66 ms, 12 5.10 ms.
    int bandR = W1 >= 64 ? 16 : W1 >= 24 ? 12 : 8;
    if (const char *env = std::getenv("SQMI_BAND_ROWS")) {   // tuning knob: 8, 12 or 16 rows per block
        const int r = std::atoi(env);
        bandR = r == 8 ? 8 : (r == 12 || r == 16) && W1 >= r ? r : bandR;
    }
    const bool packedOk = mode == 0 && !(flags & SQMI_FORCE_INT32) && W1 >= bandR && si.o <= 0 && si.e <= 0 &&
                          si.o >= -1024 && si.e >= -1024 && si.sMax <= 1024 && si.sMin >= -1024 && si.alpha + 1 <= 33 &&
                          lower_diag > -30000 && upper_diag < 30000;
    std::vector<int32_t> narrow, wide, rejected, packed;
    int64_t wNarrow = 1, wWide = 0, cells = 0;
    for (int64_t i = 0; i < N; ++i) {
        const int64_t n = setH->offsets[i + 1] - setH->offsets[i], m = setV->offsets[i + 1] - setV->offsets[i];
        if (n == 0 || m == 0) {
            rejected.push_back((int32_t)i);
            continue;
        }
        if (!bandSupported(n, m, lower_diag, upper_diag, local))
            return fail(ctx, SQMI_ERR_UNSUPPORTED,
                        "sqmi_align_score_banded: pair " + std::to_string(i) +
                            (local ? ": needs |H| >= 2, two diagonals, lower_diag <= |H| - 2 and upper_diag >= 2 - |V|"
                                   : ": needs |H| >= 2 and lower_diag < min(0, |H| - |V|) < max(0, |H| - |V|) "
                                     "< upper_diag"));
        cells += n * m;
        if (packedOk && fitsInt16(si, n, m, false)) {
            packed.push_back((int32_t)i);
            continue;
        }
        const int64_t w = std::min<int64_t>(upper_diag, n) - std::max<int64_t>(lower_diag, -m) + 1;
        if (w <= 64) {
            narrow.push_back((int32_t)i);
            wNarrow = std::max(wNarrow, w);
        } else {
            wide.push_back((int32_t)i);
            wWide = std::max(wWide, w);
        }
    }
    int seg = 1;
    while (seg < wNarrow)
        seg <<= 1;
    // Packed items: up to 128 pairs of one |V| (counting sort by |V|).
    std::vector<int32_t> bandOrder;
    std::vector<WorkItem> bandItems;
    int bandMaxN = 1;
    if (!packed.empty()) {
        int maxM = 0;
        for (int32_t id : packed)
            maxM = std::max(maxM, (int)(setV->offsets[id + 1] - setV->offsets[id]));
        std::vector<int32_t> start((size_t)maxM + 2, 0);
        for (int32_t id : packed)
            ++start[(size_t)(setV->offsets[id + 1] - setV->offsets[id]) + 1];
        for (int mm = 0; mm <= maxM; ++mm)
            start[(size_t)mm + 1] += start[(size_t)mm];
        bandOrder.resize(packed.size());
        {
            std::vector<int32_t> pos(start.begin(), start.end() - 1);
            for (int32_t id : packed)
                bandOrder[(size_t)pos[(size_t)(setV->offsets[id + 1] - setV->offsets[id])]++] = id;
        }
        for (int mm = 1; mm <= maxM; ++mm) {
            for (int32_t f = start[(size_t)mm]; f < start[(size_t)mm + 1]; f += 128) {
                const int32_t cnt = std::min<int32_t>(128, start[(size_t)mm + 1] - f);
                int mxN = 1;
                for (int32_t q = f; q < f + cnt; ++q)
                    mxN = std::max(mxN, (int)(setH->offsets[bandOrder[q] + 1] - setH->offsets[bandOrder[q]]));
                bandMaxN = std::max(bandMaxN, mxN);
                bandItems.push_back(WorkItem{f, cnt | (mm << 16), mxN, (mm + bandR) / bandR});
            }
        }
    }
    ctx->stats.pairs = N;
    ctx->stats.cells = cells;
    ctx->stats.pairs_int32 = (int64_t)(narrow.size() + wide.size());
    ctx->stats.pairs_int16 = (int64_t)packed.size();
    ctx->stats.host_plan_ms = msSince(tPlan);

    up.join();
    if (upErr != hipSuccess)
        return hipFail(ctx, upErr, "copying residues");
    if ((st = uploadInputs(ctx, setH, setV, scoring, si, scores, flags, in)) != SQMI_OK)
        return st;
    hipError_t err;
    const int wideStride = (int)((wWide + 2 + 63) & ~(int64_t)63);
    const int wideGrid = wide.empty() ? 0 : (int)std::min<int64_t>((int64_t)wide.size(), (int64_t)ctx->numCUs * 16);
    const size_t nIds = narrow.size() + wide.size() + rejected.size();
    if ((err = ensure(ctx->b.bandIds, 4 * std::max<size_t>(nIds, 1))) != hipSuccess ||
        (wideGrid && (err = ensure(ctx->b.bandScratch, (size_t)wideGrid * 4 * wideStride * 4)) != hipSuccess))
        return hipFail(ctx, err, "allocating banded buffers");
    std::vector<int32_t> all(narrow);
    all.insert(all.end(), wide.begin(), wide.end());
    all.insert(all.end(), rejected.begin(), rejected.end());
    int32_t *dIds = static_cast<int32_t *>(ctx->b.bandIds.p);
    if (!all.empty() &&
        (err = hipMemcpyAsync(dIds, all.data(), 4 * all.size(), hipMemcpyHostToDevice, stream)) != hipSuccess)
        return hipFail(ctx, err, "uploading banded pair lists");
    int bandGrid = 0, bandRowColsN = 0, bandLds = 0;
    if (!bandItems.empty()) {
        // Workgroups of 4 waves per CU: the occupancy limit, but no more than
        // keep the row-buffer rings of a CU's waves within its share of L2
        // (at least one).  On 1 M pairs of 150 bp, band [-16, 16] (rings of
        // 32 KB per wave), 1, 2 and 3 workgroups per CU ran 1.92, 2.09 and
        // 2.61 ms: the rings of more waves than L2 holds go to memory.
        // Rings of up to 64 columns (128 KB per workgroup) live in LDS
        // (SQMI_BAND_LDS=0 keeps them in global memory), which also limits
        // the CU to the workgroups whose rings fit.
        bandRowColsN = bandRingCols(bandMaxN, lower_diag, upper_diag, bandR);
        const char *ldsEnv = std::getenv("SQMI_BAND_LDS");
        bandLds = (ldsEnv && ldsEnv[0] == '0') ? 0 : bandLdsRingBytes(bandR, si.dna, bandRowColsN);
        const int64_t ringBytes = (int64_t)64 * bandRowColsN * 8;   // per wave
        const int occ = bandLds ? occupancyBand16(bandR, si.dna, bandLds)
                                : (int)std::max<int64_t>(1, std::min<int64_t>(occupancyBand16(bandR, si.dna, 0),
                                                                              ctx->l2PerCU / (4 * ringBytes)));
        const int64_t nIt = (int64_t)bandItems.size();
        bandGrid = std::max(1, (int)std::min<int64_t>((int64_t)ctx->numCUs * occ, (nIt + 3) / 4));
        if ((err = ensure(ctx->b.order, 4 * bandOrder.size())) != hipSuccess ||
            (err = ensure(ctx->b.items, sizeof(WorkItem) * bandItems.size())) != hipSuccess ||
            (err = ensure(ctx->b.queue, 64)) != hipSuccess ||
            (err = ensure(ctx->b.rowbuf, bandLds ? 64 : (size_t)bandGrid * 4 * 64 * bandRowColsN * 8)) != hipSuccess)
            return hipFail(ctx, err, "allocating packed banded buffers");
        if ((err = hipMemcpyAsync(ctx->b.order.p, bandOrder.data(), 4 * bandOrder.size(), hipMemcpyHostToDevice,
                                  stream)) != hipSuccess ||
            (err = hipMemcpyAsync(ctx->b.items.p, bandItems.data(), sizeof(WorkItem) * bandItems.size(),
                                  hipMemcpyHostToDevice, stream)) != hipSuccess)
            return hipFail(ctx, err, "uploading packed banded work");
    }
    if ((err = hipEventRecord(ctx->ev0, stream)) != hipSuccess)
        return hipFail(ctx, err, "event");
    if (!bandItems.empty()) {
        if ((err = launchZero32(static_cast<uint32_t *>(ctx->b.queue.p), 16, stream)) != hipSuccess)
            return hipFail(ctx, err, "launching queue reset");
        FwdArgs f{};
        f.hRes = in.dH;
        f.vRes = in.dV;
        f.hOff = static_cast<const int64_t *>(ctx->b.hOff.p);
        f.vOff = static_cast<const int64_t *>(ctx->b.vOff.p);
        f.order = static_cast<const int32_t *>(ctx->b.order.p);
        f.items = static_cast<const WorkItem *>(ctx->b.items.p);
        f.nItems = (int32_t)bandItems.size();
        f.queue = static_cast<uint32_t *>(ctx->b.queue.p);
        f.rowbuf = static_cast<uint32_t *>(ctx->b.rowbuf.p);
        f.rowCols = bandRowColsN;
        f.bandCols = bandRowColsN;
        f.scores = in.dScores;
        f.gapOpen = si.o;
        f.gapExtend = si.e;
        f.mode = mode;
        f.alpha = si.alpha;
        f.tTable = static_cast<const int16_t *>(ctx->b.tTable.p);
        f.dnaProfile = static_cast<const uint32_t *>(ctx->b.dnaProf.p);
        f.bandLo = (int32_t)lower_diag;
        f.bandUp = (int32_t)upper_diag;
        if ((err = launchBand16(bandR, si.dna, f, bandGrid, bandLds, stream)) != hipSuccess)
            return hipFail(ctx, err, "launching packed banded kernel");
    }
    BandedArgs a{};
    a.hRes = in.dH;
    a.vRes = in.dV;
    a.hOff = static_cast<const int64_t *>(ctx->b.hOff.p);
    a.vOff = static_cast<const int64_t *>(ctx->b.vOff.p);
    a.scores = in.dScores;
    a.gapOpen = si.o;
    a.gapExtend = si.e;
    a.mode = mode;
    a.alpha = si.alpha;
    a.table = static_cast<const int32_t *>(ctx->b.table32.p);
    a.lower = lower_diag;
    a.upper = upper_diag;
    a.scratch = static_cast<int32_t *>(ctx->b.bandScratch.p);
    a.wideStride = wideStride;
    if (!narrow.empty()) {
        a.ids = dIds;
        a.nIds = (int32_t)narrow.size();
        if ((err = launchBanded(a, seg, 0, stream)) != hipSuccess)
            return hipFail(ctx, err, "launching banded kernel");
    }
    if (!wide.empty()) {
        a.ids = dIds + narrow.size();
        a.nIds = (int32_t)wide.size();
        if ((err = launchBanded(a, 0, wideGrid, stream)) != hipSuccess)
            return hipFail(ctx, err, "launching banded kernel");
    }
    if (!rejected.empty() &&
        (err = launchFillScores(in.dScores, dIds + narrow.size() + wide.size(), (int)rejected.size(), INT32_MIN,
                                stream)) != hipSuccess)
        return hipFail(ctx, err, "launching fill kernel");
    if ((err = hipEventRecord(ctx->ev1, stream)) != hipSuccess)
        return hipFail(ctx, err, "event");
    if (!(flags & SQMI_SCORES_ON_DEVICE) &&
        (err = hipMemcpyAsync(scores, in.dScores, 4 * N, hipMemcpyDeviceToHost, stream)) != hipSuccess)
        return hipFail(ctx, err, "copying scores");
    if ((err = hipStreamSynchronize(stream)) != hipSuccess)
        return hipFail(ctx, err, "banded kernels");
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess)
        ctx->stats.device_ms = ms;
    ctx->stats.total_ms = msSince(tStart);
    return SQMI_OK;
}

sqmi_status sqmi_align_local_rows(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                                  const sqmi_scoring *scoring, int32_t *scores, int32_t *counts, uint8_t *ops,
                                  const int64_t *op_offsets, int64_t *op_lengths, int64_t *begins, uint32_t flags)
{
    if (!ops || !begins)
        return ctx ? fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "ops or begins is null") : SQMI_ERR_INVALID_ARGUMENT;
    return alignBackward(ctx, setH, setV, scoring, SQMI_LOCAL, scores, counts, ops, op_offsets, op_lengths, begins,
                         flags);
}

sqmi_status sqmi_align_rows_banded(sqmi_context *ctx, const sqmi_sequences *setH, const sqmi_sequences *setV,
                                   const sqmi_scoring *scoring, uint32_t mode, int64_t lower_diag,
                                   int64_t upper_diag, int32_t *scores, int32_t *counts, uint8_t *ops,
                                   const int64_t *op_offsets, int64_t *op_lengths, int64_t *begins, uint32_t flags)
{
    if (!ctx)
        return SQMI_ERR_INVALID_ARGUMENT;
    ctx->lastError.clear();
    sqmi_status st;
    if ((st = validate(ctx, setH, setV, scoring, mode)) != SQMI_OK)
        return st;
    if (flags & ~(uint32_t)(SQMI_RESIDUES_ON_DEVICE | SQMI_SCORES_ON_DEVICE | SQMI_ROWS_PACKED))
        return fail(ctx, SQMI_ERR_UNSUPPORTED,
                    "sqmi_align_rows_banded: flags other than the *_ON_DEVICE ones and SQMI_ROWS_PACKED");
    if (lower_diag > upper_diag)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "lower_diag > upper_diag");
    const bool local = (mode & SQMI_LOCAL) != 0;
    if (local && ops && !begins)
        return fail(ctx, SQMI_ERR_INVALID_ARGUMENT, "begins is null");
    for (int64_t i = 0; i < setH->count; ++i) {
        const int64_t n = setH->offsets[i + 1] - setH->offsets[i], m = setV->offsets[i + 1] - setV->offsets[i];
        if (n != 0 && m != 0 && !bandSupported(n, m, lower_diag, upper_diag, local))
            return fail(ctx, SQMI_ERR_UNSUPPORTED,
                        "sqmi_align_rows_banded: pair " + std::to_string(i) + " is outside the supported bands");
    }
    const int64_t band[2] = {lower_diag, upper_diag};
    return alignBackward(ctx, setH, setV, scoring, local ? (uint32_t)SQMI_LOCAL : mode, scores, counts, ops,
                         op_offsets, op_lengths, local ? begins : nullptr, flags, band);
}

}  // extern "C"
