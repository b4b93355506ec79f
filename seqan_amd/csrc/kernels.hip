// kernels.hip -- gfx950 kernels of the batched affine-gap alignment engine
// other than the packed forward (fwd16.hpp): the int32 per-pair sweep, the
// traceback walks and count kernels, the fills; and the dispatch to the
// packed forward's instantiation units (fwd16.hip).
#include <algorithm>
#include <type_traits>

#include "fwd16_parts.hpp"
#include "kernels.hpp"
#include "packed16.hpp"
#include "seqan_amd.h"

namespace sqmi {
namespace {
// ---------------------------------------------------------------------------
// int32, one thread per pair: the textbook column sweep with the reference's
// border and tracking rules.  Used for pairs whose values could leave the
// int16 window.
// ---------------------------------------------------------------------------
constexpr int kNeg32 = INT32_MIN / 2;   // DPCellDefaultInfinity (dp_cell.h:141-146)

__device__ __forceinline__ bool tracked32(uint32_t mode, int col, int cell)
{
    if (mode & SQMI_LOCAL)
        return true;
    if (cell == 2)
        return col == 2 || (mode & SQMI_FREE_BOTTOM);
    return col == 2 && (mode & SQMI_FREE_RIGHT);
}

__global__ __launch_bounds__(256) void kSimple32(Simple32Args a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nIds)
        return;
    const int pid = a.ids[t];
    const int64_t ho = a.hOff[pid], vo = a.vOff[pid];
    const int64_t n = a.hOff[pid + 1] - ho, m = a.vOff[pid + 1] - vo;
    const uint8_t *h = a.hRes + ho;
    const uint8_t *v = a.vRes + vo;
    int32_t *H = a.scratch + 2 * a.scratchOff[t];
    int32_t *E = H + (m + 1);
    const int32_t o = a.gapOpen, e = a.gapExtend;
    const uint32_t mode = a.mode;
    const bool local = (mode & SQMI_LOCAL) != 0;
    const int A = a.alpha;
    int32_t best = kNeg32;
    {
        int32_t F = kNeg32;
        H[0] = 0;
        E[0] = kNeg32;
        if (tracked32(mode, 0, 0) && H[0] > best)
            best = H[0];
        for (int64_t i = 1; i <= m; ++i) {
            if (local || (mode & SQMI_FREE_LEFT)) {
                H[i] = 0;
            } else {
                F = max(F + e, H[i - 1] + o);
                H[i] = F;
            }
            E[i] = kNeg32;
            if (tracked32(mode, 0, i == m ? 2 : 1) && H[i] > best)
                best = H[i];
        }
    }
    for (int64_t j = 1; j <= n; ++j) {
        const int col = (j == n) ? 2 : 1;
        const int hj = min((int)h[j - 1], A - 1);
        const int32_t *row = a.table + hj * A;
        int32_t diag = H[0];
        int32_t F = kNeg32;
        if (local || (mode & SQMI_FREE_TOP)) {
            H[0] = 0;
        } else {
            E[0] = max(E[0] + e, H[0] + o);
            H[0] = E[0];
        }
        if (tracked32(mode, col, 0) && H[0] > best)
            best = H[0];
        int32_t up = H[0];
        for (int64_t i = 1; i <= m; ++i) {
            const int32_t d = diag + row[min((int)v[i - 1], A - 1)];
            diag = H[i];
            const int32_t ei = max(E[i] + e, H[i] + o);
            E[i] = ei;
            F = max(F + e, up + o);
            int32_t x = max(d, max(F, ei));
            if (local)
                x = max(x, 0);
            H[i] = x;
            up = x;
            if (tracked32(mode, col, i == m ? 2 : 1) && x > best)
                best = x;
        }
    }
    a.scores[pid] = best;
}

__global__ void kFillScores(int32_t *scores, const int32_t *ids, int n, int32_t value)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n)
        scores[ids[t]] = value;
}

__global__ void kZero32(uint32_t *p, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        p[t] = 0;
}

// Exclusive prefix sums of the row lengths (one workgroup of 1024 threads:
// thread t sums a contiguous chunk, the chunk sums are scanned in LDS, and
// each thread writes its chunk's prefixes).  start[n] is the total.
__global__ __launch_bounds__(1024) void kScanLengths(const int64_t *len, int64_t *start, int64_t n)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t chunk = (n + 1023) / 1024, b = min(n, t * chunk), e = min(n, b + chunk);
    int64_t sum = 0;
    for (int64_t i = b; i < e; ++i)
        sum += len[i];
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int64_t add = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int64_t run = part[t] - sum;
    for (int64_t i = b; i < e; ++i) {
        start[i] = run;
        run += len[i];
    }
    if (t == 1023)
        start[n] = part[1023];
}

// Rows from the ends of their segments (row i ends at src + off[i + 1]) to
// dst + off[0] + start[i], back to back: one wavefront per row.
__global__ __launch_bounds__(256) void kPackRows(const uint8_t *src, const int64_t *off, const int64_t *len,
                                                 const int64_t *start, uint8_t *dst, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n)
        return;
    const int64_t L = len[i];
    const uint8_t *s = src + off[i + 1] - L;
    uint8_t *d = dst + off[0] + start[i];
    for (int64_t k = threadIdx.x & 63; k < L; k += 64)
        d[k] = s[k];
}

// ---------------------------------------------------------------------------
// Traceback.  The walk follows _computeTraceback / _doTracebackGoDiagonal /
// _doTracebackGoVertical / _doTracebackGoHorizontal (dp_traceback_impl.h:
// 115-546) over trace bits with the encoding of the trace records (bit 0: E
// opened, 1: F opened, 2: not diagonal, 3: E beats F), and counts the
// alignment the way computeAlignmentStats reads its rows
// (evaluate_alignment.h:215-300): one substitution count per aligned pair
// (index h * alpha + v), and per maximal run of gaps in one row one open and
// run-1 extensions.  A run is maximal in either reading direction, so the
// counts can be taken while walking backwards.
//
// With gap open == gap extend the reference runs its linear-gap algorithm
// (_usesAffineGaps, dp_setup.h:244-260; dp_formula_linear.h:62-106).  Its
// cells hold the same values and the same own direction (diagonal wins ties,
// then vertical), which the trace bits record; but it keeps no gap matrices,
// so every gap step re-reads the direction of the cell it reaches, and the
// end cell is taken as it is (no _correctTraceValue for linear gaps,
// dp_algorithm_impl.h:1180-1188).
// ---------------------------------------------------------------------------
constexpr uint32_t kBitEOpen = 1, kBitFOpen = 2, kBitNotDiag = 4, kBitHori = 8;
// Local alignment (byte trace only): the cell did not beat 0, so its trace
// value is NONE and a walk that reaches it stops there.
constexpr uint32_t kBitStop = 16;

__device__ __forceinline__ int dirOf(uint32_t bits)   // 0 diagonal, 1 vertical, 2 horizontal
{
    return (bits & kBitNotDiag) ? ((bits & kBitHori) ? 2 : 1) : 0;
}

template <class Bits>
__device__ __forceinline__ void walkCounts(int n, int m, const uint8_t *h, const uint8_t *v, int A, int score, int endE,
                                           int endF, Bits &bits, int32_t *cnt, uint8_t *opEnd = nullptr,
                                           int64_t *opLen = nullptr, int i0 = -1, int j0 = -1,
                                           bool local = false, int64_t *begins = nullptr, bool linear = false,
                                           const int32_t *lTab = nullptr, int go = 0, int ge = 0)
{
    // Local walks over packed trace records (lTab given) carry no stop bit:
    // they track H along the path instead (the score at the end cell, minus
    // each diagonal's substitution score, minus gapOpen or gapExtend per gap
    // character by the record's open bit) and stop on the first cell where
    // it is 0, which is where the reference's trace value is NONE.
    int cur = score;
    // The walk starts at the end cell (i0, j0), (m, n) unless free end gaps
    // or a local alignment moved it; in global modes the cells after it are
    // trailing gaps of the rows.  A local walk stops at the first cell whose
    // trace value is NONE (kBitStop) and emits neither trailing nor leading
    // gaps (TraceTail_ / TraceHead_, dp_profile.h:482-492).
    int i = i0 < 0 ? m : i0, j = j0 < 0 ? n : j0;
    if (local) {
        m = i;
        n = j;
    }
    int prev = 0;   // last emitted operation: 0 none / aligned pair, 1 gap in H (vertical), 2 gap in V
    int opens = 0, exts = 0;
    // Rows: one byte per alignment column, written backwards from opEnd
    // ('M' both residues, 'V' gap in the H row, 'H' gap in the V row).
    int64_t nOps = 0;
    auto emit = [&](uint8_t op) {
        if (opEnd)
            opEnd[-1 - nOps] = op;
        ++nOps;
    };
    // Trailing gaps after the end cell: the rest of V, then of H.
    for (int k = i; k < m; ++k) {
        if (prev == 1) ++exts; else ++opens;
        prev = 1;
        emit('V');
    }
    for (int k = j; k < n; ++k) {
        if (prev == 2) ++exts; else ++opens;
        prev = 2;
        emit('H');
    }
    // End cell: F, then E, win ties against the recorded direction
    // (_correctTraceValue, dp_algorithm_impl.h:1190-1207), and a gap start
    // there follows the gap matrix (PreferGapsAtEnd_, dp_traceback_impl.h:115-122).
    int dir;
    bool walk = true;
    {
        const uint32_t b0 = (i > 0 && j > 0) ? bits(i, j) : kBitStop;
        dir = linear ? dirOf(b0) : (endF == score) ? 1 : (endE == score) ? 2 : dirOf(b0);
        if (local && (lTab ? cur <= 0 : (b0 & kBitStop) != 0))
            walk = false;
    }
    while (walk && i > 0 && j > 0) {
        if (dir == 0) {
            const int hc = min((int)h[j - 1], A - 1), vc = min((int)v[i - 1], A - 1);
            cnt[hc * A + vc] += 1;
            if (lTab)
                cur -= lTab[hc * A + vc];
            emit('M');
            prev = 0;
            --i;
            --j;
        } else if (dir == 1) {
            for (;;) {
                const bool open = (bits(i, j) & kBitFOpen) != 0;
                if (prev == 1) ++exts; else ++opens;
                prev = 1;
                cur -= (open || linear) ? go : ge;
                emit('V');
                const bool stop = linear || open || i == 1;
                --i;
                if (stop)
                    break;
            }
        } else {
            for (;;) {
                const bool open = (bits(i, j) & kBitEOpen) != 0;
                if (prev == 2) ++exts; else ++opens;
                prev = 2;
                cur -= (open || linear) ? go : ge;
                emit('H');
                const bool stop = linear || open || j == 1;
                --j;
                if (stop)
                    break;
            }
        }
        if (i > 0 && j > 0) {
            if (local && lTab && cur == 0)
                break;
            const uint32_t b = bits(i, j);
            dir = dirOf(b);
            if (local && !lTab && (b & kBitStop))
                break;
        }
    }
    if (local) {
        // The window starts where the walk stopped (dp_traceback_adaptor.h:
        // 87-93); an empty window keeps both begins at 0.
        if (begins && nOps > 0) {
            begins[0] = j;
            begins[1] = i;
        }
        i = j = 0;
    }
    // Leading gaps down to the origin: first the column, then the row
    // (dp_traceback_impl.h:522-545).
    if (i > 0) {
        if (prev == 1) exts += i; else { ++opens; exts += i - 1; }
        prev = 1;
        for (int k = 0; k < i; ++k)
            emit('V');
    }
    if (j > 0) {
        if (prev == 2) exts += j; else { ++opens; exts += j - 1; }
        for (int k = 0; k < j; ++k)
            emit('H');
    }
    cnt[A * A] += opens;
    cnt[A * A + 1] += exts;
    if (opLen)
        *opLen = nOps;
}

// Trace bits of cell (i, j) of one pair from the packed records of
// kFwdTrace16, with the last record read kept (moves up inside a lane reuse
// it).  The walk is latency-bound (each move depends on the last record
// read); plain per-move reads measured faster than prefetching a lane's next
// records, batching the residue loads or a branch-free move loop.
struct Bits16 {
    const uint4 *trace;
    TraceLoc loc;
    int64_t key = -1;
    uint4 w = make_uint4(0, 0, 0, 0);
    __device__ __forceinline__ uint32_t operator()(int i, int j)
    {
        const int r = loc.p + i;
        const int b = r / kTraceRows, rr = r % kTraceRows, l = rr / 16, k = rr % 16;
        const int64_t idx = (loc.rec0 + (int64_t)b * loc.steps + (j + l)) * 64 + loc.lane0 + l;
        if (idx != key) {
            w = trace[idx];
            key = idx;
        }
        const uint32_t ef = k < 8 ? w.x : w.y, dx = k < 8 ? w.z : w.w;
        const int sh = (k & 7) + 8 * loc.half;
        return ((ef >> sh) & 1u) | (((ef >> (sh + 16)) & 1u) << 1) | (((dx >> sh) & 1u) << 2) |
               (((dx >> (sh + 16)) & 1u) << 3);
    }
};

// Trace bytes of a banded pair (kBandedTrace): row i - 1, lane j - i - lo.
struct BandBits {
    const uint8_t *trace;
    int S;
    int lo;
    int w;   // diagonals of the clipped band (<= S)
    int n;
    __device__ __forceinline__ uint32_t operator()(int i, int j) const
    {
        // An optimal path stays inside the band (cells outside are -inf).
        // Cells outside the clipped band or past column n read as a stop,
        // so that no walk can step onto trace bytes the trace kernel did
        // not write (the lanes of a segment beyond the band's width).
        const int c = j - i - lo;
        if (c < 0 || c >= w || j > n)
            return kBitStop;
        return trace[(int64_t)(i - 1) * S + c];
    }
};

__global__ __launch_bounds__(64) void kBandedWalk(BandedWalkArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nIds)
        return;
    const int pid = a.ids[t];
    const int A = a.alpha, S = A * A + 2;
    const int64_t ho = a.hOff[pid], vo = a.vOff[pid];
    const int n = (int)(a.hOff[pid + 1] - ho), m = (int)(a.vOff[pid + 1] - vo);
    int32_t *cnt = a.counts + (int64_t)pid * S;
    const int lo = (int)(a.lower > -(int64_t)m ? a.lower : -(int64_t)m);   // the band clipped to the matrix
    const int up = (int)(a.upper < (int64_t)n ? a.upper : (int64_t)n);
    BandBits bits{a.trace + a.traceOff[t], a.segLanes, lo, min(up - lo + 1, a.segLanes), n};
    walkCounts(n, m, a.hRes + ho, a.vRes + vo, A, a.scores[pid], a.endEF[2 * pid], a.endEF[2 * pid + 1], bits, cnt,
               a.ops ? a.ops + a.opOff[pid + 1] : nullptr, a.ops ? a.opLen + pid : nullptr, a.endIJ[2 * pid],
               a.endIJ[2 * pid + 1], a.local != 0, a.begins ? a.begins + 2 * pid : nullptr, a.linear != 0);
}

constexpr int kWalkThreads = 64;
// Counts live in LDS while walking when alpha^2 + 2 <= kWalkLdsStride.
constexpr int kWalkLdsStride = 51;

// One thread per pair.  LDS: counts accumulate in LDS, else in the output
// row itself (two kernels rather than one pointer of unknown address space).
template <bool LDS>
__global__ __launch_bounds__(kWalkThreads) void kWalk16(WalkArgs a)
{
    extern __shared__ int32_t sCnt[];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nPairs)
        return;
    const int pid = a.ids[t];
    const int A = a.alpha, S = A * A + 2;
    const int64_t ho = a.hOff[pid], vo = a.vOff[pid];
    const int n = (int)(a.hOff[pid + 1] - ho), m = (int)(a.vOff[pid + 1] - vo);
    int32_t *out = a.counts + (int64_t)pid * S;
    int32_t *cnt;
    if constexpr (LDS)
        cnt = sCnt + threadIdx.x * S;
    else
        cnt = out;
    for (int q = 0; q < S; ++q)
        cnt[q] = 0;
    Bits16 bits{a.trace, a.locs[t]};
    walkCounts(n, m, a.hRes + ho, a.vRes + vo, A, a.scores[pid], a.endEF[2 * pid], a.endEF[2 * pid + 1], bits, cnt,
               a.ops ? a.ops + a.opOff[pid + 1] : nullptr, a.ops ? a.opLen + pid : nullptr,
               a.endIJ ? a.endIJ[2 * pid] : -1, a.endIJ ? a.endIJ[2 * pid + 1] : -1, a.local != 0,
               a.begins ? a.begins + 2 * pid : nullptr, a.linear != 0, a.local ? a.table : nullptr, a.gapOpen,
               a.gapExtend);
    if constexpr (LDS)
        for (int q = 0; q < S; ++q)
            out[q] = cnt[q];
}

// int32 forward with trace bytes (same encoding), column-major with column
// stride m + 1, followed by the same walk.  For pairs outside the int16 window.
__global__ __launch_bounds__(256) void kCounts32(Counts32Args a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nIds)
        return;
    const int pid = a.ids[t];
    const int A = a.alpha, S = A * A + 2;
    const int64_t ho = a.hOff[pid], vo = a.vOff[pid];
    const int n = (int)(a.hOff[pid + 1] - ho), m = (int)(a.vOff[pid + 1] - vo);
    const uint8_t *h = a.hRes + ho;
    const uint8_t *v = a.vRes + vo;
    uint8_t *base = a.scratch + a.scratchOff[t];
    int32_t *H = reinterpret_cast<int32_t *>(base);
    int32_t *E = H + (m + 1);
    uint8_t *T = base + 8 * (int64_t)(m + 1);
    const int64_t CS = m + 1;
    const int32_t o = a.gapOpen, e = a.gapExtend;
    int32_t *cnt = a.counts + (int64_t)pid * S;
    for (int k = 0; k < S; ++k)
        cnt[k] = 0;
    // Free end gaps (AlignConfig, dp_setup.h:57-182): a free first row /
    // column is 0; the result is the first strict maximum over the tracked
    // cells in column order: (m, n), the last row if it is free, the last
    // column if it is free (dp_scout.h:169-210, dp_meta_info.h:178-221).
    // Local: both borders are 0, a cell that does not beat 0 is 0 with trace
    // NONE (dp_formula_affine.h:110-118), and every cell is tracked
    // (dp_scout_simd.h:427-440).
    const bool local = (a.mode & SQMI_LOCAL) != 0;
    const bool freeTop = local || (a.mode & SQMI_FREE_TOP) != 0, freeLeft = local || (a.mode & SQMI_FREE_LEFT) != 0;
    const bool freeRight = local || (a.mode & SQMI_FREE_RIGHT) != 0;
    const bool freeBottom = local || (a.mode & SQMI_FREE_BOTTOM) != 0;
    // Banded (globalAlignment / localAlignment with lowerDiag, upperDiag,
    // global_alignment_banded.h:152-181, local_alignment_banded.h:242-259):
    // column j holds rows [j - upper, j - lower] only; the rest are -inf,
    // which is what the reference's band-edge recursions give
    // (dp_formula_affine.h:129-230).  Rows below the band keep the -inf
    // they start with; rows above it are never read again.
    const bool banded = a.banded != 0;
    const int64_t lo = a.lower, hi = a.upper;
    int32_t best = kNeg32, bestE = kNeg32, bestF = kNeg32;
    int bi = m, bj = n;
    auto track = [&](bool lastCol, int i, int j, int32_t hv, int32_t ev, int32_t fv) {
        const bool tracked = local || ((i == m) ? (lastCol || freeBottom) : (lastCol && freeRight));
        if (tracked && hv > best) {
            best = hv;
            bestE = ev;
            bestF = fv;
            bi = i;
            bj = j;
        }
    };
    {
        int32_t F = kNeg32;
        H[0] = (banded && (lo > 0 || hi < 0)) ? kNeg32 : 0;
        E[0] = kNeg32;
        for (int i = 1; i <= m; ++i) {
            if (banded && (-i < lo || -i > hi)) {
                H[i] = kNeg32;
                E[i] = kNeg32;
                F = kNeg32;
                continue;
            }
            if (freeLeft) {
                H[i] = 0;
            } else {
                F = max(F + e, H[i - 1] + o);
                H[i] = F;
            }
            E[i] = kNeg32;
            track(false, i, 0, H[i], E[i], freeLeft ? kNeg32 : F);
        }
    }
    for (int j = 1; j <= n; ++j) {
        const int hj = min((int)h[j - 1], A - 1);
        const int32_t *row = a.table + hj * A;
        int32_t diag = H[0];
        int32_t F = kNeg32;
        if (banded && (j < lo || j > hi)) {
            H[0] = kNeg32;
            E[0] = kNeg32;
        } else if (freeTop) {
            H[0] = 0;
        } else {
            E[0] = max(E[0] + e, H[0] + o);
            H[0] = E[0];
        }
        track(j == n, 0, j, H[0], E[0], kNeg32);
        int32_t up = H[0];
        uint8_t *Tc = T + (int64_t)j * CS;
        int i0 = 1, i1 = m;
        if (banded) {
            i0 = (int)max<int64_t>(1, j - hi);
            i1 = (int)min<int64_t>(m, j - lo);
            if (i0 > 1) {
                diag = H[i0 - 1];   // column j - 1's value: (i0 - 1, j - 1) is in the band
                up = kNeg32;
            }
        }
        for (int i = i0; i <= i1; ++i) {
            const int32_t d = diag + row[min((int)v[i - 1], A - 1)];
            diag = H[i];
            const int32_t extE = E[i] + e, opnE = H[i] + o;
            uint32_t tb = extE < opnE ? kBitEOpen : 0u;
            const int32_t ei = max(extE, opnE);
            E[i] = ei;
            const int32_t extF = F + e, opnF = up + o;
            tb |= extF < opnF ? kBitFOpen : 0u;
            F = max(extF, opnF);
            const int32_t fe = F < ei ? ei : F;
            int32_t x = d;
            if (d < fe) {
                x = fe;
                tb |= kBitNotDiag | (F < ei ? kBitHori : 0u);
            }
            if (local && !(0 < x)) {
                x = 0;
                tb = kBitStop;
            }
            Tc[i] = (uint8_t)tb;
            H[i] = x;
            up = x;
            track(j == n, i, j, x, ei, F);
        }
    }
    a.scores[pid] = best;
    // A band whose lower diagonal ends on the last row: the reference tracks
    // that cell from the next column's code path (dp_algorithm_impl.h:
    // 843-862) and its traceback then takes the cell's own direction, with
    // no preference for the vertical gap.
    if (banded && bi == m && (int64_t)bj - bi == lo)
        bestF = kNeg32;
    auto bits = [&](int i, int j) -> uint32_t { return T[(int64_t)j * CS + i]; };
    walkCounts(n, m, h, v, A, best, bestE, bestF, bits, cnt, a.ops ? a.ops + a.opOff[pid + 1] : nullptr,
               a.ops ? a.opLen + pid : nullptr, bi, bj, local, a.begins ? a.begins + 2 * pid : nullptr, o == e);
}

}  // namespace

int wavesPerBlock16() { return 4; }

namespace {

// The instantiation unit of a shape: DNA profile or table, 16 or 64 lanes.
const Fwd16Part *partOf(bool dna, int G)
{
    if (G == 16)
        return dna ? &kFwdPartDna16 : &kFwdPartTab16;
    if (G == 64)
        return dna ? &kFwdPartDna64 : &kFwdPartTab64;
    return nullptr;
}

const void *kernelFor(const Variant16 &v)
{
    const Fwd16Part *p = partOf(v.dna, v.G);
    return p ? p->kernel(v) : nullptr;
}

}  // namespace

bool variantSupported(const Variant16 &v) { return kernelFor(v) != nullptr; }

int occupancy16(const Variant16 &v)
{
    const void *k = kernelFor(v);
    if (!k)
        return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 256, 0) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

hipError_t launchFwd16(const Variant16 &v, const FwdArgs &args, int gridBlocks, hipStream_t stream,
                       int threadsPerBlock)
{
    const Fwd16Part *p = partOf(v.dna, v.G);
    return p ? p->launch(v, args, gridBlocks, stream, threadsPerBlock) : hipErrorInvalidValue;
}

int traceSteps(int maxN) { return kFwdPartDna16.traceSteps(maxN); }

hipError_t launchFwdTrace16(bool dna, uint32_t track, const FwdArgs &args, int gridBlocks, hipStream_t stream)
{
    return (dna ? kFwdPartDna16 : kFwdPartTab16).launchTrace(track, args, gridBlocks, stream);
}

int occupancyTrace16(bool dna)
{
    int blocks = 0;
    const void *k = (dna ? kFwdPartDna16 : kFwdPartTab16).traceKernel();
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 256, 0) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

hipError_t launchWalk16(const WalkArgs &args, hipStream_t stream)
{
    if (args.nPairs <= 0)
        return hipSuccess;
    const int S = args.alpha * args.alpha + 2;
    const dim3 grid((args.nPairs + kWalkThreads - 1) / kWalkThreads), block(kWalkThreads);
    if (S <= kWalkLdsStride)
        hipLaunchKernelGGL(kWalk16<true>, grid, block, (size_t)kWalkThreads * S * sizeof(int32_t), stream, args);
    else
        hipLaunchKernelGGL(kWalk16<false>, grid, block, 0, stream, args);
    return hipGetLastError();
}

hipError_t launchBandedWalk(const BandedWalkArgs &args, hipStream_t stream)
{
    if (args.nIds <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(kBandedWalk, dim3((args.nIds + 63) / 64), dim3(64), 0, stream, args);
    return hipGetLastError();
}

hipError_t launchCounts32(const Counts32Args &args, hipStream_t stream)
{
    if (args.nIds <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(kCounts32, dim3((args.nIds + 255) / 256), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launchSimple32(const Simple32Args &args, hipStream_t stream)
{
    if (args.nIds <= 0)
        return hipSuccess;
    dim3 grid((args.nIds + 255) / 256), block(256);
    hipLaunchKernelGGL(kSimple32, grid, block, 0, stream, args);
    return hipGetLastError();
}

hipError_t launchZero32(uint32_t *p, size_t n, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(kZero32, dim3((unsigned)blocks), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

hipError_t launchPackRows(const uint8_t *src, const int64_t *off, const int64_t *len, int64_t *start, uint8_t *dst,
                          int64_t n, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(kScanLengths, dim3(1), dim3(1024), 0, stream, len, start, n);
    hipLaunchKernelGGL(kPackRows, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, stream, src, off, len,
                       (const int64_t *)start, dst, n);
    return hipGetLastError();
}

hipError_t launchFillScores(int32_t *scores, const int32_t *ids, int n, int32_t value, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(kFillScores, dim3((n + 255) / 256), dim3(256), 0, stream, scores, ids, n, value);
    return hipGetLastError();
}

}  // namespace sqmi
