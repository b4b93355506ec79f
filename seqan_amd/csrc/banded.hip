// banded.hip -- banded global alignment scores (int32) on gfx950.
//
// globalAlignmentScore(setH, setV, score[, AlignConfig], lowerDiag, upperDiag)
// (include/seqan/align/global_alignment_banded.h:394-470 for one pair,
// :555-583 for sets): only cells (i, j) with lowerDiag <= j - i <= upperDiag
// exist; the others are minus infinity in H, E and F (DPBandConfig<BandOn>,
// dp_band.h).  The recurrence, the borders and the tracked cells are those of
// the unbanded kernels (dp_formula_affine.h:62-126, dp_setup.h:57-182,
// dp_scout.h:169-210); requests the reference rejects are filtered on the
// host (_checkBandProperties, dp_algorithm_impl.h:192-232).
//
// Layout: the band is swept row by row (row i = V position), one lane per
// band diagonal d = j - i, so that the diagonal and vertical predecessors of
// a row come from the previous row (same lane, and the lane of d + 1) and
// only the horizontal gap chain runs along the row.  That chain is a
// max-plus prefix scan over the lanes: with g = max(gapExtend, gapOpen) and
// X = max(diagonal + s, F) the value before the horizontal gap,
//     E(i, j) = max over q < j of X(i, q) + gapOpen + (j - 1 - q) * g,
// which is the unrolled E(j) = max(E(j-1) + gapExtend, H(j-1) + gapOpen).
//   kBandedNarrow: bands of at most 64 diagonals; a wave holds 64 / S pairs
//                  in segments of S lanes (S = the band width rounded up to a
//                  power of two), the previous row in registers; lane shifts
//                  and the scan are DPP moves (wave_shl/shr:1, row_shr,
//                  row_bcast), the substitution table sits in LDS and the
//                  residue codes of the next row are loaded one row ahead.
//   kBandedWide:   wider bands; one wave per pair, the band in chunks of 64
//                  lanes, the previous row in a per-wave global buffer.
//   kBandedTrace:  kBandedNarrow with one trace byte per cell and the end
//                  cell's position, E and F, for banded rows and counts
//                  (walked by kBandedWalk in kernels.hip).
#include "kernels.hpp"
#include "seqan_amd.h"

namespace sqmi {
namespace {

constexpr int32_t kNegB = INT32_MIN / 2;   // DPCellDefaultInfinity (dp_cell.h:141-146)

struct BandPair {
    int32_t pid, n, m;
    int32_t lo, w;      // first band diagonal (clipped to the matrix) and band width
    const uint8_t *h, *v;
    __device__ bool track(int i, int j, uint32_t mode) const
    {
        return (mode & SQMI_LOCAL) || (i == m && (j == n || (mode & SQMI_FREE_BOTTOM))) ||
               (j == n && (mode & SQMI_FREE_RIGHT));
    }
};

__device__ __forceinline__ BandPair loadPair(const BandedArgs &a, int pid)
{
    BandPair p;
    p.pid = pid;
    const int64_t ho = a.hOff[pid], vo = a.vOff[pid];
    p.n = (int32_t)(a.hOff[pid + 1] - ho);
    p.m = (int32_t)(a.vOff[pid + 1] - vo);
    const int64_t lo = a.lower > -(int64_t)p.m ? a.lower : -(int64_t)p.m;
    const int64_t up = a.upper < (int64_t)p.n ? a.upper : (int64_t)p.n;
    p.lo = (int32_t)lo;
    p.w = (int32_t)(up - lo + 1);
    p.h = a.hRes + ho;
    p.v = a.vRes + vo;
    return p;
}

// Row 0 of column j (first row of the matrix, dp_setup.h:57-182): 0 at the
// origin and under a free first row, else a gap chain growing by
// max(gapOpen, gapExtend) per character.
__device__ __forceinline__ int32_t rowZero(int j, bool freeTop, int32_t o, int32_t g)
{
    return j == 0 ? 0 : (freeTop ? 0 : o + (j - 1) * g);
}

// Inclusive prefix max over the S lanes of a segment.
template <int S>
__device__ __forceinline__ int32_t prefixMax(int32_t y, int k)
{
#pragma unroll
    for (int d = 1; d < S; d <<= 1) {
        const int32_t t = __shfl_up(y, d, S);
        if (k >= d)
            y = max(y, t);
    }
    return y;
}

// DPP lane moves; lanes without a source lane read kNegB.
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138, kDppRowShr = 0x110, kDppRowBcast15 = 0x142,
              kDppRowBcast31 = 0x143;
template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ int32_t dpp(int32_t x)
{
    return __builtin_amdgcn_update_dpp(kNegB, x, Ctrl, RowMask, 0xF, false);
}

// DPP lane move where lanes without a source lane read 0 (bound_ctrl), so
// that a max with it folds into one v_max_i32_dpp.
template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ int32_t dpp0(int32_t x)
{
    return __builtin_amdgcn_update_dpp(0, x, Ctrl, RowMask, 0xF, true);
}

// Inclusive prefix max over the S lanes of each segment (S a power of two)
// of values biased to y - kNegB >= 0 (0 stands for minus infinity): row_shr
// steps inside the 16-lane rows (masked at segment edges for S < 16), then
// the row broadcasts for segments of 32 and 64 lanes.
template <int S>
__device__ __forceinline__ int32_t prefixMaxDpp(int32_t y, int k)
{
    if constexpr (S >= 2) {
        const int32_t t = dpp0<kDppRowShr + 1>(y);
        y = (S >= 16 || k >= 1) ? max(y, t) : y;
    }
    if constexpr (S >= 4) {
        const int32_t t = dpp0<kDppRowShr + 2>(y);
        y = (S >= 16 || k >= 2) ? max(y, t) : y;
    }
    if constexpr (S >= 8) {
        const int32_t t = dpp0<kDppRowShr + 4>(y);
        y = (S >= 16 || k >= 4) ? max(y, t) : y;
    }
    if constexpr (S >= 16)
        y = max(y, dpp0<kDppRowShr + 8>(y));
    if constexpr (S >= 32)
        y = max(y, dpp0<kDppRowBcast15, 0xA>(y));
    if constexpr (S >= 64)
        y = max(y, dpp0<kDppRowBcast31, 0xC>(y));
    return y;
}

template <int S>
__device__ __forceinline__ int32_t segmentMax(int32_t y)
{
#pragma unroll
    for (int d = S / 2; d >= 1; d >>= 1)
        y = max(y, __shfl_xor(y, d, S));
    return y;
}

// Rows per residue chunk of kBandedNarrow: every kChunkRows rows a segment
// copies the V residues of the next rows and the H residues they meet into
// LDS, so the row loop reads residues from LDS only.
constexpr int kChunkRows = 64;

template <int S>
__global__ __launch_bounds__(256) void kBandedNarrow(BandedArgs a)
{
    constexpr int kPairsPerWave = 64 / S;
    constexpr int kHSpan = kChunkRows + 64;   // H columns a chunk meets: rows + band width - 1
    __shared__ int32_t sTab[32 * 32];
    __shared__ uint8_t sV[4][kPairsPerWave][kChunkRows];
    __shared__ uint8_t sH[4][kPairsPerWave][kHSpan];
    for (int x = threadIdx.x; x < a.alpha * a.alpha; x += blockDim.x)
        sTab[x] = a.table[x];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int k = lane % S, seg = lane / S;
    const int t = wave * kPairsPerWave + seg;
    const bool has = t < a.nIds;
    if (wave * kPairsPerWave >= a.nIds)
        return;   // the whole wave is past the end
    BandPair P{};
    if (has)
        P = loadPair(a, a.ids[t]);
    const int32_t o = a.gapOpen, e = a.gapExtend, g = max(o, e);
    const int A = a.alpha;
    // Local: free borders, every cell tracked, H never below 0
    // (local_alignment_banded.h, dp_formula_affine.h:110-118).
    const bool freeTop = (a.mode & (SQMI_FREE_TOP | SQMI_LOCAL)) != 0;
    const bool freeLeft = (a.mode & (SQMI_FREE_LEFT | SQMI_LOCAL)) != 0;
    const int32_t floorH = (a.mode & SQMI_LOCAL) ? 0 : kNegB;
    const bool inBand = has && k < P.w;
    const int d = P.lo + k;
    // Rows of the wave: the longest pair of its segments.
    int rows = has ? P.m : 0;
#pragma unroll
    for (int x = S; x < 64; x <<= 1)
        rows = max(rows, __shfl_xor(rows, x));
    int32_t best = kNegB;
    int32_t hPrev = kNegB, fPrev = kNegB;
    if (inBand && d >= 0 && d <= P.n) {
        hPrev = rowZero(d, freeTop, o, g);
        if (P.track(0, d, a.mode))
            best = hPrev;
    }
    const int wv = threadIdx.x >> 6;
    uint8_t *myV = sV[wv][seg];
    uint8_t *myH = sH[wv][seg];
    const bool upOut = k + 1 >= P.w;
    for (int i = 1; i <= rows; ++i) {
        const int r = (i - 1) % kChunkRows;
        if (r == 0) {
            // Chunk rows i .. i + kChunkRows - 1: V[i - 1 + x] and the H
            // columns i + lo + c (c = 0 .. kChunkRows + w - 2), zero outside
            // the sequences.  LDS operations of one wave complete in order,
            // so the wave barriers keep the compiler from moving reads across.
            __builtin_amdgcn_wave_barrier();
            for (int x = k; x < kChunkRows; x += S) {
                const int row = i + x;
                myV[x] = (has && row <= P.m) ? P.v[row - 1] : 0;
            }
            for (int c = k; c < kChunkRows + S - 1; c += S) {
                const int col = i + P.lo + c;
                myH[c] = (has && col >= 1 && col <= P.n) ? P.h[col - 1] : 0;
            }
            __builtin_amdgcn_wave_barrier();
        }
        const int j = i + d;
        const bool valid = inBand && i <= P.m && j >= 0 && j <= P.n;
        const int hb = myH[r + k], vb = myV[r];
        const int32_t sub = sTab[min(hb, A - 1) * A + min(vb, A - 1)];
        // The DPP moves run with every lane enabled (a lane reads its
        // neighbour's register only while that lane is active); the band
        // edge is masked afterwards.
        const int32_t hUpAll = dpp<kDppWaveShl1>(hPrev), fUpAll = dpp<kDppWaveShl1>(fPrev);
        const int32_t hUp = upOut ? kNegB : hUpAll;
        const int32_t fUp = upOut ? kNegB : fUpAll;
        const int32_t F0 = max(fUp + e, hUp + o);
        // Column 0: H is F, or 0 (and F unused) under a free first column.
        const bool col0 = j == 0;
        const int32_t X0 = col0 ? (freeLeft ? 0 : F0) : max(max(hPrev + sub, F0), floorH);
        const int32_t X = valid ? X0 : kNegB;
        const int32_t F = (valid && !(col0 && freeLeft)) ? F0 : kNegB;
        // The scan runs on values biased by -kNegB (cells that can matter
        // are far above kNegB; 0 then stands for minus infinity, and a
        // value derived from minus infinity never wins a cell).
        const int32_t incl = prefixMaxDpp<S>(valid ? max(X - kNegB + o - k * g, 0) : 0, k);
        const int32_t inclLeft = dpp0<kDppWaveShr1>(incl);
        const int32_t E = k == 0 ? kNegB : inclLeft + kNegB + (k - 1) * g;
        const int32_t H = valid ? (col0 ? X : max(X, E)) : kNegB;
        best = (valid && P.track(i, j, a.mode)) ? max(best, H) : best;
        hPrev = H;
        fPrev = F;
    }
    best = segmentMax<S>(best);
    if (has && k == 0)
        a.scores[P.pid] = best;
}

// kBandedNarrow plus one trace byte per cell, in the encoding of the int32
// traceback (bit 0 E opened, 1 F opened, 2 not diagonal, 3 E beats F, 16
// local NONE), and the end cell: every lane keeps the first strict maximum
// of its tracked cells (a lane's cells come in column order), then the
// segment takes the largest value, the smallest column, the smallest row.
// E and F of that cell go with it (_correctTraceValue).
template <int S>
__global__ __launch_bounds__(256) void kBandedTrace(BandedArgs a)
{
    constexpr int kPairsPerWave = 64 / S;
    constexpr int kHSpan = kChunkRows + 64;
    constexpr uint32_t kEOpen = 1, kFOpen = 2, kNotDiag = 4, kHori = 8, kStop = 16;
    __shared__ int32_t sTab[32 * 32];
    __shared__ uint8_t sV[4][kPairsPerWave][kChunkRows];
    __shared__ uint8_t sH[4][kPairsPerWave][kHSpan];
    for (int x = threadIdx.x; x < a.alpha * a.alpha; x += blockDim.x)
        sTab[x] = a.table[x];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int k = lane % S, seg = lane / S;
    const int t = wave * kPairsPerWave + seg;
    const bool has = t < a.nIds;
    if (wave * kPairsPerWave >= a.nIds)
        return;
    BandPair P{};
    if (has)
        P = loadPair(a, a.ids[t]);
    const int32_t o = a.gapOpen, e = a.gapExtend, g = max(o, e);
    const int A = a.alpha;
    const bool local = (a.mode & SQMI_LOCAL) != 0;
    const bool freeTop = (a.mode & (SQMI_FREE_TOP | SQMI_LOCAL)) != 0;
    const bool freeLeft = (a.mode & (SQMI_FREE_LEFT | SQMI_LOCAL)) != 0;
    const int32_t floorH = local ? 0 : kNegB;
    const bool inBand = has && k < P.w;
    const int d = P.lo + k;
    uint8_t *rowT = has ? a.trace + a.traceOff[t] : nullptr;
    int rows = has ? P.m : 0;
#pragma unroll
    for (int x = S; x < 64; x <<= 1)
        rows = max(rows, __shfl_xor(rows, x));
    // End cell candidate of this lane.
    int32_t bv = kNegB, be = kNegB, bf = kNegB;
    int bi = 0x7fffffff, bj = 0x7fffffff;
    int32_t hPrev = kNegB, fPrev = kNegB;
    if (inBand && d >= 0 && d <= P.n) {
        hPrev = rowZero(d, freeTop, o, g);
        if (P.track(0, d, a.mode)) {
            bv = hPrev;
            be = (d == 0 || freeTop) ? kNegB : hPrev;
            bi = 0;
            bj = d;
        }
    }
    const int wv = threadIdx.x >> 6;
    uint8_t *myV = sV[wv][seg];
    uint8_t *myH = sH[wv][seg];
    const bool upOut = k + 1 >= P.w;
    for (int i = 1; i <= rows; ++i) {
        const int r = (i - 1) % kChunkRows;
        if (r == 0) {
            __builtin_amdgcn_wave_barrier();
            for (int x = k; x < kChunkRows; x += S) {
                const int row = i + x;
                myV[x] = (has && row <= P.m) ? P.v[row - 1] : 0;
            }
            for (int c = k; c < kChunkRows + S - 1; c += S) {
                const int col = i + P.lo + c;
                myH[c] = (has && col >= 1 && col <= P.n) ? P.h[col - 1] : 0;
            }
            __builtin_amdgcn_wave_barrier();
        }
        const int j = i + d;
        const bool valid = inBand && i <= P.m && j >= 0 && j <= P.n;
        const int hb = myH[r + k], vb = myV[r];
        const int32_t sub = sTab[min(hb, A - 1) * A + min(vb, A - 1)];
        const int32_t hUpAll = dpp<kDppWaveShl1>(hPrev), fUpAll = dpp<kDppWaveShl1>(fPrev);
        const int32_t hUp = upOut ? kNegB : hUpAll;
        const int32_t fUp = upOut ? kNegB : fUpAll;
        const int32_t F0 = max(fUp + e, hUp + o);
        const bool col0 = j == 0;
        const int32_t D = hPrev + sub;
        const int32_t X0 = col0 ? (freeLeft ? 0 : F0) : max(max(D, F0), floorH);
        const int32_t X = valid ? X0 : kNegB;
        const int32_t F = (valid && !(col0 && freeLeft)) ? F0 : kNegB;
        const int32_t incl = prefixMaxDpp<S>(valid ? max(X - kNegB + o - k * g, 0) : 0, k);
        const int32_t inclLeft = dpp0<kDppWaveShr1>(incl);
        const int32_t E = k == 0 ? kNegB : inclLeft + kNegB + (k - 1) * g;
        const int32_t H = valid ? (col0 ? X : max(X, E)) : kNegB;
        // The left neighbour's H and E (lane k - 1, this row); every lane
        // runs the moves, the segment edge is masked afterwards.
        const int32_t hLeftAll = dpp<kDppWaveShr1>(H), eLeftAll = dpp<kDppWaveShr1>(k == 0 ? kNegB : E);
        const int32_t hLeft = k == 0 ? kNegB : hLeftAll, eLeft = k == 0 ? kNegB : eLeftAll;
        // Every byte of the pair's rows is written: cells outside the band
        // or the matrix, and column 0 (which the walk leaves as a leading
        // gap run), get the stop byte, so no path can read a stale byte.
        if (has && i <= P.m) {
            uint32_t tb = kStop;
            if (valid && !col0) {
                const int32_t extE = eLeft + e, opnE = hLeft + o;
                tb = extE < opnE ? kEOpen : 0u;
                const int32_t ei = max(extE, opnE);
                tb |= (fUp + e) < (hUp + o) ? kFOpen : 0u;
                const int32_t fe = F0 < ei ? ei : F0;
                int32_t x = D;
                if (D < fe) {
                    x = fe;
                    tb |= kNotDiag | (F0 < ei ? kHori : 0u);
                }
                if (local && !(0 < x))
                    tb = kStop;
            }
            rowT[(int64_t)(i - 1) * S + k] = (uint8_t)tb;
        }
        if (valid && P.track(i, j, a.mode) && H > bv) {
            bv = H;
            be = col0 ? kNegB : E;
            bf = F;
            bi = i;
            bj = j;
        }
        hPrev = H;
        fPrev = F;
    }
    // Segment reduction: larger value, then smaller column, then smaller row.
#pragma unroll
    for (int x = S / 2; x >= 1; x >>= 1) {
        const int32_t ov = __shfl_xor(bv, x, S), oe = __shfl_xor(be, x, S), of = __shfl_xor(bf, x, S);
        const int oi = __shfl_xor(bi, x, S), oj = __shfl_xor(bj, x, S);
        const bool take = ov > bv || (ov == bv && (oj < bj || (oj == bj && oi < bi)));
        bv = take ? ov : bv;
        be = take ? oe : be;
        bf = take ? of : bf;
        bi = take ? oi : bi;
        bj = take ? oj : bj;
    }
    if (has && k == 0) {
        // A band whose lower diagonal ends on the last row: the reference
        // tracks that cell from the next column's code path
        // (dp_algorithm_impl.h:843-862) and its traceback takes the cell's
        // own direction there, with no preference for the vertical gap.
        if (bi == P.m && (int64_t)bj - bi == a.lower)
            bf = kNegB;
        a.scores[P.pid] = bv;
        a.endIJ[2 * P.pid] = bi;
        a.endIJ[2 * P.pid + 1] = bj;
        a.endEF[2 * P.pid] = be;
        a.endEF[2 * P.pid + 1] = bf;
    }
}

// One wave per pair; rows of the band (H and F, double-buffered) in the
// wave's slice of `scratch`: 4 * (wmax + 2) int32.  blockDim = 64, so the
// workgroup barrier between rows orders this wave's own global writes and
// reads.
__global__ __launch_bounds__(64) void kBandedWide(BandedArgs a)
{
    const int k = threadIdx.x;
    const int32_t o = a.gapOpen, e = a.gapExtend, g = max(o, e);
    const int A = a.alpha;
    // Local: free borders, every cell tracked, H never below 0
    // (local_alignment_banded.h, dp_formula_affine.h:110-118).
    const bool freeTop = (a.mode & (SQMI_FREE_TOP | SQMI_LOCAL)) != 0;
    const bool freeLeft = (a.mode & (SQMI_FREE_LEFT | SQMI_LOCAL)) != 0;
    const int32_t floorH = (a.mode & SQMI_LOCAL) ? 0 : kNegB;
    const int stride = a.wideStride;   // wmax + 2, padded
    int32_t *buf = a.scratch + (int64_t)blockIdx.x * 4 * stride;
    for (int t = blockIdx.x; t < a.nIds; t += gridDim.x) {
        const BandPair P = loadPair(a, a.ids[t]);
        int32_t *hRow[2] = {buf, buf + stride};
        int32_t *fRow[2] = {buf + 2 * stride, buf + 3 * stride};
        int32_t best = kNegB;
        for (int c = k; c < P.w + 1; c += 64) {
            const int dd = P.lo + c;
            int32_t hv = kNegB;
            if (c < P.w && dd >= 0 && dd <= P.n) {
                hv = rowZero(dd, freeTop, o, g);
                if (P.track(0, dd, a.mode))
                    best = max(best, hv);
            }
            hRow[0][c] = hv;
            fRow[0][c] = kNegB;
        }
        __syncthreads();
        for (int i = 1; i <= P.m; ++i) {
            const int32_t *hp = hRow[(i - 1) & 1], *fp = fRow[(i - 1) & 1];
            int32_t *hn = hRow[i & 1], *fn = fRow[i & 1];
            const int vc = min((int)P.v[i - 1], A - 1);
            int32_t carry = kNegB;   // E at the chunk's first lane from the cells before it
            for (int base = 0; base < P.w; base += 64) {
                const int c = base + k;
                const int j = i + P.lo + c;
                const bool valid = c < P.w && j >= 0 && j <= P.n;
                int32_t X = kNegB, F = kNegB;
                if (valid) {
                    F = max(fp[c + 1] + e, hp[c + 1] + o);
                    if (j == 0) {
                        X = freeLeft ? 0 : F;
                        if (freeLeft)
                            F = kNegB;
                    } else {
                        const int hc = min((int)P.h[j - 1], A - 1);
                        X = max(max(hp[c] + a.table[hc * A + vc], F), floorH);
                    }
                }
                const int32_t incl = prefixMax<64>(valid ? X + o - k * g : kNegB, k);
                int32_t E = __shfl_up(incl, 1, 64) + (k - 1) * g;
                if (k == 0)
                    E = kNegB;
                E = max(E, carry + k * g);
                const int32_t H = valid ? (j == 0 ? X : max(X, E)) : kNegB;
                if (valid && P.track(i, j, a.mode))
                    best = max(best, H);
                if (c < P.w) {
                    hn[c] = H;
                    fn[c] = valid ? F : kNegB;
                }
                carry = max(max(carry + 64 * g, __shfl(incl, 63, 64) + 63 * g), kNegB);
            }
            if (k == 0)
                hn[P.w] = fn[P.w] = kNegB;   // the diagonal past the band
            __syncthreads();
        }
        best = segmentMax<64>(best);
        if (k == 0)
            a.scores[P.pid] = best;
        __syncthreads();   // the buffer is reused by the next pair
    }
}

}  // namespace

hipError_t launchBanded(const BandedArgs &args, int segLanes, int wideGrid, hipStream_t stream)
{
    if (args.nIds <= 0)
        return hipSuccess;
    if (segLanes == 0) {
        hipLaunchKernelGGL(kBandedWide, dim3(wideGrid), dim3(64), 0, stream, args);
        return hipGetLastError();
    }
    const int pairsPerWave = 64 / segLanes;
    const int waves = (args.nIds + pairsPerWave - 1) / pairsPerWave;
    const dim3 grid((waves + 3) / 4), block(256);
    switch (segLanes) {
        case 1: hipLaunchKernelGGL(kBandedNarrow<1>, grid, block, 0, stream, args); break;
        case 2: hipLaunchKernelGGL(kBandedNarrow<2>, grid, block, 0, stream, args); break;
        case 4: hipLaunchKernelGGL(kBandedNarrow<4>, grid, block, 0, stream, args); break;
        case 8: hipLaunchKernelGGL(kBandedNarrow<8>, grid, block, 0, stream, args); break;
        case 16: hipLaunchKernelGGL(kBandedNarrow<16>, grid, block, 0, stream, args); break;
        case 32: hipLaunchKernelGGL(kBandedNarrow<32>, grid, block, 0, stream, args); break;
        case 64: hipLaunchKernelGGL(kBandedNarrow<64>, grid, block, 0, stream, args); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launchBandedTrace(const BandedArgs &args, int segLanes, hipStream_t stream)
{
    if (args.nIds <= 0)
        return hipSuccess;
    const int pairsPerWave = 64 / segLanes;
    const int waves = (args.nIds + pairsPerWave - 1) / pairsPerWave;
    const dim3 grid((waves + 3) / 4), block(256);
    switch (segLanes) {
        case 1: hipLaunchKernelGGL(kBandedTrace<1>, grid, block, 0, stream, args); break;
        case 2: hipLaunchKernelGGL(kBandedTrace<2>, grid, block, 0, stream, args); break;
        case 4: hipLaunchKernelGGL(kBandedTrace<4>, grid, block, 0, stream, args); break;
        case 8: hipLaunchKernelGGL(kBandedTrace<8>, grid, block, 0, stream, args); break;
        case 16: hipLaunchKernelGGL(kBandedTrace<16>, grid, block, 0, stream, args); break;
        case 32: hipLaunchKernelGGL(kBandedTrace<32>, grid, block, 0, stream, args); break;
        case 64: hipLaunchKernelGGL(kBandedTrace<64>, grid, block, 0, stream, args); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sqmi
