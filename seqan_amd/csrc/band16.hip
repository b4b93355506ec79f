// band16.hip -- packed int16 banded global alignment, one pair couple per lane.
//
// Banded DP: only the cells with lower <= j - i <= upper exist (the
// reference's banded globalAlignmentScore, include/seqan/align/
// global_alignment_banded.h:152-181, whose band walks dp_algorithm_impl.h's
// banded tracks; every value comes from cells inside the band).  The
// recurrence, borders and end cell are those of the one-pair-per-lane
// packed form (lane.hip, G-space Gotoh with saturating adds):
//
//   E(i,j) = max(E(i,j-1) + e, G(i,j-1))
//   F(i,j) = max(F(i-1,j) + e, G(i-1,j))
//   H(i,j) = max(G(i-1,j-1) + t(h,v), E(i,j), F(i,j)),   G = H + o
//
// Every lane holds a couple of pairs (the int16 halves of its registers) and
// a wavefront 128 pairs of one length m, so that their rows line up: the
// wavefront sweeps blocks of R rows, bottom-aligned (row m is the last row
// of the last block).  A block of rows i0 .. i0 + R - 1 sweeps only the
// columns its band covers, i0 + lower .. i0 + R - 1 + upper: W - 1 + R
// columns for a band of W diagonals.  Cells outside the band are computed
// like the others but never reach a cell inside it: a cell on the lower
// diagonal (j - i = lower) takes E = -inf (its left neighbour lies outside),
// one on the upper diagonal (j - i = upper) takes F = -inf (its upper
// neighbour lies outside), and a diagonal step stays on its diagonal.  Those
// two boundaries cross the block only in its first and last R columns: there
// row k is forced when k >= u (E; u = j - i0 - lower) or k <= u - (W - 1)
// (F), one bit-select per row with a wavefront-uniform mask (the masks
// depend on k minus the column's place in its group only, so a group makes
// them once); the columns between are plain sweeps.  The saturating adds
// keep the values outside the band bounded, and the row buffer entries a
// block reads outside the columns the block above wrote only feed cells
// outside the band.
#include "kernels.hpp"
#include "packed16.hpp"
#include "seqan_amd.h"

namespace sqmi {
namespace {

// LR: the row-buffer ring lives in LDS (dynamic shared memory, one ring per
// wave) instead of global memory.
template <int R, bool DNA, bool LR = false>
struct Band16 {
    static constexpr int QD = 4;   // columns per group: one dword of residues per pair

    struct Couple {
        const uint8_t *h[2];
        const uint8_t *v[2];
        int n[2], pid[2];
    };

    __host__ __device__ static int rowColsFor(int maxN) { return (maxN + QD - 1) / QD * QD + QD + 1; }

    // Residues of a pair as aligned dwords: columns j .. j + 3 are the funnel
    // shift by (j - 1 + mis) % 4 of dwords (j - 1 + mis) / 4 and the next,
    // both clamped to the dword of the pair's last residue (bytes past n are
    // arbitrary: a column never feeds the columns left of it, so they cannot
    // reach (m, n)).
    struct Residues {
        const uint32_t *words;
        uint32_t mis;
        int lastW;
    };
    __device__ __forceinline__ static Residues residues(const uint8_t *h, int n)
    {
        const uint32_t mis = (uint32_t)(reinterpret_cast<uintptr_t>(h) & 3u);
        return Residues{reinterpret_cast<const uint32_t *>(h - mis), mis, (int)((max(n, 1) - 1 + (int)mis) >> 2)};
    }

    // One block, rows i0 .. i0 + R - 1 (both pairs of every lane share them).
    __device__ __forceinline__ static void block(const FwdArgs &a, const Couple &c, int i0, int maxN, bool first,
                                                 bool last, uint2 *rb, const uint32_t *sProf, const int16_t *sT,
                                                 s16x2 &best)
    {
        const int o = a.gapOpen, e = a.gapExtend;
        const s16x2 o2 = splat(o), e2 = splat(e), neg2 = splat(kNeg16);
        const int eo = max(e, o);   // the first column's gap chain grows by max(e, o) per row
        const int PAD = a.alpha;
        const int TW = a.alpha + 1;
        const int W1 = a.bandUp - a.bandLo;          // W - 1
        const int jS = i0 + a.bandLo;                // the block's first band column (u = 0)
        // Sweep start: the first band column (LDS rings), else rounded down
        // to a multiple of QD plus one (measured faster with global rings).
        const int jA = jS <= 1 ? 1 : LR ? jS : ((jS - 1) & ~(QD - 1)) + 1;
        const int jE = min(i0 + R - 1 + a.bandUp, maxN);           // the block's last band column
        const int jEnd = LR ? jE : (jE + QD - 1) / QD * QD;   // the last group may run past it

        // Per row: DNA the row words of both pairs, else the two row codes;
        // G and E of the column left of the sweep: column 0's border when the
        // sweep starts at column 1, else cells outside the band (-inf).
        uint32_t rowA[R], rowB[DNA ? R : 1];
        s16x2 Gk[R], Ek[R];
        // The rows' residues, all loads first (an empty lane reads the set's
        // first residue; rows above row 1 are padding): one wait per block.
        int raw[R][2];
        if (i0 >= 1) {
            // Rows i0 .. i0 + R - 1 read bytes i0 - 1 .. i0 + R - 2: R / 4 + 1
            // aligned dwords per pair (the last one clamped to the dword of
            // the last byte needed), funnel-shifted into place.
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const uint8_t *p = c.v[x] + (i0 - 1);
                const uint32_t mv = (uint32_t)(reinterpret_cast<uintptr_t>(p) & 3u);
                const uint32_t *w = reinterpret_cast<const uint32_t *>(p - mv);
                uint32_t wv[R / 4 + 1];
#pragma unroll
                for (int q = 0; q < R / 4; ++q)
                    wv[q] = w[q];
                wv[R / 4] = w[min(R / 4, (int)((mv + R - 1) >> 2))];
#pragma unroll
                for (int q = 0; q < R / 4; ++q) {
                    const uint32_t al = __builtin_amdgcn_alignbyte(wv[q + 1], wv[q], mv);
#pragma unroll
                    for (int b = 0; b < 4; ++b)
                        raw[4 * q + b][x] = (int)((al >> (8 * b)) & 0xffu);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < R; ++k) {
#pragma unroll
                for (int x = 0; x < 2; ++x)
                    raw[k][x] = c.v[x][max(i0 + k - 1, 0)];
            }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const int i = i0 + k;
            int code[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const int pad = (i >= 1 && c.pid[x] >= 0) ? 0 : 0xff;   // min() turns 0xff into PAD
                code[x] = min(raw[k][x] | pad, PAD);
            }
            const int g0 = jA > 1 ? kNeg16 : (i < 0) ? kNeg16 : (i == 0) ? o : (o + (i - 1) * eo) + o;
            if constexpr (DNA) {
                rowA[k] = sProf[8 + code[0]];
                rowB[k] = sProf[8 + code[1]];
            } else {
                rowA[k] = (uint32_t)code[0] | ((uint32_t)code[1] << 16);
            }
            Gk[k] = splat(g0);
            Ek[k] = neg2;
        }
        // The diagonal input of the first swept column, and column 0 of this
        // block's bottom row for the next block.
        // The row buffer is a ring of a.bandCols columns (a power of two of at
        // least W - 1 + R + 10): the columns a block uses, jA - 1 to the end
        // of its last group, span at most W + R + 3 (first band column
        // start) or W + R + 6 (group-aligned start) columns, so no column the
        // block writes shares a slot with one it still has to read.
        const int cm = a.bandCols - 1;
        s16x2 upGprev = first ? neg2 : as16(rb[((jA - 1) & cm) * 64].x);
        if (!last && jA == 1)
            rb[0] = make_uint2(asu(Gk[R - 1]), asu(neg2));

        const s16x2 nHalf = pack(c.n[0], c.n[1]);
        uint2 rq[QD];
#pragma unroll
        for (int u = 0; u < QD; ++u)
            rq[u] = first ? make_uint2(asu(neg2), asu(neg2)) : rb[((jA + u) & cm) * 64];
        const Residues ra = residues(c.h[0], c.n[0]), rbs = residues(c.h[1], c.n[1]);
        // Column jA is byte jA - 1 of a pair's residues: dword (jA - 1 + mis)
        // / 4 of its aligned words, at byte (jA - 1 + mis) % 4.  With LDS rings
        // (any start) that is a per-lane base wA and shift; otherwise jA - 1 is
        // a multiple of QD and the dword index is the uniform group count gS
        // (the per-lane base measured 6 % slower there).
        const int offA = jA - 1 + (int)ra.mis, offB = jA - 1 + (int)rbs.mis;
        const int wA = LR ? offA >> 2 : 0, wB = LR ? offB >> 2 : 0;
        const int gS = LR ? 0 : (jA - 1) / QD;
        const uint32_t shA = LR ? (uint32_t)offA & 3u : ra.mis, shB = LR ? (uint32_t)offB & 3u : rbs.mis;
        uint32_t a0 = ra.words[min(wA + gS, ra.lastW)], a1 = ra.words[min(wA + gS + 1, ra.lastW)];
        uint32_t b0 = rbs.words[min(wB + gS, rbs.lastW)], b1 = rbs.words[min(wB + gS + 1, rbs.lastW)];

        // The first block reads -inf for the row above; its row-buffer loads
        // still run (in bounds, the values unused), so that no branch splits
        // the column loop, and the choice is made where an entry is used.
        const uint32_t above = first ? 0u : 0xffffffffu;
        const uint32_t lastMask = last ? 0xffffffffu : 0u;
        // One group of QD columns from s0; EDGE: the group meets a band
        // boundary (some column with u < R or u >= W - 1).
        // The next group's row-buffer entries are read at the start of a
        // group and the group's bottom-row entries stored at its end, so that
        // waiting for a group's loads (vmcnt counts loads and stores in
        // order) never waits for stores issued just before them.
        auto group = [&](auto edge, int s0, uint32_t curA, uint32_t curB) {
            constexpr bool EDGE = decltype(edge)::value;
            uint2 nx[QD], outv[QD];
            // Row k of column s0 + u is forced by a wave-uniform condition on
            // c = k - u alone: one mask per c and boundary for the group.
            constexpr int NC = EDGE ? QD + R - 1 : 1;
            uint32_t fE[NC], fF[NC];
            if constexpr (EDGE) {
                const int d0 = s0 - jS;
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    fE[c] = (c - (QD - 1) >= d0) ? 0xffffffffu : 0u;        // k >= u + d0: j - i <= lower
                    fF[c] = (c - (QD - 1) <= d0 - W1) ? 0xffffffffu : 0u;   // k <= u + d0 - (W - 1): j - i >= upper
                }
            }
#pragma unroll
            for (int u = 0; u < QD; ++u)
                nx[u] = rb[((s0 + QD + u) & cm) * 64];   // read before this block writes them, next group
#pragma unroll
            for (int u = 0; u < QD; ++u) {
                const int j = s0 + u;
                const uint2 up = make_uint2(bfi(above, rq[u].x, asu(neg2)), bfi(above, rq[u].y, asu(neg2)));
                uint32_t colSel = 0, pA = 0, pB = 0;
                if constexpr (DNA) {
                    // [h_A, 0x0c, h_B + 4, 0x0c]: byte h_A of the pair-A row
                    // word, byte h_B of the pair-B one, zero high bytes.
                    const uint32_t pick = (uint32_t)u | (0x0cu << 8) | ((uint32_t)(4 + u) << 16) | (0x0cu << 24);
                    colSel = __builtin_amdgcn_perm(curB, curA, pick) | 0x0c040c00u;
                } else {
                    const uint32_t after = asu((nHalf - splat(j)) >> splat(15));   // 0xffff where j > n
                    const int rA = (int)((curA >> (8 * u)) & 0xffu), rB = (int)((curB >> (8 * u)) & 0xffu);
                    pA = (uint32_t)(((after & 0xffffu) ? PAD : min(rA, PAD)) * TW);
                    pB = (uint32_t)(((after >> 16) ? PAD : min(rB, PAD)) * TW);
                }
                auto tOf = [&](int k) -> uint32_t {
                    if constexpr (DNA) {
                        return __builtin_amdgcn_perm(rowB[k], rowA[k], colSel);
                    } else {
                        const uint32_t lo = (uint16_t)sT[(int)pA + (int)(rowA[k] & 0xffffu)];
                        const uint32_t hi = (uint16_t)sT[(int)pB + (int)(rowA[k] >> 16)];
                        return lo | (hi << 16);
                    }
                };
                s16x2 uG = as16(up.x), uF = as16(up.y);
                const s16x2 upG = uG;
                s16x2 D = addSat(upGprev, as16(tOf(0)));
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    Ek[k] = vmax(addSat(Ek[k], e2), Gk[k]);
                    s16x2 F = vmax(addSat(uF, e2), uG);
                    if constexpr (EDGE) {
                        Ek[k] = as16(bfi(fE[k - u + QD - 1], asu(neg2), asu(Ek[k])));
                        F = as16(bfi(fF[k - u + QD - 1], asu(neg2), asu(F)));
                    }
                    const s16x2 H = vmax(vmax(D, Ek[k]), F);
                    if (k + 1 < R)
                        D = addSat(Gk[k], as16(tOf(k + 1)));
                    const s16x2 Gn = addSat(H, o2);
                    Gk[k] = Gn;
                    uG = Gn;
                    uF = F;
                }
                upGprev = upG;
                outv[u] = make_uint2(asu(uG), asu(uF));
                // The end cell (m, n), last block: 0xffff in the halves where j == n.
                const s16x2 rem = nHalf - splat(j);
                const uint32_t eq = ((uint16_t)rem.x == 0 ? 0xffffu : 0u) | ((uint16_t)rem.y == 0 ? 0xffff0000u : 0u);
                best = as16(bfi(eq & lastMask, asu(uG), asu(best)));
            }
            // The last block's stores are never read as values inside the band.
#pragma unroll
            for (int u = 0; u < QD; ++u) {
                rb[((s0 + u) & cm) * 64] = outv[u];
                rq[u] = nx[u];
            }
        };
        int s0 = jA, g = gS;
        // The group's residue bytes; the next group's dwords are loaded
        // before a group and shifted after it, so the loop head waits for
        // nothing (a wait there would also cover the last group's stores).
        uint32_t curA = __builtin_amdgcn_alignbyte(a1, a0, shA);
        uint32_t curB = __builtin_amdgcn_alignbyte(b1, b0, shB);
        // Consecutive groups share a dword: one new dword per pair and group,
        // loaded two groups before its use.
        uint32_t a2 = ra.words[min(wA + gS + 2, ra.lastW)], b2 = rbs.words[min(wB + gS + 2, rbs.lastW)];
        auto next = [&](auto edge) {
            const uint32_t a3 = ra.words[min(wA + g + 3, ra.lastW)];
            const uint32_t b3 = rbs.words[min(wB + g + 3, rbs.lastW)];
            group(edge, s0, curA, curB);
            curA = __builtin_amdgcn_alignbyte(a2, a1, shA);
            curB = __builtin_amdgcn_alignbyte(b2, b1, shB);
            a1 = a2;
            b1 = b2;
            a2 = a3;
            b2 = b3;
            s0 += QD;
            ++g;
        };
        // Three loops over the groups: the lower boundary's columns, the
        // plain columns, the upper boundary's columns.
        while (s0 <= jEnd && s0 - jS < R)
            next(std::true_type{});
        while (s0 <= jEnd && s0 - jS + QD - 1 < W1)
            next(std::false_type{});
        while (s0 <= jEnd)
            next(std::true_type{});
        // The next block reads this block's row-buffer entries from the
        // same lane: order the stores before those loads.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }

    // Work item: `first` .. first + (count & 0xffff) of `order`, all of length
    // m = count >> 16; nblocks = (m + R) / R; maxN the largest n.
    __device__ __forceinline__ static bool wellFormed(const FwdArgs &a, const WorkItem &w)
    {
        const int pairs = w.count & 0xffff, m = (int)((uint32_t)w.count >> 16);
        return !(pairs < 1 || pairs > 128 || w.maxN < 1 || a.bandCols < 1 || (a.bandCols & (a.bandCols - 1)) != 0 ||
                 a.bandCols > a.rowCols || (a.bandCols < rowColsFor(w.maxN) && a.bandCols < a.bandUp - a.bandLo + R + 10) ||
                 m < 1 || w.nblocks != (m + R) / R);
    }

    __device__ __forceinline__ static void run(const FwdArgs &a)
    {
        __shared__ uint32_t sProf[16];
        __shared__ int16_t sT[33 * 33];
        if constexpr (DNA) {
            if (threadIdx.x < 16)
                sProf[threadIdx.x] = a.dnaProfile[threadIdx.x];
        } else {
            const int nT = (a.alpha + 1) * (a.alpha + 1);
            for (int i = threadIdx.x; i < nT; i += blockDim.x)
                sT[i] = a.tTable[i];
        }
        __syncthreads();

        const int lane = threadIdx.x & 63;
        const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
        const size_t slot = (size_t)blockIdx.x * (blockDim.x >> 6) + wave;
        uint2 *rb;
        if constexpr (LR) {
            extern __shared__ uint2 sRing[];
            rb = sRing + (size_t)wave * a.bandCols * 64 + lane;
        } else {
            rb = reinterpret_cast<uint2 *>(a.rowbuf) + slot * (size_t)a.rowCols * 64 + lane;
        }
        for (;;) {
            // Branch-free fetch: every lane takes part (lane 0 adds 1).
            int it = (int)atomicAdd(a.queue, lane == 0 ? 1u : 0u);
            it = __builtin_amdgcn_readfirstlane(it);
            if (it >= a.nItems)
                break;
            const WorkItem w = a.items[it];
            if (!wellFormed(a, w))
                continue;
            const int pairs = w.count & 0xffff, m = (int)((uint32_t)w.count >> 16);
            Couple c;
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const int sidx = 2 * lane + x;
                const int pid = sidx < pairs ? a.order[w.first + sidx] : -1;
                int64_t ho = 0, vo = 0;
                int n = 0;
                if (pid >= 0) {
                    ho = a.hOff[pid];
                    vo = a.vOff[pid];
                    n = (int)(a.hOff[pid + 1] - ho);
                }
                c.h[x] = a.hRes + ho;
                c.v[x] = a.vRes + vo;
                c.n[x] = n;
                c.pid[x] = pid;
            }
            s16x2 best = splat(kNeg16);
            const int p = w.nblocks * R - 1 - m;   // item row of row 0
            for (int b = 0; b < w.nblocks; ++b)
                block(a, c, b * R - p, w.maxN, b == 0, b == w.nblocks - 1, rb, sProf, sT, best);
            if (c.pid[0] >= 0)
                a.scores[c.pid[0]] = (int)best.x - a.gapOpen;
            if (c.pid[1] >= 0)
                a.scores[c.pid[1]] = (int)best.y - a.gapOpen;
        }
    }
};

template <int R, bool DNA, bool LR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(R >= 12 ? 3 : 4))) void kBand16(FwdArgs a)
{
    Band16<R, DNA, LR>::run(a);
}

const void *bandKernel(int R, bool dna, bool lds)
{
    if (R == 16) {
        if (lds)
            return dna ? (const void *)&kBand16<16, true, true> : (const void *)&kBand16<16, false, true>;
        return dna ? (const void *)&kBand16<16, true, false> : (const void *)&kBand16<16, false, false>;
    }
    if (R == 12) {
        if (lds)
            return dna ? (const void *)&kBand16<12, true, true> : (const void *)&kBand16<12, false, true>;
        return dna ? (const void *)&kBand16<12, true, false> : (const void *)&kBand16<12, false, false>;
    }
    if (R == 8) {
        if (lds)
            return dna ? (const void *)&kBand16<8, true, true> : (const void *)&kBand16<8, false, true>;
        return dna ? (const void *)&kBand16<8, true, false> : (const void *)&kBand16<8, false, false>;
    }
    return nullptr;
}

}  // namespace

int bandRowCols(int maxN) { return Band16<8, true>::rowColsFor(maxN); }

int bandRingCols(int maxN, int64_t lower, int64_t upper, int R)
{
    // The smallest power of two of at least W - 1 + R + 10 columns; when that
    // reaches the pair's whole width the buffer is linear (a power of two
    // no smaller than rowColsFor).
    const int64_t need = std::min<int64_t>(upper - lower + R + 10, (int64_t)bandRowCols(maxN));
    int c = 1;
    while (c < need)
        c <<= 1;
    return c;
}

int bandLdsRingBytes(int R, bool dna, int bandCols)
{
    // Four waves' rings of bandCols columns x 64 lanes x 8 bytes, when they
    // leave room for the static tables in the 160 KB of LDS per CU.
    const int64_t bytes = (int64_t)4 * 64 * bandCols * 8;
    if (bytes > 128 * 1024)
        return 0;
    const void *k = bandKernel(R, dna, true);
    if (!k || hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
        return 0;
    return (int)bytes;
}

int occupancyBand16(int R, bool dna, int ringLdsBytes)
{
    const void *k = bandKernel(R, dna, ringLdsBytes > 0);
    if (!k)
        return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 256, (size_t)ringLdsBytes) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

hipError_t launchBand16(int R, bool dna, const FwdArgs &args, int gridBlocks, int ringLdsBytes, hipStream_t stream)
{
    const void *k = bandKernel(R, dna, ringLdsBytes > 0);
    if (!k)
        return hipErrorInvalidValue;
    FwdArgs copy = args;
    void *params[] = {&copy};
    return hipLaunchKernel(k, dim3(gridBlocks), dim3(256), params, (size_t)ringLdsBytes, stream);
}

}  // namespace sqmi
