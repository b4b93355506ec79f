// lane.hip -- the one-pair-per-lane form of the packed int16 forward.
//
// The reference's inter-sequence form puts one pair in every SIMD lane and
// pads a batch of 8 / 16 pairs to the batch's maximum lengths
// (include/seqan/align/dp_align_simd_helper.h:147-221, 277-347).  Here every
// lane of a wavefront holds a couple of pairs (the low and high int16 halves
// of its registers), so one wavefront sweeps 128 pairs, and the planner fills
// a wavefront with pairs of nearly one shape (same number of R-row blocks,
// widths sorted), which keeps the padding near zero.
//
// Recurrence, borders and result cells are those of the packed family
// (kernels.hip, G-space Gotoh of include/seqan/align/dp_formula_affine.h:
// 62-126 with the tracking rules of dp_meta_info.h:178-221 and
// dp_scout.h:169-210):
//
//   E(i,j) = max(E(i,j-1) + e, G(i,j-1))
//   F(i,j) = max(F(i-1,j) + e, G(i-1,j))
//   H(i,j) = max(G(i-1,j-1) + t(h,v), E(i,j), F(i,j))   [local: max(., 0)]
//   G(i,j) = H(i,j) + o,   t = s - o
//
// A lane sweeps its couple in blocks of R rows (rows bottom-aligned: row m
// is the last row of the last block, the rows above row 0 are virtual), all
// columns of one block before the next.  Adds saturate (v_pk_add_i16 with
// clamp), so the virtual rows sink towards -32768 and stay below every real
// value without a floor; only a free first row needs one (rows up to row 0
// floored at 0, in the first block).  No value crosses lanes, so there is no
// skew and no lane shift; the bottom row of a block (G, F per column) goes
// to a per-wave row buffer laid out column-major across the 64 lanes (one
// coalesced 512-byte load and store per column), read back by the next
// block.  The row buffer and the residues are loaded one group of QD = 4
// columns ahead of the sweep; the residues as one aligned dword window per
// pair and group.
//
// Substitution scores.  DNA: each row keeps, per pair, the word of t(h, v)
// over h = 0..3 for its residue v (the padding code: zero bytes); each
// column builds one v_perm selector for both pairs from its two residues,
// and one v_perm per row then yields t for both halves.  Other alphabets:
// the (alpha + 1)^2 table in LDS, two reads per cell.
#include "kernels.hpp"
#include "packed16.hpp"
#include "seqan_amd.h"

namespace sqmi {
namespace {

template <int R, bool DNA, uint32_t TRK>
struct Lane16 {
    static constexpr int QD = 4;   // columns per group: one dword of residues per pair

    struct Couple {
        const uint8_t *h[2];
        const uint8_t *v[2];
        int n[2], m[2], p[2], pid[2];
    };

    __device__ __forceinline__ static int padded(int Nw) { return (Nw + QD - 1) / QD * QD; }

    // Row-buffer columns a work item of width Nw needs (one group of prefetch
    // past the padded width).
    __host__ __device__ static int rowColsFor(int Nw) { return (Nw + QD - 1) / QD * QD + QD + 1; }

    // Residues of a pair, read as aligned dwords: group g (columns 4g + 1
    // .. 4g + 4, bytes h[4g] .. h[4g + 3]) is the funnel shift by `mis` of
    // dwords g and g + 1 of `words` (the dword-aligned start of the pair's
    // residues).  Both indexes are clamped to `lastW`, the dword holding the
    // pair's last residue, so no read leaves the pages of the pair's own
    // residues; bytes for columns past n are then arbitrary and are masked
    // (local, non-DNA) or cannot reach a result (a column never feeds the
    // columns left of it).
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

    // One block of R rows: column 0, then columns 1 .. Nw (rounded up to QD;
    // the extra columns cannot change a result).  FIRST: block 0, with no
    // row above it (-inf).  TOPFLOOR: block 0 under a free first row, whose
    // rows up to row 0 are floored at 0 (H(0, j) = 0; the virtual rows above
    // it then hold 0 too, which leaves row 1 exactly as the border would).
    template <bool FIRST, bool TOPFLOOR>
    __device__ __forceinline__ static void block(const FwdArgs &a, const Couple &c, int b, int Nw, bool last,
                                                 uint2 *rb, const uint32_t *sProf, const int16_t *sT, s16x2 &best)
    {
        const int o = a.gapOpen, e = a.gapExtend;
        const s16x2 o2 = splat(o), e2 = splat(e), neg2 = splat(kNeg16);
        const bool leftFree = (a.mode & (SQMI_FREE_LEFT | SQMI_LOCAL)) != 0;
        const int eo = max(e, o);   // the first column's gap chain grows by max(e, o) per row
        const int PAD = a.alpha;
        const int TW = a.alpha + 1;
        const int rowbase = b * R;

        // Per row: DNA the row words of both pairs, else the two row codes.
        uint32_t rowA[R], rowB[DNA ? R : 1];
        s16x2 Gk[R], Ek[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            int code[2], g0[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const int i = rowbase + k - c.p[x];
                int cc = PAD;
                if (i >= 1) {
                    const int idx = min(i - 1, max(c.m[x] - 1, 0));
                    cc = min((int)c.v[x][idx], PAD);
                }
                code[x] = cc;
                // G at column 0: virtual rows -inf, H(0,0) = 0, the first column.
                g0[x] = (i < 0) ? kNeg16 : (i == 0) ? o : ((leftFree ? 0 : o + (i - 1) * eo) + o);
            }
            if constexpr (DNA) {
                rowA[k] = sProf[8 + code[0]];
                rowB[k] = sProf[8 + code[1]];
            } else {
                rowA[k] = (uint32_t)code[0] | ((uint32_t)code[1] << 16);
            }
            Gk[k] = pack(g0[0], g0[1]);
            Ek[k] = neg2;
        }
        // TOPFLOOR: row k is floored at 0 in the halves whose row 0 lies at or
        // below it.  The threshold is made opaque in every step, so that the
        // compiler recomputes the R masks instead of keeping them live.
        s16x2 top1 = pack(c.p[0] + 1, c.p[1] + 1);

        // Column 0 of the previous block's bottom row (the diagonal input of
        // column 1), then this block's own column-0 entry.
        s16x2 upGprev = neg2;
        if constexpr (!FIRST)
            upGprev = as16(rb[0].x);
        if (!last)
            rb[0] = make_uint2(asu(Gk[R - 1]), asu(neg2));
        if constexpr (TRK & kTrackBottom) {
            if (last)
                best = vmax(best, Gk[R - 1]);   // H(m, 0) is a tracked cell
        }

        const int NwP = padded(Nw);
        const s16x2 nHalf = pack(c.n[0], c.n[1]);
        // Row-buffer entries of the next QD columns (blocks after the first),
        // and the residue dwords of the next group, are always in flight.
        uint2 rq[QD];
#pragma unroll
        for (int u = 0; u < QD; ++u)
            rq[u] = FIRST ? make_uint2(asu(neg2), asu(neg2)) : rb[(1 + u) * 64];
        const Residues ra = residues(c.h[0], c.n[0]), rbs = residues(c.h[1], c.n[1]);
        uint32_t a0 = ra.words[min(0, ra.lastW)], a1 = ra.words[min(1, ra.lastW)];
        uint32_t b0 = rbs.words[min(0, rbs.lastW)], b1 = rbs.words[min(1, rbs.lastW)];
        for (int s0 = 1, g = 0; s0 <= NwP; s0 += QD, ++g) {
            const uint32_t curA = __builtin_amdgcn_alignbyte(a1, a0, ra.mis);
            const uint32_t curB = __builtin_amdgcn_alignbyte(b1, b0, rbs.mis);
            // The next group's residue dwords (past the width: unused reads
            // of the clamped dwords).
            a0 = ra.words[min(g + 1, ra.lastW)];
            a1 = ra.words[min(g + 2, ra.lastW)];
            b0 = rbs.words[min(g + 1, rbs.lastW)];
            b1 = rbs.words[min(g + 2, rbs.lastW)];
#pragma unroll
            for (int u = 0; u < QD; ++u) {
                const int j = s0 + u;
                const uint2 up = rq[u];
                // Column j + QD: its row-buffer entry is read before this
                // block writes it (QD steps from now); past the item's width
                // it is an unused in-bounds read.
                if constexpr (!FIRST)
                    rq[u] = rb[(j + QD) * 64];
                // 0xffff in the halves where j > n.
                auto afterMask = [&]() -> uint32_t { return asu((nHalf - splat(j)) >> splat(15)); };
                uint32_t colSel = 0, pA = 0, pB = 0;
                if constexpr (DNA) {
                    // [h_A, 0x0c, h_B + 4, 0x0c]: byte h_A of the pair-A row
                    // word, byte h_B of the pair-B one, zero high bytes.
                    const uint32_t pick = (uint32_t)u | (0x0cu << 8) | ((uint32_t)(4 + u) << 16) | (0x0cu << 24);
                    colSel = __builtin_amdgcn_perm(curB, curA, pick) | 0x0c040c00u;
                    if constexpr (TRK & kTrackLocal)
                        colSel = bfi(afterMask(), 0x0c0c0c0cu, colSel);   // padding columns: t = 0
                } else {
                    const uint32_t after = afterMask();
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
                if constexpr (TOPFLOOR)
                    asm volatile("" : "+v"(top1));
                // The diagonal term of row k comes from the old G of row k - 1
                // before it is overwritten.
                s16x2 D = addSat(upGprev, as16(tOf(0)));
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    Ek[k] = vmax(addSat(Ek[k], e2), Gk[k]);
                    const s16x2 F = vmax(addSat(uF, e2), uG);
                    s16x2 H = vmax(vmax(D, Ek[k]), F);
                    if constexpr (TRK & kTrackLocal) {
                        H = vmax(H, splat(0));
                    } else if constexpr (TOPFLOOR) {
                        // 0 in the halves with k <= p (row k is row 0 or above it).
                        const uint32_t atOrAbove = asu(subSat(splat(k), top1) >> splat(15));
                        H = vmax(H, as16(~atOrAbove & 0x80008000u));
                    }
                    if (k + 1 < R)
                        D = addSat(Gk[k], as16(tOf(k + 1)));
                    const s16x2 Gn = addSat(H, o2);
                    Gk[k] = Gn;
                    if constexpr (TRK & kTrackLocal)
                        best = vmax(best, Gn);
                    uG = Gn;
                    uF = F;
                }
                upGprev = upG;
                if (!last)
                    rb[j * 64] = make_uint2(asu(uG), asu(uF));
                if constexpr (TRK & (kTrackEnd | kTrackBottom | kTrackRight)) {
                    // 0xffff in the halves where j == n.
                    auto eqMaskN = [&]() -> uint32_t {
                        const s16x2 rem = nHalf - splat(j);
                        return ((uint16_t)rem.x == 0 ? 0xffffu : 0u) | ((uint16_t)rem.y == 0 ? 0xffff0000u : 0u);
                    };
                    if constexpr (TRK & kTrackBottom) {
                        if (last)
                            best = vmax(best, as16(bfi(afterMask(), asu(neg2), asu(uG))));
                    }
                    if constexpr (TRK & kTrackEnd) {
                        if (last)
                            best = as16(bfi(eqMaskN(), asu(uG), asu(best)));
                    }
                    if constexpr (TRK & kTrackRight) {
                        const uint32_t eq = eqMaskN();
                        if (eq != 0) {
                            s16x2 cm = Gk[0];
#pragma unroll
                            for (int k = 1; k < R; ++k)
                                cm = vmax(cm, Gk[k]);
                            best = vmax(best, as16(bfi(eq, asu(cm), asu(neg2))));
                        }
                    }
                }
            }
        }
        // The next block reads this block's row-buffer entries from the
        // same lane: order the stores before those loads.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }

    __device__ __forceinline__ static bool wellFormed(const FwdArgs &a, const WorkItem &w)
    {
        // A malformed descriptor is skipped rather than allowed to run past
        // the row buffer or loop without bound.
        return !(w.count < 1 || w.count > 128 || w.maxN < 1 || rowColsFor(w.maxN) > a.rowCols || w.nblocks < 1 ||
                 w.nblocks > (1 << 20) / R + 1);
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
        uint2 *rb = reinterpret_cast<uint2 *>(a.rowbuf) + slot * (size_t)a.rowCols * 64 + lane;
        for (;;) {
            // Branch-free: every lane takes part (lane 0 adds 1), so no
            // lane-0 branch can be merged with another around the back edge
            // and this readfirstlane always runs with lane 0 active.
            int it = (int)atomicAdd(a.queue, lane == 0 ? 1u : 0u);
            it = __builtin_amdgcn_readfirstlane(it);
            if (it >= a.nItems)
                break;
            const WorkItem w = a.items[it];
            const bool ok = wellFormed(a, w);
            Couple c;
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const int sidx = 2 * lane + x;
                const int pid = (ok && sidx < w.count) ? a.order[w.first + sidx] : -1;
                int64_t ho = 0, vo = 0;
                int n = 0, m = 0;
                if (pid >= 0) {
                    ho = a.hOff[pid];
                    vo = a.vOff[pid];
                    n = (int)(a.hOff[pid + 1] - ho);
                    m = (int)(a.vOff[pid + 1] - vo);
                }
                c.h[x] = a.hRes + ho;
                c.v[x] = a.vRes + vo;
                c.n[x] = n;
                c.m[x] = m;
                c.pid[x] = pid;
                c.p[x] = w.nblocks * R - 1 - m;
            }
            s16x2 best = (TRK & kTrackLocal) ? splat(a.gapOpen) : splat(kNeg16);
            const int nb = ok ? w.nblocks : 0;
            const bool topFree = !(TRK & kTrackLocal) && (a.mode & SQMI_FREE_TOP) != 0;
            if (nb > 0) {
                if (topFree)
                    block<true, true>(a, c, 0, w.maxN, nb == 1, rb, sProf, sT, best);
                else
                    block<true, false>(a, c, 0, w.maxN, nb == 1, rb, sProf, sT, best);
            }
            for (int b = 1; b < nb; ++b)
                block<false, false>(a, c, b, w.maxN, b == nb - 1, rb, sProf, sT, best);
            if (c.pid[0] >= 0)
                a.scores[c.pid[0]] = (int)best.x - a.gapOpen;
            if (c.pid[1] >= 0)
                a.scores[c.pid[1]] = (int)best.y - a.gapOpen;
        }
    }
};

// Register budgets without spills in the column loops: four waves per SIMD
// (128 VGPRs) for 16 rows per lane, three for local alignment (its running
// maximum and padding masks), two for 32 rows.
template <int R, uint32_t TRK>
constexpr int laneWaves() { return R >= 32 ? 2 : (TRK & kTrackLocal) ? 3 : 4; }

template <int R, bool DNA, uint32_t TRK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(laneWaves<R, TRK>()))) void kLane16(FwdArgs a)
{
    Lane16<R, DNA, TRK>::run(a);
}

template <int R, bool DNA>
const void *laneKernel(uint32_t track)
{
    switch (track) {
        case kTrackEnd: return (const void *)&kLane16<R, DNA, kTrackEnd>;
        case kTrackBottom: return (const void *)&kLane16<R, DNA, kTrackBottom>;
        case kTrackRight: return (const void *)&kLane16<R, DNA, kTrackRight>;
        case kTrackBottom | kTrackRight: return (const void *)&kLane16<R, DNA, kTrackBottom | kTrackRight>;
        case kTrackLocal: return (const void *)&kLane16<R, DNA, kTrackLocal>;
        default: return nullptr;
    }
}

const void *laneKernelFor(const Variant16 &v)
{
    if (v.G != 1)
        return nullptr;
    if (v.R == 16)
        return v.dna ? laneKernel<16, true>(v.track) : laneKernel<16, false>(v.track);
    if (v.R == 32)
        return v.dna ? laneKernel<32, true>(v.track) : laneKernel<32, false>(v.track);
    return nullptr;
}

}  // namespace

bool laneSupported(const Variant16 &v) { return laneKernelFor(v) != nullptr; }

int laneRowCols(int maxN, int R)
{
    return R >= 32 ? Lane16<32, true, kTrackEnd>::rowColsFor(maxN) : Lane16<16, true, kTrackEnd>::rowColsFor(maxN);
}

int occupancyLane16(const Variant16 &v)
{
    const void *k = laneKernelFor(v);
    if (!k)
        return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 256, 0) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

hipError_t launchLane16(const Variant16 &v, const FwdArgs &args, int gridBlocks, hipStream_t stream)
{
    const void *k = laneKernelFor(v);
    if (!k)
        return hipErrorInvalidValue;
    FwdArgs copy = args;
    void *params[] = {&copy};
    return hipLaunchKernel(k, dim3(gridBlocks), dim3(256), params, 0, stream);
}

}  // namespace sqmi
