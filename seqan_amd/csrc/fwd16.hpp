// fwd16.hpp -- the packed int16 forward kernels (kFwd16, kFwdTrace16) of the
// batched affine-gap alignment engine, for gfx950.  Instantiated per DNA
// profile / substitution table and group size by fwd16.hip.
//
// The recurrence is SeqAn 2's Gotoh formula (include/seqan/align/
// dp_formula_affine.h:62-126) with its borders (dp_formula_affine.h:244-330,
// dp_formula.h:289-309) and result-cell rules (dp_meta_info.h:178-221,
// dp_scout.h:169-210).  It is computed here in "G-space": every lane keeps
// G = H + gapOpen instead of H, because H + gapOpen is the only H-derived term
// both gap recurrences need:
//
//   E(i,j) = max(E(i,j-1) + e, G(i,j-1))
//   F(i,j) = max(F(i-1,j) + e, G(i-1,j))
//   H(i,j) = max(G(i-1,j-1) + t(h,v), E(i,j), F(i,j))   [local: max(.,0)]
//   G(i,j) = H(i,j) + o          with t(h,v) = s(h,v) - o >= 0
//
// so one cell is 9 VALU ops: the score lookup (one v_perm_b32 for DNA), one
// add for the diagonal, add+max for E, add+max for F, two maxes for H and
// one add for G.  Two different pairs share every 32-bit register: the low
// and high int16 halves (v_pk_add_u16 / v_pk_max_i16).
//
// Layout of the packed kernel.  A wavefront takes one work item: up to
// 2*64/G pairs that need the same number of row blocks.  G consecutive lanes
// (a "group") sweep one pair-couple; each lane owns R consecutive rows of a
// block of G*R rows and the group walks the columns as a skewed wavefront:
// at step s lane l computes column j = s - l, receiving the bottom cell of
// lane l-1 (column j, one step old) through a lane shift (DPP row_shr:1,
// G = 16).  The first lane of the group reads the
// bottom row of the previous block from a global row buffer that the last
// lane writes.  Rows are bottom-aligned: row m of every pair is the last row
// of the last block, so the rows above a pair's row 0 are "virtual" (floored
// at the -inf sentinel) and the border row emerges from the recurrence
// itself (TOP free: floored at 0).
#pragma once

#include <algorithm>
#include <type_traits>

#include "kernels.hpp"
#include "packed16.hpp"
#include "seqan_amd.h"

namespace sqmi {
namespace {

// Free-end traceback (trace kernels with bottom / right tracking): the first
// strict maximum of the bottom row in column order and of the last column in
// row order, with E and F of that cell (dp_scout.h:169-210; the end-cell rule
// of dp_algorithm_impl.h:1190-1207 needs them).  Values in G = H + gapOpen,
// E and F in H; positions per half (pair) in int32.
// Local alignment uses the b fields for the best cell over every cell, with
// its item row in bRow.
struct EndTrack {
    s16x2 bV, bE, bF;   // bottom row, held by lane G - 1 of the last block
    int bJ[2];          // its column
    int bRow[2];        // local: its item row
    s16x2 rV, rE, rF;   // last column, every lane over its own rows
    int rRow[2];        // its item row
};

// Four sign bytes from two packed differences: v_perm_b32 selectors 8..11
// replicate bits 15, 31 (of the second operand) and 47, 63 (of the first)
// into whole bytes, giving bytes {sign(t0.x), sign(t0.y), sign(t1.x),
// sign(t1.y)} as 0x00 / 0xFF.
__device__ __forceinline__ uint32_t signBytes(s16x2 t0, s16x2 t1)
{
    return __builtin_amdgcn_perm(asu(t1), asu(t0), 0x0B0A0908u);
}

// Lane l of every group of G lanes receives v from lane l-1; lane 0 of the
// group receives `first`.  One DPP move whose lanes without a source keep
// the old value (`first`): row_shr:1 for rows of 16 lanes, wave_shr:1 for
// the whole wavefront.
template <int G>
__device__ __forceinline__ uint32_t shiftIn(uint32_t v, uint32_t first)
{
    static_assert(G == 16 || G == 64, "groups of 16 lanes or whole wavefronts");
    constexpr int ctrl = G == 16 ? 0x111 : 0x138;
    return (uint32_t)__builtin_amdgcn_update_dpp((int)first, (int)v, ctrl, 0xf, 0xf, false);
}

template <int G>
__device__ __forceinline__ s16x2 groupMax(s16x2 x)
{
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_xor((int)asu(x), d, G);
        x = vmax(x, as16(y));
    }
    return x;
}

__device__ __forceinline__ uint64_t rbLoad(const uint32_t *p)
{
    return __hip_atomic_load(reinterpret_cast<const uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void rbStore(uint32_t *p, uint64_t v)
{
    __hip_atomic_store(reinterpret_cast<uint64_t *>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Progress words of the hand-off between the waves of a workgroup: block b
// has made columns 0..col of its bottom row available.  Tags of later blocks
// compare greater, so one word serves a slot's successive blocks.
__device__ __forceinline__ uint32_t progTag(int b, int col) { return ((uint32_t)b << 16) | (uint32_t)(col + 1); }

// Spins (wave-uniformly: every lane reads the same word) until *w >= target.
// Only waves of the same workgroup publish these words, so the producer is
// resident and the wait ends.
__device__ __forceinline__ void waitAtLeast(uint32_t *w, uint32_t target)
{
    while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target)
        __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ void publish(uint32_t *w, uint32_t v)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if ((threadIdx.x & 63) == 0)
        __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Producer progress words of the cooperative hand-off, one per wave of the
// workgroup (the wave's row buffer is the slot its blocks write).
constexpr int kCoopWaves = 4;
__shared__ uint32_t gProd[kCoopWaves];

// Per-group LDS ring of the profile words of both pairs of a couple, by
// column: 4 G entries per group, entry x holds column x - (G - 1) (lane l
// reads entry s - l + G - 1 at step s).  Refilled every 2 G steps from
// residues loaded one refill earlier, so the column sweep issues no residue
// loads.  The columns live during one refill period span 3 G - 1 entries.
// (256 / G) groups of 4 G entries, each followed by a copy of its first 4.
constexpr int kRingWords = 4 * 256 + (256 / 16) * 4;
__shared__ uint2 gRing[kRingWords];

// Substitution table of the table kernels in LDS, codes h, v < kTabMaxCodes:
// word (h, v / 2) holds the two v of the word in its halves; with C copies,
// copy c of the word sits at byte ((h * VW + v / 2) * C + c) * 4 (VW words
// per h).  Lane c of a half-wave reads copy c mod C, so with C = 32 every
// lane reads its own bank and the lookups are free of bank conflicts.
constexpr int kTabWordsPerRow = (kTabMaxCodes + 1) / 2;
constexpr int tabBytes(int C) { return kTabMaxCodes * kTabWordsPerRow * 4 * C; }
template <int C>
__device__ __forceinline__ uint32_t tabColOffset(int h, int TW)
{
    return (uint32_t)(h * ((TW + 1) / 2) * 4 * C);
}
template <int C>
__device__ __forceinline__ uint32_t tabRowOffset(int v, int lane)
{
    return (uint32_t)((v >> 1) * 4 * C + (v & 1) * 2 + 4 * (lane & (C - 1)));
}

// Per-lane state of one pair couple (lo = pair A, hi = pair B).
struct Couple {
    const uint8_t *h[2];
    const uint8_t *v[2];
    int n[2], m[2], p[2], pid[2];
};

// TB: also write the trace bits of every cell (global alignment only).  Per
// cell four decisions, which is what dp_formula_affine.h:62-126 and
// _maxScore (dp_formula.h:165-201) record, ties resolved the same way:
//   E  E(i,j) opened a gap:    E(i,j-1) + e <  G(i,j-1)   (TraceBitMap_ HORIZONTAL_OPEN)
//   F  F(i,j) opened a gap:    F(i-1,j) + e <  G(i-1,j)   (VERTICAL_OPEN)
//   D  H(i,j) is not diagonal: D < max(E, F)              (not DIAGONAL)
//   X  E beats F:              F < E                      (MAX_FROM_HORIZONTAL_MATRIX)
// Per lane and column step one 16-byte record; bit kk of each byte is row
// kk (x, z: rows 0-7; y, w: rows 8-15) and the bytes are
//   x, y: E of pair A, E of pair B, F of pair A, F of pair B
//   z, w: D of pair A, D of pair B, X of pair A, X of pair B.
// Records of block b, step s of a work item sit at
// trace[(traceOff[item] + b * steps + s) * 64 + lane]: one coalesced 1 KB
// store per wave and step.
template <int G, int R, bool DNA, uint32_t TRK, bool TB = false>
struct Fwd16 {
    static constexpr int BR = G * R;
    static constexpr int GROUPS = 64 / G;
    static constexpr int QD = 4;   // residue prefetch depth (steps)
    // Table copies in LDS: one per bank for the R = 16 table kernels (two
    // waves per SIMD, so the 50 KB fit twice per CU); one for the others.
    static constexpr int TC = (!DNA && R >= 16) ? 32 : 1;
    // LFL (local scores): every value of the sweep is held shifted by
    // -32768 - gapOpen, so that G = gapOpen (H = 0) is -32768, and all adds
    // saturate.  Then G = H + gapOpen, saturated, is G of max(H, 0): the
    // local floor costs no instruction.  In local alignment every E, F and G
    // is >= gapOpen and a D below 0 only matters through the floor, so
    // values clamped at -32768 (-inf included) give the exact result; the
    // shift leaves the int16 range below the window's bound unused.
    static constexpr bool LFL = (TRK & kTrackLocal) && !TB;
    // U3 (local table kernels with every t >= 0): the table holds t << 16 in
    // 32-bit words, TU copies; pair A reads the high half (t, zero-extended),
    // pair B the word, and D = G + tA + (tB << 16) is one 32-bit add of
    // three operands.  The sum never carries between the halves: G + t is a
    // D of the sweep, below 0 in the shifted representation (the int16
    // window), and the padding codes' t is 0 (the cells they reach stay at
    // or below a real cell's value, so the local maximum is unchanged).
    static constexpr bool U3 = !DNA && LFL && (TRK & kLocalAdd3) != 0;
    static constexpr int TU = R >= 16 ? 16 : 1;
    static constexpr int kTabLds = DNA ? 16 : U3 ? kTabMaxCodes * kTabMaxCodes * 4 * TU : tabBytes(TC);
    static_assert(!TB || R == 16, "trace records hold 16 rows");

    __host__ __device__ __forceinline__ static int stepsFor(int Nw) { return ((Nw + G + QD - 1) / QD) * QD; }

    // Long work items are shared by the NW waves of a workgroup: wave w
    // sweeps blocks w, w + NW, ...  Block b writes its bottom row into the
    // row buffer of its wave and publishes its progress in gProd[b % NW];
    // block b + 1 reads that buffer once the progress covers the columns.
    // Block b only ever waits for block b - 1, so the waits end (block 0
    // waits for nothing).  Reusing a wave's buffer for block b + NW needs no
    // wait: block b + NW reaches column j only after blocks b + NW - 1, ...,
    // b + 1 have, and block b + 1 read column j of block b before publishing
    // its own progress past it.
    static constexpr int NW = kCoopWaves;
    static constexpr int PUB = 4;   // publish progress every PUB chunks of QD steps

    // One block of rows: column 0 init plus the skewed sweep over columns.
    // HEAD: block 0 (rows above the pairs' row 0 are floored).  last: the
    // block holding row m (result tracking, no row-buffer output).  rbIn: row
    // buffer holding block b - 1's bottom row; rbOut: where this block's
    // bottom row goes (the same buffer unless coop).  coop: prodIn / prodOut
    // are the progress words of blocks b - 1 and b.  b, last and coop are
    // wave-uniform; keeping them runtime values keeps the kernel to two
    // copies of the column loop (HEAD or not), which fit the instruction
    // cache.
    template <bool HEAD, int RB>
    __device__ __forceinline__ static void block(const FwdArgs &a, const Couple &c, int b, int Nw, int l,
                                                 const bool last, const bool coop, const uint32_t *rbIn,
                                                 uint32_t *rbOut, uint32_t *prodIn, uint32_t *prodOut,
                                                 const uint32_t *sProf, const char *sT, s16x2 &best,
                                                 uint4 *trItem, s16x2 &endE, s16x2 &endF, EndTrack &et)
    {
        const int o = a.gapOpen, e = a.gapExtend;
        const s16x2 o2 = splat(o), e2 = splat(e), neg2 = splat(LFL ? -32768 : kNeg16);
        const bool leftFree = (a.mode & (SQMI_FREE_LEFT | SQMI_LOCAL)) != 0;
        const bool topFree = (a.mode & SQMI_FREE_TOP) != 0;
        const int eo = max(e, o);   // first-column gap chain grows by max(e, o) per row
        const int PAD = a.alpha;
        const int TW = a.alpha + 1;
        // RB rows per lane: block 0 may hold fewer rows than the others
        // (its bottom G * RB rows of the block's G * R); the pairs' rows end
        // at the last block's bottom row either way.
        const int rowbase = (HEAD ? BR - G * RB : b * BR) + l * RB;

        // Rows above a pair's row 0 (block 0 only) are virtual.  The trace
        // kernels floor H of every row of block 0 (rowK = 0 for row 0 under a
        // free top, else -inf).  The score kernels instead compute block 0
        // with saturating adds: the virtual rows start at -inf and every term
        // adds a value <= 0 to them (e, o, and -256 for the row padding
        // code), so they stay in [-32768, kNeg16] and never wrap; row 0 under
        // a free top extends its horizontal gap by 0 from E = 0 (rowK = 0,
        // else e), which keeps E(0, j) = H(0, j) = 0.  Real rows never reach
        // the saturation bounds (the int16 window), so their values are the
        // wrapping adds' values.
        constexpr bool SAT = HEAD && !TB;
        // ROWT: the score kernels take the bottom-row result (free last row:
        // the maximum over row m, j <= n; global: cell (m, n)) from the
        // bottom row that the last block stores like every other block,
        // after its sweep, instead of tracking it per column.
        constexpr bool ROWT = !TB && !(TRK & kTrackLocal) && (TRK & (kTrackBottom | kTrackEnd)) != 0;
        uint32_t sel[RB];
        s16x2 Gk[RB], Ek[RB], rowK[RB];
#pragma unroll
        for (int k = 0; k < RB; ++k) {
            const int r = rowbase + k;
            int code[2], g0[2], fl[2], e0[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const int i = r - c.p[x];
                int cc = PAD;
                if (i >= 1) {
                    const int idx = min(i - 1, max(c.m[x] - 1, 0));
                    cc = min((int)c.v[x][idx], PAD);
                }
                code[x] = cc;
                // G at column 0: virtual rows -inf, border H(0,0)=0, first column.
                g0[x] = (i < 0) ? kNeg16 : (i == 0) ? o : ((leftFree ? 0 : o + (i - 1) * eo) + o);
                fl[x] = (i == 0 && topFree) ? 0 : (TB ? kNeg16 : e);
                e0[x] = (SAT && i == 0 && topFree) ? 0 : kNeg16;
                if constexpr (LFL) {
                    g0[x] = -32768;   // G = gapOpen (H = 0) on column 0 and every -inf
                    e0[x] = -32768;
                }
            }
            if constexpr (DNA) {
                const uint32_t sa = code[0] < 4 ? (uint32_t)(code[0] | 0x0c00) : 0x0d0cu;
                const uint32_t sb = code[1] < 4 ? (uint32_t)((code[1] + 4) | 0x0c00) : 0x0d0cu;
                sel[k] = sa | (sb << 16);
            } else {
                // Byte offsets of the row codes in the table, this lane's
                // copy included.
                const int lane = (int)threadIdx.x;
                if constexpr (U3) {
                    // Word (h, v) copy c at ((h * TW + v) * TU + c) * 4; pair A
                    // reads its high half.
                    const uint32_t cp = 4u * (uint32_t)(lane & (TU - 1));
                    sel[k] = ((uint32_t)code[0] * 4 * TU + cp + 2) | (((uint32_t)code[1] * 4 * TU + cp) << 16);
                } else {
                    sel[k] = tabRowOffset<TC>(code[0], lane) | (tabRowOffset<TC>(code[1], lane) << 16);
                }
            }
            Gk[k] = pack(g0[0], g0[1]);
            Ek[k] = pack(e0[0], e0[1]);
            rowK[k] = pack(fl[0], fl[1]);
        }

        // Row-buffer entry for step 0 (read before this block's column-0 write).
        uint64_t rbv = (uint64_t)asu(neg2) << 32 | asu(neg2);
        const int steps = stepsFor(Nw);
        int known = -1;   // coop: columns of block b - 1 known to be available
        if (coop && b > 0) {
            waitAtLeast(prodIn, progTag(b - 1, 0));
            known = 0;
        }
        if (b > 0)
            rbv = rbLoad(rbIn);
        if ((!last || ROWT) && l == G - 1)
            rbStore(rbOut, ((uint64_t)asu(neg2) << 32) | asu(Gk[RB - 1]));
        if (coop && !last)
            publish(prodOut, progTag(b, 0));

        if constexpr (TRK & kTrackBottom) {
            if (last && l == G - 1) {
                best = vmax(best, Gk[RB - 1]);   // H(m, 0) is a tracked cell
                if constexpr (TB) {
                    // (m, 0) is the first bottom cell in column order.
                    const uint32_t gt = gtMask(Gk[RB - 1], et.bV);
                    et.bV = as16(bfi(gt, asu(Gk[RB - 1]), asu(et.bV)));
                    et.bE = as16(bfi(gt, asu(neg2), asu(et.bE)));
                    et.bF = as16(bfi(gt, asu(neg2), asu(et.bF)));
                    if (gt & 0xffffu)
                        et.bJ[0] = 0;
                    if (gt >> 16)
                        et.bJ[1] = 0;
                }
            }
        }

        uint32_t botG = asu(Gk[RB - 1]), botF = asu(neg2);
        s16x2 upGprev = neg2;
        const s16x2 remBase = pack(c.n[0] + l, c.n[1] + l);
        uint4 *trBlock = nullptr;
        if constexpr (TB)
            trBlock = trItem + (size_t)b * steps * 64;

        auto add = [](s16x2 x, s16x2 y) -> s16x2 {
            if constexpr (SAT || LFL)
                return addSat(x, y);
            else
                return x + y;
        };

        // One column of the lane's RB rows, step s = s0 + u of a chunk.  pA /
        // pB: the column's profile (DNA: t bytes for v = 0..3) or table row
        // offset, for pair A / B.  CHK: the lane's column j is checked
        // against 1..Nw (the chunks at the skewed ends of the sweep); in the
        // other chunks every lane's column lies in 1..Nw.  rq: the chunk's
        // row-buffer entries (column s0 + 1 + u at rq + 2 u); wq: this lane's
        // bottom-row entries (column s0 + u - l at wq + 2 u).
        // Table scores of row k for column words a / b (pair A / B): two LDS
        // reads into the halves of one register.
        auto look = [&](uint32_t a, uint32_t b, int k) -> uint32_t {
            s16x2 r;
            r.x = *reinterpret_cast<const int16_t *>(sT + a + (sel[k] & 0xffffu));
            r.y = *reinterpret_cast<const int16_t *>(sT + b + (sel[k] >> 16));
            return asu(r);
        };
        // Table kernels, steady chunks: the rows' scores of the next column
        // (words nA / nB) are read one step ahead into tq, so that the LDS
        // latency of the lookups is hidden behind a column of arithmetic.
        uint32_t tq[DNA || U3 ? 1 : RB];
        // U3: the lookups run DL rows ahead, through DL slots (tA, tB << 16);
        // the last rows of a steady step read the next column's first rows.
        constexpr int DL = RB % 4 == 0 ? 4 : RB % 2 == 0 ? 2 : 1;
        uint32_t qa[U3 ? DL : 1], qb[U3 ? DL : 1];
        auto ldA = [&](uint32_t col, int k) -> uint32_t {
            return *reinterpret_cast<const uint16_t *>(sT + col + (sel[k] & 0xffffu));
        };
        auto ldB = [&](uint32_t col, int k) -> uint32_t {
            return *reinterpret_cast<const uint32_t *>(sT + col + (sel[k] >> 16));
        };
        auto step = [&](auto chk, int s, int u, uint32_t pA, uint32_t pB, uint32_t nA, uint32_t nB,
                        const uint32_t *rq, uint32_t *wq) {
            constexpr bool CHK = decltype(chk)::value;
            const int j = s - l;
            const s16x2 upG = as16(shiftIn<G>(botG, (uint32_t)rbv));
            const s16x2 upF = as16(shiftIn<G>(botF, (uint32_t)(rbv >> 32)));
            // Prefetch the next row-buffer entry (column s+1 for the group
            // head; block 0's head reads the constant -inf row).
            if constexpr (!HEAD) {
                if (!CHK || s + 1 <= Nw)
                    rbv = rbLoad(rq + 2 * u);
            }
            const s16x2 diagTop = upGprev;
            upGprev = upG;
            if (!CHK || (j >= 1 && j <= Nw)) {
                s16x2 uG = upG, uF = upF;
                uint32_t accEF0 = 0, accEF1 = 0, accDX0 = 0, accDX1 = 0;
                // t(h, v) - gapOpen of row k for both pairs.
                auto tOf = [&](int k) -> uint32_t {
                    if constexpr (DNA) {
                        return __builtin_amdgcn_perm(pB, pA, sel[k]);
                    } else if constexpr (!CHK && !U3) {
                        const uint32_t t = tq[k];
                        tq[k] = look(nA, nB, k);
                        return t;
                    } else {
                        return look(pA, pB, k);
                    }
                };
                // The diagonal term of row k from G(i-1, j-1).
                auto dOf = [&](s16x2 g, int k) -> s16x2 {
                    if constexpr (U3) {
                        uint32_t ta, tb;
                        if constexpr (!CHK) {
                            ta = qa[k % DL];
                            tb = qb[k % DL];
                            const int kn = k + DL;
                            qa[k % DL] = kn < RB ? ldA(pA, kn) : ldA(nA, kn - RB);
                            qb[k % DL] = kn < RB ? ldB(pB, kn) : ldB(nB, kn - RB);
                        } else {
                            ta = ldA(pA, k);
                            tb = ldB(pB, k);
                        }
                        return as16(asu(g) + ta + tb);
                    } else {
                        return add(g, as16(tOf(k)));
                    }
                };
                // D: the diagonal term of row k, G(i-1, j-1) + t, taken from
                // the old G of row k - 1 before it is overwritten (no copies
                // of the column's G values).
                s16x2 D = dOf(diagTop, 0);
#pragma unroll
                for (int k = 0; k < RB; ++k) {
                    s16x2 F, H;
                    if constexpr (TB) {
                        const s16x2 Eext = Ek[k] + e2, Fext = uF + e2;
                        const uint32_t ef = signBytes(subSat(Eext, Gk[k]), subSat(Fext, uG));
                        Ek[k] = vmax(Eext, Gk[k]);
                        F = vmax(Fext, uG);
                        const s16x2 fe = vmax(Ek[k], F);
                        const uint32_t dx = signBytes(subSat(D, fe), subSat(F, Ek[k]));
                        H = vmax(D, fe);
                        // Insert bit kk of every byte.  Written as v_bfi_b32 statements
                        // so that each row's bits are taken while its values are live
                        // (left to itself the compiler defers all 16 rows to one OR tree
                        // at the end of the column and needs ~250 VGPRs).
                        const uint32_t bit = 0x01010101u << (k & 7);
                        if (k < 8) {
                            asm volatile("v_bfi_b32 %0, %1, %2, %0" : "+v"(accEF0) : "s"(bit), "v"(ef));
                            asm volatile("v_bfi_b32 %0, %1, %2, %0" : "+v"(accDX0) : "s"(bit), "v"(dx));
                        } else {
                            asm volatile("v_bfi_b32 %0, %1, %2, %0" : "+v"(accEF1) : "s"(bit), "v"(ef));
                            asm volatile("v_bfi_b32 %0, %1, %2, %0" : "+v"(accDX1) : "s"(bit), "v"(dx));
                        }
                    } else {
                        Ek[k] = vmax(add(Ek[k], SAT ? rowK[k] : e2), Gk[k]);
                        F = vmax(add(uF, e2), uG);
                        H = vmax(vmax(D, Ek[k]), F);
                    }
                    if constexpr ((TRK & kTrackLocal) && TB) {
                        H = vmax(H, splat(0));
                    } else if constexpr (HEAD && TB) {
                        H = vmax(H, rowK[k]);
                    }
                    if (k + 1 < RB)
                        D = dOf(Gk[k], k + 1);
                    // LFL: the saturating add floors G at -32768, i.e. H at 0.
                    const s16x2 Gn = add(H, o2);
                    Gk[k] = Gn;
                    if constexpr ((TRK & kTrackLocal) && !TB && !U3) {
                        // Taken in the row, as a statement: left to itself the
                        // compiler defers a chunk's 4 x RB maxima to its end
                        // and keeps every G of the chunk live for them (the
                        // DNA R = 16 kernel then spills in its loops).
                        uint32_t bw = asu(best);
                        asm volatile("v_pk_max_i16 %0, %0, %1" : "+v"(bw) : "v"(asu(Gn)));
                        best = as16(bw);
                    } else if constexpr ((TRK & kTrackLocal) && !TB) {
                        best = vmax(best, Gn);
                    }
                    uG = Gn;
                    uF = F;
                }
                botG = asu(uG);
                botF = asu(uF);
                if constexpr ((TRK & kTrackLocal) && TB) {
                    // Local traceback: the first strict maximum over every
                    // cell in column order.  The column's maximum first; the
                    // rows are scanned only when it reaches the best so far
                    // (equal values count when they lie in an earlier column:
                    // a later block sweeps columns the earlier ones passed).
                    s16x2 sm = Gk[0];
#pragma unroll
                    for (int k = 1; k < RB; ++k)
                        sm = vmax(sm, Gk[k]);
                    best = vmax(best, sm);
                    if (gtMask(sm + splat(1), et.bV) != 0) {
                        // The column is taken (per half) when its maximum beats
                        // the best so far, or equals it in an earlier column;
                        // then its first row holding the maximum is the cell.
                        const uint32_t earlier = (j < et.bJ[0] ? 0xffffu : 0u) | (j < et.bJ[1] ? 0xffff0000u : 0u);
                        const uint32_t take = gtMask(sm, et.bV) | (eqMask(sm, et.bV) & earlier);
                        if (take != 0) {
                            uint32_t open = 0xffffffffu;   // halves whose first maximal row is still ahead
                            uint32_t rowK = 0;
                            s16x2 selE = neg2, selF = neg2;
                            s16x2 Fr = vmax(upF + e2, upG);
#pragma unroll
                            for (int k = 0; k < RB; ++k) {
                                if (k > 0)
                                    Fr = vmax(Fr + e2, Gk[k - 1]);
                                // 0xffff in halves where Gk[k] < sm.
                                const uint32_t below = asu(subSat(Gk[k], sm) >> splat(15));
                                const uint32_t first = ~below & open;
                                open &= below;
                                rowK = bfi(first, (uint32_t)k * 0x10001u, rowK);
                                selE = as16(bfi(first, asu(Ek[k]), asu(selE)));
                                selF = as16(bfi(first, asu(Fr), asu(selF)));
                            }
                            et.bV = as16(bfi(take, asu(sm), asu(et.bV)));
                            et.bE = as16(bfi(take, asu(selE), asu(et.bE)));
                            et.bF = as16(bfi(take, asu(selF), asu(et.bF)));
                            if (take & 0xffffu) {
                                et.bJ[0] = j;
                                et.bRow[0] = rowbase + (int)(rowK & 0xffffu);
                            }
                            if (take >> 16) {
                                et.bJ[1] = j;
                                et.bRow[1] = rowbase + (int)(rowK >> 16);
                            }
                        }
                    }
                }
                if constexpr ((TRK & kTrackRight) || (!ROWT && (TRK & (kTrackEnd | kTrackBottom)))) {
                    const s16x2 rem = remBase - splat(s);   // n - j per half
                    // eq = 0xffff in halves where rem == 0.
                    const uint16_t r0 = (uint16_t)rem.x, r1 = (uint16_t)rem.y;
                    const uint32_t eq = (r0 == 0 ? 0xffffu : 0u) | (r1 == 0 ? 0xffff0000u : 0u);
                    if constexpr ((TRK & kTrackBottom) && !ROWT) {
                        if (last) {
                            const uint32_t after = asu(rem >> splat(15));   // 0xffff where j > n
                            const s16x2 cand = as16(bfi(after, asu(neg2), asu(uG)));
                            if (l == G - 1) {
                                best = vmax(best, cand);
                                if constexpr (TB) {
                                    // With a free last column, (m, n) is the last cell
                                    // of column n and is tracked there.
                                    const uint32_t skip = (TRK & kTrackRight) ? (after | eq) : after;
                                    const s16x2 cb = as16(bfi(skip, asu(neg2), asu(uG)));
                                    const uint32_t gt = gtMask(cb, et.bV);
                                    et.bV = as16(bfi(gt, asu(cb), asu(et.bV)));
                                    et.bE = as16(bfi(gt, asu(Ek[RB - 1]), asu(et.bE)));
                                    et.bF = as16(bfi(gt, asu(uF), asu(et.bF)));
                                    if (gt & 0xffffu)
                                        et.bJ[0] = j;
                                    if (gt >> 16)
                                        et.bJ[1] = j;
                                }
                            }
                        }
                    }
                    if constexpr (TRK & (kTrackEnd | kTrackRight)) {
                        if constexpr ((TRK & kTrackEnd) && !ROWT) {
                            if (last && l == G - 1) {
                                best = as16(bfi(eq, asu(uG), asu(best)));
                                if constexpr (TB) {
                                    endE = as16(bfi(eq, asu(Ek[RB - 1]), asu(endE)));
                                    endF = as16(bfi(eq, asu(uF), asu(endF)));
                                }
                            }
                        }
                        if constexpr (TRK & kTrackRight) {
                            if (eq != 0) {
                                s16x2 cm = Gk[0];
#pragma unroll
                                for (int k = 1; k < RB; ++k)
                                    cm = vmax(cm, Gk[k]);
                                best = vmax(best, as16(bfi(eq, asu(cm), asu(neg2))));
                                if constexpr (TB) {
                                    // The lane's rows of column n in row order; F of
                                    // each row again from the column above it.
                                    s16x2 Fr = vmax(upF + e2, upG);
#pragma unroll
                                    for (int k = 0; k < RB; ++k) {
                                        if (k > 0)
                                            Fr = vmax(Fr + e2, Gk[k - 1]);
                                        const s16x2 cr = as16(bfi(eq, asu(Gk[k]), asu(neg2)));
                                        const uint32_t gt = gtMask(cr, et.rV);
                                        et.rV = as16(bfi(gt, asu(cr), asu(et.rV)));
                                        et.rE = as16(bfi(gt, asu(Ek[k]), asu(et.rE)));
                                        et.rF = as16(bfi(gt, asu(Fr), asu(et.rF)));
                                        if (gt & 0xffffu)
                                            et.rRow[0] = rowbase + k;
                                        if (gt >> 16)
                                            et.rRow[1] = rowbase + k;
                                    }
                                }
                            }
                        }
                    }
                }
                if ((!last || ROWT) && l == G - 1)
                    rbStore(wq + 2 * u, ((uint64_t)botF << 32) | botG);
                if constexpr (TB)
                    trBlock[(size_t)s * 64] = make_uint4(accEF0, accEF1, accDX0, accDX1);
            }
        };

        // Residue byte of pair x at column jj (the address is clamped so the
        // load is always in bounds), and its code (padding outside 1..n).
        auto hraw = [&](int x, int jj) -> uint32_t {
            return c.h[x][min(max(jj, 1), max(c.n[x], 1)) - 1];
        };
        auto hcode = [&](int x, int jj, uint32_t raw) -> int {
            return (jj >= 1 && jj <= c.n[x]) ? min((int)raw, PAD) : PAD;
        };
        auto prof = [&](int code) -> uint32_t {
            if constexpr (DNA)
                return sProf[code];
            else if constexpr (U3)
                return (uint32_t)(code * TW * 4 * TU);
            else
                return tabColOffset<TC>(code, TW);
        };
        // The couple's profile words come from the group's LDS ring.  A
        // refill for steps [s, s + kRefill) writes columns s + 2 l and
        // s + 2 l + 1 of lane l; their residues were loaded a refill earlier.
        // Entries 0 .. QD - 1 are mirrored after the ring, so that a chunk
        // reads its QD entries from one base without a wrap.
        constexpr int kRing = 4 * G, kRefill = 2 * G, kStride = kRing + QD;
        static_assert(kRefill % QD == 0 && (kRing & (kRing - 1)) == 0, "ring refill covers 2 columns per lane");
        static_assert((256 / G) * kStride <= kRingWords, "ring layout");
        uint2 *ring = gRing + (threadIdx.x / G) * kStride;
        uint32_t hr[4];
        auto refillLoad = [&](int s) {
            const int j0 = s + 2 * l;
            hr[0] = hraw(0, j0);
            hr[1] = hraw(0, j0 + 1);
            hr[2] = hraw(1, j0);
            hr[3] = hraw(1, j0 + 1);
        };
        auto refillStore = [&](int s) {
            const int j0 = s + 2 * l;
            const int x0 = (s + G - 1 + 2 * l) & (kRing - 1), x1 = (x0 + 1) & (kRing - 1);
            const uint2 w0 = make_uint2(prof(hcode(0, j0, hr[0])), prof(hcode(1, j0, hr[2])));
            const uint2 w1 = make_uint2(prof(hcode(0, j0 + 1, hr[1])), prof(hcode(1, j0 + 1, hr[3])));
            ring[x0] = w0;
            ring[x1] = w1;
            if (x0 < QD)
                ring[x0 + kRing] = w0;
            if (x1 < QD)
                ring[x1 + kRing] = w1;
            __builtin_amdgcn_wave_barrier();
        };
        refillLoad(0);
        refillStore(0);
        refillLoad(kRefill);
        uint2 pp = ring[(G - 1 - l) & (kRing - 1)];
        // One chunk of QD steps from s0: the cooperative wait, the ring
        // refill, the steps, the cooperative progress.
        auto chunk = [&](auto chk, int s0) {
            if (coop) {
                // This chunk reads columns up to s0 + QD of block b - 1.
                const int need = min(s0 + QD, Nw);
                if (b > 0 && need > known) {
                    waitAtLeast(prodIn, progTag(b - 1, need));
                    known = (int)(__hip_atomic_load(prodIn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & 0xFFFFu) - 1;
                }
            }
            if ((s0 & (kRefill - 1)) == kRefill - QD) {
                // The ring for the next kRefill steps (read from the next step
                // on), and the residues of the refill after it.
                refillStore(s0 + QD);
                refillLoad(s0 + QD + kRefill);
            }
            const uint2 *rr = ring + ((s0 + G - l) & (kRing - 1));   // rr[u]: the words of step s0 + u + 1
            const uint32_t *rq = rbIn + 2 * (s0 + 1);
            uint32_t *wq = rbOut + 2 * (s0 - l);
#pragma unroll
            for (int u = 0; u < QD; ++u) {
                const uint2 ppN = rr[u];
                step(chk, s0 + u, u, pp.x, pp.y, ppN.x, ppN.y, rq, wq);
                pp = ppN;
            }
            if (coop) {
                // Columns up to s0 + QD - G of this block are written.
                if (!last && ((s0 / QD) % PUB == PUB - 1 || s0 + QD >= steps))
                    publish(prodOut, progTag(b, max(0, min(s0 + QD - G, Nw))));
            }
        };
        // Three loops, so that the steady chunks keep their registers from
        // chunk to chunk: the head (some lane's column below 1), the steady
        // chunks (every lane's column in 1..Nw), the tail (some lane past
        // Nw).  The trace kernels check every chunk.
        int s0 = 0;
        if constexpr (!TB) {
            for (; s0 < min(G, steps); s0 += QD)
                chunk(std::true_type{}, s0);
            if constexpr (U3) {
                if (s0 + QD - 1 <= Nw) {
#pragma unroll
                    for (int k = 0; k < DL; ++k) {
                        qa[k] = ldA(pp.x, k);   // the first steady step's first rows
                        qb[k] = ldB(pp.y, k);
                    }
                }
            } else if constexpr (!DNA) {
                if (s0 + QD - 1 <= Nw) {
#pragma unroll
                    for (int k = 0; k < RB; ++k)
                        tq[k] = look(pp.x, pp.y, k);   // the first steady step's scores
                }
            }
            for (; s0 + QD - 1 <= Nw; s0 += QD)
                chunk(std::false_type{}, s0);
        }
        for (; s0 < steps; s0 += QD)
            chunk(std::true_type{}, s0);
        // Make this block's row-buffer writes visible to the next block.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if constexpr (ROWT) {
            if (last) {
                // Row m of both pairs: G(m, j) is the low word of row-buffer
                // entry j (column 0 came from the block's init).
                if constexpr (TRK & kTrackBottom) {
                    // The lanes share the columns 1..Nw; a half takes the
                    // columns up to its pair's n (-32768 beyond).
                    const s16x2 nn = pack(c.n[0], c.n[1]);
                    for (int j = 1 + l; j <= Nw; j += G) {
                        const s16x2 g = as16((uint32_t)rbLoad(rbOut + 2 * j));
                        const uint32_t lim = asu((nn - splat(j)) >> splat(15)) ^ 0x7fff7fffu;
                        best = vmax(best, vmin(g, as16(lim)));
                    }
                } else {
                    // The end cell (m, n) of each pair.
                    if (l == G - 1) {
                        const uint32_t ga = (uint32_t)rbLoad(rbOut + 2 * c.n[0]);
                        const uint32_t gb = (uint32_t)rbLoad(rbOut + 2 * c.n[1]);
                        best = as16((ga & 0xffffu) | (gb & 0xffff0000u));
                    }
                }
            }
        }
    }

    __device__ __forceinline__ static bool wellFormed(const FwdArgs &a, const WorkItem &w)
    {
        // A malformed descriptor is skipped rather than allowed to run past
        // the row buffer or loop without bound.
        const int pairs = w.count & 0xffff;
        return !(pairs < 1 || pairs > 2 * GROUPS || w.maxN < 1 || w.maxN + G + 2 > a.rowCols ||
                 w.nblocks < 1 || w.nblocks > (1 << 20) / BR + 1);
    }

    __device__ __forceinline__ static Couple couple(const FwdArgs &a, const WorkItem &w, int grp)
    {
        Couple c;
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const int sidx = 2 * grp + x;
            const int pid = sidx < (w.count & 0xffff) ? a.order[w.first + sidx] : -1;
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
            c.p[x] = w.nblocks * BR - 1 - m;
        }
        return c;
    }

    __device__ __forceinline__ static void run(const FwdArgs &a)
    {
        __shared__ uint32_t sProf[8];
        __shared__ __align__(16) char sT[kTabLds];
        __shared__ uint32_t sBest[NW][GROUPS];
        __shared__ int sItem;
        if constexpr (DNA) {
            if (threadIdx.x < 5)
                sProf[threadIdx.x] = a.dnaProfile[threadIdx.x];
        } else if constexpr (U3) {
            // Word (h, v) of copy cc at byte ((h * TW + v) * TU + cc) * 4:
            // t << 16, t = 0 for the padding code.
            const int TW = a.alpha + 1;
            for (int i = threadIdx.x; i < TW * TW * TU; i += blockDim.x) {
                const int w = i / TU, h = w / TW, v = w % TW;
                const uint32_t t = (h == a.alpha || v == a.alpha) ? 0u : (uint16_t)a.tTable[w];
                *reinterpret_cast<uint32_t *>(sT + 4 * i) = t << 16;
            }
        } else {
            // Word (h, v / 2) of copy cc at byte ((h * VW + v / 2) * TC + cc) * 4.
            const int TW = a.alpha + 1, VW = (TW + 1) / 2;
            for (int i = threadIdx.x; i < TW * VW * TC; i += blockDim.x) {
                const int w = i / TC, h = w / VW, v0 = 2 * (w % VW);
                const uint32_t lo = (uint16_t)a.tTable[h * TW + v0];
                const uint32_t hi = v0 + 1 < TW ? (uint16_t)a.tTable[h * TW + v0 + 1] : 0u;
                *reinterpret_cast<uint32_t *>(sT + 4 * i) = lo | (hi << 16);
            }
        }
        __syncthreads();

        const int lane = threadIdx.x & 63;
        const int l = lane % G;
        const int grp = lane / G;
        const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
        uint32_t *rowbuf = a.rowbuf + ((size_t)(blockIdx.x * (blockDim.x >> 6) + wave) * GROUPS + grp) *
                                          (size_t)a.rowCols * 2;

        // Items [0, nLong) (the ones with the most row blocks) go one per
        // workgroup, their blocks spread over its NW waves; the rest go one
        // per wave.  The long phase runs first and ends for all waves of the
        // workgroup at once (they read the same item index).
        const int nLong = (!TB && blockDim.x == 64 * NW) ? a.nLong : 0;
        const uint32_t *rbPrev = nLong > 0 ? a.rowbuf + ((size_t)(blockIdx.x * NW + (wave + NW - 1) % NW) * GROUPS +
                                                         grp) * (size_t)a.rowCols * 2
                                           : rowbuf;
        bool longPhase = nLong > 0;
        for (;;) {
            int it;
            bool coop;
            if (longPhase) {
                if (threadIdx.x == 0)
                    sItem = (int)atomicAdd(a.queueLong, 1u);
                if (threadIdx.x < NW)
                    gProd[threadIdx.x] = 0;
                __syncthreads();
                it = sItem;
                __syncthreads();   // every wave has read sItem before thread 0 writes it again
                if (it >= nLong) {
                    longPhase = false;
                    continue;
                }
                coop = true;
            } else {
                // Every lane takes part in the atomic (lane 0 adds 1, the
                // others 0) and the wave takes lane 0's old value.  A
                // branch on lane == 0 here could be merged with the
                // lane-0 store at the end of the loop (the same predicate
                // when G = 64) into divergent paths around the loop's back
                // edge, where this readfirstlane would then read a lane
                // that holds no item index.
                it = (int)atomicAdd(a.queue, lane == 0 ? 1u : 0u);
                it = __builtin_amdgcn_readfirstlane(it) + a.nLong;
                if (it >= a.nItems)
                    break;
                coop = false;
            }
            const WorkItem w = a.items[it];
            const bool ok = wellFormed(a, w);
            const Couple c = couple(a, w, grp);
            s16x2 best = LFL ? splat(-32768) : (TRK & kTrackLocal) ? splat(a.gapOpen) : splat(kNeg16);
            s16x2 endE = splat(kNeg16), endF = splat(kNeg16);
            EndTrack et;
            et.bV = et.bE = et.bF = et.rV = et.rE = et.rF = splat(kNeg16);
            if constexpr (TRK & kTrackLocal)
                et.bV = splat(a.gapOpen + 1);   // G of H = 1: a cell must beat 0 (taken while bJ is unset)
            et.bJ[0] = et.bJ[1] = (TRK & kTrackLocal) ? 0x7fffffff : 0;
            et.bRow[0] = et.bRow[1] = 0;
            et.rRow[0] = et.rRow[1] = 0x7fffffff;
            const int Nw = w.maxN;
            uint4 *trItem = nullptr;
            if constexpr (TB)
                trItem = a.trace + (size_t)a.traceOff[it] * 64 + lane;
            // Blocks of this wave: all of them (one item per wave) or wave,
            // wave + NW, ... (coop: block b reads the buffer of wave
            // (b - 1) % NW, its progress in gProd[(b - 1) % NW], and publishes
            // in gProd[b % NW]).
            const int b0 = coop ? wave : 0, bs = coop ? NW : 1;
            const uint32_t *rbIn = coop ? rbPrev : rowbuf;
            uint32_t *prodIn = &gProd[(wave + NW - 1) % NW], *prodOut = &gProd[wave];
            int b = ok ? b0 : w.nblocks;
            if (b == 0) {
                // Block 0 holds 1..4 quarters of the block's rows (bits 16..17
                // of the item's count; 0: all four).
                auto head = [&](auto rb) {
                    block<true, decltype(rb)::value>(a, c, 0, Nw, l, w.nblocks == 1, coop, rbIn, rowbuf, prodIn,
                                                     prodOut, sProf, sT, best, trItem, endE, endF, et);
                };
                if constexpr (TB || R % 4 != 0) {
                    head(std::integral_constant<int, R>{});
                } else {
                    switch ((w.count >> 16) & 3) {
                        case 1: head(std::integral_constant<int, R / 4>{}); break;
                        case 2: head(std::integral_constant<int, R / 2>{}); break;
                        case 3: head(std::integral_constant<int, 3 * R / 4>{}); break;
                        default: head(std::integral_constant<int, R>{}); break;
                    }
                }
                b += bs;
            }
            for (; b < w.nblocks; b += bs)
                block<false, R>(a, c, b, Nw, l, b == w.nblocks - 1, coop, rbIn, rowbuf, prodIn, prodOut, sProf, sT,
                                best, trItem, endE, endF, et);
            if constexpr (TB && TRK == kTrackEnd) {
                // E and F of the end cell (m, n), for the end-cell rule of
                // the traceback (dp_algorithm_impl.h:1190-1207).
                if (ok && l == G - 1) {
                    if (c.pid[0] >= 0) {
                        a.endEF[2 * c.pid[0]] = (int)endE.x;
                        a.endEF[2 * c.pid[0] + 1] = (int)endF.x;
                    }
                    if (c.pid[1] >= 0) {
                        a.endEF[2 * c.pid[1]] = (int)endE.y;
                        a.endEF[2 * c.pid[1] + 1] = (int)endF.y;
                    }
                }
            } else if constexpr (TB && (TRK & kTrackLocal)) {
                // Local: reduce the lanes' candidates (larger value, then
                // earlier column, then earlier row); no cell above 0 leaves
                // the end cell at (0, 0), an empty window.
                int bv[2] = {(int)et.bV.x, (int)et.bV.y}, bj[2] = {et.bJ[0], et.bJ[1]};
                int br[2] = {et.bRow[0], et.bRow[1]};
                int be[2] = {(int)et.bE.x, (int)et.bE.y}, bf[2] = {(int)et.bF.x, (int)et.bF.y};
#pragma unroll
                for (int d = 1; d < G; d <<= 1) {
#pragma unroll
                    for (int x = 0; x < 2; ++x) {
                        const int ov = __shfl_xor(bv[x], d, G), oj = __shfl_xor(bj[x], d, G);
                        const int orr = __shfl_xor(br[x], d, G);
                        const int oe = __shfl_xor(be[x], d, G), of = __shfl_xor(bf[x], d, G);
                        const bool take = ov > bv[x] || (ov == bv[x] && (oj < bj[x] || (oj == bj[x] && orr < br[x])));
                        bv[x] = take ? ov : bv[x];
                        bj[x] = take ? oj : bj[x];
                        br[x] = take ? orr : br[x];
                        be[x] = take ? oe : be[x];
                        bf[x] = take ? of : bf[x];
                    }
                }
                if (ok && l == G - 1) {
#pragma unroll
                    for (int x = 0; x < 2; ++x) {
                        const int pid = c.pid[x];
                        if (pid < 0)
                            continue;
                        const bool any = bj[x] != 0x7fffffff;
                        a.endEF[2 * pid] = be[x];
                        a.endEF[2 * pid + 1] = bf[x];
                        a.endIJ[2 * pid] = any ? br[x] - c.p[x] : 0;
                        a.endIJ[2 * pid + 1] = any ? bj[x] : 0;
                    }
                }
            } else if constexpr (TB) {
                // Free ends: the end cell is the first strict maximum over the
                // tracked cells in column order, the bottom row's cells (m, j),
                // j < n, before column n's cells (i, n).  The last column's
                // candidates are spread over the group's lanes: reduce them
                // (larger value, then smaller row), then lane G - 1, which
                // holds the bottom row's, picks between the two.
                int rv[2] = {(int)et.rV.x, (int)et.rV.y}, rr[2] = {et.rRow[0], et.rRow[1]};
                int re[2] = {(int)et.rE.x, (int)et.rE.y}, rf[2] = {(int)et.rF.x, (int)et.rF.y};
#pragma unroll
                for (int d = 1; d < G; d <<= 1) {
#pragma unroll
                    for (int x = 0; x < 2; ++x) {
                        const int ov = __shfl_xor(rv[x], d, G), orr = __shfl_xor(rr[x], d, G);
                        const int oe = __shfl_xor(re[x], d, G), of = __shfl_xor(rf[x], d, G);
                        const bool take = ov > rv[x] || (ov == rv[x] && orr < rr[x]);
                        rv[x] = take ? ov : rv[x];
                        rr[x] = take ? orr : rr[x];
                        re[x] = take ? oe : re[x];
                        rf[x] = take ? of : rf[x];
                    }
                }
                const int bv[2] = {(int)et.bV.x, (int)et.bV.y}, be[2] = {(int)et.bE.x, (int)et.bE.y};
                const int bf[2] = {(int)et.bF.x, (int)et.bF.y};
                if (ok && l == G - 1) {
#pragma unroll
                    for (int x = 0; x < 2; ++x) {
                        const int pid = c.pid[x];
                        if (pid < 0)
                            continue;
                        bool bottom = (TRK & kTrackBottom) != 0;
                        if constexpr ((TRK & kTrackBottom) && (TRK & kTrackRight))
                            bottom = bv[x] >= rv[x];
                        a.endEF[2 * pid] = bottom ? be[x] : re[x];
                        a.endEF[2 * pid + 1] = bottom ? bf[x] : rf[x];
                        a.endIJ[2 * pid] = bottom ? c.m[x] : rr[x] - c.p[x];
                        a.endIJ[2 * pid + 1] = bottom ? et.bJ[x] : c.n[x];
                    }
                }
            }
            best = groupMax<G>(best);
            if (coop) {
                // Combine the waves' results of the item.
                if (l == 0)
                    sBest[wave][grp] = asu(best);
                __syncthreads();
                if (wave == 0) {
#pragma unroll
                    for (int v = 1; v < NW; ++v)
                        best = vmax(best, as16(sBest[v][grp]));
                }
                __syncthreads();
            }
            if (ok && l == 0 && (!coop || wave == 0)) {
                // H = G - gapOpen; LFL: G = stored + 32768 + gapOpen.
                const int toH = LFL ? 32768 : -a.gapOpen;
                if (c.pid[0] >= 0)
                    a.scores[c.pid[0]] = (int)best.x + toH;
                if (c.pid[1] >= 0)
                    a.scores[c.pid[1]] = (int)best.y + toH;
            }
        }
    }
};

// Waves per SIMD the register budget is set for: three (168 VGPRs) for the
// DNA kernels and the R = 8 table kernels; two (256 VGPRs) for the R = 16
// table kernels, whose table-row offsets and scores read one column ahead
// would spill at three.
template <int R, bool DNA>
constexpr int kFwdWaves = (!DNA && R >= 16) ? 2 : 3;

template <int G, int R, bool DNA, uint32_t TRK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFwdWaves<R, DNA>))) void kFwd16(FwdArgs a)
{
    Fwd16<G, R, DNA, TRK>::run(a);
}

// Forward pass with trace records (global alignment under any AlignConfig,
// G = 16, R = 16).
template <bool DNA, uint32_t TRK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void kFwdTrace16(FwdArgs a)
{
    Fwd16<16, 16, DNA, TRK, true>::run(a);
}

template <int G, int R, bool DNA>
hipError_t launchTrk(const Variant16 &v, const FwdArgs &args, int gridBlocks, hipStream_t stream,
                     int threadsPerBlock)
{
    dim3 grid(gridBlocks), block(threadsPerBlock);
    switch (v.track) {
        case kTrackEnd:
            hipLaunchKernelGGL((kFwd16<G, R, DNA, kTrackEnd>), grid, block, 0, stream, args);
            break;
        case kTrackBottom:
            hipLaunchKernelGGL((kFwd16<G, R, DNA, kTrackBottom>), grid, block, 0, stream, args);
            break;
        case kTrackRight:
            hipLaunchKernelGGL((kFwd16<G, R, DNA, kTrackRight>), grid, block, 0, stream, args);
            break;
        case kTrackBottom | kTrackRight:
            hipLaunchKernelGGL((kFwd16<G, R, DNA, kTrackBottom | kTrackRight>), grid, block, 0, stream, args);
            break;
        case kTrackLocal:
            if constexpr (!DNA) {
                if (args.tNonNeg) {
                    hipLaunchKernelGGL((kFwd16<G, R, DNA, kTrackLocal | kLocalAdd3>), grid, block, 0, stream, args);
                    break;
                }
            }
            hipLaunchKernelGGL((kFwd16<G, R, DNA, kTrackLocal>), grid, block, 0, stream, args);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int G, int R, bool DNA>
const void *kernelTrk(uint32_t track)
{
    switch (track) {
        case kTrackEnd: return (const void *)&kFwd16<G, R, DNA, kTrackEnd>;
        case kTrackBottom: return (const void *)&kFwd16<G, R, DNA, kTrackBottom>;
        case kTrackRight: return (const void *)&kFwd16<G, R, DNA, kTrackRight>;
        case kTrackBottom | kTrackRight: return (const void *)&kFwd16<G, R, DNA, kTrackBottom | kTrackRight>;
        case kTrackLocal: return (const void *)&kFwd16<G, R, DNA, kTrackLocal>;
        default: return nullptr;
    }
}

// The compiled (G, R) family of this template: a pair couple per group of
// 16 lanes, or per wavefront (G = 64: the short-batch shapes, four times the
// lanes per pair for the same rows).  The one-pair-per-lane form is a
// separate kernel (lane.hip).
#define SQMI_FOR_EACH_GR(X) X(16, 16) X(16, 8) X(64, 4) X(64, 8) X(64, 16)

}  // namespace
}  // namespace sqmi
