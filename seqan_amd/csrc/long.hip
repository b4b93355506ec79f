// long.hip -- int32 kernels for long pairs (global alignment, DNA profile):
// the tiled forward that keeps every tile border as a checkpoint, and the
// checkpointed traceback that recomputes one tile per pair at a time.
//
// The recurrence, borders and traceback decisions are those of the packed
// kernels (kernels.hip), in int32: SeqAn 2's Gotoh formula
// (include/seqan/align/dp_formula_affine.h:62-126), the global borders
// (dp_formula_affine.h:244-330), the trace decisions with the tie rules of
// _maxScore (dp_formula.h:165-201), and the walk of dp_traceback_impl.h:
// 115-546 counted like computeAlignmentStats (evaluate_alignment.h:215-300).
//
// A tile is kTileRows rows (one row block of the pair, bottom-aligned) by
// tileC columns.  One group of 16 lanes sweeps a tile: lane l owns rows
// 16 l .. 16 l + 15 of the block and handles column c0 - 1 + s - l at step s
// (a skewed wavefront; the cell above the lane's first row comes from lane
// l - 1 by DPP row_shr:1, lane 0 reads the block above from the row
// checkpoint).  Inputs: the left column (G = H + gapOpen and E of the
// block's rows at column c0 - 1) and the top row (G and F of the block
// above at columns c0 - 1 .. c1).  Outputs: the bottom row and the right
// column (forward), or 8-byte trace records per lane and step (traceback).
#include "kernels.hpp"
#include "seqan_amd.h"

namespace sqmi {
namespace {

constexpr int kNegL = -(1 << 29);   // minus infinity; real values stay far above
constexpr int TG = 16;              // lanes per group
constexpr int TR = kTileRows / TG;  // rows per lane
constexpr int TQD = 4;              // steps per unrolled loop body
constexpr int TK = 32;              // steps per refill of a group's LDS rings
constexpr int TRING = 64;           // ring entries (two refills)

// Per-group LDS rings of the tile sweep.  top[s % TRING]: the block above's
// bottom row (G, F) at column c0 - 1 + s (lane 0's top input at step s);
// pr[x % TRING]: the substitution profile word of column c0 - 1 + x - 15
// (lane l reads entry s - l + 15 at step s).  Both are refilled every TK
// steps from loads issued TK steps earlier, so the sweep itself issues no
// global loads.
struct GroupRings {
    int2 top[TRING];
    uint32_t pr[TRING];
};

__device__ __forceinline__ int shiftIn16(int v, int first)
{
    return __builtin_amdgcn_update_dpp(first, v, 0x111, 0xf, 0xf, false);
}

__device__ __forceinline__ int subSat32(int a, int b) { return __builtin_elementwise_sub_sat(a, b); }

// Bit k of 32-bit mask pairs: (1 << k) | (1 << (16 + k)).
__device__ __forceinline__ uint32_t rowBits(int k) { return 0x00010001u << k; }

struct TileGeom {
    int b, t;            // row block and column tile
    int c0, c1;          // columns c0 .. c1 of the tile (1-based)
    bool active;         // group has a tile to sweep
};

// One tile sweep for the lane's group.  TRACE: write trace records instead
// of the forward outputs.  HEAD: the tile may hold rows above the pair's row
// 0 (block 0), which are floored at -inf (row 0 at 0 under a free top).
// LOCAL: Smith-Waterman (every H floored at 0, free borders), the maximum G
// of the tile's cells folded into the block's slot of a.best.  The other AlignConfigs
// (forward only): a free left border starts every row at H = 0; the free
// last row and column are reduced from the checkpoints afterwards.
template <bool HEAD, bool TRACE, bool LOCAL = false>
__device__ __forceinline__ void tileSweep(const TileArgs &a, const LongPair &P, const TileGeom &g, int l,
                                          const uint32_t *sProf, GroupRings *rg, uint2 *rec)
{
    const int o = a.gapOpen, e = a.gapExtend, eo = max(o, e);
    const bool leftFree = !TRACE && (LOCAL || (a.mode & SQMI_FREE_LEFT));
    const bool topFree = !TRACE && (a.mode & SQMI_FREE_TOP);
    const int C = a.tileC;
    const uint8_t *h = a.hRes + a.hOff[P.pid];
    const uint8_t *v = a.vRes + a.vOff[P.pid];
    const int n = P.n, m = P.m;
    const int rowbase = g.b * kTileRows + l * TR;
    int32_t *ck = a.ck;
    const int64_t rowIn = P.rowck + ((int64_t)(g.b - 1) * (n + 1)) * 2;   // block b - 1's bottom row
    const int64_t rowOut = P.rowck + ((int64_t)g.b * (n + 1)) * 2;
    const int64_t colIn = P.colck + (((int64_t)g.b * P.ntiles + (g.t - 1)) * kTileRows + l * TR) * 2;
    const int64_t colOut = P.colck + (((int64_t)g.b * P.ntiles + g.t) * kTileRows + l * TR) * 2;

    uint32_t sel[TR];
    int Gk[TR], Ek[TR];
    int floorK[HEAD && !LOCAL ? TR : 1];
#pragma unroll
    for (int k = 0; k < TR; ++k) {
        const int i = rowbase + k - P.p;
        int code = 4;
        if (i >= 1 && i <= m)
            code = min((int)v[i - 1], 4);
        sel[k] = code < 4 ? (0x0c0c0c00u | (uint32_t)code) : 0x0d0d0d0cu;   // pad: t = -256
        if constexpr (HEAD && !LOCAL)
            floorK[k] = (i == 0 && topFree) ? 0 : kNegL;
        if (g.t == 0) {
            // Column 0: -inf above row 0 (local: 0), H(0,0) = 0, then the gap
            // chain growing by max(o, e) per row (free left border: H = 0).
            Gk[k] = (i < 0) ? (LOCAL ? o : kNegL) : (i == 0 || leftFree) ? o : (o + (i - 1) * eo) + o;
            Ek[k] = kNegL;
        } else if (g.active) {
            Gk[k] = ck[colIn + 2 * k];
            Ek[k] = ck[colIn + 2 * k + 1];
        } else {
            Gk[k] = kNegL;
            Ek[k] = kNegL;
        }
    }
    // Top input of the group head at column c0 - 1 + s (the bottom row of
    // block b - 1).  The loads are unconditional: columns past c1 read column
    // c1, and block 0 (no block above) reads its own output row; those values
    // are replaced by -inf.
    const int32_t *topRow = ck + (g.b > 0 ? rowIn : rowOut);
    const bool topReal = g.b > 0 && g.active;
    auto topLoad = [&](int col) -> int2 { return *reinterpret_cast<const int2 *>(topRow + 2 * min(col, g.c1)); };
    if (!TRACE && g.active && g.t == 0 && l == TG - 1)
        *reinterpret_cast<int2 *>(ck + rowOut) = make_int2(Gk[TR - 1], kNegL);   // column 0 of the bottom row

    int botG = Gk[TR - 1], botF = kNegL;
    int upGprev = kNegL;
    int bestG = o;   // local: H = 0, the empty alignment
    const int width = g.c1 - g.c0 + 1;
    const int steps = ((width + TG + TQD - 1) / TQD) * TQD;
    // Ring refills: the refill for steps [s, s + TK) covers columns
    // c0 - 1 + s + 2 l and + 1 of lane l (top entries s + 2 l, profile
    // entries s + 15 + 2 l); residue indexes are clamped into H (columns
    // outside it are never computed).
    auto hLoad = [&](int col) -> uint32_t { return (uint32_t)h[min(max(col, 1), n) - 1]; };
    int2 t0, t1;
    uint32_t h0, h1;
    auto refillLoad = [&](int s) {
        const int cb = g.c0 - 1 + s + 2 * l;
        t0 = topLoad(cb);
        t1 = topLoad(cb + 1);
        h0 = hLoad(cb);
        h1 = hLoad(cb + 1);
    };
    auto refillStore = [&](int s) {
        rg->top[(s + 2 * l) & (TRING - 1)] = t0;
        rg->top[(s + 2 * l + 1) & (TRING - 1)] = t1;
        rg->pr[(s + TG - 1 + 2 * l) & (TRING - 1)] = sProf[min(h0, 4u)];
        rg->pr[(s + TG + 2 * l) & (TRING - 1)] = sProf[min(h1, 4u)];
        __builtin_amdgcn_wave_barrier();
    };
    refillLoad(0);
    refillStore(0);
    refillLoad(TK);
    uint32_t prof = rg->pr[(TG - 1 - l) & (TRING - 1)];
    int2 topv = rg->top[0];

    auto step = [&](int s, uint32_t pr, int2 top) {
        const int jj = g.c0 - 1 + s - l;
        const bool topOk = topReal && g.c0 - 1 + s <= g.c1;
        const int upG = shiftIn16(botG, topOk ? top.x : kNegL);
        const int upF = shiftIn16(botF, topOk ? top.y : kNegL);
        const int diagTop = upGprev;
        upGprev = upG;
        if (jj >= g.c0 && jj <= g.c1) {
            int uG = upG, uF = upF;
            uint32_t accEF = 0, accDX = 0;
            // D: the diagonal term of row k, H(i-1, j-1) + s(h, v), taken
            // from the old G of row k - 1 before it is overwritten.
            int D = diagTop + (int)__builtin_amdgcn_perm(0u, pr, sel[0]);
#pragma unroll
            for (int k = 0; k < TR; ++k) {
                const int Eext = Ek[k] + e, Fext = uF + e;
                const int Gleft = Gk[k];
                const int E = max(Eext, Gleft);
                const int F = max(Fext, uG);
                int H;
                if constexpr (TRACE) {
                    // Sign bytes of the decisions: E opened, F opened (bits k,
                    // 16 + k of accEF); not diagonal, E beats F (accDX).
                    const uint32_t ef = __builtin_amdgcn_perm((uint32_t)subSat32(Fext, uG),
                                                              (uint32_t)subSat32(Eext, Gleft), 0x0b0b0909u);
                    const int fe = max(E, F);
                    const uint32_t dx = __builtin_amdgcn_perm((uint32_t)subSat32(F, E),
                                                              (uint32_t)subSat32(D, fe), 0x0b0b0909u);
                    H = max(D, fe);
                    const uint32_t bit = rowBits(k);
                    asm volatile("v_bfi_b32 %0, %1, %2, %0" : "+v"(accEF) : "s"(bit), "v"(ef));
                    asm volatile("v_bfi_b32 %0, %1, %2, %0" : "+v"(accDX) : "s"(bit), "v"(dx));
                } else {
                    H = max(D, max(E, F));
                }
                if constexpr (LOCAL)
                    H = max(H, 0);
                else if constexpr (HEAD)
                    H = max(H, floorK[k]);
                if (k + 1 < TR)
                    D = Gleft + (int)__builtin_amdgcn_perm(0u, pr, sel[k + 1]);
                Ek[k] = E;
                Gk[k] = H + o;
                if constexpr (LOCAL)
                    bestG = max(bestG, Gk[k]);
                uG = Gk[k];
                uF = F;
            }
            botG = uG;
            botF = uF;
            if (g.active) {
                if constexpr (TRACE) {
                    rec[s * TG + l] = make_uint2(accEF, accDX);
                } else {
                    if (l == TG - 1)
                        *reinterpret_cast<int2 *>(ck + rowOut + 2 * jj) = make_int2(botG, botF);
                    // The right column (the last tile's too: a free last
                    // column is reduced from it).
                    if (jj == g.c1) {
#pragma unroll
                        for (int k = 0; k < TR; ++k)
                            *reinterpret_cast<int2 *>(ck + colOut + 2 * k) = make_int2(Gk[k], Ek[k]);
                    }
                    // Cell (m, n): the last row of the last block.
                    if (l == TG - 1 && jj == n && g.b == P.nblocks - 1) {
                        a.scores[P.pid] = botG - o;
                        a.endEF[2 * P.pid] = Ek[TR - 1];
                        a.endEF[2 * P.pid + 1] = botF;
                    }
                }
            }
        }
    };
    for (int s0 = 0; s0 < steps; s0 += TQD) {
        if ((s0 & (TK - 1)) == TK - TQD) {
            // Rings for the next TK steps (read from the next step on), and
            // the loads for the refill after it.
            refillStore(s0 + TQD);
            refillLoad(s0 + TQD + TK);
        }
#pragma unroll
        for (int u = 0; u < TQD; ++u) {
            const int s = s0 + u;
            const uint32_t prN = rg->pr[(s + 1 + TG - 1 - l) & (TRING - 1)];
            const int2 topN = rg->top[(s + 1) & (TRING - 1)];
            step(s, prof, topv);
            prof = prN;
            topv = topN;
        }
    }
    if constexpr (LOCAL) {
        // The group's maximum into the block's slot: one tile of a block per
        // launch, so the slot has no other writer (its first tile starts it).
#pragma unroll
        for (int d = 1; d < TG; d <<= 1)
            bestG = max(bestG, __shfl_xor(bestG, d, TG));
        if (g.active && l == 0) {
            int32_t *slot = a.best + P.bestck + g.b;
            *slot = g.t == 0 ? bestG : max(*slot, bestG);
        }
    }
}

__device__ __forceinline__ void loadProfile(const TileArgs &a, uint32_t *sProf)
{
    if (threadIdx.x < 5)
        sProf[threadIdx.x] = a.dnaProfile[threadIdx.x];
    __syncthreads();
}

// Forward: the tiles of anti-diagonal a.diag.  Each wave takes four
// consecutive row blocks of one pair (one per 16-lane group).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void kTileFwd32(TileArgs a)
{
    __shared__ uint32_t sProf[8];
    __shared__ GroupRings sRings[256 / TG];
    loadProfile(a, sProf);
    const int lane = threadIdx.x & 63, l = lane % TG, grp = lane / TG;
    const int w = blockIdx.x * (blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    const int pi = w / a.wavesPerPair;
    if (pi >= a.nPairs)
        return;
    const LongPair P = a.pairs[pi];
    const int b0 = (w % a.wavesPerPair) * 4;
    TileGeom g;
    g.b = b0 + grp;
    g.t = a.diag - g.b;
    g.active = g.b < P.nblocks && g.t >= 0 && g.t < P.ntiles;
    if (!__any(g.active))
        return;   // the whole wave is idle on this diagonal
    if (!g.active) {
        g.b = min(g.b, P.nblocks - 1);
        g.t = min(max(g.t, 0), P.ntiles - 1);
    }
    g.c0 = g.t * a.tileC + 1;
    g.c1 = min((g.t + 1) * a.tileC, P.n);
    if (a.mode & SQMI_LOCAL) {
        tileSweep<false, false, true>(a, P, g, l, sProf, &sRings[threadIdx.x / TG], nullptr);
    } else if (b0 == 0) {
        tileSweep<true, false>(a, P, g, l, sProf, &sRings[threadIdx.x / TG], nullptr);
    } else {
        tileSweep<false, false>(a, P, g, l, sProf, &sRings[threadIdx.x / TG], nullptr);
    }
}

// ---------------------------------------------------------------------------
// Packed forward: two pairs per lane (low / high int16 half), values held
// relative to a per-tile base.  Within one tile every value lies within
// (256 rows + 2 x tileC columns + skew) x D of the base, D = max|s| + |o| + |e|
// (one row or column moves H by at most D; the columns past a shorter
// pair's end lose at most |o| each), so the host takes this kernel when
// 800 x D stays inside the int16 window; -inf is kNeg16 as in the packed
// kernels.  The checkpoints stay int32 (converted at the ring refills and
// at the tile's right edge), so the traceback recomputes tiles exactly as
// after the int32 forward.

typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s16x2 as16(uint32_t x) { return __builtin_bit_cast(s16x2, x); }
__device__ __forceinline__ uint32_t asu(s16x2 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ s16x2 vmax(s16x2 a, s16x2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ s16x2 splat(int x)
{
    s16x2 r = {(short)x, (short)x};
    return r;
}
__device__ __forceinline__ uint32_t packu(int lo, int hi) { return (uint32_t)(uint16_t)lo | ((uint32_t)hi << 16); }
// Saturating packed add (v_pk_add_i16 with clamp): the rows above a pair's
// row 0 sink towards -32768 instead of wrapping, so no block needs a floor.
__device__ __forceinline__ s16x2 adds(s16x2 a, s16x2 b) { return __builtin_elementwise_add_sat(a, b); }

// int32 value -> int16 relative to base (-inf stays -inf), and back.
__device__ __forceinline__ int toRel(int v, int base)
{
    return v <= kNegL / 2 ? kNeg16 : min(max(v - base, kNeg16), 32767);
}
__device__ __forceinline__ int fromRel(int r, int base) { return r <= kNeg16 + kGuard16 ? kNegL : r + base; }

// Per-group LDS rings of the packed sweep: top[s]: the block above's bottom
// row (G, F, both halves packed, relative) at column c0 - 1 + s; pr[x]:
// the profile words (pair A, pair B) of column c0 - 1 + x - 15; out[s]: the
// group's bottom row (G, F packed) of step s, converted to int32 and
// written to the row checkpoints at the next refill.
struct PairRings {
    uint2 top[TRING];
    uint2 pr[TRING];
    uint2 out[TRING];
};

__device__ __forceinline__ void tileSweep16(const TileArgs &a, const LongPair (&P)[2], const bool (&act)[2], int b, int t,
                                            int l, const uint32_t *sProf, PairRings *rg)
{
    const int o = a.gapOpen, e = a.gapExtend, eo = max(o, e);
    const s16x2 o2 = splat(o), e2 = splat(e), neg2 = splat(kNeg16);
    const int C = a.tileC;
    const int c0 = t * C + 1;
    int32_t *ck = a.ck;
    int c1[2], base[2];
    int64_t rowIn[2], rowOut[2], colIn[2], colOut[2];
    const uint8_t *h[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        c1[x] = min((t + 1) * C, P[x].n);
        h[x] = a.hRes + a.hOff[P[x].pid];
        // Inactive halves point at block 0's own checkpoints (in bounds; values unused).
        const int bb = act[x] ? b : 0, tt = act[x] ? t : 0;
        rowIn[x] = P[x].rowck + (bb > 0 ? (int64_t)(bb - 1) * (P[x].n + 1) * 2 : 0);
        rowOut[x] = P[x].rowck + (int64_t)bb * (P[x].n + 1) * 2;
        colIn[x] = P[x].colck + (((int64_t)bb * P[x].ntiles + max(tt - 1, 0)) * kTileRows + l * TR) * 2;
        colOut[x] = P[x].colck + (((int64_t)bb * P[x].ntiles + tt) * kTileRows + l * TR) * 2;
        // Base: G at the block's last row (always a real row), column c0 - 1.
        base[x] = 0;
        if (act[x]) {
            const int iLast = b * kTileRows + kTileRows - 1 - P[x].p;
            base[x] = t == 0 ? ((iLast == 0) ? o : (o + (iLast - 1) * eo) + o)
                             : ck[P[x].colck + (((int64_t)b * P[x].ntiles + (t - 1)) * kTileRows + kTileRows - 1) * 2];
        }
    }
    const int c1max = max(act[0] ? c1[0] : 0, act[1] ? c1[1] : 0);

    uint32_t sel[TR];
    s16x2 Gk[TR], Ek[TR];
    const int rowbase = b * kTileRows + l * TR;
#pragma unroll
    for (int k = 0; k < TR; ++k) {
        int code[2], g[2], ev[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const int i = rowbase + k - P[x].p;
            const uint8_t *v = a.vRes + a.vOff[P[x].pid];
            code[x] = (i >= 1 && i <= P[x].m) ? min((int)v[i - 1], 4) : 4;
            if (!act[x]) {
                g[x] = kNeg16;
                ev[x] = kNeg16;
            } else if (t == 0) {
                // Column 0: -inf above row 0, H(0,0) = 0, then the gap chain.
                g[x] = (i < 0) ? kNeg16 : toRel((i == 0) ? o : (o + (i - 1) * eo) + o, base[x]);
                ev[x] = kNeg16;
            } else {
                g[x] = toRel(ck[colIn[x] + 2 * k], base[x]);
                ev[x] = toRel(ck[colIn[x] + 2 * k + 1], base[x]);
            }
        }
        const uint32_t sa = code[0] < 4 ? (uint32_t)(code[0] | 0x0c00) : 0x0d0cu;
        const uint32_t sb = code[1] < 4 ? (uint32_t)((code[1] + 4) | 0x0c00) : 0x0d0cu;
        sel[k] = sa | (sb << 16);
        Gk[k] = as16(packu(g[0], g[1]));
        Ek[k] = as16(packu(ev[0], ev[1]));
    }
    if (t == 0 && l == TG - 1) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
            if (act[x])   // column 0 of the bottom row
                *reinterpret_cast<int2 *>(ck + rowOut[x]) =
                    make_int2(fromRel(x ? Gk[TR - 1].y : Gk[TR - 1].x, base[x]), kNegL);
    }

    // Ring refills (the same scheme as tileSweep): lane l covers columns
    // c0 - 1 + s + 2 l and + 1 of both halves.
    int2 tv[2][2];
    uint32_t hv[2][2];
    auto refillLoad = [&](int s) {
        const int cb = c0 - 1 + s + 2 * l;
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int col = min(cb + y, max(c1[x], 0));
                tv[x][y] = *reinterpret_cast<const int2 *>(ck + rowIn[x] + 2 * col);
                hv[x][y] = h[x][min(max(cb + y, 1), P[x].n) - 1];
            }
    };
    auto refillStore = [&](int s) {
        const int cb = c0 - 1 + s + 2 * l;
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            int tg[2], tf[2];
            uint32_t pw[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const bool real = act[x] && b > 0 && cb + y <= c1[x];
                tg[x] = real ? toRel(tv[x][y].x, base[x]) : kNeg16;
                tf[x] = real ? toRel(tv[x][y].y, base[x]) : kNeg16;
                const int jj = cb + y;
                pw[x] = sProf[(jj >= 1 && jj <= P[x].n) ? min(hv[x][y], 4u) : 4u];
            }
            rg->top[(s + 2 * l + y) & (TRING - 1)] = make_uint2(packu(tg[0], tg[1]), packu(tf[0], tf[1]));
            rg->pr[(s + TG - 1 + 2 * l + y) & (TRING - 1)] = make_uint2(pw[0], pw[1]);
        }
        __builtin_amdgcn_wave_barrier();
    };
    // Bottom-row outputs of steps [f, s) to the row checkpoints.
    auto flushOut = [&](int f, int s) {
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int st = f + 2 * l + y;
            const int jj = c0 - 1 + st - (TG - 1);
            if (st < s && jj >= c0) {
                const uint2 w = rg->out[st & (TRING - 1)];
#pragma unroll
                for (int x = 0; x < 2; ++x)
                    if (act[x] && jj <= c1[x]) {
                        const int gr = x ? as16(w.x).y : as16(w.x).x, fr = x ? as16(w.y).y : as16(w.y).x;
                        *reinterpret_cast<int2 *>(ck + rowOut[x] + 2 * jj) =
                            make_int2(fromRel(gr, base[x]), fromRel(fr, base[x]));
                    }
            }
        }
    };

    int endCol[2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
        endCol[x] = (act[x] && b == P[x].nblocks - 1 && c1[x] == P[x].n) ? P[x].n : -1;
    uint32_t botG = asu(Gk[TR - 1]), botF = asu(neg2);
    s16x2 upGprev = neg2;
    const int width = max(c1max - c0 + 1, 0);
    const int steps = ((width + TG + TQD - 1) / TQD) * TQD;
    refillLoad(0);
    refillStore(0);
    refillLoad(TK);
    uint2 pp = rg->pr[(TG - 1 - l) & (TRING - 1)];
    uint2 topv = rg->top[0];
    int flushed = 0;

    auto step = [&](int s, uint2 pr, uint2 top) {
        const int jj = c0 - 1 + s - l;
        const s16x2 upG = as16((uint32_t)shiftIn16((int)botG, (int)top.x));
        const s16x2 upF = as16((uint32_t)shiftIn16((int)botF, (int)top.y));
        const s16x2 diagTop = upGprev;
        upGprev = upG;
        if (jj >= c0 && jj <= c1max) {
            s16x2 uG = upG, uF = upF;
            // D of row k from the old G of row k - 1 (no register copies).
            s16x2 D = adds(diagTop, as16(__builtin_amdgcn_perm(pr.y, pr.x, sel[0])));
#pragma unroll
            for (int k = 0; k < TR; ++k) {
                const s16x2 E = vmax(adds(Ek[k], e2), Gk[k]);
                const s16x2 F = vmax(adds(uF, e2), uG);
                const s16x2 H = vmax(vmax(D, E), F);
                if (k + 1 < TR)
                    D = adds(Gk[k], as16(__builtin_amdgcn_perm(pr.y, pr.x, sel[k + 1])));
                Ek[k] = E;
                Gk[k] = adds(H, o2);
                uG = Gk[k];
                uF = F;
            }
            botG = asu(uG);
            botF = asu(uF);
            if (l == TG - 1)
                rg->out[s & (TRING - 1)] = make_uint2(botG, botF);
            // Cell (m, n): the last row of the last block.
            if (l == TG - 1 && (jj == endCol[0] || jj == endCol[1])) {
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    if (jj == endCol[y]) {
                        a.scores[P[y].pid] = fromRel(y ? uG.y : uG.x, base[y]) - o;
                        a.endEF[2 * P[y].pid] = fromRel(y ? Ek[TR - 1].y : Ek[TR - 1].x, base[y]);
                        a.endEF[2 * P[y].pid + 1] = fromRel(y ? uF.y : uF.x, base[y]);
                    }
            }
        }
    };
    for (int s0 = 0; s0 < steps; s0 += TQD) {
        if ((s0 & (TK - 1)) == TK - TQD) {
            refillStore(s0 + TQD);
            refillLoad(s0 + TQD + TK);
            flushOut(flushed, s0);
            flushed = s0;
        }
#pragma unroll
        for (int u = 0; u < TQD; ++u) {
            const int s = s0 + u;
            const uint2 prN = rg->pr[(s + 1 + TG - 1 - l) & (TRING - 1)];
            const uint2 topN = rg->top[(s + 1) & (TRING - 1)];
            step(s, pp, topv);
            pp = prN;
            topv = topN;
        }
    }
    flushOut(flushed, steps);
    // Right column of a full-width tile (the column the sweep ended on).
#pragma unroll
    for (int x = 0; x < 2; ++x)
        if (act[x] && t + 1 < P[x].ntiles && c1[x] == c1max) {
#pragma unroll
            for (int k = 0; k < TR; ++k)
                *reinterpret_cast<int2 *>(ck + colOut[x] + 2 * k) =
                    make_int2(fromRel(x ? Gk[k].y : Gk[k].x, base[x]), fromRel(x ? Ek[k].y : Ek[k].x, base[x]));
        }
}

// Packed forward of the tiles of anti-diagonal a.diag: wave w takes four
// consecutive row blocks (one per 16-lane group) of the pair couple
// (2 c, 2 c + 1) of the chunk.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void kTileFwd16(TileArgs a)
{
    __shared__ uint32_t sProf[8];
    __shared__ PairRings sRings[256 / TG];
    loadProfile(a, sProf);
    const int lane = threadIdx.x & 63, l = lane % TG, grp = lane / TG;
    const int w = blockIdx.x * (blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    const int cp = w / a.wavesPerPair;
    if (2 * cp >= a.nPairs)
        return;
    const bool hasB = 2 * cp + 1 < a.nPairs;
    const LongPair P[2] = {a.pairs[2 * cp], a.pairs[hasB ? 2 * cp + 1 : 2 * cp]};
    const int b0 = (w % a.wavesPerPair) * 4;
    const int b = b0 + grp, t = a.diag - b;
    bool act[2];
    act[0] = b < P[0].nblocks && t >= 0 && t < P[0].ntiles;
    act[1] = hasB && b < P[1].nblocks && t >= 0 && t < P[1].ntiles;
    if (!__any(act[0] || act[1]))
        return;   // the whole wave is idle on this diagonal
    // Idle groups sweep a clamped tile of pair A with both halves inactive.
    const int bq = (act[0] || act[1]) ? b : min(b, P[0].nblocks - 1);
    const int tq = (act[0] || act[1]) ? t : min(max(t, 0), P[0].ntiles - 1);
    tileSweep16(a, P, act, bq, tq, l, sProf, &sRings[threadIdx.x / TG]);
}

// Streaming packed forward: one persistent launch for every tile of the
// chunk.  Waves claim items (four consecutive row blocks of a couple on one
// tile anti-diagonal) through one counter, in anti-diagonal order, and a
// tile (b, t) waits until tiles (b - 1, t) and (b, t - 1) are done: those
// lie on the previous anti-diagonal, so their items were claimed earlier by
// waves that are running, and the unfinished claimed item with the lowest
// index always has all its inputs.  No wave waits on an item that no
// running wave holds, whatever the residency.  A tile's checkpoints are
// published by an agent-scope release fence and a relaxed agent-scope store
// of the row block's progress (tiles done); a consumer reads the progress
// with relaxed agent-scope loads and then fences with acquire, so the
// checkpoints written on another XCD are visible.  A wait is bounded: past
// kSpinLimit polls (about a second) the wave sets a.error (the host reports
// a device error) and goes on, and once a.error is set no wave waits any
// more, so the launch drains at once.
constexpr int kSpinLimit = 1 << 20;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void kTileFwd16S(TileArgs a)
{
    __shared__ uint32_t sProf[8];
    __shared__ PairRings sRings[256 / TG];
    loadProfile(a, sProf);
    const int lane = threadIdx.x & 63, l = lane % TG, grp = lane / TG;
    for (;;) {
        // Branch-free claim: every lane takes part (lane 0 adds 1).
        int it = (int)atomicAdd(a.queue, lane == 0 ? 1u : 0u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= a.nItems)
            break;
        // The range holding item `it` (last range starting at or before it).
        int lo = 0, hi = a.nRanges - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.ranges[mid].x <= it)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int4 rg = a.ranges[lo];
        const int cp = rg.y, d = rg.z, wb = rg.w + (it - rg.x);
        const bool hasB = 2 * cp + 1 < a.nPairs;
        const LongPair P[2] = {a.pairs[2 * cp], a.pairs[hasB ? 2 * cp + 1 : 2 * cp]};
        const int b = wb * 4 + grp, t = d - b;
        bool act[2];
        act[0] = b < P[0].nblocks && t >= 0 && t < P[0].ntiles;
        act[1] = hasB && b < P[1].nblocks && t >= 0 && t < P[1].ntiles;
        const bool any = act[0] || act[1];
        uint32_t *prog = a.progress + (size_t)cp * a.progStride;
        for (int spin = 0;; ++spin) {
            bool ready = true;
            if (any && b > 0)
                ready = __hip_atomic_load(prog + b - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (uint32_t)(t + 1);
            if (any && t > 0)
                ready = ready && __hip_atomic_load(prog + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (uint32_t)t;
            if (__all(ready))
                break;
            if (__hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
                break;
            if (spin >= kSpinLimit) {
                if (lane == 0)
                    atomicOr(a.error, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(4);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        // Idle groups sweep a clamped tile of pair A with both halves inactive.
        const int bq = any ? b : min(b, P[0].nblocks - 1);
        const int tq = any ? t : min(max(t, 0), P[0].ntiles - 1);
        tileSweep16(a, P, act, bq, tq, l, sProf, &sRings[threadIdx.x / TG]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (any && l == 0)
            __hip_atomic_store(prog + b, (uint32_t)(t + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Walk state of one long pair: the cell (i, j) reached, the walk mode
// (0: choose from the cell's H decision, 1 / 2: inside a vertical /
// horizontal gap), the last operation, gap opens / extensions, done flag.
enum { WS_I = 0, WS_J, WS_MODE, WS_PREV, WS_OPENS, WS_EXTS, WS_DONE, WS_NOPS };

__device__ __forceinline__ TileGeom walkTile(const TileArgs &a, const LongPair &P, const int32_t *ws)
{
    TileGeom g;
    const int i = ws[WS_I], j = ws[WS_J];
    g.active = ws[WS_DONE] == 0 && i > 0 && j > 0;
    const int ii = g.active ? i : P.m, jj = g.active ? j : P.n;
    g.b = (P.p + ii) / kTileRows;
    g.t = (jj - 1) / a.tileC;
    g.c0 = g.t * a.tileC + 1;
    g.c1 = min((g.t + 1) * a.tileC, P.n);
    return g;
}

// Traceback rounds.  A round recomputes, with trace records, the window of
// kRoundBlocks x kRoundTiles tiles whose bottom-right tile holds the walk's
// position (every tile of it only needs checkpoints, so all of them run at
// once), then walks the pair through the window.  The walk leaves a window
// through its top or left side, so each round moves it up kRoundBlocks row
// blocks or left kRoundTiles tiles.

// Window slot of tile (b, t) for a window whose bottom-right tile is (b0, t0).
__device__ __forceinline__ int windowSlot(int b0, int t0, int b, int t) { return (b0 - b) * kRoundTiles + (t0 - t); }

// Step 1: one 16-lane group per (pair, window slot).  Workgroup w serves
// pairs with the same w mod 8 (the walk of pair p runs in workgroup p, and
// workgroups are spread over the 8 XCDs by id mod 8, so both kernels of a
// pair share one L2).
constexpr int kTraceGroupsPerWG = 16;
constexpr int kTraceWGsPerPair = kRoundSlots / kTraceGroupsPerWG;
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void kTileTrace32(TileArgs a)
{
    __shared__ uint32_t sProf[8];
    __shared__ GroupRings sRings[256 / TG];
    loadProfile(a, sProf);
    const int lane = threadIdx.x & 63, l = lane % TG;
    const int w = blockIdx.x;
    const int pi = 8 * (w / (8 * kTraceWGsPerPair)) + (w % 8);
    const int slot = ((w / 8) % kTraceWGsPerPair) * kTraceGroupsPerWG + (int)(threadIdx.x / TG);
    const bool exists = pi < a.nPairs;
    const int pc = exists ? pi : a.nPairs - 1;
    const LongPair P = a.pairs[pc];
    TileGeom g = walkTile(a, P, a.walk + (int64_t)pc * kWalkStateInts);
    g.b -= slot / kRoundTiles;
    g.t -= slot % kRoundTiles;
    g.active = g.active && exists && g.b >= 0 && g.t >= 0;
    if (!__any(g.active))
        return;
    if (!g.active) {
        g.b = max(g.b, 0);
        g.t = max(g.t, 0);
    }
    g.c0 = g.t * a.tileC + 1;
    g.c1 = min((g.t + 1) * a.tileC, P.n);
    uint2 *rec = a.trace + ((int64_t)pc * kRoundSlots + slot) * kTileRecSteps * TG;
    tileSweep<true, true>(a, P, g, l, sProf, &sRings[threadIdx.x / TG], rec);
}

__global__ void kTileWalkInit(TileArgs a)
{
    const int pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= a.nPairs)
        return;
    const LongPair P = a.pairs[pi];
    int32_t *ws = a.walk + (int64_t)pi * kWalkStateInts;
    const int score = a.scores[P.pid], endE = a.endEF[2 * P.pid], endF = a.endEF[2 * P.pid + 1];
    ws[WS_I] = P.m;
    ws[WS_J] = P.n;
    // End cell: F, then E, win ties (dp_algorithm_impl.h:1190-1207,
    // PreferGapsAtEnd_ dp_traceback_impl.h:115-122); the linear-gap walk
    // (gap open == gap extend) takes the cell's own direction.
    ws[WS_MODE] = (a.gapOpen == a.gapExtend) ? 0 : (endF == score) ? 1 : (endE == score) ? 2 : 0;
    ws[WS_PREV] = 0;
    ws[WS_OPENS] = 0;
    ws[WS_EXTS] = 0;
    ws[WS_DONE] = 0;
    ws[WS_NOPS] = 0;
    const int S = a.alpha * a.alpha + 2;
    for (int k = 0; k < S; ++k)
        a.counts[(int64_t)P.pid * S + k] = 0;
}

// Step 2: one workgroup per pair walks it through the window.  The records
// of half a tile (128 rows: lanes 0-7 or 8-15 of the sweep) are staged in
// LDS by the whole workgroup; thread 0 walks them, and the next half tile is
// staged when the walk leaves this one.
constexpr int kWalkWGThreads = 256;
constexpr int kHalfLanes = TG / 2;
__global__ __launch_bounds__(kWalkWGThreads) void kTileWalk32(TileArgs a)
{
    __shared__ uint2 sRec[kTileRecSteps * kHalfLanes];
    __shared__ uint8_t sH[kLongTileC], sV[kTileRows / 2];   // residues of the half tile's columns and rows
    __shared__ int32_t sCnt[16];
    __shared__ int32_t sUnit[4];   // b, t, half, stop
    const int pi = blockIdx.x;
    if (pi >= a.nPairs)
        return;
    const LongPair P = a.pairs[pi];
    int32_t *ws = a.walk + (int64_t)pi * kWalkStateInts;
    const TileGeom g0 = walkTile(a, P, ws);
    if (!g0.active)
        return;   // finished (the same for every thread of the workgroup)
    const int A = a.alpha, S = A * A + 2;
    const bool linear = a.gapOpen == a.gapExtend;
    const uint8_t *h = a.hRes + a.hOff[P.pid];
    const uint8_t *v = a.vRes + a.vOff[P.pid];
    const int tid = threadIdx.x;
    if (tid < 16)
        sCnt[tid] = 0;
    int i = 0, j = 0, mode = 0, prev = 0, opens = 0, exts = 0, nOps = 0;
    if (tid == 0) {
        i = ws[WS_I];
        j = ws[WS_J];
        mode = ws[WS_MODE];
        prev = ws[WS_PREV];
        opens = ws[WS_OPENS];
        exts = ws[WS_EXTS];
        nOps = ws[WS_NOPS];
    }
    // Rows: written backwards from the end of the pair's op segment.
    uint8_t *opEnd = a.ops ? a.ops + a.opOff[P.pid + 1] : nullptr;
    auto emit = [&](uint8_t op) {
        if (opEnd)
            opEnd[-1 - (int64_t)nOps] = op;
        ++nOps;
    };
    for (;;) {
        if (tid == 0) {
            // The unit (tile, half) holding (i, j), or stop: finished or
            // outside the window.
            const int b = (P.p + i) / kTileRows, t = (j - 1) / a.tileC;
            const int half = ((P.p + i) % kTileRows) / (kTileRows / 2);
            const bool inside = i > 0 && j > 0 && b <= g0.b && b > g0.b - kRoundBlocks && t <= g0.t &&
                                t > g0.t - kRoundTiles;
            sUnit[0] = b;
            sUnit[1] = t;
            sUnit[2] = half;
            sUnit[3] = inside ? 0 : 1;
        }
        __syncthreads();
        const int ub = sUnit[0], ut = sUnit[1], uh = sUnit[2];
        if (sUnit[3])
            break;
        // Stage the half tile: records rec[s * 16 + 8 * half + 0..7] for all steps.
        const uint2 *rec = a.trace + ((int64_t)pi * kRoundSlots + windowSlot(g0.b, g0.t, ub, ut)) * kTileRecSteps * TG;
        const uint4 *src = reinterpret_cast<const uint4 *>(rec);
        uint4 *dst = reinterpret_cast<uint4 *>(sRec);
        for (int x = tid; x < kTileRecSteps * kHalfLanes / 2; x += kWalkWGThreads) {
            const int st = x / (kHalfLanes / 2), c = x % (kHalfLanes / 2);
            dst[x] = src[st * (TG / 2) + uh * (kHalfLanes / 2) + c];
        }
        const int c0 = ut * a.tileC + 1;
        const int rowTop = ub * kTileRows + uh * (kTileRows / 2) - P.p;   // first pair row of the half
        for (int x = tid; x < a.tileC; x += kWalkWGThreads)
            sH[x] = h[min(c0 + x, P.n) - 1];
        if (tid < kTileRows / 2)
            sV[tid] = v[min(max(rowTop + tid, 1), P.m) - 1];
        __syncthreads();
        if (tid == 0) {
            while (i > 0 && j > 0 && i >= rowTop && j >= c0) {
                const int r = P.p + i - ub * kTileRows;
                const int l = r / TR, k = r % TR;
                const uint2 w = sRec[(j - (c0 - 1) + l) * kHalfLanes + (l - uh * kHalfLanes)];
                const uint32_t sh = (uint32_t)k;
                const uint32_t bits = ((w.x >> sh) & 1u) | (((w.x >> (sh + 16)) & 1u) << 1) |
                                      (((w.y >> sh) & 1u) << 2) | (((w.y >> (sh + 16)) & 1u) << 3);
                const int d = mode == 0 ? ((bits & 4u) ? ((bits & 8u) ? 2 : 1) : 0) : mode;
                if (d == 0) {
                    const int hc = min((int)sH[j - c0], A - 1), vc = min((int)sV[i - rowTop], A - 1);
                    sCnt[hc * A + vc] += 1;
                    emit('M');
                    prev = 0;
                    --i;
                    --j;
                    mode = 0;
                } else {
                    // One gap character; the gap goes on while the cell's gap
                    // extends, down to row / column 1 (dp_traceback_impl.h:220-341).
                    // Linear gaps (gap open == gap extend) take one step per
                    // cell and re-read the next cell's own direction.
                    const bool vert = d == 1;
                    const bool open = (bits & (vert ? 2u : 1u)) != 0;
                    if (prev == d) ++exts; else ++opens;
                    prev = d;
                    emit(vert ? 'V' : 'H');
                    const bool stop = linear || open || (vert ? i == 1 : j == 1);
                    if (vert)
                        --i;
                    else
                        --j;
                    mode = stop ? 0 : d;
                }
            }
        }
        __syncthreads();   // sRec is restaged next
    }
    if (tid == 0) {
        int32_t *cnt = a.counts + (int64_t)P.pid * S;
        for (int k = 0; k < A * A; ++k)
            if (sCnt[k])
                cnt[k] += sCnt[k];
        if (i == 0 || j == 0) {
            // Leading gaps down to the origin (dp_traceback_impl.h:522-545).
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
            if (a.ops)
                a.opLen[P.pid] = nOps;
            ws[WS_DONE] = 1;
        }
        ws[WS_I] = i;
        ws[WS_J] = j;
        ws[WS_MODE] = mode;
        ws[WS_PREV] = prev;
        ws[WS_OPENS] = opens;
        ws[WS_EXTS] = exts;
        ws[WS_NOPS] = nOps;
    }
}

}  // namespace

// Results of long pairs under free end gaps or local alignment, one
// workgroup per pair: the maximum G over the cells the AlignConfig tracks
// (dp_scout.h:169-210; the reference's own scouts, restated in the oracle):
// the last row (free bottom: the last block's bottom-row checkpoint, columns
// 0 .. n), the last column (free right: the last tile's right-column
// checkpoints, rows 0 .. m), or every cell (local: the maxima of the row
// blocks the tile sweeps kept).  Score = G - gapOpen.
__global__ __launch_bounds__(256) void kLongResult(TileArgs a)
{
    __shared__ int red[256 / 64];
    const int pi = blockIdx.x;
    if (pi >= a.nPairs)
        return;
    const LongPair P = a.pairs[pi];
    const int32_t *ck = a.ck;
    const bool local = (a.mode & SQMI_LOCAL) != 0;
    int best = kNegL;
    if (local) {
        for (int b = threadIdx.x; b < P.nblocks; b += blockDim.x)
            best = max(best, a.best[P.bestck + b]);
    }
    if (!local && (a.mode & SQMI_FREE_BOTTOM)) {
        const int64_t row = P.rowck + (int64_t)(P.nblocks - 1) * (P.n + 1) * 2;
        for (int j = threadIdx.x; j <= P.n; j += blockDim.x)
            best = max(best, ck[row + 2 * j]);
    }
    if (!local && (a.mode & SQMI_FREE_RIGHT)) {
        for (int r = threadIdx.x; r < P.nblocks * kTileRows; r += blockDim.x) {
            const int i = r - P.p;
            if (i >= 0 && i <= P.m)
                best = max(best, ck[P.colck + (((int64_t)(r / kTileRows) * P.ntiles + (P.ntiles - 1)) * kTileRows +
                                               r % kTileRows) * 2]);
        }
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
        best = max(best, __shfl_xor(best, d, 64));
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            best = max(best, red[w]);
        if (local || (a.mode & (SQMI_FREE_BOTTOM | SQMI_FREE_RIGHT)))
            a.scores[P.pid] = best - a.gapOpen;
    }
}

hipError_t launchLongResult(const TileArgs &args, hipStream_t stream)
{
    if (args.nPairs <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(kLongResult, dim3(args.nPairs), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launchTileFwd32(const TileArgs &args, hipStream_t stream)
{
    const int waves = args.nPairs * args.wavesPerPair;
    if (waves <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(kTileFwd32, dim3((waves + 3) / 4), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launchTileFwd16(const TileArgs &args, hipStream_t stream)
{
    const int waves = (args.nPairs + 1) / 2 * args.wavesPerPair;
    if (waves <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(kTileFwd16, dim3((waves + 3) / 4), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launchTileFwd16Stream(const TileArgs &args, int gridBlocks, hipStream_t stream)
{
    if (args.nItems <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(kTileFwd16S, dim3(gridBlocks), dim3(256), 0, stream, args);
    return hipGetLastError();
}

int occupancyTileFwd16Stream()
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void *)&kTileFwd16S, 256, 0) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

hipError_t launchTileTrace32(const TileArgs &args, hipStream_t stream)
{
    if (args.nPairs <= 0)
        return hipSuccess;
    const int wgs = ((args.nPairs + 7) / 8) * 8 * kTraceWGsPerPair;
    hipLaunchKernelGGL(kTileTrace32, dim3(wgs), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launchTileWalkInit(const TileArgs &args, hipStream_t stream)
{
    if (args.nPairs <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(kTileWalkInit, dim3((args.nPairs + 63) / 64), dim3(64), 0, stream, args);
    return hipGetLastError();
}

hipError_t launchTileWalk32(const TileArgs &args, hipStream_t stream)
{
    if (args.nPairs <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(kTileWalk32, dim3(args.nPairs), dim3(kWalkWGThreads), 0, stream, args);
    return hipGetLastError();
}

}  // namespace sqmi
