/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * A plain scalar restatement of the reference's affine DP, written to follow
 * these reference files (paths relative to the SeqAn 2 source tree):
 *
 *   recurrence, trace bits      include/seqan/align/dp_formula_affine.h:62-126
 *   borders (Zero/Horizontal/   include/seqan/align/dp_formula_affine.h:244-330,
 *            Vertical)          include/seqan/align/dp_formula.h:289-309
 *   tie rule of _maxScore       include/seqan/align/dp_formula.h:165-201
 *   which cells are tracked     include/seqan/align/dp_meta_info.h:178-221
 *   column-major order, -inf    include/seqan/align/dp_algorithm_impl.h:334-579,
 *                               include/seqan/align/dp_cell.h:137-146,
 *                               include/seqan/align/dp_cell_affine.h:58-64
 *   strict '>' best-cell scout  include/seqan/align/dp_scout.h:169-210
 *   empty input -> INT_MIN      include/seqan/align/dp_algorithm_impl.h:1399-1400
 *   end-cell correction         include/seqan/align/dp_algorithm_impl.h:1190-1207
 *   traceback walk              include/seqan/align/dp_traceback_impl.h:115-546
 *   gap open/extend counting    include/seqan/align/evaluate_alignment.h:215-300
 */
#include "oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
static int omp_get_max_threads_wrapper(void) { return omp_get_max_threads(); }
#endif

/* DPCellDefaultInfinity: numeric_limits<int>::lowest() / 2 (dp_cell.h:141-146). */
#define NEG_INF (INT32_MIN / 2)

/* TraceBitMap_ values (dp_profile.h:121-135). */
enum {
    TB_NONE = 0, TB_DIAG = 1, TB_HORI = 2, TB_VERT = 4, TB_HORI_OPEN = 8,
    TB_VERT_OPEN = 16, TB_MAX_HORI = 32, TB_MAX_VERT = 64
};

static inline int32_t max2(int32_t a, int32_t b) { return a < b ? b : a; }

static inline int32_t subst(const orc_scoring *sc, uint8_t h, uint8_t v)
{
    return sc->table[(int)h * sc->alpha + (int)v];
}

/* Per-cell tracking rule of DPMetaColumn_<FullColumn> (dp_meta_info.h:178-221).
 * col: 0 initial, 1 inner, 2 final; cell: 0 first, 1 inner, 2 last. */
static int tracked(uint32_t mode, int col, int cell)
{
    if (mode & ORC_LOCAL) return 1;
    if (cell == 2) return col == 2 || (mode & ORC_FREE_BOTTOM);
    return col == 2 && (mode & ORC_FREE_RIGHT);
}

/* ------------------------------------------------------------------------ */
/* Score only: one sparse column of H and E, F carried down the column.      */
/* ------------------------------------------------------------------------ */
int32_t orc_score(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                  const orc_scoring *sc, uint32_t mode)
{
    if (n <= 0 || m <= 0) return INT32_MIN;
    const int local = (mode & ORC_LOCAL) != 0;
    const int32_t o = sc->gap_open, e = sc->gap_extend;
    int32_t *H = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    int32_t *E = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    int32_t best = NEG_INF;

    /* Initial column: (0,0) Zero; rows >= 1 Vertical or Zero. */
    {
        int32_t F = NEG_INF;
        H[0] = 0;
        E[0] = NEG_INF;
        if (tracked(mode, 0, 0) && H[0] > best) best = H[0];
        for (int64_t i = 1; i <= m; ++i) {
            if (local || (mode & ORC_FREE_LEFT)) {
                H[i] = 0;            /* Zero: E, F untouched (still -inf) */
            } else {
                F = max2(F + e, H[i - 1] + o);
                H[i] = F;
            }
            E[i] = NEG_INF;
            if (tracked(mode, 0, i == m ? 2 : 1) && H[i] > best) best = H[i];
        }
    }
    for (int64_t j = 1; j <= n; ++j) {
        const int col = (j == n) ? 2 : 1;
        const uint8_t hj = h[j - 1];
        int32_t diag = H[0];
        int32_t F = NEG_INF;
        /* First cell: Horizontal, or Zero when the first row is free. */
        if (local || (mode & ORC_FREE_TOP)) {
            H[0] = 0;
        } else {
            E[0] = max2(E[0] + e, H[0] + o);
            H[0] = E[0];
        }
        if (tracked(mode, col, 0) && H[0] > best) best = H[0];
        int32_t up = H[0];
        for (int64_t i = 1; i <= m; ++i) {
            const int32_t d = diag + subst(sc, hj, v[i - 1]);
            diag = H[i];
            E[i] = max2(E[i] + e, H[i] + o);
            F = max2(F + e, up + o);
            int32_t x = max2(d, max2(F, E[i]));
            if (local) x = max2(x, 0);
            H[i] = x;
            up = x;
            if (tracked(mode, col, i == m ? 2 : 1) && x > best) best = x;
        }
    }
    free(H);
    free(E);
    return best;
}

/* ------------------------------------------------------------------------ */
/* Banded score (global with free-end bits): only cells with                  */
/* lower <= j - i <= upper exist; the others are -inf.  Band settings the     */
/* reference rejects give INT32_MIN: a band missing the matrix, or missing    */
/* (0, 0) / (m, n) where those ends are not free                             */
/* (_checkBandProperties, dp_algorithm_impl.h:192-232).                       */
/* ------------------------------------------------------------------------ */
int32_t orc_score_banded(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                         const orc_scoring *sc, uint32_t mode, int64_t lower, int64_t upper)
{
    const int local = (mode & ORC_LOCAL) != 0;
    if (local) mode = ORC_LOCAL | ORC_FREE_TOP | ORC_FREE_LEFT | ORC_FREE_RIGHT | ORC_FREE_BOTTOM;
    if (n <= 0 || m <= 0) return INT32_MIN;
    if (upper < -m || lower > n) return INT32_MIN;
    /* Local alignment has free ends on every side (IsFreeEndGap_,
     * dp_profile.h:604-612), so only the intersection check applies. */
    if (upper < 0 && !(mode & ORC_FREE_LEFT)) return INT32_MIN;
    if (lower > 0 && !(mode & ORC_FREE_TOP)) return INT32_MIN;
    if (upper + m < n && !(mode & ORC_FREE_BOTTOM)) return INT32_MIN;
    if (lower + m > n && !(mode & ORC_FREE_RIGHT)) return INT32_MIN;
    const int32_t o = sc->gap_open, e = sc->gap_extend;
    int32_t *H = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    int32_t *E = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    int32_t best = NEG_INF;
#define INBAND(I, J) ((J) - (I) >= lower && (J) - (I) <= upper)
    {
        int32_t F = NEG_INF;
        H[0] = 0;
        E[0] = NEG_INF;
        if (!INBAND(0, 0)) H[0] = NEG_INF;
        if (tracked(mode, 0, 0) && H[0] > best) best = H[0];
        for (int64_t i = 1; i <= m; ++i) {
            if (!INBAND(i, 0)) {
                H[i] = NEG_INF;
                F = NEG_INF;
            } else if (mode & ORC_FREE_LEFT) {
                H[i] = 0;
            } else {
                F = max2(F + e, H[i - 1] + o);
                H[i] = F;
            }
            E[i] = NEG_INF;
            if (tracked(mode, 0, i == m ? 2 : 1) && H[i] > best) best = H[i];
        }
    }
    for (int64_t j = 1; j <= n; ++j) {
        const int col = (j == n) ? 2 : 1;
        const uint8_t hj = h[j - 1];
        int32_t diag = H[0];
        int32_t F = NEG_INF;
        if (!INBAND(0, j)) {
            H[0] = NEG_INF;
            E[0] = NEG_INF;
        } else if (mode & ORC_FREE_TOP) {
            H[0] = 0;
        } else {
            E[0] = max2(E[0] + e, H[0] + o);
            H[0] = E[0];
        }
        if (tracked(mode, col, 0) && H[0] > best) best = H[0];
        int32_t up = H[0];
        for (int64_t i = 1; i <= m; ++i) {
            if (!INBAND(i, j)) {
                diag = H[i];
                H[i] = NEG_INF;
                E[i] = NEG_INF;
                F = NEG_INF;
                up = NEG_INF;
                continue;
            }
            const int32_t d = diag + subst(sc, hj, v[i - 1]);
            diag = H[i];
            E[i] = max2(E[i] + e, H[i] + o);
            F = max2(F + e, up + o);
            int32_t x = max2(d, max2(F, E[i]));
            if (local) x = max2(x, 0);
            H[i] = x;
            up = x;
            if (tracked(mode, col, i == m ? 2 : 1) && x > best) best = x;
        }
    }
#undef INBAND
    free(H);
    free(E);
    return best;
}

void orc_score_batch(const uint8_t *h_res, const int64_t *h_off,
                     const uint8_t *v_res, const int64_t *v_off, int64_t count,
                     const orc_scoring *sc, uint32_t mode, int32_t *scores,
                     int threads)
{
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads_wrapper();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
    for (int64_t k = 0; k < count; ++k)
        scores[k] = orc_score(h_res + h_off[k], h_off[k + 1] - h_off[k],
                              v_res + v_off[k], v_off[k + 1] - v_off[k], sc, mode);
    (void)threads;
}

/* ------------------------------------------------------------------------ */
/* Traceback: full trace matrix (column-major, (n+1) x (m+1) bytes).          */
/* ------------------------------------------------------------------------ */
/* Banded tracebacks (globalAlignment / localAlignment with lowerDiag,   */
/* upperDiag, global_alignment_banded.h:152-181, local_alignment_banded.h: */
/* 242-280) run the same recurrence with every cell outside               */
/* lower <= j - i <= upper at -inf: the reference's band-edge cells leave   */
/* out the direction that comes from outside the band                      */
/* (RecursionDirectionUpperDiagonal / LowerDiagonal,                       */
/* dp_formula_affine.h:129-230), which is what a -inf predecessor gives.   */
static int32_t traceback_impl(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                              const orc_scoring *sc, uint32_t mode, int32_t *counts,
                              char *ops, int64_t *n_ops, int64_t *begins,
                              int banded, int64_t lower, int64_t upper)
{
    const int A = sc->alpha;
    const int local = (mode & ORC_LOCAL) != 0;
    if (counts) memset(counts, 0, sizeof(int32_t) * (size_t)(A * A + 2));
    if (n_ops) *n_ops = 0;
    if (begins) begins[0] = begins[1] = 0;
    if (n <= 0 || m <= 0) return INT32_MIN;
    if (banded) {
        /* The band checks of orc_score_banded. */
        const uint32_t fm = local ? (ORC_FREE_TOP | ORC_FREE_LEFT | ORC_FREE_RIGHT | ORC_FREE_BOTTOM) : mode;
        if (upper < -m || lower > n) return INT32_MIN;
        if (upper < 0 && !(fm & ORC_FREE_LEFT)) return INT32_MIN;
        if (lower > 0 && !(fm & ORC_FREE_TOP)) return INT32_MIN;
        if (upper + m < n && !(fm & ORC_FREE_BOTTOM)) return INT32_MIN;
        if (lower + m > n && !(fm & ORC_FREE_RIGHT)) return INT32_MIN;
    }
#define OUT(I, J) (banded && ((J) - (I) < lower || (J) - (I) > upper))
    const int32_t o = sc->gap_open, e = sc->gap_extend;
    const int64_t R = m + 1;
    uint8_t *T = (uint8_t *)malloc((size_t)(n + 1) * (size_t)R);
    int32_t *H = (int32_t *)malloc(sizeof(int32_t) * (size_t)R);
    int32_t *E = (int32_t *)malloc(sizeof(int32_t) * (size_t)R);
    if (!T || !H || !E) { free(T); free(H); free(E); return INT32_MIN; }
    int32_t best = NEG_INF, bestE = NEG_INF, bestF = NEG_INF;
    int64_t bi = 0, bj = 0;

#define TRACK(col, cell, I, J, HV, EV, FV)                                    \
    do {                                                                       \
        if (tracked(mode, col, cell) && (HV) > best) {                         \
            best = (HV); bestE = (EV); bestF = (FV); bi = (I); bj = (J);       \
        }                                                                      \
    } while (0)

    /* Initial column. */
    {
        int32_t F = NEG_INF;
        H[0] = OUT(0, 0) ? NEG_INF : 0; E[0] = NEG_INF; T[0] = TB_NONE;
        TRACK(0, 0, 0, 0, H[0], E[0], F);
        for (int64_t i = 1; i <= m; ++i) {
            if (OUT(i, 0)) {
                H[i] = NEG_INF;
                E[i] = NEG_INF;
                F = NEG_INF;
                T[i] = TB_NONE;
                continue;
            }
            if (local || (mode & ORC_FREE_LEFT)) {
                H[i] = 0;
                T[i] = TB_NONE;
            } else {
                /* Vertical recursion: open on ties lost to extend (left arg). */
                const int32_t ext = F + e, opn = H[i - 1] + o;
                T[i] = (uint8_t)((ext < opn ? TB_VERT_OPEN : TB_VERT) | TB_MAX_VERT);
                F = max2(ext, opn);
                H[i] = F;
            }
            E[i] = NEG_INF;
            TRACK(0, i == m ? 2 : 1, i, 0, H[i], E[i], (local || (mode & ORC_FREE_LEFT)) ? NEG_INF : F);
        }
    }
    for (int64_t j = 1; j <= n; ++j) {
        const int col = (j == n) ? 2 : 1;
        uint8_t *Tc = T + j * R;
        const uint8_t hj = h[j - 1];
        int32_t diag = H[0];
        int32_t F = NEG_INF;
        if (OUT(0, j)) {
            H[0] = NEG_INF;
            E[0] = NEG_INF;
            Tc[0] = TB_NONE;
        } else if (local || (mode & ORC_FREE_TOP)) {
            H[0] = 0;
            Tc[0] = TB_NONE;
        } else {
            const int32_t ext = E[0] + e, opn = H[0] + o;
            Tc[0] = (uint8_t)((ext < opn ? TB_HORI_OPEN : TB_HORI) | TB_MAX_HORI);
            E[0] = max2(ext, opn);
            H[0] = E[0];
        }
        TRACK(col, 0, 0, j, H[0], E[0], F);
        int32_t up = H[0];
        for (int64_t i = 1; i <= m; ++i) {
            if (OUT(i, j)) {
                diag = H[i];
                H[i] = NEG_INF;
                E[i] = NEG_INF;
                F = NEG_INF;
                up = NEG_INF;
                Tc[i] = TB_NONE;
                continue;
            }
            const int32_t d = diag + subst(sc, hj, v[i - 1]);
            diag = H[i];
            int32_t ext = E[i] + e, opn = H[i] + o;
            uint8_t tv = (uint8_t)(ext < opn ? TB_HORI_OPEN : TB_HORI);
            E[i] = max2(ext, opn);
            ext = F + e; opn = up + o;
            tv |= (uint8_t)(ext < opn ? TB_VERT_OPEN : TB_VERT);
            F = max2(ext, opn);
            /* tmp2 = max(F, E): F is the left argument, ties go to F. */
            const int32_t fe = (F < E[i]) ? E[i] : F;
            const uint8_t tfe = (F < E[i]) ? TB_MAX_HORI : TB_MAX_VERT;
            /* H = max(diag, max(F,E)): diagonal is the left argument. */
            int32_t x;
            if (d < fe) { x = fe; tv |= tfe; } else { x = d; tv |= TB_DIAG; }
            /* Local: a cell that does not beat 0 becomes 0 and its whole
             * trace value NONE (dp_formula_affine.h:110-118, the single-
             * trace _maxScore of dp_formula.h:181-196). */
            if (local && !(0 < x)) { x = 0; tv = TB_NONE; }
            Tc[i] = tv;
            H[i] = x;
            up = x;
            TRACK(col, i == m ? 2 : 1, i, j, x, E[i], F);
        }
    }
#undef TRACK

    /* gap_open == gap_extend: the reference runs its linear-gap algorithm
     * (_usesAffineGaps, dp_setup.h:244-260; dp_formula_linear.h:62-106).
     * Its cell values are these, and so is each cell's own direction
     * (diagonal wins ties, then vertical), but it has no gap matrices to
     * follow: every step re-reads the direction of the cell it reaches, and
     * the end cell is taken as it is (no _correctTraceValue for linear gaps,
     * dp_algorithm_impl.h:1180-1188; PreferGapsAtEnd_ is false,
     * dp_traceback_impl.h:115-122). */
    const int linear = (o == e);
    uint8_t tv = T[bj * R + bi];
    if (!linear) {
        /* A band whose lower diagonal ends on the last row: that cell is
         * tracked from the next column's code path (dp_algorithm_impl.h:
         * 843-862), and the reference's traceback takes its own direction
         * there, with no preference for the vertical gap. */
        if (banded && bi == m && bj - bi == lower)
            bestF = NEG_INF;
        /* _correctTraceValue: at the best cell prefer the gap matrices. */
        if (bestF == best) {
            tv = (uint8_t)((tv & ~TB_DIAG) | TB_MAX_VERT);
        } else if (bestE == best) {
            tv = (uint8_t)((tv & ~TB_DIAG) | TB_MAX_HORI);
        }
        /* _retrieveInitialTraceDirection with PreferGapsAtEnd_ (affine). */
        if (tv & TB_MAX_VERT)
            tv &= (uint8_t)(TB_VERT | TB_VERT_OPEN | TB_MAX_VERT);
        else if (tv & TB_MAX_HORI)
            tv &= (uint8_t)(TB_HORI | TB_HORI_OPEN | TB_MAX_HORI);
    }

    /* Walk back, emitting ops in reverse order. */
    char *rev = (char *)malloc((size_t)(n + m + 2));
    int64_t nr = 0;
    /* Tail: cells after the best cell are free end gaps (global only;
     * TraceTail_ is false for local, dp_profile.h:482-483). */
    if (!local) {
        for (int64_t i = m; i > bi; --i) rev[nr++] = 'V';
        for (int64_t j = n; j > bj; --j) rev[nr++] = 'H';
    }
    int64_t ci = bi, cj = bj;
#define TV(J, I) T[(J) * R + (I)]
    while (!(cj <= 0 || ci <= 0) && tv != TB_NONE) {
        if (linear) {
            /* One step per cell (dp_traceback_impl.h:253-259, 332-338). */
            if (tv & TB_DIAG) {
                rev[nr++] = 'M';
                --ci; --cj;
            } else if (tv & TB_MAX_VERT) {
                rev[nr++] = 'V';
                --ci;
            } else {
                rev[nr++] = 'H';
                --cj;
            }
            tv = TV(cj, ci);
        } else if (tv & TB_DIAG) {
            rev[nr++] = 'M';
            --ci; --cj;
            tv = TV(cj, ci);
        } else if ((tv & TB_MAX_VERT) && (tv & TB_VERT)) {
            while ((!(tv & TB_VERT_OPEN) || (tv & TB_VERT)) && ci != 1) {
                rev[nr++] = 'V';
                --ci;
                tv = TV(cj, ci);
            }
            rev[nr++] = 'V';
            --ci;
            tv = TV(cj, ci);
        } else if ((tv & TB_MAX_VERT) && (tv & TB_VERT_OPEN)) {
            rev[nr++] = 'V';
            --ci;
            tv = TV(cj, ci);
        } else if ((tv & TB_MAX_HORI) && (tv & TB_HORI)) {
            while ((!(tv & TB_HORI_OPEN) || (tv & TB_HORI)) && cj != 1) {
                rev[nr++] = 'H';
                --cj;
                tv = TV(cj, ci);
            }
            rev[nr++] = 'H';
            --cj;
            tv = TV(cj, ci);
        } else if ((tv & TB_MAX_HORI) && (tv & TB_HORI_OPEN)) {
            rev[nr++] = 'H';
            --cj;
            tv = TV(cj, ci);
        } else {
            break; /* NONE */
        }
    }
#undef TV
    /* Head: leading gaps up to the origin (global only, TraceHead_,
     * dp_profile.h:491-492).  Local: the window starts where the walk
     * stopped (the begin positions of _adaptTraceSegmentsTo,
     * dp_traceback_adaptor.h:87-93); no window at all when it is empty. */
    if (!local) {
        for (int64_t i = ci; i > 0; --i) rev[nr++] = 'V';
        for (int64_t j = cj; j > 0; --j) rev[nr++] = 'H';
    } else if (begins && nr > 0) {
        begins[0] = cj;
        begins[1] = ci;
    }

    /* computeAlignmentStats: walk the rows left to right. */
    if (counts) {
        int64_t hi = local ? cj : 0, vi = local ? ci : 0;
        int open0 = 0, open1 = 0;
        for (int64_t k = nr - 1; k >= 0; --k) {
            const char op = rev[k];
            if (op == 'V') {              /* gap in row 0 (H) */
                if (!open0) counts[A * A] += 1; else counts[A * A + 1] += 1;
                open0 = 1; open1 = 0; ++vi;
            } else if (op == 'H') {       /* gap in row 1 (V) */
                if (!open1) counts[A * A] += 1; else counts[A * A + 1] += 1;
                open1 = 1; open0 = 0; ++hi;
            } else {
                counts[(int)h[hi] * A + (int)v[vi]] += 1;
                open0 = open1 = 0; ++hi; ++vi;
            }
        }
    }
    if (ops) {
        for (int64_t k = 0; k < nr; ++k) ops[k] = rev[nr - 1 - k];
    }
    if (n_ops) *n_ops = nr;
    free(rev);
    free(T);
    free(H);
    free(E);
    return best;
#undef OUT
}

int32_t orc_traceback(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                      const orc_scoring *sc, uint32_t mode, int32_t *counts,
                      char *ops, int64_t *n_ops, int64_t *begins)
{
    return traceback_impl(h, n, v, m, sc, mode, counts, ops, n_ops, begins, 0, 0, 0);
}

int32_t orc_traceback_banded(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                             const orc_scoring *sc, uint32_t mode, int64_t lower, int64_t upper,
                             int32_t *counts, char *ops, int64_t *n_ops, int64_t *begins)
{
    return traceback_impl(h, n, v, m, sc, mode, counts, ops, n_ops, begins, 1, lower, upper);
}

void orc_traceback_batch(const uint8_t *h_res, const int64_t *h_off,
                         const uint8_t *v_res, const int64_t *v_off, int64_t count,
                         const orc_scoring *sc, uint32_t mode, int32_t *scores,
                         int32_t *counts, int threads)
{
    const int64_t stride = (int64_t)sc->alpha * sc->alpha + 2;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads_wrapper();
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
#endif
    for (int64_t k = 0; k < count; ++k)
        scores[k] = orc_traceback(h_res + h_off[k], h_off[k + 1] - h_off[k],
                                  v_res + v_off[k], v_off[k + 1] - v_off[k], sc, mode,
                                  counts ? counts + k * stride : NULL, NULL, NULL, NULL);
    (void)threads;
}

void orc_rows_batch_banded(const uint8_t *h_res, const int64_t *h_off,
                           const uint8_t *v_res, const int64_t *v_off, int64_t count,
                           const orc_scoring *sc, uint32_t mode, int64_t lower, int64_t upper,
                           int32_t *scores, int32_t *counts, char *ops, const int64_t *op_off,
                           int64_t *op_len, int64_t *begins, int threads)
{
    const int64_t stride = (int64_t)sc->alpha * sc->alpha + 2;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads_wrapper();
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
#endif
    for (int64_t k = 0; k < count; ++k)
        scores[k] = orc_traceback_banded(h_res + h_off[k], h_off[k + 1] - h_off[k],
                                         v_res + v_off[k], v_off[k + 1] - v_off[k], sc, mode, lower, upper,
                                         counts ? counts + k * stride : NULL, ops + op_off[k], op_len + k,
                                         begins ? begins + 2 * k : NULL);
    (void)threads;
}

void orc_rows_batch(const uint8_t *h_res, const int64_t *h_off,
                    const uint8_t *v_res, const int64_t *v_off, int64_t count,
                    const orc_scoring *sc, uint32_t mode, int32_t *scores,
                    int32_t *counts, char *ops, const int64_t *op_off, int64_t *op_len,
                    int64_t *begins, int threads)
{
    const int64_t stride = (int64_t)sc->alpha * sc->alpha + 2;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads_wrapper();
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
#endif
    for (int64_t k = 0; k < count; ++k)
        scores[k] = orc_traceback(h_res + h_off[k], h_off[k + 1] - h_off[k],
                                  v_res + v_off[k], v_off[k + 1] - v_off[k], sc, mode,
                                  counts ? counts + k * stride : NULL, ops + op_off[k], op_len + k,
                                  begins ? begins + 2 * k : NULL);
    (void)threads;
}

/* ------------------------------------------------------------------------ */
/* Brute force: enumerate every alignment as an op string and score it with  */
/* per-row gap runs (open for the first gap character, extend for the rest). */
/* Free end gaps strip one leading / one trailing run of the free kind; local */
/* takes the best window of any alignment (or 0).                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    const uint8_t *h, *v;
    int64_t n, m;
    const orc_scoring *sc;
    uint32_t mode;
    char ops[32];
    int32_t best;
} bf_state;

static int32_t bf_window(const bf_state *s, int lo, int hi, int64_t h0, int64_t v0)
{
    /* A run of k gap characters in one row costs open + (k-1)*max(open, extend):
     * the recurrence E = max(E + e, H + o) may re-open a gap right after a gap. */
    const int32_t ext = max2(s->sc->gap_extend, s->sc->gap_open);
    int32_t sum = 0;
    int open0 = 0, open1 = 0;
    int64_t hi_ = h0, vi = v0;
    for (int k = lo; k < hi; ++k) {
        const char op = s->ops[k];
        if (op == 'V') {
            sum += open0 ? ext : s->sc->gap_open;
            open0 = 1; open1 = 0; ++vi;
        } else if (op == 'H') {
            sum += open1 ? ext : s->sc->gap_open;
            open1 = 1; open0 = 0; ++hi_;
        } else {
            sum += subst(s->sc, s->h[hi_], s->v[vi]);
            open0 = open1 = 0; ++hi_; ++vi;
        }
    }
    return sum;
}

static void bf_eval(bf_state *s, int len)
{
    if (s->mode & ORC_LOCAL) {
        int32_t best = 0;
        int64_t h0 = 0, v0 = 0;
        for (int lo = 0; lo < len; ++lo) {
            for (int hi = lo + 1; hi <= len; ++hi) {
                const int32_t x = bf_window(s, lo, hi, h0, v0);
                if (x > best) best = x;
            }
            if (s->ops[lo] != 'V') ++h0;
            if (s->ops[lo] != 'H') ++v0;
        }
        if (best > s->best) s->best = best;
        return;
    }
    int lo = 0, hi = len;
    int64_t h0 = 0, v0 = 0;
    if (len > 0 && s->ops[0] == 'H' && (s->mode & ORC_FREE_TOP)) {
        while (lo < hi && s->ops[lo] == 'H') { ++lo; ++h0; }
    } else if (len > 0 && s->ops[0] == 'V' && (s->mode & ORC_FREE_LEFT)) {
        while (lo < hi && s->ops[lo] == 'V') { ++lo; ++v0; }
    }
    if (hi > lo && s->ops[hi - 1] == 'H' && (s->mode & ORC_FREE_BOTTOM)) {
        while (hi > lo && s->ops[hi - 1] == 'H') --hi;
    } else if (hi > lo && s->ops[hi - 1] == 'V' && (s->mode & ORC_FREE_RIGHT)) {
        while (hi > lo && s->ops[hi - 1] == 'V') --hi;
    }
    const int32_t x = bf_window(s, lo, hi, h0, v0);
    if (x > s->best) s->best = x;
}

static void bf_rec(bf_state *s, int64_t i, int64_t j, int len)
{
    if (i == s->m && j == s->n) { bf_eval(s, len); return; }
    if (i < s->m && j < s->n) { s->ops[len] = 'M'; bf_rec(s, i + 1, j + 1, len + 1); }
    if (j < s->n) { s->ops[len] = 'H'; bf_rec(s, i, j + 1, len + 1); }
    if (i < s->m) { s->ops[len] = 'V'; bf_rec(s, i + 1, j, len + 1); }
}

int32_t orc_brute_force(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                        const orc_scoring *sc, uint32_t mode)
{
    if (n <= 0 || m <= 0) return INT32_MIN;
    if (n > 8 || m > 8) return INT32_MIN;
    bf_state s;
    s.h = h; s.v = v; s.n = n; s.m = m; s.sc = sc; s.mode = mode;
    s.best = INT32_MIN;
    bf_rec(&s, 0, 0, 0);
    return s.best;
}
