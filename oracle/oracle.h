/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of SeqAn 2's affine-gap (Gotoh) pairwise alignment as used by
 * globalAlignmentScore / localAlignmentScore / globalAlignment.  It exists to
 * check the GPU library; nothing in the product links or calls it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Mode bits mirror AlignConfig<TOP, LEFT, RIGHT, BOTTOM>
 * (include/seqan/align/align_config.h:92-94) as mapped by SubstituteAlignConfig_
 * (include/seqan/align/dp_setup.h:57-182): TOP = first row free, LEFT = first
 * column free, RIGHT = last column free, BOTTOM = last row free; LOCAL selects
 * LocalAlignment_ (include/seqan/align/local_alignment_unbanded.h:554-586).
 */
#ifndef SEQAN_AMD_ORACLE_H
#define SEQAN_AMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_FREE_TOP = 1,
    ORC_FREE_LEFT = 2,
    ORC_FREE_RIGHT = 4,
    ORC_FREE_BOTTOM = 8,
    ORC_LOCAL = 16
};

typedef struct {
    int32_t alpha;          /* alphabet size; table is alpha x alpha            */
    const int32_t *table;   /* score(h, v) = table[h * alpha + v]               */
    int32_t gap_open;       /* cost of the first gap character (negative)       */
    int32_t gap_extend;     /* cost of each further gap character               */
} orc_scoring;

/* Score of one pair; INT32_MIN when either sequence is empty. */
int32_t orc_score(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                  const orc_scoring *sc, uint32_t mode);

/* Scores of a batch given as concatenated residues + (count+1) offsets.
 * Runs on up to `threads` OpenMP threads (<= 0: all). */
void orc_score_batch(const uint8_t *h_res, const int64_t *h_off,
                     const uint8_t *v_res, const int64_t *v_off, int64_t count,
                     const orc_scoring *sc, uint32_t mode, int32_t *scores,
                     int threads);

/* Traceback of one pair under SeqAn's single-trace, gaps-left tie-break.
 * counts receives alpha*alpha substitution counts (index h*alpha+v), then the
 * number of gap opens and of gap extensions (computeAlignmentStats semantics).
 * ops (optional, capacity n+m) receives the path from start to end as
 * 'M' (diagonal), 'H' (horizontal: H residue vs gap) and 'V' (vertical:
 * gap vs V residue); *n_ops its length.  With ORC_LOCAL the ops are those of
 * the local window and begins (optional, 2 values) receives the positions in
 * H and in V where it starts (0, 0 for an empty window).  Sizes whose
 * (n+1)(m+1) trace matrix fits in memory.  Returns the score. */
int32_t orc_traceback(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                      const orc_scoring *sc, uint32_t mode, int32_t *counts,
                      char *ops, int64_t *n_ops, int64_t *begins);

/* Banded score: cells with lower <= j - i <= upper only; band settings the
 * reference rejects give INT32_MIN.  Global modes, or ORC_LOCAL (the band
 * need only meet the matrix). */
int32_t orc_score_banded(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                         const orc_scoring *sc, uint32_t mode, int64_t lower, int64_t upper);

/* orc_traceback inside the band lower <= j - i <= upper (cells outside at
 * -inf); band settings the reference rejects give INT32_MIN, as in
 * orc_score_banded. */
int32_t orc_traceback_banded(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                             const orc_scoring *sc, uint32_t mode, int64_t lower, int64_t upper,
                             int32_t *counts, char *ops, int64_t *n_ops, int64_t *begins);

/* orc_rows_batch for banded alignments. */
void orc_rows_batch_banded(const uint8_t *h_res, const int64_t *h_off,
                           const uint8_t *v_res, const int64_t *v_off, int64_t count,
                           const orc_scoring *sc, uint32_t mode, int64_t lower, int64_t upper,
                           int32_t *scores, int32_t *counts, char *ops, const int64_t *op_off,
                           int64_t *op_len, int64_t *begins, int threads);

/* Batch form of orc_traceback (counts stride alpha*alpha + 2). */
void orc_traceback_batch(const uint8_t *h_res, const int64_t *h_off,
                         const uint8_t *v_res, const int64_t *v_off, int64_t count,
                         const orc_scoring *sc, uint32_t mode, int32_t *scores,
                         int32_t *counts, int threads);

/* Batch form returning the rows too: pair k's ops (as orc_traceback) at
 * ops + op_off[k] (capacity >= n + m + 2), their number in op_len[k], its
 * local window's begin positions at begins + 2k (optional). */
void orc_rows_batch(const uint8_t *h_res, const int64_t *h_off,
                    const uint8_t *v_res, const int64_t *v_off, int64_t count,
                    const orc_scoring *sc, uint32_t mode, int32_t *scores,
                    int32_t *counts, char *ops, const int64_t *op_off, int64_t *op_len,
                    int64_t *begins, int threads);

/* Exhaustive maximum over all alignments (tiny inputs only, n, m <= 8). */
int32_t orc_brute_force(const uint8_t *h, int64_t n, const uint8_t *v, int64_t m,
                        const orc_scoring *sc, uint32_t mode);

#ifdef __cplusplus
}
#endif

#endif
