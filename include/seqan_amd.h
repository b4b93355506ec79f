/*
 * seqan_amd.h -- C ABI of the MI355X (gfx950) batched pairwise-alignment
 * engine.  Plain C types only; every pointer is a plain address and every size
 * an integer.  All entry points return an sqmi_status and never throw.
 *
 * The entry points replace these reference interfaces (SeqAn 2, paths relative
 * to its source tree):
 *
 *   sqmi_align_score   <- globalAlignmentScore(ExecutionPolicy<Parallel,Vectorial>,
 *                           setH, setV, score [, AlignConfig<T,L,R,B>])
 *                           include/seqan/align_parallel/parallel_align_interface.h:332-343
 *                         localAlignmentScore(ExecutionPolicy<Parallel,Vectorial>,
 *                           setH, setV, score)
 *                           include/seqan/align_parallel/parallel_align_interface.h:354-365
 *   sqmi_align_counts  <- globalAlignment(ExecutionPolicy<Parallel,Vectorial>,
 *                           gapsSetH, gapsSetV, score)
 *                           include/seqan/align_parallel/parallel_align_interface.h:288-299
 *                         followed by computeAlignmentStats on the rows
 *                           include/seqan/align/evaluate_alignment.h:215-300
 *   sqmi_align_rows    <- the same globalAlignment call, returning the rows
 *                           its traceback writes into the Gaps objects
 *                           (dp_traceback_impl.h:115-546, gaps set through
 *                           include/seqan/align/dp_traceback_adaptor.h)
 *   sqmi_align_score_banded <- globalAlignmentScore(ExecutionPolicy<Parallel,Vectorial>,
 *                           setH, setV, score [, AlignConfig], lowerDiag, upperDiag)
 *                           (the executor forwards to the banded set overload,
 *                           include/seqan/align/global_alignment_banded.h:555-583;
 *                           one pair: :394-470), and with SQMI_LOCAL
 *                           localAlignmentScore(..., lowerDiag, upperDiag)
 *                           (local_alignment_banded.h:317-383)
 *   sqmi_align_rows_banded <- globalAlignment(gapsH, gapsV, score [, AlignConfig],
 *                           lowerDiag, upperDiag)
 *                           include/seqan/align/global_alignment_banded.h:152-181
 *                           (sets: :655-760), and localAlignment(gapsH, gapsV,
 *                           score, lowerDiag, upperDiag)
 *                           include/seqan/align/local_alignment_banded.h:242-280
 *   sqmi_align_local_rows <- localAlignment(ExecutionPolicy<Parallel,Vectorial>,
 *                           gapsSetH, gapsSetV, score)
 *                           include/seqan/align_parallel/parallel_align_interface.h:312-321,
 *                           set overload local_alignment_unbanded.h:367-396:
 *                           the rows of the local window and its clipping
 *                           (dp_traceback_adaptor.h:60-117)
 *
 * Semantics follow the reference exactly: affine gaps where a gap of k
 * characters costs gap_open + (k-1)*gap_extend (Score<int,Simple> constructor
 * order match, mismatch, gapExtend, gapOpen, include/seqan/score/score_simple.h:94-97),
 * free end gaps per AlignConfig<TOP,LEFT,RIGHT,BOTTOM> (dp_setup.h:57-182),
 * and an empty sequence yields INT32_MIN (dp_algorithm_impl.h:1399-1400).
 */
#ifndef SEQAN_AMD_H
#define SEQAN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SQMI_ABI_VERSION 7   /* 2: sqmi_stats.pairs_long, pairs_long_packed; sqmi_align_rows
                                3: sqmi_align_local_rows; SQMI_LOCAL in sqmi_align_counts;
                                   sqmi_align_score_banded
                                4: sqmi_align_rows_banded
                                5: sqmi_stats.pairs_lane; one-pair-per-lane shapes (G = 1);
                                   sqmi_align_score_noalloc
                                6: sqmi_stats.replayed: identical sqmi_align_score calls relaunch
                                   the previous call's HIP graph
                                7: SQMI_ROWS_PACKED */

typedef enum sqmi_status {
    SQMI_OK = 0,
    SQMI_ERR_INVALID_ARGUMENT = 1,  /* null pointer, bad offsets, bad scoring    */
    SQMI_ERR_LENGTH_MISMATCH = 2,   /* |setH| != |setV|; the reference asserts,
                                       parallel_align_interface.h:111           */
    SQMI_ERR_UNSUPPORTED = 3,       /* request outside the library's scope      */
    SQMI_ERR_NO_DEVICE = 4,         /* no usable gfx950 device                  */
    SQMI_ERR_DEVICE = 5,            /* HIP runtime error (see sqmi_last_error)  */
    SQMI_ERR_OUT_OF_MEMORY = 6
} sqmi_status;

/* Alignment mode: AlignConfig<TOP, LEFT, RIGHT, BOTTOM> bits, or LOCAL. */
enum {
    SQMI_FREE_TOP = 1,      /* first row free: leading gaps in V are free     */
    SQMI_FREE_LEFT = 2,     /* first column free: leading gaps in H are free  */
    SQMI_FREE_RIGHT = 4,    /* last column free: trailing gaps in H are free  */
    SQMI_FREE_BOTTOM = 8,   /* last row free: trailing gaps in V are free     */
    SQMI_LOCAL = 16         /* Smith-Waterman (localAlignmentScore)           */
};

/* Call flags. */
enum {
    SQMI_RESIDUES_ON_DEVICE = 1,  /* residues pointers are device addresses   */
    SQMI_SCORES_ON_DEVICE = 2,    /* scores/counts pointers are device addresses */
    SQMI_FORCE_INT32 = 4,         /* skip the packed int16 kernels             */
    SQMI_PLAIN_PER_PAIR = 8,      /* one kernel launch per pair (baseline form) */
    SQMI_ROWS_PACKED = 0x1000000  /* rows calls: rows back to back in pair order */
};
/* Tuning knob: force the packed kernel shape instead of the planner's choice
 * (0 means automatic): lanes per pair group G and rows per lane R, G = 16
 * with R in {8, 16} (a group of 16 lanes sweeps a pair couple as a skewed
 * wavefront) or G = 1 with R in {16, 32} (one pair couple per lane).  The
 * automatic choice follows the shape table built from
 * seqan_amd/tuning/shapes.json, per length bin and number of pairs. */
#define SQMI_FORCE_SHAPE(G, R) ((((uint32_t)(G) & 0xFFu) << 8) | (((uint32_t)(R) & 0xFFu) << 16))

/* A set of sequences: concatenated residue ordinals (Dna: A0 C1 G2 T3;
 * AminoAcid: SeqAn ordinal order ABCDEFGHIJKLMNOPQRSTUVWYZX*) and count+1
 * offsets.  Offsets always live in host memory; residues in host or device
 * memory (SQMI_RESIDUES_ON_DEVICE).  Residues must be < alphabet_size. */
typedef struct sqmi_sequences {
    const uint8_t *residues;
    const int64_t *offsets;
    int64_t count;
} sqmi_sequences;

/* Scoring scheme: a substitution table over the alphabet (row = residue of
 * the horizontal sequence H, column = residue of V), plus affine gap costs.
 * Score<int,Simple>(match, mismatch, ext, open) is the table with match on the
 * diagonal and mismatch elsewhere; Blosum62(ext, open) is sqmi_blosum62(). */
typedef struct sqmi_scoring {
    int32_t alphabet_size;      /* 1..32                                        */
    int32_t gap_open;           /* cost of the first gap character              */
    int32_t gap_extend;         /* cost of every further gap character          */
    const int32_t *table;       /* alphabet_size^2 entries, host memory         */
} sqmi_scoring;

/* Counts layout of sqmi_align_counts: per pair, alphabet_size^2 substitution
 * counts (index h*alphabet_size + v), then gap opens, then gap extensions.
 * For each pair sum(count * parameter) equals the score. */
#define SQMI_COUNTS_STRIDE(alpha) ((alpha) * (alpha) + 2)

/* Per-call statistics (filled by the last call on the context). */
typedef struct sqmi_stats {
    int64_t pairs;              /* pairs in the call                            */
    int64_t cells;              /* sum of |H_i| * |V_i|                         */
    int64_t pairs_int16;        /* pairs computed by the packed int16 kernels   */
    int64_t pairs_int32;        /* pairs computed in int32 (range guard)        */
    int64_t work_items;         /* wavefront work items launched                */
    int64_t padded_cells;       /* cells computed including padding            */
    double host_plan_ms;        /* host planning (bucketing) time               */
    double device_ms;           /* device time of the kernels (HIP events)      */
    double total_ms;            /* wall time of the call                        */
    int64_t pairs_long;         /* of pairs_int32: long pairs, tiled kernels    */
    int64_t pairs_long_packed;  /* of pairs_long: forward in packed int16       */
    int64_t pairs_lane;         /* of pairs_int16: one pair couple per lane      */
    int64_t replayed;           /* 1: the call relaunched the previous call's    */
                                /* graph (same lengths, scoring, mode, flags and */
                                /* device addresses; host_plan_ms = key time)    */
} sqmi_stats;

typedef struct sqmi_context sqmi_context;

/* Creates a context on HIP device `device` (owns scratch, stream, tuning). */
sqmi_status sqmi_create(int device, sqmi_context **ctx);
sqmi_status sqmi_destroy(sqmi_context *ctx);

/* Use the given hipStream_t for subsequent calls (NULL: the context's own). */
sqmi_status sqmi_set_stream(sqmi_context *ctx, void *hip_stream);

/* Optimal alignment score of every pair (setH[i], setV[i]).  Blocks until the
 * scores are written.  scores: count int32 (host, or device with
 * SQMI_SCORES_ON_DEVICE).  The call's kernel launches are captured in a HIP
 * graph; the next call on the context with the same lengths (offsets),
 * scoring, mode and flags, the same device addresses (residues and scores on
 * the device) and the same planner switches skips planning and relaunches
 * that graph (the residues themselves may differ).  Any other call on the
 * context drops it; SQMI_GRAPH=0 in the environment turns this off. */
sqmi_status sqmi_align_score(sqmi_context *ctx, const sqmi_sequences *set_h,
                             const sqmi_sequences *set_v, const sqmi_scoring *scoring,
                             uint32_t mode, int32_t *scores, uint32_t flags);

/* Score plus the exact subgradient of the score with respect to the scoring
 * parameters: substitution counts, gap opens and gap extensions along the
 * optimal path chosen by the reference traceback (single trace, gaps left,
 * gaps preferred at the end cell), counted on the rows like
 * computeAlignmentStats (free end gaps included; for SQMI_LOCAL the rows of
 * the local window).  Global alignment with any free-end bits, or SQMI_LOCAL
 * (free-end bits beside it are ignored, as in sqmi_align_score).  Pairs whose
 * values fit int16 run on the packed trace kernel in every mode. */
sqmi_status sqmi_align_counts(sqmi_context *ctx, const sqmi_sequences *set_h,
                              const sqmi_sequences *set_v, const sqmi_scoring *scoring,
                              uint32_t mode, int32_t *scores, int32_t *counts,
                              uint32_t flags);

/* sqmi_align_counts plus the alignment rows: one byte per alignment column,
 * 'M' (a residue of H over a residue of V), 'V' (a gap in the H row: a
 * residue of V alone) or 'H' (a gap in the V row: a residue of H alone), in
 * left-to-right order.  Pair i may use ops[op_offsets[i] .. op_offsets[i+1]),
 * which must hold at least |H_i| + |V_i| bytes; its rows end at
 * op_offsets[i+1] and start at op_offsets[i+1] - op_lengths[i] (the bytes
 * before them are unspecified).  op_offsets: count + 1 int64 in host memory;
 * ops and op_lengths (count int64) are host memory, or device memory with
 * SQMI_SCORES_ON_DEVICE.  With SQMI_ROWS_PACKED the rows are instead written
 * back to back in pair order (packed on the device): row i starts at
 * op_offsets[0] + op_lengths[0] + ... + op_lengths[i-1]; op_offsets still
 * bound the space as above, and the bytes after the last row are
 * unspecified (host buffers: left as they were).  An empty pair gets
 * INT32_MIN, zero counts and no
 * rows.  Global modes as sqmi_align_counts (SQMI_LOCAL: sqmi_align_local_rows). */
sqmi_status sqmi_align_rows(sqmi_context *ctx, const sqmi_sequences *set_h,
                            const sqmi_sequences *set_v, const sqmi_scoring *scoring,
                            uint32_t mode, int32_t *scores, int32_t *counts,
                            uint8_t *ops, const int64_t *op_offsets, int64_t *op_lengths,
                            uint32_t flags);

/* Local alignment (Smith-Waterman with affine gaps): score, counts and the
 * rows of the local window, laid out as in sqmi_align_rows, plus begins
 * (count x 2 int64, host memory or device memory with SQMI_SCORES_ON_DEVICE):
 * the positions in H and in V of the window's first residue.  The window
 * covers H[begins[2i], begins[2i] + #M + #H) and V[begins[2i+1],
 * begins[2i+1] + #M + #V).  The end cell is the first strict maximum in
 * column order and the walk stops at the first cell that does not beat 0,
 * as in the reference.  A pair whose best score is 0 gets an empty window at
 * (0, 0), zero counts and score 0; an empty pair gets INT32_MIN. */
sqmi_status sqmi_align_local_rows(sqmi_context *ctx, const sqmi_sequences *set_h,
                                  const sqmi_sequences *set_v, const sqmi_scoring *scoring,
                                  int32_t *scores, int32_t *counts, uint8_t *ops,
                                  const int64_t *op_offsets, int64_t *op_lengths,
                                  int64_t *begins, uint32_t flags);

/* Banded alignment scores: only cells (i, j) with lower_diag <= j - i <=
 * upper_diag exist (i a position in V, j in H), under the free-end bits of
 * mode, or SQMI_LOCAL.  Global modes: every non-empty pair must have
 * |H| >= 2 and lower_diag < min(0, |H| - |V|), max(0, |H| - |V|) < upper_diag:
 * the band holds both corner diagonals with a margin (the reference accepts
 * only bands holding both corners unless ends are free,
 * dp_algorithm_impl.h:192-232, and outside this domain its special band
 * paths depart from the banded recurrence).  SQMI_LOCAL: |H| >= 2, at least
 * two diagonals inside the matrix, lower_diag <= |H| - 2 and upper_diag >=
 * 2 - |V|.  Otherwise the call returns SQMI_ERR_UNSUPPORTED.  Empty pairs get
 * INT32_MIN.  Every pair follows the
 * reference's banded DP for one pair (global_alignment_banded.h:394-470);
 * the reference's vectorised set path agrees with it when all H and all V
 * have one length each, the case its documentation supports.  Flags: only
 * the *_ON_DEVICE ones. */
sqmi_status sqmi_align_score_banded(sqmi_context *ctx, const sqmi_sequences *set_h,
                                    const sqmi_sequences *set_v, const sqmi_scoring *scoring,
                                    uint32_t mode, int64_t lower_diag, int64_t upper_diag,
                                    int32_t *scores, uint32_t flags);

/* Banded alignment rows: sqmi_align_rows (global modes) or
 * sqmi_align_local_rows (SQMI_LOCAL) restricted to the band of
 * sqmi_align_score_banded, with the same supported bands.  ops may be NULL
 * for scores and counts only (op_offsets, op_lengths and begins are then
 * not read); begins is required with SQMI_LOCAL and ops, and not used in
 * global modes.  Paths follow the reference's banded traceback for one pair,
 * including its end cell on a band whose lower diagonal ends on the last
 * row, where it takes the cell's own direction.  Flags: only the
 * *_ON_DEVICE ones. */
sqmi_status sqmi_align_rows_banded(sqmi_context *ctx, const sqmi_sequences *set_h,
                                   const sqmi_sequences *set_v, const sqmi_scoring *scoring,
                                   uint32_t mode, int64_t lower_diag, int64_t upper_diag,
                                   int32_t *scores, int32_t *counts, uint8_t *ops,
                                   const int64_t *op_offsets, int64_t *op_lengths,
                                   int64_t *begins, uint32_t flags);

/* sqmi_align_score without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as sqmi_align_score.  Every buffer the call
 * needs beyond the caller's inputs and outputs (device copies of host
 * residues, offsets and scores, work lists, row buffers, int32 scratch) is
 * carved from one device array of SQMI_NOALLOC_SCRATCH_BYTES per GPU that the
 * library holds for the life of the process; the call makes no hipMalloc or
 * hipFree.  A request that needs more returns SQMI_ERR_OUT_OF_MEMORY before
 * anything is copied or launched; pairs on the long-pair route and
 * SQMI_PLAIN_PER_PAIR return SQMI_ERR_UNSUPPORTED.  All calls on one GPU
 * share its array and leave their contents in it: calls that overlap on one
 * GPU (other threads, other contexts) race. */
#define SQMI_NOALLOC_SCRATCH_BYTES (64u << 20)
sqmi_status sqmi_align_score_noalloc(sqmi_context *ctx, const sqmi_sequences *set_h,
                                     const sqmi_sequences *set_v, const sqmi_scoring *scoring,
                                     uint32_t mode, int32_t *scores, uint32_t flags);

sqmi_status sqmi_get_stats(const sqmi_context *ctx, sqmi_stats *out);

/* Host-only dry run of the planner: validates the request exactly as
 * sqmi_align_score does and reports how the pairs would be routed (int16 /
 * int32 / empty) and bucketed into work items.  Needs no device.  (No
 * counterpart in the reference; the reference has no planning step.) */
sqmi_status sqmi_plan_stats(const sqmi_sequences *set_h, const sqmi_sequences *set_v,
                            const sqmi_scoring *scoring, uint32_t mode, uint32_t flags,
                            sqmi_stats *out);

/* The reference's Blosum62 table (27 x 27, AminoAcid ordinal order). */
const int32_t *sqmi_blosum62(void);

const char *sqmi_status_string(sqmi_status status);
/* Text of the last error on this context ("" when none). */
const char *sqmi_last_error(const sqmi_context *ctx);
int sqmi_abi_version(void);

/* ------------------------------------------------------------------------ */
/* Synthetic workloads (benchmarking and tests; no counterpart in the        */
/* reference).  Pairs are generated independently from (seed, index), so any */
/* window [first, first+count) is reproducible on its own.                   */
/* ------------------------------------------------------------------------ */
enum {
    SQMI_WL_DNA150 = 1,         /* 150 bp, 5% subst, 1% single-base indels     */
    SQMI_WL_DNA_LOGNORMAL = 2,  /* log-normal 100..10k bp, 8% subst, 2% indels */
    SQMI_WL_PROTEIN = 3,        /* 50..1000 aa vs 50..2000 aa, 30% homologs    */
    SQMI_WL_DNA1K = 4,          /* 1 kbp, 10% subst, 2% indels                 */
    SQMI_WL_DNA100K = 5         /* 100 kbp, 12% subst, 3% geometric indels     */
};
typedef struct sqmi_workload sqmi_workload;
/* length_override > 0 replaces the configured H length (DNA configs). */
sqmi_status sqmi_workload_create(int config, uint64_t seed, int64_t first, int64_t count,
                                 int64_t length_override, sqmi_workload **out);
sqmi_status sqmi_workload_sizes(const sqmi_workload *w, int64_t *h_total, int64_t *v_total);
sqmi_status sqmi_workload_copy(const sqmi_workload *w, uint8_t *h_res, int64_t *h_off,
                               uint8_t *v_res, int64_t *v_off);
sqmi_status sqmi_workload_destroy(sqmi_workload *w);

#ifdef __cplusplus
}
#endif

#endif /* SEQAN_AMD_H */
