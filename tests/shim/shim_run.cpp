// Runs the SeqAn 2 drop-in overloads (include/seqan_amd_seqan2.hpp) against
// the reference's own CPU path in the same program.
//
// For each case it prints one line "case <name> <status>", where status is
//   match     the overload ran on the GPU and equals the reference's scores,
//   mismatch  it ran and differs (the first differing pair is printed),
//   nodevice  the overload was taken and reported SQMI_ERR_NO_DEVICE,
//   error     any other failure.
// The reference side is the same call under ExecutionPolicy<Parallel, Serial>,
// which the overloads leave to the reference templates
// (parallel_align_interface.h:332-365).
#include "seqan_amd_seqan2.hpp"

#include <cstdio>
#include <random>
#include <string>

using namespace seqan2;

namespace {

template <typename TAlphabet>
void randomPairs(StringSet<String<TAlphabet>> & setH, StringSet<String<TAlphabet>> & setV, unsigned count,
                 unsigned lo, unsigned hi, unsigned seed)
{
    std::mt19937 rng(seed);
    unsigned const alpha = ValueSize<TAlphabet>::VALUE;
    for (unsigned i = 0; i < count; ++i)
    {
        String<TAlphabet> h, v;
        unsigned const n = lo + rng() % (hi - lo + 1);
        for (unsigned k = 0; k < n; ++k)
            appendValue(h, TAlphabet(rng() % alpha));
        for (unsigned k = 0; k < n; ++k)
        {
            unsigned const r = rng() % 100;
            if (r < 8)
                appendValue(v, TAlphabet(rng() % alpha));
            else if (r < 10)
                continue;
            else if (r < 12)
            {
                appendValue(v, TAlphabet(rng() % alpha));
                appendValue(v, h[k]);
            }
            else
                appendValue(v, h[k]);
        }
        if (i % 17 == 3)
            clear(v);  // empty sequences score INT_MIN in both paths
        appendValue(setH, h);
        appendValue(setV, v);
    }
}

// Pairs of one length with substitutions only: collections the reference's
// banded set path supports (one H length, one V length).
template <typename TAlphabet>
void equalPairs(StringSet<String<TAlphabet>> & setH, StringSet<String<TAlphabet>> & setV, unsigned count,
                unsigned len, unsigned seed)
{
    std::mt19937 rng(seed);
    unsigned const alpha = ValueSize<TAlphabet>::VALUE;
    for (unsigned i = 0; i < count; ++i)
    {
        String<TAlphabet> h, v;
        for (unsigned k = 0; k < len; ++k)
        {
            appendValue(h, TAlphabet(rng() % alpha));
            appendValue(v, rng() % 100 < 12 ? TAlphabet(rng() % alpha) : h[k]);
        }
        appendValue(setH, h);
        appendValue(setV, v);
    }
}

template <typename TFn, typename TRef>
void runCase(char const * name, TFn && gpu, TRef && ref)
{
    String<int> expect = ref();
    try
    {
        String<int> got = gpu();
        if (length(got) != length(expect))
        {
            std::printf("case %s mismatch length %zu vs %zu\n", name, (size_t)length(got), (size_t)length(expect));
            return;
        }
        for (size_t i = 0; i < length(got); ++i)
            if (got[i] != expect[i])
            {
                std::printf("case %s mismatch pair %zu: %d vs %d\n", name, i, got[i], expect[i]);
                return;
            }
        std::printf("case %s match %zu\n", name, (size_t)length(got));
    }
    catch (seqan_amd::Error const & e)
    {
        std::printf("case %s %s (%s)\n", name, e.status() == SQMI_ERR_NO_DEVICE ? "nodevice" : "error", e.what());
    }
}

// One row of an alignment as text: '-' for a gap, 'x' for a residue.
template <typename TGaps>
std::string rowText(TGaps & gaps)
{
    std::string out;
    for (auto it = begin(gaps); it != end(gaps); ++it)
        out += isGap(it) ? '-' : 'x';
    return out;
}

// globalAlignment on StringSets of Gaps: scores and both rows of every pair
// against the reference's own globalAlignment (non-empty pairs only).
template <typename TAlphabet, typename TScore, typename TGpu, typename TCpu, typename... TConfig>
void runRowsCase(char const * name, StringSet<String<TAlphabet>> const & setH,
                 StringSet<String<TAlphabet>> const & setV, TScore const & sc, TGpu const & ex, TCpu const & cpu,
                 TConfig const &... cfg)
{
    using TGaps = Gaps<String<TAlphabet> const>;
    StringSet<TGaps> gH, gV, rH, rV;
    for (size_t i = 0; i < length(setH); ++i)
    {
        if (length(setH[i]) == 0 || length(setV[i]) == 0)
            continue;
        appendValue(gH, TGaps(setH[i]));
        appendValue(gV, TGaps(setV[i]));
        appendValue(rH, TGaps(setH[i]));
        appendValue(rV, TGaps(setV[i]));
    }
    String<int> expect = globalAlignment(cpu, rH, rV, sc, cfg...);
    try
    {
        String<int> got = globalAlignment(ex, gH, gV, sc, cfg...);
        for (size_t i = 0; i < length(expect); ++i)
            if (got[i] != expect[i] || rowText(gH[i]) != rowText(rH[i]) || rowText(gV[i]) != rowText(rV[i]))
            {
                std::printf("case %s mismatch pair %zu: %d vs %d\n", name, i, got[i], expect[i]);
                return;
            }
        std::printf("case %s match %zu\n", name, (size_t)length(got));
    }
    catch (seqan_amd::Error const & e)
    {
        std::printf("case %s %s (%s)\n", name, e.status() == SQMI_ERR_NO_DEVICE ? "nodevice" : "error", e.what());
    }
}

// localAlignment on StringSets of Gaps: scores, both clipped rows and the
// clipping of every pair against the reference's own localAlignment.
template <typename TAlphabet, typename TScore, typename TGpu, typename TCpu>
void runLocalRowsCase(char const * name, StringSet<String<TAlphabet>> const & setH,
                      StringSet<String<TAlphabet>> const & setV, TScore const & sc, TGpu const & ex, TCpu const & cpu)
{
    using TGaps = Gaps<String<TAlphabet> const>;
    StringSet<TGaps> gH, gV, rH, rV;
    for (size_t i = 0; i < length(setH); ++i)
    {
        if (length(setH[i]) == 0 || length(setV[i]) == 0)
            continue;
        appendValue(gH, TGaps(setH[i]));
        appendValue(gV, TGaps(setV[i]));
        appendValue(rH, TGaps(setH[i]));
        appendValue(rV, TGaps(setV[i]));
    }
    String<int> expect = localAlignment(cpu, rH, rV, sc);
    auto same = [](TGaps & a, TGaps & b) {
        return rowText(a) == rowText(b) && clippedBeginPosition(a) == clippedBeginPosition(b) &&
               clippedEndPosition(a) == clippedEndPosition(b) && beginPosition(a) == beginPosition(b) &&
               endPosition(a) == endPosition(b);
    };
    try
    {
        String<int> got = localAlignment(ex, gH, gV, sc);
        for (size_t i = 0; i < length(expect); ++i)
            if (got[i] != expect[i] || !same(gH[i], rH[i]) || !same(gV[i], rV[i]))
            {
                std::printf("case %s mismatch pair %zu: %d vs %d\n", name, i, got[i], expect[i]);
                return;
            }
        std::printf("case %s match %zu\n", name, (size_t)length(got));
    }
    catch (seqan_amd::Error const & e)
    {
        std::printf("case %s %s (%s)\n", name, e.status() == SQMI_ERR_NO_DEVICE ? "nodevice" : "error", e.what());
    }
}

// Banded globalAlignment / localAlignment on StringSets of Gaps against the
// reference's banded function for one pair (global_alignment_banded.h:
// 152-181, local_alignment_banded.h:242-259), pair by pair.
template <typename TAlphabet, typename TScore, typename TGpu, typename... TConfig>
void runBandedRowsCaseImpl(char const * name, bool local, bool viaExecutor, StringSet<String<TAlphabet>> const & setH,
                           StringSet<String<TAlphabet>> const & setV, TScore const & sc, TGpu const & ex, int lo,
                           int up, TConfig const &... cfg)
{
    using TGaps = Gaps<String<TAlphabet> const>;
    StringSet<TGaps> gH, gV, rH, rV;
    String<int> expect;
    for (size_t i = 0; i < length(setH); ++i)
    {
        appendValue(gH, TGaps(setH[i]));
        appendValue(gV, TGaps(setV[i]));
        appendValue(rH, TGaps(setH[i]));
        appendValue(rV, TGaps(setV[i]));
    }
    if (viaExecutor)
        expect = local ? localAlignment<Parallel, Vectorial>(ex, rH, rV, sc, lo, up)
                       : globalAlignment<Parallel, Vectorial>(ex, rH, rV, sc, cfg..., lo, up);
    else
        for (size_t i = 0; i < length(rH); ++i)
            appendValue(expect, local ? localAlignment(rH[i], rV[i], sc, lo, up)
                                      : globalAlignment(rH[i], rV[i], sc, cfg..., lo, up));
    auto same = [](TGaps & a, TGaps & b) {
        return rowText(a) == rowText(b) && clippedBeginPosition(a) == clippedBeginPosition(b) &&
               clippedEndPosition(a) == clippedEndPosition(b) && beginPosition(a) == beginPosition(b) &&
               endPosition(a) == endPosition(b);
    };
    try
    {
        String<int> got = local ? localAlignment(ex, gH, gV, sc, lo, up) : globalAlignment(ex, gH, gV, sc, cfg..., lo, up);
        for (size_t i = 0; i < length(expect); ++i)
            if (got[i] != expect[i] || !same(gH[i], rH[i]) || !same(gV[i], rV[i]))
            {
                std::printf("case %s mismatch pair %zu: %d vs %d\n", name, i, got[i], expect[i]);
                return;
            }
        std::printf("case %s match %zu\n", name, (size_t)length(got));
    }
    catch (seqan_amd::Error const & e)
    {
        std::printf("case %s %s (%s)\n", name, e.status() == SQMI_ERR_NO_DEVICE ? "nodevice" : "error", e.what());
    }
}

// Banded rows on a single pair each (the reference's banded function for one
// pair), or, viaExecutor, on the whole sets through the reference's own
// ExecutionPolicy<Parallel, Vectorial> executor.
template <typename TAlphabet, typename TScore, typename TGpu, typename... TConfig>
void runBandedRowsCase(char const * name, bool local, StringSet<String<TAlphabet>> const & setH,
                       StringSet<String<TAlphabet>> const & setV, TScore const & sc, TGpu const & ex, int lo, int up,
                       TConfig const &... cfg)
{
    runBandedRowsCaseImpl(name, local, false, setH, setV, sc, ex, lo, up, cfg...);
}

}  // namespace

int main()
{
    std::setvbuf(stdout, nullptr, _IOLBF, 0);   // one line per case, as it finishes
    ExecutionPolicy<Parallel, Vectorial> ex;
    ExecutionPolicy<Parallel, Serial> cpu;
    setNumThreads(cpu, 4);

    StringSet<String<Dna>> dh, dv;
    randomPairs(dh, dv, 300, 1, 400, 7);
    Score<int, Simple> dna(2, -3, -1, -5);
    AlignConfig<true, false, false, true> semi;
    AlignConfig<false, true, true, false> overlap;

    runCase("dna_global_nonconst", [&] { return globalAlignmentScore(ex, dh, dv, dna); },
            [&] { return globalAlignmentScore(cpu, dh, dv, dna); });
    StringSet<String<Dna>> const & cdh = dh;
    StringSet<String<Dna>> const & cdv = dv;
    runCase("dna_semiglobal_const", [&] { return globalAlignmentScore(ex, cdh, cdv, dna, semi); },
            [&] { return globalAlignmentScore(cpu, cdh, cdv, dna, semi); });
    runCase("dna_overlap_rvalue",
            [&] { return globalAlignmentScore(ex, StringSet<String<Dna>>(dh), StringSet<String<Dna>>(dv),
                                              Score<int, Simple>(2, -3, -1, -5), AlignConfig<false, true, true, false>()); },
            [&] { return globalAlignmentScore(cpu, dh, dv, dna, overlap); });
    runCase("dna_local", [&] { return localAlignmentScore(ex, dh, dv, dna); },
            [&] { return localAlignmentScore(cpu, dh, dv, dna); });

    StringSet<String<Dna>, Owner<ConcatDirect<>>> ch, cv;
    for (size_t i = 0; i < length(dh); ++i)
    {
        appendValue(ch, dh[i]);
        appendValue(cv, dv[i]);
    }
    runCase("dna_concat_direct", [&] { return globalAlignmentScore(ex, ch, cv, dna, semi); },
            [&] { return globalAlignmentScore(cpu, dh, dv, dna, semi); });

    StringSet<String<Dna5>> nh, nv;
    randomPairs(nh, nv, 200, 1, 300, 11);
    Score<int, Simple> lin(1, -1, -1, -1);  // gapOpen == gapExtend: the reference's linear-gap dispatch
    runCase("dna5_linear", [&] { return globalAlignmentScore(ex, nh, nv, lin); },
            [&] { return globalAlignmentScore(cpu, nh, nv, lin); });

    StringSet<String<AminoAcid>> ph, pv;
    randomPairs(ph, pv, 200, 1, 500, 13);
    Blosum62 b62(-1, -11);
    runCase("protein_local_blosum62", [&] { return localAlignmentScore(ex, ph, pv, b62); },
            [&] { return localAlignmentScore(cpu, ph, pv, b62); });
    runCase("protein_global_blosum62", [&] { return globalAlignmentScore(ex, ph, pv, b62); },
            [&] { return globalAlignmentScore(cpu, ph, pv, b62); });
    runCase("protein_semiglobal_blosum62", [&] { return globalAlignmentScore(ex, ph, pv, b62, semi); },
            [&] { return globalAlignmentScore(cpu, ph, pv, b62, semi); });

    // Other matrices: PAM120 and one selected at run time.
    Pam120 pam(-2, -10);
    runCase("protein_global_pam120", [&] { return globalAlignmentScore(ex, ph, pv, pam); },
            [&] { return globalAlignmentScore(cpu, ph, pv, pam); });
    SelectableAminoAcidMatrix sel(-1, -8);
    setScoreMatrixById(sel, AminoAcidScoreMatrixID::BLOSUM80);
    runCase("protein_local_selectable_blosum80", [&] { return localAlignmentScore(ex, ph, pv, sel); },
            [&] { return localAlignmentScore(cpu, ph, pv, sel); });

    // The wavefront policy returns std::vector<int>.
    ExecutionPolicy<WavefrontAlignment<BlockOffsetOptimization>, Vectorial> wave;
    auto asString = [](std::vector<int> const & v) {
        String<int> s;
        for (int x : v)
            appendValue(s, x);
        return s;
    };
    runCase("wavefront_global_semi", [&] { return asString(globalAlignmentScore(wave, dh, dv, dna, semi)); },
            [&] { return globalAlignmentScore(cpu, dh, dv, dna, semi); });
    runCase("wavefront_local_blosum62", [&] { return asString(localAlignmentScore(wave, ph, pv, b62)); },
            [&] { return localAlignmentScore(cpu, ph, pv, b62); });

    // Banded scores (the executor forwards lowerDiag, upperDiag to the banded
    // set overload, global_alignment_banded.h:555-583).
    StringSet<String<Dna>> eh, ev;
    equalPairs(eh, ev, 200, 120, 21);
    runCase("dna_banded", [&] { return globalAlignmentScore(ex, eh, ev, dna, -5, 5); },
            [&] { return globalAlignmentScore(cpu, eh, ev, dna, -5, 5); });
    runCase("dna_banded_semiglobal", [&] { return globalAlignmentScore(ex, eh, ev, dna, semi, -8, 12); },
            [&] { return globalAlignmentScore(cpu, eh, ev, dna, semi, -8, 12); });
    StringSet<String<AminoAcid>> eph, epv;
    equalPairs(eph, epv, 96, 100, 22);
    runCase("protein_banded_blosum62", [&] { return globalAlignmentScore(ex, eph, epv, b62, -70, 70); },
            [&] { return globalAlignmentScore(cpu, eph, epv, b62, -70, 70); });
    runCase("protein_local_banded_blosum62", [&] { return localAlignmentScore(ex, eph, epv, b62, -20, 30); },
            [&] { return localAlignmentScore(cpu, eph, epv, b62, -20, 30); });
    // Bands the library does not take (dp_algorithm_impl.h:192-232, 595-700):
    // the overloads hand them to the reference's executor, so they must give
    // what the reference's own ExecutionPolicy<Parallel, Vectorial> call
    // gives (its template named explicitly: the overloads are not viable
    // with these template arguments).  A one-diagonal band, a band whose edge
    // is the corner diagonal under free ends, and a local band that only
    // reaches the corner column.
    runCase("dna_banded_one_diagonal", [&] { return globalAlignmentScore(ex, eh, ev, dna, 0, 0); },
            [&] { return globalAlignmentScore<Parallel, Vectorial>(ex, eh, ev, dna, 0, 0); });
    runCase("dna_banded_corner_semiglobal", [&] { return globalAlignmentScore(ex, eh, ev, dna, semi, 0, 12); },
            [&] { return globalAlignmentScore<Parallel, Vectorial>(ex, eh, ev, dna, semi, 0, 12); });
    runCase("dna_local_banded_corner", [&] { return localAlignmentScore(ex, eh, ev, dna, 119, 125); },
            [&] { return localAlignmentScore<Parallel, Vectorial>(ex, eh, ev, dna, 119, 125); });

    runRowsCase("dna_rows", dh, dv, dna, ex, cpu);
    runRowsCase("protein_rows_blosum62", ph, pv, b62, ex, cpu);
    runRowsCase("dna_rows_semiglobal", dh, dv, dna, ex, cpu, semi);
    runRowsCase("dna_rows_overlap", dh, dv, dna, ex, cpu, overlap);
    runLocalRowsCase("dna_local_rows", dh, dv, dna, ex, cpu);
    // Gap open == gap extend (lin): the reference's linear-gap traceback.
    runRowsCase("dna_rows_linear_gaps", dh, dv, lin, ex, cpu);
    runRowsCase("dna_rows_linear_gaps_semiglobal", dh, dv, lin, ex, cpu, semi);
    runLocalRowsCase("dna_local_rows_linear_gaps", dh, dv, lin, ex, cpu);
    runLocalRowsCase("protein_local_rows_blosum62", ph, pv, b62, ex, cpu);
    // Banded rows (globalAlignment / localAlignment with lowerDiag, upperDiag).
    runBandedRowsCase("dna_banded_rows", false, eh, ev, dna, ex, -5, 5);
    runBandedRowsCase("dna_banded_rows_semiglobal", false, eh, ev, dna, ex, -8, 12, semi);
    runBandedRowsCase("dna_local_banded_rows", true, eh, ev, dna, ex, -6, 9);
    runBandedRowsCase("protein_local_banded_rows_blosum62", true, eph, epv, b62, ex, -20, 30);
    // A rows call the library does not take (one H of a single residue):
    // the overloads hand the whole sets to the reference's executor, so they
    // must give what its own ExecutionPolicy<Parallel, Vectorial> call gives.
    // (No corner-band rows case: the reference's vectorised banded traceback,
    // which such calls fall back to, does not return for a one-diagonal band
    // and returns only now and then for a band ending on the corner diagonal.)
    StringSet<String<Dna>> sh = eh;
    resize(sh[7], 1);
    runBandedRowsCaseImpl("dna_banded_rows_single_residue_h", false, true, sh, ev, dna, ex, -5, 5);

    // |setH| != |setV| raises instead of asserting.
    StringSet<String<Dna>> shorter = dv;
    eraseBack(shorter);
    try
    {
        globalAlignmentScore(ex, dh, shorter, dna);
        std::printf("case length_mismatch error (no exception)\n");
    }
    catch (seqan_amd::Error const & e)
    {
        std::printf("case length_mismatch %s\n", e.status() == SQMI_ERR_LENGTH_MISMATCH ? "match 0" : "error");
    }
    return 0;
}
