// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference SeqAn 2 library (header-only, compiled from the
// reference source tree by oracle/Makefile into oracle/_ref/) on inputs written
// by the tests or bench.py, so that the GPU results can be compared with the
// reference's own CPU path and timed beside it.  This file is ours; it only
// calls the reference's public API:
//
//   globalAlignmentScore(ExecutionPolicy<Parallel,Vectorial>, ...)
//       include/seqan/align_parallel/parallel_align_interface.h:332-343
//   localAlignmentScore(ExecutionPolicy<Parallel,Vectorial>, ...)
//       include/seqan/align_parallel/parallel_align_interface.h:354-365
//   globalAlignment(ExecutionPolicy<Parallel,Vectorial>, gapsH, gapsV, score)
//       include/seqan/align_parallel/parallel_align_interface.h:288-299
//   globalAlignment(Gaps, Gaps, score)   (scalar, for very long pairs)
//       include/seqan/align/global_alignment_unbanded.h:281-306
//   localAlignment(ExecutionPolicy<Parallel,Vectorial>, gapsH, gapsV, score)
//       include/seqan/align_parallel/parallel_align_interface.h:312-321
//   localAlignment(Gaps, Gaps, score)    (scalar)
//       include/seqan/align/local_alignment_unbanded.h:262-273
//   globalAlignment(Gaps, Gaps, score, AlignConfig, lowerDiag, upperDiag)
//       include/seqan/align/global_alignment_banded.h:152-181
//   localAlignment(Gaps, Gaps, score, lowerDiag, upperDiag)
//       include/seqan/align/local_alignment_banded.h:242-259
//   computeAlignmentStats(stats, row0, row1, score)
//       include/seqan/align/evaluate_alignment.h:215-300
//
// Input file (little endian):
//   char[4] "SQRI", u32 op (0 score, 1 counts-simd, 2 counts-scalar, 3 table,
//   4 banded score, global or LOCAL: i64 lower_diag, i64 upper_diag follow
//   the count; 5 banded rows, pair by pair, the same two values follow),
//   u32 mode (1 TOP, 2 LEFT, 4 RIGHT, 8 BOTTOM, 16 LOCAL),
//   u32 alphabet (0 Dna, 1 AminoAcid), i32 match, i32 mismatch,
//   i32 gap_extend, i32 gap_open, u32 matrix (0 Simple, 1 Blosum62),
//   i32 threads, i32 repeats, i64 count,
//   i64 h_off[count+1], i64 v_off[count+1], u8 h_res[h_off[count]],
//   u8 v_res[v_off[count]]   (residue ordinals).
// Output file: i64 count, f64 best_seconds, i32 simd_lanes, i32 threads,
//   i32 scores[count], then for ops 1/2/5: i32 counts[count][A*A+2],
//   i64 op_off[count+1], u8 ops[op_off[count]]: the rows of each alignment
//   as one byte per column, 'M' (both rows hold a residue), 'V' (a gap in
//   the H row) or 'H' (a gap in the V row); with mode LOCAL the rows are
//   those of the clipped local window, followed by i64 begins[count][2]:
//   the source positions in H and in V where the window starts.
#include <seqan/align.h>
#include <seqan/align_parallel.h>
#include <seqan/score.h>
#include <seqan/sequence.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

using namespace seqan2;

namespace {

struct RefHeader
{
    uint32_t op, mode, alphabet;
    int32_t match, mismatch, gapExtend, gapOpen;
    uint32_t matrix;
    int32_t threads, repeats;
    int64_t count;
    int64_t lowerDiag = 0, upperDiag = 0;   // op 4 only
};

struct RefInput
{
    RefHeader hd;
    std::vector<int64_t> hOff, vOff;
    std::vector<uint8_t> hRes, vRes;
};

[[noreturn]] void die(char const * msg)
{
    std::fprintf(stderr, "ref_driver: %s\n", msg);
    std::exit(2);
}

template <typename T>
void readOrDie(std::FILE * f, T * dst, size_t n)
{
    if (n && std::fread(dst, sizeof(T), n, f) != n)
        die("short read");
}

RefInput readInput(char const * path)
{
    std::FILE * f = std::fopen(path, "rb");
    if (!f)
        die("cannot open input");
    char magic[4];
    readOrDie(f, magic, 4);
    if (std::memcmp(magic, "SQRI", 4) != 0)
        die("bad magic");
    RefInput in;
    RefHeader & h = in.hd;
    readOrDie(f, &h.op, 1);
    readOrDie(f, &h.mode, 1);
    readOrDie(f, &h.alphabet, 1);
    readOrDie(f, &h.match, 1);
    readOrDie(f, &h.mismatch, 1);
    readOrDie(f, &h.gapExtend, 1);
    readOrDie(f, &h.gapOpen, 1);
    readOrDie(f, &h.matrix, 1);
    readOrDie(f, &h.threads, 1);
    readOrDie(f, &h.repeats, 1);
    readOrDie(f, &h.count, 1);
    if (h.op == 4 || h.op == 5)
    {
        readOrDie(f, &h.lowerDiag, 1);
        readOrDie(f, &h.upperDiag, 1);
    }
    in.hOff.resize(h.count + 1);
    in.vOff.resize(h.count + 1);
    readOrDie(f, in.hOff.data(), in.hOff.size());
    readOrDie(f, in.vOff.data(), in.vOff.size());
    in.hRes.resize(in.hOff.back());
    in.vRes.resize(in.vOff.back());
    readOrDie(f, in.hRes.data(), in.hRes.size());
    readOrDie(f, in.vRes.data(), in.vRes.size());
    std::fclose(f);
    return in;
}

template <typename TAlphabet>
void buildSet(StringSet<String<TAlphabet>> & set, std::vector<uint8_t> const & res,
              std::vector<int64_t> const & off)
{
    size_t const count = off.size() - 1;
    resize(set, count);
    for (size_t k = 0; k < count; ++k)
    {
        String<TAlphabet> & s = set[k];
        resize(s, off[k + 1] - off[k]);
        for (int64_t p = off[k]; p < off[k + 1]; ++p)
        {
            TAlphabet c;
            c.value = res[p];   // residues arrive as ordinal values
            if (ordValue(c) != res[p])
                die("residue ordinal out of range for alphabet");
            s[p - off[k]] = c;
        }
    }
}

template <typename TExec, typename TSetH, typename TSetV, typename TScore>
String<int> runScore(TExec const & ex, TSetH const & setH, TSetV const & setV, TScore const & sc, uint32_t mode)
{
    if (mode & 16u)
        return localAlignmentScore(ex, setH, setV, sc);
    switch (mode & 15u)
    {
        case 0: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<false, false, false, false>());
        case 1: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<true, false, false, false>());
        case 2: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<false, true, false, false>());
        case 3: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<true, true, false, false>());
        case 4: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<false, false, true, false>());
        case 5: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<true, false, true, false>());
        case 6: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<false, true, true, false>());
        case 7: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<true, true, true, false>());
        case 8: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<false, false, false, true>());
        case 9: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<true, false, false, true>());
        case 10: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<false, true, false, true>());
        case 11: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<true, true, false, true>());
        case 12: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<false, false, true, true>());
        case 13: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<true, false, true, true>());
        case 14: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<false, true, true, true>());
        default: return globalAlignmentScore(ex, setH, setV, sc, AlignConfig<true, true, true, true>());
    }
}

// The AlignConfig of the mode bits (1 TOP, 2 LEFT, 4 RIGHT, 8 BOTTOM), passed
// to fn as a value.
template <typename TFn>
auto withConfig(uint32_t mode, TFn && fn)
{
    switch (mode & 15u)
    {
        case 0: return fn(AlignConfig<false, false, false, false>());
        case 1: return fn(AlignConfig<true, false, false, false>());
        case 2: return fn(AlignConfig<false, true, false, false>());
        case 3: return fn(AlignConfig<true, true, false, false>());
        case 4: return fn(AlignConfig<false, false, true, false>());
        case 5: return fn(AlignConfig<true, false, true, false>());
        case 6: return fn(AlignConfig<false, true, true, false>());
        case 7: return fn(AlignConfig<true, true, true, false>());
        case 8: return fn(AlignConfig<false, false, false, true>());
        case 9: return fn(AlignConfig<true, false, false, true>());
        case 10: return fn(AlignConfig<false, true, false, true>());
        case 11: return fn(AlignConfig<true, true, false, true>());
        case 12: return fn(AlignConfig<false, false, true, true>());
        case 13: return fn(AlignConfig<true, false, true, true>());
        case 14: return fn(AlignConfig<false, true, true, true>());
        default: return fn(AlignConfig<true, true, true, true>());
    }
}

// Substitution / gap counts of one alignment, read from its two rows.
template <typename TGaps, typename TScore>
void rowCounts(TGaps const & row0, TGaps const & row1, TScore const & sc, int alpha, int32_t * out)
{
    std::memset(out, 0, sizeof(int32_t) * (alpha * alpha + 2));
    AlignmentStats stats;
    computeAlignmentStats(stats, row0, row1, sc);
    out[alpha * alpha] = static_cast<int32_t>(stats.numGapOpens);
    out[alpha * alpha + 1] = static_cast<int32_t>(stats.numGapExtensions);
    auto it0 = begin(row0);
    auto it1 = begin(row1);
    for (; it0 != end(row0); ++it0, ++it1)
        if (!isGap(it0) && !isGap(it1))
            ++out[ordValue(convert<typename Value<typename Source<TGaps>::Type>::Type>(*it0)) * alpha +
                  ordValue(convert<typename Value<typename Source<TGaps>::Type>::Type>(*it1))];
}

template <typename TAlphabet, typename TScore>
int run(RefInput const & in, TScore const & sc, int alpha, std::FILE * out)
{
    RefHeader const & h = in.hd;
    StringSet<String<TAlphabet>> setH, setV;
    buildSet(setH, in.hRes, in.hOff);
    buildSet(setV, in.vRes, in.vOff);
    int threads = h.threads > 0 ? h.threads : static_cast<int>(std::thread::hardware_concurrency());
    ExecutionPolicy<Parallel, Vectorial> ex;
    setNumThreads(ex, threads);
    int32_t lanes = static_cast<int32_t>(LENGTH<typename SimdVector<int>::Type>::VALUE);

    std::vector<int32_t> scores(h.count, 0);
    std::vector<int32_t> counts;
    std::vector<int64_t> opOff;
    std::vector<uint8_t> ops;
    std::vector<int64_t> begins;
    double best = 1e300;
    if (h.op == 4)
    {
        // Banded globalAlignmentScore(setH, setV, score, AlignConfig, lowerDiag,
        // upperDiag) (global_alignment_banded.h:555-583), SIMD over the sets;
        // with mode LOCAL localAlignmentScore(setH, setV, score, lowerDiag,
        // upperDiag) (local_alignment_banded.h:344-383).
        bool const local = (h.mode & 16u) != 0;
        int const lo = static_cast<int>(h.lowerDiag), up = static_cast<int>(h.upperDiag);
        auto t0 = std::chrono::steady_clock::now();
        if (h.repeats < 0)
        {
            // repeats == -1: the scalar banded function, pair by pair
            // (global_alignment_banded.h:394-470, local_alignment_banded.h:317-342).
            for (int64_t k = 0; k < h.count; ++k)
                scores[k] = local ? localAlignmentScore(setH[k], setV[k], sc, lo, up)
                                  : withConfig(h.mode, [&](auto cfg) {
                                        return globalAlignmentScore(setH[k], setV[k], sc, cfg, lo, up);
                                    });
        }
        else
        {
            String<int> res = local ? localAlignmentScore(setH, setV, sc, lo, up)
                                    : withConfig(h.mode, [&](auto cfg) {
                                          return globalAlignmentScore(setH, setV, sc, cfg, lo, up);
                                      });
            for (int64_t k = 0; k < h.count; ++k)
                scores[k] = res[k];
        }
        auto t1 = std::chrono::steady_clock::now();
        best = std::chrono::duration<double>(t1 - t0).count();
    }
    else if (h.op == 0)
    {
        // One untimed warm-up call on the first pairs (at most 4,096): page
        // faults of the first allocations and the thread team's start-up stay
        // outside the timed calls.
        {
            int64_t const w = std::min<int64_t>(h.count, 4096);
            StringSet<String<TAlphabet>> wH, wV;
            for (int64_t k = 0; k < w; ++k)
            {
                appendValue(wH, setH[k]);
                appendValue(wV, setV[k]);
            }
            String<int> warm = runScore(ex, wH, wV, sc, h.mode);
            (void)warm;
        }
        int reps = h.repeats > 0 ? h.repeats : 1;
        for (int r = 0; r < reps; ++r)
        {
            auto t0 = std::chrono::steady_clock::now();
            String<int> res = runScore(ex, setH, setV, sc, h.mode);
            auto t1 = std::chrono::steady_clock::now();
            best = std::min(best, std::chrono::duration<double>(t1 - t0).count());
            for (int64_t k = 0; k < h.count; ++k)
                scores[k] = res[k];
        }
    }
    else if (h.op == 1 || h.op == 2 || h.op == 5)
    {
        bool const local = (h.mode & 16u) != 0;
        if (local && (h.mode & 15u))
            die("LOCAL takes no free end gaps");
        counts.assign(h.count * (alpha * alpha + 2), 0);
        using TGaps = Gaps<String<TAlphabet>>;
        StringSet<TGaps> gH, gV;
        resize(gH, h.count);
        resize(gV, h.count);
        for (int64_t k = 0; k < h.count; ++k)
        {
            gH[k] = TGaps(setH[k]);
            gV[k] = TGaps(setV[k]);
        }
        if (h.op == 1)
        {
            // Untimed warm-up on the first pairs (at most 256), as for op 0.
            int64_t const w = std::min<int64_t>(h.count, 256);
            StringSet<TGaps> wH, wV;
            resize(wH, w);
            resize(wV, w);
            for (int64_t k = 0; k < w; ++k)
            {
                wH[k] = TGaps(setH[k]);
                wV[k] = TGaps(setV[k]);
            }
            String<int> warm = local ? localAlignment(ex, wH, wV, sc)
                                     : withConfig(h.mode, [&](auto cfg) { return globalAlignment(ex, wH, wV, sc, cfg); });
            (void)warm;
        }
        auto t0 = std::chrono::steady_clock::now();
        if (h.op == 1)
        {
            String<int> res = local ? localAlignment(ex, gH, gV, sc)
                                    : withConfig(h.mode, [&](auto cfg) { return globalAlignment(ex, gH, gV, sc, cfg); });
            for (int64_t k = 0; k < h.count; ++k)
                scores[k] = res[k];
        }
        else if (h.op == 2)
        {
            for (int64_t k = 0; k < h.count; ++k)
                scores[k] = local ? localAlignment(gH[k], gV[k], sc)
                                  : withConfig(h.mode, [&](auto cfg) { return globalAlignment(gH[k], gV[k], sc, cfg); });
        }
        else
        {
            int const lo = static_cast<int>(h.lowerDiag), up = static_cast<int>(h.upperDiag);
            for (int64_t k = 0; k < h.count; ++k)
                scores[k] = local ? localAlignment(gH[k], gV[k], sc, lo, up)
                                  : withConfig(h.mode, [&](auto cfg) {
                                        return globalAlignment(gH[k], gV[k], sc, cfg, lo, up);
                                    });
        }
        auto t1 = std::chrono::steady_clock::now();
        best = std::chrono::duration<double>(t1 - t0).count();
        opOff.assign(1, 0);
        for (int64_t k = 0; k < h.count; ++k)
        {
            rowCounts(gH[k], gV[k], sc, alpha, counts.data() + k * (alpha * alpha + 2));
            auto it0 = begin(gH[k]);
            auto it1 = begin(gV[k]);
            for (; it0 != end(gH[k]); ++it0, ++it1)
                ops.push_back(isGap(it0) ? 'V' : isGap(it1) ? 'H' : 'M');
            opOff.push_back(static_cast<int64_t>(ops.size()));
            if (local)
            {
                begins.push_back(static_cast<int64_t>(beginPosition(gH[k])));
                begins.push_back(static_cast<int64_t>(beginPosition(gV[k])));
            }
        }
    }
    else
    {
        die("unknown op");
    }
    std::fwrite(&h.count, sizeof(int64_t), 1, out);
    std::fwrite(&best, sizeof(double), 1, out);
    std::fwrite(&lanes, sizeof(int32_t), 1, out);
    std::fwrite(&threads, sizeof(int32_t), 1, out);
    std::fwrite(scores.data(), sizeof(int32_t), scores.size(), out);
    if (!counts.empty())
    {
        std::fwrite(counts.data(), sizeof(int32_t), counts.size(), out);
        std::fwrite(opOff.data(), sizeof(int64_t), opOff.size(), out);
        std::fwrite(ops.data(), 1, ops.size(), out);
        std::fwrite(begins.data(), sizeof(int64_t), begins.size(), out);
    }
    return 0;
}

}  // namespace

int main(int argc, char ** argv)
{
    if (argc == 2 && std::strcmp(argv[1], "--blosum62") == 0)
    {
        // Print the reference's Blosum62 table in AminoAcid ordinal order.
        Blosum62 b(-1, -11);
        for (int i = 0; i < 27; ++i)
        {
            for (int j = 0; j < 27; ++j)
            {
                AminoAcid x, y;
                x.value = static_cast<uint8_t>(i);
                y.value = static_cast<uint8_t>(j);
                std::printf("%d%c", score(b, x, y), j == 26 ? '\n' : ' ');
            }
        }
        return 0;
    }
    if (argc == 2 && std::strcmp(argv[1], "--lanes") == 0)
    {
        std::printf("%d\n", static_cast<int>(LENGTH<typename SimdVector<int>::Type>::VALUE));
        return 0;
    }
    if (argc != 3)
        die("usage: ref_driver <input> <output> | --blosum62 | --lanes");
    RefInput in = readInput(argv[1]);
    std::FILE * out = std::fopen(argv[2], "wb");
    if (!out)
        die("cannot open output");
    int rc;
    if (in.hd.alphabet == 0)
    {
        if (in.hd.matrix != 0)
            die("Dna uses the Simple score");
        Score<int, Simple> sc(in.hd.match, in.hd.mismatch, in.hd.gapExtend, in.hd.gapOpen);
        rc = run<Dna>(in, sc, 4, out);
    }
    else if (in.hd.matrix == 1)
    {
        Blosum62 sc(in.hd.gapExtend, in.hd.gapOpen);
        rc = run<AminoAcid>(in, sc, 27, out);
    }
    else
    {
        Score<int, Simple> sc(in.hd.match, in.hd.mismatch, in.hd.gapExtend, in.hd.gapOpen);
        rc = run<AminoAcid>(in, sc, 27, out);
    }
    std::fclose(out);
    return rc;
}
