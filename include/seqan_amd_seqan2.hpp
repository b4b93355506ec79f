/*
 * seqan_amd_seqan2.hpp -- drop-in overloads for SeqAn 2 callers.
 *
 * Including this header after (or instead of) <seqan/align_parallel.h> routes
 *
 *   globalAlignmentScore(ExecutionPolicy<Parallel, Vectorial>, setH, setV, score [, AlignConfig<T,L,R,B>])
 *   globalAlignmentScore(ExecutionPolicy<Parallel, Vectorial>, setH, setV, score [, AlignConfig],
 *                        lowerDiag, upperDiag)   (banded; see sqmi_align_score_banded for the bands taken)
 *   localAlignmentScore(ExecutionPolicy<Parallel, Vectorial>, setH, setV, score, lowerDiag, upperDiag)
 *   localAlignmentScore(ExecutionPolicy<Parallel, Vectorial>, setH, setV, score)
 *   globalAlignment(ExecutionPolicy<Parallel, Vectorial>, gapsSetH, gapsSetV, score [, AlignConfig<T,L,R,B>])
 *   localAlignment(ExecutionPolicy<Parallel, Vectorial>, gapsSetH, gapsSetV, score)
 *   globalAlignmentScore(ExecutionPolicy<WavefrontAlignment<S>, V>, setH, setV, score [, AlignConfig])
 *   localAlignmentScore(ExecutionPolicy<WavefrontAlignment<S>, V>, setH, setV, score)
 *
 * to the MI355X library (libseqan_amd.so, C ABI in seqan_amd.h) whenever
 *   - setH and setV are StringSets over one of Dna, Dna5, Rna, Rna5, AminoAcid,
 *     both over the same alphabet (for globalAlignment: StringSets of Gaps
 *     over such sequences), and
 *   - score is Score<int, Simple> or Score<int, ScoreMatrix<A, Spec>> with A
 *     the sets' alphabet (Blosum62 and the other built-in matrices included).
 * Every other combination (another policy, another alphabet, a non-int score,
 * an algorithm tag) resolves to the reference templates as before, and a
 * banded call with a band the library does not take (sqmi_align_score_banded
 * lists the ones it takes) runs the reference's executor from inside the
 * overload.  The results are the reference's, value for value.
 *
 * The reference entry points these overloads take precedence over:
 *   globalAlignmentScore(ExecutionPolicy<TPar,TVec> const &, TArgs && ...)
 *       include/seqan/align_parallel/parallel_align_interface.h:332-343
 *   localAlignmentScore(ExecutionPolicy<TPar,TVec> const &, TArgs && ...)
 *       include/seqan/align_parallel/parallel_align_interface.h:354-365
 *   globalAlignment(ExecutionPolicy<TPar,TVec> const &, TArgs && ...)
 *       include/seqan/align_parallel/parallel_align_interface.h:288-299
 *   localAlignment(ExecutionPolicy<TPar,TVec> const &, TArgs && ...)
 *       include/seqan/align_parallel/parallel_align_interface.h:312-321
 * globalAlignment and localAlignment write the rows of the reference's
 * traceback into the Gaps objects (the way the reference does: clear gaps and
 * clipping, then insert the gap runs through Gaps iterators; localAlignment
 * also clips both rows to the local window) and return the scores.
 * The wavefront policy (intra-pair parallelism for long sequences,
 * parallel_align_interface.h:367-420) returns std::vector<int> like the
 * reference; its block size and parallel-alignment settings are not used:
 * long pairs take the library's tiled long-pair kernels by themselves.
 * Both take forwarding references; so do the overloads below, which makes the
 * conversion sequences of the two candidates identical for const, non-const
 * and rvalue arguments alike.  Partial ordering then prefers these overloads
 * because their first parameter, ExecutionPolicy<Parallel, Vectorial>, is
 * more specialised than ExecutionPolicy<TPar, TVec>.
 *
 * Differences a caller can observe:
 *   - errors throw seqan_amd::Error (the reference asserts in debug builds
 *     and is undefined otherwise, e.g. on |setH| != |setV|,
 *     parallel_align_interface.h:111);
 *   - numThreads(ex) is not used; each calling thread owns one library
 *     context (device seqan_amd::device(), stream and scratch), created on its
 *     first call and destroyed when the thread ends.
 *
 * Needs C++17.  Link with -lseqan_amd.  Define SEQAN_AMD_NO_SEQAN2_OVERLOADS to keep the
 * helpers of namespace seqan_amd without the overloads.  Define
 * SEQAN_AMD_SEQAN2_PROBE to declare the overloads deleted instead: a call
 * that resolves to one of them is then ill-formed, which lets a test detect
 * at compile time which candidate overload resolution picks.
 */
#ifndef SEQAN_AMD_SEQAN2_HPP
#define SEQAN_AMD_SEQAN2_HPP

#include <seqan/align.h>
#include <seqan/align_parallel.h>
#include <seqan/score.h>
#include <seqan/sequence.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "seqan_amd.h"

namespace seqan_amd {

/* Raised for every status other than SQMI_OK. */
class Error : public std::runtime_error
{
public:
    Error(sqmi_status status, std::string const & what) : std::runtime_error(what), status_(status) {}
    sqmi_status status() const noexcept { return status_; }

private:
    sqmi_status status_;
};

/* HIP device used by contexts created after the call (default 0). */
inline int & device_slot() { static int dev = 0; return dev; }
inline int device() { return device_slot(); }
inline void set_device(int dev) { device_slot() = dev; }

namespace detail {

inline void check(sqmi_status st, sqmi_context const * ctx)
{
    if (st == SQMI_OK)
        return;
    std::string msg = std::string("seqan_amd: ") + sqmi_status_string(st);
    if (ctx != nullptr && sqmi_last_error(ctx)[0] != '\0')
        msg += std::string(": ") + sqmi_last_error(ctx);
    throw Error(st, msg);
}

struct ThreadContext
{
    sqmi_context * ctx = nullptr;
    ~ThreadContext() { if (ctx != nullptr) sqmi_destroy(ctx); }
};

inline sqmi_context * thread_context()
{
    thread_local ThreadContext holder;
    if (holder.ctx == nullptr)
        check(sqmi_create(device(), &holder.ctx), nullptr);
    return holder.ctx;
}

/* Alphabets the library packs: one byte per residue, value == ordinal. */
template <typename TAlphabet> struct IsGpuAlphabet : std::false_type {};
template <> struct IsGpuAlphabet<seqan2::Dna> : std::true_type {};
template <> struct IsGpuAlphabet<seqan2::Dna5> : std::true_type {};
template <> struct IsGpuAlphabet<seqan2::Rna> : std::true_type {};
template <> struct IsGpuAlphabet<seqan2::Rna5> : std::true_type {};
template <> struct IsGpuAlphabet<seqan2::AminoAcid> : std::true_type {};

/* Alphabet of a StringSet<String<A, ...>, ...>; void for anything else. */
template <typename T> struct SetAlphabet { using Type = void; };
template <typename TString, typename TSpec>
struct SetAlphabet<seqan2::StringSet<TString, TSpec>>
{
    using Type = typename seqan2::Value<TString>::Type;
};

/* Score schemes the library evaluates exactly like the reference. */
template <typename TScore, typename TAlphabet> struct IsGpuScore : std::false_type {};
template <typename TAlphabet>
struct IsGpuScore<seqan2::Score<int, seqan2::Simple>, TAlphabet> : std::true_type {};
template <typename TAlphabet, typename TSpec>
struct IsGpuScore<seqan2::Score<int, seqan2::ScoreMatrix<TAlphabet, TSpec>>, TAlphabet> : std::true_type {};

template <typename T> struct IsAlignConfig : std::false_type {};
template <bool TOP, bool LEFT, bool RIGHT, bool BOTTOM, typename TSpec>
struct IsAlignConfig<seqan2::AlignConfig<TOP, LEFT, RIGHT, BOTTOM, TSpec>> : std::true_type
{
    static constexpr uint32_t mode = (TOP ? SQMI_FREE_TOP : 0u) | (LEFT ? SQMI_FREE_LEFT : 0u) |
                                     (RIGHT ? SQMI_FREE_RIGHT : 0u) | (BOTTOM ? SQMI_FREE_BOTTOM : 0u);
};

/* Alphabet of a StringSet<Gaps<TSequence, ...>, ...>; void for anything else. */
template <typename T> struct GapsSetAlphabet { using Type = void; };
template <typename TSequence, typename TGapsSpec, typename TSetSpec>
struct GapsSetAlphabet<seqan2::StringSet<seqan2::Gaps<TSequence, TGapsSpec>, TSetSpec>>
{
    using Type = typename seqan2::Value<TSequence>::Type;
};

template <typename TGapsSetH, typename TGapsSetV, typename TScore>
struct IsGpuRowsCall
{
    using AlphaH = typename GapsSetAlphabet<std::decay_t<TGapsSetH>>::Type;
    using AlphaV = typename GapsSetAlphabet<std::decay_t<TGapsSetV>>::Type;
    static constexpr bool value = std::is_same<AlphaH, AlphaV>::value && IsGpuAlphabet<AlphaH>::value &&
                                  IsGpuScore<std::decay_t<TScore>, AlphaH>::value;
};

template <typename TSetH, typename TSetV, typename TScore>
struct IsGpuCall
{
    using AlphaH = typename SetAlphabet<std::decay_t<TSetH>>::Type;
    using AlphaV = typename SetAlphabet<std::decay_t<TSetV>>::Type;
    static constexpr bool value = std::is_same<AlphaH, AlphaV>::value && IsGpuAlphabet<AlphaH>::value &&
                                  IsGpuScore<std::decay_t<TScore>, AlphaH>::value;
};

/* Concatenated residue ordinals and count+1 offsets of the sequences
 * get(0) ... get(count - 1) (get may return a reference or, for sets that
 * hand out infixes, a value). */
template <typename TAlphabet, typename TGet>
void packWith(int64_t count, TGet && get, std::vector<uint8_t> & residues, std::vector<int64_t> & offsets)
{
    offsets.resize(count + 1);
    offsets[0] = 0;
    for (int64_t i = 0; i < count; ++i)
        offsets[i + 1] = offsets[i] + static_cast<int64_t>(seqan2::length(get(i)));
    residues.resize(static_cast<size_t>(offsets[count]));
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < count; ++i)
    {
        auto const & seq = get(i);
        uint8_t * out = residues.data() + offsets[i];
        int64_t const len = offsets[i + 1] - offsets[i];
        using TIter = typename seqan2::Iterator<std::decay_t<decltype(seq)> const, seqan2::Standard>::Type;
        if constexpr (std::is_pointer<TIter>::value && sizeof(TAlphabet) == 1)
        {
            if (len > 0)
                std::memcpy(out, &*seqan2::begin(seq, seqan2::Standard()), static_cast<size_t>(len));
        }
        else
        {
            for (int64_t k = 0; k < len; ++k)
                out[k] = static_cast<uint8_t>(seqan2::ordValue(static_cast<TAlphabet>(seq[k])));
        }
    }
}

/* A StringSet's sequences. */
template <typename TSet>
void pack(TSet const & set, std::vector<uint8_t> & residues, std::vector<int64_t> & offsets)
{
    packWith<typename SetAlphabet<TSet>::Type>(
        static_cast<int64_t>(seqan2::length(set)), [&](int64_t i) -> decltype(auto) { return set[i]; }, residues,
        offsets);
}

/* The source sequences of a StringSet of Gaps. */
template <typename TGapsSet>
void packSources(TGapsSet & set, std::vector<uint8_t> & residues, std::vector<int64_t> & offsets)
{
    packWith<typename GapsSetAlphabet<TGapsSet>::Type>(
        static_cast<int64_t>(seqan2::length(set)), [&](int64_t i) -> decltype(auto) { return seqan2::source(set[i]); },
        residues, offsets);
}

/* Rows of one alignment into its two Gaps objects: clear gaps and clipping,
 * then walk both rows with Gaps iterators, inserting each gap run ('V': a
 * run in the H row, 'H': in the V row).  A local window (begins given)
 * starts at source positions begins[0] / begins[1] and the clipping ends
 * after its last column; an empty window clips both rows to nothing
 * (the Gaps state dp_traceback_adaptor.h:60-117 leaves). */
template <typename TGapsH, typename TGapsV>
void write_rows(TGapsH & gapsH, TGapsV & gapsV, uint8_t const * row, int64_t len,
                int64_t const * begins = nullptr)
{
    seqan2::clearGaps(gapsH);
    seqan2::clearClipping(gapsH);
    seqan2::clearGaps(gapsV);
    seqan2::clearClipping(gapsV);
    if (begins && len == 0)
    {
        seqan2::setClippedBeginPosition(gapsH, 0);
        seqan2::setClippedEndPosition(gapsH, 0);
        seqan2::setClippedBeginPosition(gapsV, 0);
        seqan2::setClippedEndPosition(gapsV, 0);
        return;
    }
    if (begins)
    {
        seqan2::setBeginPosition(gapsH, begins[0]);
        seqan2::setBeginPosition(gapsV, begins[1]);
    }
    auto itH = seqan2::begin(gapsH, seqan2::Standard());
    auto itV = seqan2::begin(gapsV, seqan2::Standard());
    for (int64_t k = 0; k < len;)
    {
        uint8_t const op = row[k];
        int64_t run = 1;
        while (k + run < len && row[k + run] == op)
            ++run;
        if (op == 'V')
            seqan2::insertGaps(itH, run);
        else if (op == 'H')
            seqan2::insertGaps(itV, run);
        seqan2::goFurther(itH, run);
        seqan2::goFurther(itV, run);
        k += run;
    }
    if (begins)
    {
        seqan2::setClippedEndPosition(gapsH, seqan2::position(itH) + seqan2::clippedBeginPosition(gapsH));
        seqan2::setClippedEndPosition(gapsV, seqan2::position(itV) + seqan2::clippedBeginPosition(gapsV));
    }
}

/* Substitution table over the alphabet, row = H residue, column = V residue;
 * entries are the reference's own score(scheme, h, v) values. */
template <typename TAlphabet, typename TScore>
sqmi_scoring make_scoring(TScore const & scheme, std::vector<int32_t> & table)
{
    int const alpha = static_cast<int>(seqan2::ValueSize<TAlphabet>::VALUE);
    table.resize(static_cast<size_t>(alpha) * alpha);
    for (int h = 0; h < alpha; ++h)
        for (int v = 0; v < alpha; ++v)
            table[h * alpha + v] = seqan2::score(scheme, TAlphabet(h), TAlphabet(v));
    sqmi_scoring sc;
    sc.alphabet_size = alpha;
    sc.gap_open = seqan2::scoreGapOpen(scheme);
    sc.gap_extend = seqan2::scoreGapExtend(scheme);
    sc.table = table.data();
    return sc;
}

template <typename TSetH, typename TSetV, typename TScore>
seqan2::String<int> align_scores(TSetH const & setH, TSetV const & setV, TScore const & scheme, uint32_t mode)
{
    using TAlphabet = typename SetAlphabet<TSetH>::Type;
    if (seqan2::length(setH) != seqan2::length(setV))
        throw Error(SQMI_ERR_LENGTH_MISMATCH, "seqan_amd: setH and setV differ in length");
    seqan2::String<int> result;
    if (seqan2::length(setH) == 0)
        return result;
    std::vector<uint8_t> resH, resV;
    std::vector<int64_t> offH, offV;
    pack(setH, resH, offH);
    pack(setV, resV, offV);
    std::vector<int32_t> table;
    sqmi_scoring const sc = make_scoring<TAlphabet>(scheme, table);
    sqmi_sequences const h{resH.data(), offH.data(), static_cast<int64_t>(offH.size()) - 1};
    sqmi_sequences const v{resV.data(), offV.data(), static_cast<int64_t>(offV.size()) - 1};
    seqan2::resize(result, seqan2::length(setH));
    sqmi_context * ctx = thread_context();
    check(sqmi_align_score(ctx, &h, &v, &sc, mode, &result[0], 0u), ctx);
    return result;
}

// Banded calls: the library takes the bands sqmi_align_score_banded
// documents and answers SQMI_ERR_UNSUPPORTED, before touching any output,
// for every other one.  Those calls go to `fallback`, which runs the
// reference's own executor (parallel_align_interface.h:100-145) on the same
// arguments, so a program that works with SeqAn keeps working, value for
// value.
template <typename TSetH, typename TSetV, typename TScore, typename TFallback>
seqan2::String<int> align_scores_banded(TSetH const & setH, TSetV const & setV, TScore const & scheme, uint32_t mode,
                                        int64_t lowerDiag, int64_t upperDiag, TFallback && fallback)
{
    using TAlphabet = typename SetAlphabet<TSetH>::Type;
    if (seqan2::length(setH) != seqan2::length(setV))
        throw Error(SQMI_ERR_LENGTH_MISMATCH, "seqan_amd: setH and setV differ in length");
    seqan2::String<int> result;
    if (seqan2::length(setH) == 0)
        return result;
    std::vector<uint8_t> resH, resV;
    std::vector<int64_t> offH, offV;
    pack(setH, resH, offH);
    pack(setV, resV, offV);
    std::vector<int32_t> table;
    sqmi_scoring const sc = make_scoring<TAlphabet>(scheme, table);
    sqmi_sequences const h{resH.data(), offH.data(), static_cast<int64_t>(offH.size()) - 1};
    sqmi_sequences const v{resV.data(), offV.data(), static_cast<int64_t>(offV.size()) - 1};
    seqan2::resize(result, seqan2::length(setH));
    sqmi_context * ctx = thread_context();
    sqmi_status const st = sqmi_align_score_banded(ctx, &h, &v, &sc, mode, lowerDiag, upperDiag, &result[0], 0u);
    if (st == SQMI_ERR_UNSUPPORTED)
        return fallback();
    check(st, ctx);
    return result;
}

template <typename TSetH, typename TSetV, typename TScore>
std::vector<int> align_scores_vector(TSetH const & setH, TSetV const & setV, TScore const & scheme, uint32_t mode)
{
    seqan2::String<int> s = align_scores(setH, setV, scheme, mode);
    return std::vector<int>(seqan2::begin(s, seqan2::Standard()), seqan2::end(s, seqan2::Standard()));
}

template <typename TGapsSetH, typename TGapsSetV, typename TScore>
seqan2::String<int> align_rows(TGapsSetH & gapsH, TGapsSetV & gapsV, TScore const & scheme, uint32_t mode)
{
    using TAlphabet = typename GapsSetAlphabet<TGapsSetH>::Type;
    if (seqan2::length(gapsH) != seqan2::length(gapsV))
        throw Error(SQMI_ERR_LENGTH_MISMATCH, "seqan_amd: gapsSetH and gapsSetV differ in length");
    seqan2::String<int> result;
    int64_t const count = static_cast<int64_t>(seqan2::length(gapsH));
    if (count == 0)
        return result;
    std::vector<uint8_t> resH, resV;
    std::vector<int64_t> offH, offV;
    packSources(gapsH, resH, offH);
    packSources(gapsV, resV, offV);
    std::vector<int32_t> table;
    sqmi_scoring const sc = make_scoring<TAlphabet>(scheme, table);
    sqmi_sequences const h{resH.data(), offH.data(), count};
    sqmi_sequences const v{resV.data(), offV.data(), count};
    std::vector<int64_t> opOff(count + 1, 0), opLen(count, 0);
    for (int64_t i = 0; i < count; ++i)
        opOff[i + 1] = opOff[i] + (offH[i + 1] - offH[i]) + (offV[i + 1] - offV[i]);
    std::vector<uint8_t> ops(static_cast<size_t>(opOff[count]) + 1);
    std::vector<int32_t> counts(static_cast<size_t>(count) * (sc.alphabet_size * sc.alphabet_size + 2));
    seqan2::resize(result, count);
    sqmi_context * ctx = thread_context();
    check(sqmi_align_rows(ctx, &h, &v, &sc, mode, &result[0], counts.data(), ops.data(), opOff.data(),
                          opLen.data(), 0u),
          ctx);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t i = 0; i < count; ++i)
        write_rows(gapsH[i], gapsV[i], ops.data() + opOff[i + 1] - opLen[i], opLen[i]);
    return result;
}

template <typename TGapsSetH, typename TGapsSetV, typename TScore>
seqan2::String<int> align_local_rows(TGapsSetH & gapsH, TGapsSetV & gapsV, TScore const & scheme)
{
    using TAlphabet = typename GapsSetAlphabet<TGapsSetH>::Type;
    if (seqan2::length(gapsH) != seqan2::length(gapsV))
        throw Error(SQMI_ERR_LENGTH_MISMATCH, "seqan_amd: gapsSetH and gapsSetV differ in length");
    seqan2::String<int> result;
    int64_t const count = static_cast<int64_t>(seqan2::length(gapsH));
    if (count == 0)
        return result;
    std::vector<uint8_t> resH, resV;
    std::vector<int64_t> offH, offV;
    packSources(gapsH, resH, offH);
    packSources(gapsV, resV, offV);
    std::vector<int32_t> table;
    sqmi_scoring const sc = make_scoring<TAlphabet>(scheme, table);
    sqmi_sequences const h{resH.data(), offH.data(), count};
    sqmi_sequences const v{resV.data(), offV.data(), count};
    std::vector<int64_t> opOff(count + 1, 0), opLen(count, 0), begins(2 * count, 0);
    for (int64_t i = 0; i < count; ++i)
        opOff[i + 1] = opOff[i] + (offH[i + 1] - offH[i]) + (offV[i + 1] - offV[i]);
    std::vector<uint8_t> ops(static_cast<size_t>(opOff[count]) + 1);
    std::vector<int32_t> counts(static_cast<size_t>(count) * (sc.alphabet_size * sc.alphabet_size + 2));
    seqan2::resize(result, count);
    sqmi_context * ctx = thread_context();
    check(sqmi_align_local_rows(ctx, &h, &v, &sc, &result[0], counts.data(), ops.data(), opOff.data(),
                                opLen.data(), begins.data(), 0u),
          ctx);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t i = 0; i < count; ++i)
        write_rows(gapsH[i], gapsV[i], ops.data() + opOff[i + 1] - opLen[i], opLen[i], begins.data() + 2 * i);
    return result;
}

template <typename TGapsSetH, typename TGapsSetV, typename TScore, typename TFallback>
seqan2::String<int> align_rows_banded(TGapsSetH & gapsH, TGapsSetV & gapsV, TScore const & scheme, uint32_t mode,
                                      int64_t lower, int64_t upper, TFallback && fallback)
{
    using TAlphabet = typename GapsSetAlphabet<TGapsSetH>::Type;
    if (seqan2::length(gapsH) != seqan2::length(gapsV))
        throw Error(SQMI_ERR_LENGTH_MISMATCH, "seqan_amd: gapsSetH and gapsSetV differ in length");
    seqan2::String<int> result;
    int64_t const count = static_cast<int64_t>(seqan2::length(gapsH));
    if (count == 0)
        return result;
    bool const local = (mode & SQMI_LOCAL) != 0;
    std::vector<uint8_t> resH, resV;
    std::vector<int64_t> offH, offV;
    packSources(gapsH, resH, offH);
    packSources(gapsV, resV, offV);
    std::vector<int32_t> table;
    sqmi_scoring const sc = make_scoring<TAlphabet>(scheme, table);
    sqmi_sequences const h{resH.data(), offH.data(), count};
    sqmi_sequences const v{resV.data(), offV.data(), count};
    std::vector<int64_t> opOff(count + 1, 0), opLen(count, 0), begins(2 * count, 0);
    for (int64_t i = 0; i < count; ++i)
        opOff[i + 1] = opOff[i] + (offH[i + 1] - offH[i]) + (offV[i + 1] - offV[i]);
    std::vector<uint8_t> ops(static_cast<size_t>(opOff[count]) + 1);
    std::vector<int32_t> counts(static_cast<size_t>(count) * (sc.alphabet_size * sc.alphabet_size + 2));
    seqan2::resize(result, count);
    sqmi_context * ctx = thread_context();
    sqmi_status const st = sqmi_align_rows_banded(ctx, &h, &v, &sc, mode, lower, upper, &result[0], counts.data(),
                                                  ops.data(), opOff.data(), opLen.data(),
                                                  local ? begins.data() : nullptr, 0u);
    if (st == SQMI_ERR_UNSUPPORTED)
        return fallback();   // the Gaps objects are untouched until here
    check(st, ctx);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t i = 0; i < count; ++i)
        write_rows(gapsH[i], gapsV[i], ops.data() + opOff[i + 1] - opLen[i], opLen[i],
                   local ? begins.data() + 2 * i : nullptr);
    return result;
}

}  // namespace detail
}  // namespace seqan_amd

#ifndef SEQAN_AMD_NO_SEQAN2_OVERLOADS
#ifdef SEQAN_AMD_SEQAN2_PROBE
#define SEQAN_AMD_OVERLOAD_BODY(...) = delete;
#else
#define SEQAN_AMD_OVERLOAD_BODY(...) { return __VA_ARGS__; }
#endif
namespace seqan2 {

template <typename TSetH, typename TSetV, typename TScore,
          std::enable_if_t<seqan_amd::detail::IsGpuCall<TSetH, TSetV, TScore>::value, int> = 0>
inline String<int>
globalAlignmentScore(ExecutionPolicy<Parallel, Vectorial> const & /*execPolicy*/,
                     TSetH && setH, TSetV && setV, TScore && scoringScheme)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_scores(setH, setV, scoringScheme, 0u))

template <typename TSetH, typename TSetV, typename TScore, typename TConfig,
          std::enable_if_t<seqan_amd::detail::IsGpuCall<TSetH, TSetV, TScore>::value &&
                           seqan_amd::detail::IsAlignConfig<std::decay_t<TConfig>>::value, int> = 0>
inline String<int>
globalAlignmentScore(ExecutionPolicy<Parallel, Vectorial> const & /*execPolicy*/,
                     TSetH && setH, TSetV && setV, TScore && scoringScheme, TConfig && /*alignConfig*/)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_scores(setH, setV, scoringScheme,
                                                        seqan_amd::detail::IsAlignConfig<std::decay_t<TConfig>>::mode))

template <typename TSetH, typename TSetV, typename TScore, typename TLower, typename TUpper,
          std::enable_if_t<seqan_amd::detail::IsGpuCall<TSetH, TSetV, TScore>::value &&
                           std::is_integral<std::decay_t<TLower>>::value &&
                           std::is_integral<std::decay_t<TUpper>>::value, int> = 0>
inline String<int>
localAlignmentScore(ExecutionPolicy<Parallel, Vectorial> const & execPolicy,
                    TSetH && setH, TSetV && setV, TScore && scoringScheme, TLower && lowerDiag, TUpper && upperDiag)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_scores_banded(
    setH, setV, scoringScheme, static_cast<uint32_t>(SQMI_LOCAL), static_cast<int64_t>(lowerDiag),
    static_cast<int64_t>(upperDiag), [&]() {
        return seqan2::localAlignmentScore<Parallel, Vectorial>(execPolicy, setH, setV, scoringScheme, lowerDiag,
                                                                upperDiag);
    }))

// Banded: the executor forwards (setH, setV, score [, config], lowerDiag,
// upperDiag) to global_alignment_banded.h:555-583.  Calls with a pair whose
// band sqmi_align_score_banded does not take run the reference's executor
// instead: explicit template arguments <Parallel, Vectorial> make these
// overloads non-viable (their first template parameters are the sets), so
// the call resolves to the reference's template.
template <typename TSetH, typename TSetV, typename TScore, typename TLower, typename TUpper,
          std::enable_if_t<seqan_amd::detail::IsGpuCall<TSetH, TSetV, TScore>::value &&
                           std::is_integral<std::decay_t<TLower>>::value &&
                           std::is_integral<std::decay_t<TUpper>>::value, int> = 0>
inline String<int>
globalAlignmentScore(ExecutionPolicy<Parallel, Vectorial> const & execPolicy,
                     TSetH && setH, TSetV && setV, TScore && scoringScheme, TLower && lowerDiag, TUpper && upperDiag)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_scores_banded(
    setH, setV, scoringScheme, 0u, static_cast<int64_t>(lowerDiag), static_cast<int64_t>(upperDiag), [&]() {
        return seqan2::globalAlignmentScore<Parallel, Vectorial>(execPolicy, setH, setV, scoringScheme, lowerDiag,
                                                                 upperDiag);
    }))

template <typename TSetH, typename TSetV, typename TScore, typename TConfig, typename TLower, typename TUpper,
          std::enable_if_t<seqan_amd::detail::IsGpuCall<TSetH, TSetV, TScore>::value &&
                           seqan_amd::detail::IsAlignConfig<std::decay_t<TConfig>>::value &&
                           std::is_integral<std::decay_t<TLower>>::value &&
                           std::is_integral<std::decay_t<TUpper>>::value, int> = 0>
inline String<int>
globalAlignmentScore(ExecutionPolicy<Parallel, Vectorial> const & execPolicy,
                     TSetH && setH, TSetV && setV, TScore && scoringScheme, TConfig && alignConfig,
                     TLower && lowerDiag, TUpper && upperDiag)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_scores_banded(
    setH, setV, scoringScheme, seqan_amd::detail::IsAlignConfig<std::decay_t<TConfig>>::mode,
    static_cast<int64_t>(lowerDiag), static_cast<int64_t>(upperDiag), [&]() {
        return seqan2::globalAlignmentScore<Parallel, Vectorial>(execPolicy, setH, setV, scoringScheme, alignConfig,
                                                                 lowerDiag, upperDiag);
    }))

template <typename TSetH, typename TSetV, typename TScore,
          std::enable_if_t<seqan_amd::detail::IsGpuCall<TSetH, TSetV, TScore>::value, int> = 0>
inline String<int>
localAlignmentScore(ExecutionPolicy<Parallel, Vectorial> const & /*execPolicy*/,
                    TSetH && setH, TSetV && setV, TScore && scoringScheme)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_scores(setH, setV, scoringScheme, static_cast<uint32_t>(SQMI_LOCAL)))

// The wavefront policy.  The reference's global overloads take the sets and
// the score by const reference; these take StringSets, which partial
// ordering prefers.
template <typename TWaveSpec, typename TVec, typename TStrH, typename TSpecH, typename TStrV, typename TSpecV,
          typename TScore,
          std::enable_if_t<seqan_amd::detail::IsGpuCall<StringSet<TStrH, TSpecH>, StringSet<TStrV, TSpecV>,
                                                        TScore>::value, int> = 0>
inline std::vector<int>
globalAlignmentScore(ExecutionPolicy<WavefrontAlignment<TWaveSpec>, TVec> const & /*execPolicy*/,
                     StringSet<TStrH, TSpecH> const & setH, StringSet<TStrV, TSpecV> const & setV,
                     TScore const & scoringScheme)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_scores_vector(setH, setV, scoringScheme, 0u))

template <typename TWaveSpec, typename TVec, typename TStrH, typename TSpecH, typename TStrV, typename TSpecV,
          typename TScore, typename TConfig,
          std::enable_if_t<seqan_amd::detail::IsGpuCall<StringSet<TStrH, TSpecH>, StringSet<TStrV, TSpecV>,
                                                        TScore>::value &&
                           seqan_amd::detail::IsAlignConfig<TConfig>::value, int> = 0>
inline std::vector<int>
globalAlignmentScore(ExecutionPolicy<WavefrontAlignment<TWaveSpec>, TVec> const & /*execPolicy*/,
                     StringSet<TStrH, TSpecH> const & setH, StringSet<TStrV, TSpecV> const & setV,
                     TScore const & scoringScheme, TConfig const & /*alignConfig*/)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_scores_vector(setH, setV, scoringScheme,
                                                               seqan_amd::detail::IsAlignConfig<TConfig>::mode))

// The reference's local wavefront overload takes forwarding references, so
// this one does too (and wins by its first parameter being non-variadic).
template <typename TWaveSpec, typename TVec, typename TSetH, typename TSetV, typename TScore,
          std::enable_if_t<seqan_amd::detail::IsGpuCall<TSetH, TSetV, TScore>::value, int> = 0>
inline std::vector<int>
localAlignmentScore(ExecutionPolicy<WavefrontAlignment<TWaveSpec>, TVec> const & /*execPolicy*/,
                    TSetH && setH, TSetV && setV, TScore && scoringScheme)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_scores_vector(setH, setV, scoringScheme,
                                                               static_cast<uint32_t>(SQMI_LOCAL)))

template <typename TGapsSetH, typename TGapsSetV, typename TScore,
          std::enable_if_t<seqan_amd::detail::IsGpuRowsCall<TGapsSetH, TGapsSetV, TScore>::value, int> = 0>
inline String<int>
globalAlignment(ExecutionPolicy<Parallel, Vectorial> const & /*execPolicy*/,
                TGapsSetH && gapsSetH, TGapsSetV && gapsSetV, TScore && scoringScheme)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_rows(gapsSetH, gapsSetV, scoringScheme, 0u))

template <typename TGapsSetH, typename TGapsSetV, typename TScore, typename TConfig,
          std::enable_if_t<seqan_amd::detail::IsGpuRowsCall<TGapsSetH, TGapsSetV, TScore>::value &&
                           seqan_amd::detail::IsAlignConfig<std::decay_t<TConfig>>::value, int> = 0>
inline String<int>
globalAlignment(ExecutionPolicy<Parallel, Vectorial> const & /*execPolicy*/,
                TGapsSetH && gapsSetH, TGapsSetV && gapsSetV, TScore && scoringScheme, TConfig && /*alignConfig*/)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_rows(gapsSetH, gapsSetV, scoringScheme,
                                                      seqan_amd::detail::IsAlignConfig<std::decay_t<TConfig>>::mode))

template <typename TGapsSetH, typename TGapsSetV, typename TScore,
          std::enable_if_t<seqan_amd::detail::IsGpuRowsCall<TGapsSetH, TGapsSetV, TScore>::value, int> = 0>
inline String<int>
localAlignment(ExecutionPolicy<Parallel, Vectorial> const & /*execPolicy*/,
               TGapsSetH && gapsSetH, TGapsSetV && gapsSetV, TScore && scoringScheme)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_local_rows(gapsSetH, gapsSetV, scoringScheme))

// Banded rows: the executor forwards (gapsSetH, gapsSetV, score [, config],
// lowerDiag, upperDiag) to global_alignment_banded.h:655-760 and
// local_alignment_banded.h:261-280.  Calls with a pair whose band
// sqmi_align_rows_banded does not take run the reference's executor, as the
// banded scores do.
template <typename TGapsSetH, typename TGapsSetV, typename TScore, typename TLower, typename TUpper,
          std::enable_if_t<seqan_amd::detail::IsGpuRowsCall<TGapsSetH, TGapsSetV, TScore>::value &&
                           std::is_integral<std::decay_t<TLower>>::value &&
                           std::is_integral<std::decay_t<TUpper>>::value, int> = 0>
inline String<int>
globalAlignment(ExecutionPolicy<Parallel, Vectorial> const & execPolicy,
                TGapsSetH && gapsSetH, TGapsSetV && gapsSetV, TScore && scoringScheme, TLower && lowerDiag,
                TUpper && upperDiag)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_rows_banded(
    gapsSetH, gapsSetV, scoringScheme, 0u, static_cast<int64_t>(lowerDiag), static_cast<int64_t>(upperDiag), [&]() {
        return seqan2::globalAlignment<Parallel, Vectorial>(execPolicy, gapsSetH, gapsSetV, scoringScheme,
                                                            lowerDiag, upperDiag);
    }))

template <typename TGapsSetH, typename TGapsSetV, typename TScore, typename TConfig, typename TLower,
          typename TUpper,
          std::enable_if_t<seqan_amd::detail::IsGpuRowsCall<TGapsSetH, TGapsSetV, TScore>::value &&
                           seqan_amd::detail::IsAlignConfig<std::decay_t<TConfig>>::value &&
                           std::is_integral<std::decay_t<TLower>>::value &&
                           std::is_integral<std::decay_t<TUpper>>::value, int> = 0>
inline String<int>
globalAlignment(ExecutionPolicy<Parallel, Vectorial> const & execPolicy,
                TGapsSetH && gapsSetH, TGapsSetV && gapsSetV, TScore && scoringScheme, TConfig && alignConfig,
                TLower && lowerDiag, TUpper && upperDiag)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_rows_banded(
    gapsSetH, gapsSetV, scoringScheme, seqan_amd::detail::IsAlignConfig<std::decay_t<TConfig>>::mode,
    static_cast<int64_t>(lowerDiag), static_cast<int64_t>(upperDiag), [&]() {
        return seqan2::globalAlignment<Parallel, Vectorial>(execPolicy, gapsSetH, gapsSetV, scoringScheme,
                                                            alignConfig, lowerDiag, upperDiag);
    }))

template <typename TGapsSetH, typename TGapsSetV, typename TScore, typename TLower, typename TUpper,
          std::enable_if_t<seqan_amd::detail::IsGpuRowsCall<TGapsSetH, TGapsSetV, TScore>::value &&
                           std::is_integral<std::decay_t<TLower>>::value &&
                           std::is_integral<std::decay_t<TUpper>>::value, int> = 0>
inline String<int>
localAlignment(ExecutionPolicy<Parallel, Vectorial> const & execPolicy,
               TGapsSetH && gapsSetH, TGapsSetV && gapsSetV, TScore && scoringScheme, TLower && lowerDiag,
               TUpper && upperDiag)
SEQAN_AMD_OVERLOAD_BODY(seqan_amd::detail::align_rows_banded(
    gapsSetH, gapsSetV, scoringScheme, static_cast<uint32_t>(SQMI_LOCAL), static_cast<int64_t>(lowerDiag),
    static_cast<int64_t>(upperDiag), [&]() {
        return seqan2::localAlignment<Parallel, Vectorial>(execPolicy, gapsSetH, gapsSetV, scoringScheme, lowerDiag,
                                                           upperDiag);
    }))

}  // namespace seqan2
#undef SEQAN_AMD_OVERLOAD_BODY
#endif  // SEQAN_AMD_NO_SEQAN2_OVERLOADS

#endif  // SEQAN_AMD_SEQAN2_HPP
