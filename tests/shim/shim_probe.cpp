// Compile-time check of overload resolution for include/seqan_amd_seqan2.hpp.
//
// Built with -DSEQAN_AMD_SEQAN2_PROBE, which declares the library's overloads
// deleted: a call expression that resolves to one of them is ill-formed, so
// "not callable" below means "the library overload wins" and "callable" means
// "the reference template (parallel_align_interface.h:332-365) wins".  The file
// only has to compile (-fsyntax-only); every assertion is static.
#define SEQAN_AMD_SEQAN2_PROBE 1
#include "seqan_amd_seqan2.hpp"

#include <type_traits>
#include <utility>

using namespace seqan2;

namespace {

template <typename...> using void_t = void;

template <typename, typename... A> struct GlobalCallable : std::false_type {};
template <typename... A>
struct GlobalCallable<void_t<decltype(globalAlignmentScore(std::declval<A>()...))>, A...> : std::true_type {};

template <typename, typename... A> struct LocalCallable : std::false_type {};
template <typename... A>
struct LocalCallable<void_t<decltype(localAlignmentScore(std::declval<A>()...))>, A...> : std::true_type {};

template <typename, typename... A> struct RowsCallable : std::false_type {};
template <typename... A>
struct RowsCallable<void_t<decltype(globalAlignment(std::declval<A>()...))>, A...> : std::true_type {};

template <typename, typename... A> struct LocalRowsCallable : std::false_type {};
template <typename... A>
struct LocalRowsCallable<void_t<decltype(localAlignment(std::declval<A>()...))>, A...> : std::true_type {};

template <typename... A> constexpr bool gpuGlobal() { return !GlobalCallable<void, A...>::value; }
template <typename... A> constexpr bool gpuLocalRows() { return !LocalRowsCallable<void, A...>::value; }
template <typename... A> constexpr bool gpuRows() { return !RowsCallable<void, A...>::value; }
template <typename... A> constexpr bool gpuLocal() { return !LocalCallable<void, A...>::value; }

using ExPV = ExecutionPolicy<Parallel, Vectorial>;
using ExPS = ExecutionPolicy<Parallel, Serial>;
using ExSV = ExecutionPolicy<Serial, Vectorial>;
using DnaSet = StringSet<String<Dna>>;
using Dna5Set = StringSet<String<Dna5>>;
using RnaSet = StringSet<String<Rna>>;
using AaSet = StringSet<String<AminoAcid>>;
using DnaConcat = StringSet<String<Dna>, Owner<ConcatDirect<>>>;
using CharSet = StringSet<String<char>>;
using Simple = Score<int, seqan2::Simple>;
using DnaGaps = StringSet<Gaps<String<Dna>>>;
using AaGaps = StringSet<Gaps<String<AminoAcid>>>;
using Semi = AlignConfig<true, false, false, true>;

// Non-const lvalues, const lvalues, rvalues and mixtures all reach the library.
static_assert(gpuGlobal<ExPV const &, DnaSet &, DnaSet &, Simple &>(), "non-const lvalues");
static_assert(gpuGlobal<ExPV const &, DnaSet const &, DnaSet const &, Simple const &>(), "const lvalues");
static_assert(gpuGlobal<ExPV, DnaSet, DnaSet, Simple>(), "rvalues");
static_assert(gpuGlobal<ExPV &, DnaSet const &, DnaSet &, Simple>(), "mixed");
static_assert(gpuGlobal<ExPV &, DnaSet &, DnaSet &, Simple &, Semi &>(), "config, non-const");
static_assert(gpuGlobal<ExPV const &, DnaSet const &, DnaSet const &, Simple const &, Semi const &>(), "config, const");
static_assert(gpuGlobal<ExPV, DnaSet, DnaSet, Simple, Semi>(), "config, rvalue");
static_assert(gpuGlobal<ExPV, DnaSet, DnaSet, Simple, AlignConfig<>>(), "default config");
static_assert(gpuGlobal<ExPV, DnaSet, DnaSet, Simple, AlignConfig<true, true, true, true>>(), "overlap config");
static_assert(gpuGlobal<ExPV, Dna5Set &, Dna5Set &, Simple &>(), "Dna5");
static_assert(gpuGlobal<ExPV, RnaSet &, RnaSet &, Simple &>(), "Rna");
static_assert(gpuGlobal<ExPV, DnaConcat &, DnaConcat &, Simple &>(), "ConcatDirect sets");
static_assert(gpuGlobal<ExPV, AaSet &, AaSet &, Blosum62 &>(), "Blosum62");
static_assert(gpuGlobal<ExPV, AaSet const &, AaSet const &, Blosum62 const &, Semi const &>(), "Blosum62 + config");
static_assert(gpuGlobal<ExPV, AaSet &, AaSet &, Simple &>(), "protein + Simple");
static_assert(gpuLocal<ExPV const &, AaSet &, AaSet &, Blosum62 &>(), "local, non-const");
static_assert(gpuLocal<ExPV const &, AaSet const &, AaSet const &, Blosum62 const &>(), "local, const");
static_assert(gpuLocal<ExPV, AaSet, AaSet, Blosum62>(), "local, rvalues");
static_assert(gpuLocal<ExPV, DnaSet &, DnaSet &, Simple &>(), "local DNA");

// The wavefront policy (parallel_align_interface.h:367-420).
using ExWave = ExecutionPolicy<WavefrontAlignment<BlockOffsetOptimization>, Vectorial>;
using ExWaveS = ExecutionPolicy<WavefrontAlignment<>, Serial>;
static_assert(gpuGlobal<ExWave const &, DnaSet const &, DnaSet const &, Simple const &>(), "wavefront, const");
static_assert(gpuGlobal<ExWave &, DnaSet &, DnaSet &, Simple &>(), "wavefront, non-const");
static_assert(gpuGlobal<ExWave const &, DnaSet &, DnaSet &, Simple &, Semi &>(), "wavefront + config");
static_assert(gpuGlobal<ExWaveS const &, AaSet &, AaSet &, Blosum62 &>(), "wavefront, serial vectorisation");
static_assert(gpuLocal<ExWave const &, AaSet &, AaSet &, Blosum62 &>(), "wavefront local");
static_assert(std::is_same<decltype(globalAlignmentScore(std::declval<ExPS const &>(), std::declval<DnaSet &>(),
                                                         std::declval<DnaSet &>(), std::declval<Simple &>())),
                           String<int>>::value, "reference return type (Parallel)");

// Runtime-selected and other built-in matrices.
static_assert(gpuGlobal<ExPV const &, AaSet &, AaSet &, Pam120 &>(), "PAM120");
static_assert(gpuLocal<ExPV const &, AaSet &, AaSet &, SelectableAminoAcidMatrix &>(), "selectable matrix");

// globalAlignment on StringSets of Gaps (parallel_align_interface.h:288-299).
static_assert(gpuRows<ExPV const &, DnaGaps &, DnaGaps &, Simple &>(), "rows, DNA");
static_assert(gpuRows<ExPV, DnaGaps &, DnaGaps &, Simple const &>(), "rows, const score");
static_assert(gpuRows<ExPV const &, AaGaps &, AaGaps &, Blosum62 &>(), "rows, protein");
static_assert(gpuRows<ExPV const &, DnaGaps &, DnaGaps &, Simple &, Semi &>(), "rows, AlignConfig");
// localAlignment on StringSets of Gaps (parallel_align_interface.h:312-321).
static_assert(gpuLocalRows<ExPV const &, DnaGaps &, DnaGaps &, Simple &>(), "local rows, DNA");
static_assert(gpuLocalRows<ExPV, AaGaps &, AaGaps &, Blosum62 const &>(), "local rows, protein");
static_assert(gpuRows<ExPV const &, DnaGaps &, DnaGaps &, Simple &, int, int>(), "banded rows");
static_assert(gpuRows<ExPV const &, DnaGaps &, DnaGaps &, Simple &, Semi &, int, int>(), "banded rows, AlignConfig");
static_assert(gpuLocalRows<ExPV const &, AaGaps &, AaGaps &, Blosum62 &, int, int>(), "banded local rows");
static_assert(!gpuRows<ExPS const &, DnaGaps &, DnaGaps &, Simple &, int, int>(), "banded rows, Parallel, Serial");
static_assert(!gpuLocalRows<ExPS const &, AaGaps &, AaGaps &, Blosum62 &>(), "local rows, Parallel, Serial");

// Banded scores (lowerDiag, upperDiag after the score or the AlignConfig).
static_assert(gpuGlobal<ExPV const &, DnaSet &, DnaSet &, Simple &, int, int>(), "banded");
static_assert(gpuGlobal<ExPV const &, DnaSet const &, DnaSet const &, Simple const &, Semi const &, int, int>(),
              "banded + config");
static_assert(gpuGlobal<ExPV, AaSet &, AaSet &, Blosum62 &, long, long>(), "banded, long diagonals");
static_assert(gpuLocal<ExPV const &, AaSet &, AaSet &, Blosum62 &, int, int>(), "local banded");

// Everything else stays with the reference templates.
static_assert(!gpuGlobal<ExPS const &, DnaSet &, DnaSet &, Simple &, int, int>(), "banded, Parallel, Serial");
static_assert(!gpuRows<ExPS const &, DnaGaps &, DnaGaps &, Simple &>(), "rows, Parallel, Serial");
static_assert(!gpuRows<ExPV const &, DnaGaps &, DnaGaps &, Simple &, Semi &, Gotoh &>(), "rows, algorithm tag");
static_assert(!gpuGlobal<ExPS const &, DnaSet &, DnaSet &, Simple &>(), "Parallel, Serial");
static_assert(!gpuGlobal<ExSV const &, DnaSet &, DnaSet &, Simple &>(), "Serial, Vectorial");
static_assert(!gpuGlobal<ExPV const &, CharSet &, CharSet &, Simple &>(), "char alphabet");
static_assert(!gpuGlobal<ExPV const &, DnaSet &, DnaSet &, Simple &, Semi &, Gotoh &>(), "algorithm tag");
static_assert(!gpuLocal<ExPS const &, AaSet &, AaSet &, Blosum62 &>(), "local, Parallel, Serial");
static_assert(!gpuLocal<ExPV const &, CharSet &, CharSet &, Simple &>(), "local, char alphabet");
static_assert(!gpuGlobal<ExWave const &, CharSet &, CharSet &, Simple &>(), "wavefront, char alphabet");

// The traits the overloads are constrained on.
static_assert(!seqan_amd::detail::IsGpuCall<DnaSet, Dna5Set, Simple>::value, "alphabets must agree");
static_assert(!seqan_amd::detail::IsGpuCall<DnaSet, DnaSet, Blosum62>::value, "matrix alphabet must agree");
static_assert(!seqan_amd::detail::IsGpuCall<DnaSet, DnaSet, Score<short, seqan2::Simple>>::value, "int scores only");
static_assert(seqan_amd::detail::IsAlignConfig<Semi>::mode == (SQMI_FREE_TOP | SQMI_FREE_BOTTOM), "mode bits");
static_assert(seqan_amd::detail::IsAlignConfig<AlignConfig<false, true, true, false>>::mode ==
              (SQMI_FREE_LEFT | SQMI_FREE_RIGHT), "mode bits");

}  // namespace

int main() { return 0; }
