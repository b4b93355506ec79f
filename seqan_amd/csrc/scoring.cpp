// scoring.cpp -- scoring presets.
//
// BLOSUM62 (Henikoff & Henikoff 1992) in the 27-letter AminoAcid ordinal order
// of the reference (ABCDEFGHIJKLMNOPQRSTUVWYZX*, basic/alphabet_residue.h:549-556;
// table in score/score_matrix_data.h).  The reference extends the standard
// 24-letter NCBI matrix with J (round-half-away-from-zero mean of the I and L
// scores) and with O and U scoring like X.  We build the 27x27 table from the
// NCBI matrix with that rule once, at load time.
#include "seqan_amd.h"

#include <cmath>

namespace {

// NCBI BLOSUM62, letter order ARNDCQEGHILKMFPSTWYVBZX*.
const char kNcbiOrder[] = "ARNDCQEGHILKMFPSTWYVBZX*";
const signed char kNcbi[24][24] = {
    { 4,-1,-2,-2, 0,-1,-1, 0,-2,-1,-1,-1,-1,-2,-1, 1, 0,-3,-2, 0,-2,-1, 0,-4},
    {-1, 5, 0,-2,-3, 1, 0,-2, 0,-3,-2, 2,-1,-3,-2,-1,-1,-3,-2,-3,-1, 0,-1,-4},
    {-2, 0, 6, 1,-3, 0, 0, 0, 1,-3,-3, 0,-2,-3,-2, 1, 0,-4,-2,-3, 3, 0,-1,-4},
    {-2,-2, 1, 6,-3, 0, 2,-1,-1,-3,-4,-1,-3,-3,-1, 0,-1,-4,-3,-3, 4, 1,-1,-4},
    { 0,-3,-3,-3, 9,-3,-4,-3,-3,-1,-1,-3,-1,-2,-3,-1,-1,-2,-2,-1,-3,-3,-2,-4},
    {-1, 1, 0, 0,-3, 5, 2,-2, 0,-3,-2, 1, 0,-3,-1, 0,-1,-2,-1,-2, 0, 3,-1,-4},
    {-1, 0, 0, 2,-4, 2, 5,-2, 0,-3,-3, 1,-2,-3,-1, 0,-1,-3,-2,-2, 1, 4,-1,-4},
    { 0,-2, 0,-1,-3,-2,-2, 6,-2,-4,-4,-2,-3,-3,-2, 0,-2,-2,-3,-3,-1,-2,-1,-4},
    {-2, 0, 1,-1,-3, 0, 0,-2, 8,-3,-3,-1,-2,-1,-2,-1,-2,-2, 2,-3, 0, 0,-1,-4},
    {-1,-3,-3,-3,-1,-3,-3,-4,-3, 4, 2,-3, 1, 0,-3,-2,-1,-3,-1, 3,-3,-3,-1,-4},
    {-1,-2,-3,-4,-1,-2,-3,-4,-3, 2, 4,-2, 2, 0,-3,-2,-1,-2,-1, 1,-4,-3,-1,-4},
    {-1, 2, 0,-1,-3, 1, 1,-2,-1,-3,-2, 5,-1,-3,-1, 0,-1,-3,-2,-2, 0, 1,-1,-4},
    {-1,-1,-2,-3,-1, 0,-2,-3,-2, 1, 2,-1, 5, 0,-2,-1,-1,-1,-1, 1,-3,-1,-1,-4},
    {-2,-3,-3,-3,-2,-3,-3,-3,-1, 0, 0,-3, 0, 6,-4,-2,-2, 1, 3,-1,-3,-3,-1,-4},
    {-1,-2,-2,-1,-3,-1,-1,-2,-2,-3,-3,-1,-2,-4, 7,-1,-1,-4,-3,-2,-2,-1,-2,-4},
    { 1,-1, 1, 0,-1, 0, 0, 0,-1,-2,-2, 0,-1,-2,-1, 4, 1,-3,-2,-2, 0, 0, 0,-4},
    { 0,-1, 0,-1,-1,-1,-1,-2,-2,-1,-1,-1,-1,-2,-1, 1, 5,-2,-2, 0,-1,-1, 0,-4},
    {-3,-3,-4,-4,-2,-2,-3,-2,-2,-3,-2,-3,-1, 1,-4,-3,-2,11, 2,-3,-4,-3,-2,-4},
    {-2,-2,-2,-3,-2,-1,-2,-3, 2,-1,-1,-2,-1, 3,-3,-2,-2, 2, 7,-1,-3,-2,-1,-4},
    { 0,-3,-3,-3,-1,-2,-2,-3,-3, 3, 1,-2, 1,-1,-2,-2, 0,-3,-1, 4,-3,-2,-1,-4},
    {-2,-1, 3, 4,-3, 0, 1,-1, 0,-3,-4, 0,-3,-3,-2, 0,-1,-4,-3,-3, 4, 1,-1,-4},
    {-1, 0, 0, 1,-3, 3, 4,-2, 0,-3,-3, 1,-1,-3,-1, 0,-1,-3,-2,-2, 1, 4,-1,-4},
    { 0,-1,-1,-1,-2,-1,-1,-1,-1,-1,-1,-1,-1,-1,-2, 0, 0,-2,-1,-1,-1,-1,-1,-4},
    {-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4, 1},
};

// SeqAn AminoAcid ordinal order.
const char kSeqanOrder[] = "ABCDEFGHIJKLMNOPQRSTUVWYZX*";

int ncbiIndex(char c)
{
    if (c == 'O' || c == 'U')
        c = 'X';
    for (int i = 0; i < 24; ++i)
        if (kNcbiOrder[i] == c)
            return i;
    return -1;
}

// Score of two letters where J is the I/L mean (round half away from zero).
int letterScore(char a, char b)
{
    if (a == 'J' && b == 'J')
    {
        int iJ = letterScore('I', 'J'), lJ = letterScore('L', 'J');
        return static_cast<int>(std::lround((iJ + lJ) / 2.0));
    }
    if (a == 'J')
        return static_cast<int>(std::lround((letterScore('I', b) + letterScore('L', b)) / 2.0));
    if (b == 'J')
        return static_cast<int>(std::lround((letterScore(a, 'I') + letterScore(a, 'L')) / 2.0));
    return kNcbi[ncbiIndex(a)][ncbiIndex(b)];
}

struct Blosum62Table
{
    int32_t v[27 * 27];
    Blosum62Table()
    {
        for (int i = 0; i < 27; ++i)
            for (int j = 0; j < 27; ++j)
                v[i * 27 + j] = letterScore(kSeqanOrder[i], kSeqanOrder[j]);
    }
};

const Blosum62Table kBlosum62;

}  // namespace

extern "C" const int32_t *sqmi_blosum62(void)
{
    return kBlosum62.v;
}
