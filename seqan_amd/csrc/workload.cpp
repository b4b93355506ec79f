// workload.cpp -- seeded synthetic pair sets for benchmarks and tests.
//
// Each pair is generated from its own random stream, seeded by
// splitmix64(seed, pair index), so any window of pairs can be produced alone
// and in parallel.  Random streams use xoshiro256** (Blackman & Vigna).
//
//   1 DNA150         H uniform ACGT, 150 bp; V = H with 5% substitutions and
//                    1% single-base indels (half insertions, half deletions).
//   2 DNA_LOGNORMAL  |H| = clamp(round(LogNormal(ln 300, 1.0)), 100, 10000);
//                    V = H with 8% substitutions, 2% indels of geometric
//                    length (p = 0.5).
//   3 PROTEIN        |H| ~ U[50,1000], |V| ~ U[50,2000], residues drawn from
//                    the Robinson & Robinson (1991) background frequencies;
//                    30% of pairs carry, at a random offset of V, a homolog of
//                    a segment of H at ~40% identity, substitutions drawn with
//                    weight p(b) * exp(lambda * B62(a,b)), lambda = 0.3176.
//   4 DNA1K          1,000 bp, 10% substitutions, 2% single-base indels.
//   5 DNA100K        100,000 bp, 12% substitutions, 3% indels of geometric
//                    length (p = 0.3).
#include "seqan_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

struct sqmi_workload
{
    std::vector<uint8_t> h, v;
    std::vector<int64_t> hOff, vOff;
};

namespace {

struct Rng
{
    uint64_t s[4];

    static uint64_t splitmix(uint64_t & x)
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    Rng(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
        for (auto & w : s)
            w = splitmix(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // Uniform in [0, 1) with 53 random bits.
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    // Uniform integer in [lo, hi].
    int64_t range(int64_t lo, int64_t hi)
    {
        return lo + static_cast<int64_t>(next() % static_cast<uint64_t>(hi - lo + 1));
    }
    // Standard normal (Box-Muller, one draw).
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300)
            u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    // Geometric length >= 1 with success probability p.
    int64_t geometric(double p)
    {
        int64_t k = 1;
        while (uniform() >= p && k < 1000)
            ++k;
        return k;
    }
};

// Mutates H into V: per position, an indel event with probability indel,
// otherwise a copy that is substituted with probability sub.
void mutateDna(Rng & r, std::vector<uint8_t> const & h, std::vector<uint8_t> & v, double sub, double indel,
               double geomP)
{
    v.clear();
    v.reserve(h.size() + h.size() / 8 + 8);
    size_t p = 0;
    while (p < h.size())
    {
        if (r.uniform() < indel)
        {
            int64_t len = geomP >= 1.0 ? 1 : r.geometric(geomP);
            if (r.uniform() < 0.5)
            {
                p += static_cast<size_t>(len);   // deletion
                continue;
            }
            for (int64_t k = 0; k < len; ++k)    // insertion before position p
                v.push_back(static_cast<uint8_t>(r.next() & 3));
        }
        uint8_t c = h[p++];
        if (r.uniform() < sub)
            c = static_cast<uint8_t>((c + 1 + r.range(0, 2)) & 3);
        v.push_back(c);
    }
    if (v.empty())
        v.push_back(static_cast<uint8_t>(r.next() & 3));
}

// Robinson & Robinson (1991) amino-acid background frequencies (per 1000),
// keyed by SeqAn AminoAcid ordinals of the 20 standard residues.
struct Background
{
    uint8_t code[20];
    double freq[20];
    double cum[20];
    Background()
    {
        const char letters[] = "ACDEFGHIKLMNPQRSTVWY";
        const double f[] = {78.05, 19.25, 53.64, 62.95, 38.56, 73.77, 21.99, 51.42, 57.44, 90.19,
                            22.43, 44.87, 52.03, 42.64, 51.29, 71.20, 58.41, 64.41, 13.30, 32.16};
        const char order[] = "ABCDEFGHIJKLMNOPQRSTUVWYZX*";
        double total = 0;
        for (int i = 0; i < 20; ++i)
            total += f[i];
        double c = 0;
        for (int i = 0; i < 20; ++i)
        {
            code[i] = static_cast<uint8_t>(std::strchr(order, letters[i]) - order);
            freq[i] = f[i] / total;
            c += freq[i];
            cum[i] = c;
        }
        cum[19] = 1.0;
    }
    uint8_t draw(Rng & r) const
    {
        double u = r.uniform();
        int i = 0;
        while (i < 19 && u >= cum[i])
            ++i;
        return code[i];
    }
};

const Background & background()
{
    static const Background bg;
    return bg;
}

// Conditional substitution tables q(b | a) ~ p(b) exp(lambda s(a,b)), b != a.
struct Substitution
{
    double cum[27][20];
    Substitution()
    {
        const Background & bg = background();
        const int32_t * b62 = sqmi_blosum62();
        const double lambda = 0.3176;
        for (int a = 0; a < 27; ++a)
        {
            double w[20], tot = 0;
            for (int k = 0; k < 20; ++k)
            {
                int b = bg.code[k];
                w[k] = (b == a) ? 0.0 : bg.freq[k] * std::exp(lambda * b62[a * 27 + b]);
                tot += w[k];
            }
            double c = 0;
            for (int k = 0; k < 20; ++k)
            {
                c += w[k] / tot;
                cum[a][k] = c;
            }
            cum[a][19] = 1.0;
        }
    }
    uint8_t draw(Rng & r, uint8_t a) const
    {
        double u = r.uniform();
        int k = 0;
        while (k < 19 && u >= cum[a][k])
            ++k;
        return background().code[k];
    }
};

const Substitution & substitution()
{
    static const Substitution s;
    return s;
}

void makePair(int config, uint64_t seed, int64_t index, int64_t lenOverride, std::vector<uint8_t> & h,
              std::vector<uint8_t> & v)
{
    Rng r(seed, static_cast<uint64_t>(index) + 1);
    h.clear();
    switch (config)
    {
        case SQMI_WL_DNA150:
        case SQMI_WL_DNA1K:
        case SQMI_WL_DNA100K:
        {
            int64_t len = config == SQMI_WL_DNA150 ? 150 : config == SQMI_WL_DNA1K ? 1000 : 100000;
            if (lenOverride > 0)
                len = lenOverride;
            h.resize(len);
            for (auto & c : h)
                c = static_cast<uint8_t>(r.next() & 3);
            if (config == SQMI_WL_DNA150)
                mutateDna(r, h, v, 0.05, 0.01, 1.0);
            else if (config == SQMI_WL_DNA1K)
                mutateDna(r, h, v, 0.10, 0.02, 1.0);
            else
                mutateDna(r, h, v, 0.12, 0.03, 0.3);
            break;
        }
        case SQMI_WL_DNA_LOGNORMAL:
        {
            double x = std::exp(std::log(300.0) + 1.0 * r.normal());
            int64_t len = static_cast<int64_t>(std::llround(x));
            len = std::min<int64_t>(10000, std::max<int64_t>(100, len));
            if (lenOverride > 0)
                len = lenOverride;
            h.resize(len);
            for (auto & c : h)
                c = static_cast<uint8_t>(r.next() & 3);
            mutateDna(r, h, v, 0.08, 0.02, 0.5);
            break;
        }
        case SQMI_WL_PROTEIN:
        default:
        {
            const Background & bg = background();
            int64_t qlen = r.range(50, 1000), tlen = r.range(50, 2000);
            h.resize(qlen);
            v.resize(tlen);
            for (auto & c : h)
                c = bg.draw(r);
            for (auto & c : v)
                c = bg.draw(r);
            if (r.uniform() < 0.30)
            {
                int64_t maxSeg = std::min(qlen, tlen);
                int64_t seg = r.range(std::min<int64_t>(30, maxSeg), maxSeg);
                int64_t qs = r.range(0, qlen - seg), ts = r.range(0, tlen - seg);
                const Substitution & sub = substitution();
                for (int64_t k = 0; k < seg; ++k)
                {
                    uint8_t a = h[qs + k];
                    v[ts + k] = (r.uniform() < 0.40) ? a : sub.draw(r, a);
                }
            }
            break;
        }
    }
}

}  // namespace

extern "C" sqmi_status sqmi_workload_create(int config, uint64_t seed, int64_t first, int64_t count,
                                            int64_t lengthOverride, sqmi_workload ** out)
{
    if (!out || count < 0 || first < 0 || config < 1 || config > 5)
        return SQMI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    sqmi_workload * w = new (std::nothrow) sqmi_workload();
    if (!w)
        return SQMI_ERR_OUT_OF_MEMORY;
    // Generate in parallel chunks, then concatenate in index order.
    const int64_t chunk = 4096;
    const int64_t nChunks = (count + chunk - 1) / chunk;
    std::vector<std::vector<uint8_t>> hc(nChunks), vc(nChunks);
    std::vector<std::vector<int64_t>> hl(nChunks), vl(nChunks);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < nChunks; ++c)
    {
        std::vector<uint8_t> h, v;
        int64_t b = c * chunk, e = std::min(count, b + chunk);
        for (int64_t i = b; i < e; ++i)
        {
            makePair(config, seed, first + i, lengthOverride, h, v);
            hc[c].insert(hc[c].end(), h.begin(), h.end());
            vc[c].insert(vc[c].end(), v.begin(), v.end());
            hl[c].push_back(static_cast<int64_t>(h.size()));
            vl[c].push_back(static_cast<int64_t>(v.size()));
        }
    }
    w->hOff.assign(1, 0);
    w->vOff.assign(1, 0);
    w->hOff.reserve(count + 1);
    w->vOff.reserve(count + 1);
    size_t hTot = 0, vTot = 0;
    for (int64_t c = 0; c < nChunks; ++c)
    {
        hTot += hc[c].size();
        vTot += vc[c].size();
    }
    w->h.reserve(hTot);
    w->v.reserve(vTot);
    for (int64_t c = 0; c < nChunks; ++c)
    {
        w->h.insert(w->h.end(), hc[c].begin(), hc[c].end());
        w->v.insert(w->v.end(), vc[c].begin(), vc[c].end());
        for (int64_t x : hl[c])
            w->hOff.push_back(w->hOff.back() + x);
        for (int64_t x : vl[c])
            w->vOff.push_back(w->vOff.back() + x);
        std::vector<uint8_t>().swap(hc[c]);
        std::vector<uint8_t>().swap(vc[c]);
    }
    *out = w;
    return SQMI_OK;
}

extern "C" sqmi_status sqmi_workload_sizes(const sqmi_workload * w, int64_t * hTotal, int64_t * vTotal)
{
    if (!w || !hTotal || !vTotal)
        return SQMI_ERR_INVALID_ARGUMENT;
    *hTotal = static_cast<int64_t>(w->h.size());
    *vTotal = static_cast<int64_t>(w->v.size());
    return SQMI_OK;
}

extern "C" sqmi_status sqmi_workload_copy(const sqmi_workload * w, uint8_t * hRes, int64_t * hOff, uint8_t * vRes,
                                          int64_t * vOff)
{
    if (!w || !hRes || !hOff || !vRes || !vOff)
        return SQMI_ERR_INVALID_ARGUMENT;
    std::memcpy(hRes, w->h.data(), w->h.size());
    std::memcpy(vRes, w->v.data(), w->v.size());
    std::memcpy(hOff, w->hOff.data(), w->hOff.size() * sizeof(int64_t));
    std::memcpy(vOff, w->vOff.data(), w->vOff.size() * sizeof(int64_t));
    return SQMI_OK;
}

extern "C" sqmi_status sqmi_workload_destroy(sqmi_workload * w)
{
    delete w;
    return SQMI_OK;
}
